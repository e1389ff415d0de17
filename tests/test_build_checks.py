"""Build-time invariants of the gfx950 code: no kernel calls a device function
(every function in every code object has a kernel descriptor)."""
import os

import pytest

from nullpath_amd import build as nb


def test_code_objects_hold_kernels_only():
    if not all(os.path.exists(os.path.join(nb.OBJ_DIR, s + ".o")) for s in nb.LIB_SOURCES):
        pytest.fail("library objects missing: run __graft_entry__.build() first")
    nb.check_no_device_calls()
    kernels, others = nb.device_functions(os.path.join(nb.OBJ_DIR, "kerr_render.hip.o"))
    assert others == [] and len(kernels) >= 16
    assert any("kerr_render_capped" in k for k in kernels)
