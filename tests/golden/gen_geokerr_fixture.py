"""Copy the reference's one stored Fortran geokerr_eval_lambda output into
tests/golden/ as an .npz.

Source: geokerr_port/validation/cache/geokerr/bfc7797cdaa2b7bb.bin, the
on-disk cache gkval_execute_geokerr_fortran writes
(geokerr_port/validation/gkval_adapters.c:505-620): a header {u32 magic
'GKCH', u32 version, u64 S, f64 lam0, f64 dlam, gkval_inputs_t {u0, mu0, uf,
a, lam(=Lz), q2, t0}, i32 tpmi, i32 tpri} followed by the fields u, mu, t,
phi, affine as S doubles each.  The fail-window files under
validation/runs/*/fail_windows/1/ hold 32-sample slices of the same u array.
"""
import os
import struct
import sys

import numpy as np

SRC = "/root/reference/geokerr_port/validation/cache/geokerr/bfc7797cdaa2b7bb.bin"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "geokerr_eval_lambda_fortran.npz")


def main() -> int:
    with open(SRC, "rb") as f:
        raw = f.read()
    magic, version, s = struct.unpack_from("<IIQ", raw, 0)
    if magic != 0x474B4348 or version != 1:
        print("unexpected header", hex(magic), version)
        return 1
    lam0, dlam = struct.unpack_from("<dd", raw, 16)
    inputs = struct.unpack_from("<7d", raw, 32)
    tpmi, tpri = struct.unpack_from("<ii", raw, 88)
    fields = np.frombuffer(raw, dtype="<f8", offset=96, count=5 * s).reshape(5, s)
    np.savez_compressed(DST, inputs=np.array(inputs), lam0=lam0, dlam=dlam, s=s,
                        tpmi=tpmi, tpri=tpri, u=fields[0], mu=fields[1], t=fields[2],
                        phi=fields[3], affine=fields[4])
    print("wrote", DST, "S =", s, "inputs (u0, mu0, uf, a, Lz, Q2, t0) =", inputs)
    return 0


if __name__ == "__main__":
    sys.exit(main())
