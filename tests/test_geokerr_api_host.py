"""Host-side behaviour of the geokerr_ref_api and port-solver drop-ins in
libnullpath.so that needs no GPU: argument checks and geokerr_phi_omega (a
host closed form) against the oracle."""
import ctypes as C

import numpy as np

import nullpath_amd as npa


def test_eval_lambda_rejects_nonpositive_samples():
    # returns before touching the device, as the reference does for S <= 0
    r = npa.geokerr_eval_lambda(0.02, 0.3, 0.5, 0.9, 6.0, 10.0, 0.0, 0.0, 0.1, 0)
    assert r["info"] == 1
    assert npa.lib().np_status_string(-5) == b"entry point disabled in this build"


def test_port_solver_rejects_bad_arguments():
    # -1 before any device call, as geokerr_complete.cu:451-453 returns
    pd = C.POINTER(C.c_double)
    buf = C.cast((C.c_double * 4)(), pd)
    null = C.cast(None, pd)
    for sym in ("geokerr_complete_stream_cuda", "geokerr_stream_cuda"):
        code, _ = npa.geokerr_port_stream(0.02, 0.5, 0.3, 0.9, 6.0, 10.0, 0.0, 0.1, 0, symbol=sym)
        assert code == -1
        fn = getattr(npa.lib(), sym)
        assert fn(0.02, 0.5, 0.3, 0.9, 6.0, 10.0, 0.0, 0.1, 4, buf, null, buf, buf, buf) == -1
        assert fn(0.02, 0.5, 0.3, 0.9, 6.0, 10.0, 0.0, 0.1, -3, buf, buf, buf, buf, buf) == -1


def test_geor_batch_argument_checks():
    # rejected on the host before any device call; n = 0 is a no-op
    L = npa.lib()
    assert L.np_geor_batch(*([None] * 10), 1, *([None] * 5)) == -1
    assert L.np_geor_batch(*([None] * 10), -1, *([None] * 5)) == -1
    assert L.np_geor_batch(*([None] * 10), 0, *([None] * 5)) == 0


def test_geomu_camera_batch_mixed_argument_checks():
    L = npa.lib()
    assert L.np_geomu_camera_batch_mixed(*([None] * 5), 1e-4, 1, *([None] * 6)) == -1
    assert L.np_geomu_camera_batch_mixed(*([None] * 5), 1e-4, -1, *([None] * 6)) == -1
    assert L.np_geomu_camera_batch_mixed(*([None] * 5), 1e-4, 0, *([None] * 6)) == 0


def test_geokerr_batch_argument_checks():
    # the whole-driver entry point checks its switches and pointers on the
    # host; n = 0 with valid switches is a no-op
    from nullpath_amd._lib import GeokerrOpts, GeokerrRays, GeokerrSamples
    L = npa.lib()
    r, s = GeokerrRays(), GeokerrSamples()
    ok = GeokerrOpts(nup=8, npts=4, kext=1, phit=1, usegeor=0, mufill=1, offset=0.5)
    assert L.np_geokerr_batch(None, C.byref(r), 0, C.byref(s), None) == -1
    assert L.np_geokerr_batch(C.byref(ok), None, 0, C.byref(s), None) == -1
    assert L.np_geokerr_batch(C.byref(ok), C.byref(r), -1, C.byref(s), None) == -1
    assert L.np_geokerr_batch(C.byref(ok), C.byref(r), 0, C.byref(s), None) == 0
    assert L.np_geokerr_batch(C.byref(ok), C.byref(r), 5, C.byref(s), None) == -1   # NULL arrays
    for bad in (dict(nup=0), dict(npts=-1), dict(kext=-1), dict(nup=1, npts=0, kext=1)):
        o = GeokerrOpts(nup=8, npts=4, kext=1, phit=1, usegeor=0, mufill=1, offset=0.5)
        for k, v in bad.items():
            setattr(o, k, v)
        assert L.np_geokerr_batch(C.byref(o), C.byref(r), 0, C.byref(s), None) == -1, bad


def test_phi_omega_host_matches_oracle(oracle):
    rng = np.random.default_rng(17)
    for _ in range(500):
        args = (10 ** rng.uniform(-4, -0.4), rng.uniform(-0.99, 0.99), rng.uniform(-0.99, 0.99),
                rng.uniform(-8, 8), rng.uniform(0, 30))
        got = npa.geokerr_phi_omega(*args)
        want = oracle.geokerr_phi_omega(*args)
        assert abs(got - want) <= 1e-13 * max(1.0, abs(want)), (args, got, want)
    # the floors: u0 = 0 puts r at 1e12, mu0 = 1 floors 1 - mu0^2 at 1e-30
    assert npa.geokerr_phi_omega(0.0, 0.2, 0.5, 2.0, 1.0) == oracle.geokerr_phi_omega(0.0, 0.2, 0.5, 2.0, 1.0)
    assert npa.geokerr_phi_omega(0.1, 1.0, 0.5, 2.0, 1.0) == oracle.geokerr_phi_omega(0.1, 1.0, 0.5, 2.0, 1.0)
