"""GPU parity of GEOR (np_geor_batch) against the CPU oracle, and the
GEOMU -> GEOR round trip run entirely on the GPU."""
import numpy as np
import pytest
import torch

import nullpath_amd as npa
from test_oracle_geor import physical_cases


def _t(a, dtype=np.float64):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to("cuda")


def _inputs(oracle, n, seed):
    """Physical starts with GEOMU's (muf, tpm) from the oracle, plus the
    equatorial-crossing requests muf = 0 after 0 and 1 mu turning points."""
    rows = []
    for c in physical_cases(n, seed):
        kind, u0, uf, mu0, a, l, q2, tpr, su, sm = c
        g = oracle.geomu(u0, uf, mu0, a, l, q2, tpr, su, sm)
        u0 = g["u0_used"]
        rows += [(u0, mu0, g["muf"], a, l, q2, g["tpm"], su, sm), (u0, mu0, 0.0, a, l, q2, 0, su, sm),
                 (u0, mu0, 0.0, a, l, q2, 1, su, sm)]
    return [np.array(col) for col in zip(*rows)]


@pytest.mark.gpu
def test_geor_batch_matches_oracle(oracle, device):
    u0, mu0, muf, a, l, q2, tpm, su, sm = _inputs(oracle, 1500, 41)
    n = u0.size
    print(f"[gpu] np_geor_batch n={n}", flush=True)
    g_uf, g_imu, g_tpr, g_nc = (t.cpu().numpy() for t in npa.geor_batch(
        _t(u0), _t(mu0), _t(muf), _t(a), _t(l), _t(q2), _t(su), _t(sm), _t(tpm, np.int32)))
    o_uf, o_imu, o_tpr, o_nc = oracle.geor_batch(u0, mu0, muf, a, l, q2, tpm, su, sm)
    assert np.array_equal(g_nc, o_nc)
    assert np.array_equal(g_tpr, o_tpr)
    assert np.array_equal(np.isnan(g_uf), np.isnan(o_uf))
    assert np.array_equal(g_uf == -1.0, o_uf == -1.0)
    ok = np.isfinite(o_uf) & (o_uf != -1.0)
    assert ok.sum() > n // 2
    rel = np.abs(g_uf[ok] - o_uf[ok]) / np.abs(o_uf[ok])
    assert rel.max() <= 1e-9, (rel.max(), int(np.argmax(rel)))
    fin = np.isfinite(o_imu)
    assert np.all(np.abs(g_imu[fin] - o_imu[fin]) <= 1e-11 * np.maximum(1.0, np.abs(o_imu[fin])))
    # the equatorial requests cover both rows that cross the equator and
    # rows that cannot reach it
    eq = muf == 0.0
    assert (g_uf[eq] > 0).any() and (g_uf[eq] == -1.0).any()


@pytest.mark.gpu
def test_geomu_geor_round_trip_on_gpu(device):
    cases = physical_cases(2000, 43, ("general", "a0"))
    cols = [np.array(c) for c in zip(*cases)]
    kind, u0, uf, mu0, a, l, q2, tpr, su, sm = cols
    print(f"[gpu] np_geomu_batch -> np_geor_batch n={u0.size}", flush=True)
    muf, tpm, ncase, iu = npa.geomu_batch(_t(u0), _t(uf), _t(mu0), _t(a), _t(l), _t(q2), _t(su), _t(sm),
                                          _t(tpr, np.int32))
    r_uf, r_imu, r_tpr, r_nc = npa.geor_batch(_t(u0), _t(mu0), muf, _t(a), _t(l), _t(q2), _t(su), _t(sm), tpm)
    r_uf, r_nc, ncase = r_uf.cpu().numpy(), r_nc.cpu().numpy(), ncase.cpu().numpy()
    assert np.array_equal(r_nc, ncase)
    # GEOMU stops at a radial turning point it would pass (tpr = 0 rays end
    # there at most): compare with uf where the ray does not reach it
    rel = np.abs(r_uf - uf) / uf
    assert np.median(rel) < 1e-12 and (rel < 1e-8).mean() > 0.99, (np.median(rel), (rel < 1e-8).mean())
