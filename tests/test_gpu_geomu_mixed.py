"""np_geomu_camera_batch_mixed -- GEOMU with the quartic's roots from a
float deflation polished in double, FP64 root finding for near-degenerate
arguments -- against the FP64 kernel (np_geomu_camera_batch, itself pinned
to the oracle in test_gpu_geokerr) on the GEOKERR batch distribution
(alpha, beta ~ U[-12, 12] M, a ~ U[0, 0.998], mu0 ~ U[0.05, 0.99],
u ~ U(0, u+), from u0 = 1e-4).

The polished roots are FP64 roots; they may differ from the FP64 path's in
the last bits, so the Table-1 case and mu turning points must agree and
mu_f agree to 1e-10 (absolute; |mu| <= 1)."""
import numpy as np
import pytest
import torch

import nullpath_amd as npa


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    t = lambda x: torch.tensor(x, dtype=torch.float64, device="cuda")
    a = rng.uniform(0.0, 0.998, n)
    uplus = 1.0 / (1.0 + np.sqrt(1.0 - a * a))
    return [t(rng.uniform(-12, 12, n)), t(rng.uniform(-12, 12, n)), t(a), t(rng.uniform(0.05, 0.99, n)),
            t(rng.uniform(0.0, 1.0, n) * uplus)]


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [11, 12])
def test_mixed_roots_match_fp64_geomu(device, seed):
    n = 1 << 18
    args = _rays(n, seed)
    print(f"[gpu] np_geomu_camera_batch(_mixed) n={n} seed={seed}", flush=True)
    m64, t64, c64, i64 = [x.cpu().numpy() for x in npa.geomu_camera_batch(*args, 1e-4)]
    mm, tm, cm, im, fb = [x.cpu().numpy() for x in npa.geomu_camera_batch_mixed(*args, 1e-4)]
    same_case = (c64 == cm) & (t64 == tm)
    dm = np.abs(mm - m64)
    close = same_case & ((dm <= 1e-10) | (np.isnan(mm) & np.isnan(m64)))
    print(f"[gpu] mixed vs fp64: fp64-root fallbacks {fb.mean():.4%}, case/tpm differ {np.mean(~same_case):.2e}, "
          f"|dmu| median {np.nanmedian(dm):.2e} max {np.nanmax(dm):.2e}, outside 1e-10 {np.mean(~close):.2e}",
          flush=True)
    assert np.mean(same_case) >= 0.9999
    assert np.mean(close) >= 0.9999
    assert set(np.unique(fb)) <= {0, 1}
    # the fallback rays ran the FP64 root finder: identical results there
    sel = fb == 1
    assert np.array_equal(c64[sel], cm[sel])
