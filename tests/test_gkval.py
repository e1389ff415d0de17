"""The acceptance tool tests/gkval.py on the CPU: its ULP distance against
counting np.nextafter steps, its quantization against exact rounding of the
mantissa, the Merkle tree's shape, and whole runs with the oracle as both
solvers -- one run with a mismatch planted in the port's output (found at
the planted sample, with its fail window), one interrupted and resumed
(the same ledger as an uninterrupted run), one on the adaptive grid."""
import json
import math
import os

import numpy as np
import pytest

import gkval as G


def test_ulp_distance_counts_nextafter_steps():
    rng = np.random.default_rng(3)
    for _ in range(200):
        x = rng.normal() * 10.0 ** rng.integers(-300, 300)
        k = int(rng.integers(0, 40))
        y = x
        for _ in range(k):
            y = np.nextafter(y, np.inf)
        assert int(G.ulp_distance(x, y)) == k and int(G.ulp_distance(y, x)) == k
    # across zero: -tiny, -0, +0, +tiny are consecutive; equal values are 0 apart
    tiny = np.nextafter(0.0, 1.0)
    assert int(G.ulp_distance(-tiny, tiny)) == 3
    assert int(G.ulp_distance(-0.0, 0.0)) == 0
    big = np.iinfo(np.uint64).max
    assert int(G.ulp_distance(np.nan, 1.0)) == big and int(G.ulp_distance(np.inf, -np.inf)) == big
    assert int(G.ulp_distance(np.inf, np.inf)) == 0


def test_accept_needs_both_tolerance_and_ulps():
    a = np.array([1.0, 1.0, 1.0, 0.0])
    b = np.array([1.0 + 2.2e-16, 1.0 + 1e-12, 1.0 + 4.4e-16, 1e-300])
    ok = G.accept(a, b, atol=1e-12, rtol=1e-12, max_ulps=1)
    assert ok.tolist() == [True, False, False, False]


def test_quantize_rounds_mantissa_to_nearest_even():
    from fractions import Fraction
    rng = np.random.default_rng(5)
    x = rng.normal(size=300) * 10.0 ** rng.integers(-20, 20, 300)
    for lsb in (1, 8, 20, 52):
        q = G.quantize(x, lsb).view(np.float64)
        unit = [Fraction(float(np.spacing(abs(v)))) * 2 ** lsb for v in x]
        for xi, qi, u in zip(x, q, unit):
            bits = int(np.float64(qi).view(np.uint64))
            assert bits & ((1 << lsb) - 1) == 0
            # within half a kept unit (a carry into the exponent doubles the unit)
            assert abs(Fraction(float(qi)) - Fraction(float(xi))) <= u
    # ties go to the even kept mantissa
    one = np.float64(1.0).view(np.uint64)
    tie_even = np.array([one + np.uint64(0x80)], dtype=np.uint64).view(np.float64)    # ...0 1000 0000
    tie_odd = np.array([one + np.uint64(0x180)], dtype=np.uint64).view(np.float64)    # ...1 1000 0000
    assert int(G.quantize(tie_even, 8)[0]) == int(one)
    assert int(G.quantize(tie_odd, 8)[0]) == int(one) + 0x200
    with pytest.raises(ValueError):
        G.quantize(np.array([np.nan]), 8)


def test_merkle_shape():
    h = [bytes([i]) * 32 for i in range(5)]
    n = lambda x, y: G._h(b"node", x, y)
    assert G.merkle(h[:1]) == h[0]
    assert G.merkle(h[:2]) == n(h[0], h[1])
    assert G.merkle(h[:3]) == n(n(h[0], h[1]), h[2])
    assert G.merkle(h) == n(n(n(h[0], h[1]), n(h[2], h[3])), h[4])


def _args(run_dir, **kw):
    argv = ["--run-dir", str(run_dir), "--cases", "8", "--size", "2048", "--delta", "0.4", "--batch", "2",
            "--atol", "1e-12", "--rtol", "1e-12", "--max-ulps", "2"]
    for k, v in kw.items():
        argv += [f"--{k.replace('_', '-')}"] + ([] if v is True else [str(v)])
    return G.parse(argv)


@pytest.fixture(scope="module")
def solvers(oracle):
    ref, periods, omega = G._oracle_solvers()
    return ref, periods, omega


def _ledger(run_dir):
    with open(os.path.join(run_dir, "ledger.ndjson")) as f:
        return [json.loads(line) for line in f]


def test_run_finds_planted_mismatch(tmp_path, solvers):
    ref, _, _ = solvers

    def port(c, lam0, dlam, s):
        out = {k: (v.copy() if hasattr(v, "copy") else v) for k, v in ref(c, lam0, dlam, s).items()}
        if c.shape[0] and np.any(np.all(c == G.make_cases(8, 1)[3], axis=1)):
            i = int(np.nonzero(np.all(c == G.make_cases(8, 1)[3], axis=1))[0][0])
            out["t"][i, 1500] *= 1.0 + 1e-9
        return out

    code = G.run(_args(tmp_path), ref, port)
    assert code == 1
    led = _ledger(tmp_path)
    assert [r["case_id"] for r in led] == list(range(8))
    bad = [r for r in led if not r["success"]]
    assert len(bad) == 1 and bad[0]["case_id"] == 3
    assert bad[0]["first_fail_field"] == "t" and bad[0]["first_fail_index"] == 1500
    w = G.read_fail_window(os.path.join(tmp_path, bad[0]["fail_window"]))
    assert w["field"] == "t" and w["start"] <= 1500 < w["end"] and w["end"] - w["start"] == G.WINDOW
    k = 1500 - w["start"]
    assert w["port"][k] == w["ref"][k] * (1.0 + 1e-9) and np.array_equal(np.delete(w["port"], k), np.delete(w["ref"], k))
    good = [r for r in led if r["success"]]
    assert all(len(r["case_root"]) == 64 and set(r["field_roots"]) == set(G.FIELDS) for r in good)
    with open(os.path.join(tmp_path, "cursor.json")) as f:
        assert json.load(f)["case_id"] == 8


def test_resume_gives_the_uninterrupted_ledger(tmp_path, solvers):
    ref, _, _ = solvers
    full, part = tmp_path / "full", tmp_path / "part"
    assert G.run(_args(full), ref, ref) == 0
    assert G.run(_args(part, stop_after=4), ref, ref) == 0
    assert len(_ledger(part)) == 4
    # a resume with other settings is refused
    assert G.run(_args(part, resume=True, max_ulps=3), ref, ref) == 2
    assert G.run(_args(part, resume=True), ref, ref) == 0
    assert _ledger(part) == _ledger(full)


def test_adaptive_grid(tmp_path, solvers, oracle):
    ref, periods, omega = solvers
    assert G.run(_args(tmp_path, grid="adaptive", periods=1.5, cases=4), ref, ref, periods, omega) == 0
    led = _ledger(tmp_path)
    cases = G.make_cases(4, 1)
    for r, c in zip(led, cases):
        tr, tth, info = oracle.geokerr_estimate_periods(c[0], c[1], c[3], c[4], c[5])
        dlam = float(np.uint64(r["grid"]["dlam_bits"]).view(np.float64))
        # generic cases (Q2 > 0, u0 != uf): lambda_max = K max(Tr, Ttheta), or K
        # when neither period was found (both -1)
        tchar = max(tr, tth) if info == 0 and max(tr, tth) > 0 else 1.0
        assert math.isclose(dlam * 2047, 1.5 * tchar, rel_tol=1e-12)
