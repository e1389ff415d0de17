"""gkval -- acceptance of the GPU GEOKERR (np_geokerr_eval_lambda_batch) against
the oracle's FP64 restatement at scale, in the manner of the reference's
validation harness (geokerr_port/validation/gkval_core.c:262-295 acceptance,
gkval_validator.c:227-417 per-case loop, gkval_io.c cursor / ledger / fail
windows, gkval_adapters.c:96-137 adaptive lambda grid).  Test infrastructure:
it drives the oracle, so it lives with the tests.

For every case (u0, mu0, uf, a, L, Q2, t0) both solvers resample the
geodesic onto lambda_k = k dlam, k < S.  Each field (u, mu, t, phi, affine;
phi unwrapped first) is walked in segments of 1024 samples; a sample passes
when |a - b| <= atol + rtol max(|a|, |b|) AND its distance in ULPs is at most
max_ulps.  The first failing sample ends the case: a window of 32 samples
around it goes to fail_windows/<case>/<field>_<start>_<end>.bin.  Passing
segments of the reference are quantized (round to nearest even at mantissa
bit quant_lsb) and hashed (BLAKE2b-256 with "leaf"/"node" domain prefixes)
into a Merkle root per field and a case root.  A case also fails on a NaN in
either solver or an affine drift beyond atol + rtol.  One NDJSON line per case
goes to ledger.ndjson; cursor.json (written atomically) names the next case,
so --resume continues an interrupted run.

Run directory layout: manifest.json, ledger.ndjson, cursor.json,
fail_windows/.  Exit status 0 when every case passed, 1 when any failed, 2 on
a usage error (a manifest that does not match on --resume).

    python tests/gkval.py --run-dir DIR [--cases N --seed S] [--size S]
        [--grid uniform|adaptive] [--delta D | --periods K]
        [--atol A --rtol R --max-ulps U --quant-lsb Q | --profile bitexact]
        [--fields u,mu,t,phi,affine] [--batch B] [--resume]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import math
import os
import struct
import sys

import numpy as np

FIELDS = ("u", "mu", "t", "phi", "affine")
SEG = 1024
WINDOW = 32
FAIL_MAGIC = 0x474B5057   # "GKPW"
FAIL_VERSION = 1
VERSION = "1.0"


# ---------------- acceptance ----------------
def ordered_bits(x: np.ndarray) -> np.ndarray:
    """IEEE doubles mapped to unsigned integers in the doubles' order, so
    that integer differences are ULP distances."""
    u = np.asarray(x, dtype=np.float64).view(np.uint64)
    sign = np.uint64(1) << np.uint64(63)
    return np.where(u & sign, ~u, u ^ sign)


def ulp_distance(a, b) -> np.ndarray:
    """ULP distance of two double arrays; NaN, or different infinities,
    count as the largest distance."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    oa, ob = ordered_bits(a), ordered_bits(b)
    d = np.where(oa > ob, oa - ob, ob - oa)
    big = np.uint64(np.iinfo(np.uint64).max)
    bad = np.isnan(a) | np.isnan(b) | ((np.isinf(a) | np.isinf(b)) & (a != b))
    d = np.where(bad, big, d)
    return np.where(a == b, np.uint64(0), d)


def accept(a, b, atol: float, rtol: float, max_ulps: int) -> np.ndarray:
    """Per-sample acceptance: within atol + rtol max(|a|, |b|) and within
    max_ulps ULPs."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    with np.errstate(invalid="ignore"):
        ok = np.abs(a - b) <= atol + rtol * np.maximum(np.abs(a), np.abs(b))
    return ok & (ulp_distance(a, b) <= np.uint64(max_ulps))


def unwrap(phi: np.ndarray) -> np.ndarray:
    """Sequential unwrapping as the harness does it: each sample moved by
    2 pi when it jumps more than pi from its (already unwrapped) predecessor."""
    out = np.array(phi, dtype=np.float64, copy=True)
    for i in range(1, out.size):
        d = out[i] - out[i - 1]
        if d > math.pi:
            out[i] -= 2.0 * math.pi
        elif d < -math.pi:
            out[i] += 2.0 * math.pi
    return out


# ---------------- quantization and hashing ----------------
def quantize(x: np.ndarray, lsb: int) -> np.ndarray:
    """Round the mantissa to nearest even at bit `lsb` (0 = the bits
    themselves); the exponent carries when the mantissa overflows."""
    if not 0 <= lsb <= 52:
        raise ValueError("quant_lsb must be in 0..52")
    bits = np.ascontiguousarray(x, dtype=np.float64).view(np.uint64).copy()
    if np.any((bits >> np.uint64(52)) & np.uint64(0x7FF) == np.uint64(0x7FF)):
        raise ValueError("NaN or Inf cannot be quantized")
    if lsb == 0:
        return bits
    one = np.uint64(1)
    mask = (one << np.uint64(lsb)) - one
    half = one << np.uint64(lsb - 1)
    tail = bits & mask
    base = bits & ~mask
    up = (tail > half) | ((tail == half) & (((base >> np.uint64(lsb)) & one) == one))
    # adding one unit of the kept part carries into the exponent as needed
    return np.where(up, base + (one << np.uint64(lsb)), base)


def _h(*parts: bytes) -> bytes:
    h = hashlib.blake2b(digest_size=32)
    for p in parts:
        h.update(p)
    return h.digest()


def leaf_hash(run_id: str, case_id: int, field: int, start: int, end: int, q: np.ndarray) -> bytes:
    return _h(b"leaf", run_id.encode(), struct.pack("<QIQQ", case_id, field, start, end),
              np.ascontiguousarray(q, dtype="<u8").tobytes())


def merkle(hashes: list[bytes]) -> bytes:
    """Pairwise "node" hashes level by level; an odd last node moves up."""
    if not hashes:
        return bytes(32)
    level = list(hashes)
    while len(level) > 1:
        level = [_h(b"node", level[i], level[i + 1]) if i + 1 < len(level) else level[i]
                 for i in range(0, len(level), 2)]
    return level[0]


# ---------------- files ----------------
def write_atomic(path: str, text: str) -> None:
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        f.write(text)
        f.flush()
        os.fsync(f.fileno())
    os.replace(tmp, path)


def write_fail_window(run_dir: str, case_id: int, field: int, start: int, ref: np.ndarray,
                      port: np.ndarray) -> str:
    d = os.path.join(run_dir, "fail_windows", str(case_id))
    os.makedirs(d, exist_ok=True)
    end = start + ref.size
    path = os.path.join(d, f"{FIELDS[field]}_{start}_{end}.bin")
    with open(path, "wb") as f:
        f.write(struct.pack("<IIHQQQ", FAIL_MAGIC, FAIL_VERSION, field, start, end, ref.size))
        f.write(np.ascontiguousarray(ref, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(port, dtype="<f8").tobytes())
    return path


def read_fail_window(path: str):
    with open(path, "rb") as f:
        raw = f.read()
    magic, version, field, start, end, size = struct.unpack_from("<IIHQQQ", raw, 0)
    if magic != FAIL_MAGIC or version != FAIL_VERSION:
        raise ValueError("not a fail window")
    off = struct.calcsize("<IIHQQQ")
    data = np.frombuffer(raw, dtype="<f8", offset=off, count=2 * size)
    return {"field": FIELDS[field], "start": start, "end": end, "ref": data[:size], "port": data[size:]}


# ---------------- cases and grids ----------------
def make_cases(n: int, seed: int) -> np.ndarray:
    """Seeded geodesics: camera-like starts u0 ~ U[1e-3, 0.02], targets
    uf ~ U(0.1, 0.9), a ~ U[0, 0.99], L ~ U[-7, 7], Q2 ~ U[0.5, 40], mu0
    inside the allowed band of M(mu), t0 = 0.  Columns u0, mu0, uf, a, L, Q2, t0."""
    rng = np.random.default_rng(seed)
    out = np.zeros((n, 7))
    i = 0
    while i < n:
        a = rng.uniform(0.0, 0.99)
        l = rng.uniform(-7.0, 7.0)
        q2 = rng.uniform(0.5, 40.0)
        mu0 = rng.uniform(-0.95, 0.95)
        if q2 + (a * a - q2 - l * l) * mu0 ** 2 - a * a * mu0 ** 4 < 0.0:
            continue
        out[i] = (rng.uniform(1e-3, 0.02), mu0, rng.uniform(0.1, 0.9), a, l, q2, 0.0)
        i += 1
    return out


def adaptive_dlam(case, size: int, k_periods: float, periods, omega) -> float:
    """The harness's adaptive grid (gkval_adapters.c:96-137): lambda_max =
    K T_char with T_char the radial period for an equatorial orbit, the
    polar one for u0 = uf, 2 pi / |Omega_phi| for both (as ordered there,
    that branch follows the first two), else the larger; dlam =
    lambda_max / (S - 1).  `periods` = (Tr, Ttheta, info)."""
    u0, mu0, uf, a, l, q2, _ = case
    tr, tth, info = periods
    if info != 0:
        tr, tth = 1.0, 1.0
    eps = 1e-14
    if abs(q2) < eps and tr > 0:
        tchar = tr
    elif abs(uf - u0) < eps and tth > 0:
        tchar = tth
    elif abs(q2) < eps and abs(uf - u0) < eps:
        w = omega
        tchar = 2.0 * math.pi / abs(w if abs(w) >= eps else 1.0)
    else:
        tchar = max(tr, tth)
    if tchar <= 0.0:
        tchar = 1.0
    return k_periods * tchar / (size - 1)


# ---------------- the run ----------------
def validate_case(run_id, case_id, ref, port, fields, tol, quant_lsb, run_dir):
    """Compare one case's fields; returns the ledger record."""
    atol, rtol, max_ulps = tol
    rec = {"run_id": run_id, "case_id": case_id, "success": True}
    roots = []
    nan_count = 0
    lam_drift = 0.0
    for fi, name in enumerate(FIELDS):
        if name not in fields:
            continue
        a = np.asarray(ref[name], dtype=np.float64)
        b = np.asarray(port[name], dtype=np.float64)
        if name == "phi":
            a, b = unwrap(a), unwrap(b)
        nan_count += int(np.sum(~np.isfinite(a) | ~np.isfinite(b)))
        if name == "affine":
            with np.errstate(invalid="ignore"):
                lam_drift = max(lam_drift, float(np.nanmax(np.abs(a - b), initial=0.0)))
        ok = accept(a, b, atol, rtol, max_ulps)
        leaves = []
        for start in range(0, a.size, SEG):
            seg = slice(start, min(start + SEG, a.size))
            bad = np.nonzero(~ok[seg])[0]
            if bad.size:
                j = start + int(bad[0])
                w0 = max(0, j - WINDOW // 2)
                w1 = min(a.size, w0 + WINDOW)
                w0 = max(0, w1 - WINDOW)
                rec.update(success=False, first_fail_field=name, first_fail_index=j,
                           ulps=int(min(int(ulp_distance(a[j], b[j])), 2 ** 63 - 1)),
                           ref_value=float(a[j]), port_value=float(b[j]),
                           fail_window=os.path.relpath(write_fail_window(run_dir, case_id, fi, w0, a[w0:w1],
                                                                         b[w0:w1]), run_dir))
                return rec
            leaves.append(leaf_hash(run_id, case_id, fi, seg.start, seg.stop, quantize(a[seg], quant_lsb)))
        rec.setdefault("field_roots", {})[name] = merkle(leaves).hex()
        roots.append(merkle(leaves))
    rec["case_root"] = merkle(roots).hex()
    rec["nan_count"] = nan_count
    rec["max_lambda_drift"] = lam_drift
    if nan_count or lam_drift > atol + rtol:
        rec["success"] = False
        rec["invariant_failure"] = "nan" if nan_count else "lambda_drift"
    return rec


def run(args, ref_solver, port_solver, periods_fn=None, omega_fn=None) -> int:
    """The validation loop over batches of cases.  ref_solver / port_solver:
    callables (cases[n, 7], lam0[n], dlam[n], S) -> dict of [n, S] arrays
    u, mu, t, phi, affine and [n] tpmi, tpri."""
    os.makedirs(os.path.join(args.run_dir, "fail_windows"), exist_ok=True)
    tol = (0.0, 0.0, 0) if args.profile == "bitexact" else (args.atol, args.rtol, args.max_ulps)
    qlsb = 0 if args.profile == "bitexact" else args.quant_lsb
    fields = tuple(f for f in args.fields.split(",") if f)
    manifest = {"version": VERSION, "run_id": args.run_id,
                "grid": {"type": args.grid, "S": args.size, "delta": args.delta, "periods": args.periods},
                "tolerances": {"atol": tol[0], "rtol": tol[1], "max_ulps": tol[2]}, "quant_lsb": qlsb,
                "fields": list(fields), "cases": args.cases, "seed": args.seed}
    mpath = os.path.join(args.run_dir, "manifest.json")
    cpath = os.path.join(args.run_dir, "cursor.json")
    lpath = os.path.join(args.run_dir, "ledger.ndjson")
    first = 0
    if args.resume and os.path.exists(mpath):
        with open(mpath) as f:
            if json.load(f) != manifest:
                print("gkval: manifest differs from the run being resumed", file=sys.stderr)
                return 2
        if os.path.exists(cpath):
            with open(cpath) as f:
                first = int(json.load(f)["case_id"])
    else:
        write_atomic(mpath, json.dumps(manifest, indent=2) + "\n")
        for p in (cpath, lpath):
            if os.path.exists(p):
                os.remove(p)
    cases = make_cases(args.cases, args.seed)
    failed = 0
    with open(lpath, "a") as ledger:
        for b0 in range(first, args.cases, args.batch):
            c = cases[b0:b0 + args.batch]
            n = len(c)
            lam0 = np.zeros(n)
            if args.grid == "adaptive":
                per = periods_fn(c)
                om = omega_fn(c)
                dlam = np.array([adaptive_dlam(c[i], args.size, args.periods, (per[0][i], per[1][i], per[2][i]),
                                               om[i]) for i in range(n)])
            else:
                dlam = np.full(n, args.delta)
            ref = ref_solver(c, lam0, dlam, args.size)
            port = port_solver(c, lam0, dlam, args.size)
            for i in range(n):
                cid = b0 + i
                rec = validate_case(args.run_id, cid, {k: ref[k][i] for k in FIELDS},
                                    {k: port[k][i] for k in FIELDS}, fields, tol, qlsb, args.run_dir)
                rec["grid"] = {"S": args.size, "lam0_bits": int(np.float64(lam0[i]).view(np.uint64)),
                               "dlam_bits": int(np.float64(dlam[i]).view(np.uint64))}
                rec["tpmi"] = int(ref["tpmi"][i])
                rec["tpri"] = int(ref["tpri"][i])
                rec["port_tp_match"] = bool(int(port["tpmi"][i]) == rec["tpmi"] and int(port["tpri"][i]) == rec["tpri"])
                if not rec["port_tp_match"]:
                    rec["success"] = False
                failed += not rec["success"]
                ledger.write(json.dumps(rec, separators=(",", ":")) + "\n")
            ledger.flush()
            os.fsync(ledger.fileno())
            write_atomic(cpath, json.dumps({"case_id": b0 + n, "field": 0, "start": 0}) + "\n")
            if args.stop_after and b0 + n - first >= args.stop_after:
                break
    return 1 if failed else 0


def parse(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--run-dir", required=True)
    ap.add_argument("--run-id", default="gkval")
    ap.add_argument("--cases", type=int, default=256)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--grid", choices=("uniform", "adaptive"), default="uniform")
    ap.add_argument("--delta", type=float, default=1e-3)
    ap.add_argument("--periods", type=float, default=2.0, help="adaptive grid: lambda_max = K T_char")
    ap.add_argument("--atol", type=float, default=1e-12)
    ap.add_argument("--rtol", type=float, default=1e-12)
    ap.add_argument("--max-ulps", type=int, default=2)
    ap.add_argument("--quant-lsb", type=int, default=8)
    ap.add_argument("--profile", choices=("default", "bitexact"), default="default")
    ap.add_argument("--fields", default=",".join(FIELDS))
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--resume", action="store_true")
    ap.add_argument("--stop-after", type=int, default=0, help="stop after this many cases (for resume tests)")
    return ap.parse_args(argv)


def _oracle_solvers():
    from oracle import np_oracle as O

    def ref(c, lam0, dlam, s):
        out = {k: np.zeros((len(c), s)) for k in FIELDS}
        out["tpmi"] = np.zeros(len(c), np.int64)
        out["tpri"] = np.zeros(len(c), np.int64)
        # per-case dlam: group equal values (uniform grids are one group)
        for d in np.unique(dlam):
            sel = np.nonzero(dlam == d)[0]
            o = O.geokerr_eval_lambda_batch(*(c[sel, j] for j in range(7)), float(lam0[sel[0]]), float(d), s)
            for k in FIELDS:
                out[k][sel] = o[k]
            out["tpmi"][sel] = o["tpmi"]
            out["tpri"][sel] = o["tpri"]
        return out

    def periods(c):
        tr, tth, info = [], [], []
        for row in c:
            p = O.geokerr_estimate_periods(row[0], row[1], row[3], row[4], row[5])
            tr.append(p[0]); tth.append(p[1]); info.append(p[2])
        return np.array(tr), np.array(tth), np.array(info)

    def omega(c):
        return np.array([O.geokerr_phi_omega(row[0], row[1], row[3], row[4], row[5]) for row in c])

    return ref, periods, omega


def _gpu_solver():
    import torch

    import nullpath_amd as npa

    def port(c, lam0, dlam, s):
        t = lambda x: torch.tensor(np.ascontiguousarray(x), dtype=torch.float64, device="cuda")
        g = npa.geokerr_eval_lambda_batch(*(t(c[:, j]) for j in range(7)), t(lam0), t(dlam), s)
        torch.cuda.synchronize()
        return {k: v.cpu().numpy() for k, v in g.items()}

    return port


def main(argv=None) -> int:
    args = parse(argv)
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.dirname(here))
    ref, periods, omega = _oracle_solvers()
    return run(args, ref, _gpu_solver(), periods, omega)


if __name__ == "__main__":
    sys.exit(main())
