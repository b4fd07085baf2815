"""Autotuner of the complex GEMM's tile choice.

``qepw_gemm`` has a handful of tile variants (TM x TN output tiles per
workgroup) and splits K over workgroups for the tall A^H B products over the
plane waves.  The built-in choice is a heuristic; this module times every
variant (and, for the split products, a range of split counts) on the
device for the shape classes a Davidson solve produces, and writes the
winners as rules into a JSON table that ``load_table`` installs in the
library (``qepw_gemm_set_rule``).  The table committed next to this module
(``tuning/gemm_gfx950.json``) is loaded automatically with the library.

Shape classes (per precision):

* ``N`` products with m > 4096 rows: the npw x n blocks of Ritz vectors,
  residuals, basis updates and projector sums -- buckets of n (columns) and
  k (the subspace or projector count);
* ``C`` products with k > 2048: the A^H B projections over npw -- buckets of
  m and n (the subspace and block sizes).

Each bucket is timed at its upper end with the medium workload's npw.
``python -m qe_python_amd.autotune --out PATH`` runs the sweep (on a GPU).
"""

from __future__ import annotations

import argparse
import ctypes
import json
import time
from pathlib import Path

from . import _native as nat

TABLE = Path(__file__).resolve().parent / "tuning" / "gemm_gfx950.json"
NPW = 23847
INF = 1 << 40
# (lo, hi) buckets, timed at hi (or at TIME_AT for the open-ended last one)
N_BUCKETS = [(2, 32), (33, 64), (65, 128), (129, 256), (257, 512)]
K_BUCKETS = [(1, 128), (129, 256), (257, 384), (385, INF)]
K_TIME_AT = {INF: 512}
M_BUCKETS = [(1, 128), (129, 256), (257, 384), (385, INF)]
SPLIT_TARGETS = (256, 512, 1024, 2048)   # workgroups aimed at when splitting K


def _ranges(*pairs):
    flat = [v for pair in pairs for v in pair]
    return (ctypes.c_int64 * 6)(*flat)


def _plan(lib, prec, op, m, n, k):
    out = (ctypes.c_int64 * 4)()
    nat.check(lib.qepw_gemm_plan(prec, op, m, n, k, out), "qepw_gemm_plan")
    return tuple(int(v) for v in out)


def _candidates(lib, prec, op, m, n, k):
    """(variant, splits) pairs worth timing for one shape."""
    nvar = int(lib.qepw_gemm_variant_count())
    out = []
    for v in range(nvar):
        probe = (ctypes.c_int64 * 6)(m, m, n, n, k, k)
        lib.qepw_gemm_clear_rules()
        nat.check(lib.qepw_gemm_set_rule(prec, op, probe, v, 1), "qepw_gemm_set_rule")
        _, _, tm, tn = _plan(lib, prec, op, m, n, k)
        tiles = ((m + tm - 1) // tm) * ((n + tn - 1) // tn)
        splits = {1}
        if op & nat.OP_C and k > 2048:
            for target in SPLIT_TARGETS:
                s = max(1, min(-(-target // tiles), k // 128))
                splits.add(s)
        out.extend((v, s) for s in sorted(splits))
    lib.qepw_gemm_clear_rules()
    return out


def _time(torch, lib, prec, op, m, n, k, variant, splits, reps, device):
    from . import device as dv
    dtype = torch.complex128 if prec == nat.C128 else torch.complex64
    a_cols, a_rows = (m, k) if op & nat.OP_C else (k, m)
    gen = torch.Generator(device=device)
    gen.manual_seed(7)
    a = torch.randn((a_cols, a_rows), dtype=dtype, device=device, generator=gen)
    b = torch.randn((n, k), dtype=dtype, device=device, generator=gen)
    c = torch.zeros((n, m), dtype=dtype, device=device)
    ws = dv.Workspace(device)
    lib.qepw_gemm_clear_rules()
    nat.check(lib.qepw_gemm_set_rule(prec, op, _ranges((m, m), (n, n), (k, k)), variant,
                                     splits), "qepw_gemm_set_rule")
    dv.forget_gemm_plans()
    try:
        for _ in range(2):
            dv.gemm(op, m, n, k, 1.0, a, b, 0.0, c, ws)
        stream = torch.cuda.current_stream(device)
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record(stream)
        for _ in range(reps):
            dv.gemm(op, m, n, k, 1.0, a, b, 0.0, c, ws)
        t1.record(stream)
        torch.cuda.synchronize(device)
        return t0.elapsed_time(t1) / reps
    finally:
        lib.qepw_gemm_clear_rules()
        dv.forget_gemm_plans()


def shape_classes():
    """(op, (m_lo, m_hi), (n_lo, n_hi), (k_lo, k_hi), (m, n, k) timed)."""
    out = []
    for n_lo, n_hi in N_BUCKETS:
        for k_lo, k_hi in K_BUCKETS:
            k = K_TIME_AT.get(k_hi, k_hi)
            out.append((nat.OP_N, (4097, INF), (n_lo, n_hi), (k_lo, k_hi), (NPW, n_hi, k)))
    for m_lo, m_hi in M_BUCKETS:
        for n_lo, n_hi in N_BUCKETS:
            m = K_TIME_AT.get(m_hi, m_hi)
            out.append((nat.OP_C, (m_lo, m_hi), (n_lo, n_hi), (2049, INF), (m, n_hi, NPW)))
    return out


def tune(reps: int = 10, device=None, log=print):
    """Time every candidate of every shape class in both precisions; return
    the table (a dict ready for ``json.dump``)."""
    import torch
    device = torch.device(device or "cuda")
    lib = nat.load()
    rules = []
    for prec, pname in ((nat.C64, "fp32"), (nat.C128, "fp64")):
        for op, mr, nr, kr, (m, n, k) in shape_classes():
            lib.qepw_gemm_clear_rules()
            base = _plan(lib, prec, op, m, n, k)
            times = {}
            for variant, splits in _candidates(lib, prec, op, m, n, k):
                times[(variant, splits)] = _time(torch, lib, prec, op, m, n, k, variant, splits,
                                                 reps, device)
            best = min(times, key=times.get)
            default = times.get((base[0], base[1]))
            if default is None:
                default = _time(torch, lib, prec, op, m, n, k, base[0], base[1], reps, device)
            flops = 8.0 * m * n * k
            rec = {"precision": pname, "op": "C" if op & nat.OP_C else "N",
                   "m": list(mr), "n": list(nr), "k": list(kr), "timed_mnk": [m, n, k],
                   "variant": best[0], "splits": best[1], "ms": times[best],
                   "tflops": flops / (times[best] * 1e-3) / 1e12,
                   "default_ms": default}
            rules.append(rec)
            log(f"{pname} {rec['op']} m{mr} n{nr} k{kr} at {m}x{n}x{k}: variant {best[0]} "
                f"splits {best[1]} {times[best] * 1e3:.1f} us ({rec['tflops']:.1f} TFLOP/s), "
                f"built-in {default * 1e3:.1f} us")
    lib.qepw_gemm_clear_rules()
    name = torch.cuda.get_device_name(device)
    return {"device": name, "npw": NPW, "rules": rules}


def load_table(path: Path | str | None = None) -> int:
    """Install a tuned table's rules (a rule replaces the built-in choice only
    where it was measured faster).  Returns the number of rules set."""
    from . import device as dv
    path = Path(path) if path is not None else TABLE
    if not path.exists():
        return 0
    table = json.loads(path.read_text())
    lib = nat.load()
    lib.qepw_gemm_clear_rules()
    dv.forget_gemm_plans()
    count = 0
    for r in table.get("rules", []):
        if r["ms"] >= r.get("default_ms", INF) * 0.97:
            continue      # within noise of the built-in choice: keep it
        prec = nat.C128 if r["precision"] == "fp64" else nat.C64
        op = nat.OP_C if r["op"] == "C" else nat.OP_N
        ranges = _ranges(tuple(r["m"]), tuple(r["n"]), tuple(r["k"]))
        nat.check(lib.qepw_gemm_set_rule(prec, op, ranges, int(r["variant"]),
                                         int(r["splits"])), "qepw_gemm_set_rule")
        count += 1
    return count


def main():
    ap = argparse.ArgumentParser(description="tune the complex GEMM's tile choice")
    ap.add_argument("--out", default=str(TABLE))
    ap.add_argument("--reps", type=int, default=10)
    args = ap.parse_args()
    t0 = time.perf_counter()
    table = tune(args.reps)
    table["seconds"] = time.perf_counter() - t0
    Path(args.out).parent.mkdir(parents=True, exist_ok=True)
    Path(args.out).write_text(json.dumps(table, indent=1) + "\n")
    print(f"wrote {args.out} ({len(table['rules'])} shape classes)")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
