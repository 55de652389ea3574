"""GEMM autotuner keyed on (n_occ, n_virt).

For every size class of the batched workload (half filling, n_occ = n_virt = k)
and every contraction of the iteration, times the grouped GEMM over a uniform
batch of molecules of that size with each candidate tile (32x32, 64x64,
128x32) and K split, and keeps the fastest when it beats the cost heuristic by more than 3 %:

    python -m helios_amd.autotune            # on a GPU; writes data/gemm_tuning.json
    python -m helios_amd.autotune --correction
        # the BF16 correction forms of the mixed-precision solver (tiles 32x32,
        # 64x64, 128x128; keys 100 + op), merged into the same file

The library applies the table through hcc_set_gemm_tuning (capi.Context does
so by default when the file exists).
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time


HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

from helios_amd import capi, molecules, ppp  # noqa: E402

TUNING_PATH = os.path.join(HERE, "data", "gemm_tuning.json")
N_OPS = 25
KSPLITS = (1, 2, 4, 8, 16, 32, 64)


def _prep(args):
    n, seed = args
    p = ppp.prepare(molecules.random_fragment(n, seed=seed))
    return dict(C=p["C"], F=p["F"], G=p["G"], nocc=p["nocc"], eref=p["eref"])


def candidates(M, N, K):
    out = []
    for tile in (32, 64, 128):
        if tile == 128 and (N > 32 or M < 128):
            continue
        tm, tn = (128, 32) if tile == 128 else (tile, tile)
        tiles = -(-M // tm) * -(-N // tn)
        for ks in KSPLITS:
            if ks > 1 and (K // ks < 64 or tiles * ks > 4096):
                continue
            out.append((tile, ks))
    return out


def candidates_correction(M, N, K):
    out = []
    for tile in (32, 64, 256):
        if tile == 256 and (M < 128 or N < 128):
            continue
        te = 128 if tile == 256 else tile
        tiles = -(-M // te) * -(-N // te)
        for ks in KSPLITS:
            if ks > 1 and (K // ks < 64 or tiles * ks > 4096):
                continue
            out.append((tile, ks))
    return out


def tune_correction(sizes=range(5, 25), count=200, reps=5, workers=16, log=print):
    """Tile and K split of every BF16 correction-form contraction per size class."""
    entries, report = [], []
    t0 = time.time()
    with mp.Pool(workers) as pool:
        for k in sizes:
            mols = pool.map(_prep, [(2 * k, 7000 + s) for s in range(count)], chunksize=8)
            ctx = capi.Context(0, capi.HCC_MIXED, 8)
            ctx.batch_load(mols)
            gain_k = []
            for op in range(N_OPS):
                M, N, K = ctx.gemm_shape(op)
                if M <= 0 or N <= 0:
                    continue
                base = ctx.tune_gemm(100 + op, 0, 0, reps)
                best = (base, 0, 0)
                for tile, ks in candidates_correction(M, N, K):
                    t = ctx.tune_gemm(100 + op, tile, ks, reps)
                    if t < best[0]:
                        best = (t, tile, ks)
                if best[1] and best[0] < 0.97 * base:
                    entries.append([100 + op, k, k, best[1], best[2]])
                gain_k.append((op, M, N, K, base, best[0], best[1], best[2]))
            ctx.close()
            b = sum(g[4] for g in gain_k)
            a = sum(min(g[4], g[5]) for g in gain_k)
            report.append(dict(n_occ=k, n_virt=k, heuristic_ms=b, tuned_ms=a, correction=True,
                               ops=[dict(op=100 + g[0], M=g[1], N=g[2], K=g[3], heuristic_ms=g[4], best_ms=g[5],
                                         tile=g[6], ksplit=g[7]) for g in gain_k]))
            log(f"# k={k}: correction GEMMs {b:.3f} ms -> {a:.3f} ms ({count} molecules), {time.time() - t0:.0f} s")
    return entries, report


def tune(sizes=range(5, 25), count=200, reps=5, workers=16, log=print):
    entries, report = [], []
    t0 = time.time()
    with mp.Pool(workers) as pool:
        for k in sizes:
            mols = pool.map(_prep, [(2 * k, 7000 + s) for s in range(count)], chunksize=8)
            ctx = capi.Context(0, capi.HCC_FP64, 8)
            ctx.batch_load(mols)
            gain_k = []
            for op in range(N_OPS):
                M, N, K = ctx.gemm_shape(op)
                base = ctx.tune_gemm(op, 0, 0, reps)
                best = (base, 0, 0)
                for tile, ks in candidates(M, N, K):
                    t = ctx.tune_gemm(op, tile, ks, reps)
                    if t < best[0]:
                        best = (t, tile, ks)
                if best[1] and best[0] < 0.97 * base:
                    entries.append([op, k, k, best[1], best[2]])
                gain_k.append((op, M, N, K, base, best[0], best[1], best[2]))
            ctx.close()
            b = sum(g[4] for g in gain_k)
            a = sum(min(g[4], g[5]) for g in gain_k)
            report.append(dict(n_occ=k, n_virt=k, heuristic_ms=b, tuned_ms=a,
                               ops=[dict(op=g[0], M=g[1], N=g[2], K=g[3], heuristic_ms=g[4], best_ms=g[5],
                                         tile=g[6], ksplit=g[7]) for g in gain_k]))
            log(f"# k={k}: iteration GEMMs {b:.3f} ms -> {a:.3f} ms ({count} molecules), {time.time() - t0:.0f} s")
    return entries, report


def load(path=TUNING_PATH):
    """Tuning entries (op, n_occ, n_virt, tile, ksplit) or None when no file exists."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["entries"]


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--kmin", type=int, default=5)
    ap.add_argument("--kmax", type=int, default=24)
    ap.add_argument("--count", type=int, default=200)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", default=TUNING_PATH)
    ap.add_argument("--correction", action="store_true",
                    help="tune the BF16 correction forms and merge them into the existing file")
    a = ap.parse_args()
    log = lambda s: print(s, flush=True)   # noqa: E731
    sizes = range(a.kmin, a.kmax + 1)
    if a.correction:
        entries, report = tune_correction(sizes, a.count, a.reps, log=log)
        old = {"entries": [], "report": []}
        if os.path.exists(a.out):
            with open(a.out) as f:
                old = json.load(f)
        keep = [e for e in old["entries"] if e[0] < 100]
        entries = keep + entries
        report = [r for r in old.get("report", []) if not r.get("correction")] + report
    else:
        entries, report = tune(sizes, a.count, a.reps, log=log)
        if os.path.exists(a.out):   # keep the correction-form entries of an earlier run
            with open(a.out) as f:
                old = json.load(f)
            entries += [e for e in old["entries"] if e[0] >= 100]
            report += [r for r in old.get("report", []) if r.get("correction")]
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    with open(a.out, "w") as f:
        json.dump(dict(entries=entries, molecules_per_size=a.count, reps=a.reps,
                       note="(op, n_occ, n_virt, tile, ksplit); ops in iteration launch order, "
                            "100 + op: the op's BF16 correction form",
                       report=report), f, indent=0)
    print(f"# {len(entries)} entries written to {a.out}", flush=True)


if __name__ == "__main__":
    main()
