"""Benchmark: converged CCSD solves/s on the batched workload, one MI355X.

Workload: 4096 seeded planar honeycomb carbon fragments (even site counts
uniform in [10, 48], 1.40 A bonds with +-0.02 A jitter, beta = -2.4 eV,
U = 11.13 eV, Ohno-Klopman gamma, half filling), all solved together by the
batched HIP solver.  A step is one complete batched solve of the whole batch
from device-resident inputs (integral build, guess, iterations to the
threshold, energies).  A solve counts as converged only when its energy is
within 1e-6 eV of the FP64 reference energy of the same molecule (the FP64
solver at a residual threshold of 1e-10, untimed).

Prints one JSON line (see the keys below).  --gpus N>1 is not supported (one
GPU per box); across GPUs the batch would shard by molecule, with a final gather
of the per-molecule results.
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from helios_amd import autotune, capi, learned_guess, lrccsd, molecules, ppp  # noqa: E402

PEAK_FP64 = 78.6e12      # MI355X dense FP64 (matrix) FLOP/s
PEAK_BF16 = 2.5e15
PEAK_HBM = 8.0e12        # HBM3E bytes/s
# HBM bytes per dispatch of the dominant kernels on the default 4096-molecule batch
# with the default solver settings: FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (factors from a
# 1 GiB copy), two rocprofv3 --pmc passes of "bench.py --workers 0 --no-cpu --no-lr
# --skip-reference --steps 1 --warmup 0 --no-profile" summarised in
# profiles/hbm_traffic_batch4096.md.  A counter pass cannot run inside this process, so
# other workloads report null.
GEMM_TRAFFIC_BATCH4096 = 3.285e9        # gemm_f64_kernel, FP64 solve
DUAL_GEMM_TRAFFIC_BATCH4096 = 1.468e9   # gemm_dual_bf16, mixed solve


def _prep(k):
    m = molecules.batch_member(k)
    p = ppp.prepare(m)
    _, _, hk = ppp.model(m.coords, m.bonds, m.alpha, m.U, 1)
    return dict(C=p["C"], F=p["F"], G=p["G"], nocc=p["nocc"], eref=p["eref"], n=m.n, core=p["core"], huckel=hk)


def prepare_batch(nmol, workers):
    """Host SCF of every molecule.  workers = 0 prepares them in this process
    (for profiler runs: no fork after a profiler library has started the
    runtime); otherwise a pool forked before any GPU call."""
    if workers <= 0:
        return [_prep(k) for k in range(nmol)]
    with mp.Pool(workers) as pool:
        return pool.map(_prep, range(nmol), chunksize=16)


def f_iter(o, v):
    """Algorithmic contraction flops per molecule-iteration: 2o^2v^4 + 4o^4v^2 + 12o^3v^3."""
    return 2.0 * o * o * v ** 4 + 4.0 * o ** 4 * v * v + 12.0 * o ** 3 * v ** 3


def f_int(o, v):
    """Algorithmic flops of one build of the six MO integral classes:
    sum over (P, Q) of 2 P N^2 + 2 P Q N, (P, Q) in {(o^2,o^2), (o^2,ov), (o^2,v^2),
    (ov,ov), (ov,v^2), (v^2,v^2)}."""
    n = o + v
    pq = [(o * o, o * o), (o * o, o * v), (o * o, v * v), (o * v, o * v), (o * v, v * v), (v * v, v * v)]
    return sum(2.0 * P * n * n + 2.0 * P * Q * n for P, Q in pq)


def diis_bytes(o, v, iters, m):
    """Algorithmic DIIS bytes of a solve: 8 nd (2 h + 4) per iteration that
    extrapolates (every iteration but the converged last one), h the history
    length in use (min(iteration, m)), nd = ov(ov+1)/2 + ov."""
    ov = o * v
    nd = ov * (ov + 1) / 2 + ov
    total = np.zeros(len(o))
    for k in range(len(o)):
        h = np.minimum(np.arange(1, int(iters[k])), m)
        total[k] = (8.0 * nd[k] * (2 * h + 4)).sum()
    return total


def f_triples(o, v):
    """(T) flops per molecule: the connected triples w over all ordered occupied
    triples, each a v^3 x (v + o) multiply-add (the six permutations of an
    i >= j >= k item cover the ordered triples once)."""
    o = np.asarray(o, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    return 2.0 * o ** 3 * v ** 3 * (o + v)


def triples_leg(ctx, o, v):
    """(T) of the converged batch: one untimed call, one timed, one profiled."""
    ctx.batch_triples()
    t1 = time.perf_counter()
    _, et = ctx.batch_triples()
    wall = time.perf_counter() - t1
    ctx.profile(True)
    ctx.batch_triples()
    ops = ctx.profile_ops()
    ctx.profile(False)
    kms = float(ops[78])
    fl = float(f_triples(o, v).sum())
    out = {"bound": "mfma", "achieved": fl / (kms * 1e-3) / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
           "frac": fl / (kms * 1e-3) / PEAK_FP64,
           "kernel": "k_triples_mfma (FP64 matrix cores, W in LDS; k_triples for v > 24)", "kernel_ms": kms,
           "integral_gemm_ms": float(ops[58] + ops[59] + ops[62]), "setup_ew_ms": float(ops[79]),
           "wall_ms_incl_readback": wall * 1e3, "molecules": int(len(o)),
           "mean_e_t_eV": float(et.sum(1).mean()),
           "work": "2 o^3 v^3 (o + v) flops per molecule (connected triples over ordered i, j, k)"}
    print(f"# (T): kernel {kms:.1f} ms ({out['achieved']:.1f} TFLOP/s, {out['frac']:.3f} of FP64 peak), "
          f"wall {wall * 1e3:.1f} ms for {len(o)} molecules", file=sys.stderr, flush=True)
    return out


def lr_leg(mols, nroots=3):
    """Singlet LR-CCSD (three lowest roots, CIS start vectors) of a sample of
    the batch.  The subspace lives in device memory; the wall time splits
    into the sigma products (two FP64 residual passes each), the other
    vector-space calls (dot products, combinations, preconditioner) and the
    host's small eigenproblems and bookkeeping."""
    # each molecule is loaded nwork = nroots + 4 times (copies solved alike):
    # one Jacobian-product call then applies J to that many of its vectors
    copies = nroots + 4
    ctx = capi.Context(0, capi.HCC_FP64, 20)
    ctx.batch_load([m for m in mols for _ in range(copies)])
    ctx.batch_solve(capi.GUESS_MP2, 500, 1e-10)
    f_it, _ = ctx.gemm_flops(copies * len(mols))   # residual GEMM flops per loaded molecule and pass
    timers = {"sigma": [0, 0.0], "space": [0, 0.0]}

    def timed(fn, key):
        def run(*a, **kw):
            t1 = time.perf_counter()
            out = fn(*a, **kw)
            timers[key][0] += 1
            timers[key][1] += time.perf_counter() - t1
            return out
        return run

    sigma_flops = [0.0]   # residual GEMM flops of the molecules each sigma call runs (two passes)
    sigma = ctx.lr_sigma_from

    def sigma_counted(src, in_slot, out_slot, h):
        sigma_flops[0] += 2.0 * float(f_it[np.asarray(in_slot) >= 0].sum())
        return sigma(src, in_slot, out_slot, h)

    ctx.lr_sigma_from = timed(sigma_counted, "sigma")
    for name in ("lr_dots", "lr_combine", "lr_gram", "lr_combine_block", "lr_precondition", "lr_put", "lr_get",
                 "lr_space_counts"):
        setattr(ctx, name, timed(getattr(ctx, name), "space"))
    t1 = time.perf_counter()
    om, res, its, conv, ael, r0 = lrccsd.excitation_energies(
        ctx, [(m["nocc"], m["n"] - m["nocc"]) for m in mols], nroots=nroots, gammas=[m["G"] for m in mols],
        log=lambda it, active: print(f"# LR-CCSD iteration {it}: {active} molecules active", file=sys.stderr,
                                     flush=True),
        replicas=[[copies * k + j for j in range(copies)] for k in range(len(mols))])
    wall = time.perf_counter() - t1
    ctx.close()
    (ns, ts), (nv, tv) = timers["sigma"], timers["space"]
    out = {"molecules": len(mols), "roots": nroots, "wall_s": wall, "sigma_calls": ns, "sigma_s": ts,
           "sigma_ms_per_call": 1e3 * ts / max(1, ns), "space_calls": nv, "space_s": tv,
           "host_s": wall - ts - tv, "mean_iterations": float(its.mean()),
           "converged": int(conv.sum()), "mean_omega_eV": [float(x) for x in np.nanmean(om, axis=0)],
           "mean_excitation_level": [float(x) for x in np.nanmean(ael, axis=0)],
           "mean_reference_weight_r0": [float(x) for x in np.nanmean(r0, axis=0)], "start": "cis",
           "copies_per_molecule": copies}
    # a sigma call runs two residual passes over the molecules that have a vector in it
    sig_rate = sigma_flops[0] / max(ts, 1e-9)
    out["roofline_sigma"] = {"bound": "mfma", "achieved": sig_rate / 1e12, "peak": PEAK_FP64 / 1e12,
                             "unit": "TFLOP/s", "frac": sig_rate / PEAK_FP64,
                             "work": "2 x residual GEMM flops of the molecules with a vector in each sigma "
                                     "call, over the sigma calls' wall time (transfers of the slot pointers and "
                                     "elementwise steps included)"}
    print(f"# LR-CCSD: {len(mols)} molecules, {nroots} roots, {wall:.2f} s: {ns} sigma calls "
          f"({1e3 * ts / max(1, ns):.1f} ms each), {nv} vector-space calls ({tv:.2f} s), host {wall - ts - tv:.2f} s; "
          f"converged {int(conv.sum())}", file=sys.stderr, flush=True)
    return out


def copy_calibration():
    """A plain device copy of 1 GiB, three times, to calibrate the HBM traffic
    counters of a profiler pass (FETCH_SIZE, WRITE_SIZE per dispatch)."""
    import torch
    n = (1 << 30) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"copied {1 << 30} bytes x 3", flush=True)


def cpu_baseline(mols, iters_fp64, threads_all):
    """The reference CPU path (oracle/refloops.c: the reference's loop nests and
    OpenMP pattern, O(o^4 v^4) per iteration) on the host cores: on one batch
    member per even N from 10 to 24, the R
    build and one iteration are timed with 1 thread and with every host
    thread (and N = 10, 12 solved in full, whose time the parts must predict);
    every molecule of the batch is then costed as build + (its FP64 GPU
    iteration count) x iteration, with the o^8 fit of the largest sampled
    sizes beyond N = 24 -- an extrapolation, labelled as such."""
    from oracle import refloops
    sample_n = list(range(10, 26, 2))
    reps = {}
    for k, m in enumerate(mols):
        if m["n"] in sample_n and m["n"] not in reps:
            reps[m["n"]] = k
        if len(reps) == len(sample_n):
            break
    out = {}
    t_cpu = 0.0
    check = {}
    for th in (1, threads_all):
        parts = {}
        for n in sorted(reps):
            p = mols[reps[n]]
            tb, ti = refloops.time_parts(p["C"], p["F"], p["G"], p["nocc"], iters=1, threads=th)
            parts[n] = (tb, ti)
            t_cpu += (tb + ti) * th
        for n in (10, 12):
            p = mols[reps[n]]
            t0 = time.perf_counter()
            r = refloops.solve(p["C"], p["F"], p["G"], p["nocc"], p["eref"], eps=2e-7, threads=th)
            dt = time.perf_counter() - t0
            t_cpu += dt * th
            pred = parts[n][0] + r["iterations"] * parts[n][1]
            check[(th, n)] = (dt, pred, r["iterations"])
        big = [n for n in sorted(parts)][-3:]
        c_it = np.mean([parts[n][1] / (n // 2) ** 8 for n in big])
        c_b = np.mean([parts[n][0] / n ** 6 for n in big])
        total = 0.0
        for k, m in enumerate(mols):
            n = m["n"]
            tb, ti = parts[n] if n in parts else (c_b * n ** 6, c_it * (n // 2) ** 8)
            total += tb + float(iters_fp64[k]) * ti
        out[th] = (len(mols) / total, total, parts)
    v_all, tot_all, parts_all = out[threads_all]
    v_one, tot_one, _ = out[1]
    return {"value": v_all, "unit": "solves/s", "cores": int(threads_all), "kind": "port",
            "value_1_thread": v_one,
            "batch_seconds_extrapolated": tot_all, "batch_seconds_extrapolated_1_thread": tot_one,
            "full_solve_check": {f"n{n}_threads{th}": {"measured_s": round(a, 4), "predicted_s": round(b, 4),
                                                      "iterations": it}
                                 for (th, n), (a, b, it) in check.items()},
            "sample": (f"C restatement of the reference loop nests (oracle/refloops.c, gcc -O3 -fopenmp); "
                       f"R build + 1 iteration timed on one batch member per N in {sorted(reps)} with 1 and "
                       f"{threads_all} threads, full solves at N = 10, 12 to check the parts; the 4096 batch "
                       f"costed per molecule with the FP64 GPU solver's iteration counts and an o^8 fit "
                       f"above N = 24 (extrapolation); {t_cpu:.1f} thread-seconds of CPU work")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nmol", type=int, default=4096)
    ap.add_argument("--eps", type=float, default=2e-7, help="residual threshold of the timed solves")
    ap.add_argument("--diis", type=int, default=20,
                    help="DIIS vectors of both solvers (the reference's default diis-steps)")
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp64-only", action="store_true")
    ap.add_argument("--ablation", action="store_true", help="also time FP64+learned and mixed+MP2")
    ap.add_argument("--no-tuning", action="store_true", help="ignore the measured GEMM tile/split table")
    ap.add_argument("--skip-reference", action="store_true",
                    help="profiler runs only: no FP64 reference solve, solves count as converged by status alone")
    ap.add_argument("--kappa", type=float, default=0.05)
    ap.add_argument("--max-inner", type=int, default=5)
    ap.add_argument("--max-outer", type=int, default=30)
    ap.add_argument("--no-profile", action="store_true", help="skip the profiled solves (kernel attribution)")
    ap.add_argument("--no-lr", action="store_true", help="skip the LR-CCSD leg (64-molecule sample)")
    ap.add_argument("--lr-only", action="store_true",
                    help="run only the LR-CCSD leg on the first 64 batch molecules and print its JSON")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB three times (HBM counter calibration for profiler passes)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this benchmark runs on one MI355X; "
              "across GPUs the batch would shard by molecule", file=sys.stderr)
        sys.exit(2)

    if args.lr_only:
        print(json.dumps({"lr_ccsd_sample": lr_leg(prepare_batch(64, args.workers))}), flush=True)
        return

    t0 = time.time()
    mols = prepare_batch(args.nmol, args.workers)
    print(f"# prepared {len(mols)} molecules in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    o = np.array([m["nocc"] for m in mols])
    v = np.array([m["n"] - m["nocc"] for m in mols])

    # FP64 reference energies (untimed, tight threshold)
    triples = lr = None
    if args.skip_reference:
        e_ref = np.zeros((len(mols), 5))
        st_ref = np.ones(len(mols), dtype=np.int32)
    else:
        ctx = capi.Context(0, capi.HCC_FP64, args.diis)
        ctx.batch_load(mols)
        ctx.batch_solve(capi.GUESS_MP2, 500, 1e-10)
        e_ref, it_ref, st_ref = ctx.batch_results()
        triples = triples_leg(ctx, o, v)
        ctx.close()
        if not args.no_lr:
            lr = lr_leg(mols[:64])
        print(f"# reference solve: statuses {np.unique(st_ref, return_counts=True)}, "
              f"mean iterations {it_ref.mean():.1f}", file=sys.stderr, flush=True)

    # the batch's RHF on the device (the step before the path; its time is
    # reported beside the headline, which starts from the MO inputs)
    device_scf = None
    if not args.skip_reference:
        ctx = capi.Context(0, capi.HCC_FP64, args.diis)
        ctx.batch_load_scf(mols, accuracy=1e-11, maxiter=400)      # warm-up (code object load)
        t1 = time.perf_counter()
        er, its_scf, st_scf = ctx.batch_load_scf(mols, accuracy=1e-11, maxiter=400)
        scf_s = time.perf_counter() - t1
        ctx.batch_solve(capi.GUESS_MP2, 500, 1e-10)
        e_dev, _, st_dev = ctx.batch_results()
        ctx.close()
        device_scf = {"ms_incl_upload": scf_s * 1e3, "mean_iterations": float(its_scf.mean()),
                      "converged": int((st_scf == 1).sum()),
                      "max_abs_dE_ccsd_vs_host_scf_eV": float(np.abs(e_dev[:, 0] - e_ref[:, 0]).max())}
        print(f"# device SCF: {scf_s * 1e3:.1f} ms for {len(mols)} molecules, mean iterations "
              f"{its_scf.mean():.1f}, CCSD from it vs host SCF max|dE| "
              f"{device_scf['max_abs_dE_ccsd_vs_host_scf_eV']:.2e}", file=sys.stderr, flush=True)

    net = learned_guess.load_params()
    tuning = None if args.no_tuning else autotune.load()

    def timed(precision, guess):
        ctx = capi.Context(0, precision, args.diis)
        if tuning:
            ctx.set_gemm_tuning(tuning)
        if precision == capi.HCC_MIXED:
            ctx.set_mixed(True, args.kappa, args.max_inner, args.max_outer)
        ctx.set_guess_network(net)
        ctx.batch_load(mols)
        for _ in range(args.warmup):
            ctx.batch_solve(guess, 2000, args.eps)
        # timed steps: the shipped path (iteration blocks replayed as HIP graphs,
        # no per-launch events)
        t_steps = []
        for _ in range(args.steps):
            t1 = time.perf_counter()
            ctx.batch_solve(guess, 2000, args.eps)   # returns after a device synchronize
            t_steps.append(time.perf_counter() - t1)
        e, iters, st = ctx.batch_results()
        outer = inner = None
        fl_corr = None
        if precision == capi.HCC_MIXED:
            outer, inner = ctx.batch_statistics()
            fl_corr = ctx.gemm_flops_correction(len(mols))
        # one more solve, profiled (every launch group bracketed by HIP events on
        # the context's stream, graphs off): per-kernel-class device time
        gemm_ms = launches = prof_ms = None
        ops_ms = np.zeros(96)
        ops_n = np.zeros(96, dtype=np.int64)
        if not args.no_profile:
            ctx.profile(True)
            t1 = time.perf_counter()
            ctx.batch_solve(guess, 2000, args.eps)
            prof_ms = (time.perf_counter() - t1) * 1e3
            gemm_ms, launches = ctx.profile_read()
            ops_ms = ctx.profile_ops()
            ops_n = ctx.profile_launches()
            ctx.profile(False)
            ep, itp, stp = ctx.batch_results()
            assert np.array_equal(itp, iters) and np.array_equal(stp, st), "profiled solve differs"
        fl_it, fl_pre = ctx.gemm_flops(len(mols))
        ctx.close()
        de = np.abs(e[:, 0] - e_ref[:, 0]) if not args.skip_reference else np.zeros(len(mols))
        ok = (de < 1e-6) & (st == 1) & (st_ref == 1)
        step_s = float(np.mean(t_steps))
        return dict(e=e, iters=iters, st=st, de=de, ok=ok, step_s=step_s, gemm_ms=gemm_ms, launches=launches,
                    ops_ms=ops_ms, ops_n=ops_n, fl_it=fl_it, fl_pre=fl_pre, outer=outer, inner=inner, fl_corr=fl_corr,
                    solves=ok.sum() / step_s, profiled_solve_ms=prof_ms)

    r64 = timed(capi.HCC_FP64, capi.GUESS_MP2)
    print(f"# fp64: {r64['solves']:.1f} solves/s, {r64['step_s']*1e3:.1f} ms/step, "
          f"converged {int(r64['ok'].sum())}/{len(mols)}, max|dE| {r64['de'].max():.2e}", file=sys.stderr, flush=True)
    rmx = timed(capi.HCC_MIXED, capi.GUESS_LEARNED) if not args.fp64_only else None
    if rmx is not None:
        print(f"# mixed+learned: {rmx['solves']:.1f} solves/s, {rmx['step_s']*1e3:.1f} ms/step, "
              f"converged {int(rmx['ok'].sum())}/{len(mols)}, max|dE| {rmx['de'].max():.2e}, "
              f"outer {rmx['outer'].mean():.2f} inner {rmx['inner'].mean():.2f}", file=sys.stderr, flush=True)
    ablation = {}
    if args.ablation:
        for name, prec, guess in (("fp64_learned", capi.HCC_FP64, capi.GUESS_LEARNED),
                                  ("mixed_mp2", capi.HCC_MIXED, capi.GUESS_MP2)):
            r = timed(prec, guess)
            ablation[name] = {"solves_per_s": float(r["solves"]), "ms_per_step": r["step_s"] * 1e3,
                              "mean_iterations": float(r["iters"].mean()), "converged": int(r["ok"].sum())}
            print(f"# {name}: {r['solves']:.1f} solves/s, mean iterations {r['iters'].mean():.2f}",
                  file=sys.stderr, flush=True)
    head = rmx if rmx is not None else r64

    default_workload = (args.nmol == 4096 and args.eps == 2e-7 and args.kappa == 0.05 and args.diis == 20
                        and args.max_inner == 5 and args.max_outer == 30 and not args.no_tuning)
    rooflines = {}
    if not args.no_profile:
        ms = r64["ops_ms"]
        its = r64["iters"].astype(np.float64)
        # FP64 contraction GEMM: executed flops (2MNK per launched problem) and the
        # algorithmic count F_iter per molecule-iteration, over the profiled solve
        gemm_exec = float((r64["fl_it"] * its).sum() + r64["fl_pre"].sum())
        gemm_ms = float(r64["gemm_ms"])
        falg64 = float((f_iter(o, v) * its).sum())
        rfp64 = {"bound": "mfma", "achieved": falg64 / (gemm_ms * 1e-3) / 1e12, "peak": PEAK_FP64 / 1e12,
                 "unit": "TFLOP/s", "frac": falg64 / (gemm_ms * 1e-3) / PEAK_FP64,
                 "executed_tflops": gemm_exec / (gemm_ms * 1e-3) / 1e12,
                 "traffic": GEMM_TRAFFIC_BATCH4096 if default_workload else None,
                 "kernel": "gemm_f64_kernel (FP64 solver, all contractions)", "launches": int(r64["launches"]),
                 "kernel_ms": gemm_ms, "avg_launch_ms": gemm_ms / max(1, int(r64["launches"])),
                 "work": "F_iter = 2o^2v^4 + 4o^4v^2 + 12o^3v^3 per molecule-iteration"}
        rooflines["roofline_fp64_solver"] = rfp64
        # integral build: the dressed integral blocks are rebuilt every iteration
        # (GEMM tables g1, g3, g5, g8, g9, g11, g14, g17, g18 and the set-up (kc|ld))
        int_ids = [0, 3, 5, 8, 9, 11, 14, 17, 18]
        int_ms = float(ms[int_ids].sum() + ms[62] + ms[63])
        fint = float((f_int(o, v) * (its + 1)).sum())
        rooflines["roofline_integrals"] = {
            "bound": "mfma", "achieved": fint / (int_ms * 1e-3) / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
            "frac": fint / (int_ms * 1e-3) / PEAK_FP64, "kernel_ms": int_ms,
            "work": "F_int = sum (2PN^2 + 2PQN) over the six classes, per build (one per iteration + set-up)"}
        upd_ms = float(ms[64])
        upd_b = float((24.0 * (o * v) ** 2 * its).sum())
        rooflines["roofline_update"] = {
            "bound": "hbm", "achieved": upd_b / (upd_ms * 1e-3) / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
            "frac": upd_b / (upd_ms * 1e-3) / PEAK_HBM, "kernel_ms": upd_ms,
            "kernel": "k_update1 + k_update2 (FP64 solver)", "work": "24 o^2 v^2 bytes per molecule-iteration"}
        diis_ms = float(ms[65])
        dbytes = float(diis_bytes(o, v, its, args.diis).sum())
        rooflines["roofline_diis"] = {
            "bound": "hbm", "achieved": dbytes / (diis_ms * 1e-3) / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
            "frac": dbytes / (diis_ms * 1e-3) / PEAK_HBM, "kernel_ms": diis_ms,
            "kernel": "k_diis_dot + k_diis_solve + k_diis_extrap (FP64 solver)",
            "work": "8 nd (2h + 4) bytes per molecule-iteration, h = history in use"}
        roofline = rfp64
        if rmx is not None:
            mm = rmx["ops_ms"]
            corr_ms = float(mm[32:57].sum())
            # every inner solve starts with a step taken from the outer residual,
            # without correction GEMMs: a molecule evaluates them inner - (outer - 1) times
            corr_evals = (rmx["inner"] - (rmx["outer"] - 1)).astype(np.float64)
            falg = float((f_iter(o, v) * corr_evals).sum())
            fexec = float((rmx["fl_corr"] * corr_evals).sum())
            roofline = {"bound": "mfma", "achieved": falg / (corr_ms * 1e-3) / 1e12, "peak": PEAK_BF16 / 1e12,
                        "unit": "TFLOP/s", "frac": falg / (corr_ms * 1e-3) / PEAK_BF16,
                        "executed_tflops": fexec / (corr_ms * 1e-3) / 1e12,
                        "traffic": DUAL_GEMM_TRAFFIC_BATCH4096 if default_workload else None,
                        "kernel": "gemm_dual_bf16 (mixed solver, correction contractions of the inner iterations)",
                        "kernel_ms": corr_ms, "launches": int(rmx["ops_n"][32:57].sum()),
                        "avg_launch_ms": corr_ms / max(1, int(rmx["ops_n"][32:57].sum())),
                        "work": "F_iter = 2o^2v^4 + 4o^4v^2 + 12o^3v^3 per correction evaluation"}
            groups = ["update", "diis", "update_inner", "diis_inner", "e2", "e2_correction", "residual_ew",
                      "residual_ew_correction", "outer_norm", "commit", "guess", "energy", "setup_ew", "control"]
            rooflines["time_split_mixed_ms"] = {
                "correction_gemm": corr_ms, "outer_fp64_gemm": float(mm[0:25].sum()),
                "setup_gemm": float(mm[60:64].sum()),
                "elementwise": float(mm[64:96].sum()),
                "elementwise_groups": {g: float(mm[64 + k]) for k, g in enumerate(groups)},
                "profiled_solve": rmx["profiled_solve_ms"]}
            rooflines["time_split_fp64_ms"] = {
                "gemm": gemm_ms, "elementwise_groups": {g: float(ms[64 + k]) for k, g in enumerate(groups)
                                                        if ms[64 + k] > 0},
                "profiled_solve": r64["profiled_solve_ms"]}
    else:
        roofline = None
    line = {
        "metric": "converged CCSD solves/s at |dE|<1e-6 eV",
        "value": float(head["solves"]),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["step_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16+f64" if rmx is not None else "f64",
        "data": "synthetic",
        "config": {"workload": f"batch of {len(mols)} seeded honeycomb fragments, N in [10,48], "
                               + ("mixed-precision defect-correction solver, learned T2 guess" if rmx is not None
                                  else "plain FP64 batched solver, MP2 guess"),
                   "molecules": len(mols), "residual_threshold": args.eps, "diis_steps": args.diis,
                   "energies_checked_against_fp64_reference": not args.skip_reference,
                   "gemm_tuning": bool(tuning),
                   "converged": int(head["ok"].sum()), "max_abs_dE_eV": float(head["de"].max()),
                   "fp64_solves_per_s": float(r64["solves"]), "fp64_ms_per_step": r64["step_s"] * 1e3,
                   "fp64_mean_iterations": float(r64["iters"].mean()),
                   "speedup_vs_fp64": float(head["solves"] / r64["solves"]) if rmx is not None else 1.0},
        "roofline": roofline,
        "f_iter_rate_tflops": float((f_iter(o, v) * r64["iters"]).sum()) / r64["step_s"] / 1e12,
    }
    line.update(rooflines)
    if rmx is not None:
        line["config"].update({"mixed_mean_outer": float(rmx["outer"].mean()),
                               "mixed_mean_inner": float(rmx["inner"].mean()),
                               "mixed_mean_correction_evaluations":
                                   float((rmx["inner"] - (rmx["outer"] - 1)).mean()),
                               "mixed_gemm_ms": rmx["gemm_ms"]})
        line["config"]["mixed_max_abs_dE_per_molecule_eV"] = float(rmx["de"].max())
    if ablation:
        line["config"]["ablation"] = ablation
    if device_scf is not None:
        line["config"]["device_scf"] = device_scf
    if triples is not None:
        line["roofline_triples"] = triples
    if lr is not None:
        line["config"]["lr_ccsd_sample"] = lr
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(mols, r64["iters"], threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
