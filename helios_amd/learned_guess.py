"""Learned T2 guess: training driver and parameter file.

The network and its kernels live in libhelios_ccsd (csrc/hcc_mlp.hip); this
module only orchestrates.  Training data are the converged amplitudes that
the plain FP64 batched solver produces for molecules of the batched workload's
distribution, drawn with seeds disjoint from the benchmark batch (base seed
1000000 instead of 3):

    python -m helios_amd.learned_guess --train 20000     # on a GPU

writes helios_amd/data/guess_mlp.f32 (HCC_MLP_PARAMS little-endian float32)
and guess_mlp.json (how it was made, loss history).
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

from helios_amd import capi, molecules, ppp  # noqa: E402

DATA = os.path.join(HERE, "data")
PARAMS_PATH = os.path.join(DATA, "guess_mlp.f32")
TRAIN_BASE_SEED = 1_000_000


def load_params(path=PARAMS_PATH):
    """Trained parameters (float32, MLP_PARAMS); raises if the file is missing."""
    p = np.fromfile(path, dtype="<f4")
    if p.size != capi.MLP_PARAMS:
        raise ValueError(f"{path}: {p.size} parameters, expected {capi.MLP_PARAMS}")
    return p


def initial_params(seed=0):
    """He-style initial weights, zero biases, a small output layer."""
    rng = np.random.default_rng(seed)
    nf, nh = capi.MLP_FEATURES, 64
    parts = [rng.normal(0, np.sqrt(2.0 / nf), nh * nf), np.zeros(nh),
             rng.normal(0, np.sqrt(2.0 / nh), nh * nh), np.zeros(nh),
             rng.normal(0, np.sqrt(1.0 / nh), nh) * 0.1, np.zeros(1)]
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == capi.MLP_PARAMS
    return p


def _prep(k):
    m = molecules.batch_member(k, base_seed=TRAIN_BASE_SEED)
    p = ppp.prepare(m)
    return dict(C=p["C"], F=p["F"], G=p["G"], nocc=p["nocc"], eref=p["eref"], n=m.n)


def collect(trainer, count, pool, chunk=2048, per_molecule=256, eps=1e-9, log=print):
    """Solve `count` training molecules in chunks with the FP64 batched solver
    and append per_molecule samples of each converged one to the trainer.
    `pool` is a multiprocessing pool created before this process touched the
    GPU (forking after the HIP runtime is up is not supported)."""
    t0 = time.time()
    solved = 0
    for c0 in range(0, count, chunk):
        ks = range(c0, min(count, c0 + chunk))
        mols = pool.map(_prep, ks, chunksize=16)
        ctx = capi.Context(0, capi.HCC_FP64, 8)
        ctx.batch_load(mols)
        ctx.batch_solve(capi.GUESS_MP2, 500, eps)
        _, it, st = ctx.batch_results()
        trainer.collect(ctx, per_molecule, seed=c0)
        ctx.close()
        solved += int((st == 1).sum())
        log(f"# training molecules {c0 + len(ks)}/{count}: converged {int((st == 1).sum())}/{len(ks)}, "
            f"mean iterations {it.mean():.1f}, {time.time() - t0:.1f} s")
    return solved


def train(count=20000, per_molecule=256, epochs=60, batch=4096, lr=2e-3, lr_final=5e-5, seed=0,
          out=PARAMS_PATH, workers=16, log=print):
    # the workers fork here, before the trainer initialises the GPU
    with mp.Pool(workers) as pool:
        trainer = capi.Trainer(count * per_molecule)
        trainer.set_params(initial_params(seed))
        t0 = time.time()
        solved = collect(trainer, count, pool, per_molecule=per_molecule, log=log)
        t_collect = time.time() - t0
    loss0 = trainer.evaluate()
    t1 = time.time()
    hist = trainer.train(epochs, batch, lr, lr_final, seed)
    t_train = time.time() - t1
    p = trainer.params()
    loss1 = trainer.evaluate()
    os.makedirs(os.path.dirname(out), exist_ok=True)
    p.astype("<f4").tofile(out)
    meta = dict(molecules=count, converged=solved, base_seed=TRAIN_BASE_SEED, samples=trainer.count(),
                per_molecule=per_molecule, epochs=epochs, batch=batch, lr=lr, lr_final=lr_final, seed=seed,
                initial_loss=loss0, final_loss=loss1, epoch_loss=[float(x) for x in hist],
                collect_seconds=t_collect, train_seconds=t_train,
                target="y = Dp (t_ccsd - t_mp2) in eV, mean squared error")
    with open(os.path.splitext(out)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    log(f"# trained: loss {loss0:.4e} -> {loss1:.4e} in {t_train:.1f} s ({trainer.count()} samples)")
    trainer.close()
    return p, meta


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--train", type=int, default=20000, help="training molecules")
    ap.add_argument("--per-molecule", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-final", type=float, default=5e-5)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--out", default=PARAMS_PATH)
    a = ap.parse_args()
    train(a.train, a.per_molecule, a.epochs, a.batch, a.lr, a.lr_final, a.seed, a.out, a.workers,
          log=lambda s: print(s, flush=True))


if __name__ == "__main__":
    main()
