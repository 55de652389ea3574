"""Device iterator statuses and coupled-cluster options through the C ABI,
against the oracle's iterator (src/calculation/gIterator.f90:88-210, statuses
src/modules/mod_helios.f90:119-120) and its iterationCC with non-default steps
and projection (src/methods/coupledCluster/mod_cc.f90:987-1017).

* -1: an iteration cap of 1 and of 5 (energies after 5 iterations as the oracle's);
* -2: DIIS off, so the x3 overshoot of the '2-1' update diverges past a
  divergence threshold of 1e2 -- at the oracle's iteration;
* -3: a residual threshold no solve reaches, stagnation checked after 50
  iterations at 70 % improving steps;
* -4: a stop request from another thread while a solve runs; a request made
  before a solve is cleared by it; with enable-stop-iteration off it is ignored;
* t1/t2 steps 0.8/0.6 and projection type 1 step for step against the oracle,
  and their converged energies;
* a rejected load leaves no half-built batch behind (HCC_E_STATE afterwards).
"""
import ctypes
import threading
import time

import numpy as np
import pytest

from helios_amd import capi, molecules as M
from oracle import ccsd
from tests._molprep import prepare

pytestmark = pytest.mark.gpu


def single(p, diis_steps=20):
    ctx = capi.Context(0, capi.HCC_FP64, diis_steps)
    ctx.set_molecule(p["C"], p["F"], p["G"], p["nocc"], p["eref"])
    return ctx


def oracle_run(p, maxiter, eps, **kw):
    it_kw = {k: kw.pop(k) for k in list(kw) if k in ("divergence", "stagnation_after", "stagnation_percent",
                                                     "check_divergence", "check_stagnation")}
    cc = ccsd.CCSD(p["C"], p["F"], p["G"], p["nocc"], p["eref"], diis_scale=True, **kw)
    st, it, en, _ = ccsd.iterator(cc.iterate, cc.energy, maxiter, eps, **it_kw)
    return st, it, en


def test_status_iteration_cap():
    p = prepare(M.polyene(6))
    ctx = single(p)
    for cap in (1, 5):
        e, it, st = ctx.solve(capi.GUESS_REFERENCE, cap, 1e-10)
        rs, rit, ren = oracle_run(p, cap, 1e-10)
        print("cap", cap, st, it, rs, rit, e[0], ren[0])
        assert (st, it) == (-1, cap) == (rs, rit)
        assert abs(e[0] - ren[0]) < 1e-9
    ctx.close()


def test_status_diverged_without_diis():
    p = prepare(M.polyene(6))
    ctx = single(p)
    ctx.set_options(1.0, 1.0, 2, False)
    ctx.set_iterator(divergence=1e2)
    e, it, st = ctx.solve(capi.GUESS_REFERENCE, 200, 1e-10)
    rs, rit, _ = oracle_run(p, 200, 1e-10, diis=False, divergence=1e2)
    ctx.close()
    print("diverged", st, it, "oracle", rs, rit)
    assert st == -2 and rs == -2
    assert it == rit


def test_status_stagnated():
    p = prepare(M.octatetraene())
    ctx = single(p)
    ctx.set_iterator(stagnation=70.0, stagnation_after=50)
    e, it, st = ctx.solve(capi.GUESS_REFERENCE, 2000, 1e-30)
    rs, rit, _ = oracle_run(p, 2000, 1e-30, stagnation_percent=70.0, stagnation_after=50)
    ctx.close()
    print("stagnated", st, it, "oracle", rs, rit)
    assert st == -3 and rs == -3
    assert it > 50 and rit > 50


def test_status_interrupted_by_stop_request():
    ps = [prepare(M.batch_member(k)) for k in range(8)]
    ctx = capi.Context(0, capi.HCC_FP64, 8)
    ctx.set_iterator(check_stagnation=False)
    ctx.batch_load(ps)
    # a request made before a solve is cleared when the solve starts
    ctx.request_stop()
    ctx.batch_solve(capi.GUESS_MP2, 500, 1e-9)
    _, it, st = ctx.batch_results()
    assert (st == 1).all()
    # a request during a solve that would not end by itself (residual threshold 0)
    timer = threading.Timer(1.0, ctx.request_stop)
    t0 = time.time()
    timer.start()
    ctx.batch_solve(capi.GUESS_MP2, 20000, 0.0)
    took = time.time() - t0
    timer.join()
    e, it, st = ctx.batch_results()
    print("stopped after", took, "s, iterations", it, "statuses", st)
    assert (st == -4).all()
    assert (it < 20000).all() and np.isfinite(e).all()
    ctx.close()
    # enable-stop-iteration off: the request is ignored
    ctx = single(ps[0])
    ctx.set_iterator(check_stagnation=False, enable_stop=False)
    timer = threading.Timer(0.2, ctx.request_stop)
    timer.start()
    e, it, st = ctx.solve(capi.GUESS_MP2, 300, 0.0)
    timer.join()
    ctx.close()
    assert (st, it) == (-1, 300)


@pytest.mark.parametrize("s1,s2,proj", [(0.8, 0.6, 2), (1.0, 1.0, 1)])
def test_steps_and_projection_type(s1, s2, proj):
    p = prepare(M.octatetraene())
    o = p["nocc"]
    cc = ccsd.CCSD(p["C"], p["F"], p["G"], o, p["eref"], diis_scale=True, s1=s1, s2=s2, projtype=proj)
    ctx = single(p)
    ctx.set_options(s1, s2, proj, True)
    ctx.guess(capi.GUESS_REFERENCE)
    for it in range(1, 9):
        acc = ctx.iterate(it, 1e-12)
        ref = cc.iterate(it, 1e-12)
        assert abs(acc[0] - ref[0]) < 1e-9 * max(1, ref[0])
        assert abs(acc[1] - ref[1]) < 1e-9 * max(1, ref[1])
        assert abs(ctx.energy()[0] - cc.energy()[0]) < 1e-9
    e, its, st = ctx.solve(capi.GUESS_REFERENCE, 2000, 1e-10)
    rs, rit, ren = oracle_run(p, 2000, 1e-10, s1=s1, s2=s2, projtype=proj)
    ctx.close()
    print("steps", s1, s2, "projection", proj, "iterations", its, rit, "E", e[0], ren[0])
    assert st == 1 and rs == 1
    assert abs(e[0] - ren[0]) < 1e-8


def test_rejected_load_leaves_no_batch():
    ctx = capi.Context(0, capi.HCC_FP64, 8)
    p = prepare(M.octatetraene())
    bad = dict(p, nocc=p["n"])            # no virtual orbitals
    with pytest.raises(capi.HccError):
        ctx.batch_load([p, bad])
    assert ctx.lib.hcc_batch_solve(ctx.h, capi.GUESS_MP2, 10, 1e-8) == -3      # HCC_E_STATE
    d1 = np.zeros(16)
    d2 = np.zeros(256)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))   # noqa: E731
    assert ctx.lib.hcc_residual(ctx.h, dp(d1), dp(d2)) == -3
    # a good load afterwards works
    ctx.batch_load([p])
    ctx.batch_solve(capi.GUESS_MP2, 500, 1e-9)
    _, _, st = ctx.batch_results()
    ctx.close()
    assert st[0] == 1
