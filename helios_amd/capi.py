"""ctypes binding of libhelios_ccsd (include/helios_ccsd.h).

This is the binding a Python caller would add; it loads the in-tree library
and raises if it is missing -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhelios_ccsd.so")

HCC_FP64 = 0
HCC_MIXED = 1
GUESS_REFERENCE = 0
GUESS_MP2 = 1
GUESS_LEARNED = 2
GUESS_CURRENT = 3   # the amplitudes already in the context (warm start)
MLP_PARAMS = 5313
MLP_FEATURES = 16
MLP_ROW = 18

_P = ctypes.POINTER
_d = ctypes.c_double
_i = ctypes.c_int
_vp = ctypes.c_void_p
_f = ctypes.c_float
_l = ctypes.c_long

# name -> (restype, argtypes); mirrors include/helios_ccsd.h
SIGNATURES = {
    "hcc_create": (_i, [_P(_vp), _i, _i, _i]),
    "hcc_set_options": (_i, [_vp, _d, _d, _i, _i]),
    "hcc_set_mixed": (_i, [_vp, _i, _d, _i, _i]),
    "hcc_set_iterator": (_i, [_vp, _i, _d, _i, _d, _i, _i]),
    "hcc_request_stop": (_i, [_vp]),
    "hcc_batch_statistics": (_i, [_vp, _P(_i), _P(_i)]),
    "hcc_set_molecule": (_i, [_vp, _i, _i, _P(_d), _P(_d), _P(_d), _d]),
    "hcc_guess": (_i, [_vp, _i]),
    "hcc_iterate": (_i, [_vp, _i, _d, _P(_d)]),
    "hcc_energy": (_i, [_vp, _P(_d)]),
    "hcc_get_amplitudes": (_i, [_vp, _P(_d), _P(_d)]),
    "hcc_set_amplitudes": (_i, [_vp, _P(_d), _P(_d)]),
    "hcc_residual": (_i, [_vp, _P(_d), _P(_d)]),
    "hcc_solve": (_i, [_vp, _i, _i, _d, _P(_d), _P(_i), _P(_i)]),
    "hcc_batch_load": (_i, [_vp, _i, _P(_i), _P(_i), _P(_d), _P(_d), _P(_d), _P(_d)]),
    "hcc_batch_solve": (_i, [_vp, _i, _i, _d]),
    "hcc_batch_get_orbitals": (_i, [_vp, _i, _P(_d), _P(_d), _P(_d)]),
    "hcc_batch_triples": (_i, [_vp, _P(_d), _P(_d)]),
    "hcc_batch_lr_sigma": (_i, [_vp, _d, _P(_d), _P(_d), _P(_d), _P(_d)]),
    "hcc_batch_get_amplitudes": (_i, [_vp, _i, _P(_d), _P(_d)]),
    "hcc_batch_set_amplitudes": (_i, [_vp, _i, _P(_d), _P(_d)]),
    "hcc_lr_space": (_i, [_vp, _i]),
    "hcc_lr_put": (_i, [_vp, _i, _i, _P(_d)]),
    "hcc_lr_get": (_i, [_vp, _i, _i, _P(_d)]),
    "hcc_lr_dots": (_i, [_vp, _i, _P(_i), _i, _P(_d)]),
    "hcc_lr_combine": (_i, [_vp, _i, _P(_i), _P(_d), _P(_i), _P(_d)]),
    "hcc_lr_gram": (_i, [_vp, _i, _P(_i), _i, _P(_i), _P(_d)]),
    "hcc_lr_combine_block": (_i, [_vp, _i, _P(_i), _P(_d), _i, _P(_i), _i, _P(_d)]),
    "hcc_lr_precondition": (_i, [_vp, _i, _P(_i), _P(_d), _d]),
    "hcc_lr_sigma_slots": (_i, [_vp, _d, _P(_i), _P(_i)]),
    "hcc_lr_space_counts": (_i, [_vp, _P(_i)]),
    "hcc_lr_sigma_from": (_i, [_vp, _d, _P(_i), _P(_i), _P(_i)]),
    "hcc_batch_load_scf": (_i, [_vp, _i, _P(_i), _P(_i), _P(_d), _P(_d), _P(_d), _d, _i, _P(_d), _P(_i), _P(_i)]),
    "hcc_batch_results": (_i, [_vp, _P(_d), _P(_i), _P(_i)]),
    "hcc_last_timing": (_i, [_vp, _P(_d), _P(_d)]),
    "hcc_profile": (_i, [_vp, _i]),
    "hcc_profile_read": (_i, [_vp, _P(_d), _P(ctypes.c_long)]),
    "hcc_gemm_flops": (_i, [_vp, _P(_d), _P(_d)]),
    "hcc_gemm_flops_correction": (_i, [_vp, _P(_d)]),
    "hcc_profile_ops": (_i, [_vp, _P(_d)]),
    "hcc_profile_launches": (_i, [_vp, _P(_l)]),
    "hcc_destroy": (None, [_vp]),
    "hcc_solve_noalloc": (_i, [_i, _i, _i, _P(_d), _P(_d), _P(_d), _d, _i, _i, _d, _i, _P(_d), _P(_i), _P(_i),
                               _P(_d), _P(_d), ctypes.c_char_p, _i]),
    "hcc_last_error": (ctypes.c_char_p, [_vp]),
    "hcc_selftest_gemm": (_i, [_i, _i, _i, _P(_d), _P(_d), _P(_d), _d, _d, _i]),
    "hcc_selftest_gemm_correction": (_i, [_i, _i, _i, _i, _i, _P(_d), _P(_f), _P(_d), _P(_f), _P(_f), _d, _d, _i,
                                          _i, _P(_d)]),
    "hcc_set_guess_network": (_i, [_vp, _P(_f), _i]),
    "hcc_set_gemm_tuning": (_i, [_vp, _i, _P(_i)]),
    "hcc_set_mixed_history": (_i, [_vp, _i]),
    "hcc_set_mixed_diis": (_i, [_vp, _i]),
    "hcc_set_graphs": (_i, [_vp, _i]),
    "hcc_tune_gemm": (_i, [_vp, _i, _i, _i, _i, _P(_d)]),
    "hcc_gemm_shape": (_i, [_vp, _i, _i, _P(_i)]),
    "hcc_mlp_create": (_i, [_P(_vp), _i, _l]),
    "hcc_mlp_destroy": (None, [_vp]),
    "hcc_mlp_last_error": (ctypes.c_char_p, [_vp]),
    "hcc_mlp_set_params": (_i, [_vp, _P(_f), _i]),
    "hcc_mlp_get_params": (_i, [_vp, _P(_f), _i]),
    "hcc_mlp_collect": (_i, [_vp, _vp, _i, ctypes.c_uint, _P(_l)]),
    "hcc_mlp_samples": (_i, [_vp, _P(_l)]),
    "hcc_mlp_read_samples": (_i, [_vp, _l, _l, _P(_f), _P(_i)]),
    "hcc_mlp_write_samples": (_i, [_vp, _P(_f), _l]),
    "hcc_mlp_gradient": (_i, [_vp, _l, _i, _P(_d), _P(_f)]),
    "hcc_mlp_train": (_i, [_vp, _i, _i, _f, _f, ctypes.c_uint, _P(_d)]),
    "hcc_mlp_evaluate": (_i, [_vp, _P(_d)]),
    "hcc_mlp_predict": (_i, [_vp, _P(_f), _l, _P(_f)]),
}

_lib = None


def load(path=LIB_PATH):
    """Load the library (once).  Raises OSError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise OSError(f"libhelios_ccsd.so not built at {path}; run helios_amd.build")
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _dp(a):
    return a.ctypes.data_as(_P(_d))


def _ip(a):
    return a.ctypes.data_as(_P(_i))


def _fp(a):
    return a.ctypes.data_as(_P(_f))


def fortran(a):
    """Column-major float64 copy (the boundary's storage order)."""
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


class HccError(RuntimeError):
    pass


class Context:
    """Thin owner of an hcc_ctx; every method is one C-ABI call."""

    def __init__(self, device=0, precision=HCC_FP64, diis_steps=20):
        self.lib = load()
        self.h = _vp()
        rc = self.lib.hcc_create(ctypes.byref(self.h), device, precision, diis_steps)
        if rc != 0:
            raise HccError(f"hcc_create failed ({rc})")
        self.o = self.v = None
        self.nmol = 0

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.hcc_last_error(self.h)
            raise HccError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        if self.h:
            self.lib.hcc_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_options(self, t1_step=1.0, t2_step=1.0, projection_type=2, diis=True):
        self._check(self.lib.hcc_set_options(self.h, t1_step, t2_step, projection_type, int(diis)),
                    "hcc_set_options")

    def set_iterator(self, check_divergence=True, divergence=1e3, check_stagnation=True, stagnation=60.0,
                     stagnation_after=200, enable_stop=True):
        """The reference's iteration block (definebd.f90:53-58)."""
        self._check(self.lib.hcc_set_iterator(self.h, int(check_divergence), float(divergence),
                                              int(check_stagnation), float(stagnation), int(stagnation_after),
                                              int(enable_stop)), "hcc_set_iterator")

    def request_stop(self):
        """Asks a running solve (from another thread) to stop with status -4."""
        self._check(self.lib.hcc_request_stop(self.h), "hcc_request_stop")

    def set_mixed(self, bf16=True, kappa=0.05, max_inner=5, max_outer=30):
        self._check(self.lib.hcc_set_mixed(self.h, int(bf16), kappa, max_inner, max_outer), "hcc_set_mixed")

    def batch_statistics(self):
        outer = np.zeros(self.nmol, dtype=np.int32)
        inner = np.zeros(self.nmol, dtype=np.int32)
        self._check(self.lib.hcc_batch_statistics(self.h, _ip(outer), _ip(inner)), "hcc_batch_statistics")
        return outer, inner

    def set_molecule(self, C, F, G, nocc, eref):
        n = C.shape[0]
        self._C, self._F, self._G = fortran(C), fortran(F), fortran(G)
        self._check(self.lib.hcc_set_molecule(self.h, n, nocc, _dp(self._C), _dp(self._F), _dp(self._G),
                                              float(eref)), "hcc_set_molecule")
        self.o, self.v = nocc, n - nocc
        self.nmol = 1

    def guess(self, kind=GUESS_REFERENCE):
        self._check(self.lib.hcc_guess(self.h, kind), "hcc_guess")

    def iterate(self, iteration, eps):
        acc = np.zeros(5)
        self._check(self.lib.hcc_iterate(self.h, iteration, eps, _dp(acc)), "hcc_iterate")
        return acc

    def energy(self):
        e = np.zeros(5)
        self._check(self.lib.hcc_energy(self.h, _dp(e)), "hcc_energy")
        return e

    def amplitudes(self):
        """(t1[i,a], t2[i,j,a,b]) as numpy arrays (converted from Fortran order)."""
        t1 = np.zeros((self.o, self.v), order="F")
        t2 = np.zeros((self.o, self.o, self.v, self.v), order="F")
        self._check(self.lib.hcc_get_amplitudes(self.h, _dp(t1), _dp(t2)), "hcc_get_amplitudes")
        return np.ascontiguousarray(t1), np.ascontiguousarray(t2)

    def set_amplitudes(self, t1, t2):
        a, b = fortran(t1), fortran(t2)
        self._check(self.lib.hcc_set_amplitudes(self.h, _dp(a), _dp(b)), "hcc_set_amplitudes")

    def residual(self):
        d1 = np.zeros((self.o, self.v), order="F")
        d2 = np.zeros((self.o, self.o, self.v, self.v), order="F")
        self._check(self.lib.hcc_residual(self.h, _dp(d1), _dp(d2)), "hcc_residual")
        return np.ascontiguousarray(d1), np.ascontiguousarray(d2)

    def solve(self, guess=GUESS_REFERENCE, maxiter=2000, eps=1e-10):
        e = np.zeros(5)
        it = ctypes.c_int()
        st = ctypes.c_int()
        self._check(self.lib.hcc_solve(self.h, guess, maxiter, eps, _dp(e), ctypes.byref(it), ctypes.byref(st)),
                    "hcc_solve")
        return e, it.value, st.value

    def batch_load(self, mols):
        """mols: list of dicts with C, F, G (n x n), nocc, eref."""
        n = np.array([m["C"].shape[0] for m in mols], dtype=np.int32)
        no = np.array([m["nocc"] for m in mols], dtype=np.int32)
        C = np.concatenate([fortran(m["C"]).ravel(order="F") for m in mols])
        F = np.concatenate([fortran(m["F"]).ravel(order="F") for m in mols])
        G = np.concatenate([fortran(m["G"]).ravel(order="F") for m in mols])
        er = np.array([m["eref"] for m in mols], dtype=np.float64)
        self._check(self.lib.hcc_batch_load(self.h, len(mols), _ip(n), _ip(no), _dp(C), _dp(F), _dp(G), _dp(er)),
                    "hcc_batch_load")
        self.nmol = len(mols)

    def batch_load_scf(self, mols, accuracy=1e-10, maxiter=400):
        """mols: list of dicts with core, G, huckel (n x n) and nocc; the RHF runs
        on the device.  Returns (e_ref, SCF iterations, SCF status) per molecule."""
        n = np.array([m["core"].shape[0] for m in mols], dtype=np.int32)
        no = np.array([m["nocc"] for m in mols], dtype=np.int32)
        core = np.concatenate([fortran(m["core"]).ravel(order="F") for m in mols])
        G = np.concatenate([fortran(m["G"]).ravel(order="F") for m in mols])
        hk = np.concatenate([fortran(m["huckel"]).ravel(order="F") for m in mols])
        er = np.zeros(len(mols))
        it = np.zeros(len(mols), dtype=np.int32)
        st = np.zeros(len(mols), dtype=np.int32)
        self._check(self.lib.hcc_batch_load_scf(self.h, len(mols), _ip(n), _ip(no), _dp(core), _dp(G), _dp(hk),
                                                float(accuracy), int(maxiter), _dp(er), _ip(it), _ip(st)),
                    "hcc_batch_load_scf")
        self.nmol = len(mols)
        return er, it, st

    def batch_triples(self):
        """(energies (nmol x 5) as the reference's energyCC('spin-r-ccsd(t)'),
        standard (E[T], E[ST]) (nmol x 2)) of the converged amplitudes."""
        e = np.zeros((self.nmol, 5))
        et = np.zeros((self.nmol, 2))
        self._check(self.lib.hcc_batch_triples(self.h, _dp(e), _dp(et)), "hcc_batch_triples")
        return e, et

    def batch_lr_sigma(self, r1, r2, h=1e-4):
        """sigma = J r for one trial vector per loaded molecule: r1, r2 are the
        molecules' [i][a] and [i][j][a][b] blocks concatenated (float64)."""
        r1 = np.ascontiguousarray(r1, dtype=np.float64)
        r2 = np.ascontiguousarray(r2, dtype=np.float64)
        s1 = np.zeros_like(r1)
        s2 = np.zeros_like(r2)
        self._check(self.lib.hcc_batch_lr_sigma(self.h, float(h), _dp(r1), _dp(r2), _dp(s1), _dp(s2)),
                    "hcc_batch_lr_sigma")
        return s1, s2

    def batch_amplitudes(self, mol, o, v, doubles=True):
        """(t1[i,a], t2[i,j,a,b] or None) of loaded molecule `mol`."""
        t1 = np.zeros((o, v), order="F")
        t2 = np.zeros((o, o, v, v), order="F") if doubles else None
        self._check(self.lib.hcc_batch_get_amplitudes(self.h, int(mol), _dp(t1), _dp(t2) if doubles else None),
                    "hcc_batch_get_amplitudes")
        return np.ascontiguousarray(t1), (np.ascontiguousarray(t2) if doubles else None)

    def batch_set_amplitudes(self, mol, t1, t2):
        """Amplitudes t1[i,a], t2[i,j,a,b] of loaded molecule `mol` (for GUESS_CURRENT)."""
        a, b = fortran(t1), fortran(t2)
        self._check(self.lib.hcc_batch_set_amplitudes(self.h, int(mol), _dp(a), _dp(b)), "hcc_batch_set_amplitudes")

    # device-resident vector space of the LR-CCSD eigensolver
    def lr_space(self, nslots):
        self._check(self.lib.hcc_lr_space(self.h, int(nslots)), "hcc_lr_space")

    def lr_put(self, mol, slot, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        self._check(self.lib.hcc_lr_put(self.h, int(mol), int(slot), _dp(x)), "hcc_lr_put")

    def lr_get(self, mol, slot, n):
        x = np.zeros(n)
        self._check(self.lib.hcc_lr_get(self.h, int(mol), int(slot), _dp(x)), "hcc_lr_get")
        return x

    def lr_dots(self, jobs, mode=0):
        """jobs: (J, 3) ints (mol, slot a, slot b) -> J values."""
        jobs = np.ascontiguousarray(jobs, dtype=np.int32).reshape(-1, 3)
        out = np.zeros(len(jobs))
        if len(jobs):
            self._check(self.lib.hcc_lr_dots(self.h, len(jobs), _ip(jobs), int(mode), _dp(out)), "hcc_lr_dots")
        return out

    def lr_combine(self, jobs, beta, term_slot, term_coef):
        """jobs: (J, 4) ints (mol, dst, first term, count); dst = beta dst + sum coef * slot."""
        jobs = np.ascontiguousarray(jobs, dtype=np.int32).reshape(-1, 4)
        if not len(jobs):
            return
        beta = np.ascontiguousarray(beta, dtype=np.float64)
        ts = np.ascontiguousarray(term_slot if len(term_slot) else [0], dtype=np.int32)
        tc = np.ascontiguousarray(term_coef if len(term_coef) else [0.0], dtype=np.float64)
        self._check(self.lib.hcc_lr_combine(self.h, len(jobs), _ip(jobs), _dp(beta), _ip(ts), _dp(tc)),
                    "hcc_lr_combine")

    def lr_gram(self, jobs, slots):
        """jobs: (J, 5) ints (mol, na, nb, a, b) over the slot list ``slots``
        -> the jobs' na x nb Gram blocks, row-major, concatenated."""
        jobs = np.ascontiguousarray(jobs, dtype=np.int32).reshape(-1, 5)
        slots = np.ascontiguousarray(slots if len(slots) else [0], dtype=np.int32)
        out = np.zeros(int(np.sum(jobs[:, 1].astype(np.int64) * jobs[:, 2])) if len(jobs) else 0)
        if len(jobs):
            self._check(self.lib.hcc_lr_gram(self.h, len(jobs), _ip(jobs), len(slots), _ip(slots), _dp(out)),
                        "hcc_lr_gram")
        return out

    def lr_combine_block(self, jobs, beta, slots, coef):
        """jobs: (J, 6) ints (mol, m, t, d, s, c): slots[d + p] = beta slots[d + p]
        + sum_q coef[c + p t + q] slots[s + q]."""
        jobs = np.ascontiguousarray(jobs, dtype=np.int32).reshape(-1, 6)
        if not len(jobs):
            return
        beta = np.ascontiguousarray(beta, dtype=np.float64)
        slots = np.ascontiguousarray(slots if len(slots) else [0], dtype=np.int32)
        coef = np.ascontiguousarray(coef if len(coef) else [0.0], dtype=np.float64)
        self._check(self.lib.hcc_lr_combine_block(self.h, len(jobs), _ip(jobs), _dp(beta), len(slots), _ip(slots),
                                                  len(coef), _dp(coef)), "hcc_lr_combine_block")

    def lr_precondition(self, jobs, shift, floor):
        jobs = np.ascontiguousarray(jobs, dtype=np.int32).reshape(-1, 2)
        if not len(jobs):
            return
        shift = np.ascontiguousarray(shift, dtype=np.float64)
        self._check(self.lib.hcc_lr_precondition(self.h, len(jobs), _ip(jobs), _dp(shift), float(floor)),
                    "hcc_lr_precondition")

    def lr_space_counts(self, counts):
        c = np.ascontiguousarray(counts, dtype=np.int32)
        self._check(self.lib.hcc_lr_space_counts(self.h, _ip(c)), "hcc_lr_space_counts")

    def lr_sigma_from(self, src, in_slot, out_slot, h):
        m = np.ascontiguousarray(src, dtype=np.int32)
        a = np.ascontiguousarray(in_slot, dtype=np.int32)
        b = np.ascontiguousarray(out_slot, dtype=np.int32)
        self._check(self.lib.hcc_lr_sigma_from(self.h, float(h), _ip(m), _ip(a), _ip(b)), "hcc_lr_sigma_from")

    def lr_sigma_slots(self, in_slot, out_slot, h):
        a = np.ascontiguousarray(in_slot, dtype=np.int32)
        b = np.ascontiguousarray(out_slot, dtype=np.int32)
        self._check(self.lib.hcc_lr_sigma_slots(self.h, float(h), _ip(a), _ip(b)), "hcc_lr_sigma_slots")

    def batch_orbitals(self, mol, n):
        """(C, F_mo, e_ref) of loaded molecule `mol` with n sites."""
        C = np.zeros((n, n), order="F")
        F = np.zeros((n, n), order="F")
        er = ctypes.c_double()
        self._check(self.lib.hcc_batch_get_orbitals(self.h, mol, _dp(C), _dp(F), ctypes.byref(er)),
                    "hcc_batch_get_orbitals")
        return np.ascontiguousarray(C), np.ascontiguousarray(F), er.value

    def batch_solve(self, guess=GUESS_MP2, maxiter=200, eps=1e-7):
        self._check(self.lib.hcc_batch_solve(self.h, guess, maxiter, eps), "hcc_batch_solve")

    def batch_results(self):
        e = np.zeros((self.nmol, 5))
        it = np.zeros(self.nmol, dtype=np.int32)
        st = np.zeros(self.nmol, dtype=np.int32)
        self._check(self.lib.hcc_batch_results(self.h, _dp(e), _ip(it), _ip(st)), "hcc_batch_results")
        return e, it, st

    def profile(self, enable=True):
        self._check(self.lib.hcc_profile(self.h, int(enable)), "hcc_profile")

    def profile_read(self):
        ms = ctypes.c_double()
        n = ctypes.c_long()
        self._check(self.lib.hcc_profile_read(self.h, ctypes.byref(ms), ctypes.byref(n)), "hcc_profile_read")
        return ms.value, n.value

    def profile_ops(self):
        """Profiled device ms per slot (96 slots, see hcc_profile_ops)."""
        ms = np.zeros(96)
        self._check(self.lib.hcc_profile_ops(self.h, _dp(ms)), "hcc_profile_ops")
        return ms

    def profile_launches(self):
        """GEMM kernel dispatches per profile slot (96) since profile(True)."""
        n = np.zeros(96, dtype=np.int64)
        self._check(self.lib.hcc_profile_launches(self.h, n.ctypes.data_as(ctypes.POINTER(ctypes.c_long))),
                    "hcc_profile_launches")
        return n

    def gemm_flops(self, nmol):
        it = np.zeros(nmol)
        pre = np.zeros(nmol)
        self._check(self.lib.hcc_gemm_flops(self.h, _dp(it), _dp(pre)), "hcc_gemm_flops")
        return it, pre

    def set_mixed_history(self, keep=True):
        """Mixed solver: keep the inner DIIS history across outer points."""
        self._check(self.lib.hcc_set_mixed_history(self.h, int(keep)), "hcc_set_mixed_history")

    def set_mixed_diis(self, depth=8):
        """Mixed solver: inner DIIS depth (library default 8; 0: diis_steps) for batches loaded afterwards."""
        self._check(self.lib.hcc_set_mixed_diis(self.h, int(depth)), "hcc_set_mixed_diis")

    def set_graphs(self, enable=True):
        self._check(self.lib.hcc_set_graphs(self.h, int(enable)), "hcc_set_graphs")

    def set_gemm_tuning(self, entries):
        """entries: (op, n_occ, n_virt, tile, ksplit) rows; applies to molecules loaded afterwards."""
        self._tune = np.ascontiguousarray(np.asarray(entries, dtype=np.int32).reshape(-1, 5))
        self._check(self.lib.hcc_set_gemm_tuning(self.h, self._tune.shape[0], _ip(self._tune)),
                    "hcc_set_gemm_tuning")

    def tune_gemm(self, op, tile=0, ksplit=0, reps=5):
        ms = ctypes.c_double()
        self._check(self.lib.hcc_tune_gemm(self.h, op, tile, ksplit, reps, ctypes.byref(ms)), "hcc_tune_gemm")
        return ms.value

    def gemm_shape(self, op, mol=0):
        mnk = np.zeros(3, dtype=np.int32)
        self._check(self.lib.hcc_gemm_shape(self.h, op, mol, _ip(mnk)), "hcc_gemm_shape")
        return tuple(int(x) for x in mnk)

    def set_guess_network(self, params):
        """Parameters (MLP_PARAMS float32) of the network used by GUESS_LEARNED."""
        self._w = np.ascontiguousarray(params, dtype=np.float32).ravel()
        self._check(self.lib.hcc_set_guess_network(self.h, _fp(self._w), self._w.size), "hcc_set_guess_network")

    def gemm_flops_correction(self, nmol):
        fl = np.zeros(nmol)
        self._check(self.lib.hcc_gemm_flops_correction(self.h, _dp(fl)), "hcc_gemm_flops_correction")
        return fl

    def last_timing(self):
        ms = ctypes.c_double()
        its = ctypes.c_double()
        self._check(self.lib.hcc_last_timing(self.h, ctypes.byref(ms), ctypes.byref(its)), "hcc_last_timing")
        return ms.value, its.value


class Trainer:
    """Owner of an hcc_mlp: device sample set, parameters and GPU training."""

    def __init__(self, capacity, device=0):
        self.lib = load()
        self.h = _vp()
        rc = self.lib.hcc_mlp_create(ctypes.byref(self.h), device, int(capacity))
        if rc != 0:
            raise HccError(f"hcc_mlp_create failed ({rc})")

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.hcc_mlp_last_error(self.h)
            raise HccError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        if self.h:
            self.lib.hcc_mlp_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32).ravel()
        self._check(self.lib.hcc_mlp_set_params(self.h, _fp(p), p.size), "hcc_mlp_set_params")

    def params(self):
        p = np.zeros(MLP_PARAMS, dtype=np.float32)
        self._check(self.lib.hcc_mlp_get_params(self.h, _fp(p), p.size), "hcc_mlp_get_params")
        return p

    def collect(self, ctx, per_molecule, seed=0):
        n = ctypes.c_long()
        self._check(self.lib.hcc_mlp_collect(self.h, ctx.h, per_molecule, seed, ctypes.byref(n)), "hcc_mlp_collect")
        return n.value

    def count(self):
        n = ctypes.c_long()
        self._check(self.lib.hcc_mlp_samples(self.h, ctypes.byref(n)), "hcc_mlp_samples")
        return n.value

    def read_samples(self, first=0, count=None):
        count = self.count() - first if count is None else count
        rows = np.zeros((count, MLP_ROW), dtype=np.float32)
        where = np.zeros((count, 2), dtype=np.int32)
        self._check(self.lib.hcc_mlp_read_samples(self.h, first, count, _fp(rows), _ip(where)),
                    "hcc_mlp_read_samples")
        return rows, where

    def write_samples(self, rows):
        rows = np.ascontiguousarray(rows, dtype=np.float32)
        self._check(self.lib.hcc_mlp_write_samples(self.h, _fp(rows), rows.shape[0]), "hcc_mlp_write_samples")

    def gradient(self, first=0, count=None):
        count = self.count() - first if count is None else count
        loss = ctypes.c_double()
        g = np.zeros(MLP_PARAMS, dtype=np.float32)
        self._check(self.lib.hcc_mlp_gradient(self.h, first, count, ctypes.byref(loss), _fp(g)), "hcc_mlp_gradient")
        return loss.value, g

    def train(self, epochs, batch=4096, lr=2e-3, lr_final=1e-4, seed=0):
        hist = np.zeros(epochs)
        self._check(self.lib.hcc_mlp_train(self.h, epochs, batch, lr, lr_final, seed, _dp(hist)), "hcc_mlp_train")
        return hist

    def evaluate(self):
        loss = ctypes.c_double()
        self._check(self.lib.hcc_mlp_evaluate(self.h, ctypes.byref(loss)), "hcc_mlp_evaluate")
        return loss.value

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.zeros(X.shape[0], dtype=np.float32)
        self._check(self.lib.hcc_mlp_predict(self.h, _fp(X), X.shape[0], _fp(y)), "hcc_mlp_predict")
        return y


def solve_noalloc(C, F, G, nocc, eref, guess=GUESS_MP2, maxiter=2000, eps=1e-10, diis_steps=20, device=0,
                  amplitudes=True):
    """hcc_solve_noalloc: (energy[5], iterations, status, t1, t2); raises HccError
    with the library's message on a nonzero return."""
    lib = load()
    n = C.shape[0]
    Cf, Ff, Gf = fortran(C), fortran(F), fortran(G)
    e = np.zeros(5)
    it = ctypes.c_int()
    st = ctypes.c_int()
    o, v = nocc, n - nocc
    t1 = np.zeros((o, v), order="F") if amplitudes else None
    t2 = np.zeros((o, o, v, v), order="F") if amplitudes else None
    err = ctypes.create_string_buffer(256)
    rc = lib.hcc_solve_noalloc(device, n, nocc, _dp(Cf), _dp(Ff), _dp(Gf), float(eref), guess, maxiter, eps,
                               diis_steps, _dp(e), ctypes.byref(it), ctypes.byref(st),
                               _dp(t1) if amplitudes else None, _dp(t2) if amplitudes else None, err, 256)
    if rc != 0:
        raise HccError(f"hcc_solve_noalloc failed ({rc}): {err.value.decode()}")
    if amplitudes:
        t1, t2 = np.ascontiguousarray(t1), np.ascontiguousarray(t2)
    return e, it.value, st.value, t1, t2


def selftest_gemm_correction(A0, dA, B0, dB, Cd, alpha=1.0, beta=0.0, a_kfast=True, b_nfast=True, tile=64, reps=0):
    """Cd' = alpha (A0 dB + dA (B0 + dB)) + beta Cd on the correction GEMM kernel.
    A0, dA: m x k; B0, dB: k x n (logical shapes; dA or dB may be None).
    Returns (Cd', mean ms of reps extra launches)."""
    lib = load()
    m, k = A0.shape
    n = B0.shape[1]
    lay_a = (lambda x: x) if a_kfast else (lambda x: x.T)
    lay_b = (lambda x: x) if b_nfast else (lambda x: x.T)
    a0 = np.ascontiguousarray(lay_a(A0), dtype=np.float64)
    b0 = np.ascontiguousarray(lay_b(B0), dtype=np.float64)
    da = None if dA is None else np.ascontiguousarray(lay_a(dA), dtype=np.float32)
    db = None if dB is None else np.ascontiguousarray(lay_b(dB), dtype=np.float32)
    cd = np.ascontiguousarray(Cd, dtype=np.float32).copy()
    fp = lambda x: None if x is None else x.ctypes.data_as(_P(_f))
    ms = ctypes.c_double(0.0)
    rc = lib.hcc_selftest_gemm_correction(m, n, k, int(a_kfast), int(b_nfast), _dp(a0), fp(da), _dp(b0), fp(db),
                                          fp(cd), alpha, beta, tile, reps, ctypes.byref(ms))
    if rc != 0:
        raise HccError(f"hcc_selftest_gemm_correction failed ({rc})")
    return cd, ms.value


def selftest_gemm(A, B, C=None, alpha=1.0, beta=0.0, precision=HCC_FP64):
    lib = load()
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    m, k = A.shape
    n = B.shape[1]
    C = np.zeros((m, n)) if C is None else np.ascontiguousarray(C, dtype=np.float64).copy()
    rc = lib.hcc_selftest_gemm(m, n, k, _dp(A), _dp(B), _dp(C), alpha, beta, precision)
    if rc != 0:
        raise HccError(f"hcc_selftest_gemm failed ({rc})")
    return C
