// Per-molecule descriptors and solver state shared by the host driver and the
// elementwise kernels of libhelios_ccsd.
#pragma once
#include "hcc_internal.h"

namespace hcc {

// Device pointers of one molecule's blocks in the per-buffer arenas.  V is
// the element type of the buffers that change with the amplitudes: double for
// the solver state (Mol), float for the corrections of the mixed-precision
// solver (MolD), whose inputs and constants stay FP64.  The DIIS vectors
// follow V (MolD's alias the FP64 solver's storage, unused by the mixed
// solver); the B matrix stays FP64.
template <class V>
struct MolT {
  int n, o, v, m;          // sites, occupied, virtual, DIIS depth
  int nd;                  // packed DIIS vector length: ov + ov(ov+1)/2
  int pad0;
  double eref;
  // inputs (row-major)
  double *C, *G, *F, *Fao;
  // dressing (PT[mu][p][q] = P[mu][q][p])
  V *Xf, *Yf, *Mf, *P, *PT, *Q, *Ft, *tmp;
  // amplitudes and residuals
  V *t1, *t2, *u, *w1, *w2, *x, *Zc, *WC, *WD;
  // integrals and intermediates
  V *G4, *Wkl;
  double *Gklcd, *Lldkc, *Lt;
  V *gvvov, *gooov, *Xvv, *Xoo;
  // DIIS history: m slots of nd (amplitudes st, errors sd); B matrix m x m
  V *st, *sd;
  double *Bm;
  // guess features / MP2 amplitudes (o^2 v^2)
  double *t2mp2;
  // dressed-density diagonal and its Coulomb field (2n)
  V *dvec;
  // site-localisation overlaps of the MOs (n x n), features of the learned guess
  double *Ov;
};
using Mol = MolT<double>;
using MolD = MolT<float>;

// Per-molecule iteration state (reference iterator semantics,
// src/calculation/gIterator.f90:88-210, statuses mod_helios.f90:119-120).
struct State {
  int iter;          // iterations done
  int status;        // 0 running, 1 converged, -1 max-iter, -2 diverged, -3 stagnated, -4 stopped
  int active;        // 1 while iterating
  int successful;    // improving iterations (starts at 1)
  int diis_count;    // vectors in history
  int diis_next;     // next ring slot
  int staged;        // 1 when the current iteration staged a DIIS vector
  int inner;         // mixed solver: 1 while the inner (correction) solve runs
  int inner_its;     // mixed solver: inner iterations of the current outer step
  int outer_its;     // mixed solver: FP64 residual evaluations
  double acc0;       // mixed solver: max|d| of the FP64 residual at the outer point
  double acc[2];     // max|d1|, max|d2|
  double cur, prev;  // current / previous max accuracy
  double energy[5];
  double coef[64];   // DIIS coefficients per slot
};

struct Params {
  double eps;            // convergence threshold on max(acc)
  double kappa;          // mixed solver: inner solve stops at acc < kappa * acc0
  int max_inner;         // mixed solver: inner iterations per outer step
  int keep_diis;         // mixed solver: the inner DIIS history survives outer steps
  int outer_only;        // mixed solver: the update reads the outer residual only (dt = 0 step)
  double divergence;     // divergence threshold
  double s1, s2;         // t1/t2 steps
  int projtype;          // 1 or 2 (the reference's '2-1')
  int diis;              // DIIS enabled
  int maxiter;
  int stagnation_after;  // iterations before the stagnation test
  double stagnation_pct; // success percentage at or below which the solve stagnates
  int check_divergence, check_stagnation;
};

// block -> molecule table for ragged elementwise launches
struct Ragged {
  int* prefix;   // device, nm+1 block prefix
  int* bmol;     // device, molecule of every block
  int nblocks;
};

}  // namespace hcc
