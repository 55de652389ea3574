// Bandwidth-bound kernels of the r-ccsd solver: T1 dressing, orbital
// products, residual assembly, the denominator update (fused with the DIIS
// push and the max-|residual| reduction), DIIS dot products / solve /
// extrapolation, energy, guesses and the per-molecule iterator control.
//
// Every kernel runs over a ragged batch: workgroups are assigned to molecules
// through a block prefix table, each workgroup serving one molecule, so a
// launch covers all molecules and no workgroup straddles two.  Reductions are
// two-pass (per-workgroup partials, then one fixed-order pass per molecule):
// results are bitwise reproducible run to run.
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "hcc_batch.h"
#include "hcc_ops.h"

namespace hcc {

__device__ __forceinline__ int block_mol(const int* __restrict__ prefix, int nm, int b) {
  int lo = 0, hi = nm - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

#define RAGGED_PROLOGUE(SIZE)                                         \
  const int mol = rg.bmol[blockIdx.x];                               \
  const int lb = blockIdx.x - rg.prefix[mol];                        \
  const auto& M = mols[mol];                                         \
  const long e = (long)lb * kEwThreads + threadIdx.x;                \
  const long size_ = (SIZE);                                         \
  (void)e;                                                           \
  (void)size_;

// Index decompositions in 32-bit unsigned arithmetic (64-bit division is a
// long software sequence on the GPU): every per-molecule array stays below
// 2^28 elements, so element indices within a molecule fit.
// (i, j, a, b) of element e of an o x o x v x v array (row-major)
__device__ __forceinline__ void ijab_of(long e, int o, int v, int& i, int& j, int& a, int& b) {
  const unsigned q1 = (unsigned)e / (unsigned)v, q2 = q1 / (unsigned)v, q3 = q2 / (unsigned)o;
  b = (int)((unsigned)e - q1 * (unsigned)v);
  a = (int)(q1 - q2 * (unsigned)v);
  j = (int)(q2 - q3 * (unsigned)o);
  i = (int)q3;
}
// Loads through global-address-space pointers: pointers read from a molecule
// descriptor are generic, and generic (flat) loads also count against the
// scalar-load counter, so each wait for a scalar load would drain them.
template <class T>
using gptr = const T __attribute__((address_space(1)))*;
template <class T>
using gwptr = T __attribute__((address_space(1)))*;
template <class T>
__device__ __forceinline__ gptr<T> global_view(const T* p) {
  return (gptr<T>)p;
}
template <class T>
__device__ __forceinline__ gwptr<T> global_out(T* p) {
  return (gwptr<T>)p;
}
// quotient and remainder of 0 <= x < 2^31 by d >= 1
__device__ __forceinline__ void divmod(long x, int d, int& q, int& r) {
  const unsigned u = (unsigned)x / (unsigned)d;
  q = (int)u;
  r = (int)((unsigned)x - u * (unsigned)d);
}

// ---------------------------------------------------------------------------
// T1 dressing
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kEwThreads) void k_dress(const Mol* __restrict__ mols, const State* __restrict__ st,
                                                     Ragged rg, int nm) {
  RAGGED_PROLOGUE((long)mols[mol].n * mols[mol].n);
  if (e >= size_ || !st[mol].active) return;
  const int N = M.n, o = M.o, v = M.v;
  const int mu = e / N, p = e % N;
  const double* C = M.C + (long)mu * N;
  double xf, yf;
  if (p < o) {
    double s = C[p];
    for (int a = 0; a < v; ++a) s += C[o + a] * M.t1[p * v + a];
    xf = C[p];
    yf = s;
  } else {
    const int a = p - o;
    double s = C[p];
    for (int k = 0; k < o; ++k) s -= C[k] * M.t1[k * v + a];
    xf = s;
    yf = C[p];
  }
  M.Xf[e] = xf;
  M.Yf[e] = yf;
}

// M = Fao - gamma o dD + 2 diag(gamma diag(dD)),  dD = (Yo - Co) Co^T
__device__ __forceinline__ double dressed_density(const Mol& M, int mu, int nu) {
  const double* Y = M.Yf + (long)mu * M.n;
  const double* Cm = M.C + (long)mu * M.n;
  const double* Cn = M.C + (long)nu * M.n;
  double s = 0.0;
  for (int k = 0; k < M.o; ++k) s += (Y[k] - Cm[k]) * Cn[k];
  return s;
}

// dvec[mu] = dD[mu][mu]
__global__ __launch_bounds__(kEwThreads) void k_fock_diag(const Mol* __restrict__ mols, const State* __restrict__ st,
                                                         Ragged rg, int nm) {
  RAGGED_PROLOGUE((long)mols[mol].n);
  if (e >= size_ || !st[mol].active) return;
  M.dvec[e] = dressed_density(M, (int)e, (int)e);
}

__global__ __launch_bounds__(kEwThreads) void k_fock_m(const Mol* __restrict__ mols, const State* __restrict__ st,
                                                      Ragged rg, int nm) {
  RAGGED_PROLOGUE((long)mols[mol].n * mols[mol].n);
  if (e >= size_ || !st[mol].active) return;
  const int N = M.n;
  const int mu = e / N, nu = e % N;
  double val = M.Fao[e] - M.G[e] * dressed_density(M, mu, nu);
  if (mu == nu) {
    double s = 0.0;
    for (int l = 0; l < N; ++l) s += M.G[(long)mu * N + l] * M.dvec[l];
    val += 2.0 * s;
  }
  M.Mf[e] = val;
}

// P[mu][p*N+q] = Xf[mu][p] * Yf[mu][q] and its transpose PT[mu][p*N+q] = P[mu][q*N+p];
// kPItems elements per thread (the ragged table of these kernels counts
// kEwThreads * kPItems elements per workgroup)
__global__ __launch_bounds__(kEwThreads) void k_pouter(const Mol* __restrict__ mols, const State* __restrict__ st,
                                                      Ragged rg, int nm) {
  RAGGED_PROLOGUE((long)mols[mol].n * mols[mol].n * mols[mol].n);
  if (!st[mol].active) return;
  const int N = M.n, N2 = N * N;
  const gptr<double> X = global_view(M.Xf), Y = global_view(M.Yf);
  const gwptr<double> P = global_out(M.P), PT = global_out(M.PT);
  const long f0 = (long)lb * kPItems * kEwThreads + threadIdx.x;
  // all loads of the thread's elements first, then the stores (elements past
  // the molecule's block read element 0 and store nothing)
  double xp[kPItems], yq[kPItems], xq[kPItems], yp[kPItems];
#pragma unroll
  for (int it = 0; it < kPItems; ++it) {
    const long f = f0 + (long)it * kEwThreads;
    int mu, pq, p, q;
    divmod(f < size_ ? f : 0, N2, mu, pq);
    divmod(pq, N, p, q);
    xp[it] = X[mu * N + p];
    yq[it] = Y[mu * N + q];
    xq[it] = X[mu * N + q];
    yp[it] = Y[mu * N + p];
  }
#pragma unroll
  for (int it = 0; it < kPItems; ++it) {
    const long f = f0 + (long)it * kEwThreads;
    if (f < size_) {
      P[f] = xp[it] * yq[it];
      PT[f] = xq[it] * yp[it];
    }
  }
}

// u = 2 t2 - t2(a<->b); kEwItems elements per thread, loads before stores
template <class MT>
__global__ __launch_bounds__(kEwThreads) void k_u(const MT* __restrict__ mols, const State* __restrict__ st,
                                                 Ragged rg, int nm) {
  RAGGED_PROLOGUE((long)mols[mol].o * mols[mol].o * mols[mol].v * mols[mol].v);
  if (!st[mol].active) return;
  using V = std::remove_pointer_t<decltype(M.t2)>;
  const int v = M.v, vv = v * v;
  const gptr<V> t2 = global_view(M.t2);
  const gwptr<V> u = global_out(M.u);
  const long f0 = (long)lb * kEwItems * kEwThreads + threadIdx.x;
  V x[kEwItems], y[kEwItems];
#pragma unroll
  for (int it = 0; it < kEwItems; ++it) {
    const long f = f0 + (long)it * kEwThreads;
    const long g = f < size_ ? f : 0;
    int q, ab, a, b;
    divmod(g, vv, q, ab);
    divmod(ab, v, a, b);
    x[it] = t2[g];
    y[it] = t2[g - ab + (long)b * v + a];
  }
#pragma unroll
  for (int it = 0; it < kEwItems; ++it) {
    const long f = f0 + (long)it * kEwThreads;
    if (f < size_) u[f] = 2.0 * x[it] - (double)y[it];
  }
}

// Xvv = Ft_vv, Xoo = Ft_oo (the Fock parts of the E2 intermediates)
template <class MT>
__global__ __launch_bounds__(kEwThreads) void k_init_xx(const MT* __restrict__ mols, const State* __restrict__ st,
                                                       Ragged rg, int nm) {
  RAGGED_PROLOGUE((long)mols[mol].v * mols[mol].v + (long)mols[mol].o * mols[mol].o);
  if (e >= size_ || !st[mol].active) return;
  const int N = M.n, o = M.o, v = M.v;
  if (e < (long)v * v) {
    const int b = e / v, c = e % v;
    M.Xvv[e] = M.Ft[(long)(o + b) * N + o + c];
  } else {
    const long f = e - (long)v * v;
    const int k = f / o, j = f % o;
    M.Xoo[f] = M.Ft[(long)k * N + j];
  }
}

// w2 += x + x^T - 1/2 (Z[ai,bj] + Z[bj,ai]) - (Z[aj,bi] + Z[bi,aj]), with
// Zs[i][j][a][b] = Z[(a,i),(b,j)]
template <class MT>
__global__ __launch_bounds__(kEwThreads) void k_assemble(const MT* __restrict__ mols, const State* __restrict__ st,
                                                        Ragged rg, int nm) {
  RAGGED_PROLOGUE((long)mols[mol].o * mols[mol].o * mols[mol].v * mols[mol].v);
  if (!st[mol].active) return;
  using V = std::remove_pointer_t<decltype(M.w2)>;
  const int o = M.o, v = M.v;
  const gptr<V> x = global_view(M.x), Z = global_view(M.Zc), w2in = global_view(M.w2);
  const gwptr<V> w2 = global_out(M.w2);
  const long f0 = (long)lb * kEwItems * kEwThreads + threadIdx.x;
  // kEwItems elements per thread, loads before stores
  double val[kEwItems];
  V w[kEwItems];
#pragma unroll
  for (int it = 0; it < kEwItems; ++it) {
    const long f = f0 + (long)it * kEwThreads;
    const long g = f < size_ ? f : 0;
    int i, j, a, b;
    ijab_of(g, o, v, i, j, a, b);
    const long ij = ((long)i * o + j) * v * v, ji = ((long)j * o + i) * v * v;
    const long ab = (long)a * v + b, ba = (long)b * v + a;
    double s = (double)x[g] + x[ji + ba];
    s -= 0.5 * (Z[ij + ab] + Z[ji + ba]);
    s -= Z[ji + ab] + Z[ij + ba];
    val[it] = s;
    w[it] = w2in[g];
  }
#pragma unroll
  for (int it = 0; it < kEwItems; ++it) {
    const long f = f0 + (long)it * kEwThreads;
    if (f < size_) w2[f] = w[it] + val[it];
  }
}

// ---------------------------------------------------------------------------
// update (stepCC) fused with the DIIS push and max|d| partials
// ---------------------------------------------------------------------------
__device__ __forceinline__ long packed_index(long mm, long nn, long ne) {
  // upper triangle, row-major: pairs (mm, nn) with mm <= nn
  return mm * ne - mm * (mm - 1) / 2 + (nn - mm);
}


// (mm, nn), mm <= nn, of packed index p (the inverse of packed_index): the
// row is the smaller root of mm^2 - (2 ne + 1) mm + 2 p = 0 rounded down,
// corrected by one either way against the exact integer row starts
__device__ __forceinline__ void packed_pair(long p, long ne, int& mm, int& nn) {
  const double b = 2.0 * (double)ne + 1.0;
  long r = (long)((b - sqrt(b * b - 8.0 * (double)p)) * 0.5);
  r = r < 0 ? 0 : (r >= ne ? ne - 1 : r);
  if (r * ne - r * (r - 1) / 2 > p) --r;
  else if (r + 1 < ne && (r + 1) * ne - (r + 1) * r / 2 <= p) ++r;
  mm = (int)r;
  nn = (int)(r + (p - (r * ne - r * (r - 1) / 2)));
}

// maximum that propagates NaN (fmax drops it), so that a NaN residual can
// never read as a converged one
__device__ __forceinline__ double nmax(double a, double b) { return (a > b || a != a) ? a : b; }

__device__ __forceinline__ double block_max(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v = nmax(v, __shfl_xor(v, off));
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double r = sh[0];
  for (int k = 1; k < kEwThreads / 64; ++k) r = nmax(r, sh[k]);
  __syncthreads();
  return r;
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double r = sh[0];
  for (int k = 1; k < kEwThreads / 64; ++k) r += sh[k];
  __syncthreads();
  return r;
}

// block_sum of 4 values at once (same order per value; sh holds 4 x waves)
__device__ __forceinline__ void block_sum4(double (&v)[4], double* sh) {
  constexpr int W = kEwThreads / 64;
#pragma unroll
  for (int q = 0; q < 4; ++q)
    for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_xor(v[q], off);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) sh[q * W + w] = v[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double r = sh[q * W];
    for (int k = 1; k < W; ++k) r += sh[q * W + k];
    v[q] = r;
  }
  __syncthreads();
}

// DIIS history loops issue the loads of kDiisBatch slots before using any of
// them (a loop of load-then-accumulate waits for every load in turn).
constexpr int kDiisBatch = 8;
// sum_k coef[k] x[k nd + p] over the slots in slot order; the loads of a
// partial batch repeat the last slot (in bounds, unused)
template <class T>
__device__ __forceinline__ double diis_combine(const State& S, const T* x, long nd, long p) {
  const gptr<T> gx = global_view(x);
  const int cnt = S.diis_count;
  double s = 0.0;
  for (int k0 = 0; k0 < cnt; k0 += kDiisBatch) {
    T a[kDiisBatch];
#pragma unroll
    for (int q = 0; q < kDiisBatch; ++q) a[q] = gx[(long)min(k0 + q, cnt - 1) * nd + p];
#pragma unroll
    for (int q = 0; q < kDiisBatch; ++q)
      if (k0 + q < cnt) s += S.coef[k0 + q] * a[q];
  }
  return s;
}

// Singles: d1 = w1 + Ft_ai; t1 += s1 d1 / (F_ii - F_aa).  size o*v
__global__ __launch_bounds__(kEwThreads) void k_update1(const Mol* __restrict__ mols, const State* __restrict__ st,
                                                       Ragged rg, int nm, Params prm, double* __restrict__ part) {
  __shared__ double sh[kEwThreads / 64];
  RAGGED_PROLOGUE((long)mols[mol].o * mols[mol].v);
  if (!st[mol].active) return;
  double amax = 0.0;
  if (e < size_) {
    const int N = M.n, o = M.o, v = M.v;
    const int i = e / v, a = e % v;
    const double d1 = M.w1[e] + M.Ft[(long)(o + a) * N + i];
    M.w1[e] = d1;
    const double D = M.F[(long)i * N + i] - M.F[(long)(o + a) * N + o + a];
    const double tn = M.t1[e] + prm.s1 * d1 / D;
    M.t1[e] = tn;
    amax = fabs(d1);
    if (prm.diis) {
      const State& S = st[mol];
      M.sd[(long)S.diis_next * M.nd + e] = d1;
      M.st[(long)S.diis_next * M.nd + e] = tn;
    }
  }
  const double r = block_max(amax, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// Index set of element g of an o x o x v x v array for the pair updates: the
// pair's representative r ((i,a) <= (j,b)), the representative rx of the pair
// with a <-> b, the mirror element of g and the packed DIIS index of the pair
struct PairIx {
  long r, rx, mir, pk;
  int I, J, A, B;
  bool rep;   // g is its pair's representative
};
__device__ __forceinline__ PairIx pair_ix(long g, int o, int v) {
  PairIx q;
  int i, j, a, b;
  ijab_of(g, o, v, i, j, a, b);
  const long ne = (long)o * v, mm = (long)i * v + a, nn = (long)j * v + b;
  q.rep = mm <= nn;
  q.I = i; q.J = j; q.A = a; q.B = b;
  if (!q.rep) { q.I = j; q.J = i; q.A = b; q.B = a; }
  q.r = (((long)q.I * o + q.J) * v + q.A) * v + q.B;
  const long mb = (long)q.I * v + q.B, na = (long)q.J * v + q.A;
  q.rx = mb <= na ? (((long)q.I * o + q.J) * v + q.B) * v + q.A : (((long)q.J * o + q.I) * v + q.A) * v + q.B;
  q.mir = (((long)j * o + i) * v + b) * v + a;
  q.pk = ne + packed_index(q.rep ? mm : nn, q.rep ? nn : mm, ne);
  return q;
}

// Doubles, from the (i,a)<=(j,b) representative (I,J,A,B) of each pair, so
// that t2 stays exactly pair-symmetric:
//   d2 = 2 Omega - Omega(a<->b),  t2 += s2 (2 d2 + d2(a<->b)) / D2 = 3 s2 Omega / D2
// kEwItems elements per thread (table rgT2x), loads before stores.  The
// representative of each pair updates both of its elements and contributes the
// pair's max|d|; a non-representative element reads nothing.
__global__ __launch_bounds__(kEwThreads) void k_update2(const Mol* __restrict__ mols, const State* __restrict__ st,
                                                       Ragged rg, int nm, Params prm, double* __restrict__ part) {
  __shared__ double sh[kEwThreads / 64];
  RAGGED_PROLOGUE((long)mols[mol].o * mols[mol].o * mols[mol].v * mols[mol].v);
  if (!st[mol].active) return;
  const int N = M.n, o = M.o, v = M.v;
  const gptr<double> w2 = global_view(M.w2), F = global_view(M.F), t2in = global_view(M.t2);
  const gwptr<double> t2 = global_out(M.t2), sd = global_out(M.sd), stv = global_out(M.st);
  const long f0 = (long)lb * kEwItems * kEwThreads + threadIdx.x;
  PairIx q[kEwItems];
  double om[kEwItems], omx[kEwItems], dd[kEwItems], t[kEwItems];
#pragma unroll
  for (int it = 0; it < kEwItems; ++it) {
    const long f = f0 + (long)it * kEwThreads;
    const long g = f < size_ ? f : 0;
    q[it] = pair_ix(g, o, v);
    om[it] = omx[it] = dd[it] = t[it] = 0.0;
    if (q[it].rep) {
      om[it] = w2[q[it].r];
      omx[it] = w2[q[it].rx];
      dd[it] = F[(long)q[it].I * N + q[it].I] + F[(long)q[it].J * N + q[it].J] -
               F[(long)(o + q[it].A) * N + o + q[it].A] - F[(long)(o + q[it].B) * N + o + q[it].B];
      t[it] = t2in[g];
    }
  }
  double amax = 0.0;
  const State& S = st[mol];
#pragma unroll
  for (int it = 0; it < kEwItems; ++it) {
    const long f = f0 + (long)it * kEwThreads;
    // a non-representative element's residual is its representative's
    if (f >= size_ || !q[it].rep) continue;
    const double d2 = 2.0 * om[it] - omx[it];
    amax = nmax(amax, fabs(d2));
    const double num = prm.projtype == 2 ? 3.0 * om[it] : d2;
    if (q[it].rep) {
      const double tn = t[it] + prm.s2 * num / dd[it];
      t2[f] = tn;
      t2[q[it].mir] = tn;
      if (prm.diis) {
        sd[(long)S.diis_next * M.nd + q[it].pk] = num;
        stv[(long)S.diis_next * M.nd + q[it].pk] = tn;
      }
    }
  }
  const double rr = block_max(amax, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = rr;
}

// Reference-form residuals for the parity interface: d1 (already final in w1
// after k_update1 is NOT assumed; computed here from Omega) and
// d2 = 2 Omega_sym - Omega_sym(a<->b), both in row-major [i][a] / [i][j][a][b].
__global__ __launch_bounds__(kEwThreads) void k_contravariant(const Mol* __restrict__ mols, Ragged rg, int nm,
                                                             double* __restrict__ d1, double* __restrict__ d2,
                                                             long d1_off, long d2_off) {
  RAGGED_PROLOGUE((long)mols[mol].o * mols[mol].o * mols[mol].v * mols[mol].v + (long)mols[mol].o * mols[mol].v);
  if (e >= size_) return;
  const int N = M.n, o = M.o, v = M.v;
  const long n2 = (long)o * o * v * v;
  if (e >= n2) {
    const long f = e - n2;
    const int i = f / v, a = f % v;
    d1[d1_off + f] = M.w1[f] + M.Ft[(long)(o + a) * N + i];
    return;
  }
  int i, j, a, b;
  ijab_of(e, o, v, i, j, a, b);
  const long mm = (long)i * v + a, nn = (long)j * v + b;
  const long r = mm <= nn ? e : (((long)j * o + i) * v + b) * v + a;
  const long mb = (long)i * v + b, na = (long)j * v + a;
  const long rx = mb <= na ? (((long)i * o + j) * v + b) * v + a : (((long)j * o + i) * v + a) * v + b;
  d2[d2_off + e] = 2.0 * M.w2[r] - M.w2[rx];
}

// per-molecule finalisation of the update: accuracies, convergence, DIIS commit
// (one workgroup per molecule reduces the partial maxima)
__global__ __launch_bounds__(kEwThreads) void k_control1(const Mol* __restrict__ mols, State* __restrict__ st, int nm,
                                                        Params prm, const int* __restrict__ pre1,
                                                        const double* __restrict__ part1, const int* __restrict__ pre2,
                                                        const double* __restrict__ part2) {
  __shared__ double sh[kEwThreads / 64];
  const int mol = blockIdx.x;
  State& S = st[mol];
  if (!S.active) return;
  double a1 = 0.0, a2 = 0.0;
  for (int b = pre1[mol] + threadIdx.x; b < pre1[mol + 1]; b += kEwThreads) a1 = nmax(a1, part1[b]);
  for (int b = pre2[mol] + threadIdx.x; b < pre2[mol + 1]; b += kEwThreads) a2 = nmax(a2, part2[b]);
  a1 = block_max(a1, sh);
  a2 = block_max(a2, sh);
  if (threadIdx.x != 0) return;
  S.acc[0] = a1;
  S.acc[1] = a2;
  S.iter += 1;
  S.prev = S.cur;
  S.cur = nmax(a1, a2);
  if (S.cur < S.prev) S.successful += 1;
  S.staged = 0;
  if (S.cur < prm.eps) {
    S.status = 1;
    return;
  }
  if (prm.diis) {
    const int m = mols[mol].m;
    S.staged = 1;
    S.diis_count = S.diis_count < m ? S.diis_count + 1 : m;
  }
}

// after DIIS: divergence, iteration cap, stagnation and stop flag (gIterator.f90:165-206)
__global__ void k_control2(const Mol* __restrict__ mols, State* __restrict__ st, int* __restrict__ act, int nm,
                           Params prm, const int* __restrict__ stop) {
  const int mol = blockIdx.x * blockDim.x + threadIdx.x;
  if (mol >= nm) return;
  State& S = st[mol];
  if (!S.active) return;
  if (S.staged) S.diis_next = (S.diis_next + 1) % mols[mol].m;
  int status = 0;
  if (S.status == 1) status = 1;
  else if (stop != nullptr && stop[mol]) status = -4;
  else if (!(S.cur == S.cur) || (prm.check_divergence && S.cur > prm.divergence)) status = -2;   // NaN diverges always
  else if (S.iter + 1 > prm.maxiter) status = -1;
  else if (prm.check_stagnation && S.iter > prm.stagnation_after &&
           100.0 * S.successful / S.iter <= prm.stagnation_pct) status = -3;
  if (status != 0) {
    S.status = status;
    S.active = 0;
    act[mol] = 0;
  }
}

// ---------------------------------------------------------------------------
// DIIS
// ---------------------------------------------------------------------------
// partial dots <sd_k, sd_new> over this workgroup's chunk, for every valid slot k
template <class MT>
__global__ __launch_bounds__(kEwThreads) void k_diis_dot(const MT* __restrict__ mols, const State* __restrict__ st,
                                                        Ragged rg, int nm, double* __restrict__ part) {
  __shared__ double sh[4 * (kEwThreads / 64)];
  const int mol = rg.bmol[blockIdx.x];
  const int lb = blockIdx.x - rg.prefix[mol];
  const auto& M = mols[mol];
  const State& S = st[mol];
  const int m = M.m;
  if (!S.active || !S.staged) {
    if (threadIdx.x < m) part[(long)blockIdx.x * m + threadIdx.x] = 0.0;
    return;
  }
  const long base = (long)lb * kEwThreads * kDotItems;
  const auto gsd = global_view(M.sd);
  const auto snew = gsd + (long)S.diis_next * M.nd;
  double xn[kDotItems];
#pragma unroll
  for (int it = 0; it < kDotItems; ++it) {
    const long e = base + (long)it * kEwThreads + threadIdx.x;
    xn[it] = e < M.nd ? (double)snew[e] : 0.0;
  }
  // four slots per pass: their loads are in flight together, one reduction;
  // a partial pass loads the last slot again (in bounds, unused)
  const int cnt = S.diis_count;
  const bool full = base + (long)kEwThreads * kDotItems <= M.nd;   // no element past the vector
  for (int k0 = 0; k0 < cnt; k0 += 4) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (full) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const auto sk = gsd + (long)min(k0 + q, cnt - 1) * M.nd + base + threadIdx.x;
#pragma unroll
        for (int it = 0; it < kDotItems; ++it) acc[q] += (double)sk[it * kEwThreads] * xn[it];
      }
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const auto sk = gsd + (long)min(k0 + q, cnt - 1) * M.nd;
#pragma unroll
        for (int it = 0; it < kDotItems; ++it) {
          const long e = base + (long)it * kEwThreads + threadIdx.x;
          if (e < M.nd) acc[q] += (double)sk[e] * xn[it];
        }
      }
    }
    block_sum4(acc, sh);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (k0 + q < cnt) part[(long)blockIdx.x * m + k0 + q] = acc[q];
    }
  }
}

// finalise dots into B, build the bordered system, solve it (partial pivoting),
// scaled by the largest diagonal of B so that it stays accurate near convergence
template <class MT>
__global__ __launch_bounds__(64) void k_diis_solve(const MT* __restrict__ mols, State* __restrict__ st, int nm,
                                                   const int* __restrict__ prefix, const double* __restrict__ part) {
  const int mol = blockIdx.x;
  const auto& M = mols[mol];
  State& S = st[mol];
  if (!S.active || !S.staged) return;
  const int m = M.m, cnt = S.diis_count, snew = S.diis_next;
  // lane k sums slot k's partials in workgroup order (fixed order: reproducible)
  if ((int)threadIdx.x < cnt) {
    const int k = threadIdx.x;
    double s = 0.0;
    for (int b = prefix[mol]; b < prefix[mol + 1]; ++b) s += part[(long)b * m + k];
    M.Bm[k * m + snew] = s;
    M.Bm[snew * m + k] = s;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int k = 0; k < 64; ++k) S.coef[k] = 0.0;
  if (cnt < 2) { S.coef[snew] = 1.0; return; }
  double A[(kMaxDiis + 1) * (kMaxDiis + 2)];
  const int n = cnt + 1;
  double scale = 0.0;
  for (int k = 0; k < cnt; ++k) scale = fmax(scale, M.Bm[k * m + k]);
  if (!(scale > 0.0)) scale = 1.0;
  for (int r = 0; r < cnt; ++r) {
    for (int c = 0; c < cnt; ++c) A[r * (n + 1) + c] = M.Bm[r * m + c] / scale;
    A[r * (n + 1) + cnt] = 1.0;
    A[r * (n + 1) + n] = 0.0;
  }
  for (int c = 0; c < cnt; ++c) A[cnt * (n + 1) + c] = 1.0;
  A[cnt * (n + 1) + cnt] = 0.0;
  A[cnt * (n + 1) + n] = 1.0;
  bool ok = true;
  for (int c = 0; c < n && ok; ++c) {
    int piv = c;
    double best = fabs(A[c * (n + 1) + c]);
    for (int r = c + 1; r < n; ++r) {
      const double x = fabs(A[r * (n + 1) + c]);
      if (x > best) { best = x; piv = r; }
    }
    if (!(best > 1e-300)) { ok = false; break; }
    if (piv != c)
      for (int q = 0; q <= n; ++q) {
        const double tmp = A[c * (n + 1) + q];
        A[c * (n + 1) + q] = A[piv * (n + 1) + q];
        A[piv * (n + 1) + q] = tmp;
      }
    const double inv = 1.0 / A[c * (n + 1) + c];
    for (int r = c + 1; r < n; ++r) {
      const double f = A[r * (n + 1) + c] * inv;
      if (f == 0.0) continue;
      for (int q = c; q <= n; ++q) A[r * (n + 1) + q] -= f * A[c * (n + 1) + q];
    }
  }
  if (!ok) { S.coef[snew] = 1.0; return; }
  double x[kMaxDiis + 1];
  for (int r = n - 1; r >= 0; --r) {
    double s = A[r * (n + 1) + n];
    for (int q = r + 1; q < n; ++q) s -= A[r * (n + 1) + q] * x[q];
    x[r] = s / A[r * (n + 1) + r];
  }
  for (int k = 0; k < cnt; ++k) S.coef[k] = x[k];
}

// t = sum_k c_k st_k.  Grid: one thread per packed DIIS element (singles,
// then the pairs mm <= nn in packed order, read coalesced); each pair writes
// both mirror elements of t2.
template <class MT>
__global__ __launch_bounds__(kEwThreads) void k_diis_extrap(const MT* __restrict__ mols, const State* __restrict__ st,
                                                           Ragged rg, int nm) {
  RAGGED_PROLOGUE((long)mols[mol].nd);
  const State& S = st[mol];
  if (e >= size_ || !S.active || !S.staged || S.diis_count < 2) return;
  const int o = M.o, v = M.v;
  const long ne = (long)o * v;
  if (e < ne) {
    M.t1[e] = diis_combine(S, M.st, M.nd, e);
    return;
  }
  const long p = e;
  int mm, nn, i, a, j, b;
  packed_pair(e - ne, ne, mm, nn);
  const double s = diis_combine(S, M.st, M.nd, p);
  divmod(mm, v, i, a);
  divmod(nn, v, j, b);
  M.t2[(((long)i * o + j) * v + a) * v + b] = s;
  M.t2[(((long)j * o + i) * v + b) * v + a] = s;
}

// ---------------------------------------------------------------------------
// energy: E = eref + 2 sum F_ia t_ia + sum L_iajb (t2 + t1 t1)  (energyCC)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kEwThreads) void k_energy_part(const Mol* __restrict__ mols, Ragged rg, int nm,
                                                           double* __restrict__ part) {
  __shared__ double sh[kEwThreads / 64];
  RAGGED_PROLOGUE((long)mols[mol].o * mols[mol].o * mols[mol].v * mols[mol].v + (long)mols[mol].o * mols[mol].v);
  const int N = M.n, o = M.o, v = M.v;
  const long n2 = (long)o * o * v * v;
  double s1 = 0.0, s2 = 0.0;
  if (e < size_) {
    if (e >= n2) {
      const long f = e - n2;
      const int i = f / v, a = f % v;
      s1 = M.F[(long)i * N + o + a] * M.t1[f];
    } else {
      int i, j, a, b;
      ijab_of(e, o, v, i, j, a, b);
      s2 = M.Lt[e] * (M.t2[e] + M.t1[i * v + a] * M.t1[j * v + b]);
    }
  }
  const double r1 = block_sum(s1, sh);
  const double r2 = block_sum(s2, sh);
  if (threadIdx.x == 0) {
    part[2 * (long)blockIdx.x] = r1;
    part[2 * (long)blockIdx.x + 1] = r2;
  }
}

__global__ void k_energy_final(const Mol* __restrict__ mols, State* __restrict__ st, int nm,
                               const int* __restrict__ prefix, const double* __restrict__ part) {
  const int mol = blockIdx.x * blockDim.x + threadIdx.x;
  if (mol >= nm) return;
  double s1 = 0.0, s2 = 0.0;
  for (int b = prefix[mol]; b < prefix[mol + 1]; ++b) {
    s1 += part[2 * (long)b];
    s2 += part[2 * (long)b + 1];
  }
  State& S = st[mol];
  const double eref = mols[mol].eref;
  S.energy[0] = eref + 2.0 * s1 + s2;
  S.energy[1] = eref;
  S.energy[2] = 2.0 * s1;
  S.energy[3] = s2;
  S.energy[4] = 0.0;
}

// ---------------------------------------------------------------------------
// precomputation and guesses
// ---------------------------------------------------------------------------
// Lt[i][j][a][b] = 2(ia|jb) - (ib|ja);  Lldkc[(l,d),(k,c)] = 2(ld|kc) - (lc|kd)
// from Gklcd[k][l][c][d] = (kc|ld)
__global__ __launch_bounds__(kEwThreads) void k_make_L(const Mol* __restrict__ mols, Ragged rg, int nm) {
  RAGGED_PROLOGUE((long)mols[mol].o * mols[mol].o * mols[mol].v * mols[mol].v);
  if (e >= size_) return;
  const int o = M.o, v = M.v;
  const long ov = (long)o * v;
  const double* G = M.Gklcd;
  {
    int i, j, a, b;
    ijab_of(e, o, v, i, j, a, b);
    const long base = ((long)i * o + j) * v * v;
    M.Lt[e] = 2.0 * G[base + (long)a * v + b] - G[base + (long)b * v + a];
  }
  {
    const int c = e % v;
    const int k = (e / v) % o;
    const int d = (e / ov) % v;
    const int l = e / (ov * v);
    const long base = ((long)l * o + k) * v * v;
    M.Lldkc[e] = 2.0 * G[base + (long)d * v + c] - G[base + (long)c * v + d];
  }
}

// kind 0: the reference guess t2 = (2(ia|jb)-(ib|ja))/D (mod_cc.f90:634-650);
// kind 1: MP2 amplitudes (ia|jb)/D.  t1 = F_ia/(F_ii - F_aa) for both.
__global__ __launch_bounds__(kEwThreads) void k_guess(const Mol* __restrict__ mols, Ragged rg, int nm, int kind) {
  RAGGED_PROLOGUE((long)mols[mol].o * mols[mol].o * mols[mol].v * mols[mol].v + (long)mols[mol].o * mols[mol].v);
  if (e >= size_) return;
  const int N = M.n, o = M.o, v = M.v;
  const long n2 = (long)o * o * v * v;
  if (e >= n2) {
    const long f = e - n2;
    const int i = f / v, a = f % v;
    M.t1[f] = M.F[(long)i * N + o + a] / (M.F[(long)i * N + i] - M.F[(long)(o + a) * N + o + a]);
    return;
  }
  int i, j, a, b;
  ijab_of(e, o, v, i, j, a, b);
  const long mm = (long)i * v + a, nn = (long)j * v + b;
  if (mm > nn) return;   // the representative of the pair writes both elements
  const double D = M.F[(long)i * N + i] + M.F[(long)j * N + j] - M.F[(long)(o + a) * N + o + a] -
                   M.F[(long)(o + b) * N + o + b];
  const double num = kind == 0 ? M.Lt[e] : M.Gklcd[e];
  const long eT = (((long)j * o + i) * v + b) * v + a;
  M.t2[e] = num / D;
  M.t2[eT] = num / D;
  if (M.t2mp2 != nullptr) {
    M.t2mp2[e] = M.Gklcd[e] / D;
    M.t2mp2[eT] = M.Gklcd[e] / D;
  }
}

__global__ void k_reset_state(State* __restrict__ st, int* __restrict__ act, int nm) {
  const int mol = blockIdx.x * blockDim.x + threadIdx.x;
  if (mol >= nm) return;
  State& S = st[mol];
  S.iter = 0;
  S.status = 0;
  S.active = 1;
  S.successful = 1;
  S.diis_count = 0;
  S.diis_next = 0;
  S.staged = 0;
  S.acc[0] = S.acc[1] = -1.0;
  S.cur = 0.0;
  S.prev = 10.0;
  S.inner = 0;
  S.inner_its = 0;
  S.outer_its = 0;
  S.acc0 = 0.0;
  for (int k = 0; k < 5; ++k) S.energy[k] = 0.0;
  act[mol] = 1;
}

// stop request: every running molecule ends with status -4 (interrupted)
__global__ void k_stop_running(State* __restrict__ st, int* __restrict__ act, int nm) {
  const int mol = blockIdx.x * blockDim.x + threadIdx.x;
  if (mol >= nm) return;
  State& S = st[mol];
  if (S.status != 0) return;
  S.status = -4;
  S.active = 0;
  S.inner = 0;
  act[mol] = 0;
}

// re-arm molecules for the single-step interface (the caller runs the iterator)
__global__ void k_activate(State* __restrict__ st, int* __restrict__ act, int nm) {
  const int mol = blockIdx.x * blockDim.x + threadIdx.x;
  if (mol >= nm) return;
  st[mol].active = 1;
  st[mol].status = 0;
  act[mol] = 1;
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
#define RAGGED_LAUNCH(K, rg, ...)                                                     \
  do {                                                                              \
    if ((rg).nblocks > 0) hipLaunchKernelGGL(K, dim3((rg).nblocks), dim3(kEwThreads), 0, s, __VA_ARGS__); \
  } while (0)

void ew_dress(const Mol* mols, const State* st, Ragged rgNN, int nm, hipStream_t s) {
  RAGGED_LAUNCH(k_dress, rgNN, mols, st, rgNN, nm);
}
void ew_fock_m(const Mol* mols, const State* st, Ragged rgN, Ragged rgNN, int nm, hipStream_t s) {
  RAGGED_LAUNCH(k_fock_diag, rgN, mols, st, rgN, nm);
  RAGGED_LAUNCH(k_fock_m, rgNN, mols, st, rgNN, nm);
}
void ew_pouter(const Mol* mols, const State* st, Ragged rgN3, int nm, hipStream_t s) {
  RAGGED_LAUNCH(k_pouter, rgN3, mols, st, rgN3, nm);
}
template <class MT>
static void ew_u_t(const MT* mols, const State* st, Ragged rgT2, int nm, hipStream_t s) {
  RAGGED_LAUNCH(k_u<MT>, rgT2, mols, st, rgT2, nm);
}
template <class MT>
static void ew_init_xx_t(const MT* mols, const State* st, Ragged rgXX, int nm, hipStream_t s) {
  RAGGED_LAUNCH(k_init_xx<MT>, rgXX, mols, st, rgXX, nm);
}
template <class MT>
static void ew_assemble_t(const MT* mols, const State* st, Ragged rgT2, int nm, hipStream_t s) {
  RAGGED_LAUNCH(k_assemble<MT>, rgT2, mols, st, rgT2, nm);
}
void ew_u(const Mol* mols, const State* st, Ragged rgT2, int nm, hipStream_t s) { ew_u_t(mols, st, rgT2, nm, s); }
void ew_u(const MolD* mols, const State* st, Ragged rgT2, int nm, hipStream_t s) { ew_u_t(mols, st, rgT2, nm, s); }
void ew_init_xx(const Mol* mols, const State* st, Ragged rgXX, int nm, hipStream_t s) {
  ew_init_xx_t(mols, st, rgXX, nm, s);
}
void ew_init_xx(const MolD* mols, const State* st, Ragged rgXX, int nm, hipStream_t s) {
  ew_init_xx_t(mols, st, rgXX, nm, s);
}
void ew_assemble(const Mol* mols, const State* st, Ragged rgT2, int nm, hipStream_t s) {
  ew_assemble_t(mols, st, rgT2, nm, s);
}
void ew_assemble(const MolD* mols, const State* st, Ragged rgT2, int nm, hipStream_t s) {
  ew_assemble_t(mols, st, rgT2, nm, s);
}
void ew_update(const Mol* mols, const State* st, Ragged rgOV, Ragged rgT2, int nm, const Params& prm,
               double* part1, double* part2, hipStream_t s) {
  RAGGED_LAUNCH(k_update1, rgOV, mols, st, rgOV, nm, prm, part1);
  RAGGED_LAUNCH(k_update2, rgT2, mols, st, rgT2, nm, prm, part2);
}
void ew_contravariant(const Mol* mols, Ragged rgT2S, int nm, double* d1, double* d2, hipStream_t s) {
  RAGGED_LAUNCH(k_contravariant, rgT2S, mols, rgT2S, nm, d1, d2, 0L, 0L);
}
void ctl_after_update(const Mol* mols, State* st, int nm, const Params& prm, Ragged rgOV,
                      const double* part1, Ragged rgT2, const double* part2, hipStream_t s) {
  hipLaunchKernelGGL(k_control1, dim3(nm), dim3(kEwThreads), 0, s, mols, st, nm, prm, rgOV.prefix,
                     part1, rgT2.prefix, part2);
}
template <class MT>
static void diis_step_t(const MT* mols, State* st, int nm, Ragged rgND, Ragged rgP2, double* part, hipStream_t s) {
  RAGGED_LAUNCH(k_diis_dot<MT>, rgND, mols, st, rgND, nm, part);
  hipLaunchKernelGGL(k_diis_solve<MT>, dim3(nm), dim3(64), 0, s, mols, st, nm, rgND.prefix, part);
  RAGGED_LAUNCH(k_diis_extrap<MT>, rgP2, mols, st, rgP2, nm);
}
void diis_step(const Mol* mols, State* st, int nm, Ragged rgND, Ragged rgP2, double* part, hipStream_t s) {
  diis_step_t(mols, st, nm, rgND, rgP2, part, s);
}
void diis_step(const MolD* mols, State* st, int nm, Ragged rgND, Ragged rgP2, double* part, hipStream_t s) {
  diis_step_t(mols, st, nm, rgND, rgP2, part, s);
}
void ctl_after_diis(const Mol* mols, State* st, int* act, int nm, const Params& prm, const int* stop,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_control2, dim3((nm + 63) / 64), dim3(64), 0, s, mols, st, act, nm, prm, stop);
}
void energy(const Mol* mols, State* st, int nm, Ragged rgT2S, double* part, hipStream_t s) {
  RAGGED_LAUNCH(k_energy_part, rgT2S, mols, rgT2S, nm, part);
  hipLaunchKernelGGL(k_energy_final, dim3((nm + 63) / 64), dim3(64), 0, s, mols, st, nm, rgT2S.prefix,
                     part);
}
void make_L(const Mol* mols, Ragged rgT2, int nm, hipStream_t s) {
  RAGGED_LAUNCH(k_make_L, rgT2, mols, rgT2, nm);
}
void guess(const Mol* mols, State* st, int* act, Ragged rgT2S, int nm, int kind, hipStream_t s) {
  RAGGED_LAUNCH(k_guess, rgT2S, mols, rgT2S, nm, kind);
  hipLaunchKernelGGL(k_reset_state, dim3((nm + 63) / 64), dim3(64), 0, s, st, act, nm);
}
void reset_state(State* st, int* act, int nm, hipStream_t s) {
  hipLaunchKernelGGL(k_reset_state, dim3((nm + 63) / 64), dim3(64), 0, s, st, act, nm);
}
void activate(State* st, int* act, int nm, hipStream_t s) {
  hipLaunchKernelGGL(k_activate, dim3((nm + 63) / 64), dim3(64), 0, s, st, act, nm);
}
// post-solve passes over the whole batch: arm every molecule, keeping its
// active flag in save[] (statuses untouched); then restore the flags
__global__ void k_arm_all(State* __restrict__ st, int* __restrict__ act, int* __restrict__ save, int nm, int restore,
                          const double* const* __restrict__ sel) {
  const int mol = blockIdx.x * blockDim.x + threadIdx.x;
  if (mol >= nm) return;
  if (restore) {
    st[mol].active = save[mol];
    act[mol] = save[mol];
  } else {
    const int on = sel == nullptr || sel[mol] != nullptr;
    save[mol] = st[mol].active;
    st[mol].active = on;
    act[mol] = on;
  }
}
void arm_all(State* st, int* act, int* save, int nm, bool restore, hipStream_t s, const double* const* sel) {
  hipLaunchKernelGGL(k_arm_all, dim3((nm + 63) / 64), dim3(64), 0, s, st, act, save, nm, restore ? 1 : 0, sel);
}
// Linear response: t = t0 + h r (t0 saved in the backup arenas at the same
// offsets; r given per molecule as device pointers to its r1 [i][a] and r2
// [i][j][a][b] blocks; a null r1 leaves the molecule's amplitudes as they are)
__global__ __launch_bounds__(kEwThreads) void k_lr_perturb(const Mol* __restrict__ mols, Ragged rg, int nm,
                                                          const double* __restrict__ t1base, const double* __restrict__ t1b,
                                                          const double* __restrict__ t2base, const double* __restrict__ t2b,
                                                          const double* const* __restrict__ r1p,
                                                          const double* const* __restrict__ r2p, double h) {
  RAGGED_PROLOGUE((long)mols[mol].o * mols[mol].o * mols[mol].v * mols[mol].v);
  if (e >= size_) return;
  const double* r1 = r1p[mol];
  if (r1 == nullptr) return;
  const double* r2 = r2p[mol];
  const long n1 = (long)M.o * M.v;
  const long b2 = M.t2 - t2base;
  M.t2[e] = t2b[b2 + e] + h * r2[e];
  if (e < n1) {
    const long b1 = M.t1 - t1base;
    M.t1[e] = t1b[b1 + e] + h * r1[e];
  }
}
// Omega of the current residual pass (Omega1 = w1 + F~_ai, Omega2 = w2) into
// the per-molecule outputs (null: skipped); with `minus`, out = (out - Omega) * scale
__global__ __launch_bounds__(kEwThreads) void k_lr_omega(const Mol* __restrict__ mols, Ragged rg, int nm,
                                                        double* const* __restrict__ o1p, double* const* __restrict__ o2p,
                                                        int minus, double scale) {
  RAGGED_PROLOGUE((long)mols[mol].o * mols[mol].o * mols[mol].v * mols[mol].v);
  if (e >= size_) return;
  double* o1 = o1p[mol];
  if (o1 == nullptr) return;
  const long n1 = (long)M.o * M.v;
  const double w2 = M.w2[e];
  double* q2 = o2p[mol] + e;
  *q2 = minus ? (*q2 - w2) * scale : w2;
  if (e < n1) {
    const int N = M.n, o = M.o, v = M.v;
    const int i = (int)(e / v), a = (int)(e % v);
    const double w1 = M.w1[e] + M.Ft[(long)(o + a) * N + i];
    double* q1 = o1 + e;
    *q1 = minus ? (*q1 - w1) * scale : w1;
  }
}
void lr_perturb(const Mol* mols, Ragged rgT2, int nm, const double* t1base, const double* t1b, const double* t2base,
                const double* t2b, const double* const* r1p, const double* const* r2p, double h, hipStream_t s) {
  RAGGED_LAUNCH(k_lr_perturb, rgT2, mols, rgT2, nm, t1base, t1b, t2base, t2b, r1p, r2p, h);
}
void lr_omega(const Mol* mols, Ragged rgT2, int nm, double* const* o1p, double* const* o2p, bool minus, double scale,
              hipStream_t s) {
  RAGGED_LAUNCH(k_lr_omega, rgT2, mols, rgT2, nm, o1p, o2p, minus ? 1 : 0, scale);
}
void stop_running(State* st, int* act, int nm, hipStream_t s) {
  hipLaunchKernelGGL(k_stop_running, dim3((nm + 63) / 64), dim3(64), 0, s, st, act, nm);
}

// ---------------------------------------------------------------------------
// Mixed-precision solver: correction ("delta") forms of the elementwise steps.
// M0 holds the outer-point buffers, Md the corrections; every step is the
// exact change of its plain counterpart, linear (dressing, dressed density,
// Fock matrix) or bilinear (orbital products).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kEwThreads) void k_dress_d(const Mol* __restrict__ mols, const MolD* __restrict__ dmols,
                                                       const State* __restrict__ st, Ragged rg, int nm) {
  RAGGED_PROLOGUE((long)mols[mol].n * mols[mol].n);
  if (e >= size_ || !st[mol].active) return;
  const MolD& D = dmols[mol];
  const int N = M.n, o = M.o, v = M.v;
  const int mu = e / N, p = e % N;
  const double* C = M.C + (long)mu * N;
  double dx = 0.0, dy = 0.0;
  if (p < o) {
    for (int a = 0; a < v; ++a) dy += C[o + a] * D.t1[p * v + a];
  } else {
    const int a = p - o;
    for (int k = 0; k < o; ++k) dx -= C[k] * D.t1[k * v + a];
  }
  D.Xf[e] = dx;
  D.Yf[e] = dy;
}

__device__ __forceinline__ double dressed_density_d(const Mol& M, const MolD& D, int mu, int nu) {
  const float* dY = D.Yf + (long)mu * M.n;
  const double* Cn = M.C + (long)nu * M.n;
  double s = 0.0;
  for (int k = 0; k < M.o; ++k) s += dY[k] * Cn[k];
  return s;
}

__global__ __launch_bounds__(kEwThreads) void k_fock_diag_d(const Mol* __restrict__ mols, const MolD* __restrict__ dmols,
                                                           const State* __restrict__ st, Ragged rg, int nm) {
  RAGGED_PROLOGUE((long)mols[mol].n);
  if (e >= size_ || !st[mol].active) return;
  dmols[mol].dvec[e] = dressed_density_d(M, dmols[mol], (int)e, (int)e);
}

__global__ __launch_bounds__(kEwThreads) void k_fock_m_d(const Mol* __restrict__ mols, const MolD* __restrict__ dmols,
                                                        const State* __restrict__ st, Ragged rg, int nm) {
  RAGGED_PROLOGUE((long)mols[mol].n * mols[mol].n);
  if (e >= size_ || !st[mol].active) return;
  const MolD& D = dmols[mol];
  const int N = M.n;
  const int mu = e / N, nu = e % N;
  double val = -M.G[e] * dressed_density_d(M, D, mu, nu);
  if (mu == nu) {
    double s = 0.0;
    for (int l = 0; l < N; ++l) s += M.G[(long)mu * N + l] * D.dvec[l];
    val += 2.0 * s;
  }
  D.Mf[e] = val;
}

// dP = X0 (x) dY + dX (x) (Y0 + dY), and dPT likewise
__global__ __launch_bounds__(kEwThreads) void k_pouter_d(const Mol* __restrict__ mols, const MolD* __restrict__ dmols,
                                                        const State* __restrict__ st, Ragged rg, int nm) {
  RAGGED_PROLOGUE((long)mols[mol].n * mols[mol].n * mols[mol].n);
  if (!st[mol].active) return;
  const MolD& D = dmols[mol];
  const int N = M.n, N2 = N * N;
  const gptr<double> X0 = global_view(M.Xf), Y0 = global_view(M.Yf);
  const gptr<float> dX = global_view(D.Xf), dY = global_view(D.Yf);
  const gwptr<float> P = global_out(D.P), PT = global_out(D.PT);
  const long f0 = (long)lb * kPItems * kEwThreads + threadIdx.x;
  // all loads first, then the stores (as in k_pouter)
  double x0p[kPItems], y0q[kPItems], x0q[kPItems], y0p[kPItems];
  float dxp[kPItems], dyq[kPItems], dxq[kPItems], dyp[kPItems];
#pragma unroll
  for (int it = 0; it < kPItems; ++it) {
    const long f = f0 + (long)it * kEwThreads;
    int mu, pq, p, q;
    divmod(f < size_ ? f : 0, N2, mu, pq);
    divmod(pq, N, p, q);
    const int ip = mu * N + p, iq = mu * N + q;
    x0p[it] = X0[ip];
    y0q[it] = Y0[iq];
    x0q[it] = X0[iq];
    y0p[it] = Y0[ip];
    dxp[it] = dX[ip];
    dyq[it] = dY[iq];
    dxq[it] = dX[iq];
    dyp[it] = dY[ip];
  }
#pragma unroll
  for (int it = 0; it < kPItems; ++it) {
    const long f = f0 + (long)it * kEwThreads;
    if (f < size_) {
      P[f] = x0p[it] * dyq[it] + dxp[it] * (y0q[it] + (double)dyq[it]);
      PT[f] = x0q[it] * dyp[it] + dxq[it] * (y0p[it] + (double)dyp[it]);
    }
  }
}

// t (x) 0 for the corrections of active molecules
__global__ __launch_bounds__(kEwThreads) void k_zero_delta(const Mol* __restrict__ mols, const MolD* __restrict__ dmols,
                                                          const State* __restrict__ st, Ragged rg, int nm) {
  RAGGED_PROLOGUE((long)mols[mol].o * mols[mol].o * mols[mol].v * mols[mol].v + (long)mols[mol].o * mols[mol].v);
  if (!st[mol].active) return;
  const long n2 = (long)M.o * M.o * M.v * M.v;
  const gwptr<float> t2 = global_out(dmols[mol].t2), t1 = global_out(dmols[mol].t1);
  const long f0 = (long)lb * kEwItems * kEwThreads + threadIdx.x;
#pragma unroll
  for (int it = 0; it < kEwItems; ++it) {
    const long f = f0 + (long)it * kEwThreads;
    if (f < n2) t2[f] = 0.f;
    else if (f < size_) t1[f - n2] = 0.f;
  }
}

// t1 = 0 for every molecule (a kernel, not per-molecule memsets)
__global__ __launch_bounds__(kEwThreads) void k_zero_t1(const Mol* __restrict__ mols, Ragged rg, int nm) {
  RAGGED_PROLOGUE((long)mols[mol].o * mols[mol].v);
  if (e < size_) M.t1[e] = 0.0;
}

// t0 += dt
__global__ __launch_bounds__(kEwThreads) void k_commit_delta(const Mol* __restrict__ mols, const MolD* __restrict__ dmols,
                                                            const State* __restrict__ st, Ragged rg, int nm) {
  RAGGED_PROLOGUE((long)mols[mol].o * mols[mol].o * mols[mol].v * mols[mol].v + (long)mols[mol].o * mols[mol].v);
  if (st[mol].status != 0) return;
  const long n2 = (long)M.o * M.o * M.v * M.v;
  const MolD& D = dmols[mol];
  // t2 and t1 are one index range: element f < n2 of t2, else f - n2 of t1
  const gptr<double> t2in = global_view(M.t2), t1in = global_view(M.t1);
  const gptr<float> d2 = global_view(D.t2), d1 = global_view(D.t1);
  const gwptr<double> t2 = global_out(M.t2), t1 = global_out(M.t1);
  const long f0 = (long)lb * kEwItems * kEwThreads + threadIdx.x;
  double t[kEwItems];
  float d[kEwItems];
#pragma unroll
  for (int it = 0; it < kEwItems; ++it) {
    const long f = f0 + (long)it * kEwThreads;
    const long g = f < size_ ? f : 0;
    t[it] = g < n2 ? t2in[g] : t1in[g - n2];
    d[it] = g < n2 ? d2[g] : d1[g - n2];
  }
#pragma unroll
  for (int it = 0; it < kEwItems; ++it) {
    const long f = f0 + (long)it * kEwThreads;
    if (f < n2) t2[f] = t[it] + d[it];
    else if (f < size_) t1[f - n2] = t[it] + d[it];
  }
}

// Before t0 += dt: the stored inner DIIS amplitude vectors are corrections
// relative to t0, so each loses dt (their residuals stay valid).  Packed
// layout as in k_diis_extrap.
__global__ __launch_bounds__(kEwThreads) void k_diis_shift(const MolD* __restrict__ mols, const State* __restrict__ st,
                                                          Ragged rg, int nm) {
  RAGGED_PROLOGUE((long)mols[mol].nd);
  const State& S = st[mol];
  if (e >= size_ || S.status != 0 || S.diis_count == 0) return;
  const int o = M.o, v = M.v;
  const long ne = (long)o * v;
  const long p = e;
  double d;
  if (e < ne) {
    d = M.t1[e];
  } else {
    int mm, nn, i, a, j, b;
    packed_pair(e - ne, ne, mm, nn);
    divmod(mm, v, i, a);
    divmod(nn, v, j, b);
    d = M.t2[(((long)i * o + j) * v + a) * v + b];
  }
  const int cnt = S.diis_count;
  const gptr<float> gst = global_view(M.st);
  for (int k0 = 0; k0 < cnt; k0 += kDiisBatch) {
    float x[kDiisBatch];
#pragma unroll
    for (int q = 0; q < kDiisBatch; ++q) x[q] = gst[(long)min(k0 + q, cnt - 1) * M.nd + p];
#pragma unroll
    for (int q = 0; q < kDiisBatch; ++q)
      if (k0 + q < cnt) M.st[(long)(k0 + q) * M.nd + p] = x[q] - d;
  }
}

// ---------------------------------------------------------------------------
// The two thin contractions of the doubles residual,
//   x[i][j][a][b] += sum_c t2[i][j][a][c] Xvv[b][c] - sum_k t2[i][k][a][b] Xoo[k][j]
// (K = v and K = o, outputs v and o wide), as one pass for molecules with
// o, v <= kE2Max (the others take GEMM tables g23/g24).
typedef double e2d4 __attribute__((ext_vector_type(4)));
typedef float e2f4 __attribute__((ext_vector_type(4)));

// E2 terms (FP64 matrix cores), molecules with o, v <= kE2Max; rgOcc:
// kE2Split workgroups per (molecule, i), each over a range of a.  For each a,
// the o x v slice x[i][:][a][:] is computed in 16x16 tiles, one per wave:
//   x[i][j][a][b] += sum_c S[j][c] Xvv[b][c] - sum_k Xoo[k][j] S[k][b],
// S = t2[i][:][a][:] (LDS, zero-padded to 32 x 32, like Xvv and Xoo).
// V_MFMA_F64_16X16X4F64: lane l holds A[l % 16][l / 16] and B[l / 16][l % 16];
// D[row][col] is element row / 4 of lane col + 16 (row % 4).
constexpr int kE2PS = 34, kE2PO = 48;   // LDS pitches: [row][K] reads are bank-conflict free
__global__ __launch_bounds__(kEwThreads) void k_e2(const Mol* __restrict__ mols, const State* __restrict__ st,
                                                  Ragged rg, int nm) {
  __shared__ double sX[kE2Max * kE2PS], sO[kE2Max * kE2PO], sS[kE2Max * kE2PS];
  const int mol = rg.bmol[blockIdx.x];
  const int lb = blockIdx.x - rg.prefix[mol], i = lb / kE2Split, part = lb % kE2Split;
  const Mol& M = mols[mol];
  const int o = M.o, v = M.v;
  if (!st[mol].active || o > kE2Max || v > kE2Max || i >= o) return;
  const int na = (v + kE2Split - 1) / kE2Split, a0 = part * na, a1 = min(v, a0 + na);
  for (int q = threadIdx.x; q < kE2Max * kE2Max; q += kEwThreads) {
    const int r = q / kE2Max, c = q % kE2Max;
    sX[r * kE2PS + c] = r < v && c < v ? M.Xvv[r * v + c] : 0.0;
    sO[r * kE2PO + c] = r < o && c < o ? M.Xoo[r * o + c] : 0.0;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, l16 = lane & 15, kq = lane >> 4;
  const int BT = (v + 15) >> 4, jt = wave / BT, bt = wave % BT;
  const bool work = wave < ((o + 15) >> 4) * BT;
  const int jr = 16 * jt + l16, bc = 16 * bt + l16;
  // the next a's slice and x tile are loaded into registers while the current
  // one is computed, so no global read waits at the barriers
  constexpr int kSl = kE2Max * kE2Max / kEwThreads;
  const gptr<double> t2 = global_view(M.t2), xin = global_view(M.x);
  double ns[kSl], nx[4];
  auto fetch = [&](int a) {
#pragma unroll
    for (int r = 0; r < kSl; ++r) {
      const int q = threadIdx.x + r * kEwThreads, j = q / kE2Max, c = q % kE2Max;
      const bool in = j < o && c < v;
      ns[r] = in ? t2[(((long)i * o + j) * v + a) * v + c] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int j = 16 * jt + kq + 4 * r;
      nx[r] = work && bc < v && j < o ? xin[(((long)i * o + j) * v + a) * v + bc] : 0.0;
    }
  };
  if (a0 < a1) fetch(a0);
  for (int a = a0; a < a1; ++a) {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSl; ++r) {
      const int q = threadIdx.x + r * kEwThreads;
      sS[(q / kE2Max) * kE2PS + q % kE2Max] = ns[r];
    }
    double xc[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) xc[r] = nx[r];
    __syncthreads();
    if (a + 1 < a1) fetch(a + 1);
    if (!work) continue;
    e2d4 acc = e2d4{0.0, 0.0, 0.0, 0.0};
    for (int c0 = 0; c0 < v; c0 += 4)
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(sS[jr * kE2PS + c0 + kq], sX[bc * kE2PS + c0 + kq], acc, 0, 0, 0);
    for (int k0 = 0; k0 < o; k0 += 4)
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(-sO[(k0 + kq) * kE2PO + jr], sS[(k0 + kq) * kE2PS + bc], acc, 0, 0, 0);
    if (bc < v)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = 16 * jt + kq + 4 * r;
        if (j < o) M.x[(((long)i * o + j) * v + a) * v + bc] = xc[r] + acc[r];
      }
  }
}

// Correction form of k_e2 (FP32 matrix cores): with X1 = X0 + dX,
//   dx += sum_c (t2_0 dXvv + dt2 Xvv1) - sum_k (t2_0 dXoo + dt2 Xoo1)
// -- the exact change of the FP64 terms up to FP32 rounding.  Tiles as in
// k_e2; V_MFMA_F32_16X16X4F32 holds A and B as the FP64 instruction does and
// leaves D[row][col] in element row % 4 of lane col + 16 (row / 4).
__global__ __launch_bounds__(kEwThreads) void k_e2_d(const Mol* __restrict__ mols, const MolD* __restrict__ dmols,
                                                    const State* __restrict__ st, Ragged rg, int nm) {
  __shared__ float sXd[kE2Max * kE2PS], sX1[kE2Max * kE2PS], sOd[kE2Max * kE2PO], sO1[kE2Max * kE2PO];
  __shared__ float sS0[kE2Max * kE2PS], sSd[kE2Max * kE2PS];
  const int mol = rg.bmol[blockIdx.x];
  const int lb = blockIdx.x - rg.prefix[mol], i = lb / kE2Split, part = lb % kE2Split;
  const Mol& M = mols[mol];
  const MolD& D = dmols[mol];
  const int o = M.o, v = M.v;
  if (!st[mol].active || o > kE2Max || v > kE2Max || i >= o) return;
  const int na = (v + kE2Split - 1) / kE2Split, a0 = part * na, a1 = min(v, a0 + na);
  for (int q = threadIdx.x; q < kE2Max * kE2Max; q += kEwThreads) {
    const int r = q / kE2Max, c = q % kE2Max;
    const bool xv = r < v && c < v, xo = r < o && c < o;
    sXd[r * kE2PS + c] = xv ? D.Xvv[r * v + c] : 0.f;
    sX1[r * kE2PS + c] = xv ? (float)(M.Xvv[r * v + c] + (double)D.Xvv[r * v + c]) : 0.f;
    sOd[r * kE2PO + c] = xo ? D.Xoo[r * o + c] : 0.f;
    sO1[r * kE2PO + c] = xo ? (float)(M.Xoo[r * o + c] + (double)D.Xoo[r * o + c]) : 0.f;
  }
  // global loads (generic ones would share the LDS wait counter); the next
  // a's slice is in flight while the current one is computed
  const gptr<double> t2 = global_view(M.t2);
  const gptr<float> dt2 = global_view(D.t2), xin = global_view(D.x);
  const gwptr<float> x = global_out(D.x);
  constexpr int kSl = kE2Max * kE2Max / kEwThreads;   // staged slice elements per thread
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, l16 = lane & 15, kq = lane >> 4;
  const int BT = (v + 15) >> 4, jt = wave / BT, bt = wave % BT;
  const bool work = wave < ((o + 15) >> 4) * BT;
  const int jr = 16 * jt + l16, bc = 16 * bt + l16;
  float n0[kSl], nd[kSl], nx[4];
  auto fetch = [&](int a) {
#pragma unroll
    for (int r = 0; r < kSl; ++r) {
      const int q = threadIdx.x + r * kEwThreads, j = q / kE2Max, c = q % kE2Max;
      const bool in = j < o && c < v;
      const long e = (((long)i * o + (in ? j : 0)) * v + a) * v + (in ? c : 0);
      n0[r] = in ? (float)t2[e] : 0.f;
      nd[r] = in ? dt2[e] : 0.f;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int j = 16 * jt + 4 * kq + r;
      nx[r] = work && bc < v && j < o ? xin[(((long)i * o + j) * v + a) * v + bc] : 0.f;
    }
  };
  if (a0 < a1) fetch(a0);
  for (int a = a0; a < a1; ++a) {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSl; ++r) {
      const int q = threadIdx.x + r * kEwThreads;
      sS0[(q / kE2Max) * kE2PS + q % kE2Max] = n0[r];
      sSd[(q / kE2Max) * kE2PS + q % kE2Max] = nd[r];
    }
    float xc[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) xc[r] = nx[r];
    __syncthreads();
    if (a + 1 < a1) fetch(a + 1);
    if (!work) continue;
    e2f4 acc = e2f4{0.f, 0.f, 0.f, 0.f};
    for (int c0 = 0; c0 < v; c0 += 4) {
      const int c = c0 + kq;
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(sS0[jr * kE2PS + c], sXd[bc * kE2PS + c], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(sSd[jr * kE2PS + c], sX1[bc * kE2PS + c], acc, 0, 0, 0);
    }
    for (int k0 = 0; k0 < o; k0 += 4) {
      const int k = k0 + kq;
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(-sOd[k * kE2PO + jr], sS0[k * kE2PS + bc], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(-sO1[k * kE2PO + jr], sSd[k * kE2PS + bc], acc, 0, 0, 0);
    }
    if (bc < v)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = 16 * jt + 4 * kq + r;
        if (j < o) x[(((long)i * o + j) * v + a) * v + bc] = xc[r] + acc[r];
      }
  }
}

void ew_e2(const Mol* mols, const State* st, Ragged rgOcc, int nm, hipStream_t s) {
  RAGGED_LAUNCH(k_e2, rgOcc, mols, st, rgOcc, nm);
}
void mixed_e2(const Mol* m0, const MolD* md, const State* st, Ragged rgOcc, int nm, hipStream_t s) {
  RAGGED_LAUNCH(k_e2_d, rgOcc, m0, md, st, rgOcc, nm);
}

// mixed update, singles: Omega1 = (w1_0 incl. F~_ai) + dw1 + dF~_ai; T = t1_0 + dt1
__global__ __launch_bounds__(kEwThreads) void k_update1_mixed(const Mol* __restrict__ mols, const MolD* __restrict__ dmols,
                                                             const State* __restrict__ st, Ragged rg, int nm, Params prm,
                                                             double* __restrict__ part) {
  __shared__ double sh[kEwThreads / 64];
  RAGGED_PROLOGUE((long)mols[mol].o * mols[mol].v);
  if (!st[mol].active) return;
  double amax = 0.0;
  if (e < size_) {
    const MolD& D = dmols[mol];
    const int N = M.n, o = M.o, v = M.v;
    const int i = e / v, a = e % v;
    // at the first step of an inner solve (dt = 0) the correction residual
    // arrays hold the previous solve's values and are not read
    const double d1 = M.w1[e] + (prm.outer_only ? 0.0 : (double)D.w1[e] + (double)D.Ft[(long)(o + a) * N + i]);
    const double Dn = M.F[(long)i * N + i] - M.F[(long)(o + a) * N + o + a];
    const double tn = M.t1[e] + D.t1[e] + prm.s1 * d1 / Dn;
    const float dt = (float)(tn - M.t1[e]);
    D.t1[e] = dt;
    amax = fabs(d1);
    if (prm.diis) {
      const State& S = st[mol];
      D.sd[(long)S.diis_next * M.nd + e] = (float)d1;
      D.st[(long)S.diis_next * M.nd + e] = dt;
    }
  }
  const double r = block_max(amax, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// mixed update, doubles (pair representatives as in k_update2); DIIS stores corrections
__global__ __launch_bounds__(kEwThreads) void k_update2_mixed(const Mol* __restrict__ mols, const MolD* __restrict__ dmols,
                                                             const State* __restrict__ st, Ragged rg, int nm, Params prm,
                                                             double* __restrict__ part) {
  __shared__ double sh[kEwThreads / 64];
  RAGGED_PROLOGUE((long)mols[mol].o * mols[mol].o * mols[mol].v * mols[mol].v);
  if (!st[mol].active) return;
  const MolD& Dm = dmols[mol];
  const int N = M.n, o = M.o, v = M.v;
  const gptr<double> w2 = global_view(M.w2), F = global_view(M.F), t2 = global_view(M.t2);
  const gptr<float> dw2 = global_view(Dm.w2), dt2in = global_view(Dm.t2);
  const gwptr<float> dt2 = global_out(Dm.t2), sd = global_out(Dm.sd), stv = global_out(Dm.st);
  const long f0 = (long)lb * kEwItems * kEwThreads + threadIdx.x;
  // kEwItems elements per thread, loads before stores (as in k_update2)
  PairIx q[kEwItems];
  double om[kEwItems], omx[kEwItems], dd[kEwItems], t0[kEwItems];
  float dw[kEwItems], dwx[kEwItems], dt[kEwItems];
#pragma unroll
  for (int it = 0; it < kEwItems; ++it) {
    const long f = f0 + (long)it * kEwThreads;
    const long g = f < size_ ? f : 0;
    q[it] = pair_ix(g, o, v);
    om[it] = omx[it] = dd[it] = t0[it] = 0.0;
    dw[it] = dwx[it] = dt[it] = 0.f;
    if (q[it].rep) {
      om[it] = w2[q[it].r];
      omx[it] = w2[q[it].rx];
      // at the first step of an inner solve (dt = 0) the correction residual
      // arrays hold the previous solve's values and are not used
      dw[it] = dw2[q[it].r];
      dwx[it] = dw2[q[it].rx];
      dd[it] = F[(long)q[it].I * N + q[it].I] + F[(long)q[it].J * N + q[it].J] -
               F[(long)(o + q[it].A) * N + o + q[it].A] - F[(long)(o + q[it].B) * N + o + q[it].B];
      t0[it] = t2[g];
      dt[it] = dt2in[g];
    }
  }
  double amax = 0.0;
  const State& S = st[mol];
#pragma unroll
  for (int it = 0; it < kEwItems; ++it) {
    const long f = f0 + (long)it * kEwThreads;
    // a non-representative element's residual is its representative's
    if (f >= size_ || !q[it].rep) continue;
    const double om_ = om[it] + (prm.outer_only ? 0.0 : (double)dw[it]);
    const double d2 = 2.0 * om_ - (omx[it] + (prm.outer_only ? 0.0 : (double)dwx[it]));
    amax = nmax(amax, fabs(d2));
    const double num = prm.projtype == 2 ? 3.0 * om_ : d2;
    if (q[it].rep) {
      const double tn = t0[it] + dt[it] + prm.s2 * num / dd[it];
      const float dtn = (float)(tn - t0[it]);
      dt2[f] = dtn;
      dt2[q[it].mir] = dtn;
      if (prm.diis) {
        sd[(long)S.diis_next * M.nd + q[it].pk] = (float)num;
        stv[(long)S.diis_next * M.nd + q[it].pk] = dtn;
      }
    }
  }
  const double rr = block_max(amax, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = rr;
}

// after the FP64 residual at the outer point: acc0, convergence, inner set-up
__global__ __launch_bounds__(kEwThreads) void k_outer_control(const Mol* __restrict__ mols, State* __restrict__ st,
                                                             int* __restrict__ act, int nm, Params prm, int max_outer,
                                                             const int* __restrict__ pre1, const double* __restrict__ part1,
                                                             const int* __restrict__ pre2, const double* __restrict__ part2) {
  __shared__ double sh[kEwThreads / 64];
  const int mol = blockIdx.x;
  State& S = st[mol];
  if (S.status != 0) return;
  double a1 = 0.0, a2 = 0.0;
  for (int b = pre1[mol] + threadIdx.x; b < pre1[mol + 1]; b += kEwThreads) a1 = nmax(a1, part1[b]);
  for (int b = pre2[mol] + threadIdx.x; b < pre2[mol + 1]; b += kEwThreads) a2 = nmax(a2, part2[b]);
  a1 = block_max(a1, sh);
  a2 = block_max(a2, sh);
  if (threadIdx.x != 0) return;
  S.acc[0] = a1;
  S.acc[1] = a2;
  S.acc0 = nmax(a1, a2);
  S.cur = S.acc0;
  S.iter += 1;
  S.outer_its += 1;
  S.inner = 0;
  S.inner_its = 0;
  if (!prm.keep_diis) {
    S.diis_count = 0;
    S.diis_next = 0;
  }
  S.staged = 0;
  if (S.acc0 < prm.eps) { S.status = 1; act[mol] = 0; return; }
  if (prm.check_divergence && !(S.acc0 <= prm.divergence)) { S.status = -2; act[mol] = 0; return; }
  if (!(S.acc0 == S.acc0)) { S.status = -2; act[mol] = 0; return; }   // NaN diverges whatever the checks
  if (S.outer_its >= max_outer || S.iter + 1 > prm.maxiter) { S.status = -1; act[mol] = 0; return; }
  S.inner = 1;
}

// inner-iteration control: stop at acc < kappa*acc0 (or below eps/2, or at the cap)
__global__ __launch_bounds__(kEwThreads) void k_inner_control(const MolD* __restrict__ mols, State* __restrict__ st,
                                                             int* __restrict__ act, int nm, Params prm,
                                                             const int* __restrict__ pre1, const double* __restrict__ part1,
                                                             const int* __restrict__ pre2, const double* __restrict__ part2) {
  __shared__ double sh[kEwThreads / 64];
  const int mol = blockIdx.x;
  State& S = st[mol];
  if (!S.active) return;
  double a1 = 0.0, a2 = 0.0;
  for (int b = pre1[mol] + threadIdx.x; b < pre1[mol + 1]; b += kEwThreads) a1 = nmax(a1, part1[b]);
  for (int b = pre2[mol] + threadIdx.x; b < pre2[mol + 1]; b += kEwThreads) a2 = nmax(a2, part2[b]);
  a1 = block_max(a1, sh);
  a2 = block_max(a2, sh);
  if (threadIdx.x != 0) return;
  S.acc[0] = a1;
  S.acc[1] = a2;
  S.cur = nmax(a1, a2);
  S.iter += 1;
  S.inner_its += 1;
  S.staged = 0;
  const bool done = S.cur < prm.kappa * S.acc0 || S.cur < 0.5 * prm.eps || S.inner_its >= prm.max_inner ||
                    !(S.cur == S.cur);
  if (done) {
    S.inner = 0;
    S.active = 0;
    act[mol] = 0;
    return;
  }
  if (prm.diis) {
    const int m = mols[mol].m;
    S.staged = 1;
    S.diis_count = S.diis_count < m ? S.diis_count + 1 : m;
  }
}

// DIIS slot advance for the inner phase (k_control2 without the iterator tests)
__global__ void k_inner_advance(const MolD* __restrict__ mols, State* __restrict__ st, int nm) {
  const int mol = blockIdx.x * blockDim.x + threadIdx.x;
  if (mol >= nm) return;
  State& S = st[mol];
  if (S.active && S.staged) S.diis_next = (S.diis_next + 1) % mols[mol].m;
}

// phase masks: 0 = outer (every running molecule), 1 = inner (molecules still solving)
__global__ void k_set_phase(State* __restrict__ st, int* __restrict__ act, int nm, int phase) {
  const int mol = blockIdx.x * blockDim.x + threadIdx.x;
  if (mol >= nm) return;
  State& S = st[mol];
  const int a = S.status == 0 && (phase == 0 || S.inner);
  S.active = a;
  act[mol] = a;
}

void mixed_dress(const Mol* m0, const MolD* md, const State* st, Ragged rgNN, Ragged rgN, Ragged rgN3, int nm,
                 hipStream_t s) {
  RAGGED_LAUNCH(k_dress_d, rgNN, m0, md, st, rgNN, nm);
  RAGGED_LAUNCH(k_fock_diag_d, rgN, m0, md, st, rgN, nm);
  RAGGED_LAUNCH(k_fock_m_d, rgNN, m0, md, st, rgNN, nm);
  RAGGED_LAUNCH(k_pouter_d, rgN3, m0, md, st, rgN3, nm);
}
void mixed_zero(const Mol* m0, const MolD* md, const State* st, Ragged rgT2S, int nm, hipStream_t s) {
  RAGGED_LAUNCH(k_zero_delta, rgT2S, m0, md, st, rgT2S, nm);
}
// BF16 copy of an FP64 array (round to nearest even)
__global__ __launch_bounds__(kEwThreads) void k_to_bf16(const double* __restrict__ src, __bf16* __restrict__ dst,
                                                       long n) {
  for (long e = (long)blockIdx.x * kEwThreads + threadIdx.x; e < n; e += (long)gridDim.x * kEwThreads)
    dst[e] = (__bf16)(float)src[e];
}

// BF16 copies of the running molecules' blocks of an FP64 arena: workgroup b
// belongs to molecule k with pre[k] <= b < pre[k + 1] (molecules that are not
// running have empty ranges) and converts 4 * kEwThreads elements of the
// block [off[k], off[k] + len[k]) of the arena
__global__ __launch_bounds__(kEwThreads) void k_to_bf16_runs(const double* __restrict__ src, __bf16* __restrict__ dst,
                                                            const long* __restrict__ off, const long* __restrict__ len,
                                                            const int* __restrict__ pre, int nm) {
  const int k = block_mol(pre, nm, blockIdx.x);
  const long e0 = off[k] + (long)(blockIdx.x - pre[k]) * (4 * kEwThreads) + threadIdx.x;
  const long e1 = off[k] + len[k];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const long e = e0 + q * kEwThreads;
    if (e < e1) dst[e] = (__bf16)(float)src[e];
  }
}

void to_bf16_runs(const double* src, __bf16* dst, const long* off, const long* len, const int* pre, int nm, int blocks,
                  hipStream_t s) {
  if (blocks > 0)
    hipLaunchKernelGGL(k_to_bf16_runs, dim3(blocks), dim3(kEwThreads), 0, s, src, dst, off, len, pre, nm);
}

void to_bf16(const double* src, __bf16* dst, long n, hipStream_t s) {
  if (n <= 0) return;
  const long blocks = std::min<long>((n + kEwThreads - 1) / kEwThreads, 65536);
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)blocks), dim3(kEwThreads), 0, s, src, dst, n);
}

void zero_t1(const Mol* mols, Ragged rgOV, int nm, hipStream_t s) {
  RAGGED_LAUNCH(k_zero_t1, rgOV, mols, rgOV, nm);
}
void mixed_commit(const Mol* m0, const MolD* md, const State* st, Ragged rgT2S, Ragged rgP2, int nm, bool keep_diis,
                  hipStream_t s) {
  if (keep_diis) RAGGED_LAUNCH(k_diis_shift, rgP2, md, st, rgP2, nm);
  RAGGED_LAUNCH(k_commit_delta, rgT2S, m0, md, st, rgT2S, nm);
}
void mixed_update(const Mol* m0, const MolD* md, const State* st, Ragged rgOV, Ragged rgT2, int nm, const Params& prm,
                  double* part1, double* part2, hipStream_t s) {
  RAGGED_LAUNCH(k_update1_mixed, rgOV, m0, md, st, rgOV, nm, prm, part1);
  RAGGED_LAUNCH(k_update2_mixed, rgT2, m0, md, st, rgT2, nm, prm, part2);
}
void mixed_outer_control(const Mol* mols, State* st, int* act, int nm, const Params& prm, int max_outer, Ragged rgOV,
                         const double* part1, Ragged rgT2, const double* part2, hipStream_t s) {
  hipLaunchKernelGGL(k_outer_control, dim3(nm), dim3(kEwThreads), 0, s, mols, st, act, nm, prm, max_outer,
                     rgOV.prefix, part1, rgT2.prefix, part2);
}
void mixed_inner_control(const MolD* mols, State* st, int* act, int nm, const Params& prm, Ragged rgOV,
                         const double* part1, Ragged rgT2, const double* part2, hipStream_t s) {
  hipLaunchKernelGGL(k_inner_control, dim3(nm), dim3(kEwThreads), 0, s, mols, st, act, nm, prm, rgOV.prefix,
                     part1, rgT2.prefix, part2);
}
void mixed_inner_advance(const MolD* mols, State* st, int nm, hipStream_t s) {
  hipLaunchKernelGGL(k_inner_advance, dim3((nm + 63) / 64), dim3(64), 0, s, mols, st, nm);
}
void set_phase(State* st, int* act, int nm, int phase, hipStream_t s) {
  hipLaunchKernelGGL(k_set_phase, dim3((nm + 63) / 64), dim3(64), 0, s, st, act, nm, phase);
}

}  // namespace hcc
