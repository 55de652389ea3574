// (T) correction of the converged closed-shell amplitudes, FP64.
//
// For every occupied triple i >= j >= k of a molecule (the work item):
//   w_ijk^abc = sum_d (bd|ai) t_kj^cd - sum_l (ck|jl) t_il^ab,
//   W_ijk^abc = sum of w over the six simultaneous permutations of (ia),(jb),(kc),
//   V_ijk^abc = t_i^a (bj|ck) + t_j^b (ai|ck) + t_k^c (ai|bj),
//   R3(W)_abc = 4 W_abc + W_bca + W_cab - 2 (W_acb + W_bac + W_cba),
//   E[T]_ijk  = sum_abc W R3(W) / (3 D),  E[ST]_ijk = sum_abc V R3(W) / (3 D),
// D = e_i + e_j + e_k - e_a - e_b - e_c.  R3's coefficients are constant on
// the conjugacy classes of S3, so the per-triple sums are invariant under
// permutations of (i,j,k): ordered triples are covered by i >= j >= k with
// weights 6, 3 or 1.  The reference's spin-orbital loop
// (src/methods/coupledCluster/mod_cc.f90:2028-2111) evaluates -E[T] (c1) and
// -E[ST] (c2) of the same amplitudes.
//
// The kernel is persistent: a fixed grid of workgroups walks the work items
// of every molecule; each workgroup builds W for its item in its own global
// block (six permutation passes, each a v^2 x v x (v + o) product with the
// amplitude and integral slices staged in LDS and the z-run of each (x,y) row
// accumulated in registers), then reduces the energy.  Per-item partials are
// summed per molecule in a fixed order, so results are bitwise reproducible.
#include <hip/hip_runtime.h>

#include "hcc_triples.h"

namespace hcc {

constexpr int kTripThreads = 256;
constexpr int kTripMaxV = 32;

// permutation s of three positions: position k of the permuted triple takes entry kPerm[s][k]
__constant__ int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

__device__ double trip_block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = red[0];
  for (int k = 1; k < kTripThreads / 64; ++k) r += red[k];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kTripThreads) void k_triples(const TripMol* __restrict__ mols, int nm,
                                                          long total, const int2* __restrict__ items,
                                                          double* __restrict__ wscratch,
                                                          long wstride, double* __restrict__ part) {
  extern __shared__ double lds[];
  double* red = lds;                      // kTripThreads / 64
  double* Ts = red + kTripThreads / 64;   // [d][z], v x v
  double* Qs = Ts + kTripMaxV * kTripMaxV;   // [l][z], o x v
  double* W = wscratch + (long)blockIdx.x * wstride;
  for (long item = blockIdx.x; item < total; item += gridDim.x) {
    const int2 info = items[item];   // molecule, packed triple
    const TripMol& M = mols[info.x];
    const int o = M.o, v = M.v, n = M.n;
    if (v <= kTripLdsV) continue;   // the matrix-core kernel takes it
    const int packed = info.y;
    const int occ[3] = {packed & 0xff, (packed >> 8) & 0xff, (packed >> 16) & 0xff};   // i >= j >= k
    for (int s = 0; s < 6; ++s) {
      // (p, q, r) = permuted (i, j, k); the output index (x, y, z) is the same
      // permutation of (a, b, c)
      const int p = occ[kPerm[s][0]], q = occ[kPerm[s][1]], r = occ[kPerm[s][2]];
      // strides of (x, y, z) in W[a][b][c]: entry kPerm[s][m] of (a, b, c) has stride v^(2 - kPerm[s][m])
      const long stride3[3] = {(long)v * v, (long)v, 1};
      const long sx = stride3[kPerm[s][0]], sy = stride3[kPerm[s][1]], sz = stride3[kPerm[s][2]];
      __syncthreads();
      for (int e = threadIdx.x; e < v * v; e += kTripThreads) {
        const int z = e / v, d = e % v;
        Ts[d * v + z] = M.t2[(((long)r * o + q) * v + z) * v + d];            // t_rq^zd
      }
      for (int e = threadIdx.x; e < o * v; e += kTripThreads) {
        const int l = e / v, z = e % v;
        Qs[l * v + z] = M.B[(((long)q * o + l) * o + r) * v + z];            // (ql|rz) = (zr|ql)
      }
      __syncthreads();
      for (int xy = threadIdx.x; xy < v * v; xy += kTripThreads) {
        const int x = xy / v, y = xy % v;
        double acc[kTripMaxV];
#pragma unroll
        for (int z = 0; z < kTripMaxV; ++z) acc[z] = 0.0;
        const double* Arow = M.A + (((long)p * v + x) * v + y) * v;          // (yd|xp), d = 0 .. v-1
        for (int d = 0; d < v; ++d) {
          const double a = Arow[d];
#pragma unroll
          for (int z = 0; z < kTripMaxV; ++z)
            if (z < v) acc[z] += a * Ts[d * v + z];
        }
        for (int l = 0; l < o; ++l) {
          const double u = M.t2[(((long)p * o + l) * v + x) * v + y];       // t_pl^xy
#pragma unroll
          for (int z = 0; z < kTripMaxV; ++z)
            if (z < v) acc[z] -= u * Qs[l * v + z];
        }
        // (x, y, z) holds entries kPerm[s][0..2] of (a, b, c)
        const long base = x * sx + y * sy;
#pragma unroll
        for (int z = 0; z < kTripMaxV; ++z) {
          if (z >= v) break;
          const long idx = base + z * sz;
          W[idx] = (s == 0 ? 0.0 : W[idx]) + acc[z];
        }
      }
    }
    __syncthreads();
    // energy of the item
    const int i = occ[0], j = occ[1], k = occ[2];
    const double eijk = M.F[(long)i * n + i] + M.F[(long)j * n + j] + M.F[(long)k * n + k];
    double eT = 0.0, eST = 0.0;
    const long vv = (long)v * v;
    for (long e = threadIdx.x; e < vv * v; e += kTripThreads) {
      const int a = (int)(e / vv), b = (int)((e / v) % v), c = (int)(e % v);
      const double Wabc = W[e];
      const double r3 = 4.0 * Wabc + W[((long)b * v + c) * v + a] + W[((long)c * v + a) * v + b] -
                        2.0 * (W[((long)a * v + c) * v + b] + W[((long)b * v + a) * v + c] + W[((long)c * v + b) * v + a]);
      const double D = eijk - M.F[(long)(o + a) * n + o + a] - M.F[(long)(o + b) * n + o + b] -
                       M.F[(long)(o + c) * n + o + c];
      const double V = M.t1[(long)i * v + a] * M.L[(((long)j * o + k) * v + b) * v + c] +
                       M.t1[(long)j * v + b] * M.L[(((long)i * o + k) * v + a) * v + c] +
                       M.t1[(long)k * v + c] * M.L[(((long)i * o + j) * v + a) * v + b];
      eT += Wabc * r3 / (3.0 * D);
      eST += V * r3 / (3.0 * D);
    }
    eT = trip_block_sum(eT, red);
    eST = trip_block_sum(eST, red);
    const double m = (i > j && j > k) ? 6.0 : ((i == j && j == k) ? 1.0 : 3.0);
    if (threadIdx.x == 0) {
      part[2 * item] = m * eT;
      part[2 * item + 1] = m * eST;
    }
  }
}

// Items of molecules with v <= kTripLdsV on the matrix cores, W resident in
// LDS.  For permutation s the product Y[(x,y)][z] = sum_d A[p][x][y][d]
// t2[r][q][z][d] - sum_l t2[p][l][x][y] (ql|rz) is a (v^2 x v) x (v + o)
// FP64 MFMA product: A fragments stream from memory (rows (x,y), K = d
// contiguous; the t2 rows are contiguous along (x,y)), the B operands
// (t2[r][q][z][d] as [d][z], (ql|rz) as [l][z], zero-padded to 32 x 32) sit in
// LDS; each 16 x 16 tile is added into W at the permuted positions.
// V_MFMA_F64_16X16X4F64: lane t holds A[t % 16][t / 16] and B[t / 16][t % 16];
// D[row][col] is element row / 4 of lane col + 16 (row % 4).
typedef double trip_d4 __attribute__((ext_vector_type(4)));
constexpr int kTripMfmaThreads = 512;
constexpr int kTripKD = (kTripLdsV + 3) / 4;   // K steps over d
constexpr int kTripKL = kTripMaxOV / 4;        // K steps over l
// LDS (doubles): W[v^3], B operands 2 x 32 x kTripBC (reused for the L slices
// of the energy pass), orbital energies and t1 rows, reduction
// W[a][b][c] at a S2 + b S1 + c with S1 = v + 1 and S2 = v S1 + 1 (both odd):
// accesses along any one index are free of LDS bank conflicts
constexpr int kTripLdsW = kTripLdsV * (kTripLdsV * (kTripLdsV + 1) + 1);
constexpr int kTripLdsE = 4 * kTripLdsV;
// The two permutations of a pair share the A operand, so their B operands sit
// side by side as one operand of 2v columns (the first permutation's at
// columns z, the second's at v + z): ceil(2v / 16) column tiles per product
// instead of 2 ceil(v / 16), no padding columns for v = 24 (48 = 3 x 16).
constexpr int kTripBC = 48;
static_assert(kTripBC >= 2 * kTripLdsV && kTripBC % 16 == 0, "combined B columns");

__global__ __launch_bounds__(kTripMfmaThreads) void k_triples_mfma(const TripMol* __restrict__ mols, int nm,
                                                                   long total, const int2* __restrict__ items,
                                                                   double* __restrict__ part) {
  extern __shared__ double lds[];
  double* W = lds;
  double* Tb = W + kTripLdsW;                 // [d][c], 32 x kTripBC: t_rq^{z d} | t_qr^{z d}
  double* Qb = Tb + 32 * kTripBC;             // [l][c], 32 x kTripBC: (ql|rz) | (rl|qz)
  double* E = Qb + 32 * kTripBC;              // e_vir (v), t1 rows of i, j, k
  double* red = E + kTripLdsE;                // kTripMfmaThreads / 64
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, l16 = lane & 15, kq = lane >> 4;
  constexpr int nwave = kTripMfmaThreads / 64;
  int2 next = blockIdx.x < total ? items[blockIdx.x] : int2{0, 0};
  for (long item = blockIdx.x; item < total; item += gridDim.x) {
    const int2 info = next;   // molecule, packed triple; the next item's is loaded now
    if (item + gridDim.x < total) next = items[item + gridDim.x];
    const TripMol& M = mols[info.x];
    const int o = M.o, v = M.v, n = M.n;
    if (v > kTripLdsV) continue;   // uniform over the workgroup: the global-W kernel takes it
    const int packed = info.y;
    const int occ[3] = {packed & 0xff, (packed >> 8) & 0xff, (packed >> 16) & 0xff};
    const int vv = v * v, rt_n = (vv + 15) >> 4, ct_n = (2 * v + 15) >> 4;
    const int S1 = v + 1, S2 = v * S1 + 1;
    for (int e = threadIdx.x; e < v * S2; e += kTripMfmaThreads) W[e] = 0.0;
    const int i = occ[0], j = occ[1], k = occ[2];
    // operands of the next phase, loaded into registers while the current
    // phase runs: the B operands of pair h (kStg elements of each 32 x kTripBC
    // array per thread), then the L slices, virtual orbital energies and t1 rows
    constexpr int kStg = 32 * kTripBC / kTripMfmaThreads;   // elements per B array and thread
    constexpr int kStgL = (3 * kTripLdsV * kTripLdsV + kTripMfmaThreads - 1) / kTripMfmaThreads;
    static_assert(kStg * kTripMfmaThreads == 32 * kTripBC, "staging covers each B array exactly");
    static_assert(kStgL + 1 <= 2 * kStg, "the staging registers hold the energy-pass data");
    static_assert(3 * kTripLdsV * kTripLdsV <= 2 * 32 * kTripBC, "the B region holds the L slices");
    static_assert(kTripMfmaThreads >= 4 * kTripLdsV, "one energy-pass vector element per thread");
    double stg[2 * kStg];
    auto fetch_b = [&](int h) {
      const int q = occ[kPerm[2 * h][1]], r = occ[kPerm[2 * h][2]];
#pragma unroll
      for (int u = 0; u < kStg; ++u) {
        const int e = threadIdx.x + u * kTripMfmaThreads, row = e / kTripBC, c = e - row * kTripBC;
        const bool first = c < v, second = !first && c < 2 * v;
        const int z = first ? c : c - v;
        const bool dz = row < v && (first || second), lz = row < o && (first || second);
        // first: t_rq^{z d} and (ql|rz); second: t_qr^{z d} and (rl|qz); d, l = row
        const long tt = first ? (((long)r * o + q) * v + z) * v + row : (((long)q * o + r) * v + z) * v + row;
        const long bb = first ? (((long)q * o + row) * o + r) * v + z : (((long)r * o + row) * o + q) * v + z;
        stg[u] = dz ? M.t2[tt] : 0.0;
        stg[kStg + u] = lz ? M.B[bb] : 0.0;
      }
    };
    // energy-pass data: element e of [Ljk | Lik | Lij] (3 v^2 <= kStgL x threads)
    // and of [e_vir | t1_i | t1_j | t1_k] (4 v <= threads)
    auto fetch_e = [&]() {
#pragma unroll
      for (int u = 0; u < kStgL; ++u) {
        const int e = threadIdx.x + u * kTripMfmaThreads, w = e / vv, f = e - w * vv;
        const long pair = w == 0 ? (long)j * o + k : (w == 1 ? (long)i * o + k : (long)i * o + j);
        stg[u] = w < 3 ? M.L[pair * vv + f] : 0.0;
      }
      const int w = threadIdx.x / v, f = threadIdx.x - w * v;
      const int row = w == 1 ? i : (w == 2 ? j : k);
      stg[kStgL] = w == 0 ? M.F[(long)(o + f) * n + o + f] : (w < 4 ? M.t1[(long)row * v + f] : 0.0);
    };
    fetch_b(0);
    // permutations 2h and 2h + 1 share p (hence the A fragments) and swap q, r;
    // the second one's output for (x, y, z) lands where the first one's lands
    // for (x, z, y): each tile adds the first, then (after a barrier) the second
    for (int h = 0; h < 3; ++h) {
      const int s = 2 * h;
      const int p = occ[kPerm[s][0]];
      const int stride3[3] = {S2, S1, 1};
      const int sx = stride3[kPerm[s][0]], sy = stride3[kPerm[s][1]], sz = stride3[kPerm[s][2]];
      __syncthreads();
#pragma unroll
      for (int u = 0; u < kStg; ++u) {
        const int e = threadIdx.x + u * kTripMfmaThreads;
        Tb[e] = stg[u];
        Qb[e] = stg[kStg + u];
      }
      __syncthreads();
      if (h < 2) fetch_b(h + 1);
      else fetch_e();
      const double* Ap = M.A + (long)p * vv * v;     // [(x,y)][d]
      const double* Up = M.t2 + (long)p * o * vv;    // [l][(x,y)]
      // A fragments of the next tile are prefetched; the t2 rows (contiguous
      // along (x, y)) are loaded at the tile (prefetching them too measured
      // slower)
      double fa[kTripKD], fu[kTripKL], na[kTripKD];
      auto fetch = [&](int rt) {
        const int row = rt * 16 + l16;
        const bool rin = rt < rt_n && row < vv;
#pragma unroll
        for (int k = 0; k < kTripKD; ++k) {
          const int d = 4 * k + kq;
          na[k] = rin && d < v ? Ap[(long)row * v + d] : 0.0;
        }
      };
      fetch(wave);
      // every phase writes one permutation's tiles (disjoint positions): the
      // block steps through the tiles together, so the barriers line up
      const int tiles_per_wave = (rt_n + nwave - 1) / nwave;
      for (int t = 0; t < tiles_per_wave; ++t) {
        const int rt = wave + t * nwave;
        const bool has = rt < rt_n;
        // column tiles of the combined B operand: columns c < v belong to
        // permutation 2h (z = c), v <= c < 2v to permutation 2h + 1 (z = c - v)
        constexpr int kCt = kTripBC / 16;
        trip_d4 acc[kCt];
#pragma unroll
        for (int ct = 0; ct < kCt; ++ct) acc[ct] = trip_d4{0.0, 0.0, 0.0, 0.0};
        // (x, y) of the lane's four output rows rr = 16 rt + kq + 4 e, from one
        // division and carries (x < v exactly when rr < v^2)
        int rx[4], ry[4];
        {
          const int base = rt * 16 + kq;
          int x = base / v, y = base - x * v;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            rx[e] = x;
            ry[e] = y;
            y += 4;
            while (y >= v) {
              y -= v;
              ++x;
            }
          }
        }
        if (has) {
#pragma unroll
          for (int k = 0; k < kTripKD; ++k) fa[k] = na[k];
#pragma unroll
          for (int k = 0; k < kTripKL; ++k) {
            const int l = 4 * k + kq, row = rt * 16 + l16;
            fu[k] = row < vv && l < o ? -Up[(long)l * vv + row] : 0.0;
          }
          if (rt + nwave < rt_n) fetch(rt + nwave);
#pragma unroll
          for (int k = 0; k < kTripKD; ++k) {
            if (4 * k >= v) break;
            const int d = 4 * k + kq;
#pragma unroll
            for (int ct = 0; ct < kCt; ++ct) {
              if (ct >= ct_n) break;
              acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[k], Tb[d * kTripBC + 16 * ct + l16], acc[ct], 0, 0, 0);
            }
          }
#pragma unroll
          for (int k = 0; k < kTripKL; ++k) {
            if (4 * k >= o) break;
            const int l = 4 * k + kq;
#pragma unroll
            for (int ct = 0; ct < kCt; ++ct) {
              if (ct >= ct_n) break;
              acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(fu[k], Qb[l * kTripBC + 16 * ct + l16], acc[ct], 0, 0, 0);
            }
          }
          // permutation 2h: rows (x, y), columns z < v at x sx + y sy + z sz
#pragma unroll
          for (int ct = 0; ct < kCt; ++ct) {
            const int z = 16 * ct + l16;
            if (ct >= ct_n || z >= v) continue;
#pragma unroll
            for (int e = 0; e < 4; ++e)
              if (rx[e] < v) W[rx[e] * sx + ry[e] * sy + z * sz] += acc[ct][e];
          }
        }
        __syncthreads();
        // permutation 2h + 1: columns v + z of the same tiles at x sx + y sz + z sy
        if (has) {
#pragma unroll
          for (int ct = 0; ct < kCt; ++ct) {
            const int z = 16 * ct + l16 - v;
            if (ct >= ct_n || z < 0 || z >= v) continue;
#pragma unroll
            for (int e = 0; e < 4; ++e)
              if (rx[e] < v) W[rx[e] * sx + ry[e] * sz + z * sy] += acc[ct][e];
          }
        }
        __syncthreads();
      }
    }
    // energy of the item: L slices (jb|kc), (ia|kc), (ia|jb) (in the B-operand
    // region), virtual orbital energies and t1 rows of i, j, k in LDS
    __syncthreads();
    double* Ljk = Tb;
    double* Lik = Tb + vv;
    double* Lij = Tb + 2 * vv;
    double* ev = E;
    double* t1i = E + v;
    double* t1j = t1i + v;
    double* t1k = t1j + v;
#pragma unroll
    for (int u = 0; u < kStgL; ++u) {
      const int e = threadIdx.x + u * kTripMfmaThreads;
      if (e < 3 * vv) Tb[e] = stg[u];
    }
    if ((int)threadIdx.x < 4 * v) E[threadIdx.x] = stg[kStgL];
    __syncthreads();
    const double eijk = M.F[(long)i * n + i] + M.F[(long)j * n + j] + M.F[(long)k * n + k];
    double eT = 0.0, eST = 0.0;
    // element e = (a v + b) v + c; the thread's elements step by the block
    // size, whose (a, b, c) step is added with carries instead of dividing
    const int st_a = kTripMfmaThreads / vv, st_r = kTripMfmaThreads - st_a * vv, st_b = st_r / v,
              st_c = st_r - st_b * v;
    int a = (int)threadIdx.x / vv, b, c;
    {
      const int r = (int)threadIdx.x - a * vv;
      b = r / v;
      c = r - b * v;
    }
    for (int e = threadIdx.x; e < vv * v; e += kTripMfmaThreads) {
      const double Wabc = W[a * S2 + b * S1 + c];
      const double r3 = 4.0 * Wabc + W[b * S2 + c * S1 + a] + W[c * S2 + a * S1 + b] -
                        2.0 * (W[a * S2 + c * S1 + b] + W[b * S2 + a * S1 + c] + W[c * S2 + b * S1 + a]);
      const double g = r3 / (3.0 * (eijk - ev[a] - ev[b] - ev[c]));
      const double V = t1i[a] * Ljk[b * v + c] + t1j[b] * Lik[a * v + c] + t1k[c] * Lij[a * v + b];
      eT += Wabc * g;
      eST += V * g;
      c += st_c;
      b += st_b;
      a += st_a;
      if (c >= v) {
        c -= v;
        ++b;
      }
      if (b >= v) {
        b -= v;
        ++a;
      }
    }
    // block sums (fixed order)
    for (int off = 32; off > 0; off >>= 1) {
      eT += __shfl_xor(eT, off);
      eST += __shfl_xor(eST, off);
    }
    if (lane == 0) {
      red[wave] = eT;
      red[nwave + wave] = eST;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0, st = 0.0;
      for (int w = 0; w < nwave; ++w) {
        t += red[w];
        st += red[nwave + w];
      }
      const double m = (i > j && j > k) ? 6.0 : ((i == j && j == k) ? 1.0 : 3.0);
      part[2 * item] = m * t;
      part[2 * item + 1] = m * st;
    }
  }
}

// per molecule, in item order: out[2 m] = E[T], out[2 m + 1] = E[ST]
__global__ void k_triples_final(int nm, const long* __restrict__ prefix, const double* __restrict__ part,
                                double* __restrict__ out) {
  const int mol = blockIdx.x * blockDim.x + threadIdx.x;
  if (mol >= nm) return;
  double t = 0.0, st = 0.0;
  for (long it = prefix[mol]; it < prefix[mol + 1]; ++it) {
    t += part[2 * it];
    st += part[2 * it + 1];
  }
  out[2 * mol] = t;
  out[2 * mol + 1] = st;
}

size_t triples_lds_bytes() { return sizeof(double) * (kTripThreads / 64 + 2 * kTripMaxV * kTripMaxV); }
size_t triples_mfma_lds_bytes() {
  return sizeof(double) * (kTripLdsW + 2 * 32 * kTripBC + kTripLdsE + 2 * (kTripMfmaThreads / 64));
}

hipError_t launch_triples(const TripMol* mols, int nm, const long* prefix, long total, const int2* items,
                          double* wscratch, long wstride, int grid, int grid_mfma, bool any_global, double* part,
                          double* out, hipStream_t s) {
  if (total > 0 && grid_mfma > 0) {
    const hipError_t e = hipFuncSetAttribute((const void*)k_triples_mfma, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)triples_mfma_lds_bytes());
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_triples_mfma, dim3(grid_mfma), dim3(kTripMfmaThreads), triples_mfma_lds_bytes(), s, mols, nm,
                       total, items, part);
  }
  if (total > 0 && any_global)
    hipLaunchKernelGGL(k_triples, dim3(grid), dim3(kTripThreads), triples_lds_bytes(), s, mols, nm, total, items,
                       wscratch, wstride, part);
  hipLaunchKernelGGL(k_triples_final, dim3((nm + 63) / 64), dim3(64), 0, s, nm, prefix, part, out);
  return hipGetLastError();
}

}  // namespace hcc
