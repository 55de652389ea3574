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
                                                          const long* __restrict__ prefix, long total,
                                                          const int* __restrict__ tri, double* __restrict__ wscratch,
                                                          long wstride, double* __restrict__ part) {
  extern __shared__ double lds[];
  double* red = lds;                      // kTripThreads / 64
  double* Ts = red + kTripThreads / 64;   // [d][z], v x v
  double* Qs = Ts + kTripMaxV * kTripMaxV;   // [l][z], o x v
  double* W = wscratch + (long)blockIdx.x * wstride;
  for (long item = blockIdx.x; item < total; item += gridDim.x) {
    // molecule of the item (prefix: nm + 1 entries)
    int lo = 0, hi = nm - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (prefix[mid] <= item) lo = mid; else hi = mid - 1;
    }
    const TripMol& M = mols[lo];
    const int o = M.o, v = M.v, n = M.n;
    const int packed = tri[item - prefix[lo]];
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

hipError_t launch_triples(const TripMol* mols, int nm, const long* prefix, long total, const int* tri,
                          double* wscratch, long wstride, int grid, double* part, double* out, hipStream_t s) {
  if (total > 0)
    hipLaunchKernelGGL(k_triples, dim3(grid), dim3(kTripThreads), triples_lds_bytes(), s, mols, nm, prefix, total, tri,
                       wscratch, wstride, part);
  hipLaunchKernelGGL(k_triples_final, dim3((nm + 63) / 64), dim3(64), 0, s, nm, prefix, part, out);
  return hipGetLastError();
}

}  // namespace hcc
