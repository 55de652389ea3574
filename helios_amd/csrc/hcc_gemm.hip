// Grouped composite-strided GEMM on CDNA4 matrix cores.
//
// One launch runs every problem of a table (one problem per molecule, or a
// single problem): C = alpha * A B + beta * C.  Work items are 64x64 output
// tiles; 256 threads (4 waves of 64) per tile, each wave owning a 32x32
// quadrant computed with 2x2 V_MFMA_F64_16X16X4_F64 blocks.  Operand tiles
// are staged through LDS (k-major), with the next k-slab prefetched into
// registers while the current one feeds the matrix cores.
//
// Tile -> workgroup mapping is XCD-aware: consecutive workgroups are spread
// round-robin over the 8 XCDs by the dispatcher, so logical tile
// (b % 8) * ceil(T/8) + b / 8 gives each XCD one contiguous run of tiles and
// keeps a molecule's tiles (which share operands) in one L2.
#include "hcc_internal.h"

namespace hcc {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ int row_off(const Operand& o, int r) {
  return (r / o.rr) * o.srh + (r % o.rr) * o.srl;
}
__device__ __forceinline__ int col_off(const Operand& o, int c) {
  return (c / o.rc) * o.sch + (c % o.rc) * o.scl;
}

// Finds the problem owning global tile `tile` (binary search on tile_begin).
__device__ __forceinline__ int find_problem(const GemmProb* probs, int nprob, int tile) {
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (probs[mid].tile_begin <= tile) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int logical_tile(int total) {
  int t8 = (total + 7) >> 3;
  return (int)(blockIdx.x & 7) * t8 + (int)(blockIdx.x >> 3);
}

constexpr int kLdA = kBM + 4;   // LDS row pitch (doubles) of the k-major A tile
constexpr int kLdB = kBN + 4;

// Loads this thread's 4 A elements and 4 B elements of k-slab k0.
struct Frag {
  double a[4], b[4];
};

__device__ __forceinline__ void load_slab(const GemmProb& P, int m0, int n0, int k0, int t,
                                          const int (&ra)[4], const int (&cb)[4], Frag& f) {
  const double* __restrict__ A = P.A.p;
  const double* __restrict__ B = P.B.p;
  if (P.a_kfast) {
    int k = k0 + (t & 15);
    bool kin = k < P.K;
    int co = kin ? col_off(P.A, k) : 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int m = m0 + (t >> 4) + 16 * i;
      f.a[i] = (kin && m < P.M) ? A[ra[i] + co] : 0.0;
    }
  } else {
    int m = m0 + (t & 63);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int k = k0 + (t >> 6) + 4 * i;
      f.a[i] = (m < P.M && k < P.K) ? A[ra[0] + col_off(P.A, k)] : 0.0;
    }
  }
  if (P.b_nfast) {
    int n = n0 + (t & 63);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int k = k0 + (t >> 6) + 4 * i;
      f.b[i] = (n < P.N && k < P.K) ? B[row_off(P.B, k) + cb[0]] : 0.0;
    }
  } else {
    int k = k0 + (t & 15);
    bool kin = k < P.K;
    int ro = kin ? row_off(P.B, k) : 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int n = n0 + (t >> 4) + 16 * i;
      f.b[i] = (kin && n < P.N) ? B[ro + cb[i]] : 0.0;
    }
  }
}

__device__ __forceinline__ void store_slab(const GemmProb& P, int t, const Frag& f,
                                           double* As, double* Bs) {
  if (P.a_kfast) {
#pragma unroll
    for (int i = 0; i < 4; ++i) As[(t & 15) * kLdA + (t >> 4) + 16 * i] = f.a[i];
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) As[((t >> 6) + 4 * i) * kLdA + (t & 63)] = f.a[i];
  }
  if (P.b_nfast) {
#pragma unroll
    for (int i = 0; i < 4; ++i) Bs[((t >> 6) + 4 * i) * kLdB + (t & 63)] = f.b[i];
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) Bs[(t & 15) * kLdB + (t >> 4) + 16 * i] = f.b[i];
  }
}

__global__ __launch_bounds__(kThreads) void gemm_f64_kernel(const GemmProb* __restrict__ probs,
                                                            int nprob, int total_tiles,
                                                            const int* __restrict__ active) {
  __shared__ double As[kBK * kLdA];
  __shared__ double Bs[kBK * kLdB];
  const int tile = logical_tile(total_tiles);
  if (tile >= total_tiles) return;
  const GemmProb P = probs[find_problem(probs, nprob, tile)];
  if (active != nullptr && active[P.mol] == 0) return;
  const int lt = tile - P.tile_begin;
  const int m0 = (lt / P.tiles_n) * kBM;
  const int n0 = (lt % P.tiles_n) * kBN;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;

  // fixed row offsets of A and column offsets of B for this thread
  int ra[4], cb[4];
  if (P.a_kfast) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int m = m0 + (t >> 4) + 16 * i;
      ra[i] = m < P.M ? row_off(P.A, m) : 0;
    }
  } else {
    int m = m0 + (t & 63);
    ra[0] = m < P.M ? row_off(P.A, m) : 0;
    ra[1] = ra[2] = ra[3] = 0;
  }
  if (P.b_nfast) {
    int n = n0 + (t & 63);
    cb[0] = n < P.N ? col_off(P.B, n) : 0;
    cb[1] = cb[2] = cb[3] = 0;
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int n = n0 + (t >> 4) + 16 * i;
      cb[i] = n < P.N ? col_off(P.B, n) : 0;
    }
  }

  d4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

  Frag f;
  load_slab(P, m0, n0, 0, t, ra, cb, f);
  for (int k0 = 0; k0 < P.K; k0 += kBK) {
    store_slab(P, t, f, As, Bs);
    __syncthreads();
    if (k0 + kBK < P.K) load_slab(P, m0, n0, k0 + kBK, t, ra, cb, f);
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 4) {
      const int kr = kk + (lane >> 4);
      double a0 = As[kr * kLdA + wm + (lane & 15)];
      double a1 = As[kr * kLdA + wm + 16 + (lane & 15)];
      double b0 = Bs[kr * kLdB + wn + (lane & 15)];
      double b1 = Bs[kr * kLdB + wn + 16 + (lane & 15)];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }

  // epilogue: for the FP64 MFMA, D[row][col] of a 16x16 block sits in lane
  // (col + 16*(row%4)), element row/4
  double* __restrict__ Cp = P.C.p;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int col = n0 + wn + 16 * j + (lane & 15);
    if (col >= P.N) continue;
    const int co = col_off(P.C, col);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + 16 * i + (lane >> 4) + 4 * r;
        if (row >= P.M) continue;
        const int o = row_off(P.C, row) + co;
        double v = P.alpha * acc[i][j][r];
        if (P.beta != 0.0) v += P.beta * Cp[o];
        Cp[o] = v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// BF16 variant: FP64 operands are rounded to BF16 while being staged into LDS,
// multiplied on V_MFMA_F32_16X16X32_BF16 with FP32 accumulation, and the FP32
// result is combined with C in FP64 (C = alpha*acc + beta*C).  BF16 keeps
// FP32's exponent range, so no per-tile scaling is needed against overflow;
// its 8-bit significand bounds the accuracy, which the solver restores with
// FP64 residuals.  Tiles: 64x64, BK=32, one 16x16x32 MFMA step per slab.
// ---------------------------------------------------------------------------
constexpr int kBK16 = 32;
constexpr int kLd16 = kBK16 + 8;   // bf16 elements per LDS row (k contiguous)

__device__ __forceinline__ void load_slab16(const GemmProb& P, int m0, int n0, int k0, int t, double (&a)[8],
                                            double (&b)[8]) {
  const double* __restrict__ A = P.A.p;
  const double* __restrict__ B = P.B.p;
  if (P.a_kfast) {
    const int m = m0 + (t >> 2);
    const int kb = k0 + (t & 3) * 8;
    const int ro = m < P.M ? row_off(P.A, m) : 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int k = kb + q;
      a[q] = (m < P.M && k < P.K) ? A[ro + col_off(P.A, k)] : 0.0;
    }
  } else {
    const int m = m0 + (t & 63);
    const int kb = k0 + (t >> 6) * 8;
    const int ro = m < P.M ? row_off(P.A, m) : 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int k = kb + q;
      a[q] = (m < P.M && k < P.K) ? A[ro + col_off(P.A, k)] : 0.0;
    }
  }
  if (P.b_nfast) {
    const int n = n0 + (t & 63);
    const int kb = k0 + (t >> 6) * 8;
    const int co = n < P.N ? col_off(P.B, n) : 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int k = kb + q;
      b[q] = (n < P.N && k < P.K) ? B[row_off(P.B, k) + co] : 0.0;
    }
  } else {
    const int n = n0 + (t >> 2);
    const int kb = k0 + (t & 3) * 8;
    const int co = n < P.N ? col_off(P.B, n) : 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int k = kb + q;
      b[q] = (n < P.N && k < P.K) ? B[row_off(P.B, k) + co] : 0.0;
    }
  }
}

__device__ __forceinline__ void store_slab16(const GemmProb& P, int t, const double (&a)[8], const double (&b)[8],
                                             __bf16* As, __bf16* Bs) {
  // A as [m][k], B as [n][k]; k contiguous so that a lane reads its 8 k-values at once
  const int am = P.a_kfast ? (t >> 2) : (t & 63);
  const int ak = P.a_kfast ? (t & 3) * 8 : (t >> 6) * 8;
  const int bn = P.b_nfast ? (t & 63) : (t >> 2);
  const int bk = P.b_nfast ? (t >> 6) * 8 : (t & 3) * 8;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    As[am * kLd16 + ak + q] = (__bf16)(float)a[q];
    Bs[bn * kLd16 + bk + q] = (__bf16)(float)b[q];
  }
}

__global__ __launch_bounds__(kThreads) void gemm_bf16_kernel(const GemmProb* __restrict__ probs, int nprob,
                                                             int total_tiles, const int* __restrict__ active) {
  __shared__ __attribute__((aligned(16))) __bf16 As[kBM * kLd16];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[kBN * kLd16];
  const int tile = logical_tile(total_tiles);
  if (tile >= total_tiles) return;
  const GemmProb P = probs[find_problem(probs, nprob, tile)];
  if (active != nullptr && active[P.mol] == 0) return;
  const int lt = tile - P.tile_begin;
  const int m0 = (lt / P.tiles_n) * kBM;
  const int n0 = (lt % P.tiles_n) * kBN;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  double a[8], b[8];
  load_slab16(P, m0, n0, 0, t, a, b);
  for (int k0 = 0; k0 < P.K; k0 += kBK16) {
    store_slab16(P, t, a, b, As, Bs);
    __syncthreads();
    if (k0 + kBK16 < P.K) load_slab16(P, m0, n0, k0 + kBK16, t, a, b);
    const int kq = 8 * (lane >> 4);
    bf8 a0 = *reinterpret_cast<const bf8*>(&As[(wm + (lane & 15)) * kLd16 + kq]);
    bf8 a1 = *reinterpret_cast<const bf8*>(&As[(wm + 16 + (lane & 15)) * kLd16 + kq]);
    bf8 b0 = *reinterpret_cast<const bf8*>(&Bs[(wn + (lane & 15)) * kLd16 + kq]);
    bf8 b1 = *reinterpret_cast<const bf8*>(&Bs[(wn + 16 + (lane & 15)) * kLd16 + kq]);
    acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc[1][1], 0, 0, 0);
    __syncthreads();
  }
  double* __restrict__ Cp = P.C.p;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int col = n0 + wn + 16 * j + (lane & 15);
    if (col >= P.N) continue;
    const int co = col_off(P.C, col);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
        if (row >= P.M) continue;
        const int o = row_off(P.C, row) + co;
        double v = P.alpha * (double)acc[i][j][r];
        if (P.beta != 0.0) v += P.beta * Cp[o];
        Cp[o] = v;
      }
    }
  }
}

hipError_t launch_gemm_bf16(const GemmProb* probs, int nprob, int total_tiles, const int* active,
                            hipStream_t s) {
  if (total_tiles <= 0 || nprob <= 0) return hipSuccess;
  int grid = ((total_tiles + 7) / 8) * 8;
  hipLaunchKernelGGL(gemm_bf16_kernel, dim3(grid), dim3(kThreads), 0, s, probs, nprob, total_tiles, active);
  return hipGetLastError();
}

hipError_t launch_gemm_f64(const GemmProb* probs, int nprob, int total_tiles, const int* active,
                           hipStream_t s) {
  if (total_tiles <= 0 || nprob <= 0) return hipSuccess;
  int grid = ((total_tiles + 7) / 8) * 8;
  hipLaunchKernelGGL(gemm_f64_kernel, dim3(grid), dim3(kThreads), 0, s, probs, nprob,
                     total_tiles, active);
  return hipGetLastError();
}

}  // namespace hcc
