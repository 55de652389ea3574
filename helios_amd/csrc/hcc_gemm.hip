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

// Operand pointers arrive as generic pointers inside GemmProb; loads through
// them compile to FLAT instructions, which also count against the LDS
// counter, so waiting for the LDS fragment reads would wait for the next
// slab's prefetch as well.  Global-address-space views give GLOBAL loads that
// stay in flight across the matrix-core work of the current slab.
typedef __attribute__((address_space(1))) double gdouble;
__device__ __forceinline__ const gdouble* gptr(const double* p) { return (const gdouble*)p; }
__device__ __forceinline__ gdouble* gptr(double* p) { return (gdouble*)p; }

// Operand loads through buffer resources: a masked (padding) element uses a
// byte offset beyond the resource's range and reads 0 in hardware, so no
// select follows the load and nothing has to wait for it before the next
// slab's store to LDS.  Valid offsets stay below 2 GiB (the largest operand is
// far smaller).
constexpr int kRsrcBytes = 0x7ffffff0;
constexpr int kOobOffset = 0x7ffffff8;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const double* p) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)p, 0, kRsrcBytes, 0x00020000);
}
__device__ __forceinline__ double bload(__amdgpu_buffer_rsrc_t r, int byte_off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, 0, 0));
}
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc16(const __bf16* p) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)p, 0, kRsrcBytes, 0x00020000);
}
// FP32 element at an FP64 byte offset of the same array (element index = off / 8)
__device__ __forceinline__ float bload32(__amdgpu_buffer_rsrc_t r, int byte_off64) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, byte_off64 == kOobOffset ? kOobOffset
                                                                                                 : byte_off64 >> 1,
                                                                         0, 0));
}
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrcf(const float* p) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)p, 0, kRsrcBytes, 0x00020000);
}
// BF16 element at an FP64 byte offset of the same array (element index = off / 8)
__device__ __forceinline__ unsigned short bload16(__amdgpu_buffer_rsrc_t r, int byte_off64) {
  return __builtin_amdgcn_raw_buffer_load_b16(r, byte_off64 == kOobOffset ? kOobOffset : byte_off64 >> 2, 0, 0);
}

__device__ __forceinline__ int row_off(const Operand& o, int r) {
  return (r / o.rr) * o.srh + (r % o.rr) * o.srl;
}
__device__ __forceinline__ int col_off(const Operand& o, int c) {
  return (c / o.rc) * o.sch + (c % o.rc) * o.scl;
}

// Problem owning global tile `tile`: one load from the table's tile map
// (a binary search over problems would be a chain of dependent loads).
__device__ __forceinline__ int find_problem(const int* __restrict__ tile_prob, int tile) {
  return tile_prob[tile];
}

__device__ __forceinline__ int logical_tile(int total) {
  int t8 = (total + 7) >> 3;
  return (int)(blockIdx.x & 7) * t8 + (int)(blockIdx.x >> 3);
}

// ---------------------------------------------------------------------------
// FP64 kernel.  Template on the output tile (BM x BN, 32x32 per wave, 4 or 1
// waves).  LDS layouts follow the operand's unit stride so that both the
// staging stores and the MFMA fragment reads are free of bank conflicts:
//   K-contiguous operands are staged [row][k] with a 17-double pitch,
//   row-contiguous operands [k][row] with a (BM+2)-double pitch.
// The composite-index offsets along K are precomputed once per work item into
// LDS tables (chunks up to kKTab), so the slab loop does no integer division.
// ---------------------------------------------------------------------------
constexpr int kKTab = 2048;
constexpr int kPk = kBK + 1;   // pitch of the [row][k] staging layout (16-deep slabs)

// K-contiguous operands are staged [k / 16][row][k % 16] with a 17-double
// row pitch (16-deep panels), whatever the slab depth BK.
template <int BM, int BN, int BK>
struct F64Tile {
  static constexpr int kWaves = (BM / 32) * (BN / 32);
  static constexpr int kT = 64 * kWaves;
  // pitch of the [k][m] layout: == 16 (mod 32) doubles, so the fragment reads
  // of rows k and k+1 (lanes 0-15 and 16-31) fall in disjoint bank halves
  static constexpr int kPa = BM + 16;
  static constexpr int kPb = BN + 16;
  static constexpr int kLdsA = ((BK / 16) * BM * kPk > BK * kPa) ? (BK / 16) * BM * kPk : BK * kPa;
  static constexpr int kLdsB = ((BK / 16) * BN * kPk > BK * kPb) ? (BK / 16) * BN * kPk : BK * kPb;
  // offset of (row, k) in the K-contiguous layout of an R-row operand
  template <int R>
  static __device__ __forceinline__ int rk(int row, int k) { return (k >> 4) * (R * kPk) + row * kPk + (k & 15); }
  static constexpr int kNA = BM * BK / kT;    // A elements per thread per slab
  static constexpr int kNB = BN * BK / kT;
  // K-offset table length (LDS) per chunk: small enough that the table does
  // not limit the workgroups per CU
  static constexpr int kKT = BM >= 64 ? 1024 : 512;
};

#define HCC_F64_THREADS(BM_, BN_) (64 * ((BM_) / 32) * ((BN_) / 32))

// AKF: A's unit stride runs along K; BNF: B's unit stride runs along N.  Both
// are template parameters so the slab loop is straight-line code and the
// accumulators stay in AGPRs across iterations.  BK: slab depth (16 or 32;
// 32 halves the barriers per K for long contractions).
template <int BM, int BN, int BK, bool AKF, bool BNF>
__global__ __launch_bounds__(HCC_F64_THREADS(BM, BN)) void gemm_f64_kernel(const GemmProb* __restrict__ probs,
                                                                           const int* __restrict__ tile_prob,
                                                                           int total_tiles,
                                                                           const int* __restrict__ active) {
  using TL = F64Tile<BM, BN, BK>;
  constexpr int T = TL::kT;
  __shared__ double As[TL::kLdsA];
  __shared__ double Bs[TL::kLdsB];
  __shared__ int koffA[TL::kKT];
  __shared__ int koffB[TL::kKT];
  const int tile = logical_tile(total_tiles);
  if (tile >= total_tiles) return;
  const GemmProb P = probs[find_problem(tile_prob, tile)];
  if (active != nullptr && active[P.mol] == 0) return;
  const int tiles_n = (P.N + BN - 1) / BN;
  const int tiles_mn = ((P.M + BM - 1) / BM) * tiles_n;
  int lt = tile - P.tile_begin;
  const int ks = lt / tiles_mn;
  lt -= ks * tiles_mn;
  const int m0 = (lt / tiles_n) * BM;
  const int n0 = (lt % tiles_n) * BN;
  const int kbeg = ks * P.kchunk;
  const int kend = min(P.K, kbeg + P.kchunk);
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = (wave / (BN / 32)) * 32, wn = (wave % (BN / 32)) * 32;

  // fixed (non-K) offsets of this thread's staged elements; -1 marks padding
  int fa[TL::kNA], fb[TL::kNB], kqa[TL::kNA], kqb[TL::kNB];
#pragma unroll
  for (int i = 0; i < TL::kNA; ++i) {
    const int q = t + i * T;
    const int m = m0 + (AKF ? q / BK : q % BM);
    kqa[i] = AKF ? q % BK : q / BM;
    fa[i] = m < P.M ? row_off(P.A, m) : -1;
  }
#pragma unroll
  for (int i = 0; i < TL::kNB; ++i) {
    const int q = t + i * T;
    const int n = n0 + (BNF ? q % BN : q / BK);
    kqb[i] = BNF ? q / BN : q % BK;
    fb[i] = n < P.N ? col_off(P.B, n) : -1;
  }
  const __amdgpu_buffer_rsrc_t rA = rsrc(P.A.p), rB = rsrc(P.B.p);
  double ra[TL::kNA], rb[TL::kNB];

  d4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

  for (int c0 = kbeg; c0 < kend; c0 += TL::kKT) {
    const int cend = min(kend, c0 + TL::kKT);
    __syncthreads();
    for (int q = t; q < cend - c0; q += T) {
      koffA[q] = col_off(P.A, c0 + q);
      koffB[q] = row_off(P.B, c0 + q);
    }
    __syncthreads();
    // staging loads: padded elements read 0 through the out-of-range offset
    const int klast = cend - c0 - 1;
    auto load = [&](int k0) {
#pragma unroll
      for (int i = 0; i < TL::kNA; ++i) {
        const int kl = k0 - c0 + kqa[i];
        const int kv = koffA[min(kl, klast)];
        ra[i] = bload(rA, fa[i] >= 0 && kl <= klast ? (fa[i] + kv) * 8 : kOobOffset);
      }
#pragma unroll
      for (int i = 0; i < TL::kNB; ++i) {
        const int kl = k0 - c0 + kqb[i];
        const int kv = koffB[min(kl, klast)];
        rb[i] = bload(rB, fb[i] >= 0 && kl <= klast ? (fb[i] + kv) * 8 : kOobOffset);
      }
    };
    load(c0);
    for (int k0 = c0; k0 < cend; k0 += BK) {
#pragma unroll
      for (int i = 0; i < TL::kNA; ++i) {
        const int q = t + i * T;
        if (AKF) As[TL::template rk<BM>(q / BK, q % BK)] = ra[i];
        else As[(q / BM) * TL::kPa + q % BM] = ra[i];
      }
#pragma unroll
      for (int i = 0; i < TL::kNB; ++i) {
        const int q = t + i * T;
        if (BNF) Bs[(q / BN) * TL::kPb + q % BN] = rb[i];
        else Bs[TL::template rk<BN>(q / BK, q % BK)] = rb[i];
      }
      __syncthreads();
      load(min(k0 + BK, cend));   // the last prefetch is fully masked
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        const int kr = kk + (lane >> 4);
        const int ma = wm + (lane & 15), nb = wn + (lane & 15);
        const double a0 = AKF ? As[TL::template rk<BM>(ma, kr)] : As[kr * TL::kPa + ma];
        const double a1 = AKF ? As[TL::template rk<BM>(ma + 16, kr)] : As[kr * TL::kPa + ma + 16];
        const double b0 = BNF ? Bs[kr * TL::kPb + nb] : Bs[TL::template rk<BN>(nb, kr)];
        const double b1 = BNF ? Bs[kr * TL::kPb + nb + 16] : Bs[TL::template rk<BN>(nb + 16, kr)];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
      }
      __syncthreads();
    }
  }

  // epilogue: for the FP64 MFMA, D[row][col] of a 16x16 block sits in lane
  // (col + 16*(row%4)), element row/4
  if (P.ksplit > 1) {
    gdouble* __restrict__ W = gptr(P.ws + (long)ks * P.M * P.N);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = n0 + wn + 16 * j + (lane & 15);
      if (col >= P.N) continue;
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wm + 16 * i + (lane >> 4) + 4 * r;
          if (row < P.M) W[(long)row * P.N + col] = acc[i][j][r];
        }
    }
    return;
  }
  gdouble* __restrict__ Cp = gptr(P.C.p);
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
  const gdouble* __restrict__ A = gptr(P.A.p);
  const gdouble* __restrict__ B = gptr(P.B.p);
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

__global__ __launch_bounds__(kThreads) void gemm_bf16_kernel(const GemmProb* __restrict__ probs,
                                                             const int* __restrict__ tile_prob,
                                                             int total_tiles, const int* __restrict__ active) {
  __shared__ __attribute__((aligned(16))) __bf16 As[kBM * kLd16];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[kBN * kLd16];
  const int tile = logical_tile(total_tiles);
  if (tile >= total_tiles) return;
  GemmProb P = probs[find_problem(tile_prob, tile)];
  if (active != nullptr && active[P.mol] == 0) return;
  int lt = tile - P.tile_begin;
  const int tiles_mn = ((P.M + kBM - 1) / kBM) * P.tiles_n;
  const int ks = lt / tiles_mn;
  lt -= ks * tiles_mn;
  const int m0 = (lt / P.tiles_n) * kBM;
  const int n0 = (lt % P.tiles_n) * kBN;
  const int kbeg = ks * P.kchunk;
  P.K = min(P.K, kbeg + P.kchunk);
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  double a[8], b[8];
  load_slab16(P, m0, n0, kbeg, t, a, b);
  for (int k0 = kbeg; k0 < P.K; k0 += kBK16) {
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
  if (P.ksplit > 1) {
    gdouble* __restrict__ W = gptr(P.ws + (long)ks * P.M * P.N);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = n0 + wn + 16 * j + (lane & 15);
      if (col >= P.N) continue;
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
          if (row < P.M) W[(long)row * P.N + col] = (double)acc[i][j][r];
        }
    }
    return;
  }
  gdouble* __restrict__ Cp = gptr(P.C.p);
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

hipError_t launch_gemm_bf16(const GemmProb* probs, const int* tile_prob, int total_tiles, const int* active,
                            hipStream_t s) {
  if (total_tiles <= 0) return hipSuccess;
  int grid = ((total_tiles + 7) / 8) * 8;
  hipLaunchKernelGGL(gemm_bf16_kernel, dim3(grid), dim3(kThreads), 0, s, probs, tile_prob, total_tiles, active);
  return hipGetLastError();
}


// ---------------------------------------------------------------------------
// Correction-form ("dual") GEMM of the mixed-precision inner iterations:
//   C = alpha * (A0 dB + dA (B0 + dB)) + beta * C
// which is the change of A B between the outer point (A0, B0) and the current
// point (A0 + dA, B0 + dB).  The differences and the full second factor are
// formed in FP64 before rounding, so with BF16 operands the rounding error is
// proportional to the size of the correction, not of A B.  64x64 tiles, 256
// threads; FP64 path: BK 16 and 16x16x4 MFMA; BF16 path: BK 32 and
// 16x16x32 MFMA with [row][k] LDS staging.
// ---------------------------------------------------------------------------
template <bool AKF, bool BNF, bool BF>
struct DualCfg {
  static constexpr int BK = BF ? 32 : 16;
  static constexpr int T = 256;
  static constexpr int NE = 64 * BK / T;     // elements per operand per thread per slab
};

template <bool AKF, bool BNF>
__global__ __launch_bounds__(256) void gemm_dual_f64(const GemmProb* __restrict__ probs,
                                                     const int* __restrict__ tile_prob, int total_tiles,
                                                     const int* __restrict__ active) {
  constexpr int BM = 64, BN = 64, T = 256, BK = 16, NE = 4;
  constexpr int kPa = BM + 2, kPb = BN + 2;
  __shared__ double A0s[BM * kPk > BK * kPa ? BM * kPk : BK * kPa];
  __shared__ double dAs[BM * kPk > BK * kPa ? BM * kPk : BK * kPa];
  __shared__ double dBs[BN * kPk > BK * kPb ? BN * kPk : BK * kPb];
  __shared__ double Bfs[BN * kPk > BK * kPb ? BN * kPk : BK * kPb];
  const int tile = logical_tile(total_tiles);
  if (tile >= total_tiles) return;
  const GemmProb P = probs[find_problem(tile_prob, tile)];
  if (active != nullptr && active[P.mol] == 0) return;
  const int tiles_n = (P.N + BN - 1) / BN;
  const int tiles_mn = ((P.M + BM - 1) / BM) * tiles_n;
  int lt = tile - P.tile_begin;
  const int ks = lt / tiles_mn;
  lt -= ks * tiles_mn;
  const int m0 = (lt / tiles_n) * BM, n0 = (lt % tiles_n) * BN;
  const int kbeg = ks * P.kchunk, kend = min(P.K, kbeg + P.kchunk);
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const bool hasA = P.dA != nullptr, hasB = P.dB != nullptr;
  int fa[NE], fb[NE], kqa[NE], kqb[NE];
#pragma unroll
  for (int i = 0; i < NE; ++i) {
    const int q = t + i * T;
    const int m = m0 + (AKF ? q / BK : q % BM);
    kqa[i] = AKF ? q % BK : q / BM;
    fa[i] = m < P.M ? row_off(P.A, m) : -1;
    const int n = n0 + (BNF ? q % BN : q / BK);
    kqb[i] = BNF ? q / BN : q % BK;
    fb[i] = n < P.N ? col_off(P.B, n) : -1;
  }
  const gdouble* __restrict__ A0 = gptr(P.A.p);
  const float* __restrict__ dA = P.dA;
  const gdouble* __restrict__ B0 = gptr(P.B.p);
  const float* __restrict__ dB = P.dB;
  double ra0[NE], rda[NE], rdb[NE], rbf[NE];
  auto load = [&](int k0) {
#pragma unroll
    for (int i = 0; i < NE; ++i) {
      int k = k0 + kqa[i];
      bool ok = fa[i] >= 0 && k < kend;
      int o = ok ? fa[i] + col_off(P.A, k) : 0;
      const double a0 = A0[o], da = hasA ? (double)dA[o] : 0.0;
      ra0[i] = ok ? a0 : 0.0;
      rda[i] = (ok && hasA) ? da : 0.0;
      k = k0 + kqb[i];
      ok = fb[i] >= 0 && k < kend;
      o = ok ? fb[i] + row_off(P.B, k) : 0;
      const double b0 = B0[o], db = hasB ? (double)dB[o] : 0.0;
      rdb[i] = (ok && hasB) ? db : 0.0;
      rbf[i] = ok ? b0 + rdb[i] : 0.0;
    }
  };
  d4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  load(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
#pragma unroll
    for (int i = 0; i < NE; ++i) {
      const int q = t + i * T;
      const int ia = AKF ? (q / BK) * kPk + q % BK : (q / BM) * kPa + q % BM;
      const int ib = BNF ? (q / BN) * kPb + q % BN : (q / BK) * kPk + q % BK;
      A0s[ia] = ra0[i];
      dAs[ia] = rda[i];
      dBs[ib] = rdb[i];
      Bfs[ib] = rbf[i];
    }
    __syncthreads();
    load(min(k0 + BK, kend));
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kr = kk + (lane >> 4);
      const int ma = wm + (lane & 15), nb = wn + (lane & 15);
      const int ia0 = AKF ? ma * kPk + kr : kr * kPa + ma;
      const int ia1 = AKF ? (ma + 16) * kPk + kr : kr * kPa + ma + 16;
      const int ib0 = BNF ? kr * kPb + nb : nb * kPk + kr;
      const int ib1 = BNF ? kr * kPb + nb + 16 : (nb + 16) * kPk + kr;
      const double a0 = A0s[ia0], a1 = A0s[ia1], e0 = dAs[ia0], e1 = dAs[ia1];
      const double b0 = dBs[ib0], b1 = dBs[ib1], f0 = Bfs[ib0], f1 = Bfs[ib1];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(e0, f0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(e0, f1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(e1, f0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(e1, f1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
  if (P.ksplit > 1) {
    gdouble* __restrict__ W = gptr(P.ws + (long)ks * P.M * P.N);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = n0 + wn + 16 * j + (lane & 15);
      if (col >= P.N) continue;
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wm + 16 * i + (lane >> 4) + 4 * r;
          if (row < P.M) W[(long)row * P.N + col] = acc[i][j][r];
        }
    }
    return;
  }
  float* __restrict__ Cp = P.Cd;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int col = n0 + wn + 16 * j + (lane & 15);
    if (col >= P.N) continue;
    const int co = col_off(P.C, col);
#pragma unroll
    for (int i = 0; i < 2; ++i)
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

// HA / HB: the operand has a correction (dA / dB non-null for every problem
// of the table).  A constant A needs only A0 dB, a constant B only dA B0.
// S: tile scale -- 64S x 64S output tiles, 4 waves of 32S x 32S (S = 1, 2);
// larger tiles halve the operand traffic per product, which bounds this kernel.
template <int S, bool AKF, bool BNF, bool HA, bool HB>
__global__ __launch_bounds__(256) void gemm_dual_bf16(const GemmProb* __restrict__ probs,
                                                      const int* __restrict__ tile_prob, int total_tiles,
                                                      const int* __restrict__ active) {
  constexpr int BM = 64 * S, BN = 64 * S, T = 256, BK = 32, NE = 8 * S, LD = BK + 8, F = 2 * S;
  constexpr int KT = 512;   // K-offset table length per chunk
  __shared__ __attribute__((aligned(16))) __bf16 A0s[BM * LD];
  __shared__ __attribute__((aligned(16))) __bf16 dAs[BM * LD];
  __shared__ __attribute__((aligned(16))) __bf16 dBs[BN * LD];
  __shared__ __attribute__((aligned(16))) __bf16 Bfs[BN * LD];
  __shared__ int koffA[KT];
  __shared__ int koffB[KT];
  const int tile = logical_tile(total_tiles);
  if (tile >= total_tiles) return;
  const GemmProb P = probs[find_problem(tile_prob, tile)];
  if (active != nullptr && active[P.mol] == 0) return;
  const int tiles_n = (P.N + BN - 1) / BN;
  const int tiles_mn = ((P.M + BM - 1) / BM) * tiles_n;
  int lt = tile - P.tile_begin;
  const int ks = lt / tiles_mn;
  lt -= ks * tiles_mn;
  const int m0 = (lt / tiles_n) * BM, n0 = (lt % tiles_n) * BN;
  const int kbeg = ks * P.kchunk, kend = min(P.K, kbeg + P.kchunk);
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = (wave >> 1) * (32 * S), wn = (wave & 1) * (32 * S);
  // staged element i of this thread: row ml(i), k kq(i) of the slab (element t + i T of
  // the [row][k] or [k][row] order of the operand's unit stride)
  auto mlA = [&](int i) { return AKF ? t / BK + i * (T / BK) : t % BM; };
  auto kqA = [&](int i) { return AKF ? t % BK : t / BM + i * (T / BM); };
  auto nlB = [&](int i) { return BNF ? t % BN : t / BK + i * (T / BK); };
  auto kqB = [&](int i) { return BNF ? t / BN + i * (T / BN) : t % BK; };
  int fa[NE], fb[NE];
#pragma unroll
  for (int i = 0; i < NE; ++i) {
    fa[i] = m0 + mlA(i) < P.M ? row_off(P.A, m0 + mlA(i)) : -1;
    fb[i] = n0 + nlB(i) < P.N ? col_off(P.B, n0 + nlB(i)) : -1;
  }
  const __amdgpu_buffer_rsrc_t rA0 = rsrc16(P.hA), rdA = rsrcf(HA ? P.dA : nullptr);
  const __amdgpu_buffer_rsrc_t rB0 = rsrc16(P.hB), rdB = rsrcf(HB ? P.dB : nullptr);
  // staging registers, converted at the LDS store so that the loads stay in
  // flight across the current slab's matrix-core work: A0 (BF16 copy) and dB
  // feed A0 dB (needs HB), dA and B0 (BF16 copy) + dB feed dA (B0 + dB) (needs HA)
  unsigned short ra0[NE], rb0[NE];
  float rda[NE], rdb[NE];
  f4 acc[F][F];
#pragma unroll
  for (int i = 0; i < F; ++i)
#pragma unroll
    for (int j = 0; j < F; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  for (int c0 = kbeg; c0 < kend; c0 += KT) {
    const int cend = min(kend, c0 + KT);
    __syncthreads();
    for (int q = t; q < cend - c0; q += T) {
      koffA[q] = col_off(P.A, c0 + q);
      koffB[q] = row_off(P.B, c0 + q);
    }
    __syncthreads();
    // staging loads: padded elements read 0 through the out-of-range offset
    const int klast = cend - c0 - 1;
    auto load = [&](int k0) {
#pragma unroll
      for (int i = 0; i < NE; ++i) {
        int kl = k0 - c0 + kqA(i);
        int o = fa[i] >= 0 && kl <= klast ? (fa[i] + koffA[min(kl, klast)]) * 8 : kOobOffset;
        if (HB) ra0[i] = bload16(rA0, o);
        if (HA) rda[i] = bload32(rdA, o);
        kl = k0 - c0 + kqB(i);
        o = fb[i] >= 0 && kl <= klast ? (fb[i] + koffB[min(kl, klast)]) * 8 : kOobOffset;
        if (HB) rdb[i] = bload32(rdB, o);
        if (HA) rb0[i] = bload16(rB0, o);
      }
    };
    load(c0);
    for (int k0 = c0; k0 < cend; k0 += BK) {
#pragma unroll
      for (int i = 0; i < NE; ++i) {
        const int sa = mlA(i) * LD + kqA(i), sb = nlB(i) * LD + kqB(i);
        if (HB) {
          A0s[sa] = __builtin_bit_cast(__bf16, ra0[i]);
          dBs[sb] = (__bf16)rdb[i];
        }
        if (HA) {
          dAs[sa] = (__bf16)rda[i];
          const float b0 = __builtin_bit_cast(float, (unsigned)rb0[i] << 16);
          Bfs[sb] = (__bf16)(b0 + (HB ? rdb[i] : 0.f));
        }
      }
      __syncthreads();
      load(min(k0 + BK, cend));   // the last prefetch is fully masked
      const int kq = 8 * (lane >> 4);
      if (HB) {
        bf8 a[F], b[F];
#pragma unroll
        for (int f = 0; f < F; ++f) {
          a[f] = *reinterpret_cast<const bf8*>(&A0s[(wm + 16 * f + (lane & 15)) * LD + kq]);
          b[f] = *reinterpret_cast<const bf8*>(&dBs[(wn + 16 * f + (lane & 15)) * LD + kq]);
        }
#pragma unroll
        for (int i = 0; i < F; ++i)
#pragma unroll
          for (int j = 0; j < F; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
      }
      if (HA) {
        bf8 a[F], b[F];
#pragma unroll
        for (int f = 0; f < F; ++f) {
          a[f] = *reinterpret_cast<const bf8*>(&dAs[(wm + 16 * f + (lane & 15)) * LD + kq]);
          b[f] = *reinterpret_cast<const bf8*>(&Bfs[(wn + 16 * f + (lane & 15)) * LD + kq]);
        }
#pragma unroll
        for (int i = 0; i < F; ++i)
#pragma unroll
          for (int j = 0; j < F; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
      }
      __syncthreads();
    }
  }
  // epilogue: V_MFMA_F32_16X16X32_BF16 leaves D[row][col] of a 16x16 block in
  // lane col + 16 (row / 4), element row % 4
  if (P.ksplit > 1) {
    double* __restrict__ W = P.ws + (long)ks * P.M * P.N;
#pragma unroll
    for (int j = 0; j < F; ++j) {
      const int col = n0 + wn + 16 * j + (lane & 15);
      if (col >= P.N) continue;
#pragma unroll
      for (int i = 0; i < F; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
          if (row < P.M) W[(long)row * P.N + col] = (double)acc[i][j][r];
        }
    }
    return;
  }
  float* __restrict__ Cp = P.Cd;
#pragma unroll
  for (int j = 0; j < F; ++j) {
    const int col = n0 + wn + 16 * j + (lane & 15);
    if (col >= P.N) continue;
    const int co = col_off(P.C, col);
#pragma unroll
    for (int i = 0; i < F; ++i)
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

hipError_t launch_gemm_dual(const GemmProb* probs, const int* tp, int total_tiles, const int* active,
                            hipStream_t s, int flags, int bf16, int tile) {
  if (total_tiles <= 0) return hipSuccess;
  const int grid = ((total_tiles + 7) / 8) * 8;
  if (bf16) {
#define HCC_DUAL16(SC, AK, BN)                                                                          \
  switch ((flags >> 2) & 3) {                                                                           \
    case 1: hipLaunchKernelGGL((gemm_dual_bf16<SC, AK, BN, true, false>), dim3(grid), dim3(256), 0, s, probs, tp, total_tiles, active); break; \
    case 2: hipLaunchKernelGGL((gemm_dual_bf16<SC, AK, BN, false, true>), dim3(grid), dim3(256), 0, s, probs, tp, total_tiles, active); break; \
    case 3: hipLaunchKernelGGL((gemm_dual_bf16<SC, AK, BN, true, true>), dim3(grid), dim3(256), 0, s, probs, tp, total_tiles, active); break; \
    default: return hipErrorInvalidValue; /* both operands constant: not a correction product */     \
  }
#define HCC_DUAL16_S(SC)                        \
  switch (flags & 3) {                          \
    case 0: HCC_DUAL16(SC, false, false) break; \
    case 1: HCC_DUAL16(SC, true, false) break;  \
    case 2: HCC_DUAL16(SC, false, true) break;  \
    default: HCC_DUAL16(SC, true, true) break;  \
  }
    if (tile == 256) { HCC_DUAL16_S(2) }
    else { HCC_DUAL16_S(1) }
#undef HCC_DUAL16_S
#undef HCC_DUAL16
  } else {
    if (tile != 64) return hipErrorInvalidValue;
    switch (flags & 3) {
      case 0: hipLaunchKernelGGL((gemm_dual_f64<false, false>), dim3(grid), dim3(256), 0, s, probs, tp, total_tiles, active); break;
      case 1: hipLaunchKernelGGL((gemm_dual_f64<true, false>), dim3(grid), dim3(256), 0, s, probs, tp, total_tiles, active); break;
      case 2: hipLaunchKernelGGL((gemm_dual_f64<false, true>), dim3(grid), dim3(256), 0, s, probs, tp, total_tiles, active); break;
      default: hipLaunchKernelGGL((gemm_dual_f64<true, true>), dim3(grid), dim3(256), 0, s, probs, tp, total_tiles, active); break;
    }
  }
  return hipGetLastError();
}

// fixed-order sum of the split-K partials into C
__global__ __launch_bounds__(256) void splitk_reduce_kernel(const GemmProb* __restrict__ probs,
                                                            const int* __restrict__ prefix,
                                                            const int* __restrict__ block_prob,
                                                            const int* __restrict__ active) {
  const int b = blockIdx.x;
  const int lo = block_prob[b];
  const GemmProb& P = probs[lo];
  if (P.ksplit <= 1) return;
  if (active != nullptr && active[P.mol] == 0) return;
  const long e = (long)(b - prefix[lo]) * 256 + threadIdx.x;
  const long mn = (long)P.M * P.N;
  if (e >= mn) return;
  const int row = e / P.N, col = e % P.N;
  double s = 0.0;
  for (int k = 0; k < P.ksplit; ++k) s += P.ws[k * mn + e];
  const int o = row_off(P.C, row) + col_off(P.C, col);
  double v = P.alpha * s;
  if (P.Cd != nullptr) {   // FP32 correction output
    if (P.beta != 0.0) v += P.beta * P.Cd[o];
    P.Cd[o] = (float)v;
    return;
  }
  if (P.beta != 0.0) v += P.beta * P.C.p[o];
  P.C.p[o] = v;
}

hipError_t launch_splitk_reduce(const GemmProb* probs, const int* prefix, const int* block_prob, int nblocks,
                                const int* active, hipStream_t s) {
  if (nblocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(splitk_reduce_kernel, dim3(nblocks), dim3(256), 0, s, probs, prefix, block_prob, active);
  return hipGetLastError();
}

template <int BM, int BN, int BK>
static void launch_f64(const GemmProb* probs, const int* tp, int total_tiles, const int* active, hipStream_t s,
                       int grid, int flags) {
  constexpr int T = HCC_F64_THREADS(BM, BN);
  switch (flags & 3) {
    case 0: hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, BK, false, false>), dim3(grid), dim3(T), 0, s, probs, tp, total_tiles, active); break;
    case 1: hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, BK, true, false>), dim3(grid), dim3(T), 0, s, probs, tp, total_tiles, active); break;
    case 2: hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, BK, false, true>), dim3(grid), dim3(T), 0, s, probs, tp, total_tiles, active); break;
    default: hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, BK, true, true>), dim3(grid), dim3(T), 0, s, probs, tp, total_tiles, active); break;
  }
}

hipError_t launch_gemm_f64(const GemmProb* probs, const int* tile_prob, int total_tiles, const int* active,
                           hipStream_t s, int tile, int flags) {
  if (total_tiles <= 0) return hipSuccess;
  int grid = ((total_tiles + 7) / 8) * 8;
  if (tile == 32) launch_f64<32, 32, 16>(probs, tile_prob, total_tiles, active, s, grid, flags);
  else if (tile == 128) launch_f64<128, 32, 16>(probs, tile_prob, total_tiles, active, s, grid, flags);
  else launch_f64<64, 64, 16>(probs, tile_prob, total_tiles, active, s, grid, flags);
  return hipGetLastError();
}

}  // namespace hcc
