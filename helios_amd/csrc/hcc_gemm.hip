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
// r / radix for 0 <= r < kMaxIndex, 1 <= radix < kMaxRadix (exact, see Operand)
__device__ __forceinline__ int quot(int r, double inv) { return (int)(((double)r + 0.5) * inv); }
__device__ __forceinline__ int row_off(const Operand& o, int r) {
  const int q = quot(r, o.ir);
  return q * o.srh + (r - q * o.rr) * o.srl;
}
__device__ __forceinline__ int col_off(const Operand& o, int c) {
  const int q = quot(c, o.ic);
  return q * o.sch + (c - q * o.rc) * o.scl;
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
//
// 64x64 output tiles, 4 waves of 32x32, 32-deep K slabs.  Staging is the
// bound of this kernel (the matrix cores need 8 instructions per slab and
// wave), so it is kept to a few instructions per element:
//  * offsets are formed as row + K-table entry with no compare or select:
//    a padding row or K carries the sentinel kPad, and any sum that includes
//    it lands past the end of the 16- and 32-bit resources (which read 0);
//  * K-contiguous operands are staged one element per lane with the lanes
//    along K (32 lanes per row), so every load instruction is coalesced;
//  * row-contiguous operands are staged with the lanes along the rows and two
//    adjacent K per thread, so one V_CVT_PK_BF16_F32 and one 32-bit LDS store
//    serve two elements.
// Both go to LDS as [row][k] with a 40-element pitch and are read back as 8
// consecutive K per lane.
constexpr int kPad = 1 << 28;   // 2 x (element index) of a padding row or K
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc_h(const void* p, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)p, 0, bytes, 0x00020000);
}
// staged element j (0 .. 7) of thread t (of T) in an R-row x BK slab
// (R * BK = 8 T): K-contiguous operands with the lanes along K,
// row-contiguous ones with the lanes along the rows and K in adjacent pairs
template <bool KC, int BK, int R, int T>
struct Stage {
  static __device__ __forceinline__ int row(int t, int j) { return KC ? t / BK + (T / BK) * j : t % R; }
  static __device__ __forceinline__ int k(int t, int j) {
    return KC ? t % BK : 2 * (t / R) + 2 * (T / R) * (j >> 1) + (j & 1);
  }
};
constexpr int kDualBK = 32;   // 64-deep slabs measured slower (K padding of the short contractions, fewer waves)
// TS: output tile edge, 32 (2 waves of 32x16), 64 (4 waves of 32x32) or 128
// (8 waves of 32x64).  SK (short K, at most two slabs): the next slab is loaded
// after the current one's matrix-core work instead of during it; the staging
// registers are then dead during that work, which lets six waves share a SIMD
// (a one- or two-slab tile is mostly fixed latency, hidden by more tiles per CU)
template <bool AKF, bool BNF, bool HA, bool HB, int TS, bool SK>
__global__ __launch_bounds__(TS * 4) __attribute__((amdgpu_waves_per_eu(SK ? 6 : 4, 8))) void gemm_dual_bf16(
    const GemmProb* __restrict__ probs, const int* __restrict__ tile_prob, int total_tiles,
    const int* __restrict__ active) {
  constexpr int BK = kDualBK, BM = TS, BN = TS, T = TS * 4, LD = BK + 8, NE = 8;
  constexpr int FM = 2, FN = TS / 32;   // 16x16 blocks per wave along M and N
  constexpr int KT = 512;               // K-table chunk
  constexpr bool KA = AKF, KB = !BNF;   // operand K-contiguous in memory
  static_assert(BM * BK == NE * T, "staging covers the slab");
  using SA = Stage<KA, BK, BM, T>;
  using SB = Stage<KB, BK, BN, T>;
  typedef float f2 __attribute__((ext_vector_type(2)));
  typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
  __shared__ __attribute__((aligned(16))) __bf16 A0s[BM * LD];
  __shared__ __attribute__((aligned(16))) __bf16 dAs[BM * LD];
  __shared__ __attribute__((aligned(16))) __bf16 dBs[BN * LD];
  __shared__ __attribute__((aligned(16))) __bf16 Bfs[BN * LD];
  __shared__ int koffA[KT + BK];
  __shared__ int koffB[KT + BK];
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
  // waves: 2 along N (32 FN columns each), T / 128 along M (32 rows each)
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * (16 * FN);
  // 2 x row offsets (kPad for rows past the problem)
  int fa[NE], fb[NE];
#pragma unroll
  for (int j = 0; j < NE; ++j) {
    fa[j] = m0 + SA::row(t, j) < P.M ? 2 * row_off(P.A, m0 + SA::row(t, j)) : kPad;
    fb[j] = n0 + SB::row(t, j) < P.N ? 2 * col_off(P.B, n0 + SB::row(t, j)) : kPad;
  }
  // BF16 resources: byte offset 2e, range 2^28 bytes; FP32: 4e, 2^29 bytes
  // (valid element indices stay below 2^27, checked by the host)
  const __amdgpu_buffer_rsrc_t rA0 = rsrc_h(P.hA, kPad), rdA = rsrc_h(HA ? P.dA : nullptr, 2 * kPad);
  const __amdgpu_buffer_rsrc_t rB0 = rsrc_h(P.hB, kPad), rdB = rsrc_h(HB ? P.dB : nullptr, 2 * kPad);
  // staging registers, converted at the LDS store so that the loads stay in
  // flight across the current slab's matrix-core work: A0 (BF16 copy) and dB
  // feed A0 dB (needs HB), dA and B0 (BF16 copy) + dB feed dA (B0 + dB) (needs HA)
  unsigned short ra0[NE], rb0[NE];
  float rda[NE], rdb[NE];
  f4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  for (int c0 = kbeg; c0 < kend; c0 += KT) {
    const int clen = min(kend - c0, KT);
    __syncthreads();
    // 2 x K offsets of the chunk; entries past its end (read by the last,
    // partial slab and the final prefetch) carry kPad
    for (int q = t; q < KT + BK; q += T) {
      koffA[q] = q < clen ? 2 * col_off(P.A, c0 + q) : kPad;
      koffB[q] = q < clen ? 2 * row_off(P.B, c0 + q) : kPad;
    }
    __syncthreads();
    auto load = [&](int kl0) {
#pragma unroll
      for (int j = 0; j < NE; ++j) {
        const int ea = fa[j] + koffA[kl0 + SA::k(t, j)];
        if (HB) ra0[j] = __builtin_amdgcn_raw_buffer_load_b16(rA0, ea, 0, 0);
        if (HA) rda[j] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rdA, 2 * ea, 0, 0));
        const int eb = fb[j] + koffB[kl0 + SB::k(t, j)];
        if (HB) rdb[j] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rdB, 2 * eb, 0, 0));
        if (HA) rb0[j] = __builtin_amdgcn_raw_buffer_load_b16(rB0, eb, 0, 0);
      }
    };
    // B0 + dB as FP32 (B0 is a BF16 bit pattern)
    auto bsum = [&](int j) {
      const float b0 = __builtin_bit_cast(float, (unsigned)rb0[j] << 16);
      return HB ? b0 + rdb[j] : b0;
    };
    load(0);
    for (int kl = 0; kl < clen; kl += BK) {
#pragma unroll
      for (int j = 0; j < NE; ++j) {
        // A side
        if (KA) {
          const int sa = SA::row(t, j) * LD + SA::k(t, j);
          if (HB) A0s[sa] = __builtin_bit_cast(__bf16, ra0[j]);
          if (HA) dAs[sa] = (__bf16)rda[j];
        } else if ((j & 1) == 0) {
          const int sa = (SA::row(t, j) * LD + SA::k(t, j)) >> 1;
          if (HB) reinterpret_cast<unsigned*>(A0s)[sa] = (unsigned)ra0[j] | ((unsigned)ra0[j + 1] << 16);
          if (HA) reinterpret_cast<bf2*>(dAs)[sa] = __builtin_convertvector((f2{rda[j], rda[j + 1]}), bf2);
        }
        // B side
        if (KB) {
          const int sb = SB::row(t, j) * LD + SB::k(t, j);
          if (HB) dBs[sb] = (__bf16)rdb[j];
          if (HA) Bfs[sb] = (__bf16)bsum(j);
        } else if ((j & 1) == 0) {
          const int sb = (SB::row(t, j) * LD + SB::k(t, j)) >> 1;
          if (HB) reinterpret_cast<bf2*>(dBs)[sb] = __builtin_convertvector((f2{rdb[j], rdb[j + 1]}), bf2);
          if (HA) reinterpret_cast<bf2*>(Bfs)[sb] = __builtin_convertvector((f2{bsum(j), bsum(j + 1)}), bf2);
        }
      }
      __syncthreads();
      if (!SK) load(min(kl + BK, clen));   // the last prefetch reads only padding
#pragma unroll
      for (int kk = 0; kk < BK; kk += 32) {
        const int kq = kk + 8 * (lane >> 4);
        // product with A from Xa, B from Xb (LDS [row][k])
        auto product = [&](const __bf16* Xa, const __bf16* Xb) {
          bf8 a[FM], b[FN];
#pragma unroll
          for (int f = 0; f < FM; ++f) a[f] = *reinterpret_cast<const bf8*>(&Xa[(wm + 16 * f + (lane & 15)) * LD + kq]);
#pragma unroll
          for (int f = 0; f < FN; ++f) b[f] = *reinterpret_cast<const bf8*>(&Xb[(wn + 16 * f + (lane & 15)) * LD + kq]);
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        };
        if (HB) product(A0s, dBs);
        if (HA) product(dAs, Bfs);
      }
      __syncthreads();
      if (SK && kl + BK < clen) load(kl + BK);
    }
  }
  // epilogue: V_MFMA_F32_16X16X32_BF16 leaves D[row][col] of a 16x16 block in
  // lane col + 16 (row / 4), element row % 4
  if (P.ksplit > 1) {
    double* __restrict__ W = P.ws + (long)ks * P.M * P.N;
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int col = n0 + wn + 16 * j + (lane & 15);
      if (col >= P.N) continue;
#pragma unroll
      for (int i = 0; i < FM; ++i)
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
  for (int j = 0; j < FN; ++j) {
    const int col = n0 + wn + 16 * j + (lane & 15);
    if (col >= P.N) continue;
    const int co = col_off(P.C, col);
#pragma unroll
    for (int i = 0; i < FM; ++i)
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
#define HCC_DUAL16(AK, BN, TS, SK)                                                                      \
  switch ((flags >> 2) & 3) {                                                                           \
    case 1: hipLaunchKernelGGL((gemm_dual_bf16<AK, BN, true, false, TS, SK>), dim3(grid), dim3(4 * TS), 0, s, probs, tp, total_tiles, active); break; \
    case 2: hipLaunchKernelGGL((gemm_dual_bf16<AK, BN, false, true, TS, SK>), dim3(grid), dim3(4 * TS), 0, s, probs, tp, total_tiles, active); break; \
    case 3: hipLaunchKernelGGL((gemm_dual_bf16<AK, BN, true, true, TS, SK>), dim3(grid), dim3(4 * TS), 0, s, probs, tp, total_tiles, active); break; \
    default: return hipErrorInvalidValue; /* both operands constant: not a correction product */     \
  }
#define HCC_DUAL16_T(TS, SK)                           \
    switch (flags & 3) {                               \
      case 0: HCC_DUAL16(false, false, TS, SK) break;  \
      case 1: HCC_DUAL16(true, false, TS, SK) break;   \
      case 2: HCC_DUAL16(false, true, TS, SK) break;   \
      default: HCC_DUAL16(true, true, TS, SK) break;   \
    }
    // flag 16: every problem of the table has K <= 2 slabs (short-K variant)
    const bool sk = (flags & 16) != 0;
    if (tile == 64 && sk) { HCC_DUAL16_T(64, true) }
    else if (tile == 64) { HCC_DUAL16_T(64, false) }
    else if (tile == 32 && sk) { HCC_DUAL16_T(32, true) }
    else if (tile == 32) { HCC_DUAL16_T(32, false) }
    else if (tile == 256) { HCC_DUAL16_T(128, false) }
    else return hipErrorInvalidValue;
#undef HCC_DUAL16_T
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

// Compacted table: dst[i] = src[idx[i]] with tile_begin = tbeg[i] (the live
// problems of a table, their tiles renumbered consecutively)
__global__ __launch_bounds__(256) void k_compact_probs(const GemmProb* __restrict__ src, const int* __restrict__ idx,
                                                       const int* __restrict__ tbeg, int n, GemmProb* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  GemmProb p = src[idx[i]];
  p.tile_begin = tbeg[i];
  dst[i] = p;
}

// out[x] = i with prefix[i] <= x < prefix[i + 1] (prefix: n + 1 ascending entries)
__global__ __launch_bounds__(256) void k_expand_prefix(const int* __restrict__ prefix, int n, int* __restrict__ out,
                                                       int total) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= total) return;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= x) lo = mid; else hi = mid - 1;
  }
  out[x] = lo;
}

hipError_t launch_expand_prefix(const int* prefix, int n, int* out, int total, hipStream_t s) {
  if (total > 0) hipLaunchKernelGGL(k_expand_prefix, dim3((total + 255) / 256), dim3(256), 0, s, prefix, n, out, total);
  return hipGetLastError();
}

hipError_t launch_compact_table(const GemmProb* src, const int* idx, const int* tbeg, const int* rpre, int n,
                                int tiles, int red, GemmProb* dst, int* tile_prob, int* red_bprob, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_compact_probs, dim3((n + 255) / 256), dim3(256), 0, s, src, idx, tbeg, n, dst);
  if (tiles > 0) hipLaunchKernelGGL(k_expand_prefix, dim3((tiles + 255) / 256), dim3(256), 0, s, tbeg, n, tile_prob, tiles);
  if (red > 0) hipLaunchKernelGGL(k_expand_prefix, dim3((red + 255) / 256), dim3(256), 0, s, rpre, n, red_bprob, red);
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
