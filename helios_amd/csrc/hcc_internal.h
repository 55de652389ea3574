// Internal declarations of libhelios_ccsd (not part of the C ABI).
//
// Layout conventions (all device arrays row-major, one molecule after another
// in per-buffer arenas):
//   C[mu][p]        MO coefficients, column p = MO p (occupied first)
//   t1[i][a]        o x v
//   t2[i][j][a][b]  o*o*v*v, pair-symmetric: t2[i][j][a][b] == t2[j][i][b][a]
//
// A GEMM operand is addressed through "composite" indices: a row (or column)
// index r is split as r = hi * radix + lo and the element offset is
// hi * s_hi + lo * s_lo.  With one composite index per dimension every
// permuted tensor contraction of the T1-dressed CCSD residual becomes a plain
// GEMM with no transposition pass.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace hcc {

// HIP failure inside the library; the C ABI turns it into HCC_E_HIP
struct HipError {
  hipError_t e;
  const char* what;
};

#define HCHECK(x)                                        \
  do {                                                   \
    hipError_t e_ = (x);                                 \
    if (e_ != hipSuccess) throw HipError{e_, #x};        \
  } while (0)

struct Operand {
  double* p;        // base of this molecule's block
  int rr, rc;       // radices of the row and column composite indices (1 .. kMaxRadix)
  int srh, srl;     // row offset = (r / rr) * srh + (r % rr) * srl
  int sch, scl;     // col offset = (c / rc) * sch + (c % rc) * scl
  double ir, ic;    // 1 / rr and 1 / rc (FP64, round to nearest)
};
// Composite indices stay below kMaxIndex and radices below kMaxRadix, so the
// kernels' quotient floor((r + 1/2) * (1 / radix)) is exact: (r + 1/2) / radix
// lies at least 2^-17 from an integer, the FP64 rounding error is below 2^-25.
constexpr int kMaxRadix = 1 << 16;
constexpr int kMaxIndex = 1 << 27;
inline Operand make_operand(double* p, int rr, int rc, int srh, int srl, int sch, int scl) {
  return Operand{p, rr, rc, srh, srl, sch, scl, 1.0 / rr, 1.0 / rc};
}

struct GemmProb {
  Operand A, B, C;  // A: M x K, B: K x N, C: M x N
  int M, N, K;
  double alpha, beta;
  int tiles_n;      // tiles along N for this problem (for the table's tile size)
  int tile_begin;   // first global tile of this problem
  int mol;          // molecule index (for the active mask)
  int a_kfast;      // 1: A's unit stride runs along K, 0: along M
  int b_nfast;      // 1: B's unit stride runs along N, 0: along K
  int ksplit;       // number of K chunks (1: no split)
  int kchunk;       // K range of one chunk (multiple of the K tile)
  double* ws;       // split-K partials: ksplit x M x N, row-major (ksplit > 1)
  // correction ("delta") form, used by the mixed-precision inner iterations:
  // Cd += alpha * (A0 dB + dA (B0 + dB)) with A0 = A.p, B0 = B.p and the FP32
  // corrections dA, dB, Cd (Cd at C's offsets); a null dA or dB marks an
  // operand that does not change with the amplitudes
  const float* dA;
  const float* dB;
  float* Cd;
  // BF16 copies of A0 and B0 (same element offsets), refreshed at every outer
  // point; the BF16 correction kernel reads A0 and B0 from them
  const __bf16* hA;
  const __bf16* hB;
};

// Tile shape of the FP64 MFMA GEMM.
constexpr int kBM = 64, kBN = 64, kBK = 16, kThreads = 256;

// launches one grouped GEMM over all problems of a table (device memory)
// tile: 64 (64x64 tiles, 256 threads), 32 (32x32 tiles, 64 threads) or 128
// (128x32 tiles, 256 threads, for outputs at most 32 columns wide);
// flags: bit 0 = A's unit stride along K, bit 1 = B's unit stride along N
// (the same for every problem of a table)
// tile_prob: problem index of every global tile of the table
hipError_t launch_gemm_f64(const GemmProb* probs, const int* tile_prob, int total_tiles,
                           const int* active, hipStream_t s, int tile = 64, int flags = 3);
// correction-form GEMM (see GemmProb::dA/dB); bf16 = 1 rounds the operands to
// BF16 for V_MFMA_F32_16X16X32_BF16 (FP32 accumulation; tile kinds 32, 64 and
// 256 = 128x128 output tiles), 0 keeps FP64 (tile kind 64 only)
hipError_t launch_gemm_dual(const GemmProb* probs, const int* tile_prob, int total_tiles, const int* active,
                            hipStream_t s, int flags, int bf16, int tile);
// C = alpha * sum_k ws_k + beta * C for the split problems of a table;
// prefix: block prefix over the table's problems (256 elements per block),
// block_prob: problem index of every reduction block
hipError_t launch_splitk_reduce(const GemmProb* probs, const int* prefix, const int* block_prob, int nblocks,
                                const int* active, hipStream_t s);

// out[x] = i with prefix[i] <= x < prefix[i + 1] for x < total (prefix: n + 1
// ascending entries; zero-length ranges are never chosen for a valid x)
hipError_t launch_expand_prefix(const int* prefix, int n, int* out, int total, hipStream_t s);
// Compacts a table to its live problems: dst[i] = src[idx[i]] with
// tile_begin = tbeg[i], tile_prob over the renumbered tiles (tbeg: n + 1
// entries, tiles = tbeg[n]) and red_bprob over the split-K reduction blocks
// (rpre: n + 1 entries, red = rpre[n])
hipError_t launch_compact_table(const GemmProb* src, const int* idx, const int* tbeg, const int* rpre, int n,
                                int tiles, int red, GemmProb* dst, int* tile_prob, int* red_bprob, hipStream_t s);

}  // namespace hcc
