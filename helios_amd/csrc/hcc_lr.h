// Device-resident vector space of the LR-CCSD eigensolver: per molecule,
// `nslots` vectors of o v + o^2 v^2 doubles (singles [i][a], then doubles
// [i][j][a][b]) back to back; vector s of molecule m starts at
// space + base[m] + s * len[m].  The eigensolver's vector algebra runs here as
// batched jobs over all molecules; the host keeps only the projected matrices.
#pragma once
#include "hcc_batch.h"

namespace hcc {

constexpr int kLrThreads = 256;
constexpr int kLrChunk = 4 * kLrThreads;   // elements per workgroup of the elementwise jobs
constexpr int kLrDotChunk = 8192;          // elements per workgroup of the dot products

// out[j] = <a, b> (mode 0) or max_e sqrt(a_e^2 + b_e^2) (mode 1; b < 0: max |a_e|)
// for jobs j = {mol, slot a, slot b}; `prefix` (njobs + 1) is the chunk prefix
// of the jobs (ceil(len / kLrDotChunk) each), `partial` holds nchunks values;
// fixed summation order.
void lr_dots(const double* space, const long* base, const long* len, int njobs, const int* prefix, const int* jobs,
             int mode, double* partial, int nchunks, double* out, hipStream_t s);
// Block Gram products on V_MFMA_F64_16X16X4F64: block j gives
// out[j.out + a * j.ld + b] = <slot[j.a + a], slot[j.b + b]> for a < na <= 16,
// b < nb <= 64 (slot[] the device slot list).  One workgroup per kLrGramChunk
// elements of a block reads every vector of the block once; its partial sums
// (16 x 64, accumulator order) are added per block in chunk order.
struct LrGramBlk {
  int mol, na, nb, a, b, out, ld, pad;
};
constexpr int kLrGramChunk = 4096;
void lr_gram(const double* space, const long* base, const long* len, int nblk, const int* prefix,
             const LrGramBlk* blk, const int* slots, double* partial, int nchunks, double* out, hipStream_t s);
// Block combinations on V_MFMA_F64_16X16X4F64: for q < m <= 16,
// slot[d + q] = beta slot[d + q] + sum_{p < t} coef[c + q t + p] slot[tt + p]
// (slot[] the device slot list; no destination among the terms).  One
// workgroup per kLrCombChunk elements of a block.
struct LrCombBlk {
  int mol, m, t, d, tt, c, pad0, pad1;
  double beta;
};
constexpr int kLrCombChunk = 256;   // 4 waves x 64 elements
void lr_combine_block(const double* space, const long* base, const long* len, int nblk, const int* prefix,
                      const LrCombBlk* blk, const int* slots, const double* coef, int nblocks, hipStream_t s);
// dst = beta_j dst + sum_t coef[t] * slot[t] for jobs j = {mol, dst, first term, terms};
// `prefix` (njobs + 1) is the workgroup prefix of the jobs (ceil(len / kLrChunk) each).
void lr_combine(const double* space, const long* base, const long* len, int njobs, const int* prefix,
                const int4* jobs, const double* beta, const int* tslot, const double* tcoef, int nblocks,
                hipStream_t s);
// x_e /= (shift_j - D_e), |shift_j - D_e| floored at `floor`, D_e the orbital-energy
// difference of element e (e_a - e_i, or e_a + e_b - e_i - e_j), for jobs j = {mol, slot}.
void lr_precondition(const Mol* mols, const double* space, const long* base, const long* len, int njobs,
                     const int* prefix, const int2* jobs, const double* shift, double floor, int nblocks,
                     hipStream_t s);

}  // namespace hcc
