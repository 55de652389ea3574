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

// out[j] = <a, b> (mode 0) or max_e sqrt(a_e^2 + b_e^2) (mode 1; b < 0: max |a_e|)
// for jobs j = {mol, slot a, slot b}; one workgroup per job, fixed summation order.
void lr_dots(const double* space, const long* base, const long* len, int njobs, const int* jobs, int mode,
             double* out, hipStream_t s);
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
