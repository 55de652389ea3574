// Launch wrappers of the bandwidth kernels (hcc_ops.hip).
#pragma once
#include "hcc_batch.h"

namespace hcc {

constexpr int kEwThreads = 256;
constexpr int kDotItems = 8;   // elements per thread of a DIIS dot-product workgroup
constexpr int kPItems = 4;     // elements per thread of the P / PT builders (table rgN3)
// elements per thread of k_u, k_assemble, k_update2, k_update2_mixed (table rgT2x, whose
// block prefix also indexes their max|d| partials) and k_zero_delta, k_commit_delta (rgT2Sx)
constexpr int kEwItems = 4;
constexpr int kMaxDiis = 32;

void ew_dress(const Mol* mols, const State* st, Ragged rgNN, int nm, hipStream_t s);
void ew_fock_m(const Mol* mols, const State* st, Ragged rgN, Ragged rgNN, int nm, hipStream_t s);
void ew_pouter(const Mol* mols, const State* st, Ragged rgN3, int nm, hipStream_t s);
void ew_u(const Mol* mols, const State* st, Ragged rgT2, int nm, hipStream_t s);
void ew_u(const MolD* mols, const State* st, Ragged rgT2, int nm, hipStream_t s);
void ew_init_xx(const Mol* mols, const State* st, Ragged rgXX, int nm, hipStream_t s);
void ew_init_xx(const MolD* mols, const State* st, Ragged rgXX, int nm, hipStream_t s);
void ew_assemble(const Mol* mols, const State* st, Ragged rgT2, int nm, hipStream_t s);
// x += t2 Xvv^T - Xoo^T t2 (the E2 terms, both index placements) for molecules
// with o, v <= kE2Max (the others take the GEMM path); rgOcc: kE2Split blocks
// per occupied orbital
constexpr int kE2Max = 32;
constexpr int kE2Split = 4;
constexpr int kE2Slots = (kE2Max * kE2Max + kEwThreads - 1) / kEwThreads;   // staged elements per thread
void ew_e2(const Mol* mols, const State* st, Ragged rgOcc, int nm, hipStream_t s);
void mixed_e2(const Mol* m0, const MolD* md, const State* st, Ragged rgOcc, int nm, hipStream_t s);
void ew_assemble(const MolD* mols, const State* st, Ragged rgT2, int nm, hipStream_t s);
void ew_update(const Mol* mols, const State* st, Ragged rgOV, Ragged rgT2, int nm, const Params& prm,
               double* part1, double* part2, hipStream_t s);
void ew_contravariant(const Mol* mols, Ragged rgT2S, int nm, double* d1, double* d2, hipStream_t s);
void ctl_after_update(const Mol* mols, State* st, int nm, const Params& prm, Ragged rgOV,
                      const double* part1, Ragged rgT2, const double* part2, hipStream_t s);
void diis_step(const Mol* mols, State* st, int nm, Ragged rgND, Ragged rgP2, double* part,
               hipStream_t s);
void diis_step(const MolD* mols, State* st, int nm, Ragged rgND, Ragged rgP2, double* part,
               hipStream_t s);
void ctl_after_diis(const Mol* mols, State* st, int* act, int nm, const Params& prm, const int* stop,
                    hipStream_t s);
void energy(const Mol* mols, State* st, int nm, Ragged rgT2S, double* part, hipStream_t s);
void make_L(const Mol* mols, Ragged rgT2, int nm, hipStream_t s);
void guess(const Mol* mols, State* st, int* act, Ragged rgT2S, int nm, int kind, hipStream_t s);
void reset_state(State* st, int* act, int nm, hipStream_t s);
void zero_t1(const Mol* mols, Ragged rgOV, int nm, hipStream_t s);
void to_bf16(const double* src, __bf16* dst, long n, hipStream_t s);
// the same for the blocks [off[k], off[k] + len[k]) of the molecules with a
// non-empty range in the block prefix `pre` (nm + 1 entries, 4 * 256 elements
// per block; blocks = pre[nm])
void to_bf16_runs(const double* src, __bf16* dst, const long* off, const long* len, const int* pre, int nm, int blocks,
                  hipStream_t s);
void activate(State* st, int* act, int nm, hipStream_t s);
void stop_running(State* st, int* act, int nm, hipStream_t s);
// restore = false: saves the active flags and activates every molecule, or
// with `sel` (device, nm pointers) those whose pointer is set; true: restores
void arm_all(State* st, int* act, int* save, int nm, bool restore, hipStream_t s, const double* const* sel = nullptr);
void lr_perturb(const Mol* mols, Ragged rgT2, int nm, const double* t1base, const double* t1b, const double* t2base,
                const double* t2b, const double* const* r1p, const double* const* r2p, double h, hipStream_t s);
void lr_omega(const Mol* mols, Ragged rgT2, int nm, double* const* o1p, double* const* o2p, bool minus, double scale,
              hipStream_t s);

// mixed-precision solver (correction form)
void mixed_dress(const Mol* m0, const MolD* md, const State* st, Ragged rgNN, Ragged rgN, Ragged rgN3, int nm,
                 hipStream_t s);
void mixed_zero(const Mol* m0, const MolD* md, const State* st, Ragged rgT2S, int nm, hipStream_t s);
// t0 += dt (keep_diis: the stored inner DIIS corrections are shifted by -dt first)
void mixed_commit(const Mol* m0, const MolD* md, const State* st, Ragged rgT2S, Ragged rgP2, int nm, bool keep_diis,
                  hipStream_t s);
void mixed_update(const Mol* m0, const MolD* md, const State* st, Ragged rgOV, Ragged rgT2, int nm, const Params& prm,
                  double* part1, double* part2, hipStream_t s);
void mixed_outer_control(const Mol* mols, State* st, int* act, int nm, const Params& prm, int max_outer, Ragged rgOV,
                         const double* part1, Ragged rgT2, const double* part2, hipStream_t s);
void mixed_inner_control(const MolD* mols, State* st, int* act, int nm, const Params& prm, Ragged rgOV,
                         const double* part1, Ragged rgT2, const double* part2, hipStream_t s);
void mixed_inner_advance(const MolD* mols, State* st, int nm, hipStream_t s);
void set_phase(State* st, int* act, int nm, int phase, hipStream_t s);

}  // namespace hcc
