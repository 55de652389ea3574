// Learned T2 guess (hcc_mlp.hip): per-amplitude features, the 3-layer MLP on
// the matrix cores, and its training kernels.
#pragma once
#include "hcc_batch.h"

struct hcc_ctx;

namespace hcc {

constexpr int kMlpF = 16;        // features per amplitude
constexpr int kMlpH = 64;        // hidden width
constexpr int kMlpRow = 18;      // training row: 16 features, target, weight
// flat parameter layout (FP32 master copy): W1[64][16], b1, W2[64][64], b2, W3[64], b3
constexpr int kOffW1 = 0, kOffB1 = 1024, kOffW2 = 1088, kOffB2 = 5184, kOffW3 = 5248, kOffB3 = 5312;
constexpr int kMlpParams = 5313;

// Ov[p][q] = n * sum_mu C[mu][p]^2 C[mu][q]^2 for every molecule
void mlp_overlaps(const Mol* mols, Ragged rgNN, int nm, hipStream_t s);
// t2 = t_mp2 + y / Dp for every amplitude, one evaluation per amplitude pair
// (rgPairL: 1024 pairs (i a) <= (j b) per workgroup)
// occupied block pairs (i <= j) per workgroup of the learned guess: about 1024
// amplitudes per workgroup
__host__ __device__ inline int guess_pairs_per_workgroup(int v) { return v * v >= 1024 ? 1 : 1024 / (v * v); }

// learned guess: rgPairBlk has ceil(o(o+1)/2 / guess_pairs_per_workgroup(v))
// workgroups per molecule (molecules with v <= 32); rgPairL covers the
// amplitude pairs of the larger ones (any_large: some molecule has v > 32)
void mlp_guess(const Mol* mols, Ragged rgPairBlk, Ragged rgPairL, int nm, bool any_large, const float* params,
               hipStream_t s);

// device view of a loaded batch, for the trainer (hcc_solver.hip)
struct BatchView {
  const Mol* dm;
  const State* dst;
  int nm;
  int device;
  hipStream_t stream;
  Ragged rgNN;
};
bool batch_view(hcc_ctx* c, BatchView* v);

}  // namespace hcc
