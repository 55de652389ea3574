// Batched closed-shell PPP RHF on the GPU (hcc_scf.hip).
#pragma once
#include "hcc_batch.h"

namespace hcc {

// Per-molecule SCF inputs: row-major n x n core Hamiltonian and Hueckel
// matrix (gamma is the molecule's Mol::G), and an n x n global work block.
struct ScfInput {
  const double* core;
  const double* huckel;
  double* work;
};

struct ScfResult {
  double eref;      // reference (SCF) energy
  double error;     // max|Fmin| of the last iteration
  int iterations;
  int status;       // 1 converged, -1 iteration cap, -2 NaN
};

// LDS bytes of the SCF kernel for molecules of up to nmax sites
size_t scf_lds_bytes(int nmax);
// one workgroup per molecule; writes Mol::C (row-major [mu][p]), Mol::F (MO
// Fock matrix), Mol::eref and the results
hipError_t launch_rhf(Mol* mols, const ScfInput* in, int nm, int nmax, double accuracy, int maxiter, double step,
                      ScfResult* out, hipStream_t s);

}  // namespace hcc
