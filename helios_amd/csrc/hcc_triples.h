// (T) correction kernel (hcc_triples.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>

namespace hcc {

// Per-molecule operands of the (T) kernel (row-major device arrays).
struct TripMol {
  int n, o, v, pad;
  const double* A;    // (xp|yd) at [p][x][y][d]  (o v^3)
  const double* B;    // (ql|rz) at [q][l][r][z]  (o^3 v)
  const double* L;    // (ia|jb) at [i][j][a][b]  (o^2 v^2)
  const double* t1;   // [i][a]
  const double* t2;   // [i][j][a][b]
  const double* F;    // MO Fock matrix (n x n), orbital energies on its diagonal
};

constexpr int kTripMaxOV = 32;   // o, v <= 32
constexpr int kTripLdsV = 24;    // v <= 24: matrix-core kernel with W in LDS (else W in global memory)

size_t triples_lds_bytes();
size_t triples_mfma_lds_bytes();
// prefix: nm + 1 item offsets; items: per work item (molecule, occupied
// triple i >= j >= k packed as i | j << 8 | k << 16), a molecule's items
// consecutive from prefix[m]; wscratch:
// grid blocks of wstride >= v^3 doubles (v > kTripLdsV); grid_mfma: workgroups
// of the matrix-core kernel (0: no molecule with v <= kTripLdsV); any_global:
// some molecule has v > kTripLdsV; part: 2 per item; out: 2 per molecule
// (E[T], E[ST])
hipError_t launch_triples(const TripMol* mols, int nm, const long* prefix, long total, const int2* items,
                          double* wscratch, long wstride, int grid, int grid_mfma, bool any_global, double* part,
                          double* out, hipStream_t s);

}  // namespace hcc
