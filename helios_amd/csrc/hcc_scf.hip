// Batched closed-shell PPP RHF on the GPU: the step before the CCSD path.
//
// One workgroup per molecule runs the reference's SCF (src/methods/mod_scf.f90):
// Hueckel guess (:351-386: eigenvectors of the Hueckel matrix, D = C_occ C_occ^T),
// then damped-density iterations (:106-136, :446-494)
//   R = 2D - 1,  Fmin = (F R - R F) R,  error = max|Fmin|,
//   D <- D - step Fmin,  F = F(D),  F = C E C^T,  D = C_occ C_occ^T,  F = F(D)
// with the Fock matrix F(D) = core - gamma o D + 2 diag(gamma diag D)
// (:417-442), until error < accuracy.  Then the CC inputs as initCC prepares
// them (src/methods/coupledCluster/mod_cc.f90:2311-2504): MO coefficients C
// (columns ascending in energy, each with its largest-|C| entry positive), the
// MO Fock matrix C^T F C and the reference energy sum D o (core + F) +
// sum_{mu != nu} gamma / 2.
//
// The symmetric eigenproblems are solved by cyclic Jacobi with the
// round-robin pair ordering, all n/2 disjoint rotations of a step applied
// together; from the second SCF iteration on, the Jacobi starts from the
// previous eigenvectors (A = V^T F V is then nearly diagonal and a sweep or
// two suffice).  Three n x n FP64 matrices live in LDS (the Fock/work matrix,
// the eigenvectors, the density), a fourth work matrix in global memory.
#include <hip/hip_runtime.h>

#include "hcc_batch.h"
#include "hcc_scf.h"

namespace hcc {

constexpr int kScfThreads = 256;

struct ScfShared {
  double* A;   // Fock / work matrix (n x n, row-major, pitch n)
  double* V;   // eigenvectors (columns)
  double* D;   // density
  double* T;   // work (global memory, this molecule's block)
  double* ev;  // eigenvalues (n)
  double* red; // reduction scratch (kScfThreads / 64 values)
  double* cs;  // rotation cosines / sines (2 * 33)
  int* pq;     // rotation pairs (2 * 33)
  int* perm;   // eigenvalue order (n)
};

__device__ double block_max_abs(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = red[0];
  for (int k = 1; k < kScfThreads / 64; ++k) r = fmax(r, red[k]);
  __syncthreads();
  return r;
}

__device__ double block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = red[0];
  for (int k = 1; k < kScfThreads / 64; ++k) r += red[k];
  __syncthreads();
  return r;
}

// F(D) = core - gamma o D + 2 diag(gamma diag D) into out (global core, gamma)
__device__ void fock(int n, const double* __restrict__ core, const double* __restrict__ gam, const double* D,
                     double* out) {
  for (int e = threadIdx.x; e < n * n; e += kScfThreads) {
    const int i = e / n, j = e % n;
    double f = core[e] - gam[e] * D[e];
    if (i == j) {
      double s = 0.0;
      for (int l = 0; l < n; ++l) s += gam[i * n + l] * D[l * n + l];
      f += 2.0 * s;
    }
    out[e] = f;
  }
  __syncthreads();
}

// out = X Y (n x n, row-major); tx / ty transpose X / Y
__device__ void matmul(int n, const double* X, bool tx, const double* Y, bool ty, double* out) {
  for (int e = threadIdx.x; e < n * n; e += kScfThreads) {
    const int i = e / n, j = e % n;
    double s = 0.0;
    for (int k = 0; k < n; ++k) s += (tx ? X[k * n + i] : X[i * n + k]) * (ty ? Y[j * n + k] : Y[k * n + j]);
    out[e] = s;
  }
  __syncthreads();
}

// Jacobi diagonalisation of the symmetric A (destroyed: eigenvalues on its
// diagonal), accumulating the rotations into V (V <- V J).  Round-robin
// ordering over m = n/2 (n odd: one index idles each step).
__device__ void jacobi(int n, double* A, double* V, ScfShared& sh) {
  const int np = n + (n & 1);          // players, the last a dummy when n is odd
  const int m = np / 2;
  for (int sweep = 0; sweep < 40; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int e = threadIdx.x; e < n * n; e += kScfThreads) {
      const int i = e / n, j = e % n;
      if (i != j) off = fmax(off, fabs(A[e]));
      else dia = fmax(dia, fabs(A[e]));
    }
    off = block_max_abs(off, sh.red);
    dia = block_max_abs(dia, sh.red);
    if (!(off > 1e-15 * (dia + 1e-300))) break;
    for (int r = 0; r < np - 1; ++r) {
      // rotation angles of this step's pairs
      if ((int)threadIdx.x < m) {
        const int k = threadIdx.x;
        int p, q;
        if (k == 0) {
          p = 0;
          q = 1 + r;
        } else {
          p = 1 + (r + k) % (np - 1);
          q = 1 + (r - k + np - 1) % (np - 1);
        }
        if (p > q) { const int t = p; p = q; q = t; }
        double c = 1.0, s = 0.0;
        if (q < n) {
          const double apq = A[p * n + q];
          if (fabs(apq) > 1e-300) {
            const double tau = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
            const double t = tau >= 0.0 ? 1.0 / (tau + sqrt(1.0 + tau * tau)) : -1.0 / (-tau + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
          }
        }
        sh.pq[2 * k] = p;
        sh.pq[2 * k + 1] = q < n ? q : p;   // a dummy pair rotates nothing
        sh.cs[2 * k] = c;
        sh.cs[2 * k + 1] = q < n ? s : 0.0;
      }
      __syncthreads();
      // rows p, q of A <- J^T A
      for (int t = threadIdx.x; t < m * n; t += kScfThreads) {
        const int k = t / n, col = t % n;
        const int p = sh.pq[2 * k], q = sh.pq[2 * k + 1];
        if (p == q) continue;
        const double c = sh.cs[2 * k], s = sh.cs[2 * k + 1];
        const double ap = A[p * n + col], aq = A[q * n + col];
        A[p * n + col] = c * ap - s * aq;
        A[q * n + col] = s * ap + c * aq;
      }
      __syncthreads();
      // columns p, q of A and V <- A J, V J
      for (int t = threadIdx.x; t < m * n; t += kScfThreads) {
        const int k = t / n, row = t % n;
        const int p = sh.pq[2 * k], q = sh.pq[2 * k + 1];
        if (p == q) continue;
        const double c = sh.cs[2 * k], s = sh.cs[2 * k + 1];
        const double ap = A[row * n + p], aq = A[row * n + q];
        A[row * n + p] = c * ap - s * aq;
        A[row * n + q] = s * ap + c * aq;
        const double vp = V[row * n + p], vq = V[row * n + q];
        V[row * n + p] = c * vp - s * vq;
        V[row * n + q] = s * vp + c * vq;
      }
      __syncthreads();
    }
  }
}

// Reorders the eigenpairs ascending (ties by index): V's columns into T, then
// back, and the eigenvalues into evals (diag(A) sorted)
__device__ void sort_eigen(int n, double* A, double* V, double* T, double* evals, ScfShared& sh) {
  // identity first, so that ranks that collide (NaN eigenvalues) leave valid indices
  for (int i = threadIdx.x; i < n; i += kScfThreads) sh.perm[i] = i;
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += kScfThreads) {
    const double li = A[i * n + i];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double lj = A[j * n + j];
      rank += (lj < li || (lj == li && j < i)) ? 1 : 0;
    }
    sh.perm[rank] = i;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n * n; e += kScfThreads) {
    const int row = e / n, col = e % n;
    T[e] = V[row * n + sh.perm[col]];
  }
  for (int i = threadIdx.x; i < n; i += kScfThreads) evals[i] = A[sh.perm[i] * n + sh.perm[i]];
  __syncthreads();
  for (int e = threadIdx.x; e < n * n; e += kScfThreads) V[e] = T[e];
  __syncthreads();
}

// D = V_occ V_occ^T
__device__ void density(int n, int nocc, const double* V, double* D) {
  for (int e = threadIdx.x; e < n * n; e += kScfThreads) {
    const int i = e / n, j = e % n;
    double s = 0.0;
    for (int k = 0; k < nocc; ++k) s += V[i * n + k] * V[j * n + k];
    D[e] = s;
  }
  __syncthreads();
}

__global__ __launch_bounds__(kScfThreads) void k_rhf(Mol* __restrict__ mols, const ScfInput* __restrict__ in, int nm,
                                                     double accuracy, int maxiter, double step,
                                                     ScfResult* __restrict__ out) {
  extern __shared__ double lds[];
  const int mol = blockIdx.x;
  if (mol >= nm) return;
  Mol& M = mols[mol];
  const ScfInput& I = in[mol];
  const int n = M.n, nocc = M.o, n2 = n * n;
  ScfShared sh;
  sh.A = lds;
  sh.V = sh.A + n2;
  sh.D = sh.V + n2;
  sh.ev = sh.D + n2;
  sh.red = sh.ev + n;
  sh.cs = sh.red + kScfThreads / 64;
  sh.pq = reinterpret_cast<int*>(sh.cs + 2 * 33);
  sh.perm = sh.pq + 2 * 33;
  sh.T = I.work;
  double* evals = sh.ev;
  // Hueckel guess: eigenvectors of the Hueckel matrix
  for (int e = threadIdx.x; e < n2; e += kScfThreads) {
    sh.A[e] = I.huckel[e];
    sh.V[e] = (e / n == e % n) ? 1.0 : 0.0;
  }
  __syncthreads();
  jacobi(n, sh.A, sh.V, sh);
  sort_eigen(n, sh.A, sh.V, sh.T, evals, sh);
  density(n, nocc, sh.V, sh.D);
  fock(n, I.core, M.G, sh.D, sh.A);
  int it = 0, status = -1;
  double err = 0.0;
  while (true) {
    ++it;
    // T = F R - R F, R = 2D - 1; then Fmin = T R into A (F is rebuilt below)
    for (int e = threadIdx.x; e < n2; e += kScfThreads) {
      const int i = e / n, j = e % n;
      double s = 0.0;
      for (int k = 0; k < n; ++k) {
        const double rkj = 2.0 * sh.D[k * n + j] - (k == j ? 1.0 : 0.0);
        const double rik = 2.0 * sh.D[i * n + k] - (i == k ? 1.0 : 0.0);
        s += sh.A[i * n + k] * rkj - rik * sh.A[k * n + j];
      }
      sh.T[e] = s;
    }
    __syncthreads();
    double emax = 0.0;
    for (int e = threadIdx.x; e < n2; e += kScfThreads) {
      const int i = e / n, j = e % n;
      double s = 0.0;
      for (int k = 0; k < n; ++k) s += sh.T[i * n + k] * (2.0 * sh.D[k * n + j] - (k == j ? 1.0 : 0.0));
      sh.A[e] = s;
      emax = fmax(emax, fabs(s));
    }
    err = block_max_abs(emax, sh.red);   // (its syncs also order the A writes before the reads below)
    for (int e = threadIdx.x; e < n2; e += kScfThreads) sh.D[e] -= step * sh.A[e];
    __syncthreads();
    fock(n, I.core, M.G, sh.D, sh.T);                // T = F(D)
    // warm start: A = V^T F V, V kept
    matmul(n, sh.T, false, sh.V, false, sh.D);       // D (scratch) = F V
    matmul(n, sh.V, true, sh.D, false, sh.A);        // A = V^T F V
    jacobi(n, sh.A, sh.V, sh);
    sort_eigen(n, sh.A, sh.V, sh.T, evals, sh);
    density(n, nocc, sh.V, sh.D);
    fock(n, I.core, M.G, sh.D, sh.A);                // A = F(D)
    if (err < accuracy) { status = 1; break; }
    if (!(err == err)) { status = -2; break; }
    if (it + 1 > maxiter) { status = -1; break; }
  }
  // gauge: the largest-|C| entry of each orbital positive; symmetry-equivalent
  // sites tie up to rounding, so the first entry within 1e-9 of the largest decides
  for (int p = threadIdx.x; p < n; p += kScfThreads) {
    double big = 0.0;
    for (int mu = 0; mu < n; ++mu) big = fmax(big, fabs(sh.V[mu * n + p]));
    int best = 0;
    while (best < n - 1 && fabs(sh.V[best * n + p]) < big * (1.0 - 1e-9)) ++best;
    if (sh.V[best * n + p] < 0.0)
      for (int mu = 0; mu < n; ++mu) sh.V[mu * n + p] = -sh.V[mu * n + p];
  }
  __syncthreads();
  // reference energy: sum D o (core + F) + sum_{mu != nu} gamma / 2
  double e1 = 0.0, e2 = 0.0;
  for (int e = threadIdx.x; e < n2; e += kScfThreads) {
    e1 += sh.D[e] * (I.core[e] + sh.A[e]);
    if (e / n != e % n) e2 += 0.5 * M.G[e];
  }
  e1 = block_sum(e1, sh.red);
  e2 = block_sum(e2, sh.red);
  // outputs: C (row-major [mu][p]) and the MO Fock matrix C^T F C
  for (int e = threadIdx.x; e < n2; e += kScfThreads) M.C[e] = sh.V[e];
  matmul(n, sh.A, false, sh.V, false, sh.T);         // T = F C
  matmul(n, sh.V, true, sh.T, false, sh.D);          // D = C^T F C
  for (int e = threadIdx.x; e < n2; e += kScfThreads) M.F[e] = sh.D[e];
  if (threadIdx.x == 0) {
    M.eref = e1 + e2;
    out[mol].eref = e1 + e2;
    out[mol].error = err;
    out[mol].iterations = it;
    out[mol].status = status;
  }
}

size_t scf_lds_bytes(int nmax) {
  return sizeof(double) * (3 * (size_t)nmax * nmax + nmax + kScfThreads / 64 + 2 * 33) + sizeof(int) * (2 * 33 + nmax);
}

hipError_t launch_rhf(Mol* mols, const ScfInput* in, int nm, int nmax, double accuracy, int maxiter, double step,
                      ScfResult* out, hipStream_t s) {
  if (nm <= 0) return hipSuccess;
  const size_t lds = scf_lds_bytes(nmax);
  hipLaunchKernelGGL(k_rhf, dim3(nm), dim3(kScfThreads), lds, s, mols, in, nm, accuracy, maxiter, step, out);
  return hipGetLastError();
}

}  // namespace hcc
