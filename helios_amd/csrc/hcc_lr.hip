// Batched vector algebra of the LR-CCSD eigensolver on the device-resident
// vector space (hcc_lr.h).  Every kernel takes a job list over all molecules
// of the batch, so one launch serves the whole batch; the vectors stream from
// HBM once per job and nothing but job lists and scalars crosses PCIe.
#include "hcc_lr.h"

namespace hcc {

// molecule-free ragged map: job of workgroup b from the workgroup prefix
__device__ __forceinline__ int lr_job_of(const int* __restrict__ prefix, int njobs, int b) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// wave64 + LDS reduction of one value per thread, fixed order (sum or max)
template <bool kMax>
__device__ __forceinline__ double lr_block_reduce(double x, double* red) {
  for (int off = 32; off > 0; off >>= 1) {
    const double y = __shfl_down(x, off, 64);
    x = kMax ? fmax(x, y) : x + y;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = x;
  __syncthreads();
  double r = red[0];
  for (int w = 1; w < kLrThreads / 64; ++w) r = kMax ? fmax(r, red[w]) : r + red[w];
  return r;
}

__global__ __launch_bounds__(kLrThreads) void k_lr_dots(const double* __restrict__ space, const long* __restrict__ base,
                                                        const long* __restrict__ len, const int* __restrict__ jobs,
                                                        int mode, double* __restrict__ out) {
  __shared__ double red[kLrThreads / 64];
  const int j = blockIdx.x;
  const int mol = jobs[3 * j], sa = jobs[3 * j + 1], sb = jobs[3 * j + 2];
  const long n = len[mol];
  const double* __restrict__ a = space + base[mol] + (long)sa * n;
  const double* __restrict__ b = space + base[mol] + (long)(sb < 0 ? sa : sb) * n;
  double acc = 0.0;
  if (mode == 0) {
    // four independent partial sums per thread keep several loads in flight
    double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
    long e = threadIdx.x;
    for (; e + 3 * kLrThreads < n; e += 4 * kLrThreads) {
      p0 += a[e] * b[e];
      p1 += a[e + kLrThreads] * b[e + kLrThreads];
      p2 += a[e + 2 * kLrThreads] * b[e + 2 * kLrThreads];
      p3 += a[e + 3 * kLrThreads] * b[e + 3 * kLrThreads];
    }
    for (; e < n; e += kLrThreads) p0 += a[e] * b[e];
    acc = (p0 + p1) + (p2 + p3);
    acc = lr_block_reduce<false>(acc, red);
  } else {
    const bool pair = sb >= 0;
    for (long e = threadIdx.x; e < n; e += kLrThreads) {
      const double x = a[e], y = pair ? b[e] : 0.0;
      acc = fmax(acc, sqrt(x * x + y * y));
    }
    acc = lr_block_reduce<true>(acc, red);
  }
  if (threadIdx.x == 0) out[j] = acc;
}

__global__ __launch_bounds__(kLrThreads) void k_lr_combine(double* __restrict__ space, const long* __restrict__ base,
                                                           const long* __restrict__ len, const int* __restrict__ prefix,
                                                           int njobs, const int4* __restrict__ jobs,
                                                           const double* __restrict__ beta,
                                                           const int* __restrict__ tslot,
                                                           const double* __restrict__ tcoef) {
  const int j = lr_job_of(prefix, njobs, blockIdx.x);
  const int4 J = jobs[j];
  const long n = len[J.x];
  const long e0 = (long)(blockIdx.x - prefix[j]) * kLrChunk + threadIdx.x;
  double* __restrict__ dst = space + base[J.x] + (long)J.y * n;
  const double bt = beta[j];
  double acc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const long e = e0 + q * kLrThreads;
    acc[q] = (bt != 0.0 && e < n) ? bt * dst[e] : 0.0;
  }
  for (int t = 0; t < J.w; ++t) {
    const double c = tcoef[J.z + t];
    const double* __restrict__ src = space + base[J.x] + (long)tslot[J.z + t] * n;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const long e = e0 + q * kLrThreads;
      if (e < n) acc[q] += c * src[e];
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const long e = e0 + q * kLrThreads;
    if (e < n) dst[e] = acc[q];
  }
}

__global__ __launch_bounds__(kLrThreads) void k_lr_precondition(const Mol* __restrict__ mols, double* __restrict__ space,
                                                                const long* __restrict__ base,
                                                                const long* __restrict__ len,
                                                                const int* __restrict__ prefix, int njobs,
                                                                const int2* __restrict__ jobs,
                                                                const double* __restrict__ shift, double floor_) {
  const int j = lr_job_of(prefix, njobs, blockIdx.x);
  const int2 J = jobs[j];
  const Mol& M = mols[J.x];
  const long n = len[J.x];
  const int N = M.n, o = M.o, v = M.v;
  const long n1 = (long)o * v;
  const double* __restrict__ F = M.F;
  double* __restrict__ x = space + base[J.x] + (long)J.y * n;
  const double w = shift[j];
  const long e0 = (long)(blockIdx.x - prefix[j]) * kLrChunk + threadIdx.x;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const long e = e0 + q * kLrThreads;
    if (e >= n) continue;
    double d;
    if (e < n1) {
      const int i = (int)((unsigned)e / (unsigned)v), a = (int)((unsigned)e - (unsigned)i * (unsigned)v);
      d = F[(long)(o + a) * (N + 1)] - F[(long)i * (N + 1)];
    } else {
      const unsigned r = (unsigned)(e - n1);
      const unsigned q1 = r / (unsigned)v, q2 = q1 / (unsigned)v, q3 = q2 / (unsigned)o;
      const int b = (int)(r - q1 * (unsigned)v), a = (int)(q1 - q2 * (unsigned)v);
      const int jj = (int)(q2 - q3 * (unsigned)o), i = (int)q3;
      d = F[(long)(o + a) * (N + 1)] + F[(long)(o + b) * (N + 1)] - F[(long)i * (N + 1)] - F[(long)jj * (N + 1)];
    }
    double den = w - d;
    if (fabs(den) < floor_) den = copysign(floor_, den);
    x[e] /= den;
  }
}

void lr_dots(const double* space, const long* base, const long* len, int njobs, const int* jobs, int mode,
             double* out, hipStream_t s) {
  if (njobs > 0)
    hipLaunchKernelGGL(k_lr_dots, dim3(njobs), dim3(kLrThreads), 0, s, space, base, len, jobs, mode, out);
}

void lr_combine(const double* space, const long* base, const long* len, int njobs, const int* prefix,
                const int4* jobs, const double* beta, const int* tslot, const double* tcoef, int nblocks,
                hipStream_t s) {
  if (njobs > 0 && nblocks > 0)
    hipLaunchKernelGGL(k_lr_combine, dim3(nblocks), dim3(kLrThreads), 0, s, const_cast<double*>(space), base, len,
                       prefix, njobs, jobs, beta, tslot, tcoef);
}

void lr_precondition(const Mol* mols, const double* space, const long* base, const long* len, int njobs,
                     const int* prefix, const int2* jobs, const double* shift, double floor, int nblocks,
                     hipStream_t s) {
  if (njobs > 0 && nblocks > 0)
    hipLaunchKernelGGL(k_lr_precondition, dim3(nblocks), dim3(kLrThreads), 0, s, mols, const_cast<double*>(space),
                       base, len, prefix, njobs, jobs, shift, floor);
}

}  // namespace hcc
