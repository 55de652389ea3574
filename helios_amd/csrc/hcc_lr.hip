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

// Dot products (mode 0) or max-norms (mode 1) in chunks of kLrDotChunk
// elements, one workgroup per chunk, so that a long vector streams through
// many CUs; partial[chunk] is combined per job by k_lr_dots_final in chunk
// order (fixed order: the results do not depend on scheduling).
__global__ __launch_bounds__(kLrThreads) void k_lr_dots_part(const double* __restrict__ space,
                                                             const long* __restrict__ base,
                                                             const long* __restrict__ len,
                                                             const int* __restrict__ prefix, int njobs,
                                                             const int* __restrict__ jobs, int mode,
                                                             double* __restrict__ partial) {
  __shared__ double red[kLrThreads / 64];
  const int j = lr_job_of(prefix, njobs, blockIdx.x);
  const int mol = jobs[3 * j], sa = jobs[3 * j + 1], sb = jobs[3 * j + 2];
  const long n = len[mol];
  const long e0 = (long)(blockIdx.x - prefix[j]) * kLrDotChunk;
  const long e1 = min(n, e0 + kLrDotChunk);
  const double* __restrict__ a = space + base[mol] + (long)sa * n;
  const double* __restrict__ b = space + base[mol] + (long)(sb < 0 ? sa : sb) * n;
  double acc = 0.0;
  if (mode == 0) {
    // four independent partial sums per thread keep several loads in flight
    double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
    long e = e0 + threadIdx.x;
    for (; e + 3 * kLrThreads < e1; e += 4 * kLrThreads) {
      p0 += a[e] * b[e];
      p1 += a[e + kLrThreads] * b[e + kLrThreads];
      p2 += a[e + 2 * kLrThreads] * b[e + 2 * kLrThreads];
      p3 += a[e + 3 * kLrThreads] * b[e + 3 * kLrThreads];
    }
    for (; e < e1; e += kLrThreads) p0 += a[e] * b[e];
    acc = (p0 + p1) + (p2 + p3);
    acc = lr_block_reduce<false>(acc, red);
  } else {
    const bool pair = sb >= 0;
    for (long e = e0 + threadIdx.x; e < e1; e += kLrThreads) {
      const double x = a[e], y = pair ? b[e] : 0.0;
      acc = fmax(acc, sqrt(x * x + y * y));
    }
    acc = lr_block_reduce<true>(acc, red);
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

__global__ void k_lr_dots_final(const int* __restrict__ prefix, int njobs, int mode,
                                const double* __restrict__ partial, double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= njobs) return;
  double acc = 0.0;
  for (int c = prefix[j]; c < prefix[j + 1]; ++c) acc = mode == 0 ? acc + partial[c] : fmax(acc, partial[c]);
  out[j] = acc;
}

typedef double lr_d4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const double lr_gcd;
typedef __attribute__((address_space(1))) double lr_gd;

// Gram blocks: 4 waves, each taking 64 elements per step (lane group g =
// lane / 16 holds the 16 consecutive elements 16 g ...), so every lane reads
// 128 contiguous bytes of one vector per operand row.  MFMA operands: A row
// a = lane % 16 (vector slot[a0 + a]), B column b = lane % 16 of column block
// cb (slot[b0 + 16 cb + b]); the contraction index k = g, with the 16
// elements of a lane group taken one MFMA at a time.  Output block entry
// (row (lane / 16) + 4 r, column lane % 16) of block cb sits in acc[cb][r].
__global__ __launch_bounds__(kLrThreads) void k_lr_gram(const double* __restrict__ space,
                                                        const long* __restrict__ base,
                                                        const long* __restrict__ len,
                                                        const int* __restrict__ prefix, int nblk,
                                                        const LrGramBlk* __restrict__ blk,
                                                        const int* __restrict__ slots,
                                                        double* __restrict__ partial) {
  __shared__ double red[4][kLrThreads / 64][4][64];
  const int j = lr_job_of(prefix, nblk, blockIdx.x);
  const LrGramBlk B = blk[j];
  const long n = len[B.mol];
  lr_gcd* vb = (lr_gcd*)(space + base[B.mol]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, r16 = lane & 15;
  const int ncb = (B.nb + 15) >> 4;
  const bool arow = r16 < B.na;
  lr_gcd* pa = vb + (arow ? (long)slots[B.a + r16] * n : 0);
  lr_gcd* pb[4];
  bool brow[4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    brow[cb] = 16 * cb + r16 < B.nb;
    pb[cb] = vb + (brow[cb] ? (long)slots[B.b + 16 * cb + r16] * n : 0);
  }
  lr_d4 acc[4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) acc[cb] = lr_d4{0.0, 0.0, 0.0, 0.0};
  const long c0 = (long)(blockIdx.x - prefix[j]) * kLrGramChunk;
  const long c1 = min(n, c0 + kLrGramChunk);
  // 16 consecutive elements from e of row p (zero past the end or for an absent row)
  auto load16 = [&](lr_gcd* p, bool row, long e, double* x) {
    if (row && e + 16 <= n) {
      typedef double d2 __attribute__((ext_vector_type(2)));
      typedef __attribute__((address_space(1))) const d2 gd2;
      gd2* q = (gd2*)(p + e);
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const d2 y = q[i];
        x[2 * i] = y.x;
        x[2 * i + 1] = y.y;
      }
    } else {
#pragma unroll
      for (int i = 0; i < 16; ++i) x[i] = (row && e + i < n) ? p[e + i] : 0.0;
    }
  };
  for (long e0 = c0 + 64 * wave; e0 < c1; e0 += 64 * (kLrThreads / 64)) {
    const long e = e0 + 16 * g;
    double xa[16], xb[16];
    load16(pa, arow, e, xa);
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      if (cb < ncb) {
        load16(pb[cb], brow[cb], e, xb);
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[i], xb[i], acc[cb], 0, 0, 0);
      }
    }
  }
  // fixed-order sum over the waves, partial in accumulator order (cb, r, lane)
#pragma unroll
  for (int cb = 0; cb < 4; ++cb)
#pragma unroll
    for (int r = 0; r < 4; ++r) red[cb][wave][r][lane] = acc[cb][r];
  __syncthreads();
  double* __restrict__ out = partial + (long)blockIdx.x * 1024;
  for (int q = threadIdx.x; q < 1024; q += kLrThreads) {
    const int cb = q >> 8, r = (q >> 6) & 3, l = q & 63;
    double x = 0.0;
#pragma unroll
    for (int w = 0; w < kLrThreads / 64; ++w) x += red[cb][w][r][l];
    out[q] = x;
  }
}

__global__ __launch_bounds__(256) void k_lr_gram_final(const int* __restrict__ prefix, int nblk, const LrGramBlk* __restrict__ blk,
                                const double* __restrict__ partial, double* __restrict__ out) {
  const int j = blockIdx.x >> 2;
  if (j >= nblk) return;
  const LrGramBlk B = blk[j];
  const int q = (blockIdx.x & 3) * 256 + threadIdx.x;   // entry a * 64 + b
  const int a = q >> 6, b = q & 63;
  if (a >= B.na || b >= B.nb) return;
  // accumulator position of entry (a, b): block b / 16, lane b % 16 + 16 (a % 4), element a / 4
  const int idx = ((b >> 4) << 8) + ((a >> 2) << 6) + (b & 15) + 16 * (a & 3);
  double x = 0.0;
  for (int c = prefix[j]; c < prefix[j + 1]; ++c) x += partial[(long)c * 1024 + idx];
  out[B.out + (long)a * B.ld + b] = x;
}

// Block combinations: the product D (rows q < m, 16 consecutive elements per
// column block) = coef (m x t) * slots (t x elements), 4 column blocks per
// wave, K = the terms in steps of 4.  B operand: lane reads element
// e + 16 cb + lane % 16 of term slot[tt + k + lane / 16], so 16 lanes read
// 128 contiguous bytes.  Output row (lane / 16) + 4 r, element lane % 16.
__global__ __launch_bounds__(kLrThreads) void k_lr_combine_block(double* __restrict__ space,
                                                                 const long* __restrict__ base,
                                                                 const long* __restrict__ len,
                                                                 const int* __restrict__ prefix, int nblk,
                                                                 const LrCombBlk* __restrict__ blk,
                                                                 const int* __restrict__ slots,
                                                                 const double* __restrict__ coef) {
  const int j = lr_job_of(prefix, nblk, blockIdx.x);
  const LrCombBlk B = blk[j];
  const long n = len[B.mol];
  double* vb = space + base[B.mol];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, r16 = lane & 15;
  const long e0 = (long)(blockIdx.x - prefix[j]) * kLrCombChunk + 64 * wave;
  lr_d4 acc[4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) acc[cb] = lr_d4{0.0, 0.0, 0.0, 0.0};
  lr_gcd* cf = (lr_gcd*)coef + B.c;
  for (int k0 = 0; k0 < B.t; k0 += 4) {
    const int k = k0 + g;
    const bool kin = k < B.t;
    const double a = (r16 < B.m && kin) ? cf[r16 * B.t + k] : 0.0;
    lr_gcd* src = (lr_gcd*)vb + (kin ? (long)slots[B.tt + k] * n : 0);
    double x[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      const long e = e0 + 16 * cb + r16;
      x[cb] = (kin && e < n) ? src[e] : 0.0;
    }
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) acc[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, x[cb], acc[cb], 0, 0, 0);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int q = g + 4 * r;
    if (q >= B.m) continue;
    lr_gd* dst = (lr_gd*)vb + (long)slots[B.d + q] * n;
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      const long e = e0 + 16 * cb + r16;
      if (e < n) dst[e] = (B.beta != 0.0 ? B.beta * dst[e] : 0.0) + acc[cb][r];
    }
  }
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

void lr_dots(const double* space, const long* base, const long* len, int njobs, const int* prefix, const int* jobs,
             int mode, double* partial, int nchunks, double* out, hipStream_t s) {
  if (njobs <= 0 || nchunks <= 0) return;
  hipLaunchKernelGGL(k_lr_dots_part, dim3(nchunks), dim3(kLrThreads), 0, s, space, base, len, prefix, njobs, jobs,
                     mode, partial);
  hipLaunchKernelGGL(k_lr_dots_final, dim3((njobs + 255) / 256), dim3(256), 0, s, prefix, njobs, mode, partial, out);
}

void lr_gram(const double* space, const long* base, const long* len, int nblk, const int* prefix,
             const LrGramBlk* blk, const int* slots, double* partial, int nchunks, double* out, hipStream_t s) {
  if (nblk <= 0 || nchunks <= 0) return;
  hipLaunchKernelGGL(k_lr_gram, dim3(nchunks), dim3(kLrThreads), 0, s, space, base, len, prefix, nblk, blk, slots,
                     partial);
  hipLaunchKernelGGL(k_lr_gram_final, dim3(4 * nblk), dim3(256), 0, s, prefix, nblk, blk, partial, out);
}

void lr_combine_block(const double* space, const long* base, const long* len, int nblk, const int* prefix,
                      const LrCombBlk* blk, const int* slots, const double* coef, int nblocks, hipStream_t s) {
  if (nblk > 0 && nblocks > 0)
    hipLaunchKernelGGL(k_lr_combine_block, dim3(nblocks), dim3(kLrThreads), 0, s, const_cast<double*>(space), base,
                       len, prefix, nblk, blk, slots, coef);
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
