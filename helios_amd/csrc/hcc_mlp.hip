// Learned T2 guess: a 3-layer MLP (16 -> 64 -> 64 -> 1, exact-erf GELU) that
// corrects the MP2 amplitudes, t2 = t_mp2 + y / Dp, Dp = e_a + e_b - e_i - e_j.
// The network sees 16 features of each amplitude (MP2 amplitude, the two
// integrals (ia|jb) and (ib|ja), denominator, orbital energies about the
// midgap, site-localisation overlaps of the four orbitals, size) and predicts
// y = Dp (t_ccsd - t_mp2), the residual-scaled correction.
//
// Matrix-core layout.  One wave evaluates 16 amplitudes with
// V_MFMA_F32_16X16X32_BF16 in the transposed form Z^T = W X^T, so that a
// lane (sample n = lane & 15, group g = lane >> 4) holds hidden units
// 16 mb + 4 g + r (mb = 0..3, r = 0..3) of its sample after each layer.  The
// next layer's B operand must hold, per lane, 8 consecutive k-slots 8g..8g+7
// of its sample: k-slot q of step s is mapped to hidden unit
// 16 (2 s + q / 4) + 4 g + q % 4, which is exactly what the lane already
// holds, and the A operand (W2) is loaded in the same permuted k order.  No
// LDS round trip is needed between layers.  Weights and layer inputs are
// BF16, accumulation FP32; biases, GELU and the output layer are FP32.
//
// Training (GPU, deterministic): a forward kernel (same device function)
// stores the activations of a mini-batch, a per-sample backward kernel forms
// dz2 and dz1 in FP32, a weight-gradient kernel sums per-parameter over
// sample chunks, and one kernel reduces the chunks in a fixed order and
// applies Adam.  No atomics anywhere: results are bitwise reproducible.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/helios_ccsd.h"
#include "hcc_mlp.h"

namespace hcc {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int kAct = 272;   // stored activations per sample: z1, h1r, z2, h2 (64 each), x (16)
constexpr int kActZ1 = 0, kActH1 = 64, kActZ2 = 128, kActH2 = 192, kActX = 256;

__device__ __forceinline__ float bfr(float x) { return (float)(__bf16)x; }
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_prime(float z) {
  return 0.5f * (1.f + erff(z * 0.70710678118654752f)) + z * expf(-0.5f * z * z) * 0.39894228040143268f;
}

// ---------------------------------------------------------------------------
// features
// ---------------------------------------------------------------------------
// The 16 features of an amplitude from the quantities of its pair
// representative: orbital energies e_i, e_j, e_a, e_b, the midgap mu, the
// integrals K = (ia|jb) and X = (ib|ja), the site-localisation overlaps
// O_ij, O_ab, O_ia, O_jb, O_ib, O_ja and the site count.  Returns the MP2
// amplitude t and Dp.  (One formula for the guess, the training rows and
// every kernel that evaluates the network, so their features agree bitwise.)
__device__ __forceinline__ void features_of(double ei, double ej, double ea, double eb, double mu, double K, double X,
                                            double oij, double oab, double oia, double ojb, double oib, double oja,
                                            int N, float* f, double& t, double& Dp) {
  Dp = (ea + eb) - (ei + ej);
  t = -K / Dp;
  f[0] = (float)(10.0 * t);
  f[1] = (float)(K / 2.0);
  f[2] = (float)(X / 2.0);
  f[3] = (float)(Dp / 10.0);
  f[4] = (float)(2.0 / Dp);
  f[5] = (float)(((ei + ej) - 2.0 * mu) / 10.0);
  f[6] = (float)(((ea + eb) - 2.0 * mu) / 10.0);
  f[7] = (float)(fabs(ei - ej) / 10.0);
  f[8] = (float)(fabs(ea - eb) / 10.0);
  f[9] = (float)oij;
  f[10] = (float)oab;
  f[11] = (float)((oia + ojb) / 2.0);
  f[12] = (float)(fabs(oia - ojb) / 2.0);
  f[13] = (float)((oib + oja) / 2.0);
  f[14] = (float)(fabs(oib - oja) / 2.0);
  f[15] = (float)(10.0 / N);
}

// Features of amplitude (i, j, a, b) (virtual indices from 0), computed from
// the pair's representative (i a) <= (j b) so that t2 stays exactly
// pair-symmetric.  Returns the MP2 amplitude and Dp.
__device__ void amp_features(const Mol& M, int i, int j, int a, int b, float* f, double& t, double& Dp) {
  const int N = M.n, o = M.o, v = M.v;
  if (i * v + a > j * v + b) {
    int x = i; i = j; j = x;
    x = a; a = b; b = x;
  }
  const double* F = M.F;
  const double ei = F[(long)i * N + i], ej = F[(long)j * N + j];
  const double ea = F[(long)(o + a) * N + o + a], eb = F[(long)(o + b) * N + o + b];
  const double mu = 0.5 * (F[(long)(o - 1) * N + o - 1] + F[(long)o * N + o]);
  const double K = M.Gklcd[(((long)i * o + j) * v + a) * v + b];   // (ia|jb)
  const double X = M.Gklcd[(((long)i * o + j) * v + b) * v + a];   // (ib|ja)
  const double* O = M.Ov;
  features_of(ei, ej, ea, eb, mu, K, X, O[(long)i * N + j], O[(long)(o + a) * N + o + b], O[(long)i * N + o + a],
              O[(long)j * N + o + b], O[(long)i * N + o + b], O[(long)j * N + o + a], N, f, t, Dp);
}

__device__ __forceinline__ void decode(long e, int o, int v, int& i, int& j, int& a, int& b) {
  b = (int)(e % v);
  a = (int)((e / v) % v);
  j = (int)((e / ((long)v * v)) % o);
  i = (int)(e / ((long)o * v * v));
}

// ---------------------------------------------------------------------------
// network in matrix-core operand order
// ---------------------------------------------------------------------------
struct Net {
  bf8 w1[4];        // layer 1, A operand of hidden block mb: W1[16 mb + row][8 g + q] (g < 2)
  bf8 w2[4][2];     // layer 2, A operand of block mb, k-step s: W2[16 mb + row][unit(s, g, q)]
  float b1[16], b2[16], w3[16];   // this lane's units 16 mb + 4 g + r at index 4 mb + r
  float b3;
};

__device__ __forceinline__ int unit(int s, int g, int q) { return 16 * (2 * s + (q >> 2)) + 4 * g + (q & 3); }

__device__ void net_load(Net& N, const float* __restrict__ w, int lane) {
  const int row = lane & 15, g = lane >> 4;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      N.w1[mb][q] = (__bf16)(g < 2 ? w[kOffW1 + (16 * mb + row) * kMlpF + 8 * g + q] : 0.f);
#pragma unroll
      for (int s = 0; s < 2; ++s) N.w2[mb][s][q] = (__bf16)w[kOffW2 + (16 * mb + row) * kMlpH + unit(s, g, q)];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int u = 16 * mb + 4 * g + r;
      N.b1[4 * mb + r] = w[kOffB1 + u];
      N.b2[4 * mb + r] = w[kOffB2 + u];
      N.w3[4 * mb + r] = bfr(w[kOffW3 + u]);
    }
  }
  N.b3 = w[kOffB3];
}

struct Acts {
  float z1[16], h1r[16], z2[16], h2[16];
};

// Forward pass of the wave's 16 samples; x holds features 8g..8g+7 of sample
// lane & 15 (zeros for g >= 2).  Every lane returns its sample's output.
template <bool STORE>
__device__ float net_forward(const Net& N, const bf8 x, Acts& A) {
  f4 c1[4];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
    c1[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(N.w1[mb], x, f4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
  bf8 hb[2];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float z = c1[mb][r] + N.b1[4 * mb + r];
      const float h = gelu(z);
      hb[mb >> 1][4 * (mb & 1) + r] = (__bf16)h;
      if (STORE) {
        A.z1[4 * mb + r] = z;
        A.h1r[4 * mb + r] = bfr(h);
      }
    }
  f4 c2[4];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    c2[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(N.w2[mb][0], hb[0], f4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
    c2[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(N.w2[mb][1], hb[1], c2[mb], 0, 0, 0);
  }
  float p = 0.f;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float z = c2[mb][r] + N.b2[4 * mb + r];
      const float h = gelu(z);
      p += N.w3[4 * mb + r] * h;
      if (STORE) {
        A.z2[4 * mb + r] = z;
        A.h2[4 * mb + r] = h;
      }
    }
  p += __shfl_xor(p, 16);
  p += __shfl_xor(p, 32);
  return p + N.b3;
}

__device__ __forceinline__ bf8 pack_x(const float* f, int g) {
  bf8 x;
#pragma unroll
  for (int q = 0; q < 8; ++q) x[q] = (__bf16)(g == 0 ? f[q] : (g == 1 ? f[8 + q] : 0.f));
  return x;
}

// ---------------------------------------------------------------------------
// solver-side kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mlp_overlap(const Mol* __restrict__ mols, Ragged rg, int nm) {
  const int mol = rg.bmol[blockIdx.x];
  const Mol& M = mols[mol];
  const long e = (long)(blockIdx.x - rg.prefix[mol]) * 256 + threadIdx.x;
  const int N = M.n;
  if (e >= (long)N * N) return;
  const int p = (int)(e / N), q = (int)(e % N);
  double s = 0.0;
  for (int mu = 0; mu < N; ++mu) {
    const double cp = M.C[(long)mu * N + p], cq = M.C[(long)mu * N + q];
    s += (cp * cp) * (cq * cq);
  }
  M.Ov[e] = N * s;
}

// The learned guess over the occupied block pairs i <= j of a molecule
// (o(o+1)/2 pairs; each workgroup takes guess_pairs_per_workgroup(v)
// consecutive ones, about 1024 amplitudes, so the network is loaded once per
// workgroup).
// Every (a, b) of block (i, j) with i < j is a pair representative
// ((i a) <= (j b)); for i = j those with a <= b are.  The block's integrals
// (kc|ld) at [i][j][:][:], the virtual-virtual overlaps and the overlap rows
// of i and j are read once, coalesced, into LDS; the network runs once per
// representative; t2[i][j][a][b] is written directly (consecutive b per wave)
// and the mirror block t2[j][i][b][a] through an LDS transpose, so t2 stays
// exactly pair-symmetric and every global access is coalesced.
constexpr int kGuessV = 32;   // v <= 32 (larger molecules: k_mlp_guess_pairs)
// Three waves per SIMD (168 VGPRs and a few spilled registers) hide the GELU
// chain's latency better than the two that the full register set allows:
// 38 -> 30 ms for the 4096-molecule batch (four waves spill 280 bytes a lane).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_mlp_guess_blocks(const Mol* __restrict__ mols, Ragged rg, int nm,
                                                          const float* __restrict__ w) {
  __shared__ double sK[kGuessV * kGuessV], sOvv[kGuessV * kGuessV], sT[kGuessV * (kGuessV + 1)];
  __shared__ double sOi[kGuessV], sOj[kGuessV], sEv[kGuessV];
  const int mol = rg.bmol[blockIdx.x];
  const Mol& M = mols[mol];
  const int N = M.n, o = M.o, v = M.v, vv = v * v;
  if (v > kGuessV) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, n = lane & 15, g = lane >> 4;
  Net net;
  net_load(net, w, lane);
  Acts unused;
  const double* F = M.F;
  const double* O = M.Ov;
  const double mu = 0.5 * (F[(long)(o - 1) * N + o - 1] + F[(long)o * N + o]);
  for (int e = threadIdx.x; e < vv; e += 256) sOvv[e] = O[(long)(o + e / v) * N + o + e % v];
  for (int e = threadIdx.x; e < v; e += 256) sEv[e] = F[(long)(o + e) * N + o + e];
  const int npb = o * (o + 1) / 2, bpw = guess_pairs_per_workgroup(v);
  const int pb0 = (blockIdx.x - rg.prefix[mol]) * bpw, pb1 = min(npb, pb0 + bpw);
  for (int pb = pb0; pb < pb1; ++pb) {
    int j = (int)((sqrt(8.0 * pb + 1.0) - 1.0) * 0.5);
    while (j * (j + 1) / 2 > pb) --j;
    while ((j + 1) * (j + 2) / 2 <= pb) ++j;
    const int i = pb - j * (j + 1) / 2;   // i <= j < o
    const long blk = ((long)i * o + j) * vv;
    __syncthreads();   // the previous pair's LDS reads and mirror writes are done
    for (int e = threadIdx.x; e < vv; e += 256) sK[e] = M.Gklcd[blk + e];
    for (int e = threadIdx.x; e < v; e += 256) {
      sOi[e] = O[(long)i * N + o + e];
      sOj[e] = O[(long)j * N + o + e];
    }
    __syncthreads();
    const double ei = F[(long)i * N + i], ej = F[(long)j * N + j];
    const double oij = O[(long)i * N + j];
    // representatives in order: (a, b) row-major over the block (i < j), or
    // the upper triangle a <= b row by row (i = j)
    const int nrep = i < j ? vv : v * (v + 1) / 2;
    for (int base = wave * 16; base < nrep; base += 64) {
      const int q = base + n;
      const bool ok = q < nrep;
      int a = 0, b = 0;
      if (ok) {
        if (i < j) {
          a = q / v;
          b = q - a * v;
        } else {
          // row a of the upper triangle starts at a v - a (a - 1) / 2
          a = (int)(((2.0 * v + 1.0) - sqrt((2.0 * v + 1.0) * (2.0 * v + 1.0) - 8.0 * q)) * 0.5);
          while (a > 0 && a * v - a * (a - 1) / 2 > q) --a;
          while ((a + 1) * v - (a + 1) * a / 2 <= q) ++a;
          b = a + (q - (a * v - a * (a - 1) / 2));
        }
      }
      float f[kMlpF];
#pragma unroll
      for (int r = 0; r < kMlpF; ++r) f[r] = 0.f;
      double t = 0.0, Dp = 1.0;
      if (ok && g < 2)
        features_of(ei, ej, sEv[a], sEv[b], mu, sK[a * v + b], sK[b * v + a], oij, sOvv[a * v + b], sOi[a], sOj[b],
                    sOi[b], sOj[a], N, f, t, Dp);
      const float y = net_forward<false>(net, pack_x(f, g), unused);
      if (ok && g == 0) {
        const double val = t + (double)y / Dp;
        M.t2[blk + a * v + b] = val;
        if (i < j) sT[b * (kGuessV + 1) + a] = val;   // t2[j][i][b][a]
        else M.t2[blk + b * v + a] = val;
      }
    }
    if (i < j) {
      __syncthreads();
      const long mblk = ((long)j * o + i) * vv;
      for (int e = threadIdx.x; e < vv; e += 256) M.t2[mblk + e] = sT[(e / v) * (kGuessV + 1) + e % v];
    }
  }
}

// Molecules with v > 32: one network evaluation per amplitude pair; work item
// p enumerates the pairs (mm, nn) = ((i a), (j b)) with mm <= nn
// (p = nn (nn + 1) / 2 + mm), and the value is written to t2[i][j][a][b] and
// its mirror t2[j][i][b][a].  4 waves x 16 groups x 16 amplitudes per workgroup.
__global__ __launch_bounds__(256) void k_mlp_guess_pairs(const Mol* __restrict__ mols, Ragged rg, int nm,
                                                         const float* __restrict__ w) {
  const int mol = rg.bmol[blockIdx.x];
  const Mol& M = mols[mol];
  const int o = M.o, v = M.v;
  if (v <= kGuessV) return;
  const long ne = (long)o * v, npair = ne * (ne + 1) / 2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, n = lane & 15, g = lane >> 4;
  const long base = (long)(blockIdx.x - rg.prefix[mol]) * 1024 + wave * 256;
  Net net;
  net_load(net, w, lane);
  Acts unused;
  for (int grp = 0; grp < 16; ++grp) {
    const long p = base + grp * 16 + n;
    const bool ok = p < npair;
    float f[kMlpF];
#pragma unroll
    for (int q = 0; q < kMlpF; ++q) f[q] = 0.f;
    double t = 0.0, Dp = 1.0;
    long e = 0, em = 0;
    if (ok) {
      long nn = (long)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
      while (nn * (nn + 1) / 2 > p) --nn;
      while ((nn + 1) * (nn + 2) / 2 <= p) ++nn;
      const long mm = p - nn * (nn + 1) / 2;
      const int i = (int)(mm / v), a = (int)(mm % v), j = (int)(nn / v), b = (int)(nn % v);
      e = (((long)i * o + j) * v + a) * v + b;
      em = (((long)j * o + i) * v + b) * v + a;
      if (g < 2) amp_features(M, i, j, a, b, f, t, Dp);
    }
    const float y = net_forward<false>(net, pack_x(f, g), unused);
    if (ok && g == 0) {
      const double val = t + (double)y / Dp;
      M.t2[e] = val;
      M.t2[em] = val;
    }
  }
}

void mlp_overlaps(const Mol* mols, Ragged rgNN, int nm, hipStream_t s) {
  if (rgNN.nblocks > 0) hipLaunchKernelGGL(k_mlp_overlap, dim3(rgNN.nblocks), dim3(256), 0, s, mols, rgNN, nm);
}
void mlp_guess(const Mol* mols, Ragged rgPairBlk, Ragged rgPairL, int nm, bool any_large, const float* params,
               hipStream_t s) {
  if (rgPairBlk.nblocks > 0)
    hipLaunchKernelGGL(k_mlp_guess_blocks, dim3(rgPairBlk.nblocks), dim3(256), 0, s, mols, rgPairBlk, nm, params);
  if (any_large && rgPairL.nblocks > 0)
    hipLaunchKernelGGL(k_mlp_guess_pairs, dim3(rgPairL.nblocks), dim3(256), 0, s, mols, rgPairL, nm, params);
}

// ---------------------------------------------------------------------------
// training kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned mix32(unsigned h) {
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  h *= 0x846ca68bu;
  h ^= h >> 16;
  return h;
}

// training rows of the converged molecules: per_mol hashed amplitudes each
__global__ __launch_bounds__(256) void k_mlp_collect(const Mol* __restrict__ mols, const State* __restrict__ st, int nm,
                                                     int per_mol, unsigned seed, float* __restrict__ rows,
                                                     int* __restrict__ where) {
  const long gid = (long)blockIdx.x * 256 + threadIdx.x;
  if (gid >= (long)nm * per_mol) return;
  const int mol = (int)(gid / per_mol), s = (int)(gid % per_mol);
  const Mol& M = mols[mol];
  const long n2 = (long)M.o * M.o * M.v * M.v;
  const unsigned h = mix32(seed ^ mix32(0x9e3779b9u * (unsigned)(mol + 1) + 0x85ebca6bu * (unsigned)s));
  const long e = (long)(h % (unsigned long)n2);
  float* r = rows + gid * kMlpRow;
  where[2 * gid] = mol;
  where[2 * gid + 1] = (int)e;
  if (st[mol].status != 1) {
    for (int q = 0; q < kMlpRow; ++q) r[q] = 0.f;
    return;
  }
  int i, j, a, b;
  decode(e, M.o, M.v, i, j, a, b);
  float f[kMlpF];
  double t, Dp;
  amp_features(M, i, j, a, b, f, t, Dp);
  for (int q = 0; q < kMlpF; ++q) r[q] = f[q];
  r[kMlpF] = (float)(Dp * (M.t2[e] - t));
  r[kMlpF + 1] = 1.f;
}

// Forward over a mini-batch: sample sb of the batch is row (base + sb) * A + C
// mod ns (a permutation of the rows for gcd(A, ns) = 1).  STORE: keep the
// activations and dL/dy for the backward pass.  Per-workgroup loss sums
// (fixed order) go to lossp[blockIdx.x].
template <bool STORE>
__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ w, const float* __restrict__ rows,
                                                     long ns, long pa, long pc, long base, int nb,
                                                     float* __restrict__ act, float* __restrict__ dy,
                                                     double* __restrict__ lossp, float* __restrict__ yout) {
  __shared__ double sl[256];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, n = lane & 15, g = lane >> 4;
  Net net;
  net_load(net, w, lane);
  for (int grp = 0; grp < 4; ++grp) {
    const int loc = wave * 64 + grp * 16 + n;
    const int sb = blockIdx.x * 256 + loc;
    const bool ok = sb < nb;
    const long row = ok ? ((base + sb) * pa + pc) % ns : 0;
    const float* R = rows + row * kMlpRow;
    float f[kMlpF];
#pragma unroll
    for (int q = 0; q < kMlpF; ++q) f[q] = ok ? R[q] : 0.f;
    Acts A;
    const float y = net_forward<STORE>(net, pack_x(f, g), A);
    const float tg = ok ? R[kMlpF] : 0.f, wt = ok ? R[kMlpF + 1] : 0.f;
    const float d = y - tg;
    if (g == 0) sl[loc] = ok ? (double)(wt * d * d) : 0.0;
    if (yout != nullptr && ok && g == 0) yout[sb] = y;
    if (STORE && ok) {
      float* P = act + (long)sb * kAct;
#pragma unroll
      for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int u = 16 * mb + 4 * g + r;
          P[kActZ1 + u] = A.z1[4 * mb + r];
          P[kActH1 + u] = A.h1r[4 * mb + r];
          P[kActZ2 + u] = A.z2[4 * mb + r];
          P[kActH2 + u] = A.h2[4 * mb + r];
        }
      if (g < 2)
#pragma unroll
        for (int q = 0; q < 8; ++q) P[kActX + 8 * g + q] = bfr(g == 0 ? f[q] : f[8 + q]);
      if (g == 0) dy[sb] = 2.f * wt * d / (float)nb;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < 256; ++k) s += sl[k];
    lossp[blockIdx.x] = s;
  }
}

// per-sample backward: dz2 = dy w3 gelu'(z2), dz1 = (W2^T dz2) gelu'(z1)
__global__ __launch_bounds__(256) void k_mlp_backward(const float* __restrict__ w, int nb,
                                                      const float* __restrict__ act, const float* __restrict__ dy,
                                                      float* __restrict__ gact) {
  __shared__ float W2s[kMlpH * kMlpH];
  __shared__ float w3s[kMlpH];
  for (int q = threadIdx.x; q < kMlpH * kMlpH; q += 256) W2s[q] = bfr(w[kOffW2 + q]);
  if (threadIdx.x < kMlpH) w3s[threadIdx.x] = bfr(w[kOffW3 + threadIdx.x]);
  __syncthreads();
  const int sb = blockIdx.x * 256 + threadIdx.x;
  if (sb >= nb) return;
  const float* P = act + (long)sb * kAct;
  float* G = gact + (long)sb * 2 * kMlpH;
  const float d = dy[sb];
  float dz2[kMlpH];
#pragma unroll
  for (int k = 0; k < kMlpH; ++k) {
    dz2[k] = d * w3s[k] * gelu_prime(P[kActZ2 + k]);
    G[kMlpH + k] = dz2[k];
  }
  for (int j = 0; j < kMlpH; ++j) {
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < kMlpH; ++k) s += W2s[k * kMlpH + j] * dz2[k];
    G[j] = s * gelu_prime(P[kActZ1 + j]);
  }
}

// per-parameter gradient sums over one chunk of the mini-batch
__global__ __launch_bounds__(256) void k_mlp_wgrad(int nb, int chunk, const float* __restrict__ act,
                                                   const float* __restrict__ gact, const float* __restrict__ dy,
                                                   float* __restrict__ part) {
  const int p = blockIdx.y * 256 + threadIdx.x;
  if (p >= kMlpParams) return;
  const int n0 = blockIdx.x * chunk, n1 = min(nb, n0 + chunk);
  float s = 0.f;
  if (p < kOffB1) {
    const int m = p / kMlpF, f = p % kMlpF;
    for (int n = n0; n < n1; ++n) s += gact[(long)n * 128 + m] * act[(long)n * kAct + kActX + f];
  } else if (p < kOffW2) {
    const int m = p - kOffB1;
    for (int n = n0; n < n1; ++n) s += gact[(long)n * 128 + m];
  } else if (p < kOffB2) {
    const int m = (p - kOffW2) / kMlpH, k = (p - kOffW2) % kMlpH;
    for (int n = n0; n < n1; ++n) s += gact[(long)n * 128 + kMlpH + m] * act[(long)n * kAct + kActH1 + k];
  } else if (p < kOffW3) {
    const int m = p - kOffB2;
    for (int n = n0; n < n1; ++n) s += gact[(long)n * 128 + kMlpH + m];
  } else if (p < kOffB3) {
    const int k = p - kOffW3;
    for (int n = n0; n < n1; ++n) s += dy[n] * act[(long)n * kAct + kActH2 + k];
  } else {
    for (int n = n0; n < n1; ++n) s += dy[n];
  }
  part[(long)blockIdx.x * kMlpParams + p] = s;
}

// fixed-order reduction of the chunk sums, then (update) one Adam step
__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ w, float* __restrict__ m1, float* __restrict__ m2,
                                                  float* __restrict__ grad, const float* __restrict__ part, int nchunk,
                                                  float lr, float bc1, float bc2, int update) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= kMlpParams) return;
  float g = 0.f;
  for (int c = 0; c < nchunk; ++c) g += part[(long)c * kMlpParams + p];
  grad[p] = g;
  if (!update) return;
  const float a = 0.9f * m1[p] + 0.1f * g;
  const float b = 0.999f * m2[p] + 0.001f * g * g;
  m1[p] = a;
  m2[p] = b;
  w[p] -= lr * (a / bc1) / (sqrtf(b / bc2) + 1e-8f);
}

}  // namespace hcc

// ---------------------------------------------------------------------------
// trainer object and its C ABI
// ---------------------------------------------------------------------------
using namespace hcc;

struct hcc_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  float *w = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr;
  float* rows = nullptr;
  int* where = nullptr;
  long cap = 0, count = 0;
  // mini-batch buffers (grown on demand)
  int bcap = 0;
  float *act = nullptr, *gact = nullptr, *dy = nullptr, *part = nullptr, *yout = nullptr;
  double* lossp = nullptr;
  long lcap = 0;
  long step = 0;
  std::string err;
};

namespace {

constexpr int kChunk = 256;   // samples per weight-gradient chunk

template <class T>
void dalloc(T** p, size_t count) {
  if (*p) HCHECK(hipFree(*p));
  *p = nullptr;
  HCHECK(hipMalloc((void**)p, std::max<size_t>(1, count) * sizeof(T)));
}

void ensure_batch(hcc_mlp* net, int nb) {
  if (nb <= net->bcap) return;
  const int cap = std::max(nb, 256);
  dalloc(&net->act, (size_t)cap * kAct);
  dalloc(&net->gact, (size_t)cap * 2 * kMlpH);
  dalloc(&net->dy, (size_t)cap);
  dalloc(&net->yout, (size_t)cap);
  dalloc(&net->part, (size_t)((cap + kChunk - 1) / kChunk) * kMlpParams);
  net->bcap = cap;
}

void ensure_loss(hcc_mlp* net, long blocks) {
  if (blocks <= net->lcap) return;
  dalloc(&net->lossp, (size_t)blocks);
  net->lcap = blocks;
}

// one mini-batch: forward (stored), backward, gradient; update = Adam step
void batch_step(hcc_mlp* net, long ns, long pa, long pc, long base, int nb, double* lossp, bool update, float lr) {
  hipStream_t s = net->stream;
  const int blocks = (nb + 255) / 256;
  hipLaunchKernelGGL(k_mlp_forward<true>, dim3(blocks), dim3(256), 0, s, net->w, net->rows, ns, pa, pc, base, nb,
                     net->act, net->dy, lossp, (float*)nullptr);
  hipLaunchKernelGGL(k_mlp_backward, dim3(blocks), dim3(256), 0, s, net->w, nb, net->act, net->dy, net->gact);
  const int nchunk = (nb + kChunk - 1) / kChunk;
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(nchunk, (kMlpParams + 255) / 256), dim3(256), 0, s, nb, kChunk, net->act,
                     net->gact, net->dy, net->part);
  float bc1 = 1.f, bc2 = 1.f;
  if (update) {
    net->step += 1;
    bc1 = (float)(1.0 - std::pow(0.9, (double)net->step));
    bc2 = (float)(1.0 - std::pow(0.999, (double)net->step));
  }
  hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, net->w, net->m1, net->m2,
                     net->grad, net->part, nchunk, lr, bc1, bc2, update ? 1 : 0);
  HCHECK(hipGetLastError());
}

long gcd_l(long a, long b) {
  while (b) {
    long t = a % b;
    a = b;
    b = t;
  }
  return a;
}

}  // namespace

#define MLP_BEGIN                   \
  if (net == nullptr) return HCC_E_ARG; \
  try {                             \
    HCHECK(hipSetDevice(net->device));
#define MLP_END                                                          \
  }                                                                      \
  catch (const HipError& he) {                                           \
    net->err = std::string(he.what) + ": " + hipGetErrorString(he.e);   \
    return HCC_E_HIP;                                                    \
  }                                                                      \
  catch (const std::string& se) {                                        \
    net->err = se;                                                       \
    return HCC_E_ARG;                                                    \
  }                                                                      \
  catch (const std::bad_alloc&) {                                        \
    net->err = "host allocation failed";                                 \
    return HCC_E_MEMORY;                                                 \
  }                                                                      \
  return HCC_OK;

extern "C" {

int hcc_mlp_create(hcc_mlp** out, int device, long capacity) {
  if (out == nullptr || capacity < 0) return HCC_E_ARG;
  *out = nullptr;
  hcc_mlp* net = new hcc_mlp();
  net->device = device;
  int rc = [&]() -> int {
    MLP_BEGIN
    HCHECK(hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking));
    dalloc(&net->w, kMlpParams);
    dalloc(&net->m1, kMlpParams);
    dalloc(&net->m2, kMlpParams);
    dalloc(&net->grad, kMlpParams);
    HCHECK(hipMemset(net->w, 0, kMlpParams * sizeof(float)));
    HCHECK(hipMemset(net->m1, 0, kMlpParams * sizeof(float)));
    HCHECK(hipMemset(net->m2, 0, kMlpParams * sizeof(float)));
    dalloc(&net->rows, (size_t)capacity * kMlpRow);
    dalloc(&net->where, (size_t)capacity * 2);
    net->cap = capacity;
    MLP_END
  }();
  if (rc != HCC_OK) {
    hcc_mlp_destroy(net);
    return rc;
  }
  *out = net;
  return HCC_OK;
}

void hcc_mlp_destroy(hcc_mlp* net) {
  if (net == nullptr) return;
  (void)hipSetDevice(net->device);
  for (void* p : {(void*)net->w, (void*)net->m1, (void*)net->m2, (void*)net->grad, (void*)net->rows,
                  (void*)net->where, (void*)net->act, (void*)net->gact, (void*)net->dy, (void*)net->part,
                  (void*)net->yout, (void*)net->lossp})
    if (p) (void)hipFree(p);
  if (net->stream) (void)hipStreamDestroy(net->stream);
  delete net;
}

const char* hcc_mlp_last_error(const hcc_mlp* net) { return net ? net->err.c_str() : "null trainer"; }

int hcc_mlp_set_params(hcc_mlp* net, const float* params, int count) {
  if (net == nullptr || params == nullptr) return HCC_E_ARG;
  if (count != kMlpParams) {
    net->err = "parameter count must be HCC_MLP_PARAMS";
    return HCC_E_ARG;
  }
  MLP_BEGIN
  HCHECK(hipMemcpy(net->w, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
  HCHECK(hipMemset(net->m1, 0, kMlpParams * sizeof(float)));
  HCHECK(hipMemset(net->m2, 0, kMlpParams * sizeof(float)));
  net->step = 0;
  MLP_END
}

int hcc_mlp_get_params(hcc_mlp* net, float* params, int count) {
  if (net == nullptr || params == nullptr) return HCC_E_ARG;
  if (count != kMlpParams) {
    net->err = "parameter count must be HCC_MLP_PARAMS";
    return HCC_E_ARG;
  }
  MLP_BEGIN
  HCHECK(hipMemcpy(params, net->w, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
  MLP_END
}

int hcc_mlp_collect(hcc_mlp* net, hcc_ctx* ctx, int per_molecule, unsigned seed, long* added) {
  if (net == nullptr || ctx == nullptr || per_molecule < 1) return HCC_E_ARG;
  BatchView bv;
  if (!batch_view(ctx, &bv)) {
    net->err = "the context holds no solved molecules";
    return HCC_E_STATE;
  }
  if (bv.device != net->device) {
    net->err = "trainer and context are on different devices";
    return HCC_E_ARG;
  }
  const long nrow = (long)bv.nm * per_molecule;
  if (net->count + nrow > net->cap) {
    net->err = "sample capacity exceeded";
    return HCC_E_ARG;
  }
  MLP_BEGIN
  HCHECK(hipStreamSynchronize(bv.stream));
  mlp_overlaps(bv.dm, bv.rgNN, bv.nm, net->stream);
  const long blocks = (nrow + 255) / 256;
  hipLaunchKernelGGL(k_mlp_collect, dim3((unsigned)blocks), dim3(256), 0, net->stream, bv.dm, bv.dst, bv.nm,
                     per_molecule, seed, net->rows + net->count * kMlpRow, net->where + net->count * 2);
  HCHECK(hipGetLastError());
  HCHECK(hipStreamSynchronize(net->stream));
  net->count += nrow;
  if (added) *added = nrow;
  MLP_END
}

int hcc_mlp_samples(hcc_mlp* net, long* count) {
  if (net == nullptr || count == nullptr) return HCC_E_ARG;
  *count = net->count;
  return HCC_OK;
}

int hcc_mlp_read_samples(hcc_mlp* net, long first, long count, float* rows, int* where) {
  if (net == nullptr || first < 0 || count < 0 || first + count > net->count) return HCC_E_ARG;
  MLP_BEGIN
  if (rows && count)
    HCHECK(hipMemcpy(rows, net->rows + first * kMlpRow, count * kMlpRow * sizeof(float), hipMemcpyDeviceToHost));
  if (where && count)
    HCHECK(hipMemcpy(where, net->where + first * 2, count * 2 * sizeof(int), hipMemcpyDeviceToHost));
  MLP_END
}

int hcc_mlp_write_samples(hcc_mlp* net, const float* rows, long count) {
  if (net == nullptr || rows == nullptr || count < 0) return HCC_E_ARG;
  if (count > net->cap) {
    net->err = "sample capacity exceeded";
    return HCC_E_ARG;
  }
  MLP_BEGIN
  if (count) HCHECK(hipMemcpy(net->rows, rows, count * kMlpRow * sizeof(float), hipMemcpyHostToDevice));
  std::vector<int> wh((size_t)count * 2, -1);
  if (count) HCHECK(hipMemcpy(net->where, wh.data(), wh.size() * sizeof(int), hipMemcpyHostToDevice));
  net->count = count;
  MLP_END
}

int hcc_mlp_gradient(hcc_mlp* net, long first, int count, double* loss, float* grad) {
  if (net == nullptr || first < 0 || count < 1 || first + count > net->count) return HCC_E_ARG;
  MLP_BEGIN
  ensure_batch(net, count);
  const int blocks = (count + 255) / 256;
  ensure_loss(net, blocks);
  batch_step(net, net->count, 1, 0, first, count, net->lossp, false, 0.f);
  std::vector<double> lp(blocks);
  HCHECK(hipMemcpyAsync(lp.data(), net->lossp, blocks * sizeof(double), hipMemcpyDeviceToHost, net->stream));
  if (grad) HCHECK(hipMemcpyAsync(grad, net->grad, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, net->stream));
  HCHECK(hipStreamSynchronize(net->stream));
  if (loss) *loss = std::accumulate(lp.begin(), lp.end(), 0.0) / count;
  MLP_END
}

int hcc_mlp_train(hcc_mlp* net, int epochs, int batch, float lr, float lr_final, unsigned seed, double* epoch_loss) {
  if (net == nullptr || epochs < 1 || batch < 1 || !(lr > 0.f) || lr_final < 0.f) return HCC_E_ARG;
  if (net->count < 1) {
    net->err = "no training samples";
    return HCC_E_STATE;
  }
  MLP_BEGIN
  const long ns = net->count;
  const int nb = (int)std::min<long>(batch, ns);
  const long steps = (ns + nb - 1) / nb;
  const int bpb = (nb + 255) / 256;
  ensure_batch(net, nb);
  ensure_loss(net, steps * bpb);
  const long total = steps * (long)epochs;
  std::vector<double> lp((size_t)steps * bpb);
  unsigned h = seed * 2654435761u + 12345u;
  for (int ep = 0; ep < epochs; ++ep) {
    // row permutation (q * pa + pc) mod ns with gcd(pa, ns) = 1
    h = h * 1664525u + 1013904223u;
    long pa = ns > 1 ? 1 + ((long)(0.6180339887 * ns) + (long)(h % 997u)) % (ns - 1) : 1;
    while (gcd_l(pa, ns) != 1) pa = pa + 1 < ns ? pa + 1 : 1;
    h = h * 1664525u + 1013904223u;
    const long pc = (long)(h % (unsigned long)ns);
    for (long k = 0; k < steps; ++k) {
      const long base = k * nb;
      const int cnt = (int)std::min<long>(nb, ns - base);
      const double frac = (double)(ep * steps + k) / (double)std::max<long>(1, total - 1);
      const float lrk = (float)(lr_final + 0.5 * (lr - lr_final) * (1.0 + std::cos(3.14159265358979 * frac)));
      batch_step(net, ns, pa, pc, base, cnt, net->lossp + k * bpb, true, lrk);
    }
    HCHECK(hipMemcpyAsync(lp.data(), net->lossp, lp.size() * sizeof(double), hipMemcpyDeviceToHost, net->stream));
    HCHECK(hipStreamSynchronize(net->stream));
    if (epoch_loss) epoch_loss[ep] = std::accumulate(lp.begin(), lp.end(), 0.0) / (double)ns;
  }
  MLP_END
}

int hcc_mlp_evaluate(hcc_mlp* net, double* loss) {
  if (net == nullptr || loss == nullptr) return HCC_E_ARG;
  if (net->count < 1) {
    net->err = "no samples";
    return HCC_E_STATE;
  }
  MLP_BEGIN
  const long ns = net->count;
  const long blocks = (ns + 255) / 256;
  ensure_loss(net, blocks);
  hipLaunchKernelGGL(k_mlp_forward<false>, dim3((unsigned)blocks), dim3(256), 0, net->stream, net->w, net->rows, ns,
                     1L, 0L, 0L, (int)ns, (float*)nullptr, (float*)nullptr, net->lossp, (float*)nullptr);
  HCHECK(hipGetLastError());
  std::vector<double> lp(blocks);
  HCHECK(hipMemcpyAsync(lp.data(), net->lossp, blocks * sizeof(double), hipMemcpyDeviceToHost, net->stream));
  HCHECK(hipStreamSynchronize(net->stream));
  *loss = std::accumulate(lp.begin(), lp.end(), 0.0) / (double)ns;
  MLP_END
}

int hcc_mlp_predict(hcc_mlp* net, const float* features, long count, float* y) {
  if (net == nullptr || features == nullptr || y == nullptr || count < 1 || count > (1L << 30)) return HCC_E_ARG;
  MLP_BEGIN
  float* rows = nullptr;
  float* yd = nullptr;
  dalloc(&rows, (size_t)count * kMlpRow);
  dalloc(&yd, (size_t)count);
  std::vector<float> h((size_t)count * kMlpRow, 0.f);
  for (long r = 0; r < count; ++r)
    for (int q = 0; q < kMlpF; ++q) h[(size_t)r * kMlpRow + q] = features[(size_t)r * kMlpF + q];
  HCHECK(hipMemcpy(rows, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  const long blocks = (count + 255) / 256;
  double* lp = nullptr;
  dalloc(&lp, (size_t)blocks);
  hipLaunchKernelGGL(k_mlp_forward<false>, dim3((unsigned)blocks), dim3(256), 0, net->stream, net->w, rows, count,
                     1L, 0L, 0L, (int)count, (float*)nullptr, (float*)nullptr, lp, yd);
  HCHECK(hipGetLastError());
  HCHECK(hipStreamSynchronize(net->stream));
  HCHECK(hipMemcpy(y, yd, count * sizeof(float), hipMemcpyDeviceToHost));
  HCHECK(hipFree(rows));
  HCHECK(hipFree(yd));
  HCHECK(hipFree(lp));
  MLP_END
}

}  // extern "C"
