// Host driver of libhelios_ccsd: device arenas for a (ragged) batch of
// molecules, the table of grouped GEMMs that make up one CCSD iteration, the
// iteration sequence, and the C ABI of include/helios_ccsd.h.
//
// The residual is the biorthogonal CCSD residual Omega of the T1-transformed
// PPP Hamiltonian.  With ZDO integrals every dressed integral block is one
// GEMM over sites,
//   g~(pq|rs) = sum_mu P[mu][pq] Q[mu][rs],   P[mu][p*N+q] = X[mu][p] Y[mu][q],
//   Q = gamma P,   X = C(1 - t1^T) (bra), Y = C(1 + t1) (ket),
// and, gamma being symmetric, also sum_mu Q[mu][pq] P[mu][rs]; the transpose
// PT[mu][q*N+p] = P[mu][p*N+q] lets a block's GEMM take its output's unit
// stride along the column index, where the stores coalesce.  The CCSD
// contractions (particle-particle ladder, hole-hole ladder, the two ring
// intermediates and their contractions, the Fock-like intermediates) are
// GEMMs over composite indices.  The reference's contravariant residuals
// (proj-ccsd-spat-hf-1/2.f90) are d1 = Omega1 and d2 = 2 Omega2 - Omega2(a<->b);
// the update follows stepCC (mod_cc.f90:987-1017) and the DIIS follows
// pushCCVectors/newCCVectors (mod_cc.f90:1390-1423, 1514-1555) with a ring
// buffer in place of the history shift.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <set>
#include <memory>
#include <mutex>
#include <atomic>
#include <array>
#include <string>
#include <vector>

#include "../../include/helios_ccsd.h"
#include "hcc_batch.h"
#include "hcc_ops.h"
#include "hcc_mlp.h"
#include "hcc_scf.h"
#include "hcc_triples.h"
#include "hcc_lr.h"

namespace hcc {

// ---------------------------------------------------------------------------
// operand helpers
// ---------------------------------------------------------------------------
static Operand rm(double* p, int ld) { return make_operand(p, 1, 1, ld, 0, 1, 0); }
static Operand trn(double* p, int ld) { return make_operand(p, 1, 1, 1, 0, ld, 0); }
// composite: row = (r / rr, r % rr) with strides (srh, srl), col likewise
static Operand cmp(double* p, int rr, int srh, int srl, int rc, int sch, int scl) {
  if (rr < 1 || rc < 1 || rr >= kMaxRadix || rc >= kMaxRadix) throw std::string("composite radix out of range");
  return make_operand(p, rr, rc, srh, srl, sch, scl);
}

static bool unit_cols(const Operand& o) { return (o.rc == 1 ? o.sch : o.scl) == 1; }
static bool unit_rows(const Operand& o) { return (o.rr == 1 ? o.srh : o.srl) == 1; }

struct GemmTable {
  GemmProb* d = nullptr;
  int nprob = 0;
  int tiles = 0;
  int* tile_prob = nullptr;    // problem index of every tile
  int* red_prefix = nullptr;   // split-K reduction blocks per problem (null: no split)
  int* red_bprob = nullptr;    // problem index of every reduction block
  int red_blocks = 0;
  int tile = 64;               // output tile edge (64 or 32)
  int alt = -1;                // second part of the same contraction (other tile edge), index in Batch::alt
  int flags = 3;               // a_kfast | b_nfast << 1 | (dA != 0) << 2 | (dB != 0) << 3 | (K <= 64) << 4
  // compaction to the molecules still running (compact_tables): per problem
  // its molecule, tiles and split-K reduction blocks; the compacted copies of
  // the problems, the tile and reduction-block maps, and the uploaded lists
  std::vector<int> hmol, htiles, hred, hbuf;
  GemmProb* cd = nullptr;
  int *ctp = nullptr, *cbp = nullptr, *cbuf = nullptr;
  int ctiles = -1;             // tiles of the compacted copy; -1: not compacted
  int cred = 0;
};

// what a launch of a table uses: its compacted copy when there is one
struct TableView {
  const GemmProb* d;
  const int* tp;
  int tiles;
  const int* rp;
  const int* rb;
  int red;
};
static TableView view(const GemmTable& t) {
  if (t.ctiles < 0) return {t.d, t.tile_prob, t.tiles, t.red_prefix, t.red_bprob, t.red_blocks};
  return {t.cd, t.ctp, t.ctiles, t.cbuf + 2 * t.nprob + 1, t.cbp, t.cred};
}

// Tile kinds of the FP64 GEMM: 64 (64x64, 4 waves), 32 (32x32, 1 wave) and
// 128 (128x32, 4 waves stacked along M, for outputs at most 32 wide); of the
// BF16 correction GEMM: 64 (64x64, 4 waves), 32 (32x32, 2 waves) and 256
// (128x128, 8 waves).
static int tile_m(int kind) { return kind == 128 || kind == 256 ? 128 : kind; }
static int tile_n(int kind) { return kind == 128 ? 32 : (kind == 256 ? 128 : kind); }

// Tile kind of a correction-form problem.  128x128 tiles halve the operand
// staging per product and win on long contractions wherever their padding
// stays close to that of 64x64 tiles; short contractions (K < 128) are bound
// by their output stores and stay on 64x64 tiles, in one launch.  Every lane
// stages the same number of elements whatever the tile, so 32x32 tiles (half
// the staging per product of 64x64 ones) take the narrow outputs whose
// 64x64 tiles would be mostly padding.
static int choose_tile_dual(const GemmProb& p) {
  const double t64 = (double)((p.M + 63) / 64) * ((p.N + 63) / 64) * 64 * 64;
  const double t32 = (double)((p.M + 31) / 32) * ((p.N + 31) / 32) * 32 * 32;
  const double t128 = (double)((p.M + 127) / 128) * ((p.N + 127) / 128) * 128 * 128;
  if (p.K >= 128 && p.M >= 128 && p.N >= 128 && t128 <= 1.3 * t64) return 256;
  return 2.0 * t32 < t64 ? 32 : 64;
}

// Split K when a problem has few output tiles and a long K: aim for >= 16
// work items per problem with K chunks of at least 128.
static void choose_split(GemmProb& p, int ts) {
  const int tmn = ((p.M + tile_m(ts) - 1) / tile_m(ts)) * ((p.N + tile_n(ts) - 1) / tile_n(ts));
  p.ksplit = 1;
  p.kchunk = p.K;
  if (tmn >= 16 || p.K < 512) return;
  int want = (16 + tmn - 1) / tmn;
  int maxs = p.K / 128;
  int ks = std::min(want, std::max(1, maxs));
  ks = std::min(ks, 64);
  if (ks <= 1) return;
  int chunk = (p.K + ks - 1) / ks;
  chunk = (chunk + 31) / 32 * 32;      // multiple of both K tiles (16 for FP64, 32 for BF16)
  p.kchunk = chunk;
  p.ksplit = (p.K + chunk - 1) / chunk;
  if (p.ksplit <= 1) {
    p.ksplit = 1;
    p.kchunk = p.K;
  }
}

struct DevBuf {
  double* base = nullptr;
  std::vector<long> off;   // per-molecule offsets
  long total = 0;
};

// ---------------------------------------------------------------------------
// batch
// ---------------------------------------------------------------------------
struct Batch {
  int nm = 0;
  int m = 20;
  int inner_m = 20;   // mixed solver: inner DIIS depth (<= m; the ring storage is sized for m)
  std::vector<int> n, o, v;
  std::vector<double> eref;
  std::vector<Mol> hm;
  Mol* dm = nullptr;
  State* dst = nullptr;
  int* dact = nullptr;
  std::vector<void*> allocs;
  Ragged rgNN, rgN3, rgT2, rgT2S, rgOV, rgXX, rgND, rgN, rgP2, rgPairL, rgPairBlk, rgOcc;
  bool guess_large = false;   // some molecule has v > 32 (pairwise learned-guess kernel)
  Ragged rgT2x, rgT2Sx;   // rgT2, rgT2S with kEwItems elements per thread
  double *part1 = nullptr, *part2 = nullptr, *partE = nullptr, *partD = nullptr;
  std::vector<GemmTable> pre, it;
  std::vector<GemmTable> alt;   // second (other tile edge) parts of split contraction tables
  // mixed precision: BF16 copies of the outer-point operands (one per arena)
  struct Shadow {
    double* base;      // FP64 arena (molecule 0's block)
    long count;        // elements spanned by the arena
    __bf16* copy;      // BF16 copy, same element offsets
    bool variable;     // changes at every outer point
    bool used;         // an operand of some correction GEMM (refreshed only then)
    // per molecule: its block of the arena (offset from base, length), on the
    // host and the device, and the block prefix over the running molecules
    std::vector<long> hoff, hlen;
    long *doff = nullptr, *dlen = nullptr;
    std::vector<int> hpre;
    int* dpre = nullptr;
  };
  std::vector<Shadow> shadows;
  // measured tile/split choices per (contraction, n_occ, n_virt); null: cost heuristic
  const std::map<std::array<int, 3>, std::array<int, 2>>* tuning = nullptr;
  double *d1out = nullptr, *d2out = nullptr;
  // mixed precision: correction buffers and correction-form GEMM tables
  bool mixed = false;
  bool mixed_bf16 = true;   // BF16 correction GEMMs (false: FP64 exact mode, 64x64 tiles only)
  std::vector<MolD> hd;   // FP32 corrections (variable buffers) + shared FP64 inputs
  MolD* ddm = nullptr;
  std::vector<std::function<GemmProb(const Mol&)>> ops;
  std::vector<GemmTable> dit;
  // (T): undressed integral blocks (xp|yd), (ql|rz), their GEMM tables and the work items
  std::vector<double*> tA, tB;
  std::vector<GemmTable> tpre;
  TripMol* dtm = nullptr;
  long* dtprefix = nullptr;
  int2* dtri = nullptr;
  int* dtsave = nullptr;
  double *dtw = nullptr, *dtpart = nullptr, *dtout = nullptr;
  // linear response: concatenated trial/sigma vectors, amplitude backups, and
  // the per-molecule pointers (r1, r2, s1, s2; nm each) the sigma kernels read
  double *dlr_r1 = nullptr, *dlr_r2 = nullptr, *dlr_s1 = nullptr, *dlr_s2 = nullptr;
  double *dlr_t1b = nullptr, *dlr_t2b = nullptr;
  double** dlr_pc = nullptr;   // pointers into the concatenated buffers
  double** dlr_ps = nullptr;   // pointers into the vector space (rewritten per call)
  long lr_n1 = 0, lr_n2 = 0, lr_span1 = 0, lr_span2 = 0;
  // device-resident vector space of the eigensolver (hcc_lr.h): lr_slots
  // vectors of lr_len[k] doubles per molecule from element lr_base[k]; own
  // allocations (not in `allocs`), replaced when the slot count changes
  // running molecules the FP64 (it) and correction (dit) tables are compacted
  // to; 0: not compacted
  int it_live = 0, dit_live = 0;
  // the elementwise grids, compacted to the running molecules likewise
  // (compact_raggeds): each grid, its full form, host block prefix, the
  // compacted prefix and block-to-molecule map
  struct RgComp {
    Ragged* rg;
    Ragged full;
    std::vector<int> pre, hbuf;
    int *cpre = nullptr, *cbm = nullptr;
  };
  std::vector<RgComp> rgc;
  int rg_live = 0;
  // molecules whose amplitude arenas hold meaningful amplitudes (a guess, a
  // solve or hcc_batch_set_amplitudes): HCC_GUESS_CURRENT needs all of them
  std::vector<char> amp_valid;
  int lr_slots = 0;            // largest slot count of a molecule; 0: no space
  std::vector<int> lr_nslot;   // slots of each molecule
  double* dlr_space = nullptr;
  std::vector<long> lr_len, lr_base;
  long *dlr_len = nullptr, *dlr_base = nullptr;
  char* dlr_stage = nullptr;   // job lists and results of the space calls
  size_t lr_stage_cap = 0;
  long ttotal = 0, twstride = 0;
  int tgrid = 0, tgrid_mfma = 0;
  bool tglobal = false;

  // device memory: hipMalloc per buffer, or -- when `arena` is set (the
  // no-allocation entry point) -- consecutive 256-byte aligned pieces of a
  // caller-owned region of arena_cap bytes, never freed here
  char* arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;

  ~Batch() { release(); }
  void release() {
    if (arena == nullptr)
      for (void* p : allocs) (void)hipFree(p);
    allocs.clear();
    if (dlr_space) (void)hipFree(dlr_space);
    if (dlr_stage) (void)hipFree(dlr_stage);
    dlr_space = nullptr;
    dlr_stage = nullptr;
    lr_slots = 0;
    lr_stage_cap = 0;
    arena_used = 0;
    dm = nullptr;
    dst = nullptr;
    dact = nullptr;
  }
  template <class T>
  T* dalloc(size_t count) {
    void* p = nullptr;
    if (count == 0) count = 1;
    if (arena != nullptr) {
      const size_t bytes = (count * sizeof(T) + 255) / 256 * 256;
      if (bytes > arena_cap - arena_used)
        throw std::string("molecule too large for the no-allocation scratch region (") +
            std::to_string(arena_cap >> 20) + " MiB): needs more than " + std::to_string((arena_used + bytes) >> 20) +
            " MiB";
      p = arena + arena_used;
      arena_used += bytes;
      return static_cast<T*>(p);
    }
    HCHECK(hipMalloc(&p, count * sizeof(T)));
    allocs.push_back(p);
    return static_cast<T*>(p);
  }
};

static long sz_nn(int n, int, int) { return (long)n * n; }
static long sz_n3(int n, int, int) { return (long)n * n * n; }
static long sz_t2(int, int o, int v) { return (long)o * o * v * v; }
static long sz_ov(int, int o, int v) { return (long)o * v; }

static Ragged make_ragged(Batch& b, const std::function<long(int, int, int)>& size, int items = 1) {
  std::vector<int> pre(b.nm + 1, 0);
  for (int k = 0; k < b.nm; ++k) {
    long s = size(b.n[k], b.o[k], b.v[k]);
    long per = (long)kEwThreads * items;
    long blocks = std::max<long>(1, (s + per - 1) / per);
    pre[k + 1] = pre[k] + (int)blocks;
  }
  Ragged r;
  r.prefix = b.dalloc<int>(pre.size());
  HCHECK(hipMemcpy(r.prefix, pre.data(), pre.size() * sizeof(int), hipMemcpyHostToDevice));
  r.nblocks = pre.back();
  std::vector<int> bm(r.nblocks);
  for (int k = 0; k < b.nm; ++k)
    for (int q = pre[k]; q < pre[k + 1]; ++q) bm[q] = k;
  r.bmol = b.dalloc<int>(bm.size());
  HCHECK(hipMemcpy(r.bmol, bm.data(), bm.size() * sizeof(int), hipMemcpyHostToDevice));
  return r;
}

// Tile edge of one problem: 32x32 tiles (one wave) only when the padding of
// 64x64 tiles would cost more than the (measured, about 2.5x) lower per-flop
// efficiency of the small tile.
static int choose_tile(const GemmProb& p) {
  const double t64 = (double)((p.M + 63) / 64) * ((p.N + 63) / 64) * 64 * 64;
  const double t32 = (double)((p.M + 31) / 32) * ((p.N + 31) / 32) * 32 * 32;
  const double t128 = (double)((p.M + 127) / 128) * ((p.N + 31) / 32) * 128 * 32;
  if (p.N <= 32 && p.M >= 128 && t128 < t64 && t128 < 2.5 * t32) return 128;
  return 2.5 * t32 < t64 ? 32 : 64;
}

// K split of a problem into `ks` chunks (multiples of 32, both K tiles)
static void force_split(GemmProb& p, int ks) {
  p.ksplit = 1;
  p.kchunk = p.K;
  if (ks <= 1) return;
  int chunk = (p.K + ks - 1) / ks;
  chunk = (chunk + 31) / 32 * 32;
  p.kchunk = chunk;
  p.ksplit = (p.K + chunk - 1) / chunk;
  if (p.ksplit <= 1) {
    p.ksplit = 1;
    p.kchunk = p.K;
  }
}

// One launchable table from problems that share the tile edge `ts` (their
// ksplit / kchunk already chosen).
static GemmTable build_group(Batch& b, std::vector<GemmProb> probs, int ts) {
  int tiles = 0;
  long ws_total = 0;
  for (GemmProb& p : probs) {
    p.tiles_n = (p.N + tile_n(ts) - 1) / tile_n(ts);
    const int tm = (p.M + tile_m(ts) - 1) / tile_m(ts);
    p.tile_begin = tiles;
    tiles += tm * p.tiles_n * p.ksplit;
    p.a_kfast = unit_cols(p.A) ? 1 : (unit_rows(p.A) ? 0 : 1);
    p.b_nfast = unit_cols(p.B) ? 1 : (unit_rows(p.B) ? 0 : 1);
    if (p.ksplit > 1) ws_total += (long)p.ksplit * p.M * p.N;
  }
  GemmTable t;
  t.nprob = (int)probs.size();
  t.tiles = tiles;
  t.tile = ts;
  for (const GemmProb& p : probs) {
    t.hmol.push_back(p.mol);
    t.htiles.push_back(((p.M + tile_m(ts) - 1) / tile_m(ts)) * p.tiles_n * p.ksplit);
    t.hred.push_back(p.ksplit > 1 ? (int)(((long)p.M * p.N + 255) / 256) : 0);
  }
  if (!probs.empty()) {
    // the staging layout is a performance choice only (addressing is general):
    // take the largest problem's unit-stride flags for the whole table
    size_t big = 0;
    for (size_t q = 1; q < probs.size(); ++q)
      if ((double)probs[q].M * probs[q].N * probs[q].K > (double)probs[big].M * probs[big].N * probs[big].K) big = q;
    t.flags = probs[big].a_kfast | (probs[big].b_nfast << 1) | ((probs[big].dA != nullptr) << 2) |
              ((probs[big].dB != nullptr) << 3);
    // short-K variant of the correction GEMM when no problem has more than two 32-deep slabs
    int kmax = 0;
    for (const GemmProb& p : probs) kmax = std::max(kmax, p.kchunk);
    if (kmax <= 64) t.flags |= 16;
  }
  if (ws_total > 0) {
    double* ws = b.dalloc<double>(ws_total);
    long off = 0;
    std::vector<int> pre(probs.size() + 1, 0);
    for (size_t q = 0; q < probs.size(); ++q) {
      GemmProb& p = probs[q];
      int blocks = 0;
      if (p.ksplit > 1) {
        p.ws = ws + off;
        off += (long)p.ksplit * p.M * p.N;
        blocks = (int)(((long)p.M * p.N + 255) / 256);
      }
      pre[q + 1] = pre[q] + blocks;
    }
    t.red_prefix = b.dalloc<int>(pre.size());
    HCHECK(hipMemcpy(t.red_prefix, pre.data(), pre.size() * sizeof(int), hipMemcpyHostToDevice));
    t.red_blocks = pre.back();
    std::vector<int> bp(t.red_blocks);
    for (size_t q = 0; q < probs.size(); ++q)
      for (int x = pre[q]; x < pre[q + 1]; ++x) bp[x] = (int)q;
    t.red_bprob = b.dalloc<int>(std::max<size_t>(1, bp.size()));
    if (!bp.empty())
      HCHECK(hipMemcpy(t.red_bprob, bp.data(), bp.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  if (tiles > 0) {
    std::vector<int> tp(tiles);
    for (size_t q = 0; q < probs.size(); ++q) {
      const int end = q + 1 < probs.size() ? probs[q + 1].tile_begin : tiles;
      for (int x = probs[q].tile_begin; x < end; ++x) tp[x] = (int)q;
    }
    t.tile_prob = b.dalloc<int>(tp.size());
    HCHECK(hipMemcpy(t.tile_prob, tp.data(), tp.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  if (t.nprob) {
    t.d = b.dalloc<GemmProb>(probs.size());
    HCHECK(hipMemcpy(t.d, probs.data(), probs.size() * sizeof(GemmProb), hipMemcpyHostToDevice));
  }
  return t;
}

// Table of one contraction over every molecule.  Each problem gets its tile
// edge and K split from the measured tuning of its (op, n_occ, n_virt) when
// there is one, else from the cost heuristic (choose_tile / choose_split);
// problems with 32x32 tiles go to a second launch (Batch::alt).
// force_tile / force_ks (> 0) override both for every problem.
static GemmTable make_table(Batch& b, const std::function<GemmProb(int)>& build, bool allow_split = true,
                            int force_tile = 0, int op = -1, int force_ks = 0, bool dual = false) {
  const int kinds[4] = {64, 32, 128, 256};
  std::vector<GemmProb> groups[4];
  for (int k = 0; k < b.nm; ++k) {
    GemmProb p = build(k);
    if (p.M <= 0 || p.N <= 0) continue;
    if (p.M >= kMaxIndex || p.N >= kMaxIndex || p.K >= kMaxIndex) throw std::string("GEMM dimension out of range");
    p.mol = k;
    int ts = 0, ks = -1;
    if (force_tile) {
      ts = force_tile;
      if (force_ks > 0) ks = force_ks;
    } else if (b.tuning && op >= 0) {
      auto it = b.tuning->find({op, b.o[k], b.v[k]});
      if (it != b.tuning->end()) {
        ts = it->second[0];
        ks = it->second[1];
      }
    }
    if (dual) {
      if (ts != 32 && ts != 64 && ts != 256) ts = choose_tile_dual(p);
    } else if (ts != 32 && ts != 64 && ts != 128) {
      ts = choose_tile(p);
    }
    if (!allow_split) force_split(p, 1);
    else if (ks >= 1) force_split(p, ks);
    else choose_split(p, ts);
    groups[ts == 64 ? 0 : (ts == 32 ? 1 : (ts == 128 ? 2 : 3))].push_back(p);
  }
  // the first non-empty group is the table; the others hang off it through alt links
  GemmTable head;
  bool have = false;
  int last = -1;   // index in b.alt of the last linked part, -1: the head
  for (int g = 0; g < 4; ++g) {
    if (groups[g].empty()) continue;
    GemmTable t = build_group(b, groups[g], kinds[g]);
    if (!have) {
      head = t;
      have = true;
      continue;
    }
    b.alt.push_back(t);
    const int idx = (int)b.alt.size() - 1;
    if (last < 0) head.alt = idx;
    else b.alt[last].alt = idx;
    last = idx;
  }
  if (!have) head = build_group(b, {}, 64);
  return head;
}

static GemmProb prob(Operand A, Operand B, Operand C, int M, int N, int K, double alpha, double beta) {
  GemmProb p{};
  p.A = A;
  p.B = B;
  p.C = C;
  p.M = M;
  p.N = N;
  p.K = K;
  p.alpha = alpha;
  p.beta = beta;
  return p;
}

// Buffers that change with the amplitudes (the ones that get a correction twin
// in the mixed-precision solver), with their per-molecule sizes.
struct FieldDef {
  double* Mol::*member;
  float* MolD::*dmember;   // its FP32 correction twin (variable fields)
  std::function<long(int, int, int)> size;
};

#define HCC_VF(f, sz) {&Mol::f, &MolD::f, sz}
static const std::vector<FieldDef>& variable_fields() {
  static const std::vector<FieldDef> f = {
      HCC_VF(Xf, sz_nn), HCC_VF(Yf, sz_nn), HCC_VF(Mf, sz_nn), HCC_VF(Ft, sz_nn), HCC_VF(tmp, sz_nn),
      HCC_VF(P, sz_n3), HCC_VF(PT, sz_n3), HCC_VF(Q, sz_n3), HCC_VF(t1, sz_ov), HCC_VF(w1, sz_ov), HCC_VF(t2, sz_t2),
      HCC_VF(u, sz_t2), HCC_VF(w2, sz_t2), HCC_VF(x, sz_t2), HCC_VF(Zc, sz_t2), HCC_VF(WC, sz_t2),
      HCC_VF(WD, sz_t2),
      HCC_VF(G4, [](int, int, int v) { return (long)v * v * v * v; }),
      HCC_VF(Wkl, [](int, int o, int) { return (long)o * o * o * o; }),
      HCC_VF(gvvov, [](int, int o, int v) { return (long)o * v * v * v; }),
      HCC_VF(gooov, [](int, int o, int v) { return (long)o * o * o * v; }),
      HCC_VF(Xvv, [](int, int, int v) { return (long)v * v; }),
      HCC_VF(Xoo, [](int, int o, int) { return (long)o * o; }),
      HCC_VF(dvec, [](int n, int, int) { return 2L * n; }),
  };
  return f;
}
#undef HCC_VF

// Operand arenas that never change during a solve (inputs, undressed integrals).
static const std::vector<FieldDef>& constant_fields() {
  static const std::vector<FieldDef> f = {
      {&Mol::C, nullptr, sz_nn}, {&Mol::G, nullptr, sz_nn}, {&Mol::F, nullptr, sz_nn},
      {&Mol::Fao, nullptr, sz_nn}, {&Mol::Gklcd, nullptr, sz_t2}, {&Mol::Lldkc, nullptr, sz_t2},
      {&Mol::Lt, nullptr, sz_t2},
  };
  return f;
}

// BF16 copy of an outer-point operand element, or null when p lies in no arena
static const __bf16* shadow_of(Batch& b, const double* p) {
  for (auto& sh : b.shadows)
    if (p >= sh.base && p < sh.base + sh.count) {
      sh.used = true;
      return sh.copy + (p - sh.base);
    }
  return nullptr;
}

// Correction twin of an operand pointer of molecule k, or null for a constant
// input (MO coefficients, gamma, Fock, undressed integrals).
static float* delta_of(const Batch& b, int k, const double* p) {
  const Mol& H = b.hm[k];
  const MolD& D = b.hd[k];
  for (const auto& f : variable_fields()) {
    const double* base = H.*(f.member);
    const long size = f.size(H.n, H.o, H.v);
    if (p >= base && p < base + size) return D.*(f.dmember) + (p - base);
  }
  return nullptr;
}

// Allocates all arenas and fills the per-molecule descriptors.
static void allocate(Batch& b, bool keep_mp2) {
  auto arena = [&](const std::function<long(int, int, int)>& size) {
    long total = 0;
    std::vector<long> off(b.nm);
    for (int k = 0; k < b.nm; ++k) {
      off[k] = total;
      total += (size(b.n[k], b.o[k], b.v[k]) + 31) / 32 * 32;   // 256-byte aligned blocks
    }
    double* base = b.dalloc<double>(total);
    std::vector<double*> ptr(b.nm);
    for (int k = 0; k < b.nm; ++k) ptr[k] = base + off[k];
    return ptr;
  };
  const int m = b.m;
  auto C = arena(sz_nn), G = arena(sz_nn), F = arena(sz_nn), Fao = arena(sz_nn);
  auto Xf = arena(sz_nn), Yf = arena(sz_nn), Mf = arena(sz_nn), Ft = arena(sz_nn), tmp = arena(sz_nn);
  auto P = arena(sz_n3), PT = arena(sz_n3), Q = arena(sz_n3);
  auto t1 = arena(sz_ov), w1 = arena(sz_ov);
  auto t2 = arena(sz_t2), u = arena(sz_t2), w2 = arena(sz_t2), x = arena(sz_t2), Zc = arena(sz_t2);
  auto WC = arena(sz_t2), WD = arena(sz_t2), Gklcd = arena(sz_t2), Lldkc = arena(sz_t2), Lt = arena(sz_t2);
  auto G4 = arena([](int, int, int v) { return (long)v * v * v * v; });
  auto Wkl = arena([](int, int o, int) { return (long)o * o * o * o; });
  auto gvvov = arena([](int, int o, int v) { return (long)o * v * v * v; });
  auto gooov = arena([](int, int o, int v) { return (long)o * o * o * v; });
  auto Xvv = arena([](int, int, int v) { return (long)v * v; });
  auto dvec = arena([](int n, int, int) { return 2L * n; });
  auto Xoo = arena([](int, int o, int) { return (long)o * o; });
  auto Ov = arena(sz_nn);
  auto ndsize = [m](int, int o, int v) {
    long ov = (long)o * v;
    return (long)m * (ov + ov * (ov + 1) / 2);
  };
  auto st = arena(ndsize), sd = arena(ndsize);
  auto Bm = arena([m](int, int, int) { return (long)m * m; });
  std::vector<double*> mp2(b.nm, nullptr);
  if (keep_mp2) mp2 = arena(sz_t2);
  b.hm.resize(b.nm);
  for (int k = 0; k < b.nm; ++k) {
    Mol& M = b.hm[k];
    memset(&M, 0, sizeof(M));
    M.n = b.n[k];
    M.o = b.o[k];
    M.v = b.v[k];
    M.m = m;
    const long ov = (long)M.o * M.v;
    M.nd = (int)(ov + ov * (ov + 1) / 2);
    M.eref = b.eref[k];
    M.C = C[k]; M.G = G[k]; M.F = F[k]; M.Fao = Fao[k];
    M.Xf = Xf[k]; M.Yf = Yf[k]; M.Mf = Mf[k]; M.P = P[k]; M.PT = PT[k]; M.Q = Q[k]; M.Ft = Ft[k]; M.tmp = tmp[k];
    M.t1 = t1[k]; M.t2 = t2[k]; M.u = u[k]; M.w1 = w1[k]; M.w2 = w2[k]; M.x = x[k]; M.Zc = Zc[k];
    M.WC = WC[k]; M.WD = WD[k];
    M.G4 = G4[k]; M.Wkl = Wkl[k]; M.Gklcd = Gklcd[k]; M.Lldkc = Lldkc[k]; M.Lt = Lt[k];
    M.gvvov = gvvov[k]; M.gooov = gooov[k]; M.Xvv = Xvv[k]; M.Xoo = Xoo[k]; M.dvec = dvec[k];
    M.st = st[k]; M.sd = sd[k]; M.Bm = Bm[k];
    M.Ov = Ov[k];
    M.t2mp2 = mp2[k];
  }
  b.dm = b.dalloc<Mol>(b.nm);
  HCHECK(hipMemcpy(b.dm, b.hm.data(), b.nm * sizeof(Mol), hipMemcpyHostToDevice));
  if (b.mixed) {
    // corrections: FP32 twins of the variable buffers; inputs, constants and the
    // DIIS storage are shared with the outer point
    b.hd.resize(b.nm);
    for (int k = 0; k < b.nm; ++k) {
      const Mol& H = b.hm[k];
      MolD& D = b.hd[k];
      memset(&D, 0, sizeof(D));
      D.n = H.n; D.o = H.o; D.v = H.v; D.m = b.inner_m; D.nd = H.nd; D.eref = H.eref;
      D.C = H.C; D.G = H.G; D.F = H.F; D.Fao = H.Fao;
      D.Gklcd = H.Gklcd; D.Lldkc = H.Lldkc; D.Lt = H.Lt;
      // FP32 DIIS vectors in the (otherwise unused) FP64 ring of the molecule
      D.st = reinterpret_cast<float*>(H.st);
      D.sd = reinterpret_cast<float*>(H.sd);
      D.Bm = H.Bm; D.t2mp2 = H.t2mp2; D.Ov = H.Ov;
    }
    for (const auto& f : variable_fields()) {
      long total = 0;
      std::vector<long> off(b.nm);
      for (int k = 0; k < b.nm; ++k) {
        off[k] = total;
        total += (f.size(b.n[k], b.o[k], b.v[k]) + 63) / 64 * 64;   // 256-byte aligned blocks
      }
      float* base = b.dalloc<float>(total);
      for (int k = 0; k < b.nm; ++k) b.hd[k].*(f.dmember) = base + off[k];
    }
    b.ddm = b.dalloc<MolD>(b.nm);
    HCHECK(hipMemcpy(b.ddm, b.hd.data(), b.nm * sizeof(MolD), hipMemcpyHostToDevice));
    auto add_shadow = [&](const FieldDef& f, bool variable) {
      const Mol& first = b.hm[0];
      const Mol& last = b.hm[b.nm - 1];
      double* base = first.*(f.member);
      const long count = (last.*(f.member) - base) + f.size(last.n, last.o, last.v);
      Batch::Shadow sh{base, count, b.dalloc<__bf16>(count), variable, false};
      for (int k = 0; k < b.nm; ++k) {
        sh.hoff.push_back(b.hm[k].*(f.member) - base);
        sh.hlen.push_back(f.size(b.n[k], b.o[k], b.v[k]));
      }
      sh.doff = b.dalloc<long>(b.nm);
      sh.dlen = b.dalloc<long>(b.nm);
      sh.dpre = b.dalloc<int>(b.nm + 1);
      HCHECK(hipMemcpy(sh.doff, sh.hoff.data(), b.nm * sizeof(long), hipMemcpyHostToDevice));
      HCHECK(hipMemcpy(sh.dlen, sh.hlen.data(), b.nm * sizeof(long), hipMemcpyHostToDevice));
      b.shadows.push_back(std::move(sh));
    };
    for (const auto& f : variable_fields()) add_shadow(f, true);
    for (const auto& f : constant_fields()) add_shadow(f, false);
  }
  b.dst = b.dalloc<State>(b.nm);
  HCHECK(hipMemset(b.dst, 0, b.nm * sizeof(State)));
  b.dact = b.dalloc<int>(b.nm);
  HCHECK(hipMemset(b.dact, 0, b.nm * sizeof(int)));

  b.rgNN = make_ragged(b, sz_nn);
  b.rgN3 = make_ragged(b, sz_n3, kPItems);
  b.rgT2 = make_ragged(b, sz_t2);
  b.rgOcc = make_ragged(b, [](int, int o, int) { return (long)o * kE2Split * kEwThreads; });   // kE2Split blocks per i
  // learned guess: 1024 amplitude pairs (i a) <= (j b) per workgroup
  b.rgPairL = make_ragged(b, [](int, int o, int v) { const long ne = (long)o * v; return ne * (ne + 1) / 2; }, 4);
  // consecutive occupied block pairs i <= j per workgroup
  b.rgPairBlk = make_ragged(b, [](int, int o, int v) {
    const int bpw = guess_pairs_per_workgroup(v);
    return (long)((o * (o + 1) / 2 + bpw - 1) / bpw) * kEwThreads;
  });
  b.guess_large = false;
  for (int k = 0; k < b.nm; ++k) b.guess_large = b.guess_large || b.v[k] > 32;
  b.rgT2S = make_ragged(b, [](int, int o, int v) { return (long)o * o * v * v + (long)o * v; });
  b.rgT2x = make_ragged(b, sz_t2, kEwItems);
  b.rgT2Sx = make_ragged(b, [](int, int o, int v) { return (long)o * o * v * v + (long)o * v; }, kEwItems);
  b.rgOV = make_ragged(b, sz_ov);
  b.rgXX = make_ragged(b, [](int, int o, int v) { return (long)v * v + (long)o * o; });
  b.rgN = make_ragged(b, [](int n, int, int) { return (long)n; });
  // DIIS extrapolation / shift grid: one thread per packed element (singles, then pairs mm <= nn)
  b.rgP2 = make_ragged(b, [](int, int o, int v) {
    const long ov = (long)o * v;
    return ov + ov * (ov + 1) / 2;
  });
  b.rgND = make_ragged(b, [](int, int o, int v) {
    long ov = (long)o * v;
    return ov + ov * (ov + 1) / 2;
  }, kDotItems);
  b.part1 = b.dalloc<double>(b.rgOV.nblocks);
  b.part2 = b.dalloc<double>(b.rgT2.nblocks);
  b.partE = b.dalloc<double>(2 * (long)b.rgT2S.nblocks);
  b.partD = b.dalloc<double>((long)b.rgND.nblocks * m);
  b.rgc.clear();
  b.rg_live = 0;
  for (Ragged* r : {&b.rgNN, &b.rgN3, &b.rgT2, &b.rgT2S, &b.rgOV, &b.rgXX, &b.rgND, &b.rgN, &b.rgP2, &b.rgPairL,
                    &b.rgPairBlk, &b.rgOcc, &b.rgT2x, &b.rgT2Sx}) {
    Batch::RgComp c;
    c.rg = r;
    c.full = *r;
    c.pre.resize(b.nm + 1);
    HCHECK(hipMemcpy(c.pre.data(), r->prefix, (b.nm + 1) * sizeof(int), hipMemcpyDeviceToHost));
    b.rgc.push_back(std::move(c));
  }
}

// ---------------------------------------------------------------------------
// GEMM tables (see the derivation in the file header; offsets per molecule)
// ---------------------------------------------------------------------------
static void build_tables(Batch& b) {
  std::vector<Mol>& H = b.hm;
  // --- precompute: Fao = C F C^T, Gklcd = (kc|ld) from undressed P, Q
  b.pre.clear();
  b.pre.push_back(make_table(b, [&](int k) {   // tmp = C F
    const Mol& M = H[k];
    const int N = M.n;
    return prob(rm(M.C, N), rm(M.F, N), rm(M.tmp, N), N, N, N, 1.0, 0.0);
  }));
  b.pre.push_back(make_table(b, [&](int k) {   // Fao = tmp C^T
    const Mol& M = H[k];
    const int N = M.n;
    return prob(rm(M.tmp, N), trn(M.C, N), rm(M.Fao, N), N, N, N, 1.0, 0.0);
  }));
  b.pre.push_back(make_table(b, [&](int k) {   // Gklcd[k][l][c][d] = (kc|ld)
    const Mol& M = H[k];
    const int N = M.n, o = M.o, v = M.v, N2 = N * N, ov = o * v;
    return prob(cmp(M.P + o, v, N, 1, 1, N2, 0), cmp(M.Q + o, 1, N2, 0, v, N, 1),
                cmp(M.Gklcd, v, ov * v, v, v, v * v, 1), ov, ov, N, 1.0, 0.0);
  }));

  // --- one iteration
  b.it.clear();
  b.ops.clear();
  auto add = [&](const std::function<GemmProb(const Mol&)>& f) {
    b.ops.push_back(f);
    b.it.push_back(make_table(b, [&](int k) { return f(H[k]); }, true, 0, (int)b.it.size()));
  };
  // g1: Q = gamma P
  add([](const Mol& M) {
    const int N = M.n, N2 = N * N;
    return prob(rm(M.G, N), rm(M.P, N2), rm(M.Q, N2), N, N2, N, 1.0, 0.0);
  });
  // g2: Ft = Xf^T (Mf Yf)
  add([](const Mol& M) {
    const int N = M.n;
    return prob(rm(M.Mf, N), rm(M.Yf, N), rm(M.tmp, N), N, N, N, 1.0, 0.0);
  });
  add([](const Mol& M) {
    const int N = M.n;
    return prob(trn(M.Xf, N), rm(M.tmp, N), rm(M.Ft, N), N, N, N, 1.0, 0.0);
  });
  // g3: gvvov[k][c][d][a] = g~(kc|ad) = sum_mu Q[mu][k][c] PT[mu][d][a], a row-major
  // (k,c) x (d,a) product
  add([](const Mol& M) {
    const int N = M.n, o = M.o, v = M.v, N2 = N * N, ov = o * v;
    return prob(cmp(M.Q + o, v, N, 1, 1, N2, 0), cmp(M.PT + o * N + o, 1, N2, 0, v, N, 1), rm(M.gvvov, v * v),
                ov, v * v, N, 1.0, 0.0);
  });
  // g4: w1[i][a] = sum_{k,cd} u[k][i][cd] gvvov[(k,c,d)][a]
  add([](const Mol& M) {
    const int o = M.o, v = M.v, ov = o * v;
    return prob(cmp(M.u, 1, v * v, 0, v * v, ov * v, 1), rm(M.gvvov, v), rm(M.w1, v), o, v, ov * v, 1.0,
                0.0);
  });
  // g5: gooov[i][k][l][c] = g~(ki|lc)
  add([](const Mol& M) {
    const int N = M.n, o = M.o, v = M.v, N2 = N * N, ov = o * v;
    return prob(cmp(M.P, o, N, 1, 1, N2, 0), cmp(M.Q + o, 1, N2, 0, v, N, 1),
                cmp(M.gooov, o, ov, o * ov, v, v, 1), o * o, ov, N, 1.0, 0.0);
  });
  // g6: w1 -= gooov[i][(kl,c)] u[(kl),a,c]
  add([](const Mol& M) {
    const int o = M.o, v = M.v;
    return prob(rm(M.gooov, o * o * v), cmp(M.u, v, v * v, 1, 1, v, 0), rm(M.w1, v), o, v, o * o * v,
                -1.0, 1.0);
  });
  // g7: w1[(i,a)] += sum_{k,c} u[i][k][a][c] Ft[k][o+c]
  add([](const Mol& M) {
    const int N = M.n, o = M.o, v = M.v, ov = o * v;
    return prob(cmp(M.u, v, ov * v, v, v, v * v, 1), cmp(M.Ft + o, v, N, 1, 1, 0, 0),
                cmp(M.w1, 1, 1, 0, 1, 0, 0), ov, 1, ov, 1.0, 1.0);
  });
  // g8: w2[i][j][a][b] = g~(ai|bj) = sum_mu Q[mu][a][i] PT[mu][j][b]: rows (a,i),
  // columns (j,b) so that the output's unit stride (b) runs along consecutive columns
  add([](const Mol& M) {
    const int N = M.n, o = M.o, v = M.v, N2 = N * N, ov = o * v;
    return prob(cmp(M.Q + o * N, o, N, 1, 1, N2, 0), cmp(M.PT + o, 1, N2, 0, v, N, 1),
                cmp(M.w2, o, v, ov * v, v, v * v, 1), ov, ov, N, 1.0, 0.0);
  });
  // g9: G4[c][d][a][b] = g~(ac|bd) = sum_mu Q[mu][a][c] PT[mu][d][b]: rows (a,c),
  // columns (d,b)
  add([](const Mol& M) {
    const int N = M.n, o = M.o, v = M.v, N2 = N * N;
    return prob(cmp(M.Q + o * N + o, v, N, 1, 1, N2, 0), cmp(M.PT + o * N + o, 1, N2, 0, v, N, 1),
                cmp(M.G4, v, v, v * v * v, v, v * v, 1), v * v, v * v, N, 1.0, 0.0);
  });
  // g10: particle-particle ladder w2 += t2 G4
  add([](const Mol& M) {
    const int o = M.o, v = M.v;
    return prob(rm(M.t2, v * v), rm(M.G4, v * v), rm(M.w2, v * v), o * o, v * v, v * v, 1.0, 1.0);
  });
  // g11: Wkl[(k,l),(i,j)] = g~(ki|lj)
  add([](const Mol& M) {
    const int N = M.n, o = M.o, N2 = N * N;
    return prob(cmp(M.P, o, N, 1, 1, N2, 0), cmp(M.Q, 1, N2, 0, o, N, 1),
                cmp(M.Wkl, o, o * o * o, o, o, o * o, 1), o * o, o * o, N, 1.0, 0.0);
  });
  // g12: Wkl += Gklcd t2^T
  add([](const Mol& M) {
    const int o = M.o, v = M.v;
    return prob(rm(M.Gklcd, v * v), trn(M.t2, v * v), rm(M.Wkl, o * o), o * o, o * o, v * v, 1.0, 1.0);
  });
  // g13: hole-hole ladder w2 += Wkl^T t2
  add([](const Mol& M) {
    const int o = M.o, v = M.v;
    return prob(trn(M.Wkl, o * o), rm(M.t2, v * v), rm(M.w2, v * v), o * o, v * v, o * o, 1.0, 1.0);
  });
  // g14: WC[(a,i),(k,c)] = g~(ki|ac)
  add([](const Mol& M) {
    const int N = M.n, o = M.o, v = M.v, N2 = N * N, ov = o * v;
    return prob(cmp(M.P, o, N, 1, 1, N2, 0), cmp(M.Q + o * N + o, 1, N2, 0, v, N, 1),
                cmp(M.WC, o, v, ov, v, o * ov, 1), o * o, v * v, N, 1.0, 0.0);
  });
  // g15: WC -= 1/2 t2[l][i][a][d] (lc|kd)
  add([](const Mol& M) {
    const int o = M.o, v = M.v, ov = o * v;
    return prob(cmp(M.t2, o, v, v * v, v, ov * v, 1), cmp(M.Gklcd, v, ov * v, 1, v, v * v, v), rm(M.WC, ov),
                ov, ov, ov, -0.5, 1.0);
  });
  // g16: Zc[(a,i),(j,b)] = WC t2[k][j][b][c], stored at [i][j][a][b] (columns ordered
  // (j, b) so that the output's unit stride runs along consecutive columns)
  add([](const Mol& M) {
    const int o = M.o, v = M.v, ov = o * v;
    return prob(rm(M.WC, ov), cmp(M.t2, v, ov * v, 1, v, v * v, v), cmp(M.Zc, o, v, ov * v, v, v * v, 1), ov,
                ov, ov, 1.0, 0.0);
  });
  // g17: WD = 2 g~(ai|kc)
  add([](const Mol& M) {
    const int N = M.n, o = M.o, v = M.v, N2 = N * N, ov = o * v;
    return prob(cmp(M.P + o * N, o, N, 1, 1, N2, 0), cmp(M.Q + o, 1, N2, 0, v, N, 1), rm(M.WD, ov), ov, ov,
                N, 2.0, 0.0);
  });
  // g18: WD -= g~(ac|ki), as the (k,i) x (a,c) problem so that the output's
  // unit stride (c) runs along consecutive columns
  add([](const Mol& M) {
    const int N = M.n, o = M.o, v = M.v, N2 = N * N, ov = o * v;
    return prob(cmp(M.Q, o, N, 1, 1, N2, 0), cmp(M.P + o * N + o, 1, N2, 0, v, N, 1),
                cmp(M.WD, o, v, ov, v, o * ov, 1), o * o, v * v, N, -1.0, 1.0);
  });
  // g19: WD += 1/2 u[i][l][a][d] L[(l,d),(k,c)]
  add([](const Mol& M) {
    const int o = M.o, v = M.v, ov = o * v;
    return prob(cmp(M.u, o, v, ov * v, v, v * v, 1), rm(M.Lldkc, ov), rm(M.WD, ov), ov, ov, ov, 0.5, 1.0);
  });
  // g20: x[i][j][a][b] = 1/2 WD u[j][k][b][c] (columns (j, b), as in g16)
  add([](const Mol& M) {
    const int o = M.o, v = M.v, ov = o * v;
    return prob(rm(M.WD, ov), cmp(M.u, v, v * v, 1, v, ov * v, v), cmp(M.x, o, v, ov * v, v, v * v, 1), ov,
                ov, ov, 0.5, 0.0);
  });
  // g21: Xvv -= u[(kl)][b][d] Gklcd[(kl)][c][d]
  add([](const Mol& M) {
    const int o = M.o, v = M.v;
    return prob(cmp(M.u, 1, v, 0, v, v * v, 1), cmp(M.Gklcd, v, v * v, 1, 1, v, 0), rm(M.Xvv, v), v, v,
                o * o * v, -1.0, 1.0);
  });
  // g22: Xoo += Gklcd[l][k][cd] u[l][j][cd]
  add([](const Mol& M) {
    const int o = M.o, v = M.v, ov = o * v;
    return prob(cmp(M.Gklcd, 1, v * v, 0, v * v, ov * v, 1), cmp(M.u, v * v, ov * v, 1, 1, v * v, 0),
                rm(M.Xoo, o), o, o, ov * v, 1.0, 1.0);
  });
  // g23: x += t2 Xvv^T, and g24: x -= t2[i][k][ab] Xoo[k][j]; molecules with
  // o, v <= kE2Max take both in the elementwise pass k_e2 instead (empty problems here)
  add([](const Mol& M) {
    const int o = M.o, v = M.v;
    if (o <= kE2Max && v <= kE2Max) return GemmProb{};
    return prob(rm(M.t2, v), trn(M.Xvv, v), rm(M.x, v), o * o * v, v, v, 1.0, 1.0);
  });
  add([](const Mol& M) {
    const int o = M.o, v = M.v, ov = o * v;
    if (o <= kE2Max && v <= kE2Max) return GemmProb{};
    return prob(cmp(M.t2, v * v, ov * v, 1, 1, v * v, 0), rm(M.Xoo, o), cmp(M.x, v * v, ov * v, 1, 1, v * v, 0),
                ov * v, o, o, -1.0, 1.0);
  });
}

// Correction-form twin of iteration GEMM `op` for molecule k: the FP64
// problem plus its correction operands and BF16 operand copies.
static GemmProb delta_prob(Batch& b, int op, int k) {
  GemmProb p = b.ops[op](b.hm[k]);
  if (p.M <= 0 || p.N <= 0) return p;   // handled outside the GEMM tables
  p.dA = delta_of(b, k, p.A.p);
  p.dB = delta_of(b, k, p.B.p);
  p.Cd = delta_of(b, k, p.C.p);
  if (p.Cd == nullptr) throw std::string("correction table: constant output operand");
  if (p.dA == nullptr && p.dB == nullptr) throw std::string("correction table: both operands constant");
  p.hA = shadow_of(b, p.A.p);
  p.hB = shadow_of(b, p.B.p);
  if (p.hA == nullptr || p.hB == nullptr) throw std::string("correction table: operand without a BF16 copy");
  return p;
}

// Tuning keys of the correction-form tables: kDualTuneOp + op.
constexpr int kDualTuneOp = 100;

// correction-form twins of the iteration GEMMs (BF16 tiles chosen per
// problem, or FP64 64x64 tiles in exact mode)
static void build_delta_tables(Batch& b) {
  b.dit.clear();
  for (int q = 0; q < (int)b.ops.size(); ++q)
    b.dit.push_back(make_table(b, [&](int k) { return delta_prob(b, q, k); }, true, b.mixed_bf16 ? 0 : 64,
                               b.mixed_bf16 ? kDualTuneOp + q : -1, 0, b.mixed_bf16));
}

enum ItOp {
  G1_Q = 0, G2A, G2B, G3, G4_, G5, G6, G7, G8, G9, G10, G11, G12, G13, G14, G15, G16, G17, G18, G19, G20,
  G21, G22, G23, G24, N_IT_OPS
};

}  // namespace hcc

using namespace hcc;

// Profile slots (hcc_profile_ops): 0..24 iteration GEMM tables, 32..56 their
// correction forms, 60..63 set-up GEMMs, then the elementwise groups below.
constexpr int kProfSlots = 96;
static_assert(kProfSlots == 96, "profile slot arrays are sized 96");
enum ProfGroup {
  kProfUpdate = 64,        // k_update1 + k_update2 (FP64 step, DIIS push fused)
  kProfDiis = 65,          // k_diis_dot + k_diis_solve + k_diis_extrap (FP64)
  kProfUpdateMixed = 66,   // k_update1_mixed + k_update2_mixed (inner step)
  kProfDiisMixed = 67,     // DIIS over the FP32 corrections (inner)
  kProfE2 = 68,            // k_e2 (FP64)
  kProfE2Mixed = 69,       // k_e2_d (correction form)
  kProfResidualEw = 70,    // dressing, Fock, P/PT, u, X init, assembly (FP64)
  kProfResidualEwMixed = 71,
  kProfOuterNorm = 72,     // residual norms at the outer point (mixed solver)
  kProfCommit = 73,        // t0 += dt, DIIS shift, correction zeroing, BF16 copies
  kProfGuess = 74,         // guess (+ learned correction)
  kProfEnergy = 75,
  kProfPrecomputeEw = 76,
  kProfControl = 77,       // per-molecule iterator control kernels
  kProfTriples = 78,       // k_triples + k_triples_final ((T) correction)
  kProfTriplesEw = 79,     // (T) set-up: flags, undressed products, t1 save/restore
};

struct hcc_ctx {
  int device = 0;
  int precision = HCC_FP64;
  int diis_steps = 20;
  double s1 = 1.0, s2 = 1.0;
  int projtype = 2;
  int diis = 1;
  hipStream_t stream = nullptr;
  Batch* batch = nullptr;
  std::string err;
  bool have_guess = false;
  double last_ms = 0.0;
  double last_iters = 0.0;
  // mixed-precision solver options
  int mixed_bf16 = 1;
  int keep_diis = 1;   // mixed solver: inner DIIS history kept across outer steps
  int mixed_diis = 8;  // mixed solver: inner DIIS depth (0: diis_steps; at most diis_steps)
  double kappa = 0.05;
  int max_inner = 5;
  int max_outer = 30;
  // live timing of the contraction GEMM launches (HIP events on `stream`)
  bool prof = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_events;
  std::vector<int> prof_ids;
  size_t prof_used = 0;
  double prof_ms = 0.0;
  long prof_launches = 0;
  long prof_op_launches[96] = {0};   // GEMM kernel dispatches per profile slot
  double prof_op_ms[kProfSlots] = {0};
  // learned-guess network (FP32 parameters on the device), null until set
  float* guess_net = nullptr;
  // replay blocks of iterations as captured HIP graphs (off while profiling)
  int graphs = 1;
  // measured GEMM tile / K-split choices, (op, n_occ, n_virt) -> (tile edge, ksplit)
  std::map<std::array<int, 3>, std::array<int, 2>> tuning;
  // the reference's iteration block (src/input/definebd.f90:53-58)
  int chk_divergence = 1, chk_stagnation = 1, enable_stop = 1;
  double divergence = 1e3;        // divergence-threshold 3, potentiated: 10^3
  double stagnation_pct = 60.0;   // stagnation-threshold
  int stagnation_after = 200;     // stagnation-check-after
  // stop request (the reference's SIGUSR1 -> chkStopIteration(2),
  // src/modules/mod_helios.f90:441): set from any thread or a signal handler,
  // read by the whole-solve loops between iteration blocks
  std::atomic<int> stop_req{0};
  // no-allocation contexts: device memory from this region (see Batch::arena),
  // no timing events
  char* arena = nullptr;
  size_t arena_cap = 0;
  bool timing = true;
};

static_assert(std::atomic<int>::is_always_lock_free, "hcc_request_stop must be async-signal-safe");

namespace hcc {

static void launch_table(hcc_ctx* c, const GemmTable& t) {
  const TableView v = view(t);
  HCHECK(launch_gemm_f64(v.d, v.tp, v.tiles, c->batch->dact, c->stream, t.tile, t.flags));
  if (v.red > 0) HCHECK(launch_splitk_reduce(v.d, v.rp, v.rb, v.red, c->batch->dact, c->stream));
  if (t.alt >= 0) launch_table(c, c->batch->alt[t.alt]);
}

// One table compacted to the problems of molecules with act != 0: their tiles
// renumbered consecutively, so a launch dispatches no workgroup of a finished
// molecule (each would only read its molecule's flag and return).
static void compact_table(Batch& b, GemmTable& t, const std::vector<int>& act, hipStream_t s) {
  if (t.nprob == 0) return;
  const int np = t.nprob;
  if (t.cd == nullptr) {
    t.cd = b.dalloc<GemmProb>(np);
    t.ctp = b.dalloc<int>(std::max(1, t.tiles));
    t.cbp = b.dalloc<int>(std::max(1, t.red_blocks));
    t.cbuf = b.dalloc<int>(3 * (size_t)np + 2);
  }
  // the previous upload from hbuf has completed: a compaction follows a
  // synchronisation of the stream
  t.hbuf.assign(3 * (size_t)np + 2, 0);
  int* idx = t.hbuf.data();
  int* tbeg = idx + np;
  int* rpre = tbeg + np + 1;
  int n = 0, tiles = 0, red = 0;
  for (int q = 0; q < np; ++q) {
    if (act[t.hmol[q]] == 0) continue;
    idx[n] = q;
    tbeg[n] = tiles;
    rpre[n] = red;
    ++n;
    tiles += t.htiles[q];
    red += t.hred[q];
  }
  tbeg[n] = tiles;
  rpre[n] = red;
  HCHECK(hipMemcpyAsync(t.cbuf, t.hbuf.data(), t.hbuf.size() * sizeof(int), hipMemcpyHostToDevice, s));
  HCHECK(launch_compact_table(t.d, t.cbuf, t.cbuf + np, t.cbuf + 2 * np + 1, n, tiles, red, t.cd, t.ctp, t.cbp, s));
  t.ctiles = tiles;
  t.cred = red;
}

// The FP64 residual's tables (dual = false) or the correction residual's
// (dual = true), with their parts of other tile kinds, compacted to the
// molecules with act != 0 once their count has fallen to kCompactFrac of the
// count the tables were last built for.  A compaction must keep every
// molecule the device will run before the next one: the running set of a solve
// only shrinks, except that each inner phase of the mixed solver starts again
// from all molecules still running (reset = true rebuilds from that set).
// Returns whether the launches changed (captured graphs are then stale).
constexpr double kCompactFrac = 0.85;
static bool compact_tables(hcc_ctx* c, bool dual, const std::vector<int>& act, int alive, bool reset) {
  Batch& b = *c->batch;
  if (b.arena != nullptr || alive <= 0) return false;   // no-allocation contexts keep full tables
  int& last = dual ? b.dit_live : b.it_live;
  std::vector<GemmTable>& tabs = dual ? b.dit : b.it;
  auto each = [&](const std::function<void(GemmTable&)>& f) {
    for (GemmTable& t : tabs)
      for (GemmTable* p = &t; p != nullptr; p = p->alt >= 0 ? &b.alt[p->alt] : nullptr) f(*p);
  };
  if (alive > kCompactFrac * (reset || last == 0 ? b.nm : last)) {
    if (!reset || last == 0) return false;
    each([](GemmTable& t) { t.ctiles = -1; });   // the running set grew back: full tables
    last = 0;
    return true;
  }
  each([&](GemmTable& t) { compact_table(b, t, act, c->stream); });
  last = alive;
  return true;
}

// The elementwise grids compacted to the molecules with act != 0 (the running
// set, which only shrinks during a solve), by the rule of compact_tables.  A
// running molecule keeps its blocks in order, renumbered from its compacted
// prefix, so every kernel's (molecule, block within the molecule) is as
// before; finished molecules get empty ranges.  Returns whether the launches
// changed.
static bool compact_raggeds(hcc_ctx* c, const std::vector<int>& act, int alive) {
  Batch& b = *c->batch;
  if (b.arena != nullptr || alive <= 0) return false;
  if (alive > kCompactFrac * (b.rg_live > 0 ? b.rg_live : b.nm)) return false;
  for (Batch::RgComp& r : b.rgc) {
    if (r.cpre == nullptr) {
      r.cpre = b.dalloc<int>(b.nm + 1);
      r.cbm = b.dalloc<int>(std::max(1, r.full.nblocks));
    }
    // the previous upload from hbuf has completed (a compaction follows a
    // synchronisation of the stream)
    r.hbuf.assign(b.nm + 1, 0);
    int blocks = 0;
    for (int k = 0; k < b.nm; ++k) {
      r.hbuf[k] = blocks;
      if (act[k] != 0) blocks += r.pre[k + 1] - r.pre[k];
    }
    r.hbuf[b.nm] = blocks;
    HCHECK(hipMemcpyAsync(r.cpre, r.hbuf.data(), (b.nm + 1) * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HCHECK(launch_expand_prefix(r.cpre, b.nm, r.cbm, blocks, c->stream));
    *r.rg = Ragged{r.cpre, r.cbm, blocks};
  }
  b.rg_live = alive;
  return true;
}

// full tables and grids again (end of a solve: the energies and other callers
// run every molecule)
static void uncompact_tables(Batch& b) {
  for (auto* v : {&b.it, &b.dit, &b.alt})
    for (GemmTable& t : *v) t.ctiles = -1;
  b.it_live = b.dit_live = 0;
  for (Batch::RgComp& r : b.rgc) *r.rg = r.full;
  b.rg_live = 0;
}
struct UncompactGuard {
  Batch& b;
  ~UncompactGuard() { uncompact_tables(b); }
};

// GEMM kernel dispatches of a table: its parts of different tile kinds with
// tiles to run (a part compacted to no running molecule is not dispatched)
static int table_parts(const hcc_ctx* c, const GemmTable& t) {
  int n = 0;
  for (const GemmTable* p = &t; p != nullptr; p = p->alt >= 0 ? &c->batch->alt[p->alt] : nullptr)
    n += view(*p).tiles > 0;
  return n;
}

// While profiling, the launches between prof_open and prof_close are bracketed
// by a HIP event pair on the context's stream and their time is added to
// profile slot `id` (no-ops otherwise).
static void prof_open(hcc_ctx* c, int id) {
  if (!c->prof) return;
  if (c->prof_used == c->prof_events.size()) {
    hipEvent_t a, b;
    HCHECK(hipEventCreate(&a));
    HCHECK(hipEventCreate(&b));
    c->prof_events.emplace_back(a, b);
    c->prof_ids.push_back(0);
  }
  c->prof_ids[c->prof_used] = id;
  HCHECK(hipEventRecord(c->prof_events[c->prof_used].first, c->stream));
}
static void prof_close(hcc_ctx* c) {
  if (!c->prof) return;
  HCHECK(hipEventRecord(c->prof_events[c->prof_used++].second, c->stream));
}

static void gemm(hcc_ctx* c, const GemmTable& t, int id = 63) {
  if (t.nprob == 0) return;
  if (c->prof) {
    c->prof_launches += table_parts(c, t);
    c->prof_op_launches[id] += table_parts(c, t);
  }
  prof_open(c, id);
  launch_table(c, t);
  prof_close(c);
}

// folds the recorded intervals into the profile slots (call after a
// synchronize); prof_ms sums the GEMM slots only
static void prof_collect(hcc_ctx* c) {
  for (size_t q = 0; q < c->prof_used; ++q) {
    float ms = 0.f;
    HCHECK(hipEventElapsedTime(&ms, c->prof_events[q].first, c->prof_events[q].second));
    if (c->prof_ids[q] < 64) c->prof_ms += ms;
    c->prof_op_ms[c->prof_ids[q]] += ms;
  }
  c->prof_used = 0;
}

static Params params(const hcc_ctx* c, double eps, int maxiter, bool iterator_rules) {
  Params p{};
  p.eps = eps;
  p.divergence = c->divergence;
  p.s1 = c->s1;
  p.s2 = c->s2;
  p.projtype = c->projtype;
  p.diis = c->diis;
  p.maxiter = iterator_rules ? maxiter : 0x7fffffff;
  p.stagnation_after = c->stagnation_after;
  p.stagnation_pct = c->stagnation_pct;
  p.check_divergence = iterator_rules ? c->chk_divergence : 0;
  p.check_stagnation = iterator_rules ? c->chk_stagnation : 0;
  return p;
}

// A pending stop request ends every running molecule with status -4 at this
// host check (the reference's iterator tests chkStopIteration(2) once per
// iteration, gIterator.f90:165-170; here the check runs between blocks of
// iterations).  Returns true when the solve must stop.
static bool take_stop(hcc_ctx* c) {
  if (!c->enable_stop || c->stop_req.exchange(0) == 0) return false;
  Batch& b = *c->batch;
  stop_running(b.dst, b.dact, b.nm, c->stream);
  HCHECK(hipGetLastError());
  return true;
}

// Omega1 -> w1 (without the F~_ai term), Omega2 -> w2
static void residual(hcc_ctx* c) {
  Batch& b = *c->batch;
  hipStream_t s = c->stream;
  prof_open(c, kProfResidualEw);
  ew_dress(b.dm, b.dst, b.rgNN, b.nm, s);
  ew_fock_m(b.dm, b.dst, b.rgN, b.rgNN, b.nm, s);
  ew_pouter(b.dm, b.dst, b.rgN3, b.nm, s);
  prof_close(c);
  gemm(c, b.it[G1_Q], G1_Q);
  gemm(c, b.it[G2A], G2A);
  gemm(c, b.it[G2B], G2B);
  prof_open(c, kProfResidualEw);
  ew_u(b.dm, b.dst, b.rgT2x, b.nm, s);
  prof_close(c);
  for (int g = G3; g <= G20; ++g) gemm(c, b.it[g], g);
  prof_open(c, kProfResidualEw);
  ew_init_xx(b.dm, b.dst, b.rgXX, b.nm, s);
  prof_close(c);
  for (int g = G21; g <= G24; ++g) gemm(c, b.it[g], g);
  prof_open(c, kProfE2);
  ew_e2(b.dm, b.dst, b.rgOcc, b.nm, s);
  prof_close(c);
  prof_open(c, kProfResidualEw);
  ew_assemble(b.dm, b.dst, b.rgT2x, b.nm, s);
  prof_close(c);
  HCHECK(hipGetLastError());
}

static void run_iteration(hcc_ctx* c, const Params& prm, const int* stop) {
  Batch& b = *c->batch;
  hipStream_t s = c->stream;
  residual(c);
  prof_open(c, kProfUpdate);
  ew_update(b.dm, b.dst, b.rgOV, b.rgT2x, b.nm, prm, b.part1, b.part2, s);
  prof_close(c);
  prof_open(c, kProfControl);
  ctl_after_update(b.dm, b.dst, b.nm, prm, b.rgOV, b.part1, b.rgT2x, b.part2, s);
  prof_close(c);
  if (prm.diis) {
    prof_open(c, kProfDiis);
    diis_step(b.dm, b.dst, b.nm, b.rgND, b.rgP2, b.partD, s);
    prof_close(c);
  }
  prof_open(c, kProfControl);
  ctl_after_diis(b.dm, b.dst, b.dact, b.nm, prm, stop, s);
  prof_close(c);
  HCHECK(hipGetLastError());
}

static void dgemm(hcc_ctx* c, const GemmTable& t, int id) {
  if (t.nprob == 0) return;
  if (c->prof) {
    c->prof_launches += table_parts(c, t);
    c->prof_op_launches[id] += table_parts(c, t);
  }
  prof_open(c, id);
  for (const GemmTable* p = &t; p != nullptr; p = p->alt >= 0 ? &c->batch->alt[p->alt] : nullptr) {
    const TableView v = view(*p);
    HCHECK(launch_gemm_dual(v.d, v.tp, v.tiles, c->batch->dact, c->stream, p->flags, c->mixed_bf16, p->tile));
    if (v.red > 0) HCHECK(launch_splitk_reduce(v.d, v.rp, v.rb, v.red, c->batch->dact, c->stream));
  }
  prof_close(c);
}

// correction residual dOmega at (t0 + dt): every step in correction form
static void residual_delta(hcc_ctx* c) {
  Batch& b = *c->batch;
  hipStream_t s = c->stream;
  prof_open(c, kProfResidualEwMixed);
  mixed_dress(b.dm, b.ddm, b.dst, b.rgNN, b.rgN, b.rgN3, b.nm, s);
  prof_close(c);
  dgemm(c, b.dit[G1_Q], 32 + G1_Q);
  dgemm(c, b.dit[G2A], 32 + G2A);
  dgemm(c, b.dit[G2B], 32 + G2B);
  prof_open(c, kProfResidualEwMixed);
  ew_u(b.ddm, b.dst, b.rgT2x, b.nm, s);
  prof_close(c);
  for (int g = G3; g <= G20; ++g) dgemm(c, b.dit[g], 32 + g);
  prof_open(c, kProfResidualEwMixed);
  ew_init_xx(b.ddm, b.dst, b.rgXX, b.nm, s);
  prof_close(c);
  for (int g = G21; g <= G24; ++g) dgemm(c, b.dit[g], 32 + g);
  prof_open(c, kProfE2Mixed);
  mixed_e2(b.dm, b.ddm, b.dst, b.rgOcc, b.nm, s);
  prof_close(c);
  prof_open(c, kProfResidualEwMixed);
  ew_assemble(b.ddm, b.dst, b.rgT2x, b.nm, s);
  prof_close(c);
  HCHECK(hipGetLastError());
}

// integral precompute (Fao, undressed (kc|ld), L matrices)
static void precompute(hcc_ctx* c) {
  Batch& b = *c->batch;
  hipStream_t s = c->stream;
  prof_open(c, kProfPrecomputeEw);
  reset_state(b.dst, b.dact, b.nm, s);
  // t1 = 0 -> undressed orbital products
  zero_t1(b.dm, b.rgOV, b.nm, s);
  prof_close(c);
  gemm(c, b.pre[0], 60);
  gemm(c, b.pre[1], 61);
  prof_open(c, kProfPrecomputeEw);
  ew_dress(b.dm, b.dst, b.rgNN, b.nm, s);
  ew_pouter(b.dm, b.dst, b.rgN3, b.nm, s);
  prof_close(c);
  gemm(c, b.it[G1_Q], 62);
  gemm(c, b.pre[2], 63);
  prof_open(c, kProfPrecomputeEw);
  make_L(b.dm, b.rgT2, b.nm, s);
  prof_close(c);
  HCHECK(hipGetLastError());
}

// initial amplitudes of every molecule; the learned guess corrects the MP2
// one; HCC_GUESS_CURRENT keeps the amplitudes in place and only resets the
// iteration state
static void apply_guess(hcc_ctx* c, int kind) {
  Batch& b = *c->batch;
  b.amp_valid.assign(b.nm, 1);
  if (kind == HCC_GUESS_CURRENT) {
    reset_state(b.dst, b.dact, b.nm, c->stream);
    return;
  }
  prof_open(c, kProfGuess);
  guess(b.dm, b.dst, b.dact, b.rgT2S, b.nm, kind == HCC_GUESS_LEARNED ? HCC_GUESS_MP2 : kind, c->stream);
  if (kind == HCC_GUESS_LEARNED) {
    mlp_overlaps(b.dm, b.rgNN, b.nm, c->stream);
    mlp_guess(b.dm, b.rgPairBlk, b.rgPairL, b.nm, b.guess_large, c->guess_net, c->stream);
    HCHECK(hipGetLastError());
  }
  prof_close(c);
}

bool batch_view(hcc_ctx* c, BatchView* v) {
  if (c == nullptr || c->batch == nullptr || v == nullptr) return false;
  const Batch& b = *c->batch;
  v->dm = b.dm;
  v->dst = b.dst;
  v->nm = b.nm;
  v->device = c->device;
  v->stream = c->stream;
  v->rgNN = b.rgNN;
  return true;
}

// Builds the new batch privately and publishes it in c->batch only when the
// validation, the allocations, the tables and the upload have all succeeded:
// a rejected load leaves the context with no batch (HCC_E_STATE afterwards),
// never with a half-built one.
static void load(hcc_ctx* c, int nmol, const int* n, const int* nocc, const double* C, const double* F,
                 const double* G, const double* eref, bool keep_mp2) {
  delete c->batch;
  c->batch = nullptr;
  c->have_guess = false;
  std::unique_ptr<Batch> owner(new Batch());
  Batch* b = owner.get();
  b->arena = c->arena;
  b->arena_cap = c->arena_cap;
  b->nm = nmol;
  b->m = c->diis_steps;
  b->inner_m = c->mixed_diis > 0 ? std::min(c->mixed_diis, c->diis_steps) : c->diis_steps;
  b->mixed = c->precision == HCC_MIXED;
  b->mixed_bf16 = c->mixed_bf16 != 0;
  b->tuning = c->tuning.empty() ? nullptr : &c->tuning;
  for (int k = 0; k < nmol; ++k) {
    if (n[k] < 2 || nocc[k] < 1 || nocc[k] >= n[k]) throw std::string("invalid molecule dimensions");
    // GEMM operands are addressed through buffer resources with 32-bit byte
    // offsets: every per-molecule array stays below 2^28 elements
    const long nn = n[k], oo = nocc[k], vv = n[k] - nocc[k];
    const long largest = std::max({vv * vv * vv * vv, oo * oo * oo * oo, nn * nn * nn, oo * oo * vv * vv, oo * vv * vv * vv});
    if (largest >= (1L << 28)) throw std::string("molecule too large: a per-molecule array exceeds 2^28 elements");
    // the correction GEMM forms 2 x (element index) in 29 bits
    if (b->mixed && largest >= (1L << 27))
      throw std::string("molecule too large for the mixed solver: a per-molecule array exceeds 2^27 elements");
    b->n.push_back(n[k]);
    b->o.push_back(nocc[k]);
    b->v.push_back(n[k] - nocc[k]);
    b->eref.push_back(eref[k]);
  }
  allocate(*b, keep_mp2);
  build_tables(*b);
  if (b->mixed) build_delta_tables(*b);
  // upload, transposing the column-major inputs to row-major: one staging
  // image per arena (molecule blocks at their arena offsets), one copy each;
  // a null C or F is computed on the device afterwards (hcc_batch_load_scf)
  double* dsts0[3] = {b->hm[0].C, b->hm[0].F, b->hm[0].G};
  const double* srcs[3] = {C, F, G};
  const long nlast = (long)n[nmol - 1] * n[nmol - 1];
  std::vector<double> h;
  for (int q = 0; q < 3; ++q) {
    if (srcs[q] == nullptr) continue;
    double* base = dsts0[q];
    double* const* last = q == 0 ? &b->hm[nmol - 1].C : (q == 1 ? &b->hm[nmol - 1].F : &b->hm[nmol - 1].G);
    const long span = (*last - base) + nlast;
    h.assign((size_t)span, 0.0);
    long off = 0;
    for (int k = 0; k < nmol; ++k) {
      const int N = n[k];
      const Mol& M = b->hm[k];
      double* dst = q == 0 ? M.C : (q == 1 ? M.F : M.G);
      double* hp = h.data() + (dst - base);
      for (int r = 0; r < N; ++r)
        for (int cc = 0; cc < N; ++cc) hp[(size_t)r * N + cc] = srcs[q][off + r + (size_t)cc * N];
      off += (long)N * N;
    }
    HCHECK(hipMemcpy(base, h.data(), (size_t)span * sizeof(double), hipMemcpyHostToDevice));
  }
  c->batch = owner.release();
}

}  // namespace hcc

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
#define API_BEGIN try {
#define API_END                                                                     \
  }                                                                                 \
  catch (const HipError& he) {                                                      \
    ctx->err = std::string(he.what) + ": " + hipGetErrorString(he.e);              \
    return HCC_E_HIP;                                                               \
  }                                                                                 \
  catch (const std::string& se) {                                                   \
    ctx->err = se;                                                                  \
    return HCC_E_ARG;                                                               \
  }                                                                                 \
  catch (const std::bad_alloc&) {                                                   \
    ctx->err = "host allocation failed";                                            \
    return HCC_E_MEMORY;                                                            \
  }                                                                                 \
  return HCC_OK;

extern "C" {

int hcc_create(hcc_ctx** out, int device, int precision, int diis_steps) {
  if (out == nullptr) return HCC_E_ARG;
  *out = nullptr;
  if (precision != HCC_FP64 && precision != HCC_MIXED) return HCC_E_ARG;
  if (diis_steps < 2 || diis_steps > kMaxDiis) return HCC_E_ARG;
  hcc_ctx* ctx = new hcc_ctx();
  ctx->device = device;
  ctx->precision = precision;
  ctx->diis_steps = diis_steps;
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return HCC_E_HIP;
  }
  *out = ctx;
  return HCC_OK;
}

int hcc_set_options(hcc_ctx* ctx, double t1_step, double t2_step, int projection_type, int diis_enabled) {
  if (ctx == nullptr) return HCC_E_ARG;
  if (projection_type != 1 && projection_type != 2) {
    ctx->err = "projection_type must be 1 or 2";
    return HCC_E_ARG;
  }
  ctx->s1 = t1_step;
  ctx->s2 = t2_step;
  ctx->projtype = projection_type;
  ctx->diis = diis_enabled ? 1 : 0;
  return HCC_OK;
}

int hcc_set_iterator(hcc_ctx* ctx, int check_divergence, double divergence_threshold, int check_stagnation,
                     double stagnation_threshold, int stagnation_after, int enable_stop) {
  if (ctx == nullptr) return HCC_E_ARG;
  if (!(divergence_threshold > 0.0) || !(stagnation_threshold >= 0.0 && stagnation_threshold <= 100.0) ||
      stagnation_after < 0) {
    ctx->err = "invalid iterator options";
    return HCC_E_ARG;
  }
  ctx->chk_divergence = check_divergence ? 1 : 0;
  ctx->divergence = divergence_threshold;
  ctx->chk_stagnation = check_stagnation ? 1 : 0;
  ctx->stagnation_pct = stagnation_threshold;
  ctx->stagnation_after = stagnation_after;
  ctx->enable_stop = enable_stop ? 1 : 0;
  return HCC_OK;
}

int hcc_request_stop(hcc_ctx* ctx) {
  if (ctx == nullptr) return HCC_E_ARG;
  ctx->stop_req.store(1);
  return HCC_OK;
}

static int need_batch(hcc_ctx* ctx);

int hcc_set_mixed_history(hcc_ctx* ctx, int keep) {
  if (ctx == nullptr) return HCC_E_ARG;
  ctx->keep_diis = keep ? 1 : 0;
  return HCC_OK;
}

int hcc_set_mixed_diis(hcc_ctx* ctx, int depth) {
  if (ctx == nullptr) return HCC_E_ARG;
  if (depth != 0 && (depth < 2 || depth > ctx->diis_steps)) {
    ctx->err = "inner DIIS depth must be 0 or in [2, diis_steps]";
    return HCC_E_ARG;
  }
  ctx->mixed_diis = depth;
  return HCC_OK;
}

int hcc_set_mixed(hcc_ctx* ctx, int bf16, double kappa, int max_inner, int max_outer) {
  if (ctx == nullptr) return HCC_E_ARG;
  if (!(kappa > 0.0 && kappa < 1.0) || max_inner < 1 || max_outer < 1) {
    ctx->err = "invalid mixed-precision options";
    return HCC_E_ARG;
  }
  ctx->kappa = kappa;
  ctx->max_inner = max_inner;
  ctx->max_outer = max_outer;
  const int want = bf16 ? 1 : 0;
  if (want == ctx->mixed_bf16) return HCC_OK;
  ctx->mixed_bf16 = want;
  // a loaded batch's correction tables depend on the choice (tile kinds)
  if (ctx->batch == nullptr || !ctx->batch->mixed) return HCC_OK;
  API_BEGIN
  HCHECK(hipSetDevice(ctx->device));
  ctx->batch->mixed_bf16 = want != 0;
  build_delta_tables(*ctx->batch);
  API_END
}

int hcc_batch_statistics(hcc_ctx* ctx, int* outer_evaluations, int* inner_iterations) {
  if (ctx == nullptr) return HCC_E_ARG;
  if (int r = need_batch(ctx)) return r;
  API_BEGIN
  Batch& b = *ctx->batch;
  std::vector<State> S(b.nm);
  HCHECK(hipMemcpy(S.data(), b.dst, b.nm * sizeof(State), hipMemcpyDeviceToHost));
  for (int k = 0; k < b.nm; ++k) {
    if (outer_evaluations) outer_evaluations[k] = S[k].outer_its;
    if (inner_iterations) inner_iterations[k] = S[k].iter - S[k].outer_its;
  }
  API_END
}

int hcc_set_molecule(hcc_ctx* ctx, int n, int n_occ, const double* C, const double* F_mo, const double* gamma,
                     double e_ref) {
  if (ctx == nullptr) return HCC_E_ARG;
  if (C == nullptr || F_mo == nullptr || gamma == nullptr) {
    ctx->err = "null input array";
    return HCC_E_ARG;
  }
  API_BEGIN
  HCHECK(hipSetDevice(ctx->device));
  load(ctx, 1, &n, &n_occ, C, F_mo, gamma, &e_ref, false);
  precompute(ctx);
  HCHECK(hipStreamSynchronize(ctx->stream));
  API_END
}

int hcc_batch_load(hcc_ctx* ctx, int nmol, const int* n, const int* n_occ, const double* C, const double* F_mo,
                   const double* gamma, const double* e_ref) {
  if (ctx == nullptr) return HCC_E_ARG;
  if (nmol < 1 || !n || !n_occ || !C || !F_mo || !gamma || !e_ref) {
    ctx->err = "invalid batch arguments";
    return HCC_E_ARG;
  }
  API_BEGIN
  HCHECK(hipSetDevice(ctx->device));
  load(ctx, nmol, n, n_occ, C, F_mo, gamma, e_ref, false);
  API_END
}

int hcc_batch_load_scf(hcc_ctx* ctx, int nmol, const int* n, const int* n_occ, const double* core,
                       const double* gamma, const double* huckel, double scf_accuracy, int scf_maxiter,
                       double* e_ref, int* scf_iterations, int* scf_status) {
  if (ctx == nullptr) return HCC_E_ARG;
  if (nmol < 1 || !n || !n_occ || !core || !gamma || !huckel || !(scf_accuracy > 0.0) || scf_maxiter < 1) {
    ctx->err = "invalid SCF batch arguments";
    return HCC_E_ARG;
  }
  API_BEGIN
  HCHECK(hipSetDevice(ctx->device));
  int nmax = 0;
  for (int k = 0; k < nmol; ++k) nmax = std::max(nmax, n[k]);
  int lds_max = 0;
  HCHECK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
  if (nmax > 64 || scf_lds_bytes(nmax) > (size_t)lds_max)
    throw std::string("the device SCF takes molecules of up to 64 sites whose three n x n matrices fit in LDS (") +
        std::to_string(lds_max) + " bytes per workgroup)";
  std::vector<double> zero_eref(nmol, 0.0);
  load(ctx, nmol, n, n_occ, nullptr, nullptr, gamma, zero_eref.data(), false);
  Batch& b = *ctx->batch;
  // core and Hueckel matrices (row-major; column-major inputs transposed) and work blocks
  long total = 0;
  std::vector<long> off(nmol);
  for (int k = 0; k < nmol; ++k) {
    off[k] = total;
    total += (long)n[k] * n[k];
  }
  double* dcore = b.dalloc<double>(total);
  double* dhk = b.dalloc<double>(total);
  double* dwork = b.dalloc<double>(total);
  std::vector<double> hc(total), hh(total);
  for (int k = 0; k < nmol; ++k)
    for (int r = 0; r < n[k]; ++r)
      for (int c = 0; c < n[k]; ++c) {
        hc[off[k] + (long)r * n[k] + c] = core[off[k] + r + (long)c * n[k]];
        hh[off[k] + (long)r * n[k] + c] = huckel[off[k] + r + (long)c * n[k]];
      }
  HCHECK(hipMemcpy(dcore, hc.data(), total * sizeof(double), hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(dhk, hh.data(), total * sizeof(double), hipMemcpyHostToDevice));
  std::vector<ScfInput> in(nmol);
  for (int k = 0; k < nmol; ++k) in[k] = ScfInput{dcore + off[k], dhk + off[k], dwork + off[k]};
  ScfInput* din = b.dalloc<ScfInput>(nmol);
  ScfResult* dres = b.dalloc<ScfResult>(nmol);
  HCHECK(hipMemcpy(din, in.data(), nmol * sizeof(ScfInput), hipMemcpyHostToDevice));
  HCHECK(launch_rhf(b.dm, din, nmol, nmax, scf_accuracy, scf_maxiter, 0.05, dres, ctx->stream));
  std::vector<ScfResult> res(nmol);
  HCHECK(hipMemcpyAsync(res.data(), dres, nmol * sizeof(ScfResult), hipMemcpyDeviceToHost, ctx->stream));
  HCHECK(hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < nmol; ++k) {
    b.eref[k] = res[k].eref;
    b.hm[k].eref = res[k].eref;
    if (e_ref) e_ref[k] = res[k].eref;
    if (scf_iterations) scf_iterations[k] = res[k].iterations;
    if (scf_status) scf_status[k] = res[k].status;
  }
  API_END
}

static int need_batch(hcc_ctx* ctx);

// Buffers, GEMM tables and work items of the (T) correction (once per batch).
static void setup_triples(Batch& b, int device) {
  const int nm = b.nm;
  b.tA.assign(nm, nullptr);
  b.tB.assign(nm, nullptr);
  std::vector<long> prefix(nm + 1, 0);
  for (int k = 0; k < nm; ++k) {
    const long o = b.o[k], v = b.v[k];
    b.tA[k] = b.dalloc<double>(o * v * v * v);
    b.tB[k] = b.dalloc<double>(o * o * o * v);
    prefix[k + 1] = prefix[k] + o * (o + 1) * (o + 2) / 6;
  }
  b.tpre.clear();
  // (xp|yd) at [p][x][y][d]: rows (p, x) of P, columns (y, d) of Q = gamma P (undressed)
  b.tpre.push_back(make_table(b, [&](int k) {
    const Mol& M = b.hm[k];
    const int N = M.n, o = M.o, v = M.v, N2 = N * N;
    return prob(cmp(M.P + o * N, v, 1, N, 1, N2, 0), cmp(M.Q + o * N + o, 1, N2, 0, v, N, 1), rm(b.tA[k], v * v),
                o * v, v * v, N, 1.0, 0.0);
  }));
  // (ql|rz) at [q][l][r][z]: rows (q, l) of P, columns (r, z) of Q
  b.tpre.push_back(make_table(b, [&](int k) {
    const Mol& M = b.hm[k];
    const int N = M.n, o = M.o, v = M.v, N2 = N * N;
    return prob(cmp(M.P, o, N, 1, 1, N2, 0), cmp(M.Q + o, 1, N2, 0, v, N, 1), rm(b.tB[k], o * v), o * o, o * v, N,
                1.0, 0.0);
  }));
  // work items: molecule and occupied triple i >= j >= k (i ascending)
  std::vector<int2> items;
  items.reserve(prefix[nm]);
  for (int m = 0; m < nm; ++m)
    for (int i = 0; i < b.o[m]; ++i)
      for (int j = 0; j <= i; ++j)
        for (int k = 0; k <= j; ++k) items.push_back(int2{m, i | (j << 8) | (k << 16)});
  std::vector<TripMol> tm(nm);
  for (int k = 0; k < nm; ++k) {
    const Mol& M = b.hm[k];
    tm[k] = TripMol{M.n, M.o, M.v, 0, b.tA[k], b.tB[k], M.Gklcd, M.t1, M.t2, M.F};
  }
  b.ttotal = prefix[nm];
  // molecules with v <= kTripLdsV go to the matrix-core kernel (W in LDS, one
  // workgroup per CU); the others keep W in a global block per workgroup
  int vglob = 0;
  bool small = false;
  for (int k = 0; k < nm; ++k) {
    if (b.v[k] > kTripLdsV) vglob = std::max(vglob, b.v[k]);
    else small = true;
  }
  b.tglobal = vglob > 0;
  int cus = 0;
  HCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  b.tgrid_mfma = small ? std::max(1, cus) : 0;
  b.twstride = ((long)vglob * vglob * vglob + 31) / 32 * 32;
  b.tgrid = b.tglobal ? (int)std::max<long>(1, std::min<long>(b.ttotal, 1024)) : 1;
  b.dtm = b.dalloc<TripMol>(nm);
  b.dtprefix = b.dalloc<long>(nm + 1);
  b.dtri = b.dalloc<int2>(items.size());
  b.dtw = b.dalloc<double>((size_t)b.tgrid * b.twstride);
  b.dtpart = b.dalloc<double>(2 * std::max<long>(1, b.ttotal));
  b.dtout = b.dalloc<double>(2 * (long)nm);
  b.dtsave = b.dalloc<int>(nm);
  HCHECK(hipMemcpy(b.dtm, tm.data(), nm * sizeof(TripMol), hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(b.dtprefix, prefix.data(), (nm + 1) * sizeof(long), hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(b.dtri, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice));
}

int hcc_batch_triples(hcc_ctx* ctx, double* energies, double* e_t) {
  if (ctx == nullptr) return HCC_E_ARG;
  if (int r = need_batch(ctx)) return r;
  Batch& b = *ctx->batch;
  for (int k = 0; k < b.nm; ++k)
    if (b.o[k] > kTripMaxOV || b.v[k] > kTripMaxOV) {
      ctx->err = "the (T) correction takes molecules with up to 32 occupied and 32 virtual orbitals";
      return HCC_E_ARG;
    }
  if (!ctx->have_guess) {
    ctx->err = "(T) needs amplitudes: solve first";
    return HCC_E_STATE;
  }
  API_BEGIN
  HCHECK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  if (b.dtm == nullptr) setup_triples(b, ctx->device);
  // undressed orbital products (t1 = 0), with t1 kept in w1 meanwhile: the
  // t1 and w1 arenas have the same per-molecule offsets
  const Mol& first = b.hm[0];
  const Mol& last = b.hm[b.nm - 1];
  const size_t span = (size_t)(last.t1 - first.t1) + (size_t)last.o * last.v;
  prof_open(ctx, kProfTriplesEw);
  HCHECK(hipMemcpyAsync(first.w1, first.t1, span * sizeof(double), hipMemcpyDeviceToDevice, s));
  arm_all(b.dst, b.dact, b.dtsave, b.nm, false, s);
  zero_t1(b.dm, b.rgOV, b.nm, s);
  ew_dress(b.dm, b.dst, b.rgNN, b.nm, s);
  ew_pouter(b.dm, b.dst, b.rgN3, b.nm, s);
  prof_close(ctx);
  gemm(ctx, b.it[G1_Q], 62);
  gemm(ctx, b.tpre[0], 58);
  gemm(ctx, b.tpre[1], 59);
  prof_open(ctx, kProfTriplesEw);
  HCHECK(hipMemcpyAsync(first.t1, first.w1, span * sizeof(double), hipMemcpyDeviceToDevice, s));
  prof_close(ctx);
  prof_open(ctx, kProfTriples);
  HCHECK(launch_triples(b.dtm, b.nm, b.dtprefix, b.ttotal, b.dtri, b.dtw, b.twstride, b.tgrid, b.tgrid_mfma, b.tglobal,
                        b.dtpart, b.dtout, s));
  prof_close(ctx);
  prof_open(ctx, kProfTriplesEw);
  arm_all(b.dst, b.dact, b.dtsave, b.nm, true, s);
  prof_close(ctx);
  std::vector<double> out(2 * (size_t)b.nm);
  std::vector<State> S(b.nm);
  HCHECK(hipMemcpyAsync(out.data(), b.dtout, out.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HCHECK(hipMemcpyAsync(S.data(), b.dst, b.nm * sizeof(State), hipMemcpyDeviceToHost, s));
  HCHECK(hipStreamSynchronize(s));
  if (ctx->prof) prof_collect(ctx);
  for (int k = 0; k < b.nm; ++k) {
    const double eT = out[2 * k], eST = out[2 * k + 1];
    if (e_t) {
      e_t[2 * k] = eT;
      e_t[2 * k + 1] = eST;
    }
    if (energies) {
      // energyCC('spin-r-ccsd(t)'): c1 = -E[T], c2 = -E[ST]; c3 = sum t2 F_kc = 0 (F_kc = 0 for the RHF reference)
      const double c1 = -eT, c2 = -eST;
      energies[5 * k + 0] = S[k].energy[0] + c1 + c2;
      energies[5 * k + 1] = S[k].energy[1];
      energies[5 * k + 2] = S[k].energy[2];
      energies[5 * k + 3] = S[k].energy[3];
      energies[5 * k + 4] = c1;
    }
  }
  API_END
}

// Buffers of the linear-response sigma products (once per batch).
static void setup_lr(Batch& b) {
  const int nm = b.nm;
  std::vector<double*> pc(4 * (size_t)nm);
  long n1 = 0, n2 = 0;
  for (int k = 0; k < nm; ++k) {
    n1 += (long)b.o[k] * b.v[k];
    n2 += (long)b.o[k] * b.o[k] * b.v[k] * b.v[k];
  }
  const Mol& first = b.hm[0];
  const Mol& last = b.hm[nm - 1];
  b.lr_n1 = n1;
  b.lr_n2 = n2;
  b.lr_span1 = (long)(last.t1 - first.t1) + (long)last.o * last.v;
  b.lr_span2 = (long)(last.t2 - first.t2) + (long)last.o * last.o * last.v * last.v;
  b.dlr_r1 = b.dalloc<double>(n1);
  b.dlr_r2 = b.dalloc<double>(n2);
  b.dlr_s1 = b.dalloc<double>(n1);
  b.dlr_s2 = b.dalloc<double>(n2);
  b.dlr_t1b = b.dalloc<double>(b.lr_span1);
  b.dlr_t2b = b.dalloc<double>(b.lr_span2);
  long a1 = 0, a2 = 0;
  for (int k = 0; k < nm; ++k) {
    pc[k] = b.dlr_r1 + a1;
    pc[nm + k] = b.dlr_r2 + a2;
    pc[2 * nm + k] = b.dlr_s1 + a1;
    pc[3 * nm + k] = b.dlr_s2 + a2;
    a1 += (long)b.o[k] * b.v[k];
    a2 += (long)b.o[k] * b.o[k] * b.v[k] * b.v[k];
  }
  b.dlr_pc = b.dalloc<double*>(pc.size());
  b.dlr_ps = b.dalloc<double*>(pc.size());
  HCHECK(hipMemcpy(b.dlr_pc, pc.data(), pc.size() * sizeof(double*), hipMemcpyHostToDevice));
  if (b.dtsave == nullptr) b.dtsave = b.dalloc<int>(nm);
}

static int lr_ready(hcc_ctx* ctx) {
  if (int r = need_batch(ctx)) return r;
  if (!ctx->have_guess) {
    ctx->err = "linear response needs amplitudes: solve first";
    return HCC_E_STATE;
  }
  return HCC_OK;
}

// sigma = (Omega(t0 + h r) - Omega(t0 - h r)) / 2h for every molecule whose r1
// pointer is set; ptrs (device) = r1 (nm), r2 (nm), sigma1 (nm), sigma2 (nm).
// Only those molecules are active during the two residual passes (the others'
// GEMM tiles and elementwise blocks return at once).  The amplitudes and the
// active flags are restored.
static void lr_sigma_run(hcc_ctx* ctx, double h, double* const* ptrs) {
  Batch& b = *ctx->batch;
  hipStream_t s = ctx->stream;
  const int nm = b.nm;
  const Mol& first = b.hm[0];
  const double* const* r1p = ptrs;
  const double* const* r2p = ptrs + nm;
  HCHECK(hipMemcpyAsync(b.dlr_t1b, first.t1, b.lr_span1 * sizeof(double), hipMemcpyDeviceToDevice, s));
  HCHECK(hipMemcpyAsync(b.dlr_t2b, first.t2, b.lr_span2 * sizeof(double), hipMemcpyDeviceToDevice, s));
  arm_all(b.dst, b.dact, b.dtsave, nm, false, s, r1p);
  lr_perturb(b.dm, b.rgT2, nm, first.t1, b.dlr_t1b, first.t2, b.dlr_t2b, r1p, r2p, h, s);
  residual(ctx);
  lr_omega(b.dm, b.rgT2, nm, ptrs + 2 * nm, ptrs + 3 * nm, false, 1.0, s);
  lr_perturb(b.dm, b.rgT2, nm, first.t1, b.dlr_t1b, first.t2, b.dlr_t2b, r1p, r2p, -h, s);
  residual(ctx);
  lr_omega(b.dm, b.rgT2, nm, ptrs + 2 * nm, ptrs + 3 * nm, true, 0.5 / h, s);
  HCHECK(hipMemcpyAsync(first.t1, b.dlr_t1b, b.lr_span1 * sizeof(double), hipMemcpyDeviceToDevice, s));
  HCHECK(hipMemcpyAsync(first.t2, b.dlr_t2b, b.lr_span2 * sizeof(double), hipMemcpyDeviceToDevice, s));
  arm_all(b.dst, b.dact, b.dtsave, nm, true, s);
  HCHECK(hipGetLastError());
}

int hcc_batch_lr_sigma(hcc_ctx* ctx, double h, const double* r1, const double* r2, double* s1, double* s2) {
  if (ctx == nullptr || r1 == nullptr || r2 == nullptr || s1 == nullptr || s2 == nullptr) return HCC_E_ARG;
  if (int r = lr_ready(ctx)) return r;
  if (!(h > 0.0)) {
    ctx->err = "the difference step must be positive";
    return HCC_E_ARG;
  }
  API_BEGIN
  HCHECK(hipSetDevice(ctx->device));
  Batch& b = *ctx->batch;
  hipStream_t s = ctx->stream;
  if (b.dlr_pc == nullptr) setup_lr(b);
  HCHECK(hipMemcpyAsync(b.dlr_r1, r1, b.lr_n1 * sizeof(double), hipMemcpyHostToDevice, s));
  HCHECK(hipMemcpyAsync(b.dlr_r2, r2, b.lr_n2 * sizeof(double), hipMemcpyHostToDevice, s));
  lr_sigma_run(ctx, h, b.dlr_pc);
  HCHECK(hipMemcpyAsync(s1, b.dlr_s1, b.lr_n1 * sizeof(double), hipMemcpyDeviceToHost, s));
  HCHECK(hipMemcpyAsync(s2, b.dlr_s2, b.lr_n2 * sizeof(double), hipMemcpyDeviceToHost, s));
  HCHECK(hipStreamSynchronize(s));
  API_END
}

// ---------------------------------------------------------------------------
// Device-resident vector space of the eigensolver (hcc_lr.h)
// ---------------------------------------------------------------------------
static int lr_space_impl(hcc_ctx* ctx, const std::vector<int>& counts) {
  API_BEGIN
  HCHECK(hipSetDevice(ctx->device));
  Batch& b = *ctx->batch;
  if (b.dlr_pc == nullptr) setup_lr(b);
  const int nm = b.nm;
  b.lr_len.assign(nm, 0);
  b.lr_base.assign(nm, 0);
  long total = 0;
  for (int k = 0; k < nm; ++k) {
    b.lr_len[k] = (long)b.o[k] * b.v[k] + (long)b.o[k] * b.o[k] * b.v[k] * b.v[k];
    b.lr_base[k] = total;
    total += (b.lr_len[k] * counts[k] + 31) / 32 * 32;   // 256-byte aligned molecule regions
  }
  if (b.dlr_space) HCHECK(hipFree(b.dlr_space));
  b.dlr_space = nullptr;
  b.lr_slots = 0;
  b.lr_nslot.clear();
  HCHECK(hipMalloc(&b.dlr_space, (size_t)std::max<long>(total, 32) * sizeof(double)));
  HCHECK(hipMemsetAsync(b.dlr_space, 0, (size_t)std::max<long>(total, 32) * sizeof(double), ctx->stream));
  if (b.dlr_len == nullptr) {
    b.dlr_len = b.dalloc<long>(nm);
    b.dlr_base = b.dalloc<long>(nm);
  }
  HCHECK(hipMemcpy(b.dlr_len, b.lr_len.data(), nm * sizeof(long), hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(b.dlr_base, b.lr_base.data(), nm * sizeof(long), hipMemcpyHostToDevice));
  HCHECK(hipStreamSynchronize(ctx->stream));
  b.lr_nslot = counts;
  b.lr_slots = *std::max_element(counts.begin(), counts.end());
  API_END
}

int hcc_lr_space(hcc_ctx* ctx, int nslots) {
  if (ctx == nullptr) return HCC_E_ARG;
  if (int r = lr_ready(ctx)) return r;
  if (nslots < 1) {
    ctx->err = "the vector space needs at least one slot";
    return HCC_E_ARG;
  }
  return lr_space_impl(ctx, std::vector<int>(ctx->batch->nm, nslots));
}

int hcc_lr_space_counts(hcc_ctx* ctx, const int* nslots) {
  if (ctx == nullptr || nslots == nullptr) return HCC_E_ARG;
  if (int r = lr_ready(ctx)) return r;
  std::vector<int> counts(nslots, nslots + ctx->batch->nm);
  if (*std::min_element(counts.begin(), counts.end()) < 0 || *std::max_element(counts.begin(), counts.end()) < 1) {
    ctx->err = "slot counts must be non-negative, and some molecule needs a slot";
    return HCC_E_ARG;
  }
  return lr_space_impl(ctx, counts);
}

static int lr_space_ready(hcc_ctx* ctx) {
  if (int r = lr_ready(ctx)) return r;
  if (ctx->batch->lr_slots == 0) {
    ctx->err = "no vector space: call hcc_lr_space first";
    return HCC_E_STATE;
  }
  return HCC_OK;
}

static bool lr_slot_ok(const Batch& b, int mol, int slot) {
  return mol >= 0 && mol < b.nm && slot >= 0 && slot < b.lr_nslot[mol];
}

// staging region of at least `bytes` (grown by doubling; contents not kept)
static char* lr_stage(Batch& b, size_t bytes) {
  if (bytes > b.lr_stage_cap) {
    if (b.dlr_stage) HCHECK(hipFree(b.dlr_stage));
    b.dlr_stage = nullptr;
    size_t cap = std::max<size_t>(bytes, 2 * b.lr_stage_cap);
    HCHECK(hipMalloc(&b.dlr_stage, cap));
    b.lr_stage_cap = cap;
  }
  return b.dlr_stage;
}
static size_t lr_align(size_t x) { return (x + 255) / 256 * 256; }

int hcc_lr_put(hcc_ctx* ctx, int mol, int slot, const double* x) {
  if (ctx == nullptr || x == nullptr) return HCC_E_ARG;
  if (int r = lr_space_ready(ctx)) return r;
  Batch& b = *ctx->batch;
  if (!lr_slot_ok(b, mol, slot)) {
    ctx->err = "molecule or slot out of range";
    return HCC_E_ARG;
  }
  API_BEGIN
  HCHECK(hipSetDevice(ctx->device));
  HCHECK(hipMemcpyAsync(b.dlr_space + b.lr_base[mol] + (long)slot * b.lr_len[mol], x, b.lr_len[mol] * sizeof(double),
                        hipMemcpyHostToDevice, ctx->stream));
  HCHECK(hipStreamSynchronize(ctx->stream));
  API_END
}

int hcc_lr_get(hcc_ctx* ctx, int mol, int slot, double* x) {
  if (ctx == nullptr || x == nullptr) return HCC_E_ARG;
  if (int r = lr_space_ready(ctx)) return r;
  Batch& b = *ctx->batch;
  if (!lr_slot_ok(b, mol, slot)) {
    ctx->err = "molecule or slot out of range";
    return HCC_E_ARG;
  }
  API_BEGIN
  HCHECK(hipSetDevice(ctx->device));
  HCHECK(hipMemcpyAsync(x, b.dlr_space + b.lr_base[mol] + (long)slot * b.lr_len[mol], b.lr_len[mol] * sizeof(double),
                        hipMemcpyDeviceToHost, ctx->stream));
  HCHECK(hipStreamSynchronize(ctx->stream));
  API_END
}

int hcc_lr_dots(hcc_ctx* ctx, int njobs, const int* jobs, int mode, double* out) {
  if (ctx == nullptr || njobs < 0 || (njobs > 0 && (jobs == nullptr || out == nullptr)) || (mode != 0 && mode != 1))
    return HCC_E_ARG;
  if (int r = lr_space_ready(ctx)) return r;
  Batch& b = *ctx->batch;
  std::vector<int> prefix(njobs + 1, 0);
  for (int j = 0; j < njobs; ++j) {
    const int mol = jobs[3 * j], sa = jobs[3 * j + 1], sb = jobs[3 * j + 2];
    if (!lr_slot_ok(b, mol, sa) || !(lr_slot_ok(b, mol, sb) || (mode == 1 && sb == -1))) {
      ctx->err = "dot job " + std::to_string(j) + ": molecule or slot out of range";
      return HCC_E_ARG;
    }
    prefix[j + 1] = prefix[j] + (int)((b.lr_len[mol] + kLrDotChunk - 1) / kLrDotChunk);
  }
  if (njobs == 0) return HCC_OK;
  API_BEGIN
  HCHECK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int nchunks = prefix[njobs];
  const size_t o_pre = 0, o_job = lr_align(sizeof(int) * (njobs + 1)), o_out = o_job + lr_align(3 * sizeof(int) * njobs),
               o_part = o_out + lr_align(sizeof(double) * njobs), total = o_part + sizeof(double) * (nchunks + 1);
  char* st = lr_stage(b, total);
  HCHECK(hipMemcpyAsync(st + o_pre, prefix.data(), sizeof(int) * (njobs + 1), hipMemcpyHostToDevice, s));
  HCHECK(hipMemcpyAsync(st + o_job, jobs, 3 * sizeof(int) * (size_t)njobs, hipMemcpyHostToDevice, s));
  double* dout = reinterpret_cast<double*>(st + o_out);
  lr_dots(b.dlr_space, b.dlr_base, b.dlr_len, njobs, reinterpret_cast<int*>(st + o_pre),
          reinterpret_cast<int*>(st + o_job), mode, reinterpret_cast<double*>(st + o_part), nchunks, dout, s);
  HCHECK(hipGetLastError());
  HCHECK(hipMemcpyAsync(out, dout, sizeof(double) * (size_t)njobs, hipMemcpyDeviceToHost, s));
  HCHECK(hipStreamSynchronize(s));
  API_END
}

int hcc_lr_combine(hcc_ctx* ctx, int njobs, const int* jobs, const double* beta, const int* term_slot,
                   const double* term_coef) {
  if (ctx == nullptr || njobs < 0 || (njobs > 0 && (jobs == nullptr || beta == nullptr))) return HCC_E_ARG;
  if (int r = lr_space_ready(ctx)) return r;
  Batch& b = *ctx->batch;
  long nterms = 0;
  std::vector<int> prefix(njobs + 1, 0);
  for (int j = 0; j < njobs; ++j) {
    const int mol = jobs[4 * j], dst = jobs[4 * j + 1], first = jobs[4 * j + 2], cnt = jobs[4 * j + 3];
    if (!lr_slot_ok(b, mol, dst) || first < 0 || cnt < 0 || (cnt > 0 && (term_slot == nullptr || term_coef == nullptr))) {
      ctx->err = "combine job " + std::to_string(j) + ": invalid molecule, slot or term range";
      return HCC_E_ARG;
    }
    for (int t = first; t < first + cnt; ++t)
      if (!lr_slot_ok(b, mol, term_slot[t]) || term_slot[t] == dst) {
        ctx->err = "combine job " + std::to_string(j) + ": term slot out of range or equal to the destination";
        return HCC_E_ARG;
      }
    nterms = std::max<long>(nterms, (long)first + cnt);
    prefix[j + 1] = prefix[j] + (int)((b.lr_len[mol] + kLrChunk - 1) / kLrChunk);
  }
  if (njobs == 0) return HCC_OK;
  API_BEGIN
  HCHECK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const size_t o_pre = 0, o_job = lr_align(sizeof(int) * (njobs + 1)), o_beta = o_job + lr_align(sizeof(int4) * njobs),
               o_ts = o_beta + lr_align(sizeof(double) * njobs), o_tc = o_ts + lr_align(sizeof(int) * (nterms + 1)),
               total = o_tc + sizeof(double) * (nterms + 1);
  char* st = lr_stage(b, total);
  HCHECK(hipMemcpyAsync(st + o_pre, prefix.data(), sizeof(int) * (njobs + 1), hipMemcpyHostToDevice, s));
  HCHECK(hipMemcpyAsync(st + o_job, jobs, sizeof(int4) * njobs, hipMemcpyHostToDevice, s));
  HCHECK(hipMemcpyAsync(st + o_beta, beta, sizeof(double) * njobs, hipMemcpyHostToDevice, s));
  if (nterms > 0) {
    HCHECK(hipMemcpyAsync(st + o_ts, term_slot, sizeof(int) * nterms, hipMemcpyHostToDevice, s));
    HCHECK(hipMemcpyAsync(st + o_tc, term_coef, sizeof(double) * nterms, hipMemcpyHostToDevice, s));
  }
  lr_combine(b.dlr_space, b.dlr_base, b.dlr_len, njobs, reinterpret_cast<int*>(st + o_pre),
             reinterpret_cast<int4*>(st + o_job), reinterpret_cast<double*>(st + o_beta),
             reinterpret_cast<int*>(st + o_ts), reinterpret_cast<double*>(st + o_tc), prefix[njobs], s);
  HCHECK(hipGetLastError());
  HCHECK(hipStreamSynchronize(s));
  API_END
}

int hcc_lr_precondition(hcc_ctx* ctx, int njobs, const int* jobs, const double* shift, double floor) {
  if (ctx == nullptr || njobs < 0 || (njobs > 0 && (jobs == nullptr || shift == nullptr)) || !(floor > 0.0))
    return HCC_E_ARG;
  if (int r = lr_space_ready(ctx)) return r;
  Batch& b = *ctx->batch;
  std::vector<int> prefix(njobs + 1, 0);
  for (int j = 0; j < njobs; ++j) {
    if (!lr_slot_ok(b, jobs[2 * j], jobs[2 * j + 1])) {
      ctx->err = "precondition job " + std::to_string(j) + ": molecule or slot out of range";
      return HCC_E_ARG;
    }
    prefix[j + 1] = prefix[j] + (int)((b.lr_len[jobs[2 * j]] + kLrChunk - 1) / kLrChunk);
  }
  if (njobs == 0) return HCC_OK;
  API_BEGIN
  HCHECK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const size_t o_job = lr_align(sizeof(int) * (njobs + 1)), o_sh = o_job + lr_align(sizeof(int2) * njobs),
               total = o_sh + sizeof(double) * njobs;
  char* st = lr_stage(b, total);
  HCHECK(hipMemcpyAsync(st, prefix.data(), sizeof(int) * (njobs + 1), hipMemcpyHostToDevice, s));
  HCHECK(hipMemcpyAsync(st + o_job, jobs, sizeof(int2) * njobs, hipMemcpyHostToDevice, s));
  HCHECK(hipMemcpyAsync(st + o_sh, shift, sizeof(double) * njobs, hipMemcpyHostToDevice, s));
  lr_precondition(b.dm, b.dlr_space, b.dlr_base, b.dlr_len, njobs, reinterpret_cast<int*>(st),
                  reinterpret_cast<int2*>(st + o_job), reinterpret_cast<double*>(st + o_sh), floor, prefix[njobs], s);
  HCHECK(hipGetLastError());
  HCHECK(hipStreamSynchronize(s));
  API_END
}

// slots[first .. first + count) lie in the list and name slots of `mol`
static bool lr_slot_run_ok(const Batch& b, int mol, int nslot, const int* slots, int first, int count) {
  if (first < 0 || count < 0 || (long)first + count > nslot) return false;
  for (int q = first; q < first + count; ++q)
    if (!lr_slot_ok(b, mol, slots[q])) return false;
  return true;
}

int hcc_lr_gram(hcc_ctx* ctx, int njobs, const int* jobs, int nslot, const int* slots, double* out) {
  if (ctx == nullptr || njobs < 0 || nslot < 0 || (njobs > 0 && (jobs == nullptr || out == nullptr)) ||
      (nslot > 0 && slots == nullptr))
    return HCC_E_ARG;
  if (int r = lr_space_ready(ctx)) return r;
  Batch& b = *ctx->batch;
  std::vector<LrGramBlk> blk;
  std::vector<int> prefix(1, 0);
  long at = 0;
  for (int j = 0; j < njobs; ++j) {
    const int mol = jobs[5 * j], na = jobs[5 * j + 1], nb = jobs[5 * j + 2], a = jobs[5 * j + 3], c = jobs[5 * j + 4];
    if (mol < 0 || mol >= b.nm || !lr_slot_run_ok(b, mol, nslot, slots, a, na) ||
        !lr_slot_run_ok(b, mol, nslot, slots, c, nb)) {
      ctx->err = "gram job " + std::to_string(j) + ": molecule, slot or slot-list range invalid";
      return HCC_E_ARG;
    }
    const int chunks = (int)((b.lr_len[mol] + kLrGramChunk - 1) / kLrGramChunk);
    for (int p = 0; p < na; p += 16)
      for (int q = 0; q < nb; q += 64) {
        blk.push_back(LrGramBlk{mol, std::min(16, na - p), std::min(64, nb - q), a + p, c + q, (int)(at + (long)p * nb + q),
                                nb, 0});
        prefix.push_back(prefix.back() + chunks);
      }
    at += (long)na * nb;
    if (at > (1L << 30)) {
      ctx->err = "gram: output too large";
      return HCC_E_ARG;
    }
  }
  if (blk.empty()) return HCC_OK;
  API_BEGIN
  HCHECK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int nblk = (int)blk.size(), nchunks = prefix[nblk];
  const size_t o_blk = lr_align(sizeof(int) * (nblk + 1)), o_sl = o_blk + lr_align(sizeof(LrGramBlk) * nblk),
               o_out = o_sl + lr_align(sizeof(int) * (nslot + 1)), o_part = o_out + lr_align(sizeof(double) * at),
               total = o_part + sizeof(double) * 1024 * (size_t)nchunks;
  char* st = lr_stage(b, total);
  HCHECK(hipMemcpyAsync(st, prefix.data(), sizeof(int) * (nblk + 1), hipMemcpyHostToDevice, s));
  HCHECK(hipMemcpyAsync(st + o_blk, blk.data(), sizeof(LrGramBlk) * nblk, hipMemcpyHostToDevice, s));
  HCHECK(hipMemcpyAsync(st + o_sl, slots, sizeof(int) * nslot, hipMemcpyHostToDevice, s));
  double* dout = reinterpret_cast<double*>(st + o_out);
  lr_gram(b.dlr_space, b.dlr_base, b.dlr_len, nblk, reinterpret_cast<int*>(st), reinterpret_cast<LrGramBlk*>(st + o_blk),
          reinterpret_cast<int*>(st + o_sl), reinterpret_cast<double*>(st + o_part), nchunks, dout, s);
  HCHECK(hipGetLastError());
  HCHECK(hipMemcpyAsync(out, dout, sizeof(double) * at, hipMemcpyDeviceToHost, s));
  HCHECK(hipStreamSynchronize(s));
  API_END
}

int hcc_lr_combine_block(hcc_ctx* ctx, int njobs, const int* jobs, const double* beta, int nslot, const int* slots,
                         int ncoef, const double* coef) {
  if (ctx == nullptr || njobs < 0 || nslot < 0 || ncoef < 0 || (njobs > 0 && (jobs == nullptr || beta == nullptr)) ||
      (nslot > 0 && slots == nullptr) || (ncoef > 0 && coef == nullptr))
    return HCC_E_ARG;
  if (int r = lr_space_ready(ctx)) return r;
  Batch& b = *ctx->batch;
  std::vector<LrCombBlk> blk;
  std::vector<int> prefix(1, 0);
  for (int j = 0; j < njobs; ++j) {
    const int mol = jobs[6 * j], m = jobs[6 * j + 1], t = jobs[6 * j + 2], d = jobs[6 * j + 3], sr = jobs[6 * j + 4],
              c = jobs[6 * j + 5];
    bool ok = mol >= 0 && mol < b.nm && lr_slot_run_ok(b, mol, nslot, slots, d, m) &&
              lr_slot_run_ok(b, mol, nslot, slots, sr, t) && c >= 0 && (long)c + (long)m * t <= ncoef;
    // distinct destinations, none among the terms (each element has one writer
    // and no term is read after it was written)
    for (int p = 0; ok && p < m; ++p) {
      for (int q = p + 1; ok && q < m; ++q) ok = slots[d + p] != slots[d + q];
      for (int q = 0; ok && q < t; ++q) ok = slots[d + p] != slots[sr + q];
    }
    if (!ok) {
      ctx->err = "combine-block job " + std::to_string(j) +
                 ": invalid molecule, slot, slot-list or coefficient range, repeated destination, or destination among "
                 "the terms";
      return HCC_E_ARG;
    }
    const int chunks = (int)((b.lr_len[mol] + kLrCombChunk - 1) / kLrCombChunk);
    for (int p = 0; p < m; p += 16) {
      LrCombBlk x{mol, std::min(16, m - p), t, d + p, sr, c + p * t, 0, 0, beta[j]};
      blk.push_back(x);
      prefix.push_back(prefix.back() + chunks);
    }
  }
  if (blk.empty()) return HCC_OK;
  API_BEGIN
  HCHECK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int nblk = (int)blk.size();
  const size_t o_blk = lr_align(sizeof(int) * (nblk + 1)), o_sl = o_blk + lr_align(sizeof(LrCombBlk) * nblk),
               o_cf = o_sl + lr_align(sizeof(int) * (nslot + 1)), total = o_cf + sizeof(double) * (ncoef + 1);
  char* st = lr_stage(b, total);
  HCHECK(hipMemcpyAsync(st, prefix.data(), sizeof(int) * (nblk + 1), hipMemcpyHostToDevice, s));
  HCHECK(hipMemcpyAsync(st + o_blk, blk.data(), sizeof(LrCombBlk) * nblk, hipMemcpyHostToDevice, s));
  if (nslot > 0) HCHECK(hipMemcpyAsync(st + o_sl, slots, sizeof(int) * nslot, hipMemcpyHostToDevice, s));
  if (ncoef > 0) HCHECK(hipMemcpyAsync(st + o_cf, coef, sizeof(double) * ncoef, hipMemcpyHostToDevice, s));
  lr_combine_block(b.dlr_space, b.dlr_base, b.dlr_len, nblk, reinterpret_cast<int*>(st),
                   reinterpret_cast<LrCombBlk*>(st + o_blk), reinterpret_cast<int*>(st + o_sl),
                   reinterpret_cast<double*>(st + o_cf), prefix[nblk], s);
  HCHECK(hipGetLastError());
  HCHECK(hipStreamSynchronize(s));
  API_END
}

// sigma of molecule k's trial vector, slot in_slot[k] of molecule src[k] (k
// when src is null), into slot out_slot[k] of molecule src[k]
static int lr_sigma_from(hcc_ctx* ctx, double h, const int* src, const int* in_slot, const int* out_slot) {
  if (int r = lr_space_ready(ctx)) return r;
  if (!(h > 0.0)) {
    ctx->err = "the difference step must be positive";
    return HCC_E_ARG;
  }
  Batch& b = *ctx->batch;
  const int nm = b.nm;
  std::vector<double*> ptrs(4 * (size_t)nm, nullptr);
  std::set<std::pair<int, int>> written;
  for (int k = 0; k < nm; ++k) {
    if (in_slot[k] < 0) continue;
    const int m = src != nullptr ? src[k] : k;
    const bool same = m >= 0 && m < nm && b.n[m] == b.n[k] && b.o[m] == b.o[k] && b.v[m] == b.v[k];
    if (!same || !lr_slot_ok(b, m, in_slot[k]) || !lr_slot_ok(b, m, out_slot[k]) || in_slot[k] == out_slot[k] ||
        !written.insert({m, out_slot[k]}).second) {
      ctx->err = "sigma of molecule " + std::to_string(k) +
                 ": source molecule of other dimensions, slots out of range or equal, or a sigma slot written twice";
      return HCC_E_ARG;
    }
    const long n1 = (long)b.o[k] * b.v[k];
    double* r = b.dlr_space + b.lr_base[m] + (long)in_slot[k] * b.lr_len[m];
    double* q = b.dlr_space + b.lr_base[m] + (long)out_slot[k] * b.lr_len[m];
    ptrs[k] = r;
    ptrs[nm + k] = r + n1;
    ptrs[2 * nm + k] = q;
    ptrs[3 * nm + k] = q + n1;
  }
  // an input slot that some molecule also writes would be read after its write
  for (int k = 0; k < nm; ++k)
    if (in_slot[k] >= 0 && written.count({src != nullptr ? src[k] : k, in_slot[k]})) {
      ctx->err = "sigma of molecule " + std::to_string(k) + ": its input slot is another product's output";
      return HCC_E_ARG;
    }
  API_BEGIN
  HCHECK(hipSetDevice(ctx->device));
  HCHECK(hipMemcpyAsync(b.dlr_ps, ptrs.data(), ptrs.size() * sizeof(double*), hipMemcpyHostToDevice, ctx->stream));
  lr_sigma_run(ctx, h, b.dlr_ps);
  HCHECK(hipStreamSynchronize(ctx->stream));
  API_END
}

int hcc_lr_sigma_slots(hcc_ctx* ctx, double h, const int* in_slot, const int* out_slot) {
  if (ctx == nullptr || in_slot == nullptr || out_slot == nullptr) return HCC_E_ARG;
  return lr_sigma_from(ctx, h, nullptr, in_slot, out_slot);
}

int hcc_lr_sigma_from(hcc_ctx* ctx, double h, const int* src, const int* in_slot, const int* out_slot) {
  if (ctx == nullptr || src == nullptr || in_slot == nullptr || out_slot == nullptr) return HCC_E_ARG;
  return lr_sigma_from(ctx, h, src, in_slot, out_slot);
}

int hcc_batch_get_orbitals(hcc_ctx* ctx, int mol, double* C, double* F_mo, double* e_ref) {
  if (ctx == nullptr) return HCC_E_ARG;
  if (int r = need_batch(ctx)) return r;
  Batch& b = *ctx->batch;
  if (mol < 0 || mol >= b.nm) {
    ctx->err = "invalid molecule index";
    return HCC_E_ARG;
  }
  API_BEGIN
  const int N = b.n[mol];
  std::vector<double> h((size_t)N * N);
  double* src[2] = {b.hm[mol].C, b.hm[mol].F};
  double* dst[2] = {C, F_mo};
  for (int q = 0; q < 2; ++q) {
    if (dst[q] == nullptr) continue;
    HCHECK(hipMemcpy(h.data(), src[q], h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int r = 0; r < N; ++r)
      for (int c = 0; c < N; ++c) dst[q][r + (size_t)c * N] = h[(size_t)r * N + c];
  }
  if (e_ref) *e_ref = b.eref[mol];
  API_END
}

static int need_batch(hcc_ctx* ctx) {
  if (ctx->batch == nullptr) {
    ctx->err = "no molecule loaded";
    return HCC_E_STATE;
  }
  return HCC_OK;
}

static int check_guess(hcc_ctx* ctx, int kind) {
  if (kind != HCC_GUESS_REFERENCE && kind != HCC_GUESS_MP2 && kind != HCC_GUESS_LEARNED &&
      kind != HCC_GUESS_CURRENT) {
    ctx->err = "unsupported guess kind";
    return HCC_E_ARG;
  }
  if (kind == HCC_GUESS_CURRENT) {
    const Batch& b = *ctx->batch;
    bool all = (int)b.amp_valid.size() == b.nm;
    for (size_t k = 0; all && k < b.amp_valid.size(); ++k) all = b.amp_valid[k] != 0;
    if (!all) {
      ctx->err = "HCC_GUESS_CURRENT needs amplitudes for every molecule (a solve, a guess or "
                 "hcc_batch_set_amplitudes)";
      return HCC_E_STATE;
    }
  }
  if (kind == HCC_GUESS_LEARNED && ctx->guess_net == nullptr) {
    ctx->err = "no guess network set (hcc_set_guess_network)";
    return HCC_E_STATE;
  }
  return HCC_OK;
}

int hcc_set_guess_network(hcc_ctx* ctx, const float* params, int count) {
  if (ctx == nullptr || params == nullptr) return HCC_E_ARG;
  if (count != HCC_MLP_PARAMS) {
    ctx->err = "parameter count must be HCC_MLP_PARAMS";
    return HCC_E_ARG;
  }
  API_BEGIN
  HCHECK(hipSetDevice(ctx->device));
  if (ctx->guess_net == nullptr) HCHECK(hipMalloc((void**)&ctx->guess_net, HCC_MLP_PARAMS * sizeof(float)));
  HCHECK(hipMemcpy(ctx->guess_net, params, HCC_MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice));
  API_END
}

int hcc_guess(hcc_ctx* ctx, int kind) {
  if (ctx == nullptr) return HCC_E_ARG;
  if (int r = need_batch(ctx)) return r;
  if (int r = check_guess(ctx, kind)) return r;
  API_BEGIN
  Batch& b = *ctx->batch;
  apply_guess(ctx, kind);
  HCHECK(hipStreamSynchronize(ctx->stream));
  ctx->have_guess = true;
  API_END
}

int hcc_iterate(hcc_ctx* ctx, int iteration, double eps, double accuracy[5]) {
  if (ctx == nullptr || accuracy == nullptr) return HCC_E_ARG;
  if (int r = need_batch(ctx)) return r;
  if (!ctx->have_guess) {
    ctx->err = "hcc_guess must precede hcc_iterate";
    return HCC_E_STATE;
  }
  (void)iteration;
  API_BEGIN
  Batch& b = *ctx->batch;
  activate(b.dst, b.dact, b.nm, ctx->stream);
  Params prm = params(ctx, eps, 0, false);
  run_iteration(ctx, prm, nullptr);
  State S;
  HCHECK(hipMemcpyAsync(&S, b.dst, sizeof(State), hipMemcpyDeviceToHost, ctx->stream));
  HCHECK(hipStreamSynchronize(ctx->stream));
  accuracy[0] = S.acc[0];
  accuracy[1] = S.acc[1];
  accuracy[2] = accuracy[3] = accuracy[4] = -1.0;
  API_END
}

int hcc_energy(hcc_ctx* ctx, double energy_out[5]) {
  if (ctx == nullptr || energy_out == nullptr) return HCC_E_ARG;
  if (int r = need_batch(ctx)) return r;
  API_BEGIN
  Batch& b = *ctx->batch;
  energy(b.dm, b.dst, b.nm, b.rgT2S, b.partE, ctx->stream);
  State S;
  HCHECK(hipMemcpyAsync(&S, b.dst, sizeof(State), hipMemcpyDeviceToHost, ctx->stream));
  HCHECK(hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < 5; ++k) energy_out[k] = S.energy[k];
  API_END
}

// t1, t2 of molecule 0 into the reference's column-major layouts (either may be null)
static void copy_amplitudes_out(hcc_ctx* ctx, double* t1, double* t2, int mol = 0) {
  Batch& b = *ctx->batch;
  const int o = b.o[mol], v = b.v[mol];
  std::vector<double> h1((size_t)o * v), h2((size_t)o * o * v * v);
  HCHECK(hipMemcpyAsync(h1.data(), b.hm[mol].t1, h1.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
  if (t2) HCHECK(hipMemcpyAsync(h2.data(), b.hm[mol].t2, h2.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
  HCHECK(hipStreamSynchronize(ctx->stream));
  if (t1)
    for (int i = 0; i < o; ++i)
      for (int a = 0; a < v; ++a) t1[i + (size_t)a * o] = h1[(size_t)i * v + a];
  if (t2)
    for (int i = 0; i < o; ++i)
      for (int j = 0; j < o; ++j)
        for (int a = 0; a < v; ++a)
          for (int bb = 0; bb < v; ++bb)
            t2[i + (size_t)o * (j + (size_t)o * (a + (size_t)v * bb))] =
                h2[(((size_t)i * o + j) * v + a) * v + bb];
}

int hcc_get_amplitudes(hcc_ctx* ctx, double* t1, double* t2) {
  if (ctx == nullptr) return HCC_E_ARG;
  if (int r = need_batch(ctx)) return r;
  API_BEGIN
  copy_amplitudes_out(ctx, t1, t2);
  API_END
}

int hcc_batch_get_amplitudes(hcc_ctx* ctx, int mol, double* t1, double* t2) {
  if (ctx == nullptr) return HCC_E_ARG;
  if (int r = need_batch(ctx)) return r;
  if (mol < 0 || mol >= ctx->batch->nm) {
    ctx->err = "molecule index out of range";
    return HCC_E_ARG;
  }
  API_BEGIN
  HCHECK(hipSetDevice(ctx->device));
  copy_amplitudes_out(ctx, t1, t2, mol);
  API_END
}

static void copy_amplitudes_in(hcc_ctx* ctx, int mol, const double* t1, const double* t2) {
  Batch& b = *ctx->batch;
  const int o = b.o[mol], v = b.v[mol];
  std::vector<double> h1((size_t)o * v), h2((size_t)o * o * v * v);
  for (int i = 0; i < o; ++i)
    for (int a = 0; a < v; ++a) h1[(size_t)i * v + a] = t1[i + (size_t)a * o];
  for (int i = 0; i < o; ++i)
    for (int j = 0; j < o; ++j)
      for (int a = 0; a < v; ++a)
        for (int bb = 0; bb < v; ++bb)
          h2[(((size_t)i * o + j) * v + a) * v + bb] = t2[i + (size_t)o * (j + (size_t)o * (a + (size_t)v * bb))];
  HCHECK(hipMemcpyAsync(b.hm[mol].t1, h1.data(), h1.size() * 8, hipMemcpyHostToDevice, ctx->stream));
  HCHECK(hipMemcpyAsync(b.hm[mol].t2, h2.data(), h2.size() * 8, hipMemcpyHostToDevice, ctx->stream));
  HCHECK(hipStreamSynchronize(ctx->stream));
  if ((int)b.amp_valid.size() != b.nm) b.amp_valid.assign(b.nm, 0);
  b.amp_valid[mol] = 1;
}

int hcc_batch_set_amplitudes(hcc_ctx* ctx, int mol, const double* t1, const double* t2) {
  if (ctx == nullptr || t1 == nullptr || t2 == nullptr) return HCC_E_ARG;
  if (int r = need_batch(ctx)) return r;
  if (mol < 0 || mol >= ctx->batch->nm) {
    ctx->err = "molecule index out of range";
    return HCC_E_ARG;
  }
  API_BEGIN
  HCHECK(hipSetDevice(ctx->device));
  copy_amplitudes_in(ctx, mol, t1, t2);
  API_END
}

int hcc_set_amplitudes(hcc_ctx* ctx, const double* t1, const double* t2) {
  if (ctx == nullptr || t1 == nullptr || t2 == nullptr) return HCC_E_ARG;
  if (int r = need_batch(ctx)) return r;
  API_BEGIN
  Batch& b = *ctx->batch;
  copy_amplitudes_in(ctx, 0, t1, t2);
  if (!ctx->have_guess) {
    reset_state(b.dst, b.dact, b.nm, ctx->stream);
    HCHECK(hipStreamSynchronize(ctx->stream));
  }
  ctx->have_guess = true;
  API_END
}

int hcc_residual(hcc_ctx* ctx, double* d1, double* d2) {
  if (ctx == nullptr || d1 == nullptr || d2 == nullptr) return HCC_E_ARG;
  if (int r = need_batch(ctx)) return r;
  API_BEGIN
  Batch& b = *ctx->batch;
  activate(b.dst, b.dact, b.nm, ctx->stream);
  residual(ctx);
  const int o = b.o[0], v = b.v[0];
  const size_t n1 = (size_t)o * v, n2 = (size_t)o * o * v * v;
  if (b.d1out == nullptr) {
    b.d1out = b.dalloc<double>(n1);
    b.d2out = b.dalloc<double>(n2);
  }
  ew_contravariant(b.dm, b.rgT2S, b.nm, b.d1out, b.d2out, ctx->stream);
  std::vector<double> h1(n1), h2(n2);
  HCHECK(hipMemcpyAsync(h1.data(), b.d1out, n1 * 8, hipMemcpyDeviceToHost, ctx->stream));
  HCHECK(hipMemcpyAsync(h2.data(), b.d2out, n2 * 8, hipMemcpyDeviceToHost, ctx->stream));
  HCHECK(hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < o; ++i)
    for (int a = 0; a < v; ++a) d1[i + (size_t)a * o] = h1[(size_t)i * v + a];
  for (int i = 0; i < o; ++i)
    for (int j = 0; j < o; ++j)
      for (int a = 0; a < v; ++a)
        for (int bb = 0; bb < v; ++bb)
          d2[i + (size_t)o * (j + (size_t)o * (a + (size_t)v * bb))] = h2[(((size_t)i * o + j) * v + a) * v + bb];
  API_END
}

// Executable graph of the launches `body` enqueues on stream s (captured, not
// run).  The body must not synchronize, copy to the host or memset.
struct GraphExec {
  hipGraphExec_t e = nullptr;
  ~GraphExec() {
    if (e) (void)hipGraphExecDestroy(e);
  }
  void launch(hipStream_t s) { HCHECK(hipGraphLaunch(e, s)); }
  void reset() {
    if (e) (void)hipGraphExecDestroy(e);
    e = nullptr;
  }
};

static void capture(GraphExec& ge, hipStream_t s, const std::function<void()>& body) {
  hipGraph_t g = nullptr;
  HCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    body();
  } catch (...) {
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  HCHECK(hipStreamEndCapture(s, &g));
  const hipError_t r = hipGraphInstantiate(&ge.e, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HCHECK(r);
}

static int run_solve(hcc_ctx* ctx, int guess_kind, int maxiter, double eps, bool build_integrals) {
  Batch& b = *ctx->batch;
  hipStream_t s = ctx->stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ctx->timing) {
    HCHECK(hipEventCreate(&e0));
    HCHECK(hipEventCreate(&e1));
    HCHECK(hipEventRecord(e0, s));
  }
  if (build_integrals) precompute(ctx);
  apply_guess(ctx, guess_kind);
  Params prm = params(ctx, eps, maxiter, true);
  std::vector<int> act(b.nm);
  const int check_every = 4;
  int it = 0;
  double iters_total = 0.0;
  ctx->stop_req.store(0);   // each iterator run starts unflagged (gIterator.f90:57)
  // blocks of check_every iterations replay one captured graph: every kernel
  // reads its molecule's state on the device, so the launches never change
  GraphExec block;
  const bool graphs = ctx->graphs && !ctx->prof && maxiter >= check_every;
  auto capture_block = [&]() {
    capture(block, s, [&]() {
      for (int q = 0; q < check_every; ++q) run_iteration(ctx, prm, nullptr);
    });
  };
  UncompactGuard uncompact{b};
  if (graphs) capture_block();
  while (it < maxiter) {
    if (graphs && it + check_every <= maxiter) {
      block.launch(s);
      it += check_every;
    } else {
      for (int q = 0; q < check_every && it < maxiter; ++q, ++it) run_iteration(ctx, prm, nullptr);
    }
    HCHECK(hipMemcpyAsync(act.data(), b.dact, b.nm * sizeof(int), hipMemcpyDeviceToHost, s));
    HCHECK(hipStreamSynchronize(s));
    if (ctx->prof) prof_collect(ctx);
    int alive = 0;
    for (int a : act) alive += a;
    if (alive == 0 || take_stop(ctx)) break;
    const bool tables = compact_tables(ctx, false, act, alive, false);
    if ((compact_raggeds(ctx, act, alive) || tables) && graphs) {
      block.reset();
      capture_block();
    }
  }
  uncompact_tables(b);
  prof_open(ctx, kProfEnergy);
  energy(b.dm, b.dst, b.nm, b.rgT2S, b.partE, s);
  prof_close(ctx);
  float ms = 0.f;
  if (ctx->timing) {
    HCHECK(hipEventRecord(e1, s));
    HCHECK(hipEventSynchronize(e1));
    HCHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  } else {
    HCHECK(hipStreamSynchronize(s));
  }
  if (ctx->prof) prof_collect(ctx);
  ctx->last_ms = ms;
  std::vector<State> S(b.nm);
  HCHECK(hipMemcpy(S.data(), b.dst, b.nm * sizeof(State), hipMemcpyDeviceToHost));
  for (auto& x : S) iters_total += x.iter;
  ctx->last_iters = iters_total;
  ctx->have_guess = true;
  return HCC_OK;
}

// BF16 copies of the outer-point operands (variable arenas; constants too when asked)
// With `act` (the running molecules, `alive` of them) and fewer than
// kCompactFrac of the molecules running, the variable copies are refreshed for
// the running molecules' blocks only: the others' copies are not read again in
// this solve.
static void refresh_shadows(hcc_ctx* c, bool constants, const std::vector<int>* act = nullptr, int alive = 0) {
  Batch& b = *c->batch;
  const bool runs = act != nullptr && alive < kCompactFrac * b.nm;
  for (auto& sh : b.shadows) {
    if (!sh.used || !(sh.variable || constants)) continue;
    if (!runs || !sh.variable) {
      to_bf16(sh.base, sh.copy, sh.count, c->stream);
      continue;
    }
    // the previous upload from hpre has completed (a refresh follows a
    // synchronisation of the stream)
    sh.hpre.assign(b.nm + 1, 0);
    int blocks = 0;
    for (int k = 0; k < b.nm; ++k) {
      sh.hpre[k] = blocks;
      if ((*act)[k] != 0) blocks += (int)((sh.hlen[k] + 4 * kEwThreads - 1) / (4 * kEwThreads));
    }
    sh.hpre[b.nm] = blocks;
    HCHECK(hipMemcpyAsync(sh.dpre, sh.hpre.data(), (b.nm + 1) * sizeof(int), hipMemcpyHostToDevice, c->stream));
    to_bf16_runs(sh.base, sh.copy, sh.doff, sh.dlen, sh.dpre, b.nm, blocks, c->stream);
  }
  HCHECK(hipGetLastError());
}

// Defect-correction solve: FP64 residual at the outer point t0, then inner
// iterations on the correction dt whose residual is Omega(t0) + dOmega with
// dOmega from correction-form GEMMs (BF16 operands, FP32 accumulation), until
// max|d| < kappa * max|d(t0)|; then t0 += dt.  The FP64 residual at each outer
// point decides convergence, so the answer is the FP64 fixed point.
constexpr int kInnerBlock = 2;

static int run_solve_mixed(hcc_ctx* ctx, int guess_kind, int maxiter, double eps, bool build_integrals) {
  Batch& b = *ctx->batch;
  hipStream_t s = ctx->stream;
  hipEvent_t e0, e1;
  HCHECK(hipEventCreate(&e0));
  HCHECK(hipEventCreate(&e1));
  HCHECK(hipEventRecord(e0, s));
  if (build_integrals) precompute(ctx);
  apply_guess(ctx, guess_kind);
  Params prm = params(ctx, eps, maxiter, true);
  prm.kappa = ctx->kappa;
  prm.max_inner = ctx->max_inner;
  prm.keep_diis = ctx->keep_diis;
  // The inner solve is a fixed-point iteration on the correction equation,
  // not the reference's iteration history: it takes the Jacobi step dt = Omega/D
  // (the '2-1' update (2 d2 + d2^T)/D is 3 Omega/D, hence 1/3 there).
  prm.s1 = 1.0;
  prm.s2 = ctx->projtype == 2 ? 1.0 / 3.0 : 1.0;
  Params prm0 = prm;   // residual-norm pass at the outer point: no step, no DIIS
  prm0.s1 = prm0.s2 = 0.0;
  prm0.diis = 0;
  std::vector<int> act(b.nm);
  ctx->stop_req.store(0);   // each iterator run starts unflagged (gIterator.f90:57)
  bool stopped = false;
  GraphExec inner_block;
  UncompactGuard uncompact{b};
  // outer = true: after the outer control (the running set, which the next
  // inner phase and the next outer residual start from); false: after inner
  // steps (the molecules still in their inner solve)
  auto alive = [&](bool outer) {
    HCHECK(hipMemcpyAsync(act.data(), b.dact, b.nm * sizeof(int), hipMemcpyDeviceToHost, s));
    HCHECK(hipStreamSynchronize(s));
    if (ctx->prof) prof_collect(ctx);
    int n = 0;
    for (int a : act) n += a;
    if (n > 0 && take_stop(ctx)) {
      stopped = true;
      return 0;
    }
    if (outer) {
      compact_tables(ctx, false, act, n, false);
      if (compact_raggeds(ctx, act, n)) inner_block.reset();
    }
    if (compact_tables(ctx, true, act, n, outer)) inner_block.reset();
    return n;
  };
  auto inner_iteration = [&]() {
    residual_delta(ctx);
    prof_open(ctx, kProfUpdateMixed);
    mixed_update(b.dm, b.ddm, b.dst, b.rgOV, b.rgT2x, b.nm, prm, b.part1, b.part2, s);
    prof_close(ctx);
    prof_open(ctx, kProfControl);
    mixed_inner_control(b.ddm, b.dst, b.dact, b.nm, prm, b.rgOV, b.part1, b.rgT2x, b.part2, s);
    prof_close(ctx);
    if (prm.diis) {
      prof_open(ctx, kProfDiisMixed);
      diis_step(b.ddm, b.dst, b.nm, b.rgND, b.rgP2, b.partD, s);
      mixed_inner_advance(b.ddm, b.dst, b.nm, s);
      prof_close(ctx);
    }
  };
  // The first step of every inner solve starts from dt = 0, whose correction
  // residual is zero: it takes the update from the outer FP64 residual
  // without evaluating the correction contractions.
  Params prm1 = prm;
  prm1.outer_only = 1;
  auto first_step = [&]() {
    prof_open(ctx, kProfUpdateMixed);
    mixed_update(b.dm, b.ddm, b.dst, b.rgOV, b.rgT2x, b.nm, prm1, b.part1, b.part2, s);
    prof_close(ctx);
    prof_open(ctx, kProfControl);
    mixed_inner_control(b.ddm, b.dst, b.dact, b.nm, prm1, b.rgOV, b.part1, b.rgT2x, b.part2, s);
    prof_close(ctx);
    if (prm1.diis) {
      prof_open(ctx, kProfDiisMixed);
      diis_step(b.ddm, b.dst, b.nm, b.rgND, b.rgP2, b.partD, s);
      mixed_inner_advance(b.ddm, b.dst, b.nm, s);
      prof_close(ctx);
    }
  };
  // blocks of kInnerBlock inner iterations replay one captured graph (see
  // run_solve); the host checks between blocks whether any molecule is still
  // in its inner solve, so short blocks cut the iterations that run after the
  // last molecule has finished
  const bool graphs = ctx->graphs && !ctx->prof;
  for (int outer = 0; outer < ctx->max_outer + 1; ++outer) {
    set_phase(b.dst, b.dact, b.nm, 0, s);
    residual(ctx);
    prof_open(ctx, kProfOuterNorm);
    ew_update(b.dm, b.dst, b.rgOV, b.rgT2x, b.nm, prm0, b.part1, b.part2, s);
    mixed_outer_control(b.dm, b.dst, b.dact, b.nm, prm, ctx->max_outer, b.rgOV, b.part1, b.rgT2x, b.part2, s);
    prof_close(ctx);
    const int n_running = alive(true);
    if (n_running == 0 || stopped) break;
    prof_open(ctx, kProfCommit);
    set_phase(b.dst, b.dact, b.nm, 1, s);
    if (ctx->mixed_bf16) refresh_shadows(ctx, outer == 0, &act, n_running);
    mixed_zero(b.dm, b.ddm, b.dst, b.rgT2Sx, b.nm, s);
    prof_close(ctx);
    first_step();
    for (int inner = 1; inner < ctx->max_inner;) {
      if (graphs && inner + kInnerBlock <= ctx->max_inner) {
        if (!inner_block.e) capture(inner_block, s, [&]() {
            for (int q = 0; q < kInnerBlock; ++q) inner_iteration();
          });
        inner_block.launch(s);
        inner += kInnerBlock;
      } else {
        for (int q = 0; q < kInnerBlock && inner < ctx->max_inner; ++q, ++inner) inner_iteration();
      }
      if (alive(false) == 0) break;
    }
    if (stopped) break;
    prof_open(ctx, kProfCommit);
    mixed_commit(b.dm, b.ddm, b.dst, b.rgT2Sx, b.rgP2, b.nm, prm.keep_diis && prm.diis, s);
    prof_close(ctx);
  }
  uncompact_tables(b);
  set_phase(b.dst, b.dact, b.nm, 0, s);
  prof_open(ctx, kProfEnergy);
  energy(b.dm, b.dst, b.nm, b.rgT2S, b.partE, s);
  prof_close(ctx);
  HCHECK(hipEventRecord(e1, s));
  HCHECK(hipEventSynchronize(e1));
  if (ctx->prof) prof_collect(ctx);
  float ms = 0.f;
  HCHECK(hipEventElapsedTime(&ms, e0, e1));
  ctx->last_ms = ms;
  std::vector<State> S(b.nm);
  HCHECK(hipMemcpy(S.data(), b.dst, b.nm * sizeof(State), hipMemcpyDeviceToHost));
  double its = 0.0;
  for (auto& x : S) its += x.iter;
  ctx->last_iters = its;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  ctx->have_guess = true;
  return HCC_OK;
}

int hcc_solve(hcc_ctx* ctx, int guess_kind, int maxiter, double eps, double energy_out[5], int* iterations,
              int* status) {
  if (ctx == nullptr) return HCC_E_ARG;
  if (int r = need_batch(ctx)) return r;
  if (int r = check_guess(ctx, guess_kind)) return r;
  if (maxiter < 1) {
    ctx->err = "maxiter must be positive";
    return HCC_E_ARG;
  }
  API_BEGIN
  if (ctx->precision == HCC_MIXED) run_solve_mixed(ctx, guess_kind, maxiter, eps, false);
  else run_solve(ctx, guess_kind, maxiter, eps, false);
  State S;
  HCHECK(hipMemcpy(&S, ctx->batch->dst, sizeof(State), hipMemcpyDeviceToHost));
  if (energy_out)
    for (int k = 0; k < 5; ++k) energy_out[k] = S.energy[k];
  if (iterations) *iterations = S.iter;
  if (status) *status = S.status;
  API_END
}

int hcc_batch_solve(hcc_ctx* ctx, int guess_kind, int maxiter, double eps) {
  if (ctx == nullptr) return HCC_E_ARG;
  if (int r = need_batch(ctx)) return r;
  if (int r = check_guess(ctx, guess_kind)) return r;
  if (maxiter < 1) {
    ctx->err = "maxiter must be positive";
    return HCC_E_ARG;
  }
  API_BEGIN
  if (ctx->precision == HCC_MIXED) run_solve_mixed(ctx, guess_kind, maxiter, eps, true);
  else run_solve(ctx, guess_kind, maxiter, eps, true);
  API_END
}

int hcc_batch_results(hcc_ctx* ctx, double* energies, int* iterations, int* status) {
  if (ctx == nullptr) return HCC_E_ARG;
  if (int r = need_batch(ctx)) return r;
  API_BEGIN
  Batch& b = *ctx->batch;
  std::vector<State> S(b.nm);
  HCHECK(hipMemcpy(S.data(), b.dst, b.nm * sizeof(State), hipMemcpyDeviceToHost));
  for (int k = 0; k < b.nm; ++k) {
    if (energies)
      for (int q = 0; q < 5; ++q) energies[5 * k + q] = S[k].energy[q];
    if (iterations) iterations[k] = S[k].iter;
    if (status) status[k] = S[k].status;
  }
  API_END
}

int hcc_last_timing(hcc_ctx* ctx, double* solve_ms, double* iterations_total) {
  if (ctx == nullptr) return HCC_E_ARG;
  if (solve_ms) *solve_ms = ctx->last_ms;
  if (iterations_total) *iterations_total = ctx->last_iters;
  return HCC_OK;
}

int hcc_profile(hcc_ctx* ctx, int enable) {
  if (ctx == nullptr) return HCC_E_ARG;
  ctx->prof = enable != 0;
  ctx->prof_used = 0;
  ctx->prof_ms = 0.0;
  ctx->prof_launches = 0;
  for (double& x : ctx->prof_op_ms) x = 0.0;
  for (long& x : ctx->prof_op_launches) x = 0;
  return HCC_OK;
}

int hcc_profile_ops(hcc_ctx* ctx, double* ms96) {
  if (ctx == nullptr || ms96 == nullptr) return HCC_E_ARG;
  for (int k = 0; k < kProfSlots; ++k) ms96[k] = ctx->prof_op_ms[k];
  return HCC_OK;
}

int hcc_profile_launches(hcc_ctx* ctx, long* n96) {
  if (ctx == nullptr || n96 == nullptr) return HCC_E_ARG;
  for (int k = 0; k < kProfSlots; ++k) n96[k] = ctx->prof_op_launches[k];
  return HCC_OK;
}

int hcc_profile_read(hcc_ctx* ctx, double* gemm_ms, long* gemm_launches) {
  if (ctx == nullptr) return HCC_E_ARG;
  if (gemm_ms) *gemm_ms = ctx->prof_ms;
  if (gemm_launches) *gemm_launches = ctx->prof_launches;
  return HCC_OK;
}

int hcc_gemm_flops_correction(hcc_ctx* ctx, double* per_inner_iteration) {
  if (ctx == nullptr || per_inner_iteration == nullptr) return HCC_E_ARG;
  if (int r = need_batch(ctx)) return r;
  API_BEGIN
  Batch& b = *ctx->batch;
  for (int k = 0; k < b.nm; ++k) per_inner_iteration[k] = 0.0;
  std::vector<GemmProb> h;
  std::function<void(const GemmTable&)> add = [&](const GemmTable& t) {
    h.resize(t.nprob);
    if (t.nprob) HCHECK(hipMemcpy(h.data(), t.d, t.nprob * sizeof(GemmProb), hipMemcpyDeviceToHost));
    for (const auto& p : h)   // A0 dB and dA (B0 + dB): one product per changing operand
      per_inner_iteration[p.mol] += 2.0 * p.M * (double)p.N * p.K * ((p.dA != nullptr) + (p.dB != nullptr));
    if (t.alt >= 0) add(b.alt[t.alt]);
  };
  for (const auto& t : b.dit) add(t);
  API_END
}

int hcc_set_graphs(hcc_ctx* ctx, int enable) {
  if (ctx == nullptr) return HCC_E_ARG;
  ctx->graphs = enable ? 1 : 0;
  return HCC_OK;
}

int hcc_set_gemm_tuning(hcc_ctx* ctx, int count, const int* entries) {
  if (ctx == nullptr || count < 0 || (count > 0 && entries == nullptr)) return HCC_E_ARG;
  std::map<std::array<int, 3>, std::array<int, 2>> m;
  for (int q = 0; q < count; ++q) {
    const int* e = entries + 5 * q;
    const bool dual = e[0] >= kDualTuneOp;
    const bool tile_ok = dual ? (e[3] == 32 || e[3] == 64 || e[3] == 256) : (e[3] == 32 || e[3] == 64 || e[3] == 128);
    if (e[0] < 0 || e[1] < 1 || e[2] < 1 || !tile_ok || e[4] < 1 || e[4] > 256) {
      ctx->err = "invalid tuning entry (op, n_occ, n_virt, tile 32|64|128 -- 32|64|256 for correction ops "
                 "100 + op --, ksplit 1..256)";
      return HCC_E_ARG;
    }
    m[{e[0], e[1], e[2]}] = {e[3], e[4]};
  }
  ctx->tuning.swap(m);
  return HCC_OK;
}

int hcc_gemm_shape(hcc_ctx* ctx, int op, int mol, int* mnk) {
  if (ctx == nullptr || mnk == nullptr) return HCC_E_ARG;
  if (int r = need_batch(ctx)) return r;
  Batch& b = *ctx->batch;
  if (op < 0 || op >= (int)b.ops.size() || mol < 0 || mol >= b.nm) {
    ctx->err = "invalid op or molecule";
    return HCC_E_ARG;
  }
  const GemmProb p = b.ops[op](b.hm[mol]);
  mnk[0] = p.M;
  mnk[1] = p.N;
  mnk[2] = p.K;
  return HCC_OK;
}

int hcc_tune_gemm(hcc_ctx* ctx, int op, int tile, int ksplit, int reps, double* ms) {
  if (ctx == nullptr || ms == nullptr || reps < 1) return HCC_E_ARG;
  if (int r = need_batch(ctx)) return r;
  Batch& b = *ctx->batch;
  const bool dual = op >= kDualTuneOp;
  const int q = dual ? op - kDualTuneOp : op;
  const bool tile_ok = tile == 0 || tile == 32 || tile == 64 || (dual ? tile == 256 : tile == 128);
  if (q < 0 || q >= (int)b.ops.size() || !tile_ok || ksplit < 0 || (dual && !(b.mixed && b.mixed_bf16))) {
    ctx->err = "invalid tuning request (correction ops 100 + op need a mixed-precision BF16 context)";
    return HCC_E_ARG;
  }
  API_BEGIN
  HCHECK(hipSetDevice(ctx->device));
  const size_t nalloc = b.allocs.size(), nalt = b.alt.size();
  GemmTable t;
  if (dual)
    t = tile == 0 ? b.dit[q]
                  : make_table(b, [&](int k) { return delta_prob(b, q, k); }, true, tile, -1, ksplit > 0 ? ksplit : 1,
                               true);
  else
    t = tile == 0 ? b.it[q]
                  : make_table(b, [&](int k) { return b.ops[q](b.hm[k]); }, true, tile, -1, ksplit > 0 ? ksplit : 1);
  auto run = [&]() {
    if (!dual) {
      launch_table(ctx, t);
      return;
    }
    for (const GemmTable* p = &t; p != nullptr; p = p->alt >= 0 ? &b.alt[p->alt] : nullptr) {
      if (p->nprob == 0) continue;
      HCHECK(launch_gemm_dual(p->d, p->tile_prob, p->tiles, b.dact, ctx->stream, p->flags, true, p->tile));
      if (p->red_blocks > 0)
        HCHECK(launch_splitk_reduce(p->d, p->red_prefix, p->red_bprob, p->red_blocks, b.dact, ctx->stream));
    }
  };
  reset_state(b.dst, b.dact, b.nm, ctx->stream);   // every molecule active
  hipEvent_t e0, e1;
  HCHECK(hipEventCreate(&e0));
  HCHECK(hipEventCreate(&e1));
  run();
  HCHECK(hipEventRecord(e0, ctx->stream));
  for (int r = 0; r < reps; ++r) run();
  HCHECK(hipEventRecord(e1, ctx->stream));
  HCHECK(hipEventSynchronize(e1));
  float el = 0.f;
  HCHECK(hipEventElapsedTime(&el, e0, e1));
  *ms = el / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (size_t q = nalloc; q < b.allocs.size(); ++q) HCHECK(hipFree(b.allocs[q]));
  b.allocs.resize(nalloc);
  b.alt.resize(nalt);
  ctx->have_guess = false;
  API_END
}

int hcc_gemm_flops(hcc_ctx* ctx, double* per_iteration, double* per_solve_setup) {
  if (ctx == nullptr) return HCC_E_ARG;
  if (int r = need_batch(ctx)) return r;
  std::vector<GemmProb> h;
  API_BEGIN
  Batch& b = *ctx->batch;
  std::function<void(const GemmTable&, std::vector<double>&)> table_flops = [&](const GemmTable& t,
                                                                                std::vector<double>& acc) {
    h.resize(t.nprob);
    if (t.nprob)
      HCHECK(hipMemcpy(h.data(), t.d, t.nprob * sizeof(GemmProb), hipMemcpyDeviceToHost));
    for (const auto& p : h) acc[p.mol] += 2.0 * p.M * (double)p.N * p.K;
    if (t.alt >= 0) table_flops(b.alt[t.alt], acc);
  };
  std::vector<double> it(b.nm, 0.0), pre(b.nm, 0.0);
  for (const auto& t : b.it) table_flops(t, it);
  for (const auto& t : b.pre) table_flops(t, pre);
  table_flops(b.it[G1_Q], pre);
  for (int k = 0; k < b.nm; ++k) {
    if (per_iteration) per_iteration[k] = it[k];
    if (per_solve_setup) per_solve_setup[k] = pre[k];
  }
  API_END
}

void hcc_destroy(hcc_ctx* ctx) {
  if (ctx == nullptr) return;
  (void)hipSetDevice(ctx->device);
  for (auto& ev : ctx->prof_events) {
    (void)hipEventDestroy(ev.first);
    (void)hipEventDestroy(ev.second);
  }
  delete ctx->batch;
  if (ctx->guess_net) (void)hipFree(ctx->guess_net);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* hcc_last_error(const hcc_ctx* ctx) {
  if (ctx == nullptr) return "null context";
  return ctx->err.c_str();
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Solve without device allocation (hcc_solve_noalloc).  Every device buffer of
// the solve -- the molecule's arenas, descriptors, block tables, GEMM tables
// and partial results -- is carved, 256-byte aligned, from the file-scope
// array `scratch`, whose address on each device comes from
// hipGetSymbolAddress once per device (the array's name is only a host-side
// stand-in for it).  Calls on one device share the array and a context, so
// they must not overlap; each call rebuilds everything it reads, and the state
// and the partial results it accumulates are cleared by the solve itself.
// ---------------------------------------------------------------------------
// (no `used` attribute: it would keep clang from exporting the array under the
// name the host registers, and the runtime's lookup by that name would fail)
constexpr size_t kNoAllocBytes = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
struct NoAllocDevice {
  char* ptr = nullptr;       // this device's copy of `scratch`
  hcc_ctx* ctx = nullptr;    // null stream, no graphs, no timing events
};
std::mutex g_noalloc_mu;
std::map<int, NoAllocDevice> g_noalloc;

void noalloc_message(char* err, int len, const std::string& m) {
  if (err != nullptr && len > 0) snprintf(err, (size_t)len, "%s", m.c_str());
}
}  // namespace

extern "C" {

int hcc_solve_noalloc(int device, int n, int n_occ, const double* C, const double* F_mo, const double* gamma,
                      double e_ref, int guess_kind, int maxiter, double eps, int diis_steps, double energy_out[5],
                      int* iterations, int* status, double* t1, double* t2, char* err, int err_len) {
  noalloc_message(err, err_len, "");
  if (C == nullptr || F_mo == nullptr || gamma == nullptr || energy_out == nullptr) {
    noalloc_message(err, err_len, "null input or energy array");
    return HCC_E_ARG;
  }
  if (guess_kind != HCC_GUESS_REFERENCE && guess_kind != HCC_GUESS_MP2) {
    noalloc_message(err, err_len, "guess must be HCC_GUESS_REFERENCE or HCC_GUESS_MP2");
    return HCC_E_ARG;
  }
  if (maxiter < 1 || diis_steps < 2 || diis_steps > kMaxDiis) {
    noalloc_message(err, err_len, "maxiter must be positive and diis_steps within 2..32");
    return HCC_E_ARG;
  }
  NoAllocDevice* d = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    d = &g_noalloc[device];
    if (d->ptr == nullptr) {
      if (hipSetDevice(device) != hipSuccess) {
        noalloc_message(err, err_len, "hipSetDevice failed");
        return HCC_E_HIP;
      }
      void* p = nullptr;
      const hipError_t r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (r != hipSuccess || p == nullptr) {
        noalloc_message(err, err_len, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(r));
        return HCC_E_HIP;
      }
      hcc_ctx* c = new hcc_ctx();
      c->device = device;
      c->stream = nullptr;
      c->graphs = 0;
      c->timing = false;
      c->arena = static_cast<char*>(p);
      c->arena_cap = kNoAllocBytes;
      d->ptr = static_cast<char*>(p);
      d->ctx = c;
    }
  }
  hcc_ctx* ctx = d->ctx;
  ctx->diis_steps = diis_steps;
  int rc = HCC_OK;
  try {
    HCHECK(hipSetDevice(device));
    load(ctx, 1, &n, &n_occ, C, F_mo, gamma, &e_ref, false);
    precompute(ctx);
    run_solve(ctx, guess_kind, maxiter, eps, false);
    State S;
    HCHECK(hipMemcpy(&S, ctx->batch->dst, sizeof(State), hipMemcpyDeviceToHost));
    for (int k = 0; k < 5; ++k) energy_out[k] = S.energy[k];
    if (iterations) *iterations = S.iter;
    if (status) *status = S.status;
    if (t1 || t2) copy_amplitudes_out(ctx, t1, t2);
  } catch (const HipError& he) {
    noalloc_message(err, err_len, std::string(he.what) + ": " + hipGetErrorString(he.e));
    rc = HCC_E_HIP;
  } catch (const std::string& se) {
    noalloc_message(err, err_len, se);
    rc = HCC_E_ARG;
  } catch (const std::bad_alloc&) {
    noalloc_message(err, err_len, "host allocation failed");
    rc = HCC_E_MEMORY;
  }
  delete ctx->batch;   // its pieces of the scratch region are not freed
  ctx->batch = nullptr;
  return rc;
}

int hcc_selftest_gemm(int m, int n, int k, const double* A, const double* B, double* C, double alpha, double beta,
                      int precision) {
  if (m < 1 || n < 1 || k < 1 || !A || !B || !C) return HCC_E_ARG;
  if (m >= kMaxIndex || n >= kMaxIndex || k >= kMaxIndex) return HCC_E_ARG;
  if (precision != HCC_FP64) return HCC_E_ARG;
  double *dA = nullptr, *dB = nullptr, *dC = nullptr;
  GemmProb* dp = nullptr;
  int* dtp = nullptr;
  int rc = HCC_OK;
  try {
    HCHECK(hipMalloc(&dA, sizeof(double) * m * k));
    HCHECK(hipMalloc(&dB, sizeof(double) * k * n));
    HCHECK(hipMalloc(&dC, sizeof(double) * m * n));
    HCHECK(hipMalloc(&dp, sizeof(GemmProb)));
    HCHECK(hipMemcpy(dA, A, sizeof(double) * m * k, hipMemcpyHostToDevice));
    HCHECK(hipMemcpy(dB, B, sizeof(double) * k * n, hipMemcpyHostToDevice));
    HCHECK(hipMemcpy(dC, C, sizeof(double) * m * n, hipMemcpyHostToDevice));
    GemmProb p = prob(rm(dA, k), rm(dB, n), rm(dC, n), m, n, k, alpha, beta);
    p.tiles_n = (n + kBN - 1) / kBN;
    p.tile_begin = 0;
    p.a_kfast = 1;
    p.b_nfast = 1;
    p.ksplit = 1;
    p.kchunk = k;
    const int tiles = ((m + kBM - 1) / kBM) * p.tiles_n;
    HCHECK(hipMemcpy(dp, &p, sizeof(p), hipMemcpyHostToDevice));
    HCHECK(hipMalloc(&dtp, sizeof(int) * tiles));
    HCHECK(hipMemset(dtp, 0, sizeof(int) * tiles));
    HCHECK(launch_gemm_f64(dp, dtp, tiles, nullptr, nullptr, 64, 3));
    HCHECK(hipDeviceSynchronize());
    HCHECK(hipMemcpy(C, dC, sizeof(double) * m * n, hipMemcpyDeviceToHost));
  } catch (const HipError&) {
    rc = HCC_E_HIP;
  }
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dC);
  (void)hipFree(dp);
  (void)hipFree(dtp);
  return rc;
}

int hcc_selftest_gemm_correction(int m, int n, int k, int a_kfast, int b_nfast, const double* A0, const float* dA,
                                 const double* B0, const float* dB, float* Cd, double alpha, double beta, int tile,
                                 int reps, double* ms) {
  if (m < 1 || n < 1 || k < 1 || !A0 || !B0 || !Cd || (!dA && !dB) || reps < 0) return HCC_E_ARG;
  if (tile != 32 && tile != 64 && tile != 128) return HCC_E_ARG;
  const int kind = tile == 128 ? 256 : tile;
  if ((long)m * k >= (1L << 27) || (long)k * n >= (1L << 27) || (long)m * n >= (1L << 27)) return HCC_E_ARG;
  double *a0 = nullptr, *b0 = nullptr;
  __bf16 *ha = nullptr, *hb = nullptr;
  float *da = nullptr, *db = nullptr, *cd = nullptr;
  GemmProb* dp = nullptr;
  int* dtp = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = HCC_OK;
  try {
    const long na = (long)m * k, nb = (long)k * n, nc = (long)m * n;
    HCHECK(hipMalloc(&a0, sizeof(double) * na));
    HCHECK(hipMalloc(&b0, sizeof(double) * nb));
    HCHECK(hipMalloc(&ha, sizeof(__bf16) * na));
    HCHECK(hipMalloc(&hb, sizeof(__bf16) * nb));
    HCHECK(hipMalloc(&cd, sizeof(float) * nc));
    HCHECK(hipMemcpy(a0, A0, sizeof(double) * na, hipMemcpyHostToDevice));
    HCHECK(hipMemcpy(b0, B0, sizeof(double) * nb, hipMemcpyHostToDevice));
    HCHECK(hipMemcpy(cd, Cd, sizeof(float) * nc, hipMemcpyHostToDevice));
    if (dA) {
      HCHECK(hipMalloc(&da, sizeof(float) * na));
      HCHECK(hipMemcpy(da, dA, sizeof(float) * na, hipMemcpyHostToDevice));
    }
    if (dB) {
      HCHECK(hipMalloc(&db, sizeof(float) * nb));
      HCHECK(hipMemcpy(db, dB, sizeof(float) * nb, hipMemcpyHostToDevice));
    }
    to_bf16(a0, ha, na, nullptr);
    to_bf16(b0, hb, nb, nullptr);
    // A: m x k row-major (K contiguous) or its transpose stored k x m; B: k x n
    // row-major (N contiguous) or its transpose stored n x k
    GemmProb p = prob(a_kfast ? rm(a0, k) : trn(a0, m), b_nfast ? rm(b0, n) : trn(b0, k), rm(nullptr, n), m, n, k,
                      alpha, beta);
    p.dA = da;
    p.dB = db;
    p.Cd = cd;
    p.hA = ha;
    p.hB = hb;
    p.tiles_n = (n + tile - 1) / tile;
    p.tile_begin = 0;
    p.a_kfast = a_kfast ? 1 : 0;
    p.b_nfast = b_nfast ? 1 : 0;
    p.ksplit = 1;
    p.kchunk = k;
    const int tiles = ((m + tile - 1) / tile) * p.tiles_n;
    const int flags = (a_kfast ? 1 : 0) | (b_nfast ? 2 : 0) | (dA ? 4 : 0) | (dB ? 8 : 0) | (k <= 64 && tile != 128 ? 16 : 0);
    HCHECK(hipMalloc(&dp, sizeof(GemmProb)));
    HCHECK(hipMemcpy(dp, &p, sizeof(p), hipMemcpyHostToDevice));
    HCHECK(hipMalloc(&dtp, sizeof(int) * tiles));
    HCHECK(hipMemset(dtp, 0, sizeof(int) * tiles));
    HCHECK(launch_gemm_dual(dp, dtp, tiles, nullptr, nullptr, flags, 1, kind));
    HCHECK(hipDeviceSynchronize());
    HCHECK(hipMemcpy(Cd, cd, sizeof(float) * nc, hipMemcpyDeviceToHost));
    if (ms) *ms = 0.0;
    if (reps > 0) {
      HCHECK(hipEventCreate(&e0));
      HCHECK(hipEventCreate(&e1));
      HCHECK(hipEventRecord(e0, nullptr));
      for (int r = 0; r < reps; ++r) HCHECK(launch_gemm_dual(dp, dtp, tiles, nullptr, nullptr, flags, 1, kind));
      HCHECK(hipEventRecord(e1, nullptr));
      HCHECK(hipEventSynchronize(e1));
      float t = 0.f;
      HCHECK(hipEventElapsedTime(&t, e0, e1));
      if (ms) *ms = t / reps;
    }
  } catch (const HipError&) {
    rc = HCC_E_HIP;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (void* q : {(void*)a0, (void*)b0, (void*)ha, (void*)hb, (void*)da, (void*)db, (void*)cd, (void*)dp, (void*)dtp})
    (void)hipFree(q);
  return rc;
}

}  // extern "C"
