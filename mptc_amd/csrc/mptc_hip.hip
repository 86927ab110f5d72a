// libmptc_hip.so: HIP kernels for gfx950 (batch_kernels.h, dxt_device.h) and
// the C-ABI of include/mptc.h.
//
// One encode step = one frame of each of up to kMaxFrames streams.  The frames
// are described in device memory, so every kernel of the step covers the whole
// batch in one launch, and the step's launch sequence is the same for every
// step of a given batch shape: it is captured once as three HIP graphs
// (prefix: counters, fit, reference tables, inter search, intra set-up and
// dense pass; one batch of fixpoint passes; suffix: commit and unique
// palettes) and replayed.  Between batches of passes the host reads one
// counter to learn whether the intra words reached their fixpoint; one batch
// stays queued while it does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/mptc.h"
#include "net_infer.h"
#include "prune_kernels.h"
#include "framegen.h"

using namespace mptc;

namespace {

constexpr int kMaxS = 63;

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                             \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      return fail(MPTC_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// stb__PrepareOptTable (Include/stb_dxt.h:111-137)
void prepare_opt_table(uint8_t table[256][2], const uint8_t* expand, int size) {
  for (int i = 0; i < 256; ++i) {
    int best = 256;
    for (int mn = 0; mn < size; ++mn)
      for (int mx = 0; mx < size; ++mx) {
        const int mine = expand[mn], maxe = expand[mx];
        const int e = std::abs((2 * maxe + mine) / 3 - i) + std::abs(maxe - mine) * 3 / 100;
        if (e < best) {
          table[i][0] = (uint8_t)mx;
          table[i][1] = (uint8_t)mn;
          best = e;
        }
      }
  }
}

// ToFiveBits / ToSixBits (codec/dxt_image.cpp:72-121)
uint8_t snap_bits(uint8_t x, uint8_t keep, uint8_t half, int shift) {
  uint8_t base = (uint8_t)(x & keep);
  uint8_t high = (uint8_t)(base + half);
  uint8_t low = base == 0 ? base : (uint8_t)(base - half);
  base = (uint8_t)(base | (base >> shift));
  high = (uint8_t)(high | (high >> shift));
  low = (uint8_t)(low | (low >> shift));
  auto ad = [](uint8_t a, uint8_t b) { return (uint8_t)(a > b ? a - b : b - a); };
  const uint8_t db = ad(x, base), dh = ad(x, high), dl = ad(x, low);
  if (db <= dh) return db < dl ? base : low;
  return high;
}

constexpr int kMaxDevices = 64;  // HIP device ordinals this library serves

int init_tables(int device) {
  static std::mutex mu;
  static bool done[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices) return fail(MPTC_EINVAL, "device ordinal above 63");
  std::lock_guard<std::mutex> lk(mu);
  if (done[device]) return MPTC_OK;
  uint8_t e5[32], e6[64], om5[256][2], om6[256][2], snap[512];
  for (int i = 0; i < 32; ++i) e5[i] = (uint8_t)((i << 3) | (i >> 2));
  for (int i = 0; i < 64; ++i) e6[i] = (uint8_t)((i << 2) | (i >> 4));
  prepare_opt_table(om5, e5, 32);
  prepare_opt_table(om6, e6, 64);
  for (int x = 0; x < 256; ++x) {
    snap[x] = snap_bits((uint8_t)x, 0xF8, 4, 5);
    snap[256 + x] = snap_bits((uint8_t)x, 0xFC, 2, 6);
  }
  HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_omatch5), om5, sizeof(om5)));
  HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_omatch6), om6, sizeof(om6)));
  HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_snap), snap, sizeof(snap)));
  done[device] = true;
  return MPTC_OK;
}

enum KernelId {
  KF_FIT, KF_INTER, KF_PREP, KF_DENSE, KF_INTRA_COMBINE, KF_APPLY, KF_REEVAL,
  KF_TILE, KF_COMMIT, KF_UNIQUE, KF_NET_INPUT, KF_NET_FWD, KF_PROPOSE, KF_TOP_INTRA, KF_LPASS, KF_NET_L12,
  KF_NET_L3, KF_PRUNED_INTER, KF_PRUNED_INTRA, KF_FOLD_INTRA, KF_N
};
const char* kKernelNames[KF_N] = {"fit",          "inter",
                                  "intra_prep",   "intra_dense", "intra_combine",
                                  "intra_apply",  "intra_reeval", "intra_tile", "commit",
                                  "unique",       "net_input",   "net_forward", "propose_inter",
                                  "top_intra",    "lintra_pass",  "net_l12",     "net_l3",
                                  "pruned_inter", "pruned_intra", "fold_intra"};
constexpr int kMaxChunks = 8;  // window-row chunks per block in the search kernels

constexpr int kReevalGrid = 1024;  // workgroups of the re-evaluation pass (grid-stride; 1024 measured faster than 4096)
constexpr int kListGrid = 256;     // workgroups of the list-apply pass (grid-stride)

struct Graphs {
  hipGraphExec_t prefix = nullptr, passes = nullptr, suffix = nullptr;
  void destroy() {
    if (prefix) (void)hipGraphExecDestroy(prefix);
    if (passes) (void)hipGraphExecDestroy(passes);
    if (suffix) (void)hipGraphExecDestroy(suffix);
    prefix = passes = suffix = nullptr;
  }
};
// The step's graphs for a batch with inter frames (v[1]) and for a batch of
// intra frames only (v[0]), which skips the inter search.
struct GraphSet {
  Graphs v[2];
  void destroy() {
    v[0].destroy();
    v[1].destroy();
  }
};

// One proposal network: fp32 master parameters (layout of mlp_kernels.h),
// their bf16 copy, Adam moments and gradient.
struct NetState {
  float* w = nullptr;
  uint16_t* wt = nullptr;  // bf16 training copies (kTrSize, pads zero)
  float *m = nullptr, *v = nullptr, *g = nullptr;
  int64_t t = 0;           // Adam steps taken
  bool ready = false;      // parameters set
  float* prior = nullptr;  // kNetLogits fixed scores
  bool use_prior = false;
  float* base = nullptr;   // kNetOut fixed scores added to the logits (seed outputs 0)
  bool use_base = false;
  float* sim = nullptr;    // kSimBins scores by index difference, added in the search's ranking
  bool use_sim = false;
  uint16_t* pk = nullptr;  // packed inference image (net_infer.h)
};

// Network workspace over `rows` blocks: the inputs of a forward pass (X) and,
// for training, the batch's feature-major activations and gradients
// (kActRows x ld bf16, ld = rows rounded up to 32), the k_train_rows
// workgroups' bias column sums and the weight-gradient partials of up to
// kMaxSplits batch splits.
struct NetWork {
  int rows = 0;
  bool train = false;
  uint16_t *X = nullptr, *act = nullptr;
  float *colpart = nullptr, *split = nullptr;
  void release() {
    void* b[] = {X, act, colpart, split};
    for (void* p : b)
      if (p) (void)hipFree(p);
    *this = NetWork();
  }
};

}  // namespace

struct mptc_ctx {
  mptc_config cfg;
  int alloc_S = 0;    // search area the E buffer is sized for
  int max_total = 0;  // blocks per batch
  int max_frames = 1;
  size_t nbs = 0;     // E row stride (blocks)
  hipStream_t stream = nullptr;
  // Every call queues its work on one stream (the caller's or `stream`) and
  // records done_ev there; a call on another stream first waits for it, so
  // work on the context's buffers is ordered whatever streams callers use.
  hipEvent_t done_ev = nullptr;
  hipStream_t done_stream = nullptr;  // the stream done_ev was last recorded on (null: never)
  mptc_tuning tune{256, 16, 2, 0, 4, 1, 0};  // the fastest 1080p shape measured on MI355X
  // device buffers (one entry per block of the batch unless noted)
  FrameDev* d_fr = nullptr;  // kMaxFrames descriptors
  FrameDev* h_fr = nullptr;  // pinned staging ring: 4 x kMaxFrames
  int fr_slot = 0;
  uint8_t *d_sid = nullptr, *d_tsid = nullptr, *d_need = nullptr;
  uint64_t *d_own = nullptr, *d_inter = nullptr, *d_intra = nullptr;
  uint32_t *d_rw = nullptr, *d_cw = nullptr, *d_cwn = nullptr, *d_dirty = nullptr, *d_tiles = nullptr;
  float4 *d_rtab = nullptr, *d_ctab = nullptr;
  int *d_orig = nullptr, *d_chg = nullptr;
  uint2* d_E = nullptr;  // S^2 groups of four int16 error deltas per block
  uint4* d_summ = nullptr;  // kMaxChunks selection summaries per block
  DevStats* d_stats = nullptr;
  DevStats* h_stats = nullptr;  // pinned
  uint32_t* h_conv = nullptr;   // pinned, device-mapped: changed-word count of the last finished pass
  uint32_t* d_conv = nullptr;   // its device address
  std::vector<hipEvent_t> batch_ev;
  // staging of the single-frame host-memory entry points
  uint8_t* d_rgb = nullptr;
  uint64_t *d_ref = nullptr, *d_phys = nullptr;
  uint8_t *d_motion = nullptr, *d_eps = nullptr, *d_kind = nullptr;
  uint32_t *d_unique = nullptr, *d_nunique = nullptr;
  // batch shape the maps and graphs were built for
  std::vector<int> shape;
  mptc_tuning shape_tune{0, 0, 0, 0, 0, 0, 0};
  int n_frames = 0, total = 0, n_wtiles = 0, n_tiles = 0, n_itiles = 0;
  GraphSet graphs;
  bool any_inter = true;  // the batch holds an inter frame
  int last_passes = 0, last_total = 0, last_intra = 0;
  // learned proposals
  int mode = MPTC_SEARCH_EXHAUSTIVE, top_k = 8;
  NetState net[2];
  NetWork nw;  // training and standalone forward passes
  NetWork ew;  // the encode step's forward passes (max_total rows)
  int n_cu = 256;             // compute units of the device
  uint4* d_h2p = nullptr;     // layer-2 activations of the step (net_infer.h tiles)
  int16_t* d_pinter = nullptr;  // inter proposals, kMaxTopK per block
  int16_t* d_prop = nullptr;    // intra proposals, kMaxTopK per block
  // pruned exhaustive search (prune_kernels.h), allocated on first use
  BoundTab *d_rbt = nullptr, *d_cbt = nullptr;  // bound tables of reference / current words
  uint32_t* d_pm = nullptr;                      // passing sets of the intra windows (kPassWords per block)
  bool record = false;          // exhaustive steps record acceptance masks
  bool last_recorded = false;   // the last step did
  uint16_t* d_amask[2] = {nullptr, nullptr};  // inter, intra: 16 rows of 16 offset bits per block
  uint32_t* d_seen = nullptr;
  int* d_errs = nullptr;
  uint8_t* d_tflag = nullptr;  // learned intra passes: tiles to redo, by pass parity (2 x max_total)
  unsigned long long* d_simcnt = nullptr;  // per network: tries, then accepts, by index difference
  // profiling
  bool prof = false;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> spare;
  int64_t prof_n[KF_N] = {};
  double prof_ms[KF_N] = {};
};

namespace {

template <typename F>
int launch(mptc_ctx* c, hipStream_t s, int kid, F&& f) {
  hipEvent_t a = nullptr, b = nullptr;
  if (c->prof) {
    if (!c->spare.empty()) {
      a = c->spare.back().first;
      b = c->spare.back().second;
      c->spare.pop_back();
    } else {
      HIP_TRY(hipEventCreate(&a));
      HIP_TRY(hipEventCreate(&b));
    }
    HIP_TRY(hipEventRecord(a, s));
  }
  f();
  HIP_TRY(hipGetLastError());
  if (c->prof) {
    HIP_TRY(hipEventRecord(b, s));
    c->pending.push_back({kid, {a, b}});
  }
  return MPTC_OK;
}

int collect_profile(mptc_ctx* c) {
  static const bool trace = std::getenv("MPTC_PROFILE_TRACE") != nullptr;
  for (auto& p : c->pending) {
    HIP_TRY(hipEventSynchronize(p.second.second));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, p.second.first, p.second.second));
    if (trace) std::fprintf(stderr, "mptc-trace %s %.4f\n", kKernelNames[p.first], ms);
    c->prof_n[p.first] += 1;
    c->prof_ms[p.first] += ms;
    c->spare.push_back(p.second);
  }
  c->pending.clear();
  return MPTC_OK;
}

// Orders work about to be queued on `s` after the context's previous call,
// which may have used another stream.
int join_stream(mptc_ctx* c, hipStream_t s) {
  if (c->done_stream && c->done_stream != s) HIP_TRY(hipStreamWaitEvent(s, c->done_ev, 0));
  return MPTC_OK;
}

// Marks the end of a call's work on `s` (see join_stream).
int leave_stream(mptc_ctx* c, hipStream_t s) {
  HIP_TRY(hipEventRecord(c->done_ev, s));
  c->done_stream = s;
  return MPTC_OK;
}

int check_params(int S, int thr) {
  if (S < 1 || S > kMaxS) return fail(MPTC_EINVAL, "search_area must be in 1..63");
  if (thr < -32768 || thr > 30000) return fail(MPTC_EINVAL, "err_threshold must be in -32768..30000");
  return MPTC_OK;
}

int check_tuning(const mptc_tuning& t) {
  if (t.wg_size != 64 && t.wg_size != 128 && t.wg_size != 256)
    return fail(MPTC_EINVAL, "wg_size must be 64, 128 or 256");
  if (t.tile_w != 64 && t.tile_w != 32 && t.tile_w != 16 && t.tile_w != 8)
    return fail(MPTC_EINVAL, "tile_w must be 64, 32, 16 or 8");
  if (t.pass_batch < 1 || t.pass_batch > 16) return fail(MPTC_EINVAL, "pass_batch must be in 1..16");
  if (t.win_chunks < 1 || t.win_chunks > kMaxChunks) return fail(MPTC_EINVAL, "win_chunks must be in 1..8");
  if (t.prune != 0 && t.prune != 1) return fail(MPTC_EINVAL, "prune must be 0 or 1");
  if (t.itile_w != 0 && t.itile_w != 8 && t.itile_w != 16 && t.itile_w != 32 && t.itile_w != 64)
    return fail(MPTC_EINVAL, "itile_w must be 0, 8, 16, 32 or 64");
  return MPTC_OK;
}

// Window-row chunks actually used: the largest count up to the requested one
// that divides the 2S window rows and keeps every chunk a whole number of
// four-position E groups.
int chunks_for(int S, int want) {
  for (int k = std::min(want, kMaxChunks); k > 1; --k)
    if ((2 * S) % k == 0 && ((2 * S / k) * 2 * S) % 4 == 0) return k;
  return 1;
}

template <typename T>
int dalloc(T** p, size_t n) {
  if (hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (n ? n : 1)) != hipSuccess)
    return fail(MPTC_ENOMEM, "hipMalloc failed");
  return MPTC_OK;
}

inline unsigned grid_for(int n, int wg) { return (unsigned)std::max(1, (n + wg - 1) / wg); }

// ---------------------------------------------------------------------------
// Proposal networks: buffers, GEMM launches, forward and backward passes.
// ---------------------------------------------------------------------------
int net_alloc(NetState& n) {
  if (n.w) return MPTC_OK;
  int rc;
  if ((rc = dalloc(&n.w, kNetParams)) || (rc = dalloc(&n.wt, kTrSize)) || (rc = dalloc(&n.m, kNetParams)) ||
      (rc = dalloc(&n.v, kNetParams)) || (rc = dalloc(&n.g, kNetParams)) || (rc = dalloc(&n.prior, kNetLogits)) ||
      (rc = dalloc(&n.pk, kPkSize)) || (rc = dalloc(&n.base, kNetOut)) || (rc = dalloc(&n.sim, kSimBins)))
    return rc;
  if (hipMemset(n.base, 0, sizeof(float) * kNetOut) != hipSuccess ||
      hipMemset(n.wt, 0, sizeof(uint16_t) * kTrSize) != hipSuccess)  // the copies' pads stay zero
    return fail(MPTC_EHIP, "hipMemset failed");
  return MPTC_OK;
}

void net_free(NetState& n) {
  void* b[] = {n.w, n.wt, n.m, n.v, n.g, n.prior, n.pk, n.base, n.sim};
  for (void* p : b)
    if (p) (void)hipFree(p);
  n = NetState();
}

constexpr int kMaxSplits = 16;  // batch splits of the weight gradients
constexpr int kTrainRT = 1;     // k_train_rows: 16 x kTrainRT batch rows per workgroup

int work_ensure(NetWork& w, int rows, bool train) {
  if (w.rows >= rows && (w.train || !train)) return MPTC_OK;
  w.release();
  const size_t r = (size_t)rows, ld = (r + 31) / 32 * 32;
  int rc;
  if ((rc = dalloc(&w.X, r * kNetIn))) return rc;
  if (train && ((rc = dalloc(&w.act, ld * kActRows)) || (rc = dalloc(&w.colpart, ld / 16 * kBiasCols)) ||
                (rc = dalloc(&w.split, (size_t)kMaxSplits * kWgPart))))
    return rc;
  w.rows = rows;
  w.train = train;
  return MPTC_OK;
}

// The gradients of one training step over `rows` batch rows (pair idx[r], or
// r): n.g = d(mean loss)/d(parameters), the mean loss added to *loss (device),
// and, when adam.lr > 0, the Adam step on n.  Three launches: the rows'
// forward, loss and backward products (k_train_rows), the weight gradients
// over about 256-row batch splits (k_train_wgrad), and their fixed-order sums
// with the bias sums and the Adam step (k_grad_reduce).
int net_grad_q(mptc_ctx* c, hipStream_t s, NetState& n, NetWork& w, int rows, const uint8_t* px,
               const int32_t* idx, const uint32_t* target, const uint8_t* eps, float seed_w, float* loss,
               const AdamStep& adam) {
  const int ld = (rows + 31) / 32 * 32, nwg = ld / (16 * kTrainRT);
  const int splits = std::max(1, std::min(kMaxSplits, ld / 256));
  const int kspan = ((ld + splits - 1) / splits + 31) / 32 * 32;
  const int nwb = (int)grid_for(kNetParams, 256);
  return launch(c, s, KF_NET_FWD, [&] {
    hipLaunchKernelGGL(k_train_rows<kTrainRT>, dim3(nwg), dim3(256), 0, s, rows, ld, px, idx, target, eps, seed_w, n.wt, n.w,
                       n.use_base ? n.base : nullptr, w.act, w.colpart, loss);
    hipLaunchKernelGGL(k_train_wgrad, dim3(kWgTiles, splits), dim3(256), 0, s, w.act, ld, kspan, w.split);
    hipLaunchKernelGGL(k_grad_reduce, dim3(nwb + (kBiasCols + 3) / 4), dim3(256), 0, s, w.split, splits,
                       w.colpart, nwg, nwb, n.g, n.w, n.m, n.v, n.wt, adam);
  });
}

int net_input_q(mptc_ctx* c, hipStream_t s, const uint8_t* px, const int32_t* idx, int rows, uint16_t* X) {
  return launch(c, s, KF_NET_INPUT, [&] {
    hipLaunchKernelGGL(k_net_input, dim3(grid_for(rows * kNetIn, 256)), dim3(256), 0, s, px, idx, rows, X);
  });
}

// One frame of a batch, host view.
struct FrameIn {
  const uint8_t* rgb;
  int w, h, is_intra;
  const uint64_t* ref;
  uint64_t* phys;
  uint8_t* motion;
  uint32_t* unique;
  uint32_t* n_unique;
  uint8_t* eps;
  uint8_t* kind;
};

// Width of the intra fixpoint tiles: itile_w, or 8 where itile_w is 0 (the
// plain passes then use no tiles; the learned passes always do).
int itile_eff(const mptc_tuning& t) { return t.itile_w ? t.itile_w : 8; }
int itile_log(const mptc_ctx* c) { return __builtin_ctz((unsigned)itile_eff(c->tune)); }

// The device descriptor of one frame of a batch whose earlier frames hold
// `base` blocks, `tb` unique tiles, `wt` wave tiles of width tw and `lt`
// learned-pass tiles of width ltw; advances the four counts past this frame.
FrameDev frame_dev(const FrameIn& x, int tw, int ltw, int& base, int& tb, int& wt, int& lt) {
  const int th = 64 / tw, lth = 64 / ltw;
  FrameDev d{};
  d.rgb = x.rgb; d.ref = x.ref; d.phys_out = x.phys; d.motion = x.motion; d.unique = x.unique;
  d.n_unique = x.n_unique; d.eps = x.eps; d.kind = x.kind;
  d.width = x.w; d.bw = x.w / 4; d.bh = x.h / 4; d.nb = d.bw * d.bh;
  d.base = base; d.is_intra = x.is_intra ? 1 : 0;
  d.tile_base = tb; d.n_tiles = (d.nb + kUniqueTile - 1) / kUniqueTile;
  d.tiles_x = (d.bw + tw - 1) / tw;
  d.wtile_base = wt;
  d.itile_base = lt;
  d.itiles_x = (d.bw + ltw - 1) / ltw;
  lt += d.itiles_x * ((d.bh + lth - 1) / lth);
  base += d.nb;
  tb += d.n_tiles;
  wt += d.tiles_x * ((d.bh + th - 1) / th);
  return d;
}

// Lays out a batch (block, wave-tile and unique-tile offsets) and, when its
// shape changed, rebuilds the block/tile frame maps and drops the graphs.
int prepare_batch(mptc_ctx* c, hipStream_t s, int n, const FrameIn* in) {
  if (n < 1 || n > c->max_frames) return fail(MPTC_EINVAL, "batch size out of range");
  const int tw = c->tune.tile_w;
  std::vector<int> shape;
  FrameDev* h = c->h_fr + (size_t)c->fr_slot * kMaxFrames;
  int base = 0, tb = 0, wt = 0, lt = 0;
  for (int f = 0; f < n; ++f) {
    const FrameIn& x = in[f];
    if (x.w <= 0 || x.h <= 0 || (x.w & 3) || (x.h & 3))
      return fail(MPTC_EINVAL, "width and height must be positive multiples of 4");
    if (!x.rgb || !x.phys || !x.motion || !x.unique || !x.n_unique) return fail(MPTC_EINVAL, "null argument");
    if (!x.is_intra && !x.ref) return fail(MPTC_EINVAL, "inter frame needs the previous frame's blocks");
    const FrameDev d = frame_dev(x, tw, itile_eff(c->tune), base, tb, wt, lt);
    if (base > c->max_total) return fail(MPTC_EINVAL, "batch holds more blocks than the context allows");
    h[f] = d;
    shape.push_back(x.w);
    shape.push_back(x.h);
  }
  c->fr_slot = (c->fr_slot + 1) & 3;
  HIP_TRY(hipMemcpyAsync(c->d_fr, h, sizeof(FrameDev) * n, hipMemcpyHostToDevice, s));
  const bool same = shape == c->shape && std::memcmp(&c->tune, &c->shape_tune, sizeof(mptc_tuning)) == 0;
  if (!same) {
    // frame index of every wave tile
    std::vector<uint8_t> tsid(std::max(wt, 1));
    for (int f = 0; f < n; ++f) {
      const int cnt = h[f].tiles_x * ((h[f].bh + (64 / tw) - 1) / (64 / tw));
      std::fill(tsid.begin() + h[f].wtile_base, tsid.begin() + h[f].wtile_base + cnt, (uint8_t)f);
    }
    if (c->d_tsid) (void)hipFree(c->d_tsid);
    c->d_tsid = nullptr;
    int rc = dalloc(&c->d_tsid, tsid.size());
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(c->d_tsid, tsid.data(), tsid.size(), hipMemcpyHostToDevice, s));
    int maxnb = 1;
    for (int f = 0; f < n; ++f) maxnb = std::max(maxnb, h[f].nb);
    hipLaunchKernelGGL(k_fill_sid, dim3(grid_for(maxnb, 256), n), dim3(256), 0, s, c->d_fr, n, c->d_sid);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));  // the tsid staging vector is released on return
    c->graphs.destroy();
    c->shape = shape;
    c->shape_tune = c->tune;
  }
  c->n_frames = n;
  c->any_inter = false;
  for (int f = 0; f < n; ++f) c->any_inter = c->any_inter || !in[f].is_intra;
  c->total = base;
  c->n_wtiles = wt;
  c->n_tiles = tb;
  c->n_itiles = lt;
  return MPTC_OK;
}

// The pruned exhaustive search applies to small windows and thresholds, where
// the bound rules out most candidates (thresholds 0..kPruneMaxThr).
constexpr int kPruneMaxThr = 4;

bool use_prune(const mptc_ctx* c) {
  return c->tune.prune && c->mode == MPTC_SEARCH_EXHAUSTIVE && !c->record && c->cfg.search_area <= kPruneMaxS &&
         c->cfg.err_threshold >= 0 && c->cfg.err_threshold <= kPruneMaxThr;
}

int prune_ensure(mptc_ctx* c) {
  if (c->d_rbt) return MPTC_OK;
  const size_t nb = (size_t)c->max_total;
  int rc;
  if ((rc = dalloc(&c->d_rbt, nb)) || (rc = dalloc(&c->d_cbt, nb)) || (rc = dalloc(&c->d_pm, nb * kPassWords)))
    return rc;
  return MPTC_OK;
}

// Fused inference over w.X -> H2 tiles (d_h2p) -> either Y (row-major) or
// the top-kk admitted offsets of each block into prop.
int infer_q(mptc_ctx* c, hipStream_t s, const NetState& n, const uint16_t* X, int rows, uint4* h2p, float* Y,
            int net, int kk, const uint8_t* need, int16_t* prop) {
  const int n_tiles = (rows + 31) / 32;
  const unsigned grid = (unsigned)std::max(1, std::min(c->n_cu, (n_tiles + kInferWaves - 1) / kInferWaves));
  const int S = c->cfg.search_area;
  int rc = launch(c, s, KF_NET_L12, [&] {
    const unsigned g12 = (unsigned)std::max(1, std::min(c->n_cu, (n_tiles + kL12Waves - 1) / kL12Waves));
    hipLaunchKernelGGL(k_net_l12, dim3(g12), dim3(64 * kL12Waves), 0, s, X, rows, n.pk, h2p);
  });
  if (rc) return rc;
  return launch(c, s, KF_NET_L3, [&] {
    auto kern = Y ? k_net_l3<4, 9> : kk <= 4 ? k_net_l3<4, 8> : kk <= 8 ? k_net_l3<8, 8> : k_net_l3<16, 8>;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * kInferWaves), 0, s, h2p, rows, n.pk, Y, c->d_fr, c->d_sid, net,
                       S, kk, need, prop, (!Y && n.use_sim) ? n.sim : nullptr, net == 0 ? c->d_rw : c->d_cw,
                       c->d_own);
  });
}

// Proposals of network `net` for the batch's blocks (need: the blocks that
// search; null = the blocks of inter frames) into prop.  The network input
// is built once per step (x_done).
int proposals_q(mptc_ctx* c, hipStream_t s, int net, const uint8_t* need, int16_t* prop, bool& x_done) {
  const NetState& n = c->net[net];
  const int total = c->total;
  if (n.use_prior)
    return launch(c, s, KF_TOP_INTRA, [&] {
      hipLaunchKernelGGL(k_top_scores, dim3(grid_for(total, 256)), dim3(256), 0, s, c->d_fr, c->d_sid, total, net,
                         c->cfg.search_area, c->top_k, n.prior, 1L, 0L, need, prop, n.use_sim ? n.sim : nullptr,
                         net == 0 ? c->d_rw : c->d_cw, c->d_own);
    });
  if (!x_done) {
    int rc = launch(c, s, KF_NET_INPUT, [&] {
      hipLaunchKernelGGL(k_frame_net_input, dim3(grid_for(total, 256)), dim3(256), 0, s, c->d_fr, c->d_sid, total,
                         c->ew.X);
    });
    if (rc) return rc;
    x_done = true;
  }
  return infer_q(c, s, n, c->ew.X, total, c->d_h2p, nullptr, net, c->top_k, need, prop);
}

// report: the pass writes its count of changed words for the host (the last
// pass of a batch); the others do not.
int lintra_pass_q(mptc_ctx* c, hipStream_t s, int par, int first, bool report) {
  const int S = c->cfg.search_area, thr = c->cfg.err_threshold, kk = c->top_k;
  return launch(c, s, KF_LPASS, [&] {
    auto kern = kk <= 4 ? k_lintra_pass<4> : kk <= 8 ? k_lintra_pass<8> : k_lintra_pass<16>;
    hipLaunchKernelGGL(kern, dim3(grid_for(c->n_itiles * 64, 256)), dim3(256), 0, s, c->d_fr, c->n_frames,
                       c->n_itiles, itile_log(c), S, thr, kk, c->d_prop, c->d_own, c->d_need, c->d_cw, c->d_seen, c->d_errs,
                       c->d_intra, c->d_stats, par, first, report ? c->d_conv : nullptr, c->d_tflag,
                       (long)c->max_total);
  });
}

// The three launch sequences of a step.  `fit` = false starts from
// caller-provided fitted blocks already in d_own.
int queue_prefix(mptc_ctx* c, hipStream_t s, bool fit) {
  const int total = c->total, wg = c->tune.wg_size, tw = c->tune.tile_w;
  const int S = c->cfg.search_area, thr = c->cfg.err_threshold;
  int rc;
  // counters, fit, reference tables and (learned modes) the network input, in one launch
  const bool want_x = c->mode != MPTC_SEARCH_EXHAUSTIVE && c->ew.X &&
                      (!c->net[MPTC_NET_INTER].use_prior ||
                       (c->mode == MPTC_SEARCH_LEARNED && !c->net[MPTC_NET_INTRA].use_prior));
  bool x_done = want_x;
  bool prep_done = false;  // need and the starting intra words already set
  const bool prune = use_prune(c);
  if ((rc = launch(c, s, KF_FIT, [&] {
         const int n_st = (int)(sizeof(DevStats) / 4);
         hipLaunchKernelGGL(k_block_prep, dim3(grid_for(std::max(total, n_st + c->n_tiles), wg)), dim3(wg), 0, s,
                            c->d_fr, c->d_sid, total, fit ? c->d_own : nullptr, c->d_rw, c->d_rtab,
                            want_x ? c->ew.X : nullptr, reinterpret_cast<uint32_t*>(c->d_stats), n_st, c->d_tiles,
                            c->n_tiles, prune ? c->d_rbt : nullptr);
       })))
    return rc;
  const int K = chunks_for(S, c->tune.win_chunks);
  // the exhaustive inter searches leave chunk summaries, from which
  // k_intra_prep takes the inter decisions
  const uint4* inter_summ = nullptr;
  int orig_ready = 0;
  const unsigned gwk = grid_for(c->n_wtiles * K * 64, wg);
  const unsigned gwk256 = grid_for(c->n_wtiles * K * 64, 256);
  if (!c->any_inter) {
    // a batch of intra frames only: no inter search (k_intra_prep reads no inter decision)
  } else if (prune) {
    // inter: bound (tables from k_block_prep), exact evaluation of the
    // survivors, fold (prune_kernels.h)
    if ((rc = launch(c, s, KF_PRUNED_INTER, [&] {
           hipLaunchKernelGGL(k_pruned_search<false>, dim3(gwk256), dim3(256), 0, s, c->d_fr, c->d_tsid, c->n_wtiles,
                              tw, K, S, thr, c->d_own, c->d_need, c->d_rw, c->d_rtab, c->d_rbt, c->d_orig, nullptr,
                              c->nbs, nullptr, c->d_summ, c->d_stats, nullptr, nullptr, nullptr, 0);
         })))
      return rc;
    inter_summ = c->d_summ;  // k_intra_prep takes the inter decisions
    orig_ready = 1;          // and the own errors the inter search wrote
  } else if (c->mode == MPTC_SEARCH_EXHAUSTIVE) {
    if ((rc = launch(c, s, KF_INTER, [&] {
           hipLaunchKernelGGL(k_inter, dim3(gwk), dim3(wg), 0, s, c->d_fr, c->d_tsid, c->n_wtiles, tw, K, S,
                              c->d_own, c->d_rw, c->d_rtab, c->d_summ, c->nbs, c->d_stats, thr,
                              c->record ? c->d_amask[0] : nullptr);
         })))
      return rc;
    inter_summ = c->d_summ;  // k_intra_prep takes the inter decisions
  } else {
    if ((rc = proposals_q(c, s, MPTC_NET_INTER, nullptr, c->d_pinter, x_done))) return rc;
    // the learned intra search needs only k_intra_prep's flags and starting
    // words, which the inter evaluation sets for every block of the batch
    prep_done = c->mode == MPTC_SEARCH_LEARNED;
    if ((rc = launch(c, s, KF_PROPOSE, [&] {
           hipLaunchKernelGGL(k_eval_inter, dim3(grid_for(total, wg)), dim3(wg), 0, s, c->d_fr, c->d_sid, total, S,
                              thr, c->top_k, c->d_pinter, c->d_own, c->d_rw, c->d_rtab, c->d_inter,
                              prep_done ? c->d_need : nullptr, prep_done ? c->d_cw : nullptr, c->d_stats);
         })))
      return rc;
  }
  if (!prep_done && (rc = launch(c, s, KF_PREP, [&] {
         const bool lrn = c->mode == MPTC_SEARCH_LEARNED;  // its intra passes use no tables, errors or flags
         hipLaunchKernelGGL(k_intra_prep, dim3(grid_for(total, wg)), dim3(wg), 0, s, c->d_fr, c->d_sid, total,
                            c->d_own, c->d_inter, c->d_need, c->d_cw, lrn ? nullptr : c->d_ctab,
                            lrn ? nullptr : c->d_dirty, lrn ? nullptr : c->d_orig, prune ? c->d_cbt : nullptr,
                            prune ? reinterpret_cast<uint4*>(c->d_pm) : nullptr, K, S, thr, c->d_rw, inter_summ,
                            c->nbs, orig_ready);
       })))
    return rc;
  if (c->mode == MPTC_SEARCH_LEARNED) {
    if ((rc = proposals_q(c, s, MPTC_NET_INTRA, c->d_need, c->d_prop, x_done))) return rc;
    return lintra_pass_q(c, s, 0, 1, false);
  }
  if (prune) {  // bound tables and cleared passing sets from k_intra_prep
    // with 1, 2 or 4 chunks every tile's chunks are waves of one workgroup,
    // and the search also takes the intra decisions (k_fold_intra's work)
    const int fold = 4 % K == 0 ? 1 : 0;
    if ((rc = launch(c, s, KF_PRUNED_INTRA, [&] {
           hipLaunchKernelGGL(k_pruned_search<true>, dim3(gwk256), dim3(256), 0, s, c->d_fr, c->d_tsid, c->n_wtiles,
                              tw, K, S, thr, c->d_own, c->d_need, c->d_cw, c->d_ctab, c->d_cbt, c->d_orig, c->d_E,
                              c->nbs, c->d_pm, nullptr, c->d_stats, c->d_intra, c->d_cwn, c->d_chg, fold);
         })))
      return rc;
    if (!fold && (rc = launch(c, s, KF_FOLD_INTRA, [&] {
           hipLaunchKernelGGL(k_fold_intra, dim3(grid_for(total, 256)), dim3(256), 0, s, c->d_fr, c->d_sid,
                              total, S, thr, c->d_own, c->d_need, c->d_cw, c->d_E, c->nbs, c->d_pm, c->d_intra,
                              c->d_cwn, c->d_chg, c->d_stats);
         })))
      return rc;
    return launch(c, s, KF_APPLY, [&] {
      hipLaunchKernelGGL(k_intra_apply, dim3(kListGrid), dim3(256), 0, s, c->d_cwn, c->d_cw, c->d_ctab, c->d_chg,
                         c->d_stats);
    });
  }
  if ((rc = launch(c, s, KF_DENSE, [&] {
         hipLaunchKernelGGL(k_intra_dense, dim3(gwk), dim3(wg), 0, s, c->d_fr, c->d_tsid, c->n_wtiles, tw, K, S,
                            c->d_own, c->d_need, c->d_cw, c->d_ctab, c->d_orig, c->d_E, c->nbs, c->d_summ, c->nbs,
                            c->d_stats);
       })))
    return rc;
  if ((rc = launch(c, s, KF_INTRA_COMBINE, [&] {
         hipLaunchKernelGGL(k_intra_combine, dim3(grid_for(total, wg)), dim3(wg), 0, s, c->d_fr, c->d_sid, total, K,
                            S, thr, c->d_own, c->d_need, c->d_cw, c->d_summ, c->nbs, c->d_intra, c->d_cwn, c->d_chg,
                            c->d_stats);
       })))
    return rc;
  return launch(c, s, KF_APPLY, [&] {
    hipLaunchKernelGGL(k_intra_apply, dim3(kListGrid), dim3(256), 0, s, c->d_cwn, c->d_cw, c->d_ctab, c->d_chg,
                       c->d_stats);
  });
}

// An even number of fixpoint passes, so that every batch starts and ends at
// list parity 0.
int passes_per_batch(const mptc_ctx* c) { return 2 * ((c->tune.pass_batch + 1) / 2); }

int queue_passes(mptc_ctx* c, hipStream_t s) {
  const int S = c->cfg.search_area, thr = c->cfg.err_threshold;
  const int np = passes_per_batch(c);
  int rc;
  if (c->mode == MPTC_SEARCH_LEARNED) {
    // parities 1, 0, ..., 0: the batch's last pass counts into n_changed[0]
    for (int p = 0; p < np; ++p)
      if ((rc = lintra_pass_q(c, s, (p + 1) & 1, 0, p == np - 1))) return rc;
    return MPTC_OK;
  }
  // the pruned search keeps passing sets beside E (prune_kernels.h)
  uint32_t* pm = use_prune(c) ? c->d_pm : nullptr;
  // itile_w 0: re-fold passes (no skip in k_intra_reeval; skip_tile stays 0)
  const int tw_log = itile_log(c);
  for (int p = 0; p < np; ++p) {
    const int par = p & 1;
    if ((rc = launch(c, s, KF_REEVAL, [&] {
           hipLaunchKernelGGL(k_intra_reeval, dim3(kReevalGrid), dim3(256), 0, s, c->d_fr, c->d_sid, S, thr, c->d_own,
                              c->d_need, c->d_cw, c->d_ctab, c->d_chg, c->d_E, c->nbs, c->d_dirty, c->d_orig,
                              c->d_intra, c->d_stats, par, pm, tw_log);
         })))
      return rc;
    if ((rc = launch(c, s, KF_TILE, [&] {
           if (c->tune.itile_w)
             hipLaunchKernelGGL(k_intra_tile_pass, dim3(grid_for(c->n_itiles * 64, 256)), dim3(256), 0, s, c->d_fr,
                                c->n_frames, c->n_itiles, tw_log, S, thr, c->d_own, c->d_need, c->d_cw, c->d_ctab,
                                c->d_E, c->nbs, c->d_dirty, c->d_orig, c->d_intra, c->d_chg, c->d_stats, par ^ 1,
                                p == np - 1 ? c->d_conv : nullptr, pm);
           else
             hipLaunchKernelGGL(k_intra_select, dim3(grid_for(c->total, 256)), dim3(256), 0, s, c->d_fr, c->d_sid,
                                c->total, S, thr, c->d_own, c->d_cw, c->d_ctab, c->d_E, c->nbs, c->d_dirty, c->d_intra,
                                c->d_chg, c->d_stats, par ^ 1, p == np - 1 ? c->d_conv : nullptr, pm);
         })))
      return rc;
  }
  return MPTC_OK;
}

int queue_suffix(mptc_ctx* c, hipStream_t s) {
  const int total = c->total, wg = c->tune.wg_size;
  int rc;
  if (c->record && c->mode == MPTC_SEARCH_EXHAUSTIVE &&
      (rc = launch(c, s, KF_COMMIT, [&] {
         hipLaunchKernelGGL(k_intra_amask, dim3(grid_for(total, 256)), dim3(256), 0, s, c->d_fr, c->d_sid, total,
                            c->cfg.search_area, c->cfg.err_threshold, c->d_need, c->d_E, c->nbs, c->d_amask[1]);
       })))
    return rc;
  if ((rc = launch(c, s, KF_COMMIT, [&] {
         hipLaunchKernelGGL(k_commit, dim3(grid_for(total, wg)), dim3(wg), 0, s, c->d_fr, c->d_sid, total, c->d_own,
                            c->d_inter, c->d_need, c->d_intra, c->d_tiles, c->d_stats);
       })))
    return rc;
  return launch(c, s, KF_UNIQUE, [&] {
    hipLaunchKernelGGL(k_unique, dim3(c->n_tiles), dim3(256), 0, s, c->d_fr, c->n_frames, c->d_own, c->d_tiles,
                       c->d_stats);
  });
}

template <typename Q>
int capture(hipStream_t s, hipGraphExec_t* exec, Q&& q) {
  HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int rc = q();
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(s, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(MPTC_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  const hipError_t e2 = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e2 != hipSuccess) return fail(MPTC_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e2));
  return MPTC_OK;
}

// One encode step over the prepared batch.
int run_step(mptc_ctx* c, hipStream_t s, bool fit) {
  const bool graphs = c->tune.use_graph && !c->prof && fit;
  Graphs& gr = c->graphs.v[c->any_inter ? 1 : 0];
  int rc;
  // The last pass of each batch writes its count of changed words to h_conv
  // (0 when it had nothing to do).  Passes of an earlier step that are still
  // queued can only write 0 (they follow a pass that changed nothing), and
  // they run before this step's passes on the ordered stream.
  *reinterpret_cast<volatile uint32_t*>(c->h_conv) = 0u;
  if (graphs && !gr.prefix) {
    if ((rc = capture(s, &gr.prefix, [&] { return queue_prefix(c, s, true); }))) return rc;
    if ((rc = capture(s, &gr.passes, [&] { return queue_passes(c, s); }))) return rc;
    if ((rc = capture(s, &gr.suffix, [&] { return queue_suffix(c, s); }))) return rc;
  }
  if (graphs) {
    HIP_TRY(hipGraphLaunch(gr.prefix, s));
  } else if ((rc = queue_prefix(c, s, fit))) {
    return rc;
  }
  if (c->batch_ev.empty()) {
    c->batch_ev.resize(2);
    for (auto& e : c->batch_ev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  const int np = passes_per_batch(c);
  int passes = 0;
  auto queue_batch = [&](int slot) -> int {
    if (graphs) {
      HIP_TRY(hipGraphLaunch(gr.passes, s));
    } else {
      const int r = queue_passes(c, s);
      if (r) return r;
    }
    passes += np;
    HIP_TRY(hipEventRecord(c->batch_ev[slot], s));
    return MPTC_OK;
  };
  int slot = 0;
  if ((rc = queue_batch(slot))) return rc;
  for (;;) {
    // the raster-triangular system settles within one pass per block
    if (passes > c->total + 4 * np) return fail(MPTC_ESTATE, "intra fixpoint did not converge");
    if ((rc = queue_batch(slot ^ 1))) return rc;
    // poll (a blocking wait adds its wake-up latency to every batch)
    hipError_t q;
    while ((q = hipEventQuery(c->batch_ev[slot])) == hipErrorNotReady) {
    }
    if (q != hipSuccess) return fail(MPTC_EHIP, std::string("hipEventQuery: ") + hipGetErrorString(q));
    // The last pass of every batch writes its count of changed words to
    // h_conv, so once the batch is done h_conv holds the count of its last
    // pass or of a later batch's last pass.  A pass that changes nothing
    // proves the fixpoint (every later pass is empty too); the batch queued
    // after it then runs empty.
    if (*reinterpret_cast<volatile uint32_t*>(c->h_conv) == 0) break;
    slot ^= 1;
  }
  if (graphs) {
    HIP_TRY(hipGraphLaunch(gr.suffix, s));
  } else if ((rc = queue_suffix(c, s))) {
    return rc;
  }
  c->last_passes = passes;
  c->last_total = c->total;
  c->last_recorded = c->record && c->mode == MPTC_SEARCH_EXHAUSTIVE;
  return MPTC_OK;
}

int encode_batch(mptc_ctx* c, hipStream_t s, int n, const FrameIn* in, bool fit) {
  // the previous step's suffix (commit, unique palettes) may still read the
  // work buffers on another stream
  int rc = join_stream(c, s);
  if (rc) return rc;
  if ((rc = prepare_batch(c, s, n, in))) return rc;
  if (use_prune(c) && (rc = prune_ensure(c))) return rc;
  c->last_intra = n == 1 ? in[0].is_intra : 0;
  if ((rc = run_step(c, s, fit))) return rc;
  return leave_stream(c, s);
}

int one_frame_dims(mptc_ctx* c, int w, int h) {
  if (w <= 0 || h <= 0 || (w & 3) || (h & 3))
    return fail(MPTC_EINVAL, "width and height must be positive multiples of 4");
  if ((long long)(w / 4) * (h / 4) > (long long)(c->cfg.max_width / 4) * (c->cfg.max_height / 4))
    return fail(MPTC_EINVAL, "frame larger than the context's maximum");
  return MPTC_OK;
}

}  // namespace

extern "C" {

const char* mptc_last_error(void) { return g_err.c_str(); }

int32_t mptc_create(const mptc_config* cfg, mptc_ctx** out) {
  if (!cfg || !out) return fail(MPTC_EINVAL, "null argument");
  *out = nullptr;
  if (cfg->max_width <= 0 || cfg->max_height <= 0 || (cfg->max_width & 3) || (cfg->max_height & 3))
    return fail(MPTC_EINVAL, "max_width/max_height must be positive multiples of 4");
  int rc = check_params(cfg->search_area, cfg->err_threshold);
  if (rc) return rc;
  const int frames = cfg->max_batch_frames <= 0 ? 1 : cfg->max_batch_frames;
  if (frames > kMaxFrames) return fail(MPTC_EINVAL, "max_batch_frames above 64");
  const long long one = (long long)(cfg->max_width / 4) * (cfg->max_height / 4);
  const long long total = std::max(one, cfg->max_batch_blocks > 0 ? (long long)cfg->max_batch_blocks : one * frames);
  if (total > (1ll << 30)) return fail(MPTC_EINVAL, "batch too large");
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (cfg->device < 0 || cfg->device >= ndev) return fail(MPTC_EINVAL, "no such HIP device");
  HIP_TRY(hipSetDevice(cfg->device));
  if ((rc = init_tables(cfg->device))) return rc;
  mptc_ctx* c = new mptc_ctx();
  c->cfg = *cfg;
  {
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, cfg->device) == hipSuccess && cu > 0)
      c->n_cu = cu;
  }
  c->alloc_S = cfg->search_area;
  c->max_frames = frames;
  c->max_total = (int)total;
  const int nb = c->max_total;
  c->nbs = (size_t)((nb + 63) / 64) * 64;
  auto cleanup = [&](int code) {
    mptc_destroy(c);
    return code;
  };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(MPTC_EHIP, "hipStreamCreate failed"));
  if (hipEventCreateWithFlags(&c->done_ev, hipEventDisableTiming) != hipSuccess)
    return cleanup(fail(MPTC_EHIP, "hipEventCreate failed"));
  const int max_tiles = nb / kUniqueTile + frames + 1;
  const size_t groups = (size_t)cfg->search_area * cfg->search_area;
  const size_t one_nb = (size_t)one;
  if ((rc = dalloc(&c->d_fr, kMaxFrames)) || (rc = dalloc(&c->d_sid, nb)) || (rc = dalloc(&c->d_need, nb)) ||
      (rc = dalloc(&c->d_own, nb)) || (rc = dalloc(&c->d_inter, nb)) || (rc = dalloc(&c->d_intra, nb)) ||
      (rc = dalloc(&c->d_rw, nb)) || (rc = dalloc(&c->d_cw, nb)) || (rc = dalloc(&c->d_cwn, nb)) ||
      (rc = dalloc(&c->d_dirty, nb)) || (rc = dalloc(&c->d_tiles, max_tiles)) || (rc = dalloc(&c->d_rtab, nb)) ||
      (rc = dalloc(&c->d_ctab, nb)) || (rc = dalloc(&c->d_orig, nb)) || (rc = dalloc(&c->d_chg, nb)) ||
      (rc = dalloc(&c->d_E, groups * c->nbs)) || (rc = dalloc(&c->d_summ, kMaxChunks * c->nbs)) || (rc = dalloc(&c->d_stats, 1)) ||
      (rc = dalloc(&c->d_rgb, one_nb * 48)) || (rc = dalloc(&c->d_ref, one_nb)) || (rc = dalloc(&c->d_phys, one_nb)) ||
      (rc = dalloc(&c->d_motion, one_nb * 2)) || (rc = dalloc(&c->d_eps, one_nb * 6)) ||
      (rc = dalloc(&c->d_kind, one_nb)) || (rc = dalloc(&c->d_unique, one_nb)) || (rc = dalloc(&c->d_nunique, 1)))
    return cleanup(rc);
  if (hipHostMalloc(reinterpret_cast<void**>(&c->h_stats), sizeof(DevStats), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&c->h_conv), 64 * sizeof(uint32_t),
                    hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
      hipHostGetDevicePointer(reinterpret_cast<void**>(&c->d_conv), c->h_conv, 0) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&c->h_fr), 4 * kMaxFrames * sizeof(FrameDev), hipHostMallocDefault) !=
          hipSuccess)
    return cleanup(fail(MPTC_ENOMEM, "hipHostMalloc failed"));
  std::memset(c->h_stats, 0, sizeof(DevStats));
  std::memset(c->h_conv, 0, 64 * sizeof(uint32_t));
  *out = c;
  return MPTC_OK;
}

void mptc_destroy(mptc_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->cfg.device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->done_stream) (void)hipEventSynchronize(c->done_ev);  // the last call may have used a caller's stream
  if (c->done_ev) (void)hipEventDestroy(c->done_ev);
  c->graphs.destroy();
  for (auto& p : c->pending) { (void)hipEventDestroy(p.second.first); (void)hipEventDestroy(p.second.second); }
  for (auto& p : c->spare) { (void)hipEventDestroy(p.first); (void)hipEventDestroy(p.second); }
  for (hipEvent_t e : c->batch_ev) (void)hipEventDestroy(e);
  void* bufs[] = {c->d_fr,   c->d_sid,   c->d_tsid,   c->d_need,  c->d_own,   c->d_inter,  c->d_intra,
                  c->d_rw,   c->d_cw,    c->d_cwn,    c->d_dirty, c->d_tiles, c->d_rtab,   c->d_ctab,
                  c->d_orig, c->d_chg,   c->d_E, c->d_summ,     c->d_stats, c->d_rgb,    c->d_ref,
                  c->d_phys, c->d_motion, c->d_eps,   c->d_kind,  c->d_unique, c->d_nunique};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  net_free(c->net[0]);
  net_free(c->net[1]);
  c->nw.release();
  c->ew.release();
  void* lbufs[] = {c->d_h2p, c->d_pinter, c->d_prop, c->d_seen, c->d_errs, c->d_amask[0], c->d_amask[1],
                   c->d_rbt,  c->d_cbt,  c->d_pm,  c->d_tflag, c->d_simcnt};
  for (void* p : lbufs)
    if (p) (void)hipFree(p);
  if (c->h_stats) (void)hipHostFree(c->h_stats);
  if (c->h_conv) (void)hipHostFree(c->h_conv);
  if (c->h_fr) (void)hipHostFree(c->h_fr);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int32_t mptc_set_params(mptc_ctx* c, int32_t S, int32_t thr) {
  if (!c) return fail(MPTC_EINVAL, "null context");
  int rc = check_params(S, thr);
  if (rc) return rc;
  if (S > c->alloc_S) return fail(MPTC_EINVAL, "search_area above the one the context was created with");
  if (S != c->cfg.search_area || thr != c->cfg.err_threshold) c->graphs.destroy();
  c->cfg.search_area = S;
  c->cfg.err_threshold = thr;
  return MPTC_OK;
}

int32_t mptc_get_tuning(mptc_ctx* c, mptc_tuning* out) {
  if (!c || !out) return fail(MPTC_EINVAL, "null argument");
  *out = c->tune;
  return MPTC_OK;
}

int32_t mptc_set_tuning(mptc_ctx* c, const mptc_tuning* t) {
  if (!c || !t) return fail(MPTC_EINVAL, "null argument");
  int rc = check_tuning(*t);
  if (rc) return rc;
  c->tune = *t;
  return MPTC_OK;
}

int32_t mptc_fit(mptc_ctx* c, const uint8_t* rgb, int32_t w, int32_t h, uint64_t* phys_out) {
  if (!c || !rgb || !phys_out) return fail(MPTC_EINVAL, "null argument");
  int rc = one_frame_dims(c, w, h);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(c->cfg.device));
  hipStream_t s = c->stream;
  const int nb = (w / 4) * (h / 4);
  if ((rc = join_stream(c, s))) return rc;
  HIP_TRY(hipMemcpyAsync(c->d_rgb, rgb, (size_t)w * h * 3, hipMemcpyHostToDevice, s));
  FrameIn in{c->d_rgb, w, h, 1, nullptr, c->d_phys, c->d_motion, c->d_unique, c->d_nunique, nullptr, nullptr};
  if ((rc = prepare_batch(c, s, 1, &in))) return rc;
  if ((rc = launch(c, s, KF_FIT, [&] {
         hipLaunchKernelGGL(k_fit, dim3(grid_for(nb, 256)), dim3(256), 0, s, c->d_fr, c->d_sid, nb, c->d_own);
       })))
    return rc;
  HIP_TRY(hipMemcpyAsync(phys_out, c->d_own, (size_t)nb * 8, hipMemcpyDeviceToHost, s));
  if ((rc = leave_stream(c, s))) return rc;
  HIP_TRY(hipStreamSynchronize(s));
  return MPTC_OK;
}

int32_t mptc_reencode(mptc_ctx* c, const uint8_t* rgb, int32_t w, int32_t h, int32_t is_intra,
                      const uint64_t* ref_phys, uint64_t* phys_inout, uint8_t* motion_out, uint32_t* unique_out,
                      uint32_t* n_unique_out, uint8_t* eps_out, uint8_t* kind_out) {
  if (!c || !rgb || !phys_inout || !motion_out || !unique_out || !n_unique_out)
    return fail(MPTC_EINVAL, "null argument");
  if (!is_intra && !ref_phys) return fail(MPTC_EINVAL, "inter frame needs ref_phys");
  int rc = one_frame_dims(c, w, h);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(c->cfg.device));
  const int nb = (w / 4) * (h / 4);
  hipStream_t s = c->stream;
  if ((rc = join_stream(c, s))) return rc;
  HIP_TRY(hipMemcpyAsync(c->d_rgb, rgb, (size_t)w * h * 3, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(c->d_own, phys_inout, (size_t)nb * 8, hipMemcpyHostToDevice, s));
  if (!is_intra) HIP_TRY(hipMemcpyAsync(c->d_ref, ref_phys, (size_t)nb * 8, hipMemcpyHostToDevice, s));
  FrameIn in{c->d_rgb,    w, h, is_intra, is_intra ? nullptr : c->d_ref, c->d_phys, c->d_motion, c->d_unique,
             c->d_nunique, eps_out ? c->d_eps : nullptr, kind_out ? c->d_kind : nullptr};
  if ((rc = encode_batch(c, s, 1, &in, false))) return rc;
  uint32_t n_unique = 0;
  HIP_TRY(hipMemcpyAsync(&n_unique, c->d_nunique, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(phys_inout, c->d_phys, (size_t)nb * 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(motion_out, c->d_motion, (size_t)nb * 2, hipMemcpyDeviceToHost, s));
  if (eps_out) HIP_TRY(hipMemcpyAsync(eps_out, c->d_eps, (size_t)nb * 6, hipMemcpyDeviceToHost, s));
  if (kind_out) HIP_TRY(hipMemcpyAsync(kind_out, c->d_kind, (size_t)nb, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  HIP_TRY(hipMemcpy(unique_out, c->d_unique, (size_t)n_unique * 4, hipMemcpyDeviceToHost));
  *n_unique_out = n_unique;
  return MPTC_OK;
}

int32_t mptc_encode_batch_device(mptc_ctx* c, int32_t n, const mptc_frame_io* frames, void* hip_stream) {
  if (!c || !frames) return fail(MPTC_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(c->cfg.device));
  if (n < 1 || n > c->max_frames) return fail(MPTC_EINVAL, "batch size out of range");
  FrameIn in[kMaxFrames];
  for (int f = 0; f < n; ++f) {
    const mptc_frame_io& x = frames[f];
    in[f] = FrameIn{x.d_rgb,        x.width,        x.height,       x.is_intra, x.d_ref_phys, x.d_phys_out,
                    x.d_motion_out, x.d_unique_out, x.d_n_unique_out, x.d_eps_out, x.d_kind_out};
  }
  hipStream_t s = hip_stream ? static_cast<hipStream_t>(hip_stream) : c->stream;
  return encode_batch(c, s, n, in, true);
}

int32_t mptc_encode_frame_device(mptc_ctx* c, const uint8_t* d_rgb, int32_t w, int32_t h, int32_t is_intra,
                                 const uint64_t* d_ref_phys, uint64_t* d_phys_out, uint8_t* d_motion_out,
                                 uint32_t* d_unique_out, uint32_t* d_n_unique_out, uint8_t* d_eps_out,
                                 uint8_t* d_kind_out, void* hip_stream) {
  if (!c) return fail(MPTC_EINVAL, "null context");
  int rc = one_frame_dims(c, w, h);
  if (rc) return rc;
  const mptc_frame_io io{d_rgb,          w, h, is_intra, d_ref_phys, d_phys_out, d_motion_out, d_unique_out,
                         d_n_unique_out, d_eps_out, d_kind_out};
  return mptc_encode_batch_device(c, 1, &io, hip_stream);
}

int32_t mptc_autotune(mptc_ctx* c, int32_t n, const mptc_frame_io* frames, int32_t steps, mptc_tuning* best,
                      double* best_ms) {
  if (!c || !frames || steps < 1) return fail(MPTC_EINVAL, "bad argument");
  HIP_TRY(hipSetDevice(c->cfg.device));
  const mptc_tuning keep = c->tune;
  mptc_tuning win = keep;
  double win_ms = 1e30;
  const int wgs[] = {64, 128, 256};
  const int tws[] = {64, 32, 16, 8};
  const int chunks[] = {1, 2, 4, 8};
  // the pruned search is tried where it applies (both give the same output)
  mptc_tuning probe = keep;
  probe.prune = 1;
  c->tune = probe;
  const int n_prune = use_prune(c) ? 2 : 1;
  // times the shape t (warm-up: maps and graphs; then `steps` encodes) and keeps the fastest
  auto trial = [&](const mptc_tuning& t) -> int {
    c->tune = t;
    int rc = mptc_encode_batch_device(c, n, frames, nullptr);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    const auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < steps; ++i)
      if ((rc = mptc_encode_batch_device(c, n, frames, nullptr))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / steps;
    if (ms < win_ms) {
      win_ms = ms;
      win = t;
    }
    return MPTC_OK;
  };
  int rc = MPTC_OK;
  for (int pr = 0; pr < n_prune && !rc; ++pr)
    for (int graph = 0; graph < 2 && !rc; ++graph)
      for (int wg : wgs)
        for (int tw : tws)
          for (int kc : chunks)
            if (!rc)
              rc = trial(mptc_tuning{wg, tw, keep.pass_batch, graph, kc, n_prune == 2 ? pr : keep.prune,
                                     keep.itile_w});
  // then the intra fixpoint tile on the winning shape
  const mptc_tuning shape = win;
  for (int itw : {0, 8, 16, 32, 64})
    if (!rc && itw != shape.itile_w) {
      mptc_tuning t = shape;
      t.itile_w = itw;
      rc = trial(t);
    }
  if (rc) {
    c->tune = keep;
    return rc;
  }
  c->tune = win;
  if (best) *best = win;
  if (best_ms) *best_ms = win_ms;
  return MPTC_OK;
}

int32_t mptc_synchronize(mptc_ctx* c) {
  if (!c) return fail(MPTC_EINVAL, "null context");
  HIP_TRY(hipSetDevice(c->cfg.device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (c->done_stream) HIP_TRY(hipEventSynchronize(c->done_ev));  // the last call's stream
  return MPTC_OK;
}

int32_t mptc_get_stats(mptc_ctx* c, mptc_frame_stats* out) {
  if (!c || !out) return fail(MPTC_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(c->cfg.device));
  HIP_TRY(hipDeviceSynchronize());
  DevStats d;
  HIP_TRY(hipMemcpy(&d, c->d_stats, sizeof(d), hipMemcpyDeviceToHost));
  out->n_blocks = c->last_total;
  out->n_inter = out->n_intra = out->inter_candidates = out->intra_candidates = out->reeval_candidates = 0;
  out->exact_inter_candidates = out->exact_intra_candidates = 0;
  for (int i = 0; i < kStatSlots; ++i) {
    out->n_inter += d.n_inter[i * kSlot32];
    out->n_intra += d.n_intra[i * kSlot32];
    out->inter_candidates += (int64_t)d.inter_cands[i * kSlot64];
    out->intra_candidates += (int64_t)d.intra_cands[i * kSlot64];
    out->reeval_candidates += (int64_t)d.reeval_cands[i * kSlot64];
    out->exact_inter_candidates += (int64_t)d.exact_inter[i * kSlot64];
    out->exact_intra_candidates += (int64_t)d.exact_intra[i * kSlot64];
  }
  out->n_unique = d.n_unique;
  out->intra_iterations = (int32_t)d.active_passes + 1;
  out->is_intra = c->last_intra;
  return MPTC_OK;
}

int32_t mptc_profile_enable(mptc_ctx* c, int32_t enable) {
  if (!c) return fail(MPTC_EINVAL, "null context");
  HIP_TRY(hipSetDevice(c->cfg.device));
  int rc = collect_profile(c);
  if (rc) return rc;
  for (int i = 0; i < KF_N; ++i) {
    c->prof_n[i] = 0;
    c->prof_ms[i] = 0.0;
  }
  c->prof = enable != 0;
  return MPTC_OK;
}

int32_t mptc_profile_read(mptc_ctx* c, const char* name, int64_t* launches, double* total_ms) {
  if (!c || !name || !launches || !total_ms) return fail(MPTC_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(c->cfg.device));
  int rc = collect_profile(c);
  if (rc) return rc;
  for (int i = 0; i < KF_N; ++i)
    if (std::strcmp(name, kKernelNames[i]) == 0) {
      *launches = c->prof_n[i];
      *total_ms = c->prof_ms[i];
      return MPTC_OK;
    }
  return fail(MPTC_EINVAL, std::string("unknown kernel name: ") + name);
}

// ---------------------------------------------------------------------------
// Learned proposals
// ---------------------------------------------------------------------------
static int check_net(mptc_ctx* c, int net) {
  if (!c) return fail(MPTC_EINVAL, "null context");
  if (net != MPTC_NET_INTER && net != MPTC_NET_INTRA) return fail(MPTC_EINVAL, "net must be 0 (inter) or 1 (intra)");
  HIP_TRY(hipSetDevice(c->cfg.device));
  // The network entry points run on the context's stream and synchronise it
  // before they return; a step queued on a caller's stream may still be
  // reading the weights or writing the outputs they read.
  if (c->done_stream && c->done_stream != c->stream) HIP_TRY(hipEventSynchronize(c->done_ev));
  return net_alloc(c->net[net]);
}

static int upload_params(mptc_ctx* c, NetState& n, const float* params) {
  hipStream_t s = c->stream;
  HIP_TRY(hipMemcpyAsync(n.w, params, sizeof(float) * kNetParams, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_train_copies, dim3(grid_for(kNetParams, 256)), dim3(256), 0, s, n.w, kNetParams, n.wt);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemsetAsync(n.m, 0, sizeof(float) * kNetParams, s));
  HIP_TRY(hipMemsetAsync(n.v, 0, sizeof(float) * kNetParams, s));
  hipLaunchKernelGGL(k_pack_infer, dim3(grid_for(kL3Rows * kNetHid, 256)), dim3(256), 0, s, n.w,
                     n.use_base ? n.base : nullptr, n.pk);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s));
  n.t = 0;
  n.ready = true;
  return MPTC_OK;
}

int32_t mptc_net_init(mptc_ctx* c, int32_t net, uint64_t seed) {
  int rc = check_net(c, net);
  if (rc) return rc;
  std::vector<float> p(kNetParams, 0.0f);
  std::mt19937_64 rng(seed);
  auto fill = [&](int off, int fan_in, int count, float gain) {
    std::normal_distribution<float> nd(0.0f, std::sqrt(gain / (float)fan_in));
    for (int i = 0; i < count; ++i) p[off + i] = nd(rng);
  };
  fill(kOffW1, kNetIn, kNetIn * kNetHid, 2.0f);
  fill(kOffW2, kNetHid, kNetHid * kNetHid, 2.0f);
  fill(kOffW3, kNetHid, kNetHid * kNetOut, 1.0f);
  return upload_params(c, c->net[net], p.data());
}

int32_t mptc_net_get_params(mptc_ctx* c, int32_t net, float* out) {
  int rc = check_net(c, net);
  if (rc) return rc;
  if (!out) return fail(MPTC_EINVAL, "null argument");
  if (!c->net[net].ready) return fail(MPTC_ESTATE, "network parameters not set");
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(out, c->net[net].w, sizeof(float) * kNetParams, hipMemcpyDeviceToHost));
  return MPTC_OK;
}

int32_t mptc_net_set_params(mptc_ctx* c, int32_t net, const float* params) {
  int rc = check_net(c, net);
  if (rc) return rc;
  if (!params) return fail(MPTC_EINVAL, "null argument");
  return upload_params(c, c->net[net], params);
}

int32_t mptc_net_forward(mptc_ctx* c, int32_t net, const uint8_t* d_px, int32_t n, float* d_out) {
  int rc = check_net(c, net);
  if (rc) return rc;
  if (!d_px || !d_out || n < 1) return fail(MPTC_EINVAL, "bad argument");
  if (!c->net[net].ready) return fail(MPTC_ESTATE, "network parameters not set");
  hipStream_t s = c->stream;
  uint4* h2p = nullptr;
  if ((rc = dalloc(&h2p, (size_t)(n + 31) / 32 * (kH2TileBytes / 16)))) return rc;
  if ((rc = work_ensure(c->nw, n, false)) || (rc = net_input_q(c, s, d_px, nullptr, n, c->nw.X)) ||
      (rc = infer_q(c, s, c->net[net], c->nw.X, n, h2p, d_out, net, 1, nullptr, nullptr))) {
    (void)hipStreamSynchronize(s);
    (void)hipFree(h2p);
    return rc;
  }
  const hipError_t e = hipStreamSynchronize(s);
  (void)hipFree(h2p);
  HIP_TRY(e);
  return MPTC_OK;
}

int32_t mptc_net_gradients(mptc_ctx* c, int32_t net, const uint8_t* d_px, const uint32_t* d_target,
                           const uint8_t* d_eps, int32_t n, float seed_w, float* grad_out, float* loss_out) {
  int rc = check_net(c, net);
  if (rc) return rc;
  if (!d_px || !d_target || !d_eps || !grad_out || n < 1) return fail(MPTC_EINVAL, "bad argument");
  NetState& ns = c->net[net];
  if (!ns.ready) return fail(MPTC_ESTATE, "network parameters not set");
  hipStream_t s = c->stream;
  float* d_loss = nullptr;
  if ((rc = dalloc(&d_loss, 1))) return rc;
  HIP_TRY(hipMemsetAsync(d_loss, 0, sizeof(float), s));
  if ((rc = work_ensure(c->nw, n, true)) ||
      (rc = net_grad_q(c, s, ns, c->nw, n, d_px, nullptr, d_target, d_eps, seed_w, d_loss, AdamStep{}))) {
    (void)hipFree(d_loss);
    return rc;
  }
  float loss = 0.0f;
  const hipError_t e1 = hipMemcpyAsync(grad_out, ns.g, sizeof(float) * kNetParams, hipMemcpyDeviceToHost, s);
  const hipError_t e2 = hipMemcpyAsync(&loss, d_loss, sizeof(float), hipMemcpyDeviceToHost, s);
  const hipError_t e3 = hipStreamSynchronize(s);
  (void)hipFree(d_loss);
  HIP_TRY(e1);
  HIP_TRY(e2);
  HIP_TRY(e3);
  if (loss_out) *loss_out = loss;
  return MPTC_OK;
}

int32_t mptc_net_train(mptc_ctx* c, int32_t net, const uint8_t* d_px, const uint32_t* d_target, const uint8_t* d_eps,
                       int32_t n_pairs, const mptc_train_cfg* cfg, float* loss_out) {
  int rc = check_net(c, net);
  if (rc) return rc;
  if (!d_px || !d_target || !d_eps || !cfg || n_pairs < 1) return fail(MPTC_EINVAL, "bad argument");
  if (cfg->batch < 1 || cfg->batch > (1 << 22) || cfg->steps < 1 || !(cfg->lr > 0.0f))
    return fail(MPTC_EINVAL, "bad training configuration");
  NetState& ns = c->net[net];
  if (!ns.ready) return fail(MPTC_ESTATE, "network parameters not set");
  const int batch = std::min(cfg->batch, n_pairs);
  const float b1 = cfg->beta1 > 0.0f ? cfg->beta1 : 0.9f, b2 = cfg->beta2 > 0.0f ? cfg->beta2 : 0.999f;
  const float epsv = cfg->eps > 0.0f ? cfg->eps : 1e-8f;
  hipStream_t s = c->stream;
  if ((rc = work_ensure(c->nw, batch, true))) return rc;
  int32_t* d_perm = nullptr;
  float* d_losses = nullptr;
  auto done = [&](int code) {
    (void)hipStreamSynchronize(s);
    if (d_perm) (void)hipFree(d_perm);
    if (d_losses) (void)hipFree(d_losses);
    return code;
  };
  if ((rc = dalloc(&d_perm, n_pairs)) || (rc = dalloc(&d_losses, cfg->steps))) return done(rc);
  if (hipMemsetAsync(d_losses, 0, sizeof(float) * cfg->steps, s) != hipSuccess)
    return done(fail(MPTC_EHIP, "hipMemsetAsync failed"));
  std::vector<int32_t> perm(n_pairs);
  for (int i = 0; i < n_pairs; ++i) perm[i] = i;
  std::mt19937_64 rng(cfg->seed);
  int pos = n_pairs;
  for (int step = 0; step < cfg->steps; ++step) {
    if (pos + batch > n_pairs) {
      // the previous epoch's upload must be done before the host order changes
      if (hipStreamSynchronize(s) != hipSuccess) return done(fail(MPTC_EHIP, "hipStreamSynchronize failed"));
      std::shuffle(perm.begin(), perm.end(), rng);
      if (hipMemcpyAsync(d_perm, perm.data(), sizeof(int32_t) * n_pairs, hipMemcpyHostToDevice, s) != hipSuccess)
        return done(fail(MPTC_EHIP, "hipMemcpyAsync failed"));
      pos = 0;
    }
    const int32_t* idx = d_perm + pos;
    pos += batch;
    float lr = cfg->lr;
    if (cfg->lr_final > 0.0f && cfg->steps > 1) {
      const double u = (double)step / (double)(cfg->steps - 1);
      lr = (float)(cfg->lr_final + 0.5 * (cfg->lr - cfg->lr_final) * (1.0 + std::cos(3.141592653589793 * u)));
    }
    ns.t += 1;
    const AdamStep adam{lr, b1, b2, (float)(1.0 - std::pow((double)b1, (double)ns.t)),
                        (float)(1.0 - std::pow((double)b2, (double)ns.t)), epsv};
    if ((rc = net_grad_q(c, s, ns, c->nw, batch, d_px, idx, d_target, d_eps, cfg->seed_weight, d_losses + step,
                         adam)))
      return done(rc);
  }
  hipLaunchKernelGGL(k_pack_infer, dim3(grid_for(kL3Rows * kNetHid, 256)), dim3(256), 0, s, ns.w,
                     ns.use_base ? ns.base : nullptr, ns.pk);
  if (hipGetLastError() != hipSuccess) return done(fail(MPTC_EHIP, "k_pack_infer launch failed"));
  if (loss_out &&
      hipMemcpyAsync(loss_out, d_losses, sizeof(float) * cfg->steps, hipMemcpyDeviceToHost, s) != hipSuccess)
    return done(fail(MPTC_EHIP, "hipMemcpyAsync failed"));
  return done(MPTC_OK);
}

int32_t mptc_record_pairs(mptc_ctx* c, int32_t enable) {
  if (!c) return fail(MPTC_EINVAL, "null context");
  HIP_TRY(hipSetDevice(c->cfg.device));
  if (enable && !c->d_amask[0]) {
    int rc;
    const size_t nb = (size_t)c->max_total * 16;
    if ((rc = dalloc(&c->d_amask[0], nb)) || (rc = dalloc(&c->d_amask[1], nb)) ||
        (rc = dalloc(&c->d_simcnt, 4 * kSimBins)))
      return rc;
  }
  if (enable) {  // a new recording starts the index-similarity statistics afresh
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemset(c->d_simcnt, 0, sizeof(unsigned long long) * 4 * kSimBins));
  }
  if ((enable != 0) != c->record) c->graphs.destroy();
  c->record = enable != 0;
  c->last_recorded = false;
  return MPTC_OK;
}

int32_t mptc_collect_pairs(mptc_ctx* c, int32_t net, uint8_t* d_px, uint32_t* d_target, uint8_t* d_eps_out,
                           uint32_t* d_count, int32_t capacity) {
  int rc = check_net(c, net);
  if (rc) return rc;
  if (!d_px || !d_target || !d_eps_out || !d_count || capacity < 0) return fail(MPTC_EINVAL, "bad argument");
  if (!c->last_recorded) return fail(MPTC_ESTATE, "the last step did not record acceptance sets");
  hipStream_t s = c->stream;  // check_net waited for a step queued on a caller's stream
  hipLaunchKernelGGL(k_collect_masks, dim3(grid_for(c->total, 256)), dim3(256), 0, s, c->d_fr, c->d_sid, c->total,
                     net, c->cfg.search_area, c->d_need, c->d_amask[net], c->d_own, d_px, d_target, d_eps_out,
                     d_count, capacity);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_sim_counts, dim3(grid_for(c->total, 256)), dim3(256), 0, s, c->d_fr, c->d_sid, c->total, net,
                     c->cfg.search_area, c->d_need, c->d_amask[net], c->d_own, c->d_rw, c->d_inter,
                     c->d_simcnt + 2 * kSimBins * net);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s));
  return MPTC_OK;
}

int32_t mptc_set_search_mode(mptc_ctx* c, int32_t mode, int32_t top_k) {
  if (!c) return fail(MPTC_EINVAL, "null context");
  if (mode < MPTC_SEARCH_EXHAUSTIVE || mode > MPTC_SEARCH_LEARNED) return fail(MPTC_EINVAL, "unknown search mode");
  if (top_k < 1 || top_k > kMaxTopK) return fail(MPTC_EINVAL, "top_k must be in 1..16");
  HIP_TRY(hipSetDevice(c->cfg.device));
  if (mode != MPTC_SEARCH_EXHAUSTIVE) {
    const int needed = mode == MPTC_SEARCH_LEARNED ? 2 : 1;
    for (int i = 0; i < needed; ++i)
      if (!c->net[i].ready && !c->net[i].use_prior)
        return fail(MPTC_ESTATE, "learned search needs trained network parameters or a prior");
    int rc;
    const size_t nb = (size_t)c->max_total;
    if ((rc = work_ensure(c->ew, c->max_total, false))) return rc;
    if (!c->d_h2p && ((rc = dalloc(&c->d_h2p, (nb + 31) / 32 * (kH2TileBytes / 16))) ||
                      (rc = dalloc(&c->d_pinter, nb * kMaxTopK)) || (rc = dalloc(&c->d_prop, nb * kMaxTopK)) ||
                      (rc = dalloc(&c->d_seen, nb * kMaxTopK)) || (rc = dalloc(&c->d_errs, nb * kMaxTopK))))
      return rc;
    if (!c->d_tflag) {  // one flag per tile and parity; a tile holds at least one block
      if ((rc = dalloc(&c->d_tflag, 2 * nb))) return rc;
      HIP_TRY(hipMemset(c->d_tflag, 0, 2 * nb));
    }
  }
  if (mode != c->mode || top_k != c->top_k) c->graphs.destroy();
  c->mode = mode;
  c->top_k = top_k;
  return MPTC_OK;
}

int32_t mptc_get_proposals(mptc_ctx* c, int32_t net, int16_t* out, int64_t n_blocks) {
  int rc = check_net(c, net);
  if (rc) return rc;
  if (!out || n_blocks < 0 || n_blocks > c->max_total) return fail(MPTC_EINVAL, "bad argument");
  const int16_t* src = net == MPTC_NET_INTER ? c->d_pinter : c->d_prop;
  if (!src) return fail(MPTC_ESTATE, "no learned step has run");
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(out, src, sizeof(int16_t) * (size_t)n_blocks * c->top_k, hipMemcpyDeviceToHost));
  return MPTC_OK;
}

int32_t mptc_net_set_base(mptc_ctx* c, int32_t net, const float* scores) {
  int rc = check_net(c, net);
  if (rc) return rc;
  NetState& n = c->net[net];
  hipStream_t s = c->stream;
  HIP_TRY(hipMemsetAsync(n.base, 0, sizeof(float) * kNetOut, s));
  if (scores) HIP_TRY(hipMemcpyAsync(n.base, scores, sizeof(float) * kNetLogits, hipMemcpyHostToDevice, s));
  n.use_base = scores != nullptr;
  if (n.ready) {  // the inference image carries the base in its layer-3 biases
    hipLaunchKernelGGL(k_pack_infer, dim3(grid_for(kL3Rows * kNetHid, 256)), dim3(256), 0, s, n.w,
                       n.use_base ? n.base : nullptr, n.pk);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipStreamSynchronize(s));
  c->graphs.destroy();
  return MPTC_OK;
}

int32_t mptc_net_set_similarity(mptc_ctx* c, int32_t net, const float* scores) {
  int rc = check_net(c, net);
  if (rc) return rc;
  NetState& n = c->net[net];
  if (scores) {
    HIP_TRY(hipMemcpyAsync(n.sim, scores, sizeof(float) * kSimBins, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  n.use_sim = scores != nullptr;
  c->graphs.destroy();
  return MPTC_OK;
}

int32_t mptc_get_similarity_counts(mptc_ctx* c, int32_t net, uint64_t* out) {
  int rc = check_net(c, net);
  if (rc) return rc;
  if (!out) return fail(MPTC_EINVAL, "null argument");
  if (!c->d_simcnt) return fail(MPTC_ESTATE, "no recording has been made");
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(out, c->d_simcnt + 2 * kSimBins * net, sizeof(uint64_t) * 2 * kSimBins, hipMemcpyDeviceToHost));
  return MPTC_OK;
}

int32_t mptc_set_prior(mptc_ctx* c, int32_t net, const float* scores) {
  int rc = check_net(c, net);
  if (rc) return rc;
  NetState& n = c->net[net];
  if (scores) {
    HIP_TRY(hipMemcpy(n.prior, scores, sizeof(float) * kNetLogits, hipMemcpyHostToDevice));
    n.use_prior = true;
  } else {
    if (c->mode == MPTC_SEARCH_LEARNED || (c->mode == MPTC_SEARCH_LEARNED_INTER && net == MPTC_NET_INTER))
      if (!n.ready) return fail(MPTC_ESTATE, "network parameters not set");
    n.use_prior = false;
  }
  c->graphs.destroy();
  return MPTC_OK;
}

}  // extern "C"

// Sum of squared RGB differences between a frame and the decoding of its
// physical blocks (what a decoder of the stream shows): the numerator of
// DXTImage::PSNR (codec/dxt_image.cpp:363-383), taken over physical blocks.
// One lane per block, a wave sum, one 64-bit atomic per wave.
__global__ void __launch_bounds__(256) k_frame_sse(const uint8_t* __restrict__ rgb, int w, int nb,
                                                   const uint64_t* __restrict__ phys,
                                                   unsigned long long* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long sse = 0;
  if (k < nb) {
    const int bw = w >> 2;
    BlockPx px;
    load_block(rgb, w, k % bw, k / bw, px);
    const uint64_t p = phys[k];
    uint32_t pal[4];
    phys_palette(p, pal);
    sse = block_sse(px, (uint32_t)(p >> 32), pal);
  }
  for (int o = 32; o > 0; o >>= 1) sse += __shfl_down(sse, o);
  if ((threadIdx.x & 63) == 0 && sse) atomicAdd(out, sse);
}

__global__ void k_zero_u64(unsigned long long* p) { *p = 0; }

extern "C" {

int32_t mptc_frame_sse(const uint8_t* d_rgb, int32_t w, int32_t h, const uint64_t* d_phys, uint64_t* d_sse_out,
                       void* hip_stream) {
  if (!d_rgb || !d_phys || !d_sse_out) return fail(MPTC_EINVAL, "null argument");
  if (w <= 0 || h <= 0 || (w & 3) || (h & 3)) return fail(MPTC_EINVAL, "width and height must be positive multiples of 4");
  const long nbl = (long)(w >> 2) * (h >> 2);
  if (nbl > (1L << 30)) return fail(MPTC_EINVAL, "frame too large");
  const int nb = (int)nbl;
  hipStream_t s = static_cast<hipStream_t>(hip_stream);
  auto* out = reinterpret_cast<unsigned long long*>(d_sse_out);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, out);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_frame_sse, dim3(grid_for(nb, 256)), dim3(256), 0, s, d_rgb, w, nb, d_phys, out);
  HIP_TRY(hipGetLastError());
  return MPTC_OK;
}

int32_t mptc_synth_frames(const void* cfg, int32_t first, int32_t n, uint8_t* rgb_out) {
  if (!cfg || !rgb_out || n < 0) return fail(MPTC_EINVAL, "bad argument");
  const mptc_synth_cfg* sc = static_cast<const mptc_synth_cfg*>(cfg);
  if (sc->width <= 0 || sc->height <= 0) return fail(MPTC_EINVAL, "bad frame size");
  const size_t fb = (size_t)sc->width * sc->height * 3;
  for (int f = 0; f < n; ++f) mptc_synth_frame(sc, first + f, rgb_out + fb * f);
  return MPTC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Entry point that allocates nothing: mptc_encode_frame_noalloc.
//
// The work buffers of one frame of up to kNaMaxBlocks blocks at search areas
// up to kNaMaxS live in `scratch`, a device global: every device this library
// runs on holds its own copy of it (about 53 MiB) until the process ends.  On
// a device's first call the entry point looks up that copy's address
// (hipGetSymbolAddress) and lays the buffers out in it; host code never uses
// the array's name as an address.  Calls on one device are serialised (one
// mutex per device) and each returns once its frame is done, so calls never
// overlap on the array; each call's first kernel clears the counters the
// frame needs (nothing depends on what the previous call left).
// ---------------------------------------------------------------------------
namespace {

constexpr int kNaMaxS = 8;
constexpr int kNaMaxBlocks = 65536;  // e.g. 1280x720 (57,600 blocks) and every smaller frame
constexpr size_t kNaNbs = kNaMaxBlocks;  // E row stride, a multiple of 64

enum NaBuf {
  NA_FR, NA_SID, NA_TSID, NA_NEED, NA_OWN, NA_INTER, NA_INTRA, NA_RW, NA_CW, NA_CWN, NA_DIRTY, NA_TILES,
  NA_RTAB, NA_CTAB, NA_ORIG, NA_CHG, NA_E, NA_SUMM, NA_STATS, NA_RBT, NA_CBT, NA_PM, NA_CONV, NA_N
};

constexpr size_t na_align(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

// Bytes of each work buffer (the sizes mptc_create allocates for one frame
// of kNaMaxBlocks blocks at S = kNaMaxS).
constexpr size_t na_bytes(int b) {
  switch (b) {
    case NA_FR: return sizeof(FrameDev);
    case NA_SID: case NA_TSID: case NA_NEED: return kNaNbs;  // a wave tile holds at least one block
    case NA_OWN: case NA_INTER: case NA_INTRA: return 8 * kNaNbs;
    case NA_RW: case NA_CW: case NA_CWN: case NA_DIRTY: case NA_ORIG: case NA_CHG: return 4 * kNaNbs;
    case NA_TILES: return 4 * (kNaMaxBlocks / kUniqueTile + 2);
    case NA_RTAB: case NA_CTAB: return sizeof(float4) * kNaNbs;
    case NA_E: return static_cast<size_t>(kNaMaxS) * kNaMaxS * kNaNbs * sizeof(uint2);
    case NA_SUMM: return static_cast<size_t>(kMaxChunks) * kNaNbs * sizeof(uint4);
    case NA_STATS: return sizeof(DevStats);
    case NA_RBT: case NA_CBT: return sizeof(BoundTab) * kNaNbs;
    case NA_PM: return sizeof(uint32_t) * kPassWords * kNaNbs;
    case NA_CONV: return 64 * sizeof(uint32_t);
    default: return 0;
  }
}

constexpr size_t na_offset(int b) {
  size_t off = 0;
  for (int i = 0; i < b; ++i) off = na_align(off + na_bytes(i));
  return off;
}

constexpr size_t kNaScratchBytes = na_offset(NA_N);
static_assert(kNaScratchBytes <= (size_t)56 << 20, "scratch stays below 56 MiB");

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNaScratchBytes];

namespace {

__global__ void __launch_bounds__(256) k_zero_bytes(uint8_t* __restrict__ p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

struct NoAllocDev {
  std::mutex mu;           // calls on one device run one at a time
  bool ready = false;      // buffers laid out in this device's copy of `scratch`
  mptc_ctx ctx;            // work-buffer pointers into `scratch`; no allocation of its own
  FrameDev h_fr{};         // staging of the frame descriptor
  uint32_t h_conv[64] = {};  // host side of the context's status words (the kernels write the device side)
  uint32_t h_changed = 0;    // changed words of the last fixpoint pass, copied back
};

NoAllocDev g_na[kMaxDevices];

// Once per device, with the device current: the address of this device's copy
// of `scratch` and the buffer layout in it.
int na_setup(int device, NoAllocDev& nd) {
  int rc = init_tables(device);
  if (rc) return rc;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(MPTC_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  unsigned char* base = static_cast<unsigned char*>(p);
  auto at = [&](int b) { return static_cast<void*>(base + na_offset(b)); };
  mptc_ctx& c = nd.ctx;
  c.cfg = mptc_config{device, 4, 4, kNaMaxS, 0, 1, kNaMaxBlocks};
  c.alloc_S = kNaMaxS;
  c.max_total = kNaMaxBlocks;
  c.max_frames = 1;
  c.nbs = kNaNbs;
  c.d_fr = static_cast<FrameDev*>(at(NA_FR));
  c.d_sid = static_cast<uint8_t*>(at(NA_SID));
  c.d_tsid = static_cast<uint8_t*>(at(NA_TSID));
  c.d_need = static_cast<uint8_t*>(at(NA_NEED));
  c.d_own = static_cast<uint64_t*>(at(NA_OWN));
  c.d_inter = static_cast<uint64_t*>(at(NA_INTER));
  c.d_intra = static_cast<uint64_t*>(at(NA_INTRA));
  c.d_rw = static_cast<uint32_t*>(at(NA_RW));
  c.d_cw = static_cast<uint32_t*>(at(NA_CW));
  c.d_cwn = static_cast<uint32_t*>(at(NA_CWN));
  c.d_dirty = static_cast<uint32_t*>(at(NA_DIRTY));
  c.d_tiles = static_cast<uint32_t*>(at(NA_TILES));
  c.d_rtab = static_cast<float4*>(at(NA_RTAB));
  c.d_ctab = static_cast<float4*>(at(NA_CTAB));
  c.d_orig = static_cast<int*>(at(NA_ORIG));
  c.d_chg = static_cast<int*>(at(NA_CHG));
  c.d_E = static_cast<uint2*>(at(NA_E));
  c.d_summ = static_cast<uint4*>(at(NA_SUMM));
  c.d_stats = static_cast<DevStats*>(at(NA_STATS));
  c.d_rbt = static_cast<BoundTab*>(at(NA_RBT));
  c.d_cbt = static_cast<BoundTab*>(at(NA_CBT));
  c.d_pm = static_cast<uint32_t*>(at(NA_PM));
  c.d_conv = static_cast<uint32_t*>(at(NA_CONV));
  c.h_conv = nd.h_conv;
  c.tune = mptc_tuning{256, 16, 2, 0, 4, 1, 0};  // direct launches
  int cu = 0;
  if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cu > 0) c.n_cu = cu;
  nd.ready = true;
  return MPTC_OK;
}

}  // namespace

extern "C" {

int32_t mptc_encode_frame_noalloc(int32_t device, const uint8_t* d_rgb, int32_t w, int32_t h, int32_t search_area,
                                  int32_t err_threshold, int32_t is_intra, const uint64_t* d_ref_phys,
                                  uint64_t* d_phys_out, uint8_t* d_motion_out, uint32_t* d_unique_out,
                                  uint32_t* d_n_unique_out, uint8_t* d_eps_out, uint8_t* d_kind_out,
                                  void* hip_stream) {
  if (device < 0 || device >= kMaxDevices) return fail(MPTC_EINVAL, "no such HIP device");
  if (w <= 0 || h <= 0 || (w & 3) || (h & 3))
    return fail(MPTC_EINVAL, "width and height must be positive multiples of 4");
  if ((long long)(w / 4) * (h / 4) > kNaMaxBlocks)
    return fail(MPTC_EINVAL, "frame larger than the allocation-free maximum of 65536 blocks");
  if (search_area < 1 || search_area > kNaMaxS)
    return fail(MPTC_EINVAL, "search_area must be in 1..8 for the allocation-free entry point");
  int rc = check_params(search_area, err_threshold);
  if (rc) return rc;
  if (!d_rgb || !d_phys_out || !d_motion_out || !d_unique_out || !d_n_unique_out)
    return fail(MPTC_EINVAL, "null argument");
  if (!is_intra && !d_ref_phys) return fail(MPTC_EINVAL, "inter frame needs the previous frame's blocks");
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device >= ndev) return fail(MPTC_EINVAL, "no such HIP device");
  NoAllocDev& nd = g_na[device];
  std::lock_guard<std::mutex> lk(nd.mu);
  HIP_TRY(hipSetDevice(device));
  if (!nd.ready && (rc = na_setup(device, nd))) return rc;
  mptc_ctx* c = &nd.ctx;
  hipStream_t s = static_cast<hipStream_t>(hip_stream);
  c->cfg.search_area = search_area;
  c->cfg.err_threshold = err_threshold;
  const FrameIn in{d_rgb,          w, h, is_intra, is_intra ? nullptr : d_ref_phys, d_phys_out, d_motion_out,
                   d_unique_out, d_n_unique_out, d_eps_out, d_kind_out};
  int base = 0, tb = 0, wt = 0, lt = 0;
  nd.h_fr = frame_dev(in, c->tune.tile_w, itile_eff(c->tune), base, tb, wt, lt);
  HIP_TRY(hipMemcpyAsync(c->d_fr, &nd.h_fr, sizeof(FrameDev), hipMemcpyHostToDevice, s));
  // one frame: every block and wave tile belongs to frame 0; the pass count
  // word cleared (a step whose first search changes nothing has no pass that
  // writes it)
  hipLaunchKernelGGL(k_fill_sid, dim3(grid_for(base, 256), 1), dim3(256), 0, s, c->d_fr, 1, c->d_sid);
  hipLaunchKernelGGL(k_zero_bytes, dim3(grid_for(wt, 256)), dim3(256), 0, s, c->d_tsid, wt);
  hipLaunchKernelGGL(k_zero_bytes, dim3(1), dim3(256), 0, s, reinterpret_cast<uint8_t*>(c->d_conv),
                     (int)(64 * sizeof(uint32_t)));
  HIP_TRY(hipGetLastError());
  c->n_frames = 1;
  c->total = base;
  c->n_wtiles = wt;
  c->n_tiles = tb;
  c->n_itiles = lt;
  c->any_inter = !is_intra;
  c->last_intra = is_intra;
  if ((rc = queue_prefix(c, s, true))) return rc;
  // fixpoint passes in batches; after each batch the count of changed words
  // of its last pass comes back by a plain copy (no event, no mapped memory)
  const int np = passes_per_batch(c);
  for (int passes = 0;; passes += np) {
    if (passes > c->total + 4 * np) return fail(MPTC_ESTATE, "intra fixpoint did not converge");
    if ((rc = queue_passes(c, s))) return rc;
    HIP_TRY(hipMemcpyAsync(&nd.h_changed, c->d_conv, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (nd.h_changed == 0) break;
  }
  if ((rc = queue_suffix(c, s))) return rc;
  HIP_TRY(hipStreamSynchronize(s));  // the frame is done when the call returns
  return MPTC_OK;
}

}  // extern "C"

