// Host side of the kNN wrapper feature-selection engine: the C ABI of
// include/speedynn.h over HIP.  One selection runs as a fixed sequence of
// kernels per step on one stream (gather candidates, search, losses, arg-min,
// update); every step reads the device-resident state, so the host issues all
// steps without synchronising and reads the results once at the end.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/speedynn.h"
#include "common.h"
#include "engine.h"
#include "kernels.h"

// Device work memory of snn_select_device_noalloc: one fixed array on every
// GPU the library runs on, held until the process ends.  Its host-side name
// is only a stand-in: host code uses the device address hipGetSymbolAddress
// gives for the current device (snn_create keeps it per device), never
// `scratch` itself.  Calls on one GPU share it, so overlapping calls race;
// each call initialises every part it reads.
constexpr size_t kFixedScratchBytes = (size_t)SNN_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) char scratch[kFixedScratchBytes];

using snn::DevBuf;
using snn::fail;
using snn::Problem;
using snn::SelState;

namespace {

thread_local std::string g_err = "no error";

int validate_losses(const int* loss_functions, int num_loss) {
  if (num_loss < 1) return fail(SNN_EINVAL, "at least one loss function is required");
  for (int l = 0; l < num_loss; ++l) {
    const int c = loss_functions[l];
    if (c < 0 || c > 8) return fail(SNN_EINVAL, "unknown loss function code " + std::to_string(c));
  }
  return SNN_OK;
}

}  // namespace

namespace snn {

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int validate_problem(int64_t n, int d, const int* k_vals, int num_k, const int* loss_functions,
                     int num_loss, const snn_params* prm, Problem* out) {
  if (!prm) return fail(SNN_EINVAL, "null params");
  if (n < 2 || d < 1) return fail(SNN_EINVAL, "need n >= 2 points and d >= 1 features");
  if (n > (int64_t)INT32_MAX) return fail(SNN_EINVAL, "n exceeds the 32-bit neighbour index");
  if (prm->method != 0 && prm->method != 1) return fail(SNN_EINVAL, "method must be 0 or 1");
  if (prm->model_type != 0 && prm->model_type != 1)
    return fail(SNN_EINVAL, "model_type must be 0 (regression) or 1 (classification)");
  if (prm->form != SNN_FORM_PLAIN && prm->form != SNN_FORM_PRECISION)
    return fail(SNN_EINVAL, "form must be SNN_FORM_PLAIN or SNN_FORM_PRECISION");
  if (prm->k_fold < 1 || prm->k_fold > n)
    return fail(SNN_EINVAL, "k_fold must be in [1, n] (k_fold > n divides by zero in the reference)");
  if (num_k < 1 || !k_vals) return fail(SNN_EINVAL, "at least one k value is required");
  int kmax = 0;
  for (int m = 0; m < num_k; ++m) {
    if (k_vals[m] < 1) return fail(SNN_EINVAL, "k values must be >= 1");
    kmax = k_vals[m] > kmax ? k_vals[m] : kmax;
  }
  if ((int64_t)kmax * 64 * 12 > ((int64_t)512 << 20))
    return fail(SNN_ENOTSUP, "k values above 699050 are not supported");
  if (!loss_functions) return fail(SNN_EINVAL, "null loss function list");
  int rc = validate_losses(loss_functions, num_loss);
  if (rc) return rc;
  if (!prm->auto_feature && prm->num_features < 1)
    return fail(SNN_EINVAL, "num_features must be > 0");
  if (!prm->auto_feature && prm->num_features > d)
    return fail(SNN_EINVAL, "num_features must be <= the number of features");

  Problem pb{};
  pb.n = n;
  pb.npf = n / prm->k_fold;
  pb.d = d;
  pb.cpad = (d + 63) / 64 * 64;
  if (n * (int64_t)pb.cpad >= ((int64_t)1 << 31))
    return fail(SNN_ENOTSUP, "n * padded d must stay below 2^31 (32-bit gather offsets)");
  pb.nk = num_k;
  pb.nl = num_loss;
  pb.kmax = kmax;
  pb.method = prm->method;
  pb.model_type = prm->model_type;
  pb.auto_feature = prm->auto_feature ? 1 : 0;
  const bool fwd = prm->method == 1;
  pb.num_desired = prm->auto_feature ? (fwd ? d : 0) : prm->num_features;
  pb.max_steps = fwd ? pb.num_desired : d - pb.num_desired;
  *out = pb;
  return SNN_OK;
}

}  // namespace snn

namespace {

struct Timing {
  std::vector<hipEvent_t> ev;
  hipEvent_t make() {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    ev.push_back(e);
    return e;
  }
  ~Timing() {
    for (auto e : ev) hipEventDestroy(e);
  }
};

double elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (a && b && hipEventElapsedTime(&ms, a, b) == hipSuccess) return ms;
  return 0.0;
}

// The context's event pool, created on first use (hipSetDevice done).
int ensure_events(snn_ctx* ctx) {
  EventPool& ep = ctx->events;
  if (ep.ready) return SNN_OK;
  ep.ring.assign((size_t)EventPool::kRingSteps * EventPool::kMarks, nullptr);
  for (hipEvent_t& e : ep.ring) HIPCHK(hipEventCreate(&e));
  for (hipEvent_t& e : ep.tune) HIPCHK(hipEventCreate(&e));
  for (hipEvent_t& e : ep.span) HIPCHK(hipEventCreate(&e));
  ep.ready = true;
  return SNN_OK;
}

void destroy_events(snn_ctx* ctx) {
  EventPool& ep = ctx->events;
  for (hipEvent_t e : ep.ring)
    if (e) hipEventDestroy(e);
  for (hipEvent_t e : ep.tune)
    if (e) hipEventDestroy(e);
  for (hipEvent_t e : ep.span)
    if (e) hipEventDestroy(e);
  ep = EventPool{};
}

// Autotuner of the pruned search over its tile shapes.  Which is faster -- a
// wave per (query, candidate block), or a workgroup of W = 4, 8 or 16 waves
// per (query, W candidate blocks) sharing the row scan (ceil(blocks / W)
// workgroups per query) -- depends on the data (how much the candidate
// blocks' bounds differ: a workgroup waits for its loosest block) and on the
// step (early steps keep many points, late ones few), not only on the shape.
// The first step of a shape (n, d, list size, direction) and step range
// [2^b, 2^(b+1)) runs every applicable kernel on the same step, timed with
// HIP events, and the fastest serves the rest of that range.  With nc kernels to try, kernel r serves the
// queries r, r + nc, r + 2 nc, ... of that step (the order of the kernels
// rotates with the range, so no kernel always pays for cold caches): the
// parts are disjoint and cover every query once, and every kernel gives the
// same bits, so the step costs about one run, not nc.  (Below 64 nc points
// each kernel runs the whole step instead; only the last run counts visits.)
// Choices persist: SPEEDYNN_TUNE_FILE names a file of choices read at first
// use and rewritten after every new choice; without it the file tuned.json
// beside the library is read (never written), so a process starts with the
// choices committed there.  SPEEDYNN_TUNING=off keeps the rule by direction;
// so does a call without events (`tune_ev` null), which runs a kernel already
// chosen for the key.
std::mutex g_tune_mu;
using TuneKey = std::tuple<int64_t, int, int, int, int>;  // n, d, list size, mult, range
std::map<TuneKey, int> g_tuned;
// Row-list capacity per shape (n, d, list size, mult): n / divisor entries per
// row (default 8).  A longer cap keeps lists for rows whose list would overflow
// (the search then scans the list, not the row) at the price of writing them;
// which wins depends on the data, so it is a recorded choice, measured per
// bench shape, not tuned online.
using ShapeKey = std::tuple<int64_t, int, int, int>;
std::map<ShapeKey, int> g_listcap;
bool g_tune_loaded = false;
std::string g_tune_write;  // file rewritten after new choices (empty: none)

std::string library_dir() {
  Dl_info info{};
  if (!dladdr(reinterpret_cast<void*>(&library_dir), &info) || !info.dli_fname) return ".";
  std::string p = info.dli_fname;
  const size_t k = p.rfind('/');
  return k == std::string::npos ? "." : p.substr(0, k);
}

// One choice per line: {"n": N, "d": D, "list": K, "mult": M, "range": B, "kernel": V}.
void load_tuned_locked() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  const char* e = std::getenv("SPEEDYNN_TUNE_FILE");
  std::string path = e && *e ? std::string(e) : library_dir() + "/tuned.json";
  if (e && *e) g_tune_write = path;
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return;
  char line[512];
  while (std::fgets(line, sizeof line, f)) {
    long long n = 0;
    int d = 0, K = 0, mult = 0, b = 0, v = 0;
    if (std::sscanf(line, " { \"n\": %lld, \"d\": %d, \"list\": %d, \"mult\": %d, \"range\": %d, "
                          "\"kernel\": %d",
                    &n, &d, &K, &mult, &b, &v) == 6 &&
        v >= 1 && v <= 7)
      g_tuned[std::make_tuple((int64_t)n, d, K, mult, b)] = v;
    else if (std::sscanf(line, " { \"n\": %lld, \"d\": %d, \"list\": %d, \"mult\": %d, "
                               "\"list_cap_div\": %d",
                         &n, &d, &K, &mult, &v) == 5 &&
             v >= 2 && v <= 64)
      g_listcap[std::make_tuple((int64_t)n, d, K, mult)] = v;
  }
  std::fclose(f);
}

void save_tuned_locked() {
  if (g_tune_write.empty()) return;
  const std::string tmp = g_tune_write + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "w");
  if (!f) return;
  std::fprintf(f, "[\n");
  size_t i = 0;
  const size_t total = g_tuned.size() + g_listcap.size();
  for (const auto& kv : g_tuned) {
    const TuneKey& k = kv.first;
    std::fprintf(f, "  {\"n\": %lld, \"d\": %d, \"list\": %d, \"mult\": %d, \"range\": %d, "
                    "\"kernel\": %d}%s\n",
                 (long long)std::get<0>(k), std::get<1>(k), std::get<2>(k), std::get<3>(k),
                 std::get<4>(k), kv.second, ++i < total ? "," : "");
  }
  for (const auto& kv : g_listcap) {
    const ShapeKey& k = kv.first;
    std::fprintf(f, "  {\"n\": %lld, \"d\": %d, \"list\": %d, \"mult\": %d, "
                    "\"list_cap_div\": %d}%s\n",
                 (long long)std::get<0>(k), std::get<1>(k), std::get<2>(k), std::get<3>(k),
                 kv.second, ++i < total ? "," : "");
  }
  std::fprintf(f, "]\n");
  std::fclose(f);
  std::rename(tmp.c_str(), g_tune_write.c_str());
}

// Entries per row list for a shape: n / divisor (SPEEDYNN_LIST_CAP_DIV, else a
// recorded choice, else 8), at least 256, at most n.
int64_t list_cap(int64_t n, int d, int K, int mult) {
  static const int forced = [] {
    const char* e = std::getenv("SPEEDYNN_LIST_CAP_DIV");
    const int v = e ? std::atoi(e) : 0;
    return v >= 2 && v <= 64 ? v : 0;
  }();
  int div = 8;
  if (forced) {
    div = forced;
  } else {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    load_tuned_locked();
    auto it = g_listcap.find(std::make_tuple(n, d, K, mult));
    if (it != g_listcap.end()) div = it->second;
  }
  return std::min<int64_t>(n, std::max<int64_t>(256, n / div));
}

// Variant 4: the walk along the first chosen feature's sorted column, for a
// forward selection's second step (rank != null); it has a step range of its own.
hipError_t launch_variant(const snn::SearchProblem& sl, int variant, const int* rank,
                          hipStream_t s) {
  if (variant == 4) return snn::launch_search_walk(sl, rank, s);
  return snn::launch_search_pruned(sl, variant, s);
}

// The pruned kernel a step runs: `variant` (0: the rule by direction), or
// tuning trials (`tune`) where the shape and step range have no recorded
// choice and events exist to time them.
struct PrunedPlan {
  int variant = 0;
  bool tune = false;
  int cands[6] = {1, 0, 0, 0, 0, 0};
  int nc = 1;
  int bucket = 0;
  TuneKey key;
};

PrunedPlan plan_pruned(const snn::SearchProblem& sl, int d, int step, const int* rank,
                       bool events) {
  static const bool off = [] {
    const char* e = std::getenv("SPEEDYNN_TUNING");
    return e && !std::strcmp(e, "off");
  }();
  PrunedPlan pp;
  while ((2 << pp.bucket) <= step + 1) ++pp.bucket;
  if (rank) pp.bucket = 64;
  pp.key = std::make_tuple(sl.n, d, snn::search_list_size(sl.kmax), sl.mult, pp.bucket);
  // the tile shapes that apply: a wave per (query, candidate block), and the
  // shared row scan in workgroups of W = 4, 8, 16 waves (one query x W
  // candidate blocks; a width is tried where more than half of its waves
  // hold candidate blocks), then the second forward step's walk
  const int nblk = snn::live_blocks(sl);
  for (int W = 4; W <= 16; W *= 2)
    if (snn::pruned_width_applies(sl, W) && (W == 4 || 2 * nblk > W))
      pp.cands[pp.nc++] = W == 4 ? 5 : W == 8 ? 6 : 7;
  if (rank) pp.cands[pp.nc++] = 4;
  // SPEEDYNN_PRUNED=wave|shared|split|walk|shared4|shared8|shared16 runs that
  // kernel wherever it applies (read per call, for measurements and the
  // per-kernel tests); shared and split take the width by rule
  const char* fe = std::getenv("SPEEDYNN_PRUNED");
  if (fe && *fe) {
    const int f = !std::strcmp(fe, "wave") ? 1 : !std::strcmp(fe, "shared") ? 2
                : !std::strcmp(fe, "split") ? 3 : !std::strcmp(fe, "walk") ? 4
                : !std::strcmp(fe, "shared4") ? 5 : !std::strcmp(fe, "shared8") ? 6
                : !std::strcmp(fe, "shared16") ? 7 : 0;
    if (f == 2 && snn::pruned_shared_applies(sl)) pp.variant = 2;
    else if (f == 3 && snn::pruned_split_applies(sl)) pp.variant = 3;
    else if (f == 4 && rank) pp.variant = 4;
    else if (f >= 5 && snn::pruned_width_applies(sl, 4 << (f - 5))) pp.variant = f;
    else if (f != 4) pp.variant = 1;  // the per-wave kernel applies everywhere
    return pp;
  }
  if (!off && pp.nc > 1) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    load_tuned_locked();
    auto it = g_tuned.find(pp.key);
    if (it != g_tuned.end()) {
      pp.variant = it->second;
      if (pp.variant == 4 && !rank) pp.variant = 1;
    } else if (events) {
      pp.tune = true;
    }
  }
  return pp;
}

int search_pruned_tuned(const snn::SearchProblem& sl, int d, int step, const int* rank,
                        hipStream_t s, int* used, hipEvent_t* tune_ev) {
  PrunedPlan pp = plan_pruned(sl, d, step, rank, tune_ev != nullptr);
  int variant = pp.variant;
  if (!pp.tune) {
    HIPCHK(launch_variant(sl, variant, rank, s));
    int v = variant ? variant : (sl.mult > 0 || !snn::pruned_shared_applies(sl) ? 1 : 2);
    if (v >= 5 && !snn::pruned_width_applies(sl, 4 << (v - 5))) v = 1;
    if (v == 3 && !snn::pruned_split_applies(sl)) v = 1;
    *used |= 1 << (v - 1);
    return SNN_OK;
  }
  int* cands = pp.cands;
  const int nc = pp.nc, bucket = pp.bucket;
  const TuneKey key = pp.key;
  // the kernels' order rotates with the range
  {
    int rot[6];
    for (int i = 0; i < nc; ++i) rot[i] = cands[(i + bucket) % nc];
    for (int i = 0; i < nc; ++i) cands[i] = rot[i];
  }
  // every applicable kernel on its part of this step's queries
  hipEvent_t* ev = tune_ev;
  const bool parts = sl.n >= 64 * (int64_t)nc;
  HIPCHK(hipEventRecord(ev[0], s));
  for (int i = 0; i < nc; ++i) {
    snn::SearchProblem run = sl;
    if (parts) {
      run.q0 = i;
      run.qstep = nc;
      run.nq = (sl.n - i + nc - 1) / nc;
    } else if (i + 1 < nc) {
      run.visited = nullptr;
    }
    HIPCHK(launch_variant(run, cands[i], rank, s));
    HIPCHK(hipEventRecord(ev[i + 1], s));
    *used |= 1 << (cands[i] - 1);
  }
  HIPCHK(hipEventSynchronize(ev[nc]));
  variant = cands[0];
  double t[6];
  double best = t[0] = elapsed(ev[0], ev[1]);
  for (int i = 1; i < nc; ++i) {
    t[i] = elapsed(ev[i], ev[i + 1]);
    if (t[i] < best) {
      best = t[i];
      variant = cands[i];
    }
  }
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tuned[key] = variant;
    save_tuned_locked();
  }
  return SNN_OK;
}

// Device buffers of one selection.
struct Work {
  float *XT = nullptr, *D = nullptr, *Dlo = nullptr, *Xc = nullptr, *P = nullptr;
  float *err = nullptr, *val = nullptr, *errors = nullptr, *cand_err = nullptr;
  int *bestk = nullptr, *cand = nullptr, *sel = nullptr, *ordering = nullptr;
  int *kvals = nullptr, *losses = nullptr, *perm = nullptr, *iperm = nullptr;
  SelState* st = nullptr;
  double* key = nullptr;
  float *M = nullptr, *skey = nullptr, *med = nullptr, *tauw = nullptr, *frange = nullptr;
  int *lidx = nullptr, *lcount = nullptr;
  float* lbound = nullptr;
  float *lval = nullptr, *llb = nullptr;
  unsigned char* Mq = nullptr;  // M rounded up to 8 bits (n * n), then the row scales (n ints)
  int* Mexp = nullptr;
  int *seeds = nullptr, *sidx = nullptr, *rank = nullptr;
  float *long_d = nullptr, *long_lab = nullptr;
  int* long_i = nullptr;
  unsigned long long* visited = nullptr;
  void* sort_tmp = nullptr;
};

// Binds each requested buffer either to the context's grow-only device
// buffer of that role (arena null: hipMalloc when it must grow) or to a
// 256-byte aligned range of a caller-provided arena of `cap` bytes (no
// allocation; the request fails when the ranges do not fit).
class Binder {
 public:
  Binder(char* arena, size_t cap) : arena_(arena), cap_(cap) {}
  template <class T>
  void add(snn::DevBuf& buf, size_t bytes, T** slot) {
    items_.push_back({&buf, bytes ? bytes : 256, reinterpret_cast<void**>(slot)});
  }
  size_t total() const {
    size_t off = 0;
    for (const Item& it : items_) off = (off + 255) / 256 * 256 + it.bytes;
    return off;
  }
  int bind() {
    if (!arena_) {
      for (const Item& it : items_) {
        HIPCHK(it.buf->ensure(it.bytes));
        *it.slot = it.buf->p;
      }
      return SNN_OK;
    }
    const size_t need = total();
    if (need > cap_)
      return fail(SNN_EINVAL, "problem needs " + std::to_string(need) +
                                  " bytes of device work memory; the fixed scratch holds " +
                                  std::to_string(cap_));
    size_t off = 0;
    for (const Item& it : items_) {
      off = (off + 255) / 256 * 256;
      *it.slot = arena_ + off;
      off += it.bytes;
    }
    return SNN_OK;
  }

 private:
  struct Item {
    snn::DevBuf* buf;
    size_t bytes;
    void** slot;
  };
  char* arena_;
  size_t cap_;
  std::vector<Item> items_;
};

// Phase times of the steps issued so far, read from the event ring as the
// steps' slots come up for reuse (and at the end).
struct PhaseTimes {
  double search = 0, update = 0, tally = 0, bound = 0, search_first = 0;
};

int retire_step(snn_ctx* ctx, int t, int issued, hipEvent_t end, PhaseTimes* pt, bool trace) {
  constexpr int R = EventPool::kRingSteps, M = EventPool::kMarks;
  const hipEvent_t* m = &ctx->events.ring[(size_t)(t % R) * M];
  const hipEvent_t next = t + 1 < issued ? ctx->events.ring[(size_t)((t + 1) % R) * M] : end;
  HIPCHK(hipEventSynchronize(next));
  const double tally = elapsed(m[0], m[1]) + elapsed(m[3], m[4]);
  const double bound = elapsed(m[1], m[2]);
  const double search = elapsed(m[2], m[3]);
  const double update = elapsed(m[4], next);
  pt->tally += tally;
  pt->bound += bound;
  pt->search += search;
  pt->update += update;
  if (t == 0) pt->search_first = search;
  if (trace)
    std::fprintf(stderr, "speedynn step %d: search %.3f ms, bound %.3f, update %.3f, tally %.3f\n",
                 t, search, bound, update, tally);
  return SNN_OK;
}

// One selection.  arena == null: device buffers from the context (allocated
// on growth), phase timing and autotuning on.  arena != null: every device
// buffer is a range of the arena (cap bytes), no allocation, no events (the
// summary's times stay 0) and no tuning trials.
int select_impl(snn_ctx* ctx, const float* dX, int64_t n, int d, const float* dy,
                const int* k_vals, int num_k, const int* loss_functions, int num_loss,
                const snn_params* prm, int* selected_dimensions, int* selected_ordering,
                float* selected_errors, float* candidate_errors, snn_summary* summary,
                char* arena = nullptr, size_t cap = 0) {
  if (!ctx) return fail(SNN_EINVAL, "null context");
  Problem pb{};
  int rc = snn::validate_problem(n, d, k_vals, num_k, loss_functions, num_loss, prm, &pb);
  if (rc) return rc;
  const int kmax = pb.kmax;
  const bool fwd = prm->method == 1;
  // precision form: exact pruned search over D in the reference's bits; the
  // compensation term Dlo of D is kept only on request (nothing in the
  // selection reads it)
  // k values above 64: one search kernel with its lists in global memory
  // serves every step of either form (the pruned kernels keep their lists in
  // registers)
  const bool long_list = snn::search_list_size(kmax) < 0;
  const bool pruned = prm->form == SNN_FORM_PRECISION && !long_list;
  const bool comp = prm->form == SNN_FORM_PRECISION && prm->compensated;
  const bool timed = arena == nullptr;

  HIPCHK(hipSetDevice(ctx->device));
  const size_t nn = (size_t)n * (size_t)n;
  // precision form, forward: the first step runs over sorted feature columns
  const bool sorted_first = pruned && fwd;
  size_t sort_bytes = 0;
  if (sorted_first) {
    sort_bytes = snn::sort_columns_temp_bytes(n, d);
    if (!sort_bytes) return fail(SNN_EDEVICE, "segmented sort: no temporary storage size");
  }
  const bool median = snn::needs_median_scratch(loss_functions, num_loss);
  Work w;
  Binder bd(arena, cap);
  bd.add(ctx->XT, sizeof(float) * (size_t)n * d, &w.XT);
  bd.add(ctx->D, sizeof(float) * nn, &w.D);
  if (comp) bd.add(ctx->Dlo, sizeof(float) * nn, &w.Dlo);
  bd.add(ctx->Xc, sizeof(float) * (size_t)n * pb.cpad, &w.Xc);
  bd.add(ctx->P, sizeof(float) * (size_t)num_k * pb.cpad * n, &w.P);
  bd.add(ctx->err, sizeof(float) * (size_t)num_loss * num_k * pb.cpad, &w.err);
  bd.add(ctx->val, sizeof(float) * (size_t)num_loss * pb.cpad, &w.val);
  bd.add(ctx->bestk, sizeof(int) * pb.cpad, &w.bestk);
  bd.add(ctx->cand, sizeof(int) * d, &w.cand);
  bd.add(ctx->sel, sizeof(int) * d, &w.sel);
  bd.add(ctx->ordering, sizeof(int) * d, &w.ordering);
  bd.add(ctx->errors, sizeof(float) * (size_t)num_loss * d, &w.errors);
  bd.add(ctx->kvals, sizeof(int) * num_k, &w.kvals);
  bd.add(ctx->losses, sizeof(int) * num_loss, &w.losses);
  bd.add(ctx->state, sizeof(SelState), &w.st);
  bd.add(ctx->perm, sizeof(int) * d, &w.perm);
  bd.add(ctx->key, sizeof(double) * d, &w.key);
  // exact pruning with bounds from 128 seeds per query (backward also keeps M,
  // the max single-feature term, for its lower bound) and warm bounds from
  // the previous step's K-th distances
  if (pruned && !fwd) bd.add(ctx->M, sizeof(float) * nn, &w.M);
  if (pruned) {
    bd.add(ctx->seeds, sizeof(int) * (size_t)n * 128, &w.seeds);
    bd.add(ctx->visited, sizeof(unsigned long long) * snn::kVisitSlots * snn::kVisitStride,
           &w.visited);
    bd.add(ctx->tauw, sizeof(float) * (size_t)n * d, &w.tauw);
    bd.add(ctx->iperm, sizeof(int) * d, &w.iperm);
    bd.add(ctx->frange, sizeof(float) * 2, &w.frange);
  }
  // precision form: the fused update writes each row's list of points the
  // next search may visit (list_cap: n/8 entries per row unless recorded otherwise)
  const int64_t rows_cap =
      (pruned && n % 4 == 0) ? list_cap(n, d, snn::search_list_size(kmax), fwd ? 1 : -1) : 0;
  if (rows_cap) {
    bd.add(ctx->lidx, sizeof(int) * (size_t)n * rows_cap, &w.lidx);
    bd.add(ctx->lval, sizeof(float) * (size_t)n * rows_cap, &w.lval);
    if (!fwd) {
      bd.add(ctx->llb, sizeof(float) * (size_t)n * rows_cap, &w.llb);
      bd.add(ctx->Mq, (size_t)nn + sizeof(int) * (size_t)n, &w.Mq);
    }
    bd.add(ctx->lcount, sizeof(int) * (size_t)n, &w.lcount);
    bd.add(ctx->lbound, sizeof(float) * (size_t)n, &w.lbound);
  }
  if (sorted_first) {
    bd.add(ctx->skey, sizeof(float) * (size_t)n * d, &w.skey);
    bd.add(ctx->sidx, sizeof(int) * (size_t)n * d, &w.sidx);
    bd.add(ctx->sort_tmp, sort_bytes, &w.sort_tmp);
    bd.add(ctx->rank, sizeof(int) * (size_t)n, &w.rank);
  }
  // median losses: sort space for NaN-holding error vectors
  if (median) bd.add(ctx->med, sizeof(float) * 2 * (size_t)n * pb.cpad * num_k, &w.med);
  if (candidate_errors) bd.add(ctx->cand_err, sizeof(float) * (size_t)d * d, &w.cand_err);
  const int64_t long_waves = long_list ? snn::long_search_waves(n, pb.cpad, kmax) : 0;
  if (long_list) {
    const size_t slots = (size_t)long_waves * (size_t)kmax * 64;
    bd.add(ctx->long_d, sizeof(float) * slots, &w.long_d);
    bd.add(ctx->long_i, sizeof(int) * slots, &w.long_i);
    bd.add(ctx->long_lab, sizeof(float) * slots, &w.long_lab);
  }
  rc = bd.bind();
  if (rc) return rc;
  if (w.Mq) w.Mexp = reinterpret_cast<int*>(w.Mq + nn);  // nn % 16 == 0 (n % 4 == 0)
  if (timed) {
    rc = ensure_events(ctx);
    if (rc) return rc;
  }

  hipStream_t s = ctx->stream;
  HIPCHK(hipMemcpyAsync(w.kvals, k_vals, sizeof(int) * num_k, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(w.losses, loss_functions, sizeof(int) * num_loss, hipMemcpyHostToDevice,
                        s));
  if (w.cand_err) HIPCHK(snn::launch_fill(w.cand_err, (int64_t)d * d, std::nanf(""), s));
  if (pruned) {
    HIPCHK(snn::launch_fill_u64(w.visited, snn::kVisitSlots * snn::kVisitStride, s));
    HIPCHK(snn::launch_fill(w.tauw, (int64_t)n * d, snn::kMaxF, s));  // no warm bound yet
    if (w.lcount) HIPCHK(snn::launch_fill_i32(w.lcount, n, -1, s));   // no row lists yet
  }

  EventPool& ep = ctx->events;
  hipEvent_t e0 = timed ? ep.span[0] : nullptr, e1 = timed ? ep.span[1] : nullptr;
  hipEvent_t e2 = timed ? ep.span[2] : nullptr;
  if (timed) HIPCHK(hipEventRecord(e0, s));
  SelState* st = w.st;
  float* D = w.D;
  float* Dlo = w.Dlo;
  HIPCHK(snn::launch_transpose(dX, w.XT, n, d, s));
  // candidate slots: ascending feature index (plain), grouped by scale (pruned)
  HIPCHK(snn::launch_feature_order(w.XT, n, d, w.key, w.perm, w.iperm, pruned, s));
  HIPCHK(snn::launch_init_state(st, pb, w.cand, w.sel, w.ordering, w.errors, w.perm, s));
  if (sorted_first)
    HIPCHK(snn::launch_sort_columns(w.XT, n, d, st, w.skey, w.sidx, w.sort_tmp, sort_bytes, s));
  else if (pruned)  // the warm bounds need finite distances
    HIPCHK(snn::launch_check_finite(w.XT, (int64_t)n * d, st, s));
  if (fwd) {
    HIPCHK(snn::launch_fill(D, (int64_t)nn, 0.0f, s));
    if (Dlo) HIPCHK(snn::launch_fill(Dlo, (int64_t)nn, 0.0f, s));
  } else {
    HIPCHK(snn::launch_gram_exact(w.XT, D, Dlo, pruned ? w.M : nullptr, n, d, s));
    if (w.Mq) HIPCHK(snn::launch_quantize_m(w.M, w.Mq, w.Mexp, n, s));
  }
  if (timed) HIPCHK(hipEventRecord(e1, s));

  snn::SearchProblem sl{};
  sl.D = D;
  sl.Xc = w.Xc;
  sl.y = dy;
  sl.kvals = w.kvals;
  sl.st = st;
  sl.P = w.P;
  sl.n = n;
  sl.npf = pb.npf;
  sl.cpad = pb.cpad;
  sl.nk = num_k;
  sl.kmax = kmax;
  sl.mult = fwd ? 1 : -1;
  sl.model_type = prm->model_type;
  sl.zero = 0;
  sl.cand = w.cand;
  sl.d = d;
  if (sorted_first) {
    sl.sorted = 1;
    sl.skey = w.skey;
    sl.sidx = w.sidx;
  }
  if (pruned) {
    sl.M = fwd ? nullptr : w.M;
    sl.nseed = snn::backward_seed_count(kmax);
    sl.seeds = w.seeds;
    sl.visited = w.visited;
    sl.tauw = w.tauw;
    sl.iperm = w.iperm;
    sl.XT = w.XT;
    sl.frange = w.frange;
    if (rows_cap) sl.rows = snn::RowList{w.lidx, w.lval, w.llb, w.lcount, w.lbound, rows_cap};
  }

  // per step: [a] gather [b] bound seeds [b2] search [c] losses + arg-min
  // [u] update [next a]; the marks live in the context's event ring
  static const bool trace = std::getenv("SPEEDYNN_TRACE_STEPS") != nullptr;
  constexpr int R = EventPool::kRingSteps, MK = EventPool::kMarks;
  PhaseTimes pt;
  int used_variant = 0;
  // backward: M is recomputed over the remaining candidates each time their
  // number has halved (d/2, d/4, ...), keeping the bound fl(D - M) tight
  int m_refresh = d / 2;
  // precision form: each update also writes the next step's seeds; the first
  // pruned step of a backward run scans them
  const bool fused_seeds = pruned;
  // graph mode (params.use_graph; not with a caller's arena): a step's
  // launches are captured once per distinct sequence and replayed; a step
  // whose pruned kernel is still to be tuned runs on the stream.  Graph mode
  // records no phase times.
  const bool graphs = prm->use_graph && timed;
  const bool phases = timed && !graphs;
  const int lpaths = snn::loss_paths(loss_functions, num_loss);

  // one step's launches; mk: its five phase marks (null: none), tune_ev: the
  // tuner's events (null: no trials)
  // deferred writes of D (backward precision form with row lists): the
  // update after an even step leaves its term out of D in memory, the next
  // search's row reads subtract it, and the update after it applies both
  // terms and writes D -- every other n^2 write saved, every value the same.
  // SPEEDYNN_DEFER=off writes D after every step.
  static const bool defer_off = [] {
    const char* e = std::getenv("SPEEDYNN_DEFER");
    return e && !std::strcmp(e, "off");
  }();
  const bool defer_on = pruned && !fwd && rows_cap > 0 && !comp && !defer_off;
  auto search_defers = [&](int step) { return defer_on && step % 2 == 1 ? 1 : 0; };
  auto update_mode = [&](int step) { return defer_on ? (step % 2 == 0 ? 1 : 2) : 0; };
  // candidate blocks with live candidates at a step (each step moves one
  // feature out of the candidate list, or the selection has stopped)
  auto blocks_at = [&](int step) {
    const int live = (d - step + 63) / 64, all = pb.cpad / 64;
    return live < 1 ? 1 : (live < all ? live : all);
  };
  auto issue = [&](int step, bool refresh, hipEvent_t* mk, hipEvent_t* tune_ev) -> int {
    auto mark = [&](int i) { return mk ? hipEventRecord(mk[i], s) : hipSuccess; };
    sl.nblk = blocks_at(step);
    sl.defer = search_defers(step);
    HIPCHK(mark(0));
    HIPCHK(snn::launch_compact(dX, w.cand, st, w.Xc, n, d, pb.cpad, s));
    HIPCHK(mark(1));
    const bool pruned_step = pruned && (step > 0 || !fwd);
    if (refresh) {
      HIPCHK(snn::launch_max_square(w.Xc, st, w.M, n, pb.cpad, s));
      if (w.Mq) HIPCHK(snn::launch_quantize_m(w.M, w.Mq, w.Mexp, n, s));
    }
    // seeds: the previous step's update wrote them where it was fused
    if (pruned_step && !(fused_seeds && step > 0))
      HIPCHK(snn::launch_seed_scan(D, st, w.seeds, n, pb.npf, s));
    HIPCHK(mark(2));
    if (pruned_step) {
      // the second forward step may walk the first chosen feature's sorted column
      const int* rank = nullptr;
      if (sorted_first && step == 1) {
        HIPCHK(snn::launch_rank(w.sidx, st, w.rank, n, s));
        rank = w.rank;
      }
      int rc2 = search_pruned_tuned(sl, d, step, rank, s, &used_variant, tune_ev);
      if (rc2) return rc2;
    } else if (long_list) {
      HIPCHK(snn::launch_search_long(sl, w.long_d, w.long_i, w.long_lab, long_waves, s));
    } else {
      // first forward step (D = 0): sorted columns; the dense search runs
      // instead where X is not finite (each kernel checks the state's flag)
      if (sorted_first) HIPCHK(snn::launch_search_sorted(sl, d, s));
      snn::SearchProblem sz = sl;
      sz.zero = pruned ? 1 : 0;
      HIPCHK(snn::launch_search_dense(sz, s));
    }
    HIPCHK(mark(3));
    HIPCHK(snn::launch_loss(w.P, dy, w.losses, st, w.err, n, pb.cpad, num_k, num_loss, d, w.med,
                            lpaths, s));
    HIPCHK(snn::launch_select(st, pb, w.kvals, w.err, w.cand, w.sel, w.ordering, w.errors, w.val,
                              w.bestk, w.cand_err, w.perm, s));
    // the forward warm bound's largest square comes from the chosen column
    if (pruned && fwd && step + 1 < pb.max_steps)
      HIPCHK(snn::launch_col_range(w.XT, st, n, w.frange, s));
    HIPCHK(mark(4));
    // the last step's choice is not applied to D: no later search reads it
    // (the state's pending flag is reset when the next selection starts)
    if (step + 1 < pb.max_steps) {
      if (fused_seeds)
        HIPCHK(snn::launch_update_seeds(D, Dlo, w.XT, n, st, fwd ? 1 : -1, w.seeds, pb.npf,
                                        sl.rows, w.tauw, w.frange, d, w.cand, w.iperm, w.Mq, w.Mexp,
                                        s, update_mode(step)));
      else
        HIPCHK(snn::launch_update(D, Dlo, w.XT, n, st, -1, fwd ? 1 : -1, s));
    }
    return SNN_OK;
  };

  // what every captured launch embeds: buffers, inputs, shapes and options
  std::vector<size_t> gbase;
  if (graphs) {
    const void* ptrs[] = {dX, dy, w.XT, w.D, w.Dlo, w.Xc, w.P, w.err, w.val, w.errors,
                          w.cand_err, w.bestk, w.cand, w.sel, w.ordering, w.kvals, w.losses,
                          w.perm, w.iperm, w.st, w.key, w.M, w.skey, w.med, w.tauw, w.frange,
                          w.lidx, w.lcount, w.lbound, w.lval, w.llb, w.Mq, w.seeds, w.sidx, w.rank,
                          w.visited, w.long_d, w.long_i, w.long_lab};
    for (const void* q : ptrs) gbase.push_back((size_t)q);
    const int64_t vals[] = {n, d, pb.npf, pb.cpad, num_k, num_loss, kmax, pb.method,
                            pb.model_type, pb.auto_feature, pb.max_steps, pb.num_desired,
                            prm->form, comp, rows_cap, long_waves, lpaths};
    for (int64_t v : vals) gbase.push_back((size_t)v);
    for (int l = 0; l < num_loss; ++l) gbase.push_back((size_t)loss_functions[l]);
    if (ctx->step_graphs.map.size() > 256) ctx->step_graphs.clear();  // bound the cache
  }

  int issued = 0;
  for (int step = 0; step < pb.max_steps; ++step) {
    const bool refresh = pruned && !fwd && step > 0 && d - step <= m_refresh;
    StepGraphs::Entry* ge = nullptr;
    if (graphs) {
      const bool pruned_step = pruned && (step > 0 || !fwd);
      const int* rank = (sorted_first && step == 1) ? w.rank : nullptr;
      PrunedPlan pp;
      sl.nblk = blocks_at(step);
      sl.defer = search_defers(step);
      if (pruned_step) pp = plan_pruned(sl, d, step, rank, true);
      if (!pp.tune) {
        std::vector<size_t> key = gbase;
        const size_t shape[] = {(size_t)(step == 0), (size_t)refresh, (size_t)(rank != nullptr),
                                (size_t)(step + 1 < pb.max_steps), (size_t)pp.variant,
                                (size_t)sl.nblk, (size_t)sl.defer, (size_t)update_mode(step)};
        key.insert(key.end(), shape, shape + 8);
        auto it = ctx->step_graphs.map.find(key);
        if (it == ctx->step_graphs.map.end()) {
          const int before = used_variant;
          used_variant = 0;
          HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
          const int rc3 = issue(step, refresh, nullptr, nullptr);
          hipGraph_t g = nullptr;
          const hipError_t ec = hipStreamEndCapture(s, &g);
          if (rc3) {
            if (g) hipGraphDestroy(g);
            return rc3;
          }
          HIPCHK(ec);
          StepGraphs::Entry en;
          en.used = used_variant;
          used_variant = before;
          const hipError_t ei = hipGraphInstantiate(&en.exec, g, nullptr, nullptr, 0);
          hipGraphDestroy(g);
          HIPCHK(ei);
          it = ctx->step_graphs.map.emplace(std::move(key), en).first;
        }
        ge = &it->second;
      }
    }
    if (ge) {
      HIPCHK(hipGraphLaunch(ge->exec, s));
      used_variant |= ge->used;
    } else {
      hipEvent_t* mk = nullptr;
      if (phases) {
        if (step >= R) {  // the slot's previous step: read its times before reuse
          rc = retire_step(ctx, step - R, step, nullptr, &pt, trace);
          if (rc) return rc;
        }
        mk = &ep.ring[(size_t)(step % R) * MK];
      }
      rc = issue(step, refresh, mk, timed ? ep.tune : nullptr);
      if (rc) return rc;
    }
    if (refresh) m_refresh = (d - step) / 2;
    issued = step + 1;
  }
  if (timed) HIPCHK(hipEventRecord(e2, s));
  HIPCHK(hipStreamSynchronize(s));
  if (phases) {
    for (int t = issued > R ? issued - R : 0; t < issued; ++t) {
      rc = retire_step(ctx, t, issued, e2, &pt, trace);
      if (rc) return rc;
    }
  }

  SelState hs{};
  HIPCHK(hipMemcpy(&hs, st, sizeof(SelState), hipMemcpyDeviceToHost));
  if (selected_dimensions)
    HIPCHK(hipMemcpy(selected_dimensions, w.sel, sizeof(int) * d, hipMemcpyDeviceToHost));
  if (selected_ordering)
    HIPCHK(hipMemcpy(selected_ordering, w.ordering, sizeof(int) * d, hipMemcpyDeviceToHost));
  if (selected_errors)
    HIPCHK(hipMemcpy(selected_errors, w.errors, sizeof(float) * (size_t)num_loss * d,
                     hipMemcpyDeviceToHost));
  if (candidate_errors)
    HIPCHK(hipMemcpy(candidate_errors, w.cand_err, sizeof(float) * (size_t)d * d,
                     hipMemcpyDeviceToHost));

  if (summary) {
    std::memset(summary, 0, sizeof(*summary));
    summary->num_final = hs.num_final;
    summary->optimal_k = hs.best_k;
    summary->steps = hs.step;
    summary->pruned_kernel = used_variant;
    if (timed) {
      summary->ms_total = elapsed(e0, e2);
      summary->ms_distance = elapsed(e0, e1);
      summary->ms_search = pt.search;
      summary->ms_update = pt.update;
      summary->ms_tally = pt.tally;
      summary->ms_bound = pt.bound;
      summary->ms_search_first = pt.search_first;
    }
    // candidates evaluated per step: fwd d, d-1, ...; bwd d, d-1, ...
    int64_t ev = 0, pruned_full = 0;
    for (int t = 0; t < hs.step; ++t) {
      ev += (int64_t)(d - t) * n;
      if (pruned && (t > 0 || !fwd)) pruned_full += (int64_t)(d - t) * n * (n - pb.npf);
    }
    summary->evals = ev;
    if (pruned) {
      unsigned long long done = 0;
      std::vector<unsigned long long> slots((size_t)snn::kVisitSlots * snn::kVisitStride);
      HIPCHK(hipMemcpy(slots.data(), w.visited, sizeof(unsigned long long) * slots.size(),
                       hipMemcpyDeviceToHost));
      for (int i = 0; i < snn::kVisitSlots; ++i) done += slots[(size_t)i * snn::kVisitStride];
      summary->pruned_point_evals = pruned_full - (int64_t)done;
    }
  }
  if (hs.error) return fail(SNN_ESTATE, "no candidate had a finite validation error");
  return SNN_OK;
}

// Device address of `scratch` per device ordinal (null: not resolved yet).
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_addr;

std::mutex g_default_mu;
snn_ctx* g_default = nullptr;
int g_default_form = SNN_FORM_PRECISION;
std::vector<int> g_sel, g_order;
std::vector<float> g_errors;
int g_best_k = 0;

snn_ctx* default_ctx() {
  if (!g_default) {
    int dev = 0;
    if (const char* e = std::getenv("SPEEDYNN_DEVICE")) dev = std::atoi(e);
    if (snn_create(&g_default, dev) != SNN_OK) g_default = nullptr;
  }
  return g_default;
}

int train_common(int method, float* X, int n, int d, float* Y, int nY, int* kvals, int nk,
                 int num_features, int kfold, int* losses, int nl, int auto_feature,
                 int model_type) {
  std::lock_guard<std::mutex> lock(g_default_mu);
  if (nY != n) return fail(SNN_EINVAL, "Ytrain must have one label per row of Xtrain");
  snn_ctx* ctx = default_ctx();
  if (!ctx) return SNN_EDEVICE;
  snn_params p{};
  p.method = method;
  p.model_type = model_type;
  p.k_fold = kfold;
  p.num_features = num_features;
  p.auto_feature = auto_feature;
  p.form = g_default_form;
  std::vector<int> sel(d), order(d);
  std::vector<float> errs((size_t)nl * d);
  snn_summary sm{};
  int rc = snn_select(ctx, X, n, d, Y, kvals, nk, losses, nl, &p, sel.data(), order.data(),
                      errs.data(), nullptr, &sm);
  if (rc != SNN_OK) return rc;
  g_sel.swap(sel);
  g_order.swap(order);
  g_errors.swap(errs);
  g_best_k = sm.optimal_k;
  return sm.num_final;
}

}  // namespace

extern "C" {

const char* snn_last_error(void) { return g_err.c_str(); }

void snn_set_default_form(int form) { g_default_form = form; }

int snn_create(snn_ctx** out, int device) {
  if (!out) return fail(SNN_EINVAL, "null output pointer");
  *out = nullptr;
  int count = 0;
  HIPCHK(hipGetDeviceCount(&count));
  if (device < 0 || device >= count)
    return fail(SNN_EDEVICE, "HIP device " + std::to_string(device) + " not present (" +
                                 std::to_string(count) + " visible)");
  HIPCHK(hipSetDevice(device));
  void* fixed = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_addr.size() < count) g_scratch_addr.resize(count, nullptr);
    if (!g_scratch_addr[device]) {
      hipError_t es = hipGetSymbolAddress(&g_scratch_addr[device], HIP_SYMBOL(scratch));
      if (es != hipSuccess || !g_scratch_addr[device]) {
        g_scratch_addr[device] = nullptr;
        return fail(SNN_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") +
                                     hipGetErrorString(es));
      }
    }
    fixed = g_scratch_addr[device];
  }
  snn_ctx* c = new snn_ctx();
  c->device = device;
  c->fixed_scratch = fixed;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(SNN_EDEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  *out = c;
  return SNN_OK;
}

void snn_destroy(snn_ctx* c) {
  if (!c) return;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  DevBuf* bufs[] = {&c->X, &c->y, &c->XT, &c->D, &c->Dlo, &c->Xc, &c->P, &c->err, &c->val,
                    &c->bestk, &c->cand, &c->sel, &c->ordering, &c->errors, &c->cand_err,
                    &c->kvals, &c->losses, &c->state, &c->scratch, &c->seeds,
                    &c->visited, &c->perm, &c->key, &c->M, &c->med, &c->skey,
                    &c->sidx, &c->sort_tmp, &c->rank, &c->tauw, &c->iperm, &c->frange,
                    &c->lidx, &c->lval, &c->llb, &c->lcount, &c->Mq, &c->long_d,
                    &c->long_i, &c->long_lab, &c->lbound};
  for (DevBuf* b : bufs) b->release();
  destroy_events(c);
  c->step_graphs.clear();
  c->batch.drop_graphs();
  c->batch.arena.release();
  c->batch.meta.release();
  hipStreamDestroy(c->stream);
  delete c;
}

int snn_select_device(snn_ctx* ctx, const float* X_device, int64_t n, int d,
                      const float* y_device, const int* k_vals, int num_k,
                      const int* loss_functions, int num_loss, const snn_params* params,
                      int* selected_dimensions, int* selected_ordering, float* selected_errors,
                      float* candidate_errors, snn_summary* summary) {
  if (!X_device || !y_device) return fail(SNN_EINVAL, "null input");
  return select_impl(ctx, X_device, n, d, y_device, k_vals, num_k, loss_functions, num_loss,
                     params, selected_dimensions, selected_ordering, selected_errors,
                     candidate_errors, summary);
}

int snn_select_device_noalloc(snn_ctx* ctx, const float* X_device, int64_t n, int d,
                              const float* y_device, const int* k_vals, int num_k,
                              const int* loss_functions, int num_loss, const snn_params* params,
                              int* selected_dimensions, int* selected_ordering,
                              float* selected_errors, float* candidate_errors,
                              snn_summary* summary) {
  if (!ctx) return fail(SNN_EINVAL, "null context");
  if (!X_device || !y_device) return fail(SNN_EINVAL, "null input");
  if (!ctx->fixed_scratch) return fail(SNN_EDEVICE, "no fixed work array on this device");
  return select_impl(ctx, X_device, n, d, y_device, k_vals, num_k, loss_functions, num_loss,
                     params, selected_dimensions, selected_ordering, selected_errors,
                     candidate_errors, summary, static_cast<char*>(ctx->fixed_scratch),
                     kFixedScratchBytes);
}

int snn_select(snn_ctx* ctx, const float* X, int64_t n, int d, const float* y,
               const int* k_vals, int num_k, const int* loss_functions, int num_loss,
               const snn_params* params, int* selected_dimensions, int* selected_ordering,
               float* selected_errors, float* candidate_errors, snn_summary* summary) {
  if (!ctx) return fail(SNN_EINVAL, "null context");
  if (!X || !y) return fail(SNN_EINVAL, "null input");
  if (n < 1 || d < 1) return fail(SNN_EINVAL, "empty input");
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(ctx->X.ensure(sizeof(float) * (size_t)n * d));
  HIPCHK(ctx->y.ensure(sizeof(float) * (size_t)n));
  HIPCHK(hipMemcpyAsync(ctx->X.p, X, sizeof(float) * (size_t)n * d, hipMemcpyHostToDevice,
                        ctx->stream));
  HIPCHK(hipMemcpyAsync(ctx->y.p, y, sizeof(float) * (size_t)n, hipMemcpyHostToDevice,
                        ctx->stream));
  return select_impl(ctx, ctx->X.as<float>(), n, d, ctx->y.as<float>(), k_vals, num_k,
                     loss_functions, num_loss, params, selected_dimensions, selected_ordering,
                     selected_errors, candidate_errors, summary);
}

int snn_distance_rows(snn_ctx* ctx, const float* X, int64_t n, int d, int method,
                      const int* order, int s, int form, const int64_t* rows, int nrows,
                      float* rows_out, float* lo_out) {
  if (!ctx || !X || (!rows && nrows > 0) || !rows_out) return fail(SNN_EINVAL, "null argument");
  if (n < 1 || d < 1 || s < 0 || s > d) return fail(SNN_EINVAL, "bad sizes");
  for (int t = 0; t < s; ++t)
    if (order[t] < 0 || order[t] >= d) return fail(SNN_EINVAL, "feature index out of range");
  for (int r = 0; r < nrows; ++r)
    if (rows[r] < 0 || rows[r] >= n) return fail(SNN_EINVAL, "row index out of range");
  if (form != SNN_FORM_PLAIN && form != SNN_FORM_PRECISION && form != SNN_FORM_MFMA)
    return fail(SNN_EINVAL, "unknown form");
  const bool comp = form == SNN_FORM_PRECISION;
  HIPCHK(hipSetDevice(ctx->device));
  const size_t nn = (size_t)n * (size_t)n;
  HIPCHK(ctx->X.ensure(sizeof(float) * (size_t)n * d));
  HIPCHK(ctx->XT.ensure(sizeof(float) * (size_t)n * d));
  HIPCHK(ctx->D.ensure(sizeof(float) * nn));
  if (comp) HIPCHK(ctx->Dlo.ensure(sizeof(float) * nn));
  if (form == SNN_FORM_MFMA) HIPCHK(ctx->scratch.ensure(snn::gram_mfma_scratch_bytes(n, d)));
  hipStream_t st = ctx->stream;
  float* D = ctx->D.as<float>();
  float* Dlo = comp ? ctx->Dlo.as<float>() : nullptr;
  HIPCHK(hipMemcpyAsync(ctx->X.p, X, sizeof(float) * (size_t)n * d, hipMemcpyHostToDevice, st));
  HIPCHK(snn::launch_transpose(ctx->X.as<float>(), ctx->XT.as<float>(), n, d, st));
  if (method == 1) {
    HIPCHK(snn::launch_fill(D, (int64_t)nn, 0.0f, st));
    if (Dlo) HIPCHK(snn::launch_fill(Dlo, (int64_t)nn, 0.0f, st));
  } else if (form == SNN_FORM_MFMA) {
    HIPCHK(snn::launch_gram_mfma(ctx->X.as<float>(), ctx->XT.as<float>(), n, d, D,
                                 ctx->scratch.p, st));
  } else {
    HIPCHK(snn::launch_gram_exact(ctx->XT.as<float>(), D, Dlo, nullptr, n, d, st));
  }
  for (int t = 0; t < s; ++t)
    HIPCHK(snn::launch_update(D, Dlo, ctx->XT.as<float>(), n, nullptr, order[t],
                              method == 1 ? 1 : -1, st));
  for (int r = 0; r < nrows; ++r) {
    HIPCHK(hipMemcpyAsync(rows_out + (size_t)r * n, D + (size_t)rows[r] * n, sizeof(float) * n,
                          hipMemcpyDeviceToHost, st));
    if (lo_out && Dlo)
      HIPCHK(hipMemcpyAsync(lo_out + (size_t)r * n, Dlo + (size_t)rows[r] * n, sizeof(float) * n,
                            hipMemcpyDeviceToHost, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  return SNN_OK;
}

int snn_distance_matrix(snn_ctx* ctx, const float* X_device, int64_t n, int d, int form,
                        float* D_device, double* ms) {
  if (!ctx || !X_device || !D_device) return fail(SNN_EINVAL, "null argument");
  if (n < 1 || d < 1) return fail(SNN_EINVAL, "bad sizes");
  if (form != SNN_FORM_PLAIN && form != SNN_FORM_MFMA)
    return fail(SNN_EINVAL, "form must be SNN_FORM_PLAIN or SNN_FORM_MFMA");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  HIPCHK(ctx->XT.ensure(sizeof(float) * (size_t)n * d));
  if (form == SNN_FORM_MFMA) HIPCHK(ctx->scratch.ensure(snn::gram_mfma_scratch_bytes(n, d)));
  Timing tm;
  hipEvent_t e0 = tm.make(), e1 = tm.make();
  HIPCHK(snn::launch_transpose(X_device, ctx->XT.as<float>(), n, d, st));
  HIPCHK(hipEventRecord(e0, st));
  if (form == SNN_FORM_MFMA)
    HIPCHK(snn::launch_gram_mfma(X_device, ctx->XT.as<float>(), n, d, D_device, ctx->scratch.p,
                                 st));
  else
    HIPCHK(snn::launch_gram_exact(ctx->XT.as<float>(), D_device, nullptr, nullptr, n, d, st));
  HIPCHK(hipEventRecord(e1, st));
  HIPCHK(hipStreamSynchronize(st));
  if (ms) *ms = elapsed(e0, e1);
  return SNN_OK;
}

int snn_loss(snn_ctx* ctx, const float* labels, const float* predictions, int64_t n,
             int loss_function, float* out) {
  if (!ctx || !labels || !predictions || !out) return fail(SNN_EINVAL, "null argument");
  if (n < 1) return fail(SNN_EINVAL, "empty input");
  int rc = validate_losses(&loss_function, 1);
  if (rc) return rc;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(ctx->scratch.ensure(sizeof(float) * (size_t)(4 * n + 64)));
  float* dy = ctx->scratch.as<float>();
  float* dp = dy + n;
  float* dout = dp + n;
  float* dmed = dout + 64;
  HIPCHK(hipMemcpyAsync(dy, labels, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(dp, predictions, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(snn::launch_loss_single(dy, dp, n, loss_function, dout, dmed, ctx->stream));
  HIPCHK(hipMemcpyAsync(out, dout, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return SNN_OK;
}

int train_backward_selection(float* Xtrain, int nXtrain, int dXtrain, float* Ytrain, int nYtrain,
                             int* kvals, int num_kvals, int num_features, int kfold,
                             int* loss_functions, int num_loss_functions, int auto_feature,
                             int model_type) {
  return train_common(0, Xtrain, nXtrain, dXtrain, Ytrain, nYtrain, kvals, num_kvals,
                      num_features, kfold, loss_functions, num_loss_functions, auto_feature,
                      model_type);
}

int train_forward_selection(float* Xtrain, int nXtrain, int dXtrain, float* Ytrain, int nYtrain,
                            int* kvals, int num_kvals, int num_features, int kfold,
                            int* loss_functions, int num_loss_functions, int auto_feature,
                            int model_type) {
  return train_common(1, Xtrain, nXtrain, dXtrain, Ytrain, nYtrain, kvals, num_kvals,
                      num_features, kfold, loss_functions, num_loss_functions, auto_feature,
                      model_type);
}

void get_selected_dimensions(int* out, int num) {
  for (int i = 0; i < num; ++i) out[i] = i < (int)g_sel.size() ? g_sel[i] : 0;
}

void get_selected_ordering(int* out, int num) {
  for (int i = 0; i < num; ++i) out[i] = i < (int)g_order.size() ? g_order[i] : 0;
}

void get_selected_errors(float* out, int num) {
  for (int i = 0; i < num; ++i) out[i] = i < (int)g_errors.size() ? g_errors[i] : 0.0f;
}

int get_final_optimal_model_k(void) { return g_best_k; }

float get_loss_all(float* labels, int num_labels, float* predictions, int num_predictions,
                   int loss_function) {
  std::lock_guard<std::mutex> lock(g_default_mu);
  if (num_labels != num_predictions) {
    fail(SNN_EINVAL, "labels and predictions differ in length");
    return std::nanf("");
  }
  snn_ctx* ctx = default_ctx();
  float out = std::nanf("");
  if (!ctx || snn_loss(ctx, labels, predictions, num_labels, loss_function, &out) != SNN_OK)
    return std::nanf("");
  return out;
}

}  // extern "C"
