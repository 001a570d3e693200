// Batched selection driver: many independent selection problems advance one
// step at a time together.  Per step, each problem's candidate gather, and
// after the search its losses, arg-min and distance update, are small kernels
// on the context's stream; the candidate search of all problems of one
// direction is ONE persistent launch over a shared work queue (search.hip).
// The whole step can be captured in a hipGraph and replayed: every kernel
// reads the device-resident state of its problem, so a step's launches do not
// depend on the host.  Graphs are kept while the batch keeps its shapes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/speedynn.h"
#include "common.h"
#include "engine.h"
#include "kernels.h"

using snn::DevBuf;
using snn::fail;
using snn::Problem;
using snn::SelState;

namespace {

// Byte offsets of one problem's buffers in the batch arena.
struct Slot {
  int index;  // position in the caller's array
  Problem pb;
  int64_t work;
  size_t X, y, XT, D, Dlo, M, Xc, P, err, val, bestk, cand, sel, ordering, errors, kvals, losses,
      state, perm, key, seeds, med, skey, sidx, sort_tmp;
  bool has_med;      // median losses: sort space for NaN-holding error vectors
  int loss_paths;    // which loss kernels the problem's losses need
  bool sorted;       // forward, precision form: sorted columns for the first step
  size_t sort_bytes;
  bool comp;         // keep D's compensation term
  bool fused;        // precision form: the update writes the next step's seeds
  bool on_device;    // the caller's X and y are device pointers
  const float* dX;   // X and y on the device (the caller's, or copies in the arena)
  const float* dy;
};

size_t take(size_t& off, size_t bytes) {
  const size_t o = off;
  off += (bytes + 255) / 256 * 256;
  return o;
}

template <class T>
T* at(const DevBuf& b, size_t off) {
  return reinterpret_cast<T*>(static_cast<char*>(b.p) + off);
}

// Problems of one direction and one list size: a contiguous range of the
// device problem array (largest work first, so the queue hands out the long
// items early) served by one persistent launch.
struct Group {
  int first = 0, count = 0;
  int mult = 1;
  int K = 0;
  int grid_dense = 0, grid_pruned = 0;
  size_t dense_start = 0, pruned_start = 0;  // prefix arrays in meta
};

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

struct Plan {
  snn_ctx* ctx;
  std::vector<Slot> slots;  // device order: groups contiguous
  std::vector<Group> groups;
  bool pruned;
  bool dense_first;  // some forward problem holds non-finite X: dense first step for it
  size_t probs_off, counter_off, visited_off;
  std::vector<snn::SearchProblem> hp;  // host copies of the device descriptors

  const DevBuf& A() const { return ctx->batch.arena; }
  const DevBuf& Mt() const { return ctx->batch.meta; }

  // Launches of one selection step (step > 0 differs only for a forward
  // selection in the precision form: its first step, D = 0, runs over sorted
  // columns -- one launch per problem, a short step -- later ones run the
  // pruned search).
  // ev[0]: gather done, ev[1]: bound seeds done, ev[2]: search done (each may
  // be null).
  int issue_step(bool first_step, hipStream_t s, const hipEvent_t* ev) const {
    for (const Slot& sl : slots) {
      const Problem& pb = sl.pb;
      HIPCHK(snn::launch_compact(sl.dX, at<int>(A(), sl.cand),
                                 at<SelState>(A(), sl.state), at<float>(A(), sl.Xc), pb.n, pb.d,
                                 pb.cpad, s));
    }
    if (ev) HIPCHK(hipEventRecord(ev[0], s));
    auto* probs = at<snn::SearchProblem>(Mt(), probs_off);
    auto* counter = at<unsigned long long>(Mt(), counter_off);
    auto pruned_group = [&](const Group& g) { return pruned && !(g.mult > 0 && first_step); };
    // seeds: the previous step's fused update wrote them, except before a
    // backward problem's first step and for problems without the fused update
    for (const Group& g : groups) {
      if (!pruned_group(g)) continue;
      for (int i = g.first; i < g.first + g.count; ++i) {
        const Slot& sl = slots[i];
        if (sl.fused && !(first_step && g.mult < 0)) continue;
        HIPCHK(snn::launch_seed_scan(at<float>(A(), sl.D), at<SelState>(A(), sl.state),
                                     at<int>(A(), sl.seeds), sl.pb.n, sl.pb.npf, s));
      }
    }
    if (ev) HIPCHK(hipEventRecord(ev[1], s));
    for (const Group& g : groups) {
      if (!g.count) continue;
      if (pruned_group(g)) {
        HIPCHK(snn::launch_batch_search(probs + g.first, at<int64_t>(Mt(), g.pruned_start),
                                        g.count, counter, 2, g.mult, g.K, g.grid_pruned, s));
      } else if (pruned && g.mult > 0) {
        for (int i = g.first; i < g.first + g.count; ++i)
          HIPCHK(snn::launch_search_sorted(hp[i], slots[i].pb.d, s));
        if (dense_first)
          HIPCHK(snn::launch_batch_search(probs + g.first, at<int64_t>(Mt(), g.dense_start),
                                          g.count, counter, 1, g.mult, g.K, g.grid_dense, s));
      } else {
        HIPCHK(snn::launch_batch_search(probs + g.first, at<int64_t>(Mt(), g.dense_start),
                                        g.count, counter, 0, g.mult, g.K, g.grid_dense, s));
      }
    }
    if (ev) HIPCHK(hipEventRecord(ev[2], s));
    for (const Slot& sl : slots) {
      const Problem& pb = sl.pb;
      SelState* st = at<SelState>(A(), sl.state);
      HIPCHK(snn::launch_loss(at<float>(A(), sl.P), sl.dy, at<int>(A(), sl.losses),
                              st, at<float>(A(), sl.err), pb.n, pb.cpad, pb.nk, pb.nl, pb.d,
                              sl.has_med ? at<float>(A(), sl.med) : nullptr, sl.loss_paths, s));
      HIPCHK(snn::launch_select(st, pb, at<int>(A(), sl.kvals), at<float>(A(), sl.err),
                                at<int>(A(), sl.cand), at<int>(A(), sl.sel),
                                at<int>(A(), sl.ordering), at<float>(A(), sl.errors),
                                at<float>(A(), sl.val), at<int>(A(), sl.bestk), nullptr,
                                at<int>(A(), sl.perm), s));
      const int mult = pb.method == 1 ? 1 : -1;
      float* Dlo = sl.comp ? at<float>(A(), sl.Dlo) : nullptr;
      if (sl.fused)
        HIPCHK(snn::launch_update_seeds(at<float>(A(), sl.D), Dlo, at<float>(A(), sl.XT), pb.n,
                                        st, mult, at<int>(A(), sl.seeds), pb.npf, snn::RowList{},
                                        nullptr, nullptr, pb.d, nullptr, nullptr, nullptr,
                                        nullptr, s));
      else
        HIPCHK(snn::launch_update(at<float>(A(), sl.D), Dlo, at<float>(A(), sl.XT), pb.n, st, -1,
                                  mult, s));
    }
    return SNN_OK;
  }

  int capture(bool first_step, hipStream_t s, hipGraphExec_t* out) const {
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = issue_step(first_step, s, nullptr);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    HIPCHK(e);
    const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    HIPCHK(ei);
    return SNN_OK;
  }
};

}  // namespace

extern "C" int snn_select_batch(snn_ctx* ctx, int num_problems, snn_problem* probs,
                                snn_summary* summary) {
  if (!ctx) return fail(SNN_EINVAL, "null context");
  if (num_problems < 1 || !probs) return fail(SNN_EINVAL, "no problems");
  if (summary) std::memset(summary, 0, sizeof(*summary));

  // ---- validate; invalid problems are reported and left out
  std::vector<Slot> valid;
  int form = -1, use_graph = 0;
  std::string first_err;
  for (int p = 0; p < num_problems; ++p) {
    snn_problem& q = probs[p];
    q.status = SNN_OK;
    q.num_final = q.optimal_k = q.steps = 0;
    Problem pb{};
    int rc = (q.X && q.y) ? snn::validate_problem(q.n, q.d, q.k_vals, q.num_k, q.loss_functions,
                                                  q.num_loss, &q.params, &pb)
                          : fail(SNN_EINVAL, "null input");
    if (!rc && snn::search_list_size(pb.kmax) < 0)
      rc = fail(SNN_ENOTSUP, "batched selections take k values up to 64 (snn_select takes more)");
    if (!rc && form >= 0 &&
        (q.params.form != form || (q.params.use_graph != 0) != (use_graph != 0)))
      rc = fail(SNN_EINVAL, "form and use_graph must agree across the batch");
    if (rc) {
      q.status = rc;
      if (first_err.empty()) first_err = "problem " + std::to_string(p) + ": " + snn_last_error();
      continue;
    }
    if (form < 0) {
      form = q.params.form;
      use_graph = q.params.use_graph != 0;
    }
    Slot s{};
    s.index = p;
    s.pb = pb;
    s.work = pb.n * pb.n * (int64_t)pb.cpad;
    valid.push_back(s);
  }
  if (valid.empty()) return fail(SNN_EINVAL, first_err);

  Plan plan{};
  plan.ctx = ctx;
  plan.pruned = form == SNN_FORM_PRECISION;
  const bool pruned = plan.pruned;
  HIPCHK(hipSetDevice(ctx->device));

  // ---- device order: groups of one direction and list size, each by work descending
  for (int dir = 1; dir >= 0; --dir) {
    for (int K = 1; K <= 64; ++K) {
      if (snn::search_list_size(K) != K) continue;  // compiled list sizes only
      Group g;
      g.first = (int)plan.slots.size();
      g.mult = dir == 1 ? 1 : -1;
      g.K = K;
      std::vector<Slot> part;
      for (const Slot& s : valid)
        if (s.pb.method == dir && snn::search_list_size(s.pb.kmax) == K) part.push_back(s);
      if (part.empty()) continue;
      std::stable_sort(part.begin(), part.end(),
                       [](const Slot& a, const Slot& b) { return a.work > b.work; });
      for (const Slot& s : part) plan.slots.push_back(s);
      g.count = (int)part.size();
      g.grid_dense = snn::batch_grid(pruned && dir == 1 ? 1 : 0, g.mult, K);
      g.grid_pruned = snn::batch_grid(2, g.mult, K);
      plan.groups.push_back(g);
    }
  }

  // ---- arena layout
  size_t off = 0;
  for (Slot& s : plan.slots) {
    const Problem& pb = s.pb;
    const size_t n = (size_t)pb.n, d = (size_t)pb.d, cpad = (size_t)pb.cpad;
    s.on_device = probs[s.index].inputs_on_device != 0;
    s.X = s.on_device ? 0 : take(off, sizeof(float) * n * d);
    s.y = s.on_device ? 0 : take(off, sizeof(float) * n);
    s.XT = take(off, sizeof(float) * n * d);
    s.D = take(off, sizeof(float) * n * n);
    s.comp = pruned && probs[s.index].params.compensated;
    s.Dlo = s.comp ? take(off, sizeof(float) * n * n) : 0;
    s.M = (pruned && pb.method == 0) ? take(off, sizeof(float) * n * n) : 0;
    s.Xc = take(off, sizeof(float) * n * cpad);
    s.P = take(off, sizeof(float) * pb.nk * cpad * n);
    s.err = take(off, sizeof(float) * pb.nl * pb.nk * cpad);
    s.val = take(off, sizeof(float) * pb.nl * cpad);
    s.bestk = take(off, sizeof(int) * cpad);
    s.cand = take(off, sizeof(int) * d);
    s.sel = take(off, sizeof(int) * d);
    s.ordering = take(off, sizeof(int) * d);
    s.errors = take(off, sizeof(float) * pb.nl * d);
    s.kvals = take(off, sizeof(int) * pb.nk);
    s.losses = take(off, sizeof(int) * pb.nl);
    s.state = take(off, sizeof(SelState));
    s.perm = take(off, sizeof(int) * d);
    s.key = take(off, sizeof(double) * d);
    s.seeds = pruned ? take(off, sizeof(int) * n * 128) : 0;
    s.fused = pruned;
    const snn_problem& q = probs[s.index];
    s.has_med = snn::needs_median_scratch(q.loss_functions, q.num_loss);
    s.loss_paths = snn::loss_paths(q.loss_functions, q.num_loss);
    s.med = s.has_med ? take(off, sizeof(float) * 2 * n * cpad * pb.nk) : 0;
    s.sorted = pruned && pb.method == 1;
    s.sort_bytes = s.sorted ? snn::sort_columns_temp_bytes(pb.n, pb.d) : 0;
    if (s.sorted && !s.sort_bytes) return fail(SNN_EDEVICE, "segmented sort: no temporary size");
    s.skey = s.sorted ? take(off, sizeof(float) * n * d) : 0;
    s.sidx = s.sorted ? take(off, sizeof(int) * n * d) : 0;
    s.sort_tmp = s.sorted ? take(off, s.sort_bytes) : 0;
  }
  const size_t arena_bytes = off;
  const int B = (int)plan.slots.size();
  size_t moff = 0;
  plan.probs_off = take(moff, sizeof(snn::SearchProblem) * B);
  for (Group& g : plan.groups) {
    g.dense_start = take(moff, sizeof(int64_t) * (g.count + 1));
    g.pruned_start = take(moff, sizeof(int64_t) * (g.count + 1));
  }
  plan.counter_off = take(moff, sizeof(unsigned long long));
  plan.visited_off = take(moff, sizeof(unsigned long long) * snn::kVisitSlots * snn::kVisitStride);

  HIPCHK(hipSetDevice(ctx->device));
  snn::BatchCache& cache = ctx->batch;
  const void* old_arena = cache.arena.p;
  const void* old_meta = cache.meta.p;
  HIPCHK(cache.arena.ensure(arena_bytes));
  HIPCHK(cache.meta.ensure(moff));
  hipStream_t s = ctx->stream;
  for (Slot& sl : plan.slots) {
    const snn_problem& q = probs[sl.index];
    sl.dX = sl.on_device ? q.X : at<float>(cache.arena, sl.X);
    sl.dy = sl.on_device ? q.y : at<float>(cache.arena, sl.y);
  }

  // ---- host copies of the search descriptors and work queues
  std::vector<snn::SearchProblem>& hp = plan.hp;
  hp.assign(B, snn::SearchProblem{});
  std::vector<char> hmeta(moff, 0);
  for (int i = 0; i < B; ++i) {
    const Slot& sl = plan.slots[i];
    const Problem& pb = sl.pb;
    snn::SearchProblem& a = hp[i];
    a.sorted = sl.sorted ? 1 : 0;
    a.skey = sl.sorted ? at<float>(cache.arena, sl.skey) : nullptr;
    a.sidx = sl.sorted ? at<int>(cache.arena, sl.sidx) : nullptr;
    a.cand = at<int>(cache.arena, sl.cand);
    a.D = at<float>(cache.arena, sl.D);
    a.M = (pruned && pb.method == 0) ? at<float>(cache.arena, sl.M) : nullptr;
    a.Xc = at<float>(cache.arena, sl.Xc);
    a.y = sl.dy;
    a.kvals = at<int>(cache.arena, sl.kvals);
    a.st = at<SelState>(cache.arena, sl.state);
    a.seeds = pruned ? at<int>(cache.arena, sl.seeds) : nullptr;
    a.P = at<float>(cache.arena, sl.P);
    a.visited = pruned ? at<unsigned long long>(cache.meta, plan.visited_off) : nullptr;
    a.n = pb.n;
    a.npf = pb.npf;
    a.cpad = pb.cpad;
    a.nk = pb.nk;
    a.kmax = pb.kmax;
    a.mult = pb.method == 1 ? 1 : -1;
    a.model_type = pb.model_type;
    a.zero = 0;
    a.nseed = snn::backward_seed_count(pb.kmax);
  }
  std::memcpy(hmeta.data() + plan.probs_off, hp.data(), sizeof(snn::SearchProblem) * B);
  for (const Group& g : plan.groups) {
    int64_t* ds = reinterpret_cast<int64_t*>(hmeta.data() + g.dense_start);
    int64_t* ps = reinterpret_cast<int64_t*>(hmeta.data() + g.pruned_start);
    ds[0] = ps[0] = 0;
    for (int i = 0; i < g.count; ++i) {
      ds[i + 1] = ds[i] + snn::batch_dense_items(hp[g.first + i]);
      ps[i + 1] = ps[i] + snn::batch_pruned_items(hp[g.first + i]);
    }
  }
  HIPCHK(hipMemcpyAsync(cache.meta.p, hmeta.data(), moff, hipMemcpyHostToDevice, s));

  // ---- inputs and per-problem start state
  Timing tm;
  hipEvent_t e0 = tm.make(), e1 = tm.make(), e2 = tm.make();
  for (const Slot& sl : plan.slots) {
    const snn_problem& q = probs[sl.index];
    const Problem& pb = sl.pb;
    if (!sl.on_device) {
      HIPCHK(hipMemcpyAsync(at<float>(cache.arena, sl.X), q.X, sizeof(float) * pb.n * pb.d,
                            hipMemcpyHostToDevice, s));
      HIPCHK(hipMemcpyAsync(at<float>(cache.arena, sl.y), q.y, sizeof(float) * pb.n,
                            hipMemcpyHostToDevice, s));
    }
    HIPCHK(hipMemcpyAsync(at<int>(cache.arena, sl.kvals), q.k_vals, sizeof(int) * pb.nk,
                          hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(at<int>(cache.arena, sl.losses), q.loss_functions, sizeof(int) * pb.nl,
                          hipMemcpyHostToDevice, s));
  }
  HIPCHK(hipEventRecord(e0, s));
  if (pruned)
    HIPCHK(snn::launch_fill_u64(at<unsigned long long>(cache.meta, plan.visited_off),
                                snn::kVisitSlots * snn::kVisitStride, s));
  for (const Slot& sl : plan.slots) {
    const Problem& pb = sl.pb;
    const int64_t nn = pb.n * pb.n;
    float* XT = at<float>(cache.arena, sl.XT);
    float* D = at<float>(cache.arena, sl.D);
    float* Dlo = sl.comp ? at<float>(cache.arena, sl.Dlo) : nullptr;
    HIPCHK(snn::launch_transpose(sl.dX, XT, pb.n, pb.d, s));
    HIPCHK(snn::launch_feature_order(XT, pb.n, pb.d, at<double>(cache.arena, sl.key),
                                     at<int>(cache.arena, sl.perm), nullptr, pruned, s));
    HIPCHK(snn::launch_init_state(at<SelState>(cache.arena, sl.state), pb,
                                  at<int>(cache.arena, sl.cand), at<int>(cache.arena, sl.sel),
                                  at<int>(cache.arena, sl.ordering),
                                  at<float>(cache.arena, sl.errors), at<int>(cache.arena, sl.perm),
                                  s));
    if (sl.sorted)
      HIPCHK(snn::launch_sort_columns(XT, pb.n, pb.d, at<SelState>(cache.arena, sl.state),
                                      at<float>(cache.arena, sl.skey),
                                      at<int>(cache.arena, sl.sidx),
                                      at<void>(cache.arena, sl.sort_tmp), sl.sort_bytes, s));
    if (pb.method == 1) {
      HIPCHK(snn::launch_fill(D, nn, 0.0f, s));
      if (Dlo) HIPCHK(snn::launch_fill(Dlo, nn, 0.0f, s));
    } else {
      HIPCHK(snn::launch_gram_exact(XT, D, Dlo, pruned ? at<float>(cache.arena, sl.M) : nullptr,
                                    pb.n, pb.d, s));
    }
  }
  HIPCHK(hipEventRecord(e1, s));
  // does any sorted problem hold non-finite X (then the dense first step runs
  // for it; the sorted and dense kernels each check the problem's flag)?
  plan.dense_first = false;
  for (const Slot& sl : plan.slots) {
    if (!sl.sorted) continue;
    SelState hs{};
    HIPCHK(hipMemcpyAsync(&hs, at<SelState>(cache.arena, sl.state), sizeof(hs),
                          hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    plan.dense_first |= hs.nonfinite != 0;
  }

  // ---- steps
  int max_steps = 0;
  for (const Slot& sl : plan.slots) max_steps = std::max(max_steps, sl.pb.max_steps);
  double ms_search = 0.0, ms_search_first = 0.0, ms_bound = 0.0;  // without graphs only
  if (use_graph) {
    std::vector<size_t> key{(size_t)cache.arena.p, (size_t)cache.meta.p, (size_t)form,
                            (size_t)B, (size_t)plan.dense_first};
    // everything a captured launch embeds: shapes, problem order, device pointers
    for (const Slot& sl : plan.slots) {
      const Problem& pb = sl.pb;
      const int64_t f[] = {sl.index, pb.n, pb.npf, pb.d, pb.nk, pb.nl, pb.kmax, pb.method,
                           pb.model_type, pb.auto_feature, pb.max_steps, pb.num_desired,
                           sl.comp, sl.has_med, sl.loss_paths, (int64_t)(size_t)sl.dX,
                           (int64_t)(size_t)sl.dy};
      for (int64_t v : f) key.push_back((size_t)v);
    }
    if (key != cache.key || old_arena != cache.arena.p || old_meta != cache.meta.p ||
        !cache.first || !cache.rest) {
      cache.drop_graphs();
      cache.key.clear();
      int rc = plan.capture(true, s, &cache.first);
      if (!rc) rc = plan.capture(false, s, &cache.rest);
      if (rc) {
        cache.drop_graphs();
        return rc;
      }
      cache.key = key;
    }
    for (int t = 0; t < max_steps; ++t) HIPCHK(hipGraphLaunch(t == 0 ? cache.first : cache.rest, s));
  } else {
    // per step: gather done, seeds done, search done (phase split of the summary)
    std::vector<hipEvent_t> marks;
    for (int t = 0; t < max_steps; ++t) {
      const hipEvent_t ev[3] = {tm.make(), tm.make(), tm.make()};
      int rc = plan.issue_step(t == 0, s, ev);
      if (rc) return rc;
      marks.insert(marks.end(), ev, ev + 3);
    }
    HIPCHK(hipEventRecord(e2, s));
    HIPCHK(hipStreamSynchronize(s));
    for (size_t i = 0; i + 2 < marks.size(); i += 3) {
      ms_bound += elapsed(marks[i], marks[i + 1]);
      const double sm = elapsed(marks[i + 1], marks[i + 2]);
      ms_search += sm;
      if (i == 0) ms_search_first = sm;
    }
  }
  if (use_graph) {
    HIPCHK(hipEventRecord(e2, s));
    HIPCHK(hipStreamSynchronize(s));
  }

  // ---- results
  int64_t evals = 0, pruned_full = 0;
  for (const Slot& sl : plan.slots) {
    snn_problem& q = probs[sl.index];
    const Problem& pb = sl.pb;
    SelState hs{};
    HIPCHK(hipMemcpy(&hs, at<SelState>(cache.arena, sl.state), sizeof(hs), hipMemcpyDeviceToHost));
    if (q.selected_dimensions)
      HIPCHK(hipMemcpy(q.selected_dimensions, at<int>(cache.arena, sl.sel), sizeof(int) * pb.d,
                       hipMemcpyDeviceToHost));
    if (q.selected_ordering)
      HIPCHK(hipMemcpy(q.selected_ordering, at<int>(cache.arena, sl.ordering), sizeof(int) * pb.d,
                       hipMemcpyDeviceToHost));
    if (q.selected_errors)
      HIPCHK(hipMemcpy(q.selected_errors, at<float>(cache.arena, sl.errors),
                       sizeof(float) * pb.nl * pb.d, hipMemcpyDeviceToHost));
    q.num_final = hs.num_final;
    q.optimal_k = hs.best_k;
    q.steps = hs.step;
    if (hs.error) {
      q.status = SNN_ESTATE;
      if (first_err.empty())
        first_err = "problem " + std::to_string(sl.index) +
                    ": no candidate had a finite validation error";
    }
    for (int t = 0; t < hs.step; ++t) {
      evals += (int64_t)(pb.d - t) * pb.n;
      if (pruned && (t > 0 || pb.method == 0))
        pruned_full += (int64_t)(pb.d - t) * pb.n * (pb.n - pb.npf);
    }
  }
  if (summary) {
    summary->steps = max_steps;
    summary->ms_total = elapsed(e0, e2);
    summary->ms_distance = elapsed(e0, e1);
    summary->ms_search = ms_search;
    summary->ms_search_first = ms_search_first;
    summary->ms_bound = ms_bound;
    summary->evals = evals;
    if (pruned) {
      unsigned long long done = 0;
      std::vector<unsigned long long> slots((size_t)snn::kVisitSlots * snn::kVisitStride);
      HIPCHK(hipMemcpy(slots.data(), at<unsigned long long>(cache.meta, plan.visited_off),
                       sizeof(unsigned long long) * slots.size(), hipMemcpyDeviceToHost));
      for (int i = 0; i < snn::kVisitSlots; ++i) done += slots[(size_t)i * snn::kVisitStride];
      summary->pruned_point_evals = pruned_full - (int64_t)done;
    }
  }
  if (!first_err.empty()) {
    bool invalid = false;
    for (int p = 0; p < num_problems; ++p) invalid |= probs[p].status != SNN_OK && probs[p].status != SNN_ESTATE;
    return fail(invalid ? SNN_EINVAL : SNN_ESTATE, first_err);
  }
  return SNN_OK;
}
