// Internal definitions shared by the host-side drivers (engine.hip: one
// selection; batch.hip: many selections per launch).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/speedynn.h"
#include "common.h"

namespace snn {

// Records the message of the calling thread's last error; returns code.
int fail(int code, const std::string& msg);

#define HIPCHK(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return ::snn::fail(e_ == hipErrorOutOfMemory ? SNN_ENOMEM : SNN_EDEVICE,              \
                         std::string(#expr) + " failed: " + hipGetErrorString(e_));         \
  } while (0)

// Grow-only device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t want) {
    if (want <= bytes && p) return hipSuccess;
    if (p) {
      hipFree(p);
      p = nullptr;
      bytes = 0;
    }
    if (want == 0) want = 256;
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) bytes = want;
    return e;
  }
  void release() {
    if (p) hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

// Checks one problem's arguments the way snn_select does and fills pb.
int validate_problem(int64_t n, int d, const int* k_vals, int num_k, const int* loss_functions,
                     int num_loss, const snn_params* prm, Problem* pb);

// Cached batch launch state (graphs are rebuilt when the batch changes).
struct BatchCache {
  DevBuf arena, meta;
  std::vector<size_t> key;  // shapes and arena addresses the graphs were built for
  hipGraphExec_t first = nullptr, rest = nullptr;
  void drop_graphs() {
    if (first) hipGraphExecDestroy(first);
    if (rest) hipGraphExecDestroy(rest);
    first = rest = nullptr;
  }
};

}  // namespace snn

// Phase-timing events of one context, created once and reused by every
// call: a ring of kRingSteps steps x 5 marks (a step's slots are reused
// kRingSteps steps later, after its times have been read), the tuner's
// trial marks and the call's start/matrix/end marks.
struct EventPool {
  static constexpr int kRingSteps = 64;
  static constexpr int kMarks = 5;
  std::vector<hipEvent_t> ring;  // kRingSteps * kMarks
  hipEvent_t tune[7] = {};
  hipEvent_t span[3] = {};
  bool ready = false;
};

// Step graphs of single selections: the launches of one step, captured once
// per distinct launch sequence (key: every value the launches embed) and
// replayed; `used` is the pruned-kernel bit mask of the captured step.
struct StepGraphs {
  struct Entry {
    hipGraphExec_t exec = nullptr;
    int used = 0;
  };
  std::map<std::vector<size_t>, Entry> map;
  void clear() {
    for (auto& kv : map)
      if (kv.second.exec) hipGraphExecDestroy(kv.second.exec);
    map.clear();
  }
};

struct snn_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  snn::DevBuf X, y, XT, D, Dlo, Xc, P, err, val, bestk, cand, sel, ordering, errors, cand_err,
      kvals, losses, state, scratch, seeds, visited, perm, key, M, med, skey, sidx, sort_tmp,
      rank, tauw, iperm, frange, lidx, lval, llb, lcount, Mq, long_d, long_i, long_lab, lbound;
  snn::BatchCache batch;
  EventPool events;
  StepGraphs step_graphs;
  void* fixed_scratch = nullptr;  // device address of the fixed work array on this device
};
