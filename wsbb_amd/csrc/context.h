// Library-private state behind the wsbb_ctx handle.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/wsbb_amd.h"
#include "device.h"
#include "instance.h"

struct wsbb_ctx;

namespace wsbb {

void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);

// Grow-only device buffer owned by a context.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  hipError_t reserve(size_t need) {
    if (need <= bytes) return hipSuccess;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&ptr, need);
    if (e == hipSuccess) bytes = need;
    return e;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as(size_t byte_offset = 0) const {
    return reinterpret_cast<T*>(static_cast<char*>(ptr) + byte_offset);
  }
};

struct SearchState;    // search.hip
struct SelectorState;  // selector.hip

struct Launch {
  int groups = 1;
  int grid_cap = 0;  // 0: automatic
  int int16 = 1;     // two nodes per lane in int16 halves: 0 never, 1 for
                     // large batches, 2 whenever the instance fits
};

}  // namespace wsbb

struct wsbb_ctx {
  int device = 0;
  int num_cus = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  wsbb::Instance inst;
  wsbb::DevTables T{};
  wsbb::DevBuf tables;  // every instance table in one allocation
  wsbb::DevBuf batch;   // staging for host-buffer batch calls
  wsbb::Launch launch;
  int pool_capacity = 1 << 24;  // device node-pool capacity (nodes)
  int profile_every = 0;         // time every k-th round's bound launch in wsbb_solve
  bool search_graph = false;     // replay search rounds from a captured hipGraph
  int rounds_per_sync = 4;       // rounds queued between host status checks
  wsbb_trace_fn trace_fn = nullptr;  // per-round search trace (one round per sync)
  void* trace_user = nullptr;
  // autotuned shapes by batch size class (log2 of tiles), -1 = not tuned
  int tuned_groups[32];
  wsbb::SearchState* search = nullptr;
  wsbb::SelectorState* selector = nullptr;

  int grid_cap() const {
    if (launch.grid_cap > 0) return launch.grid_cap;
    return num_cus * 16;  // 16 blocks per CU: a 2^20-node batch, one per four tiles
  }
  int groups_for(int N) const;
  // two nodes per lane pays once a batch gives every SIMD two 128-node tiles;
  // smaller launches are latency-bound and prefer more wavefronts
  int int16_for(long long N) const {
    if (launch.int16 == 2) return 1;
    return launch.int16 == 1 && N >= 2LL * 128 * 4 * num_cus;
  }
};
