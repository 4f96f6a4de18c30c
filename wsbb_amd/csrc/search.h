// Device-resident branch-and-bound driver (search.hip).
#pragma once
#include <cstdint>

struct wsbb_ctx;

namespace wsbb {

// A saved node pool (position-major permutations [n][count]) with the
// incumbent and counters it was saved with.
struct ResumeState {
  const uint8_t* permT;
  const int32_t* l1;
  const int32_t* l2;
  const int32_t* cost;
  int count;
  const int64_t* stats;      // WSBB_NSTATS counters to continue from, or null
  const int32_t* best_perm;  // permutation of the incumbent, or null
};

// Runs the search to completion (or budget) with the device node pool;
// the context's rounds_per_sync rounds are queued between host checks of the
// pool state.  With `resume` the pool, incumbent (ub0) and counters are
// loaded from it instead of starting from the root.
int solve(wsbb_ctx* ctx, int kind, int pool_size, int max_batches, int32_t ub0,
          int64_t budget, int32_t* best, int32_t* best_perm, int64_t* stats,
          const ResumeState* resume = nullptr);

}  // namespace wsbb
