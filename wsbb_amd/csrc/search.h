// Device-resident branch-and-bound driver (search.hip).
#pragma once
#include <cstdint>

struct wsbb_ctx;

namespace wsbb {

// Runs the search to completion (or budget) with the device node pool;
// rounds_per_sync rounds are queued between host checks of the pool state.
int solve(wsbb_ctx* ctx, int kind, int pool_size, int max_batches, int32_t ub0,
          int64_t budget, int32_t* best, int32_t* best_perm, int64_t* stats,
          int cap_nodes, int rounds_per_sync);

}  // namespace wsbb
