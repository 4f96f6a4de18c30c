// Learned pair selector hooks used by the search (selector.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

struct wsbb_ctx;

namespace wsbb {

// MLP scores of N nodes (features [5m][ld]) -> per-64-node-tile pair masks
// [ntiles][ceil(P/32)] holding the tile's top-k pairs; optional logits.
hipError_t launch_mlp_select(wsbb_ctx* ctx, const float* feat, int ld, int N, const int* Ndev,
                             uint32_t* tile_sel, float* scores);
int selector_topk(const wsbb_ctx* ctx);
int selector_margin(const wsbb_ctx* ctx);
bool selector_ready(const wsbb_ctx* ctx);
int selector_nsw(const wsbb_ctx* ctx);
int selector_log_stride(const wsbb_ctx* ctx);
void selector_log_buffers(const wsbb_ctx* ctx, float** feat, int** label, int** count, int* cap);

}  // namespace wsbb
