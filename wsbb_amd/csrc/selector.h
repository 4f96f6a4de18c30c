// Learned pair selector hooks used by the search (selector.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

struct wsbb_ctx;

namespace wsbb {

// MLP scores of N nodes (features as BF16 B fragments
// [2 ceil(N/64)][ceil(5m/8)][64], see FeatArgs) -> per-64-node-tile pair
// masks [ntiles][ceil(P/32)] holding the tile's top-k pairs; optional logits
// [P][ld].
// Tile subset of a mixed batch (several instances): score only the listed
// tiles, whose valid node counts are cnt[tile]; feature blocks are fstride
// steps apart and tile masks tstride words apart (words past the instance's
// pairs are zeroed).
struct MlpTiles {
  const int* list;
  int nlist;
  const int* cnt;
  int fstride, tstride;
};
hipError_t launch_mlp_select(wsbb_ctx* ctx, const uint2* xfrag, int N, const int* Ndev,
                             uint32_t* tile_sel, float* scores, int ld,
                             const MlpTiles* tiles = nullptr);
int selector_fsteps(const wsbb_ctx* ctx);
int selector_topk(const wsbb_ctx* ctx);
int selector_margin(const wsbb_ctx* ctx);
bool selector_ready(const wsbb_ctx* ctx);
int selector_nsw(const wsbb_ctx* ctx);
int selector_log_stride(const wsbb_ctx* ctx);
void selector_log_buffers(const wsbb_ctx* ctx, float** feat, int** label, int** count, int* cap);

}  // namespace wsbb
