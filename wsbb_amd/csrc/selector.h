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
// With probs set, only the per-node BF16 softmax probabilities are written
// (node-major rows of selector_ldp(ctx) entries, padding pairs 0).
hipError_t launch_mlp_select(wsbb_ctx* ctx, const uint2* xfrag, int N, const int* Ndev,
                             uint32_t* tile_sel, float* scores, int ld,
                             const MlpTiles* tiles = nullptr, uint16_t* probs = nullptr);
// Pair subsets of 64-child tiles from the probabilities of their parents:
// tile t covers children [64t, 64t + 64) of the Nc live ones (children of a
// parent are consecutive, parent p owns [p_off[p], p_off[p+1]), child c's
// parent is desc[c] >> 9); the mask holds the top-k pairs of the sum over the
// tile's parents of (children in the tile) x (parent's probabilities).
hipError_t launch_parent_tiles(wsbb_ctx* ctx, const uint16_t* probs, const int* p_off,
                               const int* desc, const int* Nc_dev, int Nc_max,
                               uint32_t* tile_sel);
int selector_ldp(const wsbb_ctx* ctx);
bool selector_parents(const wsbb_ctx* ctx);
int selector_fsteps(const wsbb_ctx* ctx);
int selector_topk(const wsbb_ctx* ctx);
int selector_margin(const wsbb_ctx* ctx);
bool selector_ready(const wsbb_ctx* ctx);
int selector_nsw(const wsbb_ctx* ctx);
int selector_log_stride(const wsbb_ctx* ctx);
void selector_log_buffers(const wsbb_ctx* ctx, float** feat, int** label, int** count, int* cap);

}  // namespace wsbb
