// Launchers for the node bound kernels (bound.hip).
#pragma once
#include "device.h"

namespace wsbb {

// groups (A.groups): 1, 2, 4, 8 or 16 waves per 64-node tile (8 and 16 only
// for plain LB2 with up to 16 machines, else 4); grid_cap bounds the number
// of blocks (each loops over its share of tiles).
hipError_t launch_bound(const DevTables& T, const BatchSource& src,
                        const BoundArgs& A, int grid_cap, hipStream_t stream);
hipError_t launch_bound(const DevTables& T, const ChildSource& src,
                        const BoundArgs& A, int grid_cap, hipStream_t stream);
// plain LB2 of listed children (kind 0, no optional outputs)
hipError_t launch_bound(const DevTables& T, const ListSource& src,
                        const BoundArgs& A, int grid_cap, hipStream_t stream);

// Selector features as BF16 MFMA B-operand fragments (k_features in
// bound.hip): frag[(b * fsteps + t) * 64 + lane] for 32-node block b and
// feature step t (8 features), fsteps = ceil(5m / 8); optionally also in
// FP32, feature f of node k at f32[f * ld + k].
struct FeatArgs {
  int N;
  const int* Ndev;  // device node count (<= N), or null
  float scale;      // features are divided by this
  int fsteps;
  uint2* frag;      // [ceil(N/32)][fsteps][64], or null
  float* f32 = nullptr;  // [5m][ld] features in FP32, or null
  int ld = 0;
};
hipError_t launch_features(const DevTables& T, const BatchSource& src, const FeatArgs& A,
                           int grid_cap, hipStream_t stream);
hipError_t launch_features(const DevTables& T, const ChildSource& src, const FeatArgs& A,
                           int grid_cap, hipStream_t stream);

}  // namespace wsbb
