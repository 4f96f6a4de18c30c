// Launchers for the node bound kernels (bound.hip).
#pragma once
#include "device.h"

namespace wsbb {

// groups (A.groups) must be 1, 2 or 4; grid_cap bounds the number of
// 256-thread blocks (each loops over its share of 64-node tiles).
hipError_t launch_bound(const DevTables& T, const BatchSource& src,
                        const BoundArgs& A, int grid_cap, hipStream_t stream);
hipError_t launch_bound(const DevTables& T, const ChildSource& src,
                        const BoundArgs& A, int grid_cap, hipStream_t stream);

}  // namespace wsbb
