// Wave-level pair selection shared by the selector and the bound kernel.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.h"

namespace wsbb {

// Top-k of a wave's non-negative candidates without sorting.  Candidate
// (i, lane) has number 64 i + lane; the bit pattern of a non-negative float
// orders like its value, so the k-th largest value T is built bit by bit
// from ballot counts (31 rounds of compares and scalar popcounts, no lane
// exchanges), then sel[i] holds the candidates above T and, of those equal
// to T, the lowest-numbered ones until k are chosen.
template <int NI>
__device__ __forceinline__ void wave_topk(const float (&v)[NI], const bool (&ok)[NI], int k,
                                          unsigned long long (&sel)[NI]) {
  uint32_t u[NI];
#pragma unroll
  for (int i = 0; i < NI; ++i) u[i] = v[i] > 0.f ? __float_as_uint(v[i]) : 0u;
  uint32_t T = 0;
  for (int bit = 30; bit >= 0; --bit) {
    const uint32_t cand = T | (1u << bit);
    int cnt = 0;
#pragma unroll
    for (int i = 0; i < NI; ++i) cnt += __popcll(__ballot(ok[i] && u[i] >= cand));
    if (cnt >= k) T = cand;
  }
  int rem = k;
  unsigned long long eq[NI];
#pragma unroll
  for (int i = 0; i < NI; ++i) {
    sel[i] = __ballot(ok[i] && u[i] > T);
    rem -= __popcll(sel[i]);
    eq[i] = __ballot(ok[i] && u[i] == T);
  }
#pragma unroll
  for (int i = 0; i < NI; ++i)
    while (rem > 0 && eq[i]) {
      const unsigned long long low = eq[i] & (~eq[i] + 1ull);
      sel[i] |= low;
      eq[i] ^= low;
      --rem;
    }
}

// Pair subset of the 64-child tile [c0, c1) from parent-level predictions.
// Children of a parent are consecutive, so the tile's parents are the range
// from the first child's parent to the last child's (desc >> 9); each adds
// its BF16 probability row weighted by its number of children in the tile
// (eight rows in flight at once), then the top-k of the sums (ties: lower
// pair index).  sel[i] holds pairs 64 i .. 64 i + 63.  desc and p_off were
// written by earlier launches, so they are read through the scalar path.
template <int NI>
__device__ __forceinline__ void parent_tile_mask(const uint16_t* probs, int ldp, int P, int topk,
                                                 const int* p_off, const int* desc, int c0,
                                                 int c1, int lane,
                                                 unsigned long long (&sel)[NI]) {
  const int pa = ld_const(desc, c0) >> 9;
  const int pb = ld_const(desc, c1 - 1) >> 9;
  float agg[NI];
#pragma unroll
  for (int i = 0; i < NI; ++i) agg[i] = 0.f;
  for (int p0 = pa; p0 <= pb; p0 += 8) {
    float w[8], x[8][NI];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int p = p0 + j;
      const bool in = p <= pb;
      const int lo = in ? max(ld_const(p_off, p), c0) : 0;
      const int hi = in ? min(ld_const(p_off, p + 1), c1) : 0;
      w[j] = static_cast<float>(max(hi - lo, 0));
      const uint16_t* row = probs + static_cast<size_t>(in ? p : pa) * ldp;
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        const int s = 64 * i + lane;
        // clamped index: every load of the batch is issued before the sums
        const float v = __uint_as_float(static_cast<uint32_t>(row[min(s, ldp - 1)]) << 16);
        x[j][i] = s < ldp ? v : 0.f;
      }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int i = 0; i < NI; ++i) agg[i] += w[j] * x[j][i];
  }
  bool ok[NI];
#pragma unroll
  for (int i = 0; i < NI; ++i) ok[i] = 64 * i + lane < P;
  wave_topk<NI>(agg, ok, min(topk, P), sel);
}

}  // namespace wsbb
