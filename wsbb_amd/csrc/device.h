// Device-side data layouts shared by the bound, search and selector kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "instance.h"

namespace wsbb {

// Loads from tables that no kernel writes while they are read: going through
// the constant address space lets wave-uniform loads use the scalar unit.
__device__ __forceinline__ int ld_const(const int* p, size_t i) {
  return ((const __attribute__((address_space(4))) int*)(p))[i];  // address-space cast
}
__device__ __forceinline__ int2 ld_const2(const int2* p, size_t i) {
  const int* q = reinterpret_cast<const int*>(p) + 2 * i;
  return make_int2(ld_const(q, 0), ld_const(q, 1));
}
__device__ __forceinline__ int4 ld_const4(const int4* p, size_t i) {
  const int* q = reinterpret_cast<const int*>(p) + 4 * i;
  return make_int4(ld_const(q, 0), ld_const(q, 1), ld_const(q, 2), ld_const(q, 3));
}
__device__ __forceinline__ uint32_t ld_const_u32(const uint32_t* p, size_t i) {
  return static_cast<uint32_t>(ld_const(reinterpret_cast<const int*>(p), i));
}
__device__ __forceinline__ uint2 ld_const2u(const uint2* p, size_t i) {
  const int* q = reinterpret_cast<const int*>(p) + 2 * i;
  return make_uint2(static_cast<uint32_t>(ld_const(q, 0)), static_cast<uint32_t>(ld_const(q, 1)));
}

// Instance tables resident in HBM (built once per instance on the host).
struct DevTables {
  int n, m, P;
  int n4;                // Johnson steps per pair, padded to a multiple of 4
  const int32_t* p;      // [m][n]
  const int2* pairs;     // [P] (machine a, machine b)
  const int4* steps;     // [P][n4] {job, p_a, p_b, kBias + lag + p_b}; pad = 0,
                         // plus 32 zero steps after the last pair
  const uint2* wsel;     // [P][n4] masks selecting the job's 32-bit mask word:
                         // x = all ones iff bit 0 of (job >> 5), y = bit 1
  const int32_t* R0;     // [m]
  const int32_t* Q0;     // [m]
  const int32_t* ptot;   // [m]
  const int* steps32;    // [P][n4] int32 offset-form records of rs dwords:
                         // {32 (p_a - p_b), 32 (lag + p_b) + (job & 31)
                         //  [, word mask of bit 0 of job >> 5, of bit 1]};
                         // pad = 0, plus 32 zero steps after the last pair
  int rs;                // dwords per steps32 record: 2 (n <= 32), 3 (n <= 64), 4
  const uint2* steps16;  // [P][n4] two-node form {job, (p_a - p_b mod 2^16)
                         // | (2^14 + lag + p_b) << 16}; pad = 0, plus 16
                         // zero steps; null when bound values may reach 2^14
};

// What a bound launch computes and where it writes.
struct BoundArgs {
  int N;                 // nodes (launch size when Ndev is set)
  const int* Ndev;       // device-side node count (<= N), or null
  int ld;                // leading dimension of pairv/sel/feat (0: node count)
  int kind;              // 0 = LB2, 1 = LB1
  int best;              // LB2 returns the first pair value above best (-1: off)
  int groups;            // waves that split the pair range of one 64-node tile
  int32_t* out;          // [N] bound (leaves: makespan)
  int32_t* argmax;       // [N] first pair reaching the LB2 value, or null
  int32_t* pairv;        // [P][N] every pair value, or null
  const uint32_t* sel;   // [ceil(P/32)][N] pair subset per node, or null
  float* feat;           // [5m][N] selector features, or null
  float feat_scale;      // features are divided by this
  const uint32_t* tile_sel;  // [ntiles][nsw] learned pair subset per 64-node
                             // tile, or null
  int nsw;                   // words per tile mask
  int margin;                // with tile_sel: nodes whose subset bound is in
                             // [inc - margin, inc) get full LB2 (-1: never)
  const int* inc_dev;        // device incumbent for the margin test
  long long* stats;          // with tile_sel: [0] nodes, [1] full-LB2
                             // fallbacks, [2] node-pair evaluations
  int pk = 1;                // plain LB2 may run two nodes per lane (int16)
  // plain LB2 over pairs whose steps fit one scalar batch loads the next
  // pair's table lines ahead (set for launches of fewer waves than SIMDs,
  // where nothing else hides a scalar-cache miss)
  int touch = 0;
  // optional timestamps of the bound kernel's own dispatch (taken by the
  // launch itself, no separate marker packets in the stream)
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
};

// Processing times staged in LDS job-major, [n][MM] with machines m..MM-1
// zero: the head/tail recurrences read a job's machines as MM/4 16-byte
// loads from one address instead of MM loads from MM rows.
template <int MM>
__device__ __forceinline__ void stage_times(const DevTables& T, int32_t* sp) {
  static_assert(MM % 4 == 0, "machine rows are read four at a time");
  const int n = T.n, m = T.m;
  for (int i = threadIdx.x; i < MM * n; i += blockDim.x) {
    const int j = i / MM, k = i - j * MM;
    sp[i] = k < m ? T.p[k * n + j] : 0;
  }
}
template <int MM>
__device__ __forceinline__ void job_row(const int32_t* sp, int j, int (&pk)[MM]) {
  const int4* rp = reinterpret_cast<const int4*>(sp + j * MM);
#pragma unroll
  for (int k4 = 0; k4 < MM / 4; ++k4) {
    const int4 v = rp[k4];
    pk[4 * k4] = v.x;
    pk[4 * k4 + 1] = v.y;
    pk[4 * k4 + 2] = v.z;
    pk[4 * k4 + 3] = v.w;
  }
}

template <int NW>
__device__ __forceinline__ void set_job_bit(uint32_t (&sm)[NW], int j) {
#pragma unroll
  for (int w = 0; w < NW; ++w) sm[w] |= ((j >> 5) == w) ? (1u << (j & 31)) : 0u;
}

// Head/tail recurrences by one job (rows m..MM-1 of the times are zero, so
// the padded machines need no guards): heads r of a prefix extended by job j
// at its end, tails q of a suffix extended by job j at its front.
template <int MM>
__device__ __forceinline__ void head_step(int (&r)[MM], const int (&pk)[MM]) {
  r[0] += pk[0];
#pragma unroll
  for (int k = 1; k < MM; ++k) r[k] = max(r[k], r[k - 1]) + pk[k];
}
template <int MM>
__device__ __forceinline__ void tail_step(int (&q)[MM], const int (&pk)[MM]) {
#pragma unroll
  for (int k = MM - 1; k >= 0; --k) q[k] = max(q[k], (k + 1 < MM) ? q[k + 1] : 0) + pk[k];
}

__device__ __forceinline__ int perm_at(const uint8_t* col, int stride, int sw0,
                                       int sw1, int i) {
  const int src = (i == sw0) ? sw1 : ((i == sw1) ? sw0 : i);
  return col[static_cast<size_t>(src) * stride];
}

// Node source: explicit position-major nodes.
struct BatchSource {
  static constexpr bool kState = false;  // no parent state: heads/tails from the permutation
  static constexpr bool kSiblings = false;  // consecutive nodes need not share free jobs
  static constexpr bool kPrefetch = false;
  const uint8_t* permT;  // [n][N]
  const int32_t* l1;
  const int32_t* l2;
  int N;
  __device__ __forceinline__ bool load(int k, int& a, int& b, const uint8_t*& col,
                                       int& stride, int& sw0, int& sw1) const {
    if (k >= N) return false;
    a = l1[k];
    b = l2[k];
    col = permT + k;
    stride = N;
    sw0 = sw1 = -1;
    return true;
  }
  __device__ __forceinline__ int slot(int k) const { return k; }
};

// Node source: children described by (parent, side, position) over a
// position-major parent array (decompose_beginend child semantics).
struct ChildSource {
  static constexpr bool kState = true;
  // consecutive children are a few sibling parents' children: a tile's free
  // jobs are a small union (the two-node form scans only those)
  static constexpr bool kSiblings = true;
  // a child's descriptor can be read before anything else is known (any k
  // below the capacity N is inside the array), so the bound kernel issues
  // its first tile's descriptors with its staging loads
  static constexpr bool kPrefetch = true;
  const uint8_t* ppermT;  // parents [n][PN]
  const int32_t* pl1;
  const int32_t* pl2;
  const int32_t* desc;    // [N] parent << 9 | side << 8 | j
  int PN, N;
  // optional parent state (heads [MM][PN], tails [MM][PN] of the parent's
  // prefix and suffix, scheduled-job masks [NW][PN]): a child's heads and
  // tails are then its parent's extended by the one or two jobs it places
  const int32_t* hr = nullptr;
  const int32_t* hq = nullptr;
  const uint32_t* hm = nullptr;
  const int32_t* hw = nullptr;  // [MM][PN] work of the parent's scheduled jobs
  __device__ __forceinline__ bool has_state() const { return hr != nullptr; }
  // heads, tails, scheduled mask and scheduled work of child k from its
  // parent's state; sp holds the processing times job-major in LDS
  __device__ __forceinline__ int key(int k) const { return desc[k]; }
  template <int MM, int NW>
  __device__ __forceinline__ void state(int k, const int32_t* sp, int (&r)[MM], int (&q)[MM],
                                        uint32_t (&sm)[NW], int (&sc)[MM]) const {
    state_key<MM, NW>(desc[k], sp, r, q, sm, sc);
  }
  // the same from the child's descriptor d
  template <int MM, int NW>
  __device__ __forceinline__ void state_key(int d, const int32_t* sp, int (&r)[MM], int (&q)[MM],
                                            uint32_t (&sm)[NW], int (&sc)[MM]) const {
    const int par = d >> 9, side = (d >> 8) & 1, j = d & 255;
    const int pa = pl1[par], pb = pl2[par];
#pragma unroll
    for (int kk = 0; kk < MM; ++kk) {
      r[kk] = hr[static_cast<size_t>(kk) * PN + par];
      q[kk] = hq[static_cast<size_t>(kk) * PN + par];
      sc[kk] = hw[static_cast<size_t>(kk) * PN + par];
    }
#pragma unroll
    for (int w = 0; w < NW; ++w) sm[w] = hm[static_cast<size_t>(w) * PN + par];
    const uint8_t* col = ppermT + par;
    const int job = col[static_cast<size_t>(j) * PN];  // placed at l1 + 1 or l2 - 1
    set_job_bit<NW>(sm, job);
    int pk[MM];
    job_row<MM>(sp, job, pk);
    if (side == 0)
      head_step<MM>(r, pk);
    else
      tail_step<MM>(q, pk);
#pragma unroll
    for (int kk = 0; kk < MM; ++kk) sc[kk] += pk[kk];
    // the last free job goes to the end of the prefix (limites_set)
    const int a = pa + (side == 0), b = pb - (side == 1);
    if (b - a == 2) {
      const int sw0 = side == 0 ? pa + 1 : pb - 1;
      const int j2 = perm_at(col, PN, sw0, j, a + 1);
      set_job_bit<NW>(sm, j2);
      job_row<MM>(sp, j2, pk);
      head_step<MM>(r, pk);
#pragma unroll
      for (int kk = 0; kk < MM; ++kk) sc[kk] += pk[kk];
    }
  }
  __device__ __forceinline__ bool load(int k, int& a, int& b, const uint8_t*& col,
                                       int& stride, int& sw0, int& sw1) const {
    if (k >= N) return false;
    load_key(desc[k], a, b, col, stride, sw0, sw1);
    return true;
  }
  // limits and permutation view of the child with descriptor d
  __device__ __forceinline__ void load_key(int d, int& a, int& b, const uint8_t*& col,
                                           int& stride, int& sw0, int& sw1) const {
    const int par = d >> 9, side = (d >> 8) & 1, j = d & 255;
    const int pa = pl1[par], pb = pl2[par];
    // limites_set / permutation_set (bb_problem_calculator.cpp:40-48, 66-74)
    a = pa + (side == 0);
    b = pb - (side == 1);
    if (b - a == 2) a += 1;
    sw0 = side == 0 ? pa + 1 : pb - 1;
    sw1 = j;
    col = ppermT + par;
    stride = PN;
  }
  __device__ __forceinline__ int slot(int k) const { return k; }
};

// Node source: the children listed in list[0..*count) (a compacted subset of
// a round's children); results go to the listed child's slot.
struct ListSource {
  static constexpr bool kState = true;
  static constexpr bool kSiblings = false;
  static constexpr bool kPrefetch = false;
  ChildSource base;
  const int32_t* list;
  const int* count;
  __device__ __forceinline__ bool has_state() const { return base.has_state(); }
  template <int MM, int NW>
  __device__ __forceinline__ void state(int k, const int32_t* sp, int (&r)[MM], int (&q)[MM],
                                        uint32_t (&sm)[NW], int (&sc)[MM]) const {
    base.state<MM, NW>(list[k], sp, r, q, sm, sc);
  }
  __device__ __forceinline__ bool load(int k, int& a, int& b, const uint8_t*& col,
                                       int& stride, int& sw0, int& sw1) const {
    if (k >= *count) return false;
    return base.load(list[k], a, b, col, stride, sw0, sw1);
  }
  __device__ __forceinline__ int slot(int k) const { return list[k]; }
};


}  // namespace wsbb
