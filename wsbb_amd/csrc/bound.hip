// Node bound kernels for the permutation flowshop on gfx950.
//
// One lane bounds one node.  All lanes of a wavefront walk the same machine
// pair and the same Johnson step at the same time, so every per-step constant
// (job, p_a, p_b, lag) is wave-uniform and arrives through scalar loads into
// SGPRs; per lane a step is one bit extract from the node's unscheduled-job
// mask, three 24-bit multiply-adds and a max.  The two-node int16 form
// (bound_kernel_pk) streams its step table eight steps per scalar load,
// with the next eight in flight while the current ones compute.  Heads,
// tails and the job mask live in VGPRs (read with wave-uniform indices); the
// processing-time matrix and each lane's permutation are staged in LDS for
// the per-lane gathers of the head/tail recurrences.  `groups` waves may
// split the pair range of one 64-node tile (pairs-per-wavefront), combining
// through LDS.
//
// Semantics follow bound_flowshop::bornes_calculer (src/bound_flowshop.cpp
// 929-951): a leaf (limite2 - limite1 == 1) yields its makespan
// (criteres_calculer 897-909); otherwise LB2 = max(0, max_s v_s) over the
// pairs in order (borneInfMakespan 121-168), with v_s built from the head
// start (initCmax 79-90), the Johnson scan over unscheduled jobs
// (heuristiqueCmax 64-77) and the tail combine (cmaxFin 92-100, 147-148).
#include "bound.h"

#include <hip/hip_ext.h>

#include <type_traits>

namespace wsbb {

namespace {


// float -> bfloat16, round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf_dev(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}

__device__ __forceinline__ uint32_t bitsel(uint32_t mask, uint32_t a, uint32_t b) {
  return (mask & a) | (~mask & b);  // v_bfi_b32
}

// Johnson steps in the int32 offset form.  With S the running sum of p_b
// over the unscheduled jobs passed, the scan keeps D = T0 - S and
// U = t1 - S, both scaled by 32; an unscheduled job's step t0 += p_a;
// t1 = max(t1, t0 + lag) + p_b (heuristiqueCmax, bound_flowshop.cpp:64-77)
// becomes
//   D += 32 (p_a - p_b);  U = max(U, D + 32 (lag + p_b))
// and a scheduled job (flag f = 0) leaves D and gives max(U, D) = U, since
// T0 <= t1 throughout.  The second constant also carries the job's bit
// offset in its five low bits (a bit-field extract reads only those), so a
// step is one extract (plus a word select for n > 32), two 24-bit
// multiply-adds and a max; the offset adds less than 32 to a candidate,
// which U >> 5 drops.  At the end S = rem(b): t1 = (U >> 5) + rem(b) and
// T0 = (D >> 5) + rem(b).
template <int NW>
struct Rec32 {
  // dwords per step: the two constants, then one word mask per selection
  // level (every dword of a batch is used, so it loads as wide scalar loads)
  static constexpr int RS = NW == 1 ? 2 : (NW == 2 ? 3 : 4);
  static constexpr int GW = 4 * RS;  // dwords per four-step group
  // groups per scalar-load batch (one wait): up to 48 SGPRs of records for
  // n <= 32, 64 above (measured best on 20- and 50-job instances)
  static constexpr int CH = (NW == 1 ? 48 : 64) / GW;
};

template <int NW, int CH>
struct Chunk32 {
  int w[CH * Rec32<NW>::GW];
  // every dword of the batch is made live here, so the batch loads as wide
  // scalar loads with one wait instead of being split behind the group
  // branches that follow
  __device__ __forceinline__ void load(const int* p) {
#pragma unroll
    for (int i = 0; i < CH * Rec32<NW>::GW; ++i) w[i] = ld_const(p, i);
#pragma unroll
    for (int i = 0; i < CH * Rec32<NW>::GW; ++i) asm volatile("" ::"s"(w[i]));
  }
};

template <int NW>
__device__ __forceinline__ void step32(const int* r, const uint32_t (&um)[NW], int& D, int& U) {
  uint32_t word;
  if constexpr (NW == 1) {
    word = um[0];
  } else if constexpr (NW == 2) {
    word = bitsel(static_cast<uint32_t>(r[2]), um[1], um[0]);
  } else {
    const uint32_t w01 = bitsel(static_cast<uint32_t>(r[2]), um[1], um[0]);
    const uint32_t w23 = bitsel(static_cast<uint32_t>(r[2]), um[3], um[2]);
    word = bitsel(static_cast<uint32_t>(r[3]), w23, w01);
  }
  const uint32_t f = __builtin_amdgcn_ubfe(word, static_cast<uint32_t>(r[1]), 1u);
  D = __mul24(static_cast<int>(f), r[0]) + D;
  U = max(U, static_cast<int>(__umul24(f, static_cast<uint32_t>(r[1])) + static_cast<uint32_t>(D)));
}

template <int NW>
__device__ __forceinline__ void group32(const int* g, const uint32_t (&um)[NW], int& D, int& U) {
#pragma unroll
  for (int k = 0; k < 4; ++k) step32<NW>(g + k * Rec32<NW>::RS, um, D, U);
}

// One pair's scan over its gpp groups, CH groups per scalar-load batch
// (a batch may read past the pair's end: the table holds 32 zero steps
// after the last pair).
template <int NW>
__device__ __forceinline__ void scan_pair32(const DevTables& T, int s, const uint32_t (&um)[NW],
                                            int& D, int& U) {
  constexpr int GW = Rec32<NW>::GW, CH = Rec32<NW>::CH;
  const int gpp = T.n4 >> 2;
  const int* gp = T.steps32 + static_cast<size_t>(s) * gpp * GW;
  for (int k = 0; k < gpp; k += CH) {
    Chunk32<NW, CH> c;
    c.load(gp + static_cast<size_t>(k) * GW);
#pragma unroll
    for (int h = 0; h < CH; ++h)
      if (k + h < gpp) group32<NW>(c.w + h * GW, um, D, U);
  }
}

// The same scan for a pair whose steps fit one batch (gpp <= CH), which also
// issues scalar loads of the next pair's table lines (`s_next` >= 0) once
// its own batch is in: the lines are in the scalar cache when the next
// pair's batch is loaded, instead of each pair waiting a full miss.  At one
// or two waves per SIMD (a small search round) that miss is not hidden by
// other waves.  The loaded words are only made live after the next batch's
// own wait (the scalar loads complete out of order, so that wait covers them
// anyway) and never used otherwise.
template <int NW>
__device__ __forceinline__ void scan_pair32_next(const DevTables& T, int s, int s_next,
                                                 const uint32_t (&um)[NW], int& D, int& U,
                                                 int (&t)[4]) {
  constexpr int GW = Rec32<NW>::GW, CH = Rec32<NW>::CH;
  const int gpp = T.n4 >> 2;
  Chunk32<NW, CH> c;
  c.load(T.steps32 + static_cast<size_t>(s) * gpp * GW);
  asm volatile("" ::"s"(t[0]), "s"(t[1]), "s"(t[2]), "s"(t[3]));
  if (s_next >= 0) {
    // every 64-byte line the next batch reads (it starts anywhere in a line
    // and spans CH * GW dwords; the table has read-ahead padding after the
    // last pair)
    const int* nx = T.steps32 + static_cast<size_t>(s_next) * gpp * GW;
    t[0] = ld_const(nx, 0);
    t[1] = ld_const(nx, 16);
    t[2] = ld_const(nx, 32);
    t[3] = ld_const(nx, CH * GW - 1);
  }
#pragma unroll
  for (int h = 0; h < CH; ++h)
    if (h < gpp) group32<NW>(c.w + h * GW, um, D, U);
}

// MODE 0: plain LB2 (optionally with the `best` early exit) and nothing
// else, so none of the optional outputs holds scalar registers in the pair
// loop; MODE 1: LB2 with the optional outputs, pair masks and tile subsets;
// MODE 2: MODE 1 plus LB1 and the selector features; MODE 3: MODE 0 over the
// pairs of each tile's subset (tile_sel) only.  Modes 0 and 3 write their
// result to the source's slot of the node.
template <int MM, int NW, int MODE, int BLOCK, class Src>
__global__ __launch_bounds__(BLOCK) void bound_kernel(DevTables T, Src src, BoundArgs A) {
  constexpr bool LEAN = MODE == 0 || MODE == 3, EXTRA = MODE == 2;
  extern __shared__ int32_t lds[];
  const int n = T.n, m = T.m, P = T.P;
  // rows m..MM-1 of the staged times are zero: the recurrences then run
  // over all MM machines without per-machine guards, and the padded heads
  // and tails never exceed the real ones
  const int mn4 = MM * (n + 3);
  int32_t* sp = lds;  // [n][MM] times, then rows n, n+1, n+2: R0, Q0, row totals
  int4* part = reinterpret_cast<int4*>(lds + mn4);   // [blockDim] partials
  // [n][64 * tiles per block] bytes: the permutation columns of the block's
  // tiles, staged once per tile so the head/tail recurrences read LDS
  // instead of waiting on one global load per position
  uint8_t* pl = reinterpret_cast<uint8_t*>(part + blockDim.x);
  const int G = A.groups;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int tpb = (blockDim.x >> 6) / G;  // 64-node tiles per block
  const int tib = wave / G;
  // the live count and (children of a round) the first tile's descriptors
  // load with the staging below instead of after its barrier: two memory
  // round trips fewer in front of a search round's bound
  const int N = (A.Ndev != nullptr) ? *A.Ndev : A.N;
  int key0 = 0;
  if constexpr (Src::kPrefetch) {
    const int node0 = (blockIdx.x * tpb + tib) * 64 + lane;
    if (node0 < A.N) key0 = src.key(node0);
  }
  stage_times<MM>(T, sp);
  for (int i = threadIdx.x; i < 3 * MM; i += blockDim.x) {
    const int c = i / MM, k = i - c * MM;
    const int32_t* row = c == 0 ? T.R0 : (c == 1 ? T.Q0 : T.ptot);  // zero past machine m
    sp[(n + c) * MM + k] = row[k];
  }
  __syncthreads();

  // wave-uniform by construction; readfirstlane lets the compiler keep the
  // pair loop, its table loads and the register-array indices scalar
  const int grp = __builtin_amdgcn_readfirstlane(wave - tib * G);
  const int s_lo = (A.kind == 0) ? (grp * P) / G : 0;
  const int s_hi = (A.kind == 0) ? ((grp + 1) * P) / G : 0;
  const size_t ld = A.ld > 0 ? static_cast<size_t>(A.ld) : static_cast<size_t>(N);
  const int ntiles = (N + 63) >> 6;
  const int nbt = (ntiles + tpb - 1) / tpb;
  int inc = 0;
  if constexpr (!LEAN) inc = (A.inc_dev != nullptr) ? *A.inc_dev : 0;

  uint32_t valid_bits[NW];
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    const int hi = n - 32 * w;
    valid_bits[w] = hi >= 32 ? 0xffffffffu : (hi > 0 ? ((1u << hi) - 1u) : 0u);
  }

  for (int bt = blockIdx.x; bt < nbt; bt += gridDim.x) {
    const int tile = __builtin_amdgcn_readfirstlane(bt * tpb + tib);
    const int node = tile * 64 + lane;
    int l1 = -1, l2 = n, stride = 0, sw0 = -1, sw1 = -1;
    const uint8_t* col = nullptr;
    // the live count (device-side in a search round) bounds the last tile;
    // a source's own limit is its capacity
    bool valid;
    int key = 0;  // descriptor of the child whose state this lane builds
    if constexpr (Src::kPrefetch) {
      valid = node < N;
      if (bt == static_cast<int>(blockIdx.x)) key = key0;
      if (bt != static_cast<int>(blockIdx.x) || !valid) key = src.key(valid ? node : N - 1);
      if (valid) src.load_key(key, l1, l2, col, stride, sw0, sw1);
    } else {
      valid = node < N && src.load(node, l1, l2, col, stride, sw0, sw1);
    }
    if (!valid) {
      l1 = -1;
      l2 = n;
    }
    const int d1 = l1 + 1, d2 = n - l2;
    // children of a search round carry their parent's heads, tails and job
    // mask: a child extends them by the one or two jobs it places, and no
    // permutation is staged.  Otherwise the tile's first wave stages the
    // permutations (all loads of a chunk are issued before the first LDS
    // store waits); the other waves of the tile read them after the barrier.
    bool from_parent = false;
    if constexpr (Src::kState && !EXTRA) from_parent = src.has_state();
    const uint8_t* mine = pl + tib * 64 + lane;
    const int bstride = 64 * tpb;
    // (sources with a parent state stage only when it is absent: smaller
    // chunks keep that rarely used path from setting the kernel's register
    // count, which decides how many waves fit per SIMD)
    constexpr int kChunk = Src::kState ? 4 : 16;
    if (!from_parent && grp == 0 && valid) {
      for (int i0 = 0; i0 < n; i0 += kChunk) {
        uint32_t v[kChunk];
#pragma unroll
        for (int u = 0; u < kChunk; ++u)
          if (i0 + u < n) v[u] = col[static_cast<size_t>(i0 + u) * stride];
#pragma unroll
        for (int u = 0; u < kChunk; ++u)
          if (i0 + u < n) pl[(i0 + u) * bstride + tib * 64 + lane] = static_cast<uint8_t>(v[u]);
      }
    }
    if (G > 1) __syncthreads();

    // heads of the prefix (set_tempsMachines_retardDebut 841-856) and tails
    // of the suffix.  The reference runs one forward pass per start machine
    // (set_tempsMachinesFin_tempsJobFin 876-895); the reversed recurrence
    // below yields the same longest-path values for all machines at once.
    // q[k] = tail of machine k; q[k] for k >= m stays 0, which is the
    // boundary of the reversed recurrence.
    // r: heads of the prefix; ew: remaining work of the free jobs (row total
    // minus the scheduled jobs) + tail of the suffix (Q0 without suffix),
    // what a pair ending on machine k adds to t1 - S.  The tails are folded
    // into ew as soon as the suffix is done, so at most three per-machine
    // arrays are live at once.  qx keeps the tails for the features.
    int r[MM], ew[MM];
    int qx[EXTRA ? MM : 1];
    uint32_t sm[NW];
    if (from_parent) {
      // every lane reads a child's state (padding lanes the tile's last
      // child, their results unused), so the arrays are assigned on this
      // path without a zero fill held in registers beside them
      if constexpr (Src::kState) {
        int q[MM], sc[MM];
        if constexpr (Src::kPrefetch)
          src.template state_key<MM, NW>(key, sp, r, q, sm, sc);
        else
          src.template state<MM, NW>(min(node, N - 1), sp, r, q, sm, sc);
        int tot[MM], q0[MM];
        job_row<MM>(sp, n + 2, tot);
        job_row<MM>(sp, n + 1, q0);
#pragma unroll
        for (int k = 0; k < MM; ++k) {
          ew[k] = tot[k] - sc[k] + (d2 ? q[k] : q0[k]);
          if constexpr (EXTRA) qx[k] = q[k];
        }
      }
    } else {
      // a zero the compiler cannot hoist: hoisted, the fill below became a
      // block of registers held across the tile loop
      int zero;
      asm volatile("v_mov_b32 %0, 0" : "=v"(zero));
      job_row<MM>(sp, n + 2, ew);  // row totals
#pragma unroll
      for (int k = 0; k < MM; ++k) r[k] = zero;
#pragma unroll
      for (int w = 0; w < NW; ++w) sm[w] = 0;
      {
        int q[MM];
#pragma unroll
        for (int k = 0; k < MM; ++k) q[k] = zero;
        // the next job index is read while the current row is processed
        int jn = d2 > 0 ? perm_at(mine, bstride, sw0, sw1, n - 1) : 0;
        for (int i = n - 1; i >= l2; --i) {
          const int j = jn;
          if (i - 1 >= l2) jn = perm_at(mine, bstride, sw0, sw1, i - 1);
          set_job_bit<NW>(sm, j);
          int pk[MM];
          job_row<MM>(sp, j, pk);
          tail_step<MM>(q, pk);
#pragma unroll
          for (int k = 0; k < MM; ++k) ew[k] -= pk[k];
        }
        int q0[MM];
        job_row<MM>(sp, n + 1, q0);
#pragma unroll
        for (int k = 0; k < MM; ++k) {
          ew[k] += d2 ? q[k] : q0[k];
          if constexpr (EXTRA) qx[k] = q[k];
        }
      }
      int jn = d1 > 0 ? perm_at(mine, bstride, sw0, sw1, 0) : 0;
      for (int i = 0; i < d1; ++i) {
        const int j = jn;
        if (i + 1 < d1) jn = perm_at(mine, bstride, sw0, sw1, i + 1);
        set_job_bit<NW>(sm, j);
        int pk[MM];
        job_row<MM>(sp, j, pk);
        head_step<MM>(r, pk);
#pragma unroll
        for (int k = 0; k < MM; ++k) ew[k] -= pk[k];
      }
    }
    // a leaf has no free job (ew = tail): its makespan max_k r[k] + q[k]
    // (criteres_calculer 897-909), or r[m-1] with an empty suffix
    const bool leaf = (l2 - l1 == 1);
    int leafv = 0;
#pragma unroll
    for (int k = 0; k < MM; ++k)
      leafv = max(leafv, r[k] + (d2 ? ew[k] : 0));

    // st32 = 32 x the pair start clock (heads, or R0 without prefix: initCmax
    // 79-90).  Tables read zero past machine m, like the padded heads.
    //
    // Tail combine (cmaxFin 92-100, borneInfMakespan 146-148): without
    // suffix v = t1 + Q0[b] = (U >> 5) + ew[b]; with one v = max(t1 + q[b],
    // T0 + q[a]), and after the scan T0 = start(a) + rem(a) whatever the
    // pair's order, so T0 + q[a] = (st32[a] >> 5) + ew[a], known when the
    // pair starts.
    int st32[MM];
    {
      int r0[MM];
      job_row<MM>(sp, n, r0);
#pragma unroll
      for (int k = 0; k < MM; ++k) st32[k] = (d1 ? r[k] : r0[k]) * (1 << kScaleShift);
    }
    // a pair's one-machine term (machine a), or 0 without suffix
    auto one_of = [&](int a) { return d2 ? (st32[a] >> kScaleShift) + ew[a] : 0; };
    auto pair_value = [&](int one, int b, int U) { return max((U >> kScaleShift) + ew[b], one); };

    // one-machine bound and selector features: head (or R0) + remaining
    // work + tail (or Q0) per machine
    int lb1 = 0;
    if constexpr (EXTRA) {
#pragma unroll
      for (int k = 0; k < MM; ++k)
        if (k < m) {
          const int head = st32[k] >> kScaleShift;
          const int tail = d2 ? qx[k] : sp[(n + 1) * MM + k];
          const int rem = ew[k] - tail;
          lb1 = max(lb1, head + ew[k]);
          if (A.feat != nullptr && valid) {
            const float fs = 1.0f / A.feat_scale;
            A.feat[(0 * m + k) * ld + node] = head * fs;
            A.feat[(1 * m + k) * ld + node] = tail * fs;
            A.feat[(2 * m + k) * ld + node] = rem * fs;
            A.feat[(3 * m + k) * ld + node] = (head + rem) * fs;
            A.feat[(4 * m + k) * ld + node] = (rem + tail) * fs;
          }
        }
    }

    uint32_t um[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) um[w] = ~sm[w] & valid_bits[w];

    int lb = 0, fx = -1, am = -1;
    const bool use_best = A.best != -1;
    const uint32_t* tsel = nullptr;
    if constexpr (!LEAN || MODE == 3)
      if (A.tile_sel != nullptr) tsel = A.tile_sel + static_cast<size_t>(tile) * A.nsw;

    bool take = true;
    long long fallbacks = 0;
    int npairs = 0;  // pairs scanned by this wave
    if constexpr (MODE == 0) {
      // every pair of [s_lo, s_hi), in lexicographic order (remplirMachine
      // 500-508): the next pair's machines follow without a load
      npairs = s_hi - s_lo;
      if (s_lo < s_hi) {
        const int2 ab0 = ld_const2(T.pairs, s_lo);
        int a = ab0.x, b = ab0.y;
        if (A.touch && (T.n4 >> 2) <= Rec32<NW>::CH) {
          int t[4] = {0, 0, 0, 0};
          for (int s = s_lo; s < s_hi; ++s) {
            int D = st32[a], U = st32[b];
            const int one = one_of(a);
            scan_pair32_next<NW>(T, s, s + 1 < s_hi ? s + 1 : -1, um, D, U, t);
            lb = max(lb, pair_value(one, b, U));
            if (++b == m) {
              ++a;
              b = a + 1;
            }
          }
        } else {
          for (int s = s_lo; s < s_hi; ++s) {
            int D = st32[a], U = st32[b];
            const int one = one_of(a);
            scan_pair32<NW>(T, s, um, D, U);
            lb = max(lb, pair_value(one, b, U));
            if (++b == m) {
              ++a;
              b = a + 1;
            }
          }
        }
      }
    } else if constexpr (MODE == 3) {
      // the set bits of the tile's mask words in order (the words are
      // wave-uniform scalars)
      npairs = s_hi - s_lo;
      uint32_t word = 0;
      int wi = -1;
      for (;;) {
        while (word == 0) {
          if (++wi >= A.nsw) break;
          word = ld_const_u32(tsel, wi);
        }
        if (word == 0) break;
        const int s = 32 * wi + __builtin_ctz(word);
        word &= word - 1;
        if (s < s_lo) continue;  // another group's share of the pairs
        if (s >= s_hi) break;
        const int2 ab = ld_const2(T.pairs, s);
        int D = st32[ab.x], U = st32[ab.y];
        const int one = one_of(ab.x);
        scan_pair32<NW>(T, s, um, D, U);
        lb = max(lb, pair_value(one, ab.y, U));
      }
    } else {
    // Pass 0 scans all pairs of [s_lo, s_hi), or the tile's learned subset;
    // pass 1 (learned subset only) scans the complement for the nodes whose
    // partial bound lands within `margin` below the incumbent (full LB2).
    const int npass = (tsel != nullptr && A.margin >= 0) ? 2 : 1;
    for (int ps = 0; ps < npass; ++ps) {
      const uint32_t want = ps == 0 ? 1u : 0u;
      if (ps == 1) {
        take = valid && !leaf && lb < inc && lb >= inc - A.margin;
        fallbacks = __popcll(__ballot(take));
        if (fallbacks == 0) break;
      }
      for (int s = s_lo; s < s_hi; ++s) {
        if (tsel != nullptr) {
          if (((ld_const_u32(tsel, s >> 5) >> (s & 31)) & 1u) != want) continue;
        } else if (A.sel != nullptr) {
          take = valid && ((A.sel[static_cast<size_t>(s >> 5) * ld + node] >> (s & 31)) & 1u);
          if (!__any(take)) continue;
        }
        ++npairs;
        const int2 ab = ld_const2(T.pairs, s);
        int D = st32[ab.x], U = st32[ab.y];
        const int one = one_of(ab.x);
        scan_pair32<NW>(T, s, um, D, U);
        const int v = pair_value(one, ab.y, U);
        if (A.pairv != nullptr && valid) A.pairv[static_cast<size_t>(s) * ld + node] = v;
        if (take) {
          if (v > lb) {
            lb = v;
            am = s;
          }
          if (use_best && fx < 0 && lb > A.best) fx = lb;
        }
      }
    }

    }

    if (G > 1) {
      part[threadIdx.x] = make_int4(lb, fx, am, 0);
      __syncthreads();
      if (grp == 0) {
        int blb = 0, bfx = -1, bam = -1;
        for (int g = 0; g < G; ++g) {
          const int4 e = part[(tib * G + g) * 64 + lane];
          if (bfx < 0 && e.y >= 0) bfx = e.y;
          if (e.x > blb) {
            blb = e.x;
            bam = e.z;
          }
        }
        lb = blb;
        fx = bfx;
        am = bam;
      }
      __syncthreads();
    }

    if (grp == 0 && valid) {
      int res;
      if (leaf) res = leafv;
      else if (A.kind == 1) res = lb1;
      else res = fx >= 0 ? fx : lb;
      if constexpr (LEAN) {
        A.out[src.slot(node)] = res;
      } else {
        A.out[node] = res;
        if (A.argmax != nullptr) A.argmax[node] = leaf ? -1 : am;
      }
    }
    if (!LEAN && A.stats != nullptr && tsel != nullptr && grp == 0) {
      const int nvalid = __popcll(__ballot(valid));
      if (lane == 0) {
        atomicAdd(reinterpret_cast<unsigned long long*>(&A.stats[0]),
                  static_cast<unsigned long long>(nvalid));
        if (fallbacks)
          atomicAdd(reinterpret_cast<unsigned long long*>(&A.stats[1]),
                    static_cast<unsigned long long>(fallbacks));
        atomicAdd(reinterpret_cast<unsigned long long*>(&A.stats[2]),
                  static_cast<unsigned long long>(nvalid) * static_cast<unsigned long long>(npairs));
      }
    }
  }
}

// ---- two nodes per lane in int16 halves ------------------------------------
// Used for plain LB2 (no optional outputs, no `best` exit) when every bound
// value of the instance stays below 2^14: lane l of a 128-node tile bounds
// nodes 128 t + l (low halves) and 128 t + 64 + l (high halves).  Heads,
// tails and remaining work are built for both nodes at once in the int16
// halves; a Johnson step is then two bit extracts (plus word selects for
// n > 32), one pack, two v_pk_mad_u16 and a v_pk_max_i16 for both nodes
// (the offset form below).  The arithmetic is exact: every value, including
// the clock biased by -2^14 and the offsets by at most the remaining work,
// fits a signed 16-bit half while the path bound stays below 2^14.
//
// Tail combine.  After the scan T0 = head(a) + the remaining work of
// machine a, whatever the pair's Johnson order, so the term T0 + q[a] of
// cmaxFin is the one-machine bound of machine a; every machine a <= m - 2
// is the first machine of some pair, so with a suffix LB2 = max(max_a
// (head(a) + rem(a) + q[a]), max_s (t1_s + q[b_s])), and without one
// max_s (t1_s + Q0[b_s]) (borneInfMakespan 146-148).
typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ s16x2 as_s16(uint32_t x) { return __builtin_bit_cast(s16x2, x); }
__device__ __forceinline__ uint32_t as_u32(s16x2 x) { return __builtin_bit_cast(uint32_t, x); }

// four steps for two nodes, in offset form.  With S the running sum of p_b
// over the unscheduled jobs passed so far, the scan keeps u = t1 - S and
// D = T0 - S; an unscheduled job's step t0 += p_a; t1 = max(t1, t0 + lag)
// + p_b becomes
//   D += p_a - p_b;  u = max(u, D + 2^14 + lag + p_b)
// and a scheduled job (f = 0) leaves D and gives max(u, D) = u, since
// D <= u whenever T0 <= t1.  At the end S = rem(b), so t1 = u + rem(b).
// The flags: the job's bit of each node's 32-bit mask word (word chosen by
// the step's select masks, as in the one-node form), packed into the halves.
template <int NW>
__device__ __forceinline__ void step1_pk(const uint2 rec, const uint32_t (&uA)[NW],
                                         const uint32_t (&uB)[NW], s16x2& D, s16x2& u) {
  {
    const uint32_t j = rec.x;  // job: bit offset in its low 5 bits, word above
    uint32_t wA, wB;
    if constexpr (NW == 1) {
      wA = uA[0];
      wB = uB[0];
    } else {
      // word-select masks from the job's bits 5 and 6 (scalar ops)
      const uint32_t s5 = static_cast<uint32_t>(-static_cast<int>((j >> 5) & 1u));
      wA = bitsel(s5, uA[1], uA[0]);
      wB = bitsel(s5, uB[1], uB[0]);
      if constexpr (NW == 4) {
        const uint32_t s6 = static_cast<uint32_t>(-static_cast<int>((j >> 6) & 1u));
        wA = bitsel(s6, bitsel(s5, uA[3], uA[2]), wA);
        wB = bitsel(s6, bitsel(s5, uB[3], uB[2]), wB);
      }
    }
    const uint32_t fA = __builtin_amdgcn_ubfe(wA, j, 1u);
    const uint32_t fB = __builtin_amdgcn_ubfe(wB, j, 1u);
    const s16x2 f = as_s16((fB << 16) | fA);
    // one word holds both constants: p_a - p_b (low half), 2^14 + lag + p_b
    // (high half); each multiply-add reads its half in both lanes
    const s16x2 c = as_s16(rec.y);
    D = f * c.xx + D;
    const s16x2 z = f * c.yy + D;
    u = __builtin_elementwise_max(u, z);
  }
}

template <int NW>
__device__ __forceinline__ void step4_pk(const uint2 (&rec)[4], const uint32_t (&uA)[NW],
                                         const uint32_t (&uB)[NW], s16x2& D, s16x2& u) {
#pragma unroll
  for (int k = 0; k < 4; ++k) step1_pk<NW>(rec[k], uA, uB, D, u);
}

// Free-job compaction for tiles of sibling children (instances above 32
// jobs).  The children of a search round's 128-node tile belong to a few
// sibling parents, so the jobs free in at least one of them (the tile's
// union) are typically about half of n; a step of any other job has f = 0
// in both halves of every lane and changes nothing.  The wave filters its pairs' steps by the union, in
// chunks of kCompPairs(n4) pairs: every pair keeps exactly |union| steps, in
// its Johnson order, so pair k of a chunk starts at k |union| in the wave's
// LDS list, and the scan reads each step from LDS (one address for the
// whole wave) instead of the shared table.
constexpr int kCompPasses = 5;  // 64-step filter passes per chunk
__device__ __forceinline__ int comp_pairs(int n4) { return max(1, 64 * kCompPasses / n4); }
// uint2 records of a wave's compaction area (8 of read-ahead slack)
__device__ __host__ __forceinline__ int comp_area(int n4) {
  return (64 * kCompPasses / n4 > 1 ? 64 * kCompPasses / n4 : 1) * n4 + 8;
}

template <int MM, int NW, class Src>
__global__ __launch_bounds__(256, (MM <= 20 ? 3 : 2)) void bound_kernel_pk(DevTables T, Src src, BoundArgs A) {
  extern __shared__ int32_t lds[];
  const int n = T.n, m = T.m, P = T.P;
  const int mn4 = (MM * (n + 1) + 3) & ~3;
  int32_t* sp = lds;                                           // [n + 1][MM], row n zero
  uint32_t* part = reinterpret_cast<uint32_t*>(lds + mn4);     // [blockDim]
  uint8_t* pl = reinterpret_cast<uint8_t*>(part + blockDim.x);  // [n][128 tpb]
  stage_times<MM>(T, sp);
  for (int i = threadIdx.x; i < MM; i += blockDim.x) sp[n * MM + i] = 0;
  __syncthreads();
  const int G = A.groups;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int tpb = (blockDim.x >> 6) / G;
  // free-job compaction: sibling tiles of instances above 32 jobs (on 20
  // jobs the filter and the per-pair LDS latency cost more than the ~40% of
  // steps it removes; on 50 it saves a fifth of the solve)
  constexpr bool kCompact = Src::kSiblings && NW >= 2;
  // compaction only, after pl: the pairs' machines, then each wave's
  // compacted step list
  int2* s_pairs = reinterpret_cast<int2*>(pl + static_cast<size_t>(n) * 128 * tpb);
  uint2* cs = reinterpret_cast<uint2*>(s_pairs + P) +
              static_cast<size_t>(wave) * comp_area(T.n4);
  if constexpr (kCompact) {
    for (int i = threadIdx.x; i < P; i += blockDim.x) s_pairs[i] = T.pairs[i];
    __syncthreads();
  }
  const int tib = wave / G;
  const int grp = __builtin_amdgcn_readfirstlane(wave - tib * G);
  const int s_lo = (grp * P) / G, s_hi = ((grp + 1) * P) / G;
  const int N = (A.Ndev != nullptr) ? *A.Ndev : A.N;
  const int ntiles = (N + 127) >> 7;
  const int nbt = (ntiles + tpb - 1) / tpb;
  const int bst = 128 * tpb;
  uint32_t valid_bits[NW];
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    const int hi = n - 32 * w;
    valid_bits[w] = hi >= 32 ? 0xffffffffu : (hi > 0 ? ((1u << hi) - 1u) : 0u);
  }
  // the two-node form runs the recurrences itself (parent state, which
  // bound_kernel uses, measured slower here: two nodes' state per lane)
  constexpr bool from_parent = false;
  for (int bt = blockIdx.x; bt < nbt; bt += gridDim.x) {
    const int tile = __builtin_amdgcn_readfirstlane(bt * tpb + tib);
    int node[2], l1[2], l2[2], sw0[2], sw1[2];
    bool valid[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      node[h] = tile * 128 + h * 64 + lane;
      int stride = 0;
      const uint8_t* col = nullptr;
      l1[h] = -1;
      l2[h] = n;
      sw0[h] = sw1[h] = -1;
      valid[h] = node[h] < N && src.load(node[h], l1[h], l2[h], col, stride, sw0[h], sw1[h]);
      if (!valid[h]) {
        l1[h] = -1;
        l2[h] = n;
      }
      if (!from_parent && grp == 0 && valid[h]) {
        for (int i0 = 0; i0 < n; i0 += 16) {
          uint32_t v[16];
#pragma unroll
          for (int u = 0; u < 16; ++u)
            if (i0 + u < n) v[u] = col[static_cast<size_t>(i0 + u) * stride];
#pragma unroll
          for (int u = 0; u < 16; ++u)
            if (i0 + u < n) pl[(i0 + u) * bst + tib * 128 + h * 64 + lane] = static_cast<uint8_t>(v[u]);
        }
      }
    }
    if (G > 1) __syncthreads();
    // rp: start clocks (heads, or R0 without prefix); ep: what a pair's t1
    // still needs at its end, rem + (q with a suffix, Q0 without);
    // l1m: the one-machine terms of machines 0..m-2 (with a suffix)
    // Heads, tails and remaining work of both nodes at once, one int16 half
    // each.  At a position only one node uses, the other half reads the
    // zero row n: a zero step leaves heads (non-decreasing over machines)
    // and tails (non-increasing) unchanged.
    uint32_t rp[MM], ep[MM], uA[NW], uB[NW];
    uint32_t rP[MM], cP[MM];  // heads; tails (0 without suffix) + work
    uint32_t smA[NW], smB[NW];
#pragma unroll
    for (int k = 0; k < MM; ++k) rP[k] = cP[k] = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) smA[w] = smB[w] = 0;
    const uint8_t* mA = pl + tib * 128 + lane;
    const uint8_t* mB = mA + 64;
    const int d1A = l1[0] + 1, d1B = l1[1] + 1;
    // both halves' job rows at positions iA, iB packed (job n: zero row)
    auto rows = [&](int jA, int jB, uint32_t (&pP)[MM]) {
      const int4* ra = reinterpret_cast<const int4*>(sp + jA * MM);
      const int4* rb = reinterpret_cast<const int4*>(sp + jB * MM);
#pragma unroll
      for (int k4 = 0; k4 < MM / 4; ++k4) {
        const int4 a = ra[k4], b = rb[k4];
        pP[4 * k4] = static_cast<uint32_t>(a.x) | (static_cast<uint32_t>(b.x) << 16);
        pP[4 * k4 + 1] = static_cast<uint32_t>(a.y) | (static_cast<uint32_t>(b.y) << 16);
        pP[4 * k4 + 2] = static_cast<uint32_t>(a.z) | (static_cast<uint32_t>(b.z) << 16);
        pP[4 * k4 + 3] = static_cast<uint32_t>(a.w) | (static_cast<uint32_t>(b.w) << 16);
      }
    };
    // tails over the suffixes
    for (int i = n - 1; i >= min(l2[0], l2[1]); --i) {
      const bool aA = i >= l2[0], aB = i >= l2[1];
      const int pa = perm_at(mA, bst, sw0[0], sw1[0], i), pb = perm_at(mB, bst, sw0[1], sw1[1], i);
      if (aA) set_job_bit<NW>(smA, pa);
      if (aB) set_job_bit<NW>(smB, pb);
      uint32_t pP[MM];
      rows(aA ? pa : n, aB ? pb : n, pP);
#pragma unroll
      for (int k = MM - 1; k >= 0; --k) {
        const s16x2 nx = as_s16(k + 1 < MM ? cP[k + 1] : 0u);
        cP[k] = as_u32(__builtin_elementwise_max(as_s16(cP[k]), nx) + as_s16(pP[k]));
      }
    }
    // remaining work: the free positions (padding lanes read none)
    {
      const int fA = valid[0] ? l2[0] : d1A, fB = valid[1] ? l2[1] : d1B;
      for (int i = min(d1A, d1B); i < max(fA, fB); ++i) {
        const bool aA = i >= d1A && i < fA, aB = i >= d1B && i < fB;
        const int pa = perm_at(mA, bst, sw0[0], sw1[0], i), pb = perm_at(mB, bst, sw0[1], sw1[1], i);
        uint32_t pP[MM];
        rows(aA ? pa : n, aB ? pb : n, pP);
#pragma unroll
        for (int k = 0; k < MM; ++k) cP[k] = as_u32(as_s16(cP[k]) + as_s16(pP[k]));
      }
    }
    // heads over the prefixes
    for (int i = 0; i < max(d1A, d1B); ++i) {
      const bool aA = i < d1A, aB = i < d1B;
      const int pa = perm_at(mA, bst, sw0[0], sw1[0], i), pb = perm_at(mB, bst, sw0[1], sw1[1], i);
      if (aA) set_job_bit<NW>(smA, pa);
      if (aB) set_job_bit<NW>(smB, pb);
      uint32_t pP[MM];
      rows(aA ? pa : n, aB ? pb : n, pP);
      rP[0] = as_u32(as_s16(rP[0]) + as_s16(pP[0]));
#pragma unroll
      for (int k = 1; k < MM; ++k)
        rP[k] = as_u32(__builtin_elementwise_max(as_s16(rP[k]), as_s16(rP[k - 1])) + as_s16(pP[k]));
    }
    // per-half masks: no prefix (R0 instead of heads), a suffix (tails and
    // one-machine terms count)
    const uint32_t nd1 = (d1A == 0 ? 0xffffu : 0u) | (d1B == 0 ? 0xffff0000u : 0u);
    const uint32_t d2m = (l2[0] < n ? 0xffffu : 0u) | (l2[1] < n ? 0xffff0000u : 0u);
    s16x2 lvP = as_s16(0u), l1P = as_s16(0u);
#pragma unroll
    for (int k = 0; k < MM; ++k) {
      // a leaf has no free position (c = tail): its makespan
      lvP = __builtin_elementwise_max(lvP, as_s16(rP[k]) + as_s16(cP[k]));
      const uint32_t r0 = static_cast<uint32_t>(ld_const(T.R0, k)) * 0x10001u;  // zero past m
      const uint32_t q0 = static_cast<uint32_t>(ld_const(T.Q0, k)) * 0x10001u;
      rp[k] = (r0 & nd1) | (rP[k] & ~nd1);
      ep[k] = as_u32(as_s16(cP[k]) + as_s16(q0 & ~d2m));
      if (k < m - 1) l1P = __builtin_elementwise_max(l1P, as_s16(rp[k]) + as_s16(cP[k]));
    }
    const uint32_t l1m = as_u32(l1P) & d2m;
    const int leafv0 = static_cast<int>(as_u32(lvP) & 0xffffu);
    const int leafv1 = static_cast<int>(as_u32(lvP) >> 16);
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      uA[w] = ~smA[w] & valid_bits[w];
      uB[w] = ~smB[w] & valid_bits[w];
    }
    s16x2 lb = as_s16(grp == 0 ? l1m : 0u);
    const s16x2 bias = as_s16(static_cast<uint32_t>(kBias16) * 0x10001u);
    if constexpr (kCompact) {
      // the tile's union of free jobs (valid nodes only), wave-uniform
      uint32_t un[NW];
      int c = 0;  // steps kept per pair
#pragma unroll
      for (int w = 0; w < NW; ++w) {
        uint32_t x = (valid[0] ? uA[w] : 0u) | (valid[1] ? uB[w] : 0u);
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) x |= static_cast<uint32_t>(__shfl_xor(static_cast<int>(x), d, 64));
        un[w] = __builtin_amdgcn_readfirstlane(x);
        c += __popc(un[w]);
      }
      const int n4 = T.n4, cpp = comp_pairs(n4);
      const int kin0 = lane % n4;            // step index in its pair, first pass
      const int r64 = 64 % n4;                // and its advance per pass
      // a chunk's table loads (64 steps per pass, at most kCompPasses; the
      // steps of a chunk stay inside its pairs, so nothing past the table is
      // read); the next chunk's are in flight while this one is scanned
      uint2 rec[kCompPasses];
      auto load_chunk = [&](int sc0) {
        const int t0 = sc0 * n4, t1 = (sc0 + min(cpp, s_hi - sc0)) * n4;
#pragma unroll
        for (int ps = 0; ps < kCompPasses; ++ps) {
          const int tt = t0 + 64 * ps + lane;
          rec[ps] = tt < t1 ? T.steps16[tt] : make_uint2(0u, 0u);
        }
      };
      if (s_lo < s_hi) load_chunk(s_lo);
      for (int sc = s_lo; sc < s_hi; sc += cpp) {
        const int np = min(cpp, s_hi - sc);
        const int t0 = sc * n4, t1 = (sc + np) * n4;
        __builtin_amdgcn_wave_barrier();  // the previous chunk's reads first
        // pair k's list starts at k c4 (c4 = c rounded up to four steps);
        // the rounding slots hold zero steps, which change nothing
        const int c4 = (c + 3) & ~3;
        const float inv_c = 1.0f / static_cast<float>(max(c, 1));
        for (int i = lane; i < np * c4; i += 64) cs[i] = make_uint2(0u, 0u);
        __builtin_amdgcn_wave_barrier();
        int carry = 0, kin = kin0;
#pragma unroll
        for (int ps = 0; ps < kCompPasses; ++ps) {
          const int tt = t0 + 64 * ps + lane;
          const uint32_t j = rec[ps].x;
          uint32_t wd;
          if constexpr (NW == 1) {
            wd = un[0];
          } else if constexpr (NW == 2) {
            wd = (j >> 5) & 1u ? un[1] : un[0];
          } else {
            const uint32_t wi = (j >> 5) & 3u;
            wd = wi == 0 ? un[0] : wi == 1 ? un[1] : wi == 2 ? un[2] : un[3];
          }
          const bool keep = tt < t1 && kin < n && ((wd >> (j & 31u)) & 1u);
          const unsigned long long bal = __ballot(keep);
          if (keep) {
            // kept step K of the chunk: pair K / c, step K % c of that pair
            // (K < 320, so the float quotient is exact after the floor)
            const int K = carry + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(bal >> 32),
                                                            __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(bal), 0u));
            const int pk = static_cast<int>((static_cast<float>(K) + 0.5f) * inv_c);
            cs[pk * c4 + (K - pk * c)] = rec[ps];
          }
          carry += __popcll(bal);
          kin += r64;
          if (kin >= n4) kin -= n4;
        }
        if (sc + cpp < s_hi) load_chunk(sc + cpp);
        __builtin_amdgcn_wave_barrier();  // the wave's LDS writes before its reads
        // LDS only below (reads complete in order, so each wait is partial):
        // the pairs' machines, and groups of four steps read one group ahead
        const int2* pq = s_pairs + sc;
        int2 abn = pq[0];
        for (int pp = 0; pp < np; ++pp) {
          const int a = __builtin_amdgcn_readfirstlane(abn.x);
          const int b = __builtin_amdgcn_readfirstlane(abn.y);
          abn = pq[min(pp + 1, np - 1)];
          s16x2 D = as_s16(rp[a]) - bias;  // machine-a clocks, biased below zero
          s16x2 u = as_s16(rp[b]);
          const uint2* q = cs + pp * c4;
          uint2 ga[4], gb[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) ga[i] = q[i];  // reads past the list stay in the slack
          for (int k = 0; k < c4; k += 8) {
#pragma unroll
            for (int i = 0; i < 4; ++i) gb[i] = q[k + 4 + i];
            __builtin_amdgcn_sched_barrier(0);
            step4_pk<NW>(ga, uA, uB, D, u);
            if (k + 4 >= c4) break;
#pragma unroll
            for (int i = 0; i < 4; ++i) ga[i] = q[k + 8 + i];
            __builtin_amdgcn_sched_barrier(0);
            step4_pk<NW>(gb, uA, uB, D, u);
          }
          lb = __builtin_elementwise_max(lb, u + as_s16(ep[b]));
        }
      }
    }
    // the pairs' step groups are contiguous in the table: walk them as one
    // stream; the read-ahead past the last group stays inside the table's
    // padding steps
    const int gpp = T.n4 >> 2;  // step groups per pair
    const int g_lo = s_lo * gpp, g_hi = s_hi * gpp;
    if (!kCompact && g_lo < g_hi) {
      const uint2* gp = T.steps16 + static_cast<size_t>(g_lo) * 4;
      uint2 recA[4], recB[4];
      auto load_group = [&](const uint2* at, uint2 (&rc)[4]) {
#pragma unroll
        for (int q = 0; q < 4; ++q) rc[q] = ld_const2u(at, q);
      };
      int s = s_lo, kg = 0;
      int2 ab = ld_const2(T.pairs, s);
      s16x2 D = as_s16(rp[ab.x]) - bias;  // machine-a clocks, biased below zero
      s16x2 u = as_s16(rp[ab.y]);
      // at the end of a pair: its value, then the next pair's start clocks
      auto end_group = [&]() {
        if (++kg == gpp) {
          lb = __builtin_elementwise_max(lb, u + as_s16(ep[ab.y]));
          kg = 0;
          ab = ld_const2(T.pairs, min(++s, s_hi - 1));
          D = as_s16(rp[ab.x]) - bias;
          u = as_s16(rp[ab.y]);
        }
      };
      // eight steps (two groups) per scalar load: chunk c + 1 is loaded
      // while chunk c computes
      uint2 recC[4], recD[4];
      load_group(gp, recA);
      load_group(gp + 4, recB);
      for (int g = g_lo;;) {
        load_group(gp + 8, recC);
        load_group(gp + 12, recD);
        step4_pk<NW>(recA, uA, uB, D, u);
        end_group();
        if (++g == g_hi) break;
        step4_pk<NW>(recB, uA, uB, D, u);
        end_group();
        if (++g == g_hi) break;
        load_group(gp + 16, recA);
        load_group(gp + 20, recB);
        step4_pk<NW>(recC, uA, uB, D, u);
        end_group();
        if (++g == g_hi) break;
        step4_pk<NW>(recD, uA, uB, D, u);
        end_group();
        if (++g == g_hi) break;
        gp += 16;
      }
    }
    if (G > 1) {
      part[threadIdx.x] = as_u32(lb);
      __syncthreads();
      if (grp == 0)
        for (int g = 1; g < G; ++g) lb = __builtin_elementwise_max(lb, as_s16(part[(tib * G + g) * 64 + lane]));
      __syncthreads();
    }
    if (grp == 0) {
      const uint32_t lbu = as_u32(lb);
#pragma unroll
      for (int h = 0; h < 2; ++h)
        if (valid[h]) {
          const int res = (l2[h] - l1[h] == 1) ? (h ? leafv1 : leafv0)
                                               : static_cast<int>((lbu >> (16 * h)) & 0xffffu);
          A.out[src.slot(node[h])] = res;
        }
    }
  }
}

template <int MM, int NW, class Src>
hipError_t launch_pk(const DevTables& T, const Src& src, const BoundArgs& A, int grid_cap,
                     hipStream_t stream) {
  const int block = 256;
  const int tpb = (block / 64) / A.groups;
  const int ntiles = (A.N + 127) / 128;
  int grid = (ntiles + tpb - 1) / tpb;
  if (grid > grid_cap) grid = grid_cap;
  if (grid < 1) grid = 1;
  size_t shmem = sizeof(int32_t) * (((MM * (T.n + 1) + 3) & ~3) + block) +
                 static_cast<size_t>(T.n) * 128 * tpb;
  if (Src::kSiblings && NW >= 2)
    shmem += sizeof(int2) * T.P + sizeof(uint2) * comp_area(T.n4) * (block / 64);
  hipExtLaunchKernelGGL((bound_kernel_pk<MM, NW, Src>), dim3(grid), dim3(block), shmem, stream,
                        A.ev_start, A.ev_stop, 0, T, src, A);
  return hipGetLastError();
}

template <class Src>
hipError_t launch_pk_src(const DevTables& T, const Src& src, const BoundArgs& A, int grid_cap,
                         hipStream_t stream) {
  const int nw = T.n <= 32 ? 1 : (T.n <= 64 ? 2 : 4);
  const int mm = T.m <= 8 ? 8 : (T.m <= 12 ? 12 : (T.m <= 16 ? 16 : (T.m <= 20 ? 20 : 24)));
#define WSBB_PK(MMv, NWv) \
  if (mm == MMv && nw == NWv) return launch_pk<MMv, NWv, Src>(T, src, A, grid_cap, stream);
  WSBB_PK(8, 1) WSBB_PK(8, 2) WSBB_PK(8, 4)
  WSBB_PK(12, 1) WSBB_PK(12, 2) WSBB_PK(12, 4)
  WSBB_PK(16, 1) WSBB_PK(16, 2) WSBB_PK(16, 4)
  WSBB_PK(20, 1) WSBB_PK(20, 2) WSBB_PK(20, 4)
  WSBB_PK(24, 1) WSBB_PK(24, 2) WSBB_PK(24, 4)
#undef WSBB_PK
  return hipErrorInvalidValue;
}

// Selector features of every node as BF16 B-operand fragments in the
// 8-wide MFMA layout (two of them make one 32x32x16 operand).  5m features per node: per machine the head (R0
// without prefix), the tail (Q0 without suffix), the remaining work of the
// free jobs, head + work and work + tail, in that order of groups, divided by
// `scale`.  Lane half*32 + c of the fragment of 32-node block b and feature
// step t holds features 8t + 4half .. 8t + 4half + 3 of node 32b + c.  One
// wave per 64-node tile; the remaining work is the row total minus the
// scheduled jobs, which the head/tail recurrences visit anyway.
//
// A launch over a round's parents is small (a few hundred waves), so the
// kernel is written for latency: processing times sit in LDS job-major (a
// job's machines are one 16-byte read per four), the next job index is read
// while the current one is processed, and with EXACT (m == MM) the feature
// loop is unrolled over constant indices.  Otherwise the per-machine values
// pass through the wave's LDS area and a block table gives, per feature, the
// one or two rows it sums, so that loop has no branches and no selects.
template <int MM, bool EXACT, class Src>
__global__ __launch_bounds__(256) void k_features(DevTables T, Src src, FeatArgs A) {
  extern __shared__ int32_t lds[];
  static_assert(MM % 4 == 0, "machine rows are read four at a time");
  const int n = T.n, m = T.m;
  int32_t* sp = lds;  // [n][MM] job-major, machines m..MM-1 zero
  uint8_t* pl = reinterpret_cast<uint8_t*>(lds + MM * n);  // [n][256]
  stage_times<MM>(T, sp);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // feature f = g m + k sums rows ra + rb of hv (row 3 MM is zero): one entry
  // ra | rb << 16 per feature, zero rows for the padding features
  constexpr int kRows = 3 * MM + 1, kZero = 3 * MM, kFmax = 8 * ((5 * MM + 7) / 8);
  int32_t* tab = reinterpret_cast<int32_t*>(pl + n * 256);
  if (!EXACT && threadIdx.x < kFmax) {
    const int f = threadIdx.x, g = f / m, k = f % m;
    const int ra = g == 0 ? k : g == 1 ? MM + k : g == 2 || g == 4 ? 2 * MM + k : g == 3 ? k : kZero;
    const int rb = g == 3 ? 2 * MM + k : g == 4 ? MM + k : kZero;
    tab[f] = ra | (rb << 16);
  }
  __syncthreads();
  // per wave: head, tail, remaining work and a zero row [3 MM + 1][64]
  int32_t* hv = tab + kFmax + static_cast<size_t>(wave) * kRows * 64;
  if (!EXACT) hv[kZero * 64 + lane] = 0;
  const int N = A.Ndev ? *A.Ndev : A.N;
  const int ntiles = (N + 63) >> 6;
  const float sc = 1.0f / A.scale;
  for (int tile0 = blockIdx.x * 4; tile0 < ntiles; tile0 += gridDim.x * 4) {
    const int tile = __builtin_amdgcn_readfirstlane(tile0 + wave);
    if (tile >= ntiles) continue;
    const int node = tile * 64 + lane;
    int l1 = -1, l2 = n, stride = 0, sw0 = -1, sw1 = -1;
    const uint8_t* col = nullptr;
    const bool valid = node < N && src.load(node, l1, l2, col, stride, sw0, sw1);
    if (!valid) {
      l1 = -1;
      l2 = n;
    }
    const int d1 = l1 + 1, d2 = n - l2;
    if (valid) {
      for (int i0 = 0; i0 < n; i0 += 32) {
        uint32_t v[32];
#pragma unroll
        for (int u = 0; u < 32; ++u)
          if (i0 + u < n) v[u] = col[static_cast<size_t>(i0 + u) * stride];
#pragma unroll
        for (int u = 0; u < 32; ++u)
          if (i0 + u < n) pl[(i0 + u) * 256 + threadIdx.x] = static_cast<uint8_t>(v[u]);
      }
    }
    const uint8_t* mine = pl + threadIdx.x;
    int r[MM], q[MM], sc_[MM];
#pragma unroll
    for (int k = 0; k < MM; ++k) r[k] = q[k] = sc_[k] = 0;
    int jn = d1 > 0 ? perm_at(mine, 256, sw0, sw1, 0) : 0;
    for (int i = 0; i < d1; ++i) {
      const int j = jn;
      if (i + 1 < d1) jn = perm_at(mine, 256, sw0, sw1, i + 1);
      int pk[MM];
      job_row<MM>(sp, j, pk);
      r[0] += pk[0];
      sc_[0] += pk[0];
#pragma unroll
      for (int k = 1; k < MM; ++k) {
        r[k] = max(r[k], r[k - 1]) + pk[k];
        sc_[k] += pk[k];
      }
    }
    jn = d2 > 0 ? perm_at(mine, 256, sw0, sw1, n - 1) : 0;
    for (int i = n - 1; i >= l2; --i) {
      const int j = jn;
      if (i - 1 >= l2) jn = perm_at(mine, 256, sw0, sw1, i - 1);
      int pk[MM];
      job_row<MM>(sp, j, pk);
#pragma unroll
      for (int k = MM - 1; k >= 0; --k) {
        q[k] = max(q[k], (k + 1 < MM) ? q[k + 1] : 0) + pk[k];
        sc_[k] += pk[k];
      }
    }
    // per machine: head, tail, remaining work
#pragma unroll
    for (int k = 0; k < MM; ++k) {
      r[k] = d1 ? r[k] : ld_const(T.R0, k);  // tables are zero past machine m
      q[k] = d2 ? q[k] : ld_const(T.Q0, k);
      sc_[k] = ld_const(T.ptot, k) - sc_[k];
    }
    const int b = node >> 5, c = lane & 31;
    const float vs = valid ? sc : 0.f;
    if constexpr (EXACT) {
      // the FP32 copy (logging only) is a separate instance of the loop, so
      // the common case carries no per-feature test
      auto emit = [&](auto with_f32) {
        float* fo = A.f32 + node;  // advanced by ld per feature
        uint2* fr = A.frag + static_cast<size_t>(b) * A.fsteps * 64 + c;
#pragma unroll
        for (int t = 0; t < (5 * MM + 7) / 8; ++t)
#pragma unroll
          for (int half = 0; half < 2; ++half) {
            uint32_t w[2] = {0u, 0u};
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const int f = 8 * t + 4 * half + i, g = f / MM, k = f % MM;
              if (f < 5 * MM) {
                const int v = g == 0 ? r[k] : g == 1 ? q[k] : g == 2 ? sc_[k] : g == 3 ? r[k] + sc_[k] : sc_[k] + q[k];
                const float x = static_cast<float>(v) * vs;
                if constexpr (decltype(with_f32)::value) {
                  if (valid) *fo = x;
                  fo += A.ld;
                }
                w[i >> 1] |= static_cast<uint32_t>(f2bf_dev(x)) << (16 * (i & 1));
              }
            }
            // lanes 0-31 and 32-63 belong to consecutive 32-node blocks
            if (A.frag) fr[t * 64 + half * 32] = make_uint2(w[0], w[1]);
          }
      };
      if (A.f32)
        emit(std::true_type{});
      else
        emit(std::false_type{});
    } else {
      // each lane writes and reads back only its own column
#pragma unroll
      for (int k = 0; k < MM; ++k) {
        hv[k * 64 + lane] = r[k];
        hv[(MM + k) * 64 + lane] = q[k];
        hv[(2 * MM + k) * 64 + lane] = sc_[k];
      }
      const int F = 5 * m;
      for (int t = 0; t < A.fsteps; ++t)
#pragma unroll
        for (int half = 0; half < 2; ++half) {
          uint32_t w[2] = {0u, 0u};
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int f = 8 * t + 4 * half + i;
            const int e = tab[f];
            const float x = static_cast<float>(hv[(e & 0xffff) * 64 + lane] + hv[(e >> 16) * 64 + lane]) * vs;
            if (A.f32 && valid && f < F) A.f32[static_cast<size_t>(f) * A.ld + node] = x;
            w[i >> 1] |= static_cast<uint32_t>(f2bf_dev(x)) << (16 * (i & 1));
          }
          if (A.frag)
            A.frag[(static_cast<size_t>(b) * A.fsteps + t) * 64 + half * 32 + c] = make_uint2(w[0], w[1]);
        }
    }
  }
}

template <int MM, bool EXACT, class Src>
hipError_t launch_feat_mm(const DevTables& T, const Src& src, const FeatArgs& A, int grid_cap,
                          hipStream_t stream) {
  const int ntiles = (A.N + 63) / 64;
  int grid = (ntiles + 3) / 4;
  if (grid > grid_cap) grid = grid_cap;
  if (grid < 1) grid = 1;
  const size_t shmem = sizeof(int32_t) * MM * T.n + static_cast<size_t>(T.n) * 256 +
                       (EXACT ? 0 : sizeof(int32_t) * (8 * ((5 * MM + 7) / 8) + 4 * (3 * MM + 1) * 64));
  hipLaunchKernelGGL((k_features<MM, EXACT, Src>), dim3(grid), dim3(256), shmem, stream, T, src, A);
  return hipGetLastError();
}

template <class Src>
hipError_t launch_feat_src(const DevTables& T, const Src& src, const FeatArgs& A, int grid_cap,
                           hipStream_t stream) {
  if (A.N <= 0) return hipSuccess;
  if (A.fsteps != (5 * T.m + 7) / 8) return hipErrorInvalidValue;
  switch (T.m) {
    case 8: return launch_feat_mm<8, true, Src>(T, src, A, grid_cap, stream);
    case 12: return launch_feat_mm<12, true, Src>(T, src, A, grid_cap, stream);
    case 16: return launch_feat_mm<16, true, Src>(T, src, A, grid_cap, stream);
    case 20: return launch_feat_mm<20, true, Src>(T, src, A, grid_cap, stream);
    case 24: return launch_feat_mm<24, true, Src>(T, src, A, grid_cap, stream);
    default: break;
  }
  const int mm = T.m <= 8 ? 8 : (T.m <= 12 ? 12 : (T.m <= 16 ? 16 : (T.m <= 20 ? 20 : 24)));
  switch (mm) {
    case 8: return launch_feat_mm<8, false, Src>(T, src, A, grid_cap, stream);
    case 12: return launch_feat_mm<12, false, Src>(T, src, A, grid_cap, stream);
    case 16: return launch_feat_mm<16, false, Src>(T, src, A, grid_cap, stream);
    case 20: return launch_feat_mm<20, false, Src>(T, src, A, grid_cap, stream);
    default: return launch_feat_mm<24, false, Src>(T, src, A, grid_cap, stream);
  }
}

template <int MM, int NW, int MODE, int BLOCK, class Src>
hipError_t launch_mm_nw(const DevTables& T, const Src& src, const BoundArgs& A,
                        int grid_cap, hipStream_t stream) {
  const int block = BLOCK;
  const int tpb = (block / 64) / A.groups;
  const int ntiles = (A.N + 63) / 64;
  int grid = (ntiles + tpb - 1) / tpb;
  if (grid > grid_cap) grid = grid_cap;
  if (grid < 1) grid = 1;
  const size_t shmem = sizeof(int32_t) * (MM * (T.n + 3) + 4 * block) +
                       static_cast<size_t>(T.n) * 64 * tpb;
  hipExtLaunchKernelGGL((bound_kernel<MM, NW, MODE, BLOCK, Src>), dim3(grid), dim3(block), shmem,
                        stream, A.ev_start, A.ev_stop, 0, T, src, A);
  return hipGetLastError();
}

template <class Src>
hipError_t launch_src(const DevTables& T, const Src& src, const BoundArgs& A,
                      int grid_cap, hipStream_t stream) {
  const int nw = T.n <= 32 ? 1 : (T.n <= 64 ? 2 : 4);
  const int mm = T.m <= 8 ? 8 : (T.m <= 12 ? 12 : (T.m <= 16 ? 16 : (T.m <= 20 ? 20 : 24)));
  const bool bare = A.kind == 0 && A.argmax == nullptr && A.pairv == nullptr &&
                   A.sel == nullptr && A.stats == nullptr && A.feat == nullptr;
  // the `best` early exit takes the first running maximum above best in
  // pair order, which the general mode follows pair by pair
  const int mode = (A.kind == 1 || A.feat != nullptr)           ? 2
                   : (bare && A.best == -1 && A.tile_sel == nullptr) ? 0
                   : (bare && A.best == -1 && A.margin < 0)          ? 3
                                                                     : 1;
  // 8 or 16 waves per tile (1024-thread blocks) for small batches; the lean
  // variants for up to 16 machines fit four such waves per SIMD
  if (mode == 0 && A.best == -1 && T.steps16 != nullptr && A.groups <= 4 && A.pk != 0)
    return launch_pk_src(T, src, A, grid_cap, stream);
  bool wide = A.groups > 4;
  if (wide && (mode != 0 || mm > 16 || std::is_same<Src, ListSource>::value)) {
    // the general and wide-machine variants stop at four waves per tile
    BoundArgs A4 = A;
    A4.groups = 4;
    return launch_src(T, src, A4, grid_cap, stream);
  }
#define WSBB_DISPATCH(MMv, NWv)                                                         \
  if (mm == MMv && nw == NWv) {                                                         \
    if constexpr (MMv <= 16)                                                            \
      if (wide) return launch_mm_nw<MMv, NWv, 0, 1024, Src>(T, src, A, grid_cap, stream); \
    if constexpr (std::is_same<Src, ListSource>::value)                                   \
      return launch_mm_nw<MMv, NWv, 0, 256, Src>(T, src, A, grid_cap, stream);            \
    else                                                                                \
      return mode == 0   ? launch_mm_nw<MMv, NWv, 0, 256, Src>(T, src, A, grid_cap, stream) \
             : mode == 3 ? launch_mm_nw<MMv, NWv, 3, 256, Src>(T, src, A, grid_cap, stream) \
             : mode == 1 ? launch_mm_nw<MMv, NWv, 1, 256, Src>(T, src, A, grid_cap, stream) \
                         : launch_mm_nw<MMv, NWv, 2, 256, Src>(T, src, A, grid_cap, stream); \
  }
  WSBB_DISPATCH(8, 1) WSBB_DISPATCH(8, 2) WSBB_DISPATCH(8, 4)
  WSBB_DISPATCH(12, 1) WSBB_DISPATCH(12, 2) WSBB_DISPATCH(12, 4)
  WSBB_DISPATCH(16, 1) WSBB_DISPATCH(16, 2) WSBB_DISPATCH(16, 4)
  WSBB_DISPATCH(20, 1) WSBB_DISPATCH(20, 2) WSBB_DISPATCH(20, 4)
  WSBB_DISPATCH(24, 1) WSBB_DISPATCH(24, 2) WSBB_DISPATCH(24, 4)
#undef WSBB_DISPATCH
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_features(const DevTables& T, const BatchSource& src, const FeatArgs& A,
                           int grid_cap, hipStream_t stream) {
  return launch_feat_src(T, src, A, grid_cap, stream);
}

hipError_t launch_features(const DevTables& T, const ChildSource& src, const FeatArgs& A,
                           int grid_cap, hipStream_t stream) {
  return launch_feat_src(T, src, A, grid_cap, stream);
}

hipError_t launch_bound(const DevTables& T, const BatchSource& src,
                        const BoundArgs& A, int grid_cap, hipStream_t stream) {
  if (A.N <= 0) return hipSuccess;
  return launch_src(T, src, A, grid_cap, stream);
}

hipError_t launch_bound(const DevTables& T, const ChildSource& src,
                        const BoundArgs& A, int grid_cap, hipStream_t stream) {
  if (A.N <= 0) return hipSuccess;
  return launch_src(T, src, A, grid_cap, stream);
}

hipError_t launch_bound(const DevTables& T, const ListSource& src,
                        const BoundArgs& A, int grid_cap, hipStream_t stream) {
  if (A.N <= 0) return hipSuccess;
  if (A.kind != 0 || A.argmax || A.pairv || A.sel || A.tile_sel || A.stats || A.feat)
    return hipErrorInvalidValue;  // listed nodes get plain LB2 only
  return launch_src(T, src, A, grid_cap, stream);
}

}  // namespace wsbb
