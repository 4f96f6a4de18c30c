// Device-resident branch-and-bound for the permutation flowshop.
//
// The node pool is a LIFO stack in HBM (structure of arrays: position-major
// uint8 permutations, limits, bound).  Pushing the surviving children of the
// popped parents last-parent-first and, within a parent, worst-rank-first
// makes the stack pop nodes in exactly the order of the reference's
// priority_queue over lexicographic ranks (tree.cpp:37-46): two pool nodes
// are never ancestor and descendant, so every child of a popped node sorts
// before every other pool node.
//
// One round (all kernels on the context's stream):
//   select   pop parents into <= max_batches sub-batches, each reserving
//            2*(n-depth) children below pool_size-1 (explorer.cu:166-208,
//            problems::push_back1 problems.cpp:64-78); drop parents whose
//            bound is not below the incumbent (tree::bound, tree.cpp:247-250)
//   expand   copy parents out of the stack, write child descriptors
//            (problems_generate3 / leaves_generate3, decompose_beginend.cpp
//            114-138, in the reference's BEGIN-then-END order)
//   bound    LB of every child (bound.hip)
//   rank     per parent: branching side (choise, decompose_beginend.cpp
//            39-50 with best = 99999), ranks by (bound, job) (problems::ranks,
//            problems.cpp:46-62 with select_worst), leaf minimum
//   settle   incumbent seen by each parent in pop order: the running
//            minimum of earlier parents' leaves (solutions::insert
//            strictly-better rule, solutions.cpp:45-65; prune loop
//            gpu.cu:317-394); children of the chosen side with bound below
//            it are kept; prefix sums of kept children, overflow check
//   push     write survivors onto the stack
#include <hip/hip_runtime.h>

#include <climits>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/wsbb_amd.h"
#include "bound.h"
#include "context.h"
#include "search.h"
#include "selector.h"

namespace wsbb {

namespace {

enum : int { kRunning = 0, kDone = 1, kBudget = 2, kOverflow = 3 };
enum : int {
  kStDecomposed = 0, kStBounded, kStLeaves, kStPrunedPop, kStPrunedIns,
  kStImproved, kStMaxPool, kStRounds, kStSubset, kStFallback, kStPairs, kStWork
};
constexpr int kSideBest = 99999;  // decompose_beginend.h:26, never updated
constexpr int kSelWin = 16384;  // candidates per select window after the first
constexpr int kSelWin0 = 4096;  // candidates in the first window
constexpr int kSettleK = 16;   // consecutive parents per settle thread
constexpr int kRankGrid = 4096;  // at most this many rank blocks (4 per settle thread)
constexpr int kSettleChunk = 1024 * kSettleK;  // parents per chunk of the one-block settle
constexpr int kChainK = 8;     // consecutive parents per thread of the chained settle
constexpr int kChainChunk = 1024 * kChainK;    // parents per block of the chained settle
static_assert(kRankGrid == 4 * 1024, "each of the settle block's threads sums four rank blocks' counts");
constexpr int kSelChain = 8;   // blocks of the chained select (large pools)
constexpr int kLink = 8;       // ints per hand-off: flag, stop, carry, nb, np, nc, pruned, consumed
constexpr int kSettleChain = 16;  // blocks of the chained settle (large pools)
// largest cost a select key holds: bounds stay below 2^19, and a larger
// host-given cost compares with every incumbent as this one does
constexpr int kKeyCost = (1 << 23) - 1;
__host__ __device__ inline int select_key(int l1, int l2, int cost, int n) {
  return ((cost < kKeyCost ? cost : kKeyCost) << 8) | ((l1 + 1) + (n - l2));
}

struct Ctrl {
  int status;
  int stack_size;
  int inc;
  int n_parents;
  int n_children;
  int push_base;
  int push_total;
  int have_perm;
  int n_fb;  // children of the round listed for the full-LB2 fallback
  int sel_ticket;  // blocks of the chained select started this round
  int set_ticket;  // blocks of the chained settle started this round
  int pad[5];
  long long stats[WSBB_NSTATS];
  uint8_t best_perm[kMaxJobs];
};

struct SArgs {
  int n, m, P, CAP, PCAP, CCAP, pool_m1, max_batches;
  int count_pairs;  // 1: every child evaluates all P pairs (plain LB2);
                    // 2: topk pairs per child (learned subsets)
  int topk;
  long long budget;
  Ctrl* ctrl;
  uint8_t* s_perm;  // [n][CAP]
  int* s_l1;
  int* s_l2;
  int* s_cost;
  int* s_key;       // [CAP] what select reads of a node, one word:
                    // min(cost, kKeyCost) << 8 | depth
  int* sel_link;    // [kSelChain][kLink] hand-off records of the chained select
  int* set_link;    // [kSettleChain][kLink] records of the chained settle
  uint8_t* p_perm;  // [n][PCAP]
  int* p_l1;
  int* p_l2;
  int* p_cost;
  int* p_sel;       // stack index of parent
  int* p_off;       // [PCAP+1] first child
  int* p_side;
  int* p_nch;       // children on the chosen side
  int* p_leafmin;
  int* p_leafmax;   // cost of the leaf a simple parent inserts first
  int* p_leafwin;
  int* p_kept0;     // chosen-side children with bound below the round's
                    // starting incumbent
  int* p_kept;
  int* p_koff;
  int* c_desc;
  int* c_bound;
  int* r_order;     // [PCAP][n] per parent, by rank: the chosen-side child index
  int* r_cost;      // [PCAP][n] per parent, by rank: that child's bound
  int* r_leaves;    // [kRankGrid] leaf children counted by each rank block
  int* p_hr;        // [MM][PCAP + 1] heads of the parent's prefix
  int* p_hq;        // [MM][PCAP + 1] tails of the parent's suffix
  uint32_t* p_hm;   // [NW][PCAP + 1] the parent's scheduled jobs
  int* p_hw;        // [MM][PCAP + 1] work of the parent's scheduled jobs
  // selector features of the parents written by expand (m a multiple of
  // four): BF16 MFMA fragments and/or FP32 rows [5m][x_ld], or null
  uint2* x_frag;
  float* x_f32;
  int x_ld, x_fsteps;
  float x_inv_scale;
};

struct OpAdd {
  __device__ int operator()(int a, int b) const { return a + b; }
};
struct OpMin {
  __device__ int operator()(int a, int b) const { return a < b ? a : b; }
};

// A kernel reads its control fields, then issues the loads it can make at
// any index (clamped to the capacity) before it knows the round's parent
// count, then calls this: loads written before it stay before it.  Without
// it the compiler sinks a load whose value is only used after a branch on
// the parent count below that branch, and the kernel makes two dependent
// memory round trips instead of one.  No instruction and no wait: the
// barrier only orders memory operations in the compiler.
__device__ __forceinline__ void issue_before_branch() { asm volatile("" ::: "memory"); }

// Wave-wide scans on DPP lane moves: row shifts by 1, 2, 4, 8 scan each
// 16-lane row, the row broadcasts of lanes 15 and 31 carry the row totals
// upward (GFX9 DPP controls 0x111-0x118, 0x142, 0x143; wave shift right by
// one is 0x138).  Each step is a VALU move, not an LDS round trip.  Every
// lane of the wave must be active.
#define WSBB_DPP(v, ident, ctrl, rmask) __builtin_amdgcn_update_dpp((ident), (v), (ctrl), (rmask), 0xf, false)
template <class Op>
__device__ __forceinline__ int wave_incl_scan(int v, int ident, Op op) {
  v = op(v, WSBB_DPP(v, ident, 0x111, 0xf));
  v = op(v, WSBB_DPP(v, ident, 0x112, 0xf));
  v = op(v, WSBB_DPP(v, ident, 0x114, 0xf));
  v = op(v, WSBB_DPP(v, ident, 0x118, 0xf));
  v = op(v, WSBB_DPP(v, ident, 0x142, 0xa));
  v = op(v, WSBB_DPP(v, ident, 0x143, 0xc));
  return v;
}
// value of the lane below (ident in lane 0)
__device__ __forceinline__ int wave_shr1(int v, int ident) { return WSBB_DPP(v, ident, 0x138, 0xf); }
// reduction over the wave, in every lane (wave-uniform)
template <class Op>
__device__ __forceinline__ int wave_reduce(int v, int ident, Op op) {
  return __builtin_amdgcn_readlane(wave_incl_scan(v, ident, op), 63);
}
struct OpMax {
  __device__ int operator()(int a, int b) const { return a > b ? a : b; }
};

// exclusive scan of one value per thread over a 1024-thread block; *total
// receives the reduction of all values (every thread)
template <class Op>
__device__ __forceinline__ int block_exscan1024(int v, int ident, Op op, int* s_w16,
                                                int* total) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int incl = wave_incl_scan(v, ident, op);
  const int excl_w = wave_shr1(incl, ident);
  if (lane == 63) s_w16[wave] = incl;
  __syncthreads();
  if (wave == 0) {
    const int x = lane < 16 ? s_w16[lane] : ident;
    const int xi = wave_incl_scan(x, ident, op);
    const int xe = wave_shr1(xi, ident);
    if (lane < 16) s_w16[16 + lane] = xe;  // exclusive wave prefixes
    if (lane == 15) s_w16[32] = xi;
  }
  __syncthreads();
  const int res = op(s_w16[16 + wave], excl_w);
  *total = s_w16[32];
  __syncthreads();
  return res;
}
__device__ __forceinline__ int block_exscan1024(int v, int* s_w16, int* total) {
  return block_exscan1024(v, 0, OpAdd(), s_w16, total);
}

// two exclusive sums over a 1024-thread block sharing one set of barriers
// (s_w needs 80 ints)
__device__ __forceinline__ void block_exscan1024_2(int a, int b, int* s_w, int* ea, int* eb,
                                                   int* ta, int* tb) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ia = wave_incl_scan(a, 0, OpAdd()), ib = wave_incl_scan(b, 0, OpAdd());
  const int xa = wave_shr1(ia, 0), xb = wave_shr1(ib, 0);
  if (lane == 63) {
    s_w[wave] = ia;
    s_w[40 + wave] = ib;
  }
  __syncthreads();
  if (wave == 0) {
    const int va = lane < 16 ? s_w[lane] : 0, vb = lane < 16 ? s_w[40 + lane] : 0;
    const int sa = wave_incl_scan(va, 0, OpAdd()), sb = wave_incl_scan(vb, 0, OpAdd());
    const int pa = wave_shr1(sa, 0), pb = wave_shr1(sb, 0);
    if (lane < 16) {
      s_w[16 + lane] = pa;
      s_w[56 + lane] = pb;
    }
    if (lane == 15) {
      s_w[32] = sa;
      s_w[72] = sb;
    }
  }
  __syncthreads();
  *ea = s_w[16 + wave] + xa;
  *eb = s_w[56 + wave] + xb;
  *ta = s_w[32];
  *tb = s_w[72];
  __syncthreads();
}

// first index i in [lo, hi) with a[i] >= thr for nondecreasing a (hi if
// none), by the whole wave: each round probes the last element of 64 equal
// buckets and keeps the first bucket that reaches thr, until at most 64
// candidates remain; a last ballot picks the first
// LDS index of candidate i in the select window: one pad word per 16
// entries keeps a thread's 16 consecutive candidates out of its neighbours'
// banks
__device__ __forceinline__ int sel_pad(int i) { return i + (i >> 4); }

__device__ __forceinline__ int wave_first_at_least(const int* a, int lo, int hi, int thr,
                                                   int lane) {
  const int hi0 = hi;
  if (hi - lo <= 0) return hi0;
  while (hi - lo > 64) {
    const int stride = (hi - lo + 63) >> 6;
    int idx = lo + (lane + 1) * stride - 1;
    idx = idx < hi - 1 ? idx : hi - 1;
    const unsigned long long b1 = __ballot(a[sel_pad(idx)] >= thr);
    if (b1 == 0) return hi0;
    const int f = __ffsll(static_cast<long long>(b1)) - 1;
    lo += f * stride;
    hi = min(lo + stride, hi);
  }
  const int j = lo + lane;
  const unsigned long long b2 = __ballot(j < hi && a[sel_pad(j)] >= thr);
  return b2 ? lo + __ffsll(static_cast<long long>(b2)) - 1 : hi0;
}

// ---- select: one block ------------------------------------------------------
// Candidates are taken from the top of the stack in windows of kSelWin.  The
// inclusive prefix sums P of the reservations 2(n - depth) (0 for parents
// whose bound is not below the incumbent) locate every sub-batch boundary: a
// sub-batch that opened at candidate b (with carried reservation c) closes in
// front of the first candidate i with c + P[i] - P[b-1] >= pool_size - 1,
// which is the parent push_back1 turns away and the next sub-batch starts
// with.  Wave 0 walks the boundaries with wave-wide searches; the taken
// parents and their first-child offsets are then written by all threads.
// One window of KPER * 1024 candidates; the carried state is block-uniform.
struct SelState {
  int carry, nb, np, nc, pruned, consumed, stop;
};

// The part of a window that needs no carried state: the words of the
// candidates [consumed, consumed + win) and their prefix sums within the
// window (in LDS, for the boundary search); word[] keeps each thread's own.
template <int KPER>
__device__ __forceinline__ void window_words(const SArgs& S, int top, int inc, int consumed,
                                             int win, int* s_p, int* s_w16, int (&word)[KPER]) {
  const int tid = threadIdx.x, n = S.n;
  // coalesced loads: candidate i = t + 1024 k is the i-th from the top; its
  // reservation w (push_back1) and children c share one LDS word (both
  // at most 2n); one key word per candidate (cost and depth); every load is
  // issued before the first use (indices past the window are clamped to a
  // valid slot and their words zeroed)
  {
    int key[KPER];
#pragma unroll
    for (int k = 0; k < KPER; ++k) {
      const int i = tid + 1024 * k;
      key[k] = S.s_key[top - 1 - (consumed + min(i, win - 1))];
    }
#pragma unroll
    for (int k = 0; k < KPER; ++k) {
      const int i = tid + 1024 * k;
      const int depth = key[k] & 255, u = n - depth;
      const int w = ((key[k] >> 8) < inc) ? 2 * u : 0;  // reservation of push_back1
      const int c = (depth == n - 2) ? 2 : 2 * u;       // children actually generated
      s_p[sel_pad(i)] = i < win ? (w | (c << 16)) : 0;
    }
  }
  __syncthreads();
  // blocked inclusive scan: thread t owns candidates [KPER t, KPER t + KPER)
  // and keeps their words in registers; the prefix sums go back to LDS for
  // the boundary search
  int loc[KPER], tsum = 0;
#pragma unroll
  for (int k = 0; k < KPER; ++k) {
    word[k] = s_p[sel_pad(KPER * tid + k)];
    tsum += word[k] & 0xffff;
    loc[k] = tsum;
  }
  int total;
  const int off = block_exscan1024(tsum, s_w16, &total);
#pragma unroll
  for (int k = 0; k < KPER; ++k) s_p[sel_pad(KPER * tid + k)] = off + loc[k];
  __syncthreads();
}

// The part that needs the carried state.  window_bounds: the sub-batch
// boundaries from the carried reservation (wave 0); returns the end of the
// round's candidates in the window (win unless the last sub-batch fills in
// it) and carries the reservation, sub-batch count and stop on.
template <int KPER>
__device__ __forceinline__ int window_bounds(const SArgs& S, int win, int* s_p, int* s_state,
                                             SelState& st) {
  const int tid = threadIdx.x, lane = tid & 63;
  if (tid < 64) {
    int end = win, stop = 0, base = 0, cur = st.carry, nbw = st.nb;
    int i = wave_first_at_least(s_p, 0, win, S.pool_m1 - cur, lane);
    while (i < win) {
      ++nbw;
      if (nbw == S.max_batches) {
        stop = 1;
        end = i;
        break;
      }
      base = i > 0 ? s_p[sel_pad(i - 1)] : 0;
      cur = 0;
      i = wave_first_at_least(s_p, i + 1, win, base + S.pool_m1, lane);
    }
    if (lane == 0) {
      const int lastp = s_p[sel_pad(win - 1)];
      s_state[0] = end;
      s_state[1] = stop;
      s_state[2] = cur + (stop ? 0 : lastp - base);  // reservation carried on
      s_state[3] = nbw;
    }
  }
  __syncthreads();
  const int end = s_state[0];
  st.stop = s_state[1];
  st.carry = s_state[2];
  st.nb = s_state[3];
  return end;
}

// window_take: the parents taken in [0, end) (candidates whose prefix steps
// up, w > 0) get their stack slots and first children from the running
// counts; the taken ones' slot and child offsets within the window are a
// block scan, or come precomputed (pre, for end == win)
struct TakeScan {
  int slot, coff, ntaken, nchild;
};
template <int KPER>
__device__ __forceinline__ TakeScan window_scan(int end, int* s_w16, const int (&word)[KPER]) {
  const int tid = threadIdx.x;
  int mine = 0, mc = 0;
#pragma unroll
  for (int k = 0; k < KPER; ++k) {
    const bool t = KPER * tid + k < end && (word[k] & 0xffff) != 0;
    mine += t;
    mc += t ? (word[k] >> 16) : 0;
  }
  TakeScan r;
  block_exscan1024_2(mine, mc, s_w16, &r.slot, &r.coff, &r.ntaken, &r.nchild);
  return r;
}
template <int KPER>
__device__ __forceinline__ void window_write(const SArgs& S, int top, int end, const TakeScan& r,
                                             const SelState& st, const int (&word)[KPER]) {
  const int tid = threadIdx.x;
  int slot = r.slot + st.np, coff = r.coff + st.nc;
#pragma unroll
  for (int k = 0; k < KPER; ++k) {
    const int i = KPER * tid + k;
    if (i < end && (word[k] & 0xffff) != 0) {
      if (slot < S.PCAP) {
        S.p_sel[slot] = top - 1 - (st.consumed + i);
        S.p_off[slot] = coff;
      }
      ++slot;
      coff += word[k] >> 16;
    }
  }
}
__device__ __forceinline__ void window_count(SelState& st, int end, const TakeScan& r) {
  st.np += r.ntaken;
  st.nc += r.nchild;
  st.pruned += end - r.ntaken;
  st.consumed += end;
}

template <int KPER>
__device__ __forceinline__ void window_settle(const SArgs& S, int top, int win, int* s_p,
                                              int* s_w16, int* s_state, SelState& st,
                                              const int (&word)[KPER]) {
  const int end = window_bounds<KPER>(S, win, s_p, s_state, st);
  const TakeScan r = window_scan<KPER>(end, s_w16, word);
  window_write<KPER>(S, top, end, r, st, word);
  window_count(st, end, r);
}

template <int KPER>
__device__ __forceinline__ void select_window(const SArgs& S, int top, int inc, int* s_p,
                                              int* s_w16, int* s_state, SelState& st) {
  const int win = min(KPER * 1024, top - st.consumed);
  int word[KPER];
  window_words<KPER>(S, top, inc, st.consumed, win, s_p, s_w16, word);
  window_settle<KPER>(S, top, win, s_p, s_w16, s_state, st, word);
}

// the end of a select: overflow and budget checks, then the round's
// parents, children and counters (one thread)
__device__ void select_commit(const SArgs& S, Ctrl* C, const SelState& st, int top,
                              long long decomposed0) {
  const int np = st.np, nc = st.nc, pruned = st.pruned, consumed = st.consumed;
  C->n_fb = 0;
  // Overflow is decided here, before anything is committed: the round
  // keeps at most nc children on top of the top - consumed nodes left
  // below its parents, so if those fit, settle and push cannot run out of
  // stack.  An overflowing round leaves the stack, incumbent and counters
  // exactly as they were before it, and a search resumed from that state
  // (with more capacity) visits the nodes an uninterrupted one would.
  if (np > S.PCAP || nc > S.CCAP || static_cast<long long>(top - consumed) + nc > S.CAP) {
    C->status = kOverflow;
    C->n_parents = 0;
    C->n_children = 0;
    return;
  }
  if (S.budget > 0 && np > 0 && decomposed0 + np > S.budget) {
    // stop before this round: the stack keeps every candidate, so a
    // resumed search continues exactly where this one stopped
    C->status = kBudget;
    C->n_parents = 0;
    C->n_children = 0;
    return;
  }
  auto add = [&](int k, long long v) {
    atomicAdd(reinterpret_cast<unsigned long long*>(&C->stats[k]),
              static_cast<unsigned long long>(v));
  };
  add(kStRounds, 1);
  S.p_off[np] = nc;
  if (pruned) add(kStPrunedPop, pruned);
  C->stack_size = top - consumed;
  if (np == 0) {
    C->status = kDone;
    C->n_parents = 0;
    C->n_children = 0;
  } else {
    C->n_parents = np;
    C->n_children = nc;
    add(kStDecomposed, np);
    add(kStBounded, nc);
    if (S.count_pairs == 1) add(kStPairs, static_cast<long long>(nc) * S.P);
    if (S.count_pairs == 2) {
      add(kStPairs, static_cast<long long>(nc) * S.topk);
      add(kStSubset, nc);
    }
  }
}

__global__ __launch_bounds__(1024) void k_select(SArgs S) {
  __shared__ int s_p[kSelWin + kSelWin / 16];  // candidate words, then prefix sums
  __shared__ int s_w16[80];
  __shared__ int s_state[4];
  Ctrl* C = S.ctrl;
  const int tid = threadIdx.x;
  // every control field in one round of loads; the counter updates at the
  // end are atomics whose results are not waited on
  const int status = C->status, top = C->stack_size, inc = C->inc;
  const long long decomposed0 = C->stats[kStDecomposed];
  // a round after the search ended (or with nothing to pop) does nothing;
  // the fallback count is cleared with the others so that the statistics
  // kernel of such a round adds zero
  if (status != kRunning) {
    if (tid == 0) {
      C->n_parents = 0;
      C->n_children = 0;
      C->n_fb = 0;
    }
    return;
  }
  if (top == 0) {
    if (tid == 0) {
      C->status = kDone;
      C->n_parents = 0;
      C->n_children = 0;
      C->n_fb = 0;
    }
    return;
  }
  // block-uniform carried state: reservation of the open sub-batch, closed
  // sub-batches, parents taken, children, pruned candidates, consumed; a
  // small first window for small pools (most of their rounds end in it),
  // then large ones
  SelState st{0, 0, 0, 0, 0, 0, 0};
  // a round reserves about pool_size x sub-batches / (2u) parents: large
  // pools start with the large window
  if (static_cast<long long>(S.pool_m1) * S.max_batches <= 16LL * kSelWin0)
    select_window<kSelWin0 / 1024>(S, top, inc, s_p, s_w16, s_state, st);
  while (!st.stop && top - st.consumed > 0)
    select_window<kSelWin / 1024>(S, top, inc, s_p, s_w16, s_state, st);
  if (tid == 0) select_commit(S, C, st, top, decomposed0);
}

// ---- select for large pools: windows on kSelChain blocks ------------------
// Block t (in the order blocks take tickets) turns window t (kSelWin
// candidates) into words and prefix sums at once, then waits for block
// t - 1's hand-off record (carried reservation, sub-batches, parents,
// children, pruned, consumed, or a stop), settles its window from it and
// hands on.  Every block that takes a ticket writes its hand-off on every
// path (a stop when it has nothing to do), and block t waits only on
// t - 1, which took its ticket earlier: the waits form a chain down to
// ticket 0, which waits on nothing, so every wait ends.  The last block
// continues alone if the round needs more windows; the block where the
// round ends commits it (select_commit).
__global__ __launch_bounds__(1024) void k_select_chain(SArgs S) {
  constexpr int KPER = kSelWin / 1024;
  __shared__ int s_p[kSelWin + kSelWin / 16];
  __shared__ int s_w16[80];
  __shared__ int s_state[4];
  __shared__ int s_ticket;
  __shared__ int s_link[kLink];
  Ctrl* C = S.ctrl;
  const int tid = threadIdx.x;
  const int status = C->status, top = C->stack_size, inc = C->inc;
  const long long decomposed0 = C->stats[kStDecomposed];
  if (tid == 0) s_ticket = atomicAdd(&C->sel_ticket, 1);
  __syncthreads();
  const int t = s_ticket;
  // tickets run 0..kSelChain-1 while expand clears the count after every
  // select; a block beyond them (never the case) has no record to write
  if (t >= kSelChain) return;
  int* mine = S.sel_link + t * kLink;
  auto hand_stop = [&]() {  // thread 0: this block hands on a stop
    mine[1] = 1;
    __hip_atomic_store(mine, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  };
  if (status != kRunning || top == 0) {
    // as k_select: nothing to pop
    if (tid == 0) {
      if (t == 0) {
        if (status == kRunning) C->status = kDone;
        C->n_parents = 0;
        C->n_children = 0;
        C->n_fb = 0;
      }
      hand_stop();
    }
    return;
  }
  const int w0 = t * kSelWin;
  if (w0 >= top) {  // no candidates in this block's window
    if (tid == 0) hand_stop();
    return;
  }
  const int win = min(kSelWin, top - w0);
  int word[KPER];
  window_words<KPER>(S, top, inc, w0, win, s_p, s_w16, word);
  // the window's parents and children as if the round does not stop in it
  // (it stops only in the window where its last sub-batch fills): scanned
  // before the wait, off the chain of hand-offs
  const TakeScan whole = window_scan<KPER>(win, s_w16, word);
  SelState st{0, 0, 0, 0, 0, w0, 0};
  if (t > 0) {
    if (tid == 0) {
      const int* prev = S.sel_link + (t - 1) * kLink;
      while (__hip_atomic_load(prev, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
        __builtin_amdgcn_s_sleep(2);
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      for (int i = 1; i < kLink; ++i) s_link[i] = prev[i];
    }
    __syncthreads();
    if (s_link[1]) {  // the round ended before this window: pass that on
      if (tid == 0) hand_stop();
      return;
    }
    // every earlier window was taken whole, so st.consumed == w0
    st.carry = s_link[2];
    st.nb = s_link[3];
    st.np = s_link[4];
    st.nc = s_link[5];
    st.pruned = s_link[6];
    st.consumed = s_link[7];
  }
  auto hand_on = [&](bool done) {  // thread 0
    mine[1] = done ? 1 : 0;
    mine[2] = st.carry;
    mine[3] = st.nb;
    mine[4] = st.np;
    mine[5] = st.nc;
    mine[6] = st.pruned;
    mine[7] = st.consumed;
    __hip_atomic_store(mine, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  };
  const int end = window_bounds<KPER>(S, win, s_p, s_state, st);
  bool handed = false;
  if (end == win) {
    // the whole window is the round's: the next block gets its record
    // before this one writes its parents (into slots of its own)
    const SelState at = st;
    window_count(st, win, whole);
    if (t != kSelChain - 1) {
      if (tid == 0) hand_on(top - st.consumed <= 0);
      handed = true;
    }
    window_write<KPER>(S, top, win, whole, at, word);
  } else {
    const TakeScan r = window_scan<KPER>(end, s_w16, word);
    window_write<KPER>(S, top, end, r, st, word);
    window_count(st, end, r);
  }
  if (t == kSelChain - 1)
    while (!st.stop && top - st.consumed > 0)
      select_window<KPER>(S, top, inc, s_p, s_w16, s_state, st);
  const bool done = st.stop || top - st.consumed <= 0;
  if (tid == 0) {
    if (!handed) hand_on(done);
    if (done) select_commit(S, C, st, top, decomposed0);
  }
}

// ---- expand: parents out of the stack, child descriptors ------------------
// A thread per parent copies its permutation column (consecutive parents sit
// at consecutive stack slots, so every position row is read coalesced),
// writes its children's descriptors and computes the parent's heads, tails
// and scheduled-job mask once: the bound kernel extends them by the one or
// two jobs each child places instead of rerunning the recurrences over the
// whole prefix and suffix per child.  Written for latency (a round has a few
// thousand parents): one batch of loads per column chunk, no lane exchanges.
// small blocks: a round's parents (a few thousand) spread over many CUs
// instead of filling a few
constexpr int kExpandBlock = 64;

template <int MM, int NW, bool FEAT>
__global__ __launch_bounds__(kExpandBlock) void k_expand(SArgs S, DevTables T, int want_state) {
  extern __shared__ int32_t lds[];
  int32_t* sp = lds;                                          // [n][MM] job-major
  uint8_t* pl = reinterpret_cast<uint8_t*>(lds + MM * T.n);    // [n][block] own columns
  const Ctrl* C = S.ctrl;
  // the grid covers the parent capacity, so a thread has at most one
  // parent: its stack slot and first child are loaded (at a clamped index,
  // so in bounds whatever the round's count) together with the parent count
  // and before the times are staged, and the memory round trips overlap
  const int np = C->n_parents;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int pc = p < S.PCAP ? p : S.PCAP - 1;
  const int si_any = S.p_sel[pc], off_any = S.p_off[pc];
  issue_before_branch();
  // the chained select of the next round starts from cleared hand-offs
  // (this kernel runs after every select and before the next)
  if (blockIdx.x == 0) {
    if (threadIdx.x < kSelChain) S.sel_link[threadIdx.x * kLink] = 0;
    if (threadIdx.x == 0) S.ctrl->sel_ticket = 0;
  }
  // blocks past the round's parents leave before staging
  if (static_cast<int>(blockIdx.x * blockDim.x) >= np) return;
  const bool mine_p = p < np;
  const int si = mine_p ? si_any : 0;
  const int off = mine_p ? off_any : 0;
  stage_times<MM>(T, sp);
  __syncthreads();
  const int n = S.n, m = S.m, P = S.P;
  const int lane = threadIdx.x & 63;
  const size_t PH = static_cast<size_t>(S.PCAP);  // rows of the parent state, as ChildSource reads them
  long long work = 0;
  if (mine_p) {
    const int l1 = S.s_l1[si], l2 = S.s_l2[si], cost = S.s_cost[si];
    // loads of a chunk all issue before its stores (the arrays may alias as
    // far as the compiler knows)
    for (int i0 = 0; i0 < n; i0 += 32) {
      uint32_t v[32];  // one register per byte: every load in flight at once
#pragma unroll
      for (int u = 0; u < 32; ++u)
        if (i0 + u < n) v[u] = S.s_perm[static_cast<size_t>(i0 + u) * S.CAP + si];
#pragma unroll
      for (int u = 0; u < 32; ++u)
        if (i0 + u < n) {
          S.p_perm[static_cast<size_t>(i0 + u) * S.PCAP + p] = static_cast<uint8_t>(v[u]);
          pl[(i0 + u) * kExpandBlock + threadIdx.x] = static_cast<uint8_t>(v[u]);
        }
    }
    S.p_l1[p] = l1;
    S.p_l2[p] = l2;
    S.p_cost[p] = cost;
    const int tag = p << 9, u = l2 - l1 - 1;
    // algorithmic bound work of the children: per child with prefix d1,
    // suffix d2 and u free jobs, P(n + 4u + 4) + 2m d1 + m(m+1) d2 int ops
    const long long d1 = l1 + 1, d2 = n - l2;
    const long long mm1 = static_cast<long long>(m) * (m + 1);
    if (d1 + d2 == n - 2) {  // simple parent: two leaves, BEGIN side
      S.c_desc[off] = tag | (l1 + 1);
      S.c_desc[off + 1] = tag | (l1 + 2);
      work += 2 * (P * (n + 4LL) + 2LL * m * (d1 + 2) + mm1 * d2);
    } else {                 // BEGIN children then END children
      for (int c = 0; c < u; ++c) S.c_desc[off + c] = tag | (l1 + 1 + c);
      for (int c = 0; c < u; ++c) S.c_desc[off + u + c] = tag | (1 << 8) | (l1 + 1 + c);
      const long long base = P * (n + 4 * (u - 1LL) + 4);
      work += u * (base + 2LL * m * (d1 + 1) + mm1 * d2) + u * (base + 2LL * m * d1 + mm1 * (d2 + 1));
    }
    if (want_state) {
      // the parent's heads and tails (no R0/Q0 substitution: that depends on
      // the child's own prefix and suffix), scheduled jobs and their work
      int r[MM], q[MM], wk[MM];
      uint32_t sm[NW];
#pragma unroll
      for (int k = 0; k < MM; ++k) r[k] = q[k] = wk[k] = 0;
#pragma unroll
      for (int w = 0; w < NW; ++w) sm[w] = 0;
      const uint8_t* mine = pl + threadIdx.x;
      int jn = l1 >= 0 ? mine[0] : 0;
      for (int i = 0; i <= l1; ++i) {
        const int j = jn;
        if (i + 1 <= l1) jn = mine[(i + 1) * kExpandBlock];
        set_job_bit<NW>(sm, j);
        int pk[MM];
        job_row<MM>(sp, j, pk);
        head_step<MM>(r, pk);
#pragma unroll
        for (int k = 0; k < MM; ++k) wk[k] += pk[k];
      }
      jn = l2 < n ? mine[(n - 1) * kExpandBlock] : 0;
      for (int i = n - 1; i >= l2; --i) {
        const int j = jn;
        if (i - 1 >= l2) jn = mine[(i - 1) * kExpandBlock];
        set_job_bit<NW>(sm, j);
        int pk[MM];
        job_row<MM>(sp, j, pk);
        tail_step<MM>(q, pk);
#pragma unroll
        for (int k = 0; k < MM; ++k) wk[k] += pk[k];
      }
#pragma unroll
      for (int k = 0; k < MM; ++k) {
        S.p_hr[k * PH + p] = r[k];
        S.p_hq[k * PH + p] = q[k];
        S.p_hw[k * PH + p] = wk[k];
      }
#pragma unroll
      for (int w = 0; w < NW; ++w) S.p_hm[w * PH + p] = sm[w];
      // selector features of the parent (the feature kernel's definition and
      // fragment layout; used with m == MM so every index is a constant): per
      // machine head (or R0), tail (or Q0), remaining work, head + work,
      // work + tail, scaled
      if constexpr (FEAT) {
        int h[MM], tl[MM], rw[MM];
#pragma unroll
        for (int k = 0; k < MM; ++k) {
          h[k] = l1 >= 0 ? r[k] : ld_const(T.R0, k);
          tl[k] = l2 < n ? q[k] : ld_const(T.Q0, k);
          rw[k] = ld_const(T.ptot, k) - wk[k];
        }
        const int b = p >> 5, c = p & 31;
#pragma unroll
        for (int t = 0; t < (5 * MM + 7) / 8; ++t)
#pragma unroll
          for (int half = 0; half < 2; ++half) {
            uint32_t w2[2] = {0u, 0u};
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const int f = 8 * t + 4 * half + i, g = f / MM, k = f % MM;
              if (f < 5 * MM) {
                const int v = g == 0 ? h[k] : g == 1 ? tl[k] : g == 2 ? rw[k] : g == 3 ? h[k] + rw[k] : rw[k] + tl[k];
                const float x = static_cast<float>(v) * S.x_inv_scale;
                if (S.x_f32) S.x_f32[static_cast<size_t>(f) * S.x_ld + p] = x;
                uint32_t u = __float_as_uint(x);
                u += 0x7FFFu + ((u >> 16) & 1u);  // BF16, round to nearest even
                w2[i >> 1] |= (u >> 16) << (16 * (i & 1));
              }
            }
            if (S.x_frag)
              S.x_frag[(static_cast<size_t>(b) * S.x_fsteps + t) * 64 + half * 32 + c] =
                  make_uint2(w2[0], w2[1]);
          }
      }
    }
  }
  for (int d = 32; d >= 1; d >>= 1) work += __shfl_xor(work, d, 64);
  if (lane == 0 && work)
    atomicAdd(reinterpret_cast<unsigned long long*>(&S.ctrl->stats[kStWork]),
              static_cast<unsigned long long>(work));
}

hipError_t launch_expand(const wsbb_ctx* ctx, const SArgs& A, int want_state, hipStream_t st) {
  const int m = ctx->T.m, n = ctx->T.n;
  const int mm = m <= 8 ? 8 : (m <= 12 ? 12 : (m <= 16 ? 16 : (m <= 20 ? 20 : 24)));
  const int nw = n <= 32 ? 1 : (n <= 64 ? 2 : 4);
  const dim3 grid((A.PCAP + kExpandBlock - 1) / kExpandBlock), block(kExpandBlock);
  const size_t shmem = sizeof(int32_t) * mm * n + static_cast<size_t>(n) * kExpandBlock;
  // feature output only where the constant-index form applies (m == MM)
  const bool feat = (A.x_frag != nullptr || A.x_f32 != nullptr);
  if (feat && m != mm) return hipErrorInvalidValue;
#define WSBB_EXPAND(MMv, NWv)                                                                \
  if (mm == MMv && nw == NWv) {                                                              \
    if (feat)                                                                                \
      hipLaunchKernelGGL((k_expand<MMv, NWv, true>), grid, block, shmem, st, A, ctx->T,      \
                         want_state);                                                        \
    else                                                                                     \
      hipLaunchKernelGGL((k_expand<MMv, NWv, false>), grid, block, shmem, st, A, ctx->T,     \
                         want_state);                                                        \
    return hipGetLastError();                                                                \
  }
  WSBB_EXPAND(8, 1) WSBB_EXPAND(8, 2) WSBB_EXPAND(8, 4)
  WSBB_EXPAND(12, 1) WSBB_EXPAND(12, 2) WSBB_EXPAND(12, 4)
  WSBB_EXPAND(16, 1) WSBB_EXPAND(16, 2) WSBB_EXPAND(16, 4)
  WSBB_EXPAND(20, 1) WSBB_EXPAND(20, 2) WSBB_EXPAND(20, 4)
  WSBB_EXPAND(24, 1) WSBB_EXPAND(24, 2) WSBB_EXPAND(24, 4)
#undef WSBB_EXPAND
  return hipErrorInvalidValue;
}

// ---- settle: incumbent per parent, kept children, survivor offsets --------
// One 1024-thread block.  Parents in chunks of 16384 (thread t owns 16 consecutive
// parents, read and written with 16-byte vector accesses; a round of a
// 65536-child pool fits one chunk).  The incumbent a parent's children are
// pruned against is the running minimum of the leaves of the parents popped
// before it (solutions::insert strictly-better rule, solutions.cpp:45-65;
// prune loop gpu.cu:317-394).  Its kept children are the ranks whose bound is
// below that incumbent: counted by the rank kernel against the round's
// starting incumbent, searched again (in the parent's ranked bounds) only for
// parents after an improvement.  Prefix sums of the kept counts place the
// survivors on the stack.
template <int K>
__device__ __forceinline__ void ld16(const int* a, int (&v)[K]) {
  const int4* q = reinterpret_cast<const int4*>(a);
#pragma unroll
  for (int i = 0; i < K / 4; ++i) {
    const int4 x = q[i];
    v[4 * i] = x.x;
    v[4 * i + 1] = x.y;
    v[4 * i + 2] = x.z;
    v[4 * i + 3] = x.w;
  }
}
template <int K>
__device__ __forceinline__ void st16(int* a, const int (&v)[K]) {
  int4* q = reinterpret_cast<int4*>(a);
#pragma unroll
  for (int i = 0; i < K / 4; ++i) q[i] = make_int4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
}

// The chained settle's records: block t publishes, in S.set_link[t][f..f+2],
// the flag and two values of phase f (0: its chunk's leaf minimum; 4: its
// kept count and last improving parent).  chain_wait has wave 0 of block t
// wait for phase f of every block before it (lane s for block s; all of them
// took their tickets earlier and wait only on blocks before themselves, so
// every wait ends) and combine their values: the minimum of the first value
// (MIN) or their sum, with the maximum of the second, into s_pre[0..1].
__device__ __forceinline__ void chain_put(const SArgs& S, int t, int f, int a, int b) {
  int* rec = S.set_link + t * kLink;
  rec[f + 1] = a;
  rec[f + 2] = b;
  __hip_atomic_store(rec + f, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
template <bool MIN>
__device__ __forceinline__ void chain_wait(const SArgs& S, int t, int f, int* s_pre) {
  const int lane = threadIdx.x;
  if (lane < 64) {
    if (lane < t) {
      const int* flag = S.set_link + lane * kLink + f;
      while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
        __builtin_amdgcn_s_sleep(2);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    int a = MIN ? INT_MAX : 0, b = -1;
    if (lane < t) {
      a = S.set_link[lane * kLink + f + 1];
      b = S.set_link[lane * kLink + f + 2];
    }
    a = MIN ? wave_reduce(a, INT_MAX, OpMin()) : wave_reduce(a, 0, OpAdd());
    b = wave_reduce(b, -1, OpMax());
    if (lane == 0) {
      s_pre[0] = a;
      s_pre[1] = b;
    }
  }
  __syncthreads();
}

// np > 0 parents; s_w16 holds 40 ints, s_last one, s_pre two.  One block
// settles every chunk in order (CHAIN false), or block t of the chained
// settle settles chunk t, taking the incumbent and the survivor offset its
// chunk starts from out of the records of the blocks before it; the block
// of the last chunk commits the round.
template <bool CHAIN>
__device__ void settle_round(const SArgs& S, int np, int t, int* s_w16, int* s_last, int* s_pre) {
  constexpr int K = CHAIN ? kChainK : kSettleK;  // parents per thread
  constexpr int CHUNK = 1024 * K;
  Ctrl* C = S.ctrl;
  const int n = S.n, tid = threadIdx.x;
  const int inc0 = C->inc;
  const int sbase = C->stack_size;  // read with the other control fields
  const bool final = !CHAIN || t == (np - 1) / CHUNK;
  if (tid == 0) *s_last = -1;
  int run_min = inc0, run_kept = 0;  // block-uniform carries over chunks
  int pre_last = -1;                 // last improving parent of earlier blocks
  int improved = 0, last = -1;
  int pruned = 0;  // per thread: at most K parents' children per chunk, few chunks
  // the rank blocks' leaf counts, four per thread (added once per round)
  int leaves = 0;
  if (final) {
    const int4 q = reinterpret_cast<const int4*>(S.r_leaves)[tid];
    leaves = q.x + q.y + q.z + q.w;
  }
  const int c_first = CHAIN ? t * CHUNK : 0;
  const int c_stop = CHAIN ? min(np, c_first + CHUNK) : np;
  for (int c0 = c_first; c0 < c_stop; c0 += CHUNK) {
    const int base = c0 + K * tid;  // arrays hold PCAP + 1 + K entries
    const bool any = base < np;
    int lmin[K], kept[K], k0[K], nc[K];
    uint32_t inner = 0;  // bit k: parent base + k has more than two leaves' depth left
    if (any) {
      int a1[K], a2[K];
      ld16(S.p_l1 + base, a1);
      ld16(S.p_l2 + base, a2);
#pragma unroll
      for (int k = 0; k < K; ++k) inner |= static_cast<uint32_t>(((a1[k] + 1) + (n - a2[k])) != n - 2) << k;
      ld16(S.p_leafmin + base, lmin);
      ld16(S.p_kept0 + base, k0);
      ld16(S.p_nch + base, nc);
    }
    int tmin = INT_MAX;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (!(base + k < np)) lmin[k] = INT_MAX;
      tmin = min(tmin, lmin[k]);
    }
    int cmin;
    const int xmin = block_exscan1024(tmin, INT_MAX, OpMin(), s_w16, &cmin);
    if constexpr (CHAIN) {
      if (tid == 0) chain_put(S, t, 0, cmin, 0);
      chain_wait<true>(S, t, 0, s_pre);
      run_min = min(inc0, s_pre[0]);
    }
    int ib = min(run_min, xmin);
    int tkept = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int p = base + k;
      kept[k] = 0;
      if (p < np) {
        if (lmin[k] < ib) {
          // two leaves, inserted worst first: the first may improve, then
          // the second only if strictly better still
          const int lmax = S.p_leafmax[p];
          improved += (lmax < ib) + (lmin[k] < min(ib, lmax));
          last = p;
        }
        if ((inner >> k) & 1u) {
          int kp;
          if (ib == inc0) {
            kp = k0[k];
          } else {
            const int* rc = S.r_cost + static_cast<size_t>(p) * n;
            int lo = 0, hi = nc[k];  // first rank whose bound is >= ib
            while (lo < hi) {
              const int mid = (lo + hi) >> 1;
              if (rc[mid] < ib) lo = mid + 1;
              else hi = mid;
            }
            kp = lo;
          }
          kept[k] = kp;
          pruned += nc[k] - kp;
        }
        tkept += kept[k];
      }
      ib = min(ib, lmin[k]);
    }
    if (last >= 0) atomicMax(s_last, last);  // final after the scan's barriers
    int csum;
    const int xsum = block_exscan1024(tkept, s_w16, &csum);
    if constexpr (CHAIN) {
      if (tid == 0) chain_put(S, t, 4, csum, *s_last);
      chain_wait<false>(S, t, 4, s_pre);
      run_kept = s_pre[0];
      pre_last = s_pre[1];
    }
    int koff = run_kept + xsum;
    if (any) {
      st16(S.p_kept + base, kept);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int x = kept[k];
        kept[k] = koff;  // becomes the survivor offset
        koff += x;
      }
      st16(S.p_koff + base, kept);
    }
    run_min = min(run_min, cmin);
    run_kept += csum;
  }
  for (int d = 32; d >= 1; d >>= 1) {
    improved += __shfl_xor(improved, d, 64);
    pruned += __shfl_xor(pruned, d, 64);
    leaves += __shfl_xor(leaves, d, 64);
  }
  if ((tid & 63) == 0) {
    if (improved)
      atomicAdd(reinterpret_cast<unsigned long long*>(&C->stats[kStImproved]),
                static_cast<unsigned long long>(improved));
    if (pruned)
      atomicAdd(reinterpret_cast<unsigned long long*>(&C->stats[kStPrunedIns]),
                static_cast<unsigned long long>(pruned));
    if (leaves)
      atomicAdd(reinterpret_cast<unsigned long long*>(&C->stats[kStLeaves]),
                static_cast<unsigned long long>(leaves));
  }
  if (!final) return;  // the block of the last chunk commits the round
  __syncthreads();
  const int bl = max(pre_last, *s_last);
  if (tid == 0) {
    if (bl >= 0) C->inc = run_min;
    C->push_base = sbase;
    C->push_total = run_kept;
    // fits: select admitted the round only with room for all its children
    C->stack_size = sbase + run_kept;
    atomicMax(reinterpret_cast<unsigned long long*>(&C->stats[kStMaxPool]),
              static_cast<unsigned long long>(sbase + run_kept));
  }
  // the leaf that set the incumbent: parent permutation with its swap.  It
  // is a child of leaf parent bl, whose children are numbered from p_off on
  // the BEGIN side (child i places the job at position l1 + 1 + i, as expand
  // wrote them), so the winner's fields come in one round of loads
  if (bl >= 0 && tid < n) {
    const int bc = S.p_leafwin[bl], boff = S.p_off[bl], bl1 = S.p_l1[bl];
    const int p = bl, sw0 = bl1 + 1, j = bl1 + 1 + (bc - boff);
    const int i = tid;
    const int src = (i == sw0) ? j : ((i == j) ? sw0 : i);
    C->best_perm[i] = S.p_perm[static_cast<size_t>(src) * S.PCAP + p];
    if (i == 0) C->have_perm = 1;
  }
}

// ---- rank: one wave per parent -------------------------------------------
// one DPP step of the leaf minimum: cost ascending, then job descending
template <int CTRL, int RMASK>
__device__ __forceinline__ void leaf_step(int& lmin, int& lwin, int& lwjob) {
  const int om = __builtin_amdgcn_update_dpp(INT_MAX, lmin, CTRL, RMASK, 0xf, false);
  const int oj = __builtin_amdgcn_update_dpp(-1, lwjob, CTRL, RMASK, 0xf, false);
  const int ow = __builtin_amdgcn_update_dpp(-1, lwin, CTRL, RMASK, 0xf, false);
  if (om < lmin || (om == lmin && oj > lwjob)) {
    lmin = om;
    lwin = ow;
    lwjob = oj;
  }
}

// Four parents per block and pass.  A parent's children and the jobs at its
// free positions are read in one round of loads (its child count is 2u, or 2
// at depth n - 2, as expand numbered them).  The chosen side's ranks go to
// the parent's ranked rows (child index and bound by rank), which settle and
// push read without the child descriptors.
constexpr int kRankWaves = 4;

__global__ __launch_bounds__(256) void k_rank(SArgs S) {
  __shared__ int s_b[kRankWaves][2 * kMaxJobs];  // children's bounds, BEGIN then END
  __shared__ int s_job[kRankWaves][kMaxJobs];    // job at free position l1 + 1 + i
  Ctrl* C = S.ctrl;
  const int np = C->n_parents;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int n = S.n;
  const int inc0 = C->inc;
  // leaf children of the block's parents: one plain store per block (the
  // settle kernel sums the blocks' counts); an atomic per wave into one
  // counter serialises when a round is all leaf parents (ten thousand waves)
  __shared__ int s_leaves;
  if (threadIdx.x == 0) s_leaves = 0;
  // the chained settle that follows starts from cleared records
  if (blockIdx.x == 0) {
    if (threadIdx.x < kSettleChain * kLink) S.set_link[threadIdx.x] = 0;
    if (threadIdx.x == 0) C->set_ticket = 0;
  }
  for (int pb = blockIdx.x * kRankWaves; pb < np; pb += gridDim.x * kRankWaves) {
    const int p = pb + wave;
    const bool live = p < np;
    int off = 0, l1 = 0, u = 0, simple = 0;
    if (live) {
      off = S.p_off[p];
      l1 = S.p_l1[p];
      const int l2 = S.p_l2[p];
      u = l2 - l1 - 1;
      simple = ((l1 + 1) + (n - l2)) == n - 2;
      const int cnt = simple ? 2 : 2 * u;
      for (int c = lane; c < cnt; c += 64) s_b[wave][c] = S.c_bound[off + c];
      for (int i = lane; i < u; i += 64)
        s_job[wave][i] = S.p_perm[static_cast<size_t>(l1 + 1 + i) * S.PCAP + p];
    }
    __syncthreads();
    if (live) {
      const int cnt = simple ? 2 : 2 * u;
      int c0 = 0, c1 = 0, s0 = 0, s1 = 0;
      for (int c = lane; c < cnt; c += 64) {
        const int b = s_b[wave][c];
        const int sd = simple ? 0 : (c >= u);
        const int big = b >= kSideBest;
        if (sd == 0) {
          c0 += big;
          s0 += big ? 0 : b;
        } else {
          c1 += big;
          s1 += big ? 0 : b;
        }
      }
      c0 = wave_reduce(c0, 0, OpAdd());
      c1 = wave_reduce(c1, 0, OpAdd());
      s0 = wave_reduce(s0, 0, OpAdd());
      s1 = wave_reduce(s1, 0, OpAdd());
      const int side = (c0 > c1) ? 0 : (c1 > c0) ? 1 : (s0 > s1) ? 0 : 1;
      const int nch = simple ? (side == 0 ? 2 : 0) : u;
      // the i-th child of either side (and of a simple parent) places the
      // job at position l1 + 1 + i
      const int* sb = &s_b[wave][simple ? 0 : side * u];
      int* ro = S.r_order + static_cast<size_t>(p) * n;
      int* rc = S.r_cost + static_cast<size_t>(p) * n;
      int lmin = INT_MAX, lwin = -1, lwjob = -1, lmax = INT_MIN, below = 0;
      for (int i = lane; i < nch; i += 64) {
        const int cc = sb[i], jj = s_job[wave][i];
        int rk = 0;
        for (int k = 0; k < nch; ++k) {
          const int ck = sb[k];
          rk += (ck < cc) || (ck == cc && s_job[wave][k] < jj);
        }
        ro[rk] = i;
        rc[rk] = cc;
        below += cc < inc0;
        if (simple) {
          // leaves are inserted worst first (cost desc, job desc) and only a
          // strictly better one replaces the incumbent: among equal minima
          // the one with the larger job id wins
          if (cc < lmin || (cc == lmin && jj > lwjob)) {
            lmin = cc;
            lwin = off + i;
            lwjob = jj;
          }
          lmax = max(lmax, cc);
        }
      }
      if (simple) {
        // scan of (cost ascending, job descending) with the child index as
        // payload, same DPP steps as wave_incl_scan
        leaf_step<0x111, 0xf>(lmin, lwin, lwjob);
        leaf_step<0x112, 0xf>(lmin, lwin, lwjob);
        leaf_step<0x114, 0xf>(lmin, lwin, lwjob);
        leaf_step<0x118, 0xf>(lmin, lwin, lwjob);
        leaf_step<0x142, 0xa>(lmin, lwin, lwjob);
        leaf_step<0x143, 0xc>(lmin, lwin, lwjob);
        lmin = __builtin_amdgcn_readlane(lmin, 63);
        lwin = __builtin_amdgcn_readlane(lwin, 63);
        lwjob = __builtin_amdgcn_readlane(lwjob, 63);
        lmax = wave_reduce(lmax, INT_MIN, OpMax());
      }
      below = wave_reduce(below, 0, OpAdd());
      if (lane == 0) {
        S.p_kept0[p] = below;
        S.p_side[p] = side;
        S.p_nch[p] = nch;
        S.p_leafmin[p] = simple ? lmin : INT_MAX;
        S.p_leafmax[p] = (simple && nch > 0) ? lmax : INT_MAX;
        S.p_leafwin[p] = simple ? lwin : -1;
        if (simple && nch > 0) atomicAdd(&s_leaves, nch);
      }
    }
    __syncthreads();
  }
  __syncthreads();
  if (threadIdx.x == 0) S.r_leaves[blockIdx.x] = s_leaves;
}

__global__ __launch_bounds__(1024) void k_settle(SArgs S) {
  __shared__ int s_w16[40];
  __shared__ int s_last, s_pre[2];
  const int np = S.ctrl->n_parents;
  if (np == 0) return;
  settle_round<false>(S, np, 0, s_w16, &s_last, s_pre);
}

// The chained settle for large pools: kSettleChain blocks, block t (in the
// order blocks take tickets) settles parents [t, t + 1) kChainChunk.  Its
// chunk's leaf minimum and, once its incumbents are known, its kept count
// are published at once and combined with those of the blocks before it
// (chain_wait), so the chunks run side by side and only wait for two
// records each.  A block whose chunk starts past the round's parents leaves
// at once: only blocks after it would wait for it, and they have nothing
// to settle either.
__global__ __launch_bounds__(1024) void k_settle_chain(SArgs S) {
  __shared__ int s_w16[40];
  __shared__ int s_last, s_pre[2], s_ticket;
  const int np = S.ctrl->n_parents;
  if (threadIdx.x == 0) s_ticket = atomicAdd(&S.ctrl->set_ticket, 1);
  __syncthreads();
  const int t = s_ticket;
  // tickets run 0..kSettleChain-1 (the rank kernel clears the count and the
  // records before every settle)
  if (t >= kSettleChain || np == 0 || t * kChainChunk >= np) return;
  settle_round<true>(S, np, t, s_w16, &s_last, s_pre);
}

// ---- push survivors: one wave per parent ------------------------------------
// A parent's kept children are its first `kept` ranks; the rank-k child goes
// to stack slot base + total - 1 - (koff + k), so the round's first parent's
// best child ends on top.  A parent's fields, its permutation column and its
// ranked rows are read in one round of loads (rows are read whole, n entries,
// whatever the kept count); lane k then writes the rank-k child.
__global__ __launch_bounds__(256) void k_push(SArgs S) {
  __shared__ int s_col[4][kMaxJobs];
  const Ctrl* C = S.ctrl;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int n = S.n;
  const int stride = gridDim.x * 4;
  // a parent's loads depend only on its index: the first pass's (at an
  // index clamped to the capacity) go out with the control loads
  bool first = true;
  int kept = 0, koff = 0, side = 0, pl1 = 0, pl2 = 0;
  // positions / ranks lane and lane + 64 (n <= 128)
  int ord[2] = {0, 0}, cst[2] = {0, 0};
  uint32_t col[2] = {0u, 0u};  // a register per byte: no wait before the control loads
  auto load = [&](int p) {
    kept = S.p_kept[p];
    koff = S.p_koff[p];
    side = S.p_side[p];
    pl1 = S.p_l1[p];
    pl2 = S.p_l2[p];
    const int* ro = S.r_order + static_cast<size_t>(p) * n;
    const int* rc = S.r_cost + static_cast<size_t>(p) * n;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int k = lane + 64 * t;
      if (k < n) {
        col[t] = S.p_perm[static_cast<size_t>(k) * S.PCAP + p];
        ord[t] = ro[k];
        cst[t] = rc[k];
      }
    }
  };
  // after an overflow or at the end of the search select left no parents
  const int np = C->n_parents, base = C->push_base, total = C->push_total;
  {
    const int p0 = blockIdx.x * 4 + wave;
    load(p0 < S.PCAP ? p0 : S.PCAP - 1);
  }
  issue_before_branch();
  for (int pb = blockIdx.x * 4; pb < np; pb += stride) {
    const int p = pb + wave;
    const bool live = p < np;
    if (!first && live) load(p);
    first = false;
    if (live) {
#pragma unroll
      for (int t = 0; t < 2; ++t)
        if (lane + 64 * t < n) s_col[wave][lane + 64 * t] = static_cast<int>(col[t]);
    }
    __syncthreads();
    if (live) {
      const int sw0 = side == 0 ? pl1 + 1 : pl2 - 1;
      int l1 = pl1 + (side == 0), l2 = pl2 - (side == 1);
      if (l2 - l1 == 2) l1 += 1;  // the last free job is placed too (limites_set)
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int k = lane + 64 * t;
        if (k < kept) {
          const int j = pl1 + 1 + ord[t];
          const int slot = base + total - 1 - (koff + k);
          for (int i = 0; i < n; ++i) {
            const int src = (i == sw0) ? j : ((i == j) ? sw0 : i);
            S.s_perm[static_cast<size_t>(i) * S.CAP + slot] = static_cast<uint8_t>(s_col[wave][src]);
          }
          S.s_l1[slot] = l1;
          S.s_l2[slot] = l2;
          S.s_cost[slot] = cst[t];
          S.s_key[slot] = select_key(l1, l2, cst[t], n);
        }
      }
    }
    __syncthreads();
  }
}

// Children whose learned-subset bound lies in [incumbent - margin,
// incumbent) and that are not leaves are listed for full LB2 (the order of
// the list does not matter: every listed child gets its own exact value).
__global__ __launch_bounds__(256) void k_fallback_list(SArgs S, int margin, int* list) {
  Ctrl* C = S.ctrl;
  const int nc = C->n_children, inc = C->inc;
  const int lane = threadIdx.x & 63;
  for (int c0 = blockIdx.x * blockDim.x; c0 < nc; c0 += gridDim.x * blockDim.x) {
    const int c = c0 + threadIdx.x;
    bool take = false;
    if (c < nc) {
      const int b = S.c_bound[c];
      if (b < inc && b >= inc - margin) {
        const int p = S.c_desc[c] >> 9;
        take = S.p_l2[p] - S.p_l1[p] - 1 > 2;  // parents with u <= 2 have leaf children
      }
    }
    const unsigned long long bal = __ballot(take);
    if (bal == 0) continue;
    int base = 0;
    if (lane == 0) base = atomicAdd(&C->n_fb, __popcll(bal));
    base = __shfl(base, 0, 64);
    if (take) list[base + __popcll(bal & ((1ull << lane) - 1ull))] = c;
  }
}

__global__ void k_fallback_stats(SArgs S) {
  Ctrl* C = S.ctrl;
  C->stats[kStFallback] += C->n_fb;
  C->stats[kStPairs] += static_cast<long long>(C->n_fb) * S.P;
}

// Records of the plain solver for the pair selector: every stride-th child
// of the round (features and LB2-maximising pair) at a slot fixed by its
// index, so the log does not depend on thread timing.  With desc set the
// features are the child's parent's (column desc[c] >> 9).
struct LogArgs {
  const Ctrl* ctrl;
  const float* feat;  // [F][ld]
  const int* label;   // [children]
  const int* desc;    // child descriptors, or null
  int ld, F, stride, cap;
  float* log_feat;    // [F][cap]
  int* log_label;     // [cap]
  int* log_count;
};

__global__ __launch_bounds__(256) void k_log(LogArgs L) {
  const int nc = L.ctrl->n_children;
  const int base = *L.log_count;
  const int nrec = (nc + L.stride - 1) / L.stride;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nrec; r += gridDim.x * blockDim.x) {
    const int slot = base + r;
    if (slot >= L.cap) break;
    const int c = r * L.stride;
    const int col = L.desc ? (L.desc[c] >> 9) : c;
    for (int f = 0; f < L.F; ++f)
      L.log_feat[static_cast<size_t>(f) * L.cap + slot] = L.feat[static_cast<size_t>(f) * L.ld + col];
    L.log_label[slot] = L.label[c];
  }
}

__global__ void k_log_advance(LogArgs L) {
  const int nc = L.ctrl->n_children;
  const int nrec = (nc + L.stride - 1) / L.stride;
  *L.log_count = min(L.cap, *L.log_count + nrec);
}


}  // namespace

struct SearchState {
  int n = 0, CAP = 0, PCAP = 0, CCAP = 0;
  DevBuf mem;
  SArgs args{};
  Ctrl* host_ctrl = nullptr;  // pinned, two copies: batches of rounds alternate
  hipEvent_t copied[2] = {nullptr, nullptr};  // control copy of each batch landed
  // per-round selector buffers: logged features [5m][CCAP] and labels, BF16
  // feature fragments, tile masks, the fallback list
  DevBuf aux;
  float* r_feat = nullptr;
  int* r_label = nullptr;
  uint32_t* r_tsel = nullptr;
  uint2* r_xfrag = nullptr;
  int* r_fbl = nullptr;
  uint16_t* r_prob = nullptr;  // per-parent BF16 probabilities [PCAP][ldp]
  // rounds_per_sync rounds captured once per solve and replayed
  hipGraphExec_t graph = nullptr;
  // optional timing of every bound launch (pairs of events per round)
  std::vector<hipEvent_t> ev;
  double bound_ms = 0;
  long long bound_launches = 0;
  // the open session: round arguments and search kind of later rounds
  bool open = false;
  int kind = WSBB_LB2;
  SArgs sess{};
  DevBuf tmp;  // staging for moving pool nodes down (wsbb_pool_pop from the bottom)
};

static int ensure_aux(wsbb_ctx* ctx) {
  SearchState* S = ctx->search;
  if (S->r_feat) return WSBB_OK;
  const int F = 5 * ctx->T.m, nsw = (ctx->T.P + 31) / 32, fsteps = (F + 7) / 8;
  const size_t ntiles = (static_cast<size_t>(S->CCAP) + 63) / 64;
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t bf = al(sizeof(float) * F * static_cast<size_t>(S->CCAP));
  const size_t bl = al(sizeof(int) * static_cast<size_t>(S->CCAP));
  const size_t bt = al(sizeof(uint32_t) * ntiles * nsw);
  const size_t bx = al(sizeof(uint2) * 2 * ntiles * fsteps * 64);
  const size_t bp = al(sizeof(uint16_t) * static_cast<size_t>(S->PCAP) * 32 * ((ctx->T.P + 31) / 32));
  hipError_t e = S->aux.reserve(bf + 2 * bl + bt + bx + bp);
  if (e != hipSuccess) {
    hip_fail(e, "hipMalloc(selector round buffers)");
    return WSBB_ENOMEM;
  }
  S->r_feat = S->aux.as<float>(0);
  S->r_label = S->aux.as<int>(bf);
  S->r_fbl = S->aux.as<int>(bf + bl);
  S->r_tsel = S->aux.as<uint32_t>(bf + 2 * bl);
  S->r_xfrag = S->aux.as<uint2>(bf + 2 * bl + bt);
  S->r_prob = S->aux.as<uint16_t>(bf + 2 * bl + bt + bx);
  return WSBB_OK;
}

void destroy_search(wsbb_ctx* ctx) {
  if (!ctx->search) return;
  for (hipEvent_t e : ctx->search->ev) (void)hipEventDestroy(e);
  if (ctx->search->graph) (void)hipGraphExecDestroy(ctx->search->graph);
  ctx->search->tmp.release();
  ctx->search->aux.release();
  ctx->search->mem.release();
  if (ctx->search->host_ctrl) (void)hipHostFree(ctx->search->host_ctrl);
  for (hipEvent_t& e : ctx->search->copied)
    if (e) (void)hipEventDestroy(e);
  delete ctx->search;
  ctx->search = nullptr;
}

static int ensure_search(wsbb_ctx* ctx, int pool_size, int max_batches, int cap_nodes) {
  const int n = ctx->T.n;
  const int PCAP = max_batches * ((pool_size - 2) / 4 + 1);
  const int CCAP = max_batches * pool_size;
  SearchState* S = ctx->search;
  if (S && S->PCAP >= PCAP && S->CCAP >= CCAP && S->CAP == cap_nodes) return WSBB_OK;
  destroy_search(ctx);
  S = new SearchState();
  ctx->search = S;
  S->n = n;
  S->CAP = cap_nodes;
  S->PCAP = PCAP;
  S->CCAP = CCAP;
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += al(bytes);
    return o;
  };
  const size_t o_ctrl = take(sizeof(Ctrl));
  const size_t o_sperm = take(size_t(n) * cap_nodes);
  const size_t o_sl1 = take(4 * size_t(cap_nodes));
  const size_t o_sl2 = take(4 * size_t(cap_nodes));
  const size_t o_scost = take(4 * size_t(cap_nodes));
  const size_t o_skey = take(4 * size_t(cap_nodes));
  const size_t o_link = take(4 * size_t(kSelChain) * kLink);
  const size_t o_slink = take(4 * size_t(kSettleChain) * kLink);
  const size_t o_pperm = take(size_t(n) * PCAP);
  // parent arrays: PCAP + 1 entries plus room for the settle kernel's
  // 16-wide vector accesses past the last parent
  size_t o_pint[13];
  for (int i = 0; i < 13; ++i) o_pint[i] = take(4 * size_t(PCAP + 1 + kSettleK));
  size_t o_cint[2];
  for (int i = 0; i < 2; ++i) o_cint[i] = take(4 * size_t(CCAP));
  // ranked rows: a parent has at most n chosen-side children
  const size_t o_rorder = take(4 * size_t(PCAP) * n);
  const size_t o_rcost = take(4 * size_t(PCAP) * n);
  const size_t o_rleaves = take(4 * size_t(kRankGrid));
  // parent state rows: padded machine count, job-mask words
  const int m = ctx->T.m;
  const int mmp = m <= 8 ? 8 : (m <= 12 ? 12 : (m <= 16 ? 16 : (m <= 20 ? 20 : 24)));
  const int nwp = n <= 32 ? 1 : (n <= 64 ? 2 : 4);
  const size_t o_hr = take(4 * size_t(mmp) * (PCAP + 1));
  const size_t o_hq = take(4 * size_t(mmp) * (PCAP + 1));
  const size_t o_hm = take(4 * size_t(nwp) * (PCAP + 1));
  const size_t o_hw = take(4 * size_t(mmp) * (PCAP + 1));
  hipError_t e = S->mem.reserve(off);
  if (e != hipSuccess) {
    destroy_search(ctx);
    hip_fail(e, "hipMalloc(search pool)");
    return WSBB_ENOMEM;
  }
  e = hipHostMalloc(reinterpret_cast<void**>(&S->host_ctrl), 2 * sizeof(Ctrl), hipHostMallocDefault);
  for (int i = 0; i < 2 && e == hipSuccess; ++i)
    e = hipEventCreateWithFlags(&S->copied[i], hipEventDisableTiming);
  if (e != hipSuccess) {
    destroy_search(ctx);
    return hip_fail(e, "hipHostMalloc(ctrl)");
  }
  SArgs& A = S->args;
  A.n = n;
  A.m = ctx->T.m;
  A.P = ctx->T.P;
  A.CAP = cap_nodes;
  A.PCAP = PCAP;
  A.CCAP = CCAP;
  A.ctrl = S->mem.as<Ctrl>(o_ctrl);
  A.s_perm = S->mem.as<uint8_t>(o_sperm);
  A.s_l1 = S->mem.as<int>(o_sl1);
  A.s_l2 = S->mem.as<int>(o_sl2);
  A.s_cost = S->mem.as<int>(o_scost);
  A.s_key = S->mem.as<int>(o_skey);
  A.sel_link = S->mem.as<int>(o_link);
  A.set_link = S->mem.as<int>(o_slink);
  // the hand-off flags start cleared (afterwards expand and rank clear them)
  e = hipMemset(A.sel_link, 0, 4 * size_t(kSelChain) * kLink);
  if (e == hipSuccess) e = hipMemset(A.set_link, 0, 4 * size_t(kSettleChain) * kLink);
  if (e != hipSuccess) {
    destroy_search(ctx);
    return hip_fail(e, "hipMemset(select hand-off)");
  }
  A.p_perm = S->mem.as<uint8_t>(o_pperm);
  int** pints[13] = {&A.p_l1,      &A.p_l2,      &A.p_cost,    &A.p_sel,
                     &A.p_off,     &A.p_side,    &A.p_nch,     &A.p_leafmin, &A.p_leafmax,
                     &A.p_leafwin, &A.p_kept0,   &A.p_kept,    &A.p_koff};
  for (int i = 0; i < 13; ++i) *pints[i] = S->mem.as<int>(o_pint[i]);
  A.c_desc = S->mem.as<int>(o_cint[0]);
  A.c_bound = S->mem.as<int>(o_cint[1]);
  A.r_order = S->mem.as<int>(o_rorder);
  A.r_cost = S->mem.as<int>(o_rcost);
  A.r_leaves = S->mem.as<int>(o_rleaves);
  // rank blocks past the grid never write their count
  e = hipMemset(A.r_leaves, 0, 4 * size_t(kRankGrid));
  if (e != hipSuccess) {
    destroy_search(ctx);
    return hip_fail(e, "hipMemset(rank counts)");
  }
  A.p_hr = S->mem.as<int>(o_hr);
  A.p_hq = S->mem.as<int>(o_hq);
  A.p_hm = S->mem.as<uint32_t>(o_hm);
  A.p_hw = S->mem.as<int>(o_hw);
  return WSBB_OK;
}

static hipError_t launch_round(wsbb_ctx* ctx, const SArgs& A, int kind, hipEvent_t ev0,
                               hipEvent_t ev1) {
  hipStream_t st = ctx->stream;
  const int rblocks = (A.PCAP + kRankWaves - 1) / kRankWaves;
  const int pblocks = (A.PCAP + 3) / 4;
  // large pools: the chained select over several CUs
  if (static_cast<long long>(A.pool_m1) * A.max_batches > 16LL * kSelWin0)
    hipLaunchKernelGGL(k_select_chain, dim3(kSelChain), dim3(1024), 0, st, A);
  else
    hipLaunchKernelGGL(k_select, dim3(1), dim3(1024), 0, st, A);
  // parent state for the int32 bound kernel (learned subsets, logging,
  // rounds below the two-node form's batch size); the two-node form of plain
  // LB2 runs its own recurrences
  const bool pk_round = kind == WSBB_LB2 && selector_log_stride(ctx) == 0 &&
                        ctx->T.steps16 != nullptr && ctx->int16_for(A.CCAP);
  const bool want_state = !pk_round;
  // per-parent selector features come from expand when m is a multiple of
  // four (its constant-index form), else from the feature kernel
  SArgs Ax = A;
  const bool per_parent = selector_parents(ctx);
  const bool feat_in_expand = per_parent && ctx->T.m % 4 == 0 &&
                              (kind == WSBB_LEARNED || selector_log_stride(ctx) > 0);
  if (feat_in_expand) {
    SearchState* S0 = ctx->search;
    Ax.x_frag = kind == WSBB_LEARNED ? S0->r_xfrag : nullptr;
    Ax.x_f32 = kind == WSBB_LEARNED ? nullptr : S0->r_feat;
    Ax.x_ld = A.PCAP;
    Ax.x_fsteps = selector_fsteps(ctx);
    Ax.x_inv_scale = 1.0f / static_cast<float>(ctx->inst.path_bound);
  }
  hipError_t e = launch_expand(ctx, Ax, (want_state || feat_in_expand) ? 1 : 0, st);
  if (e != hipSuccess) return e;
  ChildSource src{A.p_perm, A.p_l1, A.p_l2, A.c_desc, A.PCAP, A.CCAP};
  if (want_state) {
    src.hr = A.p_hr;
    src.hq = A.p_hq;
    src.hm = A.p_hm;
    src.hw = A.p_hw;
  }
  SearchState* SS = ctx->search;
  const int log_stride = selector_log_stride(ctx);
  BoundArgs B{};
  B.N = A.CCAP;
  B.Ndev = &A.ctrl->n_children;
  B.kind = kind == WSBB_LB1 ? 1 : 0;
  B.best = -1;
  // a round's live child count is often well below CCAP: split pairs over
  // more waves than a full batch of CCAP would need (measured on 20x10:
  // four waves per tile beat one, two, eight and sixteen)
  {
    // in the kernel's own tiles (128 nodes in the two-node form)
    const int tiles = pk_round ? (A.CCAP + 127) / 128 : (A.CCAP + 63) / 64;
    const int simds = 4 * ctx->num_cus;
    B.groups = ctx->launch.groups > 1 ? ctx->launch.groups
               : tiles < 2 * simds    ? 4
               : tiles < 4 * simds    ? 2
                                      : 1;
    // table lines loaded a pair ahead: measured on 20x10 to pay at one
    // 4096-child pool per round (256 waves) and to cost at ten (2560)
    B.touch = static_cast<long long>(tiles) * B.groups < simds;
  }
  B.out = A.c_bound;
  B.feat_scale = static_cast<float>(ctx->inst.path_bound);
  B.pk = ctx->int16_for(A.CCAP);
  const BatchSource parents{A.p_perm, A.p_l1, A.p_l2, A.PCAP};
  const float fscale = static_cast<float>(ctx->inst.path_bound);
  if (kind == WSBB_LEARNED) {
    // BF16 features, MLP pair subset per 64-child tile, LB2 over the subset;
    // children close to the incumbent then get full LB2.  Per parent: the
    // parents are scored and each tile sums its parents' probabilities.
    if (per_parent) {
      FeatArgs Fa{A.PCAP, &A.ctrl->n_parents, fscale, selector_fsteps(ctx), SS->r_xfrag};
      if (!feat_in_expand &&
          (e = launch_features(ctx->T, parents, Fa, ctx->grid_cap(), st)) != hipSuccess)
        return e;
      if ((e = launch_mlp_select(ctx, SS->r_xfrag, A.PCAP, &A.ctrl->n_parents, nullptr, nullptr, 0,
                                 nullptr, SS->r_prob)) != hipSuccess)
        return e;
      if ((e = launch_parent_tiles(ctx, SS->r_prob, A.p_off, A.c_desc, &A.ctrl->n_children,
                                   A.CCAP, SS->r_tsel)) != hipSuccess)
        return e;
    } else {
      FeatArgs Fa{A.CCAP, &A.ctrl->n_children, fscale, selector_fsteps(ctx), SS->r_xfrag};
      if ((e = launch_features(ctx->T, src, Fa, ctx->grid_cap(), st)) != hipSuccess) return e;
      if ((e = launch_mlp_select(ctx, SS->r_xfrag, A.CCAP, &A.ctrl->n_children, SS->r_tsel,
                                 nullptr, 0)) != hipSuccess)
        return e;
    }
    B.groups = 1;
    B.tile_sel = SS->r_tsel;
    B.nsw = selector_nsw(ctx);
    B.margin = -1;
  } else if (log_stride > 0) {
    B.argmax = SS->r_label;
    B.ld = A.CCAP;
    if (per_parent) {
      FeatArgs Fa{A.PCAP, &A.ctrl->n_parents, fscale, selector_fsteps(ctx), nullptr,
                  SS->r_feat, A.PCAP};
      if (!feat_in_expand &&
          (e = launch_features(ctx->T, parents, Fa, ctx->grid_cap(), st)) != hipSuccess)
        return e;
    } else {
      B.feat = SS->r_feat;
    }
  }
  // the bound kernel's own dispatch carries the timing events
  B.ev_start = ev0;
  B.ev_stop = ev1;
  e = launch_bound(ctx->T, src, B, ctx->grid_cap(), st);
  if (e != hipSuccess) return e;
  if (kind == WSBB_LEARNED && selector_margin(ctx) >= 0) {
    hipLaunchKernelGGL(k_fallback_list, dim3((A.CCAP + 255) / 256), dim3(256), 0, st, A,
                       selector_margin(ctx), SS->r_fbl);
    ListSource ls{src, SS->r_fbl, &A.ctrl->n_fb};
    BoundArgs Bf{};
    Bf.N = A.CCAP;
    Bf.Ndev = &A.ctrl->n_fb;
    Bf.kind = 0;
    Bf.best = -1;
    Bf.groups = B.groups > 1 ? B.groups : 4;
    Bf.margin = -1;
    Bf.out = A.c_bound;
    Bf.pk = ctx->int16_for(A.CCAP);
    if ((e = launch_bound(ctx->T, ls, Bf, ctx->grid_cap(), st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_fallback_stats, dim3(1), dim3(1), 0, st, A);
  }
  if (kind != WSBB_LEARNED && log_stride > 0) {
    LogArgs L{};
    L.ctrl = A.ctrl;
    L.feat = SS->r_feat;
    L.label = SS->r_label;
    L.ld = per_parent ? A.PCAP : A.CCAP;
    L.desc = per_parent ? A.c_desc : nullptr;
    L.F = 5 * ctx->T.m;
    L.stride = log_stride;
    selector_log_buffers(ctx, &L.log_feat, &L.log_label, &L.log_count, &L.cap);
    if (L.cap > 0) {
      hipLaunchKernelGGL(k_log, dim3(64), dim3(256), 0, st, L);
      hipLaunchKernelGGL(k_log_advance, dim3(1), dim3(1), 0, st, L);
    }
  }
  hipLaunchKernelGGL(k_rank, dim3(rblocks < kRankGrid ? rblocks : kRankGrid), dim3(256), 0, st, A);
  // more parents than one settle chunk: the chunks on kSettleChain blocks
  if (A.PCAP > kSettleChunk && A.PCAP <= kSettleChain * kChainChunk)
    hipLaunchKernelGGL(k_settle_chain, dim3(kSettleChain), dim3(1024), 0, st, A);
  else
    hipLaunchKernelGGL(k_settle, dim3(1), dim3(1024), 0, st, A);
  hipLaunchKernelGGL(k_push, dim3(pblocks < 4096 ? pblocks : 4096), dim3(256), 0, st, A);
  return hipGetLastError();
}

// final state of a solve from its last control block
static int finish_solve(const wsbb_ctx* ctx, const Ctrl& H, int32_t* best, int32_t* best_perm,
                        int64_t* stats) {
  const int n = ctx->T.n;
  if (best) *best = H.inc;
  if (best_perm) {
    for (int i = 0; i < n; ++i) best_perm[i] = H.have_perm ? H.best_perm[i] : -1;
  }
  if (stats)
    for (int i = 0; i < WSBB_NSTATS; ++i) stats[i] = H.stats[i];
  if (H.status == kOverflow) {
    set_error("device node pool overflow; raise the pool capacity");
    return WSBB_EOVERFLOW;
  }
  if (H.status == kBudget) {
    set_error("node budget exhausted");
    return WSBB_EBUDGET;
  }
  return WSBB_OK;
}

// Opens a search session on the context: the pool shape of later rounds,
// their buffers and kernel arguments.  The pool's contents are left as they
// are (callers reset the control block).
static int open_session(wsbb_ctx* ctx, int kind, int pool_size, int max_batches) {
  const int n = ctx->T.n;
  if (pool_size - 1 <= 2 * n || max_batches < 1 || max_batches > 64 ||
      static_cast<long long>(pool_size) * max_batches > (1 << 23)) {
    set_error("pool_size must exceed 2n+1 and pool_size*max_batches must be <= 2^23");
    return WSBB_EINVAL;
  }
  if (kind != WSBB_LB2 && kind != WSBB_LB1 && kind != WSBB_LEARNED) {
    set_error("search kind must be WSBB_LB2, WSBB_LB1 or WSBB_LEARNED");
    return WSBB_EINVAL;
  }
  if (kind == WSBB_LEARNED && !selector_ready(ctx)) {
    set_error("learned search needs a pair selector (wsbb_selector_init / train)");
    return WSBB_EINVAL;
  }
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  int rc = ensure_search(ctx, pool_size, max_batches, ctx->pool_capacity);
  if (rc) return rc;
  SearchState* S = ctx->search;
  if (kind == WSBB_LEARNED || selector_log_stride(ctx) > 0) {
    rc = ensure_aux(ctx);
    if (rc) return rc;
  }
  SArgs A = S->args;
  A.pool_m1 = pool_size - 1;
  A.max_batches = max_batches;
  A.budget = 0;
  A.count_pairs = kind == WSBB_LB2 ? 1 : kind == WSBB_LEARNED ? 2 : 0;
  A.topk = kind == WSBB_LEARNED ? std::min(selector_topk(ctx), ctx->T.P) : 0;
  S->sess = A;
  S->kind = kind;
  S->open = true;
  return WSBB_OK;
}

static int need_session(const wsbb_ctx* ctx) {
  if (!ctx->search || !ctx->search->open) {
    set_error("no open search: call wsbb_pool_open or wsbb_solve first");
    return WSBB_EINVAL;
  }
  return WSBB_OK;
}

// the control block, synchronously (the context's stream is drained)
static int read_ctrl(wsbb_ctx* ctx, Ctrl* h) {
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h, ctx->search->sess.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  return e == hipSuccess ? WSBB_OK : hip_fail(e, "control block read");
}
static int write_ctrl(wsbb_ctx* ctx, const Ctrl& h) {
  hipError_t e = hipMemcpyAsync(ctx->search->sess.ctrl, &h, sizeof(Ctrl), hipMemcpyHostToDevice,
                                ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  return e == hipSuccess ? WSBB_OK : hip_fail(e, "control block write");
}

// An empty pool with incumbent inc, the given counters (or zeros) and
// incumbent permutation (or none).
static int reset_pool(wsbb_ctx* ctx, int32_t inc, const int64_t* stats, const int32_t* perm) {
  const int n = ctx->T.n;
  Ctrl h{};
  h.status = kRunning;
  h.stack_size = 0;
  h.inc = inc;
  if (stats)
    for (int i = 0; i < WSBB_NSTATS; ++i) h.stats[i] = stats[i];
  if (perm && perm[0] >= 0) {
    uint64_t seen[2] = {0, 0};
    for (int i = 0; i < n; ++i) {
      const int j = perm[i];
      if (j < 0 || j >= n || ((seen[j >> 6] >> (j & 63)) & 1u)) {
        set_error("incumbent permutation is not a permutation of 0..n-1");
        return WSBB_EINVAL;
      }
      seen[j >> 6] |= uint64_t(1) << (j & 63);
      h.best_perm[i] = static_cast<uint8_t>(j);
    }
    h.have_perm = 1;
  }
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  return write_ctrl(ctx, h);
}

// Pushes count nodes (position-major permT [n][count], l1, l2, cost) on top
// of the pool, bottom-first: node count - 1 ends on top and is taken first.
// Every pool node the device makes has at least two free jobs (a child left
// with one is completed by limites_set, and leaves are evaluated, never
// stored), and each column is a permutation of 0..n-1.
static int push_nodes(wsbb_ctx* ctx, const uint8_t* permT, const int32_t* l1, const int32_t* l2,
                      const int32_t* cost, int count) {
  const int n = ctx->T.n;
  for (int k = 0; k < count; ++k) {
    const int a = l1[k], b = l2[k];
    bool ok = a >= -1 && b <= n && b - a >= 3 && cost[k] >= 0;
    uint64_t seen[2] = {0, 0};
    for (int i = 0; ok && i < n; ++i) {
      const int j = permT[static_cast<size_t>(i) * count + k];
      ok = j < n && !((seen[j >> 6] >> (j & 63)) & 1u);
      if (ok) seen[j >> 6] |= uint64_t(1) << (j & 63);
    }
    if (!ok) {
      set_error("pool node " + std::to_string(k) +
                " has fewer than two free jobs, a negative bound or is not a permutation of 0..n-1");
      return WSBB_EINVAL;
    }
  }
  SearchState* S = ctx->search;
  Ctrl h{};
  int rc = read_ctrl(ctx, &h);
  if (rc) return rc;
  const int top = h.stack_size;
  if (static_cast<long long>(top) + count > S->CAP) {
    set_error("pool push beyond the device pool capacity (" + std::to_string(S->CAP) + " nodes)");
    return WSBB_EOVERFLOW;
  }
  if (count == 0) return WSBB_OK;
  const SArgs& A = S->sess;
  hipError_t e = hipMemcpy2DAsync(A.s_perm + top, S->CAP, permT, count, count, n,
                                  hipMemcpyHostToDevice, ctx->stream);
  // select's key words, built here (the stream is synchronised before the
  // host vector goes out of scope)
  std::vector<int32_t> key(count);
  for (int k = 0; k < count; ++k) key[k] = select_key(l1[k], l2[k], cost[k], n);
  const std::pair<int*, const int32_t*> rows[] = {{A.s_l1, l1}, {A.s_l2, l2}, {A.s_cost, cost}};
  for (const auto& r : rows)
    if (e == hipSuccess)
      e = hipMemcpyAsync(r.first + top, r.second, 4 * size_t(count), hipMemcpyHostToDevice,
                         ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(A.s_key + top, key.data(), 4 * size_t(count), hipMemcpyHostToDevice,
                       ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(e, "pool push");
  h.stack_size = top + count;
  if (h.status == kDone) h.status = kRunning;
  return write_ctrl(ctx, h);
}

// Takes up to max nodes off the pool, from the top (the next nodes the
// search would decompose) or from the bottom (the shallowest: the largest
// subtrees); they are returned bottom-first, so pushing them back restores
// the pool.
static int pop_nodes(wsbb_ctx* ctx, int max, int from_bottom, uint8_t* permT, int32_t* l1,
                     int32_t* l2, int32_t* cost, int* count) {
  const int n = ctx->T.n;
  SearchState* S = ctx->search;
  Ctrl h{};
  int rc = read_ctrl(ctx, &h);
  if (rc) return rc;
  const int size = h.stack_size;
  const int k = std::max(0, std::min(max, size));
  *count = k;
  if (k == 0) return WSBB_OK;
  const SArgs& A = S->sess;
  const int first = from_bottom ? 0 : size - k;
  hipError_t e = hipMemcpy2DAsync(permT, k, A.s_perm + first, S->CAP, k, n, hipMemcpyDeviceToHost,
                                  ctx->stream);
  const std::pair<int32_t*, const int*> rows[] = {{l1, A.s_l1}, {l2, A.s_l2}, {cost, A.s_cost}};
  for (const auto& r : rows)
    if (e == hipSuccess)
      e = hipMemcpyAsync(r.first, r.second + first, 4 * size_t(k), hipMemcpyDeviceToHost,
                         ctx->stream);
  if (e == hipSuccess && from_bottom && size > k) {
    // the nodes above move down by k through a staging buffer (the source
    // and destination ranges may overlap)
    const size_t rest = size - k;
    e = S->tmp.reserve((n + 12) * rest);
    if (e == hipSuccess) {
      uint8_t* tp = S->tmp.as<uint8_t>(0);
      e = hipMemcpy2DAsync(tp, rest, A.s_perm + k, S->CAP, rest, n, hipMemcpyDeviceToDevice,
                           ctx->stream);
      if (e == hipSuccess)
        e = hipMemcpy2DAsync(A.s_perm, S->CAP, tp, rest, rest, n, hipMemcpyDeviceToDevice,
                             ctx->stream);
      int* ti = reinterpret_cast<int*>(tp + static_cast<size_t>(n) * rest);
      int* cols[] = {A.s_l1, A.s_l2, A.s_cost, A.s_key};
      for (int c = 0; c < 4 && e == hipSuccess; ++c) {
        e = hipMemcpyAsync(ti, cols[c] + k, 4 * rest, hipMemcpyDeviceToDevice, ctx->stream);
        if (e == hipSuccess)
          e = hipMemcpyAsync(cols[c], ti, 4 * rest, hipMemcpyDeviceToDevice, ctx->stream);
      }
    }
  }
  if (e != hipSuccess) return hip_fail(e, "pool pop");
  h.stack_size = size - k;
  return write_ctrl(ctx, h);
}

// Runs rounds of the open session until the pool is empty (or the budget
// stops it, or it would overflow), or after max_rounds rounds (<= 0: no
// limit).  budget bounds the decomposed parents counted so far (<= 0:
// none).  *last receives the final control block.
static int run_rounds(wsbb_ctx* ctx, long long max_rounds, int64_t budget, Ctrl* last) {
  SearchState* S = ctx->search;
  const int kind = S->kind;
  Ctrl h{};
  int rc = read_ctrl(ctx, &h);
  if (rc) return rc;
  if (h.status != kRunning) {  // a stop by budget or overflow, or an emptied pool
    h.status = kRunning;
    if ((rc = write_ctrl(ctx, h))) return rc;
  }
  SArgs A = S->sess;
  A.budget = budget;
  hipError_t e;
  const long long limit = max_rounds > 0 ? max_rounds : LLONG_MAX;
  if (ctx->trace_fn) {
    // one round at a time, each followed by host copies of its parents,
    // children and bounds handed to the trace callback
    const int n = ctx->T.n;
    std::vector<uint8_t> permT;
    std::vector<int32_t> l1, l2, cost, desc, bound;
    Ctrl& H = S->host_ctrl[0];
    for (long long round = 0; round < limit; ++round) {
      if ((e = launch_round(ctx, A, kind, nullptr, nullptr)) != hipSuccess)
        return hip_fail(e, "round launch");
      e = hipMemcpyAsync(&H, A.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, ctx->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) return hip_fail(e, "trace control copy");
      const int np = H.n_parents, nc = H.n_children;
      if (np > 0) {
        permT.resize(static_cast<size_t>(n) * np);
        l1.resize(np);
        l2.resize(np);
        cost.resize(np);
        desc.resize(nc);
        bound.resize(nc);
        e = hipMemcpy2DAsync(permT.data(), np, A.p_perm, A.PCAP, np, n, hipMemcpyDeviceToHost,
                             ctx->stream);
        const std::pair<int32_t*, const int*> rows[] = {{l1.data(), A.p_l1}, {l2.data(), A.p_l2},
                                                        {cost.data(), A.p_cost}};
        for (const auto& r : rows)
          if (e == hipSuccess)
            e = hipMemcpyAsync(r.first, r.second, 4 * size_t(np), hipMemcpyDeviceToHost,
                               ctx->stream);
        if (e == hipSuccess && nc > 0)
          e = hipMemcpyAsync(desc.data(), A.c_desc, 4 * size_t(nc), hipMemcpyDeviceToHost,
                             ctx->stream);
        if (e == hipSuccess && nc > 0)
          e = hipMemcpyAsync(bound.data(), A.c_bound, 4 * size_t(nc), hipMemcpyDeviceToHost,
                             ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return hip_fail(e, "trace copy");
        ctx->trace_fn(ctx->trace_user, round, np, permT.data(), l1.data(), l2.data(),
                      cost.data(), nc, desc.data(), bound.data(), H.inc);
      }
      if (H.status != kRunning) break;
    }
    *last = H;
    return WSBB_OK;
  }
  const int rps = ctx->rounds_per_sync;
  const int every = ctx->profile_every;
  const bool timing = every > 0;
  if (timing && static_cast<int>(S->ev.size()) < 4 * rps) {
    while (static_cast<int>(S->ev.size()) < 4 * rps) {
      hipEvent_t ev;
      if ((e = hipEventCreate(&ev)) != hipSuccess) return hip_fail(e, "hipEventCreate");
      S->ev.push_back(ev);
    }
  }
  S->bound_ms = 0;
  S->bound_launches = 0;
  // hipGraph replay: the kernels of rps rounds are captured once per call
  // (their arguments do not change during it; all round state lives in
  // device memory) and replayed until the pool state says stop
  const bool use_graph = ctx->search_graph && !timing;
  if (S->graph) {
    (void)hipGraphExecDestroy(S->graph);
    S->graph = nullptr;
  }
  if (use_graph) {
    hipGraph_t g = nullptr;
    if ((e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
      return hip_fail(e, "hipStreamBeginCapture");
    hipError_t le = hipSuccess;
    for (int r = 0; r < rps && le == hipSuccess; ++r)
      le = launch_round(ctx, A, kind, nullptr, nullptr);
    e = hipStreamEndCapture(ctx->stream, &g);
    if (le != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return hip_fail(le, "round capture");
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&S->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      S->graph = nullptr;
      return hip_fail(e, "hipGraphInstantiate");
    }
  }
  // Batches of up to rps rounds, each followed by a copy of the control
  // block into its own pinned slot.  The next batch is queued before the
  // host waits on the current one, so the GPU never idles on a host round
  // trip; once a batch reports the search stopped, the batch behind it only
  // repeats finished rounds (each returns at once) and is drained.  Timed
  // rounds: every `every`-th round of the call, counted across batches.
  long long queued = 0, base[2] = {0, 0};
  int nq[2] = {0, 0};
  auto enqueue = [&](int slot) -> hipError_t {
    hipError_t le;
    const int k = static_cast<int>(std::min<long long>(rps, limit - queued));
    nq[slot] = k;
    if (k <= 0) return hipSuccess;
    base[slot] = queued;
    queued += k;
    if (use_graph && k == rps) {
      if ((le = hipGraphLaunch(S->graph, ctx->stream)) != hipSuccess) return le;
    } else {
      for (int r = 0; r < k; ++r) {
        hipEvent_t* ev = timing && (base[slot] + r) % every == 0
                             ? &S->ev[2 * (slot * rps + r)]
                             : nullptr;
        if ((le = launch_round(ctx, A, kind, ev ? ev[0] : nullptr, ev ? ev[1] : nullptr)) !=
            hipSuccess)
          return le;
      }
    }
    le = hipMemcpyAsync(S->host_ctrl + slot, A.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost,
                        ctx->stream);
    if (le == hipSuccess) le = hipEventRecord(S->copied[slot], ctx->stream);
    return le;
  };
  // rounds of a batch that did work: the first (rounds counter growth) ones
  long long prev_rounds = h.stats[kStRounds];
  auto collect = [&](int slot) -> hipError_t {
    const hipError_t le = hipEventSynchronize(S->copied[slot]);
    if (le != hipSuccess) return le;
    const long long live = S->host_ctrl[slot].stats[kStRounds] - prev_rounds;
    prev_rounds = S->host_ctrl[slot].stats[kStRounds];
    if (!timing) return le;
    for (int r = 0; r < nq[slot] && r < live; ++r) {
      if ((base[slot] + r) % every != 0) continue;
      float t = 0;
      const int i = 2 * (slot * rps + r);
      if (hipEventElapsedTime(&t, S->ev[i], S->ev[i + 1]) == hipSuccess) {
        S->bound_ms += t;
        ++S->bound_launches;
      }
    }
    return le;
  };
  int cur = 0;
  if ((e = enqueue(cur)) != hipSuccess) return hip_fail(e, "round launch");
  for (;;) {
    if ((e = enqueue(cur ^ 1)) != hipSuccess) return hip_fail(e, "round launch");
    const bool next = nq[cur ^ 1] > 0;
    if ((e = collect(cur)) != hipSuccess) return hip_fail(e, "search sync");
    if (S->host_ctrl[cur].status != kRunning) {
      // the batch behind repeats finished rounds: wait for it, keep this
      // batch's (identical) control block and leave its timings out
      if (next && (e = hipEventSynchronize(S->copied[cur ^ 1])) != hipSuccess)
        return hip_fail(e, "search sync");
      break;
    }
    if (!next) break;  // the round limit
    cur ^= 1;
  }
  *last = S->host_ctrl[cur];
  return WSBB_OK;
}

int solve(wsbb_ctx* ctx, int kind, int pool_size, int max_batches, int32_t ub0,
          int64_t budget, int32_t* best, int32_t* best_perm, int64_t* stats,
          const ResumeState* resume) {
  int rc = open_session(ctx, kind, pool_size, max_batches);
  if (rc) return rc;
  const int n = ctx->T.n;
  if (resume == nullptr) {
    // the root (create_raw_problem 9-22, tree::add_root tree.cpp:170-176)
    std::vector<uint8_t> root(n);
    for (int i = 0; i < n; ++i) root[i] = static_cast<uint8_t>(i);
    const int32_t a = -1, b = n, c = 0;
    if ((rc = reset_pool(ctx, ub0, nullptr, nullptr))) return rc;
    if ((rc = push_nodes(ctx, root.data(), &a, &b, &c, 1))) return rc;
  } else {
    if (resume->count < 0 || resume->count > ctx->search->CAP) {
      set_error("saved pool larger than the device pool capacity");
      return WSBB_EOVERFLOW;
    }
    if ((rc = reset_pool(ctx, ub0, resume->stats, resume->best_perm))) return rc;
    if ((rc = push_nodes(ctx, resume->permT, resume->l1, resume->l2, resume->cost,
                         resume->count)))
      return rc;
  }
  Ctrl last{};
  if ((rc = run_rounds(ctx, 0, budget, &last))) return rc;
  return finish_solve(ctx, last, best, best_perm, stats);
}

}  // namespace wsbb

extern "C" int wsbb_search_profile(wsbb_ctx* ctx, int enable, double* bound_ms,
                                   int64_t* bound_launches) {
  if (!ctx) {
    wsbb::set_error("null context");
    return WSBB_EINVAL;
  }
  if (enable >= 0) ctx->profile_every = enable;
  if (bound_ms) *bound_ms = ctx->search ? ctx->search->bound_ms : 0.0;
  if (bound_launches) *bound_launches = ctx->search ? ctx->search->bound_launches : 0;
  return WSBB_OK;
}

extern "C" int wsbb_solve(wsbb_ctx* ctx, int kind, int pool_size, int max_batches,
                          int32_t ub0, int64_t budget, int32_t* best,
                          int32_t* best_perm, int64_t* stats) {
  if (!ctx) {
    wsbb::set_error("null context");
    return WSBB_EINVAL;
  }
  return wsbb::solve(ctx, kind, pool_size, max_batches, ub0, budget, best, best_perm, stats);
}

extern "C" int wsbb_search_state(wsbb_ctx* ctx, int32_t* inc, int64_t* stats,
                                 int32_t* best_perm, int* count, uint8_t* permT,
                                 int32_t* l1, int32_t* l2, int32_t* cost, int cap) {
  using namespace wsbb;
  if (!ctx || !count) {
    set_error("null context or count");
    return WSBB_EINVAL;
  }
  if (!ctx->search) {
    set_error("no search state: call wsbb_solve first");
    return WSBB_EINVAL;
  }
  SearchState* S = ctx->search;
  const int n = ctx->T.n;
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  Ctrl h{};
  if (e == hipSuccess) e = hipMemcpy(&h, S->args.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "search state read");
  *count = h.stack_size;
  if (inc) *inc = h.inc;
  if (stats)
    for (int i = 0; i < WSBB_NSTATS; ++i) stats[i] = h.stats[i];
  if (best_perm)
    for (int i = 0; i < n; ++i) best_perm[i] = h.have_perm ? h.best_perm[i] : -1;
  if (!permT && !l1 && !l2 && !cost) return WSBB_OK;
  if (cap < h.stack_size || !permT || !l1 || !l2 || !cost) {
    set_error("pool buffers missing or smaller than the pool (see count)");
    return WSBB_EINVAL;
  }
  const int cnt = h.stack_size;
  for (int i = 0; i < n && e == hipSuccess && cnt > 0; ++i)
    e = hipMemcpy(permT + static_cast<size_t>(i) * cnt, S->args.s_perm + static_cast<size_t>(i) * S->CAP,
                  cnt, hipMemcpyDeviceToHost);
  if (e == hipSuccess && cnt > 0) e = hipMemcpy(l1, S->args.s_l1, 4 * size_t(cnt), hipMemcpyDeviceToHost);
  if (e == hipSuccess && cnt > 0) e = hipMemcpy(l2, S->args.s_l2, 4 * size_t(cnt), hipMemcpyDeviceToHost);
  if (e == hipSuccess && cnt > 0) e = hipMemcpy(cost, S->args.s_cost, 4 * size_t(cnt), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "search pool read");
  return WSBB_OK;
}

extern "C" int wsbb_solve_resume(wsbb_ctx* ctx, int kind, int pool_size, int max_batches,
                                 int64_t budget, const uint8_t* permT, const int32_t* l1,
                                 const int32_t* l2, const int32_t* cost, int count,
                                 int32_t inc, const int64_t* stats_in,
                                 const int32_t* best_perm_in, int32_t* best,
                                 int32_t* best_perm, int64_t* stats) {
  if (!ctx) {
    wsbb::set_error("null context");
    return WSBB_EINVAL;
  }
  if (count < 0 || (count > 0 && (!permT || !l1 || !l2 || !cost))) {
    wsbb::set_error("bad saved pool");
    return WSBB_EINVAL;
  }
  const wsbb::ResumeState R{permT, l1, l2, cost, count, stats_in, best_perm_in};
  return wsbb::solve(ctx, kind, pool_size, max_batches, inc, budget, best, best_perm, stats, &R);
}

extern "C" int wsbb_search_trace(wsbb_ctx* ctx, wsbb_trace_fn fn, void* user) {
  if (!ctx) {
    wsbb::set_error("null context");
    return WSBB_EINVAL;
  }
  ctx->trace_fn = fn;
  ctx->trace_user = fn ? user : nullptr;
  return WSBB_OK;
}

extern "C" int wsbb_pool_open(wsbb_ctx* ctx, int kind, int pool_size, int max_batches) {
  if (!ctx) {
    wsbb::set_error("null context");
    return WSBB_EINVAL;
  }
  int rc = wsbb::open_session(ctx, kind, pool_size, max_batches);
  if (rc) return rc;
  return wsbb::reset_pool(ctx, INT32_MAX, nullptr, nullptr);
}

extern "C" int wsbb_pool_seed_root(wsbb_ctx* ctx) {
  if (!ctx) {
    wsbb::set_error("null context");
    return WSBB_EINVAL;
  }
  if (int rc = wsbb::need_session(ctx)) return rc;
  const int n = ctx->T.n;
  std::vector<uint8_t> root(n);
  for (int i = 0; i < n; ++i) root[i] = static_cast<uint8_t>(i);
  const int32_t a = -1, b = n, c = 0;
  return wsbb::push_nodes(ctx, root.data(), &a, &b, &c, 1);
}

extern "C" int wsbb_pool_push(wsbb_ctx* ctx, const uint8_t* permT, const int32_t* l1,
                              const int32_t* l2, const int32_t* cost, int count) {
  if (!ctx || count < 0 || (count > 0 && (!permT || !l1 || !l2 || !cost))) {
    wsbb::set_error("null context or bad node buffers");
    return WSBB_EINVAL;
  }
  if (int rc = wsbb::need_session(ctx)) return rc;
  return wsbb::push_nodes(ctx, permT, l1, l2, cost, count);
}

extern "C" int wsbb_pool_pop(wsbb_ctx* ctx, int max, int from_bottom, uint8_t* permT,
                             int32_t* l1, int32_t* l2, int32_t* cost, int* count) {
  if (!ctx || max < 0 || !count || (max > 0 && (!permT || !l1 || !l2 || !cost))) {
    wsbb::set_error("null context, negative max or bad node buffers");
    return WSBB_EINVAL;
  }
  if (int rc = wsbb::need_session(ctx)) return rc;
  return wsbb::pop_nodes(ctx, max, from_bottom, permT, l1, l2, cost, count);
}

extern "C" int wsbb_pool_size(wsbb_ctx* ctx, int* size) {
  if (!ctx || !size) {
    wsbb::set_error("null context or size");
    return WSBB_EINVAL;
  }
  if (int rc = wsbb::need_session(ctx)) return rc;
  wsbb::Ctrl h{};
  if (int rc = wsbb::read_ctrl(ctx, &h)) return rc;
  *size = h.stack_size;
  return WSBB_OK;
}

extern "C" int wsbb_incumbent_get(wsbb_ctx* ctx, int32_t* inc, int32_t* perm) {
  if (!ctx || !inc) {
    wsbb::set_error("null context or incumbent");
    return WSBB_EINVAL;
  }
  if (int rc = wsbb::need_session(ctx)) return rc;
  wsbb::Ctrl h{};
  if (int rc = wsbb::read_ctrl(ctx, &h)) return rc;
  *inc = h.inc;
  if (perm)
    for (int i = 0; i < ctx->T.n; ++i) perm[i] = h.have_perm ? h.best_perm[i] : -1;
  return WSBB_OK;
}

extern "C" int wsbb_incumbent_set(wsbb_ctx* ctx, int32_t inc, const int32_t* perm) {
  if (!ctx) {
    wsbb::set_error("null context");
    return WSBB_EINVAL;
  }
  if (int rc = wsbb::need_session(ctx)) return rc;
  const int n = ctx->T.n;
  wsbb::Ctrl h{};
  if (int rc = wsbb::read_ctrl(ctx, &h)) return rc;
  if (inc >= h.inc) return WSBB_OK;  // only a strictly better value replaces it
  h.inc = inc;
  h.have_perm = 0;
  if (perm) {
    uint64_t seen[2] = {0, 0};
    for (int i = 0; i < n; ++i) {
      const int j = perm[i];
      if (j < 0 || j >= n || ((seen[j >> 6] >> (j & 63)) & 1u)) {
        wsbb::set_error("incumbent permutation is not a permutation of 0..n-1");
        return WSBB_EINVAL;
      }
      seen[j >> 6] |= uint64_t(1) << (j & 63);
      h.best_perm[i] = static_cast<uint8_t>(j);
    }
    h.have_perm = 1;
  }
  return wsbb::write_ctrl(ctx, h);
}

extern "C" int wsbb_pool_run(wsbb_ctx* ctx, int rounds, int64_t budget, int* size,
                             int64_t* stats) {
  if (!ctx) {
    wsbb::set_error("null context");
    return WSBB_EINVAL;
  }
  if (int rc = wsbb::need_session(ctx)) return rc;
  wsbb::Ctrl last{};
  if (int rc = wsbb::run_rounds(ctx, rounds, budget, &last)) return rc;
  if (size) *size = last.stack_size;
  return wsbb::finish_solve(ctx, last, nullptr, nullptr, stats);
}

extern "C" int wsbb_search_options(wsbb_ctx* ctx, int use_graph, int rounds_per_sync) {
  if (!ctx) {
    wsbb::set_error("null context");
    return WSBB_EINVAL;
  }
  if (rounds_per_sync == 0 || rounds_per_sync > 64) {
    wsbb::set_error("rounds_per_sync must be in 1..64 (-1 keeps the current value)");
    return WSBB_EINVAL;
  }
  if (use_graph >= 0) ctx->search_graph = use_graph != 0;
  if (rounds_per_sync > 0) ctx->rounds_per_sync = rounds_per_sync;
  return WSBB_OK;
}
