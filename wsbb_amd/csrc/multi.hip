// Mixed batches: nodes of several instances (different n and m) bounded in
// one launch.
//
// Every 64-node tile holds nodes of one instance only; a tile map gives each
// tile its instance, its first node and its node count, and every instance
// keeps its own context (tables, pair selector).  The bound kernel runs four
// waves per tile (the pair range split four ways, like the single-instance
// kernel), loads the tile's tables through scalar loads and restages the
// processing-time matrix in LDS only when the instance changes from one tile
// to the next.  With learned pair selection each tile is bounded over the
// top-k pairs its instance's MLP chose; a node whose subset bound lies within
// `margin` below its instance's incumbent gets the remaining pairs too
// (full LB2).  Per node the result equals bound_flowshop::bornes_calculer
// (src/bound_flowshop.cpp:929-951) of its own instance, or the LB2 over the
// tile's subset.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/wsbb_amd.h"
#include "bound.h"
#include "context.h"
#include "selector.h"

struct wsbb_multi {
  int device = 0;
  int num_cus = 0;
  std::vector<wsbb_ctx*> ctx;
  int maxn = 0, maxm = 0, nsw = 0, fstride = 0;
  wsbb::DevBuf tabs;   // DevTables per instance
  wsbb::DevBuf batch;  // resident batch: nodes, tile map, masks, features, outputs
  // resident batch layout
  int N = 0, ntiles = 0;
  uint8_t* permT = nullptr;  // [maxn][N]
  int32_t *l1 = nullptr, *l2 = nullptr, *out = nullptr;
  int *tile_inst = nullptr, *tile_node0 = nullptr, *tile_cnt = nullptr, *tile_list = nullptr;
  int32_t* inc = nullptr;  // [count]
  uint32_t* tsel = nullptr;
  uint2* xfrag = nullptr;
  long long* stats = nullptr;
  // per machine-count class: the tiles of its instances (in tile_list) and
  // the instance whose selector scores them
  std::vector<int> class_m, class_first, class_count, class_rep;
  std::vector<int> offsets;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace wsbb {

namespace {

struct MultiArgs {
  const DevTables* tabs;
  const int* tile_inst;
  const int* tile_node0;
  const int* tile_cnt;
  int ntiles, maxn, N;
  const uint8_t* permT;  // [maxn][N]
  const int32_t* l1;
  const int32_t* l2;
  int32_t* out;
  const uint32_t* tile_sel;  // [ntiles][nsw] (learned), or null
  int nsw;
  const int32_t* inc;        // [instances] incumbents for the fallback
  int margin;                // -1: no fallback
  long long* stats;          // [3] nodes, fallbacks, node-pair evaluations
  float scale_unused;
};

// the tile's tables, word by word through scalar loads (inst is uniform)
__device__ __forceinline__ DevTables tile_tables(const DevTables* tabs, int inst) {
  static_assert(sizeof(DevTables) % 4 == 0, "word copy");
  DevTables T;
  const int* src = reinterpret_cast<const int*>(tabs + inst);
  int* dst = reinterpret_cast<int*>(&T);
#pragma unroll
  for (int i = 0; i < static_cast<int>(sizeof(DevTables) / 4); ++i) dst[i] = ld_const(src, i);
  return T;
}

// four Johnson steps, as in bound.hip
__device__ __forceinline__ uint32_t bsel(uint32_t mask, uint32_t a, uint32_t b) {
  return (mask & a) | (~mask & b);
}

template <int NW>
__device__ __forceinline__ int scan_pair(const DevTables& T, int s, const uint32_t (&um)[NW],
                                         const int* r, const int* q, int d2) {
  const int2 ab = ld_const2(T.pairs, s);
  int T0 = r[ab.x] - kBias;
  int t1 = r[ab.y];
  const size_t base = static_cast<size_t>(s) * T.n4;
  for (int k = 0; k < T.n4; k += 4) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int4 qq = ld_const4(T.steps, base + k + u);
      uint32_t word;
      if constexpr (NW == 1) {
        word = um[0];
      } else {
        const uint2 sl = ld_const2u(T.wsel, base + k + u);
        if constexpr (NW == 2) {
          word = bsel(sl.x, um[1], um[0]);
        } else {
          const uint32_t w01 = bsel(sl.x, um[1], um[0]);
          const uint32_t w23 = bsel(sl.x, um[3], um[2]);
          word = bsel(sl.y, w23, w01);
        }
      }
      const uint32_t f = __builtin_amdgcn_ubfe(word, static_cast<uint32_t>(qq.x), 1u);
      T0 = static_cast<int>(__umul24(f, static_cast<uint32_t>(qq.y)) + T0);
      const int x = static_cast<int>(__umul24(f, static_cast<uint32_t>(qq.z)) + t1);
      const int y = static_cast<int>(__umul24(f, static_cast<uint32_t>(qq.w)) + T0);
      t1 = max(x, y);
    }
  }
  const int qa = q[ab.x], qb = q[ab.y], q0b = ld_const(T.Q0, ab.y);
  return d2 ? max(t1 + qb, T0 + kBias + qa) : t1 + q0b;
}

template <int MM, int NW, bool SUBSET>
__global__ __launch_bounds__(256) void k_multi_bound(MultiArgs A) {
  extern __shared__ int32_t lds[];
  const int mn4 = (MM * A.maxn + 3) & ~3;
  int32_t* sp = lds;                                 // [MM][n] of the staged instance
  int4* part = reinterpret_cast<int4*>(lds + mn4);   // [256]
  uint8_t* pl = reinterpret_cast<uint8_t*>(part + 256);  // [maxn][64]
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int grp = __builtin_amdgcn_readfirstlane(wave);
  int staged = -1;
  for (int bt = blockIdx.x; bt < A.ntiles; bt += gridDim.x) {
    const int inst = __builtin_amdgcn_readfirstlane(A.tile_inst[bt]);
    const DevTables T = tile_tables(A.tabs, inst);
    const int n = T.n, m = T.m, P = T.P;
    if (inst != staged) {
      __syncthreads();
      for (int i = threadIdx.x; i < MM * n; i += blockDim.x) sp[i] = i < m * n ? T.p[i] : 0;
      staged = inst;
    }
    const int node0 = __builtin_amdgcn_readfirstlane(A.tile_node0[bt]);
    const int cnt = __builtin_amdgcn_readfirstlane(A.tile_cnt[bt]);
    const int node = node0 + lane;
    const bool valid = lane < cnt;
    int l1 = -1, l2 = n;
    if (valid) {
      l1 = A.l1[node];
      l2 = A.l2[node];
    }
    if (grp == 0 && valid) {
      for (int i0 = 0; i0 < n; i0 += 16) {
        uint32_t v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u)
          if (i0 + u < n) v[u] = A.permT[static_cast<size_t>(i0 + u) * A.N + node];
#pragma unroll
        for (int u = 0; u < 16; ++u)
          if (i0 + u < n) pl[(i0 + u) * 64 + lane] = static_cast<uint8_t>(v[u]);
      }
    }
    __syncthreads();
    const int d1 = l1 + 1, d2 = n - l2;
    int r[MM], q[MM];
    uint32_t sm[NW];
#pragma unroll
    for (int k = 0; k < MM; ++k) r[k] = q[k] = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) sm[w] = 0;
    for (int i = 0; i < d1; ++i) {
      const int j = pl[i * 64 + lane];
#pragma unroll
      for (int w = 0; w < NW; ++w) sm[w] |= ((j >> 5) == w) ? (1u << (j & 31)) : 0u;
      r[0] += sp[j];
#pragma unroll
      for (int k = 1; k < MM; ++k) r[k] = max(r[k], r[k - 1]) + sp[k * n + j];
    }
    for (int i = n - 1; i >= l2; --i) {
      const int j = pl[i * 64 + lane];
#pragma unroll
      for (int w = 0; w < NW; ++w) sm[w] |= ((j >> 5) == w) ? (1u << (j & 31)) : 0u;
#pragma unroll
      for (int k = MM - 1; k >= 0; --k) q[k] = max(q[k], (k + 1 < MM) ? q[k + 1] : 0) + sp[k * n + j];
    }
    const bool leaf = l2 - l1 == 1;
    int leafv = 0;
#pragma unroll
    for (int k = 0; k < MM; ++k) leafv = max(leafv, r[k] + q[k]);
    uint32_t um[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const int hi = n - 32 * w;
      const uint32_t vb = hi >= 32 ? 0xffffffffu : (hi > 0 ? ((1u << hi) - 1u) : 0u);
      um[w] = ~sm[w] & vb;
    }
#pragma unroll
    for (int k = 0; k < MM; ++k)
      if (d1 == 0) r[k] = ld_const(T.R0, k);  // zero past machine m
    const int s_lo = (grp * P) >> 2, s_hi = ((grp + 1) * P) >> 2;
    const uint32_t* tsel = SUBSET ? A.tile_sel + static_cast<size_t>(bt) * A.nsw : nullptr;
    int lb = 0, npairs = 0;
    for (int s = s_lo; s < s_hi; ++s) {
      if constexpr (SUBSET)
        if (!((ld_const_u32(tsel, s >> 5) >> (s & 31)) & 1u)) continue;
      ++npairs;
      lb = max(lb, scan_pair<NW>(T, s, um, r, q, d2));
    }
    // combine the four waves; every wave reads the tile's subset bound
    part[threadIdx.x] = make_int4(lb, npairs, 0, 0);
    __syncthreads();
    int tlb = 0, tpairs = 0;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int4 e = part[g * 64 + lane];
      tlb = max(tlb, e.x);
      tpairs += e.y;
    }
    __syncthreads();
    long long fallbacks = 0;
    if constexpr (SUBSET) {
      if (A.margin >= 0 && A.inc != nullptr) {
        const int inc = A.inc[inst];
        const bool take = valid && !leaf && tlb < inc && tlb >= inc - A.margin;
        const unsigned long long bal = __ballot(take);
        if (bal) {  // wave-uniform: the complement pairs for the listed nodes
          fallbacks = __popcll(bal);
          int lb2 = 0;
          for (int s = s_lo; s < s_hi; ++s) {
            if ((ld_const_u32(tsel, s >> 5) >> (s & 31)) & 1u) continue;
            const int v = scan_pair<NW>(T, s, um, r, q, d2);
            lb2 = max(lb2, v);
          }
          part[threadIdx.x] = make_int4(take ? lb2 : 0, 0, 0, 0);
        } else {
          part[threadIdx.x] = make_int4(0, 0, 0, 0);
        }
        // every wave of the tile computed the same ballot; combine
        __syncthreads();
#pragma unroll
        for (int g = 0; g < 4; ++g) tlb = max(tlb, part[g * 64 + lane].x);
        __syncthreads();
      }
    }
    if (grp == 0 && valid) A.out[node] = leaf ? leafv : tlb;
    if (A.stats != nullptr && grp == 0 && lane == 0) {
      atomicAdd(reinterpret_cast<unsigned long long*>(&A.stats[0]), static_cast<unsigned long long>(cnt));
      atomicAdd(reinterpret_cast<unsigned long long*>(&A.stats[1]), static_cast<unsigned long long>(fallbacks));
      const unsigned long long pe = static_cast<unsigned long long>(cnt) * tpairs +
                                    static_cast<unsigned long long>(fallbacks) * (P - tpairs);
      atomicAdd(reinterpret_cast<unsigned long long*>(&A.stats[2]), pe);
    }
  }
}

// selector features of a mixed batch: as k_features (bound.hip), with each
// tile's own instance; one tile per wave, each wave staging its own
// instance's times
struct MultiFeatArgs {
  const DevTables* tabs;
  const int* tile_inst;
  const int* tile_node0;
  const int* tile_cnt;
  int ntiles, maxn, N, fstride;
  const uint8_t* permT;
  const int32_t* l1;
  const int32_t* l2;
  const float* scale;  // [instances]
  uint2* frag;         // [2 ntiles][fstride][64]
};

__device__ __forceinline__ uint16_t f2bf_m(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}

template <int MM>
__global__ __launch_bounds__(256) void k_multi_features(MultiFeatArgs A) {
  extern __shared__ int32_t lds[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int mn4 = (MM * A.maxn + 3) & ~3;
  int32_t* sp = lds + wave * mn4;  // this wave's staged instance
  uint8_t* pl = reinterpret_cast<uint8_t*>(lds + 4 * mn4) + wave * A.maxn * 64;
  int staged = -1;
  for (int t0 = blockIdx.x * 4; t0 < A.ntiles; t0 += gridDim.x * 4) {
    const int tile = __builtin_amdgcn_readfirstlane(t0 + wave);
    if (tile >= A.ntiles) continue;
    const int inst = __builtin_amdgcn_readfirstlane(A.tile_inst[tile]);
    const DevTables T = tile_tables(A.tabs, inst);
    const int n = T.n, m = T.m;
    if (inst != staged) {
      for (int i = lane; i < MM * n; i += 64) sp[i] = i < m * n ? T.p[i] : 0;
      staged = inst;
    }
    const int node0 = __builtin_amdgcn_readfirstlane(A.tile_node0[tile]);
    const int cnt = __builtin_amdgcn_readfirstlane(A.tile_cnt[tile]);
    const int node = node0 + lane;
    const bool valid = lane < cnt;
    int l1 = -1, l2 = n;
    if (valid) {
      l1 = A.l1[node];
      l2 = A.l2[node];
      for (int i0 = 0; i0 < n; i0 += 16) {
        uint32_t v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u)
          if (i0 + u < n) v[u] = A.permT[static_cast<size_t>(i0 + u) * A.N + node];
#pragma unroll
        for (int u = 0; u < 16; ++u)
          if (i0 + u < n) pl[(i0 + u) * 64 + lane] = static_cast<uint8_t>(v[u]);
      }
    }
    const int d1 = l1 + 1, d2 = n - l2;
    int r[MM], q[MM], sc_[MM];
#pragma unroll
    for (int k = 0; k < MM; ++k) r[k] = q[k] = sc_[k] = 0;
    for (int i = 0; i < d1; ++i) {
      const int j = pl[i * 64 + lane];
      r[0] += sp[j];
      sc_[0] += sp[j];
#pragma unroll
      for (int k = 1; k < MM; ++k) {
        const int pk = sp[k * n + j];
        r[k] = max(r[k], r[k - 1]) + pk;
        sc_[k] += pk;
      }
    }
    for (int i = n - 1; i >= l2; --i) {
      const int j = pl[i * 64 + lane];
#pragma unroll
      for (int k = MM - 1; k >= 0; --k) {
        const int pk = sp[k * n + j];
        q[k] = max(q[k], (k + 1 < MM) ? q[k + 1] : 0) + pk;
        sc_[k] += pk;
      }
    }
#pragma unroll
    for (int k = 0; k < MM; ++k) {
      r[k] = d1 ? r[k] : ld_const(T.R0, k);
      q[k] = d2 ? q[k] : ld_const(T.Q0, k);
      sc_[k] = ld_const(T.ptot, k) - sc_[k];
    }
    const float sc = 1.0f / A.scale[inst];
    const int fsteps = (5 * m + 7) / 8;
    const int blk = 2 * tile + (lane >> 5), c = lane & 31;
    int g = 0, k = 0;
    for (int t = 0; t < A.fstride; ++t)
      for (int half = 0; half < 2; ++half) {
        uint32_t w[2] = {0u, 0u};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          float x = 0.f;
          if (t < fsteps && g < 5) {
            const int h = r[k], tl = q[k], rw = sc_[k];
            const int v = g == 0 ? h : g == 1 ? tl : g == 2 ? rw : g == 3 ? h + rw : rw + tl;
            x = valid ? static_cast<float>(v) * sc : 0.f;
            if (++k == m) {
              k = 0;
              ++g;
            }
          }
          w[i >> 1] |= static_cast<uint32_t>(f2bf_m(x)) << (16 * (i & 1));
        }
        A.frag[(static_cast<size_t>(blk) * A.fstride + t) * 64 + half * 32 + c] = make_uint2(w[0], w[1]);
      }
  }
}

template <int MM, int NW>
hipError_t launch_multi_mm(const MultiArgs& A, bool subset, int grid, hipStream_t st) {
  const size_t shmem = sizeof(int32_t) * ((MM * A.maxn + 3) & ~3) + sizeof(int4) * 256 +
                       static_cast<size_t>(A.maxn) * 64;
  if (subset)
    hipLaunchKernelGGL((k_multi_bound<MM, NW, true>), dim3(grid), dim3(256), shmem, st, A);
  else
    hipLaunchKernelGGL((k_multi_bound<MM, NW, false>), dim3(grid), dim3(256), shmem, st, A);
  return hipGetLastError();
}

int mm_class(int m) { return m <= 8 ? 8 : m <= 12 ? 12 : m <= 16 ? 16 : m <= 20 ? 20 : 24; }

hipError_t launch_multi(const MultiArgs& A, int maxm, bool subset, int grid, hipStream_t st) {
  const int nw = A.maxn <= 32 ? 1 : (A.maxn <= 64 ? 2 : 4);
#define WSBB_MULTI(MMv, NWv) \
  if (mm_class(maxm) == MMv && nw == NWv) return launch_multi_mm<MMv, NWv>(A, subset, grid, st);
  WSBB_MULTI(8, 1) WSBB_MULTI(8, 2) WSBB_MULTI(8, 4)
  WSBB_MULTI(12, 1) WSBB_MULTI(12, 2) WSBB_MULTI(12, 4)
  WSBB_MULTI(16, 1) WSBB_MULTI(16, 2) WSBB_MULTI(16, 4)
  WSBB_MULTI(20, 1) WSBB_MULTI(20, 2) WSBB_MULTI(20, 4)
  WSBB_MULTI(24, 1) WSBB_MULTI(24, 2) WSBB_MULTI(24, 4)
#undef WSBB_MULTI
  return hipErrorInvalidValue;
}

hipError_t launch_multi_features(const MultiFeatArgs& A, int maxm, int grid, hipStream_t st) {
  const size_t shmem = sizeof(int32_t) * 4 * ((mm_class(maxm) * A.maxn + 3) & ~3) +
                       static_cast<size_t>(4) * A.maxn * 64;
  switch (mm_class(maxm)) {
#define WSBB_MF(MMv) \
  case MMv: hipLaunchKernelGGL(k_multi_features<MMv>, dim3(grid), dim3(256), shmem, st, A); break;
    WSBB_MF(8) WSBB_MF(12) WSBB_MF(16) WSBB_MF(20) WSBB_MF(24)
#undef WSBB_MF
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace
}  // namespace wsbb

using namespace wsbb;

namespace {

#define MULTI_HIP(call, what)                         \
  do {                                                \
    hipError_t e_ = (call);                           \
    if (e_ != hipSuccess) return hip_fail(e_, what);  \
  } while (0)

int run_multi(wsbb_multi* M, int kind, int margin, bool have_tsel) {
  hipStream_t st = M->ctx[0]->stream;
  const int grid = std::max(1, std::min(M->ntiles, 8 * M->num_cus));
  if (kind == WSBB_LEARNED && !have_tsel) {
    MultiFeatArgs F{};
    F.tabs = M->tabs.as<const DevTables>();
    F.tile_inst = M->tile_inst;
    F.tile_node0 = M->tile_node0;
    F.tile_cnt = M->tile_cnt;
    F.ntiles = M->ntiles;
    F.maxn = M->maxn;
    F.N = M->N;
    F.fstride = M->fstride;
    F.permT = M->permT;
    F.l1 = M->l1;
    F.l2 = M->l2;
    F.scale = reinterpret_cast<const float*>(M->tile_list + M->ntiles);  // per-instance scales
    F.frag = M->xfrag;
    MULTI_HIP(launch_multi_features(F, M->maxm, std::max(1, std::min((M->ntiles + 3) / 4, 4 * M->num_cus)), st),
              "multi features launch");
    for (size_t c = 0; c < M->class_m.size(); ++c) {
      MlpTiles tl{M->tile_list + M->class_first[c], M->class_count[c], M->tile_cnt, M->fstride, M->nsw};
      MULTI_HIP(launch_mlp_select(M->ctx[M->class_rep[c]], M->xfrag, 64 * M->ntiles, nullptr, M->tsel,
                                  nullptr, 0, &tl),
                "multi mlp launch");
    }
  }
  MultiArgs A{};
  A.tabs = M->tabs.as<const DevTables>();
  A.tile_inst = M->tile_inst;
  A.tile_node0 = M->tile_node0;
  A.tile_cnt = M->tile_cnt;
  A.ntiles = M->ntiles;
  A.maxn = M->maxn;
  A.N = M->N;
  A.permT = M->permT;
  A.l1 = M->l1;
  A.l2 = M->l2;
  A.out = M->out;
  A.tile_sel = kind == WSBB_LEARNED ? M->tsel : nullptr;
  A.nsw = M->nsw;
  A.inc = M->inc;
  A.margin = margin;
  A.stats = M->stats;
  MULTI_HIP(launch_multi(A, M->maxm, kind == WSBB_LEARNED, grid, st), "multi bound launch");
  return WSBB_OK;
}

}  // namespace

extern "C" {

int wsbb_multi_create(int count, const int32_t* const* p, const int* n, const int* m, int device,
                      wsbb_multi** out) {
  if (!out || count < 1 || !p || !n || !m) {
    set_error("bad multi-instance arguments");
    return WSBB_EINVAL;
  }
  *out = nullptr;
  auto* M = new wsbb_multi();
  M->device = device;
  std::vector<DevTables> host(count);
  for (int i = 0; i < count; ++i) {
    wsbb_ctx* c = nullptr;
    const int rc = wsbb_create(p[i], n[i], m[i], device, &c);
    if (rc != WSBB_OK) {
      wsbb_multi_destroy(M);
      return rc;
    }
    M->ctx.push_back(c);
    host[i] = c->T;
    M->maxn = std::max(M->maxn, n[i]);
    M->maxm = std::max(M->maxm, m[i]);
    M->nsw = std::max(M->nsw, (c->T.P + 31) / 32);
    M->fstride = std::max(M->fstride, (5 * m[i] + 7) / 8);
  }
  M->num_cus = M->ctx[0]->num_cus;
  hipError_t e = M->tabs.reserve(sizeof(DevTables) * count);
  if (e == hipSuccess)
    e = hipMemcpy(M->tabs.ptr, host.data(), sizeof(DevTables) * count, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipEventCreate(&M->ev0);
  if (e == hipSuccess) e = hipEventCreate(&M->ev1);
  if (e != hipSuccess) {
    wsbb_multi_destroy(M);
    return hip_fail(e, "multi tables");
  }
  *out = M;
  return WSBB_OK;
}

int wsbb_multi_destroy(wsbb_multi* M) {
  if (!M) return WSBB_OK;
  (void)hipSetDevice(M->device);
  if (!M->ctx.empty()) (void)hipStreamSynchronize(M->ctx[0]->stream);
  M->batch.release();
  M->tabs.release();
  if (M->ev0) (void)hipEventDestroy(M->ev0);
  if (M->ev1) (void)hipEventDestroy(M->ev1);
  for (wsbb_ctx* c : M->ctx) wsbb_destroy(c);
  delete M;
  return WSBB_OK;
}

int wsbb_multi_context(wsbb_multi* M, int i, wsbb_ctx** out) {
  if (!M || !out || i < 0 || i >= static_cast<int>(M->ctx.size())) {
    set_error("bad instance index");
    return WSBB_EINVAL;
  }
  *out = M->ctx[i];
  return WSBB_OK;
}

int wsbb_multi_load(wsbb_multi* M, const uint8_t* permT, const int32_t* l1, const int32_t* l2,
                    const int32_t* offsets, int N) {
  if (!M || !permT || !l1 || !l2 || !offsets || N < 1) {
    set_error("bad multi batch arguments");
    return WSBB_EINVAL;
  }
  const int count = static_cast<int>(M->ctx.size());
  if (offsets[0] != 0 || offsets[count] != N) {
    set_error("offsets must run from 0 to N");
    return WSBB_EINVAL;
  }
  // tile map: instance i's nodes [offsets[i], offsets[i+1]) in 64-node tiles,
  // tiles grouped by machine-count class (for the per-class MLP launches)
  std::vector<int> tinst, tnode0, tcnt;
  for (int i = 0; i < count; ++i) {
    if (offsets[i + 1] < offsets[i]) {
      set_error("offsets must not decrease");
      return WSBB_EINVAL;
    }
    const int ni = M->ctx[i]->T.n;
    for (int k = offsets[i]; k < offsets[i + 1]; ++k) {
      bool ok = l1[k] >= -1 && l2[k] <= ni && l1[k] < l2[k];
      for (int r = 0; ok && r < ni; ++r) ok = permT[static_cast<size_t>(r) * N + k] < ni;
      if (!ok) {
        set_error("node " + std::to_string(k) + " has bad limits or job ids for its instance");
        return WSBB_EINVAL;
      }
    }
    for (int k = offsets[i]; k < offsets[i + 1]; k += 64) {
      tinst.push_back(i);
      tnode0.push_back(k);
      tcnt.push_back(std::min(64, offsets[i + 1] - k));
    }
  }
  const int ntiles = static_cast<int>(tinst.size());
  M->class_m.clear();
  M->class_first.clear();
  M->class_count.clear();
  M->class_rep.clear();
  std::vector<int> tlist;
  for (int i = 0; i < count; ++i) {
    const int mi = M->ctx[i]->T.m;
    if (std::find(M->class_m.begin(), M->class_m.end(), mi) != M->class_m.end()) continue;
    M->class_m.push_back(mi);
    M->class_first.push_back(static_cast<int>(tlist.size()));
    M->class_rep.push_back(i);
    for (int t = 0; t < ntiles; ++t)
      if (M->ctx[tinst[t]]->T.m == mi) tlist.push_back(t);
    M->class_count.push_back(static_cast<int>(tlist.size()) - M->class_first.back());
  }
  std::vector<float> scale(count);
  for (int i = 0; i < count; ++i)
    scale[i] = static_cast<float>(M->ctx[i]->inst.path_bound > 0 ? M->ctx[i]->inst.path_bound : 1);
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  size_t off = 0;
  auto take = [&](size_t b) {
    const size_t o = off;
    off += al(b);
    return o;
  };
  const size_t o_perm = take(static_cast<size_t>(M->maxn) * N);
  const size_t o_l1 = take(4 * size_t(N)), o_l2 = take(4 * size_t(N)), o_out = take(4 * size_t(N));
  const size_t o_ti = take(4 * size_t(ntiles)), o_tn = take(4 * size_t(ntiles)),
               o_tc = take(4 * size_t(ntiles));
  const size_t o_tl = take(4 * size_t(ntiles) + 4 * size_t(count));  // list, then scales
  const size_t o_inc = take(4 * size_t(count));
  const size_t o_ts = take(4 * size_t(ntiles) * M->nsw);
  const size_t o_xf = take(sizeof(uint2) * 2 * size_t(ntiles) * M->fstride * 64);
  const size_t o_st = take(8 * 4);
  MULTI_HIP(hipSetDevice(M->device), "hipSetDevice");
  MULTI_HIP(M->batch.reserve(off), "hipMalloc(multi batch)");
  M->N = N;
  M->ntiles = ntiles;
  M->permT = M->batch.as<uint8_t>(o_perm);
  M->l1 = M->batch.as<int32_t>(o_l1);
  M->l2 = M->batch.as<int32_t>(o_l2);
  M->out = M->batch.as<int32_t>(o_out);
  M->tile_inst = M->batch.as<int>(o_ti);
  M->tile_node0 = M->batch.as<int>(o_tn);
  M->tile_cnt = M->batch.as<int>(o_tc);
  M->tile_list = M->batch.as<int>(o_tl);
  M->inc = M->batch.as<int32_t>(o_inc);
  M->tsel = M->batch.as<uint32_t>(o_ts);
  M->xfrag = M->batch.as<uint2>(o_xf);
  M->stats = M->batch.as<long long>(o_st);
  hipStream_t st = M->ctx[0]->stream;
  // rows past an instance's n are never read
  for (int r = 0; r < M->maxn; ++r)
    MULTI_HIP(hipMemcpyAsync(M->permT + static_cast<size_t>(r) * N, permT + static_cast<size_t>(r) * N, N,
                             hipMemcpyHostToDevice, st), "H2D perm");
  MULTI_HIP(hipMemcpyAsync(M->l1, l1, 4 * size_t(N), hipMemcpyHostToDevice, st), "H2D l1");
  MULTI_HIP(hipMemcpyAsync(M->l2, l2, 4 * size_t(N), hipMemcpyHostToDevice, st), "H2D l2");
  MULTI_HIP(hipMemcpyAsync(M->tile_inst, tinst.data(), 4 * size_t(ntiles), hipMemcpyHostToDevice, st), "H2D map");
  MULTI_HIP(hipMemcpyAsync(M->tile_node0, tnode0.data(), 4 * size_t(ntiles), hipMemcpyHostToDevice, st), "H2D map");
  MULTI_HIP(hipMemcpyAsync(M->tile_cnt, tcnt.data(), 4 * size_t(ntiles), hipMemcpyHostToDevice, st), "H2D map");
  MULTI_HIP(hipMemcpyAsync(M->tile_list, tlist.data(), 4 * size_t(ntiles), hipMemcpyHostToDevice, st), "H2D list");
  MULTI_HIP(hipMemcpyAsync(M->tile_list + ntiles, scale.data(), 4 * size_t(count), hipMemcpyHostToDevice, st),
            "H2D scales");
  M->offsets.assign(offsets, offsets + count + 1);
  MULTI_HIP(hipStreamSynchronize(st), "sync");
  return WSBB_OK;
}

int wsbb_multi_run(wsbb_multi* M, int kind, const int32_t* inc, int margin,
                   const uint32_t* tile_sel, int reps, int32_t* out, int64_t* stats, float* ms) {
  if (!M || M->ntiles == 0) {
    set_error("no batch loaded (wsbb_multi_load)");
    return WSBB_EINVAL;
  }
  if ((kind != WSBB_LB2 && kind != WSBB_LEARNED) || reps < 1 || margin < -1) {
    set_error("kind must be WSBB_LB2 or WSBB_LEARNED, reps >= 1, margin >= -1");
    return WSBB_EINVAL;
  }
  const int count = static_cast<int>(M->ctx.size());
  if (kind == WSBB_LEARNED && !tile_sel)
    for (int rep : M->class_rep)
      if (!selector_ready(M->ctx[rep])) {
        set_error("learned multi bound needs the pair selector of instance " + std::to_string(rep));
        return WSBB_EINVAL;
      }
  hipStream_t st = M->ctx[0]->stream;
  MULTI_HIP(hipSetDevice(M->device), "hipSetDevice");
  std::vector<int32_t> incv(count, INT32_MAX);
  if (inc) incv.assign(inc, inc + count);
  MULTI_HIP(hipMemcpyAsync(M->inc, incv.data(), 4 * size_t(count), hipMemcpyHostToDevice, st), "H2D inc");
  if (tile_sel)
    MULTI_HIP(hipMemcpyAsync(M->tsel, tile_sel, 4 * size_t(M->ntiles) * M->nsw, hipMemcpyHostToDevice, st),
              "H2D tile masks");
  MULTI_HIP(hipMemsetAsync(M->stats, 0, 8 * 4, st), "stats clear");
  MULTI_HIP(hipEventRecord(M->ev0, st), "event");
  for (int r = 0; r < reps; ++r) {
    const int rc = run_multi(M, kind, inc ? margin : -1, tile_sel != nullptr);
    if (rc != WSBB_OK) return rc;
  }
  MULTI_HIP(hipEventRecord(M->ev1, st), "event");
  MULTI_HIP(hipEventSynchronize(M->ev1), "sync");
  if (ms) {
    float t = 0;
    MULTI_HIP(hipEventElapsedTime(&t, M->ev0, M->ev1), "elapsed");
    *ms = t / reps;
  }
  if (out) MULTI_HIP(hipMemcpy(out, M->out, 4 * size_t(M->N), hipMemcpyDeviceToHost), "D2H out");
  if (stats) {
    long long h[3];
    MULTI_HIP(hipMemcpy(h, M->stats, sizeof(h), hipMemcpyDeviceToHost), "D2H stats");
    for (int i = 0; i < 3; ++i) stats[i] = h[i];
  }
  return WSBB_OK;
}

}  // extern "C"
