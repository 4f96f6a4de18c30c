// C ABI: contexts, instance tables, batch bounding, device memory, timing and
// launch-shape autotuning.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/wsbb_amd.h"
#include "bound.h"
#include "context.h"

namespace wsbb {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int hip_fail(hipError_t e, const char* what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return WSBB_EHIP;
}

void destroy_search(wsbb_ctx* ctx);    // search.hip
void destroy_selector(wsbb_ctx* ctx);  // selector.hip

// Staging of wsbb_bound_batch_noalloc: a file-scope device array, so that
// entry point never allocates.  Each GPU the library runs on holds its own
// copy until the process ends.  Host code must not take the array's name as
// a device address: its address on each device is read once, with that
// device current, by hipGetSymbolAddress and kept in g_scratch_ptr.
constexpr size_t kScratchBytes = size_t(8) << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];
static std::mutex g_scratch_mu;
static uint8_t* g_scratch_ptr[kMaxDevices] = {};

// with `device` current: resolve that device's copy of the array once
static int scratch_setup(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) {
    set_error("device index beyond the no-allocation staging table");
    return WSBB_EINVAL;
  }
  if (g_scratch_ptr[device]) return WSBB_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
  g_scratch_ptr[device] = static_cast<uint8_t*>(ptr);
  return WSBB_OK;
}

static uint8_t* scratch_of(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return (device >= 0 && device < kMaxDevices) ? g_scratch_ptr[device] : nullptr;
}

}  // namespace wsbb

using namespace wsbb;

int wsbb_ctx::groups_for(int N) const {
  if (launch.groups > 1) return launch.groups;
  const int tiles = (N + 63) / 64;
  int cls = 0;
  while ((1 << (cls + 1)) <= tiles && cls < 31) ++cls;
  if (tuned_groups[cls] > 0) return tuned_groups[cls];
  // enough wavefronts to give every SIMD of every CU at least one
  const int want = num_cus * 4;
  if (tiles >= want) return 1;
  if (tiles * 2 >= want) return 2;
  return 4;
}

#define WSBB_CHECK_CTX(ctx)                 \
  do {                                      \
    if ((ctx) == nullptr) {                 \
      set_error("null context");            \
      return WSBB_EINVAL;                   \
    }                                       \
  } while (0)

#define WSBB_HIP(call, what)                          \
  do {                                                \
    hipError_t e_ = (call);                           \
    if (e_ != hipSuccess) return hip_fail(e_, what);  \
  } while (0)

extern "C" {

int wsbb_version(void) { return 1; }

const char* wsbb_last_error(void) { return g_last_error.c_str(); }

int wsbb_device_count(int* count) {
  if (!count) return WSBB_EINVAL;
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  *count = (e == hipSuccess) ? c : 0;
  return WSBB_OK;
}

int wsbb_taillard(int id, int* n, int* m, int32_t* p) {
  long seed;
  int nn, mm;
  if (!taillard_dims(id, &nn, &mm, &seed)) {
    set_error("Taillard id must be in [1, 120]");
    return WSBB_EINVAL;
  }
  if (n) *n = nn;
  if (m) *m = mm;
  if (p) taillard_generate(nn, mm, seed, p);
  return WSBB_OK;
}

int wsbb_generate(int n, int m, long seed, int32_t* p) {
  if (n < 1 || m < 1 || !p) return WSBB_EINVAL;
  taillard_generate(n, m, seed, p);
  return WSBB_OK;
}

int wsbb_generate_correlated(int n, int m, uint64_t seed, int32_t* p) {
  if (n < 1 || m < 1 || !p) return WSBB_EINVAL;
  correlated_generate(n, m, seed, p);
  return WSBB_OK;
}

int wsbb_generate_machine_correlated(int n, int m, uint64_t seed, int32_t* p) {
  if (n < 1 || m < 1 || !p) return WSBB_EINVAL;
  machine_correlated_generate(n, m, seed, p);
  return WSBB_OK;
}

int wsbb_random_nodes(int n, int N, uint64_t seed0, uint8_t* permT,
                      int32_t* l1, int32_t* l2) {
  if (n < 2 || n > 255 || N < 0 || !permT || !l1 || !l2) return WSBB_EINVAL;
  random_nodes(n, N, seed0, permT, l1, l2);
  return WSBB_OK;
}

int wsbb_precompute(const int32_t* p, int n, int m, int32_t* pair_a,
                    int32_t* pair_b, int32_t* lag, int32_t* johnson,
                    int32_t* R0, int32_t* Q0) {
  if (!p) return WSBB_EINVAL;
  Instance I;
  std::string err = I.build(p, n, m);
  if (!err.empty()) {
    set_error(err);
    return WSBB_EINVAL;
  }
  auto cp = [](int32_t* dst, const std::vector<int32_t>& v) {
    if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(int32_t));
  };
  cp(pair_a, I.pair_a);
  cp(pair_b, I.pair_b);
  cp(lag, I.lag);
  cp(johnson, I.johnson);
  cp(R0, I.R0);
  cp(Q0, I.Q0);
  return WSBB_OK;
}

int wsbb_create(const int32_t* p, int n, int m, int device, wsbb_ctx** out) {
  if (!p || !out) {
    set_error("null argument");
    return WSBB_EINVAL;
  }
  *out = nullptr;
  // the instance is checked (host only) before any device is touched
  auto* ctx = new wsbb_ctx();
  for (int& g : ctx->tuned_groups) g = -1;
  std::string err = ctx->inst.build(p, n, m);
  if (!err.empty()) {
    set_error(err);
    delete ctx;
    return WSBB_EINVAL;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_error("no HIP device available");
    delete ctx;
    return WSBB_ENODEV;
  }
  if (device < 0 || device >= ndev) {
    set_error("device index out of range");
    delete ctx;
    return WSBB_EINVAL;
  }
  ctx->device = device;
  auto fail = [&](hipError_t e, const char* what) {
    int rc = hip_fail(e, what);
    wsbb_destroy(ctx);
    return rc;
  };
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return fail(e, "hipSetDevice");
  if (int rc = scratch_setup(device)) {
    wsbb_destroy(ctx);
    return rc;
  }
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) return fail(e, "hipGetDeviceProperties");
  ctx->num_cus = prop.multiProcessorCount;
  if ((e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess)
    return fail(e, "hipStreamCreate");
  if ((e = hipEventCreate(&ctx->ev_start)) != hipSuccess) return fail(e, "hipEventCreate");
  if ((e = hipEventCreate(&ctx->ev_stop)) != hipSuccess) return fail(e, "hipEventCreate");

  // Pack every table into one upload.  steps[s][k] = {job, p_a, p_b,
  // kBias + lag + p_b} in Johnson order, wsel[s][k] the masks that pick the
  // job's 32-bit word of a node's job mask; padding steps are all zero
  // (no-ops), and 32 more follow the last pair for the scans' read-ahead.
  const Instance& I = ctx->inst;
  const int P = I.P, n4 = (n + 3) & ~3;
  const size_t nsteps = static_cast<size_t>(P) * n4 + 32;
  std::vector<int4> steps(nsteps, int4{0, 0, 0, 0});
  std::vector<uint2> wsel(nsteps, uint2{0u, 0u});
  for (int s = 0; s < P; ++s)
    for (int k = 0; k < n; ++k) {
      const int j = I.johnson[s * n + k];
      const int pa = I.p[I.pair_a[s] * n + j], pb = I.p[I.pair_b[s] * n + j];
      const size_t at = static_cast<size_t>(s) * n4 + k;
      steps[at] = int4{j, pa, pb, kBias + I.lag[s * n + j] + pb};
      wsel[at] = uint2{((j >> 5) & 1) ? 0xffffffffu : 0u, ((j >> 6) & 1) ? 0xffffffffu : 0u};
    }
  // two nodes per lane in int16 halves when every bound value (<= the path
  // bound) and every step constant stays below 2^14
  const bool two16 = I.path_bound < kBias16 - 1;
  std::vector<uint2> steps16(two16 ? nsteps : 0, uint2{0u, 0u});
  if (two16)
    for (int s = 0; s < P; ++s)
      for (int k = 0; k < n; ++k) {
        const size_t at = static_cast<size_t>(s) * n4 + k;
        const int4 w = steps[at];
        const int K = kBias16 + (w.w - kBias);  // 2^14 + lag + p_b
        const uint32_t dab = static_cast<uint32_t>(w.y - w.z) & 0xffffu;  // p_a - p_b
        steps16[at] = uint2{static_cast<uint32_t>(w.x), dab | (static_cast<uint32_t>(K) << 16)};
      }
  // int32 offset form: per step 32 (p_a - p_b) and 32 (lag + p_b) plus the
  // job's bit offset (the scan's clocks carry five low bits that never reach
  // a bound: every value below kPathLimit keeps 32 x value + 31 in 24 bits),
  // and for n > 32 the word-select masks of the job
  const int rs = n <= 32 ? 2 : (n <= 64 ? 3 : 4);
  std::vector<int32_t> steps32(nsteps * rs, 0);
  for (int s = 0; s < P; ++s)
    for (int k = 0; k < n; ++k) {
      const size_t at = static_cast<size_t>(s) * n4 + k;
      const int4 w = steps[at];
      const int j = w.x, pa = w.y, pb = w.z, lag = w.w - kBias - pb;
      int32_t* rec = &steps32[at * rs];
      rec[0] = (pa - pb) * (1 << kScaleShift);
      rec[1] = (lag + pb) * (1 << kScaleShift) + (j & 31);
      if (rs >= 3) rec[2] = static_cast<int32_t>(wsel[at].x);
      if (rs == 4) rec[3] = static_cast<int32_t>(wsel[at].y);
    }
  std::vector<int2> pairs(P);
  for (int s = 0; s < P; ++s) pairs[s] = int2{I.pair_a[s], I.pair_b[s]};
  const size_t o_steps = 0;
  const size_t o_wsel = o_steps + steps.size() * sizeof(int4);
  auto al256 = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t o_s32 = al256(o_wsel + wsel.size() * sizeof(uint2));
  const size_t o_s16 = al256(o_s32 + steps32.size() * sizeof(int32_t));
  const size_t o_pairs = al256(o_s16 + steps16.size() * sizeof(uint2));
  const size_t o_p = (o_pairs + pairs.size() * sizeof(int2) + 255) & ~size_t(255);
  const size_t o_R0 = o_p + ((I.p.size() * 4 + 255) & ~size_t(255));
  const size_t o_Q0 = o_R0 + 256;
  const size_t o_tot = o_Q0 + 256;
  const size_t total = o_tot + 256;
  if ((e = ctx->tables.reserve(total)) != hipSuccess) return fail(e, "hipMalloc(tables)");
  std::vector<char> blob(total, 0);
  std::memcpy(blob.data() + o_steps, steps.data(), steps.size() * sizeof(int4));
  std::memcpy(blob.data() + o_wsel, wsel.data(), wsel.size() * sizeof(uint2));
  std::memcpy(blob.data() + o_s32, steps32.data(), steps32.size() * sizeof(int32_t));
  if (two16) std::memcpy(blob.data() + o_s16, steps16.data(), steps16.size() * sizeof(uint2));
  std::memcpy(blob.data() + o_pairs, pairs.data(), pairs.size() * sizeof(int2));
  std::memcpy(blob.data() + o_p, I.p.data(), I.p.size() * 4);
  std::memcpy(blob.data() + o_R0, I.R0.data(), m * 4);
  std::memcpy(blob.data() + o_Q0, I.Q0.data(), m * 4);
  std::memcpy(blob.data() + o_tot, I.ptot.data(), m * 4);
  if ((e = hipMemcpy(ctx->tables.ptr, blob.data(), total, hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e, "hipMemcpy(tables)");
  DevTables& T = ctx->T;
  T.n = n;
  T.m = m;
  T.P = P;
  T.n4 = n4;
  T.steps = ctx->tables.as<const int4>(o_steps);
  T.wsel = ctx->tables.as<const uint2>(o_wsel);
  T.steps32 = ctx->tables.as<const int>(o_s32);
  T.rs = rs;
  T.pairs = ctx->tables.as<const int2>(o_pairs);
  T.p = ctx->tables.as<const int32_t>(o_p);
  T.R0 = ctx->tables.as<const int32_t>(o_R0);
  T.Q0 = ctx->tables.as<const int32_t>(o_Q0);
  T.ptot = ctx->tables.as<const int32_t>(o_tot);
  T.steps16 = two16 ? ctx->tables.as<const uint2>(o_s16) : nullptr;
  *out = ctx;
  return WSBB_OK;
}

int wsbb_destroy(wsbb_ctx* ctx) {
  if (!ctx) return WSBB_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  destroy_search(ctx);
  destroy_selector(ctx);
  ctx->tables.release();
  ctx->batch.release();
  if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
  if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return WSBB_OK;
}

int wsbb_info(const wsbb_ctx* ctx, int* n, int* m, int* pairs, int* num_cus) {
  WSBB_CHECK_CTX(ctx);
  if (n) *n = ctx->T.n;
  if (m) *m = ctx->T.m;
  if (pairs) *pairs = ctx->T.P;
  if (num_cus) *num_cus = ctx->num_cus;
  return WSBB_OK;
}

static int check_nodes_host(const wsbb_ctx* ctx, const uint8_t* permT,
                            const int32_t* l1, const int32_t* l2, int N) {
  const int n = ctx->T.n;
  for (int k = 0; k < N; ++k) {
    if (l1[k] < -1 || l2[k] > n || l2[k] <= l1[k]) {
      set_error("node " + std::to_string(k) + ": limits out of range");
      return WSBB_EINVAL;
    }
  }
  for (size_t i = 0; i < static_cast<size_t>(n) * N; ++i)
    if (permT[i] >= n) {
      set_error("job id out of range at position-major index " + std::to_string(i));
      return WSBB_EINVAL;
    }
  return WSBB_OK;
}

int wsbb_bound_batch_ex(wsbb_ctx* ctx, const uint8_t* permT, const int32_t* l1,
                        const int32_t* l2, int N, int kind, int best,
                        int32_t* out, int32_t* argmax, int32_t* pairv,
                        const uint32_t* sel, float* feat) {
  WSBB_CHECK_CTX(ctx);
  if (N < 0 || (N > 0 && (!permT || !l1 || !l2 || !out)) || (kind != 0 && kind != 1)) {
    set_error("bad batch arguments");
    return WSBB_EINVAL;
  }
  if (N == 0) return WSBB_OK;
  int rc = check_nodes_host(ctx, permT, l1, l2, N);
  if (rc) return rc;
  const int n = ctx->T.n, m = ctx->T.m, P = ctx->T.P;
  const int nsw = (P + 31) / 32;
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t b_perm = al(size_t(n) * N), b_lim = al(size_t(4) * N);
  const size_t o_perm = 0, o_l1 = b_perm, o_l2 = o_l1 + b_lim, o_out = o_l2 + b_lim;
  const size_t o_am = o_out + b_lim;
  const size_t o_pv = o_am + (argmax ? b_lim : 0);
  const size_t o_sel = o_pv + (pairv ? al(size_t(4) * P * N) : 0);
  const size_t o_feat = o_sel + (sel ? al(size_t(4) * nsw * N) : 0);
  const size_t total = o_feat + (feat ? al(size_t(4) * 5 * m * N) : 0);
  WSBB_HIP(hipSetDevice(ctx->device), "hipSetDevice");
  WSBB_HIP(ctx->batch.reserve(total), "hipMalloc(batch)");
  hipStream_t st = ctx->stream;
  WSBB_HIP(hipMemcpyAsync(ctx->batch.as<char>(o_perm), permT, size_t(n) * N, hipMemcpyHostToDevice, st), "H2D perm");
  WSBB_HIP(hipMemcpyAsync(ctx->batch.as<char>(o_l1), l1, size_t(4) * N, hipMemcpyHostToDevice, st), "H2D l1");
  WSBB_HIP(hipMemcpyAsync(ctx->batch.as<char>(o_l2), l2, size_t(4) * N, hipMemcpyHostToDevice, st), "H2D l2");
  if (sel)
    WSBB_HIP(hipMemcpyAsync(ctx->batch.as<char>(o_sel), sel, size_t(4) * nsw * N, hipMemcpyHostToDevice, st), "H2D sel");
  BatchSource src{ctx->batch.as<uint8_t>(o_perm), ctx->batch.as<int32_t>(o_l1),
                  ctx->batch.as<int32_t>(o_l2), N};
  BoundArgs A{};
  A.N = N;
  A.kind = kind;
  A.best = best;
  A.groups = ctx->groups_for(N);
  A.pk = ctx->int16_for(N);
  A.out = ctx->batch.as<int32_t>(o_out);
  A.argmax = argmax ? ctx->batch.as<int32_t>(o_am) : nullptr;
  A.pairv = pairv ? ctx->batch.as<int32_t>(o_pv) : nullptr;
  A.sel = sel ? ctx->batch.as<const uint32_t>(o_sel) : nullptr;
  A.feat = feat ? ctx->batch.as<float>(o_feat) : nullptr;
  A.feat_scale = static_cast<float>(ctx->inst.path_bound > 0 ? ctx->inst.path_bound : 1);
  WSBB_HIP(launch_bound(ctx->T, src, A, ctx->grid_cap(), st), "bound kernel launch");
  WSBB_HIP(hipMemcpyAsync(out, A.out, size_t(4) * N, hipMemcpyDeviceToHost, st), "D2H out");
  if (argmax) WSBB_HIP(hipMemcpyAsync(argmax, A.argmax, size_t(4) * N, hipMemcpyDeviceToHost, st), "D2H argmax");
  if (pairv) WSBB_HIP(hipMemcpyAsync(pairv, A.pairv, size_t(4) * P * N, hipMemcpyDeviceToHost, st), "D2H pairv");
  if (feat) WSBB_HIP(hipMemcpyAsync(feat, A.feat, size_t(4) * 5 * m * N, hipMemcpyDeviceToHost, st), "D2H feat");
  WSBB_HIP(hipStreamSynchronize(st), "hipStreamSynchronize");
  return WSBB_OK;
}

int wsbb_bound_tiles(wsbb_ctx* ctx, const uint8_t* permT, const int32_t* l1,
                     const int32_t* l2, int N, const uint32_t* tile_sel, int margin,
                     int32_t inc, int32_t* out, int64_t* stats) {
  WSBB_CHECK_CTX(ctx);
  if (N <= 0 || !permT || !l1 || !l2 || !tile_sel || !out) {
    set_error("bad tile-bound arguments");
    return WSBB_EINVAL;
  }
  int rc = check_nodes_host(ctx, permT, l1, l2, N);
  if (rc) return rc;
  const int n = ctx->T.n, nsw = (ctx->T.P + 31) / 32, ntiles = (N + 63) / 64;
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t o_perm = 0, o_l1 = al(size_t(n) * N), o_l2 = o_l1 + al(4 * size_t(N));
  const size_t o_out = o_l2 + al(4 * size_t(N)), o_ts = o_out + al(4 * size_t(N));
  const size_t o_inc = o_ts + al(4 * size_t(ntiles) * nsw), o_st = o_inc + 256;
  const size_t total = o_st + 256;
  WSBB_HIP(hipSetDevice(ctx->device), "hipSetDevice");
  WSBB_HIP(ctx->batch.reserve(total), "hipMalloc(batch)");
  hipStream_t st = ctx->stream;
  long long zero[4] = {0, 0, 0, 0};
  WSBB_HIP(hipMemcpyAsync(ctx->batch.as<char>(o_perm), permT, size_t(n) * N, hipMemcpyHostToDevice, st), "H2D");
  WSBB_HIP(hipMemcpyAsync(ctx->batch.as<char>(o_l1), l1, 4 * size_t(N), hipMemcpyHostToDevice, st), "H2D");
  WSBB_HIP(hipMemcpyAsync(ctx->batch.as<char>(o_l2), l2, 4 * size_t(N), hipMemcpyHostToDevice, st), "H2D");
  WSBB_HIP(hipMemcpyAsync(ctx->batch.as<char>(o_ts), tile_sel, 4 * size_t(ntiles) * nsw, hipMemcpyHostToDevice, st), "H2D");
  WSBB_HIP(hipMemcpyAsync(ctx->batch.as<char>(o_inc), &inc, 4, hipMemcpyHostToDevice, st), "H2D");
  WSBB_HIP(hipMemcpyAsync(ctx->batch.as<char>(o_st), zero, sizeof(zero), hipMemcpyHostToDevice, st), "H2D");
  BatchSource src{ctx->batch.as<uint8_t>(o_perm), ctx->batch.as<int32_t>(o_l1),
                  ctx->batch.as<int32_t>(o_l2), N};
  BoundArgs A{};
  A.N = N;
  A.kind = 0;
  A.best = -1;
  A.groups = 1;
  A.out = ctx->batch.as<int32_t>(o_out);
  A.tile_sel = ctx->batch.as<const uint32_t>(o_ts);
  A.nsw = nsw;
  A.margin = margin;
  A.inc_dev = ctx->batch.as<const int>(o_inc);
  A.stats = ctx->batch.as<long long>(o_st);
  WSBB_HIP(launch_bound(ctx->T, src, A, ctx->grid_cap(), st), "bound kernel launch");
  WSBB_HIP(hipMemcpyAsync(out, A.out, 4 * size_t(N), hipMemcpyDeviceToHost, st), "D2H");
  WSBB_HIP(hipMemcpyAsync(zero, A.stats, sizeof(zero), hipMemcpyDeviceToHost, st), "D2H");
  WSBB_HIP(hipStreamSynchronize(st), "hipStreamSynchronize");
  if (stats)
    for (int i = 0; i < 3; ++i) stats[i] = zero[i];
  return WSBB_OK;
}

int wsbb_bound_batch(wsbb_ctx* ctx, const uint8_t* permT, const int32_t* l1,
                     const int32_t* l2, int N, int kind, int best, int32_t* out) {
  return wsbb_bound_batch_ex(ctx, permT, l1, l2, N, kind, best, out, nullptr,
                             nullptr, nullptr, nullptr);
}

int wsbb_bound_batch_noalloc(wsbb_ctx* ctx, const uint8_t* permT, const int32_t* l1,
                             const int32_t* l2, int N, int kind, int best, int32_t* out) {
  WSBB_CHECK_CTX(ctx);
  if (N < 0 || (N > 0 && (!permT || !l1 || !l2 || !out)) || (kind != 0 && kind != 1)) {
    set_error("bad batch arguments");
    return WSBB_EINVAL;
  }
  if (N == 0) return WSBB_OK;
  const int n = ctx->T.n;
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t o_l1 = al(size_t(n) * N), o_l2 = o_l1 + al(size_t(4) * N);
  const size_t o_out = o_l2 + al(size_t(4) * N), total = o_out + al(size_t(4) * N);
  if (total > kScratchBytes) {
    // the largest N whose staging fits (each segment rounded up to 256 B)
    const long long nmax = static_cast<long long>(kScratchBytes - 4 * 256) / (n + 12);
    set_error("wsbb_bound_batch_noalloc: " + std::to_string(N) + " nodes of " +
              std::to_string(n) + " jobs need " + std::to_string(total) +
              " bytes of staging; the fixed buffer holds " + std::to_string(kScratchBytes) +
              " (about " + std::to_string(nmax) + " nodes)");
    return WSBB_EINVAL;
  }
  int rc = check_nodes_host(ctx, permT, l1, l2, N);
  if (rc) return rc;
  uint8_t* base = scratch_of(ctx->device);
  if (!base) {
    set_error("no-allocation staging was not set up on this device");
    return WSBB_EHIP;
  }
  WSBB_HIP(hipSetDevice(ctx->device), "hipSetDevice");
  hipStream_t st = ctx->stream;
  WSBB_HIP(hipMemcpyAsync(base, permT, size_t(n) * N, hipMemcpyHostToDevice, st), "H2D perm");
  WSBB_HIP(hipMemcpyAsync(base + o_l1, l1, size_t(4) * N, hipMemcpyHostToDevice, st), "H2D l1");
  WSBB_HIP(hipMemcpyAsync(base + o_l2, l2, size_t(4) * N, hipMemcpyHostToDevice, st), "H2D l2");
  BatchSource src{base, reinterpret_cast<const int32_t*>(base + o_l1),
                  reinterpret_cast<const int32_t*>(base + o_l2), N};
  BoundArgs A{};
  A.N = N;
  A.kind = kind;
  A.best = best;
  A.groups = ctx->groups_for(N);
  A.pk = ctx->int16_for(N);
  A.out = reinterpret_cast<int32_t*>(base + o_out);
  WSBB_HIP(launch_bound(ctx->T, src, A, ctx->grid_cap(), st), "bound kernel launch");
  WSBB_HIP(hipMemcpyAsync(out, A.out, size_t(4) * N, hipMemcpyDeviceToHost, st), "D2H out");
  WSBB_HIP(hipStreamSynchronize(st), "hipStreamSynchronize");
  return WSBB_OK;
}

int wsbb_bound_device(wsbb_ctx* ctx, const uint8_t* d_permT, const int32_t* d_l1,
                      const int32_t* d_l2, int N, int kind, int32_t* d_out) {
  WSBB_CHECK_CTX(ctx);
  if (N < 0 || (N > 0 && (!d_permT || !d_l1 || !d_l2 || !d_out)) || (kind != 0 && kind != 1)) {
    set_error("bad batch arguments");
    return WSBB_EINVAL;
  }
  BatchSource src{d_permT, d_l1, d_l2, N};
  BoundArgs A{};
  A.N = N;
  A.kind = kind;
  A.best = -1;
  A.groups = ctx->groups_for(N);
  A.pk = ctx->int16_for(N);
  A.out = d_out;
  WSBB_HIP(launch_bound(ctx->T, src, A, ctx->grid_cap(), ctx->stream), "bound kernel launch");
  return WSBB_OK;
}

int wsbb_malloc(wsbb_ctx* ctx, size_t bytes, void** d_ptr) {
  WSBB_CHECK_CTX(ctx);
  if (!d_ptr) return WSBB_EINVAL;
  WSBB_HIP(hipSetDevice(ctx->device), "hipSetDevice");
  hipError_t e = hipMalloc(d_ptr, bytes ? bytes : 1);
  if (e != hipSuccess) {
    hip_fail(e, "hipMalloc");
    return WSBB_ENOMEM;
  }
  return WSBB_OK;
}

int wsbb_free(wsbb_ctx* ctx, void* d_ptr) {
  WSBB_CHECK_CTX(ctx);
  if (d_ptr) WSBB_HIP(hipFree(d_ptr), "hipFree");
  return WSBB_OK;
}

int wsbb_h2d(wsbb_ctx* ctx, void* d_dst, const void* src, size_t bytes) {
  WSBB_CHECK_CTX(ctx);
  WSBB_HIP(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), "H2D");
  WSBB_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return WSBB_OK;
}

int wsbb_d2h(wsbb_ctx* ctx, void* dst, const void* d_src, size_t bytes) {
  WSBB_CHECK_CTX(ctx);
  WSBB_HIP(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream), "D2H");
  WSBB_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return WSBB_OK;
}

int wsbb_synchronize(wsbb_ctx* ctx) {
  WSBB_CHECK_CTX(ctx);
  WSBB_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return WSBB_OK;
}

int wsbb_timer_start(wsbb_ctx* ctx) {
  WSBB_CHECK_CTX(ctx);
  WSBB_HIP(hipEventRecord(ctx->ev_start, ctx->stream), "hipEventRecord");
  return WSBB_OK;
}

int wsbb_timer_stop(wsbb_ctx* ctx, float* ms) {
  WSBB_CHECK_CTX(ctx);
  WSBB_HIP(hipEventRecord(ctx->ev_stop, ctx->stream), "hipEventRecord");
  WSBB_HIP(hipEventSynchronize(ctx->ev_stop), "hipEventSynchronize");
  float t = 0;
  WSBB_HIP(hipEventElapsedTime(&t, ctx->ev_start, ctx->ev_stop), "hipEventElapsedTime");
  if (ms) *ms = t;
  return WSBB_OK;
}

int wsbb_set_launch(wsbb_ctx* ctx, int groups, int grid_blocks) {
  WSBB_CHECK_CTX(ctx);
  if (groups != 0 && groups != 1 && groups != 2 && groups != 4 && groups != 8 && groups != 16) {
    set_error("groups must be 1, 2, 4, 8 or 16 (0 = automatic)");
    return WSBB_EINVAL;
  }
  if (grid_blocks < 0) return WSBB_EINVAL;
  ctx->launch.groups = groups;
  ctx->launch.grid_cap = grid_blocks;
  return WSBB_OK;
}

int wsbb_set_int16(wsbb_ctx* ctx, int mode) {
  WSBB_CHECK_CTX(ctx);
  if (mode < 0 || mode > 2) {
    set_error("int16 mode must be 0 (never), 1 (large batches) or 2 (always when it fits)");
    return WSBB_EINVAL;
  }
  ctx->launch.int16 = mode;
  return WSBB_OK;
}

int wsbb_autotune(wsbb_ctx* ctx, int N, int reps, int32_t* shapes, float* ms,
                  int* count) {
  WSBB_CHECK_CTX(ctx);
  if (N <= 0 || reps <= 0) return WSBB_EINVAL;
  const int n = ctx->T.n;
  std::vector<uint8_t> permT(size_t(n) * N);
  std::vector<int32_t> l1(N), l2(N);
  random_nodes(n, N, 0xA070u, permT.data(), l1.data(), l2.data());
  void *dp = nullptr, *d1 = nullptr, *d2 = nullptr, *dout = nullptr;
  int rc = WSBB_OK;
  auto cleanup = [&]() {
    if (dp) (void)hipFree(dp);
    if (d1) (void)hipFree(d1);
    if (d2) (void)hipFree(d2);
    if (dout) (void)hipFree(dout);
  };
  WSBB_HIP(hipSetDevice(ctx->device), "hipSetDevice");
  if (hipMalloc(&dp, permT.size()) != hipSuccess || hipMalloc(&d1, 4 * size_t(N)) != hipSuccess ||
      hipMalloc(&d2, 4 * size_t(N)) != hipSuccess || hipMalloc(&dout, 4 * size_t(N)) != hipSuccess) {
    cleanup();
    set_error("hipMalloc(autotune)");
    return WSBB_ENOMEM;
  }
  (void)hipMemcpy(dp, permT.data(), permT.size(), hipMemcpyHostToDevice);
  (void)hipMemcpy(d1, l1.data(), 4 * size_t(N), hipMemcpyHostToDevice);
  (void)hipMemcpy(d2, l2.data(), 4 * size_t(N), hipMemcpyHostToDevice);
  // 8 and 16 waves per tile run in 1024-thread blocks (up to 16 machines)
  const int cand_groups[5] = {1, 2, 4, 8, 16};
  const int n_groups = ctx->T.m <= 16 ? 5 : 3;
  const int cand_grid[3] = {ctx->num_cus * 4, ctx->num_cus * 8, ctx->num_cus * 16};
  int best_g = 1, best_grid = 0, idx = 0;
  float best_t = 1e30f;
  const Launch saved = ctx->launch;
  for (int gi = 0; gi < n_groups; ++gi)
    for (int grid : cand_grid) {
      const int g = cand_groups[gi];
      ctx->launch.groups = g;
      ctx->launch.grid_cap = grid;
      BatchSource src{static_cast<uint8_t*>(dp), static_cast<int32_t*>(d1),
                      static_cast<int32_t*>(d2), N};
      BoundArgs A{};
      A.N = N;
      A.kind = 0;
      A.best = -1;
      A.groups = g;
      A.pk = ctx->int16_for(N);
      A.out = static_cast<int32_t*>(dout);
      hipError_t e = launch_bound(ctx->T, src, A, grid, ctx->stream);  // warm-up
      if (e == hipSuccess) e = hipEventRecord(ctx->ev_start, ctx->stream);
      for (int r = 0; r < reps && e == hipSuccess; ++r)
        e = launch_bound(ctx->T, src, A, grid, ctx->stream);
      if (e == hipSuccess) e = hipEventRecord(ctx->ev_stop, ctx->stream);
      if (e == hipSuccess) e = hipEventSynchronize(ctx->ev_stop);
      float t = 0;
      if (e == hipSuccess) e = hipEventElapsedTime(&t, ctx->ev_start, ctx->ev_stop);
      if (e != hipSuccess) {
        rc = hip_fail(e, "autotune launch");
        break;
      }
      t /= reps;
      if (shapes) {
        shapes[2 * idx] = g;
        shapes[2 * idx + 1] = grid;
      }
      if (ms) ms[idx] = t;
      ++idx;
      if (t < best_t) {
        best_t = t;
        best_g = g;
        best_grid = grid;
      }
    }
  cleanup();
  ctx->launch = saved;
  if (rc != WSBB_OK) return rc;
  if (count) *count = idx;
  const int tiles = (N + 63) / 64;
  int cls = 0;
  while ((1 << (cls + 1)) <= tiles && cls < 31) ++cls;
  ctx->tuned_groups[cls] = best_g;
  if (ctx->launch.grid_cap == 0) ctx->launch.grid_cap = best_grid;
  return WSBB_OK;
}

}  // extern "C"

namespace {
__global__ __launch_bounds__(256) void k_copy16(uint4* __restrict__ dst,
                                                const uint4* __restrict__ src, size_t n16) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n16;
       i += size_t(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}
// Integer VALU throughput probe: eight independent chains of one VALU
// instruction per thread, written as inline asm so the count per iteration
// is exact (kind 0 v_add_u32, 1 v_mad_u32_u24, 2 v_max_i32: the operations of
// the Johnson scan).  Each thread stores its chains' XOR so nothing is
// dead.  One lane-op per instruction and lane.
template <int KIND>
__global__ __launch_bounds__(256) void k_valu_probe(uint32_t* __restrict__ out, int iters) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t a0 = t, a1 = t ^ 1u, a2 = t ^ 2u, a3 = t ^ 3u, a4 = t ^ 4u, a5 = t ^ 5u, a6 = t ^ 6u,
           a7 = t ^ 7u;
  const uint32_t b = (t & 7u) + 1u, c = (t >> 3) & 7u;
  for (int i = 0; i < iters; ++i) {
#define WSBB_P8(INS)                                             \
  asm volatile(INS : "+v"(a0) : "v"(b), "v"(c));                  \
  asm volatile(INS : "+v"(a1) : "v"(b), "v"(c));                  \
  asm volatile(INS : "+v"(a2) : "v"(b), "v"(c));                  \
  asm volatile(INS : "+v"(a3) : "v"(b), "v"(c));                  \
  asm volatile(INS : "+v"(a4) : "v"(b), "v"(c));                  \
  asm volatile(INS : "+v"(a5) : "v"(b), "v"(c));                  \
  asm volatile(INS : "+v"(a6) : "v"(b), "v"(c));                  \
  asm volatile(INS : "+v"(a7) : "v"(b), "v"(c));
    if constexpr (KIND == 0) {
      WSBB_P8("v_add_u32 %0, %0, %1")
      WSBB_P8("v_add_u32 %0, %0, %2")
    } else if constexpr (KIND == 1) {
      WSBB_P8("v_mad_u32_u24 %0, %1, %2, %0")
      WSBB_P8("v_mad_u32_u24 %0, %2, %1, %0")
    } else {
      WSBB_P8("v_max_i32 %0, %0, %1")
      WSBB_P8("v_max_i32 %0, %2, %0")
    }
#undef WSBB_P8
  }
  out[t] = a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7;
}
}  // namespace

extern "C" {

int wsbb_valu_probe(wsbb_ctx* ctx, int kind, int iters, void* d_out, size_t out_bytes,
                    double* lane_ops) {
  WSBB_CHECK_CTX(ctx);
  if (kind < 0 || kind > 2 || iters < 1 || lane_ops == nullptr || d_out == nullptr) {
    set_error("valu probe: kind 0..2, iters >= 1, output buffer and lane_ops required");
    return WSBB_EINVAL;
  }
  // eight blocks of four waves per CU: eight waves per SIMD
  const int grid = ctx->num_cus * 8, block = 256;
  if (out_bytes < sizeof(uint32_t) * size_t(grid) * block) {
    set_error("valu probe: the output buffer needs " +
              std::to_string(sizeof(uint32_t) * size_t(grid) * block) + " bytes");
    return WSBB_EINVAL;
  }
  uint32_t* out = static_cast<uint32_t*>(d_out);
  if (kind == 0) hipLaunchKernelGGL(k_valu_probe<0>, dim3(grid), dim3(block), 0, ctx->stream, out, iters);
  if (kind == 1) hipLaunchKernelGGL(k_valu_probe<1>, dim3(grid), dim3(block), 0, ctx->stream, out, iters);
  if (kind == 2) hipLaunchKernelGGL(k_valu_probe<2>, dim3(grid), dim3(block), 0, ctx->stream, out, iters);
  WSBB_HIP(hipGetLastError(), "valu probe launch");
  *lane_ops = 16.0 * double(iters) * grid * block;
  return WSBB_OK;
}

int wsbb_copy_device(wsbb_ctx* ctx, void* d_dst, const void* d_src, size_t bytes) {
  WSBB_CHECK_CTX(ctx);
  if ((bytes & 15) || (reinterpret_cast<uintptr_t>(d_dst) & 15) ||
      (reinterpret_cast<uintptr_t>(d_src) & 15)) {
    set_error("copy needs 16-byte aligned pointers and sizes");
    return WSBB_EINVAL;
  }
  const size_t n16 = bytes / 16;
  const int grid = static_cast<int>(std::min<size_t>((n16 + 255) / 256, size_t(ctx->num_cus) * 16));
  hipLaunchKernelGGL(k_copy16, dim3(grid > 0 ? grid : 1), dim3(256), 0, ctx->stream,
                     static_cast<uint4*>(d_dst), static_cast<const uint4*>(d_src), n16);
  WSBB_HIP(hipGetLastError(), "copy launch");
  return WSBB_OK;
}

int wsbb_set_pool_capacity(wsbb_ctx* ctx, int nodes) {
  WSBB_CHECK_CTX(ctx);
  if (nodes < 1024) {
    set_error("pool capacity must be at least 1024 nodes");
    return WSBB_EINVAL;
  }
  ctx->pool_capacity = nodes;
  return WSBB_OK;
}

}  // extern "C"
