// Learned machine-pair selector: a two-layer MLP (5m features -> 128 hidden,
// ReLU -> P pair scores) with BF16 weights and FP32 accumulation.
//
// Inference (the search's hot path) runs on MFMA.  It computes the transposed
// product: weights are the A operand and node features the B operand of
// v_mfma_f32_32x32x16_bf16, so the layer-1 accumulator of a lane holds exactly
// the hidden units that lane must supply as layer-2 B operand; the two layers
// chain in registers.  One wavefront handles one 64-node tile, turns the
// scores into per-node softmax probabilities, sums them over the tile and
// keeps the top-k pairs as the tile's pair subset (a wave-uniform subset keeps
// the Johnson scan on the scalar path).  The max over any subset of pairs is
// still a lower bound.
//
// Training is plain FP32 VALU work on small matrices: forward, softmax cross
// entropy against the LB2-maximising pair logged by the plain solver,
// backward, Adam on FP32 master weights; forward passes use the BF16-rounded
// weights and activations the inference kernel sees.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/wsbb_amd.h"
#include "context.h"
#include "selector.h"
#include "tiles.h"

namespace wsbb {

namespace {

constexpr int kHidden = 128;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t f2bf(float f) {  // round to nearest even
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
  uint32_t u = static_cast<uint32_t>(h) << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
__host__ __device__ inline float bfround(float f) { return bf2f(f2bf(f)); }

// ---------------------------------------------------------------- inference
constexpr int kMaxFsteps = 15;  // ceil(5m / 8) for m <= 24

struct MlpArgs {
  int N, P, nsw, topk, fsteps;
  int fstride;              // feature steps between 32-node blocks of xfrag
  int wstride;              // uint2 words of LDS per wave (features / sums)
  int tstride;              // words between tiles of tile_sel (>= nsw)
  const int* Ndev;          // device node count, or null
  const int* tile_list;     // tiles to score (nlist of them), or null: all
  int nlist;
  const int* tile_cnt;      // [ntiles] nodes per tile (with tile_list), or null
  const uint2* xfrag;       // [2 ntiles][fstride][64] feature B fragments
  const uint2* frag1;       // [4][fsteps][64] A fragments of W1^T
  const uint2* frag2;       // [PB][16][64]    A fragments of W2^T, stored right
                            // after frag1 (one allocation)
  const float* b1;          // [128]
  const float* b2;          // [PB*32], padding = -inf
  uint32_t* tile_sel;       // [ntiles][nsw]
  float* scores;            // [P][ld] logits, or null
  int ld;
  uint16_t* probs;          // [N][ldp] BF16 softmax probabilities (OUT == 2)
  int ldp;                  // multiple of 4, >= 32 PB
};

// One operand of v_mfma_f32_32x32x16_bf16 from two 8-wide fragments of the
// same lane (four BF16 values each, the 32x32x8 layout the feature and
// weight writers produce).  A lane's eight k values of the 16-wide step are
// its four of step 2t followed by its four of step 2t + 1; A and B operands
// are both built this way, so the product sums the same 16 k terms, only in
// another order.
__device__ __forceinline__ bf16x8 cat_bf16(uint2 lo, uint2 hi) {
  const u32x4 v = {lo.x, lo.y, hi.x, hi.y};
  return __builtin_bit_cast(bf16x8, v);
}

// One persistent block per CU with eight waves (four when the LDS cannot hold
// eight waves' areas), one 64-node tile per wave and iteration; the weight
// fragments of both layers are staged in LDS once per block, and each
// block's feature fragments pass through the wave's LDS area so that every
// feature load of a block is in flight at once.  OUT: 0 tile masks, 1 logits
// and tile masks, 2 per-node probabilities only (node-major BF16 rows).
template <int PB, int OUT>
__global__ __launch_bounds__(512) void k_mlp_select(MlpArgs A) {
  extern __shared__ uint2 s_w[];
  // the grid is sized for the largest batch: blocks without work leave
  // before staging the weights.  Work items go wave-major over the blocks
  // (item = wave * grid + block), so a small batch spreads over every CU
  // instead of filling the first blocks' waves
  {
    const int N0 = A.Ndev ? *A.Ndev : A.N;
    const int nw0 = A.tile_list ? A.nlist : OUT == 2 ? (N0 + 31) >> 5 : (N0 + 63) >> 6;
    if (static_cast<int>(blockIdx.x) >= nw0) return;
  }
  const int n1 = 4 * A.fsteps * 64, n2 = PB * 16 * 64;
  // 16 loads in flight per thread before their LDS stores: the layer-2
  // fragments follow the layer-1 ones in memory, and indices past the end
  // are clamped (their values are not stored)
  for (int i0 = 0; i0 < n1 + n2; i0 += 16 * blockDim.x) {
    uint2 v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int i = i0 + u * blockDim.x + threadIdx.x;
      v[u] = A.frag1[min(i, n1 + n2 - 1)];
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int i = i0 + u * blockDim.x + threadIdx.x;
      if (i < n1 + n2) s_w[i] = v[u];
    }
  }
  __syncthreads();
  const uint2* w1 = s_w;
  const uint2* w2 = s_w + n1;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nwaves = blockDim.x >> 6;
  // per wave: the feature fragments of one 32-node block [fsteps][64] for
  // layer 1, then (reused) the probabilities of 32 pairs [32][33]
  uint2* s_x = s_w + n1 + n2 + static_cast<size_t>(wave) * A.wstride;
  float* s_p = reinterpret_cast<float*>(s_x);
  const int N = A.Ndev ? *A.Ndev : A.N;
  const int ntiles = (N + 63) >> 6;
  const int col = lane & 31, half = lane >> 5;
  // probabilities only: a work item is one 32-node block
  const int nwork = A.tile_list ? A.nlist : OUT == 2 ? (N + 31) >> 5 : ntiles;
  for (int w0 = wave * gridDim.x + blockIdx.x; w0 < nwork; w0 += gridDim.x * nwaves) {
    const int wi = __builtin_amdgcn_readfirstlane(w0);  // no block barrier below
    const int tile = A.tile_list ? __builtin_amdgcn_readfirstlane(A.tile_list[wi]) : wi;
    const int tcnt = A.tile_cnt ? A.tile_cnt[tile] : 64;
    float agg[PB];  // tile sums of the probabilities of pairs 32 bp + col
#pragma unroll
    for (int bp = 0; bp < PB; ++bp) agg[bp] = 0.f;
    for (int nb = 0; nb < (OUT == 2 ? 1 : 2); ++nb) {
      const int blk = OUT == 2 ? wi : tile * 2 + nb;
      const int node = blk * 32 + col;
      const bool valid = A.tile_list ? (nb * 32 + col < tcnt) : (node < N);
      // the block's feature fragments go to LDS with every load in flight at
      // once (one memory latency per block instead of one per step)
      const uint2* xp = A.xfrag + static_cast<size_t>(blk) * A.fstride * 64 + lane;
      __builtin_amdgcn_sched_barrier(0);
      {
        uint2 xv[kMaxFsteps];
#pragma unroll
        for (int t = 0; t < kMaxFsteps; ++t)
          if (t < A.fsteps) xv[t] = xp[t * 64];
#pragma unroll
        for (int t = 0; t < kMaxFsteps; ++t)
          if (t < A.fsteps) s_x[t * 64 + lane] = xv[t];
      }
      __builtin_amdgcn_sched_barrier(0);
      // layer 1: H^T[128 x 32] = W1^T[128 x F] . X^T[F x 32]
      // the accumulators start from the biases: row 32bi + 8q + 4half + r of
      // the lane sits in acc1[bi][4q + r]
      f32x16 acc1[4];
#pragma unroll
      for (int bi = 0; bi < 4; ++bi)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc1[bi][v] = A.b1[32 * bi + 8 * (v >> 2) + 4 * half + (v & 3)];
      // 16 features per step; an odd last step pairs with zeros (in both
      // operands, so nothing past the fragments is read)
      const uint2 z2 = make_uint2(0u, 0u);
#pragma unroll 1
      for (int t = 0; t < A.fsteps; t += 2) {
        const bool two = t + 1 < A.fsteps;
        const bf16x8 b = cat_bf16(s_x[t * 64 + lane], two ? s_x[(t + 1) * 64 + lane] : z2);
#pragma unroll
        for (int bi = 0; bi < 4; ++bi) {
          const uint2* wr = w1 + (bi * A.fsteps + t) * 64 + lane;
          acc1[bi] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cat_bf16(wr[0], two ? wr[64] : z2), b,
                                                             acc1[bi], 0, 0, 0);
        }
      }
      __builtin_amdgcn_wave_barrier();  // s_x is reused for s_p below
      // bias, ReLU, BF16: lane holds hidden 32bi + 8q + 4half + r for its node,
      // which is element r of the layer-2 B fragment of 8-wide step 4bi + q,
      // so element 4(q & 1) + r of 16-wide step 2bi + q/2 (packed two to a word)
      uint32_t hw[8][4];
#pragma unroll
      for (int bi = 0; bi < 4; ++bi)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
          for (int r = 0; r < 4; r += 2) {
            const uint32_t lo = f2bf(fmaxf(acc1[bi][4 * q + r], 0.f));
            const uint32_t hi = f2bf(fmaxf(acc1[bi][4 * q + r + 1], 0.f));
            hw[2 * bi + (q >> 1)][2 * (q & 1) + (r >> 1)] = lo | (hi << 16);
          }
      // layer 2: Z^T[PB*32 x 32] = W2^T . H^T
      f32x16 acc2[PB];
#pragma unroll
      for (int bp = 0; bp < PB; ++bp)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc2[bp][v] = A.b2[32 * bp + 8 * (v >> 2) + 4 * half + (v & 3)];
      // fully unrolled (hb is indexed statically); the next step's weight
      // fragments are read while this step multiplies, and the scheduling
      // barrier keeps the compiler from hoisting further reads
      uint2 wa[PB][2], wb[PB][2];
#pragma unroll
      for (int bp = 0; bp < PB; ++bp) {
        wa[bp][0] = w2[(bp * 16) * 64 + lane];
        wa[bp][1] = w2[(bp * 16 + 1) * 64 + lane];
      }
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        if (t + 1 < 8) {
#pragma unroll
          for (int bp = 0; bp < PB; ++bp) {
            wb[bp][0] = w2[(bp * 16 + 2 * t + 2) * 64 + lane];
            wb[bp][1] = w2[(bp * 16 + 2 * t + 3) * 64 + lane];
          }
        }
        const u32x4 hv = {hw[t][0], hw[t][1], hw[t][2], hw[t][3]};
        const bf16x8 hb = __builtin_bit_cast(bf16x8, hv);
#pragma unroll
        for (int bp = 0; bp < PB; ++bp)
          acc2[bp] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cat_bf16(wa[bp][0], wa[bp][1]), hb,
                                                             acc2[bp], 0, 0, 0);
#pragma unroll
        for (int bp = 0; bp < PB; ++bp) {
          wa[bp][0] = wb[bp][0];
          wa[bp][1] = wb[bp][1];
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      // logits; this lane's pairs are 32bp + 8(v/4) + 4half + v%4
      if constexpr (OUT == 1) {
        int P = A.P;
        asm volatile("" : "+s"(P));  // keep the bounds tests out of the loop preheader
#pragma unroll
        for (int bp = 0; bp < PB; ++bp)
#pragma unroll
          for (int v = 0; v < 16; ++v) {
            const int s = 32 * bp + 8 * (v >> 2) + 4 * half + (v & 3);
            if (valid && s < P) A.scores[static_cast<size_t>(s) * A.ld + node] = acc2[bp][v];
          }
      }
      float mx = -INFINITY;
#pragma unroll
      for (int bp = 0; bp < PB; ++bp)
#pragma unroll
        for (int v = 0; v < 16; ++v) mx = fmaxf(mx, acc2[bp][v]);
      mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
      float sum = 0.f;
#pragma unroll
      for (int bp = 0; bp < PB; ++bp)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const float e = __expf(acc2[bp][v] - mx);
          acc2[bp][v] = e;
          sum += e;
        }
      sum += __shfl_xor(sum, 32, 64);
      const float inv = valid ? 1.f / sum : 0.f;
      if constexpr (OUT == 2) {
        // pairs 32bp + 8q + 4half .. +3 of this lane's node: one 8-byte store
        if (valid) {
          uint16_t* row = A.probs + static_cast<size_t>(node) * A.ldp + 4 * half;
#pragma unroll
          for (int bp = 0; bp < PB; ++bp)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              uint2 w;
              w.x = f2bf(acc2[bp][4 * q] * inv) |
                    (static_cast<uint32_t>(f2bf(acc2[bp][4 * q + 1] * inv)) << 16);
              w.y = f2bf(acc2[bp][4 * q + 2] * inv) |
                    (static_cast<uint32_t>(f2bf(acc2[bp][4 * q + 3] * inv)) << 16);
              *reinterpret_cast<uint2*>(row + 32 * bp + 8 * q) = w;
            }
        }
        __builtin_amdgcn_wave_barrier();  // s_x of the next block
        continue;
      }
      // tile sums, 32 pairs at a time through this wave's LDS rows (the
      // wave's LDS operations complete in order)
#pragma unroll
      for (int bp = 0; bp < PB; ++bp) {
#pragma unroll
        for (int v = 0; v < 16; ++v)
          s_p[(8 * (v >> 2) + 4 * half + (v & 3)) * 33 + col] = acc2[bp][v] * inv;
        __builtin_amdgcn_wave_barrier();
        float a = 0.f;
#pragma unroll
        for (int c = 0; c < 16; ++c) a += s_p[col * 33 + half * 16 + c];
        a += __shfl_xor(a, 32, 64);
        agg[bp] += a;
        __builtin_amdgcn_wave_barrier();
      }
    }
    if constexpr (OUT == 2) continue;
    // top-k pairs of the tile (ties: lower pair index); lanes 0-31 hold the
    // candidates 32 bp + col, padding and the upper half excluded
    bool ok[PB];
#pragma unroll
    for (int bp = 0; bp < PB; ++bp) ok[bp] = half == 0 && 32 * bp + col < A.P;
    unsigned long long sel[PB];
    wave_topk<PB>(agg, ok, min(A.topk, A.P), sel);
    // word bp of the mask: pairs 32 bp .. 32 bp + 31 from lanes 0-31
#pragma unroll
    for (int bp = 0; bp < PB; ++bp)
      if (lane == 0 && bp < A.nsw)
        A.tile_sel[static_cast<size_t>(tile) * A.tstride + bp] = static_cast<uint32_t>(sel[bp]);
    if (lane == 0)
      for (int w = A.nsw; w < A.tstride; ++w) A.tile_sel[static_cast<size_t>(tile) * A.tstride + w] = 0u;
  }
}

// Pair subsets of child tiles from parent-level predictions
// (parent_tile_mask), one wave per tile.  The learned search computes the
// same masks inside its subset bound kernel; this launch serves the
// per-parent inference entry point.
struct ParentTileArgs {
  const uint16_t* probs;  // [parents][ldp] BF16
  int ldp, P, nsw, topk;
  const int* p_off;       // [parents + 1] first child of each parent
  const int* desc;        // [children] parent << 9 | ...
  const int* Nc;          // live children
  uint32_t* tile_sel;     // [tiles][nsw]
};

template <int NI>
__global__ __launch_bounds__(256) void k_parent_tiles(ParentTileArgs A) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nc = *A.Nc;
  const int ntiles = (nc + 63) >> 6;
  for (int t0 = blockIdx.x * 4; t0 < ntiles; t0 += gridDim.x * 4) {
    const int tile = __builtin_amdgcn_readfirstlane(t0 + wave);
    if (tile >= ntiles) continue;
    const int c0 = tile * 64, c1 = min(c0 + 64, nc);
    unsigned long long sel[NI];
    parent_tile_mask<NI>(A.probs, A.ldp, A.P, A.topk, A.p_off, A.desc, c0, c1, lane, sel);
    // candidates 64 i .. 64 i + 63 = mask words 2i and 2i + 1
    if (lane == 0) {
      uint32_t* ts = A.tile_sel + static_cast<size_t>(tile) * A.nsw;
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        if (2 * i < A.nsw) ts[2 * i] = static_cast<uint32_t>(sel[i]);
        if (2 * i + 1 < A.nsw) ts[2 * i + 1] = static_cast<uint32_t>(sel[i] >> 32);
      }
    }
  }
}

// float features [F][ld] -> B fragments [2 ntiles][fsteps][64]
__global__ void k_feat_pack(const float* feat, int ld, int N, int F, int fsteps, uint2* xfrag) {
  const int nblk = 2 * ((N + 63) / 64);
  const long long total = static_cast<long long>(nblk) * fsteps * 64;
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < total;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    const int lane = static_cast<int>(i & 63);
    const long long bt = i >> 6;
    const int t = static_cast<int>(bt % fsteps);
    const int blk = static_cast<int>(bt / fsteps);
    const int node = blk * 32 + (lane & 31);
    uint16_t v[4];
    for (int r = 0; r < 4; ++r) {
      const int f = 8 * t + 4 * (lane >> 5) + r;
      v[r] = (node < N && f < F) ? f2bf(feat[static_cast<size_t>(f) * ld + node]) : 0;
    }
    xfrag[i] = make_uint2(v[0] | (static_cast<uint32_t>(v[1]) << 16),
                          v[2] | (static_cast<uint32_t>(v[3]) << 16));
  }
}

// ---------------------------------------------------------------- training
struct TrainArgs {
  int B, F, P;             // minibatch rows, features, pairs
  const float* feat;       // [F][ldf] records
  int ldf;
  const int* label;        // [records]
  const int* idx;          // [B] record index of each row
  const float* W1;         // [F][128] masters
  const float* b1;
  const float* W2;         // [128][P]
  const float* b2;
  float* X;                // [B][F] bf16-rounded inputs
  float* A1;               // [B][128] pre-activations
  float* Hb;               // [B][128] bf16-rounded activations
  float* Z;                // [B][P]
  float* dZ;               // [B][P]
  float* dH;               // [B][128]
  float* loss;             // [B]
  float* gW1;
  float* gb1;
  float* gW2;
  float* gb2;
  int nvalid;              // rows with a label (>= 0)
  float* part;             // split-K partial sums of the weight gradients
};

// Weight gradients are batch reductions C[m][n] = sum_b A[b][m] Bm[b][n]
// (A = activations or inputs, Bm = output gradients), with an implicit row
// m = M of ones in A so that C[M][n] is the bias gradient.  64 x 64 output
// tiles, the batch split into chunks of kWgradK rows (one block each) that
// stage 32-row slices of A and Bm in LDS; each thread accumulates a 4 x 4
// block.  Partial sums go to part[split][M + 1][N] and are added in split
// order by k_tr_wreduce, so the result does not depend on timing.
constexpr int kWgradK = 256;

__global__ __launch_bounds__(256) void k_tr_wgrad(const float* A, int M, const float* Bm, int N,
                                                  int K, float* part) {
  __shared__ float As[32][68];
  __shared__ float Bs[32][68];
  const int t = threadIdx.x, ty = t >> 4, tx = t & 15;
  const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64, split = blockIdx.z;
  const int kb = split * kWgradK, ke = min(kb + kWgradK, K);
  float acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.f;
  for (int k0 = kb; k0 < ke; k0 += 32) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int e = t + 256 * r, k = e >> 6, c = e & 63;
      const int kk = k0 + k, m = m0 + c, n = n0 + c;
      const bool kin = kk < ke;
      As[k][c] = !kin ? 0.f : m < M ? A[static_cast<size_t>(kk) * M + m] : (m == M ? 1.f : 0.f);
      Bs[k][c] = kin && n < N ? Bm[static_cast<size_t>(kk) * N + n] : 0.f;
    }
    __syncthreads();
#pragma unroll 8
    for (int k = 0; k < 32; ++k) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[k][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[k][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
    }
    __syncthreads();
  }
  float* out = part + static_cast<size_t>(split) * (M + 1) * N;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + ty * 4 + i;
    if (m > M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = n0 + tx * 4 + j;
      if (n < N) out[static_cast<size_t>(m) * N + n] = acc[i][j];
    }
  }
}

// C = sum over splits; rows 0..M-1 to W, row M to the bias gradient
__global__ void k_tr_wreduce(const float* part, int M, int N, int splits, float* W, float* bias) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int tot = (M + 1) * N;
  if (i >= tot) return;
  float g = 0.f;
  for (int sp = 0; sp < splits; ++sp) g += part[static_cast<size_t>(sp) * tot + i];
  if (i < M * N)
    W[i] = g;
  else
    bias[i - M * N] = g;
}

__global__ void k_tr_in(TrainArgs T) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= T.B * T.F) return;
  const int b = i / T.F, f = i - b * T.F;
  T.X[i] = bfround(T.feat[static_cast<size_t>(f) * T.ldf + T.idx[b]]);
}

__global__ void k_tr_hidden(TrainArgs T) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= T.B * kHidden) return;
  const int b = i / kHidden, h = i - b * kHidden;
  float a = 0.f;
  for (int f = 0; f < T.F; ++f) a += T.X[b * T.F + f] * bfround(T.W1[f * kHidden + h]);
  a += T.b1[h];
  T.A1[i] = a;
  T.Hb[i] = bfround(fmaxf(a, 0.f));
}

__global__ void k_tr_out(TrainArgs T) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= T.B * T.P) return;
  const int b = i / T.P, s = i - b * T.P;
  float z = 0.f;
  for (int h = 0; h < kHidden; ++h) z += T.Hb[b * kHidden + h] * bfround(T.W2[h * T.P + s]);
  T.Z[i] = z + T.b2[s];
}

// one wave per row: lanes over the pairs, wave reductions for max and sum
__global__ __launch_bounds__(256) void k_tr_softmax(TrainArgs T) {
  const int b = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (b >= T.B) return;
  const int y = T.label[T.idx[b]];
  const float* z = T.Z + static_cast<size_t>(b) * T.P;
  float* dz = T.dZ + static_cast<size_t>(b) * T.P;
  if (y < 0) {
    for (int s = lane; s < T.P; s += 64) dz[s] = 0.f;
    if (lane == 0) T.loss[b] = 0.f;
    return;
  }
  float mx = -INFINITY;
  for (int s = lane; s < T.P; s += 64) mx = fmaxf(mx, z[s]);
  for (int d = 32; d >= 1; d >>= 1) mx = fmaxf(mx, __shfl_xor(mx, d, 64));
  float sum = 0.f;
  for (int s = lane; s < T.P; s += 64) sum += expf(z[s] - mx);
  for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, 64);
  const float inv = 1.f / sum, scale = 1.f / static_cast<float>(T.nvalid);
  for (int s = lane; s < T.P; s += 64) dz[s] = (expf(z[s] - mx) * inv - (s == y ? 1.f : 0.f)) * scale;
  if (lane == 0) T.loss[b] = (logf(sum) + mx - z[y]) * scale;
}

__global__ void k_tr_dH(TrainArgs T) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= T.B * kHidden) return;
  const int b = i / kHidden, h = i - b * kHidden;
  float g = 0.f;
  if (T.A1[i] > 0.f)
    for (int s = 0; s < T.P; ++s) g += T.dZ[static_cast<size_t>(b) * T.P + s] * bfround(T.W2[h * T.P + s]);
  T.dH[i] = g;
}

__global__ void k_adam(float* w, float* m, float* v, const float* g, int n, float lr,
                       float b1c, float b2c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  const float mi = beta1 * m[i] + (1.f - beta1) * g[i];
  const float vi = beta2 * v[i] + (1.f - beta2) * g[i] * g[i];
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / b1c) / (sqrtf(vi / b2c) + eps);
}

// inference fragments from the masters
__global__ void k_frags(const float* W1, const float* W2, int F, int Fsteps, int P, int PB,
                        uint2* frag1, uint2* frag2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n1 = 4 * Fsteps * 64, n2 = PB * 16 * 64;
  if (i < n1) {
    const int lane = i & 63, t = (i >> 6) % Fsteps, bi = (i >> 6) / Fsteps;
    const int h = 32 * bi + (lane & 31);
    uint16_t v[4];
    for (int r = 0; r < 4; ++r) {
      const int f = 8 * t + 4 * (lane >> 5) + r;
      v[r] = f < F ? f2bf(W1[f * kHidden + h]) : 0;
    }
    frag1[i] = make_uint2(v[0] | (static_cast<uint32_t>(v[1]) << 16),
                          v[2] | (static_cast<uint32_t>(v[3]) << 16));
  } else if (i < n1 + n2) {
    const int j = i - n1;
    const int lane = j & 63, t = (j >> 6) & 15, bp = j >> 10;
    const int s = 32 * bp + (lane & 31);
    uint16_t v[4];
    for (int r = 0; r < 4; ++r) {
      const int h = 8 * t + 4 * (lane >> 5) + r;
      v[r] = s < P ? f2bf(W2[h * P + s]) : 0;
    }
    frag2[j] = make_uint2(v[0] | (static_cast<uint32_t>(v[1]) << 16),
                          v[2] | (static_cast<uint32_t>(v[3]) << 16));
  }
}

__global__ void k_pad_b2(const float* b2, float* b2p, int P, int Pp) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < Pp) b2p[s] = s < P ? b2[s] : -INFINITY;
}

}  // namespace

struct SelectorState {
  int F = 0, Fsteps = 0, P = 0, PB = 0, nsw = 0;
  int topk = 16, margin = 0;
  // the network predicts, from a parent's features, the pairs that maximise
  // LB2 over its children (logging records the parent's features with each
  // child's label; learned rounds score parents, not children)
  bool parents = false;
  // parameters: masters (W1, b1, W2, b2), Adam moments, gradients
  DevBuf params, frags;
  float *W1, *b1, *W2, *b2, *mW1, *mb1, *mW2, *mb2, *vW1, *vb1, *vW2, *vb2;
  float *gW1, *gb1, *gW2, *gb2, *b2pad;
  uint2 *frag1, *frag2;
  bool frags_fresh = false;
  long long adam_t = 0;
  // records logged by the plain solver
  DevBuf log;
  int log_cap = 0, log_stride = 0;
  float* log_feat = nullptr;   // [F][log_cap]
  int* log_label = nullptr;    // [log_cap]
  int* log_count = nullptr;    // device counter
  // training scratch
  DevBuf scratch;
  int scratch_B = 0;
  // inference scratch for host-buffer calls
  DevBuf io, xio;
};

void destroy_selector(wsbb_ctx* ctx) {
  if (!ctx->selector) return;
  SelectorState* S = ctx->selector;
  S->params.release();
  S->frags.release();
  S->log.release();
  S->scratch.release();
  S->io.release();
  S->xio.release();
  delete S;
  ctx->selector = nullptr;
}

static int param_count(const SelectorState& S) {
  return S.F * kHidden + kHidden + kHidden * S.P + S.P;
}

static int ensure_selector(wsbb_ctx* ctx) {
  if (ctx->selector) return WSBB_OK;
  auto* S = new SelectorState();
  ctx->selector = S;
  S->F = 5 * ctx->T.m;
  S->Fsteps = (S->F + 7) / 8;
  S->P = ctx->T.P;
  S->PB = (S->P + 31) / 32;  // blocks of 32 pair scores
  S->nsw = (S->P + 31) / 32;
  const int np = param_count(*S);
  const size_t per = (static_cast<size_t>(np) * 4 + 255) & ~size_t(255);
  hipError_t e = S->params.reserve(per * 4 + 4 * 32 * S->PB + 256);
  if (e != hipSuccess) {
    destroy_selector(ctx);
    return hip_fail(e, "hipMalloc(selector)");
  }
  float* base[4];
  for (int k = 0; k < 4; ++k) base[k] = S->params.as<float>(per * k);
  auto split = [&](float* p, float** w1, float** bb1, float** w2, float** bb2) {
    *w1 = p;
    *bb1 = p + S->F * kHidden;
    *w2 = *bb1 + kHidden;
    *bb2 = *w2 + kHidden * S->P;
  };
  split(base[0], &S->W1, &S->b1, &S->W2, &S->b2);
  split(base[1], &S->mW1, &S->mb1, &S->mW2, &S->mb2);
  split(base[2], &S->vW1, &S->vb1, &S->vW2, &S->vb2);
  split(base[3], &S->gW1, &S->gb1, &S->gW2, &S->gb2);
  S->b2pad = S->params.as<float>(per * 4);
  const size_t f1 = static_cast<size_t>(4) * S->Fsteps * 64, f2 = static_cast<size_t>(S->PB) * 16 * 64;
  e = S->frags.reserve((f1 + f2) * sizeof(uint2));
  if (e != hipSuccess) {
    destroy_selector(ctx);
    return hip_fail(e, "hipMalloc(selector frags)");
  }
  S->frag1 = S->frags.as<uint2>(0);
  S->frag2 = S->frags.as<uint2>(f1 * sizeof(uint2));
  return WSBB_OK;
}

static hipError_t refresh_frags(wsbb_ctx* ctx) {
  SelectorState* S = ctx->selector;
  if (S->frags_fresh) return hipSuccess;
  const int n = 4 * S->Fsteps * 64 + S->PB * 16 * 64;
  hipLaunchKernelGGL(k_frags, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, S->W1, S->W2,
                     S->F, S->Fsteps, S->P, S->PB, S->frag1, S->frag2);
  hipLaunchKernelGGL(k_pad_b2, dim3((32 * S->PB + 255) / 256), dim3(256), 0, ctx->stream, S->b2,
                     S->b2pad, S->P, 32 * S->PB);
  S->frags_fresh = true;
  return hipGetLastError();
}

hipError_t launch_mlp_select(wsbb_ctx* ctx, const uint2* xfrag, int N, const int* Ndev,
                             uint32_t* tile_sel, float* scores, int ld, const MlpTiles* tl,
                             uint16_t* probs) {
  SelectorState* S = ctx->selector;
  hipError_t e = refresh_frags(ctx);
  if (e != hipSuccess) return e;
  MlpArgs A{};
  A.N = N;
  A.Ndev = Ndev;
  A.P = S->P;
  A.nsw = S->nsw;
  A.topk = S->topk;
  A.fsteps = S->Fsteps;
  A.fstride = tl ? tl->fstride : S->Fsteps;
  A.tstride = tl ? tl->tstride : S->nsw;
  A.tile_list = tl ? tl->list : nullptr;
  A.nlist = tl ? tl->nlist : 0;
  A.tile_cnt = tl ? tl->cnt : nullptr;
  A.xfrag = xfrag;
  A.frag1 = S->frag1;
  A.frag2 = S->frag2;
  A.b1 = S->b1;
  A.b2 = S->b2pad;
  A.tile_sel = tile_sel;
  A.scores = scores;
  A.ld = ld;
  A.probs = probs;
  A.ldp = 32 * S->PB;
  const int ntiles = tl ? tl->nlist : probs ? (N + 31) / 32 : (N + 63) / 64;
  if (ntiles <= 0) return hipSuccess;
  // one persistent block per CU: the weights are staged once per block, and
  // eight waves (two per SIMD) share them when the LDS holds their areas and
  // there is work for two waves per SIMD
  A.wstride = std::max(S->Fsteps * 64, (32 * 33 + 1) / 2);
  const size_t wbytes = sizeof(uint2) * (4 * S->Fsteps * 64 + S->PB * 16 * 64);
  const int waves = wbytes + sizeof(uint2) * 8 * A.wstride <= 160 * 1024 &&
                            ntiles >= 8 * ctx->num_cus
                        ? 8
                        : 4;
  const int grid = std::max(1, std::min((ntiles + waves - 1) / waves, ctx->num_cus));
  const size_t shmem = wbytes + sizeof(uint2) * waves * A.wstride;
  switch (S->PB) {
#define WSBB_PB(v)                                                                          \
  case v:                                                                                   \
    if (probs)                                                                              \
      hipLaunchKernelGGL((k_mlp_select<v, 2>), dim3(grid), dim3(64 * waves), shmem, ctx->stream, A); \
    else if (scores)                                                                        \
      hipLaunchKernelGGL((k_mlp_select<v, 1>), dim3(grid), dim3(64 * waves), shmem, ctx->stream, A); \
    else                                                                                    \
      hipLaunchKernelGGL((k_mlp_select<v, 0>), dim3(grid), dim3(64 * waves), shmem, ctx->stream, A); \
    break;
    WSBB_PB(1) WSBB_PB(2) WSBB_PB(3) WSBB_PB(4) WSBB_PB(5) WSBB_PB(6) WSBB_PB(7) WSBB_PB(8)
    WSBB_PB(9)
#undef WSBB_PB
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_parent_tiles(wsbb_ctx* ctx, const uint16_t* probs, const int* p_off,
                               const int* desc, const int* Nc_dev, int Nc_max,
                               uint32_t* tile_sel) {
  SelectorState* S = ctx->selector;
  ParentTileArgs A{probs, 32 * S->PB, S->P, S->nsw, S->topk, p_off, desc, Nc_dev, tile_sel};
  const int ntiles = (Nc_max + 63) / 64;
  if (ntiles <= 0) return hipSuccess;
  const int grid = std::max(1, std::min((ntiles + 3) / 4, ctx->grid_cap()));
  switch ((S->P + 63) / 64) {
#define WSBB_NI(v)                                                                         \
  case v:                                                                                  \
    hipLaunchKernelGGL((k_parent_tiles<v>), dim3(grid), dim3(256), 0, ctx->stream, A);   \
    break;
    WSBB_NI(1) WSBB_NI(2) WSBB_NI(3) WSBB_NI(4) WSBB_NI(5)
#undef WSBB_NI
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

int selector_fsteps(const wsbb_ctx* ctx) { return ctx->selector ? ctx->selector->Fsteps : 0; }
int selector_ldp(const wsbb_ctx* ctx) { return ctx->selector ? 32 * ctx->selector->PB : 0; }
bool selector_parents(const wsbb_ctx* ctx) { return ctx->selector && ctx->selector->parents; }

int selector_topk(const wsbb_ctx* ctx) { return ctx->selector ? ctx->selector->topk : 0; }
int selector_margin(const wsbb_ctx* ctx) { return ctx->selector ? ctx->selector->margin : 0; }
bool selector_ready(const wsbb_ctx* ctx) { return ctx->selector != nullptr; }
int selector_nsw(const wsbb_ctx* ctx) { return (ctx->T.P + 31) / 32; }

// logging hooks used by the search
int selector_log_stride(const wsbb_ctx* ctx) {
  return ctx->selector ? ctx->selector->log_stride : 0;
}
void selector_log_buffers(const wsbb_ctx* ctx, float** feat, int** label, int** count,
                          int* cap) {
  const SelectorState* S = ctx->selector;
  *feat = S->log_feat;
  *label = S->log_label;
  *count = S->log_count;
  *cap = S->log_cap;
}

static size_t wgrad_part_floats(const SelectorState* S, int B) {
  const size_t splits = (static_cast<size_t>(B) + kWgradK - 1) / kWgradK;
  return splits * std::max(static_cast<size_t>(kHidden + 1) * S->P, static_cast<size_t>(S->F + 1) * kHidden);
}

static hipError_t ensure_scratch(SelectorState* S, int B) {
  if (S->scratch_B >= B) return hipSuccess;
  const size_t bytes = sizeof(float) * (static_cast<size_t>(B) * (S->F + 3 * kHidden + 2 * S->P + 1)) +
                       sizeof(int) * B + 1024 + sizeof(float) * wgrad_part_floats(S, B) + 256;
  hipError_t e = S->scratch.reserve(bytes);
  if (e == hipSuccess) S->scratch_B = B;
  return e;
}

static TrainArgs train_args(SelectorState* S, int B) {
  TrainArgs T{};
  T.B = B;
  T.F = S->F;
  T.P = S->P;
  T.feat = S->log_feat;
  T.ldf = S->log_cap;
  T.label = S->log_label;
  T.W1 = S->W1;
  T.b1 = S->b1;
  T.W2 = S->W2;
  T.b2 = S->b2;
  float* p = S->scratch.as<float>();
  T.X = p;
  p += static_cast<size_t>(B) * S->F;
  T.A1 = p;
  p += static_cast<size_t>(B) * kHidden;
  T.Hb = p;
  p += static_cast<size_t>(B) * kHidden;
  T.Z = p;
  p += static_cast<size_t>(B) * S->P;
  T.dZ = p;
  p += static_cast<size_t>(B) * S->P;
  T.dH = p;
  p += static_cast<size_t>(B) * kHidden;
  T.loss = p;
  p += B;
  T.idx = reinterpret_cast<int*>(p);
  T.part = reinterpret_cast<float*>(reinterpret_cast<uintptr_t>(T.idx + B + 64) & ~uintptr_t(255));
  T.gW1 = S->gW1;
  T.gb1 = S->gb1;
  T.gW2 = S->gW2;
  T.gb2 = S->gb2;
  return T;
}

static void launch_fwd_bwd(wsbb_ctx* ctx, const TrainArgs& T) {
  hipStream_t st = ctx->stream;
  auto g = [](long long n) { return dim3(static_cast<unsigned>((n + 255) / 256)); };
  hipLaunchKernelGGL(k_tr_in, g(1LL * T.B * T.F), dim3(256), 0, st, T);
  hipLaunchKernelGGL(k_tr_hidden, g(1LL * T.B * kHidden), dim3(256), 0, st, T);
  hipLaunchKernelGGL(k_tr_out, g(1LL * T.B * T.P), dim3(256), 0, st, T);
  hipLaunchKernelGGL(k_tr_softmax, dim3((T.B + 3) / 4), dim3(256), 0, st, T);
  const int splits = (T.B + kWgradK - 1) / kWgradK;
  // gW2 = Hb^T dZ, gb2 = column sums of dZ
  hipLaunchKernelGGL(k_tr_wgrad, dim3((kHidden + 1 + 63) / 64, (T.P + 63) / 64, splits), dim3(256), 0,
                     st, T.Hb, kHidden, T.dZ, T.P, T.B, T.part);
  hipLaunchKernelGGL(k_tr_wreduce, g(1LL * (kHidden + 1) * T.P), dim3(256), 0, st, T.part, kHidden,
                     T.P, splits, T.gW2, T.gb2);
  hipLaunchKernelGGL(k_tr_dH, g(1LL * T.B * kHidden), dim3(256), 0, st, T);
  // gW1 = X^T dH, gb1 = column sums of dH
  hipLaunchKernelGGL(k_tr_wgrad, dim3((T.F + 1 + 63) / 64, (kHidden + 63) / 64, splits), dim3(256), 0,
                     st, T.X, T.F, T.dH, kHidden, T.B, T.part);
  hipLaunchKernelGGL(k_tr_wreduce, g(1LL * (T.F + 1) * kHidden), dim3(256), 0, st, T.part, T.F,
                     kHidden, splits, T.gW1, T.gb1);
}

}  // namespace wsbb

using namespace wsbb;

#define SEL_CHECK(ctx)                         \
  do {                                         \
    if (!(ctx)) {                              \
      set_error("null context");               \
      return WSBB_EINVAL;                      \
    }                                          \
    int rc_ = ensure_selector(ctx);            \
    if (rc_) return rc_;                       \
  } while (0)

#define SEL_HIP(call, what)                            \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return hip_fail(e_, what);   \
  } while (0)

extern "C" {

int wsbb_selector_init(wsbb_ctx* ctx, uint64_t seed) {
  SEL_CHECK(ctx);
  SelectorState* S = ctx->selector;
  const int np = param_count(*S);
  std::vector<float> h(np, 0.f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd(0.f, 1.f);
  const float s1 = std::sqrt(2.f / S->F), s2 = std::sqrt(1.f / kHidden);
  for (int i = 0; i < S->F * kHidden; ++i) h[i] = nd(rng) * s1;
  float* w2 = h.data() + S->F * kHidden + kHidden;
  for (int i = 0; i < kHidden * S->P; ++i) w2[i] = nd(rng) * s2;
  SEL_HIP(hipMemcpy(S->W1, h.data(), sizeof(float) * np, hipMemcpyHostToDevice), "H2D params");
  std::vector<float> z(np, 0.f);
  SEL_HIP(hipMemcpy(S->mW1, z.data(), sizeof(float) * np, hipMemcpyHostToDevice), "H2D adam");
  SEL_HIP(hipMemcpy(S->vW1, z.data(), sizeof(float) * np, hipMemcpyHostToDevice), "H2D adam");
  S->adam_t = 0;
  S->frags_fresh = false;
  return WSBB_OK;
}

int wsbb_selector_params(wsbb_ctx* ctx, int set, float* W1, float* b1, float* W2, float* b2) {
  SEL_CHECK(ctx);
  SelectorState* S = ctx->selector;
  struct {
    float* host;
    float* dev;
    size_t n;
  } parts[4] = {{W1, S->W1, size_t(S->F) * kHidden},
                {b1, S->b1, size_t(kHidden)},
                {W2, S->W2, size_t(kHidden) * S->P},
                {b2, S->b2, size_t(S->P)}};
  for (auto& pt : parts) {
    if (!pt.host) continue;
    if (set)
      SEL_HIP(hipMemcpy(pt.dev, pt.host, pt.n * 4, hipMemcpyHostToDevice), "H2D params");
    else
      SEL_HIP(hipMemcpy(pt.host, pt.dev, pt.n * 4, hipMemcpyDeviceToHost), "D2H params");
  }
  if (set) S->frags_fresh = false;
  return WSBB_OK;
}

int wsbb_selector_config(wsbb_ctx* ctx, int topk, int margin) {
  SEL_CHECK(ctx);
  if (topk < 1 || margin < -1) {
    set_error("topk must be >= 1 and margin >= -1 (-1: no full-LB2 fallback)");
    return WSBB_EINVAL;
  }
  ctx->selector->topk = topk;
  ctx->selector->margin = margin;
  return WSBB_OK;
}

int wsbb_selector_parents(wsbb_ctx* ctx, int on) {
  SEL_CHECK(ctx);
  if (on != 0 && on != 1) {
    set_error("parents must be 0 or 1");
    return WSBB_EINVAL;
  }
  ctx->selector->parents = on != 0;
  return WSBB_OK;
}

int wsbb_selector_log(wsbb_ctx* ctx, int stride, int capacity) {
  SEL_CHECK(ctx);
  SelectorState* S = ctx->selector;
  if (stride < 0 || capacity < 0) return WSBB_EINVAL;
  S->log_stride = stride;
  if (capacity != S->log_cap) {
    S->log.release();
    S->log_cap = 0;
    if (capacity > 0) {
      const size_t bytes = sizeof(float) * static_cast<size_t>(S->F) * capacity +
                           sizeof(int) * static_cast<size_t>(capacity) + 256;
      hipError_t e = S->log.reserve(bytes);
      if (e != hipSuccess) return hip_fail(e, "hipMalloc(log)");
      S->log_cap = capacity;
      S->log_feat = S->log.as<float>();
      S->log_label = reinterpret_cast<int*>(S->log_feat + static_cast<size_t>(S->F) * capacity);
      S->log_count = S->log_label + capacity;
    }
  }
  if (S->log_cap > 0) SEL_HIP(hipMemsetAsync(S->log_count, 0, sizeof(int), ctx->stream), "memset");
  SEL_HIP(hipStreamSynchronize(ctx->stream), "sync");
  return WSBB_OK;
}

int wsbb_selector_records(wsbb_ctx* ctx, float* feat, int32_t* label, int* count) {
  SEL_CHECK(ctx);
  SelectorState* S = ctx->selector;
  int c = 0;
  if (S->log_cap > 0)
    SEL_HIP(hipMemcpy(&c, S->log_count, sizeof(int), hipMemcpyDeviceToHost), "D2H count");
  c = std::min(c, S->log_cap);
  if (count) *count = c;
  if (c > 0 && feat)
    for (int f = 0; f < S->F; ++f)
      SEL_HIP(hipMemcpy(feat + static_cast<size_t>(f) * c, S->log_feat + static_cast<size_t>(f) * S->log_cap,
                        sizeof(float) * c, hipMemcpyDeviceToHost),
              "D2H feat");
  if (c > 0 && label)
    SEL_HIP(hipMemcpy(label, S->log_label, sizeof(int) * c, hipMemcpyDeviceToHost), "D2H label");
  return WSBB_OK;
}

int wsbb_selector_set_records(wsbb_ctx* ctx, const float* feat, const int32_t* label, int count) {
  SEL_CHECK(ctx);
  SelectorState* S = ctx->selector;
  if (count < 0 || (count > 0 && (!feat || !label))) return WSBB_EINVAL;
  int rc = wsbb_selector_log(ctx, S->log_stride, count);
  if (rc) return rc;
  for (int f = 0; f < S->F; ++f)
    SEL_HIP(hipMemcpy(S->log_feat + static_cast<size_t>(f) * S->log_cap, feat + static_cast<size_t>(f) * count,
                      sizeof(float) * count, hipMemcpyHostToDevice),
            "H2D feat");
  SEL_HIP(hipMemcpy(S->log_label, label, sizeof(int) * count, hipMemcpyHostToDevice), "H2D label");
  SEL_HIP(hipMemcpy(S->log_count, &count, sizeof(int), hipMemcpyHostToDevice), "H2D count");
  return WSBB_OK;
}

// one forward + backward over the records idx[0..B) (host indices into the
// record store); gradients and mean loss are copied out
int wsbb_selector_grad(wsbb_ctx* ctx, const int32_t* idx, int B, float* loss, float* gW1,
                       float* gb1, float* gW2, float* gb2) {
  SEL_CHECK(ctx);
  SelectorState* S = ctx->selector;
  if (B <= 0 || !idx || S->log_cap == 0) return WSBB_EINVAL;
  int c = 0;
  SEL_HIP(hipMemcpy(&c, S->log_count, sizeof(int), hipMemcpyDeviceToHost), "D2H count");
  std::vector<int> lab(std::min(c, S->log_cap));
  if (!lab.empty())
    SEL_HIP(hipMemcpy(lab.data(), S->log_label, sizeof(int) * lab.size(), hipMemcpyDeviceToHost), "D2H");
  int nvalid = 0;
  for (int b = 0; b < B; ++b) {
    if (idx[b] < 0 || idx[b] >= static_cast<int>(lab.size())) {
      set_error("record index out of range");
      return WSBB_EINVAL;
    }
    nvalid += lab[idx[b]] >= 0;
  }
  SEL_HIP(ensure_scratch(S, B), "hipMalloc(train scratch)");
  TrainArgs T = train_args(S, B);
  T.nvalid = std::max(nvalid, 1);
  SEL_HIP(hipMemcpyAsync(const_cast<int*>(T.idx), idx, sizeof(int) * B, hipMemcpyHostToDevice, ctx->stream), "H2D idx");
  launch_fwd_bwd(ctx, T);
  SEL_HIP(hipGetLastError(), "train launch");
  std::vector<float> l(B);
  SEL_HIP(hipMemcpyAsync(l.data(), T.loss, sizeof(float) * B, hipMemcpyDeviceToHost, ctx->stream), "D2H loss");
  if (gW1) SEL_HIP(hipMemcpyAsync(gW1, S->gW1, sizeof(float) * S->F * kHidden, hipMemcpyDeviceToHost, ctx->stream), "D2H");
  if (gb1) SEL_HIP(hipMemcpyAsync(gb1, S->gb1, sizeof(float) * kHidden, hipMemcpyDeviceToHost, ctx->stream), "D2H");
  if (gW2) SEL_HIP(hipMemcpyAsync(gW2, S->gW2, sizeof(float) * kHidden * S->P, hipMemcpyDeviceToHost, ctx->stream), "D2H");
  if (gb2) SEL_HIP(hipMemcpyAsync(gb2, S->gb2, sizeof(float) * S->P, hipMemcpyDeviceToHost, ctx->stream), "D2H");
  SEL_HIP(hipStreamSynchronize(ctx->stream), "sync");
  if (loss) {
    double s = 0;
    for (float x : l) s += x;
    *loss = static_cast<float>(s);
  }
  return WSBB_OK;
}

int wsbb_selector_train(wsbb_ctx* ctx, int epochs, int batch, float lr, uint64_t seed,
                        float* epoch_loss) {
  SEL_CHECK(ctx);
  SelectorState* S = ctx->selector;
  if (epochs < 1 || batch < 1 || !(lr > 0.f)) return WSBB_EINVAL;
  int c = 0;
  if (S->log_cap > 0)
    SEL_HIP(hipMemcpy(&c, S->log_count, sizeof(int), hipMemcpyDeviceToHost), "D2H count");
  c = std::min(c, S->log_cap);
  std::vector<int> lab(c);
  if (c) SEL_HIP(hipMemcpy(lab.data(), S->log_label, sizeof(int) * c, hipMemcpyDeviceToHost), "D2H");
  std::vector<int> rows;
  for (int i = 0; i < c; ++i)
    if (lab[i] >= 0) rows.push_back(i);
  if (rows.empty()) {
    set_error("no labelled records to train on");
    return WSBB_EINVAL;
  }
  batch = std::min<int>(batch, static_cast<int>(rows.size()));
  SEL_HIP(ensure_scratch(S, batch), "hipMalloc(train scratch)");
  TrainArgs T = train_args(S, batch);
  T.nvalid = batch;
  std::mt19937_64 rng(seed);
  const int np = param_count(*S);
  std::vector<float> lbuf(batch);
  for (int ep = 0; ep < epochs; ++ep) {
    std::shuffle(rows.begin(), rows.end(), rng);
    double tot = 0;
    int nb = 0;
    for (size_t start = 0; start + batch <= rows.size(); start += batch) {
      SEL_HIP(hipMemcpyAsync(const_cast<int*>(T.idx), rows.data() + start, sizeof(int) * batch,
                             hipMemcpyHostToDevice, ctx->stream),
              "H2D idx");
      launch_fwd_bwd(ctx, T);
      ++S->adam_t;
      const float b1c = 1.f - std::pow(0.9f, static_cast<float>(S->adam_t));
      const float b2c = 1.f - std::pow(0.999f, static_cast<float>(S->adam_t));
      hipLaunchKernelGGL(k_adam, dim3((np + 255) / 256), dim3(256), 0, ctx->stream, S->W1, S->mW1,
                         S->vW1, S->gW1, np, lr, b1c, b2c);
      SEL_HIP(hipGetLastError(), "train launch");
      SEL_HIP(hipMemcpyAsync(lbuf.data(), T.loss, sizeof(float) * batch, hipMemcpyDeviceToHost, ctx->stream),
              "D2H loss");
      SEL_HIP(hipStreamSynchronize(ctx->stream), "sync");
      for (float x : lbuf) tot += x;
      ++nb;
    }
    if (epoch_loss) epoch_loss[ep] = nb ? static_cast<float>(tot / nb) : 0.f;
  }
  S->frags_fresh = false;
  return WSBB_OK;
}

int wsbb_selector_infer(wsbb_ctx* ctx, const float* feat, int N, float* scores, uint32_t* tile_sel) {
  SEL_CHECK(ctx);
  SelectorState* S = ctx->selector;
  if (N <= 0 || !feat) return WSBB_EINVAL;
  const int ntiles = (N + 63) / 64;
  const size_t fb = sizeof(float) * static_cast<size_t>(S->F) * N;
  const size_t sb = sizeof(float) * static_cast<size_t>(S->P) * N;
  const size_t tb = sizeof(uint32_t) * static_cast<size_t>(ntiles) * S->nsw;
  SEL_HIP(S->io.reserve(fb + sb + tb + 512), "hipMalloc(io)");
  float* dfeat = S->io.as<float>();
  float* dsc = S->io.as<float>((fb + 255) & ~size_t(255));
  uint32_t* dts = S->io.as<uint32_t>(((fb + 255) & ~size_t(255)) + ((sb + 255) & ~size_t(255)));
  SEL_HIP(hipMemcpyAsync(dfeat, feat, fb, hipMemcpyHostToDevice, ctx->stream), "H2D feat");
  const size_t xb = sizeof(uint2) * static_cast<size_t>(2 * ntiles) * S->Fsteps * 64;
  SEL_HIP(S->xio.reserve(xb), "hipMalloc(feature fragments)");
  uint2* dx = S->xio.as<uint2>();
  hipLaunchKernelGGL(k_feat_pack, dim3(std::min<long long>((xb / 8 + 255) / 256, 4096)), dim3(256), 0,
                     ctx->stream, dfeat, N, N, S->F, S->Fsteps, dx);
  SEL_HIP(hipGetLastError(), "pack launch");
  SEL_HIP(launch_mlp_select(ctx, dx, N, nullptr, dts, scores ? dsc : nullptr, N, nullptr),
          "mlp launch");
  if (scores) SEL_HIP(hipMemcpyAsync(scores, dsc, sb, hipMemcpyDeviceToHost, ctx->stream), "D2H scores");
  if (tile_sel) SEL_HIP(hipMemcpyAsync(tile_sel, dts, tb, hipMemcpyDeviceToHost, ctx->stream), "D2H sel");
  SEL_HIP(hipStreamSynchronize(ctx->stream), "sync");
  return WSBB_OK;
}

int wsbb_selector_infer_parents(wsbb_ctx* ctx, const float* feat, int Np, const int32_t* counts,
                                float* probs, uint32_t* tile_sel) {
  SEL_CHECK(ctx);
  SelectorState* S = ctx->selector;
  if (Np <= 0 || !feat || !counts) return WSBB_EINVAL;
  std::vector<int> off(Np + 1, 0);
  for (int p = 0; p < Np; ++p) {
    if (counts[p] < 0) return WSBB_EINVAL;
    off[p + 1] = off[p] + counts[p];
  }
  const int Nc = off[Np];
  if (Nc <= 0 || Nc > (1 << 30)) return WSBB_EINVAL;
  std::vector<int> desc(Nc);
  for (int p = 0; p < Np; ++p)
    for (int c = off[p]; c < off[p + 1]; ++c) desc[c] = p << 9;
  const int ldp = 32 * S->PB, ntiles = (Nc + 63) / 64;
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t fb = al(sizeof(float) * static_cast<size_t>(S->F) * Np);
  const size_t pb = al(sizeof(uint16_t) * static_cast<size_t>(Np) * ldp);
  const size_t ob = al(sizeof(int) * (static_cast<size_t>(Np) + 2));
  const size_t db = al(sizeof(int) * static_cast<size_t>(Nc));
  const size_t tb = al(sizeof(uint32_t) * static_cast<size_t>(ntiles) * S->nsw);
  SEL_HIP(S->io.reserve(fb + pb + ob + db + tb), "hipMalloc(io)");
  float* dfeat = S->io.as<float>();
  uint16_t* dprob = S->io.as<uint16_t>(fb);
  int* doff = S->io.as<int>(fb + pb);
  int* ddesc = S->io.as<int>(fb + pb + ob);
  uint32_t* dts = S->io.as<uint32_t>(fb + pb + ob + db);
  off.push_back(Nc);  // live child count after the offsets
  SEL_HIP(hipMemcpyAsync(dfeat, feat, sizeof(float) * S->F * static_cast<size_t>(Np),
                         hipMemcpyHostToDevice, ctx->stream), "H2D feat");
  SEL_HIP(hipMemcpyAsync(doff, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice, ctx->stream),
          "H2D offsets");
  SEL_HIP(hipMemcpyAsync(ddesc, desc.data(), sizeof(int) * Nc, hipMemcpyHostToDevice, ctx->stream),
          "H2D desc");
  const size_t xb = sizeof(uint2) * static_cast<size_t>(2 * ((Np + 63) / 64)) * S->Fsteps * 64;
  SEL_HIP(S->xio.reserve(xb), "hipMalloc(feature fragments)");
  uint2* dx = S->xio.as<uint2>();
  hipLaunchKernelGGL(k_feat_pack, dim3(std::min<long long>((xb / 8 + 255) / 256, 4096)), dim3(256), 0,
                     ctx->stream, dfeat, Np, Np, S->F, S->Fsteps, dx);
  SEL_HIP(hipGetLastError(), "pack launch");
  SEL_HIP(launch_mlp_select(ctx, dx, Np, nullptr, nullptr, nullptr, 0, nullptr, dprob), "mlp launch");
  SEL_HIP(launch_parent_tiles(ctx, dprob, doff, ddesc, doff + Np + 1, Nc, dts), "tiles launch");
  std::vector<uint16_t> hp(static_cast<size_t>(Np) * ldp);
  SEL_HIP(hipMemcpyAsync(hp.data(), dprob, sizeof(uint16_t) * hp.size(), hipMemcpyDeviceToHost, ctx->stream),
          "D2H probs");
  if (tile_sel)
    SEL_HIP(hipMemcpyAsync(tile_sel, dts, sizeof(uint32_t) * static_cast<size_t>(ntiles) * S->nsw,
                           hipMemcpyDeviceToHost, ctx->stream), "D2H sel");
  SEL_HIP(hipStreamSynchronize(ctx->stream), "sync");
  if (probs)
    for (int p = 0; p < Np; ++p)
      for (int s = 0; s < S->P; ++s) probs[static_cast<size_t>(p) * S->P + s] = bf2f(hp[static_cast<size_t>(p) * ldp + s]);
  return WSBB_OK;
}

}  // extern "C"
