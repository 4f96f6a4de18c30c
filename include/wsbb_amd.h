/*
 * wsbb_amd.h -- C ABI of the MI355X (gfx950) flowshop bounding engine.
 *
 * Plain C types only: host pointers are caller-owned; pointers named d_* are
 * device (HBM) pointers obtained from wsbb_malloc.  Every call returns
 * WSBB_OK (0) or a negative error code and never exits the process;
 * wsbb_last_error() returns the calling thread's last message.  One context
 * owns one instance, one HIP stream and its device memory; a context is not
 * re-entrant (use one per host thread), like the reference's bound object
 * (src/bound_flowshop.cpp:921, shared `courant`).
 *
 * Node encoding (reference raw_bb_problem fields, bb_problem_calculator.cpp
 * 101-144): a permutation of the n jobs plus limite1 (index of the last job
 * fixed at the front, -1 if none) and limite2 (index of the first job fixed
 * at the back, n if none).  Batches are position-major: permT[i*N + k] is
 * position i of node k (uint8 job ids).
 */
#ifndef WSBB_AMD_H
#define WSBB_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WSBB_OK 0
#define WSBB_EINVAL (-1)   /* bad argument */
#define WSBB_EHIP (-2)     /* HIP runtime error */
#define WSBB_ENOMEM (-3)   /* allocation failed */
#define WSBB_EOVERFLOW (-4) /* device node pool full */
#define WSBB_EBUDGET (-5)  /* node budget exhausted before proof */
#define WSBB_ENODEV (-6)   /* no usable GPU */

#define WSBB_LB2 0     /* Johnson two-machine bound over all machine pairs */
#define WSBB_LB1 1     /* one-machine bound (no counterpart in the reference) */
#define WSBB_LEARNED 2 /* LB2 over learned pair subsets (search only) */

#define WSBB_NSTATS 12
/* stats[] of wsbb_solve:
 *  0 decomposed parents   1 children bounded     2 leaves reached
 *  3 pruned at pop        4 pruned at insertion  5 incumbent improvements
 *  6 max pool size        7 rounds               8 subset-bounded children
 *  9 full-LB2 fallbacks  10 pairs evaluated
 * 11 algorithmic bound work of the bounded children, in int32 ops:
 *    per child P(n + 4u + 4) + 2m d1 + m(m+1) d2 (u free jobs, d1/d2
 *    prefix/suffix lengths) */

int wsbb_version(void);
const char *wsbb_last_error(void);
int wsbb_device_count(int *count);

/* ---- instances (host only, no GPU) ---------------------------------------
 * Taillard generator, src/instance_flowshop.cpp:55-96 (unif 74-82,
 * generate_instance 84-89).  p is machine-major [m][n]; pass p = NULL to
 * query n and m only. */
int wsbb_taillard(int id, int *n, int *m, int32_t *p);
int wsbb_generate(int n, int m, long seed, int32_t *p);
/* job-correlated instance: p[k][j] = a_j + U[0,20], a_j ~ U[1,80] */
int wsbb_generate_correlated(int n, int m, uint64_t seed, int32_t *p);
/* machine-correlated instance: p[k][j] = b_k + U[0,20], b_k ~ U[1,80]
 * (splitmix64 from seed: the m values b_k, then the times machine-major) */
int wsbb_generate_machine_correlated(int n, int m, uint64_t seed, int32_t *p);
/* synthetic node batch: node k draws from splitmix64 seeded seed0 + k:
 * d1 ~ U{0..n-1}, d2 ~ U{0..n-1-d1}, Fisher-Yates permutation;
 * l1 = d1 - 1, l2 = n - d2. */
int wsbb_random_nodes(int n, int N, uint64_t seed0, uint8_t *permT,
                      int32_t *l1, int32_t *l2);
/* Instance precompute on the host, replacing bound_flowshop::initialiserVar
 * (src/bound_flowshop.cpp:556-576): pairs (remplirMachine 500-508), lags
 * (remplirLag 486-498), Johnson orders (Johnson 627-643 with the quicksort
 * tie order of 597-625), R0/Q0 (remplirTempsArriverDepart 510-538).
 * Sizes: pair_a/pair_b [P], lag/johnson [P*n], R0/Q0 [m], P = m(m-1)/2.
 * Any output pointer may be NULL. */
int wsbb_precompute(const int32_t *p, int n, int m, int32_t *pair_a,
                    int32_t *pair_b, int32_t *lag, int32_t *johnson,
                    int32_t *R0, int32_t *Q0);

/* ---- context ---------------------------------------------------------------
 * Replaces bound_flowshop::set_instance -> init (src/bound_flowshop.cpp
 * 911-927, 953-956) together with gpu::init / allocate_on_device /
 * copy_to_device (src/gpu.cu:126-162, 620-642): precomputes the instance
 * tables on the host and uploads them to `device`.  Limits: 2 <= n <= 128,
 * 2 <= m <= 24, processing times in [0, 65535], and the path bound
 * sum_j max_k p[k][j] + sum_k max_j p[k][j] (which no bound value exceeds)
 * below 2^19 = 524288. */
typedef struct wsbb_ctx wsbb_ctx;
int wsbb_create(const int32_t *p, int n, int m, int device, wsbb_ctx **out);
int wsbb_destroy(wsbb_ctx *ctx);
int wsbb_info(const wsbb_ctx *ctx, int *n, int *m, int *pairs, int *num_cus);

/* ---- batch bounding ----------------------------------------------------------
 * Per node equals bound_flowshop::bornes_calculer(perm, limite1, limite2,
 * couts, best) -> couts[0] (src/bound_flowshop.cpp:929-951): the makespan
 * for a leaf (limite2 - limite1 == 1), else LB2 (kind WSBB_LB2; with
 * best != -1 the first running maximum above best, 154-161) or LB1.
 * Replaces the H2D copy + Evaluate_ON_GPU + D2H of gpu::decompose_on_cpu
 * (src/gpu.cu:211-234; kernel src/bound_flowshop_gpu.cu:619-674).
 * Host-buffer form: copies in, bounds, copies out, synchronises. */
int wsbb_bound_batch(wsbb_ctx *ctx, const uint8_t *permT, const int32_t *l1,
                     const int32_t *l2, int N, int kind, int best,
                     int32_t *out);
/* Host-buffer form with the optional outputs used by tests and the pair
 * selector: argmax [N] (first pair reaching LB2; the index idxMoinsBonCmax
 * of borneInfMakespan 150-153), pairv [P*N] (every pair value, [s][k]),
 * sel [ceil(P/32)*N] (restricts LB2 to a per-node pair subset), feat
 * [5m*N] (selector features).  NULL disables each. */
int wsbb_bound_batch_ex(wsbb_ctx *ctx, const uint8_t *permT,
                        const int32_t *l1, const int32_t *l2, int N, int kind,
                        int best, int32_t *out, int32_t *argmax,
                        int32_t *pairv, const uint32_t *sel, float *feat);
/* The learned search's bound, host-buffer form: LB2 restricted, for every
 * 64-node tile t, to the pairs set in tile_sel [t*ceil(P/32) ..]; nodes whose
 * subset bound lies in [inc - margin, inc) (margin >= 0) get full LB2.
 * stats [3] (may be NULL): nodes, full-LB2 fallbacks, node-pair
 * evaluations.  No counterpart in the reference. */
int wsbb_bound_tiles(wsbb_ctx *ctx, const uint8_t *permT, const int32_t *l1,
                     const int32_t *l2, int N, const uint32_t *tile_sel,
                     int margin, int32_t inc, int32_t *out, int64_t *stats);
/* No counterpart in the reference: wsbb_bound_batch for callers that cannot
 * allocate device memory during the call (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * values and errors as wsbb_bound_batch, but the device staging of the
 * nodes and bounds is a fixed 8 MiB buffer that the library holds on each
 * GPU from the first wsbb_create on it until the process ends: at most
 * about 8 MiB / (n + 12) nodes per call (WSBB_EINVAL with the limit in the
 * message beyond that).  Nothing is allocated or freed.  Every call on one
 * GPU, from any context, uses that GPU's buffer: calls that overlap in time
 * on one GPU race (serialise them). */
int wsbb_bound_batch_noalloc(wsbb_ctx *ctx, const uint8_t *permT,
                             const int32_t *l1, const int32_t *l2, int N,
                             int kind, int best, int32_t *out);
/* Device-resident form: all pointers in HBM, asynchronous on the context's
 * stream (call wsbb_synchronize before reading d_out). */
int wsbb_bound_device(wsbb_ctx *ctx, const uint8_t *d_permT,
                      const int32_t *d_l1, const int32_t *d_l2, int N,
                      int kind, int32_t *d_out);

/* ---- device memory, stream and timing -------------------------------------- */
int wsbb_malloc(wsbb_ctx *ctx, size_t bytes, void **d_ptr);
int wsbb_free(wsbb_ctx *ctx, void *d_ptr);
int wsbb_h2d(wsbb_ctx *ctx, void *d_dst, const void *src, size_t bytes);
int wsbb_d2h(wsbb_ctx *ctx, void *dst, const void *d_src, size_t bytes);
int wsbb_synchronize(wsbb_ctx *ctx);
/* Plain streaming device-to-device copy (16-byte aligned), asynchronous on
 * the context's stream.  A measurement helper with no counterpart in the
 * reference: it calibrates hardware byte counters against a known size. */
int wsbb_copy_device(wsbb_ctx *ctx, void *d_dst, const void *d_src,
                     size_t bytes);
/* Integer VALU throughput probe, asynchronous on the context's stream: eight
 * blocks of 256 threads per CU, each thread running 16 x iters independent
 * instructions of one kind (0 v_add_u32, 1 v_mad_u32_u24, 2 v_max_i32, the
 * Johnson scan's operations); *lane_ops receives the lane-operations the
 * launch executes, d_out (num_cus x 8192 bytes) the threads' results.  A
 * measurement helper with no counterpart in the reference: it checks the
 * int32 lane-op peak the bound kernels are measured against. */
int wsbb_valu_probe(wsbb_ctx *ctx, int kind, int iters, void *d_out,
                    size_t out_bytes, double *lane_ops);
/* HIP events recorded on the context's stream around a timed region */
int wsbb_timer_start(wsbb_ctx *ctx);
int wsbb_timer_stop(wsbb_ctx *ctx, float *ms);

/* ---- launch shape and autotuning --------------------------------------------
 * groups: waves that split the pairs of one 64-node tile (1, 2, 4, 8 or
 * 16; 8 and 16 apply to plain LB2 with at most 16 machines, else 4);
 * grid_blocks: cap on 256-thread blocks (0 = automatic).  wsbb_autotune
 * times every shape on a synthetic batch of N nodes and keeps the fastest
 * for batches of that size class; shapes/times are written if non-NULL. */
int wsbb_set_launch(wsbb_ctx *ctx, int groups, int grid_blocks);
int wsbb_autotune(wsbb_ctx *ctx, int N, int reps, int32_t *shapes,
                  float *ms, int *count);
/* Plain LB2 (no optional outputs, best = -1) may run two nodes per lane in
 * int16 halves when every bound value of the instance stays below 2^14:
 * mode 1 (default) for batches of at least 2 x 128 x (4 x CUs) nodes, 2 for
 * every batch, 0 never.  Both forms give identical values. */
int wsbb_set_int16(wsbb_ctx *ctx, int mode);

/* ---- device-resident branch-and-bound -------------------------------------------
 * Replaces explorer::_explore_cpu_gpu (src/explorer.cu:154-305) with
 * gpu::decompose_on_cpu (src/gpu.cu:180-255) and gpu::prune (257-408): the
 * node pool, parent selection, child generation (decompose_beginend,
 * src/decompose_beginend.cpp:26-138), bounding, branching-side choice
 * (choise 39-50), ranking (problems::ranks, src/problems.cpp:46-62 with
 * select_worst), incumbent update (solutions::insert, src/solutions.cpp
 * 45-65) and pruning all stay on the device.  Parents are popped depth-first
 * into up to max_batches sub-batches whose reserved child counts
 * 2*(n-depth) stay below pool_size-1 (problems::push_back1, problems.cpp
 * 64-78).  ub0: initial incumbent (INT32_MAX = none).  budget: maximum
 * decomposed parents (<= 0: none).  kind: WSBB_LB2, WSBB_LB1 or
 * WSBB_LEARNED (needs wsbb_selector_init).  Outputs: best makespan,
 * best_perm [n] (may be NULL), stats [WSBB_NSTATS] (may be NULL).
 * Returns WSBB_EBUDGET (with the incumbent so far) when the budget ran out. */
int wsbb_solve(wsbb_ctx *ctx, int kind, int pool_size, int max_batches,
               int32_t ub0, int64_t budget, int32_t *best,
               int32_t *best_perm, int64_t *stats);
/* Optional HIP-event timing of the bound-kernel launches of the next
 * wsbb_solve calls: enable = k > 0 times the bound launch of every k-th round
 * (events taken by the launch itself), 0 turns timing off, -1 only reads.
 * Reports the summed kernel time and the number of timed launches of the
 * last solve (rounds after the search ended are not counted).  No
 * counterpart in the reference (its statistics time whole GPU phases,
 * explorer.cu:219-249). */
int wsbb_search_profile(wsbb_ctx *ctx, int enable, double *bound_ms,
                        int64_t *bound_launches);
/* Search trace, for tests and debugging.  With fn set, later wsbb_solve /
 * wsbb_solve_resume calls run one round at a time (no graph, no rounds
 * queued ahead) and, after every round that took parents, call fn on the
 * calling thread with host copies of that round:
 *   parents in pop order: permT [n][np] (position-major), l1, l2, cost [np];
 *   children in generation order: desc [nc], bound [nc], where
 *     desc = parent << 9 | side << 8 | j is the child of parent `parent`
 *     that swaps position j with position l1 + 1 (side 0, BEGIN) or l2 - 1
 *     (side 1, END), then applies limites_set (l2 - l1 == 2 -> l1 + 1);
 *     both sides are listed, BEGIN children first (decompose_beginend.cpp
 *     26-37, 114-138);
 *   inc: the incumbent after the round.
 * The buffers are valid during the call only.  fn = NULL turns it off.
 * The reference prints per-pool statistics instead (explorer.cu:219-249). */
typedef void (*wsbb_trace_fn)(void *user, int64_t round, int np,
                              const uint8_t *permT, const int32_t *l1,
                              const int32_t *l2, const int32_t *cost, int nc,
                              const int32_t *desc, const int32_t *bound,
                              int32_t inc);
int wsbb_search_trace(wsbb_ctx *ctx, wsbb_trace_fn fn, void *user);
/* Device node-pool capacity in nodes (default 2^24) for later solves.  A
 * round that could leave more nodes than that on the pool is not started:
 * the solve returns WSBB_EOVERFLOW with the pool, incumbent and counters as
 * they were before that round, so wsbb_search_state followed by
 * wsbb_solve_resume with a larger capacity continues exactly. */
int wsbb_set_pool_capacity(wsbb_ctx *ctx, int nodes);
/* Launch options of later solves: use_graph = 1 captures the kernels of
 * rounds_per_sync rounds in a hipGraph once per solve and replays it, 0
 * launches them directly (-1 keeps the setting; default 0: both measure
 * the same, the rounds being bound by the device); rounds_per_sync
 * rounds run between host checks of the pool state (1..64, -1 keeps it;
 * default 4).  Bound-kernel timing (wsbb_search_profile) launches directly.
 * No counterpart in the reference (its host loop synchronises every pool,
 * explorer.cu:216-250). */
int wsbb_search_options(wsbb_ctx *ctx, int use_graph, int rounds_per_sync);
/* Checkpoint of the last solve's device state: incumbent, counters, the
 * incumbent's permutation (-1s if none) and the node pool left when it
 * stopped (budget, overflow) -- the count of pool nodes in *count, and with
 * buffers of cap >= count nodes: permT [n][count] (position-major), l1, l2,
 * cost [count].  Counterpart of solutions::save (src/solutions.cpp:97-108),
 * extended to the pool. */
int wsbb_search_state(wsbb_ctx *ctx, int32_t *inc, int64_t *stats,
                      int32_t *best_perm, int *count, uint8_t *permT,
                      int32_t *l1, int32_t *l2, int32_t *cost, int cap);
/* Continues a search from a saved pool (same layout as wsbb_search_state),
 * incumbent inc, counters stats_in (may be NULL) and incumbent permutation
 * best_perm_in (may be NULL); other arguments and outputs as wsbb_solve.  A
 * search stopped by its budget and resumed here visits exactly the nodes an
 * uninterrupted search visits.  Counterpart of solutions::init /
 * readOptimalSolution (src/solutions.cpp:24-33, 69-86), extended to the
 * pool. */
int wsbb_solve_resume(wsbb_ctx *ctx, int kind, int pool_size, int max_batches,
                      int64_t budget, const uint8_t *permT, const int32_t *l1,
                      const int32_t *l2, const int32_t *cost, int count,
                      int32_t inc, const int64_t *stats_in,
                      const int32_t *best_perm_in, int32_t *best,
                      int32_t *best_perm, int64_t *stats);

/* ---- live node pool -------------------------------------------------------------
 * The device search driven round by round from the host, with nodes moved
 * in and out of the pool between rounds.  In the reference the GPU peer's
 * thread inserts nodes received from other peers into its tree between
 * pools (Node::recvWorkInWorkStealingGPU3, src/work_stealing/node.cpp
 * 1240-1283; the GPU thread protocol.cpp:937-1033), gives nodes away
 * (node.cpp:479-576) and adopts better incumbents broadcast by others
 * (node.cpp:94-152).  One context is driven by one host thread; several
 * threads each drive their own context and exchange nodes and incumbents
 * through these calls.  A pool node is (permutation, limite1, limite2,
 * bound) with at least two free jobs; node batches are position-major
 * (permT [n][count]) and bottom-first: the last node of a pushed batch ends
 * on top of the pool and is decomposed first.
 *
 * wsbb_pool_open: a new session of the given kind and pool shape (as
 * wsbb_solve) with an empty pool, no incumbent (INT32_MAX) and zero
 * counters.  wsbb_solve leaves its own session open on its final state.
 * wsbb_pool_seed_root pushes the root (tree::add_root, tree.cpp:170-176). */
int wsbb_pool_open(wsbb_ctx *ctx, int kind, int pool_size, int max_batches);
int wsbb_pool_seed_root(wsbb_ctx *ctx);
/* WSBB_EINVAL for a node with fewer than two free jobs, a negative bound or
 * a column that is not a permutation; WSBB_EOVERFLOW beyond the pool
 * capacity. */
int wsbb_pool_push(wsbb_ctx *ctx, const uint8_t *permT, const int32_t *l1,
                   const int32_t *l2, const int32_t *cost, int count);
/* Takes up to max nodes off the pool (*count of them): from the top, the
 * nodes the search would decompose next (tree::take, tree.cpp:257-266), or
 * with from_bottom != 0 the bottom, the shallowest nodes and so the largest
 * subtrees (what a peer gives away).  Returned bottom-first; pushing them
 * back restores the pool. */
int wsbb_pool_pop(wsbb_ctx *ctx, int max, int from_bottom, uint8_t *permT,
                  int32_t *l1, int32_t *l2, int32_t *cost, int *count);
int wsbb_pool_size(wsbb_ctx *ctx, int *size);
/* The incumbent and its permutation ([n], -1s if unknown; may be NULL).
 * wsbb_incumbent_set replaces it only by a strictly smaller value
 * (solutions::insert, src/solutions.cpp:45-65); perm may be NULL. */
int wsbb_incumbent_get(wsbb_ctx *ctx, int32_t *inc, int32_t *perm);
int wsbb_incumbent_set(wsbb_ctx *ctx, int32_t inc, const int32_t *perm);
/* Runs up to `rounds` rounds (<= 0: until the pool is empty) of the open
 * session; budget bounds the session's decomposed parents (<= 0: none).
 * *size (may be NULL) receives the pool size after the last round, stats
 * [WSBB_NSTATS] (may be NULL) the session's counters.  An empty pool means
 * every node pushed so far is resolved against the incumbent.  Returns
 * WSBB_EBUDGET or WSBB_EOVERFLOW (pool unchanged by the round that did not
 * fit) like wsbb_solve; a later call continues. */
int wsbb_pool_run(wsbb_ctx *ctx, int rounds, int64_t budget, int *size,
                  int64_t *stats);

/* ---- learned machine-pair selector -------------------------------------------
 * No counterpart in the reference (its pair order ordoSomme is fixed,
 * bound_flowshop.cpp:130 is commented out).  A two-layer MLP (5m features:
 * per machine head, tail, remaining work, head+work, work+tail; 128 hidden
 * ReLU units; one score per machine pair), BF16 weights with FP32
 * accumulation, trained on records (node features, LB2-maximising pair)
 * logged by plain solves.  With kind WSBB_LEARNED the search bounds each
 * 64-child wavefront tile over the top-k pairs of its summed predicted
 * probabilities; a child whose subset bound lies in [incumbent - margin,
 * incumbent) gets the remaining pairs too (full LB2).  Any subset bound is
 * a valid lower bound, so the proven optimum is unchanged.
 * Record layout: feat [5m][count] (float, scaled by the instance's path
 * bound), label [count] (pair index, -1 for leaves). */
int wsbb_selector_init(wsbb_ctx *ctx, uint64_t seed);
/* set = 1 uploads, 0 downloads the FP32 master weights: W1 [5m][128],
 * b1 [128], W2 [128][P], b2 [P]; NULL pointers are skipped. */
int wsbb_selector_params(wsbb_ctx *ctx, int set, float *W1, float *b1,
                         float *W2, float *b2);
/* topk pairs per 64-child tile; children whose subset bound lies in
 * [incumbent - margin, incumbent) get full LB2 (margin -1: none) */
int wsbb_selector_config(wsbb_ctx *ctx, int topk, int margin);
/* on = 1: the network works per parent.  Logging then records, for every
 * stride-th child, its parent's features with the child's label; learned
 * rounds score the round's parents, and a 64-child tile takes the top-k
 * pairs of its parents' probabilities weighted by their children in the
 * tile.  on = 0 (default): features and scores are the children's own. */
int wsbb_selector_parents(wsbb_ctx *ctx, int on);
/* log every stride-th child of later plain solves into a device store of
 * `capacity` records (stride 0 stops logging); clears the store */
int wsbb_selector_log(wsbb_ctx *ctx, int stride, int capacity);
int wsbb_selector_records(wsbb_ctx *ctx, float *feat, int32_t *label,
                          int *count);
int wsbb_selector_set_records(wsbb_ctx *ctx, const float *feat,
                              const int32_t *label, int count);
/* one forward/backward pass (softmax cross-entropy, mean over labelled
 * rows) over records idx[0..B): loss and gradients of the masters */
int wsbb_selector_grad(wsbb_ctx *ctx, const int32_t *idx, int B, float *loss,
                       float *gW1, float *gb1, float *gW2, float *gb2);
/* Adam training on the stored records; epoch_loss [epochs] may be NULL */
int wsbb_selector_train(wsbb_ctx *ctx, int epochs, int batch, float lr,
                        uint64_t seed, float *epoch_loss);
/* MFMA inference: scores [P][N] logits (may be NULL) and the per-tile
 * top-k masks tile_sel [ceil(N/64)][ceil(P/32)] (may be NULL) */
int wsbb_selector_infer(wsbb_ctx *ctx, const float *feat, int N,
                        float *scores, uint32_t *tile_sel);
/* per-parent inference: features feat [5m][Np] of Np parents with counts[p]
 * consecutive children each; probs [Np][P] (BF16 softmax probabilities as
 * float, may be NULL) and the masks tile_sel [ceil(sum counts / 64)]
 * [ceil(P/32)] of the 64-child tiles (may be NULL) */
int wsbb_selector_infer_parents(wsbb_ctx *ctx, const float *feat, int Np,
                                const int32_t *counts, float *probs,
                                uint32_t *tile_sel);

/* ---- mixed batches of several instances -------------------------------------
 * No counterpart in the reference (one instance per process there).  A
 * multi-instance handle owns one context per instance (reachable with
 * wsbb_multi_context, e.g. to train or load its pair selector).  A batch
 * holds the nodes of instance i at [offsets[i], offsets[i+1]); permT is
 * [max n][N] position-major (rows past an instance's n are not read).
 * wsbb_multi_load uploads and checks a batch once; wsbb_multi_run bounds it
 * `reps` times, each time in one bound launch over 64-node tiles that never
 * mix instances:
 *   kind WSBB_LB2:     every node's bornes_calculer value for its instance;
 *   kind WSBB_LEARNED: LB2 over each tile's top-k pairs chosen by the MLP of
 *                      the first instance with the tile's machine count (one
 *                      MLP launch per machine count), or over the masks in
 *                      tile_sel [ntiles][ceil(Pmax/32)] if given (tiles in
 *                      instance order, 64 nodes each); with inc [count] and
 *                      margin >= 0, nodes whose subset bound lies in
 *                      [inc - margin, inc) of their instance get full LB2.
 * Outputs: out [N] (may be NULL), stats [3] = nodes, full-LB2 fallbacks,
 * node-pair evaluations over all reps (may be NULL), ms = device time per
 * rep (may be NULL). */
typedef struct wsbb_multi wsbb_multi;
int wsbb_multi_create(int count, const int32_t *const *p, const int *n,
                      const int *m, int device, wsbb_multi **out);
int wsbb_multi_destroy(wsbb_multi *multi);
int wsbb_multi_context(wsbb_multi *multi, int i, wsbb_ctx **out);
int wsbb_multi_load(wsbb_multi *multi, const uint8_t *permT, const int32_t *l1,
                    const int32_t *l2, const int32_t *offsets, int N);
int wsbb_multi_run(wsbb_multi *multi, int kind, const int32_t *inc, int margin,
                   const uint32_t *tile_sel, int reps, int32_t *out,
                   int64_t *stats, float *ms);

#ifdef __cplusplus
}
#endif
#endif /* WSBB_AMD_H */
