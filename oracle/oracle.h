/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's CPU path for the permutation
 * flowshop (Cmax) branch-and-bound: Taillard generator, instance precompute,
 * per-node heads/tails, LB2 (Johnson two-machine bound over all machine
 * pairs), leaf makespan, the build-defined LB1, begin-end branching with the
 * branching-side rule and child ranks, and the two search loops (the CPU DFS
 * explorer and the batched GPU-path explorer).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker / CPU baseline.  The product
 * (wsbb_amd/) never links, loads or calls it.
 *
 * Every function cites the reference file:line it follows.  Pinning (what
 * checks each function beyond the restatement itself) is listed in
 * tests/test_oracle.py; the Taillard generator is also pinned against the
 * reference's own instance_flowshop.cpp compiled into oracle/_ref/.
 */
#ifndef WSBB_ORACLE_H
#define WSBB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXN 255 /* perm entries are stored as uint8 */
#define ORC_MAXM 64

/* --- instance generation (src/instance_flowshop.cpp:13-96) --- */
long orc_unif(long *seed, long low, long high);
int orc_taillard_dims(int id, int *n, int *m, long *seed);
void orc_generate(int n, int m, long seed, int *p /* [m*n] machine-major */);
/* job-correlated generator: p[k][j] = a_j + U[0,20], a_j ~ U[1,80] (splitmix64) */
void orc_generate_correlated(int n, int m, uint64_t seed, int *p);
/* machine-correlated generator: p[k][j] = b_k + U[0,20], b_k ~ U[1,80] */
void orc_generate_machine_correlated(int n, int m, uint64_t seed, int *p);
uint64_t orc_splitmix64(uint64_t *state);
/* random node batch: node k uses splitmix64 stream seeded with seed0+k;
 * d1~U{0..n-1}, d2~U{0..n-1-d1}, Fisher-Yates permutation.
 * permT is position-major [n][N] uint8, l1/l2 int32 [N]. */
void orc_random_nodes(int n, int N, uint64_t seed0, uint8_t *permT, int *l1,
                      int *l2);

/* --- instance precompute (src/bound_flowshop.cpp:486-576, 579-643) --- */
typedef struct orc_inst orc_inst;
orc_inst *orc_create(const int *p, int n, int m);
void orc_destroy(orc_inst *h);
int orc_pairs(const orc_inst *h);
/* copies out ma[P], mb[P], lag[P*n], johnson[P*n], R0[m], Q0[m] */
void orc_tables(const orc_inst *h, int *ma, int *mb, int *lag, int *johnson,
                int *R0, int *Q0);

/* --- per-node bounds (src/bound_flowshop.cpp:64-168, 841-951) --- */
enum { ORC_LB2 = 0, ORC_LB1 = 1 };
int orc_bound(const orc_inst *h, const int *perm, int l1, int l2, int best);
int orc_lb1(const orc_inst *h, const int *perm, int l1, int l2);
int orc_cmax(const orc_inst *h, const int *perm);
void orc_heads_tails(const orc_inst *h, const int *perm, int l1, int l2,
                     int *r, int *q);
/* per-pair values v_s (the quantities LB2 takes the max of); returns LB2 */
int orc_pair_values(const orc_inst *h, const int *perm, int l1, int l2,
                    int *v);
/* LB2 restricted to the pairs with sel[s] != 0 (max(0, max_{s in sel} v_s)) */
int orc_bound_subset(const orc_inst *h, const int *perm, int l1, int l2,
                     const uint8_t *sel);
/* batch over position-major nodes; kind = ORC_LB2 / ORC_LB1; nthreads >= 1 */
void orc_bound_batch(const orc_inst *h, const uint8_t *permT, const int *l1,
                     const int *l2, int N, int kind, int *out, int nthreads);

/* --- branching (src/decompose_beginend.cpp:26-107,
 *     src/bb_problem_calculator.cpp:25-87, src/problems.cpp:13-62,
 *     src/select_worst.cpp:7-12) ---
 * Expands one node exactly as decompose_beginend::operator().  Returns the
 * number of children of the chosen side in the order the reference leaves
 * them in `ns` (sorted by select_worst).  For each child c:
 * cperm[c*n..], cl1[c], cl2[c], cdepth[c], ccost[c], cjob[c], crank[c],
 * and *side = 0 (BEGIN) / 1 (END); nbounded receives 2u (or 2).       */
int orc_decompose(const orc_inst *h, const int *perm, int l1, int l2,
                  int depth, int kind, int *cperm, int *cl1, int *cl2,
                  int *cdepth, int *ccost, int *cjob, int *crank, int *side,
                  int *nbounded);

/* --- search loops --- */
/* stats layout (int64):
 *  [0] decomposed parents  [1] children bounded  [2] leaves reached
 *  [3] pruned at pop       [4] pruned at insertion  [5] incumbent updates
 *  [6] max pool size       [7] rounds (batched explorer only)           */
#define ORC_NSTATS 8
/* CPU explorer (src/explorer.cu:92-148): DFS by ranks, prune at pop.
 * order: 0 = lexicographic-ranks heap (tree.cpp:37-46), 1 = LIFO stack.
 * ub0: initial incumbent (INT32_MAX = +inf).  budget: max decompositions
 * (<=0: unlimited).  Returns the best makespan (ub0 if none found), or -1
 * when the budget was exhausted before the pool emptied.              */
int orc_solve_dfs(const orc_inst *h, int kind, int order, int ub0,
                  int64_t budget, int64_t *stats, int *best_perm);
/* batched GPU-path explorer (src/explorer.cu:154-305, src/gpu.cu:180-408):
 * pops parents into up to max_batches sub-batches whose reserved child
 * counts 2*(n-depth) stay below pool_size-1 (problems.cpp:64-78), bounds
 * every child of both sides, then prunes parent by parent.              */
int orc_solve_batched(const orc_inst *h, int kind, int pool_size,
                      int max_batches, int ub0, int64_t budget,
                      int64_t *stats, int *best_perm);
/* the same with the pool order chosen: 0 = lexicographic-ranks heap (the
 * reference's priority_queue), 1 = LIFO stack (the default above) */
int orc_solve_batched_ex(const orc_inst *h, int kind, int pool_size,
                         int max_batches, int ub0, int64_t budget, int order,
                         int64_t *stats, int *best_perm);

/* all-cores CPU explorer (the CPU baseline's time-to-optimum): the DFS
 * explorer's rules over independent subtrees of a sequentially expanded
 * top, nthreads threads, one shared incumbent.  Same optimum as
 * orc_solve_dfs; node counts depend on thread timing (stats summed). */
int orc_solve_parallel(const orc_inst *h, int kind, int nthreads, int ub0,
                       int64_t *stats, int *best_perm);

/* --- search trace (test infrastructure) ---
 * Between orc_trace_begin and orc_trace_end the explorers append one record
 * of 2n + 8 int32 per event to buf (at most cap_records are stored; the
 * count returned by orc_trace_end includes those past the capacity):
 * [type, round, l1, l2, depth, cost, 0, perm[n], ranks[n+1]]; round is the
 * decomposition count (DFS) or the round number (batched).  A leaf met by
 * the batched prune loop carries its parent's ranks. */
enum { ORC_TR_TAKEN = 0, ORC_TR_PRUNED = 1, ORC_TR_LEAF = 2, ORC_TR_REPUSH = 3 };
void orc_trace_begin(int32_t *buf, int64_t cap_records, int n);
int64_t orc_trace_end(void);

#ifdef __cplusplus
}
#endif
#endif
