/* Drives the oracle (oracle/oracle.c) through every entry point on small
 * instances for the sanitizer build (tests/test_sanitizers.py): generators,
 * precompute, per-node bounds, batches on several threads, decomposition,
 * the DFS explorer in both pool orders with a trace, the batched explorer,
 * the parallel explorer.  Exits 0 when the explorers agree. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../oracle/oracle.h"

int main(void) {
  int fails = 0;
  for (int id = 1; id <= 12; id += 11) {
    int n = 0, m = 0;
    long seed = 0;
    orc_taillard_dims(id, &n, &m, &seed);
    int *p = (int *)malloc(sizeof(int) * n * m);
    orc_generate(n, m, seed, p);
    orc_inst *h = orc_create(p, n, m);
    int N = 3000;
    uint8_t *permT = (uint8_t *)malloc((size_t)n * N);
    int *l1 = (int *)malloc(sizeof(int) * N), *l2 = (int *)malloc(sizeof(int) * N);
    int *o1 = (int *)malloc(sizeof(int) * N), *o4 = (int *)malloc(sizeof(int) * N);
    orc_random_nodes(n, N, 0x5EED, permT, l1, l2);
    orc_bound_batch(h, permT, l1, l2, N, ORC_LB2, o1, 1);
    orc_bound_batch(h, permT, l1, l2, N, ORC_LB2, o4, 4);
    for (int k = 0; k < N; k++) fails += o1[k] != o4[k];
    orc_bound_batch(h, permT, l1, l2, N, ORC_LB1, o1, 3);
    int P = orc_pairs(h);
    int *v = (int *)malloc(sizeof(int) * (P + 1));
    int perm[ORC_MAXN], r[ORC_MAXM], q[ORC_MAXM];
    uint8_t sel[4096] = {0};
    for (int s = 0; s < P; s += 3) sel[s] = 1;
    for (int k = 0; k < 50; k++) {
      for (int i = 0; i < n; i++) perm[i] = permT[(size_t)i * N + k];
      fails += orc_pair_values(h, perm, l1[k], l2[k], v) != orc_bound(h, perm, l1[k], l2[k], 999999);
      orc_heads_tails(h, perm, l1[k], l2[k], r, q);
      fails += orc_bound_subset(h, perm, l1[k], l2[k], sel) > orc_bound(h, perm, l1[k], l2[k], 999999);
    }
    int *cperm = (int *)malloc(sizeof(int) * 2 * n * n), *ca = (int *)malloc(sizeof(int) * 12 * n);
    int side = 0, nb = 0;
    for (int i = 0; i < n; i++) perm[i] = i;
    orc_decompose(h, perm, -1, n, 0, ORC_LB2, cperm, ca, ca + 2 * n, ca + 4 * n, ca + 6 * n,
                  ca + 8 * n, ca + 10 * n, &side, &nb);
    int64_t st[ORC_NSTATS], st2[ORC_NSTATS];
    int bp[ORC_MAXN];
    int32_t *tr = (int32_t *)malloc(sizeof(int32_t) * (2 * n + 8) * 20000);
    orc_trace_begin(tr, 20000, n);
    int b0 = orc_solve_dfs(h, ORC_LB2, 0, 0x7fffffff, id == 1 ? 0 : 3000, st, bp);
    orc_trace_end();
    int b1 = orc_solve_dfs(h, ORC_LB2, 1, 0x7fffffff, id == 1 ? 0 : 3000, st2, bp);
    fails += b0 != b1;
    int b2 = orc_solve_batched_ex(h, ORC_LB2, 300, 3, 0x7fffffff, id == 1 ? 0 : 2000, 0, st, bp);
    int b3 = orc_solve_batched(h, ORC_LB2, 300, 3, 0x7fffffff, id == 1 ? 0 : 2000, st2, bp);
    fails += b2 != b3;
    if (id == 1) fails += orc_solve_parallel(h, ORC_LB2, 4, 0x7fffffff, st, bp) != b0;
    printf("instance %d: dfs %d batched %d\n", id, b0, b2);
    free(tr);
    free(cperm);
    free(ca);
    free(v);
    free(o1);
    free(o4);
    free(l1);
    free(l2);
    free(permT);
    orc_destroy(h);
    free(p);
  }
  int pc[20 * 5];
  orc_generate_correlated(20, 5, 3, pc);
  orc_generate_machine_correlated(20, 5, 3, pc);
  printf(fails ? "MISMATCH\n" : "ok\n");
  return fails ? 1 : 0;
}
