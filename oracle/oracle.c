/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Plain-C restatement of the reference's CPU branch-and-bound path for the
 * permutation flowshop.  The restatement follows the reference's own
 * algorithm step by step (including its quicksort tie order and its
 * m-pass tail computation) so that it can serve as the bit-exact checker
 * of the HIP product and as the timed CPU baseline.  Reference citations are
 * relative to the reference source tree.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <limits.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Instance generation: src/instance_flowshop.cpp                            */
/* ------------------------------------------------------------------------ */

/* time_seeds[] of src/instance_flowshop.cpp:13-52 (Taillard's seeds). */
static const long k_time_seeds[120] = {
    873654221,  379008056,  1866992158, 216771124,  495070989,  402959317,
    1369363414, 2021925980, 573109518,  88325120,   587595453,  1401007982,
    873136276,  268827376,  1634173168, 691823909,  73807235,   1273398721,
    2065119309, 1672900551, 479340445,  268827376,  1958948863, 918272953,
    555010963,  2010851491, 1519833303, 1748670931, 1923497586, 1829909967,
    1328042058, 200382020,  496319842,  1203030903, 1730708564, 450926852,
    1303135678, 1273398721, 587288402,  248421594,  1958948863, 575633267,
    655816003,  1977864101, 93805469,   1803345551, 49612559,   1899802599,
    2013025619, 578962478,  1539989115, 691823909,  655816003,  1315102446,
    1949668355, 1923497586, 1805594913, 1861070898, 715643788,  464843328,
    896678084,  1179439976, 1122278347, 416756875,  267829958,  1835213917,
    1328833962, 1418570761, 161033112,  304212574,  1539989115, 655816003,
    960914243,  1915696806, 2013025619, 1168140026, 1923497586, 167698528,
    1528387973, 993794175,  450926852,  1462772409, 1021685265, 83696007,
    508154254,  1861070898, 26482542,   444956424,  2115448041, 118254244,
    471503978,  1215892992, 135346136,  1602504050, 160037322,  551454346,
    519485142,  383947510,  1968171878, 540872513,  2013025619, 475051709,
    914834335,  810642687,  1019331795, 2056065863, 1342855162, 1325809384,
    1988803007, 765656702,  1368624604, 450181436,  1927888393, 1759567256,
    606425239,  19268348,   1298201670, 2041736264, 379756761,  28837162};

/* instance_flowshop::unif, src/instance_flowshop.cpp:74-82: Park-Miller
 * minimal standard generator (Schrage's factorisation), float-cast to [0,1). */
long orc_unif(long *seed, long low, long high) {
  const long m = 2147483647, a = 16807, b = 127773, c = 2836;
  long k = (*seed) / b;
  *seed = a * ((*seed) % b) - k * c;
  if (*seed < 0) *seed += m;
  float q = (float)(*seed) / (float)m; /* float division, as in the source */
  double v = (double)q;                /* stored into a double there */
  return low + (long)(v * (double)(high - low + 1));
}

/* get_job_number / get_machine_number, src/instance_flowshop.cpp:55-71 */
int orc_taillard_dims(int id, int *n, int *m, long *seed) {
  if (id < 1 || id > 120) return -1;
  int N = 20, M = 5;
  if (id > 110) N = 500;
  else if (id > 90) N = 200;
  else if (id > 60) N = 100;
  else if (id > 30) N = 50;
  if (id > 110) M = 20;
  else if (id > 100) M = 20;
  else if (id > 90) M = 10;
  else if (id > 80) M = 20;
  else if (id > 70) M = 10;
  else if (id > 60) M = 5;
  else if (id > 50) M = 20;
  else if (id > 40) M = 10;
  else if (id > 30) M = 5;
  else if (id > 20) M = 20;
  else if (id > 10) M = 10;
  *n = N;
  *m = M;
  *seed = k_time_seeds[id - 1];
  return 0;
}

/* generate_instance, src/instance_flowshop.cpp:84-89: machine-major stream */
void orc_generate(int n, int m, long seed, int *p) {
  long s = seed;
  for (int k = 0; k < m; k++)
    for (int j = 0; j < n; j++) p[k * n + j] = (int)orc_unif(&s, 1, 99);
}

uint64_t orc_splitmix64(uint64_t *state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void orc_generate_correlated(int n, int m, uint64_t seed, int *p) {
  uint64_t st = seed;
  int a[ORC_MAXN + 1];
  for (int j = 0; j < n; j++) a[j] = 1 + (int)(orc_splitmix64(&st) % 80);
  for (int k = 0; k < m; k++)
    for (int j = 0; j < n; j++)
      p[k * n + j] = a[j] + (int)(orc_splitmix64(&st) % 21);
}

void orc_generate_machine_correlated(int n, int m, uint64_t seed, int *p) {
  uint64_t st = seed;
  int b[ORC_MAXM];
  for (int k = 0; k < m; k++) b[k] = 1 + (int)(orc_splitmix64(&st) % 80);
  for (int k = 0; k < m; k++)
    for (int j = 0; j < n; j++)
      p[k * n + j] = b[k] + (int)(orc_splitmix64(&st) % 21);
}

void orc_random_nodes(int n, int N, uint64_t seed0, uint8_t *permT, int *l1,
                      int *l2) {
  int perm[ORC_MAXN + 1];
  for (int k = 0; k < N; k++) {
    uint64_t st = seed0 + (uint64_t)k;
    int d1 = (int)(orc_splitmix64(&st) % (uint64_t)n);
    int d2 = (int)(orc_splitmix64(&st) % (uint64_t)(n - d1));
    for (int i = 0; i < n; i++) perm[i] = i;
    for (int i = n - 1; i > 0; i--) {
      int r = (int)(orc_splitmix64(&st) % (uint64_t)(i + 1));
      int t = perm[i];
      perm[i] = perm[r];
      perm[r] = t;
    }
    for (int i = 0; i < n; i++) permT[(size_t)i * N + k] = (uint8_t)perm[i];
    l1[k] = d1 - 1;
    l2[k] = n - d2;
  }
}

/* ------------------------------------------------------------------------ */
/* Instance precompute: bound_flowshop::initialiserVar and helpers           */
/* ------------------------------------------------------------------------ */

struct orc_inst {
  int n, m, P;
  int *p;       /* tempsJob [m][n] */
  int *ma, *mb; /* machine[0], machine[1] [P] */
  int *lag;     /* tempsLag [P][n] */
  int *john;    /* tabJohnson [P][n] */
  int *R0, *Q0; /* minTempsArr, minTempsDep [m] */
};

/* quicksort state: pluspetit[0] (group) and pluspetit[1] (key) */
typedef struct {
  const int *grp, *key;
} orc_sortkeys;

/* estInf / estSup, src/bound_flowshop.cpp:579-595 */
static int est_inf(const orc_sortkeys *s, int i, int j) {
  if (s->grp[i] == s->grp[j]) {
    if (s->grp[i] == 1) return s->key[i] < s->key[j];
    return s->key[i] > s->key[j];
  }
  return s->grp[i] < s->grp[j];
}
static int est_sup(const orc_sortkeys *s, int i, int j) {
  if (s->grp[i] == s->grp[j]) {
    if (s->grp[i] == 1) return s->key[i] > s->key[j];
    return s->key[i] < s->key[j];
  }
  return s->grp[i] > s->grp[j];
}
/* partionner / quicksort, src/bound_flowshop.cpp:597-625 (Hoare partition,
 * pivot = first element's job, recursion on [deb,k] and [k+1,fin]). */
static int partition_(const orc_sortkeys *s, int *ordo, int deb, int fin) {
  int d = deb - 1, f = fin + 1, pivot = ordo[deb];
  do {
    do f--; while (est_sup(s, ordo[f], pivot));
    do d++; while (est_inf(s, ordo[d], pivot));
    if (d < f) {
      int t = ordo[d];
      ordo[d] = ordo[f];
      ordo[f] = t;
    }
  } while (d < f);
  return f;
}
static void quicksort_(const orc_sortkeys *s, int *ordo, int deb, int fin) {
  if (fin - deb > 0) {
    int k = partition_(s, ordo, deb, fin);
    quicksort_(s, ordo, deb, k);
    quicksort_(s, ordo, k + 1, fin);
  }
}

orc_inst *orc_create(const int *p, int n, int m) {
  if (n < 2 || n > ORC_MAXN || m < 1 || m > ORC_MAXM) return NULL;
  orc_inst *h = (orc_inst *)calloc(1, sizeof(orc_inst));
  h->n = n;
  h->m = m;
  h->P = m * (m - 1) / 2; /* initSomme, bound_flowshop.cpp:359-363 */
  int P = h->P;
  h->p = (int *)malloc(sizeof(int) * m * n);
  memcpy(h->p, p, sizeof(int) * m * n);
  h->ma = (int *)malloc(sizeof(int) * (P ? P : 1));
  h->mb = (int *)malloc(sizeof(int) * (P ? P : 1));
  h->lag = (int *)malloc(sizeof(int) * (P ? P : 1) * n);
  h->john = (int *)malloc(sizeof(int) * (P ? P : 1) * n);
  h->R0 = (int *)malloc(sizeof(int) * m);
  h->Q0 = (int *)malloc(sizeof(int) * m);

  /* remplirTempsArriverDepart, bound_flowshop.cpp:510-538 */
  {
    int done_arr[ORC_MAXM] = {0}, done_dep[ORC_MAXM] = {0};
    int arr[ORC_MAXM], dep[ORC_MAXM];
    h->Q0[m - 1] = 0;
    h->R0[0] = 0;
    for (int i = 0; i < n; i++) {
      arr[0] = 0;
      dep[m - 1] = 0;
      for (int k = 1; k < m; k++) {
        int mach = m - k - 1;
        arr[k] = arr[k - 1] + p[(k - 1) * n + i];
        dep[mach] = dep[mach + 1] + p[(mach + 1) * n + i];
        if (!done_arr[k] || h->R0[k] > arr[k]) {
          done_arr[k] = 1;
          h->R0[k] = arr[k];
        }
        if (!done_dep[k] || h->Q0[mach] > dep[mach]) {
          done_dep[k] = 1;
          h->Q0[mach] = dep[mach];
        }
      }
    }
  }
  /* remplirMachine, bound_flowshop.cpp:500-508 */
  {
    int c = 0;
    for (int i = 0; i < m - 1; i++)
      for (int j = i + 1; j < m; j++) {
        h->ma[c] = i;
        h->mb[c] = j;
        c++;
      }
  }
  /* remplirLag, bound_flowshop.cpp:486-498 */
  for (int s = 0; s < P; s++)
    for (int j = 0; j < n; j++) {
      int L = 0;
      for (int k = h->ma[s] + 1; k < h->mb[s]; k++) L += p[k * n + j];
      h->lag[s * n + j] = L;
    }
  /* remplirTabJohnson + Johnson, bound_flowshop.cpp:540-547, 627-643 */
  {
    int grp[ORC_MAXN], key[ORC_MAXN];
    orc_sortkeys sk = {grp, key};
    for (int s = 0; s < P; s++) {
      int a = h->ma[s], b = h->mb[s];
      int *ordo = h->john + s * n;
      for (int i = 0; i < n; i++) {
        ordo[i] = i;
        if (p[a * n + i] < p[b * n + i]) {
          grp[i] = 1;
          key[i] = p[a * n + i] + h->lag[s * n + i];
        } else {
          grp[i] = 2;
          key[i] = p[b * n + i] + h->lag[s * n + i];
        }
      }
      quicksort_(&sk, ordo, 0, n - 1);
    }
  }
  return h;
}

void orc_destroy(orc_inst *h) {
  if (!h) return;
  free(h->p);
  free(h->ma);
  free(h->mb);
  free(h->lag);
  free(h->john);
  free(h->R0);
  free(h->Q0);
  free(h);
}

int orc_pairs(const orc_inst *h) { return h->P; }

void orc_tables(const orc_inst *h, int *ma, int *mb, int *lag, int *johnson,
                int *R0, int *Q0) {
  int P = h->P, n = h->n, m = h->m;
  if (ma) memcpy(ma, h->ma, sizeof(int) * P);
  if (mb) memcpy(mb, h->mb, sizeof(int) * P);
  if (lag) memcpy(lag, h->lag, sizeof(int) * P * n);
  if (johnson) memcpy(johnson, h->john, sizeof(int) * P * n);
  if (R0) memcpy(R0, h->R0, sizeof(int) * m);
  if (Q0) memcpy(Q0, h->Q0, sizeof(int) * m);
}

/* ------------------------------------------------------------------------ */
/* Per-node computations                                                     */
/* ------------------------------------------------------------------------ */

typedef struct {
  int r[ORC_MAXM];    /* tempsMachines: heads of the prefix */
  int q[ORC_MAXM];    /* tempsMachinesFin: tails of the suffix */
  uint8_t job[ORC_MAXN]; /* job[]: 0 = unscheduled */
  int d1, d2;         /* nbAffectDebut, nbAffectFin */
} orc_node_state;

static inline int imax(int a, int b) { return a > b ? a : b; }

/* criteres_calculer, bound_flowshop.cpp:897-909 */
int orc_cmax(const orc_inst *h, const int *perm) {
  int t[ORC_MAXM];
  int n = h->n, m = h->m;
  for (int k = 0; k < m; k++) t[k] = 0;
  for (int j = 0; j < n; j++) {
    int jb = perm[j];
    t[0] = t[0] + h->p[jb];
    for (int k = 1; k < m; k++) t[k] = imax(t[k], t[k - 1]) + h->p[k * n + jb];
  }
  return t[m - 1];
}

/* set_tempsMachines_retardDebut 841-856, set_job_jobFin 858-868,
 * set_nombres 870-874, set_tempsMachinesFin_tempsJobFin 876-895 */
static void node_state(const orc_inst *h, const int *perm, int l1, int l2,
                       orc_node_state *st) {
  int n = h->n, m = h->m;
  const int *p = h->p;
  for (int k = 0; k < m; k++) st->r[k] = 0;
  for (int j = 0; j <= l1; j++) {
    int jb = perm[j];
    st->r[0] += p[jb];
    for (int k = 1; k < m; k++)
      st->r[k] = imax(st->r[k], st->r[k - 1]) + p[k * n + jb];
  }
  for (int j = 0; j <= l1; j++) st->job[perm[j]] = (uint8_t)1;
  for (int j = l1 + 1; j < l2; j++) st->job[perm[j]] = 0;
  for (int j = l2; j < n; j++) st->job[perm[j]] = (uint8_t)1;
  st->d1 = l1 + 1;
  st->d2 = n - l2;
  int tmp[ORC_MAXM];
  for (int j = m - 1; j >= 0; j--) {
    for (int k = j; k < m; k++) tmp[k] = 0;
    for (int k = n - st->d2; k < n; k++) {
      int jb = perm[k];
      tmp[j] += p[j * n + jb];
      for (int l = j + 1; l < m; l++) {
        if (tmp[l - 1] > tmp[l])
          tmp[l] = tmp[l - 1] + p[l * n + jb];
        else
          tmp[l] += p[l * n + jb];
      }
    }
    st->q[j] = tmp[m - 1];
  }
}

void orc_heads_tails(const orc_inst *h, const int *perm, int l1, int l2,
                     int *r, int *q) {
  orc_node_state st;
  node_state(h, perm, l1, l2, &st);
  memcpy(r, st.r, sizeof(int) * h->m);
  memcpy(q, st.q, sizeof(int) * h->m);
}

/* one pair: initCmax 79-90, heuristiqueCmax 64-77, cmaxFin 92-100 and the
 * nbAffectFin==0 branch of borneInfMakespan 140-149 */
static int pair_value(const orc_inst *h, const orc_node_state *st, int s) {
  int n = h->n;
  const int *p = h->p;
  int a = h->ma[s], b = h->mb[s];
  int t0, t1;
  if (st->d1 != 0) {
    t0 = st->r[a];
    t1 = st->r[b];
  } else {
    t0 = h->R0[a];
    t1 = h->R0[b];
  }
  const int *jo = h->john + s * n;
  const int *lg = h->lag + s * n;
  for (int k = 0; k < n; k++) {
    int jb = jo[k];
    if (st->job[jb] == 0) {
      t0 += p[a * n + jb];
      if (t1 > t0 + lg[jb])
        t1 += p[b * n + jb];
      else
        t1 = t0 + lg[jb] + p[b * n + jb];
    }
  }
  if (st->d2 != 0) {
    if (t1 + st->q[b] > t0 + st->q[a]) return t1 + st->q[b];
    return t0 + st->q[a];
  }
  return t1 + h->Q0[b];
}

/* bornes_calculer 929-946 -> calculBorne 807-817 -> borneInfMakespan
 * 121-168 (ordoSomme[l] = l, initNbFois 549-554; lambda1 = 0, lambda2 = 1,
 * init 925-926; early exit when the running max exceeds `best`, 154-161). */
int orc_bound(const orc_inst *h, const int *perm, int l1, int l2, int best) {
  if (l2 - l1 == 1) return orc_cmax(h, perm);
  orc_node_state st;
  node_state(h, perm, l1, l2, &st);
  int worst = 0;
  for (int s = 0; s < h->P; s++) {
    int v = pair_value(h, &st, s);
    if (v > worst) worst = v;
    if (worst > best && best != -1) return worst;
  }
  return worst;
}

int orc_pair_values(const orc_inst *h, const int *perm, int l1, int l2,
                    int *v) {
  orc_node_state st;
  node_state(h, perm, l1, l2, &st);
  int worst = 0;
  for (int s = 0; s < h->P; s++) {
    v[s] = pair_value(h, &st, s);
    if (v[s] > worst) worst = v[s];
  }
  return worst;
}

int orc_bound_subset(const orc_inst *h, const int *perm, int l1, int l2,
                     const uint8_t *sel) {
  if (l2 - l1 == 1) return orc_cmax(h, perm);
  orc_node_state st;
  node_state(h, perm, l1, l2, &st);
  int worst = 0;
  for (int s = 0; s < h->P; s++)
    if (sel[s]) {
      int v = pair_value(h, &st, s);
      if (v > worst) worst = v;
    }
  return worst;
}

/* LB1 has no counterpart in the reference (its only other bound is a
 * tardiness bound that is commented out, bound_flowshop.cpp:808).  It is the
 * textbook one-machine bound written with the reference's head/tail
 * conventions: head = r[k] if a prefix exists else R0[k]; tail = q[k] if a
 * suffix exists else Q0[k]; LB1 = max_k head + sum_{unscheduled} p[k][j] +
 * tail.  Leaves return the makespan, as bornes_calculer does. */
int orc_lb1(const orc_inst *h, const int *perm, int l1, int l2) {
  if (l2 - l1 == 1) return orc_cmax(h, perm);
  orc_node_state st;
  node_state(h, perm, l1, l2, &st);
  int n = h->n, m = h->m, best = 0;
  for (int k = 0; k < m; k++) {
    int sum = 0;
    for (int i = l1 + 1; i < l2; i++) sum += h->p[k * n + perm[i]];
    int head = st.d1 ? st.r[k] : h->R0[k];
    int tail = st.d2 ? st.q[k] : h->Q0[k];
    best = imax(best, head + sum + tail);
  }
  return best;
}

static int node_bound(const orc_inst *h, int kind, const int *perm, int l1,
                      int l2) {
  if (kind == ORC_LB1) return orc_lb1(h, perm, l1, l2);
  return orc_bound(h, perm, l1, l2, 999999); /* bound_flowshop.cpp:948-951 */
}

typedef struct {
  const orc_inst *h;
  const uint8_t *permT;
  const int *l1, *l2;
  int N, kind, lo, hi;
  int *out;
} batch_job;

static void *batch_worker(void *arg) {
  batch_job *b = (batch_job *)arg;
  int n = b->h->n;
  int perm[ORC_MAXN];
  for (int k = b->lo; k < b->hi; k++) {
    for (int i = 0; i < n; i++) perm[i] = b->permT[(size_t)i * b->N + k];
    b->out[k] = node_bound(b->h, b->kind, perm, b->l1[k], b->l2[k]);
  }
  return NULL;
}

void orc_bound_batch(const orc_inst *h, const uint8_t *permT, const int *l1,
                     const int *l2, int N, int kind, int *out, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  batch_job jobs[256];
  int chunk = (N + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; t++) {
    jobs[t] = (batch_job){h, permT, l1, l2, N, kind, t * chunk,
                          (t + 1) * chunk < N ? (t + 1) * chunk : N, out};
    if (jobs[t].lo > jobs[t].hi) jobs[t].lo = jobs[t].hi;
  }
  if (nthreads == 1) {
    batch_worker(&jobs[0]);
    return;
  }
  for (int t = 0; t < nthreads; t++)
    pthread_create(&th[t], NULL, batch_worker, &jobs[t]);
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
}

/* ------------------------------------------------------------------------ */
/* Branching                                                                 */
/* ------------------------------------------------------------------------ */

typedef struct {
  int perm[ORC_MAXN];
  int l1, l2, depth, cost, job;
} orc_child;

/* create_raw_problem3 (bb_problem_calculator.cpp:25-33) = limites_set 40-48,
 * depth_set 59-64, permutation_set 66-74; ranks are handled by the caller. */
static void make_child(const orc_inst *h, const int *perm, int l1, int l2,
                       int depth, int j, int side, orc_child *c) {
  int n = h->n;
  c->l1 = l1;
  c->l2 = l2;
  if (side == 0) c->l1++;
  else c->l2--;
  if (c->l2 - c->l1 == 2) c->l1++;
  c->depth = (depth == n - 2) ? n : depth + 1;
  c->job = perm[j];
  memcpy(c->perm, perm, sizeof(int) * n);
  int t = (side == 0) ? l1 + 1 : l2 - 1;
  int tmp = c->perm[t];
  c->perm[t] = c->perm[j];
  c->perm[j] = tmp;
}

/* select_worst (select_worst.cpp:7-12): n1 before n2 iff cost higher, then
 * job higher.  Used to sort a side; ranks count from the back (problems.cpp
 * 46-62), so rank 0 = lowest (cost, job). */
static int worst_first(const orc_child *a, const orc_child *b) {
  if (a->cost != b->cost) return b->cost < a->cost;
  return a->job > b->job;
}
static void sort_side(orc_child **v, int cnt) { /* insertion sort, cnt <= n */
  for (int i = 1; i < cnt; i++) {
    orc_child *x = v[i];
    int k = i - 1;
    while (k >= 0 && worst_first(x, v[k])) {
      v[k + 1] = v[k];
      k--;
    }
    v[k + 1] = x;
  }
}

/* problems::coupe / sum (problems.cpp:13-33) and decompose_beginend::choise
 * (decompose_beginend.cpp:39-50) with best = 99999 (decompose_beginend.h:26,
 * never updated).  Returns 0 = BEGIN, 1 = END. */
static int choose_side(orc_child *const *beg, int nb, orc_child *const *end,
                       int ne) {
  const int best = 99999;
  int c0 = 0, c1 = 0, s0 = 0, s1 = 0;
  for (int i = 0; i < nb; i++) {
    if (beg[i]->cost >= best) c0++;
    else s0 += beg[i]->cost;
  }
  for (int i = 0; i < ne; i++) {
    if (end[i]->cost >= best) c1++;
    else s1 += end[i]->cost;
  }
  if (c0 > c1) return 0;
  if (c1 > c0) return 1;
  return (s0 > s1) ? 0 : 1;
}

/* generates and bounds both sides of one parent (or the two leaves of a
 * simple parent), picks the side; returns the chosen side's children sorted
 * by select_worst in `out` and their count. */
typedef struct {
  orc_child kids[2 * ORC_MAXN];
  orc_child *beg[ORC_MAXN], *end[ORC_MAXN];
  int nb, ne, side, nbounded;
} orc_expansion;

static void expand(const orc_inst *h, int kind, const int *perm, int l1,
                   int l2, int depth, orc_expansion *x) {
  int n = h->n, c = 0;
  x->nb = x->ne = 0;
  if (depth == n - 2) { /* simple: leaves_generate 70-89 */
    for (int t = 1; t <= 2; t++) {
      orc_child *k = &x->kids[c++];
      make_child(h, perm, l1, l2, depth, l1 + t, 0, k);
      k->depth = n;
      k->cost = node_bound(h, kind, k->perm, k->l1, k->l2);
      x->beg[x->nb++] = k;
    }
  } else { /* problems_generate 26-37, 52-68 */
    for (int side = 0; side < 2; side++)
      for (int j = l1 + 1; j < l2; j++) {
        orc_child *k = &x->kids[c++];
        make_child(h, perm, l1, l2, depth, j, side, k);
        k->cost = node_bound(h, kind, k->perm, k->l1, k->l2);
        if (side == 0) x->beg[x->nb++] = k;
        else x->end[x->ne++] = k;
      }
  }
  x->nbounded = c;
  x->side = choose_side(x->beg, x->nb, x->end, x->ne);
  if (x->side == 0) sort_side(x->beg, x->nb);
  else sort_side(x->end, x->ne);
}

int orc_decompose(const orc_inst *h, const int *perm, int l1, int l2,
                  int depth, int kind, int *cperm, int *cl1, int *cl2,
                  int *cdepth, int *ccost, int *cjob, int *crank, int *side,
                  int *nbounded) {
  orc_expansion *x = (orc_expansion *)malloc(sizeof(orc_expansion));
  expand(h, kind, perm, l1, l2, depth, x);
  orc_child **v = x->side == 0 ? x->beg : x->end;
  int cnt = x->side == 0 ? x->nb : x->ne;
  for (int i = 0; i < cnt; i++) {
    memcpy(cperm + i * h->n, v[i]->perm, sizeof(int) * h->n);
    cl1[i] = v[i]->l1;
    cl2[i] = v[i]->l2;
    cdepth[i] = v[i]->depth;
    ccost[i] = v[i]->cost;
    cjob[i] = v[i]->job;
    crank[i] = cnt - 1 - i; /* reverse iteration assigns rank 0 to the last */
  }
  *side = x->side;
  *nbounded = x->nbounded;
  free(x);
  return cnt;
}

/* ------------------------------------------------------------------------ */
/* Node pool                                                                 */
/* ------------------------------------------------------------------------ */
/* Node record: header + perm[n] + ranks[n+1] (bytes).  The CPU explorer's
 * priority_queue<comparator_depth> (tree.h:76, tree.cpp:37-46) pops the node
 * with the lexicographically smallest ranks[].  Two nodes in the pool are
 * never ancestor and descendant, so their rank strings differ within their
 * common depth; consequently all children of the popped node sort before
 * every other pool node, and a LIFO stack that pushes children in
 * descending rank order pops nodes in exactly the same sequence.  order=0
 * keeps the heap (the faithful form), order=1 the stack. */

typedef struct {
  int16_t l1, l2, depth;
  int32_t cost;
} node_hdr;

typedef struct {
  int n, stride;
  unsigned char *mem;
  int64_t cap, size;  /* arena slots */
  int32_t *heap;      /* heap mode: slot indices */
  int64_t hsize, hcap;
  int32_t *freel;
  int64_t nfree, fcap;
  int order;
} pool_t;

#define NODE(pl, i) ((pl)->mem + (size_t)(i) * (pl)->stride)
#define NHDR(pl, i) ((node_hdr *)NODE(pl, i))
#define NPERM(pl, i) (NODE(pl, i) + sizeof(node_hdr))
#define NRANK(pl, i) (NODE(pl, i) + sizeof(node_hdr) + (pl)->n)

static void pool_init(pool_t *pl, int n, int order) {
  memset(pl, 0, sizeof(*pl));
  pl->n = n;
  pl->stride = (int)((sizeof(node_hdr) + n + n + 1 + 7) & ~7u);
  pl->cap = 1024;
  pl->mem = (unsigned char *)malloc((size_t)pl->cap * pl->stride);
  pl->order = order;
  if (order == 0) {
    pl->hcap = 1024;
    pl->heap = (int32_t *)malloc(sizeof(int32_t) * pl->hcap);
    pl->fcap = 1024;
    pl->freel = (int32_t *)malloc(sizeof(int32_t) * pl->fcap);
  }
}
static void pool_free(pool_t *pl) {
  free(pl->mem);
  free(pl->heap);
  free(pl->freel);
}
static int64_t pool_count(const pool_t *pl) {
  return pl->order == 0 ? pl->hsize : pl->size;
}
/* lexicographic rank comparison: returns <0 if a before b */
static int rank_cmp(const pool_t *pl, int64_t a, int64_t b) {
  const unsigned char *ra = NRANK(pl, a), *rb = NRANK(pl, b);
  for (int d = 0; d <= pl->n; d++)
    if (ra[d] != rb[d]) return (int)ra[d] - (int)rb[d];
  return 0;
}
static int64_t pool_alloc_slot(pool_t *pl) {
  if (pl->order == 0 && pl->nfree > 0) return pl->freel[--pl->nfree];
  if (pl->size == pl->cap) {
    pl->cap *= 2;
    pl->mem = (unsigned char *)realloc(pl->mem, (size_t)pl->cap * pl->stride);
  }
  return pl->size++;
}
static void heap_push(pool_t *pl, int32_t slot) {
  if (pl->hsize == pl->hcap) {
    pl->hcap *= 2;
    pl->heap = (int32_t *)realloc(pl->heap, sizeof(int32_t) * pl->hcap);
  }
  int64_t i = pl->hsize++;
  while (i > 0) {
    int64_t par = (i - 1) / 2;
    if (rank_cmp(pl, slot, pl->heap[par]) >= 0) break;
    pl->heap[i] = pl->heap[par];
    i = par;
  }
  pl->heap[i] = slot;
}
static int32_t heap_pop(pool_t *pl) {
  int32_t top = pl->heap[0];
  int32_t last = pl->heap[--pl->hsize];
  int64_t i = 0, n = pl->hsize;
  for (;;) {
    int64_t c = 2 * i + 1;
    if (c >= n) break;
    if (c + 1 < n && rank_cmp(pl, pl->heap[c + 1], pl->heap[c]) < 0) c++;
    if (rank_cmp(pl, pl->heap[c], last) >= 0) break;
    pl->heap[i] = pl->heap[c];
    i = c;
  }
  if (n > 0) pl->heap[i] = last;
  return top;
}
static void free_slot(pool_t *pl, int32_t slot) {
  if (pl->nfree == pl->fcap) {
    pl->fcap *= 2;
    pl->freel = (int32_t *)realloc(pl->freel, sizeof(int32_t) * pl->fcap);
  }
  pl->freel[pl->nfree++] = slot;
}

/* node popped from the pool, copied out */
typedef struct {
  int perm[ORC_MAXN];
  unsigned char ranks[ORC_MAXN + 1];
  int l1, l2, depth, cost;
} popped;

static int pool_pop(pool_t *pl, popped *o) {
  int64_t slot;
  if (pl->order == 0) {
    if (pl->hsize == 0) return 0;
    slot = heap_pop(pl);
  } else {
    if (pl->size == 0) return 0;
    slot = --pl->size;
  }
  node_hdr *hd = NHDR(pl, slot);
  o->l1 = hd->l1;
  o->l2 = hd->l2;
  o->depth = hd->depth;
  o->cost = hd->cost;
  for (int i = 0; i < pl->n; i++) o->perm[i] = NPERM(pl, slot)[i];
  memcpy(o->ranks, NRANK(pl, slot), pl->n + 1);
  if (pl->order == 0) free_slot(pl, (int32_t)slot);
  return 1;
}
/* pushes one node; in stack mode nodes must be pushed from last to first */
static void pool_push(pool_t *pl, const int *perm, int l1, int l2, int depth,
                      int cost, const unsigned char *ranks) {
  int64_t slot = pool_alloc_slot(pl);
  node_hdr *hd = NHDR(pl, slot);
  hd->l1 = (int16_t)l1;
  hd->l2 = (int16_t)l2;
  hd->depth = (int16_t)depth;
  hd->cost = cost;
  for (int i = 0; i < pl->n; i++) NPERM(pl, slot)[i] = (unsigned char)perm[i];
  memcpy(NRANK(pl, slot), ranks, pl->n + 1);
  if (pl->order == 0) heap_push(pl, (int32_t)slot);
}
/* repush a node that was just popped (stack: on top again; heap: re-insert) */
static void pool_repush(pool_t *pl, const popped *o) {
  pool_push(pl, o->perm, o->l1, o->l2, o->depth, o->cost, o->ranks);
}

static void push_root(pool_t *pl, int n) {
  /* create_raw_problem 9-22 and tree::add_root (tree.cpp:170-176) */
  int perm[ORC_MAXN];
  unsigned char ranks[ORC_MAXN + 1];
  for (int j = 0; j < n; j++) perm[j] = j;
  memset(ranks, 0, sizeof(ranks));
  pool_push(pl, perm, -1, n, 0, 0, ranks);
}

/* ------------------------------------------------------------------------ */
/* Search trace (test infrastructure): one record of 2n + 8 int32 per event, */
/* [type, round, l1, l2, depth, cost, 0, perm[n], ranks[n+1]], in the order  */
/* the explorer meets them.  Types: ORC_TR_TAKEN (popped and decomposed),    */
/* ORC_TR_PRUNED (popped, bound not below the incumbent), ORC_TR_LEAF (leaf  */
/* popped by the DFS explorer, or leaf child met by the batched explorer's   */
/* prune loop, which never stores leaves), ORC_TR_REPUSH (popped by the      */
/* batched explorer's collection loop, did not fit the sub-batch, pushed     */
/* back).  Not re-entrant: one traced search at a time.                      */
/* ------------------------------------------------------------------------ */
static int32_t *g_trace;
static int64_t g_trace_cap, g_trace_len;
static int g_trace_n;

void orc_trace_begin(int32_t *buf, int64_t cap_records, int n) {
  g_trace = buf;
  g_trace_cap = cap_records;
  g_trace_len = 0;
  g_trace_n = n;
}
int64_t orc_trace_end(void) {
  int64_t r = g_trace_len;
  g_trace = NULL;
  g_trace_cap = g_trace_len = 0;
  return r;
}
static void trace_rec(int type, int64_t round, int l1, int l2, int depth,
                      int cost, const int *perm, const unsigned char *ranks) {
  if (!g_trace) return;
  if (g_trace_len < g_trace_cap) {
    int n = g_trace_n;
    int32_t *r = g_trace + (size_t)g_trace_len * (size_t)(2 * n + 8);
    r[0] = type;
    r[1] = (int32_t)round;
    r[2] = l1;
    r[3] = l2;
    r[4] = depth;
    r[5] = cost;
    r[6] = 0;
    for (int i = 0; i < n; i++) r[7 + i] = perm[i];
    for (int i = 0; i <= n; i++) r[7 + n + i] = ranks ? ranks[i] : 0;
  }
  g_trace_len++; /* counts past the capacity, so the caller can tell */
}

/* ------------------------------------------------------------------------ */
/* CPU explorer: explorer::_explore_cpu (explorer.cu:92-148)                 */
/* ------------------------------------------------------------------------ */

int orc_solve_dfs(const orc_inst *h, int kind, int order, int ub0,
                  int64_t budget, int64_t *stats, int *best_perm) {
  int n = h->n;
  int64_t st[ORC_NSTATS] = {0};
  pool_t pl;
  pool_init(&pl, n, order);
  push_root(&pl, n);
  int best = ub0;
  int have = 0;
  int exhausted = 0;
  popped cur;
  orc_expansion *x = (orc_expansion *)malloc(sizeof(orc_expansion));
  while (pool_pop(&pl, &cur)) {
    if (cur.depth == n) { /* bb_problem_calculator::leaf 54-56 */
      trace_rec(ORC_TR_LEAF, st[0], cur.l1, cur.l2, cur.depth, cur.cost,
                cur.perm, cur.ranks);
      st[2]++;
      /* solutions::insert (solutions.cpp:45-65): strictly better only */
      if (!have ? cur.cost < best : cur.cost < best) {
        best = cur.cost;
        have = 1;
        st[5]++;
        if (best_perm) memcpy(best_perm, cur.perm, sizeof(int) * n);
      }
    } else if (cur.cost < best) { /* solutions::inserable 37-43 */
      if (budget > 0 && st[0] >= budget) {
        exhausted = 1;
        break;
      }
      trace_rec(ORC_TR_TAKEN, st[0], cur.l1, cur.l2, cur.depth, cur.cost,
                cur.perm, cur.ranks);
      expand(h, kind, cur.perm, cur.l1, cur.l2, cur.depth, x);
      st[0]++;
      st[1] += x->nbounded;
      orc_child **v = x->side == 0 ? x->beg : x->end;
      int cnt = x->side == 0 ? x->nb : x->ne;
      /* ranks (problems.cpp:46-62): rank written at ranks[depth] for inner
       * children, ranks[depth-1] for leaves of a simple parent (ranks2).
       * v[] is sorted worst first, so v[i] has rank cnt-1-i; pushing v[0]
       * first leaves rank 0 on top of the stack. */
      for (int i = 0; i < cnt; i++) {
        unsigned char rk[ORC_MAXN + 1];
        memcpy(rk, cur.ranks, n + 1);
        int pos = (cur.depth == n - 2) ? v[i]->depth - 1 : v[i]->depth;
        rk[pos] = (unsigned char)(cnt - 1 - i);
        pool_push(&pl, v[i]->perm, v[i]->l1, v[i]->l2, v[i]->depth,
                  v[i]->cost, rk);
      }
      if (pool_count(&pl) > st[6]) st[6] = pool_count(&pl);
    } else {
      trace_rec(ORC_TR_PRUNED, st[0], cur.l1, cur.l2, cur.depth, cur.cost,
                cur.perm, cur.ranks);
      st[3]++;
    }
  }
  free(x);
  pool_free(&pl);
  if (stats) memcpy(stats, st, sizeof(st));
  return exhausted ? -1 : best;
}

/* ------------------------------------------------------------------------ */
/* Batched GPU-path explorer: explorer::_explore_cpu_gpu (explorer.cu:154-305)*/
/* with gpu::decompose_on_cpu (gpu.cu:180-255) and gpu::prune (257-408).     */
/* Departures from the reference (its defects, not its behaviour):           */
/*  - child reservation uses n instead of MAX_NBJOBS (problems.cpp:68);      */
/*  - both sides prune against the incumbent (gpu.cu:344 vs 381 differ only */
/*    when an initial UB is passed with -o);                                 */
/*  - leaves_generate3's writes to child[0] (decompose_beginend.cpp:131-137)*/
/*    are not reproduced: leaves get depth n.                                */
/* ------------------------------------------------------------------------ */

int orc_solve_batched(const orc_inst *h, int kind, int pool_size,
                      int max_batches, int ub0, int64_t budget,
                      int64_t *stats, int *best_perm) {
  return orc_solve_batched_ex(h, kind, pool_size, max_batches, ub0, budget, 1,
                              stats, best_perm);
}

int orc_solve_batched_ex(const orc_inst *h, int kind, int pool_size,
                         int max_batches, int ub0, int64_t budget, int order,
                         int64_t *stats, int *best_perm) {
  int n = h->n;
  if (pool_size - 1 <= 2 * n || max_batches < 1) return -2;
  int64_t st[ORC_NSTATS] = {0};
  pool_t pl;
  pool_init(&pl, n, order);
  push_root(&pl, n);
  int best = ub0;
  int exhausted = 0;
  int64_t pcap = 1024;
  popped *par = (popped *)malloc(sizeof(popped) * pcap);
  orc_expansion *xs = NULL;
  int64_t xcap = 0;
  while (pool_count(&pl) > 0) {
    st[7]++;
    /* collect up to max_batches sub-batches (explorer.cu:166-208) */
    int64_t np = 0;
    int nb = 0;
    while (pool_count(&pl) > 0 && nb < max_batches) {
      int size1 = 0, cnt = 0;
      popped cur;
      while (pool_pop(&pl, &cur)) {
        if (cur.cost < best) { /* tree::bound, tree.cpp:247-250 */
          int fils = 2 * (n - cur.depth);
          if (size1 + fils < pool_size - 1) { /* problems::push_back1 */
            trace_rec(ORC_TR_TAKEN, st[7], cur.l1, cur.l2, cur.depth,
                      cur.cost, cur.perm, cur.ranks);
            if (np == pcap) {
              pcap *= 2;
              par = (popped *)realloc(par, sizeof(popped) * pcap);
            }
            par[np++] = cur;
            size1 += fils;
            cnt++;
          } else {
            trace_rec(ORC_TR_REPUSH, st[7], cur.l1, cur.l2, cur.depth,
                      cur.cost, cur.perm, cur.ranks);
            pool_repush(&pl, &cur);
            break;
          }
        } else {
          trace_rec(ORC_TR_PRUNED, st[7], cur.l1, cur.l2, cur.depth,
                    cur.cost, cur.perm, cur.ranks);
          st[3]++;
        }
      }
      if (cnt > 0) nb++;
    }
    if (np == 0) break;
    if (budget > 0 && st[0] + np > budget) {
      exhausted = 1;
      break;
    }
    st[0] += np;
    if (np > xcap) {
      free(xs);
      xcap = np;
      xs = (orc_expansion *)malloc(sizeof(orc_expansion) * xcap);
    }
    for (int64_t i = 0; i < np; i++) {
      expand(h, kind, par[i].perm, par[i].l1, par[i].l2, par[i].depth, &xs[i]);
      st[1] += xs[i].nbounded;
    }
    /* prune parent by parent in pop order (gpu.cu:317-394) */
    int64_t *keep = (int64_t *)calloc(np, sizeof(int64_t));
    for (int64_t i = 0; i < np; i++) {
      orc_expansion *x = &xs[i];
      orc_child **v = x->side == 0 ? x->beg : x->end;
      int cnt = x->side == 0 ? x->nb : x->ne;
      int kept = 0;
      for (int c = 0; c < cnt; c++) { /* worst first */
        if (v[c]->depth == n) {
          trace_rec(ORC_TR_LEAF, st[7], v[c]->l1, v[c]->l2, v[c]->depth,
                    v[c]->cost, v[c]->perm, par[i].ranks);
          st[2]++;
          if (v[c]->cost < best) {
            best = v[c]->cost;
            st[5]++;
            if (best_perm) memcpy(best_perm, v[c]->perm, sizeof(int) * n);
          }
        } else if (v[c]->cost < best) {
          kept++;
        } else {
          st[4]++;
        }
      }
      keep[i] = kept;
    }
    /* survivors of one parent are its `kept` best-ranked children (a
     * suffix of the worst-first order); push parents last to first. */
    for (int64_t i = np - 1; i >= 0; i--) {
      orc_expansion *x = &xs[i];
      orc_child **v = x->side == 0 ? x->beg : x->end;
      int cnt = x->side == 0 ? x->nb : x->ne;
      unsigned char rk[ORC_MAXN + 1];
      for (int c = cnt - (int)keep[i]; c < cnt; c++) {
        memcpy(rk, par[i].ranks, n + 1);
        rk[v[c]->depth] = (unsigned char)(cnt - 1 - c);
        pool_push(&pl, v[c]->perm, v[c]->l1, v[c]->l2, v[c]->depth,
                  v[c]->cost, rk);
      }
    }
    free(keep);
    if (pool_count(&pl) > st[6]) st[6] = pool_count(&pl);
  }
  free(xs);
  free(par);
  pool_free(&pl);
  if (stats) memcpy(stats, st, sizeof(st));
  return exhausted ? -1 : best;
}

/* ------------------------------------------------------------------------ */
/* All-cores CPU search (CPU baseline only): the CPU explorer's rules        */
/* (explorer.cu:92-148) run by nthreads threads over independent subtrees.   */
/* The top of the tree is expanded sequentially, level by level, until the  */
/* frontier holds at least 64 nodes per thread; the frontier's nodes are     */
/* then taken one at a time (top of the stack first) by the threads, each    */
/* running the DFS explorer on its subtree with one shared incumbent (read   */
/* at every pop, lowered under a lock by strictly better leaves).  A thread  */
/* that finds the shared stack empty raises a flag, and a busy thread then   */
/* gives it the bottom (shallowest) node of its own stack.  The optimum is   */
/* the sequential one; node counts depend on when incumbents are shared.     */
/* ------------------------------------------------------------------------ */
typedef struct {
  const orc_inst *h;
  int kind, nthreads;
  pool_t *roots; /* shared subtrees (stack), taken and given under lock */
  pthread_mutex_t lock;
  pthread_cond_t cond;
  int idle, done;
  volatile int hungry; /* an idle thread waits for work */
  volatile int best;
  int have;
  int *best_perm;
  int64_t st[ORC_NSTATS];
} par_ctx;

/* the bottom node of a stack pool (its shallowest: the largest subtree) */
static void pool_take_bottom(pool_t *pl, popped *o) {
  node_hdr *hd = NHDR(pl, 0);
  o->l1 = hd->l1;
  o->l2 = hd->l2;
  o->depth = hd->depth;
  o->cost = hd->cost;
  for (int i = 0; i < pl->n; i++) o->perm[i] = NPERM(pl, 0)[i];
  memcpy(o->ranks, NRANK(pl, 0), pl->n + 1);
  memmove(pl->mem, pl->mem + pl->stride, (size_t)(pl->size - 1) * pl->stride);
  pl->size--;
}

static void *par_worker(void *arg) {
  par_ctx *P = (par_ctx *)arg;
  const orc_inst *h = P->h;
  int n = h->n;
  int64_t st[ORC_NSTATS] = {0};
  pool_t pl;
  pool_init(&pl, n, 1);
  orc_expansion *x = (orc_expansion *)malloc(sizeof(orc_expansion));
  popped cur;
  for (;;) {
    int got = 0;
    pthread_mutex_lock(&P->lock);
    for (;;) {
      if (pool_pop(P->roots, &cur)) {
        got = 1;
        break;
      }
      if (P->done) break;
      if (++P->idle == P->nthreads) { /* no thread holds work: finished */
        P->done = 1;
        pthread_cond_broadcast(&P->cond);
        break;
      }
      P->hungry = 1;
      pthread_cond_wait(&P->cond, &P->lock);
      P->idle--;
    }
    pthread_mutex_unlock(&P->lock);
    if (!got) break;
    pool_repush(&pl, &cur);
    while (pool_pop(&pl, &cur)) {
      if (P->hungry && pl.size >= 2) { /* give the shallowest node away */
        popped give;
        pool_take_bottom(&pl, &give);
        pthread_mutex_lock(&P->lock);
        pool_repush(P->roots, &give);
        P->hungry = P->idle > 1;
        pthread_cond_signal(&P->cond);
        pthread_mutex_unlock(&P->lock);
      }
      int best = __atomic_load_n(&P->best, __ATOMIC_RELAXED);
      if (cur.depth == n) {
        st[2]++;
        if (cur.cost < best) {
          pthread_mutex_lock(&P->lock);
          if (cur.cost < P->best) { /* solutions::insert: strictly better */
            __atomic_store_n(&P->best, cur.cost, __ATOMIC_RELAXED);
            P->have = 1;
            if (P->best_perm) memcpy(P->best_perm, cur.perm, sizeof(int) * n);
            st[5]++;
          }
          pthread_mutex_unlock(&P->lock);
        }
      } else if (cur.cost < best) {
        expand(h, P->kind, cur.perm, cur.l1, cur.l2, cur.depth, x);
        st[0]++;
        st[1] += x->nbounded;
        orc_child **v = x->side == 0 ? x->beg : x->end;
        int cnt = x->side == 0 ? x->nb : x->ne;
        for (int i = 0; i < cnt; i++) {
          unsigned char rk[ORC_MAXN + 1];
          memcpy(rk, cur.ranks, n + 1);
          int pos = (cur.depth == n - 2) ? v[i]->depth - 1 : v[i]->depth;
          rk[pos] = (unsigned char)(cnt - 1 - i);
          pool_push(&pl, v[i]->perm, v[i]->l1, v[i]->l2, v[i]->depth, v[i]->cost, rk);
        }
      } else {
        st[3]++;
      }
    }
  }
  free(x);
  pool_free(&pl);
  pthread_mutex_lock(&P->lock);
  for (int i = 0; i < ORC_NSTATS; i++) P->st[i] += st[i];
  pthread_mutex_unlock(&P->lock);
  return NULL;
}

int orc_solve_parallel(const orc_inst *h, int kind, int nthreads, int ub0,
                       int64_t *stats, int *best_perm) {
  int n = h->n;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  par_ctx P;
  memset(&P, 0, sizeof(P));
  P.h = h;
  P.kind = kind;
  P.nthreads = nthreads;
  P.best = ub0;
  P.best_perm = best_perm;
  pthread_mutex_init(&P.lock, NULL);
  pthread_cond_init(&P.cond, NULL);
  pool_t roots;
  pool_init(&roots, n, 1);
  push_root(&roots, n);
  P.roots = &roots;
  /* sequential top: expand level by level (each level in DFS order) until
   * the frontier holds at least 64 nodes per thread */
  orc_expansion *x = (orc_expansion *)malloc(sizeof(orc_expansion));
  int64_t cnt = 1, cap = 1024;
  popped *fr = (popped *)malloc(sizeof(popped) * cap);
  pool_pop(&roots, &fr[0]);
  while (cnt > 0 && cnt < 64 * nthreads) {
    int64_t ncnt = 0, ncap = 1024;
    popped *nf = (popped *)malloc(sizeof(popped) * ncap);
    for (int64_t f = 0; f < cnt; f++) {
      popped *cur = &fr[f];
      if (cur->depth == n) {
        P.st[2]++;
        if (cur->cost < P.best) {
          P.best = cur->cost;
          P.have = 1;
          P.st[5]++;
          if (best_perm) memcpy(best_perm, cur->perm, sizeof(int) * n);
        }
        continue;
      }
      if (cur->cost >= P.best) {
        P.st[3]++;
        continue;
      }
      expand(h, kind, cur->perm, cur->l1, cur->l2, cur->depth, x);
      P.st[0]++;
      P.st[1] += x->nbounded;
      orc_child **v = x->side == 0 ? x->beg : x->end;
      int c = x->side == 0 ? x->nb : x->ne;
      for (int i = c - 1; i >= 0; i--) { /* rank 0 (v[c-1]) first: DFS order */
        if (ncnt == ncap) {
          ncap *= 2;
          nf = (popped *)realloc(nf, sizeof(popped) * ncap);
        }
        popped *o = &nf[ncnt++];
        memcpy(o->perm, v[i]->perm, sizeof(int) * n);
        memcpy(o->ranks, cur->ranks, n + 1);
        int pos = (cur->depth == n - 2) ? v[i]->depth - 1 : v[i]->depth;
        o->ranks[pos] = (unsigned char)(c - 1 - i);
        o->l1 = v[i]->l1;
        o->l2 = v[i]->l2;
        o->depth = v[i]->depth;
        o->cost = v[i]->cost;
      }
    }
    free(fr);
    fr = nf;
    cnt = ncnt;
  }
  /* the first subtree in DFS order ends on top of the shared stack */
  for (int64_t f = cnt - 1; f >= 0; f--) pool_repush(&roots, &fr[f]);
  free(fr);
  free(x);
  pthread_t th[256];
  for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, par_worker, &P);
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  pool_free(&roots);
  pthread_cond_destroy(&P.cond);
  pthread_mutex_destroy(&P.lock);
  if (stats) memcpy(stats, P.st, sizeof(P.st));
  return P.best;
}
