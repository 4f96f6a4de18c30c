// Instance precompute for the flowshop bounds.
//
// Follows bound_flowshop::initialiserVar (src/bound_flowshop.cpp:556-576):
// machine pairs in lexicographic order (remplirMachine 500-508), lags
// (remplirLag 486-498), Johnson orders with the reference's quicksort tie
// order (Johnson 627-643, partionner/quicksort 597-625), least release and
// delivery times R0/Q0 over all jobs (remplirTempsArriverDepart 510-538).
#include "instance.h"

#include <algorithm>

namespace wsbb {

long taillard_unif(long* seed, long low, long high) {
  constexpr long kM = 2147483647, kA = 16807, kQ = 127773, kR = 2836;
  long hi = *seed / kQ;
  long s = kA * (*seed % kQ) - hi * kR;
  if (s < 0) s += kM;
  *seed = s;
  // the reference divides in float and keeps the quotient in a double
  double u = static_cast<double>(static_cast<float>(s) / static_cast<float>(kM));
  return low + static_cast<long>(u * static_cast<double>(high - low + 1));
}

static const long kTaillardSeeds[120] = {
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

bool taillard_dims(int id, int* n, int* m, long* seed) {
  if (id < 1 || id > 120) return false;
  // benchmark classes of 10 instances: (jobs, machines)
  static const int kClass[12][2] = {{20, 5},   {20, 10},  {20, 20},  {50, 5},
                                    {50, 10},  {50, 20},  {100, 5},  {100, 10},
                                    {100, 20}, {200, 10}, {200, 20}, {500, 20}};
  int c = (id - 1) / 10;
  *n = kClass[c][0];
  *m = kClass[c][1];
  *seed = kTaillardSeeds[id - 1];
  return true;
}

void taillard_generate(int n, int m, long seed, int32_t* p) {
  for (int k = 0; k < m; ++k)
    for (int j = 0; j < n; ++j)
      p[k * n + j] = static_cast<int32_t>(taillard_unif(&seed, 1, 99));
}

uint64_t splitmix64(uint64_t* state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void correlated_generate(int n, int m, uint64_t seed, int32_t* p) {
  uint64_t st = seed;
  std::vector<int32_t> a(n);
  for (int j = 0; j < n; ++j) a[j] = 1 + static_cast<int32_t>(splitmix64(&st) % 80);
  for (int k = 0; k < m; ++k)
    for (int j = 0; j < n; ++j)
      p[k * n + j] = a[j] + static_cast<int32_t>(splitmix64(&st) % 21);
}

void machine_correlated_generate(int n, int m, uint64_t seed, int32_t* p) {
  uint64_t st = seed;
  std::vector<int32_t> b(m);
  for (int k = 0; k < m; ++k) b[k] = 1 + static_cast<int32_t>(splitmix64(&st) % 80);
  for (int k = 0; k < m; ++k)
    for (int j = 0; j < n; ++j)
      p[k * n + j] = b[k] + static_cast<int32_t>(splitmix64(&st) % 21);
}

void random_nodes(int n, int N, uint64_t seed0, uint8_t* permT, int32_t* l1,
                  int32_t* l2) {
  std::vector<int> perm(n);
  for (int k = 0; k < N; ++k) {
    uint64_t st = seed0 + static_cast<uint64_t>(k);
    int d1 = static_cast<int>(splitmix64(&st) % static_cast<uint64_t>(n));
    int d2 = static_cast<int>(splitmix64(&st) % static_cast<uint64_t>(n - d1));
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int i = n - 1; i > 0; --i) {
      int r = static_cast<int>(splitmix64(&st) % static_cast<uint64_t>(i + 1));
      std::swap(perm[i], perm[r]);
    }
    for (int i = 0; i < n; ++i)
      permT[static_cast<size_t>(i) * N + k] = static_cast<uint8_t>(perm[i]);
    l1[k] = d1 - 1;
    l2[k] = n - d2;
  }
}

namespace {

// Hoare partition with the first element as pivot and the two inner scans
// strictly against the pivot's key, recursing on [lo, split] and
// [split+1, hi]: the scheme of partionner/quicksort (bound_flowshop.cpp
// 597-625), which fixes the order of jobs with equal Johnson keys.
void hoare_sort(std::vector<int>& idx, int lo, int hi,
                const std::vector<int64_t>& key) {
  while (hi - lo > 0) {
    const int64_t pivot = key[idx[lo]];
    int i = lo - 1, j = hi + 1;
    for (;;) {
      do --j; while (key[idx[j]] > pivot);
      do ++i; while (key[idx[i]] < pivot);
      if (i >= j) break;
      std::swap(idx[i], idx[j]);
    }
    hoare_sort(idx, lo, j, key);
    lo = j + 1;  // tail call on the right part
  }
}

}  // namespace

std::string Instance::build(const int32_t* p_mm, int n_, int m_) {
  if (n_ < 2 || n_ > kMaxJobs) return "number of jobs must be in [2, 128]";
  if (m_ < 2 || m_ > kMaxMachines) return "number of machines must be in [2, 24]";
  n = n_;
  m = m_;
  P = m * (m - 1) / 2;
  p.assign(p_mm, p_mm + static_cast<size_t>(m) * n);
  for (int32_t v : p)
    if (v < 0 || v > 65535) return "processing times must be in [0, 65535]";

  // Any monotone path of the job x machine grid enters each job row once and
  // steps down at most m-1 times, so this bounds every head, tail, Johnson
  // value and makespan the kernels form.
  int64_t pb = 0;
  for (int j = 0; j < n; ++j) {
    int mx = 0;
    for (int k = 0; k < m; ++k) mx = std::max(mx, p[k * n + j]);
    pb += mx;
  }
  for (int k = 0; k < m; ++k) {
    int mx = 0;
    for (int j = 0; j < n; ++j) mx = std::max(mx, p[k * n + j]);
    pb += mx;
  }
  if (pb >= kPathLimit)
    return "instance too large for the 24-bit bound arithmetic: sum over jobs of the largest "
           "time plus sum over machines of the largest time must stay below 2^19";
  path_bound = static_cast<int>(pb);

  pair_a.resize(P);
  pair_b.resize(P);
  for (int a = 0, s = 0; a < m - 1; ++a)
    for (int b = a + 1; b < m; ++b, ++s) {
      pair_a[s] = a;
      pair_b[s] = b;
    }

  lag.assign(static_cast<size_t>(P) * n, 0);
  for (int s = 0; s < P; ++s)
    for (int k = pair_a[s] + 1; k < pair_b[s]; ++k)
      for (int j = 0; j < n; ++j) lag[s * n + j] += p[k * n + j];

  // Johnson's rule for the two-machine relaxation with lags: jobs faster on
  // the first machine come first by increasing (p_a + lag), the others last
  // by decreasing (p_b + lag).  One integer key orders both groups.
  johnson.resize(static_cast<size_t>(P) * n);
  std::vector<int64_t> key(n);
  std::vector<int> idx(n);
  for (int s = 0; s < P; ++s) {
    const int a = pair_a[s], b = pair_b[s];
    for (int j = 0; j < n; ++j) {
      const int L = lag[s * n + j];
      key[j] = (p[a * n + j] < p[b * n + j])
                   ? static_cast<int64_t>(p[a * n + j] + L)
                   : (int64_t{1} << 40) - (p[b * n + j] + L);
      idx[j] = j;
    }
    hoare_sort(idx, 0, n - 1, key);
    std::copy(idx.begin(), idx.end(), johnson.begin() + static_cast<size_t>(s) * n);
  }

  // R0[k]: least time any job needs before it can reach machine k;
  // Q0[k]: least time any job needs after leaving machine k.
  R0.assign(m, 0);
  Q0.assign(m, 0);
  for (int k = 1; k < m; ++k) {
    int best = 0;
    for (int j = 0; j < n; ++j) {
      int acc = 0;
      for (int kk = 0; kk < k; ++kk) acc += p[kk * n + j];
      best = (j == 0) ? acc : std::min(best, acc);
    }
    R0[k] = best;
  }
  for (int k = 0; k < m - 1; ++k) {
    int best = 0;
    for (int j = 0; j < n; ++j) {
      int acc = 0;
      for (int kk = k + 1; kk < m; ++kk) acc += p[kk * n + j];
      best = (j == 0) ? acc : std::min(best, acc);
    }
    Q0[k] = best;
  }
  ptot.assign(m, 0);
  for (int k = 0; k < m; ++k)
    for (int j = 0; j < n; ++j) ptot[k] += p[k * n + j];
  return {};
}

}  // namespace wsbb
