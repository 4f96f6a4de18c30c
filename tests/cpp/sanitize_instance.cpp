// Drives the product's host-side instance code (wsbb_amd/csrc/instance.cpp:
// the Taillard and correlated generators, node batches and the Johnson /
// lag / R0 / Q0 precompute, including rejected shapes) for the sanitizer
// build (tests/test_sanitizers.py).  Prints "ok".
#include <cstdio>
#include <vector>

#include "../../wsbb_amd/csrc/instance.h"

int main() {
  using namespace wsbb;
  int fails = 0;
  for (int id : {1, 11, 21, 51, 81}) {
    int n = 0, m = 0;
    long seed = 0;
    if (!taillard_dims(id, &n, &m, &seed)) return 2;
    std::vector<int32_t> p(static_cast<size_t>(n) * m);
    taillard_generate(n, m, seed, p.data());
    Instance I;
    fails += !I.build(p.data(), n, m).empty();
    std::vector<uint8_t> permT(static_cast<size_t>(n) * 500);
    std::vector<int32_t> l1(500), l2(500);
    random_nodes(n, 500, 7, permT.data(), l1.data(), l2.data());
  }
  // Taillard 111 has 500 jobs: beyond the 128-job limit, refused
  {
    int n = 0, m = 0;
    long seed = 0;
    taillard_dims(111, &n, &m, &seed);
    std::vector<int32_t> big(static_cast<size_t>(n) * m);
    taillard_generate(n, m, seed, big.data());
    Instance B;
    fails += B.build(big.data(), n, m).empty();
  }
  std::vector<int32_t> q(50 * 20);
  correlated_generate(50, 20, 9, q.data());
  machine_correlated_generate(50, 20, 9, q.data());
  Instance J;
  fails += !J.build(q.data(), 50, 20).empty();
  // rejected shapes and values leave the instance unbuilt, without faults
  std::vector<int32_t> bad(200 * 30, 1);
  Instance K;
  fails += K.build(bad.data(), 1, 5).empty();
  fails += K.build(bad.data(), 129, 5).empty();
  fails += K.build(bad.data(), 20, 25).empty();
  bad[3] = -1;
  fails += K.build(bad.data(), 20, 5).empty();
  std::printf(fails ? "MISMATCH\n" : "ok\n");
  return fails ? 1 : 0;
}
