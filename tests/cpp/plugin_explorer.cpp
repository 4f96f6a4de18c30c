// The reference's CPU explorer loop (explorer::_explore_cpu, explorer.cu
// 92-148) written against the plugin interfaces of wsbb_plugins.hpp, with
// the GPU bound behind bound_abstract.  Usage: plugin_explorer <taillard id>
// Prints: "<best> <decomposed> <bounded> <leaves> <pruned>".
//
// plugin_explorer --resume <taillard id> <budget> <file>: a device search
// (4096-child pools, ten sub-batches) stopped after <budget> parents is saved
// to <file> as a pool checkpoint, then continued from that file in a fresh
// context.  Prints: "<best> <proven> <counters...>".
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "wsbb_plugins.hpp"

using namespace wsbb_plugins;

static int resume_check(int tid, long long budget, const char* path) {
  int n = 0, m = 0;
  if (wsbb_taillard(tid, &n, &m, nullptr) != WSBB_OK) return 2;
  std::vector<int32_t> p(static_cast<size_t>(n) * m);
  wsbb_taillard(tid, &n, &m, p.data());
  int32_t best = 0;
  std::vector<int32_t> perm(n);
  std::vector<int64_t> st(WSBB_NSTATS);
  {
    wsbb_ctx* ctx = nullptr;
    if (wsbb_create(p.data(), n, m, 0, &ctx) != WSBB_OK) return 3;
    const int rc = wsbb_solve(ctx, WSBB_LB2, 4096, 10, INT32_MAX, budget, &best, perm.data(), st.data());
    if (rc != WSBB_EBUDGET) return 4;
    std::ofstream f(path);
    save_pool(f, snapshot(ctx));
    wsbb_destroy(ctx);
  }
  std::ifstream f(path);
  const pool_snapshot s = load_pool(f);
  wsbb_ctx* ctx = nullptr;
  if (wsbb_create(p.data(), n, m, 0, &ctx) != WSBB_OK) return 3;
  const int rc = resume(ctx, s, WSBB_LB2, 4096, 10, 0, &best, perm.data(), st.data());
  wsbb_destroy(ctx);
  std::printf("%d %d", best, rc == WSBB_OK ? 1 : 0);
  for (int64_t v : st) std::printf(" %lld", static_cast<long long>(v));
  std::printf("\n");
  return 0;
}

int main(int argc, char** argv) {
  if (argc >= 5 && std::string(argv[1]) == "--resume")
    return resume_check(std::atoi(argv[2]), std::atoll(argv[3]), argv[4]);
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <taillard id>\n", argv[0]);
    return 2;
  }
  int n = 0, m = 0;
  if (wsbb_taillard(std::atoi(argv[1]), &n, &m, nullptr) != WSBB_OK) return 2;
  std::vector<int32_t> p(static_cast<size_t>(n) * m);
  wsbb_taillard(std::atoi(argv[1]), &n, &m, p.data());
  // the instance text stream of instance_flowshop::generate_instance
  std::stringstream data;
  data << n << ' ' << m << ' ' << 0;
  for (int32_t v : p) data << ' ' << v;
  instance_abstract inst;
  inst.size = n;
  inst.data = &data;

  bound_hip bound;
  bound.set_instance(&inst);
  decompose_beginend decompose(&bound, n);

  // DFS pool: children pushed worst first leave rank 0 on top, the order of
  // the reference's priority queue over ranks (tree.cpp:37-46)
  std::vector<raw_bb_problem*> pool{make_root(n)};
  int best = INT_MAX;
  long long decomposed = 0, bounded = 0, leaves = 0, pruned = 0;
  while (!pool.empty()) {
    raw_bb_problem* cur = pool.back();
    pool.pop_back();
    if (cur->depth == n) {  // leaf: solutions::insert, strictly better only
      ++leaves;
      if (cur->couts[0] < best) best = cur->couts[0];
    } else if (cur->couts[0] < best) {  // solutions::inserable
      problems& ns = decompose(*cur);
      ++decomposed;
      bounded += (cur->depth == n - 2) ? 2 : 2 * (n - cur->depth);
      for (raw_bb_problem* c : ns) pool.push_back(new raw_bb_problem(*c));
    } else {
      ++pruned;
    }
    delete cur;
  }
  std::printf("%d %lld %lld %lld %lld\n", best, decomposed, bounded, leaves, pruned);
  return 0;
}
