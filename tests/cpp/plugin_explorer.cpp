// The reference's CPU explorer loop (explorer::_explore_cpu, explorer.cu
// 92-148) written against the plugin interfaces of wsbb_plugins.hpp, with
// the GPU bound behind bound_abstract, wired by a composition root that
// makes the reference's own calls (pbab::pbab, pbab.cpp:49-53).
// Usage: plugin_explorer <taillard id>
// Prints: "<best> <decomposed> <bounded> <leaves> <pruned>".
//
// plugin_explorer --resume <taillard id> <budget> <file>: a device search
// (4096-child pools, ten sub-batches) stopped after <budget> parents is saved
// to <file> as a pool checkpoint, then continued from that file in a fresh
// context.  Prints: "<best> <proven> <counters...>".
//
// plugin_explorer --records <taillard id> <budget> <file>: the same search
// stopped after <budget> parents, its pool written to <file> as the
// reference's binary node records (184 bytes each), read back and pushed
// into a fresh context's live pool (gpu_pool) with the incumbent, and run
// to the end.  Prints: "<best> <records> <first decomposed> <rest decomposed>".
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "wsbb_plugins.hpp"

using namespace wsbb_plugins;

// The composition root of the reference, reduced to what its constructor
// wires between the plugins (src/pbab.cpp:30-53): the same calls in the
// same order.
class pbab {
 public:
  int size;
  instance_abstract* instance;
  bound_abstract* bound;
  decompose_beginend* decompose;
  pbab(decompose_beginend* _decompose, instance_abstract* _instance, bound_abstract* _bound)
      : size(_instance->size), instance(_instance), bound(_bound), decompose(_decompose) {
    decompose->set_pbab(this);
    decompose->set_bound(bound);
    bound->set_pbab(this);
    bound->set_instance(instance);
  }
};

static std::vector<int32_t> taillard_times(int tid, int* n, int* m) {
  if (wsbb_taillard(tid, n, m, nullptr) != WSBB_OK) return {};
  std::vector<int32_t> p(static_cast<size_t>(*n) * *m);
  wsbb_taillard(tid, n, m, p.data());
  return p;
}

static int records_check(int tid, long long budget, const char* path) {
  int n = 0, m = 0;
  const std::vector<int32_t> p = taillard_times(tid, &n, &m);
  if (p.empty()) return 2;
  int32_t best = 0;
  std::vector<int32_t> perm(n);
  std::vector<int64_t> st(WSBB_NSTATS);
  {
    wsbb_ctx* ctx = nullptr;
    if (wsbb_create(p.data(), n, m, 0, &ctx) != WSBB_OK) return 3;
    const int rc = wsbb_solve(ctx, WSBB_LB2, 4096, 10, INT32_MAX, budget, &best, perm.data(), st.data());
    if (rc != WSBB_EBUDGET) return 4;
    const std::vector<char> bytes = records_from_pool(snapshot(ctx));
    std::ofstream f(path, std::ios::binary);
    f.write(bytes.data(), static_cast<std::streamsize>(bytes.size()));
    wsbb_destroy(ctx);
  }
  std::ifstream f(path, std::ios::binary);
  const std::vector<char> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  const std::vector<raw_bb_problem> recs = deserialize_nodes(bytes.data(), bytes.size());
  bound_hip bound;
  bound.set_times(p.data(), n, m);
  gpu_pool pool(bound, 4096, 10);
  pool.set_incumbent(best, perm[0] >= 0 ? perm.data() : nullptr);
  pool.push(recs);
  std::vector<int64_t> st2(WSBB_NSTATS);
  if (pool.run(0, 0, st2.data()) != 0) return 5;
  std::printf("%d %zu %lld %lld\n", pool.incumbent(), recs.size(), static_cast<long long>(st[0]),
              static_cast<long long>(st2[0]));
  return 0;
}

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
  if (argc >= 5 && std::string(argv[1]) == "--records")
    return records_check(std::atoi(argv[2]), std::atoll(argv[3]), argv[4]);
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
  decompose_beginend decompose;
  pbab root(&decompose, &inst, &bound);

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
