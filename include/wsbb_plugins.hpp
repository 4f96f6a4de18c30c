// wsbb_plugins.hpp -- C++ face of the engine in the shape of the reference's
// operator plugins, header-only over the C ABI of wsbb_amd.h.
//
// It keeps the interfaces a caller of the reference codes against:
//   struct bound_abstract      (headers/bound_abstract.h:9-19)
//   class select_abstract      (headers/select_abstract.h:4-7)
//   select_worst               (src/select_worst.cpp:7-12)
//   class problems             (headers/problems.h:13-28, src/problems.cpp)
//   decompose_beginend         (headers/decompose_beginend.h:15-44,
//                               src/decompose_beginend.cpp:26-107)
// and adds bound_hip, the GPU bound behind bound_abstract.  A per-node
// bornes_calculer is a batch of one; bornes_calculer_batch bounds a whole
// vector in one launch, which is how decompose_beginend uses it.  It also
// reads and writes the reference's text formats (node records, bab.solutions
// blocks, the formater CSV row) and a pool checkpoint file of the device
// search built from those records.
#ifndef WSBB_PLUGINS_HPP
#define WSBB_PLUGINS_HPP

#include <algorithm>
#include <cstdint>
#include <fstream>
#include <istream>
#include <ostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "wsbb_amd.h"

namespace wsbb_plugins {

constexpr int MAX_NBJOBS = 128;  // the reference fixes 20 (tree.h:31)
constexpr int BEGIN_ORDER = 0;   // decompose_beginend.cpp:17-18
constexpr int END_ORDER = 1;

// Node record with the reference's raw_bb_problem fields
// (bb_problem_calculator.cpp:101-144).
struct raw_bb_problem {
  int couts[2];
  int couts_somme;
  int job;
  int permutation[MAX_NBJOBS];
  int ranks[MAX_NBJOBS + 1];
  int limite1, limite2, depth, begin_end;
};

// instance_abstract (headers/instance_abstract.h): the "N M seed p..." text
// stream written by instance_flowshop::generate_instance.
struct instance_abstract {
  int size = 0;
  std::stringstream* data = nullptr;
};

struct bound_abstract {
  virtual void bornes_calculer(int permutation[], int limite1, int limite2, int* couts,
                               int best) = 0;
  virtual void bornes_calculer(raw_bb_problem& p) = 0;
  virtual void set_instance(instance_abstract* instance) = 0;
  virtual ~bound_abstract() = default;
};

class select_abstract {
 public:
  virtual bool operator()(raw_bb_problem const* n1, raw_bb_problem const* n2) const = 0;
  virtual ~select_abstract() = default;
};

// cost descending, then job descending
class select_worst : public select_abstract {
 public:
  bool operator()(raw_bb_problem const* n1, raw_bb_problem const* n2) const override {
    if (n1->couts_somme != n2->couts_somme) return n2->couts_somme < n1->couts_somme;
    return n1->job > n2->job;
  }
};

class bound_hip : public bound_abstract {
 public:
  explicit bound_hip(int device = 0, int kind = WSBB_LB2) : device_(device), kind_(kind) {}
  ~bound_hip() override { wsbb_destroy(ctx_); }
  bound_hip(const bound_hip&) = delete;
  bound_hip& operator=(const bound_hip&) = delete;

  // bound_flowshop::set_instance -> init (bound_flowshop.cpp:911-927,
  // 953-956): reads N, M, seed and the machine-major times, precomputes and
  // uploads the tables.
  void set_instance(instance_abstract* instance) override {
    std::stringstream& in = *instance->data;
    in.clear();
    in.seekg(0);
    long seed = 0;
    in >> n_ >> m_ >> seed;
    std::vector<int32_t> p(static_cast<size_t>(n_) * m_);
    for (auto& v : p) in >> v;
    if (!in) throw std::runtime_error("bad instance stream");
    wsbb_destroy(ctx_);
    ctx_ = nullptr;
    check(wsbb_create(p.data(), n_, m_, device_, &ctx_));
  }
  void set_times(const int32_t* p_machine_major, int n, int m) {
    n_ = n;
    m_ = m;
    wsbb_destroy(ctx_);
    ctx_ = nullptr;
    check(wsbb_create(p_machine_major, n, m, device_, &ctx_));
  }

  // bornes_calculer(perm, l1, l2, couts, best), bound_flowshop.cpp:929-946
  void bornes_calculer(int permutation[], int limite1, int limite2, int* couts,
                       int best) override {
    std::vector<uint8_t> permT(n_);
    for (int i = 0; i < n_; ++i) permT[i] = static_cast<uint8_t>(permutation[i]);
    int32_t out = 0;
    check(wsbb_bound_batch(ctx_, permT.data(), &limite1, &limite2, 1, kind_, best, &out));
    couts[0] = out;
    couts[1] = 0;
  }
  // bornes_calculer(raw_bb_problem&), bound_flowshop.cpp:948-951
  void bornes_calculer(raw_bb_problem& p) override {
    bornes_calculer(p.permutation, p.limite1, p.limite2, p.couts, 999999);
    p.couts_somme = p.couts[0] + p.couts[1];
  }
  // every node of the vector in one launch
  void bornes_calculer_batch(std::vector<raw_bb_problem*>& nodes) {
    const int N = static_cast<int>(nodes.size());
    if (N == 0) return;
    std::vector<uint8_t> permT(static_cast<size_t>(n_) * N);
    std::vector<int32_t> l1(N), l2(N), out(N);
    for (int k = 0; k < N; ++k) {
      for (int i = 0; i < n_; ++i)
        permT[static_cast<size_t>(i) * N + k] = static_cast<uint8_t>(nodes[k]->permutation[i]);
      l1[k] = nodes[k]->limite1;
      l2[k] = nodes[k]->limite2;
    }
    check(wsbb_bound_batch(ctx_, permT.data(), l1.data(), l2.data(), N, kind_, -1, out.data()));
    for (int k = 0; k < N; ++k) {
      nodes[k]->couts[0] = out[k];
      nodes[k]->couts[1] = 0;
      nodes[k]->couts_somme = out[k];
    }
  }
  int size() const { return n_; }
  wsbb_ctx* context() { return ctx_; }

 private:
  static void check(int rc) {
    if (rc != WSBB_OK) throw std::runtime_error(std::string("wsbb: ") + wsbb_last_error());
  }
  int device_, kind_;
  int n_ = 0, m_ = 0;
  wsbb_ctx* ctx_ = nullptr;
};

// problems (src/problems.cpp): a side's children with the ranking helpers
class problems : public std::vector<raw_bb_problem*> {
 public:
  select_worst select2;
  int sum(int best) const {  // problems.cpp:13-22
    int r = 0;
    for (auto* p : *this)
      if (p->couts_somme < best) r += p->couts_somme;
    return r;
  }
  int coupe(int best) const {  // problems.cpp:24-33
    int r = 0;
    for (auto* p : *this)
      if (p->couts_somme >= best) ++r;
    return r;
  }
  // sort worst first, rank 0 to the last (problems.cpp:46-62)
  void ranks(bool leaves) {
    std::sort(begin(), end(), [this](raw_bb_problem* a, raw_bb_problem* b) { return select2(a, b); });
    int rank = 0;
    for (auto it = rbegin(); it != rend(); ++it)
      (*it)->ranks[leaves ? (*it)->depth - 1 : (*it)->depth] = rank++;
  }
  void release() {
    for (auto* p : *this) delete p;
    clear();
  }
};

// decompose_beginend::operator() (decompose_beginend.cpp:91-107): children
// of both sides (both sides bounded in one launch), the side with more
// pruned children or, on a tie, the larger bound sum wins (choise 39-50,
// best fixed at 99999), ranked; returns the chosen side, owned by the
// decomposer until the next call.
class decompose_beginend {
 public:
  decompose_beginend(bound_hip* bound, int size) : bound_(bound), size_(size) {}
  ~decompose_beginend() { ns_.release(); }

  problems& operator()(raw_bb_problem& n) {
    ns_.release();
    problems debut, fin;
    if (n.depth == size_ - 2) {  // leaves_generate 70-89
      for (int t = 1; t <= 2; ++t) debut.push_back(child(n, n.limite1 + t, BEGIN_ORDER, true));
    } else {                     // problems_generate 26-37
      for (int j = n.limite1 + 1; j < n.limite2; ++j) debut.push_back(child(n, j, BEGIN_ORDER, false));
      for (int j = n.limite1 + 1; j < n.limite2; ++j) fin.push_back(child(n, j, END_ORDER, false));
    }
    std::vector<raw_bb_problem*> all(debut.begin(), debut.end());
    all.insert(all.end(), fin.begin(), fin.end());
    bound_->bornes_calculer_batch(all);
    const bool begin_side = choise(debut, fin) == BEGIN_ORDER;
    problems& keep = begin_side ? debut : fin;
    problems& drop = begin_side ? fin : debut;
    drop.release();
    ns_.assign(keep.begin(), keep.end());
    keep.clear();
    ns_.ranks(n.depth == size_ - 2);
    return ns_;
  }

  int choise(const problems& debut, const problems& fin) const {
    const int c0 = debut.coupe(best_), c1 = fin.coupe(best_);
    if (c0 > c1) return BEGIN_ORDER;
    if (c1 > c0) return END_ORDER;
    return debut.sum(best_) > fin.sum(best_) ? BEGIN_ORDER : END_ORDER;
  }

 private:
  // create_raw_problem3 (bb_problem_calculator.cpp:25-87)
  raw_bb_problem* child(const raw_bb_problem& f, int j, int side, bool leaf) const {
    auto* c = new raw_bb_problem(f);
    c->begin_end = side;
    c->limite1 = f.limite1 + (side == BEGIN_ORDER);
    c->limite2 = f.limite2 - (side == END_ORDER);
    if (c->limite2 - c->limite1 == 2) c->limite1++;
    c->depth = (f.depth == size_ - 2 || leaf) ? size_ : f.depth + 1;
    c->job = f.permutation[j];
    const int t = side == BEGIN_ORDER ? f.limite1 + 1 : f.limite2 - 1;
    std::swap(c->permutation[t], c->permutation[j]);
    return c;
  }
  bound_hip* bound_;
  int size_;
  int best_ = 99999;
  problems ns_;
};

// ---- text formats ------------------------------------------------------------

// A node as text, in the field order and spacing of raw_bb_problem's
// operator<< / operator>> (bb_problem_calculator.cpp:101-144): couts[0]
// couts[1] couts_somme, `njobs` permutation entries (the reference writes
// MAX_NBJOBS = 20), limite1 limite2 depth job begin_end, each followed by a
// space, then a newline.
inline std::ostream& write_record(std::ostream& os, const raw_bb_problem& p, int njobs = 20) {
  os << p.couts[0] << ' ' << p.couts[1] << ' ' << p.couts_somme << ' ';
  for (int j = 0; j < njobs; ++j) os << p.permutation[j] << ' ';
  os << p.limite1 << ' ' << p.limite2 << ' ' << p.depth << ' ' << p.job << ' ' << p.begin_end
     << ' ' << '\n';
  return os;
}
inline bool read_record(std::istream& is, raw_bb_problem& p, int njobs = 20) {
  if (njobs < 0 || njobs > MAX_NBJOBS) return false;
  is >> p.couts[0] >> p.couts[1] >> p.couts_somme;
  for (int j = 0; j < njobs; ++j) is >> p.permutation[j];
  is >> p.limite1 >> p.limite2 >> p.depth >> p.job >> p.begin_end;
  return static_cast<bool>(is);
}

// One bab.solutions block (operator<< for solutions, solutions.cpp:88-95):
// the count, a space and a newline, then the records.  solutions::save
// (97-108) appends a block at every improvement.
inline void write_solutions(std::ostream& os, const std::vector<raw_bb_problem>& s,
                            int njobs = 20) {
  os << s.size() << ' ' << '\n';
  for (const auto& r : s) write_record(os, r, njobs);
}
inline void save_solutions(const std::string& path, const std::vector<raw_bb_problem>& s,
                           int njobs = 20) {
  std::ofstream f(path, std::ofstream::app);
  if (!f) throw std::runtime_error("cannot open " + path);
  write_solutions(f, s, njobs);
}
// the first block of a bab.solutions stream (readOptimalSolution 69-86)
inline std::vector<raw_bb_problem> read_solutions(std::istream& is, int njobs = 20) {
  int count = 0;
  std::vector<raw_bb_problem> out;
  if (!(is >> count) || count < 0) return out;
  for (int i = 0; i < count; ++i) {
    raw_bb_problem r{};
    if (!read_record(is, r, njobs)) throw std::runtime_error("truncated solutions block");
    out.push_back(r);
  }
  return out;
}

// One row of the formater's CSV (formater.cpp:94-116):
// "instance;makespan;nodes;time;"
inline std::string formater_line(int instance, int makespan, long long nodes, double seconds) {
  std::ostringstream os;
  os << instance << ';' << makespan << ';' << nodes << ';' << static_cast<float>(seconds) << ';';
  return os.str();
}

// ---- device pool checkpoint -------------------------------------------------
// A text file: "wsbb-pool <n> <incumbent> <count>", the counters, the
// incumbent's permutation (-1s if none), then one record per pool node
// (njobs = n; couts[0] = couts_somme = the node's bound, depth from its
// limits, job and begin_end -1), deepest-first as the stack holds them.
struct pool_snapshot {
  int n = 0;
  int32_t inc = 0;
  std::vector<int64_t> stats;
  std::vector<int32_t> best_perm;
  std::vector<uint8_t> permT;  // [n][count]
  std::vector<int32_t> l1, l2, cost;
  int count() const { return static_cast<int>(l1.size()); }
};

inline void save_pool(std::ostream& os, const pool_snapshot& s) {
  os << "wsbb-pool " << s.n << ' ' << s.inc << ' ' << s.count() << '\n';
  for (int64_t v : s.stats) os << v << ' ';
  os << '\n';
  for (int32_t v : s.best_perm) os << v << ' ';
  os << '\n';
  const int c = s.count();
  for (int k = 0; k < c; ++k) {
    raw_bb_problem r{};
    r.couts[0] = r.couts_somme = s.cost[k];
    for (int i = 0; i < s.n; ++i) r.permutation[i] = s.permT[static_cast<size_t>(i) * c + k];
    r.limite1 = s.l1[k];
    r.limite2 = s.l2[k];
    r.depth = (s.l1[k] + 1) + (s.n - s.l2[k]);
    r.job = r.begin_end = -1;
    write_record(os, r, s.n);
  }
}
inline pool_snapshot load_pool(std::istream& is) {
  pool_snapshot s;
  std::string tag;
  int count = 0;
  is >> tag >> s.n >> s.inc >> count;
  if (!is || tag != "wsbb-pool" || s.n < 2 || s.n > MAX_NBJOBS || count < 0)
    throw std::runtime_error("not a wsbb-pool file");
  s.stats.resize(WSBB_NSTATS);
  for (auto& v : s.stats) is >> v;
  s.best_perm.resize(s.n);
  for (auto& v : s.best_perm) is >> v;
  s.permT.resize(static_cast<size_t>(s.n) * count);
  s.l1.resize(count);
  s.l2.resize(count);
  s.cost.resize(count);
  for (int k = 0; k < count; ++k) {
    raw_bb_problem r{};
    if (!read_record(is, r, s.n)) throw std::runtime_error("truncated wsbb-pool file");
    for (int i = 0; i < s.n; ++i) s.permT[static_cast<size_t>(i) * count + k] = static_cast<uint8_t>(r.permutation[i]);
    s.l1[k] = r.limite1;
    s.l2[k] = r.limite2;
    s.cost[k] = r.couts[0];
  }
  return s;
}

// the last solve's state of a context, as a snapshot
inline pool_snapshot snapshot(wsbb_ctx* ctx) {
  pool_snapshot s;
  int n = 0, m = 0, P = 0, cus = 0, count = 0;
  if (wsbb_info(ctx, &n, &m, &P, &cus) != WSBB_OK) throw std::runtime_error(wsbb_last_error());
  s.n = n;
  s.stats.resize(WSBB_NSTATS);
  s.best_perm.resize(n);
  if (wsbb_search_state(ctx, &s.inc, s.stats.data(), s.best_perm.data(), &count, nullptr,
                        nullptr, nullptr, nullptr, 0) != WSBB_OK)
    throw std::runtime_error(wsbb_last_error());
  s.permT.resize(static_cast<size_t>(n) * count);
  s.l1.resize(count);
  s.l2.resize(count);
  s.cost.resize(count);
  if (wsbb_search_state(ctx, &s.inc, s.stats.data(), s.best_perm.data(), &count, s.permT.data(),
                        s.l1.data(), s.l2.data(), s.cost.data(), count) != WSBB_OK)
    throw std::runtime_error(wsbb_last_error());
  return s;
}

// continue a search from a snapshot; returns wsbb_solve_resume's code
inline int resume(wsbb_ctx* ctx, const pool_snapshot& s, int kind, int pool_size, int max_batches,
                  int64_t budget, int32_t* best, int32_t* best_perm, int64_t* stats) {
  return wsbb_solve_resume(ctx, kind, pool_size, max_batches, budget, s.permT.data(), s.l1.data(),
                           s.l2.data(), s.cost.data(), s.count(), s.inc, s.stats.data(),
                           s.best_perm.data(), best, best_perm, stats);
}

// the root of the reference's tree (create_raw_problem 9-22, add_root)
inline raw_bb_problem* make_root(int size) {
  auto* r = new raw_bb_problem();
  r->limite1 = -1;
  r->limite2 = size;
  r->depth = 0;
  r->job = -1;
  for (int j = 0; j < size; ++j) r->permutation[j] = j;
  return r;
}

}  // namespace wsbb_plugins

#endif  // WSBB_PLUGINS_HPP
