// wsbb_plugins.hpp -- C++ face of the engine in the shape of the reference's
// operator plugins, header-only over the C ABI of wsbb_amd.h.
//
// It keeps the interfaces a caller of the reference codes against:
//   struct bound_abstract      (headers/bound_abstract.h:9-19)
//   struct decompose_abstract  (headers/decompose_abstract.h:11-17)
//   class select_abstract      (headers/select_abstract.h:4-7)
//   select_worst               (src/select_worst.cpp:7-12)
//   class problems             (headers/problems.h:13-28, src/problems.cpp)
//   decompose_beginend         (headers/decompose_beginend.h:15-44,
//                               src/decompose_beginend.cpp:20-107)
// so that the reference's composition root wires them unchanged
// (pbab::pbab, src/pbab.cpp:49-53: decompose->set_pbab, decompose->set_bound,
// bound->set_pbab, bound->set_instance), and adds bound_hip, the GPU bound
// behind bound_abstract.  A per-node bornes_calculer is a batch of one;
// decompose_beginend bounds both sides of a parent in one launch when its
// bound is a bound_hip.  gpu_pool drives the device-resident search of a
// bound_hip's context with the live-pool calls.  The reference's text
// formats (node records, bab.solutions blocks, the formater CSV row), its
// binary node record of the work-stealing messages, and a pool checkpoint
// file of the device search are read and written here too.
//
// The node record holds WSBB_MAX_NBJOBS jobs: 20 as in the reference
// (MAX_NBJOBS, headers/tree.h:31), which makes the binary record 184 bytes;
// define it before including this header for larger instances (up to 128).
#ifndef WSBB_PLUGINS_HPP
#define WSBB_PLUGINS_HPP

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <istream>
#include <ostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "wsbb_amd.h"

#ifndef WSBB_MAX_NBJOBS
#define WSBB_MAX_NBJOBS 20
#endif

class pbab;     // the caller's composition root (headers/pbab.h)
class problem;  // the reference's legacy node class (headers/problem.h)

namespace wsbb_plugins {

constexpr int MAX_NBJOBS = WSBB_MAX_NBJOBS;
static_assert(MAX_NBJOBS >= 2 && MAX_NBJOBS <= 128, "WSBB_MAX_NBJOBS must be in [2, 128]");
constexpr int BEGIN_ORDER = 0;  // decompose_beginend.cpp:17-18
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
  pbab* pbb = nullptr;
  virtual void bornes_calculer(int permutation[], int limite1, int limite2, int* couts,
                               int best) = 0;
  virtual void bornes_calculer(raw_bb_problem& p) = 0;
  void set_pbab(pbab* p) { pbb = p; }

  instance_abstract* instance = nullptr;
  virtual void set_instance(instance_abstract* instance) = 0;
  virtual ~bound_abstract() = default;
};

class problems;

// decompose_abstract (headers/decompose_abstract.h:11-17).  Its operator()
// takes the reference's legacy `problem` class, which the device path does
// not use; the reference's own decompose_beginend does not derive from it
// either (decompose_beginend.h:15) and neither does the one below.
struct decompose_abstract {
  pbab* pbb = nullptr;
  virtual problems& operator()(problem& n) const = 0;

  bound_abstract* bound = nullptr;
  virtual void set_bound(bound_abstract* b) = 0;
  virtual ~decompose_abstract() = default;
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
  void set_instance(instance_abstract* inst) override {
    instance = inst;
    std::stringstream& in = *inst->data;
    in.clear();
    in.seekg(0);
    long seed = 0;
    int n = 0, m = 0;
    in >> n >> m >> seed;
    if (!in || n < 2 || n > MAX_NBJOBS || m < 2)
      throw std::runtime_error("bad instance stream or more jobs than WSBB_MAX_NBJOBS");
    std::vector<int32_t> p(static_cast<size_t>(n) * m);
    for (auto& v : p) in >> v;
    if (!in) throw std::runtime_error("bad instance stream");
    set_times(p.data(), n, m);
  }
  void set_times(const int32_t* p_machine_major, int n, int m) {
    if (n > MAX_NBJOBS) throw std::runtime_error("more jobs than WSBB_MAX_NBJOBS");
    wsbb_destroy(ctx_);
    ctx_ = nullptr;
    check(wsbb_create(p_machine_major, n, m, device_, &ctx_));
    n_ = n;
    m_ = m;
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
  int kind() const { return kind_; }
  wsbb_ctx* context() { return ctx_; }

  static void check(int rc) {
    if (rc != WSBB_OK) throw std::runtime_error(std::string("wsbb: ") + wsbb_last_error());
  }

 private:
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

// decompose_beginend (decompose_beginend.h:15-44): children of both sides
// (bounded in one launch when the bound is a bound_hip, else one
// bornes_calculer per child, problems_generate 26-37), the side with more
// children at or above `best` or, on a tie, the larger bound sum below it
// wins (choise 39-50; best stays 99999 as in the reference), ranked; the
// chosen side is returned and owned by the decomposer until the next call.
// The instance size comes from the bound (bound_hip::size) or set_size.
class decompose_beginend {
 public:
  pbab* pbb = nullptr;
  bound_abstract* bound = nullptr;
  int best;
  problems ns;

  explicit decompose_beginend(long unsigned best = 99999) : best(static_cast<int>(best)) {}
  ~decompose_beginend() { ns.release(); }
  decompose_beginend(const decompose_beginend&) = delete;
  decompose_beginend& operator=(const decompose_beginend&) = delete;

  void set_pbab(pbab* p) { pbb = p; }
  void set_bound(bound_abstract* b) { bound = b; }
  void set_size(int n) { size_ = n; }

  int choise(problems& debut, problems& fin) const {
    const int c0 = debut.coupe(best), c1 = fin.coupe(best);
    if (c0 > c1) return BEGIN_ORDER;
    if (c1 > c0) return END_ORDER;
    return debut.sum(best) > fin.sum(best) ? BEGIN_ORDER : END_ORDER;
  }

  // the children of one side, unbounded (problems_generate 26-37)
  void problems_generate(problems& out, raw_bb_problem& n, int begin_end) const {
    out.clear();
    for (int j = n.limite1 + 1; j < n.limite2; ++j) out.push_back(child(n, j, begin_end, false));
  }
  // the two leaves of a parent with two free jobs (leaves_generate 70-89)
  void leaves_generate(raw_bb_problem& n, problems& out) const {
    out.clear();
    for (int t = 1; t <= 2; ++t) out.push_back(child(n, n.limite1 + t, BEGIN_ORDER, true));
  }

  // decompose_beginend::operator() (decompose_beginend.cpp:91-107)
  problems& operator()(raw_bb_problem& n) {
    if (!bound) throw std::runtime_error("decompose_beginend: set_bound first");
    const int sz = size();
    ns.release();
    problems debut, fin;
    if (n.depth == sz - 2) {
      leaves_generate(n, debut);
    } else {
      problems_generate(debut, n, BEGIN_ORDER);
      problems_generate(fin, n, END_ORDER);
    }
    std::vector<raw_bb_problem*> all(debut.begin(), debut.end());
    all.insert(all.end(), fin.begin(), fin.end());
    if (auto* hip = dynamic_cast<bound_hip*>(bound)) {
      hip->bornes_calculer_batch(all);
    } else {
      for (raw_bb_problem* c : all) bound->bornes_calculer(*c);
    }
    const bool begin_side = choise(debut, fin) == BEGIN_ORDER;
    problems& keep = begin_side ? debut : fin;
    problems& drop = begin_side ? fin : debut;
    drop.release();
    ns.assign(keep.begin(), keep.end());
    keep.clear();
    ns.ranks(n.depth == sz - 2);
    return ns;
  }

  int size() const {
    if (size_ > 0) return size_;
    if (auto* hip = dynamic_cast<const bound_hip*>(bound)) return hip->size();
    throw std::runtime_error("decompose_beginend: instance size unknown (set_size)");
  }

 private:
  // create_raw_problem3 (bb_problem_calculator.cpp:25-87)
  raw_bb_problem* child(const raw_bb_problem& f, int j, int side, bool leaf) const {
    const int sz = size();
    auto* c = new raw_bb_problem(f);
    c->begin_end = side;
    c->limite1 = f.limite1 + (side == BEGIN_ORDER);
    c->limite2 = f.limite2 - (side == END_ORDER);
    if (c->limite2 - c->limite1 == 2) c->limite1++;
    c->depth = (f.depth == sz - 2 || leaf) ? sz : f.depth + 1;
    c->job = f.permutation[j];
    const int t = side == BEGIN_ORDER ? f.limite1 + 1 : f.limite2 - 1;
    std::swap(c->permutation[t], c->permutation[j]);
    return c;
  }
  int size_ = 0;
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
  // the record fields of write_record, with n permutation entries
  const int c = s.count();
  for (int k = 0; k < c; ++k) {
    os << s.cost[k] << ' ' << 0 << ' ' << s.cost[k] << ' ';
    for (int i = 0; i < s.n; ++i) os << int(s.permT[static_cast<size_t>(i) * c + k]) << ' ';
    os << s.l1[k] << ' ' << s.l2[k] << ' ' << (s.l1[k] + 1) + (s.n - s.l2[k]) << ' ' << -1 << ' '
       << -1 << ' ' << '\n';
  }
}
inline pool_snapshot load_pool(std::istream& is) {
  pool_snapshot s;
  std::string tag;
  int count = 0;
  is >> tag >> s.n >> s.inc >> count;
  if (!is || tag != "wsbb-pool" || s.n < 2 || s.n > 128 || count < 0)
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
    int c0 = 0, c1 = 0, sum = 0, depth = 0, job = 0, side = 0;
    is >> c0 >> c1 >> sum;
    for (int i = 0; i < s.n; ++i) {
      int v = 0;
      is >> v;
      s.permT[static_cast<size_t>(i) * count + k] = static_cast<uint8_t>(v);
    }
    is >> s.l1[k] >> s.l2[k] >> depth >> job >> side;
    s.cost[k] = c0;
  }
  if (!is) throw std::runtime_error("truncated wsbb-pool file");
  return s;
}

// ---- binary node record of the work-stealing messages ------------------------
// The reference sends nodes to other peers as consecutive records of int32
// fields in host byte order: permutation[MAX_NBJOBS], limite1, limite2,
// couts_somme, depth, begin_end, ranks[MAX_NBJOBS + 1] (the serve side,
// src/work_stealing/node.cpp:524-556), and the receiver rebuilds a node
// from each with couts[0] = couts_somme, couts[1] = 0 (node.cpp:1240-1283).
// With MAX_NBJOBS = 20 a record is 184 bytes.
constexpr size_t RECORD_BYTES = sizeof(int32_t) * (2 * MAX_NBJOBS + 6);

inline void append_record(std::vector<char>& out, const raw_bb_problem& p) {
  const size_t at = out.size();
  out.resize(at + RECORD_BYTES);
  char* w = out.data() + at;
  auto put = [&w](const int* v, int cnt) {
    std::memcpy(w, v, sizeof(int) * cnt);
    w += sizeof(int) * cnt;
  };
  put(p.permutation, MAX_NBJOBS);
  put(&p.limite1, 1);
  put(&p.limite2, 1);
  put(&p.couts_somme, 1);
  put(&p.depth, 1);
  put(&p.begin_end, 1);
  put(p.ranks, MAX_NBJOBS + 1);
}
inline std::vector<char> serialize_nodes(const std::vector<const raw_bb_problem*>& nodes) {
  std::vector<char> out;
  out.reserve(nodes.size() * RECORD_BYTES);
  for (const raw_bb_problem* p : nodes) append_record(out, *p);
  return out;
}
// the records of a message; a trailing partial record is an error
inline std::vector<raw_bb_problem> deserialize_nodes(const char* data, size_t bytes) {
  if (bytes % RECORD_BYTES != 0) throw std::runtime_error("partial node record");
  std::vector<raw_bb_problem> out(bytes / RECORD_BYTES);
  for (raw_bb_problem& p : out) {
    p = raw_bb_problem{};
    auto get = [&data](int* v, int cnt) {
      std::memcpy(v, data, sizeof(int) * cnt);
      data += sizeof(int) * cnt;
    };
    get(p.permutation, MAX_NBJOBS);
    get(&p.limite1, 1);
    get(&p.limite2, 1);
    get(&p.couts_somme, 1);
    get(&p.depth, 1);
    get(&p.begin_end, 1);
    get(p.ranks, MAX_NBJOBS + 1);
    p.couts[0] = p.couts_somme;
    p.couts[1] = 0;
    p.job = -1;
  }
  return out;
}

// A device pool as binary records: the rank strings give the reference's
// priority queue (tree.cpp:37-46) the device's pop order (ranks[0] = how
// many nodes are above it on the stack, the rest 0), depth follows from the
// limits, job and begin_end are -1.
inline std::vector<char> records_from_pool(const pool_snapshot& s) {
  if (s.n > MAX_NBJOBS) throw std::runtime_error("more jobs than WSBB_MAX_NBJOBS");
  const int c = s.count();
  std::vector<char> out;
  out.reserve(static_cast<size_t>(c) * RECORD_BYTES);
  for (int k = 0; k < c; ++k) {
    raw_bb_problem r{};
    for (int i = 0; i < s.n; ++i) r.permutation[i] = s.permT[static_cast<size_t>(i) * c + k];
    r.limite1 = s.l1[k];
    r.limite2 = s.l2[k];
    r.couts[0] = r.couts_somme = s.cost[k];
    r.depth = (s.l1[k] + 1) + (s.n - s.l2[k]);
    r.job = r.begin_end = -1;
    r.ranks[0] = c - 1 - k;
    append_record(out, r);
  }
  return out;
}
// Records (any order) as pool nodes, bottom-first: the lexicographically
// smallest rank string, which the reference's queue pops first, on top.
inline pool_snapshot pool_from_records(const std::vector<raw_bb_problem>& recs, int n) {
  if (n < 2 || n > MAX_NBJOBS) throw std::runtime_error("bad job count for node records");
  std::vector<const raw_bb_problem*> order;
  for (const auto& r : recs) order.push_back(&r);
  std::stable_sort(order.begin(), order.end(), [](const raw_bb_problem* a, const raw_bb_problem* b) {
    return std::lexicographical_compare(b->ranks, b->ranks + MAX_NBJOBS + 1, a->ranks,
                                        a->ranks + MAX_NBJOBS + 1);
  });
  pool_snapshot s;
  s.n = n;
  const int c = static_cast<int>(order.size());
  s.permT.resize(static_cast<size_t>(n) * c);
  s.l1.resize(c);
  s.l2.resize(c);
  s.cost.resize(c);
  for (int k = 0; k < c; ++k) {
    for (int i = 0; i < n; ++i)
      s.permT[static_cast<size_t>(i) * c + k] = static_cast<uint8_t>(order[k]->permutation[i]);
    s.l1[k] = order[k]->limite1;
    s.l2[k] = order[k]->limite2;
    s.cost[k] = order[k]->couts_somme;
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

// The device search of a bound_hip's context driven from the host with the
// live-pool calls (wsbb_pool_*): nodes in and out as pool snapshots or node
// records, the incumbent, rounds.  The host thread that feeds a GPU in the
// reference (protocol.cpp:937-1033) becomes a loop over these calls.
class gpu_pool {
 public:
  gpu_pool(bound_hip& bound, int pool_size = 4096, int max_batches = 1)
      : ctx_(bound.context()), n_(bound.size()) {
    bound_hip::check(wsbb_pool_open(ctx_, bound.kind(), pool_size, max_batches));
  }
  void seed_root() { bound_hip::check(wsbb_pool_seed_root(ctx_)); }
  void push(const pool_snapshot& s) {
    bound_hip::check(wsbb_pool_push(ctx_, s.permT.data(), s.l1.data(), s.l2.data(),
                                    s.cost.data(), s.count()));
  }
  void push(const std::vector<raw_bb_problem>& recs) { push(pool_from_records(recs, n_)); }
  pool_snapshot pop(int max, bool from_bottom) {
    pool_snapshot s;
    s.n = n_;
    s.permT.resize(static_cast<size_t>(n_) * std::max(max, 1));
    s.l1.resize(std::max(max, 1));
    s.l2.resize(std::max(max, 1));
    s.cost.resize(std::max(max, 1));
    int cnt = 0;
    bound_hip::check(wsbb_pool_pop(ctx_, max, from_bottom ? 1 : 0, s.permT.data(), s.l1.data(),
                                   s.l2.data(), s.cost.data(), &cnt));
    s.permT.resize(static_cast<size_t>(n_) * cnt);  // rows of cnt entries
    s.l1.resize(cnt);
    s.l2.resize(cnt);
    s.cost.resize(cnt);
    return s;
  }
  int size() {
    int sz = 0;
    bound_hip::check(wsbb_pool_size(ctx_, &sz));
    return sz;
  }
  int incumbent(std::vector<int32_t>* perm = nullptr) {
    int32_t inc = 0;
    if (perm) perm->resize(n_);
    bound_hip::check(wsbb_incumbent_get(ctx_, &inc, perm ? perm->data() : nullptr));
    return inc;
  }
  void set_incumbent(int32_t inc, const int32_t* perm = nullptr) {
    bound_hip::check(wsbb_incumbent_set(ctx_, inc, perm));
  }
  // runs up to `rounds` rounds (<= 0: until empty); returns the pool size
  int run(int rounds = 0, int64_t budget = 0, int64_t* stats = nullptr) {
    int sz = 0;
    const int rc = wsbb_pool_run(ctx_, rounds, budget, &sz, stats);
    if (rc != WSBB_OK && rc != WSBB_EBUDGET) bound_hip::check(rc);
    return sz;
  }

 private:
  wsbb_ctx* ctx_;
  int n_;
};

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
