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
// vector in one launch, which is how decompose_beginend uses it.
#ifndef WSBB_PLUGINS_HPP
#define WSBB_PLUGINS_HPP

#include <algorithm>
#include <cstdint>
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
