// Host-side flowshop instance model and the once-per-instance precompute
// that the bound kernels consume.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace wsbb {

// Offset that keeps the biased machine-a clock (t0 - kBias) negative and the
// 24-bit multiply-add operands in range; every bound value must stay below it.
constexpr int kBias = 1 << 22;
// The same offset for the two-node int16 form, used when every bound value
// of the instance stays below it.
constexpr int kBias16 = 1 << 14;
// The int32 scan keeps its clocks scaled by 32 (kScaleShift bits) with the
// job's bit offset in the low bits of a 24-bit multiply-add operand: every
// step constant and bound value must stay below this limit.
constexpr int kScaleShift = 5;
constexpr int kPathLimit = 1 << (24 - kScaleShift);
constexpr int kMaxJobs = 128;     // 4 x 32-bit job masks per node
constexpr int kMaxMachines = 24;  // register arrays for heads/tails

// Taillard's generator as the reference writes it
// (src/instance_flowshop.cpp:74-89): Park-Miller minimal standard with
// Schrage's factorisation, scaled through a float quotient.
long taillard_unif(long* seed, long low, long high);
bool taillard_dims(int id, int* n, int* m, long* seed);
void taillard_generate(int n, int m, long seed, int32_t* p);

// splitmix64 streams used for synthetic node batches and job-correlated
// instances (p[k][j] = a_j + U[0,20], a_j ~ U[1,80]).
uint64_t splitmix64(uint64_t* state);
void correlated_generate(int n, int m, uint64_t seed, int32_t* p);
// machine-correlated instance: p[k][j] = b_k + U[0,20], b_k ~ U[1,80]
void machine_correlated_generate(int n, int m, uint64_t seed, int32_t* p);
void random_nodes(int n, int N, uint64_t seed0, uint8_t* permT, int32_t* l1,
                  int32_t* l2);

struct Instance {
  int n = 0, m = 0, P = 0;
  std::vector<int32_t> p;        // [m][n] processing times, machine-major
  std::vector<int32_t> pair_a;   // [P] first machine of pair s
  std::vector<int32_t> pair_b;   // [P] second machine of pair s
  std::vector<int32_t> lag;      // [P][n] sum of p over machines strictly between
  std::vector<int32_t> johnson;  // [P][n] Johnson job order of pair s
  std::vector<int32_t> R0;       // [m] least release date of machine k
  std::vector<int32_t> Q0;       // [m] least tail after machine k
  std::vector<int32_t> ptot;     // [m] sum_j p[k][j]
  int path_bound = 0;  // >= every bound/makespan value of any node

  // Builds every table from p (machine-major).  Returns an empty string or
  // the reason the instance is rejected.
  std::string build(const int32_t* p_mm, int n, int m);
};

}  // namespace wsbb
