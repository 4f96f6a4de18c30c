// Round trips of the text formats in wsbb_plugins.hpp: node records,
// bab.solutions blocks, the formater CSV row and the pool checkpoint file.
// Prints the written texts (compared with expected strings by the test) and
// "ok" when every round trip reproduces its input.
#include <cstdio>
#include <iostream>
#include <sstream>

#include "wsbb_plugins.hpp"

using namespace wsbb_plugins;

static bool same(const raw_bb_problem& a, const raw_bb_problem& b, int njobs) {
  if (a.couts[0] != b.couts[0] || a.couts[1] != b.couts[1] || a.couts_somme != b.couts_somme)
    return false;
  for (int j = 0; j < njobs; ++j)
    if (a.permutation[j] != b.permutation[j]) return false;
  return a.limite1 == b.limite1 && a.limite2 == b.limite2 && a.depth == b.depth &&
         a.job == b.job && a.begin_end == b.begin_end;
}

int main() {
  bool ok = true;
  raw_bb_problem r{};
  r.couts[0] = 1278;
  r.couts[1] = 0;
  r.couts_somme = 1278;
  for (int j = 0; j < 20; ++j) r.permutation[j] = (j * 7) % 20;
  r.limite1 = 3;
  r.limite2 = 17;
  r.depth = 7;
  r.job = 14;
  r.begin_end = 1;
  std::ostringstream os;
  write_record(os, r);
  std::cout << os.str();
  raw_bb_problem back{};
  std::istringstream is(os.str());
  ok = ok && read_record(is, back) && same(r, back, 20);

  raw_bb_problem r2 = r;
  r2.couts[0] = r2.couts_somme = 1290;
  r2.job = 3;
  std::ostringstream sol;
  write_solutions(sol, {r, r2});
  std::cout << sol.str();
  std::istringstream sin(sol.str() + sol.str());  // two appended blocks
  auto got = read_solutions(sin);
  ok = ok && got.size() == 2 && same(got[0], r, 20) && same(got[1], r2, 20);

  std::cout << formater_line(11, 1582, 4295624, 0.5) << "\n";

  pool_snapshot s;
  s.n = 5;
  s.inc = 321;
  s.stats.assign(WSBB_NSTATS, 0);
  for (int i = 0; i < WSBB_NSTATS; ++i) s.stats[i] = 1000 + i;
  s.best_perm = {4, 3, 2, 1, 0};
  const int c = 3;
  s.l1 = {-1, 0, 1};
  s.l2 = {4, 5, 3};
  s.cost = {300, 310, 320};
  s.permT.resize(5 * c);
  for (int i = 0; i < 5; ++i)
    for (int k = 0; k < c; ++k) s.permT[i * c + k] = static_cast<uint8_t>((i + k) % 5);
  std::ostringstream ps;
  save_pool(ps, s);
  std::istringstream pin(ps.str());
  pool_snapshot t = load_pool(pin);
  ok = ok && t.n == s.n && t.inc == s.inc && t.stats == s.stats && t.best_perm == s.best_perm &&
       t.permT == s.permT && t.l1 == s.l1 && t.l2 == s.l2 && t.cost == s.cost;

  // binary node records: the record's size, a round trip, and its bytes in
  // hex for comparison with the Python codec
  for (int i = 0; i <= MAX_NBJOBS; ++i) r.ranks[i] = (i * 3) % 7;
  std::vector<char> bin = serialize_nodes({&r, &r2});
  ok = ok && RECORD_BYTES == 184 && bin.size() == 2 * RECORD_BYTES;
  auto recs = deserialize_nodes(bin.data(), bin.size());
  ok = ok && recs.size() == 2 && recs[0].couts[0] == r.couts_somme && recs[0].couts[1] == 0;
  for (int i = 0; ok && i < 20; ++i) ok = recs[0].permutation[i] == r.permutation[i];
  for (int i = 0; ok && i <= MAX_NBJOBS; ++i) ok = recs[0].ranks[i] == r.ranks[i];
  ok = ok && recs[1].limite1 == r2.limite1 && recs[1].depth == r2.depth &&
       recs[1].begin_end == r2.begin_end && recs[1].couts_somme == r2.couts_somme;
  bool threw = false;
  try {
    deserialize_nodes(bin.data(), bin.size() - 4);
  } catch (const std::runtime_error&) {
    threw = true;
  }
  ok = ok && threw;
  // a pool as records and back: same nodes, same order
  pool_snapshot back_pool = pool_from_records(
      deserialize_nodes(records_from_pool(s).data(), s.count() * RECORD_BYTES), s.n);
  ok = ok && back_pool.permT == s.permT && back_pool.l1 == s.l1 && back_pool.l2 == s.l2 &&
       back_pool.cost == s.cost;
  for (size_t i = 0; i < RECORD_BYTES; ++i) std::printf("%02x", static_cast<unsigned char>(bin[i]));
  std::printf("\n");
  std::cout << (ok ? "ok" : "MISMATCH") << "\n";
  return ok ? 0 : 1;
}
