// TEST INFRASTRUCTURE ONLY.  Driver linked against the reference's own
// src/instance_flowshop.cpp (compiled from /root/reference by oracle/Makefile,
// output in oracle/_ref/).  It prints, for each requested Taillard id, the
// reference's instance stream "N M seed p[0][0..N-1] p[1][..] ..." exactly as
// instance_flowshop::generate_instance writes it (instance_flowshop.cpp:84-89),
// and, for "unif <seed> <count>", the raw unif(seed, 1, 99) sequence.
#include <cstdlib>
#include <iostream>
#include <sstream>
#include <string>
using namespace std;
#include "instance_abstract.h"
#include "instance_flowshop.h"

int main(int argc, char **argv) {
  if (argc >= 4 && string(argv[1]) == "unif") {
    instance_flowshop inst(1);
    long seed = atol(argv[2]);
    int count = atoi(argv[3]);
    for (int i = 0; i < count; i++) cout << inst.unif(&seed, 1, 99) << ' ';
    cout << '\n';
    return 0;
  }
  for (int a = 1; a < argc; a++) {
    int id = atoi(argv[a]);
    instance_flowshop inst(id);
    cout << id << ' ' << inst.data->str() << '\n';
  }
  return 0;
}
