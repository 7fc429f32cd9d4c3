// kfvm.ex <file.init> [problem] -- command-line driver with the reference's
// contract (src/main.cpp:23-55): parse the init file, build the solver, run
// Solve().  The reference compiles its problem into kfvm.ex; here the optional
// second argument names the registry problem, as a problem:name line appended
// to the file would, so a reference init file runs unchanged.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>

#include "../../../include/kfvm.h"

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <file.init> [problem]\n", argv[0]);
    return 1;
  }
  std::printf("Setting options from input file: %s\n", argv[1]);
  kfvm_solver *s = nullptr;
  int rc;
  if (argc > 2) {
    std::ifstream f(argv[1]);
    if (!f.is_open()) {
      std::fprintf(stderr, "Error: could not open input file: %s\n", argv[1]);
      return 1;
    }
    std::stringstream ss;
    ss << f.rdbuf() << "\nproblem:name " << argv[2] << "\n";
    rc = kfvm_create_from_text(ss.str().c_str(), &s);
  } else {
    rc = kfvm_create(argv[1], &s);
  }
  if (rc != 0) {
    std::fprintf(stderr, "Error: %s\n", kfvm_last_error());
    return 1;
  }
  int64_t steps = 0;
  if (kfvm_solve(s, &steps) != 0) {
    std::fprintf(stderr, "Error: %s\n", kfvm_last_error());
    kfvm_destroy(s);
    return 1;
  }
  double t = 0, dt = 0;
  int64_t step = 0, rhs = 0, rtol = 0, runph = 0, nweno = 0;
  kfvm_get_time(s, &t, &dt, &step);
  kfvm_get_counters(s, &rhs, &rtol, &runph, &nweno);
  std::printf("Time stepping completed\n  %lld RHS evals and %lld/%lld rejected (accuracy/unphysical) steps\n",
              (long long)rhs, (long long)rtol, (long long)runph);
  std::printf("  final time %.6g after %lld steps\n", t, (long long)steps);
  kfvm_destroy(s);
  return 0;
}
