// kfvm.ex <file.init> -- command-line driver with the reference's contract
// (src/main.cpp:23-55): parse the init file, build the solver, run Solve().
#include <cstdio>
#include <cstdlib>

#include "../../../include/kfvm.h"

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <file.init>\n", argv[0]);
    return 1;
  }
  std::printf("Setting options from input file: %s\n", argv[1]);
  kfvm_solver *s = nullptr;
  if (kfvm_create(argv[1], &s) != 0) {
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
