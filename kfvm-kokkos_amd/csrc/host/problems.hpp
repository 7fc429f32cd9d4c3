// Problem registry (initial conditions) of the stage library.
#pragma once

#include <string>
#include <vector>

#include "common.hpp"

namespace kfvm {

bool problem_known(const std::string &name, Eq eq);
// Comma-separated names of the registry's problems.
std::string problem_names();
// The scheme a reference problem compiles in (dim, nq, equation, Riemann
// solver, RK table, boundary conditions; its Definitions.H and
// ProblemSetup.cpp), applied before the explicit problem: keys of an init
// file, and the defaults of its named user parameters.
void apply_problem_preset(Config &c);
// Named user: keys of the configured problem (UserParameters.H::set);
// unknown names are reported as ignored lines.
void resolve_user_keys(Config &c, std::string *warnings);
// Cell averages of the configured problem, [NV][nz][ny][nx].
void fill_initial_condition(const Config &c, std::vector<double> &U);

}  // namespace kfvm
