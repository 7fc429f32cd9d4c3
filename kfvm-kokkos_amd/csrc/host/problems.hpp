// Problem registry (initial conditions) of the stage library.
#pragma once

#include <string>
#include <vector>

#include "common.hpp"

namespace kfvm {

bool problem_known(const std::string &name, Eq eq);
// Cell averages of the configured problem, [NV][nz][ny][nx].
void fill_initial_condition(const Config &c, std::vector<double> &U);

}  // namespace kfvm
