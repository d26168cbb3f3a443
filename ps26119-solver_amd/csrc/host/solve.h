// solve.h — the library entry point of the MI355X backend (reference include/ps26119/solve.h:13).
#pragma once

#include "contract.h"

namespace ps26119::amd {

// Validates, presolves, runs r²HPDHG on the GPU, postsolves and gates the answer through the
// in-process check on the original model. Never throws.
Solution solve(const Model& model, const Options& options);

// The verification gate (reference src/core/solve.cpp:119-185, LP part).
void gate(const Model& model, const Options& options, Solution& s);

}  // namespace ps26119::amd
