// batch.h — many scenarios of one LP that share the constraint matrix (reference
// include/ps26119/batch.h:10-33, src/pdhg/batch.cpp): objective and bounds vary per scenario,
// the matrix, its scaling and ‖Ã‖₂ are computed once, and the scenarios iterate in lockstep on
// the GPU as one sparse-matrix × block-vector product per half step.
#pragma once

#include <vector>

#include "contract.h"

namespace ps26119::amd {

// One Solution per scenario, in order. Per scenario the iterates, restarts, primal weight and
// termination are those of the reference batch solver; with the engine parameter
// hip_cpu_order = 1 the sums follow the reference's order and the results are its bits.
// Options::warm_x / warm_y (original space, original sense; empty = cold) start every scenario
// from that point, projected onto its own bounds, with warm_primal_weight as the first ω.
std::vector<Solution> solve_batch(const Model& base, const std::vector<Scenario>& scenarios, const Options& options);

}  // namespace ps26119::amd
