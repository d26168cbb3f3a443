// presolve.h — the safe reductions the reference applies before every default solve
// (reference src/core/presolve.h:3-21, presolve.cpp:22-209), with primal, dual and Farkas
// postsolve. Repeated until nothing changes:
//   R1 empty row (0 ∈ [rl, ru], else infeasible)      R2 fixed column (substituted out)
//   R3 empty column at its best finite bound           R4 singleton row → column bound
// Same reductions in the same order as the reference, so a presolved model is identical.
#pragma once

#include <string>
#include <vector>

#include "contract.h"

namespace ps26119::amd {

struct PresolveResult {
  enum class Outcome { Reduced, Infeasible, Unchanged } outcome = Outcome::Unchanged;
  Model reduced;
  std::string message;
  int removed_rows = 0, removed_cols = 0, tightened_bounds = 0;
  std::vector<int> row_map, col_map;  // reduced → original
  std::vector<double> fixed_value;    // original column → value (NaN if kept)
  struct Singleton {
    int row, col;
    double a;
    bool set_lower, set_upper;
  };
  std::vector<Singleton> singletons;
};

PresolveResult presolve(const Model& model);
Solution postsolve(const Model& original, const PresolveResult& r, const Solution& reduced);
std::vector<double> postsolve_farkas(const Model& original, const PresolveResult& r, const std::vector<double>& reduced_ray);

}  // namespace ps26119::amd
