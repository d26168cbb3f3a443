// presolve.h — the safe reductions every default solve applies before the engine runs, and the
// way back (reference src/core/presolve.h:3-21 and presolve.cpp:22-209 define the reductions;
// the rounds, the undo log and the lifting here are this library's own structure):
//   an empty row is dropped (infeasible when 0 is outside its range),
//   a row with one entry becomes a bound of that column,
//   a fixed column, or one with no entries left, is substituted out,
// in rounds until nothing changes. The operations and their order are the reference's, so the
// reduced model is the same bits.
#pragma once

#include <string>
#include <vector>

#include "contract.h"

namespace ps26119::amd {

// What was done, in order; postsolve replays it backwards.
struct ReductionStep {
  enum class Kind : unsigned char { DropEmptyRow, RowToBound, SubstituteColumn } kind;
  int row = -1, col = -1;
  double coef = 0.0;   // RowToBound: the row's only coefficient
  double value = 0.0;  // SubstituteColumn: the value the column was fixed at
  bool sets_lower = false, sets_upper = false;  // RowToBound: which column bounds the row set
};

struct Presolved {
  enum class Outcome { NothingToDo, Reduced, ProvedInfeasible } outcome = Outcome::NothingToDo;
  Model reduced;
  std::string reason;               // ProvedInfeasible: the row that proves it
  int rows_dropped = 0, cols_dropped = 0, bounds_tightened = 0;
  std::vector<int> kept_rows, kept_cols;  // reduced index → original index
  std::vector<ReductionStep> log;
};

Presolved presolve(const Model& model);

// The reduced model's solution on the original model (x, y, z, activities, objective). Vectors
// the reduced solution lacks stay empty.
Solution postsolve(const Model& original, const Presolved& p, const Solution& reduced);

// A dual ray (Farkas vector) of the reduced model on the original rows, or {} for a wrong size.
std::vector<double> lift_dual_ray(const Model& original, const Presolved& p, const std::vector<double>& reduced_ray);

}  // namespace ps26119::amd
