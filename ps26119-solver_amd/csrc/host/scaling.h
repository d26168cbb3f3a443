// scaling.h — diagonal preconditioning of the LP for the first-order iteration.
//
// Same mathematics and the same floating-point operations as the reference's
// make_scaled_problem (reference src/pdhg/scaling.cpp:37-142; scaling.h:17-22 for the formulas):
// optional geometric-mean equilibration (12 sweeps, only when the entries span ≥ 10^4.5),
// Ruiz ∞-norm equilibration ×10, Pock–Chambolle α = 1, then the bound / objective scalars
//   Ã = R A C,  c̃ = os·C(sense·c),  r̃ = bs·R r,  l̃ = bs·l / C,
// so the scaled data are bit-identical to the reference's. The loops run in parallel over
// rows and over columns; column norms are taken through the transpose permutation in
// increasing-row order, which is the order the reference accumulates them in.
#pragma once

#include <vector>

#include "la.h"

namespace ps26119::amd {

struct ScalingOptions {  // reference src/pdhg/scaling.h:32-46
  int geometric_mean_iterations = 12;
  double geometric_mean_min_log10_range = 4.5;
  int ruiz_iterations = 10;
  bool pock_chambolle = true;
  bool bound_objective_rescaling = true;
};

struct ScaledProblem {
  int m = 0, n = 0, sense = 1;
  Csr A;                          // Ã (m×n), columns sorted in each row
  Csr At;                         // Ãᵀ (n×m), rows sorted in each column
  std::vector<double> a_orig;     // original values in Ã's layout (= CSR of A)
  std::vector<double> c, col_lower, col_upper, row_lower, row_upper;
  std::vector<double> row_scale, col_scale;
  double bound_scale = 1.0, obj_scale = 1.0, log10_range = 0.0;
  bool geometric_mean_applied = false;
  const Model* original = nullptr;

  // x = clip(C x̃ / bs, l, u) and y_min = R ỹ / os (reference scaling.cpp:144-155).
  void unscale_primal(const std::vector<double>& xs, std::vector<double>& x) const;
  void unscale_dual(const std::vector<double>& ys, std::vector<double>& y_min) const;
};

ScaledProblem make_scaled_problem(const Model& model, const ScalingOptions& opt);

// The same problem with unit factors: Ã = A, Ãᵀ's values in Ãᵀ's order, log10_range and
// geometric_mean_applied decided, vectors unscaled. The device path loads it, runs the sweeps on
// the device (psh_scale) and then calls finish_scaled_problem with the factors it returns.
ScaledProblem make_unscaled_problem(const Model& model, const ScalingOptions& opt);

// c̃, l̃, ũ, r̃l, r̃u and the bound / objective scalars from sp.row_scale and sp.col_scale (the tail
// of make_scaled_problem, shared so that both paths do the same arithmetic).
void finish_scaled_problem(ScaledProblem& sp, const Model& model, const ScalingOptions& opt);

}  // namespace ps26119::amd
