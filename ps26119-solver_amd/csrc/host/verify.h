// verify.h — the host-side verification stage every answer of the GPU path passes through:
//   check_solution ........ optimality proof of (x, y) on the ORIGINAL model with the
//                           verifier tolerances (reference src/core/solution_checker.cpp:61-133)
//   certified_dual_bound .. rounding-proof Lagrangian bound from any y (Neumaier–Shcherbina;
//                           reference src/core/safe_bound.cpp:16-195, implied_bounds.cpp:11-111)
//   certificate checks .... Farkas ray / primal ray verification
//                           (reference src/core/certificates.cpp:13-135)
// Results are identical to the reference's; the O(nnz) loops run in parallel where each output
// is written by one task and every sequential sum keeps the reference's order.
#pragma once

#include <atomic>
#include <string>
#include <thread>
#include <vector>

#include "contract.h"

namespace ps26119::amd {

struct CheckReport {
  double max_primal_violation = 0, max_dual_violation = 0, max_complementarity = 0;
  double primal_objective = 0, dual_objective = 0, relative_gap = 0;
  bool primal_ok = false, dual_ok = false, gap_ok = false;
  bool passed() const { return primal_ok && dual_ok && gap_ok; }
  std::string summary() const;
};

CheckReport check_solution(const Model& lp, const std::vector<double>& x, const std::vector<double>& row_dual);

struct SafeBound {
  bool finite = false;
  double bound = 0.0;
  int unbounded_row_terms = 0, unbounded_col_terms = 0;
};

// The column bounds the rows imply (a function of the model alone, sequential by definition),
// computed on a thread of its own from the moment a solve starts, so that the certified bound
// after the solve finds them ready instead of adding their time to the solve's. The model must
// outlive the job; destroying the job stops the thread early if it is still running.
class ImpliedBoundsJob {
 public:
  explicit ImpliedBoundsJob(const Model& model);
  ~ImpliedBoundsJob();
  ImpliedBoundsJob(const ImpliedBoundsJob&) = delete;
  ImpliedBoundsJob& operator=(const ImpliedBoundsJob&) = delete;
  const Model& model() const { return model_; }
  // Waits for the bounds; false when they could not be computed (the caller then computes them).
  bool wait(const std::vector<double>** lo, const std::vector<double>** up);

 private:
  const Model& model_;
  std::vector<double> lo_, up_;
  std::atomic<bool> cancel_{false};
  bool ok_ = false;
  std::thread thread_;
};

// `implied`, when given, must belong to `model`: its bounds are used instead of computing them.
SafeBound certified_dual_bound(const Model& model, const std::vector<double>& y, bool use_implied_bounds = true,
                               ImpliedBoundsJob* implied = nullptr);

struct CertificateCheck {
  bool present = false, passed = false, rigorous = false;
  double measure = 0.0, violation = 0.0;
  std::string detail;
};
CertificateCheck check_infeasibility_certificate(const Model& model, const std::vector<double>& dual_ray);
CertificateCheck check_unboundedness_certificate(const Model& model, const std::vector<double>& x,
                                                 const std::vector<double>& primal_ray);

// Row activity A x summed per row in increasing column order (the same bits as a CSC scatter).
std::vector<double> activity(const Model& m, const std::vector<double>& x);

}  // namespace ps26119::amd
