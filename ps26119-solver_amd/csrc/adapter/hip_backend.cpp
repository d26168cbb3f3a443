// hip_backend.cpp — the MI355X backend behind the reference's own engine interface,
// ps26119::pdhg::Backend (reference src/pdhg/backend.h:36-69).
//
// This file is compiled against the REFERENCE's headers (include/ and src/ of the reference
// tree), not this repository's host code: it is what a maintainer of the reference adds
// next to src/pdhg/cpu_backend.cpp to run the reference's r²HPDHG and PDLP drivers
// (src/pdhg/r2hpdhg.cpp, src/pdhg/pdlp.cpp) on an MI355X. It speaks to the GPU only
// through the psh_* C-ABI (include/ps26119_hip.h), so the reference needs no HIP headers and
// links libps26119_hip.so. The factory the reference's make_backend (src/pdhg/cpu_backend.cpp:
// 293-310) would call is make_hip_backend() below.
//
// Every reduction is asked for in PSH_RED_CPU order, so each operation returns the reference
// CPU backend's bits (src/pdhg/cpu_backend.cpp:53-221) and the reference's drivers follow the
// same trajectory on either backend.
#include <cmath>
#include <cstdint>
#include <limits>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "pdhg/backend.h"
#include "ps26119_hip.h"

namespace ps26119::pdhg {

namespace {

// int64 CSR row pointers (reference src/la/csr.h:21) as the int32 ones psh_load takes.
std::vector<int> narrow_ptr(const std::vector<std::int64_t>& ptr) {
  std::vector<int> out(ptr.size());
  for (std::size_t i = 0; i < ptr.size(); ++i) {
    if (ptr[i] > std::numeric_limits<int>::max()) throw std::length_error("hip backend: more than 2^31-1 nonzeros");
    out[i] = static_cast<int>(ptr[i]);
  }
  return out;
}

class HipBackend final : public Backend {
 public:
  HipBackend() {
    const int rc = psh_create(0, &ctx_);
    if (rc != PSH_OK) {
      const std::string why = ctx_ ? psh_error(ctx_) : "psh_create failed";
      psh_destroy(ctx_);
      ctx_ = nullptr;
      throw std::runtime_error("no usable MI355X: " + why);
    }
  }
  ~HipBackend() override { psh_destroy(ctx_); }

  std::string name() const override { return "hip"; }

  // The scaled problem (Ã and Ãᵀ in CSR, scaled vectors, factors) plus the original matrix
  // and bounds kkt() evaluates on. Ã and A share a structure (scaling rescales values in
  // place, reference src/pdhg/scaling.cpp:43-45); so do Ãᵀ and the model's CSC read as Aᵀ.
  void setup(const ScaledProblem& sp) override {
    if (!sp.original) throw std::invalid_argument("hip backend: the scaled problem has no original model");
    const Model& M = *sp.original;
    if (sp.A_orig.val.size() != sp.A.val.size() || sp.At_orig.row_ptr != sp.At.row_ptr ||
        sp.A_orig.row_ptr != sp.A.row_ptr)
      throw std::invalid_argument("hip backend: the original and scaled matrices differ in structure");
    m_ = sp.m;
    n_ = sp.n;
    const std::vector<int> a_ptr = narrow_ptr(sp.A.row_ptr), at_ptr = narrow_ptr(sp.At.row_ptr);
    std::vector<double> c_min(M.obj.size());
    for (std::size_t j = 0; j < c_min.size(); ++j) c_min[j] = M.sense * M.obj[j];
    psh_problem p{};
    p.m = sp.m;
    p.n = sp.n;
    p.nnz = sp.A.nnz();
    p.a_ptr = a_ptr.data();
    p.a_col = sp.A.col.data();
    p.a_val = sp.A.val.data();
    p.a_val_orig = sp.A_orig.val.data();
    p.at_ptr = at_ptr.data();
    p.at_col = sp.At.col.data();
    p.at_val = sp.At.val.data();
    p.ato_col = sp.At_orig.col.data();
    p.ato_val = sp.At_orig.val.data();
    p.c = sp.c.data();
    p.l = sp.col_lower.data();
    p.u = sp.col_upper.data();
    p.rl = sp.row_lower.data();
    p.ru = sp.row_upper.data();
    p.row_scale = sp.row_scale.data();
    p.col_scale = sp.col_scale.data();
    p.bound_scale = sp.bound_scale;
    p.obj_scale = sp.obj_scale;
    p.cl_o = M.col_lower.data();
    p.cu_o = M.col_upper.data();
    p.rl_o = M.row_lower.data();
    p.ru_o = M.row_upper.data();
    p.c_min = c_min.data();
    ok(psh_load(ctx_, &p));  // copies everything to the device before it returns
  }

  void set_precision(Precision p) override { ok(psh_set_fp32(ctx_, p == Precision::Mixed ? 1 : 0)); }
  Precision precision() const override { return psh_is_fp32(ctx_) ? Precision::Mixed : Precision::Fp64; }

  // Handles are dealt out in creation order over the device's six primal and six dual slots
  // (the reference drivers create six of each, src/pdhg/r2hpdhg.cpp:26-29, pdlp.cpp:24-27).
  int create(Space s) override {
    int& used = s == Space::Primal ? primal_used_ : dual_used_;
    if (used == kSlotsPerSpace) throw std::runtime_error("hip backend: more than 6 vectors in one space");
    slot_.push_back((s == Space::Primal ? PSH_X : PSH_Y) + used++);
    dual_.push_back(s == Space::Dual);
    return static_cast<int>(slot_.size()) - 1;
  }
  void upload(int v, const std::vector<double>& src) override {
    if (src.size() != size_of(v)) throw std::invalid_argument("hip backend: upload of a vector of the wrong size");
    ok(psh_upload(ctx_, slot(v), src.data()));
  }
  void download(int v, std::vector<double>& dst) override {
    dst.resize(size_of(v));
    ok(psh_download(ctx_, slot(v), dst.data()));
  }
  void fill(int v, double a) override { ok(psh_fill(ctx_, slot(v), a)); }
  void copy(int dst, int src) override { ok(psh_copy(ctx_, slot(dst), slot(src))); }
  void axpby(double a, int x, double b, int y) override { ok(psh_axpby(ctx_, a, slot(x), b, slot(y))); }
  void spmv(int x, int out) override { ok(psh_spmv(ctx_, slot(x), slot(out))); }
  void spmv_t(int y, int out) override { ok(psh_spmv_t(ctx_, slot(y), slot(out))); }
  double dot(int a, int b) override {
    double r = 0.0;
    ok(psh_dot(ctx_, slot(a), slot(b), PSH_RED_CPU, &r));
    return r;
  }
  double norm2(int a) override { return std::sqrt(dot(a, a)); }
  double diff_norm2(int a, int b) override {
    double r = 0.0;
    ok(psh_diff_norm2(ctx_, slot(a), slot(b), PSH_RED_CPU, &r));
    return r;
  }

  void primal_step(int x, int aty, double tau, int xhat) override {
    ok(psh_primal_step(ctx_, slot(x), slot(aty), tau, slot(xhat)));
  }
  void dual_step(int y, int ax, double sigma, int yhat) override {
    ok(psh_dual_step(ctx_, slot(y), slot(ax), sigma, slot(yhat)));
  }
  void r2h_primal(int x, int x0, int aty, double tau, double w, double rho, int xhat, int xbar) override {
    ok(psh_r2h_primal(ctx_, slot(x), slot(x0), slot(aty), tau, w, rho, slot(xhat), slot(xbar)));
  }
  // ybar < 0: the reflected dual point is not wanted (the reference's optional output).
  void r2h_dual(int y, int y0, int ax, double sigma, double w, double rho, int yhat, int ybar) override {
    ok(psh_r2h_dual(ctx_, slot(y), slot(y0), slot(ax), sigma, w, rho, slot(yhat), ybar < 0 ? -1 : slot(ybar)));
  }

  KktStats kkt(int x_scaled, int y_scaled) override {
    psh_kkt k{};
    ok(psh_kkt_of(ctx_, slot(x_scaled), slot(y_scaled), PSH_RED_CPU, &k));
    KktStats s;
    s.primal_residual = k.primal_residual;
    s.dual_residual = k.dual_residual;
    s.primal_obj = k.primal_obj;
    s.dual_obj = k.dual_obj;
    s.b_norm = k.b_norm;
    s.c_norm = k.c_norm;
    s.primal_max_rel = k.primal_max_rel;
    s.dual_max_rel = k.dual_max_rel;
    return s;  // obj_shift is the driver's to set
  }

  void sync() override { ok(psh_sync(ctx_)); }

 private:
  static constexpr int kSlotsPerSpace = 6;

  // Out of device memory surfaces as std::bad_alloc, which the reference's solve() reports as
  // NotSolved "out of memory"; every other failure as std::runtime_error with psh's message.
  void ok(int rc) const {
    if (rc == PSH_OK) return;
    if (rc == PSH_E_OOM) throw std::bad_alloc();
    throw std::runtime_error(std::string("hip backend: ") + psh_error(ctx_));
  }
  int slot(int v) const {
    if (v < 0 || v >= static_cast<int>(slot_.size())) throw std::out_of_range("hip backend: unknown vector handle");
    return slot_[v];
  }
  std::size_t size_of(int v) const {
    slot(v);
    return static_cast<std::size_t>(dual_[v] ? m_ : n_);
  }

  psh_ctx* ctx_ = nullptr;
  std::vector<int> slot_;     // handle → psh slot
  std::vector<bool> dual_;    // handle → lives in the dual space (size m)
  int primal_used_ = 0, dual_used_ = 0;
  int m_ = 0, n_ = 0;
};

}  // namespace

std::unique_ptr<Backend> make_hip_backend() { return std::make_unique<HipBackend>(); }

}  // namespace ps26119::pdhg
