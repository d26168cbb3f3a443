// ps26119-hip — command-line front end of the MI355X backend: the reference CLI's GPU solve and
// batch commands (reference apps/cli/main.cpp:191-355 `solve`, :366-450 `batch`) over this
// library. Same exit codes: 0 optimal, 1 limit / infeasible / unbounded, 2 usage, 3 read error,
// 4 cannot write the output, 5 numerical error / not solved.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "batch.h"
#include "model_io.h"
#include "ps26119/ps26119.h"
#include "solve.h"

using namespace ps26119;
using namespace ps26119::amd;

namespace {

constexpr const char* kName = "ps26119-hip";
constexpr int kUsage = 2, kRead = 3, kWrite = 4, kFailed = 5;

void usage(std::FILE* f) {
  std::fprintf(f,
               "%s — r2HPDHG / PDLP LP solves on an AMD Instinct MI355X\n"
               "usage:\n"
               "  %s --version\n"
               "  %s solve <file.lpm> [--algorithm auto|r2hpdhg|pdlp] [--precision fp64|mixed] [--gpu]\n"
               "        [--tol e] [--time-limit s] [--iteration-limit n] [--threads n] [--no-presolve]\n"
               "        [--warm file.sol [--warm-weight]] [--set name=value]... [--out file.sol] [-v|-vv]\n"
               "  %s batch <base.lpm> <scenario.lpm>... [--out-dir d] [--tol e] [--time-limit s]\n"
               "        [--iteration-limit n] [--precision fp64|mixed] [--set name=value]... [-v|-vv]\n"
               "        scenarios share the base matrix; their objective and bounds may differ\n"
               "exit codes: 0 optimal, 1 infeasible/unbounded/limit, 2 usage, 3 read error,\n"
               "            4 cannot write output, 5 numerical error / not solved\n",
               kName, kName, kName, kName);
}

int exit_code(Status s) {
  switch (s) {
    case Status::Optimal: return 0;
    case Status::Infeasible:
    case Status::Unbounded:
    case Status::IterationLimit:
    case Status::TimeLimit: return 1;
    default: return kFailed;
  }
}

double number(const std::string& flag, const char* s, bool positive) {
  char* end = nullptr;
  const double v = std::strtod(s, &end);
  if (end == s || *end != '\0' || !std::isfinite(v) || (positive ? !(v > 0) : !(v >= 0))) {
    std::fprintf(stderr, "%s needs a %s number, got '%s'\n", flag.c_str(), positive ? "positive" : "non-negative", s);
    std::exit(kUsage);
  }
  return v;
}

std::pair<std::string, double> name_value(const std::string& kv) {
  const auto eq = kv.find('=');
  char* end = nullptr;
  const double v = eq == std::string::npos ? 0 : std::strtod(kv.c_str() + eq + 1, &end);
  if (eq == std::string::npos || eq == 0 || end == kv.c_str() + eq + 1 || *end != '\0') {
    std::fprintf(stderr, "--set expects name=number, got '%s'\n", kv.c_str());
    std::exit(kUsage);
  }
  return {kv.substr(0, eq), v};
}

// Fails fast when an output cannot be created (a probe file is removed again).
bool writable(const std::string& path) {
  std::FILE* probe = std::fopen(path.c_str(), "r");
  const bool existed = probe != nullptr;
  if (probe) std::fclose(probe);
  std::FILE* f = std::fopen(path.c_str(), "a");
  if (!f) return false;
  std::fclose(f);
  if (!existed) std::remove(path.c_str());
  return true;
}

bool load(const std::string& path, Model& m, Names* names) {
  std::string err;
  if (path.size() < 4 || path.compare(path.size() - 4, 4, ".lpm") != 0) {
    std::fprintf(stderr, "read error: %s: expected a .lpm file (MPS is read by the reference CLI)\n", path.c_str());
    return false;
  }
  if (!read_lpm(path, m, names, err)) {
    std::fprintf(stderr, "read error: %s\n", err.c_str());
    return false;
  }
  return true;
}

void report(const Model& model, const Solution& s) {
  std::printf("model      %s  rows %d  cols %d  nnz %zu  fingerprint %s\n", model.name.c_str(), model.num_rows,
              model.num_cols, model.nnz(), s.model_fingerprint.c_str());
  std::printf("status     %s\n", to_string(s.status));
  std::printf("engine     %s (%s)\n", s.engine.c_str(), s.precision.c_str());
  std::printf("objective  %.12g\n", s.objective);
  std::printf("residuals  primal %.2e  dual %.2e  gap %.2e\n", s.primal_residual, s.dual_residual, s.gap);
  const bool claim = s.status == Status::Infeasible || s.status == Status::Unbounded;
  if (claim)
    std::printf("check      %s\n", s.check.empty() ? "NOT CERTIFIED (no certificate)"
                                                   : (s.check + "  (certificate verified on the original model)").c_str());
  else if (!s.check.empty())
    std::printf("check      %s  (in-process, original model: primal %.1e  dual %.1e  gap %.1e)\n", s.check.c_str(),
                s.check_primal, s.check_dual, s.check_gap);
  if (!std::isnan(s.certified_bound) && !claim) {
    if (std::isfinite(s.certified_bound))
      std::printf("certified  %s %.12g  (rounding-proof %s bound from y)\n", model.sense > 0 ? "optimum >=" : "optimum <=",
                  s.certified_bound, model.sense > 0 ? "lower" : "upper");
    else
      std::printf("certified  none  (y gives no finite rounding-proof bound)\n");
  }
  std::printf("iterations %lld   seconds %.3f   (setup %.3f)\n", static_cast<long long>(s.iterations), s.seconds,
              s.setup_seconds);
  if (s.iterations_to_fast >= 0)
    std::printf("to 1e-4    iterations %lld   seconds %.3f\n", static_cast<long long>(s.iterations_to_fast),
                s.seconds_to_fast);
  if (!s.message.empty()) std::printf("message    %s\n", s.message.c_str());
}

int cmd_solve(int argc, char** argv) {
  std::string path, out, warm;
  bool warm_weight = false;
  Options o;
  o.use_gpu = true;  // this program is the GPU path; --gpu is accepted for the reference's command lines
  for (int i = 0; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> const char* {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "missing value for %s\n", a.c_str());
        std::exit(kUsage);
      }
      return argv[++i];
    };
    if (a == "--algorithm") {
      const std::string v = next();
      if (v == "auto") o.algorithm = Algorithm::Auto;
      else if (v == "r2hpdhg") o.algorithm = Algorithm::R2hpdhg;
      else if (v == "pdlp") o.algorithm = Algorithm::Pdlp;
      else {
        std::fprintf(stderr, "unknown algorithm '%s' (the GPU runs r2hpdhg and pdlp)\n", v.c_str());
        return kUsage;
      }
    } else if (a == "--precision") {
      const std::string v = next();
      if (v == "fp64") o.precision = Precision::Fp64;
      else if (v == "mixed") o.precision = Precision::Mixed;
      else {
        std::fprintf(stderr, "unknown precision '%s'\n", v.c_str());
        return kUsage;
      }
    } else if (a == "--gpu") {
      o.use_gpu = true;
    } else if (a == "--tol") {
      o.tolerance = number(a, next(), true);
    } else if (a == "--time-limit") {
      o.time_limit = number(a, next(), true);
    } else if (a == "--iteration-limit") {
      o.iteration_limit = static_cast<std::int64_t>(number(a, next(), true));
    } else if (a == "--threads") {
      o.threads = static_cast<int>(number(a, next(), false));
    } else if (a == "--presolve") {
      o.presolve = true;
    } else if (a == "--no-presolve") {
      o.presolve = false;
    } else if (a == "--warm") {
      warm = next();
    } else if (a == "--warm-weight") {
      warm_weight = true;
    } else if (a == "--set") {
      o.engine_params.push_back(name_value(next()));
    } else if (a == "--out") {
      out = next();
    } else if (a == "-v") {
      o.verbosity = 1;
    } else if (a == "-vv") {
      o.verbosity = 2;
    } else if (!a.empty() && a[0] == '-') {
      std::fprintf(stderr, "unknown option %s\n", a.c_str());
      return kUsage;
    } else if (path.empty()) {
      path = a;
    } else {
      std::fprintf(stderr, "more than one input file\n");
      return kUsage;
    }
  }
  if (path.empty()) {
    usage(stderr);
    return kUsage;
  }
  if (!out.empty() && !writable(out)) {
    std::fprintf(stderr, "cannot write the solution file '%s'\n", out.c_str());
    return kWrite;
  }
  Model model;
  Names names;
  if (!load(path, model, &names)) return kRead;
  if (!warm.empty()) {
    Solution prev;
    std::string err;
    if (!read_solution(warm, prev, err)) {
      std::fprintf(stderr, "read error (warm start): %s\n", err.c_str());
      return kRead;
    }
    if (static_cast<int>(prev.x.size()) != model.num_cols || static_cast<int>(prev.y.size()) != model.num_rows) {
      std::fprintf(stderr, "warm start %s does not match the model's size\n", warm.c_str());
      return kUsage;
    }
    o.warm_x = std::move(prev.x);
    o.warm_y = std::move(prev.y);
    if (warm_weight && prev.primal_weight > 0) o.warm_primal_weight = prev.primal_weight;
  }
  const Solution s = solve(model, o);
  report(model, s);
  if (!out.empty()) {
    std::string err;
    if (!write_solution(out, model, names, s, err)) {
      std::fprintf(stderr, "%s\n", err.c_str());
      return kWrite;
    }
  }
  return exit_code(s.status);
}

int cmd_batch(int argc, char** argv) {
  std::vector<std::string> files;
  std::string out_dir;
  Options o;
  o.use_gpu = true;
  for (int i = 0; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> const char* {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "missing value for %s\n", a.c_str());
        std::exit(kUsage);
      }
      return argv[++i];
    };
    if (a == "--out-dir") out_dir = next();
    else if (a == "--tol") o.tolerance = number(a, next(), true);
    else if (a == "--time-limit") o.time_limit = number(a, next(), true);
    else if (a == "--iteration-limit") o.iteration_limit = static_cast<std::int64_t>(number(a, next(), true));
    else if (a == "--threads") o.threads = static_cast<int>(number(a, next(), false));
    else if (a == "--set") o.engine_params.push_back(name_value(next()));
    else if (a == "--precision") {
      const std::string v = next();
      if (v == "fp64") o.precision = Precision::Fp64;
      else if (v == "mixed") o.precision = Precision::Mixed;
      else {
        std::fprintf(stderr, "unknown precision '%s'\n", v.c_str());
        return kUsage;
      }
    } else if (a == "--gpu") o.use_gpu = true;
    else if (a == "-v") o.verbosity = 1;
    else if (a == "-vv") o.verbosity = 2;
    else if (!a.empty() && a[0] == '-') {
      std::fprintf(stderr, "unknown option %s\n", a.c_str());
      return kUsage;
    } else {
      files.push_back(a);
    }
  }
  if (files.size() < 2) {
    usage(stderr);
    return kUsage;
  }
  Model base;
  Names names;
  if (!load(files[0], base, &names)) return kRead;
  std::vector<Scenario> sc;
  std::vector<std::string> labels, model_names;
  for (std::size_t f = 1; f < files.size(); ++f) {
    Model s;
    if (!load(files[f], s, nullptr)) return kRead;
    if (s.num_rows != base.num_rows || s.num_cols != base.num_cols || s.sense != base.sense ||
        s.col_start != base.col_start || s.row_index != base.row_index || s.value != base.value) {
      std::fprintf(stderr, "%s: matrix or sense differs from the base model (batch needs a shared A)\n", files[f].c_str());
      return kUsage;
    }
    sc.push_back(Scenario{s.obj, s.col_lower, s.col_upper, s.row_lower, s.row_upper});
    const auto slash = files[f].find_last_of('/');
    const std::string nm = files[f].substr(slash == std::string::npos ? 0 : slash + 1);
    labels.push_back(nm.substr(0, nm.rfind('.')));
    model_names.push_back(s.name);
  }
  const std::vector<Solution> sols = solve_batch(base, sc, o);
  int worst = 0;
  for (std::size_t k = 0; k < sols.size(); ++k) {
    const Solution& s = sols[k];
    std::printf("%-28s %-14s obj %.12g  it %lld  t %.3fs%s%s\n", labels[k].c_str(), to_string(s.status), s.objective,
                static_cast<long long>(s.iterations), s.seconds, s.message.empty() ? "" : "  ", s.message.c_str());
    worst = std::max(worst, exit_code(s.status));
    if (!out_dir.empty()) {
      Model named = base;  // the file names the scenario; the matrix and names are the base's
      named.name = model_names[k];
      std::string err;
      if (!write_solution(out_dir + "/" + labels[k] + ".sol", named, names, s, err)) {
        std::fprintf(stderr, "%s\n", err.c_str());
        worst = std::max(worst, kWrite);
      }
    }
  }
  return worst;
}

int run(int argc, char** argv) {
  if (argc < 2) {
    usage(stderr);
    return kUsage;
  }
  const std::string cmd = argv[1];
  if (cmd == "--version" || cmd == "version") {
    std::printf("%s %s\n", kName, ps26119_version());
    return 0;
  }
  if (cmd == "--help" || cmd == "-h" || cmd == "help") {
    usage(stdout);
    return 0;
  }
  if (cmd == "solve") return cmd_solve(argc - 2, argv + 2);
  if (cmd == "batch") return cmd_batch(argc - 2, argv + 2);
  std::fprintf(stderr, "unknown command '%s'\n", cmd.c_str());
  usage(stderr);
  return kUsage;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const std::bad_alloc&) {
    std::fprintf(stderr, "fatal: out of memory\n");
  } catch (const std::exception& e) {
    std::fprintf(stderr, "fatal: %s\n", e.what());
  }
  return kFailed;
}
