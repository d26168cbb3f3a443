// model_io.cpp — .lpm reader and solution file writer / reader (see model_io.h).
#include "model_io.h"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace ps26119::amd {

namespace {

// Whitespace tokens over the whole text, with line numbers for messages; name sections are read
// line by line.
class Text {
 public:
  explicit Text(std::string s) : s_(std::move(s)) {}
  int line() const { return line_; }
  bool token(std::string& out) {
    skip_space_and_comments();
    if (p_ >= s_.size()) return false;
    const std::size_t b = p_;
    while (p_ < s_.size() && !std::isspace(static_cast<unsigned char>(s_[p_]))) ++p_;
    out.assign(s_, b, p_ - b);
    return true;
  }
  // The rest of the current line (after the last token), trimmed.
  std::string rest_of_line() {
    while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\t')) ++p_;
    const std::size_t b = p_;
    while (p_ < s_.size() && s_[p_] != '\n') ++p_;
    std::string r(s_, b, p_ - b);
    while (!r.empty() && (r.back() == '\r' || r.back() == ' ' || r.back() == '\t')) r.pop_back();
    return r;
  }
  // The next full line (for names: they may contain spaces).
  bool next_line(std::string& out) {
    if (p_ < s_.size() && s_[p_] == '\n') ++p_, ++line_;
    if (p_ >= s_.size()) return false;
    const std::size_t b = p_;
    while (p_ < s_.size() && s_[p_] != '\n') ++p_;
    out.assign(s_, b, p_ - b);
    if (!out.empty() && out.back() == '\r') out.pop_back();
    return true;
  }

 private:
  void skip_space_and_comments() {
    for (;;) {
      while (p_ < s_.size() && std::isspace(static_cast<unsigned char>(s_[p_]))) {
        if (s_[p_] == '\n') ++line_;
        ++p_;
      }
      if (p_ < s_.size() && s_[p_] == '#' && (p_ == 0 || s_[p_ - 1] == '\n')) {
        while (p_ < s_.size() && s_[p_] != '\n') ++p_;
        continue;
      }
      return;
    }
  }
  std::string s_;
  std::size_t p_ = 0;
  int line_ = 1;
};

bool parse_double(const std::string& t, double& v) {
  if (t == "inf" || t == "+inf" || t == "Infinity" || t == "+Infinity") return v = HUGE_VAL, true;
  if (t == "-inf" || t == "-Infinity") return v = -HUGE_VAL, true;
  char* end = nullptr;
  errno = 0;
  v = std::strtod(t.c_str(), &end);
  return end != t.c_str() && *end == '\0' && !std::isnan(v);
}

bool parse_int(const std::string& t, long long& v) {
  char* end = nullptr;
  errno = 0;
  v = std::strtoll(t.c_str(), &end, 10);
  return end != t.c_str() && *end == '\0' && errno == 0;
}

}  // namespace

bool read_lpm(const std::string& path, Model& model, Names* names, std::string& error) {
  std::ifstream in(path, std::ios::binary);
  if (!in) {
    error = "cannot open " + path;
    return false;
  }
  std::stringstream ss;
  ss << in.rdbuf();
  Text t(ss.str());
  auto fail = [&](const std::string& what) {
    error = path + ":" + std::to_string(t.line()) + ": " + what;
    return false;
  };
  std::string tok;
  if (!t.token(tok) || tok != "LPM") return fail("not an .lpm file (expected 'LPM 1')");
  if (!t.token(tok) || tok != "1") return fail("unsupported .lpm version " + tok);
  Model m;
  long long rows = -1, cols = -1, nnz = -1;
  auto numbers = [&](std::vector<double>& v, long long count, const char* what) {
    v.resize(static_cast<std::size_t>(count));
    for (long long k = 0; k < count; ++k)
      if (!t.token(tok) || !parse_double(tok, v[static_cast<std::size_t>(k)]))
        return fail(std::string("bad or missing number in ") + what);
    return true;
  };
  auto integers = [&](std::vector<int>& v, long long count, const char* what) {
    v.resize(static_cast<std::size_t>(count));
    for (long long k = 0; k < count; ++k) {
      long long x;
      if (!t.token(tok) || !parse_int(tok, x) || x < INT32_MIN || x > INT32_MAX)
        return fail(std::string("bad or missing integer in ") + what);
      v[static_cast<std::size_t>(k)] = static_cast<int>(x);
    }
    return true;
  };
  auto need_sizes = [&]() { return rows >= 0 && cols >= 0 && nnz >= 0; };
  bool ended = false;
  while (t.token(tok)) {
    if (tok == "END") {
      ended = true;
      break;
    }
    long long v;
    if (tok == "NAME") {
      m.name = t.rest_of_line();
    } else if (tok == "SENSE") {
      if (!t.token(tok) || (tok != "MIN" && tok != "MAX")) return fail("SENSE must be MIN or MAX");
      m.sense = tok == "MAX" ? -1 : 1;
    } else if (tok == "ROWS" || tok == "COLS" || tok == "NNZ") {
      const std::string key = tok;
      if (!t.token(tok) || !parse_int(tok, v) || v < 0 || v > INT32_MAX) return fail("bad " + key + " count");
      (key == "ROWS" ? rows : key == "COLS" ? cols : nnz) = v;
    } else if (tok == "OFFSET") {
      if (!t.token(tok) || !parse_double(tok, m.obj_offset) || !std::isfinite(m.obj_offset)) return fail("bad OFFSET");
    } else if (tok == "OBJ" || tok == "COL_LOWER" || tok == "COL_UPPER") {
      if (!need_sizes()) return fail(tok + " before ROWS / COLS / NNZ");
      auto& dst = tok == "OBJ" ? m.obj : tok == "COL_LOWER" ? m.col_lower : m.col_upper;
      if (!numbers(dst, cols, tok.c_str())) return false;
    } else if (tok == "ROW_LOWER" || tok == "ROW_UPPER") {
      if (!need_sizes()) return fail(tok + " before ROWS / COLS / NNZ");
      if (!numbers(tok == "ROW_LOWER" ? m.row_lower : m.row_upper, rows, tok.c_str())) return false;
    } else if (tok == "COL_START") {
      if (!need_sizes()) return fail("COL_START before ROWS / COLS / NNZ");
      if (!integers(m.col_start, cols + 1, "COL_START")) return false;
    } else if (tok == "ROW_INDEX") {
      if (!need_sizes()) return fail("ROW_INDEX before ROWS / COLS / NNZ");
      if (!integers(m.row_index, nnz, "ROW_INDEX")) return false;
    } else if (tok == "VALUE") {
      if (!need_sizes()) return fail("VALUE before ROWS / COLS / NNZ");
      if (!numbers(m.value, nnz, "VALUE")) return false;
    } else if (tok == "INTEGER") {
      if (!need_sizes()) return fail("INTEGER before ROWS / COLS / NNZ");
      std::vector<int> flags;
      if (!integers(flags, cols, "INTEGER")) return false;
      m.is_integer.assign(flags.begin(), flags.end());
    } else if (tok == "ROW_NAMES" || tok == "COL_NAMES") {
      if (!need_sizes()) return fail(tok + " before ROWS / COLS / NNZ");
      const long long count = tok == "ROW_NAMES" ? rows : cols;
      std::vector<std::string> v(static_cast<std::size_t>(count));
      t.rest_of_line();
      for (auto& s : v)
        if (!t.next_line(s)) return fail("missing names in " + tok);
      if (names) (tok == "ROW_NAMES" ? names->rows : names->cols) = std::move(v);
    } else {
      return fail("unknown section '" + tok + "'");
    }
  }
  if (!ended) return fail("missing END");
  if (!need_sizes()) return fail("missing ROWS / COLS / NNZ");
  m.num_rows = static_cast<int>(rows);
  m.num_cols = static_cast<int>(cols);
  if (m.obj.empty() && cols > 0) return fail("missing OBJ");
  if (m.col_start.empty()) return fail("missing COL_START");
  if (m.col_lower.empty()) m.col_lower.assign(static_cast<std::size_t>(cols), 0.0);
  if (m.col_upper.empty()) m.col_upper.assign(static_cast<std::size_t>(cols), HUGE_VAL);
  if (m.row_lower.empty() || m.row_upper.empty()) return fail("missing ROW_LOWER / ROW_UPPER");
  if (auto err = m.validate(); !err.empty()) return fail("invalid model: " + err);
  model = std::move(m);
  return true;
}

bool write_solution(const std::string& path, const Model& m, const Names& names, const Solution& s,
                    std::string& error) {
  std::FILE* f = std::fopen(path.c_str(), "w");
  if (!f) {
    error = "cannot open " + path + " for writing";
    return false;
  }
  const std::string fp = s.model_fingerprint.empty() ? m.fingerprint_hex() : s.model_fingerprint;
  std::fprintf(f, "PS26119-SOLUTION 1\nmodel %s\nname %s\nstatus %s\nengine %s\nprecision %s\n", fp.c_str(),
               m.name.c_str(), to_string(s.status), s.engine.c_str(), s.precision.c_str());
  std::fprintf(f, "objective %.17g\ndual_objective %.17g\n", s.objective, s.dual_objective);
  if (!std::isnan(s.certified_bound)) std::fprintf(f, "certified_bound %.17g\n", s.certified_bound);
  std::fprintf(f, "primal_residual %.6g\ndual_residual %.6g\ngap %.6g\n", s.primal_residual, s.dual_residual, s.gap);
  std::fprintf(f, "iterations %lld\nseconds %.6f\nsetup_seconds %.6f\n", static_cast<long long>(s.iterations), s.seconds,
               s.setup_seconds);
  if (!std::isnan(s.primal_weight)) std::fprintf(f, "primal_weight %.17g\n", s.primal_weight);
  if (s.iterations_to_fast >= 0)
    std::fprintf(f, "iterations_to_fast %lld\nseconds_to_fast %.6f\n", static_cast<long long>(s.iterations_to_fast),
                 s.seconds_to_fast);
  if (!s.check.empty())
    std::fprintf(f, "check %s primal %.3g dual %.3g gap %.3g\n", s.check.c_str(), s.check_primal, s.check_dual,
                 s.check_gap);
  if (!s.message.empty()) std::fprintf(f, "message %s\n", s.message.c_str());
  const int n = m.num_cols, mm = m.num_rows;
  const bool primal = static_cast<int>(s.x.size()) == n;
  const bool dual = static_cast<int>(s.z.size()) == n && static_cast<int>(s.y.size()) == mm;
  if (primal) {
    auto name = [](const std::vector<std::string>& v, int k) { return k < static_cast<int>(v.size()) ? v[k].c_str() : ""; };
    std::fprintf(f, "COLUMNS %d\n", n);
    for (int j = 0; j < n; ++j) std::fprintf(f, "%d %.17g %.17g %s\n", j, s.x[j], dual ? s.z[j] : 0.0, name(names.cols, j));
    const std::vector<double> act = static_cast<int>(s.row_activity.size()) == mm ? s.row_activity : m.row_activity(s.x);
    std::fprintf(f, "ROWS %d\n", mm);
    for (int i = 0; i < mm; ++i) std::fprintf(f, "%d %.17g %.17g %s\n", i, act[i], dual ? s.y[i] : 0.0, name(names.rows, i));
  }
  auto ray = [&](const char* key, const std::vector<double>& v) {
    if (v.empty()) return;
    std::fprintf(f, "%s %zu\n", key, v.size());
    for (std::size_t k = 0; k < v.size(); ++k) std::fprintf(f, "%zu %.17g\n", k, v[k]);
  };
  ray("DUAL_RAY", s.dual_ray);
  ray("PRIMAL_RAY", s.primal_ray);
  std::fprintf(f, "END\n");
  if (std::fclose(f) != 0) {
    error = "write failed for " + path;
    return false;
  }
  return true;
}

bool read_solution(const std::string& path, Solution& s, std::string& error) {
  std::ifstream in(path);
  if (!in) {
    error = "cannot open " + path;
    return false;
  }
  std::string line;
  if (!std::getline(in, line) || line.rfind("PS26119-SOLUTION", 0) != 0) {
    error = path + ": not a solution file";
    return false;
  }
  auto section = [&](std::vector<double>& out, long long count, int which) {
    out.assign(static_cast<std::size_t>(count), 0.0);
    for (long long k = 0; k < count; ++k) {
      if (!std::getline(in, line)) return false;
      std::istringstream ls(line);
      long long idx;
      std::string a, b;
      if (!(ls >> idx >> a >> b) || idx != k) return false;
      double va, vb;
      if (!parse_double(a, va) || !parse_double(b, vb)) return false;
      out[static_cast<std::size_t>(k)] = which == 0 ? va : vb;
    }
    return true;
  };
  while (std::getline(in, line)) {
    std::istringstream ls(line);
    std::string key;
    ls >> key;
    long long count = 0;
    if (key == "COLUMNS") {
      if (!(ls >> count) || !section(s.x, count, 0)) break;
    } else if (key == "ROWS") {
      if (!(ls >> count) || !section(s.y, count, 1)) break;
    } else if (key == "primal_weight") {
      std::string v;
      ls >> v;
      parse_double(v, s.primal_weight);
    } else if (key == "END") {
      return true;
    }
  }
  error = path + ": malformed solution file";
  return false;
}

}  // namespace ps26119::amd
