// model_io.h — the text formats around a GPU solve: the `.lpm` model format (reference
// tools/lpm.py module docstring, version 1; reader src/io/lpm_reader.cpp) and the solution file
// (reference src/io/solution_writer.cpp:8-67, read back by src/io/solution_reader.cpp for --warm).
#pragma once

#include <string>
#include <vector>

#include "contract.h"

namespace ps26119::amd {

struct Names {
  std::vector<std::string> rows, cols;  // empty when the file has none
};

// Reads a .lpm file; on failure returns false with a message naming the line.
bool read_lpm(const std::string& path, Model& model, Names* names, std::string& error);

// Writes the solution file: header keys, then COLUMNS (index, x, reduced cost, name), ROWS
// (index, activity, dual, name), certificate sections and END; doubles as %.17g.
bool write_solution(const std::string& path, const Model& model, const Names& names, const Solution& s,
                    std::string& error);

// Reads x, y (from the COLUMNS / ROWS sections) and the primal weight of a solution file.
bool read_solution(const std::string& path, Solution& s, std::string& error);

}  // namespace ps26119::amd
