// flow_io.hpp -- host side of the legacy flow-solution files (flow_io.cpp).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace gdtk {

constexpr int FLOW_GZIPTEXT = 0;
constexpr int FLOW_RAWBINARY = 1;

// Variable names of a cell record on this path, in the reference's order; `species` holds
// one name per species ("air" for the ideal-air model).
std::vector<std::string> flow_variable_list(int nsp, const std::vector<std::string>& species);
// Whitespace-separated names.
std::vector<std::string> split_names(const char* names);

// Writes a block's flow file from ncell records of vars.size() doubles each (cell order
// (k*njc + j)*nic + i).
void write_flow_file(const std::string& fn, int format, const std::string& label, double sim_time, int dims, int nic,
                     int njc, int nkc, const std::vector<std::string>& vars, const double* rec, int64_t ncell);
// Reads a block's flow file into records of `vars` (the file may list the variables in any
// order and carry others); returns the file's sim_time.  Throws on a size mismatch or a
// missing variable, as the reference does.
double read_flow_file(const std::string& fn, int format, int dims, int nic, int njc, int nkc,
                      const std::vector<std::string>& vars, std::vector<double>& rec);

}  // namespace gdtk
