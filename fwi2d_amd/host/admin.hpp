// admin.hpp -- parameter file, binary / text I/O and acquisition geometry of
// the drop-in executables (src/admin/admin.hpp:1-35, src/geometry/geometry.hpp:1-24).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace fwi {

// Value of `key` in a parameters.txt file (src/admin/admin.cpp:53-98): the
// first non-comment line whose first token is `key`; the value runs from
// after '=' to an optional '#' comment, with spaces removed.  Throws
// std::runtime_error when the file cannot be opened or the key is missing
// and no default is given.
std::string catch_parameter(const std::string &key, const std::string &file);
std::string catch_parameter(const std::string &key, const std::string &file,
                            const std::string &fallback);
int param_int(const std::string &key, const std::string &file);
float param_float(const std::string &key, const std::string &file);
int param_int(const std::string &key, const std::string &file, int fallback);
float param_float(const std::string &key, const std::string &file, float fallback);
uint64_t param_u64(const std::string &key, const std::string &file, uint64_t fallback);
// "true" / "false" in any letter case (src/admin/admin.cpp:3-11); any other
// value is an error.
bool param_bool(const std::string &key, const std::string &file, bool fallback);

// Raw little-endian float32 arrays (admin.cpp:13-37).  import throws when
// the file is missing or shorter than n values.
void import_binary_float(const std::string &path, float *array, long n);
void export_binary_float(const std::string &path, const float *array, long n);

// Lines of a text file, skipping '#' comments and blank lines (admin.cpp:39-51).
std::vector<std::string> import_text_file(const std::string &path);
std::vector<std::string> split(const std::string &s, char delimiter);

// Source / receiver positions from the SPS / RPS files named in the
// parameters ("x,z" in metres per line; src/geometry/geometry.cpp:3-43).
struct Geometry {
    int nsrc = 0, nrec = 0;
    std::vector<float> xsrc, zsrc, xrec, zrec;
    void set_parameters(const std::string &parameters);
};

}  // namespace fwi
