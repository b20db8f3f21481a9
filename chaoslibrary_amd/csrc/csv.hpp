// csv.hpp -- the reference's text output formats (what scripts/*.py parse).
//
// std::ostream prints a double in the default float field with precision p
// exactly as printf("%.{p}g"); these writers produce the same bytes with
// snprintf and large buffered writes.
#pragma once

#include <cstdio>
#include <string>

#include "../../include/chaoslibrary.h"

namespace chaos {

class TextFile {
public:
    // Opens for writing (truncate) or appending; fails with CHAOS_E_IO.
    TextFile(const std::string& path, bool append = false);
    ~TextFile();
    TextFile(const TextFile&) = delete;
    TextFile& operator=(const TextFile&) = delete;

    void put(const char* s, size_t n);
    void put(const char* s);
    void put_g(double v, int precision);  // "%.{p}g"
    void put_int(long long v);
    void close();

private:
    FILE* f_ = nullptr;
    std::string path_;
    std::string buf_;
    void flush_buf();
};

// Two header lines "<r0> <r1>\n<r2> <r3>\n" (or one for 1-D maps).
void write_range_header(TextFile& f, const double* ranges, int lines, int precision);

// Row-major n*n map in the reference layout: values separated by ", ", and
// "\n" emitted after the separator every nPts values, so rows end ", \n"
// and the file has no final newline
// (LLE/src/LLECUDA.cu:384-398, Bifurcation/src/bifurcationCUDA.cu:221-233).
void write_map_double(TextFile& f, const double* v, long long count, int nPts, int precision,
                      bool nan_to_zero);
void write_map_int(TextFile& f, const int* v, long long count, int nPts);

// chaos_options.output: which forms of a map file to write (0 = CSV).
inline bool out_csv(int output) { return output == 0 || (output & CHAOS_OUT_CSV); }
inline bool out_bin(int output) { return (output & CHAOS_OUT_BINARY) != 0; }

// One map file: optional range header lines, then count values (doubles at
// `precision`, NaN optionally written as 0, or ints) in the row layout above.
struct MapSpec {
    const double* ranges = nullptr;
    int header_lines = 0;
    int header_precision = 6;
    int nPts = 0;
    long long count = 0;
    int precision = 6;
    bool nan_to_zero = false;
    const double* dv = nullptr;  // exactly one of dv / iv
    const int* iv = nullptr;
};

void write_map_text(TextFile& f, const MapSpec& m);
// Binary sidecar: an 80-byte header (magic "CHAOSMAP", version, value type,
// count, nPts, precision, NaN rule, header lines and precision, ranges) and
// the raw values, little-endian.
void write_map_bin(const std::string& path, const MapSpec& m);

}  // namespace chaos
