// csv.hpp -- the reference's text output formats (what scripts/*.py parse).
//
// std::ostream prints a double in the default float field with precision p
// exactly as printf("%.{p}g"); these writers produce the same bytes with
// snprintf and large buffered writes.
#pragma once

#include <cstdio>
#include <string>

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

}  // namespace chaos
