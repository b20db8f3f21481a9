// csv.cpp -- byte-compatible writers for the reference's workspace/ files.
#include "csv.hpp"

#include <cerrno>
#include <cmath>
#include <cstring>

#include "runtime.hpp"

namespace chaos {

TextFile::TextFile(const std::string& path, bool append) : path_(path)
{
    f_ = std::fopen(path.c_str(), append ? "ab" : "wb");
    if (!f_) fail(CHAOS_E_IO, "cannot open output file '%s': %s", path.c_str(), std::strerror(errno));
    buf_.reserve(1 << 20);
}

TextFile::~TextFile()
{
    if (f_) {
        flush_buf();
        std::fclose(f_);
    }
}

void TextFile::flush_buf()
{
    if (!buf_.empty() && f_) {
        if (std::fwrite(buf_.data(), 1, buf_.size(), f_) != buf_.size())
            fail(CHAOS_E_IO, "write to '%s' failed", path_.c_str());
        buf_.clear();
    }
}

void TextFile::put(const char* s, size_t n)
{
    buf_.append(s, n);
    if (buf_.size() >= (1u << 20)) flush_buf();
}

void TextFile::put(const char* s) { put(s, std::strlen(s)); }

void TextFile::put_g(double v, int precision)
{
    char tmp[64];
    int n = std::snprintf(tmp, sizeof tmp, "%.*g", precision, v);
    put(tmp, (size_t)n);
}

void TextFile::put_int(long long v)
{
    char tmp[32];
    int n = std::snprintf(tmp, sizeof tmp, "%lld", v);
    put(tmp, (size_t)n);
}

void TextFile::close()
{
    if (!f_) return;
    flush_buf();
    if (std::fclose(f_) != 0) {
        f_ = nullptr;
        fail(CHAOS_E_IO, "closing '%s' failed", path_.c_str());
    }
    f_ = nullptr;
}

void write_range_header(TextFile& f, const double* ranges, int lines, int precision)
{
    for (int l = 0; l < lines; ++l) {
        f.put_g(ranges[2 * l], precision);
        f.put(" ", 1);
        f.put_g(ranges[2 * l + 1], precision);
        f.put("\n", 1);
    }
}

void write_map_double(TextFile& f, const double* v, long long count, int nPts, int precision,
                      bool nan_to_zero)
{
    int counter = 0;
    for (long long i = 0; i < count; ++i) {
        if (counter != 0) f.put(", ", 2);
        if (counter == nPts) {
            f.put("\n", 1);
            counter = 0;
        }
        double x = v[i];
        if (nan_to_zero && std::isnan(x)) f.put("0", 1);
        else f.put_g(x, precision);
        ++counter;
    }
}

void write_map_int(TextFile& f, const int* v, long long count, int nPts)
{
    int counter = 0;
    for (long long i = 0; i < count; ++i) {
        if (counter != 0) f.put(", ", 2);
        if (counter == nPts) {
            f.put("\n", 1);
            counter = 0;
        }
        f.put_int(v[i]);
        ++counter;
    }
}

}  // namespace chaos
