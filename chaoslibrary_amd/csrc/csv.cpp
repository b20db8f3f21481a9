// csv.cpp -- byte-compatible writers for the reference's workspace/ files.
#include "csv.hpp"

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

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

void write_map_text(TextFile& f, const MapSpec& m)
{
    if (m.header_lines > 0) write_range_header(f, m.ranges, m.header_lines, m.header_precision);
    if (m.dv) write_map_double(f, m.dv, m.count, m.nPts, m.precision, m.nan_to_zero);
    else write_map_int(f, m.iv, m.count, m.nPts);
}

namespace {
struct BinHeader {
    char magic[8];
    uint32_t version;
    uint32_t type;  // 0 double, 1 int32
    int64_t count;
    int32_t nPts, precision, nan_to_zero, header_lines, header_precision, pad;
    double ranges[4];
};
static_assert(sizeof(BinHeader) == 80, "map sidecar header is 80 bytes");
const char kMagic[8] = {'C', 'H', 'A', 'O', 'S', 'M', 'A', 'P'};
}  // namespace

void write_map_bin(const std::string& path, const MapSpec& m)
{
    BinHeader h{};
    std::memcpy(h.magic, kMagic, 8);
    h.version = 1;
    h.type = m.dv ? 0 : 1;
    h.count = m.count;
    h.nPts = m.nPts;
    h.precision = m.precision;
    h.nan_to_zero = m.nan_to_zero ? 1 : 0;
    h.header_lines = m.header_lines;
    h.header_precision = m.header_precision;
    for (int k = 0; k < 2 * m.header_lines && k < 4; ++k) h.ranges[k] = m.ranges[k];
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) fail(CHAOS_E_IO, "cannot open output file '%s': %s", path.c_str(), std::strerror(errno));
    const size_t es = m.dv ? sizeof(double) : sizeof(int);
    const void* data = m.dv ? (const void*)m.dv : (const void*)m.iv;
    const bool ok = std::fwrite(&h, sizeof h, 1, f) == 1 &&
                    (m.count == 0 || std::fwrite(data, es, (size_t)m.count, f) == (size_t)m.count);
    if (std::fclose(f) != 0 || !ok) fail(CHAOS_E_IO, "write to '%s' failed", path.c_str());
}

}  // namespace chaos

extern "C" int chaos_map_bin_to_csv(const char* bin_path, const char* csv_path)
{
    using namespace chaos;
    return guarded("chaos_map_bin_to_csv", [&] {
        CHAOS_REQUIRE(bin_path && csv_path, "NULL path");
        FILE* f = std::fopen(bin_path, "rb");
        if (!f) fail(CHAOS_E_IO, "cannot open '%s': %s", bin_path, std::strerror(errno));
        BinHeader h{};
        const bool hdr = std::fread(&h, sizeof h, 1, f) == 1;
        if (!hdr || std::memcmp(h.magic, kMagic, 8) != 0 || h.version != 1 || h.type > 1 || h.count < 0 ||
            h.header_lines < 0 || h.header_lines > 2) {
            std::fclose(f);
            fail(CHAOS_E_INVALID, "'%s' is not a map sidecar", bin_path);
        }
        std::vector<double> dv;
        std::vector<int> iv;
        size_t got = 0;
        if (h.type == 0) {
            dv.resize((size_t)h.count);
            got = h.count ? std::fread(dv.data(), sizeof(double), (size_t)h.count, f) : 0;
        } else {
            iv.resize((size_t)h.count);
            got = h.count ? std::fread(iv.data(), sizeof(int), (size_t)h.count, f) : 0;
        }
        std::fclose(f);
        if (got != (size_t)h.count) fail(CHAOS_E_IO, "'%s' is truncated", bin_path);
        MapSpec m;
        m.ranges = h.ranges;
        m.header_lines = h.header_lines;
        m.header_precision = h.header_precision;
        m.nPts = h.nPts;
        m.count = h.count;
        m.precision = h.precision;
        m.nan_to_zero = h.nan_to_zero != 0;
        m.dv = h.type == 0 ? dv.data() : nullptr;
        m.iv = h.type == 1 ? iv.data() : nullptr;
        TextFile t(csv_path);
        write_map_text(t, m);
        t.close();
    });
}
