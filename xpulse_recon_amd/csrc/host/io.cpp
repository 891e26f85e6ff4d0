#include "io.hpp"

#include <dlfcn.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <mutex>
#include <stdexcept>

namespace fs = std::filesystem;

namespace xpr {

// ---------------------------------------------------------------------------
// TIFF
// ---------------------------------------------------------------------------
namespace {

struct Reader {
    std::vector<uint8_t> buf;
    bool big = false;
    uint16_t u16(size_t o) const
    {
        if (o + 2 > buf.size()) throw std::runtime_error("Corrupted or missing TIFF file");
        return big ? uint16_t(buf[o] << 8 | buf[o + 1]) : uint16_t(buf[o] | buf[o + 1] << 8);
    }
    uint32_t u32(size_t o) const
    {
        if (o + 4 > buf.size()) throw std::runtime_error("Corrupted or missing TIFF file");
        return big ? (uint32_t(buf[o]) << 24 | uint32_t(buf[o + 1]) << 16 |
                      uint32_t(buf[o + 2]) << 8 | buf[o + 3])
                   : (uint32_t(buf[o]) | uint32_t(buf[o + 1]) << 8 | uint32_t(buf[o + 2]) << 16 |
                      uint32_t(buf[o + 3]) << 24);
    }
};

struct Ifd {
    TiffInfo info;
    std::vector<uint32_t> offsets, counts;
};

std::vector<uint8_t> slurp(const std::string &path, size_t limit = 0)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("Corrupted or missing TIFF file: " + path);
    std::vector<uint8_t> b;
    if (limit) {
        b.resize(limit);
        f.read((char *)b.data(), (std::streamsize)limit);
        b.resize(size_t(f.gcount()));
    } else {
        b.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    }
    return b;
}

Ifd parse(const Reader &r)
{
    if (r.buf.size() < 8) throw std::runtime_error("Corrupted or missing TIFF file");
    const uint32_t ifd = r.u32(4);
    const uint16_t n = r.u16(ifd);
    Ifd out;
    out.info.samples = 1;
    out.info.compression = 1;
    out.info.sample_format = 1;
    auto values = [&](size_t e, std::vector<uint32_t> &v) {
        const uint16_t type = r.u16(e + 2);
        const uint32_t count = r.u32(e + 4);
        const size_t sz = (type == 3) ? 2 : 4;
        size_t at = (sz * count <= 4) ? e + 8 : r.u32(e + 8);
        v.resize(count);
        for (uint32_t k = 0; k < count; ++k) v[k] = (type == 3) ? r.u16(at + 2 * k) : r.u32(at + 4 * k);
    };
    auto scalar = [&](size_t e) {
        const uint16_t type = r.u16(e + 2);
        return type == 3 ? uint32_t(r.u16(e + 8)) : r.u32(e + 8);
    };
    for (uint16_t k = 0; k < n; ++k) {
        const size_t e = ifd + 2 + 12 * size_t(k);
        switch (r.u16(e)) {
        case 256: out.info.width = scalar(e); break;
        case 257: out.info.height = scalar(e); break;
        case 258: out.info.bits = int(scalar(e)); break;
        case 259: out.info.compression = int(scalar(e)); break;
        case 273: values(e, out.offsets); break;
        case 277: out.info.samples = int(scalar(e)); break;
        case 279: values(e, out.counts); break;
        case 339: out.info.sample_format = int(scalar(e)); break;
        default: break;
        }
    }
    return out;
}

Reader load(const std::string &path, size_t limit = 0)
{
    Reader r;
    r.buf = slurp(path, limit);
    if (r.buf.size() < 8 || !((r.buf[0] == 'I' && r.buf[1] == 'I') || (r.buf[0] == 'M' && r.buf[1] == 'M')))
        throw std::runtime_error("Corrupted or missing TIFF file: " + path);
    r.big = r.buf[0] == 'M';
    if (r.u16(2) != 42) throw std::runtime_error("Corrupted or missing TIFF file: " + path);
    return r;
}

void put16(std::vector<uint8_t> &b, uint16_t v) { b.push_back(uint8_t(v)); b.push_back(uint8_t(v >> 8)); }
void put32(std::vector<uint8_t> &b, uint32_t v)
{
    for (int k = 0; k < 4; ++k) b.push_back(uint8_t(v >> (8 * k)));
}

} // namespace

TiffInfo tiff_info(const std::string &path)
{
    Reader r = load(path);
    return parse(r).info;
}

std::vector<float> tiff_read_f32(const std::string &path, int64_t *width, int64_t *height)
{
    Reader r = load(path);
    Ifd d = parse(r);
    const TiffInfo &i = d.info;
    if (i.sample_format != 3 || i.bits != 32 || i.samples != 1)
        throw std::runtime_error("Trying to load a tiff with an incompatble format (should be "
                                 "floating-point 32bits): " + path);
    if (i.compression != 1) throw std::runtime_error("compressed TIFF not supported: " + path);
    if (i.width <= 0 || i.height <= 0)
        throw std::runtime_error("Trying to load a tiff with an incorrect size: " + path);
    if (d.offsets.size() != d.counts.size() || d.offsets.empty())
        throw std::runtime_error("Corrupted or missing TIFF file: " + path);
    std::vector<float> out(size_t(i.width * i.height));
    size_t pos = 0;
    const size_t total = out.size() * 4;
    uint8_t *dst = (uint8_t *)out.data();
    for (size_t s = 0; s < d.offsets.size() && pos < total; ++s) {
        const size_t n = std::min<size_t>(d.counts[s], total - pos);
        if (size_t(d.offsets[s]) + n > r.buf.size())
            throw std::runtime_error("Corrupted or missing TIFF file: " + path);
        std::memcpy(dst + pos, r.buf.data() + d.offsets[s], n);
        pos += n;
    }
    if (pos != total) throw std::runtime_error("Corrupted or missing TIFF file: " + path);
    if (r.big)
        for (size_t k = 0; k < out.size(); ++k) {
            uint8_t *b = dst + 4 * k;
            std::swap(b[0], b[3]);
            std::swap(b[1], b[2]);
        }
    if (width) *width = i.width;
    if (height) *height = i.height;
    return out;
}

void tiff_write_f32(const std::string &path, const float *data, int64_t width, int64_t height)
{
    const uint32_t bytes = uint32_t(width * height * 4);
    const uint16_t ntags = 11;
    const uint32_t ifd = 8, data_off = ifd + 2 + 12 * ntags + 4;
    std::vector<uint8_t> h;
    h.reserve(data_off);
    h.push_back('I'); h.push_back('I');
    put16(h, 42);
    put32(h, ifd);
    put16(h, ntags);
    auto tag = [&](uint16_t id, uint16_t type, uint32_t v) {
        put16(h, id); put16(h, type); put32(h, 1);
        if (type == 3) { put16(h, uint16_t(v)); put16(h, 0); } else { put32(h, v); }
    };
    tag(256, 4, uint32_t(width));
    tag(257, 4, uint32_t(height));
    tag(258, 3, 32);
    tag(259, 3, 1);
    tag(262, 3, 1);
    tag(273, 4, data_off);
    tag(277, 3, 1);
    tag(278, 4, uint32_t(height));
    tag(279, 4, bytes);
    tag(284, 3, 1);
    tag(339, 3, 3);
    put32(h, 0);
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + path);
    f.write((const char *)h.data(), (std::streamsize)h.size());
    f.write((const char *)data, bytes);
    if (!f) throw std::runtime_error("cannot write " + path);
}

// ---------------------------------------------------------------------------
// zstd spill (libzstd is loaded at run time; its four entry points are
// declared here with their public signatures)
// ---------------------------------------------------------------------------
namespace {

struct Zstd {
    size_t (*compress)(void *, size_t, const void *, size_t, int) = nullptr;
    size_t (*decompress)(void *, size_t, const void *, size_t) = nullptr;
    size_t (*bound)(size_t) = nullptr;
    unsigned (*is_error)(size_t) = nullptr;
    bool ok = false;
};

const Zstd &zstd()
{
    static Zstd z;
    static std::once_flag once;
    std::call_once(once, [] {
        void *h = dlopen("libzstd.so.1", RTLD_NOW | RTLD_LOCAL);
        if (!h) return;
        z.compress = (decltype(z.compress))dlsym(h, "ZSTD_compress");
        z.decompress = (decltype(z.decompress))dlsym(h, "ZSTD_decompress");
        z.bound = (decltype(z.bound))dlsym(h, "ZSTD_compressBound");
        z.is_error = (decltype(z.is_error))dlsym(h, "ZSTD_isError");
        z.ok = z.compress && z.decompress && z.bound && z.is_error;
    });
    return z;
}

} // namespace

bool zstd_available() { return zstd().ok; }

void spill_write(const std::string &path, const float *data, size_t n)
{
    const Zstd &z = zstd();
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + path);
    if (z.ok) {
        std::vector<char> out(z.bound(n * 4));
        const size_t sz = z.compress(out.data(), out.size(), data, n * 4, 1);   // level 1
        if (z.is_error(sz)) throw std::runtime_error("zstd compression failed: " + path);
        f.write(out.data(), (std::streamsize)sz);
    } else {
        f.write((const char *)data, (std::streamsize)(n * 4));
    }
    if (!f) throw std::runtime_error("cannot write " + path);
}

void spill_read(const std::string &path, float *data, size_t n)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot read " + path);
    std::vector<char> in((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    const bool framed = in.size() >= 4 && (uint8_t)in[0] == 0x28 && (uint8_t)in[1] == 0xB5 &&
                        (uint8_t)in[2] == 0x2F && (uint8_t)in[3] == 0xFD;
    if (framed) {
        const Zstd &z = zstd();
        if (!z.ok) throw std::runtime_error("zstd file but libzstd.so.1 is missing: " + path);
        const size_t sz = z.decompress(data, n * 4, in.data(), in.size());
        if (z.is_error(sz) || sz != n * 4) throw std::runtime_error("zstd decompression failed: " + path);
    } else {
        if (in.size() != n * 4) throw std::runtime_error("spill file has the wrong size: " + path);
        std::memcpy(data, in.data(), n * 4);
    }
}

std::vector<std::string> list_tiffs(const std::string &dir)
{
    std::vector<std::string> out;
    for (const auto &e : fs::directory_iterator(dir)) {
        if (!e.is_regular_file()) continue;
        const std::string ext = e.path().extension().string();
        if (ext == ".tif" || ext == ".tiff") out.push_back(e.path().string());
    }
    std::sort(out.begin(), out.end());
    return out;
}

std::string numbered_path(const std::string &dir, const std::string &prefix, int64_t index,
                          int width, const std::string &ext)
{
    char num[32];
    std::snprintf(num, sizeof num, "%0*lld", width, (long long)index);
    return (fs::path(dir) / (prefix + "_" + num + "." + ext)).string();
}

} // namespace xpr
