// PNG writer for screenshots (the stbi_write_png call of
// Renderer::writeScreenshot, src/Renderer.cpp:787-788): 8-bit RGBA, filter
// type 0 on every row, the zlib stream made of uncompressed (stored) deflate
// blocks (RFC 1950/1951), chunk CRC-32 and zlib Adler-32 computed here.
#include "restirpt_host.h"
#include <algorithm>
#include <cstdio>
#include <vector>

namespace {

uint32_t crcTable(uint32_t n) {
    uint32_t c = n;
    for (int k = 0; k < 8; k++) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
    return c;
}

struct Crc32 {
    uint32_t table[256];
    Crc32() { for (uint32_t n = 0; n < 256; n++) table[n] = crcTable(n); }
    uint32_t operator()(const uint8_t* p, size_t n, uint32_t c = 0xFFFFFFFFu) const {
        for (size_t i = 0; i < n; i++) c = table[(c ^ p[i]) & 0xffu] ^ (c >> 8);
        return c;
    }
};

void be32(std::vector<uint8_t>& o, uint32_t v) {
    o.push_back((uint8_t)(v >> 24)); o.push_back((uint8_t)(v >> 16)); o.push_back((uint8_t)(v >> 8)); o.push_back((uint8_t)v);
}

void chunk(std::vector<uint8_t>& out, const char type[4], const std::vector<uint8_t>& data, const Crc32& crc) {
    be32(out, (uint32_t)data.size());
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    be32(out, crc(out.data() + start, out.size() - start) ^ 0xFFFFFFFFu);
}

}  // namespace

extern "C" rt_status rt_write_png(const char* path, const uint8_t* rgba8, uint32_t width, uint32_t height) {
    if (!path || !rgba8 || width == 0 || height == 0 || width > (1u << 24) || height > (1u << 24))
        return RT_ERR_INVALID_ARGUMENT;
    static const Crc32 crc;
    const size_t rowBytes = (size_t)width * 4;
    // raw scanlines: filter byte 0 + pixels
    std::vector<uint8_t> raw;
    raw.reserve((rowBytes + 1) * height);
    for (uint32_t y = 0; y < height; y++) {
        raw.push_back(0);
        raw.insert(raw.end(), rgba8 + rowBytes * y, rgba8 + rowBytes * (y + 1));
    }
    std::vector<uint8_t> z = {0x78, 0x01};
    for (size_t pos = 0; pos < raw.size() || pos == 0;) {
        const size_t n = std::min<size_t>(65535, raw.size() - pos);
        const bool last = pos + n == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((uint8_t)(n & 0xff)); z.push_back((uint8_t)(n >> 8));
        z.push_back((uint8_t)(~n & 0xff)); z.push_back((uint8_t)((~n >> 8) & 0xff));
        z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + n);
        pos += n;
        if (last) break;
    }
    uint32_t a = 1, b = 0;
    for (uint8_t v : raw) { a = (a + v) % 65521u; b = (b + a) % 65521u; }
    be32(z, (b << 16) | a);
    std::vector<uint8_t> ihdr;
    be32(ihdr, width); be32(ihdr, height);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});  // 8 bits, RGBA, deflate, filter 0, no interlace
    std::vector<uint8_t> out = {137, 80, 78, 71, 13, 10, 26, 10};
    chunk(out, "IHDR", ihdr, crc);
    chunk(out, "IDAT", z, crc);
    chunk(out, "IEND", {}, crc);
    FILE* f = std::fopen(path, "wb");
    if (!f) return RT_ERR_INVALID_ARGUMENT;
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    return (std::fclose(f) == 0 && ok) ? RT_OK : RT_ERR_INVALID_ARGUMENT;
}
