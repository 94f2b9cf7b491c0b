// PNG reader for albedo textures (the reference loads them with stb_image,
// zvk::HostImage::createFromFile, zvk/core/HostImage.cpp:65-96, as 4-channel
// 8-bit images): 8-bit greyscale, grey + alpha, RGB, RGBA and palette images
// (with tRNS alpha), non-interlaced; the zlib stream is inflated here (RFC 1950
// and 1951: stored, fixed and dynamic Huffman blocks; the Adler-32 and chunk
// CRCs are checked) and the five row filters undone. Output: RGBA8, row 0 at
// the top, alpha 255 where the image has none.
#include "Scene.h"
#include "restirpt_host.h"

namespace rti {
void setError(const std::string& msg);  // the calling thread's rt_last_error() message (librestirpt.so)
}
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

struct BitReader {
    const uint8_t* p;
    size_t n, pos = 0;
    uint32_t bitBuf = 0;
    int bitCnt = 0;
    uint32_t bits(int k) {
        while (bitCnt < k) {
            if (pos >= n) throw std::runtime_error("png: truncated deflate stream");
            bitBuf |= (uint32_t)p[pos++] << bitCnt;
            bitCnt += 8;
        }
        uint32_t v = bitBuf & ((1u << k) - 1u);
        bitBuf >>= k;
        bitCnt -= k;
        return v;
    }
    void alignByte() { bitBuf = 0; bitCnt = 0; }
};

// Canonical Huffman code from code lengths: counts per length and symbols in code order.
struct Huffman {
    uint16_t count[16] = {};
    std::vector<uint16_t> symbol;
    void build(const uint8_t* lengths, int n) {
        std::memset(count, 0, sizeof(count));
        for (int i = 0; i < n; i++) count[lengths[i]]++;
        count[0] = 0;
        uint16_t offs[16] = {};
        for (int l = 1; l < 16; l++) offs[l] = (uint16_t)(offs[l - 1] + count[l - 1]);
        symbol.assign(n, 0);
        for (int i = 0; i < n; i++)
            if (lengths[i]) symbol[offs[lengths[i]]++] = (uint16_t)i;
    }
    int decode(BitReader& br) const {
        int code = 0, first = 0, index = 0;
        for (int l = 1; l < 16; l++) {
            code |= (int)br.bits(1);
            int c = count[l];
            if (code - first < c) return symbol[index + (code - first)];
            index += c;
            first = (first + c) << 1;
            code <<= 1;
        }
        throw std::runtime_error("png: invalid Huffman code");
    }
};

const uint16_t kLenBase[29] = {3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31, 35, 43, 51, 59, 67, 83, 99, 115,
                               131, 163, 195, 227, 258};
const uint8_t kLenExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
const uint16_t kDistBase[30] = {1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129, 193, 257, 385, 513, 769,
                                1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
const uint8_t kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11,
                                12, 12, 13, 13};

void inflateBlock(BitReader& br, const Huffman& lit, const Huffman& dist, std::vector<uint8_t>& out) {
    for (;;) {
        int s = lit.decode(br);
        if (s < 256) { out.push_back((uint8_t)s); continue; }
        if (s == 256) return;
        s -= 257;
        if (s >= 29) throw std::runtime_error("png: invalid length symbol");
        size_t len = kLenBase[s] + br.bits(kLenExtra[s]);
        int d = dist.decode(br);
        if (d >= 30) throw std::runtime_error("png: invalid distance symbol");
        size_t back = kDistBase[d] + br.bits(kDistExtra[d]);
        if (back > out.size()) throw std::runtime_error("png: distance before the start of the stream");
        size_t from = out.size() - back;
        for (size_t k = 0; k < len; k++) out.push_back(out[from + k]);
    }
}

std::vector<uint8_t> inflateZlib(const std::vector<uint8_t>& z) {
    if (z.size() < 6) throw std::runtime_error("png: zlib stream too short");
    if ((z[0] & 0x0f) != 8 || ((z[0] << 8) | z[1]) % 31 != 0 || (z[1] & 0x20))
        throw std::runtime_error("png: unsupported zlib header");
    BitReader br{z.data() + 2, z.size() - 2};
    std::vector<uint8_t> out;
    bool last = false;
    while (!last) {
        last = br.bits(1) != 0;
        uint32_t type = br.bits(2);
        if (type == 0) {
            br.alignByte();
            if (br.pos + 4 > br.n) throw std::runtime_error("png: truncated stored block");
            uint32_t len = br.p[br.pos] | (br.p[br.pos + 1] << 8), nlen = br.p[br.pos + 2] | (br.p[br.pos + 3] << 8);
            br.pos += 4;
            if ((len ^ 0xffffu) != nlen || br.pos + len > br.n) throw std::runtime_error("png: bad stored block");
            out.insert(out.end(), br.p + br.pos, br.p + br.pos + len);
            br.pos += len;
        } else if (type == 1) {
            uint8_t l[288], d[30];
            for (int i = 0; i < 144; i++) l[i] = 8;
            for (int i = 144; i < 256; i++) l[i] = 9;
            for (int i = 256; i < 280; i++) l[i] = 7;
            for (int i = 280; i < 288; i++) l[i] = 8;
            for (int i = 0; i < 30; i++) d[i] = 5;
            Huffman lit, dist;
            lit.build(l, 288);
            dist.build(d, 30);
            inflateBlock(br, lit, dist, out);
        } else if (type == 2) {
            int hlit = (int)br.bits(5) + 257, hdist = (int)br.bits(5) + 1, hclen = (int)br.bits(4) + 4;
            static const uint8_t order[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
            uint8_t cl[19] = {};
            for (int i = 0; i < hclen; i++) cl[order[i]] = (uint8_t)br.bits(3);
            Huffman clh;
            clh.build(cl, 19);
            uint8_t lens[320] = {};
            int i = 0;
            while (i < hlit + hdist) {
                int s = clh.decode(br);
                if (s < 16) { lens[i++] = (uint8_t)s; continue; }
                int rep = 0;
                uint8_t v = 0;
                if (s == 16) {
                    if (i == 0) throw std::runtime_error("png: repeat without a previous length");
                    v = lens[i - 1];
                    rep = 3 + (int)br.bits(2);
                } else if (s == 17) {
                    rep = 3 + (int)br.bits(3);
                } else {
                    rep = 11 + (int)br.bits(7);
                }
                if (i + rep > hlit + hdist) throw std::runtime_error("png: code lengths overflow");
                while (rep--) lens[i++] = v;
            }
            Huffman lit, dist;
            lit.build(lens, hlit);
            dist.build(lens + hlit, hdist);
            inflateBlock(br, lit, dist, out);
        } else {
            throw std::runtime_error("png: invalid deflate block type");
        }
    }
    br.alignByte();
    if (br.pos + 4 > br.n) throw std::runtime_error("png: missing Adler-32");
    uint32_t a = 1, b = 0;
    for (uint8_t c : out) { a = (a + c) % 65521u; b = (b + a) % 65521u; }
    uint32_t want = ((uint32_t)br.p[br.pos] << 24) | ((uint32_t)br.p[br.pos + 1] << 16) |
                    ((uint32_t)br.p[br.pos + 2] << 8) | br.p[br.pos + 3];
    if (((b << 16) | a) != want) throw std::runtime_error("png: Adler-32 mismatch");
    return out;
}

uint32_t be32(const uint8_t* p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }

uint32_t crc32(const uint8_t* p, size_t n) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t k = 0; k < 256; k++) {
            uint32_t c = k;
            for (int j = 0; j < 8; j++) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[k] = c;
        }
        init = true;
    }
    uint32_t c = 0xFFFFFFFFu;
    for (size_t i = 0; i < n; i++) c = table[(c ^ p[i]) & 0xffu] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

int paeth(int a, int b, int c) {
    int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

}  // namespace

// Decodes a PNG file into RGBA8.
Texture readPng(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("png: cannot open " + path);
    std::vector<uint8_t> file;
    uint8_t buf[65536];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) file.insert(file.end(), buf, buf + got);
    std::fclose(f);
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (file.size() < 8 || std::memcmp(file.data(), sig, 8) != 0) throw std::runtime_error("png: not a PNG file: " + path);
    uint32_t W = 0, H = 0;
    int depth = 0, ctype = -1;
    std::vector<uint8_t> idat, plte, trns;
    size_t pos = 8;
    bool end = false;
    while (!end) {
        if (pos + 12 > file.size()) throw std::runtime_error("png: truncated chunk in " + path);
        uint32_t len = be32(&file[pos]);
        if (pos + 12 + (size_t)len > file.size()) throw std::runtime_error("png: truncated chunk in " + path);
        const uint8_t* type = &file[pos + 4];
        const uint8_t* data = &file[pos + 8];
        if (crc32(type, len + 4) != be32(data + len)) throw std::runtime_error("png: chunk CRC mismatch in " + path);
        std::string t((const char*)type, 4);
        if (t == "IHDR") {
            if (len != 13) throw std::runtime_error("png: bad IHDR");
            W = be32(data); H = be32(data + 4);
            depth = data[8]; ctype = data[9];
            if (data[10] != 0 || data[11] != 0) throw std::runtime_error("png: unsupported compression or filter method");
            if (data[12] != 0) throw std::runtime_error("png: interlaced images are not supported");
        } else if (t == "PLTE") {
            plte.assign(data, data + len);
        } else if (t == "tRNS") {
            trns.assign(data, data + len);
        } else if (t == "IDAT") {
            idat.insert(idat.end(), data, data + len);
        } else if (t == "IEND") {
            end = true;
        }
        pos += 12 + (size_t)len;
    }
    if (W == 0 || H == 0 || W > 16384 || H > 16384) throw std::runtime_error("png: bad image size in " + path);
    if (depth != 8) throw std::runtime_error("png: only 8-bit channels are supported: " + path);
    int ch = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
    if (ch == 0) throw std::runtime_error("png: unsupported colour type in " + path);
    if (ctype == 3 && (plte.empty() || plte.size() % 3)) throw std::runtime_error("png: palette image without PLTE");
    std::vector<uint8_t> raw = inflateZlib(idat);
    const size_t stride = (size_t)W * ch;
    if (raw.size() != (stride + 1) * H) throw std::runtime_error("png: image data size mismatch in " + path);
    std::vector<uint8_t> img(stride * H);
    for (uint32_t y = 0; y < H; y++) {
        const uint8_t ft = raw[y * (stride + 1)];
        const uint8_t* src = &raw[y * (stride + 1) + 1];
        uint8_t* dst = &img[y * stride];
        const uint8_t* up = y ? &img[(y - 1) * stride] : nullptr;
        for (size_t i = 0; i < stride; i++) {
            int a = i >= (size_t)ch ? dst[i - ch] : 0, b = up ? up[i] : 0, c = (up && i >= (size_t)ch) ? up[i - ch] : 0;
            int v = src[i];
            switch (ft) {
            case 0: break;
            case 1: v += a; break;
            case 2: v += b; break;
            case 3: v += (a + b) >> 1; break;
            case 4: v += paeth(a, b, c); break;
            default: throw std::runtime_error("png: invalid row filter in " + path);
            }
            dst[i] = (uint8_t)v;
        }
    }
    Texture tex;
    tex.width = W; tex.height = H;
    tex.rgba.resize((size_t)W * H * 4);
    for (size_t p = 0; p < (size_t)W * H; p++) {
        uint8_t* o = &tex.rgba[p * 4];
        const uint8_t* s = &img[p * ch];
        switch (ctype) {
        case 0: o[0] = o[1] = o[2] = s[0]; o[3] = 255; break;
        case 4: o[0] = o[1] = o[2] = s[0]; o[3] = s[1]; break;
        case 2: o[0] = s[0]; o[1] = s[1]; o[2] = s[2]; o[3] = 255; break;
        case 6: std::memcpy(o, s, 4); break;
        case 3: {
            size_t k = s[0];
            if (3 * k + 2 >= plte.size()) throw std::runtime_error("png: palette index out of range in " + path);
            o[0] = plte[3 * k]; o[1] = plte[3 * k + 1]; o[2] = plte[3 * k + 2];
            o[3] = k < trns.size() ? trns[k] : 255;
            break;
        }
        }
    }
    return tex;
}

extern "C" rt_status rt_read_png(const char* path, uint32_t* width, uint32_t* height, uint8_t* rgba8) {
    if (!path || !width || !height) return RT_ERR_INVALID_ARGUMENT;
    try {
        Texture t = readPng(path);
        *width = t.width;
        *height = t.height;
        if (rgba8) std::memcpy(rgba8, t.rgba.data(), t.rgba.size());
        return RT_OK;
    } catch (const std::exception& e) {
        rti::setError(e.what());
        return RT_ERR_INVALID_ARGUMENT;
    }
}
