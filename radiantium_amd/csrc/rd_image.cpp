// rd_image.cpp — image decoders and encoders for TextureUtility's formats
// (Radiantium.Core/TextureUtility.cs:54-203), replacing Magick.NET.
//
// The reference reads every image through Magick.NET-Q16-HDRI 11.1.2
// (Radiantium.Core/Radiantium.Core.csproj:11): pixels arrive as float
// quanta in [0, 65535] and LoadImageFromPath divides them by Quantum.Max.
// The decoders here deliver the same floats: 8/16-bit samples v / (2^depth - 1)
// (Magick's 257 * v / 65535 is the same real number, so the same float), and
// float samples (float)(65535.0 * v) / 65535.0f, the float quantum divided
// back.  OpenEXR goes through Magick's RGBA (half) interface, so FLOAT and
// UINT channels are rounded to half first, a missing alpha reads as 1 and the
// image has four channels; a luminance-only file gives R = G = B = Y.
//
// Formats: PNG (8/16-bit, palette, low-depth gray; not interlaced), OpenEXR
// scanline files (NONE, RLE, ZIPS, ZIP, PIZ compression), Radiance RGBE
// (.hdr, run-length and flat scanlines) and PFM.  The OpenEXR and Radiance
// formats are restated from their published specifications (OpenEXR file
// layout and PIZ = bitmap range + LUT + 2D Haar wavelet + canonical Huffman;
// Radiance: Ward's RGBE with the new run-length scanlines).
#include "rd_image.h"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>

#include "../../include/radiantium_hip.h"

namespace rd {

namespace {

[[noreturn]] void bad(const std::string& m) { throw ImageError(RD_ERR_INVALID, m); }
[[noreturn]] void unsupported(const std::string& m) { throw ImageError(RD_ERR_UNSUPPORTED, m); }

// A float sample as Magick's float quantum delivers it after / Quantum.Max.
inline float via_quantum(double v) { return (float)(65535.0 * v) / 65535.0f; }

uint32_t be32(const unsigned char* p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }

// ---------------------------------------------------------------- PNG
RawImage decode_png(const std::string& bytes, const std::string& name) {
    const unsigned char* d = (const unsigned char*)bytes.data();
    size_t n = bytes.size();
    size_t i = 8;
    int w = 0, h = 0, depth = 0, ctype = 0, interlace = 0;
    std::vector<unsigned char> idat, plte;
    bool seen_hdr = false;
    while (i + 8 <= n) {
        uint32_t len = be32(d + i);
        std::string type((const char*)d + i + 4, 4);
        if (i + 12 + (size_t)len > n) bad("image " + name + ": truncated PNG");
        const unsigned char* c = d + i + 8;
        if (type == "IHDR") {
            if (len < 13) bad("image " + name + ": bad IHDR");
            w = (int)be32(c); h = (int)be32(c + 4); depth = c[8]; ctype = c[9]; interlace = c[12];
            seen_hdr = true;
        } else if (type == "PLTE") {
            plte.assign(c, c + len);
        } else if (type == "IDAT") {
            idat.insert(idat.end(), c, c + len);
        } else if (type == "IEND") {
            break;
        }
        i += 12 + (size_t)len;
    }
    if (!seen_hdr || w <= 0 || h <= 0 || (size_t)w * h > (1u << 28)) bad("image " + name + ": bad PNG header");
    if (interlace != 0) unsupported("image " + name + ": interlaced PNG");
    int ch;
    switch (ctype) {
        case 0: ch = 1; break;
        case 2: ch = 3; break;
        case 3: ch = 1; break;   // palette index, expanded to RGB below
        case 4: ch = 2; break;
        case 6: ch = 4; break;
        default: bad("image " + name + ": bad PNG colour type");
    }
    bool low = depth < 8;
    if (!(depth == 8 || depth == 16 || (low && (ctype == 0 || ctype == 3) && (depth == 1 || depth == 2 || depth == 4))))
        unsupported("image " + name + ": PNG bit depth " + std::to_string(depth));
    size_t bpp_bits = (size_t)ch * depth;
    size_t stride = ((size_t)w * bpp_bits + 7) / 8;
    size_t bpp = std::max<size_t>(1, bpp_bits / 8);
    std::vector<unsigned char> raw((stride + 1) * (size_t)h);
    uLongf raw_len = raw.size();
    if (uncompress(raw.data(), &raw_len, idat.data(), idat.size()) != Z_OK || raw_len != raw.size())
        bad("image " + name + ": corrupt PNG data");
    std::vector<unsigned char> img(stride * (size_t)h);
    for (int y = 0; y < h; y++) {
        const unsigned char* src = raw.data() + (size_t)y * (stride + 1);
        unsigned char* cur = img.data() + (size_t)y * stride;
        const unsigned char* prev = y > 0 ? cur - stride : nullptr;
        int f = src[0];
        src++;
        for (size_t x = 0; x < stride; x++) {
            int a = x >= bpp ? cur[x - bpp] : 0;
            int b = prev ? prev[x] : 0;
            int c = (prev && x >= bpp) ? prev[x - bpp] : 0;
            int v = src[x];
            switch (f) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) / 2; break;
                case 4: {
                    int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
                    v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
                    break;
                }
                default: bad("image " + name + ": bad PNG filter");
            }
            cur[x] = (unsigned char)v;
        }
    }
    RawImage im;
    im.w = w; im.h = h;
    im.channels = ctype == 3 ? 3 : ch;
    im.px.assign((size_t)w * h * im.channels, 0.0f);
    const float maxv = (float)((1u << depth) - 1);
    for (int y = 0; y < h; y++) {
        const unsigned char* row = img.data() + (size_t)y * stride;
        for (int x = 0; x < w; x++) {
            auto sample = [&](int k) -> unsigned {
                if (depth == 16) return ((unsigned)row[2 * ((size_t)x * ch + k)] << 8) | row[2 * ((size_t)x * ch + k) + 1];
                if (depth == 8) return row[(size_t)x * ch + k];
                size_t bit = (size_t)x * depth;
                return (row[bit / 8] >> (8 - depth - (bit % 8))) & ((1u << depth) - 1);
            };
            float* o = im.px.data() + ((size_t)y * w + x) * im.channels;
            if (ctype == 3) {
                unsigned idx = sample(0);
                if (3 * idx + 2 >= plte.size()) bad("image " + name + ": palette index out of range");
                for (int k = 0; k < 3; k++) o[k] = (float)plte[3 * idx + k] / 255.0f;
            } else {
                for (int k = 0; k < ch; k++) o[k] = (float)sample(k) / maxv;
            }
        }
    }
    return im;
}

// ---------------------------------------------------------------- OpenEXR
struct In {
    const unsigned char* p;
    size_t n, i = 0;
    std::string name;
    void need(size_t k) const {
        if (i + k > n) bad("image " + name + ": truncated OpenEXR file");
    }
    uint8_t u8() { need(1); return p[i++]; }
    uint32_t u32() { need(4); uint32_t v = p[i] | (p[i + 1] << 8) | (p[i + 2] << 16) | ((uint32_t)p[i + 3] << 24); i += 4; return v; }
    int32_t i32() { return (int32_t)u32(); }
    uint64_t u64() { uint64_t lo = u32(); uint64_t hi = u32(); return lo | (hi << 32); }
    std::string str() {
        size_t s = i;
        while (i < n && p[i]) i++;
        if (i >= n) bad("image " + name + ": truncated OpenEXR header");
        std::string r((const char*)p + s, i - s);
        i++;
        return r;
    }
};

struct ExrChannel {
    std::string name;
    int type;    // 0 UINT, 1 HALF, 2 FLOAT
    int bytes;   // 4, 2, 4
};

enum { kExrNone = 0, kExrRle = 1, kExrZips = 2, kExrZip = 3, kExrPiz = 4 };

// The ZIP / RLE reorder: bytes were split into even / odd halves and delta
// coded (predictor + interleave of the OpenEXR specification).
void exr_unpredict(std::vector<unsigned char>& t) {
    for (size_t k = 1; k < t.size(); k++) t[k] = (unsigned char)(t[k - 1] + t[k] - 128);
    std::vector<unsigned char> o(t.size());
    const size_t half = (t.size() + 1) / 2;
    for (size_t k = 0; k < t.size(); k++) o[k] = (k & 1) ? t[half + k / 2] : t[k / 2];
    t.swap(o);
}

void exr_rle(const unsigned char* s, size_t n, std::vector<unsigned char>& out, size_t want, const std::string& name) {
    out.clear();
    size_t k = 0;
    while (k < n) {
        int c = (signed char)s[k++];
        if (c < 0) {
            size_t cnt = (size_t)(-c);
            if (k + cnt > n || out.size() + cnt > want) bad("image " + name + ": corrupt RLE block");
            out.insert(out.end(), s + k, s + k + cnt);
            k += cnt;
        } else {
            size_t cnt = (size_t)c + 1;
            if (k >= n || out.size() + cnt > want) bad("image " + name + ": corrupt RLE block");
            out.insert(out.end(), cnt, s[k++]);
        }
    }
    if (out.size() != want) bad("image " + name + ": corrupt RLE block");
}

// ---- PIZ: canonical Huffman (code lengths <= 58, zero runs in the table, a
// run-length symbol), then the inverse 2D wavelet and the value LUT.
constexpr int kHufEncSize = (1 << 16) + 1;

struct BitIn {
    const unsigned char* p;
    const unsigned char* end;
    uint64_t c = 0;
    int lc = 0;
    bool get(int nbits, uint64_t* v) {   // MSB first
        while (lc < nbits) {
            if (p >= end) return false;
            c = (c << 8) | *p++;
            lc += 8;
        }
        lc -= nbits;
        *v = (c >> lc) & ((1ull << nbits) - 1);
        return true;
    }
};

void huf_decode(const unsigned char* in, size_t n, uint16_t* out, size_t nout, const std::string& name) {
    if (n == 0) {
        if (nout != 0) bad("image " + name + ": empty PIZ Huffman block");
        return;
    }
    if (n < 20) bad("image " + name + ": short PIZ Huffman block");
    auto rd32 = [&](size_t o) { return (uint32_t)in[o] | ((uint32_t)in[o + 1] << 8) | ((uint32_t)in[o + 2] << 16) | ((uint32_t)in[o + 3] << 24); };
    const uint32_t im = rd32(0), iM = rd32(4), nbits = rd32(12);
    if (im >= (uint32_t)kHufEncSize || iM >= (uint32_t)kHufEncSize || im > iM) bad("image " + name + ": bad PIZ Huffman range");
    // code lengths, 6 bits each; 59-62: 2-5 zeros; 63 + 8 bits: 6-261 zeros
    std::vector<uint8_t> len(kHufEncSize, 0);
    BitIn t{in + 20, in + n};
    for (uint32_t s = im; s <= iM; s++) {
        uint64_t l;
        if (!t.get(6, &l)) bad("image " + name + ": truncated PIZ code table");
        if (l == 63) {
            uint64_t z;
            if (!t.get(8, &z)) bad("image " + name + ": truncated PIZ code table");
            z += 6;
            if (s + z > iM + 1) bad("image " + name + ": bad PIZ code table");
            s += (uint32_t)z - 1;
        } else if (l >= 59) {
            uint32_t z = (uint32_t)l - 59 + 2;
            if (s + z > iM + 1) bad("image " + name + ": bad PIZ code table");
            s += z - 1;
        } else {
            len[s] = (uint8_t)l;
        }
    }
    const unsigned char* data = t.p;   // the code bits start at the next whole byte
    if ((uint64_t)nbits > 8ull * (uint64_t)(in + n - data)) bad("image " + name + ": PIZ bit count past the block");
    // canonical codes: from the longest length down, codes of one length
    // consecutive in symbol order
    uint64_t count[59] = {0}, start[59] = {0};
    for (int s = 0; s < kHufEncSize; s++) count[len[s]]++;
    uint64_t c = 0;
    for (int l = 58; l > 0; l--) {
        uint64_t nc = (c + count[l]) >> 1;
        start[l] = c;
        c = nc;
    }
    std::vector<std::vector<uint32_t>> syms(59);
    for (int s = 0; s < kHufEncSize; s++)
        if (len[s] > 0) syms[len[s]].push_back((uint32_t)s);
    BitIn b{data, in + n};
    uint64_t consumed = 0;
    size_t o = 0;
    uint64_t code = 0;
    int clen = 0;
    while (consumed < nbits && o < nout) {
        uint64_t bit;
        if (!b.get(1, &bit)) bad("image " + name + ": truncated PIZ data");
        consumed++;
        code = (code << 1) | bit;
        clen++;
        if (clen > 58) bad("image " + name + ": bad PIZ code");
        if (syms[clen].empty() || code < start[clen] || code - start[clen] >= syms[clen].size()) continue;
        const uint32_t sym = syms[clen][(size_t)(code - start[clen])];
        code = 0;
        clen = 0;
        if (sym == iM) {   // run: repeat the last value the next 8 bits' times
            uint64_t cs;
            if (consumed + 8 > nbits || !b.get(8, &cs)) bad("image " + name + ": truncated PIZ run");
            consumed += 8;
            if (o == 0 || o + cs > nout) bad("image " + name + ": bad PIZ run");
            const uint16_t v = out[o - 1];
            for (uint64_t k = 0; k < cs; k++) out[o++] = v;
        } else {
            out[o++] = (uint16_t)sym;
        }
    }
    if (o != nout) bad("image " + name + ": PIZ block decodes to the wrong size");
}

inline void wdec14(uint16_t l, uint16_t h, uint16_t& a, uint16_t& b) {
    const int16_t ls = (int16_t)l, hs = (int16_t)h;
    const int hi = hs;
    const int ai = ls + (hi & 1) + (hi >> 1);
    a = (uint16_t)(int16_t)ai;
    b = (uint16_t)(int16_t)(ai - hi);
}

inline void wdec16(uint16_t l, uint16_t h, uint16_t& a, uint16_t& b) {
    const int m = l, d = h;
    const int bb = (m - (d >> 1)) & 0xffff;
    const int aa = (d + bb - 0x8000) & 0xffff;
    b = (uint16_t)bb;
    a = (uint16_t)aa;
}

// Inverse of the PIZ wavelet over an nx x ny plane (element stride ox, row
// stride oy), coarsest level first.
void wav2_decode(uint16_t* in, int nx, int ox, int ny, int oy, uint16_t mx) {
    const bool w14 = mx < (1 << 14);
    const int n = std::min(nx, ny);
    int p = 1;
    while (p <= n) p <<= 1;
    p >>= 1;
    int p2 = p;
    p >>= 1;
    auto dec = [&](uint16_t l, uint16_t h, uint16_t& a, uint16_t& b) {
        if (w14) wdec14(l, h, a, b);
        else wdec16(l, h, a, b);
    };
    while (p >= 1) {
        uint16_t* py = in;
        uint16_t* ey = in + (ptrdiff_t)oy * (ny - p2);
        const int oy1 = oy * p, oy2 = oy * p2, ox1 = ox * p, ox2 = ox * p2;
        uint16_t i00, i01, i10, i11;
        for (; py <= ey; py += oy2) {
            uint16_t* px = py;
            uint16_t* ex = py + (ptrdiff_t)ox * (nx - p2);
            for (; px <= ex; px += ox2) {
                uint16_t* p01 = px + ox1;
                uint16_t* p10 = px + oy1;
                uint16_t* p11 = p10 + ox1;
                dec(*px, *p10, i00, i10);
                dec(*p01, *p11, i01, i11);
                dec(i00, i01, *px, *p01);
                dec(i10, i11, *p10, *p11);
            }
            if (nx & p) {
                uint16_t* p10 = px + oy1;
                dec(*px, *p10, i00, *p10);
                *px = i00;
            }
        }
        if (ny & p) {
            uint16_t* px = py;
            uint16_t* ex = py + (ptrdiff_t)ox * (nx - p2);
            for (; px <= ex; px += ox2) {
                uint16_t* p01 = px + ox1;
                dec(*px, *p01, i00, *p01);
                *px = i00;
            }
        }
        p2 = p;
        p >>= 1;
    }
}

// One PIZ block (`lines` scanlines of width w) into the scanline layout
// (per line, per channel, little-endian samples).
void piz_decode(const unsigned char* in, size_t n, std::vector<unsigned char>& out, size_t raw_size, int w, int lines,
                const std::vector<ExrChannel>& chans, const std::string& name) {
    In r{in, n};
    r.name = name;
    const uint16_t lo = (uint16_t)(r.u8() | (r.u8() << 8));
    const uint16_t hi = (uint16_t)(r.u8() | (r.u8() << 8));
    std::vector<uint8_t> bitmap(8192, 0);
    if (lo <= hi) {
        if (hi >= 8192) bad("image " + name + ": bad PIZ bitmap range");
        r.need((size_t)(hi - lo + 1));
        std::memcpy(bitmap.data() + lo, in + r.i, (size_t)(hi - lo + 1));
        r.i += (size_t)(hi - lo + 1);
    }
    std::vector<uint16_t> lut(65536, 0);
    int k = 0;
    for (int v = 0; v < 65536; v++)
        if (v == 0 || (bitmap[v >> 3] & (1 << (v & 7)))) lut[k++] = (uint16_t)v;
    const uint16_t maxv = (uint16_t)(k - 1);
    const int32_t hlen = r.i32();
    if (hlen < 0) bad("image " + name + ": bad PIZ length");
    r.need((size_t)hlen);
    const size_t nvals = raw_size / 2;
    std::vector<uint16_t> tmp(nvals);
    huf_decode(in + r.i, (size_t)hlen, tmp.data(), nvals, name);
    size_t startw = 0;
    for (const ExrChannel& c : chans) {
        const int words = c.bytes / 2;
        for (int j = 0; j < words; j++) wav2_decode(tmp.data() + startw + j, w, words, lines, w * words, maxv);
        startw += (size_t)w * lines * words;
    }
    for (uint16_t& v : tmp) v = lut[v];
    out.resize(raw_size);
    size_t o = 0;
    std::vector<size_t> plane(chans.size());
    size_t acc = 0;
    for (size_t ci = 0; ci < chans.size(); ci++) { plane[ci] = acc; acc += (size_t)w * lines * (chans[ci].bytes / 2); }
    for (int l = 0; l < lines; l++)
        for (size_t ci = 0; ci < chans.size(); ci++) {
            const size_t nw = (size_t)w * (chans[ci].bytes / 2);
            for (size_t q = 0; q < nw; q++) {
                const uint16_t v = tmp[plane[ci]++];
                out[o++] = (unsigned char)(v & 0xff);
                out[o++] = (unsigned char)(v >> 8);
            }
        }
}

RawImage decode_exr(const std::string& bytes, const std::string& name) {
    In r{(const unsigned char*)bytes.data(), bytes.size()};
    r.name = name;
    if (r.u32() != 20000630u) bad("image " + name + ": not an OpenEXR file");
    const uint32_t ver = r.u32();
    if ((ver & 0xff) != 2) unsupported("image " + name + ": OpenEXR version " + std::to_string(ver & 0xff));
    if (ver & 0x200) unsupported("image " + name + ": tiled OpenEXR");
    if (ver & 0x800) unsupported("image " + name + ": deep OpenEXR");
    if (ver & 0x1000) unsupported("image " + name + ": multi-part OpenEXR");
    std::vector<ExrChannel> chans;
    int compression = -1;
    int32_t xmin = 0, ymin = 0, xmax = -1, ymax = -1;
    bool have_dw = false;
    for (;;) {
        std::string an = r.str();
        if (an.empty()) break;
        std::string at = r.str();
        const int32_t size = r.i32();
        if (size < 0) bad("image " + name + ": bad OpenEXR attribute size");
        r.need((size_t)size);
        const size_t end = r.i + (size_t)size;
        if (an == "channels" && at == "chlist") {
            while (r.i < end) {
                std::string cn = r.str();
                if (cn.empty()) break;
                ExrChannel c;
                c.name = cn;
                c.type = r.i32();
                r.u8(); r.u8(); r.u8(); r.u8();   // pLinear, reserved
                const int32_t xs = r.i32(), ys = r.i32();
                if (c.type < 0 || c.type > 2) bad("image " + name + ": bad OpenEXR pixel type");
                if (xs != 1 || ys != 1) unsupported("image " + name + ": subsampled OpenEXR channel " + cn);
                c.bytes = c.type == 1 ? 2 : 4;
                chans.push_back(c);
            }
        } else if (an == "compression" && at == "compression") {
            compression = r.u8();
        } else if (an == "dataWindow" && at == "box2i") {
            xmin = r.i32(); ymin = r.i32(); xmax = r.i32(); ymax = r.i32();
            have_dw = true;
        }
        r.i = end;
    }
    if (chans.empty() || compression < 0 || !have_dw) bad("image " + name + ": OpenEXR header lacks channels, compression or dataWindow");
    const int64_t W = (int64_t)xmax - xmin + 1, H = (int64_t)ymax - ymin + 1;
    if (W <= 0 || H <= 0 || W * H > (1 << 28)) bad("image " + name + ": bad OpenEXR data window");
    int lpb;
    switch (compression) {
        case kExrNone: case kExrRle: case kExrZips: lpb = 1; break;
        case kExrZip: lpb = 16; break;
        case kExrPiz: lpb = 32; break;
        default: unsupported("image " + name + ": OpenEXR compression " + std::to_string(compression));
    }
    const int w = (int)W, h = (int)H;
    size_t line_bytes = 0;
    for (const ExrChannel& c : chans) line_bytes += (size_t)w * c.bytes;
    const int nblocks = (h + lpb - 1) / lpb;
    std::vector<uint64_t> offs(nblocks);
    for (int b = 0; b < nblocks; b++) offs[b] = r.u64();
    // the channels Magick's RGBA interface reads
    int ci_r = -1, ci_g = -1, ci_b = -1, ci_a = -1, ci_y = -1;
    for (size_t k = 0; k < chans.size(); k++) {
        const std::string& cn = chans[k].name;
        if (cn == "R") ci_r = (int)k;
        else if (cn == "G") ci_g = (int)k;
        else if (cn == "B") ci_b = (int)k;
        else if (cn == "A") ci_a = (int)k;
        else if (cn == "Y") ci_y = (int)k;
    }
    if (ci_r < 0 && ci_g < 0 && ci_b < 0 && ci_y < 0) unsupported("image " + name + ": OpenEXR without R, G, B or Y channels");
    RawImage im;
    im.w = w; im.h = h; im.channels = 4;
    im.px.assign((size_t)w * h * 4, 0.0f);
    std::vector<bool> seen(h, false);
    std::vector<unsigned char> block, tmp;
    for (int b = 0; b < nblocks; b++) {
        r.i = (size_t)offs[b];
        const int32_t y = r.i32();
        const int32_t size = r.i32();
        if (y < ymin || y > ymax || (y - ymin) % lpb != 0 || size < 0) bad("image " + name + ": bad OpenEXR chunk");
        const int lines = std::min(lpb, ymax - y + 1);
        const size_t raw_size = line_bytes * lines;
        r.need((size_t)size);
        const unsigned char* data = r.p + r.i;
        if ((size_t)size >= raw_size || compression == kExrNone) {   // stored as is
            if ((size_t)size != raw_size) bad("image " + name + ": bad OpenEXR chunk size");
            block.assign(data, data + raw_size);
        } else if (compression == kExrRle) {
            exr_rle(data, (size_t)size, block, raw_size, name);
            exr_unpredict(block);
        } else if (compression == kExrZips || compression == kExrZip) {
            block.resize(raw_size);
            uLongf len = raw_size;
            if (uncompress(block.data(), &len, data, (uLong)size) != Z_OK || len != raw_size)
                bad("image " + name + ": corrupt ZIP chunk");
            exr_unpredict(block);
        } else {
            piz_decode(data, (size_t)size, block, raw_size, w, lines, chans, name);
        }
        // per line, per channel, w samples
        size_t o = 0;
        for (int l = 0; l < lines; l++) {
            const int row = y - ymin + l;
            seen[row] = true;
            for (size_t ci = 0; ci < chans.size(); ci++) {
                const ExrChannel& c = chans[ci];
                int slot = -1;
                if ((int)ci == ci_r) slot = 0;
                else if ((int)ci == ci_g) slot = 1;
                else if ((int)ci == ci_b) slot = 2;
                else if ((int)ci == ci_a) slot = 3;
                else if ((int)ci == ci_y && ci_r < 0 && ci_g < 0 && ci_b < 0) slot = 4;   // luminance only
                for (int x = 0; x < w; x++) {
                    const unsigned char* s = block.data() + o + (size_t)x * c.bytes;
                    if (slot < 0) continue;
                    float v;
                    if (c.type == 1) {
                        v = half_to_float((uint16_t)(s[0] | (s[1] << 8)));
                    } else {
                        uint32_t u = s[0] | (s[1] << 8) | (s[2] << 16) | ((uint32_t)s[3] << 24);
                        float f;
                        if (c.type == 2) std::memcpy(&f, &u, 4);
                        else f = (float)u;
                        v = half_to_float(float_to_half(f));   // through the half RGBA interface
                    }
                    float* p = im.px.data() + ((size_t)row * w + x) * 4;
                    const float q = via_quantum(v);
                    if (slot == 4) p[0] = p[1] = p[2] = q;
                    else p[slot] = q;
                }
                o += (size_t)w * c.bytes;
            }
        }
    }
    for (int row = 0; row < h; row++)
        if (!seen[row]) bad("image " + name + ": OpenEXR file misses scanlines");
    if (ci_a < 0)
        for (size_t k = 0; k < (size_t)w * h; k++) im.px[k * 4 + 3] = via_quantum(1.0);
    return im;
}

// ---------------------------------------------------------------- Radiance RGBE
RawImage decode_hdr(const std::string& bytes, const std::string& name) {
    const unsigned char* d = (const unsigned char*)bytes.data();
    const size_t n = bytes.size();
    size_t i = 0;
    auto line = [&]() {
        size_t s = i;
        while (i < n && d[i] != '\n') i++;
        if (i >= n) bad("image " + name + ": truncated Radiance header");
        std::string l((const char*)d + s, i - s);
        i++;
        return l;
    };
    std::string first = line();
    if (first.compare(0, 2, "#?") != 0) bad("image " + name + ": not a Radiance file");
    for (;;) {
        std::string l = line();
        if (l.empty()) break;
        if (l.compare(0, 7, "FORMAT=") == 0 && l != "FORMAT=32-bit_rle_rgbe")
            unsupported("image " + name + ": Radiance " + l);
    }
    std::string res = line();
    char ya[3] = {0}, xa[3] = {0};
    int h = 0, w = 0;
    if (std::sscanf(res.c_str(), "%2s %d %2s %d", ya, &h, xa, &w) != 4) bad("image " + name + ": bad Radiance resolution");
    const bool top_first = std::string(ya) == "-Y";
    if (!(top_first || std::string(ya) == "+Y") || std::string(xa) != "+X")
        unsupported("image " + name + ": Radiance orientation " + res);
    if (w <= 0 || h <= 0 || (size_t)w * h > (1u << 28)) bad("image " + name + ": bad Radiance size");
    RawImage im;
    im.w = w; im.h = h; im.channels = 3;
    im.px.assign((size_t)w * h * 3, 0.0f);
    std::vector<unsigned char> sl((size_t)w * 4);
    for (int y = 0; y < h; y++) {
        auto byte = [&]() -> unsigned char {
            if (i >= n) bad("image " + name + ": truncated Radiance data");
            return d[i++];
        };
        if (w >= 8 && w < 0x8000 && i + 4 <= n && d[i] == 2 && d[i + 1] == 2 && !(d[i + 2] & 0x80)) {
            if (((d[i + 2] << 8) | d[i + 3]) != w) bad("image " + name + ": Radiance scanline width");
            i += 4;
            for (int c = 0; c < 4; c++) {   // each component run-length coded on its own
                int x = 0;
                while (x < w) {
                    int cnt = byte();
                    if (cnt > 128) {
                        cnt -= 128;
                        if (x + cnt > w) bad("image " + name + ": bad Radiance run");
                        unsigned char v = byte();
                        for (int k = 0; k < cnt; k++) sl[(size_t)(x++) * 4 + c] = v;
                    } else {
                        if (cnt == 0 || x + cnt > w) bad("image " + name + ": bad Radiance run");
                        for (int k = 0; k < cnt; k++) sl[(size_t)(x++) * 4 + c] = byte();
                    }
                }
            }
        } else {   // flat pixels
            for (int x = 0; x < w; x++)
                for (int c = 0; c < 4; c++) sl[(size_t)x * 4 + c] = byte();
        }
        const int row = top_first ? y : h - 1 - y;
        for (int x = 0; x < w; x++) {
            const unsigned char* p = sl.data() + (size_t)x * 4;
            float* o = im.px.data() + ((size_t)row * w + x) * 3;
            if (p[3] != 0) {
                // Magick: QuantumRange * 2^(e - 136) * mantissa, as a float quantum
                const double g = std::pow(2.0, (double)p[3] - (128.0 + 8.0));
                for (int c = 0; c < 3; c++) o[c] = (float)(65535.0 * g * (double)p[c]) / 65535.0f;
            }
        }
    }
    return im;
}

// ---------------------------------------------------------------- PFM
RawImage decode_pfm(const std::string& bytes, const std::string& name) {
    const char* d = bytes.data();
    const size_t n = bytes.size();
    size_t i = 0;
    auto token = [&]() {
        while (i < n && std::isspace((unsigned char)d[i])) i++;
        size_t s = i;
        while (i < n && !std::isspace((unsigned char)d[i])) i++;
        if (s == i) bad("image " + name + ": truncated PFM header");
        return std::string(d + s, i - s);
    };
    std::string magic = token();
    const int ch = magic == "PF" ? 3 : (magic == "Pf" ? 1 : 0);
    if (!ch) bad("image " + name + ": not a PFM file");
    const int w = std::atoi(token().c_str()), h = std::atoi(token().c_str());
    const double scale = std::atof(token().c_str());
    if (i >= n) bad("image " + name + ": truncated PFM header");
    i++;   // the single whitespace byte before the samples
    if (w <= 0 || h <= 0 || (size_t)w * h > (1u << 28) || scale == 0.0) bad("image " + name + ": bad PFM header");
    const bool little = scale < 0;
    if ((size_t)w * h * ch * 4 > n - i) bad("image " + name + ": truncated PFM data");
    RawImage im;
    im.w = w; im.h = h; im.channels = ch == 3 ? 3 : 1;
    im.px.assign((size_t)w * h * im.channels, 0.0f);
    const unsigned char* s = (const unsigned char*)d + i;
    for (int fy = 0; fy < h; fy++) {
        const int row = h - 1 - fy;   // PFM rows run bottom to top
        for (int x = 0; x < w; x++)
            for (int c = 0; c < ch; c++) {
                const unsigned char* b = s + (((size_t)fy * w + x) * ch + c) * 4;
                uint32_t u = little ? (b[0] | (b[1] << 8) | (b[2] << 16) | ((uint32_t)b[3] << 24))
                                    : (((uint32_t)b[0] << 24) | (b[1] << 16) | (b[2] << 8) | b[3]);
                float f;
                std::memcpy(&f, &u, 4);
                im.px[((size_t)row * w + x) * im.channels + c] = via_quantum(f);
            }
    }
    return im;
}

// ---------------------------------------------------------------- writers
void put32(std::string& s, uint32_t v) { for (int k = 0; k < 4; k++) s.push_back((char)((v >> (8 * k)) & 0xff)); }
void put64(std::string& s, uint64_t v) { put32(s, (uint32_t)v); put32(s, (uint32_t)(v >> 32)); }
void attr(std::string& s, const char* name, const char* type, const std::string& val) {
    s += name; s.push_back(0);
    s += type; s.push_back(0);
    put32(s, (uint32_t)val.size());
    s += val;
}

void write_file(const char* path, const std::string& data) {
    std::FILE* f = std::fopen(path, "wb");
    if (!f) throw ImageError(RD_ERR_IO, std::string("cannot open ") + path + " for writing");
    const size_t k = std::fwrite(data.data(), 1, data.size(), f);
    const bool ok = std::fclose(f) == 0 && k == data.size();
    if (!ok) throw ImageError(RD_ERR_IO, std::string("cannot write ") + path);
}

// ColorBuffer.GetRGB (Radiantium.Core/ColorBuffer.cs:44-62) of a buffer in
// ColorBuffer layout ((x * height + y) * channels).
void get_rgb(const float* buf, int height, int channels, size_t npx, int x, int y, float* o) {
    const size_t st = ((size_t)x * height + y) * channels;
    if (channels >= 3) {
        for (int c = 0; c < 3; c++) o[c] = buf[st + c];
        return;
    }
    for (int c = 0; c < 3; c++) {   // as written: `_channel >= i ? 0 : _buffer[start + i]`
        const size_t k = st + c;
        o[c] = channels >= c ? 0.0f : (k < npx * channels ? buf[k] : 0.0f);
    }
}

}  // namespace

float half_to_float(uint16_t h) {
    const uint32_t s = (uint32_t)(h >> 15) << 31, e = (h >> 10) & 31, m = h & 1023;
    if (e == 0) {
        const float v = (float)m * 5.9604644775390625e-8f;   // m * 2^-24, exact
        return s ? -v : v;
    }
    uint32_t bits = e == 31 ? (s | 0x7f800000u | (m << 13)) : (s | ((e + 112) << 23) | (m << 13));
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

uint16_t float_to_half(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint16_t s = (uint16_t)((x >> 16) & 0x8000);
    const uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return (uint16_t)(s | 0x7c00 | (ax > 0x7f800000u ? (0x200 | ((ax >> 13) & 0x3ff)) : 0));
    if (ax >= 0x477ff000u) return (uint16_t)(s | 0x7c00);   // rounds to 65520 or more: infinity
    if (ax < 0x38800000u) {   // below 2^-14: a subnormal half (or zero), m = round(v * 2^24)
        if (ax < 0x33000000u) return s;
        const uint32_t eb = ax >> 23;
        const uint32_t mant = (ax & 0x7fffffu) | 0x800000u;
        const int shift = 126 - (int)eb;
        uint32_t m = mant >> shift;
        const uint32_t rem = mant & ((1u << shift) - 1), halfway = 1u << (shift - 1);
        if (rem > halfway || (rem == halfway && (m & 1))) m++;
        return (uint16_t)(s | m);
    }
    uint32_t r = ax - (112u << 23);
    r += 0xfffu + ((r >> 13) & 1u);
    return (uint16_t)(s | (r >> 13));
}

RawImage decode_image(const std::string& bytes, const std::string& name) {
    const unsigned char* d = (const unsigned char*)bytes.data();
    const size_t n = bytes.size();
    static const unsigned char png_sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (n >= 8 && std::memcmp(d, png_sig, 8) == 0) return decode_png(bytes, name);
    if (n >= 4 && d[0] == 0x76 && d[1] == 0x2f && d[2] == 0x31 && d[3] == 0x01) return decode_exr(bytes, name);
    if (n >= 2 && d[0] == '#' && d[1] == '?') return decode_hdr(bytes, name);
    if (n >= 2 && d[0] == 'P' && (d[1] == 'F' || d[1] == 'f')) return decode_pfm(bytes, name);
    unsupported("image " + name + ": not a PNG, OpenEXR, Radiance or PFM file");
}

std::string encode_exr(const float* buf, int width, int height, int channels, int compression) {
    if (compression != kExrNone && compression != kExrZip) throw ImageError(RD_ERR_INVALID, "EXR compression must be NONE (0) or ZIP (3)");
    const size_t npx = (size_t)width * height;
    // Magick's writer: half(float(QuantumScale * quantum)), quantum = color * Quantum.Max;
    // rows as ColorBuffer rows (SaveOpenExr's `k = i`), opaque alpha
    auto half_of = [](float c) {
        const float q = c * 65535.0f;
        return float_to_half((float)((1.0 / 65535.0) * (double)q));
    };
    const uint16_t one = float_to_half((float)((1.0 / 65535.0) * 65535.0));
    std::string hdr;
    put32(hdr, 20000630u);
    put32(hdr, 2u);
    std::string ch;
    for (const char* cn : {"A", "B", "G", "R"}) {
        ch += cn; ch.push_back(0);
        put32(ch, 1u);                               // HALF
        ch.push_back(0); ch.push_back(0); ch.push_back(0); ch.push_back(0);   // pLinear, reserved
        put32(ch, 1u); put32(ch, 1u);                // sampling
    }
    ch.push_back(0);
    attr(hdr, "channels", "chlist", ch);
    attr(hdr, "compression", "compression", std::string(1, (char)compression));
    std::string box;
    put32(box, 0); put32(box, 0); put32(box, (uint32_t)(width - 1)); put32(box, (uint32_t)(height - 1));
    attr(hdr, "dataWindow", "box2i", box);
    attr(hdr, "displayWindow", "box2i", box);
    attr(hdr, "lineOrder", "lineOrder", std::string(1, (char)0));
    std::string one_f;
    float onef = 1.0f, zero = 0.0f;
    uint32_t u;
    std::memcpy(&u, &onef, 4); put32(one_f, u);
    attr(hdr, "pixelAspectRatio", "float", one_f);
    std::string v2;
    std::memcpy(&u, &zero, 4); put32(v2, u); put32(v2, u);
    attr(hdr, "screenWindowCenter", "v2f", v2);
    attr(hdr, "screenWindowWidth", "float", one_f);
    hdr.push_back(0);
    const int lpb = compression == kExrZip ? 16 : 1;
    const int nblocks = (height + lpb - 1) / lpb;
    std::vector<std::string> chunks(nblocks);
    std::vector<float> rgb(3);
    for (int b = 0; b < nblocks; b++) {
        const int y0 = b * lpb, lines = std::min(lpb, height - y0);
        std::vector<unsigned char> raw;
        raw.reserve((size_t)lines * width * 8);
        for (int l = 0; l < lines; l++) {
            const int y = y0 + l;
            std::vector<uint16_t> a(width, one), bb(width), g(width), r(width);
            for (int x = 0; x < width; x++) {
                get_rgb(buf, height, channels, npx, x, y, rgb.data());
                r[x] = half_of(rgb[0]); g[x] = half_of(rgb[1]); bb[x] = half_of(rgb[2]);
            }
            for (const std::vector<uint16_t>* c : {&a, &bb, &g, &r})
                for (uint16_t v : *c) { raw.push_back((unsigned char)(v & 0xff)); raw.push_back((unsigned char)(v >> 8)); }
        }
        std::string data;
        if (compression == kExrZip) {
            // interleave + predictor (the inverse of exr_unpredict), then zlib
            std::vector<unsigned char> t(raw.size());
            const size_t half = (raw.size() + 1) / 2;
            for (size_t k = 0; k < raw.size(); k++) {
                if (k & 1) t[half + k / 2] = raw[k];
                else t[k / 2] = raw[k];
            }
            for (size_t k = t.size(); k-- > 1;) t[k] = (unsigned char)(t[k] - t[k - 1] + 128);
            uLongf clen = compressBound(t.size());
            std::vector<unsigned char> z(clen);
            if (compress2(z.data(), &clen, t.data(), t.size(), Z_DEFAULT_COMPRESSION) != Z_OK)
                throw ImageError(RD_ERR_IO, "zlib failed");
            if (clen < raw.size()) data.assign((const char*)z.data(), clen);
            else data.assign((const char*)raw.data(), raw.size());
        } else {
            data.assign((const char*)raw.data(), raw.size());
        }
        std::string c;
        put32(c, (uint32_t)y0);
        put32(c, (uint32_t)data.size());
        c += data;
        chunks[b] = c;
    }
    std::string out = hdr;
    uint64_t off = hdr.size() + 8ull * nblocks;
    for (int b = 0; b < nblocks; b++) { put64(out, off); off += chunks[b].size(); }
    for (int b = 0; b < nblocks; b++) out += chunks[b];
    return out;
}

std::string encode_png(const float* buf, int width, int height, int channels, bool to_srgb) {
    const size_t npx = (size_t)width * height;
    // Png32 (8-bit RGBA); Magick's ScaleQuantumToChar of clamp(color * Quantum.Max, 0, Quantum.Max);
    // PNG row i is ColorBuffer row height - 1 - i (SavePng)
    std::vector<unsigned char> raw;
    raw.reserve((size_t)height * (1 + 4 * (size_t)width));
    float rgb[3];
    for (int i = 0; i < height; i++) {
        raw.push_back(0);
        const int k = height - i - 1;
        for (int x = 0; x < width; x++) {
            get_rgb(buf, height, channels, npx, x, k, rgb);
            for (int c = 0; c < 3; c++) {
                float v = rgb[c];
                if (to_srgb)   // Color3F.ToSrgb (Radiantium.Core/Color3F.cs:91-94)
                    v = v <= 0.0031308f ? 12.92f * v : (1.0f + 0.055f) * std::pow(v, 1.0f / 2.4f) - 0.055f;
                float q = v * 65535.0f;
                q = std::min(std::max(q, 0.0f), 65535.0f);   // Math.Clamp (NaN stays NaN)
                unsigned char b;
                if (std::isnan(q) || q <= 0.0f) b = 0;
                else if (q / 257.0f >= 255.0f) b = 255;
                else b = (unsigned char)(q / 257.0f + 0.5f);
                raw.push_back(b);
            }
            raw.push_back(255);
        }
    }
    uLongf clen = compressBound(raw.size());
    std::vector<unsigned char> z(clen);
    if (compress2(z.data(), &clen, raw.data(), raw.size(), Z_DEFAULT_COMPRESSION) != Z_OK)
        throw ImageError(RD_ERR_IO, "zlib failed");
    std::string out("\x89PNG\r\n\x1a\n", 8);
    auto chunk = [&](const char* type, const std::string& data) {
        std::string c;
        for (int k = 3; k >= 0; k--) c.push_back((char)((data.size() >> (8 * k)) & 0xff));
        std::string td = std::string(type, 4) + data;
        c += td;
        uint32_t crc = (uint32_t)crc32(0, (const Bytef*)td.data(), (uInt)td.size());
        for (int k = 3; k >= 0; k--) c.push_back((char)((crc >> (8 * k)) & 0xff));
        out += c;
    };
    std::string ihdr;
    for (uint32_t v : {(uint32_t)width, (uint32_t)height})
        for (int k = 3; k >= 0; k--) ihdr.push_back((char)((v >> (8 * k)) & 0xff));
    ihdr.push_back(8); ihdr.push_back(6); ihdr.push_back(0); ihdr.push_back(0); ihdr.push_back(0);
    chunk("IHDR", ihdr);
    chunk("IDAT", std::string((const char*)z.data(), clen));
    chunk("IEND", std::string());
    return out;
}

std::string encode_pfm(const float* buf, int width, int height, int channels) {
    const size_t npx = (size_t)width * height;
    char head[64];
    std::snprintf(head, sizeof(head), "PF\n%d %d\n-1.0\n", width, height);
    std::string out(head);
    float rgb[3];
    for (int y = 0; y < height; y++)   // PFM's first row is the bottom one: ColorBuffer row 0
        for (int x = 0; x < width; x++) {
            get_rgb(buf, height, channels, npx, x, y, rgb);
            for (int c = 0; c < 3; c++) {
                uint32_t u;
                std::memcpy(&u, &rgb[c], 4);
                put32(out, u);
            }
        }
    return out;
}

}  // namespace rd

// ---------------------------------------------------------------- C ABI
namespace {
thread_local std::string g_image_error;
int image_fail(int code, const std::string& m) {
    g_image_error = m;
    return code;
}
}  // namespace

extern "C" {

const char* rd_image_last_error(void) { return g_image_error.c_str(); }

int rd_image_load(const char* path, int need_channel, int flip_y, int cast_to_linear, rd_image* out) {
    if (!path || !out) return image_fail(RD_ERR_INVALID, "null argument");
    *out = rd_image{};
    try {
        std::FILE* f = std::fopen(path, "rb");
        if (!f) return image_fail(RD_ERR_IO, std::string("cannot open ") + path);
        std::string bytes;
        char tmp[1 << 16];
        size_t k;
        while ((k = std::fread(tmp, 1, sizeof(tmp), f)) > 0) bytes.append(tmp, k);
        std::fclose(f);
        rd::RawImage im = rd::decode_image(bytes, path);
        const int ch = need_channel > 0 ? need_channel : im.channels;
        float* data = (float*)std::malloc(sizeof(float) * (size_t)im.w * im.h * ch);
        if (!data) return image_fail(RD_ERR_OOM, "out of host memory");
        // LoadImageFromPath: ColorBuffer (x, y) from file row (flip ? h - 1 - y : y)
        for (int x = 0; x < im.w; x++)
            for (int y = 0; y < im.h; y++) {
                const int row = flip_y ? im.h - y - 1 : y;
                const float* s = im.px.data() + ((size_t)row * im.w + x) * im.channels;
                float* o = data + ((size_t)x * im.h + y) * ch;
                for (int a = 0; a < ch; a++) {
                    float v = a >= im.channels ? 0.0f : s[a];
                    if (cast_to_linear)   // Color3F.ToLinear (Radiantium.Core/Color3F.cs:86-89)
                        v = v <= 0.04045f ? v * (1.0f / 12.92f) : std::pow((v + 0.055f) * (1.0f / 1.055f), 2.4f);
                    o[a] = v;
                }
            }
        out->width = im.w;
        out->height = im.h;
        out->channels = ch;
        out->data = data;
        return RD_OK;
    } catch (const rd::ImageError& e) {
        return image_fail(e.code, e.what());
    } catch (const std::bad_alloc&) {
        return image_fail(RD_ERR_OOM, "out of host memory");
    }
}

void rd_image_free(rd_image* img) {
    if (!img) return;
    std::free(img->data);
    img->data = nullptr;
}

int rd_image_save_exr(const char* path, const float* data, int width, int height, int channels, int compression) {
    if (!path || !data || width <= 0 || height <= 0 || channels <= 0) return image_fail(RD_ERR_INVALID, "bad argument");
    try {
        rd::write_file(path, rd::encode_exr(data, width, height, channels, compression));
        return RD_OK;
    } catch (const rd::ImageError& e) {
        return image_fail(e.code, e.what());
    }
}

int rd_image_save_png(const char* path, const float* data, int width, int height, int channels, int to_srgb) {
    if (!path || !data || width <= 0 || height <= 0 || channels <= 0) return image_fail(RD_ERR_INVALID, "bad argument");
    try {
        rd::write_file(path, rd::encode_png(data, width, height, channels, to_srgb != 0));
        return RD_OK;
    } catch (const rd::ImageError& e) {
        return image_fail(e.code, e.what());
    }
}

int rd_image_save_pfm(const char* path, const float* data, int width, int height, int channels) {
    if (!path || !data || width <= 0 || height <= 0 || channels <= 0) return image_fail(RD_ERR_INVALID, "bad argument");
    try {
        rd::write_file(path, rd::encode_pfm(data, width, height, channels));
        return RD_OK;
    } catch (const rd::ImageError& e) {
        return image_fail(e.code, e.what());
    }
}

}  // extern "C"
