/*
 * vgi_image.cpp — PNG decoding and Lanczos3 mips for the glTF importer
 * (what each function restates is listed in vgi_image.h).
 */
#include "vgi_image.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace vgi {
namespace {

// ---------------------------------------------------------------------------
// DEFLATE (RFC 1951): canonical Huffman codes decoded one bit at a time.
// ---------------------------------------------------------------------------
struct BitIn {
    const uint8_t* p;
    size_t n, i = 0;
    uint32_t buf = 0;
    int cnt = 0;
    bool over = false;
    uint32_t bits(int k) {  // k <= 24, least significant bit first
        while (cnt < k) {
            if (i >= n) {
                over = true;
                return 0;
            }
            buf |= (uint32_t)p[i++] << cnt;
            cnt += 8;
        }
        const uint32_t v = buf & ((1u << k) - 1u);
        buf >>= k;
        cnt -= k;
        return v;
    }
    void to_byte() {  // drop the partial byte, give back the whole ones
        i -= (size_t)(cnt / 8);
        buf = 0;
        cnt = 0;
    }
};

struct Huffman {
    uint16_t count[16];
    uint16_t symbol[320];
};

// False on an over-subscribed code; incomplete codes are accepted (a
// missing code then fails when it is read).
bool huffman_build(Huffman& h, const uint8_t* len, int n) {
    memset(h.count, 0, sizeof(h.count));
    for (int s = 0; s < n; ++s) ++h.count[len[s]];
    if (h.count[0] == n) return true;
    int left = 1;
    for (int l = 1; l < 16; ++l) {
        left <<= 1;
        left -= h.count[l];
        if (left < 0) return false;
    }
    uint16_t off[16];
    off[1] = 0;
    for (int l = 1; l < 15; ++l) off[l + 1] = (uint16_t)(off[l] + h.count[l]);
    for (int s = 0; s < n; ++s)
        if (len[s]) h.symbol[off[len[s]]++] = (uint16_t)s;
    return true;
}

int huffman_read(BitIn& b, const Huffman& h) {
    int code = 0, first = 0, index = 0;
    for (int l = 1; l < 16; ++l) {
        code |= (int)b.bits(1);
        if (b.over) return -1;
        const int c = h.count[l];
        if (code - c < first) return h.symbol[index + (code - first)];
        index += c;
        first = (first + c) << 1;
        code <<= 1;
    }
    return -1;
}

// Length and distance bases of the format (RFC 1951 section 3.2.5).
constexpr uint16_t kLenBase[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                                   31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
constexpr uint8_t kLenExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
constexpr uint16_t kDistBase[30] = {1,   2,   3,   4,   5,   7,    9,    13,   17,   25,   33,   49,   65,    97,    129,
                                    193, 257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
constexpr uint8_t kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};

bool inflate_codes(BitIn& b, const Huffman& lit, const Huffman& dist, std::vector<uint8_t>& out, size_t cap,
                   std::string& err) {
    for (;;) {
        const int s = huffman_read(b, lit);
        if (s < 0) return err = "bad literal/length code", false;
        if (s < 256) {
            if (out.size() >= cap) return err = "more data than the image holds", false;
            out.push_back((uint8_t)s);
            continue;
        }
        if (s == 256) return true;
        const int li = s - 257;
        if (li >= 29) return err = "bad length symbol", false;
        const size_t len = kLenBase[li] + b.bits(kLenExtra[li]);
        const int ds = huffman_read(b, dist);
        if (ds < 0 || ds >= 30) return err = "bad distance code", false;
        const size_t d = kDistBase[ds] + b.bits(kDistExtra[ds]);
        if (b.over) return err = "truncated stream", false;
        if (d > out.size()) return err = "distance before the start", false;
        if (out.size() + len > cap) return err = "more data than the image holds", false;
        const size_t from = out.size() - d;
        for (size_t k = 0; k < len; ++k) out.push_back(out[from + k]);
    }
}

bool inflate(BitIn& b, std::vector<uint8_t>& out, size_t cap, std::string& err) {
    static const uint8_t kOrder[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
    Huffman lit, dist;
    int last = 0;
    do {
        last = (int)b.bits(1);
        const uint32_t type = b.bits(2);
        if (b.over) return err = "truncated stream", false;
        if (type == 0) {
            b.to_byte();
            if (b.i + 4 > b.n) return err = "truncated stored block", false;
            const uint32_t len = b.p[b.i] | (b.p[b.i + 1] << 8), nlen = b.p[b.i + 2] | (b.p[b.i + 3] << 8);
            b.i += 4;
            if ((len ^ 0xFFFFu) != nlen) return err = "stored block length check", false;
            if (b.i + len > b.n) return err = "truncated stored block", false;
            if (out.size() + len > cap) return err = "more data than the image holds", false;
            out.insert(out.end(), b.p + b.i, b.p + b.i + len);
            b.i += len;
        } else if (type == 1) {
            uint8_t l[320];
            for (int s = 0; s < 144; ++s) l[s] = 8;
            for (int s = 144; s < 256; ++s) l[s] = 9;
            for (int s = 256; s < 280; ++s) l[s] = 7;
            for (int s = 280; s < 288; ++s) l[s] = 8;
            huffman_build(lit, l, 288);
            for (int s = 0; s < 30; ++s) l[s] = 5;
            huffman_build(dist, l, 30);
            if (!inflate_codes(b, lit, dist, out, cap, err)) return false;
        } else if (type == 2) {
            const int nlit = (int)b.bits(5) + 257, ndist = (int)b.bits(5) + 1, ncode = (int)b.bits(4) + 4;
            if (b.over || nlit > 286 || ndist > 30) return err = "bad dynamic block header", false;
            uint8_t l[320] = {0};
            for (int k = 0; k < ncode; ++k) l[kOrder[k]] = (uint8_t)b.bits(3);
            Huffman lens;
            if (b.over || !huffman_build(lens, l, 19)) return err = "bad code-length code", false;
            uint8_t ll[320] = {0};
            int k = 0;
            while (k < nlit + ndist) {
                int s = huffman_read(b, lens);
                if (s < 0) return err = "bad code length", false;
                if (s < 16) {
                    ll[k++] = (uint8_t)s;
                    continue;
                }
                uint8_t v = 0;
                int rep;
                if (s == 16) {
                    if (k == 0) return err = "repeat with no previous length", false;
                    v = ll[k - 1];
                    rep = 3 + (int)b.bits(2);
                } else if (s == 17) {
                    rep = 3 + (int)b.bits(3);
                } else {
                    rep = 11 + (int)b.bits(7);
                }
                if (b.over || k + rep > nlit + ndist) return err = "code lengths overrun", false;
                while (rep--) ll[k++] = v;
            }
            if (ll[256] == 0) return err = "no end-of-block code", false;
            if (!huffman_build(lit, ll, nlit) || !huffman_build(dist, ll + nlit, ndist))
                return err = "over-subscribed code", false;
            if (!inflate_codes(b, lit, dist, out, cap, err)) return false;
        } else {
            return err = "reserved block type", false;
        }
    } while (!last);
    return true;
}

// zlib (RFC 1950) around DEFLATE, with the Adler-32 check.
bool zlib_decompress(const uint8_t* p, size_t n, std::vector<uint8_t>& out, size_t cap, std::string& err) {
    if (n < 6) return err = "zlib stream too short", false;
    const uint32_t cmf = p[0], flg = p[1];
    if ((cmf & 15u) != 8u || (cmf >> 4) > 7u || ((cmf << 8) | flg) % 31u != 0 || (flg & 0x20u))
        return err = "bad zlib header", false;
    BitIn b{p + 2, n - 2};
    if (!inflate(b, out, cap, err)) return false;
    b.to_byte();
    if (b.i + 4 > b.n) return err = "missing Adler-32", false;
    uint32_t a = 1, c = 0;
    for (uint8_t v : out) {
        a = (a + v) % 65521u;
        c = (c + a) % 65521u;
    }
    const uint32_t want = ((uint32_t)b.p[b.i] << 24) | ((uint32_t)b.p[b.i + 1] << 16) | ((uint32_t)b.p[b.i + 2] << 8) |
                          b.p[b.i + 3];
    if (((c << 16) | a) != want) return err = "Adler-32 mismatch", false;
    return true;
}

uint32_t be32(const uint8_t* p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }

uint32_t crc32(const uint8_t* p, size_t n) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t k = 0; k < 256; ++k) {
            uint32_t c = k;
            for (int j = 0; j < 8; ++j) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[k] = c;
        }
        init = true;
    }
    uint32_t c = 0xFFFFFFFFu;
    for (size_t k = 0; k < n; ++k) c = table[(c ^ p[k]) & 0xFFu] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

uint8_t paeth(uint8_t a, uint8_t b, uint8_t c) {
    const int p = (int)a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    return pa <= pb && pa <= pc ? a : (pb <= pc ? b : c);
}

}  // namespace

bool is_png(const uint8_t* d, size_t n) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    return n >= 8 && memcmp(d, sig, 8) == 0;
}

bool decode_png(const uint8_t* d, size_t n, Rgba8Image& out, std::string& err, uint64_t max_texels) {
    if (!is_png(d, n)) return err = "not a PNG file", false;
    size_t pos = 8;
    uint32_t w = 0, h = 0, depth = 0, ctype = 0, interlace = 0;
    bool have_hdr = false, have_end = false;
    uint8_t plte[256][3] = {};
    uint32_t nplte = 0;
    uint8_t trns_a[256];
    memset(trns_a, 255, sizeof(trns_a));
    uint16_t key[3] = {0, 0, 0};
    bool have_key = false;
    std::vector<uint8_t> idat;
    while (pos + 12 <= n && !have_end) {
        const uint32_t len = be32(d + pos);
        const uint8_t* type = d + pos + 4;
        if (len > 0x7FFFFFFFu || pos + 12 + (size_t)len > n) return err = "truncated chunk", false;
        const uint8_t* data = d + pos + 8;
        if (crc32(type, 4 + (size_t)len) != be32(data + len)) return err = "chunk CRC mismatch", false;
        pos += 12 + (size_t)len;
        if (!have_hdr && memcmp(type, "IHDR", 4) != 0) return err = "IHDR is not first", false;
        if (memcmp(type, "IHDR", 4) == 0) {
            if (have_hdr || len != 13) return err = "bad IHDR", false;
            w = be32(data);
            h = be32(data + 4);
            depth = data[8];
            ctype = data[9];
            interlace = data[12];
            const bool ok_depth = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                                  (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                                  ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
            if (!ok_depth || data[10] != 0 || data[11] != 0 || interlace > 1) return err = "unsupported IHDR", false;
            if (w == 0 || h == 0 || w > 0x7FFFFFFFu || h > 0x7FFFFFFFu || (uint64_t)w * h > max_texels)
                return err = "image size out of range", false;
            have_hdr = true;
        } else if (memcmp(type, "PLTE", 4) == 0) {
            if (len % 3 != 0 || len / 3 > 256 || len == 0) return err = "bad PLTE", false;
            nplte = len / 3;
            memcpy(plte, data, len);
        } else if (memcmp(type, "tRNS", 4) == 0) {
            if (ctype == 3) {
                if (len > nplte) return err = "bad tRNS", false;
                memcpy(trns_a, data, len);
            } else if (ctype == 0 && len == 2) {
                key[0] = (uint16_t)((data[0] << 8) | data[1]);
                have_key = true;
            } else if (ctype == 2 && len == 6) {
                for (int c = 0; c < 3; ++c) key[c] = (uint16_t)((data[2 * c] << 8) | data[2 * c + 1]);
                have_key = true;
            } else {
                return err = "bad tRNS", false;
            }
        } else if (memcmp(type, "IDAT", 4) == 0) {
            idat.insert(idat.end(), data, data + len);
        } else if (memcmp(type, "IEND", 4) == 0) {
            have_end = true;
        } else if (!(type[0] & 0x20)) {
            return err = "unknown critical chunk", false;
        }
    }
    if (!have_hdr || !have_end) return err = "missing IHDR or IEND", false;
    if (ctype == 3 && nplte == 0) return err = "palette image without PLTE", false;
    const uint32_t channels = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : 4;
    const uint32_t bpp_bits = channels * depth;
    const uint32_t fbpp = (bpp_bits + 7) / 8;  // filter distance in bytes
    struct Pass {
        uint32_t x0, y0, dx, dy;
    };
    static const Pass kAdam7[7] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                   {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    static const Pass kWhole[1] = {{0, 0, 1, 1}};
    const Pass* passes = interlace ? kAdam7 : kWhole;
    const int npass = interlace ? 7 : 1;
    uint64_t need = 0;
    for (int k = 0; k < npass; ++k) {
        const Pass& P = passes[k];
        const uint64_t pw = w > P.x0 ? (w - P.x0 + P.dx - 1) / P.dx : 0, ph = h > P.y0 ? (h - P.y0 + P.dy - 1) / P.dy : 0;
        if (pw && ph) need += ph * (1 + (pw * bpp_bits + 7) / 8);
    }
    std::vector<uint8_t> raw;
    raw.reserve((size_t)need);
    if (!zlib_decompress(idat.data(), idat.size(), raw, (size_t)need, err)) return false;
    if (raw.size() != need) return err = "image data shorter than the image", false;
    out.width = w;
    out.height = h;
    out.texels.assign((size_t)w * h * 4, 0);
    const uint32_t maxv = (1u << depth) - 1u;
    auto sample = [&](const uint8_t* row, uint64_t idx) -> uint32_t {  // sample idx of a row
        if (depth == 8) return row[idx];
        if (depth == 16) return ((uint32_t)row[2 * idx] << 8) | row[2 * idx + 1];
        const uint64_t bit = idx * depth;
        return (row[bit >> 3] >> (8 - depth - (bit & 7))) & maxv;
    };
    auto to8 = [&](uint32_t v) -> uint8_t {
        if (depth == 8) return (uint8_t)v;
        if (depth == 16) return (uint8_t)((2 * v + 257) / 514);  // round(v / 257)
        return (uint8_t)(v * 255u / maxv);
    };
    size_t at = 0;
    std::vector<uint8_t> prev, cur;
    for (int k = 0; k < npass; ++k) {
        const Pass& P = passes[k];
        const uint32_t pw = w > P.x0 ? (w - P.x0 + P.dx - 1) / P.dx : 0, ph = h > P.y0 ? (h - P.y0 + P.dy - 1) / P.dy : 0;
        if (!pw || !ph) continue;
        const size_t rb = ((size_t)pw * bpp_bits + 7) / 8;
        prev.assign(rb, 0);
        cur.resize(rb);
        for (uint32_t y = 0; y < ph; ++y) {
            const uint8_t f = raw[at++];
            const uint8_t* src = raw.data() + at;
            at += rb;
            for (size_t i = 0; i < rb; ++i) {
                const uint8_t a = i >= fbpp ? cur[i - fbpp] : 0, b = prev[i], c = i >= fbpp ? prev[i - fbpp] : 0;
                uint8_t v = src[i];
                switch (f) {
                    case 0: break;
                    case 1: v = (uint8_t)(v + a); break;
                    case 2: v = (uint8_t)(v + b); break;
                    case 3: v = (uint8_t)(v + ((a + b) >> 1)); break;
                    case 4: v = (uint8_t)(v + paeth(a, b, c)); break;
                    default: return err = "bad filter type", false;
                }
                cur[i] = v;
            }
            const uint32_t oy = P.y0 + y * P.dy;
            for (uint32_t x = 0; x < pw; ++x) {
                uint8_t* o = out.texels.data() + 4 * ((size_t)oy * w + P.x0 + (size_t)x * P.dx);
                const uint64_t s0 = (uint64_t)x * channels;
                switch (ctype) {
                    case 0: {
                        const uint32_t g = sample(cur.data(), s0);
                        o[0] = o[1] = o[2] = to8(g);
                        o[3] = have_key && g == key[0] ? 0 : 255;
                        break;
                    }
                    case 2: {
                        const uint32_t r = sample(cur.data(), s0), g = sample(cur.data(), s0 + 1),
                                       b = sample(cur.data(), s0 + 2);
                        o[0] = to8(r), o[1] = to8(g), o[2] = to8(b);
                        o[3] = have_key && r == key[0] && g == key[1] && b == key[2] ? 0 : 255;
                        break;
                    }
                    case 3: {
                        const uint32_t i = sample(cur.data(), s0);
                        if (i >= nplte) return err = "palette index out of range", false;
                        o[0] = plte[i][0], o[1] = plte[i][1], o[2] = plte[i][2], o[3] = trns_a[i];
                        break;
                    }
                    case 4:
                        o[0] = o[1] = o[2] = to8(sample(cur.data(), s0));
                        o[3] = to8(sample(cur.data(), s0 + 1));
                        break;
                    default:
                        for (int c = 0; c < 4; ++c) o[c] = to8(sample(cur.data(), s0 + c));
                        break;
                }
            }
            std::swap(prev, cur);
            cur.resize(rb);
        }
    }
    return true;
}

// ---------------------------------------------------------------------------
// Baseline JPEG (ITU-T T.81, sequential DCT with Huffman coding)
// ---------------------------------------------------------------------------
namespace {

// zig-zag scan position -> natural (row-major) coefficient index (T.81 A.3.6)
constexpr uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                                 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                                 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                                 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct JHuff {
    bool present = false;
    uint8_t vals[256];
    int32_t mincode[17], maxcode[18], valptr[17];  // T.81 F.2.2.3 decoding tables
};

struct JComp {
    uint32_t id = 0, h = 1, v = 1, tq = 0, td = 0, ta = 0;
    uint32_t bw = 0, bh = 0;  // blocks across / down in the padded plane
    std::vector<uint8_t> plane;  // bw*8 x bh*8 samples
    int32_t pred = 0;
};

struct JBits {
    const uint8_t* p;
    size_t n, i;
    uint32_t buf = 0;
    int cnt = 0;
    bool marker = false;  // a marker was reached: zeros are fed from here on
    void fill() {
        while (cnt <= 24) {
            uint32_t b = 0;
            if (!marker && i < n) {
                b = p[i];
                if (b == 0xFF) {
                    const uint8_t nx = i + 1 < n ? p[i + 1] : 0xD9;
                    if (nx == 0x00) {
                        i += 2;
                    } else {
                        marker = true;
                        b = 0;
                    }
                } else {
                    ++i;
                }
            }
            buf |= b << (24 - cnt);
            cnt += 8;
        }
    }
    uint32_t bits(int k) {  // k <= 16, most significant bit first
        if (k == 0) return 0;
        fill();
        const uint32_t v = buf >> (32 - k);
        buf <<= k;
        cnt -= k;
        return v;
    }
    void reset_at(size_t pos) {
        i = pos;
        buf = 0;
        cnt = 0;
        marker = false;
    }
};

bool jhuff_build(JHuff& t, const uint8_t* counts, const uint8_t* vals, int nvals) {
    int total = 0;
    for (int l = 0; l < 16; ++l) total += counts[l];
    if (total > 256 || total != nvals) return false;
    memcpy(t.vals, vals, (size_t)total);
    int32_t code = 0, k = 0;
    for (int l = 1; l <= 16; ++l) {
        t.valptr[l] = k;
        t.mincode[l] = code;
        code += counts[l - 1];
        k += counts[l - 1];
        t.maxcode[l] = counts[l - 1] ? code - 1 : -1;
        if (code > (1 << l)) return false;  // over-subscribed
        code <<= 1;
    }
    t.maxcode[17] = 0x7FFFFFFF;
    t.present = true;
    return true;
}

int jhuff_decode(JBits& b, const JHuff& t) {
    int32_t code = (int32_t)b.bits(1);
    for (int l = 1; l <= 16; ++l) {
        if (code <= t.maxcode[l] && t.maxcode[l] >= 0) return t.vals[t.valptr[l] + code - t.mincode[l]];
        code = (code << 1) | (int32_t)b.bits(1);
    }
    return -1;
}

inline int32_t jextend(uint32_t v, int s) { return s && v < (1u << (s - 1)) ? (int32_t)v - (1 << s) + 1 : (int32_t)v; }

// Separable inverse DCT (T.81 A.3.3) in double, level shift, round, clamp.
void jidct(const int32_t* coef, uint8_t* out, size_t stride) {
    static double c[8][8];
    static bool init = false;
    if (!init) {
        for (int x = 0; x < 8; ++x)
            for (int u = 0; u < 8; ++u)
                c[x][u] = (u == 0 ? std::sqrt(0.5) : 1.0) * std::cos((2 * x + 1) * u * 3.14159265358979323846 / 16.0);
        init = true;
    }
    double tmp[64];
    for (int y = 0; y < 8; ++y)  // rows of the result, columns of frequency v
        for (int u = 0; u < 8; ++u) {
            double acc = 0.0;
            for (int v = 0; v < 8; ++v) acc += c[y][v] * coef[v * 8 + u];
            tmp[y * 8 + u] = acc;
        }
    for (int y = 0; y < 8; ++y)
        for (int x = 0; x < 8; ++x) {
            double acc = 0.0;
            for (int u = 0; u < 8; ++u) acc += c[x][u] * tmp[y * 8 + u];
            const double v = std::floor(acc / 4.0 + 128.5);
            out[y * stride + x] = (uint8_t)(v < 0.0 ? 0.0 : v > 255.0 ? 255.0 : v);
        }
}

inline uint16_t jbe16(const uint8_t* p) { return (uint16_t)((p[0] << 8) | p[1]); }

// Triangle-filter upsampling of one component to the image size for 2:1
// horizontal and/or vertical factors; other ratios replicate samples.
std::vector<uint8_t> upsample(const JComp& c, uint32_t hmax, uint32_t vmax, uint32_t w, uint32_t h) {
    const uint32_t sx = hmax / c.h, sy = vmax / c.v;
    const uint32_t pw = c.bw * 8, ph = c.bh * 8;
    const uint32_t cw = (w * c.h + hmax - 1) / hmax, ch = (h * c.v + vmax - 1) / vmax;  // the component's own size
    std::vector<uint8_t> out((size_t)w * h);
    auto at = [&](int64_t x, int64_t y) -> int {  // edge-replicated sample of the component's own extent
        x = std::min<int64_t>(std::max<int64_t>(x, 0), (int64_t)cw - 1);
        y = std::min<int64_t>(std::max<int64_t>(y, 0), (int64_t)ch - 1);
        (void)pw;
        (void)ph;
        return c.plane[(size_t)y * pw + (size_t)x];
    };
    if (sx == 1 && sy == 1) {
        for (uint32_t y = 0; y < h; ++y)
            for (uint32_t x = 0; x < w; ++x) out[(size_t)y * w + x] = (uint8_t)at(x, y);
    } else if (sx == 2 && sy == 1) {
        for (uint32_t y = 0; y < h; ++y)
            for (uint32_t x = 0; x < w; ++x) {
                const int64_t i = x >> 1;
                const int near = at(i, y);
                int v;
                if (x & 1)
                    v = i + 1 >= (int64_t)cw ? near : (near * 3 + at(i + 1, y) + 2) >> 2;
                else
                    v = i == 0 ? near : (near * 3 + at(i - 1, y) + 1) >> 2;
                out[(size_t)y * w + x] = (uint8_t)v;
            }
    } else if (sx == 2 && sy == 2) {
        for (uint32_t y = 0; y < h; ++y) {
            const int64_t j = y >> 1, jf = (y & 1) ? j + 1 : j - 1;  // near and far input rows
            auto colsum = [&](int64_t i) { return at(i, j) * 3 + at(i, jf); };
            for (uint32_t x = 0; x < w; ++x) {
                const int64_t i = x >> 1;
                int v;
                if (x & 1)
                    v = i + 1 >= (int64_t)cw ? (colsum(i) * 4 + 7) >> 4 : (colsum(i) * 3 + colsum(i + 1) + 7) >> 4;
                else
                    v = i == 0 ? (colsum(i) * 4 + 8) >> 4 : (colsum(i) * 3 + colsum(i - 1) + 8) >> 4;
                out[(size_t)y * w + x] = (uint8_t)v;
            }
        }
    } else {
        for (uint32_t y = 0; y < h; ++y)
            for (uint32_t x = 0; x < w; ++x) out[(size_t)y * w + x] = (uint8_t)at(x / sx, y / sy);
    }
    return out;
}

}  // namespace

bool is_jpeg(const uint8_t* d, size_t n) { return n >= 3 && d[0] == 0xFF && d[1] == 0xD8 && d[2] == 0xFF; }

int decode_jpeg(const uint8_t* d, size_t n, Rgba8Image& out, std::string& err, uint64_t max_texels) {
    if (!is_jpeg(d, n)) return err = "not a JPEG file", -1;
    uint16_t qt[4][64];
    bool have_q[4] = {false, false, false, false};
    JHuff hdc[4], hac[4];
    std::vector<JComp> comps;
    uint32_t W = 0, H = 0, hmax = 1, vmax = 1, restart = 0, mcux = 0, mcuy = 0;
    bool have_frame = false, have_scan = false;
    size_t pos = 2;
    while (pos + 4 <= n) {
        if (d[pos] != 0xFF) return err = "marker expected", -1;
        const uint8_t m = d[pos + 1];
        if (m == 0xFF) {  // fill byte
            ++pos;
            continue;
        }
        if (m == 0xD9) break;  // EOI
        if (m == 0x01 || (m >= 0xD0 && m <= 0xD7)) {
            pos += 2;
            continue;
        }
        const uint32_t len = jbe16(d + pos + 2);
        if (len < 2 || pos + 2 + len > n) return err = "truncated segment", -1;
        const uint8_t* seg = d + pos + 4;
        const uint32_t sl = len - 2;
        pos += 2 + len;
        if (m == 0xC2 || m == 0xC3 || (m >= 0xC5 && m <= 0xC7) || (m >= 0xC9 && m <= 0xCB) || (m >= 0xCD && m <= 0xCF))
            return err = "not a baseline JPEG (progressive, lossless or arithmetic-coded)", 0;
        if (m == 0xDB) {  // DQT
            for (uint32_t o = 0; o < sl;) {
                const uint32_t pq = seg[o] >> 4, tq = seg[o] & 15;
                if (tq > 3 || pq > 1 || o + 1 + 64 * (pq + 1) > sl) return err = "bad DQT", -1;
                for (int k = 0; k < 64; ++k)
                    qt[tq][k] = pq ? jbe16(seg + o + 1 + 2 * k) : seg[o + 1 + k];
                have_q[tq] = true;
                o += 1 + 64 * (pq + 1);
            }
        } else if (m == 0xC4) {  // DHT
            for (uint32_t o = 0; o < sl;) {
                if (o + 17 > sl) return err = "bad DHT", -1;
                const uint32_t tc = seg[o] >> 4, th = seg[o] & 15;
                int total = 0;
                for (int l = 0; l < 16; ++l) total += seg[o + 1 + l];
                if (tc > 1 || th > 3 || o + 17 + (uint32_t)total > sl) return err = "bad DHT", -1;
                if (!jhuff_build(tc ? hac[th] : hdc[th], seg + o + 1, seg + o + 17, total)) return err = "bad Huffman table", -1;
                o += 17 + (uint32_t)total;
            }
        } else if (m == 0xDD) {  // DRI
            if (sl < 2) return err = "bad DRI", -1;
            restart = jbe16(seg);
        } else if (m == 0xC0 || m == 0xC1) {  // SOF0 / SOF1
            if (have_frame || sl < 6) return err = "bad SOF", -1;
            if (seg[0] != 8) return err = "not 8-bit samples", 0;
            H = jbe16(seg + 1);
            W = jbe16(seg + 3);
            const uint32_t nc = seg[5];
            if (nc != 1 && nc != 3) return err = "unsupported component count", 0;
            if (W == 0 || H == 0 || (uint64_t)W * H > max_texels) return err = "image size out of range", -1;
            if (sl < 6 + 3 * nc) return err = "bad SOF", -1;
            comps.resize(nc);
            for (uint32_t k = 0; k < nc; ++k) {
                comps[k].id = seg[6 + 3 * k];
                comps[k].h = seg[7 + 3 * k] >> 4;
                comps[k].v = seg[7 + 3 * k] & 15;
                comps[k].tq = seg[8 + 3 * k];
                if (comps[k].h < 1 || comps[k].h > 4 || comps[k].v < 1 || comps[k].v > 4 || comps[k].tq > 3)
                    return err = "bad component parameters", -1;
                hmax = std::max(hmax, comps[k].h);
                vmax = std::max(vmax, comps[k].v);
            }
            mcux = (W + 8 * hmax - 1) / (8 * hmax);
            mcuy = (H + 8 * vmax - 1) / (8 * vmax);
            for (JComp& c : comps) {
                c.bw = mcux * c.h;
                c.bh = mcuy * c.v;
                c.plane.assign((size_t)c.bw * 8 * c.bh * 8, 0);
            }
            have_frame = true;
        } else if (m == 0xDA) {  // SOS: the entropy-coded data follows
            if (!have_frame || sl < 1) return err = "scan before frame", -1;
            const uint32_t ns = seg[0];
            if (ns < 1 || ns > comps.size() || sl < 1 + 2 * ns + 3) return err = "bad SOS", -1;
            std::vector<JComp*> sc;
            for (uint32_t k = 0; k < ns; ++k) {
                JComp* c = nullptr;
                for (JComp& x : comps)
                    if (x.id == seg[1 + 2 * k]) c = &x;
                if (!c) return err = "scan names an unknown component", -1;
                c->td = seg[2 + 2 * k] >> 4;
                c->ta = seg[2 + 2 * k] & 15;
                if (c->td > 3 || c->ta > 3 || !hdc[c->td].present || !hac[c->ta].present || !have_q[c->tq])
                    return err = "scan without its tables", -1;
                c->pred = 0;
                sc.push_back(c);
            }
            if (ns > 1) {
                uint32_t blocks = 0;
                for (JComp* c : sc) blocks += c->h * c->v;
                if (blocks > 10) return err = "too many blocks per MCU", -1;
            }
            JBits b{d, n, pos};
            int32_t coef[64];
            // units of the scan: MCUs (interleaved) or single blocks
            const uint32_t ucx = ns > 1 ? mcux : (uint32_t)(((uint64_t)W * sc[0]->h + 8 * hmax - 1) / (8 * hmax));
            const uint32_t ucy = ns > 1 ? mcuy : (uint32_t)(((uint64_t)H * sc[0]->v + 8 * vmax - 1) / (8 * vmax));
            const uint64_t units = (uint64_t)ucx * ucy;
            auto block = [&](JComp& c, uint32_t bx, uint32_t by) -> bool {
                memset(coef, 0, sizeof(coef));
                const int t = jhuff_decode(b, hdc[c.td]);
                if (t < 0 || t > 11) return false;
                c.pred += jextend(b.bits(t), t);
                coef[0] = c.pred * (int32_t)qt[c.tq][0];
                for (int k = 1; k < 64;) {
                    const int rs = jhuff_decode(b, hac[c.ta]);
                    if (rs < 0) return false;
                    const int r = rs >> 4, sz = rs & 15;
                    if (sz == 0) {
                        if (r != 15) break;
                        k += 16;
                        continue;
                    }
                    k += r;
                    if (k > 63 || sz > 10) return false;
                    coef[kZigzag[k]] = jextend(b.bits(sz), sz) * (int32_t)qt[c.tq][k];
                    ++k;
                }
                if (bx < c.bw && by < c.bh) jidct(coef, c.plane.data() + ((size_t)by * 8) * c.bw * 8 + (size_t)bx * 8, (size_t)c.bw * 8);
                return true;
            };
            for (uint64_t u = 0; u < units; ++u) {
                if (restart && u > 0 && u % restart == 0) {  // expect RSTn, realign, reset predictors
                    size_t q = b.i;
                    while (q + 1 < n && !(d[q] == 0xFF && d[q + 1] >= 0xD0 && d[q + 1] <= 0xD7)) ++q;
                    if (q + 1 >= n) return err = "missing restart marker", -1;
                    b.reset_at(q + 2);
                    for (JComp* c : sc) c->pred = 0;
                }
                const uint32_t ux2 = (uint32_t)(u % ucx), uy2 = (uint32_t)(u / ucx);
                if (ns > 1) {
                    for (JComp* c : sc)
                        for (uint32_t yy = 0; yy < c->v; ++yy)
                            for (uint32_t xx = 0; xx < c->h; ++xx)
                                if (!block(*c, ux2 * c->h + xx, uy2 * c->v + yy)) return err = "bad entropy-coded data", -1;
                } else if (!block(*sc[0], ux2, uy2)) {
                    return err = "bad entropy-coded data", -1;
                }
            }
            // continue after the scan's data: the next marker that is not RSTn
            size_t q = b.i;
            while (q + 1 < n && !(d[q] == 0xFF && d[q + 1] != 0x00 && !(d[q + 1] >= 0xD0 && d[q + 1] <= 0xD7))) ++q;
            pos = q;
            have_scan = true;
        } else if (m == 0xC8 || (m >= 0xF0 && m <= 0xFD) || m == 0xDC || m == 0xDE || m == 0xDF) {
            return err = "unsupported JPEG marker", 0;
        }
        // APPn, COM and other segments are skipped
    }
    if (!have_frame || !have_scan) return err = "no frame or scan", -1;
    out.width = W;
    out.height = H;
    out.texels.assign((size_t)W * H * 4, 255);
    if (comps.size() == 1) {
        const std::vector<uint8_t> y = upsample(comps[0], hmax, vmax, W, H);
        for (size_t i = 0; i < (size_t)W * H; ++i) out.texels[4 * i] = out.texels[4 * i + 1] = out.texels[4 * i + 2] = y[i];
        return 1;
    }
    const std::vector<uint8_t> Y = upsample(comps[0], hmax, vmax, W, H), Cb = upsample(comps[1], hmax, vmax, W, H),
                               Cr = upsample(comps[2], hmax, vmax, W, H);
    auto clamp8 = [](double v) { return (uint8_t)(v < 0.0 ? 0.0 : v > 255.0 ? 255.0 : v); };
    for (size_t i = 0; i < (size_t)W * H; ++i) {  // JFIF (ITU-R BT.601 full range)
        const double y = Y[i], cb = Cb[i] - 128.0, cr = Cr[i] - 128.0;
        out.texels[4 * i] = clamp8(std::floor(y + 1.402 * cr + 0.5));
        out.texels[4 * i + 1] = clamp8(std::floor(y - 0.344136 * cb - 0.714136 * cr + 0.5));
        out.texels[4 * i + 2] = clamp8(std::floor(y + 1.772 * cb + 0.5));
    }
    return 1;
}

// ---------------------------------------------------------------------------
// image 0.24 imageops::sample::resize with the Lanczos3 filter
// ---------------------------------------------------------------------------
namespace {

float sinc(float t) {
    const float a = t * 3.14159265358979323846f;
    return t == 0.0f ? 1.0f : std::sin(a) / a;
}
float lanczos3(float x) { return std::fabs(x) < 3.0f ? sinc(x) * sinc(x / 3.0f) : 0.0f; }

// Taps [left, right) and normalised weights of output index o when n inputs
// map onto m outputs.
void taps(uint32_t n, uint32_t m, uint32_t o, int64_t& left, int64_t& right, std::vector<float>& ws) {
    const float ratio = (float)n / (float)m;
    const float sratio = ratio < 1.0f ? 1.0f : ratio;
    const float support = 3.0f * sratio;
    float in = ((float)o + 0.5f) * ratio;
    left = std::min<int64_t>(std::max<int64_t>((int64_t)std::floor(in - support), 0), (int64_t)n - 1);
    right = std::min<int64_t>(std::max<int64_t>((int64_t)std::ceil(in + support), left + 1), (int64_t)n);
    in = in - 0.5f;
    ws.clear();
    float sum = 0.0f;
    for (int64_t i = left; i < right; ++i) {
        const float wgt = lanczos3(((float)i - in) / sratio);
        ws.push_back(wgt);
        sum += wgt;
    }
    for (float& wgt : ws) wgt /= sum;
}

}  // namespace

Rgba8Image resize_lanczos3(const Rgba8Image& src, uint32_t nw, uint32_t nh) {
    Rgba8Image out;
    out.width = nw;
    out.height = nh;
    if (nw == src.width && nh == src.height) {
        out.texels = src.texels;
        return out;
    }
    const uint32_t w = src.width, h = src.height;
    std::vector<float> tmp((size_t)w * nh * 4), ws;
    for (uint32_t oy = 0; oy < nh; ++oy) {
        int64_t l, r;
        taps(h, nh, oy, l, r, ws);
        for (uint32_t x = 0; x < w; ++x) {
            float t[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            for (size_t i = 0; i < ws.size(); ++i) {
                const uint8_t* p = src.texels.data() + 4 * ((size_t)(l + (int64_t)i) * w + x);
                for (int c = 0; c < 4; ++c) t[c] += (float)p[c] * ws[i];
            }
            memcpy(tmp.data() + 4 * ((size_t)oy * w + x), t, sizeof(t));
        }
    }
    out.texels.resize((size_t)nw * nh * 4);
    for (uint32_t ox = 0; ox < nw; ++ox) {
        int64_t l, r;
        taps(w, nw, ox, l, r, ws);
        for (uint32_t y = 0; y < nh; ++y) {
            float t[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            for (size_t i = 0; i < ws.size(); ++i) {
                const float* p = tmp.data() + 4 * ((size_t)y * w + (size_t)(l + (int64_t)i));
                for (int c = 0; c < 4; ++c) t[c] += p[c] * ws[i];
            }
            uint8_t* o = out.texels.data() + 4 * ((size_t)y * nw + ox);
            for (int c = 0; c < 4; ++c) o[c] = (uint8_t)std::round(std::min(std::max(t[c], 0.0f), 255.0f));
        }
    }
    return out;
}

std::vector<Rgba8Image> mip_chain(const Rgba8Image& base) {
    const uint32_t m = std::max(base.width, base.height);
    uint32_t levels = 1;
    while ((1u << (levels - 1)) < m) ++levels;  // 1 + log2(next_pow2(m))
    std::vector<Rgba8Image> out;
    out.reserve(levels);
    out.push_back(base);
    for (uint32_t l = 1; l < levels; ++l) {
        const Rgba8Image& p = out.back();
        out.push_back(resize_lanczos3(p, std::max(1u, p.width / 2), std::max(1u, p.height / 2)));
    }
    return out;
}

}  // namespace vgi
