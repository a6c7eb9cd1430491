// nr_image.cpp — JPEG, BMP and TGA decoders (see nr_image.h).
//
// JPEG is decoded per ITU-T T.81: Huffman-coded sequential (SOF0, SOF1) and
// progressive (SOF2) frames, interleaved and single-component scans, restart
// intervals, spectral selection and successive approximation.  Reconstruction
// uses the arithmetic stb_image 2.30 uses on x86-64, so the output bytes are
// the reference loader's: coefficients dequantised to 16 bits (progressive
// frames at the end, with the tables in force then), the integer inverse DCT of
// the IJG "islow" method with 12-bit constants and 16-bit intermediate rows,
// chroma upsampled with the triangle filters of the JFIF siting (3:1 weights,
// 2:1 horizontally and/or vertically; nearest neighbour for other ratios), and
// YCbCr -> RGB in 20-bit fixed point with the green Cb term truncated to its
// upper 16 bits.  Adobe APP14 transform flags and component ids 'R','G','B'
// select RGB, CMYK and YCCK data as stb selects them.
//
// BMP (Windows 3.x / OS/2 headers, V4 / V5 headers; 1, 4, 8, 16, 24, 32 bits,
// BI_BITFIELDS masks; bottom-up and top-down rows) and TGA (colour-mapped,
// true-colour and grey, raw and run-length, 15/16-bit 5:5:5 colour, 8-bit grey
// + alpha at 16 bits) follow their published layouts, with stb's channel counts
// and conversions (n-bit mask fields widened to 8 bits by bit replication,
// all-zero BMP alpha read as opaque, TGA 5-bit fields scaled by 255/31).
#include "nr_image.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>

namespace nrio {
namespace {

// Byte source: reads past the end return 0, as a file read at EOF does for stb.
struct Src {
    const uint8_t* d;
    size_t n, p = 0;
    Src(const uint8_t* d_, size_t n_) : d(d_), n(n_) {}
    int u8() { return p < n ? d[p++] : 0; }
    int be16() { int a = u8(); return (a << 8) | u8(); }
    int le16() { int a = u8(); return a | (u8() << 8); }
    uint32_t le32() { uint32_t a = (uint32_t)le16(); return a | ((uint32_t)le16() << 16); }
    void skip(long k) {
        if (k < 0) { p = n; return; }
        p = (size_t)k > n - std::min(p, n) ? n : p + (size_t)k;
    }
    bool eof() const { return p >= n; }
};

// ===========================================================================
// JPEG
// ===========================================================================

// natural (row-major) position of the k-th coefficient in zigzag order
// (T.81 figure A.6); indices past 63 (a corrupt run) land on 63
const uint8_t kZigzag[64 + 15] = {
    0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13,
    6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31,
    39, 46, 53, 60, 61, 54, 47, 55, 62, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63};

constexpr int kMarkerNone = 0xff;
constexpr int kLookBits = 9;

// Canonical Huffman table (T.81 annex C): codes of each length assigned in
// order; `lookup` resolves codes of at most kLookBits bits in one step.
struct Huffman {
    uint8_t lookup[1 << kLookBits];   // symbol index, 255 = longer code
    uint8_t length[257];
    uint8_t symbol[256];
    uint32_t limit[18];               // first code past length L, left-aligned to 16 bits
    int offset[17];                   // symbol index = code + offset[L]
};

bool huffman_build(Huffman& h, const int counts[16]) {
    int k = 0;
    for (int L = 1; L <= 16; L++)
        for (int i = 0; i < counts[L - 1]; i++) h.length[k++] = (uint8_t)L;
    h.length[k] = 0;
    uint16_t codes[256];
    uint32_t code = 0;
    k = 0;
    for (int L = 1; L <= 16; L++) {
        h.offset[L] = k - (int)code;
        if (h.length[k] == L) {
            while (h.length[k] == L) codes[k++] = (uint16_t)code++;
            if (code - 1 >= (1u << L)) return false;            // lengths overflow the code space
        }
        h.limit[L] = code << (16 - L);
        code <<= 1;
    }
    h.limit[17] = 0xffffffffu;
    std::memset(h.lookup, 255, sizeof(h.lookup));
    for (int i = 0; i < k; i++) {
        const int L = h.length[i];
        if (L > kLookBits) continue;
        const int first = codes[i] << (kLookBits - L), span = 1 << (kLookBits - L);
        for (int j = 0; j < span; j++) h.lookup[first + j] = (uint8_t)i;
    }
    return true;
}

struct Component {
    int id = 0, h = 1, v = 1, tq = 0, td = 0, ta = 0;
    int x = 0, y = 0;          // samples of the component (before upsampling)
    int w2 = 0, h2 = 0;        // padded to whole MCUs
    int dc_pred = 0;
    std::vector<uint8_t> pix;  // w2 x h2
    std::vector<int16_t> coef; // progressive: 64 per block, (w2 / 8) blocks per row
    int coef_w = 0;
};

struct Decoder {
    Src s;
    std::string err;
    uint16_t dq[4][64] = {};
    Huffman hdc[4], hac[4];
    Component comp[4];
    int n_comp = 0, W = 0, H = 0, hmax = 1, vmax = 1, mcu_x = 0, mcu_y = 0;
    bool progressive = false, jfif = false;
    int adobe = -1, rgb_ids = 0;
    int restart_interval = 0, todo = 0, eob_run = 0;
    int scan_n = 0, order[4] = {}, ss = 0, se = 0, ah = 0, al = 0;
    // entropy-coded bits: left-aligned in `bits`, `nbits` valid
    uint32_t bits = 0;
    int nbits = 0;
    int marker = kMarkerNone;
    bool ended = false;        // a marker stopped the bit stream: zeros follow

    explicit Decoder(const uint8_t* d, size_t n) : s(d, n) {}

    bool fail(const char* m) { err = m; return false; }

    void refill() {
        do {
            int b = ended ? 0 : s.u8();
            if (b == 0xff) {
                int c = s.u8();
                while (c == 0xff) c = s.u8();
                if (c != 0) { marker = c; ended = true; return; }
            }
            bits |= (uint32_t)b << (24 - nbits);
            nbits += 8;
        } while (nbits <= 24);
    }

    int decode(const Huffman& h) {
        if (nbits < 16) refill();
        const int k = h.lookup[bits >> (32 - kLookBits)];
        if (k < 255) {
            const int L = h.length[k];
            if (L > nbits) return -1;
            bits <<= L;
            nbits -= L;
            return h.symbol[k];
        }
        const uint32_t t = bits >> 16;
        int L = kLookBits + 1;
        while (t >= h.limit[L]) L++;
        if (L == 17) { nbits -= 16; return -1; }
        if (L > nbits) return -1;
        const int idx = (int)(bits >> (32 - L)) + h.offset[L];
        if (idx < 0 || idx >= 256) return -1;
        bits <<= L;
        nbits -= L;
        return h.symbol[idx];
    }

    int take(int n) {                         // n raw bits (n >= 1), 0 when the stream ran dry
        if (nbits < n) refill();
        if (nbits < n) return 0;
        const uint32_t v = bits >> (32 - n);
        bits = n == 32 ? 0 : bits << n;
        nbits -= n;
        return (int)v;
    }
    int take_extended(int n) {                // RECEIVE + EXTEND (T.81 F.2.2.1)
        if (nbits < n) refill();
        if (nbits < n) return 0;
        const int v = (int)(bits >> (32 - n));
        bits <<= n;
        nbits -= n;
        return v < (1 << (n - 1)) ? v - (1 << n) + 1 : v;
    }
    bool take_bit() {
        if (nbits < 1) refill();
        if (nbits < 1) return false;
        const bool b = (bits & 0x80000000u) != 0;
        bits <<= 1;
        nbits--;
        return b;
    }

    void reset_entropy() {
        bits = 0;
        nbits = 0;
        ended = false;
        for (auto& c : comp) c.dc_pred = 0;
        marker = kMarkerNone;
        todo = restart_interval ? restart_interval : 0x7fffffff;
        eob_run = 0;
    }

    int next_marker() {
        if (marker != kMarkerNone) { int m = marker; marker = kMarkerNone; return m; }
        int x = s.u8();
        if (x != 0xff) return kMarkerNone;
        while (x == 0xff) x = s.u8();
        return x;
    }

    // ---- segments ----
    bool marker_segment(int m) {
        if (m == kMarkerNone) return fail("expected marker");
        if (m == 0xDD) {                                                        // DRI
            if (s.be16() != 4) return fail("bad DRI length");
            restart_interval = s.be16();
            return true;
        }
        if (m == 0xDB) {                                                        // DQT
            int L = s.be16() - 2;
            while (L > 0) {
                const int q = s.u8(), prec = q >> 4, t = q & 15;
                if (prec > 1) return fail("bad DQT precision");
                if (t > 3) return fail("bad DQT table");
                for (int i = 0; i < 64; i++) dq[t][kZigzag[i]] = (uint16_t)(prec ? s.be16() : s.u8());
                L -= prec ? 129 : 65;
            }
            return L == 0 ? true : fail("bad DQT length");
        }
        if (m == 0xC4) {                                                        // DHT
            int L = s.be16() - 2;
            while (L > 0) {
                const int q = s.u8(), tc = q >> 4, th = q & 15;
                if (tc > 1 || th > 3) return fail("bad DHT header");
                int counts[16], total = 0;
                for (int i = 0; i < 16; i++) { counts[i] = s.u8(); total += counts[i]; }
                if (total > 256) return fail("bad DHT header");
                L -= 17;
                Huffman& h = tc == 0 ? hdc[th] : hac[th];
                if (!huffman_build(h, counts)) return fail("bad code lengths");
                for (int i = 0; i < total; i++) h.symbol[i] = (uint8_t)s.u8();
                L -= total;
            }
            return L == 0 ? true : fail("bad DHT length");
        }
        if ((m >= 0xE0 && m <= 0xEF) || m == 0xFE) {                            // APPn, COM
            int L = s.be16();
            if (L < 2) return fail("bad segment length");
            L -= 2;
            if (m == 0xE0 && L >= 5) {
                static const uint8_t tag[5] = {'J', 'F', 'I', 'F', 0};
                bool ok = true;
                for (int i = 0; i < 5; i++) ok &= s.u8() == tag[i];
                L -= 5;
                if (ok) jfif = true;
            } else if (m == 0xEE && L >= 12) {
                static const uint8_t tag[6] = {'A', 'd', 'o', 'b', 'e', 0};
                bool ok = true;
                for (int i = 0; i < 6; i++) ok &= s.u8() == tag[i];
                L -= 6;
                if (ok) {
                    s.u8();            // version
                    s.be16();          // flags0
                    s.be16();          // flags1
                    adobe = s.u8();    // colour transform: 0 none (RGB / CMYK), 1 YCbCr, 2 YCCK
                    L -= 6;
                }
            }
            s.skip(L);
            return true;
        }
        return fail("unknown marker");
    }

    bool frame_header() {
        const int Lf = s.be16();
        if (Lf < 11) return fail("bad SOF length");
        if (s.u8() != 8) return fail("only 8-bit JPEG samples are supported");
        H = s.be16();
        if (H == 0) return fail("JPEG with its height in a DNL marker");
        W = s.be16();
        if (W == 0) return fail("zero width");
        if ((uint32_t)H > MAX_DIM || (uint32_t)W > MAX_DIM) return fail("image too large");
        n_comp = s.u8();
        if (n_comp != 1 && n_comp != 3 && n_comp != 4) return fail("bad component count");
        if (Lf != 8 + 3 * n_comp) return fail("bad SOF length");
        rgb_ids = 0;
        for (int i = 0; i < n_comp; i++) {
            static const uint8_t rgb[3] = {'R', 'G', 'B'};
            Component& c = comp[i];
            c.id = s.u8();
            if (n_comp == 3 && c.id == rgb[i]) rgb_ids++;
            const int q = s.u8();
            c.h = q >> 4;
            c.v = q & 15;
            if (c.h < 1 || c.h > 4) return fail("bad H sampling factor");
            if (c.v < 1 || c.v > 4) return fail("bad V sampling factor");
            c.tq = s.u8();
            if (c.tq > 3) return fail("bad quantisation table index");
        }
        if ((uint64_t)W * H * n_comp > 0x7fffffffull) return fail("image too large");
        hmax = vmax = 1;
        for (int i = 0; i < n_comp; i++) { hmax = std::max(hmax, comp[i].h); vmax = std::max(vmax, comp[i].v); }
        for (int i = 0; i < n_comp; i++) {
            if (hmax % comp[i].h) return fail("bad H sampling factor");
            if (vmax % comp[i].v) return fail("bad V sampling factor");
        }
        mcu_x = (W + hmax * 8 - 1) / (hmax * 8);
        mcu_y = (H + vmax * 8 - 1) / (vmax * 8);
        for (int i = 0; i < n_comp; i++) {
            Component& c = comp[i];
            c.x = (W * c.h + hmax - 1) / hmax;
            c.y = (H * c.v + vmax - 1) / vmax;
            c.w2 = mcu_x * c.h * 8;
            c.h2 = mcu_y * c.v * 8;
            c.pix.assign((size_t)c.w2 * c.h2, 0);
            if (progressive) {
                c.coef_w = c.w2 / 8;
                c.coef.assign((size_t)c.w2 * c.h2, 0);
            }
        }
        return true;
    }

    bool scan_header() {
        const int Ls = s.be16();
        scan_n = s.u8();
        if (scan_n < 1 || scan_n > 4 || scan_n > n_comp) return fail("bad SOS component count");
        if (Ls != 6 + 2 * scan_n) return fail("bad SOS length");
        for (int i = 0; i < scan_n; i++) {
            const int id = s.u8(), q = s.u8();
            int which = 0;
            while (which < n_comp && comp[which].id != id) which++;
            if (which == n_comp) return fail("SOS names an unknown component");
            comp[which].td = q >> 4;
            if (comp[which].td > 3) return fail("bad DC table index");
            comp[which].ta = q & 15;
            if (comp[which].ta > 3) return fail("bad AC table index");
            order[i] = which;
        }
        ss = s.u8();
        se = s.u8();
        const int a = s.u8();
        ah = a >> 4;
        al = a & 15;
        if (progressive) {
            if (ss > 63 || se > 63 || ss > se || ah > 13 || al > 13) return fail("bad SOS");
        } else {
            if (ss != 0 || ah != 0 || al != 0) return fail("bad SOS");
            se = 63;
        }
        return true;
    }

    // ---- blocks ----
    bool sequential_block(int16_t* data, Component& c) {
        if (nbits < 16) refill();
        const int t = decode(hdc[c.td]);
        if (t < 0 || t > 15) return fail("bad Huffman code");
        std::memset(data, 0, 64 * sizeof(int16_t));
        const int diff = t ? take_extended(t) : 0;
        const int64_t dc = (int64_t)c.dc_pred + diff;
        if (dc < INT32_MIN || dc > INT32_MAX) return fail("bad DC delta");
        c.dc_pred = (int)dc;
        const uint16_t* q = dq[c.tq];
        const int64_t d0 = dc * q[0];
        if (d0 < INT16_MIN || d0 > INT16_MAX) return fail("DC out of range");
        data[0] = (int16_t)d0;
        int k = 1;
        do {
            const int rs = decode(hac[c.ta]);
            if (rs < 0) return fail("bad Huffman code");
            const int size = rs & 15, run = rs >> 4;
            if (size == 0) {
                if (rs != 0xf0) break;                          // end of block
                k += 16;
            } else {
                k += run;
                const int z = kZigzag[k++];
                data[z] = (int16_t)(take_extended(size) * q[z]);
            }
        } while (k < 64);
        return true;
    }

    bool dc_first_or_refine(int16_t* data, Component& c) {
        if (se != 0) return fail("DC and AC in one progressive scan");
        if (nbits < 16) refill();
        if (ah == 0) {
            std::memset(data, 0, 64 * sizeof(int16_t));
            const int t = decode(hdc[c.td]);
            if (t < 0 || t > 15) return fail("bad Huffman code");
            const int diff = t ? take_extended(t) : 0;
            const int64_t dc = (int64_t)c.dc_pred + diff;
            if (dc < INT32_MIN || dc > INT32_MAX) return fail("bad DC delta");
            c.dc_pred = (int)dc;
            const int64_t d0 = dc * (int64_t)(1 << al);
            if (d0 < INT16_MIN || d0 > INT16_MAX) return fail("DC out of range");
            data[0] = (int16_t)d0;
        } else if (take_bit()) {
            data[0] = (int16_t)(data[0] + (1 << al));
        }
        return true;
    }

    // one coefficient already non-zero in a refinement pass: its next bit
    void refine(int16_t* p, int16_t bit) {
        if (take_bit() && (*p & bit) == 0) *p = (int16_t)(*p > 0 ? *p + bit : *p - bit);
    }

    bool ac_first_or_refine(int16_t* data, const Huffman& h) {
        if (ss == 0) return fail("DC and AC in one progressive scan");
        if (ah == 0) {
            if (eob_run) { eob_run--; return true; }
            int k = ss;
            do {
                const int rs = decode(h);
                if (rs < 0) return fail("bad Huffman code");
                const int size = rs & 15, run = rs >> 4;
                if (size == 0) {
                    if (run < 15) {                               // EOBn: this and 2^n - 1 + bits more bands end
                        eob_run = 1 << run;
                        if (run) eob_run += take(run);
                        eob_run--;
                        break;
                    }
                    k += 16;
                } else {
                    k += run;
                    const int z = kZigzag[k++];
                    data[z] = (int16_t)(take_extended(size) * (1 << al));
                }
            } while (k <= se);
            return true;
        }
        const int16_t bit = (int16_t)(1 << al);
        if (eob_run) {
            eob_run--;
            for (int k = ss; k <= se; k++) {
                int16_t* p = &data[kZigzag[k]];
                if (*p != 0) refine(p, bit);
            }
            return true;
        }
        int k = ss;
        do {
            const int rs = decode(h);
            if (rs < 0) return fail("bad Huffman code");
            int size = rs & 15, run = rs >> 4, value = 0;
            if (size == 0) {
                if (run < 15) {
                    eob_run = (1 << run) - 1;
                    if (run) eob_run += take(run);
                    run = 64;                                     // refine the rest of the band, place nothing
                }                                                 // run 15: skip 16 zero-history coefficients
            } else {
                if (size != 1) return fail("bad Huffman code");
                value = take_bit() ? bit : -bit;
            }
            while (k <= se) {
                int16_t* p = &data[kZigzag[k++]];
                if (*p != 0) {
                    refine(p, bit);
                } else {
                    if (run == 0) {
                        *p = (int16_t)value;
                        break;
                    }
                    run--;
                }
            }
        } while (k <= se);
        return true;
    }

    // ---- the inverse DCT ----
    // The islow method (IJG jidctint) with 12-bit constants, evaluated as stb's
    // SSE2 path evaluates it on x86-64: the four input sums s0 +/- s4, s1 + s7,
    // s3 + s5 wrap at 16 bits, each rotation is a pair of 16 x 16 -> 32-bit
    // products with pre-added constants, 32-bit sums wrap, and each pass's outputs
    // saturate to 16 bits (for coefficients of ordinary files this is the plain
    // integer transform).  Columns keep 2 extra bits (bias 512, >> 10); rows remove
    // 2^17 with rounding and add 128 (bias 65536 + 128 << 17), then clamp to 0..255.
    static inline int32_t wrap16(int32_t x) { return (int16_t)(uint16_t)(uint32_t)x; }
    static inline int32_t add32(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
    static inline int32_t sub32(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
    static inline int16_t sat16(int32_t x) { return (int16_t)(x < -32768 ? -32768 : (x > 32767 ? 32767 : x)); }
    static inline int K(double c) { return (int)(c * 4096 + 0.5); }

    // one 8-point pass; out[k] receives output k (before saturation, after bias and shift)
    static inline void idct8(const int32_t s[8], int32_t bias, int shift, int32_t out[8]) {
        static const int c0541 = K(0.5411961f), c_1847 = K(-1.847759065f), c0765 = K(0.765366865f),
                         c1175 = K(1.175875602f), c0298 = K(0.298631336f), c2053 = K(2.053119869f),
                         c3072 = K(3.072711026f), c1501 = K(1.501321110f), c_0899 = K(-0.899976223f),
                         c_2562 = K(-2.562915447f), c_1961 = K(-1.961570560f), c_0390 = K(-0.390180644f);
        // even part
        const int32_t t2 = add32(s[2] * c0541, s[6] * (c0541 + c_1847));
        const int32_t t3 = add32(s[2] * (c0541 + c0765), s[6] * c0541);
        const int32_t t0 = wrap16(s[0] + s[4]) * 4096, t1 = wrap16(s[0] - s[4]) * 4096;
        const int32_t x0 = add32(t0, t3), x3 = sub32(t0, t3), x1 = add32(t1, t2), x2 = sub32(t1, t2);
        // odd part
        const int32_t y0 = add32(s[7] * (c_1961 + c0298), s[3] * c_1961);
        const int32_t y2 = add32(s[7] * c_1961, s[3] * (c_1961 + c3072));
        const int32_t y1 = add32(s[5] * (c_0390 + c2053), s[1] * c_0390);
        const int32_t y3 = add32(s[5] * c_0390, s[1] * (c_0390 + c1501));
        const int32_t a17 = wrap16(s[1] + s[7]), a35 = wrap16(s[3] + s[5]);
        const int32_t y4 = add32(a17 * (c1175 + c_0899), a35 * c1175);
        const int32_t y5 = add32(a17 * c1175, a35 * (c1175 + c_2562));
        const int32_t x4 = add32(y0, y4), x5 = add32(y1, y5), x6 = add32(y2, y5), x7 = add32(y3, y4);
        const int32_t e[4] = {add32(x0, bias), add32(x1, bias), add32(x2, bias), add32(x3, bias)};
        const int32_t o[4] = {x7, x6, x5, x4};
        for (int k = 0; k < 4; k++) {
            out[k] = add32(e[k], o[k]) >> shift;
            out[7 - k] = sub32(e[k], o[k]) >> shift;
        }
    }

    static void idct_block(uint8_t* out, int stride, const int16_t* d) {
        int16_t tmp[64];
        int32_t s[8], r[8];
        for (int c = 0; c < 8; c++) {                     // columns
            for (int k = 0; k < 8; k++) s[k] = d[8 * k + c];
            idct8(s, 512, 10, r);
            for (int k = 0; k < 8; k++) tmp[8 * k + c] = sat16(r[k]);
        }
        for (int row = 0; row < 8; row++, out += stride) {
            for (int k = 0; k < 8; k++) s[k] = tmp[8 * row + k];
            idct8(s, 65536 + (128 << 17), 17, r);
            for (int k = 0; k < 8; k++) out[k] = (uint8_t)(r[k] < 0 ? 0 : (r[k] > 255 ? 255 : r[k]));
        }
    }

    // after each MCU: count down to the next restart marker
    // returns 0 = go on, 1 = the scan ended (a marker that is no RSTn)
    int restart_step() {
        if (--todo > 0) return 0;
        if (nbits < 24) refill();
        if (!(marker >= 0xD0 && marker <= 0xD7)) return 1;
        reset_entropy();
        return 0;
    }

    bool scan_data() {
        reset_entropy();
        int16_t blk[64];
        if (scan_n == 1) {
            Component& c = comp[order[0]];
            const int bw = (c.x + 7) >> 3, bh = (c.y + 7) >> 3;
            for (int j = 0; j < bh; j++)
                for (int i = 0; i < bw; i++) {
                    if (!progressive) {
                        if (!sequential_block(blk, c)) return false;
                        idct_block(&c.pix[(size_t)c.w2 * j * 8 + i * 8], c.w2, blk);
                    } else {
                        int16_t* data = &c.coef[64 * ((size_t)i + (size_t)j * c.coef_w)];
                        if (ss == 0) {
                            if (!dc_first_or_refine(data, c)) return false;
                        } else if (!ac_first_or_refine(data, hac[c.ta])) {
                            return false;
                        }
                    }
                    if (restart_step()) return true;
                }
            return true;
        }
        for (int j = 0; j < mcu_y; j++)
            for (int i = 0; i < mcu_x; i++) {
                for (int k = 0; k < scan_n; k++) {
                    Component& c = comp[order[k]];
                    for (int y = 0; y < c.v; y++)
                        for (int x = 0; x < c.h; x++) {
                            const int bx = i * c.h + x, by = j * c.v + y;
                            if (!progressive) {
                                if (!sequential_block(blk, c)) return false;
                                idct_block(&c.pix[(size_t)c.w2 * by * 8 + bx * 8], c.w2, blk);
                            } else if (!dc_first_or_refine(&c.coef[64 * ((size_t)bx + (size_t)by * c.coef_w)], c)) {
                                return false;
                            }
                        }
                }
                if (restart_step()) return true;
            }
        return true;
    }

    // progressive frames: dequantise with the tables in force now, inverse DCT
    void finish_progressive() {
        for (int n = 0; n < n_comp; n++) {
            Component& c = comp[n];
            const int bw = (c.x + 7) >> 3, bh = (c.y + 7) >> 3;
            const uint16_t* q = dq[c.tq];
            for (int j = 0; j < bh; j++)
                for (int i = 0; i < bw; i++) {
                    int16_t* data = &c.coef[64 * ((size_t)i + (size_t)j * c.coef_w)];
                    for (int k = 0; k < 64; k++) data[k] = (int16_t)(data[k] * q[k]);
                    idct_block(&c.pix[(size_t)c.w2 * j * 8 + i * 8], c.w2, data);
                }
        }
    }

    // bytes after a scan's data up to the next real marker
    int skip_to_marker() {
        while (!s.eof()) {
            int x = s.u8();
            while (x == 0xff) {
                if (s.eof()) return kMarkerNone;
                x = s.u8();
                if (x != 0x00 && x != 0xff) return x;
            }
        }
        return kMarkerNone;
    }

    bool decode_all() {
        if (next_marker() != 0xD8) return fail("no SOI marker");
        int m = next_marker();
        while (!(m == 0xC0 || m == 0xC1 || m == 0xC2)) {
            if (!marker_segment(m)) return false;
            m = next_marker();
            while (m == kMarkerNone) {
                if (s.eof()) return fail("no SOF marker");
                m = next_marker();
            }
        }
        progressive = m == 0xC2;
        if (!frame_header()) return false;
        m = next_marker();
        while (m != 0xD9) {                                                     // EOI
            if (m == 0xDA) {                                                    // SOS
                if (!scan_header() || !scan_data()) return false;
                if (marker == kMarkerNone) marker = skip_to_marker();
                m = next_marker();
                if (m >= 0xD0 && m <= 0xD7) m = next_marker();
            } else if (m == 0xDC) {                                             // DNL
                const int Ld = s.be16();
                const int NL = s.be16();
                if (Ld != 4) return fail("bad DNL length");
                if (NL != H) return fail("bad DNL height");
                m = next_marker();
            } else {
                if (!marker_segment(m)) break;          // stop at what cannot be parsed; keep the image
                m = next_marker();
            }
        }
        if (progressive) finish_progressive();
        err.clear();
        return true;
    }
};

// ---- upsampling (one output row from the nearest and the far source row) ----
using Resample = const uint8_t* (*)(uint8_t* out, const uint8_t* near, const uint8_t* far, int w, int hs);

const uint8_t* up_none(uint8_t*, const uint8_t* near, const uint8_t*, int, int) { return near; }

const uint8_t* up_v2(uint8_t* out, const uint8_t* near, const uint8_t* far, int w, int) {
    for (int i = 0; i < w; i++) out[i] = (uint8_t)((3 * near[i] + far[i] + 2) >> 2);
    return out;
}

const uint8_t* up_h2(uint8_t* out, const uint8_t* in, const uint8_t*, int w, int) {
    if (w == 1) {
        out[0] = out[1] = in[0];
        return out;
    }
    out[0] = in[0];
    out[1] = (uint8_t)((3 * in[0] + in[1] + 2) >> 2);
    int i = 1;
    for (; i < w - 1; i++) {
        const int c = 3 * in[i] + 2;
        out[2 * i] = (uint8_t)((c + in[i - 1]) >> 2);
        out[2 * i + 1] = (uint8_t)((c + in[i + 1]) >> 2);
    }
    out[2 * i] = (uint8_t)((3 * in[w - 2] + in[w - 1] + 2) >> 2);
    out[2 * i + 1] = in[w - 1];
    return out;
}

const uint8_t* up_hv2(uint8_t* out, const uint8_t* near, const uint8_t* far, int w, int) {
    if (w == 1) {
        out[0] = out[1] = (uint8_t)((3 * near[0] + far[0] + 2) >> 2);
        return out;
    }
    int cur = 3 * near[0] + far[0];
    out[0] = (uint8_t)((cur + 2) >> 2);
    for (int i = 1; i < w; i++) {
        const int prev = cur;
        cur = 3 * near[i] + far[i];
        out[2 * i - 1] = (uint8_t)((3 * prev + cur + 8) >> 4);
        out[2 * i] = (uint8_t)((3 * cur + prev + 8) >> 4);
    }
    out[2 * w - 1] = (uint8_t)((cur + 2) >> 2);
    return out;
}

const uint8_t* up_nearest(uint8_t* out, const uint8_t* near, const uint8_t*, int w, int hs) {
    for (int i = 0; i < w; i++)
        for (int j = 0; j < hs; j++) out[i * hs + j] = near[i];
    return out;
}

// YCbCr -> RGB, 20-bit fixed point (constants rounded to 12 bits, shifted 8)
inline int fx(float c) { return ((int)(c * 4096.0f + 0.5f)) << 8; }
void ycc_to_rgb(uint8_t* out, const uint8_t* y, const uint8_t* cb, const uint8_t* cr, int n, int step) {
    static const int kr = fx(1.40200f), kg_cr = -fx(0.71414f), kg_cb = -fx(0.34414f), kb = fx(1.77200f);
    for (int i = 0; i < n; i++, out += step) {
        const int base = (y[i] << 20) + (1 << 19);
        const int r_ = cr[i] - 128, b_ = cb[i] - 128;
        int r = (base + r_ * kr) >> 20;
        int g = (base + r_ * kg_cr + (int)((unsigned)(b_ * kg_cb) & 0xffff0000u)) >> 20;
        int b = (base + b_ * kb) >> 20;
        out[0] = (uint8_t)(r < 0 ? 0 : (r > 255 ? 255 : r));
        out[1] = (uint8_t)(g < 0 ? 0 : (g > 255 ? 255 : g));
        out[2] = (uint8_t)(b < 0 ? 0 : (b > 255 ? 255 : b));
    }
}

// x * y / 255, rounded (CMYK and YCCK with the K channel)
inline uint8_t mul255(uint8_t x, uint8_t y) {
    const unsigned t = (unsigned)x * y + 128;
    return (uint8_t)((t + (t >> 8)) >> 8);
}

}  // namespace

// SOI after any 0xff fill bytes
bool is_jpeg(const uint8_t* d, size_t n) {
    if (n < 2 || d[0] != 0xff) return false;
    size_t i = 1;
    while (i < n && d[i] == 0xff) i++;
    return i < n && d[i] == 0xd8;
}

bool decode_jpeg(const uint8_t* d, size_t n, Raw8& out, std::string& err) {
    Decoder j(d, n);
    if (!j.decode_all()) { err = "JPEG: " + j.err; return false; }
    const int nc = j.n_comp;
    const int outc = nc >= 3 ? 3 : 1;
    const bool is_rgb = nc == 3 && (j.rgb_ids == 3 || (j.adobe == 0 && !j.jfif));
    struct Up {
        Resample fn;
        const uint8_t *line0, *line1;
        int hs, vs, ystep, ypos, w_lores;
        std::vector<uint8_t> buf;
    } up[4];
    for (int k = 0; k < nc; k++) {
        Up& u = up[k];
        const Component& c = j.comp[k];
        u.hs = j.hmax / c.h;
        u.vs = j.vmax / c.v;
        u.ystep = u.vs >> 1;
        u.w_lores = (j.W + u.hs - 1) / u.hs;
        u.ypos = 0;
        u.line0 = u.line1 = c.pix.data();
        u.buf.assign((size_t)j.W + 3, 0);
        if (u.hs == 1 && u.vs == 1) u.fn = up_none;
        else if (u.hs == 1 && u.vs == 2) u.fn = up_v2;
        else if (u.hs == 2 && u.vs == 1) u.fn = up_h2;
        else if (u.hs == 2 && u.vs == 2) u.fn = up_hv2;
        else u.fn = up_nearest;
    }
    out.C = out.Cout = outc;
    out.H = j.H;
    out.W = j.W;
    out.hwc.assign((size_t)outc * j.W * j.H, 0);
    const uint8_t* row[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int y = 0; y < j.H; y++) {
        uint8_t* o = &out.hwc[(size_t)outc * j.W * y];
        for (int k = 0; k < nc; k++) {
            Up& u = up[k];
            // the source row nearest to this output row's centre, and the one beyond it
            const bool lower = u.ystep >= (u.vs >> 1);
            row[k] = u.fn(u.buf.data(), lower ? u.line1 : u.line0, lower ? u.line0 : u.line1, u.w_lores, u.hs);
            if (++u.ystep >= u.vs) {
                u.ystep = 0;
                u.line0 = u.line1;
                if (++u.ypos < j.comp[k].y) u.line1 += j.comp[k].w2;
            }
        }
        if (outc == 3) {
            if (nc == 3) {
                if (is_rgb) {
                    for (int i = 0; i < j.W; i++) {
                        o[3 * i] = row[0][i];
                        o[3 * i + 1] = row[1][i];
                        o[3 * i + 2] = row[2][i];
                    }
                } else {
                    ycc_to_rgb(o, row[0], row[1], row[2], j.W, 3);
                }
            } else if (j.adobe == 0) {                                   // CMYK
                for (int i = 0; i < j.W; i++) {
                    const uint8_t kk = row[3][i];
                    o[3 * i] = mul255(row[0][i], kk);
                    o[3 * i + 1] = mul255(row[1][i], kk);
                    o[3 * i + 2] = mul255(row[2][i], kk);
                }
            } else if (j.adobe == 2) {                                   // YCCK
                ycc_to_rgb(o, row[0], row[1], row[2], j.W, 3);
                for (int i = 0; i < j.W; i++) {
                    const uint8_t kk = row[3][i];
                    for (int c = 0; c < 3; c++) o[3 * i + c] = mul255((uint8_t)(255 - o[3 * i + c]), kk);
                }
            } else {                                                     // YCbCr + a fourth channel, ignored
                ycc_to_rgb(o, row[0], row[1], row[2], j.W, 3);
            }
        } else {
            for (int i = 0; i < j.W; i++) o[i] = row[0][i];
        }
    }
    return true;
}

// ===========================================================================
// BMP
// ===========================================================================
namespace {

int top_bit(uint32_t z) {
    if (z == 0) return -1;
    int n = 0;
    while (z >>= 1) n++;
    return n;
}

int popcount32(uint32_t a) {
    int n = 0;
    for (; a; a &= a - 1) n++;
    return n;
}

// a mask field of `bits` bits whose top bit sits at `shift + 7`, widened to 8
// bits by repeating its bit pattern (x * (2^8 - 1) / (2^bits - 1), truncated)
int widen_field(uint32_t v, int shift, int bits) {
    static const unsigned mul[9] = {0, 0xff, 0x55, 0x49, 0x11, 0x21, 0x41, 0x81, 0x01};
    static const unsigned sh[9] = {0, 0, 0, 1, 0, 2, 4, 6, 0};
    v = shift < 0 ? v << -shift : v >> shift;
    v >>= (8 - bits);
    return (int)((v * mul[bits]) >> sh[bits]);
}

}  // namespace

bool is_bmp(const uint8_t* d, size_t n) {
    if (n < 18 || d[0] != 'B' || d[1] != 'M') return false;
    const uint32_t hsz = (uint32_t)d[14] | (uint32_t)d[15] << 8 | (uint32_t)d[16] << 16 | (uint32_t)d[17] << 24;
    return hsz == 12 || hsz == 40 || hsz == 56 || hsz == 108 || hsz == 124;
}

bool decode_bmp(const uint8_t* d, size_t n, Raw8& out, std::string& err) {
    Src s(d, n);
    auto bad = [&](const char* m) { err = std::string("BMP: ") + m; return false; };
    if (s.u8() != 'B' || s.u8() != 'M') return bad("not a BMP file");
    s.le32();                                   // file size
    s.le16();
    s.le16();
    const int32_t offset = (int32_t)s.le32();
    const uint32_t hsz = s.le32();
    if (offset < 0) return bad("bad data offset");
    if (hsz != 12 && hsz != 40 && hsz != 56 && hsz != 108 && hsz != 124) return bad("unknown header");
    int32_t Wd, Hd;
    if (hsz == 12) {
        Wd = s.le16();
        Hd = s.le16();
    } else {
        Wd = (int32_t)s.le32();
        Hd = (int32_t)s.le32();
    }
    if (s.le16() != 1) return bad("planes != 1");
    const int bpp = s.le16();
    uint32_t mr = 0, mg = 0, mb = 0, ma = 0;
    bool alpha_seen_zero_check = false;        // 32-bit BI_RGB: alpha all 0 means opaque
    int header_read = 14;                      // bytes before the info header the masks follow
    auto default_masks = [&](int compress) {
        if (compress != 0) return;
        if (bpp == 16) { mr = 31u << 10; mg = 31u << 5; mb = 31u; }
        else if (bpp == 32) { mr = 0xffu << 16; mg = 0xffu << 8; mb = 0xffu; ma = 0xffu << 24; alpha_seen_zero_check = true; }
        else mr = mg = mb = ma = 0;
    };
    if (hsz != 12) {
        const uint32_t compress = s.le32();
        if (compress == 1 || compress == 2) return bad("RLE compression is not supported");
        if (compress >= 4) return bad("JPEG/PNG compression is not supported");
        if (compress == 3 && bpp != 16 && bpp != 32) return bad("bit fields need 16 or 32 bits per pixel");
        for (int i = 0; i < 5; i++) s.le32();   // image size, resolutions, colours used / important
        if (hsz == 40 || hsz == 56) {
            if (hsz == 56) for (int i = 0; i < 4; i++) s.le32();
            if (bpp == 16 || bpp == 32) {
                if (compress == 0) {
                    default_masks(0);
                } else {
                    mr = s.le32();
                    mg = s.le32();
                    mb = s.le32();
                    header_read += 12;
                    if (mr == mg && mg == mb) return bad("bad bit fields");
                }
            }
        } else {
            mr = s.le32();
            mg = s.le32();
            mb = s.le32();
            ma = s.le32();
            if (compress != 3) default_masks((int)compress);
            for (int i = 0; i < 13; i++) s.le32();         // colour space and its endpoints, gamma
            if (hsz == 124) for (int i = 0; i < 4; i++) s.le32();
        }
    }
    const bool bottom_up = Hd > 0;
    const uint32_t W = (uint32_t)Wd, H = (uint32_t)std::abs(Hd);
    if (H > MAX_DIM || W > MAX_DIM) return bad("image too large");
    int psize = 0;
    if (hsz == 12) {
        if (bpp < 24) psize = (offset - header_read - 24) / 3;
    } else {
        if (bpp < 16) psize = (offset - header_read - (int)hsz) >> 2;
    }
    if (psize == 0) {
        const long so_far = (long)s.p;
        if (so_far <= 0 || so_far > 1024) return bad("bad header");
        if (offset < so_far || offset - so_far > 1024) return bad("bad data offset");
        s.skip(offset - so_far);
    }
    const int C = (bpp == 24 && ma == 0xff000000u) ? 3 : (ma ? 4 : 3);
    if ((uint64_t)W * H * C > 0x7fffffffull) return bad("image too large");
    out.C = out.Cout = C;
    out.W = (int)W;
    out.H = (int)H;
    out.hwc.assign((size_t)C * W * H, 0);
    uint8_t* o = out.hwc.data();
    size_t z = 0;
    unsigned all_a = alpha_seen_zero_check ? 0 : 255;
    if (bpp < 16) {
        if (psize == 0 || psize > 256) return bad("bad palette");
        uint8_t pal[256][4];
        for (int i = 0; i < psize; i++) {
            pal[i][2] = (uint8_t)s.u8();
            pal[i][1] = (uint8_t)s.u8();
            pal[i][0] = (uint8_t)s.u8();
            if (hsz != 12) s.u8();
            pal[i][3] = 255;
        }
        s.skip(offset - header_read - (long)hsz - psize * (hsz == 12 ? 3 : 4));
        int width;
        if (bpp == 1) width = (int)((W + 7) >> 3);
        else if (bpp == 4) width = (int)((W + 1) >> 1);
        else if (bpp == 8) width = (int)W;
        else return bad("bad bits per pixel");
        const int pad = (-width) & 3;
        auto put = [&](int idx) {
            o[z++] = pal[idx][0];
            o[z++] = pal[idx][1];
            o[z++] = pal[idx][2];
            if (C == 4) o[z++] = 255;
        };
        for (uint32_t j = 0; j < H; j++) {
            if (bpp == 1) {
                int bit = 7, v = s.u8();
                for (uint32_t i = 0; i < W; i++) {
                    put((v >> bit) & 1);
                    if (i + 1 == W) break;
                    if (--bit < 0) { bit = 7; v = s.u8(); }
                }
            } else {
                for (uint32_t i = 0; i < W; i += 2) {
                    int v = s.u8(), v2 = 0;
                    if (bpp == 4) { v2 = v & 15; v >>= 4; }
                    put(v);
                    if (i + 1 == W) break;
                    put(bpp == 8 ? s.u8() : v2);
                }
            }
            s.skip(pad);
        }
    } else {
        s.skip(offset - header_read - (long)hsz);
        const int width = bpp == 24 ? 3 * (int)W : (bpp == 16 ? 2 * (int)W : 0);
        const int pad = (-width) & 3;
        int easy = 0;
        if (bpp == 24) easy = 1;
        else if (bpp == 32 && mb == 0xff && mg == 0xff00 && mr == 0xff0000 && ma == 0xff000000u) easy = 2;
        int rs = 0, gs = 0, bs = 0, as = 0, rc = 0, gc = 0, bc = 0, ac = 0;
        if (!easy) {
            if (!mr || !mg || !mb) return bad("bad bit fields");
            rs = top_bit(mr) - 7; rc = popcount32(mr);
            gs = top_bit(mg) - 7; gc = popcount32(mg);
            bs = top_bit(mb) - 7; bc = popcount32(mb);
            as = top_bit(ma) - 7; ac = popcount32(ma);
            if (rc > 8 || gc > 8 || bc > 8 || ac > 8) return bad("bad bit fields");
        }
        for (uint32_t j = 0; j < H; j++) {
            for (uint32_t i = 0; i < W; i++) {
                unsigned a;
                if (easy) {
                    o[z + 2] = (uint8_t)s.u8();
                    o[z + 1] = (uint8_t)s.u8();
                    o[z] = (uint8_t)s.u8();
                    z += 3;
                    a = easy == 2 ? (unsigned)s.u8() : 255u;
                } else {
                    const uint32_t v = bpp == 16 ? (uint32_t)s.le16() : s.le32();
                    o[z++] = (uint8_t)widen_field(v & mr, rs, rc);
                    o[z++] = (uint8_t)widen_field(v & mg, gs, gc);
                    o[z++] = (uint8_t)widen_field(v & mb, bs, bc);
                    a = ma ? (unsigned)widen_field(v & ma, as, ac) : 255u;
                }
                all_a |= a;
                if (C == 4) o[z++] = (uint8_t)a;
            }
            s.skip(pad);
        }
    }
    if (C == 4 && all_a == 0)
        for (size_t i = 3; i < out.hwc.size(); i += 4) o[i] = 255;
    if (bottom_up) {
        const size_t row = (size_t)W * C;
        for (uint32_t j = 0; j < H / 2; j++) std::swap_ranges(o + j * row, o + (j + 1) * row, o + (H - 1 - j) * row);
    }
    return true;
}

// ===========================================================================
// TGA
// ===========================================================================
namespace {

// channels of a TGA pixel format: 8 grey, 16 grey+alpha (grey images) or
// 5:5:5 colour, 15 5:5:5 colour, 24 / 32 BGR(A); 0 = unsupported
int tga_channels(int bits, bool grey, bool& rgb16) {
    rgb16 = false;
    switch (bits) {
        case 8: return 1;
        case 16:
            if (grey) return 2;
            rgb16 = true;
            return 3;
        case 15: rgb16 = true; return 3;
        case 24: return 3;
        case 32: return 4;
        default: return 0;
    }
}

void tga_rgb16(Src& s, uint8_t* out) {
    const int px = s.le16();
    out[0] = (uint8_t)((((px >> 10) & 31) * 255) / 31);
    out[1] = (uint8_t)((((px >> 5) & 31) * 255) / 31);
    out[2] = (uint8_t)(((px & 31) * 255) / 31);
}

}  // namespace

bool is_tga(const uint8_t* d, size_t n) {
    Src s(d, n);
    s.u8();                                     // id length
    const int cmap = s.u8();
    if (cmap > 1) return false;
    int type = s.u8();
    if (cmap == 1) {
        if (type != 1 && type != 9) return false;
        s.skip(4);
        const int eb = s.u8();
        if (eb != 8 && eb != 15 && eb != 16 && eb != 24 && eb != 32) return false;
        s.skip(4);
    } else {
        if (type != 2 && type != 3 && type != 10 && type != 11) return false;
        s.skip(9);
    }
    if (s.le16() < 1 || s.le16() < 1) return false;
    const int bpp = s.u8();
    if (cmap == 1 && bpp != 8 && bpp != 16) return false;
    return bpp == 8 || bpp == 15 || bpp == 16 || bpp == 24 || bpp == 32;
}

bool decode_tga(const uint8_t* d, size_t n, Raw8& out, std::string& err) {
    Src s(d, n);
    auto bad = [&](const char* m) { err = std::string("TGA: ") + m; return false; };
    const int id_len = s.u8();
    const int indexed = s.u8();
    int type = s.u8();
    const int pal_start = s.le16();
    const int pal_len = s.le16();
    const int pal_bits = s.u8();
    s.le16();                                   // x origin
    s.le16();                                   // y origin
    const int W = s.le16();
    const int H = s.le16();
    const int bpp = s.u8();
    const int desc = s.u8();
    bool rle = false;
    if (type >= 8) { type -= 8; rle = true; }
    const bool flip = ((desc >> 5) & 1) == 0;   // origin at the bottom: rows stored bottom-up
    bool rgb16 = false;
    const int C = indexed ? tga_channels(pal_bits, false, rgb16) : tga_channels(bpp, type == 3, rgb16);
    if (!C) return bad("unsupported pixel format");
    if ((uint64_t)W * H * C > 0x7fffffffull) return bad("image too large");
    out.C = out.Cout = C;
    out.W = W;
    out.H = H;
    out.hwc.assign((size_t)W * H * C, 0);
    uint8_t* o = out.hwc.data();
    s.skip(id_len);
    if (!indexed && !rle && !rgb16) {
        for (int i = 0; i < H; i++) {
            const int row = flip ? H - i - 1 : i;
            uint8_t* dst = o + (size_t)row * W * C;
            const size_t want = (size_t)W * C;
            if (s.p + want > s.n) return bad("truncated pixel data");
            std::memcpy(dst, s.d + s.p, want);
            s.p += want;
        }
    } else {
        std::vector<uint8_t> pal;
        if (indexed) {
            if (pal_len == 0) return bad("empty colour map");
            s.skip(pal_start);
            pal.assign((size_t)pal_len * C, 0);
            if (rgb16) {
                for (int i = 0; i < pal_len; i++) tga_rgb16(s, &pal[(size_t)i * C]);
            } else {
                if (s.p + pal.size() > s.n) return bad("truncated colour map");
                std::memcpy(pal.data(), s.d + s.p, pal.size());
                s.p += pal.size();
            }
        }
        uint8_t px[4] = {0, 0, 0, 0};
        int count = 0;
        bool repeat = false, fetch = true;
        for (size_t i = 0; i < (size_t)W * H; i++) {
            if (rle) {
                if (count == 0) {
                    const int cmd = s.u8();
                    count = 1 + (cmd & 127);
                    repeat = (cmd >> 7) != 0;
                    fetch = true;
                } else if (!repeat) {
                    fetch = true;
                }
            } else {
                fetch = true;
            }
            if (fetch) {
                if (indexed) {
                    int idx = bpp == 8 ? s.u8() : s.le16();
                    if (idx >= pal_len) idx = 0;
                    for (int c = 0; c < C; c++) px[c] = pal[(size_t)idx * C + c];
                } else if (rgb16) {
                    tga_rgb16(s, px);
                } else {
                    for (int c = 0; c < C; c++) px[c] = (uint8_t)s.u8();
                }
                fetch = false;
            }
            for (int c = 0; c < C; c++) o[i * C + c] = px[c];
            count--;
        }
        if (flip) {
            const size_t row = (size_t)W * C;
            for (int j = 0; 2 * j < H; j++) {
                if (j == H - 1 - j) continue;
                std::swap_ranges(o + j * row, o + (j + 1) * row, o + (size_t)(H - 1 - j) * row);
            }
        }
    }
    if (C >= 3 && !rgb16)                       // BGR(A) -> RGB(A)
        for (size_t i = 0; i < (size_t)W * H; i++) std::swap(o[i * C], o[i * C + 2]);
    return true;
}

}  // namespace nrio
