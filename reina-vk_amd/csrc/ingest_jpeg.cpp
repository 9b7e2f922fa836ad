// ingest_jpeg.cpp — baseline JPEG → RGBA8, host only.
//
// The JPEG half of the stb_image calls of src/graphics/Image.cpp:10-41
// (stbi_load with 4 channels; stb is fetched at `master`,
// CMakeLists.txt:21-26). Entropy decoding follows the JPEG standard
// (ITU T.81: Huffman tables, DC prediction, restart intervals), which fixes
// the coefficients. Where the pixels depend on arithmetic, this file restates
// stb_image's published algorithm, so that the bytes are stb's:
//   * stb's integer IDCT: constants scaled by 4096 and rounded, columns first
//     (a column with only a DC term becomes 4 * DC), + 512 >> 10 after the
//     columns, + 65536 + (128 << 17) >> 17 and a clamp after the rows;
//   * chroma upsampling: 2x horizontal (3:1 triangle, edges copied), 2x
//     vertical (3 * near + far + 2 >> 2), 2x2 (3 * near + far per column,
//     then (3 * a + b + 8) >> 4 between columns, (t + 2) >> 2 at the edges),
//     nearest for other factors; output row r uses input row r / vs as "near"
//     and the row above or below as "far";
//   * YCbCr -> RGB in fixed point: y << 20 + 2^19, coefficients
//     (int(c * 4096 + 0.5) << 8), the Cb term of green masked to its upper 16
//     bits, >> 20 and a clamp; components named 'R', 'G', 'B' (or an Adobe
//     marker with transform 0 in a file without JFIF marker) are taken as RGB;
//   * grey as (y, y, y, 255).
//   * progressive files (ITU-T T.81 Annex G): every scan's coefficients
//     accumulate per block, then the blocks are dequantised and transformed
//     as above.
//   * 4-component files as stb: Adobe transform 0 (CMYK) gives each colour
//     channel times K (x * y / 255 rounded as stb's blinn_8x8), transform 2
//     (YCCK) converts YCbCr first and multiplies (255 - channel) by K, any
//     other file ignores the fourth channel.
// Lossless, arithmetic-coded, differential and 12-bit JPEGs return
// REINA_RT_ERR_UNSUPPORTED (stb refuses them too).
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/reina_rt.h"

namespace rt {
// Decodes a baseline JPEG into RGBA8 (rows flipped when `flip`); throws
// JpegError on failure.
struct JpegError {
    int code;
    std::string msg;
};
std::vector<uint8_t> decode_jpeg(const uint8_t* p, uint64_t n, bool flip, uint32_t& w, uint32_t& h);
}  // namespace rt

namespace {

using rt::JpegError;
[[noreturn]] void bad(const std::string& m) { throw JpegError{REINA_RT_ERR_INVALID_ARG, "JPEG: " + m}; }
[[noreturn]] void unsupported(const std::string& m) { throw JpegError{REINA_RT_ERR_UNSUPPORTED, "JPEG: " + m}; }

const uint8_t kDezigzag[64 + 15] = {
    0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
    41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
    30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63,
    // entries past 63 guard run lengths that overshoot a corrupt block
    63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63};

struct Huffman {
    // canonical code: for each length L, codes [first[L], first[L] + count[L])
    // map to symbols starting at index start[L]
    int count[17] = {0}, first[17] = {0}, start[17] = {0};
    uint8_t symbols[256] = {0};
    bool defined = false;
};

struct Component {
    int id = 0, h = 1, v = 1, tq = 0;
    int td = 0, ta = 0;
    int x = 0, y = 0;             // samples of the component
    int w2 = 0, h2 = 0;           // padded to whole MCUs
    int dc_pred = 0;
    std::vector<uint8_t> data;    // w2 * h2 samples
    std::vector<short> coef;      // progressive: (w2 / 8) * (h2 / 8) blocks of 64, natural order
};

struct Decoder {
    const uint8_t* p = nullptr;
    uint64_t n = 0, pos = 0;
    uint16_t dq[4][64] = {{0}};   // natural order
    bool dq_defined[4] = {false, false, false, false};
    Huffman hdc[4], hac[4];
    std::vector<Component> comp;
    int width = 0, height = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0;
    int restart = 0;
    int adobe_transform = -1;
    bool jfif = false;
    bool frame = false;
    bool progressive = false;
    int eobrun = 0;   // progressive AC scans: blocks left in the current end-of-band run
    // entropy-coded bit reader
    uint32_t bits = 0;
    int nbits = 0;
    bool hit_marker = false;

    uint8_t u8() {
        if (pos >= n) bad("truncated file");
        return p[pos++];
    }
    int u16() {
        const int a = u8();
        return (a << 8) | u8();
    }

    void define_quant(uint64_t end) {
        while (pos < end) {
            const int pq_tq = u8();
            const int pq = pq_tq >> 4, tq = pq_tq & 15;
            if (tq > 3) bad("bad quantization table id");
            for (int i = 0; i < 64; ++i) dq[tq][kDezigzag[i]] = (uint16_t)(pq ? u16() : u8());
            dq_defined[tq] = true;
        }
    }

    void define_huffman(uint64_t end) {
        while (pos < end) {
            const int tc_th = u8();
            const int tc = tc_th >> 4, th = tc_th & 15;
            if (tc > 1 || th > 3) bad("bad Huffman table id");
            Huffman& t = tc == 0 ? hdc[th] : hac[th];
            int total = 0;
            for (int l = 1; l <= 16; ++l) {
                t.count[l] = u8();
                total += t.count[l];
            }
            if (total > 256) bad("Huffman table too large");
            for (int i = 0; i < total; ++i) t.symbols[i] = u8();
            int code = 0, k = 0;
            for (int l = 1; l <= 16; ++l) {
                t.first[l] = code;
                t.start[l] = k;
                code += t.count[l];
                k += t.count[l];
                if (code > (1 << l)) bad("bad Huffman code lengths");
                code <<= 1;
            }
            t.defined = true;
        }
    }

    void start_of_frame(int marker, uint64_t end) {
        if (frame) bad("more than one frame");
        if (marker == 0xC6 || marker == 0xCA || marker == 0xCE) unsupported("differential JPEG");
        if (marker != 0xC0 && marker != 0xC1 && marker != 0xC2) unsupported("lossless or arithmetic-coded JPEG");
        progressive = marker == 0xC2;
        if (u8() != 8) unsupported("only 8-bit samples");
        height = u16();
        width = u16();
        if (height == 0) bad("zero height (DNL marker not supported)");
        if (width == 0) bad("zero width");
        if ((uint64_t)width * height > (1ull << 28)) bad("image too large");
        const int nc = u8();
        if (nc != 1 && nc != 3 && nc != 4) unsupported(std::to_string(nc) + "-component JPEG");
        comp.resize((size_t)nc);
        for (Component& c : comp) {
            c.id = u8();
            const int hv = u8();
            c.h = hv >> 4;
            c.v = hv & 15;
            c.tq = u8();
            if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4) bad("bad sampling factor");
            if (c.tq > 3) bad("bad quantization table id");
        }
        if (pos != end) bad("bad frame header length");
        for (const Component& c : comp) {
            hmax = std::max(hmax, c.h);
            vmax = std::max(vmax, c.v);
        }
        for (const Component& c : comp)
            if (hmax % c.h || vmax % c.v) unsupported("non-integer subsampling ratio");
        mcux = (width + 8 * hmax - 1) / (8 * hmax);
        mcuy = (height + 8 * vmax - 1) / (8 * vmax);
        for (Component& c : comp) {
            c.x = (width * c.h + hmax - 1) / hmax;
            c.y = (height * c.v + vmax - 1) / vmax;
            c.w2 = mcux * c.h * 8;
            c.h2 = mcuy * c.v * 8;
            c.data.assign((size_t)c.w2 * c.h2, 0);
            if (progressive) c.coef.assign((size_t)c.w2 * c.h2, 0);   // 64 per 8x8 block
        }
        frame = true;
    }

    // ---- entropy-coded data ----
    void fill() {
        while (nbits <= 24) {
            uint32_t b = 0;
            if (!hit_marker && pos < n) {
                b = p[pos];
                if (b == 0xFF) {
                    uint8_t c = pos + 1 < n ? p[pos + 1] : 0;
                    while (c == 0xFF && pos + 2 < n) {   // fill bytes
                        ++pos;
                        c = p[pos + 1];
                    }
                    if (c == 0x00) {
                        pos += 2;
                    } else {
                        hit_marker = true;   // leave the marker for the caller; feed zeros
                        b = 0;
                    }
                } else {
                    ++pos;
                }
            }
            bits |= b << (24 - nbits);
            nbits += 8;
        }
    }
    int bit() {
        if (nbits < 1) fill();
        const int b = (int)(bits >> 31);
        bits <<= 1;
        --nbits;
        return b;
    }
    int receive(int s) {   // s raw bits
        if (s == 0) return 0;
        if (nbits < s) fill();
        const int v = (int)(bits >> (32 - s));
        bits <<= s;
        nbits -= s;
        return v;
    }
    static int extend(int v, int s) { return v < (1 << (s - 1)) ? v - (1 << s) + 1 : v; }
    int decode(const Huffman& t) {
        if (!t.defined) bad("undefined Huffman table");
        int code = 0;
        for (int l = 1; l <= 16; ++l) {
            code = (code << 1) | bit();
            const int off = code - t.first[l];
            if (off >= 0 && off < t.count[l]) return t.symbols[t.start[l] + off];
        }
        bad("bad Huffman code");
    }
    void reset_bits() {
        bits = 0;
        nbits = 0;
        hit_marker = false;
    }

    void decode_block(Component& c, short* blk) {
        std::memset(blk, 0, 64 * sizeof(short));
        const uint16_t* q = dq[c.tq];
        const int t = decode(hdc[c.td]);
        if (t > 15) bad("bad DC magnitude");
        const int diff = t ? extend(receive(t), t) : 0;
        const int dc = c.dc_pred + diff;
        c.dc_pred = dc;
        blk[0] = (short)(dc * q[0]);
        int k = 1;
        while (k < 64) {
            const int rs = decode(hac[c.ta]);
            const int s = rs & 15, r = rs >> 4;
            if (s == 0) {
                if (rs != 0xF0) break;   // end of block
                k += 16;
            } else {
                k += r;
                if (k > 63) bad("AC run past the block");
                const int zig = kDezigzag[k++];
                blk[zig] = (short)(extend(receive(s), s) * q[zig]);
            }
        }
    }

    // ---- progressive scans (ITU-T T.81 Annex G.1.2): coefficients accumulate in
    // c.coef over the scans; dequantisation and the IDCT run after the last ----
    void prog_dc(Component& c, short* coef, int ah, int al) {
        if (ah == 0) {   // first scan: the DC difference, scaled by 2^Al
            const int t = decode(hdc[c.td]);
            if (t > 15) bad("bad DC magnitude");
            c.dc_pred += t ? extend(receive(t), t) : 0;
            coef[0] = (short)(c.dc_pred * (1 << al));
        } else if (bit()) {   // refinement: one bit
            coef[0] = (short)(coef[0] | (1 << al));
        }
    }
    void prog_ac_first(Component& c, short* coef, int ss, int se, int al) {
        if (eobrun > 0) { --eobrun; return; }
        for (int k = ss; k <= se;) {
            const int rs = decode(hac[c.ta]);
            const int s = rs & 15, r = rs >> 4;
            if (s == 0) {
                if (r < 15) {   // end of band: this block and 2^r - 1 + (r bits) more
                    eobrun = (1 << r) - 1 + (r ? receive(r) : 0);
                    return;
                }
                k += 16;
            } else {
                k += r;
                if (k > se) bad("AC run past the band");
                coef[kDezigzag[k++]] = (short)(extend(receive(s), s) * (1 << al));
            }
        }
    }
    void refine(short& v, int p1) {   // correction bit of an already nonzero coefficient
        if (bit() && (v & p1) == 0) v = (short)(v >= 0 ? v + p1 : v - p1);
    }
    void prog_ac_refine(Component& c, short* coef, int ss, int se, int al) {
        const int p1 = 1 << al;
        int k = ss;
        if (eobrun == 0) {
            for (; k <= se; ++k) {
                const int rs = decode(hac[c.ta]);
                int s = rs & 15, r = rs >> 4;
                int value = 0;
                if (s != 0) {
                    if (s != 1) bad("bad refinement magnitude");
                    value = bit() ? p1 : -p1;
                } else if (r != 15) {   // end of band
                    eobrun = (1 << r) + (r ? receive(r) : 0);
                    break;
                }
                // pass over r zero coefficients (refining the nonzero ones on the way),
                // then place the new one on the next zero coefficient
                for (; k <= se; ++k) {
                    short& v = coef[kDezigzag[k]];
                    if (v != 0) {
                        refine(v, p1);
                    } else {
                        if (r == 0) break;
                        --r;
                    }
                }
                if (value != 0) {
                    if (k > se) bad("AC refinement past the band");
                    coef[kDezigzag[k]] = (short)value;
                }
            }
        }
        if (eobrun > 0) {   // inside an end-of-band run: refine the nonzero coefficients left
            for (; k <= se; ++k) {
                short& v = coef[kDezigzag[k]];
                if (v != 0) refine(v, p1);
            }
            --eobrun;
        }
    }

    void scan(uint64_t end) {
        if (!frame) bad("scan before the frame header");
        const int ns = u8();
        if (ns < 1 || ns > (int)comp.size()) bad("bad scan component count");
        std::vector<Component*> sc;
        for (int i = 0; i < ns; ++i) {
            const int id = u8(), t = u8();
            Component* found = nullptr;
            for (Component& c : comp)
                if (c.id == id) found = &c;
            if (!found) bad("scan names an unknown component");
            found->td = t >> 4;
            found->ta = t & 15;
            if (found->td > 3 || found->ta > 3) bad("bad Huffman table id");
            sc.push_back(found);
        }
        const int ss = u8(), se = u8(), a = u8();
        const int ah = a >> 4, al = a & 15;
        if (!progressive && (ss != 0 || se != 63 || a != 0)) bad("baseline scan with spectral selection");
        if (progressive) {
            if (ss == 0 ? se != 0 : (se < ss || se > 63 || ns != 1)) bad("bad progressive scan bands");
            if (al > 13 || ah > 13) bad("bad successive approximation");
        }
        if (pos != end) bad("bad scan header length");
        for (Component* c : sc) {
            if (!progressive && !dq_defined[c->tq]) bad("undefined quantization table");
            c->dc_pred = 0;
        }
        eobrun = 0;
        reset_bits();
        short blk[64];
        const auto put = [&](Component& c, int bx, int by) {
            if (progressive) {
                short* coef = c.coef.data() + ((size_t)by * (c.w2 / 8) + bx) * 64;
                if (ss == 0) prog_dc(c, coef, ah, al);
                else if (ah == 0) prog_ac_first(c, coef, ss, se, al);
                else prog_ac_refine(c, coef, ss, se, al);
                return;
            }
            decode_block(c, blk);
            idct(c.data.data() + (size_t)by * 8 * c.w2 + (size_t)bx * 8, c.w2, blk);
        };
        int todo = restart ? restart : 0x7FFFFFFF;
        const auto after_mcu = [&](bool last) {
            if (--todo > 0 || last) return;
            // expect RSTn: drop the bit buffer, find the marker, reset predictions
            reset_bits();
            while (pos + 1 < n && !(p[pos] == 0xFF && p[pos + 1] >= 0xD0 && p[pos + 1] <= 0xD7)) ++pos;
            if (pos + 1 >= n) bad("missing restart marker");
            pos += 2;
            for (Component* c : sc) c->dc_pred = 0;
            eobrun = 0;
            todo = restart;
        };
        if (ns == 1) {   // non-interleaved: the component's own block grid
            Component& c = *sc[0];
            const int bw = (c.x + 7) / 8, bh = (c.y + 7) / 8;
            for (int by = 0; by < bh; ++by)
                for (int bx = 0; bx < bw; ++bx) {
                    put(c, bx, by);
                    after_mcu(by == bh - 1 && bx == bw - 1);
                }
        } else {
            for (int my = 0; my < mcuy; ++my)
                for (int mx = 0; mx < mcux; ++mx) {
                    for (Component* c : sc)
                        for (int yy = 0; yy < c->v; ++yy)
                            for (int xx = 0; xx < c->h; ++xx) put(*c, mx * c->h + xx, my * c->v + yy);
                    after_mcu(my == mcuy - 1 && mx == mcux - 1);
                }
        }
        // continue after the entropy-coded segment: the next marker
        while (pos + 1 < n && !(p[pos] == 0xFF && p[pos + 1] != 0x00 && !(p[pos + 1] >= 0xD0 && p[pos + 1] <= 0xD7)))
            ++pos;
        reset_bits();
    }

    // dequantise and transform every block once all scans are in
    void finish_progressive() {
        short blk[64];
        for (Component& c : comp) {
            if (!dq_defined[c.tq]) bad("undefined quantization table");
            const uint16_t* q = dq[c.tq];
            const int bw = c.w2 / 8, bh = c.h2 / 8;
            for (int by = 0; by < bh; ++by)
                for (int bx = 0; bx < bw; ++bx) {
                    const short* coef = c.coef.data() + ((size_t)by * bw + bx) * 64;
                    for (int i = 0; i < 64; ++i) blk[i] = (short)(coef[i] * q[i]);
                    idct(c.data.data() + (size_t)by * 8 * c.w2 + (size_t)bx * 8, c.w2, blk);
                }
        }
    }

    // ---- stb's integer IDCT ----
    static int f2f(float x) { return (int)(x * 4096 + 0.5f); }
    static uint8_t clamp(int x) { return (uint8_t)((unsigned)x > 255 ? (x < 0 ? 0 : 255) : x); }
    struct Idct1d {
        int t0, t1, t2, t3, x0, x1, x2, x3;
        Idct1d(int s0, int s1, int s2, int s3, int s4, int s5, int s6, int s7) {
            int p1, p2, p3, p4, p5;
            p2 = s2;
            p3 = s6;
            p1 = (p2 + p3) * f2f(0.5411961f);
            t2 = p1 + p3 * f2f(-1.847759065f);
            t3 = p1 + p2 * f2f(0.765366865f);
            p2 = s0;
            p3 = s4;
            t0 = (p2 + p3) * 4096;
            t1 = (p2 - p3) * 4096;
            x0 = t0 + t3;
            x3 = t0 - t3;
            x1 = t1 + t2;
            x2 = t1 - t2;
            t0 = s7;
            t1 = s5;
            t2 = s3;
            t3 = s1;
            p3 = t0 + t2;
            p4 = t1 + t3;
            p1 = t0 + t3;
            p2 = t1 + t2;
            p5 = (p3 + p4) * f2f(1.175875602f);
            t0 = t0 * f2f(0.298631336f);
            t1 = t1 * f2f(2.053119869f);
            t2 = t2 * f2f(3.072711026f);
            t3 = t3 * f2f(1.501321110f);
            p1 = p5 + p1 * f2f(-0.899976223f);
            p2 = p5 + p2 * f2f(-2.562915447f);
            p3 = p3 * f2f(-1.961570560f);
            p4 = p4 * f2f(-0.390180644f);
            t3 += p1 + p4;
            t2 += p2 + p3;
            t1 += p2 + p4;
            t0 += p1 + p3;
        }
    };
    static void idct(uint8_t* out, int stride, const short* d) {
        int val[64];
        for (int i = 0; i < 8; ++i) {
            const short* c = d + i;
            int* v = val + i;
            if (c[8] == 0 && c[16] == 0 && c[24] == 0 && c[32] == 0 && c[40] == 0 && c[48] == 0 && c[56] == 0) {
                const int dcterm = c[0] * 4;
                for (int r = 0; r < 8; ++r) v[8 * r] = dcterm;
            } else {
                Idct1d k(c[0], c[8], c[16], c[24], c[32], c[40], c[48], c[56]);
                k.x0 += 512; k.x1 += 512; k.x2 += 512; k.x3 += 512;
                v[0] = (k.x0 + k.t3) >> 10;
                v[56] = (k.x0 - k.t3) >> 10;
                v[8] = (k.x1 + k.t2) >> 10;
                v[48] = (k.x1 - k.t2) >> 10;
                v[16] = (k.x2 + k.t1) >> 10;
                v[40] = (k.x2 - k.t1) >> 10;
                v[24] = (k.x3 + k.t0) >> 10;
                v[32] = (k.x3 - k.t0) >> 10;
            }
        }
        for (int i = 0; i < 8; ++i) {
            const int* v = val + 8 * i;
            uint8_t* o = out + (size_t)i * stride;
            Idct1d k(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]);
            const int bias = 65536 + (128 << 17);
            k.x0 += bias; k.x1 += bias; k.x2 += bias; k.x3 += bias;
            o[0] = clamp((k.x0 + k.t3) >> 17);
            o[7] = clamp((k.x0 - k.t3) >> 17);
            o[1] = clamp((k.x1 + k.t2) >> 17);
            o[6] = clamp((k.x1 - k.t2) >> 17);
            o[2] = clamp((k.x2 + k.t1) >> 17);
            o[5] = clamp((k.x2 - k.t1) >> 17);
            o[3] = clamp((k.x3 + k.t0) >> 17);
            o[4] = clamp((k.x3 - k.t0) >> 17);
        }
    }
};

// ---- stb's upsampling kernels: w input samples -> w * hs output samples ----
const uint8_t* row_1(uint8_t*, const uint8_t* near, const uint8_t*, int, int) { return near; }

const uint8_t* row_v2(uint8_t* out, const uint8_t* near, const uint8_t* far, int w, int) {
    for (int i = 0; i < w; ++i) out[i] = (uint8_t)((3 * near[i] + far[i] + 2) >> 2);
    return out;
}

const uint8_t* row_h2(uint8_t* out, const uint8_t* in, const uint8_t*, int w, int) {
    if (w == 1) {
        out[0] = out[1] = in[0];
        return out;
    }
    out[0] = in[0];
    out[1] = (uint8_t)((in[0] * 3 + in[1] + 2) >> 2);
    int i;
    for (i = 1; i < w - 1; ++i) {
        const int n = 3 * in[i] + 2;
        out[i * 2 + 0] = (uint8_t)((n + in[i - 1]) >> 2);
        out[i * 2 + 1] = (uint8_t)((n + in[i + 1]) >> 2);
    }
    out[i * 2 + 0] = (uint8_t)((in[w - 2] * 3 + in[w - 1] + 2) >> 2);
    out[i * 2 + 1] = in[w - 1];
    return out;
}

const uint8_t* row_hv2(uint8_t* out, const uint8_t* near, const uint8_t* far, int w, int) {
    if (w == 1) {
        out[0] = out[1] = (uint8_t)((3 * near[0] + far[0] + 2) >> 2);
        return out;
    }
    int t1 = 3 * near[0] + far[0];
    out[0] = (uint8_t)((t1 + 2) >> 2);
    for (int i = 1; i < w; ++i) {
        const int t0 = t1;
        t1 = 3 * near[i] + far[i];
        out[i * 2 - 1] = (uint8_t)((3 * t0 + t1 + 8) >> 4);
        out[i * 2] = (uint8_t)((3 * t1 + t0 + 8) >> 4);
    }
    out[w * 2 - 1] = (uint8_t)((t1 + 2) >> 2);
    return out;
}

const uint8_t* row_generic(uint8_t* out, const uint8_t* near, const uint8_t*, int w, int hs) {
    for (int i = 0; i < w; ++i)
        for (int j = 0; j < hs; ++j) out[i * hs + j] = near[i];
    return out;
}

int fix(float x) { return ((int)(x * 4096.0f + 0.5f)) << 8; }

}  // namespace

namespace rt {

std::vector<uint8_t> decode_jpeg(const uint8_t* p, uint64_t n, bool flip, uint32_t& w_out, uint32_t& h_out) {
    Decoder d;
    d.p = p;
    d.n = n;
    if (n < 4 || p[0] != 0xFF || p[1] != 0xD8) bad("no SOI marker");
    d.pos = 2;
    bool eoi = false;
    while (!eoi) {
        // next marker (fill bytes allowed)
        uint8_t m = 0;
        do {
            if (d.u8() != 0xFF) bad("expected a marker");
            m = d.u8();
            while (m == 0xFF) m = d.u8();
        } while (m == 0x00);
        if (m == 0xD9) { eoi = true; break; }
        if (m == 0x01 || (m >= 0xD0 && m <= 0xD7)) continue;   // TEM / stray RSTn
        const uint64_t start = d.pos;
        const int len = d.u16();
        if (len < 2 || start + (uint64_t)len > n) bad("bad segment length");
        const uint64_t end = start + (uint64_t)len;
        switch (m) {
            case 0xDB: d.define_quant(end); break;
            case 0xC4: d.define_huffman(end); break;
            case 0xDD:
                d.restart = d.u16();
                break;
            case 0xDA:
                d.scan(end);
                break;
            case 0xE0:   // APP0: JFIF
                if (len >= 7 && std::memcmp(p + start + 2, "JFIF\0", 5) == 0) d.jfif = true;
                break;
            case 0xEE:   // Adobe APP14: transform flag
                if (len >= 14 && std::memcmp(p + start + 2, "Adobe", 5) == 0) d.adobe_transform = p[start + 2 + 11];
                break;
            default:
                if (m >= 0xC0 && m <= 0xCF && m != 0xC4 && m != 0xC8 && m != 0xCC) d.start_of_frame(m, end);
                break;
        }
        if (m != 0xDA) d.pos = end;
        if (m == 0xDA && d.pos >= n) break;   // data ran to the end without EOI
    }
    if (!d.frame) bad("no frame");
    if (d.progressive) d.finish_progressive();

    const int W = d.width, H = d.height;
    std::vector<uint8_t> out((size_t)W * H * 4);
    const int nc = (int)d.comp.size();
    struct Resample {
        const uint8_t* (*fn)(uint8_t*, const uint8_t*, const uint8_t*, int, int);
        int hs, vs, ystep, ypos, w_lores;
        const uint8_t *line0, *line1;
        std::vector<uint8_t> buf;
    };
    std::vector<Resample> rs((size_t)nc);
    for (int k = 0; k < nc; ++k) {
        Resample& r = rs[(size_t)k];
        const Component& c = d.comp[(size_t)k];
        r.hs = d.hmax / c.h;
        r.vs = d.vmax / c.v;
        r.ystep = r.vs >> 1;
        r.w_lores = (W + r.hs - 1) / r.hs;
        r.ypos = 0;
        r.line0 = r.line1 = c.data.data();
        r.buf.assign((size_t)W + 3, 0);
        if (r.hs == 1 && r.vs == 1) r.fn = row_1;
        else if (r.hs == 1 && r.vs == 2) r.fn = row_v2;
        else if (r.hs == 2 && r.vs == 1) r.fn = row_h2;
        else if (r.hs == 2 && r.vs == 2) r.fn = row_hv2;
        else r.fn = row_generic;
    }
    const bool is_rgb = nc == 3 && ((d.comp[0].id == 'R' && d.comp[1].id == 'G' && d.comp[2].id == 'B') ||
                                    (d.adobe_transform == 0 && !d.jfif));
    const int cr_r = fix(1.40200f), cr_g = -fix(0.71414f), cb_g = -fix(0.34414f), cb_b = fix(1.77200f);
    const uint8_t* row[4] = {nullptr, nullptr, nullptr, nullptr};
    // stb's fixed-point YCbCr -> RGB of sample i into o[0..2], alpha 255
    const auto ycc = [&](int i, uint8_t* o) {
        const int yf = (row[0][i] << 20) + (1 << 19);
        const int cr = row[2][i] - 128, cb = row[1][i] - 128;
        int rr = yf + cr * cr_r;
        int gg = yf + (cr * cr_g) + (int)((unsigned)(cb * cb_g) & 0xffff0000u);
        int bb = yf + cb * cb_b;
        rr >>= 20;
        gg >>= 20;
        bb >>= 20;
        o[0] = (uint8_t)((unsigned)rr > 255 ? (rr < 0 ? 0 : 255) : rr);
        o[1] = (uint8_t)((unsigned)gg > 255 ? (gg < 0 ? 0 : 255) : gg);
        o[2] = (uint8_t)((unsigned)bb > 255 ? (bb < 0 ? 0 : 255) : bb);
        o[3] = 255;
    };
    // x * y / 255 rounded as stb's blinn_8x8: t = x * y + 128, (t + (t >> 8)) >> 8
    const auto blinn = [](int x, int y) {
        const unsigned t = (unsigned)(x * y + 128);
        return (uint8_t)((t + (t >> 8)) >> 8);
    };
    for (int j = 0; j < H; ++j) {
        for (int k = 0; k < nc; ++k) {
            Resample& r = rs[(size_t)k];
            const bool y_bot = r.ystep >= (r.vs >> 1);
            row[k] = r.fn(r.buf.data(), y_bot ? r.line1 : r.line0, y_bot ? r.line0 : r.line1, r.w_lores, r.hs);
            if (++r.ystep >= r.vs) {
                r.ystep = 0;
                r.line0 = r.line1;
                if (++r.ypos < d.comp[(size_t)k].y) r.line1 += d.comp[(size_t)k].w2;
            }
        }
        uint8_t* o = out.data() + (size_t)(flip ? H - 1 - j : j) * W * 4;
        if (nc == 1) {
            for (int i = 0; i < W; ++i, o += 4) {
                o[0] = o[1] = o[2] = row[0][i];
                o[3] = 255;
            }
        } else if (nc == 4) {
            // Adobe transform 0: CMYK as stored (inverted by convention) times K;
            // 2: YCCK, YCbCr -> RGB, then (255 - rgb) times K; otherwise the
            // fourth channel is ignored (stb's rules)
            for (int i = 0; i < W; ++i, o += 4) {
                const int m = row[3][i];
                if (d.adobe_transform == 0) {
                    o[0] = blinn(row[0][i], m);
                    o[1] = blinn(row[1][i], m);
                    o[2] = blinn(row[2][i], m);
                    o[3] = 255;
                } else {
                    ycc(i, o);
                    if (d.adobe_transform == 2)
                        for (int c = 0; c < 3; ++c) o[c] = blinn(255 - o[c], m);
                }
            }
        } else if (is_rgb) {
            for (int i = 0; i < W; ++i, o += 4) {
                o[0] = row[0][i];
                o[1] = row[1][i];
                o[2] = row[2][i];
                o[3] = 255;
            }
        } else {
            for (int i = 0; i < W; ++i, o += 4) ycc(i, o);
        }
    }
    w_out = (uint32_t)W;
    h_out = (uint32_t)H;
    return out;
}

}  // namespace rt
