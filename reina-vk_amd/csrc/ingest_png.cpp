// ingest_png.cpp — PNG → RGBA8 (and the entry points that also route JPEG
// data to ingest_jpeg.cpp), host only.
//
// Replaces the stb_image calls of src/graphics/Image.cpp:10-41 for the
// textures Reina loads (stbi_load / stbi_load_from_memory with 4 channels
// requested; stb is fetched at `master`, CMakeLists.txt:21-26). The decode
// follows the PNG specification; the expansion to 4 channels follows stb's
// rules:
//   grey            (g, g, g, 255); bit depths 1/2/4 scaled by 255/(2^d - 1)
//   grey + alpha    (g, g, g, a)
//   RGB / RGBA      (r, g, b, 255) / (r, g, b, a)
//   palette         the PLTE entry, alpha from tRNS (255 without)
//   16-bit samples  the high byte
//   tRNS key        (grey, RGB) alpha 0 where every sample equals the key
// Loading from a path flips the rows (stbi_set_flip_vertically_on_load(true),
// Image.cpp:14); loading from memory does not (Image.cpp:28). Chunk CRCs are
// not checked (nor are they by stb). zlib (the system library) inflates.
#include <zlib.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/reina_rt.h"

namespace rt {
int fail(int code, const std::string& msg);
struct JpegError {
    int code;
    std::string msg;
};
std::vector<uint8_t> decode_jpeg(const uint8_t* p, uint64_t n, bool flip, uint32_t& w, uint32_t& h);
}
using rt::fail;

namespace {

struct Error {
    int code;
    std::string msg;
};

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

constexpr uint64_t kMaxPixels = 1ull << 28;

struct Header {
    uint32_t w = 0, h = 0;
    unsigned depth = 0, color = 0, interlace = 0;
    unsigned channels() const { return color == 0 ? 1 : color == 2 ? 3 : color == 3 ? 1 : color == 4 ? 2 : 4; }
    // bytes of one filtered row of `width` pixels (without the filter byte)
    uint64_t row_bytes(uint64_t width) const { return (width * channels() * depth + 7) / 8; }
};

std::vector<uint8_t> inflate_all(const std::vector<uint8_t>& z, uint64_t need) {
    std::vector<uint8_t> out(need);
    z_stream s;
    std::memset(&s, 0, sizeof s);
    if (inflateInit(&s) != Z_OK) throw Error{REINA_RT_ERR_OUT_OF_MEMORY, "PNG: zlib init failed"};
    s.next_in = const_cast<Bytef*>(z.data());
    s.avail_in = 0;
    uint64_t in_off = 0, out_off = 0;
    int rc = Z_OK;
    while (rc != Z_STREAM_END && out_off < need) {
        if (s.avail_in == 0) {
            const uint64_t chunk = std::min<uint64_t>(z.size() - in_off, 1u << 30);
            if (chunk == 0) break;
            s.next_in = const_cast<Bytef*>(z.data() + in_off);
            s.avail_in = (uInt)chunk;
            in_off += chunk;
        }
        const uint64_t room = std::min<uint64_t>(need - out_off, 1u << 30);
        s.next_out = out.data() + out_off;
        s.avail_out = (uInt)room;
        rc = inflate(&s, Z_NO_FLUSH);
        out_off += room - s.avail_out;
        if (rc != Z_OK && rc != Z_STREAM_END && rc != Z_BUF_ERROR) {
            inflateEnd(&s);
            throw Error{REINA_RT_ERR_INVALID_ARG, "PNG: corrupt image data (zlib)"};
        }
        if (rc == Z_BUF_ERROR && s.avail_in == 0 && in_off >= z.size()) break;
    }
    inflateEnd(&s);
    if (out_off < need) throw Error{REINA_RT_ERR_INVALID_ARG, "PNG: not enough pixels (truncated image data)"};
    return out;
}

int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

// undoes the per-row filters of one (sub)image; returns its rows without
// their filter bytes
std::vector<uint8_t> unfilter(const uint8_t* data, uint64_t rows, uint64_t rb, unsigned bpp) {
    std::vector<uint8_t> out(rows * rb);
    const std::vector<uint8_t> zero(rb, 0);
    for (uint64_t y = 0; y < rows; ++y) {
        const uint8_t* src = data + y * (rb + 1) + 1;
        const unsigned f = src[-1];
        uint8_t* cur = out.data() + y * rb;
        const uint8_t* prev = y ? out.data() + (y - 1) * rb : zero.data();
        const uint64_t head = std::min<uint64_t>(bpp, rb);
        switch (f) {
            case 0:
                std::memcpy(cur, src, rb);
                break;
            case 1:
                std::memcpy(cur, src, head);
                for (uint64_t i = bpp; i < rb; ++i) cur[i] = (uint8_t)(src[i] + cur[i - bpp]);
                break;
            case 2:
                for (uint64_t i = 0; i < rb; ++i) cur[i] = (uint8_t)(src[i] + prev[i]);
                break;
            case 3:
                for (uint64_t i = 0; i < head; ++i) cur[i] = (uint8_t)(src[i] + (prev[i] >> 1));
                for (uint64_t i = bpp; i < rb; ++i) cur[i] = (uint8_t)(src[i] + ((cur[i - bpp] + prev[i]) >> 1));
                break;
            case 4:
                for (uint64_t i = 0; i < head; ++i) cur[i] = (uint8_t)(src[i] + prev[i]);   // paeth(0, b, 0) = b
                for (uint64_t i = bpp; i < rb; ++i)
                    cur[i] = (uint8_t)(src[i] + paeth(cur[i - bpp], prev[i], prev[i - bpp]));
                break;
            default:
                throw Error{REINA_RT_ERR_INVALID_ARG, "PNG: invalid filter type"};
        }
    }
    return out;
}

// sample k (0-based, channel-major within the pixel) of a row
unsigned sample(const uint8_t* row, uint64_t k, unsigned depth) {
    switch (depth) {
        case 16: return (unsigned)row[2 * k] << 8 | row[2 * k + 1];
        case 8: return row[k];
        default: {
            const uint64_t bit = k * depth;
            const unsigned shift = 8 - depth - (unsigned)(bit & 7);
            return (row[bit >> 3] >> shift) & ((1u << depth) - 1);
        }
    }
}

ReinaImage* decode(const uint8_t* p, uint64_t n, bool flip) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (n >= 3 && p[0] == 0xFF && p[1] == 0xD8 && p[2] == 0xFF) {
        uint32_t w = 0, h = 0;
        std::vector<uint8_t> rgba;
        try {
            rgba = rt::decode_jpeg(p, n, flip, w, h);
        } catch (const rt::JpegError& e) {
            throw Error{e.code, e.msg};
        }
        ReinaImage* img = static_cast<ReinaImage*>(std::calloc(1, sizeof(ReinaImage)));
        if (!img) throw std::bad_alloc();
        img->rgba8 = static_cast<uint8_t*>(std::malloc(rgba.size()));
        if (!img->rgba8) {
            std::free(img);
            throw std::bad_alloc();
        }
        std::memcpy(img->rgba8, rgba.data(), rgba.size());
        img->width = w;
        img->height = h;
        return img;
    }
    if (n < 8 || std::memcmp(p, sig, 8) != 0) throw Error{REINA_RT_ERR_UNSUPPORTED, "not a PNG or JPEG image"};
    Header hd;
    uint8_t pal[256 * 4];
    unsigned pal_len = 0, trns_len = 0;
    unsigned key[3] = {0, 0, 0};
    bool have_key = false, first = true, ended = false;
    std::vector<uint8_t> z;
    uint64_t off = 8;
    while (!ended) {
        if (off + 8 > n) throw Error{REINA_RT_ERR_INVALID_ARG, "PNG: truncated file"};
        const uint32_t len = be32(p + off);
        const uint8_t* type = p + off + 4;
        const uint8_t* d = p + off + 8;
        if ((uint64_t)len + 12 > n - off) throw Error{REINA_RT_ERR_INVALID_ARG, "PNG: truncated chunk"};
        const std::string t(reinterpret_cast<const char*>(type), 4);
        if (first && t != "IHDR") throw Error{REINA_RT_ERR_INVALID_ARG, "PNG: first chunk is not IHDR"};
        if (t == "IHDR") {
            if (!first || len != 13) throw Error{REINA_RT_ERR_INVALID_ARG, "PNG: bad IHDR"};
            hd.w = be32(d);
            hd.h = be32(d + 4);
            hd.depth = d[8];
            hd.color = d[9];
            hd.interlace = d[12];
            if (hd.w == 0 || hd.h == 0 || hd.w > (1u << 24) || hd.h > (1u << 24) ||
                (uint64_t)hd.w * hd.h > kMaxPixels)
                throw Error{REINA_RT_ERR_INVALID_ARG, "PNG: image size out of range"};
            const unsigned dp = hd.depth, ct = hd.color;
            const bool ok_depth = dp == 1 || dp == 2 || dp == 4 || dp == 8 || dp == 16;
            const bool ok = ok_depth && (ct == 0 || ((ct == 2 || ct == 4 || ct == 6) && dp >= 8) ||
                                         (ct == 3 && dp <= 8));
            if (!ok) throw Error{REINA_RT_ERR_INVALID_ARG, "PNG: bad colour type / bit depth"};
            if (d[10] != 0 || d[11] != 0 || hd.interlace > 1)
                throw Error{REINA_RT_ERR_INVALID_ARG, "PNG: bad compression, filter or interlace method"};
            first = false;
        } else if (t == "PLTE") {
            if (len > 768 || len % 3 != 0) throw Error{REINA_RT_ERR_INVALID_ARG, "PNG: bad PLTE"};
            pal_len = len / 3;
            for (unsigned i = 0; i < pal_len; ++i) {
                pal[4 * i + 0] = d[3 * i + 0];
                pal[4 * i + 1] = d[3 * i + 1];
                pal[4 * i + 2] = d[3 * i + 2];
                pal[4 * i + 3] = 255;
            }
        } else if (t == "tRNS") {
            if (!z.empty()) throw Error{REINA_RT_ERR_INVALID_ARG, "PNG: tRNS after IDAT"};
            if (hd.color == 3) {
                if (pal_len == 0) throw Error{REINA_RT_ERR_INVALID_ARG, "PNG: tRNS before PLTE"};
                if (len > pal_len) throw Error{REINA_RT_ERR_INVALID_ARG, "PNG: bad tRNS length"};
                trns_len = len;
                for (unsigned i = 0; i < len; ++i) pal[4 * i + 3] = d[i];
            } else if (hd.color == 0 || hd.color == 2) {
                const unsigned nk = hd.color == 0 ? 1 : 3;
                if (len != 2 * nk) throw Error{REINA_RT_ERR_INVALID_ARG, "PNG: bad tRNS length"};
                for (unsigned k = 0; k < nk; ++k) key[k] = (unsigned)d[2 * k] << 8 | d[2 * k + 1];
                have_key = true;
            } else {
                throw Error{REINA_RT_ERR_INVALID_ARG, "PNG: tRNS with an alpha channel"};
            }
        } else if (t == "IDAT") {
            z.insert(z.end(), d, d + len);
        } else if (t == "IEND") {
            ended = true;
        } else if (!(type[0] & 0x20)) {
            throw Error{REINA_RT_ERR_UNSUPPORTED, "PNG: unknown critical chunk " + t};
        }
        off += 12ull + len;
    }
    if (first) throw Error{REINA_RT_ERR_INVALID_ARG, "PNG: no IHDR"};
    if (hd.color == 3 && pal_len == 0) throw Error{REINA_RT_ERR_INVALID_ARG, "PNG: palette image without PLTE"};
    if (z.empty()) throw Error{REINA_RT_ERR_INVALID_ARG, "PNG: no image data"};
    (void)trns_len;

    // the passes: the whole image, or Adam7's seven
    struct Pass { unsigned x0, y0, dx, dy; };
    static const Pass adam7[7] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                  {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    static const Pass whole[1] = {{0, 0, 1, 1}};
    const Pass* passes = hd.interlace ? adam7 : whole;
    const int npass = hd.interlace ? 7 : 1;
    uint64_t need = 0;
    for (int q = 0; q < npass; ++q) {
        const uint64_t pw = (hd.w - passes[q].x0 + passes[q].dx - 1) / passes[q].dx;
        const uint64_t ph = (hd.h - passes[q].y0 + passes[q].dy - 1) / passes[q].dy;
        if (hd.w > passes[q].x0 && hd.h > passes[q].y0 && pw && ph) need += ph * (hd.row_bytes(pw) + 1);
    }
    const std::vector<uint8_t> raw = inflate_all(z, need);

    const unsigned ch = hd.channels();
    const unsigned bpp = std::max(1u, ch * hd.depth / 8);
    const unsigned scale = hd.depth == 1 ? 255 : hd.depth == 2 ? 85 : hd.depth == 4 ? 17 : 1;
    std::vector<uint8_t> rgba((uint64_t)hd.w * hd.h * 4);
    uint64_t at = 0;
    for (int q = 0; q < npass; ++q) {
        const Pass& ps = passes[q];
        if (hd.w <= ps.x0 || hd.h <= ps.y0) continue;
        const uint64_t pw = (hd.w - ps.x0 + ps.dx - 1) / ps.dx;
        const uint64_t ph = (hd.h - ps.y0 + ps.dy - 1) / ps.dy;
        const uint64_t rb = hd.row_bytes(pw);
        const std::vector<uint8_t> rows = unfilter(raw.data() + at, ph, rb, bpp);
        at += ph * (rb + 1);
        for (uint64_t y = 0; y < ph; ++y) {
            const uint8_t* row = rows.data() + y * rb;
            const uint64_t oy = ps.y0 + y * ps.dy;
            if (hd.depth == 8 && !have_key && hd.color != 3 && !hd.interlace) {
                // 8-bit samples without a colour key: direct expansion
                uint8_t* o = rgba.data() + oy * hd.w * 4;
                if (hd.color == 6) {
                    std::memcpy(o, row, (size_t)pw * 4);
                } else {
                    for (uint64_t x = 0; x < pw; ++x, o += 4) {
                        if (hd.color == 2) { o[0] = row[3 * x]; o[1] = row[3 * x + 1]; o[2] = row[3 * x + 2]; o[3] = 255; }
                        else if (hd.color == 0) { o[0] = o[1] = o[2] = row[x]; o[3] = 255; }
                        else { o[0] = o[1] = o[2] = row[2 * x]; o[3] = row[2 * x + 1]; }
                    }
                }
                continue;
            }
            for (uint64_t x = 0; x < pw; ++x) {
                const uint64_t ox = ps.x0 + x * ps.dx;
                uint8_t* o = rgba.data() + (oy * hd.w + ox) * 4;
                unsigned s[4] = {0, 0, 0, 0};
                for (unsigned c = 0; c < ch; ++c) s[c] = sample(row, x * ch + c, hd.depth);
                const auto to8 = [&](unsigned v) -> uint8_t {
                    return hd.depth == 16 ? (uint8_t)(v >> 8) : (uint8_t)(v * scale);
                };
                switch (hd.color) {
                    case 0: {
                        const uint8_t g = to8(s[0]);
                        bool clear = false;
                        if (have_key)
                            clear = hd.depth == 16 ? s[0] == key[0] : g == (uint8_t)((key[0] & 255) * scale);
                        o[0] = o[1] = o[2] = g;
                        o[3] = clear ? 0 : 255;
                        break;
                    }
                    case 2: {
                        bool clear = have_key;
                        for (unsigned c = 0; c < 3; ++c) {
                            o[c] = to8(s[c]);
                            if (have_key) clear = clear && (hd.depth == 16 ? s[c] == key[c] : o[c] == (key[c] & 255));
                        }
                        o[3] = clear ? 0 : 255;
                        break;
                    }
                    case 3: {
                        if (s[0] >= pal_len) throw Error{REINA_RT_ERR_INVALID_ARG, "PNG: palette index out of range"};
                        std::memcpy(o, pal + 4 * s[0], 4);
                        break;
                    }
                    case 4:
                        o[0] = o[1] = o[2] = to8(s[0]);
                        o[3] = to8(s[1]);
                        break;
                    default:
                        for (unsigned c = 0; c < 4; ++c) o[c] = to8(s[c]);
                        break;
                }
            }
        }
    }
    if (flip) {
        const uint64_t stride = (uint64_t)hd.w * 4;
        std::vector<uint8_t> tmp(stride);
        for (uint64_t y = 0; y < hd.h / 2; ++y) {
            uint8_t* a = rgba.data() + y * stride;
            uint8_t* b = rgba.data() + (hd.h - 1 - y) * stride;
            std::memcpy(tmp.data(), a, stride);
            std::memcpy(a, b, stride);
            std::memcpy(b, tmp.data(), stride);
        }
    }
    ReinaImage* img = static_cast<ReinaImage*>(std::calloc(1, sizeof(ReinaImage)));
    if (!img) throw std::bad_alloc();
    img->rgba8 = static_cast<uint8_t*>(std::malloc(rgba.size()));
    if (!img->rgba8) {
        std::free(img);
        throw std::bad_alloc();
    }
    std::memcpy(img->rgba8, rgba.data(), rgba.size());
    img->width = hd.w;
    img->height = hd.h;
    return img;
}

int load(const uint8_t* p, uint64_t n, bool flip, ReinaImage** out) {
    try {
        *out = decode(p, n, flip);
    } catch (const Error& e) {
        return fail(e.code, e.msg);
    } catch (const std::bad_alloc&) {
        return fail(REINA_RT_ERR_OUT_OF_MEMORY, "host allocation failed while decoding a PNG image");
    }
    return REINA_RT_OK;
}

}  // namespace

extern "C" {

int reina_rt_image_load_memory(const uint8_t* bytes, uint64_t len, ReinaImage** out) {
    if (!out || (!bytes && len)) return fail(REINA_RT_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    return load(bytes, len, false, out);
}

int reina_rt_image_load(const char* path, ReinaImage** out) {
    if (!path || !out) return fail(REINA_RT_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    std::ifstream f(path, std::ios::binary);
    if (!f) return fail(REINA_RT_ERR_INVALID_ARG, std::string("Could not load image at path: ") + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    const std::string data = ss.str();
    const int rc = load(reinterpret_cast<const uint8_t*>(data.data()), data.size(), true, out);
    if (rc != REINA_RT_OK) return fail(rc, std::string("Could not load image at path: ") + path + ": " + reina_rt_last_error());
    return rc;
}

void reina_rt_image_free(ReinaImage* image) {
    if (!image) return;
    std::free(image->rgba8);
    std::free(image);
}

}  // extern "C"
