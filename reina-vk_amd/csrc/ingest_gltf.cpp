// ingest_gltf.cpp — glTF 2.0 (.gltf / .glb) → models, images and instances,
// host only.
//
// Replaces reina::scene::gltf::loadScene (src/scene/gltf/gltfloader.cpp:
// 65-456), which parses with fastgltf (CMakeLists.txt:35-40) and builds a
// Scene from the default scene's nodes:
//   * every mesh a node of the default scene references, every primitive of
//     it one model: POSITION, NORMAL (required), TANGENT (VEC4: bitangent =
//     cross(N, T) * w; VEC3: cross(N, T)), TEXCOORD_0 (zeros without), indices
//     (0..n-1 without);
//   * without TANGENT the reference calls MikkTSpace (gltfloader.cpp:206-219)
//     before TEXCOORD_0 is read, i.e. with every texture coordinate zero and
//     each consecutive vertex triple taken as a face. No face then has UV
//     area, no tangent group forms, and every written corner keeps
//     MikkTSpace's initial frame T = (1, 0, 0), B = (0, 1, 0); vertices past
//     the last whole triple keep zeros. This file writes that result directly;
//   * every image decoded in order (embedded, data: URI or file next to the
//     asset; PNG through ingest_png.cpp, unflipped as stbi_load_from_memory);
//   * materials: Disney (materialIdx 3) with interpolated normals, albedo =
//     baseColorFactor.rgb, metallic, roughness clamped to [0.1, 0.7],
//     emission = emissiveFactor * emissiveStrength when positive and without
//     an emissive texture, back faces culled unless doubleSided, transmission
//     -> specularTransmission (and no culling), base colour and normal
//     textures (gltfloader.cpp:366-428); ior stays 1.5 (KHR_materials_ior is
//     not among the reference's enabled extensions);
//   * one instance per node and primitive with the node's world matrix:
//     parent * matrix, or ((parent * T) * R) * S for TRS nodes.
// Models are numbered by ascending mesh index, then primitive order (the
// reference walks unordered containers, whose order its standard library
// chooses).
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/reina_rt.h"

namespace rt {
int fail(int code, const std::string& msg);
}
using rt::fail;

namespace {

struct Error {
    int code;
    std::string msg;
};
[[noreturn]] void bad(const std::string& m) { throw Error{REINA_RT_ERR_INVALID_ARG, "glTF: " + m}; }

// ---- JSON --------------------------------------------------------------------------
struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;

    const Json* get(const char* key) const {
        if (kind != Object) return nullptr;
        for (const auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    double number(const char* key, double dflt) const {
        const Json* v = get(key);
        if (!v) return dflt;
        if (v->kind != Number) bad(std::string("'") + key + "' is not a number");
        return v->num;
    }
    int64_t index(const char* key, int64_t dflt = -1) const {
        const Json* v = get(key);
        if (!v) return dflt;
        if (v->kind != Number || v->num < 0 || v->num != std::floor(v->num) || v->num > 4294967295.0)
            bad(std::string("'") + key + "' is not an index");
        return (int64_t)v->num;
    }
    const std::vector<Json>& array(const char* key) const {
        static const std::vector<Json> empty;
        const Json* v = get(key);
        if (!v) return empty;
        if (v->kind != Array) bad(std::string("'") + key + "' is not an array");
        return v->arr;
    }
};

struct JsonParser {
    const char* p;
    const char* e;
    int depth = 0;
    void ws() {
        while (p < e && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    [[noreturn]] void err(const char* m) { bad(std::string("JSON: ") + m); }
    void expect(char c) {
        ws();
        if (p >= e || *p != c) err("unexpected character");
        ++p;
    }
    static void utf8(std::string& out, uint32_t cp) {
        if (cp < 0x80) out += (char)cp;
        else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 0x3F)); }
        else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | (cp >> 18)); out += (char)(0x80 | ((cp >> 12) & 0x3F));
            out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F));
        }
    }
    uint32_t hex4() {
        if (e - p < 4) err("bad \\u escape");
        uint32_t v = 0;
        for (int i = 0; i < 4; ++i) {
            const char c = *p++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
            else err("bad \\u escape");
        }
        return v;
    }
    std::string string() {
        expect('"');
        std::string out;
        while (true) {
            if (p >= e) err("unterminated string");
            const char c = *p++;
            if (c == '"') break;
            if (c != '\\') { out += c; continue; }
            if (p >= e) err("unterminated string");
            const char x = *p++;
            switch (x) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    uint32_t cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00 && e - p >= 6 && p[0] == '\\' && p[1] == 'u') {
                        p += 2;
                        const uint32_t lo = hex4();
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    utf8(out, cp);
                    break;
                }
                default: err("bad escape");
            }
        }
        return out;
    }
    Json value() {
        if (++depth > 256) err("nesting too deep");
        ws();
        if (p >= e) err("unexpected end");
        Json v;
        if (*p == '{') {
            v.kind = Json::Object;
            ++p;
            ws();
            if (p < e && *p == '}') { ++p; --depth; return v; }
            while (true) {
                std::string k = string();
                expect(':');
                v.obj.emplace_back(std::move(k), value());
                ws();
                if (p < e && *p == ',') { ++p; continue; }
                expect('}');
                break;
            }
        } else if (*p == '[') {
            v.kind = Json::Array;
            ++p;
            ws();
            if (p < e && *p == ']') { ++p; --depth; return v; }
            while (true) {
                v.arr.push_back(value());
                ws();
                if (p < e && *p == ',') { ++p; continue; }
                expect(']');
                break;
            }
        } else if (*p == '"') {
            v.kind = Json::String;
            v.str = string();
        } else if (e - p >= 4 && std::memcmp(p, "true", 4) == 0) {
            v.kind = Json::Bool; v.b = true; p += 4;
        } else if (e - p >= 5 && std::memcmp(p, "false", 5) == 0) {
            v.kind = Json::Bool; v.b = false; p += 5;
        } else if (e - p >= 4 && std::memcmp(p, "null", 4) == 0) {
            p += 4;
        } else {
            const char* s = p;
            while (p < e && (std::strchr("+-0123456789.eE", *p) != nullptr)) ++p;
            if (s == p) err("unexpected character");
            const std::string num(s, p);
            char* end = nullptr;
            v.kind = Json::Number;
            v.num = std::strtod(num.c_str(), &end);
            if (end != num.c_str() + num.size()) err("bad number");
        }
        --depth;
        return v;
    }
};

// ---- buffers -------------------------------------------------------------------------
std::string read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) bad("cannot open " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

std::vector<uint8_t> base64(const std::string& s, size_t from) {
    static int8_t table[256];
    static bool init = false;
    if (!init) {
        std::memset(table, -1, sizeof table);
        const char* a = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
        for (int i = 0; i < 64; ++i) table[(uint8_t)a[i]] = (int8_t)i;
        init = true;
    }
    std::vector<uint8_t> out;
    out.reserve((s.size() - from) * 3 / 4);
    uint32_t acc = 0;
    int bits = 0;
    for (size_t i = from; i < s.size(); ++i) {
        const char c = s[i];
        if (c == '=') break;
        const int v = table[(uint8_t)c];
        if (v < 0) bad("bad base64 data");
        acc = (acc << 6) | (uint32_t)v;
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back((uint8_t)(acc >> bits));
        }
    }
    return out;
}

std::string percent_decode(const std::string& s) {
    std::string out;
    for (size_t i = 0; i < s.size(); ++i) {
        if (s[i] == '%' && i + 2 < s.size()) {
            out += (char)std::strtol(s.substr(i + 1, 2).c_str(), nullptr, 16);
            i += 2;
        } else {
            out += s[i];
        }
    }
    return out;
}

// the bytes a URI names: a data: URI, or a file relative to the asset
std::vector<uint8_t> load_uri(const std::string& uri, const std::string& dir) {
    if (uri.compare(0, 5, "data:") == 0) {
        const size_t comma = uri.find(',');
        if (comma == std::string::npos || uri.find(";base64") == std::string::npos || uri.find(";base64") > comma)
            bad("only base64 data: URIs are supported");
        return base64(uri, comma + 1);
    }
    if (uri.find("://") != std::string::npos) bad("non-local URI " + uri);
    if (dir.empty()) bad("external file " + uri + " referenced by an asset loaded from memory");
    const std::string data = read_file(dir + "/" + percent_decode(uri));
    return std::vector<uint8_t>(data.begin(), data.end());
}

struct Asset {
    Json doc;
    std::vector<std::vector<uint8_t>> buffers;
};

Asset parse_asset(const uint8_t* bytes, uint64_t len, const std::string& dir) {
    Asset a;
    const uint8_t* json = bytes;
    uint64_t json_len = len;
    const uint8_t* bin = nullptr;
    uint64_t bin_len = 0;
    if (len >= 12 && std::memcmp(bytes, "glTF", 4) == 0) {   // GLB container
        uint32_t hdr[3];
        std::memcpy(hdr, bytes, 12);
        if (hdr[1] != 2) bad("GLB version is not 2");
        if (hdr[2] > len) bad("truncated GLB");
        uint64_t off = 12;
        json = nullptr;
        while (off + 8 <= hdr[2]) {
            uint32_t ch[2];
            std::memcpy(ch, bytes + off, 8);
            if (off + 8 + ch[0] > hdr[2]) bad("truncated GLB chunk");
            if (ch[1] == 0x4E4F534Au && !json) {   // "JSON"
                json = bytes + off + 8;
                json_len = ch[0];
            } else if (ch[1] == 0x004E4942u && !bin) {   // "BIN\0"
                bin = bytes + off + 8;
                bin_len = ch[0];
            }
            off += 8ull + ch[0];
        }
        if (!json) bad("GLB without a JSON chunk");
    }
    JsonParser jp{reinterpret_cast<const char*>(json), reinterpret_cast<const char*>(json) + json_len};
    a.doc = jp.value();
    if (a.doc.kind != Json::Object) bad("the document is not an object");
    static const char* supported[] = {"KHR_mesh_quantization", "KHR_texture_transform", "KHR_materials_variants",
                                      "KHR_materials_transmission", "KHR_materials_clearcoat",
                                      "KHR_materials_emissive_strength"};
    for (const Json& x : a.doc.array("extensionsRequired")) {
        bool ok = false;
        for (const char* s : supported) ok = ok || x.str == s;
        if (!ok) throw Error{REINA_RT_ERR_UNSUPPORTED, "glTF: required extension " + x.str + " is not supported"};
    }
    for (const Json& b : a.doc.array("buffers")) {
        const Json* uri = b.get("uri");
        std::vector<uint8_t> data;
        if (uri) {
            data = load_uri(uri->str, dir);
        } else {
            if (!bin) bad("buffer without uri outside a GLB");
            data.assign(bin, bin + bin_len);
        }
        const uint64_t want = (uint64_t)b.number("byteLength", (double)data.size());
        if (data.size() < want) bad("buffer shorter than its byteLength");
        a.buffers.push_back(std::move(data));
    }
    return a;
}

// ---- accessors -------------------------------------------------------------------------
unsigned comp_size(int64_t ct) {
    switch (ct) {
        case 5120: case 5121: return 1;
        case 5122: case 5123: return 2;
        case 5125: case 5126: return 4;
        default: bad("bad componentType");
    }
}

unsigned type_count(const std::string& t) {
    if (t == "SCALAR") return 1;
    if (t == "VEC2") return 2;
    if (t == "VEC3") return 3;
    if (t == "VEC4") return 4;
    if (t == "MAT2") return 4;
    if (t == "MAT3") return 9;
    if (t == "MAT4") return 16;
    bad("bad accessor type " + t);
}

// one component as float: normalized integers map to max(x / max, -1)
float to_float(const uint8_t* p, int64_t ct, bool norm) {
    switch (ct) {
        case 5126: { float f; std::memcpy(&f, p, 4); return f; }
        case 5120: { int8_t v; std::memcpy(&v, p, 1); return norm ? std::max((float)v / 127.0f, -1.0f) : (float)v; }
        case 5121: { uint8_t v = *p; return norm ? std::max((float)v / 255.0f, -1.0f) : (float)v; }
        case 5122: { int16_t v; std::memcpy(&v, p, 2); return norm ? std::max((float)v / 32767.0f, -1.0f) : (float)v; }
        case 5123: { uint16_t v; std::memcpy(&v, p, 2); return norm ? std::max((float)v / 65535.0f, -1.0f) : (float)v; }
        case 5125: { uint32_t v; std::memcpy(&v, p, 4); return norm ? std::max((float)v / 4294967295.0f, -1.0f) : (float)v; }
        default: bad("bad componentType");
    }
}

uint32_t to_u32(const uint8_t* p, int64_t ct) {
    switch (ct) {
        case 5121: return *p;
        case 5123: { uint16_t v; std::memcpy(&v, p, 2); return v; }
        case 5125: { uint32_t v; std::memcpy(&v, p, 4); return v; }
        default: bad("index accessor must be unsigned 8, 16 or 32-bit");
    }
}

// (pointer to element 0, stride) of a bufferView-backed array, bounds-checked
const uint8_t* view_data(const Asset& a, int64_t view, uint64_t offset, uint64_t count, uint64_t elem, uint64_t& stride) {
    const auto& views = a.doc.array("bufferViews");
    if (view < 0 || (size_t)view >= views.size()) bad("bad bufferView index");
    const Json& v = views[(size_t)view];
    const int64_t buf = v.index("buffer");
    if (buf < 0 || (size_t)buf >= a.buffers.size()) bad("bad buffer index");
    const uint64_t voff = (uint64_t)v.number("byteOffset", 0.0), vlen = (uint64_t)v.number("byteLength", 0.0);
    stride = (uint64_t)v.number("byteStride", 0.0);
    if (stride == 0) stride = elem;
    if (voff + vlen > a.buffers[(size_t)buf].size()) bad("bufferView outside its buffer");
    if (count && offset + stride * (count - 1) + elem > vlen) bad("accessor outside its bufferView");
    return a.buffers[(size_t)buf].data() + voff + offset;
}

struct Accessor {
    uint64_t count = 0;
    unsigned ncomp = 0;
    std::vector<float> f;        // count * ncomp, when read as floats
    std::vector<uint32_t> u;     // count, when read as indices
};

Accessor read_accessor(const Asset& a, int64_t idx, bool as_index) {
    const auto& accs = a.doc.array("accessors");
    if (idx < 0 || (size_t)idx >= accs.size()) bad("bad accessor index");
    const Json& j = accs[(size_t)idx];
    Accessor out;
    const int64_t ct = j.index("componentType");
    const Json* tp = j.get("type");
    if (!tp || tp->kind != Json::String) bad("accessor without type");
    out.ncomp = type_count(tp->str);
    out.count = (uint64_t)j.index("count", 0);
    const bool norm = j.get("normalized") && j.get("normalized")->b;
    const unsigned cs = comp_size(ct);
    const uint64_t elem = (uint64_t)cs * out.ncomp;
    if (out.count > (1ull << 31)) bad("accessor too large");
    if (as_index) {
        if (out.ncomp != 1) bad("index accessor is not SCALAR");
        out.u.assign(out.count, 0);
    } else {
        out.f.assign(out.count * out.ncomp, 0.0f);
    }
    const int64_t view = j.index("bufferView");
    const auto read_elem = [&](const uint8_t* p, uint64_t i) {
        if (as_index) out.u[i] = to_u32(p, ct);
        else for (unsigned c = 0; c < out.ncomp; ++c) out.f[i * out.ncomp + c] = to_float(p + c * cs, ct, norm);
    };
    if (view >= 0) {
        uint64_t stride = 0;
        const uint8_t* base = view_data(a, view, (uint64_t)j.number("byteOffset", 0.0), out.count, elem, stride);
        for (uint64_t i = 0; i < out.count; ++i) read_elem(base + i * stride, i);
    }
    if (const Json* sp = j.get("sparse")) {
        const uint64_t n = (uint64_t)sp->index("count", 0);
        const Json* si = sp->get("indices");
        const Json* sv = sp->get("values");
        if (!si || !sv) bad("sparse accessor without indices or values");
        const int64_t ict = si->index("componentType");
        uint64_t istride = 0, vstride = 0;
        const uint8_t* ib = view_data(a, si->index("bufferView"), (uint64_t)si->number("byteOffset", 0.0), n,
                                      comp_size(ict), istride);
        const uint8_t* vb = view_data(a, sv->index("bufferView"), (uint64_t)sv->number("byteOffset", 0.0), n, elem,
                                      vstride);
        for (uint64_t k = 0; k < n; ++k) {
            const uint32_t i = to_u32(ib + k * comp_size(ict), ict);
            if (i >= out.count) bad("sparse index out of range");
            read_elem(vb + k * elem, i);
        }
    }
    return out;
}

// ---- transforms (column-major 4x4, float) ---------------------------------------------
struct M4 {
    float m[16];   // m[c * 4 + r]
};

M4 identity() {
    M4 r{};
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f;
    return r;
}

M4 mul(const M4& a, const M4& b) {
    M4 r;
    for (int c = 0; c < 4; ++c)
        for (int row = 0; row < 4; ++row)
            r.m[c * 4 + row] = ((a.m[0 * 4 + row] * b.m[c * 4 + 0] + a.m[1 * 4 + row] * b.m[c * 4 + 1]) +
                                a.m[2 * 4 + row] * b.m[c * 4 + 2]) + a.m[3 * 4 + row] * b.m[c * 4 + 3];
    return r;
}

float f32(double d) { return (float)d; }

// parent * local for one node: a matrix, or ((parent * T) * R) * S
M4 node_matrix(const Json& node, const M4& parent) {
    const auto& mat = node.array("matrix");
    if (!mat.empty()) {
        if (mat.size() != 16) bad("node matrix needs 16 numbers");
        M4 l;
        for (int k = 0; k < 16; ++k) l.m[k] = f32(mat[(size_t)k].num);
        return mul(parent, l);
    }
    float t[3] = {0, 0, 0}, q[4] = {0, 0, 0, 1}, s[3] = {1, 1, 1};
    const auto vec = [&](const char* key, float* out, size_t n) {
        const auto& v = node.array(key);
        if (v.empty()) return;
        if (v.size() != n) bad(std::string("node ") + key + " has the wrong length");
        for (size_t k = 0; k < n; ++k) out[k] = f32(v[k].num);
    };
    vec("translation", t, 3);
    vec("rotation", q, 4);
    vec("scale", s, 3);
    M4 r = parent;   // translate: column 3 = c0 tx + c1 ty + c2 tz + c3
    for (int row = 0; row < 4; ++row)
        r.m[12 + row] = ((parent.m[row] * t[0] + parent.m[4 + row] * t[1]) + parent.m[8 + row] * t[2]) + parent.m[12 + row];
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z, wx = w * x, wy = w * y,
                wz = w * z;
    M4 R = identity();
    R.m[0] = 1.0f - 2.0f * (yy + zz); R.m[1] = 2.0f * (xy + wz);        R.m[2] = 2.0f * (xz - wy);
    R.m[4] = 2.0f * (xy - wz);        R.m[5] = 1.0f - 2.0f * (xx + zz); R.m[6] = 2.0f * (yz + wx);
    R.m[8] = 2.0f * (xz + wy);        R.m[9] = 2.0f * (yz - wx);        R.m[10] = 1.0f - 2.0f * (xx + yy);
    r = mul(r, R);
    for (int row = 0; row < 4; ++row) {
        r.m[row] *= s[0];
        r.m[4 + row] *= s[1];
        r.m[8 + row] *= s[2];
    }
    return r;
}

// ---- the scene ------------------------------------------------------------------------
struct Prim {
    std::vector<float> verts, tbns, uvs;
    std::vector<uint32_t> idx;
    int64_t material = -1;
};

Prim load_primitive(const Asset& a, const Json& p) {
    Prim out;
    out.material = p.index("material");
    const Json* attrs = p.get("attributes");
    if (!attrs || attrs->kind != Json::Object) bad("primitive without attributes");
    const int64_t ip = attrs->index("POSITION"), in = attrs->index("NORMAL"), it = attrs->index("TANGENT"),
                  iuv = attrs->index("TEXCOORD_0");
    if (ip < 0) bad("primitive without POSITION");
    if (in < 0) bad("Meshes without vertex normals are not supported");
    const Accessor P = read_accessor(a, ip, false);
    const Accessor N = read_accessor(a, in, false);
    if (P.ncomp != 3 || N.ncomp != 3) bad("POSITION and NORMAL must be VEC3");
    const uint64_t n = P.count;
    if (N.count != n) bad("NORMAL count differs from POSITION count");
    std::vector<float> T(3 * n, 0.0f), B(3 * n, 0.0f);
    if (it >= 0) {
        const Accessor Ta = read_accessor(a, it, false);
        if (Ta.count != n) bad("TANGENT count differs from POSITION count");
        if (Ta.ncomp != 3 && Ta.ncomp != 4) bad("TANGENT must be VEC3 or VEC4");
        for (uint64_t i = 0; i < n; ++i) {
            const float* t = &Ta.f[i * Ta.ncomp];
            const float* nn = &N.f[3 * i];
            T[3 * i] = t[0]; T[3 * i + 1] = t[1]; T[3 * i + 2] = t[2];
            float b[3] = {nn[1] * t[2] - nn[2] * t[1], nn[2] * t[0] - nn[0] * t[2], nn[0] * t[1] - nn[1] * t[0]};
            if (Ta.ncomp == 4) for (float& c : b) c = c * t[3];
            B[3 * i] = b[0]; B[3 * i + 1] = b[1]; B[3 * i + 2] = b[2];
        }
    } else {
        // MikkTSpace over zero texture coordinates: the initial frame for the
        // vertices of every whole consecutive triple
        const uint64_t m = n / 3 * 3;
        for (uint64_t i = 0; i < m; ++i) {
            T[3 * i] = 1.0f;
            B[3 * i + 1] = 1.0f;
        }
    }
    std::vector<float> uv(2 * n, 0.0f);
    if (iuv >= 0) {
        const Accessor U = read_accessor(a, iuv, false);
        if (U.ncomp != 2) bad("TEXCOORD_0 must be VEC2");
        if (U.count != n) bad("TEXCOORD_0 count differs from POSITION count");
        uv = U.f;
    }
    const int64_t ii = p.index("indices");
    if (ii >= 0) {
        out.idx = read_accessor(a, ii, true).u;
        for (uint32_t i : out.idx)
            if (i >= n) bad("vertex index out of range");
    } else {
        out.idx.resize(n);
        for (uint64_t i = 0; i < n; ++i) out.idx[i] = (uint32_t)i;
    }
    out.verts.resize(4 * n);
    out.tbns.resize(9 * n);
    for (uint64_t i = 0; i < n; ++i) {
        out.verts[4 * i] = P.f[3 * i];
        out.verts[4 * i + 1] = P.f[3 * i + 1];
        out.verts[4 * i + 2] = P.f[3 * i + 2];
        out.verts[4 * i + 3] = 1.0f;
        for (int c = 0; c < 3; ++c) {
            out.tbns[9 * i + c] = T[3 * i + c];
            out.tbns[9 * i + 3 + c] = B[3 * i + c];
            out.tbns[9 * i + 6 + c] = N.f[3 * i + c];
        }
    }
    out.uvs = std::move(uv);
    return out;
}

ReinaMaterial default_material() {
    ReinaMaterial m{};
    m.materialIdx = 3;
    m.textureID = m.normalMapID = m.bumpMapID = -1;
    for (int c = 0; c < 3; ++c) {
        m.albedo[c] = 1.0f;
        m.sheenTint[c] = 1.0f;
        m.specularTint[c] = 1.0f;
    }
    m.ior = 1.5f;
    m.interpNormals = 1;
    return m;
}

ReinaMaterial material(const Asset& a, int64_t mi, const std::vector<int32_t>& image_ids) {
    ReinaMaterial m = default_material();
    if (mi < 0) return m;
    const auto& mats = a.doc.array("materials");
    if ((size_t)mi >= mats.size()) bad("bad material index");
    const Json& g = mats[(size_t)mi];
    const Json* pbr = g.get("pbrMetallicRoughness");
    const Json* ext = g.get("extensions");
    float base[4] = {1, 1, 1, 1}, emissive[3] = {0, 0, 0};
    const auto& bcf = pbr ? pbr->array("baseColorFactor") : std::vector<Json>();
    for (size_t k = 0; k < 4 && k < bcf.size(); ++k) base[k] = f32(bcf[k].num);
    const auto& ef = g.array("emissiveFactor");
    for (size_t k = 0; k < 3 && k < ef.size(); ++k) emissive[k] = f32(ef[k].num);
    float strength = 1.0f;
    const Json* es = ext ? ext->get("KHR_materials_emissive_strength") : nullptr;
    if (es) strength = f32(es->number("emissiveStrength", 1.0));
    const bool emissive_tex = g.get("emissiveTexture") != nullptr;
    if ((emissive[0] > 0.0f || emissive[1] > 0.0f || emissive[2] > 0.0f) && !emissive_tex)
        for (int c = 0; c < 3; ++c) m.emission[c] = emissive[c] * strength;
    m.metallic = f32(pbr ? pbr->number("metallicFactor", 1.0) : 1.0);
    const float rough = f32(pbr ? pbr->number("roughnessFactor", 1.0) : 1.0);
    m.roughness = std::fmax(std::fmin(rough, 0.7f), 0.1f);
    for (int c = 0; c < 3; ++c) m.albedo[c] = base[c];
    const Json* ds = g.get("doubleSided");
    m.cullBackface = (ds && ds->b) ? 0u : 1u;
    const Json* tr = ext ? ext->get("KHR_materials_transmission") : nullptr;
    if (tr) {
        m.specularTransmission = f32(tr->number("transmissionFactor", 0.0));
        m.cullBackface = 0;
    }
    const auto texture_image = [&](const Json* info) -> int32_t {
        const int64_t ti = info->index("index");
        const auto& texs = a.doc.array("textures");
        if (ti < 0 || (size_t)ti >= texs.size()) bad("bad texture index");
        const int64_t src = texs[(size_t)ti].index("source");
        if (src < 0) bad("texture without an image source");
        if ((size_t)src >= image_ids.size()) return -1;
        return image_ids[(size_t)src];
    };
    if (pbr && pbr->get("baseColorTexture")) m.textureID = texture_image(pbr->get("baseColorTexture"));
    if (g.get("normalTexture")) m.normalMapID = texture_image(g.get("normalTexture"));
    return m;
}

struct Loaded {
    std::vector<Prim> prims;
    std::vector<ReinaImage*> images;
    std::vector<ReinaGltfInstance> instances;
    ~Loaded() {
        for (ReinaImage* im : images) reina_rt_image_free(im);
    }
};

void walk(const Asset& a, int64_t ni, const M4& parent, int depth, std::vector<std::pair<int64_t, M4>>& out) {
    const auto& nodes = a.doc.array("nodes");
    if (ni < 0 || (size_t)ni >= nodes.size()) bad("bad node index");
    if (depth > 1024) bad("node hierarchy too deep (a cycle?)");
    const Json& n = nodes[(size_t)ni];
    const M4 m = node_matrix(n, parent);
    out.emplace_back(ni, m);
    for (const Json& c : n.array("children")) walk(a, (int64_t)c.num, m, depth + 1, out);
}

std::unique_ptr<Loaded> load(const uint8_t* bytes, uint64_t len, const std::string& dir, uint32_t flags) {
    const Asset a = parse_asset(bytes, len, dir);
    auto L = std::make_unique<Loaded>();
    const auto& scenes = a.doc.array("scenes");
    if (scenes.empty()) bad("No scenes supplied in gLTF file");
    const int64_t si = a.doc.index("scene", 0);
    if ((size_t)si >= scenes.size()) bad("bad default scene index");
    std::vector<std::pair<int64_t, M4>> visits;
    for (const Json& r : scenes[(size_t)si].array("nodes")) walk(a, (int64_t)r.num, identity(), 0, visits);

    // images, in order
    std::vector<int32_t> image_ids;
    for (const Json& im : a.doc.array("images")) {
        std::vector<uint8_t> data;
        if (const Json* uri = im.get("uri")) {
            data = load_uri(uri->str, dir);
        } else {
            uint64_t stride = 0;
            const int64_t view = im.index("bufferView");
            const auto& views = a.doc.array("bufferViews");
            if (view < 0 || (size_t)view >= views.size()) bad("image without uri or bufferView");
            const uint64_t vlen = (uint64_t)views[(size_t)view].number("byteLength", 0.0);
            const uint8_t* p = view_data(a, view, 0, 1, vlen, stride);
            data.assign(p, p + vlen);
        }
        ReinaImage* img = nullptr;
        const int rc = reina_rt_image_load_memory(data.data(), data.size(), &img);
        if (rc != REINA_RT_OK) {
            if (rc == REINA_RT_ERR_UNSUPPORTED && (flags & REINA_RT_GLTF_SKIP_UNDECODABLE_IMAGES)) {
                image_ids.push_back(-1);
                continue;
            }
            throw Error{rc, "glTF: image " + std::to_string(image_ids.size()) + ": " + reina_rt_last_error()};
        }
        image_ids.push_back((int32_t)L->images.size());
        L->images.push_back(img);
    }

    // models: the meshes the default scene uses, ascending
    const auto& meshes = a.doc.array("meshes");
    std::map<int64_t, std::vector<uint32_t>> mesh_models;
    for (const auto& v : visits) {
        const int64_t mi = a.doc.array("nodes")[(size_t)v.first].index("mesh");
        if (mi < 0) continue;
        if ((size_t)mi >= meshes.size()) bad("bad mesh index");
        mesh_models[mi];
    }
    for (auto& mm : mesh_models)
        for (const Json& p : meshes[(size_t)mm.first].array("primitives")) {
            mm.second.push_back((uint32_t)L->prims.size());
            L->prims.push_back(load_primitive(a, p));
        }
    // instances: one per node and primitive, in traversal order
    for (const auto& v : visits) {
        const int64_t mi = a.doc.array("nodes")[(size_t)v.first].index("mesh");
        if (mi < 0) continue;
        for (uint32_t model : mesh_models[mi]) {
            ReinaGltfInstance inst{};
            inst.model = model;
            std::memcpy(inst.transform, v.second.m, sizeof inst.transform);
            inst.material = material(a, L->prims[model].material, image_ids);
            L->instances.push_back(inst);
        }
    }
    return L;
}

template <class T>
T* dup(const std::vector<T>& v) {
    if (v.empty()) return nullptr;
    T* p = static_cast<T*>(std::malloc(v.size() * sizeof(T)));
    if (!p) throw std::bad_alloc();
    std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

ReinaGltfScene* to_c(Loaded& L) {
    ReinaGltfScene* s = static_cast<ReinaGltfScene*>(std::calloc(1, sizeof(ReinaGltfScene)));
    if (!s) throw std::bad_alloc();
    try {
        s->num_models = (uint32_t)L.prims.size();
        s->models = static_cast<ReinaModelData*>(std::calloc(L.prims.size() + 1, sizeof(ReinaModelData)));
        if (!s->models) throw std::bad_alloc();
        for (size_t i = 0; i < L.prims.size(); ++i) {
            const Prim& p = L.prims[i];
            ReinaModelData& m = s->models[i];
            m.vertices = dup(p.verts);   m.num_vertices = p.verts.size() / 4;
            m.indices = dup(p.idx);      m.num_indices = p.idx.size();
            m.tbns = dup(p.tbns);        m.num_tbns = p.tbns.size() / 9;
            m.tbn_indices = dup(p.idx);  m.num_tbn_indices = p.idx.size();
            m.tex_coords = dup(p.uvs);   m.num_tex_coords = p.uvs.size() / 2;
            m.tex_indices = dup(p.idx);  m.num_tex_indices = p.idx.size();
        }
        s->num_images = (uint32_t)L.images.size();
        s->images = static_cast<ReinaImage*>(std::calloc(L.images.size() + 1, sizeof(ReinaImage)));
        if (!s->images) throw std::bad_alloc();
        for (size_t i = 0; i < L.images.size(); ++i) {   // take over the pixel buffers
            s->images[i] = *L.images[i];
            L.images[i]->rgba8 = nullptr;
        }
        s->num_instances = (uint32_t)L.instances.size();
        s->instances = dup(L.instances);
    } catch (...) {
        reina_rt_gltf_free(s);
        throw;
    }
    return s;
}

int run(const uint8_t* bytes, uint64_t len, const std::string& dir, uint32_t flags, ReinaGltfScene** out) {
    try {
        std::unique_ptr<Loaded> L = load(bytes, len, dir, flags);
        *out = to_c(*L);
    } catch (const Error& e) {
        return fail(e.code, e.msg);
    } catch (const std::bad_alloc&) {
        return fail(REINA_RT_ERR_OUT_OF_MEMORY, "host allocation failed while loading a glTF asset");
    }
    return REINA_RT_OK;
}

}  // namespace

extern "C" {

int reina_rt_gltf_load(const char* path, uint32_t flags, ReinaGltfScene** out) {
    if (!path || !out) return fail(REINA_RT_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (flags & ~REINA_RT_GLTF_SKIP_UNDECODABLE_IMAGES) return fail(REINA_RT_ERR_INVALID_ARG, "unknown glTF load flags");
    std::string data;
    try {
        data = read_file(path);
    } catch (const Error& e) {
        return fail(e.code, "Failed to find glTF file: " + std::string(path));
    }
    std::string dir(path);
    const size_t slash = dir.find_last_of('/');
    dir = slash == std::string::npos ? std::string(".") : dir.substr(0, slash);
    return run(reinterpret_cast<const uint8_t*>(data.data()), data.size(), dir, flags, out);
}

int reina_rt_gltf_load_memory(const uint8_t* bytes, uint64_t len, uint32_t flags, ReinaGltfScene** out) {
    if (!out || (!bytes && len)) return fail(REINA_RT_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (flags & ~REINA_RT_GLTF_SKIP_UNDECODABLE_IMAGES) return fail(REINA_RT_ERR_INVALID_ARG, "unknown glTF load flags");
    return run(bytes, len, std::string(), flags, out);
}

void reina_rt_gltf_free(ReinaGltfScene* s) {
    if (!s) return;
    if (s->models) {
        for (uint32_t i = 0; i < s->num_models; ++i) {
            ReinaModelData& m = s->models[i];
            std::free(m.vertices); std::free(m.indices); std::free(m.tbns);
            std::free(m.tbn_indices); std::free(m.tex_coords); std::free(m.tex_indices);
        }
        std::free(s->models);
    }
    if (s->images) {
        for (uint32_t i = 0; i < s->num_images; ++i) std::free(s->images[i].rgba8);
        std::free(s->images);
    }
    std::free(s->instances);
    std::free(s);
}

}  // extern "C"
