// ingest_obj.cpp — Wavefront OBJ → one model's buffers, host only.
//
// Replaces Models::getObjData (src/scene/Models.cpp:117-175): Assimp's OBJ
// importer, post-processed with aiProcess_Triangulate | GenSmoothNormals |
// JoinIdenticalVertices | CalcTangentSpace | FlipUVs, and the copy of
// mMeshes[0] into ModelData. Assimp is a vcpkg dependency of the reference
// (CMakeLists.txt:8, version unpinned); this file restates the published
// algorithm of its 5.x releases, in the order its post-process registry runs
// the steps:
//
//   parse     ObjFileParser: v / vt / vn / f / l / p / o / g / usemtl; numbers
//             through Assimp's fast_atof (integer part + fraction scaled in
//             double, then summed in float: not always the correctly rounded
//             float of the decimal)
//   mesh      ObjFileImporter::createVertexArray: one vertex per face corner
//             (no sharing yet); a polyline of k points becomes k - 1 two-index
//             faces, a point record one-index faces
//   FlipUVs   v = 1 - v (the registry runs it first)
//   Triangulate  quads split at their concave corner (vertex 0 when convex)
//             into (s, s+1, s+2), (s, s+2, s+3); larger polygons by ear
//             clipping (the clipping order is this library's own)
//   GenSmoothNormals (only without normals): per-face normals, then every
//             corner gets the normalised sum over all corners within
//             1e-4 * bounding-box diagonal (175 degree limit = no limit)
//   CalcTangentSpace  per-face tangents from the first three corners, projected
//             per corner, then smoothed over corners within the same radius
//             whose normal, tangent and bitangent agree (0.9999 / cos 45)
//   JoinIdenticalVertices  a corner joins the unique vertex with the same
//             position whose other attributes lie within 1e-5
//
// Candidate corners are visited in ascending vertex order. Assimp visits them
// in the order of its spatial sort, whose ties (coincident corners) its
// std::sort leaves unspecified, so a smoothed sum can differ from Assimp's in
// its last bits; which corners join a group does not depend on that order.
#include <strings.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/reina_rt.h"

namespace rt {
int fail(int code, const std::string& msg);
}
using rt::fail;

namespace {

// ---- aiVector3D arithmetic (float, evaluated left to right) ------------------
struct V3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
};
inline V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator*(V3 a, float f) { return v3(a.x * f, a.y * f, a.z * f); }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline float length(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
// operator/=(f): nothing for f == 0, else multiply by 1 / f
inline V3 divide(V3 a, float f) {
    if (f == 0.0f) return a;
    const float inv = 1.0f / f;
    return v3(a.x * inv, a.y * inv, a.z * inv);
}
inline V3 normalize(V3 a) { return divide(a, length(a)); }
inline V3 normalize_safe(V3 a) {
    const float len = length(a);
    return len > 0.0f ? divide(a, len) : a;
}
inline bool special(float f) { return !std::isfinite(f); }

struct Error {
    std::string msg;
};

// ---- fast_atof (Assimp's fast_atoreal_move<float>) ---------------------------
constexpr int kFracDigits = 15;   // AI_FAST_ATOF_RELAVANT_DECIMALS
const double kNegPow10[kFracDigits + 1] = {0.0,   0.1,   0.01,  0.001, 0.0001, 1e-5,  1e-6,  1e-7,
                                           1e-8,  1e-9,  1e-10, 1e-11, 1e-12,  1e-13, 1e-14, 1e-15};

// strtoul10_64: decimal digits into a uint64; with max_digits, reads at most
// that many, stores the count, and skips the rest
uint64_t read_u64(const char*& c, unsigned* max_digits = nullptr) {
    unsigned cur = 0;
    uint64_t value = 0;
    if (*c < '0' || *c > '9') throw Error{"OBJ: cannot parse a number"};
    for (; *c >= '0' && *c <= '9'; ++c) {
        const uint64_t next = value * 10 + (uint64_t)(*c - '0');
        if (next < value) throw Error{"OBJ: number overflows 64 bits"};
        value = next;
        if (max_digits && ++cur == *max_digits) {
            while (*c >= '0' && *c <= '9') ++c;
            return value;
        }
    }
    if (max_digits) *max_digits = cur;
    return value;
}

float fast_atof(const char* c) {
    float f = 0.0f;
    const bool inv = (*c == '-');
    if (inv || *c == '+') ++c;
    if ((c[0] == 'n' || c[0] == 'N') && strncasecmp(c, "nan", 3) == 0) return std::nanf("");
    if ((c[0] == 'i' || c[0] == 'I') && strncasecmp(c, "inf", 3) == 0)
        return inv ? -INFINITY : INFINITY;
    if (!(c[0] >= '0' && c[0] <= '9') && !((c[0] == '.' || c[0] == ',') && c[1] >= '0' && c[1] <= '9'))
        throw Error{"OBJ: cannot parse a real number"};
    if (*c != '.' && *c != ',') f = (float)read_u64(c);
    if ((*c == '.' || *c == ',') && c[1] >= '0' && c[1] <= '9') {
        ++c;
        unsigned diff = kFracDigits;
        double pl = (double)read_u64(c, &diff);
        pl *= kNegPow10[diff];
        f += (float)pl;
    } else if (*c == '.') {
        ++c;
    }
    if (*c == 'e' || *c == 'E') {
        ++c;
        const bool einv = (*c == '-');
        if (einv || *c == '+') ++c;
        float e = (float)read_u64(c);
        if (einv) e = -e;
        f *= std::pow(10.0f, e);
    }
    return inv ? -f : f;
}

// ---- parsed model (ObjFile::Model) ------------------------------------------
enum class Prim : uint8_t { Polygon, Line, Point };

// a face record: its corners' v / vt / vn indices sit in the mesh's flat
// index arrays at [first, first + count) (vt / vn may hold fewer)
struct Face {
    Prim type;
    uint32_t nv, nvt, nvn;
    uint64_t v0, vt0, vn0;
};

struct Mesh {
    std::vector<Face> faces;
    std::vector<uint32_t> v, vt, vn;
    int material = -1;
    bool has_normals = false;
    uint64_t uv_refs = 0;
};

struct Object {
    std::string name;
    std::vector<size_t> meshes;
};

struct Model {
    std::vector<V3> pos, colors, normals, uvs;
    std::vector<Object> objects;
    std::vector<Mesh> meshes;
    std::vector<std::string> materials;
    int cur_object = -1, cur_mesh = -1;
    std::string cur_material, active_group;
    bool has_material = false;

    void create_mesh() {
        meshes.emplace_back();
        cur_mesh = (int)meshes.size() - 1;
        if (cur_object >= 0) objects[cur_object].meshes.push_back((size_t)cur_mesh);
    }
    void create_object(const std::string& name) {
        objects.push_back(Object{name, {}});
        cur_object = (int)objects.size() - 1;
        create_mesh();
        if (has_material) meshes[cur_mesh].material = material_index(cur_material);
    }
    int material_index(const std::string& name) const {
        for (size_t i = 0; i < materials.size(); ++i)
            if (materials[i] == name) return (int)i;
        return -1;
    }
};

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\f' || c == '\v'; }

// a word of the current line (the line buffer is NUL-terminated, so number
// parsing that starts in a word stops at its end at the latest)
struct Word {
    const char* p;
    size_t n;
    bool is(const char* s) const { return std::strlen(s) == n && std::memcmp(p, s, n) == 0; }
    std::string str() const { return std::string(p, n); }
};

// the whitespace-separated words of line[from..], stopping at a comment
void words(const std::string& line, size_t from, std::vector<Word>& out) {
    out.clear();
    size_t i = from;
    const size_t e = line.size();
    while (i < e) {
        while (i < e && is_space(line[i])) ++i;
        if (i >= e || line[i] == '#') break;
        size_t j = i;
        while (j < e && !is_space(line[j])) ++j;
        out.push_back(Word{line.data() + i, j - i});
        i = j;
    }
}

std::string rest_of_line(const std::string& line, size_t from) {
    size_t a = from, b = line.size();
    while (a < b && is_space(line[a])) ++a;
    while (b > a && is_space(line[b - 1])) --b;
    return line.substr(a, b - a);
}

bool numeric_word(const Word& w) {
    const char c = w.p[0];
    return (c >= '0' && c <= '9') || c == '-' || c == '+' || (w.n >= 3 && (strncasecmp(w.p, "nan", 3) == 0 ||
                                                                            strncasecmp(w.p, "inf", 3) == 0));
}

// the numeric words of a record (Assimp counts components this way)
size_t numbers(const std::vector<Word>& ws, float* out, size_t cap) {
    size_t n = 0;
    for (const Word& w : ws)
        if (numeric_word(w)) {
            const float f = fast_atof(w.p);
            if (n < cap) out[n] = f;
            ++n;
        }
    return n;
}

// ObjFileParser::getFace: "v", "v/vt", "v//vn", "v/vt/vn"; negative = relative
void parse_face(Model& m, const std::vector<Word>& ws, Prim type) {
    const bool have_vt = !m.uvs.empty(), have_vn = !m.normals.empty();
    if (m.cur_object < 0) m.create_object("defaultobject");
    if (m.cur_mesh < 0) m.create_mesh();
    Mesh& mesh = m.meshes[m.cur_mesh];
    Face f{type, 0, 0, 0, mesh.v.size(), mesh.vt.size(), mesh.vn.size()};
    bool has_normal = false;
    for (const Word& w : ws) {
        int slot = 0;
        size_t i = 0;
        while (i < w.n) {
            if (w.p[i] == '/') {
                ++slot;
                ++i;
                continue;
            }
            // a signed decimal integer
            size_t k = i;
            const bool neg = w.p[k] == '-';
            if (neg || w.p[k] == '+') ++k;
            if (k >= w.n || w.p[k] < '0' || w.p[k] > '9') throw Error{"OBJ: invalid face index"};
            long val = 0;
            for (; k < w.n && w.p[k] >= '0' && w.p[k] <= '9'; ++k) {
                val = val * 10 + (w.p[k] - '0');
                if (val > (1l << 40)) throw Error{"OBJ: face index out of range"};
            }
            if (neg) val = -val;
            i = k;
            int s = slot;
            if (s == 1 && !have_vt && have_vn) s = 2;   // "v/vn" in files without texture coordinates
            if (val == 0) throw Error{"OBJ: invalid face index"};
            const size_t count = s == 0 ? m.pos.size() : s == 1 ? m.uvs.size() : m.normals.size();
            const long idx = val > 0 ? val - 1 : (long)count + val;
            if (idx < 0) throw Error{"OBJ: relative face index before the first element"};
            if (s == 0) { mesh.v.push_back((uint32_t)idx); ++f.nv; }
            else if (s == 1) { mesh.vt.push_back((uint32_t)idx); ++f.nvt; }
            else if (s == 2) { mesh.vn.push_back((uint32_t)idx); ++f.nvn; has_normal = true; }
            else throw Error{"OBJ: too many indices in a face element"};
        }
    }
    if (f.nv == 0) {   // "Obj: Ignoring empty face"
        mesh.vt.resize(f.vt0);
        mesh.vn.resize(f.vn0);
        return;
    }
    mesh.uv_refs += f.nvt;
    if (has_normal) mesh.has_normals = true;
    mesh.faces.push_back(f);
}

// ObjFileParser::getMaterialDesc: a new mesh when the current one has faces
// and another material
void use_material(Model& m, const std::string& name) {
    if (m.has_material && m.cur_material == name) return;
    if (m.material_index(name) < 0) m.materials.push_back(name);
    m.cur_material = name;
    m.has_material = true;
    const int idx = m.material_index(name);
    bool need_new = m.cur_mesh < 0;
    if (!need_new) {
        const Mesh& cur = m.meshes[m.cur_mesh];
        need_new = cur.material != -1 && cur.material != idx && !cur.faces.empty();
    }
    if (need_new) m.create_mesh();
    m.meshes[m.cur_mesh].material = idx;
}

Model parse(const char* text, uint64_t len) {
    Model m;
    uint64_t pos = 0;
    std::string line;
    std::vector<Word> ws;
    float n[8];
    while (pos < len) {
        const char* nl = static_cast<const char*>(std::memchr(text + pos, '\n', len - pos));
        const uint64_t e = nl ? (uint64_t)(nl - text) : len;
        line.assign(text + pos, text + e);
        pos = e + 1;
        size_t i = 0;
        while (i < line.size() && is_space(line[i])) ++i;
        if (i >= line.size()) continue;
        size_t j = i;
        while (j < line.size() && !is_space(line[j])) ++j;
        const Word key{line.data() + i, j - i};
        if (key.is("v")) {
            words(line, j, ws);
            const size_t c = numbers(ws, n, 8);
            if (c == 3) {
                m.pos.push_back(v3(n[0], n[1], n[2]));
            } else if (c == 4) {
                if (n[3] == 0.0f) throw Error{"OBJ: invalid component in homogeneous vector (division by zero)"};
                m.pos.push_back(v3(n[0] / n[3], n[1] / n[3], n[2] / n[3]));
            } else if (c == 6) {
                if (m.colors.size() < m.pos.size()) m.colors.resize(m.pos.size());   // earlier vertices: black
                m.pos.push_back(v3(n[0], n[1], n[2]));
                m.colors.push_back(v3(n[3], n[4], n[5]));
            }
            // other component counts add no vertex (as Assimp's parser)
            if (!m.colors.empty() && m.colors.size() < m.pos.size()) m.colors.resize(m.pos.size());
        } else if (key.is("vt")) {
            words(line, j, ws);
            const size_t c = numbers(ws, n, 8);
            if (c != 2 && c != 3) throw Error{"OBJ: invalid number of components in a texture coordinate"};
            float x = n[0], y = n[1], z = c == 3 ? n[2] : 0.0f;
            if (!std::isfinite(x)) x = 0.0f;   // nan and inf coerced to 0
            if (!std::isfinite(y)) y = 0.0f;
            if (!std::isfinite(z)) z = 0.0f;
            m.uvs.push_back(v3(x, y, z));
        } else if (key.is("vn")) {
            words(line, j, ws);
            if (ws.size() < 3) throw Error{"OBJ: normal with fewer than 3 components"};
            m.normals.push_back(v3(fast_atof(ws[0].p), fast_atof(ws[1].p), fast_atof(ws[2].p)));
        } else if (key.is("f") || key.is("l") || key.is("p")) {
            words(line, j, ws);
            parse_face(m, ws, key.is("f") ? Prim::Polygon : key.is("l") ? Prim::Line : Prim::Point);
        } else if (key.is("o")) {
            // the first word names the object; an existing name makes that
            // object current again
            words(line, j, ws);
            if (!ws.empty()) {
                const std::string name = ws[0].str();
                int found = -1;
                for (size_t k = 0; k < m.objects.size() && found < 0; ++k)
                    if (m.objects[k].name == name) found = (int)k;
                if (found >= 0) m.cur_object = found;
                else m.create_object(name);
            }
        } else if (key.is("g")) {
            // groups map to objects; the name is the rest of the line
            const std::string name = rest_of_line(line, j);
            if (!name.empty() && name != m.active_group) {
                m.create_object(name);
                m.active_group = name;
            }
        } else if (key.is("usemtl")) {
            use_material(m, rest_of_line(line, j));
        }
        // mtllib, s, comments and unknown records do not change the geometry
    }
    return m;
}

// ---- aiMesh in Assimp's verbose format ----------------------------------------
struct Verbose {
    std::vector<V3> pos, nrm, uv, col, tan, bit;
    std::vector<float> col_a;   // colour alpha: 1 where copied from the file
    bool has_n = false, has_uv = false, has_col = false, has_tan = false;
    // faces: face k's corners are idx[start[k] .. start[k+1])
    std::vector<uint32_t> idx;
    std::vector<uint64_t> start{0};
    bool has_polys = false;   // triangles or larger polygons
    size_t num_faces() const { return start.size() - 1; }
    size_t face_size(size_t k) const { return (size_t)(start[k + 1] - start[k]); }
    const uint32_t* face(size_t k) const { return idx.data() + start[k]; }
    void add_face(std::initializer_list<uint32_t> c) {
        idx.insert(idx.end(), c);
        start.push_back(idx.size());
    }
};

// ObjFileImporter::createVertexArray for one mesh
Verbose build_mesh(const Model& m, const Mesh& mesh) {
    Verbose o;
    uint64_t nv = 0;
    for (const Face& f : mesh.faces) nv += f.type == Prim::Line ? 2ull * (f.nv - 1) : f.nv;
    o.has_n = !m.normals.empty() && mesh.has_normals;
    o.has_uv = !m.uvs.empty() && mesh.uv_refs != 0;
    o.has_col = !m.colors.empty();
    o.pos.resize(nv);
    if (o.has_n) o.nrm.resize(nv);
    if (o.has_uv) o.uv.resize(nv);
    if (o.has_col) {
        o.col.resize(nv);
        o.col_a.resize(nv, 0.0f);
    }
    o.idx.reserve(nv);
    o.start.reserve(mesh.faces.size() + 1);
    bool normals_ok = true, uv_ok = true;
    uint32_t next = 0;
    for (const Face& f : mesh.faces) {
        const size_t k = f.nv;
        if (f.type == Prim::Polygon && k >= 3) o.has_polys = true;
        const uint64_t line_first = o.start.size() - 1;   // first segment face of a polyline
        for (size_t c = 0; c < k; ++c) {
            const uint32_t vi = mesh.v[f.v0 + c];
            if (vi >= m.pos.size()) throw Error{"OBJ: vertex index out of range"};
            if (!std::isfinite(m.pos[vi].x) || !std::isfinite(m.pos[vi].y) || !std::isfinite(m.pos[vi].z))
                throw Error{"OBJ: non-finite vertex position"};
            o.pos[next] = m.pos[vi];
            if (o.has_n && normals_ok && c < f.nvn) {
                const uint32_t ni = mesh.vn[f.vn0 + c];
                if (ni >= m.normals.size()) normals_ok = false;
                else o.nrm[next] = m.normals[ni];
            }
            if (o.has_col) {
                o.col[next] = m.colors[vi];
                o.col_a[next] = 1.0f;
            }
            if (o.has_uv && uv_ok && c < f.nvt) {
                const uint32_t ti = mesh.vt[f.vt0 + c];
                if (ti >= m.uvs.size()) uv_ok = false;
                else o.uv[next] = m.uvs[ti];
            }
            if (f.type == Prim::Polygon) {
                o.idx.push_back(next);
            } else if (f.type == Prim::Point) {
                o.add_face({next});
            } else {
                // a polyline: segment c starts at this corner; inner corners
                // are duplicated so each segment owns its two vertices
                const bool last = c + 1 == k;
                if (!last) o.add_face({next, 0u});
                if (c > 0) {
                    if (!last) {
                        o.pos[next + 1] = o.pos[next];
                        if (o.has_n && f.nvn) o.nrm[next + 1] = o.nrm[next];
                        if (o.has_uv) o.uv[next + 1] = o.uv[next];
                        ++next;
                    }
                    const uint64_t seg = line_first + c - 1;   // the segment ending here
                    o.idx[o.start[seg] + 1] = next;
                }
            }
            ++next;
        }
        if (f.type == Prim::Polygon) o.start.push_back(o.idx.size());
    }
    return o;
}

// TriangulateProcess for n > 4: ear clipping in the polygon's best-fit plane
void clip_ears(const Verbose& o, const uint32_t* poly, size_t n, Verbose& out) {
    V3 nrm;   // Newell normal
    for (size_t i = 0; i < n; ++i) {
        const V3 a = o.pos[poly[i]], b = o.pos[poly[(i + 1) % n]];
        nrm = nrm + v3((a.y - b.y) * (a.z + b.z), (a.z - b.z) * (a.x + b.x), (a.x - b.x) * (a.y + b.y));
    }
    const float ax = std::fabs(nrm.x), ay = std::fabs(nrm.y), az = std::fabs(nrm.z);
    int drop = 2;
    if (ax >= ay && ax >= az) drop = 0;
    else if (ay >= az) drop = 1;
    const float sgn = (drop == 0 ? nrm.x : drop == 1 ? nrm.y : nrm.z) < 0.0f ? -1.0f : 1.0f;
    std::vector<float> px(n), py(n);
    for (size_t i = 0; i < n; ++i) {
        const V3 p = o.pos[poly[i]];
        if (drop == 0) { px[i] = p.y; py[i] = p.z; }
        else if (drop == 1) { px[i] = p.z; py[i] = p.x; }
        else { px[i] = p.x; py[i] = p.y; }
        px[i] *= sgn;
    }
    std::vector<size_t> ring(n);
    for (size_t i = 0; i < n; ++i) ring[i] = i;
    const auto area2 = [&](size_t a, size_t b, size_t c) {
        return (px[b] - px[a]) * (py[c] - py[a]) - (py[b] - py[a]) * (px[c] - px[a]);
    };
    size_t guard = 0;
    size_t i = 0;
    while (ring.size() > 3) {
        const size_t m = ring.size();
        const size_t a = ring[(i + m - 1) % m], b = ring[i % m], c = ring[(i + 1) % m];
        bool ear = area2(a, b, c) > 0.0f;
        for (size_t k = 0; ear && k < m; ++k) {
            const size_t p = ring[k];
            if (p == a || p == b || p == c) continue;
            if (area2(a, b, p) >= 0.0f && area2(b, c, p) >= 0.0f && area2(c, a, p) >= 0.0f) ear = false;
        }
        if (ear || guard > m) {   // no ear in a full cycle: degenerate polygon, clip anyway
            out.add_face({poly[a], poly[b], poly[c]});
            ring.erase(ring.begin() + (long)(i % m));
            guard = 0;
            if (i >= ring.size()) i = 0;
        } else {
            i = (i + 1) % m;
            ++guard;
        }
    }
    out.add_face({poly[ring[0]], poly[ring[1]], poly[ring[2]]});
}

void triangulate(Verbose& o) {
    Verbose out;   // only its face arrays are used
    out.idx.reserve(o.idx.size() * 3 / 2);
    out.start.reserve(o.num_faces() * 2 + 1);
    for (size_t k = 0; k < o.num_faces(); ++k) {
        const uint32_t* f = o.face(k);
        const size_t n = o.face_size(k);
        if (n <= 3) {
            out.idx.insert(out.idx.end(), f, f + n);
            out.start.push_back(out.idx.size());
        } else if (n == 4) {
            unsigned s = 0;
            for (unsigned i = 0; i < 4; ++i) {
                const V3 p0 = o.pos[f[(i + 3) % 4]], p1 = o.pos[f[(i + 2) % 4]], p2 = o.pos[f[(i + 1) % 4]];
                const V3 p = o.pos[f[i]];
                const V3 left = normalize(p0 - p), diag = normalize(p1 - p), right = normalize(p2 - p);
                const float angle = std::acos(dot(left, diag)) + std::acos(dot(right, diag));
                if (angle > 3.1415926538f) {   // the concave corner
                    s = i;
                    break;
                }
            }
            out.add_face({f[s], f[(s + 1) % 4], f[(s + 2) % 4]});
            out.add_face({f[s], f[(s + 2) % 4], f[(s + 3) % 4]});
        } else {
            clip_ears(o, f, n, out);
        }
    }
    o.idx.swap(out.idx);
    o.start.swap(out.start);
}

// ComputePositionEpsilon: 1e-4 * |max - min| over all positions
float position_epsilon(const Verbose& o) {
    V3 lo = v3(1e10f, 1e10f, 1e10f), hi = v3(-1e10f, -1e10f, -1e10f);
    for (const V3& p : o.pos) {
        lo = v3(lo.x < p.x ? lo.x : p.x, lo.y < p.y ? lo.y : p.y, lo.z < p.z ? lo.z : p.z);
        hi = v3(p.x < hi.x ? hi.x : p.x, p.y < hi.y ? hi.y : p.y, p.z < hi.z ? hi.z : p.z);
    }
    return length(hi - lo) * 1e-4f;
}

// open-addressing table from a 3-part key to the head of an intrusive list
// (next[] links the entries; a new entry becomes the head)
template <class Key>
struct ListTable {
    struct Slot {
        Key key;
        uint32_t head;
        bool used;
    };
    std::vector<Slot> slots;
    uint64_t mask;
    explicit ListTable(size_t n) {
        size_t cap = 16;
        while (cap < 2 * n + 16) cap <<= 1;
        slots.assign(cap, Slot{Key{}, 0u, false});
        mask = cap - 1;
    }
    static uint64_t hash(const Key& k) {
        uint64_t h = 0x9E3779B97F4A7C15ull;
        for (int i = 0; i < 3; ++i) {
            h ^= (uint64_t)k.c[i] + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
            h *= 0xBF58476D1CE4E5B9ull;
        }
        return h ^ (h >> 31);
    }
    // the slot of key (inserting it with head = none when absent)
    Slot& at(const Key& k) {
        for (uint64_t i = hash(k) & mask;; i = (i + 1) & mask) {
            Slot& s = slots[i];
            if (!s.used) {
                s.used = true;
                s.key = k;
                s.head = 0xFFFFFFFFu;
                return s;
            }
            if (s.key == k) return s;
        }
    }
    const Slot* find(const Key& k) const {
        for (uint64_t i = hash(k) & mask;; i = (i + 1) & mask) {
            const Slot& s = slots[i];
            if (!s.used) return nullptr;
            if (s.key == k) return &s;
        }
    }
};

struct CellKey {
    int64_t c[3];
    bool operator==(const CellKey& o) const { return c[0] == o.c[0] && c[1] == o.c[1] && c[2] == o.c[2]; }
};

// all corners strictly within r of a position, ascending: a uniform grid of
// cell size 16 r, so a query ball touches 1 to 8 cells
struct Neighbours {
    float r, r2;
    double cell;
    const std::vector<V3>& pos;
    ListTable<CellKey> table;
    std::vector<uint32_t> next;
    static int64_t coord(double x, double cell) {
        const double c = std::floor(x / cell);
        return (int64_t)std::max(-4.0e18, std::min(4.0e18, c));
    }
    Neighbours(const std::vector<V3>& p, float radius)
        : r(radius), r2(radius * radius), pos(p), table(p.size()), next(p.size()) {
        cell = (r > 0.0f && std::isfinite(r)) ? 16.0 * (double)r : 1.0;
        for (uint32_t i = 0; i < (uint32_t)pos.size(); ++i) {
            auto& s = table.at(CellKey{{coord(pos[i].x, cell), coord(pos[i].y, cell), coord(pos[i].z, cell)}});
            next[i] = s.head;
            s.head = i;
        }
    }
    void find(V3 q, std::vector<uint32_t>& out) const {
        out.clear();
        // a neighbour differs from q by less than r (slightly more after the
        // float rounding of its squared distance) in every coordinate
        const double m = 1.01 * (double)r;
        int64_t lo[3], hi[3];
        const double qc[3] = {q.x, q.y, q.z};
        for (int a = 0; a < 3; ++a) {
            lo[a] = coord(qc[a] - m, cell);
            hi[a] = coord(qc[a] + m, cell);
        }
        for (int64_t x = lo[0]; x <= hi[0]; ++x)
            for (int64_t y = lo[1]; y <= hi[1]; ++y)
                for (int64_t z = lo[2]; z <= hi[2]; ++z) {
                    const auto* s = table.find(CellKey{{x, y, z}});
                    if (!s) continue;
                    for (uint32_t j = s->head; j != 0xFFFFFFFFu; j = next[j]) {
                        const V3 d = pos[j] - q;
                        if (d.x * d.x + d.y * d.y + d.z * d.z < r2) out.push_back(j);
                    }
                }
        std::sort(out.begin(), out.end());
    }
};

// GenVertexNormalsProcess (no angle limit)
void gen_smooth_normals(Verbose& o) {
    if (o.has_n || !o.has_polys) return;
    const float qnan = std::numeric_limits<float>::quiet_NaN();
    std::vector<V3> face_n(o.pos.size());
    for (size_t k = 0; k < o.num_faces(); ++k) {
        const uint32_t* f = o.face(k);
        const size_t n = o.face_size(k);
        if (n < 3) {
            for (size_t c = 0; c < n; ++c) face_n[f[c]] = v3(qnan, qnan, qnan);
            continue;
        }
        const V3 p1 = o.pos[f[0]], p2 = o.pos[f[1]], p3 = o.pos[f[n - 1]];
        const V3 nv = normalize_safe(cross(p2 - p1, p3 - p1));
        for (size_t c = 0; c < n; ++c) face_n[f[c]] = nv;
    }
    const float eps = position_epsilon(o);
    Neighbours nb(o.pos, eps);
    std::vector<uint32_t> found;
    std::vector<bool> had(o.pos.size(), false);
    o.nrm.assign(o.pos.size(), V3{});
    for (uint32_t i = 0; i < (uint32_t)o.pos.size(); ++i) {
        if (had[i]) continue;
        nb.find(o.pos[i], found);
        V3 sum;
        for (uint32_t j : found)
            if (!std::isnan(face_n[j].x)) sum = sum + face_n[j];
        sum = normalize_safe(sum);
        for (uint32_t j : found) {
            o.nrm[j] = sum;
            had[j] = true;
        }
    }
    o.has_n = true;
}

// CalcTangentsProcess
void calc_tangents(Verbose& o) {
    if (!o.has_polys || !o.has_n || !o.has_uv) return;
    const size_t nv = o.pos.size();
    const float qnan = std::numeric_limits<float>::quiet_NaN();
    o.tan.assign(nv, V3{});
    o.bit.assign(nv, V3{});
    std::vector<bool> done(nv, false);
    for (size_t k = 0; k < o.num_faces(); ++k) {
        const uint32_t* f = o.face(k);
        const size_t n = o.face_size(k);
        if (n < 3) {
            for (size_t c = 0; c < n; ++c) {
                done[f[c]] = true;
                o.tan[f[c]] = v3(qnan, qnan, qnan);
                o.bit[f[c]] = v3(qnan, qnan, qnan);
            }
            continue;
        }
        const uint32_t p0 = f[0], p1 = f[1], p2 = f[2];
        const V3 v = o.pos[p1] - o.pos[p0], w = o.pos[p2] - o.pos[p0];
        float sx = o.uv[p1].x - o.uv[p0].x, sy = o.uv[p1].y - o.uv[p0].y;
        float tx = o.uv[p2].x - o.uv[p0].x, ty = o.uv[p2].y - o.uv[p0].y;
        const float dir = (tx * sy - ty * sx) < 0.0f ? -1.0f : 1.0f;
        if (sx * ty == sy * tx) {   // no UV extent: the default directions
            sx = 0.0f; sy = 1.0f;
            tx = 1.0f; ty = 0.0f;
        }
        const V3 tangent = v3((w.x * sy - v.x * ty) * dir, (w.y * sy - v.y * ty) * dir, (w.z * sy - v.z * ty) * dir);
        const V3 bitangent = v3((w.x * sx - v.x * tx) * dir, (w.y * sx - v.y * tx) * dir, (w.z * sx - v.z * tx) * dir);
        for (size_t c = 0; c < n; ++c) {
            const uint32_t p = f[c];
            const V3 nn = o.nrm[p];
            V3 lt = tangent - nn * dot(tangent, nn);
            V3 lb = (bitangent - nn * dot(bitangent, nn)) - lt * dot(bitangent, lt);
            lt = normalize_safe(lt);
            lb = normalize_safe(lb);
            const bool bad_t = special(lt.x) || special(lt.y) || special(lt.z);
            const bool bad_b = special(lb.x) || special(lb.y) || special(lb.z);
            if (bad_t != bad_b) {
                if (bad_t) lt = normalize_safe(cross(nn, lb));
                else lb = normalize_safe(cross(lt, nn));
            }
            o.tan[p] = lt;
            o.bit[p] = lb;
        }
    }
    const float eps = position_epsilon(o);
    const float limit = std::cos(45.0f * 0.0174532925f);
    const float angle_eps = 0.9999f;
    Neighbours nb(o.pos, eps);
    std::vector<uint32_t> found, group;
    for (uint32_t a = 0; a < (uint32_t)nv; ++a) {
        if (done[a]) continue;
        const V3 on = o.nrm[a], ot = o.tan[a], ob = o.bit[a];
        nb.find(o.pos[a], found);
        group.clear();
        group.push_back(a);
        for (uint32_t j : found) {
            if (done[j]) continue;
            if (dot(o.nrm[j], on) < angle_eps) continue;
            if (dot(o.tan[j], ot) < limit) continue;
            if (dot(o.bit[j], ob) < limit) continue;
            group.push_back(j);   // `a` itself passes too and is counted twice, as in Assimp
            done[j] = true;
        }
        V3 st, sb;
        for (uint32_t j : group) {
            st = st + o.tan[j];
            sb = sb + o.bit[j];
        }
        st = normalize(st);
        sb = normalize(sb);
        for (uint32_t j : group) {
            o.tan[j] = st;
            o.bit[j] = sb;
        }
    }
    o.has_tan = true;
}

struct PosKey {
    uint32_t c[3];
    bool operator==(const PosKey& o) const { return c[0] == o.c[0] && c[1] == o.c[1] && c[2] == o.c[2]; }
};

// JoinVerticesProcess
void join_vertices(Verbose& o) {
    const size_t nv = o.pos.size();
    const float sq_eps = 1e-5f * 1e-5f;
    const auto close = [sq_eps](V3 a, V3 b) {
        const V3 d = a - b;
        return !(d.x * d.x + d.y * d.y + d.z * d.z > sq_eps);
    };
    const auto equal = [&](uint32_t a, uint32_t b) {
        if (!close(o.pos[a], o.pos[b])) return false;
        if (o.has_n && !close(o.nrm[a], o.nrm[b])) return false;
        if (o.has_tan && (!close(o.tan[a], o.tan[b]) || !close(o.bit[a], o.bit[b]))) return false;
        if (o.has_uv && !close(o.uv[a], o.uv[b])) return false;
        if (o.has_col) {
            const V3 d = o.col[a] - o.col[b];
            const float da = o.col_a[a] - o.col_a[b];
            if (d.x * d.x + d.y * d.y + d.z * d.z + da * da > sq_eps) return false;
        }
        return true;
    };
    // candidates share the position (+0 and -0 alike); the most recently
    // added unique vertex is tried first
    const auto bits = [](float f) {
        uint32_t u;
        std::memcpy(&u, &f, 4);
        return f == 0.0f ? 0u : u;
    };
    ListTable<PosKey> table(nv);
    std::vector<uint32_t> unique, remap(nv), next;
    unique.reserve(nv);
    next.reserve(nv);
    for (uint32_t a = 0; a < (uint32_t)nv; ++a) {
        const V3 p = o.pos[a];
        auto& s = table.at(PosKey{{bits(p.x), bits(p.y), bits(p.z)}});
        uint32_t match = 0xFFFFFFFFu;
        for (uint32_t u = s.head; u != 0xFFFFFFFFu; u = next[u])
            if (equal(a, unique[u])) {
                match = u;
                break;
            }
        if (match == 0xFFFFFFFFu) {
            const uint32_t u = (uint32_t)unique.size();
            remap[a] = u;
            unique.push_back(a);
            next.push_back(s.head);
            s.head = u;
        } else {
            remap[a] = match;
        }
    }
    const auto gather = [&](std::vector<V3>& arr) {
        if (arr.empty()) return;
        std::vector<V3> out(unique.size());
        for (size_t i = 0; i < unique.size(); ++i) out[i] = arr[unique[i]];
        arr.swap(out);
    };
    gather(o.pos);
    gather(o.nrm);
    gather(o.uv);
    gather(o.col);
    gather(o.tan);
    gather(o.bit);
    if (!o.col_a.empty()) {
        std::vector<float> out(unique.size());
        for (size_t i = 0; i < unique.size(); ++i) out[i] = o.col_a[unique[i]];
        o.col_a.swap(out);
    }
    for (uint32_t& i : o.idx) i = remap[i];
}

template <class T>
T* dup(const std::vector<T>& v) {
    if (v.empty()) return nullptr;
    T* p = static_cast<T*>(std::malloc(v.size() * sizeof(T)));
    if (!p) throw std::bad_alloc();
    std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

ReinaModelData* load(const char* text, uint64_t len) {
    Model m = parse(text, len);
    const Mesh* first = nullptr;
    for (const Object& ob : m.objects) {
        for (size_t id : ob.meshes)
            if (!m.meshes[id].faces.empty()) {
                first = &m.meshes[id];
                break;
            }
        if (first) break;
    }
    if (!first) throw Error{"OBJ: the file has no faces (no mesh to load)"};
    Verbose o = build_mesh(m, *first);
    if (o.has_uv)
        for (V3& t : o.uv) t.y = 1.0f - t.y;   // FlipUVs
    triangulate(o);
    gen_smooth_normals(o);
    calc_tangents(o);
    join_vertices(o);

    // Models::getObjData: xyz1 vertices, mat3(T, B, N) per vertex when the
    // mesh has normals, uv when it has texture coordinates, and every face's
    // indices (all three index streams alike)
    std::vector<float> verts, tbns, uvs;
    std::vector<uint32_t> idx;
    verts.reserve(o.pos.size() * 4);
    for (size_t i = 0; i < o.pos.size(); ++i) {
        verts.insert(verts.end(), {o.pos[i].x, o.pos[i].y, o.pos[i].z, 1.0f});
        if (o.has_n) {
            const V3 t = o.has_tan ? o.tan[i] : V3{}, b = o.has_tan ? o.bit[i] : V3{}, n = o.nrm[i];
            tbns.insert(tbns.end(), {t.x, t.y, t.z, b.x, b.y, b.z, n.x, n.y, n.z});
        }
        if (o.has_uv) uvs.insert(uvs.end(), {o.uv[i].x, o.uv[i].y});
    }
    idx = o.idx;

    ReinaModelData* r = static_cast<ReinaModelData*>(std::calloc(1, sizeof(ReinaModelData)));
    if (!r) throw std::bad_alloc();
    try {
        r->vertices = dup(verts);            r->num_vertices = o.pos.size();
        r->indices = dup(idx);               r->num_indices = idx.size();
        r->tbns = dup(tbns);                 r->num_tbns = tbns.size() / 9;
        r->tbn_indices = dup(idx);           r->num_tbn_indices = idx.size();
        r->tex_coords = dup(uvs);            r->num_tex_coords = uvs.size() / 2;
        r->tex_indices = dup(idx);           r->num_tex_indices = idx.size();
    } catch (...) {
        reina_rt_model_free(r);
        throw;
    }
    return r;
}

}  // namespace

extern "C" {

int reina_rt_model_load_obj_memory(const char* text, uint64_t len, ReinaModelData** out) {
    if (!out || (!text && len)) return fail(REINA_RT_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    try {
        *out = load(text, len);
    } catch (const Error& e) {
        return fail(REINA_RT_ERR_INVALID_ARG, e.msg);
    } catch (const std::bad_alloc&) {
        return fail(REINA_RT_ERR_OUT_OF_MEMORY, "host allocation failed while loading an OBJ file");
    }
    return REINA_RT_OK;
}

int reina_rt_model_load_obj(const char* path, ReinaModelData** out) {
    if (!path || !out) return fail(REINA_RT_ERR_INVALID_ARG, "null argument");
    std::ifstream f(path, std::ios::binary);
    if (!f) return fail(REINA_RT_ERR_INVALID_ARG, std::string("Failed to load model: cannot open ") + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    const int rc = reina_rt_model_load_obj_memory(text.data(), text.size(), out);
    if (rc != REINA_RT_OK) return fail(rc, std::string("Failed to load model ") + path + ": " + reina_rt_last_error());
    return rc;
}

void reina_rt_model_free(ReinaModelData* m) {
    if (!m) return;
    std::free(m->vertices);
    std::free(m->indices);
    std::free(m->tbns);
    std::free(m->tbn_indices);
    std::free(m->tex_coords);
    std::free(m->tex_indices);
    std::free(m);
}

}  // extern "C"
