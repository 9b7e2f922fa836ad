// reina_headless.cpp — a headless Reina host driving libreina_rt through its
// C ABI: the render loop of src/Reina.cpp:296-392 without a window.
//
//   reina_headless <config.toml> <scene.rscn|scene.glb|scene.gltf> <output-dir> [push_consts.bin]
//
// config.toml uses the reference's keys (config/config.toml): [camera.dof]
// focus_dist, defocus_multiplier; [sampling] samples_per_pixel, max_bounces,
// direct_clamp, indirect_clamp; [saving] save_on_samples, save_on_times;
// [postprocessing.bloom] radius, threshold, intensity;
// [postprocessing.tonemap] exposure; plus a [headless] section for what the
// reference takes from its window and camera: width, height, eye, center,
// fov, total_samples (the run's end), nee, print_summary (default 1). Each
// frame is one reina_rt_trace (samples_per_pixel samples, sampleBatch = frame
// index, running mean as Reina::traceRays). Saving follows SaveManager and
// Clock (see SaveManager below): the image is post-processed (bloom + ACES
// tone map, reina_rt_postprocess) and written as <output-dir>/output_<N>spp.png
// or output_<T>sec.png, as Reina's save() writes its RGBA8 swapchain image
// (src/Reina.cpp:19-47); after the first frame every frame prints the clock
// summary (src/tools/Clock.cpp summary(); categories "Ray Tracing" and "Save":
// there is no display, so post-processing runs only for a save). The scene is
// a glTF asset (loaded with reina_rt_gltf_load and packed as Models::addModel /
// Scene::addInstance do, the reference's gltf::loadScene path of
// Reina.cpp:91) or the packed ReinaSceneDesc arrays
// (reina_rt.scene.save_scene); push_consts.bin, if given, replaces the camera
// and sampling push constants by a host's own 160-byte RtPushConsts.
#include <hip/hip_runtime.h>

#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/reina_rt.h"

namespace {

// ---- minimal TOML: [section.sub] headers, key = number | [n, n, ...], # comments
struct Config {
    std::map<std::string, std::vector<double>> v;
    double num(const std::string& k, double dflt) const {
        auto it = v.find(k);
        return (it == v.end() || it->second.empty()) ? dflt : it->second[0];
    }
    std::vector<double> list(const std::string& k, std::vector<double> dflt) const {
        auto it = v.find(k);
        return it == v.end() ? dflt : it->second;
    }
};

std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

Config parse_toml(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    Config c;
    std::string line, section;
    while (std::getline(f, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line = line.substr(0, hash);
        line = trim(line);
        if (line.empty()) continue;
        if (line.front() == '[') {
            section = trim(line.substr(1, line.find(']') - 1));
            continue;
        }
        const size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        std::string key = trim(line.substr(0, eq)), val = trim(line.substr(eq + 1));
        std::vector<double> nums;
        for (char& ch : val)
            if (ch == '[' || ch == ']' || ch == ',') ch = ' ';
        std::istringstream is(val);
        std::string tok;
        while (is >> tok) {
            if (tok == "true") nums.push_back(1.0);
            else if (tok == "false") nums.push_back(0.0);
            else nums.push_back(std::stod(tok));
        }
        c.v[section.empty() ? key : section + "." + key] = nums;
    }
    return c;
}

// ---- packed scene file ------------------------------------------------------
struct SceneFile {
    std::vector<float> vertices, tbns, tex_coords;
    std::vector<uint32_t> indices, tbn_indices, tex_indices;
    std::vector<ReinaInstanceProperties> props;
    std::vector<ReinaModelRange> models;
    std::vector<ReinaInstanceDesc> instances;
    std::vector<std::vector<uint8_t>> texels;   // RGBA8 per texture
    std::vector<ReinaTextureDesc> textures;
};

template <typename T>
void read_array(std::ifstream& f, std::vector<T>& out, size_t elem_bytes) {
    uint64_t n = 0;
    f.read(reinterpret_cast<char*>(&n), 8);
    if (!f || n > (1ull << 34) / elem_bytes) throw std::runtime_error("scene file: bad array size");
    out.resize(n * elem_bytes / sizeof(T));
    f.read(reinterpret_cast<char*>(out.data()), (std::streamsize)(n * elem_bytes));
    if (!f) throw std::runtime_error("scene file: truncated");
}

SceneFile load_scene(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    char magic[4];
    uint32_t version = 0;
    f.read(magic, 4);
    f.read(reinterpret_cast<char*>(&version), 4);
    if (!f || std::memcmp(magic, "RSCN", 4) != 0 || (version != 1 && version != 2))
        throw std::runtime_error("not a scene file");
    SceneFile s;
    read_array(f, s.vertices, 16);
    read_array(f, s.indices, 4);
    read_array(f, s.tbns, 36);
    read_array(f, s.tbn_indices, 4);
    read_array(f, s.tex_coords, 8);
    read_array(f, s.tex_indices, 4);
    read_array(f, s.props, sizeof(ReinaInstanceProperties));
    read_array(f, s.models, sizeof(ReinaModelRange));
    read_array(f, s.instances, sizeof(ReinaInstanceDesc));
    if (version >= 2) {   // textures: count, then width, height, RGBA8 texels each
        uint64_t n = 0;
        f.read(reinterpret_cast<char*>(&n), 8);
        if (!f || n > (1u << 20)) throw std::runtime_error("scene file: bad texture count");
        s.texels.resize(n);
        s.textures.resize(n);
        for (uint64_t t = 0; t < n; ++t) {
            uint32_t wh[2] = {0, 0};
            f.read(reinterpret_cast<char*>(wh), 8);
            if (!f || wh[0] == 0 || wh[1] == 0 || wh[0] > (1u << 16) || wh[1] > (1u << 16))
                throw std::runtime_error("scene file: bad texture size");
            s.texels[t].resize((size_t)wh[0] * wh[1] * 4);
            f.read(reinterpret_cast<char*>(s.texels[t].data()), (std::streamsize)s.texels[t].size());
            if (!f) throw std::runtime_error("scene file: truncated");
            s.textures[t] = ReinaTextureDesc{wh[0], wh[1], s.texels[t].data()};
        }
    }
    return s;
}

// ---- a glTF scene packed as Models::addModel / Scene::addInstance do -------
// (src/scene/Models.cpp:24-92: per-model offsets added to the index streams;
// src/scene/Scene.cpp:27-54: one InstanceProperties per instance)
SceneFile scene_from_gltf(const std::string& path) {
    ReinaGltfScene* g = nullptr;
    if (reina_rt_gltf_load(path.c_str(), 0, &g) != REINA_RT_OK)
        throw std::runtime_error(std::string("reina_rt_gltf_load: ") + reina_rt_last_error());
    SceneFile s;
    std::vector<ReinaModelRange> ranges;
    for (uint32_t m = 0; m < g->num_models; ++m) {
        const ReinaModelData& md = g->models[m];
        ReinaModelRange r{};
        const uint32_t vo = (uint32_t)(s.vertices.size() / 4), to = (uint32_t)(s.tbns.size() / 9),
                       xo = (uint32_t)(s.tex_coords.size() / 2);
        r.firstVertex = vo;
        r.firstNormal = to;
        r.indexOffset = (uint32_t)s.indices.size();
        r.tbnsIndexOffset = (uint32_t)s.tbn_indices.size();
        r.texIndexOffset = md.num_tex_coords ? (uint32_t)s.tex_indices.size() : 0xFFFFFFFFu;
        r.indexCount = r.tbnsIndexCount = (uint32_t)(md.num_indices / 3);
        r.texIndexCount = (uint32_t)(md.num_tex_indices / 3);
        s.vertices.insert(s.vertices.end(), md.vertices, md.vertices + 4 * md.num_vertices);
        s.tbns.insert(s.tbns.end(), md.tbns, md.tbns + 9 * md.num_tbns);
        s.tex_coords.insert(s.tex_coords.end(), md.tex_coords, md.tex_coords + 2 * md.num_tex_coords);
        for (uint64_t i = 0; i < md.num_indices; ++i) s.indices.push_back(md.indices[i] + vo);
        for (uint64_t i = 0; i < md.num_tbn_indices; ++i) s.tbn_indices.push_back(md.tbn_indices[i] + to);
        for (uint64_t i = 0; i < md.num_tex_indices; ++i)
            s.tex_indices.push_back(md.tex_indices[i] == 0xFFFFFFFFu ? 0xFFFFFFFFu : md.tex_indices[i] + xo);
        ranges.push_back(r);
    }
    for (uint32_t t = 0; t < g->num_images; ++t) {
        const ReinaImage& im = g->images[t];
        s.texels.emplace_back(im.rgba8, im.rgba8 + (size_t)im.width * im.height * 4);
    }
    for (uint32_t t = 0; t < g->num_images; ++t)
        s.textures.push_back(ReinaTextureDesc{g->images[t].width, g->images[t].height, s.texels[t].data()});
    for (uint32_t i = 0; i < g->num_instances; ++i) {
        const ReinaGltfInstance& in = g->instances[i];
        const ReinaMaterial& m = in.material;
        const ReinaModelRange& r = ranges[in.model];
        ReinaInstanceProperties p{};
        p.indicesOffset = r.indexOffset;
        std::memcpy(p.albedo, m.albedo, sizeof p.albedo);
        std::memcpy(p.emission, m.emission, sizeof p.emission);
        p.tbnsIndicesOffset = r.tbnsIndexOffset;
        p.texIndicesOffset = r.texIndexOffset;
        p.roughness = m.roughness;
        p.ior = m.ior;
        p.interpNormals = m.interpNormals ? 1 : 0;
        p.absorption = m.absorption;
        p.textureID = m.textureID;
        p.normalMapTexID = m.normalMapID;
        p.bumpMapTexID = m.bumpMapID;
        p.cullBackface = m.cullBackface ? 1u : 0u;
        p.anisotropic = m.anisotropic;
        p.subsurface = m.subsurface;
        p.clearcoatGloss = m.clearcoatGloss;
        std::memcpy(p.sheenTint, m.sheenTint, sizeof p.sheenTint);
        std::memcpy(p.specularTint, m.specularTint, sizeof p.specularTint);
        p.metallic = m.metallic;
        p.clearcoat = m.clearcoat;
        p.specularTransmission = m.specularTransmission;
        p.sheen = m.sheen;
        s.props.push_back(p);
        ReinaInstanceDesc d{};
        std::memcpy(d.transform, in.transform, sizeof d.transform);
        d.instancePropertiesID = (uint32_t)(s.props.size() - 1);
        d.materialIdx = m.materialIdx;
        d.modelID = in.model;
        s.instances.push_back(d);
    }
    s.models = ranges;
    reina_rt_gltf_free(g);
    return s;
}

bool ends_with(const std::string& s, const char* suffix) {
    const size_t n = std::strlen(suffix);
    return s.size() >= n && s.compare(s.size() - n, n, suffix) == 0;
}

// ---- camera: glm::perspective / lookAt (src/graphics/Camera.cpp:10-11), inverted
void mat_inverse(const double* m, double* inv) {   // 4x4 column-major, Gauss-Jordan
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) { a[r][c] = m[c * 4 + r]; a[r][c + 4] = r == c ? 1.0 : 0.0; }
    for (int c = 0; c < 4; ++c) {
        int p = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
        for (int k = 0; k < 8; ++k) std::swap(a[c][k], a[p][k]);
        const double d = a[c][c];
        for (int k = 0; k < 8; ++k) a[c][k] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == c) continue;
            const double f = a[r][c];
            for (int k = 0; k < 8; ++k) a[r][k] -= f * a[c][k];
        }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) inv[c * 4 + r] = a[r][c + 4];
}

void camera_matrices(const double* eye, const double* center, double fov_deg, double aspect, float* inv_view,
                     float* inv_proj) {
    double f[3] = {center[0] - eye[0], center[1] - eye[1], center[2] - eye[2]};
    double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& x : f) x /= fl;
    const double up[3] = {0.0, 1.0, 0.0};
    double s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    double sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (double& x : s) x /= sl;
    const double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
    double view[16] = {s[0], u[0], -f[0], 0, s[1], u[1], -f[1], 0, s[2], u[2], -f[2], 0,
                       -(s[0] * eye[0] + s[1] * eye[1] + s[2] * eye[2]),
                       -(u[0] * eye[0] + u[1] * eye[1] + u[2] * eye[2]),
                       (f[0] * eye[0] + f[1] * eye[1] + f[2] * eye[2]), 1};
    const double zn = 0.1, zf = 100.0, t = std::tan(fov_deg * M_PI / 360.0);
    double proj[16] = {0};
    proj[0] = 1.0 / (aspect * t);
    proj[5] = 1.0 / t;
    proj[10] = -(zf + zn) / (zf - zn);
    proj[11] = -1.0;
    proj[14] = -(2.0 * zf * zn) / (zf - zn);
    double iv[16], ip[16];
    mat_inverse(view, iv);
    mat_inverse(proj, ip);
    for (int k = 0; k < 16; ++k) { inv_view[k] = (float)iv[k]; inv_proj[k] = (float)ip[k]; }
}

// ---- PNG (RGBA8, stored deflate blocks) -----------------------------------
uint32_t crc32(const uint8_t* p, size_t n, uint32_t c = 0) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t v = i;
            for (int k = 0; k < 8; ++k) v = (v & 1) ? 0xEDB88320u ^ (v >> 1) : v >> 1;
            table[i] = v;
        }
        init = true;
    }
    c = ~c;
    for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
    return ~c;
}

void put_be32(std::vector<uint8_t>& o, uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) o.push_back((uint8_t)(v >> s));
}

void chunk(std::vector<uint8_t>& png, const char* type, const std::vector<uint8_t>& data) {
    put_be32(png, (uint32_t)data.size());
    std::vector<uint8_t> td(type, type + 4);
    td.insert(td.end(), data.begin(), data.end());
    png.insert(png.end(), td.begin(), td.end());
    put_be32(png, crc32(td.data(), td.size()));
}

void write_png(const std::string& path, const uint8_t* rgba, uint32_t W, uint32_t H) {
    std::vector<uint8_t> raw;
    raw.reserve((size_t)H * (4 * W + 1));
    for (uint32_t y = 0; y < H; ++y) {
        raw.push_back(0);   // filter: none
        raw.insert(raw.end(), rgba + (size_t)y * W * 4, rgba + (size_t)(y + 1) * W * 4);
    }
    std::vector<uint8_t> z = {0x78, 0x01};   // zlib header, no compression
    for (size_t off = 0; off < raw.size() || off == 0; off += 65535) {
        const size_t n = std::min<size_t>(65535, raw.size() - off);
        z.push_back(off + n >= raw.size() ? 1 : 0);
        z.push_back((uint8_t)(n & 0xFF));
        z.push_back((uint8_t)(n >> 8));
        z.push_back((uint8_t)(~n & 0xFF));
        z.push_back((uint8_t)((~n >> 8) & 0xFF));
        z.insert(z.end(), raw.begin() + off, raw.begin() + off + n);
        if (raw.empty()) break;
    }
    uint32_t a = 1, b = 0;   // Adler-32
    for (uint8_t x : raw) { a = (a + x) % 65521; b = (b + a) % 65521; }
    put_be32(z, (b << 16) | a);
    std::vector<uint8_t> png = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, W);
    put_be32(ihdr, H);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});   // 8-bit RGBA
    chunk(png, "IHDR", ihdr);
    chunk(png, "IDAT", z);
    chunk(png, "IEND", {});
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(png.data()), (std::streamsize)png.size());
    if (!f) throw std::runtime_error("could not write " + path);
}

// ---- SaveManager (src/tools/SaveManager.cpp) and Clock (src/tools/Clock.cpp) --
// One save per frame at most: the smallest sample threshold below the clock's
// sample count (output_<N>spp.png), else the smallest time threshold below its
// age (output_<T>sec.png). The clock counts a frame's samples from the second
// frame on (its first markFrame only starts the frame timer), so a threshold
// N fires on the first frame k with N < spp * (k - 1), and that frame's image
// holds spp * (k + 1) samples -- as in the reference.
struct SaveManager {
    std::vector<double> samples, times;   // ascending: front is next
    SaveManager(std::vector<double> s, std::vector<double> t) : samples(std::move(s)), times(std::move(t)) {
        std::sort(samples.begin(), samples.end());
        std::sort(times.begin(), times.end());
    }
    bool should_save(uint64_t n, double age, std::string& name) {
        if (!samples.empty() && (int64_t)samples.front() < (int64_t)n) {
            name = "output_" + std::to_string((int64_t)samples.front()) + "spp.png";
            samples.erase(samples.begin());
            return true;
        }
        if (!times.empty() && times.front() < age) {
            name = "output_" + std::to_string((int)times.front()) + "sec.png";
            times.erase(times.begin());
            return true;
        }
        return false;
    }
};

struct Clock {
    using tp = std::chrono::steady_clock::time_point;
    tp created = std::chrono::steady_clock::now(), last_frame{}, last_category{};
    bool frame_started = false, category_started = false;
    double frame_avg = 0.0;
    uint64_t frames = 0, samples = 0;
    std::string category;
    std::map<std::string, std::pair<double, uint64_t>> categories;   // average, recordings
    static double since(tp a, tp b) { return std::chrono::duration<double>(b - a).count(); }
    static void add(double& avg, uint64_t& n, double x) {
        avg = n == 0 ? x : (avg * (double)n + x) / (double)(n + 1);
        ++n;
    }
    double age() const { return since(created, std::chrono::steady_clock::now()); }
    uint64_t frame_count() const { return frames; }
    uint64_t sample_count() const { return samples; }
    void mark_frame(uint32_t spp) {
        const tp now = std::chrono::steady_clock::now();
        if (!frame_started) {
            frame_started = true;
            last_frame = now;
            return;
        }
        add(frame_avg, frames, since(last_frame, now));
        last_frame = now;
        samples += spp;
    }
    void mark_category(const std::string& c) {
        const tp now = std::chrono::steady_clock::now();
        categories.emplace(c, std::make_pair(0.0, (uint64_t)0));
        if (category_started) {
            auto& e = categories.at(category);
            add(e.first, e.second, since(last_category, now));
        }
        category_started = true;
        category = c;
        last_category = now;
    }
    std::string summary() const {
        std::ostringstream o;
        o << "Timer age: " << age() << "s\n";
        o << "Samples: " << samples << "\n";
        o << "Average frame time: " << frame_avg * 1000 << "ms\n";
        o << "Average time per spp: " << frame_avg * (double)frames * 1000 / (double)samples << "ms\n";
        for (const auto& e : categories) o << "Average category time | " << e.first << ": " << e.second.first * 1000 << "ms\n";
        return o.str();
    }
};

void check(int rc, const char* what) {
    if (rc != REINA_RT_OK) throw std::runtime_error(std::string(what) + ": " + reina_rt_last_error());
}

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

int run(int argc, char** argv) {
    if (argc != 4 && argc != 5) {
        std::fprintf(stderr, "usage: %s <config.toml> <scene.rscn|scene.glb|scene.gltf> <output-dir> [push_consts.bin]\n",
                     argv[0]);
        return 2;
    }
    const Config cfg = parse_toml(argv[1]);
    const std::string scene_path = argv[2];
    SceneFile sf = (ends_with(scene_path, ".glb") || ends_with(scene_path, ".gltf")) ? scene_from_gltf(scene_path)
                                                                                   : load_scene(scene_path);
    const std::string out_dir = argv[3];
    mkdir(out_dir.c_str(), 0755);   // an existing directory is fine
    const uint32_t W = (uint32_t)cfg.num("headless.width", 256), H = (uint32_t)cfg.num("headless.height", 256);
    const std::vector<double> eye = cfg.list("headless.eye", {0.0, 1.0, 3.4});
    const std::vector<double> center = cfg.list("headless.center", {0.0, 1.0, 0.0});
    if (eye.size() != 3 || center.size() != 3) throw std::runtime_error("eye / center need 3 numbers");
    const double fov = cfg.num("headless.fov", 40.0);
    const uint32_t spp = (uint32_t)cfg.num("sampling.samples_per_pixel", 8);
    const uint64_t total = (uint64_t)cfg.num("headless.total_samples", 64);
    const std::vector<double> saves = cfg.list("saving.save_on_samples", {});
    const std::vector<double> save_times = cfg.list("saving.save_on_times", {});

    ReinaSceneDesc d{};
    d.struct_size = sizeof d;
    d.vertices = sf.vertices.data(); d.num_vertices = sf.vertices.size() / 4;
    d.indices = sf.indices.data(); d.num_indices = sf.indices.size();
    d.tbns = sf.tbns.data(); d.num_tbns = sf.tbns.size() / 9;
    d.tbn_indices = sf.tbn_indices.data(); d.num_tbn_indices = sf.tbn_indices.size();
    d.tex_coords = sf.tex_coords.data(); d.num_tex_coords = sf.tex_coords.size() / 2;
    d.tex_indices = sf.tex_indices.data(); d.num_tex_indices = sf.tex_indices.size();
    d.props = sf.props.data(); d.num_props = (uint32_t)sf.props.size();
    d.models = sf.models.data(); d.num_models = (uint32_t)sf.models.size();
    d.instances = sf.instances.data(); d.num_instances = (uint32_t)sf.instances.size();
    d.textures = sf.textures.empty() ? nullptr : sf.textures.data();
    d.num_textures = (uint32_t)sf.textures.size();
    ReinaScene* scene = nullptr;
    check(reina_rt_scene_create(&d, &scene), "reina_rt_scene_create");
    if (cfg.num("headless.nee", 0) != 0) check(reina_rt_set_trace_options(scene, REINA_RT_TRACE_NEE), "options");

    RtPushConsts pc{};
    camera_matrices(eye.data(), center.data(), fov, (double)W / H, pc.invView, pc.invProjection);

    pc.sampleBatch = 0;
    pc.focusDist = (float)cfg.num("camera.dof.focus_dist", 2.2);
    pc.defocusMultiplier = (float)(cfg.num("camera.dof.defocus_multiplier", 0.0) / 100.0);   // Reina.cpp:149
    pc.directClamp = (float)cfg.num("sampling.direct_clamp", 100.0);
    pc.indirectClamp = (float)cfg.num("sampling.indirect_clamp", 10.0);
    pc.samplesPerPixel = spp;
    pc.maxBounces = (uint32_t)cfg.num("sampling.max_bounces", 8);
    if (argc == 5) {
        // the host application's own RtPushConsts (160 bytes, polyglot/raytrace.h layout)
        std::ifstream pf(argv[4], std::ios::binary);
        RtPushConsts given{};
        pf.read(reinterpret_cast<char*>(&given), sizeof given);
        if (!pf) throw std::runtime_error(std::string("cannot read 160 bytes from ") + argv[4]);
        pc = given;
        pc.sampleBatch = 0;
    }
    BloomPushConsts bloom{};
    bloom.radius = (float)cfg.num("postprocessing.bloom.radius", 5.0);
    bloom.threshold = (float)cfg.num("postprocessing.bloom.threshold", 1.0);
    bloom.intensity = (float)cfg.num("postprocessing.bloom.intensity", 0.05);
    TonemappingPushConsts tone{};
    tone.exposure = (float)cfg.num("postprocessing.tonemap.exposure", 1.0);

    hipStream_t st;
    hip_check(hipStreamCreate(&st), "hipStreamCreate");
    float *image = nullptr, *scratch = nullptr;
    uint8_t* ldr = nullptr;
    hip_check(hipMalloc(&image, (size_t)W * H * 16), "hipMalloc");
    hip_check(hipMalloc(&scratch, (size_t)W * H * 32), "hipMalloc");
    hip_check(hipMalloc(&ldr, (size_t)W * H * 4), "hipMalloc");
    std::vector<uint8_t> host((size_t)W * H * 4);
    SaveManager saver(saves, save_times);
    Clock clock;
    const bool print_summary = cfg.num("headless.print_summary", 1.0) != 0.0;
    const uint64_t frames = (total + spp - 1) / spp;
    for (uint64_t frame = 0; frame < frames; ++frame) {
        if (print_summary && clock.frame_count() != 0) std::printf("%s\n", clock.summary().c_str());
        clock.mark_category("Ray Tracing");
        check(reina_rt_trace(scene, &pc, image, W, H, st), "reina_rt_trace");
        pc.sampleBatch++;   // after recording, as Reina::traceRays (src/Reina.cpp:431-432)
        hip_check(hipStreamSynchronize(st), "sync");
        clock.mark_category("Save");
        std::string name;
        if (saver.should_save(clock.sample_count(), clock.age(), name)) {
            check(reina_rt_postprocess(image, scratch, ldr, W, H, &bloom, &tone, st), "reina_rt_postprocess");
            hip_check(hipMemcpyAsync(host.data(), ldr, host.size(), hipMemcpyDeviceToHost, st), "copy");
            hip_check(hipStreamSynchronize(st), "sync");
            write_png(out_dir + "/" + name, host.data(), W, H);
            std::printf("saved %s/%s (%llu samples per pixel in the image)\n", out_dir.c_str(), name.c_str(),
                        (unsigned long long)(frame + 1) * spp);
        }
        clock.mark_frame(spp);
    }
    hip_check(hipStreamSynchronize(st), "sync");
    std::printf("rendered %llu samples per pixel at %ux%u in %.3f s (%llu frames)\n",
                (unsigned long long)frames * spp, W, H, clock.age(), (unsigned long long)pc.sampleBatch);
    (void)hipFree(image);
    (void)hipFree(scratch);
    (void)hipFree(ldr);
    (void)hipStreamDestroy(st);
    check(reina_rt_scene_destroy(scene), "reina_rt_scene_destroy");
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        return run(argc, argv);
    } catch (const std::exception& e) {   // same error behaviour as the reference (src/main.cpp:11-14)
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
}
