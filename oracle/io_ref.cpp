// io_ref.cpp — test infrastructure only: the reference's scene-file readers,
// built from the third-party sources the reference vendors
// (nectar_render/csrc/external/stb/stb_image.h 2.30 and
// nectar_render/csrc/external/tinyobjloader/tiny_obj_loader.h 2.x), compiled
// where they lie by `make ref` into oracle/_ref/libnr_ioref.so.
//
// The glue restates, for the checker:
//   load_image_uint8 / load_image_float  (csrc/material/include/material/texture.h:8-47)
//   parse_obj_file                       (csrc/hittable/include/hittable/mesh.h:39-141)
// Only tests load this library; the product's readers are
// nectarrender_amd/csrc/host/nr_io.cpp.
#define STB_IMAGE_IMPLEMENTATION
#include "stb_image.h"
#define TINYOBJLOADER_IMPLEMENTATION
#include "tiny_obj_loader.h"

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

struct R3 {
    float v[3];
};
R3 r3(float a, float b, float c) { return R3{{a, b, c}}; }
R3 operator+(R3 a, R3 b) { return r3(a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]); }
R3 operator-(R3 a, R3 b) { return r3(a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]); }
R3 operator*(R3 a, float t) { return r3(a.v[0] * t, a.v[1] * t, a.v[2] * t); }
float dotr(R3 a, R3 b) { return a.v[0] * b.v[0] + a.v[1] * b.v[1] + a.v[2] * b.v[2]; }
R3 crossr(R3 a, R3 b) {
    return r3(a.v[1] * b.v[2] - a.v[2] * b.v[1], a.v[2] * b.v[0] - a.v[0] * b.v[2], a.v[0] * b.v[1] - a.v[1] * b.v[0]);
}
float lenr(R3 a) {
    float s = a.v[0] * a.v[0] + a.v[1] * a.v[1];
    s += a.v[2] * a.v[2];
    return sqrtf(s);
}
R3 unit(R3 a) { return a * (1.0f / lenr(a)); }      // operator/(T, float) = (1/t) * v (vector.h:250-253)

struct Vert {
    R3 p, n, t;
    float u = 0.0f, w = 0.0f;
};

struct Key {
    int v, vt, vn;
    bool operator==(const Key& o) const { return v == o.v && vt == o.vt && vn == o.vn; }
};
struct KeyHash {
    size_t operator()(const Key& k) const {
        size_t h = std::hash<int>()(k.v);
        h ^= std::hash<int>()(k.vt) + 0x9e3779b9 + (h << 6) + (h >> 2);
        h ^= std::hash<int>()(k.vn) + 0x9e3779b9 + (h << 6) + (h >> 2);
        return h;
    }
};

}  // namespace

extern "C" {

void ioref_free(void* p) { std::free(p); }

// texture.h:8-27 (C >= 3 only: grey files make the reference read past its buffer)
uint8_t* ioref_image_u8(const char* path, int* C, int* H, int* W) {
    unsigned char* data = stbi_load(path, W, H, C, 0);
    if (!data) return nullptr;
    const size_t hw = (size_t)(*W) * (*H);
    uint8_t* chw = (uint8_t*)std::malloc(hw * (*C));
    for (size_t c = 0; c < (size_t)*C; ++c)
        for (size_t i = 0; i < hw; ++i) chw[c * hw + i] = data[i * 3 + c];
    stbi_image_free(data);
    return chw;
}

// stbi_load(path, ..., 0) itself: interleaved H x W x C samples as the library
// returns them (grey and grey + alpha files included), before the glue above
uint8_t* ioref_stbi_u8(const char* path, int* C, int* H, int* W) {
    unsigned char* data = stbi_load(path, W, H, C, 0);
    if (!data) return nullptr;
    const size_t n = (size_t)(*W) * (*H) * (*C);
    uint8_t* out = (uint8_t*)std::malloc(n);
    std::memcpy(out, data, n);
    stbi_image_free(data);
    return out;
}

// texture.h:29-47
float* ioref_image_f32(const char* path, int* C, int* H, int* W) {
    float* data = stbi_loadf(path, W, H, C, 0);
    if (!data) return nullptr;
    const size_t hw = (size_t)(*W) * (*H), n = (size_t)*C;
    float* chw = (float*)std::malloc(hw * n * sizeof(float));
    for (size_t c = 0; c < n; c++)
        for (size_t i = 0; i < hw; i++) chw[c * hw + i] = data[i * n + c];
    stbi_image_free(data);
    return chw;
}

// mesh.h:39-141: returns 0 on success; *verts (n x 11) and *tris (m x 3) are malloc'd
int ioref_obj(const char* path, int* nv, int* nt, float** verts, uint32_t** tris) {
    tinyobj::ObjReaderConfig config;
    config.triangulate = true;
    tinyobj::ObjReader reader;
    if (!reader.ParseFromFile(path, config)) return 1;
    const tinyobj::attrib_t& at = reader.GetAttrib();
    const bool has_normals = !at.normals.empty();
    std::vector<Vert> vs;
    std::vector<uint32_t> ts;
    std::unordered_map<Key, uint32_t, KeyHash> seen;
    for (const tinyobj::shape_t& sh : reader.GetShapes()) {
        size_t off = 0;
        for (size_t f = 0; f < sh.mesh.num_face_vertices.size(); f++) {
            for (size_t k = 0; k < 3; k++) {
                tinyobj::index_t ix = sh.mesh.indices[off + k];
                Key key{ix.vertex_index, ix.texcoord_index, ix.normal_index};
                auto it = seen.find(key);
                if (it != seen.end()) { ts.push_back(it->second); continue; }
                Vert x;
                x.p = r3(at.vertices[3 * ix.vertex_index], at.vertices[3 * ix.vertex_index + 1],
                         at.vertices[3 * ix.vertex_index + 2]);
                x.n = ix.normal_index >= 0 ? r3(at.normals[3 * ix.normal_index], at.normals[3 * ix.normal_index + 1],
                                                at.normals[3 * ix.normal_index + 2])
                                           : r3(0, 0, 0);
                if (ix.texcoord_index >= 0) {
                    x.u = at.texcoords[2 * ix.texcoord_index];
                    x.w = at.texcoords[2 * ix.texcoord_index + 1];
                }
                x.t = r3(0, 0, 0);
                uint32_t id = (uint32_t)vs.size();
                vs.push_back(x);
                seen[key] = id;
                ts.push_back(id);
            }
            off += 3;
        }
    }
    const size_t m = ts.size() / 3;
    if (!has_normals) {
        for (size_t f = 0; f < m; f++) {
            Vert &a = vs[ts[3 * f]], &b = vs[ts[3 * f + 1]], &c = vs[ts[3 * f + 2]];
            R3 fn = unit(crossr(b.p - a.p, c.p - a.p));
            a.n = a.n + fn;
            b.n = b.n + fn;
            c.n = c.n + fn;
        }
        for (Vert& v : vs) v.n = lenr(v.n) < 1e-8f ? r3(0, 1, 0) : unit(v.n);
    }
    for (size_t f = 0; f < m; f++) {
        Vert &a = vs[ts[3 * f]], &b = vs[ts[3 * f + 1]], &c = vs[ts[3 * f + 2]];
        R3 e1 = b.p - a.p, e2 = c.p - a.p;
        float d1u = b.u - a.u, d1v = b.w - a.w, d2u = c.u - a.u, d2v = c.w - a.w;
        float den = d1u * d2v - d2u * d1v;
        R3 t = fabsf(den) < 1e-8f ? r3(1, 0, 0) : (e1 * d2v - e2 * d1v) * (1.0f / den);
        a.t = a.t + t;
        b.t = b.t + t;
        c.t = c.t + t;
    }
    for (Vert& v : vs) {
        R3 t = v.t - v.n * dotr(v.t, v.n);
        v.t = lenr(t) < 1e-8f ? r3(1, 0, 0) : unit(t);
    }
    *nv = (int)vs.size();
    *nt = (int)m;
    *verts = (float*)std::malloc(vs.size() * 11 * sizeof(float) + 4);
    *tris = (uint32_t*)std::malloc(ts.size() * sizeof(uint32_t) + 4);
    for (size_t i = 0; i < vs.size(); i++) {
        float* o = *verts + 11 * i;
        const Vert& v = vs[i];
        for (int k = 0; k < 3; k++) { o[k] = v.p.v[k]; o[3 + k] = v.n.v[k]; o[6 + k] = v.t.v[k]; }
        o[9] = v.u;
        o[10] = v.w;
    }
    std::memcpy(*tris, ts.data(), ts.size() * sizeof(uint32_t));
    return 0;
}

}  // extern "C"
