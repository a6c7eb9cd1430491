// bindings.cpp — the `_pathtracer` Python module on top of libnectar_hip.
//
// Mirrors the reference's pybind11 module (csrc/bindings/bindings.cpp:16-30):
// the same submodules (core, cuda, material, hittable, light, data, scene,
// engine, host), class names, factory names, argument names, defaults and
// error types, so nectar_render/python/* and user scene scripts import it
// unchanged.  Scene objects are plain host descriptions here; they are lowered
// into the C ABI (include/nectar_hip.h) when a scene is set on an engine, and
// every device operation goes through that ABI.
//
// Ownership is more lenient than the reference's move-only objects
// (material.h:430-436, hittable.h:92-98): materials, textures and hittables
// are shared, so every program the reference accepts is accepted and an object
// may also be reused.  C ABI failures and loader errors raise RuntimeError;
// nothing aborts the process.
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/operators.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <hip/hip_runtime_api.h>

#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <optional>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/nectar_denoise.h"
#include "../../include/nectar_hip.h"
#include "host/nr_io.h"

namespace py = pybind11;
using rvp = py::return_value_policy;

namespace {

constexpr float FMIN = 1e-8f;

void check(int rc) {
    if (rc != NR_OK) throw std::runtime_error(std::string("nectar_hip: ") + nr_last_error());
}

// ---------------------------------------------------------------------------
// vectors (csrc/core/include/core/vector.h): FMIN-biased element division,
// scalar division as multiplication by the reciprocal
// ---------------------------------------------------------------------------
template <int N, int Tag>
struct Vec {
    float e[N];
    Vec() { for (int i = 0; i < N; i++) e[i] = 0.0f; }
    explicit Vec(float s) { for (int i = 0; i < N; i++) e[i] = s; }
    float operator[](int i) const { return e[i]; }
    float& operator[](int i) { return e[i]; }
    float length_squared() const {
        float o = e[0] * e[0] + e[1] * e[1];
        if (N >= 3) o += e[2] * e[2];
        return o;
    }
    float length() const { return std::sqrt(length_squared()); }
};
using Vector2 = Vec<2, 0>;
using Vector3 = Vec<3, 1>;
using Color = Vec<3, 2>;

Vector2 v2(float x, float y) { Vector2 v; v.e[0] = x; v.e[1] = y; return v; }
Vector3 v3(float x, float y, float z) { Vector3 v; v.e[0] = x; v.e[1] = y; v.e[2] = z; return v; }
Color c3(float r, float g, float b) { Color v; v.e[0] = r; v.e[1] = g; v.e[2] = b; return v; }

template <class T> T vadd(const T& a, const T& b) { T r = a; for (int i = 0; i < (int)(sizeof(a.e) / 4); i++) r.e[i] += b.e[i]; return r; }
template <class T> T vsub(const T& a, const T& b) { T r = a; for (int i = 0; i < (int)(sizeof(a.e) / 4); i++) r.e[i] -= b.e[i]; return r; }
template <class T> T vmul(const T& a, const T& b) { T r = a; for (int i = 0; i < (int)(sizeof(a.e) / 4); i++) r.e[i] *= b.e[i]; return r; }
template <class T> T vdiv(const T& a, const T& b) { T r = a; for (int i = 0; i < (int)(sizeof(a.e) / 4); i++) r.e[i] /= (b.e[i] + FMIN); return r; }
template <class T> T sadd(const T& a, float t) { T r = a; for (auto& x : r.e) x += t; return r; }
template <class T> T ssub(const T& a, float t) { T r = a; for (auto& x : r.e) x -= t; return r; }
template <class T> T smul(const T& a, float t) { T r = a; for (auto& x : r.e) x *= t; return r; }
template <class T> T sdiv(const T& a, float t) { return smul(a, 1.0f / t); }                // operator/(T, float)
template <class T> T sdiv_assign(const T& a, float t) { return smul(a, 1.0f / (t + FMIN)); } // operator/=(float)

float dot3(const Vector3& u, const Vector3& v) { return u.e[0] * v.e[0] + u.e[1] * v.e[1] + u.e[2] * v.e[2]; }
Vector3 cross3(const Vector3& u, const Vector3& v) {
    return v3(u.e[1] * v.e[2] - u.e[2] * v.e[1], u.e[2] * v.e[0] - u.e[0] * v.e[2], u.e[0] * v.e[1] - u.e[1] * v.e[0]);
}
Vector3 neg3(const Vector3& v) { return v3(-v.e[0], -v.e[1], -v.e[2]); }
Vector3 reflect3(const Vector3& v, const Vector3& n) { return vsub(v, smul(n, 2.0f * dot3(v, n))); }   // vector.h:469-474
Vector3 refract3(const Vector3& uv, const Vector3& n, float ior) {                                   // vector.h:476-485
    float cos_theta = std::fmin(dot3(neg3(uv), n), 1.0f);
    Vector3 perp = smul(vadd(uv, smul(n, cos_theta)), ior);
    Vector3 par = smul(n, -std::sqrt(std::fabs((float)(1.0 - perp.length_squared()))));
    return vadd(perp, par);
}

// ---------------------------------------------------------------------------
// Matrix3 / rotations (csrc/core/include/core/matrix.h), Transform (transform.h)
// ---------------------------------------------------------------------------
struct Matrix3 {
    float m[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    Matrix3 operator*(const Matrix3& b) const {
        Matrix3 r;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                r.m[i][j] = 0;
                for (int k = 0; k < 3; k++) r.m[i][j] += m[i][k] * b.m[k][j];
            }
        return r;
    }
    Vector3 operator*(const Vector3& v) const {
        return v3(m[0][0] * v.e[0] + m[0][1] * v.e[1] + m[0][2] * v.e[2],
                  m[1][0] * v.e[0] + m[1][1] * v.e[1] + m[1][2] * v.e[2],
                  m[2][0] * v.e[0] + m[2][1] * v.e[1] + m[2][2] * v.e[2]);
    }
    Matrix3 T() const {
        Matrix3 r;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) r.m[i][j] = m[j][i];
        return r;
    }
};
float deg2rad(float d) { return (float)((double)d * 0.01745329); }                          // utils.h:12-14
Matrix3 rot_x(float a) { float c = cosf(a), s = sinf(a); Matrix3 r; r.m[1][1] = c; r.m[1][2] = -s; r.m[2][1] = s; r.m[2][2] = c; return r; }
Matrix3 rot_y(float a) { float c = cosf(a), s = sinf(a); Matrix3 r; r.m[0][0] = c; r.m[0][2] = s; r.m[2][0] = -s; r.m[2][2] = c; return r; }
Matrix3 rot_z(float a) { float c = cosf(a), s = sinf(a); Matrix3 r; r.m[0][0] = c; r.m[0][1] = -s; r.m[1][0] = s; r.m[1][1] = c; return r; }
Matrix3 rotation_from_euler(const Vector3& r) { return rot_z(r.e[2]) * rot_y(r.e[1]) * rot_x(r.e[0]); }

struct Transform {
    Vector3 position{0.0f}, rotation_deg{0.0f}, scale{1.0f};
    Matrix3 R;
    Transform() = default;
    Transform(const Vector3& p, const Vector3& r, const Vector3& s) : position(p), rotation_deg(r), scale(s) { rebuild(); }
    void rebuild() { R = rotation_from_euler(v3(deg2rad(rotation_deg.e[0]), deg2rad(rotation_deg.e[1]), deg2rad(rotation_deg.e[2]))); }
};

// ---------------------------------------------------------------------------
// textures, materials (csrc/material/include/material/{texture.h, material.h})
// ---------------------------------------------------------------------------
enum class TextureType { CONSTANT, IMAGE };
struct Texture {
    TextureType type = TextureType::CONSTANT;
    std::string filepath;
    Color constant_color{0.0f};
    size_t C = 0, H = 0, W = 0;
    float scale = 1.0f;
    std::shared_ptr<std::vector<uint8_t>> chw;
    static std::shared_ptr<Texture> from_color(const Color& c) {
        auto t = std::make_shared<Texture>();
        t->constant_color = c;
        t->C = 3;
        return t;
    }
    static std::shared_ptr<Texture> from_image(const std::string& path, float scale) {
        nrio::Image8 im = nrio::load_image_u8(path);
        auto t = std::make_shared<Texture>();
        t->type = TextureType::IMAGE;
        t->filepath = path;
        t->scale = scale;
        t->C = (size_t)im.C; t->H = (size_t)im.H; t->W = (size_t)im.W;
        t->chw = std::make_shared<std::vector<uint8_t>>(std::move(im.chw));
        return t;
    }
};
using TexP = std::shared_ptr<Texture>;

enum class MaterialType { Lambertian = NR_MAT_LAMBERTIAN, PBR = NR_MAT_PBR, Dielectric = NR_MAT_DIELECTRIC,
                          Emissive = NR_MAT_EMISSIVE, Isotropic = NR_MAT_ISOTROPIC };
struct Material {
    MaterialType type = MaterialType::Lambertian;
    std::vector<TexP> tex;          // tracked textures (material.h:550-556)
    Color tint{1.0f};
    float ior = 1.5f;
    float brightness = 35.0f;
    static Material lambertian(TexP t) { Material m; m.type = MaterialType::Lambertian; m.tex = {std::move(t)}; return m; }
    static Material pbr(TexP a, TexP r, TexP mt, TexP e, TexP n) {
        Material m; m.type = MaterialType::PBR; m.tex = {a, r, mt, e, n}; return m;
    }
    static Material dielectric(const Color& tint, float ior) {
        Material m; m.type = MaterialType::Dielectric; m.tint = tint; m.ior = ior; return m;
    }
    static Material emissive(TexP t, float b) { Material m; m.type = MaterialType::Emissive; m.tex = {std::move(t)}; m.brightness = b; return m; }
    static Material isotropic(TexP t) { Material m; m.type = MaterialType::Isotropic; m.tex = {std::move(t)}; return m; }
};
Material default_material() { return Material::lambertian(Texture::from_color(c3(1.0f, 0.0f, 1.0f))); }  // material.h:650-652

// ---------------------------------------------------------------------------
// hittables (csrc/hittable/include/hittable/hittable.h:148-277)
// ---------------------------------------------------------------------------
struct MeshGeom { nrio::MeshData data; std::string path; };
struct Hittable {
    int type = NR_HIT_NULL;
    int bound_type = NR_HIT_NULL;
    Transform xform;
    float radius = 0.0f, density = 0.0f;
    std::shared_ptr<Material> mat;
    std::shared_ptr<MeshGeom> mesh;
};
using HitP = std::shared_ptr<Hittable>;

HitP make_obj(int type, const Vector3& p, const Vector3& r, const Vector3& s, const Material& m) {
    auto h = std::make_shared<Hittable>();
    h->type = type;
    h->xform = Transform(p, r, s);
    h->mat = std::make_shared<Material>(m);
    return h;
}
HitP wrap(const Hittable& bound, int type, float density, const Material& m) {
    if (bound.type == NR_HIT_CONSTANT_MEDIUM || bound.type == NR_HIT_OBJECT_LIGHT)
        throw std::runtime_error("a constant medium or object light cannot wrap another wrapper object");
    if (bound.type == NR_HIT_NULL) throw std::runtime_error("bound object has no geometry");
    auto h = std::make_shared<Hittable>();
    h->type = type;
    h->bound_type = bound.type;
    h->xform = bound.xform;
    h->radius = bound.radius;
    h->density = density;
    h->mesh = bound.mesh;
    h->mat = std::make_shared<Material>(m);
    return h;
}

struct HitRecord {
    int object_id = -1, material_id = 0, prim = -1;
    Vector3 p, n, tangent;
    Vector2 uv;
    float t = 0.0f;
    bool front_face = false;
};

// ---------------------------------------------------------------------------
// skylight (csrc/light/include/skylight.h)
// ---------------------------------------------------------------------------
enum class SkylightType { Null = NR_SKY_NULL, Simple = NR_SKY_SIMPLE, HDRI = NR_SKY_HDRI };
struct SimpleSkylightConfig { Color start = c3(1.0f, 1.0f, 1.0f), end = c3(0.5f, 0.7f, 1.0f); };
struct HDRISkylightConfig {
    float rotation = 0.0f, intensity = 1.0f;
    int C = 0, H = 0, W = 0;
    std::shared_ptr<std::vector<float>> data;
};
struct Skylight {
    SkylightType type = SkylightType::Null;
    SimpleSkylightConfig cfg_simple;
    HDRISkylightConfig cfg_hdri;
    void load_hdri_file(const std::string& path) {
        if (type != SkylightType::HDRI)
            throw std::runtime_error("Skylight::update_hdri is only valid on skylights with type SkylightType::HDRI.");
        nrio::ImageF im = nrio::load_image_f32(path);
        cfg_hdri.C = im.C; cfg_hdri.H = im.H; cfg_hdri.W = im.W;
        cfg_hdri.data = std::make_shared<std::vector<float>>(std::move(im.chw));
    }
};

// ---------------------------------------------------------------------------
// scene lowering into nr_scene_desc
// ---------------------------------------------------------------------------
struct Scene {
    std::vector<HitP> objects;      // hittables then lights (scene.h:31-39)
    size_t n_hittables = 0;
    std::shared_ptr<Skylight> sky;
};

struct Lowered {
    std::vector<nr_object_desc> objs;
    std::vector<nr_material_desc> mats;
    std::vector<nr_image_desc> imgs;
    std::vector<nr_mesh_desc> meshes;
    std::vector<TexP> keep_tex;
    std::vector<std::shared_ptr<MeshGeom>> keep_mesh;
    std::shared_ptr<Skylight> keep_sky;
    nr_scene_desc desc{};
};

void lower_texture(Lowered& L, const TexP& t, nr_texture_desc& d) {
    d.image = -1;
    d.scale = 1.0f;
    for (int k = 0; k < 3; k++) d.color[k] = t ? t->constant_color.e[k] : 0.0f;
    if (!t || t->type != TextureType::IMAGE) return;
    for (size_t i = 0; i < L.keep_tex.size(); i++)
        if (L.keep_tex[i] == t) { d.image = (int)i; d.scale = t->scale; return; }
    L.keep_tex.push_back(t);
    nr_image_desc im{};
    im.C = (int32_t)t->C; im.H = (int32_t)t->H; im.W = (int32_t)t->W;
    im.data = t->chw->data();
    L.imgs.push_back(im);
    d.image = (int)L.imgs.size() - 1;
    d.scale = t->scale;
}

void lower_material(Lowered& L, const Material& m) {
    nr_material_desc d{};
    d.type = (int32_t)m.type;
    for (int k = 0; k < 5; k++) {
        TexP t = k < (int)m.tex.size() ? m.tex[k] : nullptr;
        lower_texture(L, t, d.tex[k]);
    }
    for (int k = 0; k < 3; k++) d.tint[k] = m.tint.e[k];
    d.ior = m.ior;
    d.brightness = m.brightness;
    L.mats.push_back(d);
}

// One registered object: its own material (material registry, registry.h:29-43),
// textures and mesh geometry not yet lowered are appended to L's tables.
void lower_object(Lowered& L, const HitP& h) {
    nr_object_desc o{};
    o.type = h->type;
    o.bound_type = h->bound_type;
    lower_material(L, h->mat ? *h->mat : default_material());
    o.material = (int32_t)L.mats.size() - 1;
    o.mesh = -1;
    if (h->mesh) {
        int idx = -1;
        for (size_t i = 0; i < L.keep_mesh.size(); i++)
            if (L.keep_mesh[i] == h->mesh) idx = (int)i;
        if (idx < 0) {
            L.keep_mesh.push_back(h->mesh);
            nr_mesh_desc md{};
            md.n_vertices = (int32_t)h->mesh->data.n_vertices();
            md.n_triangles = (int32_t)h->mesh->data.n_triangles();
            md.vertices = h->mesh->data.vertices.data();
            md.triangles = h->mesh->data.triangles.data();
            L.meshes.push_back(md);
            idx = (int)L.meshes.size() - 1;
        }
        o.mesh = idx;
    }
    for (int k = 0; k < 3; k++) {
        o.position[k] = h->xform.position.e[k];
        o.rotation[k] = h->xform.rotation_deg.e[k];
        o.scale[k] = h->xform.scale.e[k];
    }
    o.radius = h->radius;
    o.density = h->density;
    L.objs.push_back(o);
}

nr_sky_desc lower_sky(Lowered& L, const std::shared_ptr<Skylight>& sky) {
    nr_sky_desc sd{};
    sd.intensity = 1.0f;
    if (!sky) return sd;
    L.keep_sky = sky;
    sd.type = (int32_t)sky->type;
    for (int k = 0; k < 3; k++) { sd.start[k] = sky->cfg_simple.start.e[k]; sd.end[k] = sky->cfg_simple.end.e[k]; }
    sd.rotation = sky->cfg_hdri.rotation;
    sd.intensity = sky->cfg_hdri.intensity;
    if (sky->type == SkylightType::HDRI && sky->cfg_hdri.data) {
        sd.C = sky->cfg_hdri.C; sd.H = sky->cfg_hdri.H; sd.W = sky->cfg_hdri.W;
        sd.data = sky->cfg_hdri.data->data();
    }
    return sd;
}

std::unique_ptr<Lowered> lower_scene(const Scene& s) {
    auto L = std::make_unique<Lowered>();
    lower_material(*L, default_material());                   // slot 0 (material/registry.h:23-26)
    for (const HitP& h : s.objects) lower_object(*L, h);
    nr_sky_desc sd = lower_sky(*L, s.sky);
    L->desc.n_objects = (int32_t)L->objs.size();
    L->desc.objects = L->objs.data();
    L->desc.n_materials = (int32_t)L->mats.size();
    L->desc.materials = L->mats.data();
    L->desc.n_images = (int32_t)L->imgs.size();
    L->desc.images = L->imgs.data();
    L->desc.n_meshes = (int32_t)L->meshes.size();
    L->desc.meshes = L->meshes.data();
    L->desc.sky = sd;
    return L;
}

// ---------------------------------------------------------------------------
// camera (csrc/engine/include/engine/camera.h:68-160, 280-360)
// ---------------------------------------------------------------------------
struct CameraParams {
    Vector2 resolution;
    Vector3 position, rotation;
    uint32_t samples_per_pixel = 0;
    float focal_length = 0, focus_distance = 0, aperture = 0, sensor_width = 0, shutter_speed = 0;
    Matrix3 R;
    Vector3 uvw;
    float shutter_time = 0, aspect_ratio = 0, defocus_angle = 0, defocus_radius = 0;
    float movement_speed = 5.0f, look_sensitivity = 10.0f;

    void build() {                                                           // camera.h:124-136
        R = rotation_from_euler(v3(deg2rad(rotation.e[0]), deg2rad(rotation.e[1]), deg2rad(rotation.e[2])));
        shutter_time = 1.0f / (shutter_speed + FMIN);
        float aperture_radius = aperture / 2.0f;
        defocus_angle = 2.0f * atanf(aperture_radius / (focus_distance + FMIN));
        defocus_radius = focus_distance * tanf(defocus_angle / 2.0f);
        aspect_ratio = resolution.e[1] / resolution.e[0];
        uvw = v3(sensor_width, -sensor_width * aspect_ratio, 0.0f);
    }
    void update(const CameraParams& o) {                                     // camera.h:138-157
        Vector3 pos = smul(o.position, movement_speed);
        Vector3 rot = smul(o.rotation, look_sensitivity);
        rotation.e[1] -= rot.e[1];
        rotation.e[0] -= rot.e[0];
        rotation.e[0] = std::fmax(-89.0f, std::fmin(89.0f, rotation.e[0]));
        Matrix3 Ryx = rot_y(deg2rad(rotation.e[1])) * rot_x(deg2rad(rotation.e[0]));
        position = vadd(position, Ryx * pos);
        focal_length = o.focal_length;
        focus_distance = o.focus_distance;
        aperture = o.aperture;
        sensor_width = o.sensor_width;
        shutter_speed = o.shutter_speed;
        build();
    }
    nr_camera_desc desc(uint32_t pattern_spp) const {
        nr_camera_desc d{};
        d.resolution[0] = resolution.e[0];
        d.resolution[1] = resolution.e[1];
        for (int k = 0; k < 3; k++) { d.position[k] = position.e[k]; d.rotation[k] = rotation.e[k]; }
        d.samples_per_pixel = pattern_spp;
        d.focal_length = focal_length;
        d.focus_distance = focus_distance;
        d.aperture = aperture;
        d.sensor_width = sensor_width;
        d.shutter_speed = shutter_speed;
        return d;
    }
};

struct Camera {
    CameraParams p;
    std::function<void(CameraParams)> on_updated = [](CameraParams) {};
    explicit Camera(const CameraParams& params) : p(params) {}
    std::optional<Vector2> project_to_screen(const Vector3& w) const {       // camera.h:337-354
        Vector3 local = p.R.T() * vsub(w, p.position);
        if (local.e[2] >= 0.0f) return std::nullopt;
        float scale = -p.focal_length / local.e[2];
        float lu = local.e[0] * scale, lv = local.e[1] * scale;
        float rx = p.resolution.e[0], ry = p.resolution.e[1];
        float px = lu * (rx / p.uvw.e[0]) + (rx - 1.0f) * 0.5f;
        float py = lv * (ry / p.uvw.e[1]) + (ry - 1.0f) * 0.5f;
        return v2(px / rx, py / ry);
    }
};

// ---------------------------------------------------------------------------
// device context shared by an engine and the layers/streams it hands out
// ---------------------------------------------------------------------------
struct Ctx {
    nr_ctx* h = nullptr;
    std::mutex mu;            // one stream per context: calls are serialised
    // Readbacks (nr_readback) run beside renders; scene and camera changes exclude them.
    std::shared_mutex layout_mu;
    explicit Ctx(int device) {
        h = nr_create(device);
        if (!h) throw std::runtime_error(std::string("nectar_hip: ") + nr_last_error());
    }
    ~Ctx() { if (h) nr_destroy(h); }
};
using CtxP = std::shared_ptr<Ctx>;

// ---------------------------------------------------------------------------
// data (csrc/data/include/data/{data_object.h, render_layers.h, stream.h})
// ---------------------------------------------------------------------------
enum class LayerType { BEAUTY = NR_LAYER_BEAUTY, DIFFUSE = NR_LAYER_DIFFUSE, SPECULAR = NR_LAYER_SPECULAR,
                       OBJECT_NORMAL = NR_LAYER_OBJECT_NORMAL, WORLD_NORMAL = NR_LAYER_WORLD_NORMAL,
                       SHADOW = NR_LAYER_SHADOW, DEPTH = NR_LAYER_DEPTH, EMISSION = NR_LAYER_EMISSION,
                       OBJECT_ID = NR_LAYER_OBJECT_ID };
constexpr size_t kLayerChannels[9] = {3, 3, 3, 3, 3, 1, 1, 3, 3};              // render_layers.h:178-188

struct DeviceBuf {
    float* p = nullptr;
    explicit DeviceBuf(size_t n) {
        if (hipMalloc((void**)&p, n * sizeof(float)) != hipSuccess) throw std::runtime_error("hipMalloc failed");
        if (hipMemset(p, 0, n * sizeof(float)) != hipSuccess) throw std::runtime_error("hipMemset failed");
    }
    ~DeviceBuf() { if (p) (void)hipFree(p); }
};

struct DataObject {
    size_t C = 0, H = 0, W = 0;
    bool enabled = false;
    CtxP ctx;                  // engine-owned layer: (ctx, view, layer)
    int view = 0, layer = 0;
    std::shared_ptr<DeviceBuf> own;   // standalone RenderLayers buffer

    size_t n_pixels() const { return H * W; }
    size_t n_elements() const { return C * n_pixels(); }
    size_t n_bytes() const { return n_elements() * sizeof(float); }
    std::vector<size_t> shape() const { return {C, H, W}; }
    uintptr_t device_ptr() const {
        if (!enabled) return 0;
        if (own) return reinterpret_cast<uintptr_t>(own->p);
        return reinterpret_cast<uintptr_t>(nr_aov_device_ptr(ctx->h, view, layer));
    }
    std::vector<float> read_float() const {
        if (!enabled) throw std::runtime_error("layer is not enabled");
        std::vector<float> out(n_elements());
        if (own) {
            if (hipDeviceSynchronize() != hipSuccess ||
                hipMemcpy(out.data(), own->p, n_bytes(), hipMemcpyDeviceToHost) != hipSuccess)
                throw std::runtime_error("layer readback failed");
        } else {
            std::lock_guard<std::mutex> g(ctx->mu);
            check(nr_sync(ctx->h));
            check(nr_aov_read(ctx->h, view, layer, out.data(), out.size()));
        }
        return out;
    }
    // data_to_image_kernel (data.cu:77-92): NaN -> 0, clamp to [0,1], x255, truncate
    py::array_t<uint8_t> numpy() const {
        std::vector<float> f = read_float();
        py::array_t<uint8_t> a({C, H, W});
        uint8_t* o = a.mutable_data();
        for (size_t i = 0; i < f.size(); i++) {
            float v = f[i];
            if (std::isnan(v)) v = 0.0f;
            v = std::fmax(0.0f, std::fmin(1.0f, v)) * 255.0f;
            o[i] = (uint8_t)v;
        }
        return a;
    }
    py::array_t<float> numpy_float() const {
        std::vector<float> f = read_float();
        py::array_t<float> a({C, H, W});
        std::memcpy(a.mutable_data(), f.data(), f.size() * sizeof(float));
        return a;
    }
};

struct RenderLayersConfig {
    bool beauty = true, diffuse = true, world_normal = true, object_normal = false, specular = false,
         shadow = false, depth = false, emission = false, object_id = false;
    bool enabled(int L) const {
        switch (L) {
            case NR_LAYER_BEAUTY: return beauty;
            case NR_LAYER_DIFFUSE: return diffuse;
            case NR_LAYER_SPECULAR: return specular;
            case NR_LAYER_OBJECT_NORMAL: return object_normal;
            case NR_LAYER_WORLD_NORMAL: return world_normal;
            case NR_LAYER_SHADOW: return shadow;
            case NR_LAYER_DEPTH: return depth;
            case NR_LAYER_EMISSION: return emission;
            default: return object_id;
        }
    }
};

struct RenderLayers {
    size_t H = 0, W = 0;
    std::array<DataObject, 9> layers;
    // standalone layers (render_layers.h:178-188)
    RenderLayers(size_t h, size_t w, const RenderLayersConfig& cfg) : H(h), W(w) {
        for (int L = 0; L < 9; L++) {
            DataObject& d = layers[L];
            if (!cfg.enabled(L)) continue;
            d.C = kLayerChannels[L]; d.H = h; d.W = w; d.enabled = true;
            d.own = std::make_shared<DeviceBuf>(d.n_elements());
        }
    }
    // an engine view's running-mean AOVs
    RenderLayers(const CtxP& ctx, int view, size_t h, size_t w) : H(h), W(w) {
        for (int L = 0; L < 9; L++) {
            DataObject& d = layers[L];
            if (!nr_aov_device_ptr(ctx->h, view, L)) continue;
            d.C = kLayerChannels[L]; d.H = h; d.W = w; d.enabled = true;
            d.ctx = ctx; d.view = view; d.layer = L;
        }
    }
    DataObject& get(LayerType t) { return layers[(int)t]; }
};

enum class StreamState { ACTIVE, INACTIVE, FROZEN };
enum class TonemapMethod { REINHARD, REINHARD_EXTENDED, ACES };
enum class DenoiseFilterQuality { DEFAULT, HIGH, BALANCED, FAST };
struct StreamConfig {                                                        // stream.h:13-30
    bool apply_denoising = true, denoise_clean_auxiliaries = true;
    DenoiseFilterQuality denoise_quality = DenoiseFilterQuality::FAST;
    float denoise_input_scale = 1.0f;
    bool linear_to_gamma = true, apply_white_balance = true;
    float wb_temperature = 6800.0f, wb_tint = 1.0f;
    bool apply_tonemapping = true;
    TonemapMethod tm_method = TonemapMethod::REINHARD;
    float tm_white_point = 1.0f, tm_alpha = 1.0f;
};

// The learned U-Net in OIDN's slot (denoise.h:25-83), loaded with the shipped
// weights on the first denoised readback.
struct Denoiser {
    nr_unet* h = nullptr;
    explicit Denoiser(int device) {
        h = nr_unet_create(device, 0u);
        if (!h) throw std::runtime_error(std::string("nr_unet_create: ") + nr_last_error());
        py::object path = py::module_::import("nectarrender_amd.denoise").attr("WEIGHTS");
        py::array_t<float, py::array::c_style | py::array::forcecast> w =
            py::module_::import("numpy").attr("load")(path, py::arg("allow_pickle") = false);
        if (nr_unet_set_params(h, w.data(), (size_t)w.size()) != NR_OK) {
            nr_unet_destroy(h);
            throw std::runtime_error(std::string("U-Net weights: ") + nr_last_error());
        }
    }
    ~Denoiser() { if (h) nr_unet_destroy(h); }
};

// TransferStream (stream.h:47-279): the linked layer -> [U-Net denoise] -> white
// balance / tonemap / gamma -> [selection overlay] -> a host uint8 CHW buffer,
// on the library's readback stream (nr_readback), concurrently with rendering.
struct TransferStream {
    CtxP ctx;
    int view = 0;
    int linked = NR_LAYER_BEAUTY;
    size_t C = 3, H = 0, W = 0;
    StreamConfig cfg;
    std::atomic<StreamState> state{StreamState::INACTIVE};
    std::vector<uint8_t> buffer;
    std::mutex cfg_mu;
    std::mutex readback_mu;                      // readback_mutex (stream.h:155)
    std::shared_ptr<Denoiser> denoiser;
    std::atomic<const uint8_t*> overlay_mask{nullptr};   // device H x W mask (SceneInterface)
    Color overlay_color = c3(1.0f, 1.0f, 1.0f);

    uintptr_t buffer_ptr() { return reinterpret_cast<uintptr_t>(buffer.data()); }
    StreamConfig config() { std::lock_guard<std::mutex> g(cfg_mu); return cfg; }
    static nr_post_desc post_of(const StreamConfig& c, int layer) {
        nr_post_desc d{};
        d.linear_to_gamma = c.linear_to_gamma;
        d.apply_white_balance = c.apply_white_balance;
        d.wb_temperature = c.wb_temperature;
        d.wb_tint = c.wb_tint;
        d.apply_tonemapping = c.apply_tonemapping;
        d.tm_method = (int32_t)c.tm_method;
        d.tm_white_point = c.tm_white_point;
        d.tm_alpha = c.tm_alpha;
        d.fix_reinhard = 0;
        d.layer = layer;
        return d;
    }
    // with the GIL held: create the denoiser the configuration asks for
    void prepare() {
        if (config().apply_denoising && !denoiser) denoiser = std::make_shared<Denoiser>(0);
    }
    uintptr_t readback() {
        std::lock_guard<std::mutex> rb(readback_mu);
        if (state.load() == StreamState::FROZEN) return buffer_ptr();
        if (!ctx) throw std::runtime_error("TransferStream is not linked to an engine");
        const StreamConfig c = config();
        nr_readback_desc d{};
        d.post = post_of(c, linked);
        d.denoiser = (c.apply_denoising && denoiser) ? denoiser->h : nullptr;
        d.d_overlay = overlay_mask.load();
        for (int k = 0; k < 3; k++) d.overlay_color[k] = overlay_color.e[k];
        buffer.resize(C * H * W);
        std::shared_lock<std::shared_mutex> g(ctx->layout_mu);
        check(nr_readback(ctx->h, view, &d, buffer.data()));
        return buffer_ptr();
    }
    py::array_t<uint8_t> readback_numpy() {
        prepare();
        {
            py::gil_scoped_release r;
            readback();
        }
        py::array_t<uint8_t> a({C, H, W});
        std::memcpy(a.mutable_data(), buffer.data(), buffer.size());
        return a;
    }
    py::array_t<float> readback_float() {
        if (!ctx) throw std::runtime_error("TransferStream is not linked to an engine");
        const nr_post_desc d = post_of(config(), linked);
        py::array_t<float> a({C, H, W});
        std::lock_guard<std::mutex> g(ctx->mu);
        check(nr_post(ctx->h, view, &d, nullptr, a.mutable_data()));
        return a;
    }
};

// ---------------------------------------------------------------------------
// engine (csrc/engine/include/engine/engine.h, csrc/engine/src/engine.cpp)
// ---------------------------------------------------------------------------
enum class EngineState { IDLE, RENDERING };
enum class EngineType { PATHTRACER, VIEWPORT };
enum class RenderMode { FULL, INTERACTIVE };
enum class ReturnState { FINISHED, STOPPED, RESTARTED, WAITING };

struct EnginePollResponse {
    bool should_update_camera = false;
    CameraParams camera_params;
};

struct EngineRequests {                                                      // requests.h
    std::atomic<bool> start_{false}, stop_{false}, restart_{false}, shutdown_{false};
    void start() { start_.store(true); }
    void stop() { stop_.store(true); }
    void restart() { restart_.store(true); }
    void shutdown() { shutdown_.store(true); }
    bool start_pending() { return start_.exchange(false); }
    bool stop_pending() { return stop_.exchange(false); }
    bool restart_pending() { return restart_.exchange(false); }
    bool shutdown_pending() { return shutdown_.exchange(false); }
};

template <class F, class... A>
void call_with_gil(const F& f, A&&... a) {
    py::gil_scoped_acquire g;
    f(std::forward<A>(a)...);
}

// SceneInterface (csrc/scene/include/interface.h): picking, the selection
// outline overlaid on the stream, the outliner, and live edits (transforms,
// added objects, the sky).
struct SceneNode {                                                           // outliner.h:7-17
    size_t object_id = 0, parent_id = 0, material_id = 0;
    bool is_light = false;
    std::string name, type_name;
};
struct SceneOutline { std::vector<SceneNode> nodes; };

const char* hittable_type_name(int t) {                                      // hittable.h:17-28
    switch (t) {
        case NR_HIT_QUAD: return "Quad";
        case NR_HIT_SPHERE: return "Sphere";
        case NR_HIT_CUBE: return "Cube";
        case NR_HIT_MESH: return "Mesh";
        case NR_HIT_CONSTANT_MEDIUM: return "ConstantMedium";
        case NR_HIT_OBJECT_LIGHT: return "ObjectLight";
        default: return "Null";
    }
}

struct RenderEngine;
struct SceneInterface {
    RenderEngine* eng = nullptr;
    std::mutex mu;                        // interface_mutex
    HitRecord rec;
    std::atomic<bool> enabled{false}, teardown_pending{false};
    Color mask_color = c3(0.98f, 0.75f, 0.17f);                              // SelectionMaskConfig, interface.h:27-30
    int outline_radius = 3;

    void update();
    void query_scene(float u, float v);
    void select_scene_node(const SceneNode& n) {
        std::lock_guard<std::mutex> g(mu);
        rec.object_id = (int)n.object_id;
        rec.material_id = (int)n.material_id;
        enabled = true;
    }
    void enable() { enabled = true; }
    void disable() { teardown_pending = true; }
    bool is_enabled() const { return enabled.load(); }
    SceneOutline outline();
    HitP picked();
    Transform get_transform() { std::lock_guard<std::mutex> g(mu); return picked()->xform; }
    void set_transform(const Transform& x);
    void add_object(const HitP& h);
    std::shared_ptr<Skylight> get_skylight();
    void swap_skylight(const Skylight& sky);
    void request_skylight_update();
    Material get_material() { std::lock_guard<std::mutex> g(mu); return *picked()->mat; }
    void set_material(const Material&);
};

struct RenderEngine {
    CtxP ctx;
    Camera cam;
    uint32_t ray_depth, seed;
    uint32_t pattern_spp;        // size of the sample pattern built at construction (camera.h:296-305)
    uint32_t sample_idx = 1;
    std::shared_ptr<Scene> scene;
    std::unique_ptr<Lowered> lowered;   // the descriptors of what is on the device (tables grow with add_object)
    bool has_scene = false;
    std::shared_ptr<RenderLayers> layers;
    std::shared_ptr<TransferStream> stream;
    EngineRequests requests;
    std::atomic<EngineState> state{EngineState::IDLE};
    std::atomic<EngineType> engine_type{EngineType::PATHTRACER};
    std::atomic<RenderMode> render_mode{RenderMode::FULL};
    std::atomic<bool> show_axis_grid{true};
    std::atomic<bool> mis{false};      // NR_MODE_FAST (no counterpart in the reference)
    std::chrono::steady_clock::time_point last_poll{};

    std::function<void()> on_render_started = [] {}, on_render_finished = [] {}, on_stopped = [] {},
                          on_restarted = [] {}, on_reset = [] {}, on_shutdown = [] {};
    std::function<void(uint32_t)> on_frame_finished = [](uint32_t) {};
    std::function<EnginePollResponse()> poll_updates = [] { return EnginePollResponse{}; };
    bool frame_hook = false, poll_hook = false;
    SceneInterface iface;

    RenderEngine(const Camera& camera, uint32_t depth, uint32_t seed_)
        : ctx(std::make_shared<Ctx>(0)), cam(camera), ray_depth(depth), seed(seed_) {
        iface.eng = this;
        cam.p.build();
        pattern_spp = cam.p.samples_per_pixel;
        nr_camera_desc d = cam.p.desc(pattern_spp);
        check(nr_camera_set(ctx->h, 0, &d, seed));
        const size_t H = (size_t)cam.p.resolution.e[0], W = (size_t)cam.p.resolution.e[1];
        layers = std::make_shared<RenderLayers>(ctx, 0, H, W);
        stream = std::make_shared<TransferStream>();
        stream->ctx = ctx;
        stream->H = H;
        stream->W = W;
        stream->state = StreamState::ACTIVE;
        reset();
    }

    uint32_t n_samples() const { return cam.p.samples_per_pixel; }

    void reset() {                                                           // engine.cpp:216-220
        sample_idx = mis.load() ? 0u : 1u;       // fast mode renders all n pattern entries
        call_with_gil(on_reset);
    }

    void set_scene(const Scene& s) {                                         // engine.cpp:117-121
        auto L = lower_scene(s);
        std::lock_guard<std::mutex> g(ctx->mu);
        std::unique_lock<std::shared_mutex> lay(ctx->layout_mu);
        check(nr_scene_upload(ctx->h, &L->desc));
        scene = std::make_shared<Scene>(s);
        lowered = std::move(L);
        has_scene = true;
    }

    // Registers h last (Scene::add_hittable, scene.cpp:101-104) and uploads it with
    // what it brings (its material, new textures, a new mesh) as a partial update.
    void add_object(const HitP& h) {
        std::lock_guard<std::mutex> g(ctx->mu);
        std::unique_lock<std::shared_mutex> lay(ctx->layout_mu);
        Lowered& L = *lowered;
        const size_t i0 = L.imgs.size(), m0 = L.mats.size(), me0 = L.meshes.size();
        lower_object(L, h);
        nr_scene_desc add{};
        add.n_materials = (int32_t)(L.mats.size() - m0);
        add.materials = L.mats.data() + m0;
        add.n_images = (int32_t)(L.imgs.size() - i0);
        add.images = L.imgs.data() + i0;
        add.n_meshes = (int32_t)(L.meshes.size() - me0);
        add.meshes = L.meshes.data() + me0;
        const int rc = nr_scene_add_object(ctx->h, &L.objs.back(), &add);
        if (rc != NR_OK) {                      // keep the lowered tables equal to the device's
            L.objs.pop_back();
            L.mats.resize(m0);
            L.imgs.resize(i0);
            L.keep_tex.resize(i0);
            L.meshes.resize(me0);
            L.keep_mesh.resize(me0);
            check(rc);
        }
        scene->objects.push_back(h);
    }

    // uploads the engine's sky alone (Skylight::build in Scene::update, scene.cpp:126-127)
    void upload_sky() {
        std::lock_guard<std::mutex> g(ctx->mu);
        std::unique_lock<std::shared_mutex> lay(ctx->layout_mu);
        nr_sky_desc sd = lower_sky(*lowered, scene->sky);
        check(nr_scene_set_sky(ctx->h, &sd));
    }

    void push_camera() {
        nr_camera_desc d = cam.p.desc(pattern_spp);
        std::lock_guard<std::mutex> g(ctx->mu);
        std::unique_lock<std::shared_mutex> lay(ctx->layout_mu);
        check(nr_camera_set(ctx->h, 0, &d, seed));
    }

    ReturnState viewport() {                                                 // engine.cpp:76-96
        if (!has_scene) throw std::runtime_error("RenderEngine.viewport() called before set_scene()");
        iface.update();
        poll_gui_updates();
        if (requests.stop_pending()) return ReturnState::STOPPED;
        {
            std::lock_guard<std::mutex> g(ctx->mu);
            std::shared_lock<std::shared_mutex> lay(ctx->layout_mu);
            check(nr_viewport(ctx->h, 0, show_axis_grid.load() ? 1 : 0));
        }
        if (requests.restart_pending()) return ReturnState::RESTARTED;
        return ReturnState::WAITING;
    }

    void poll_gui_updates() {                                                // engine.cpp:170-191
        if (!poll_hook) return;
        auto now = std::chrono::steady_clock::now();
        if (now - last_poll >= std::chrono::milliseconds(16)) {
            last_poll = now;
            EnginePollResponse r;
            { py::gil_scoped_acquire g; r = poll_updates(); }
            if (r.should_update_camera) {
                cam.p.update(r.camera_params);
                call_with_gil(cam.on_updated, cam.p);
                push_camera();
                render_mode = RenderMode::INTERACTIVE;
                requests.restart();
            }
        } else if (render_mode.load() == RenderMode::INTERACTIVE) {
            render_mode = RenderMode::FULL;
        }
    }

    ReturnState render() {                                                   // engine.cpp:42-74
        if (!has_scene) throw std::runtime_error("RenderEngine.render() called before set_scene()");
        const bool per_frame = frame_hook || poll_hook;
        while (sample_idx < n_samples()) {
            iface.update();
            poll_gui_updates();
            if (requests.stop_pending()) return ReturnState::STOPPED;
            const uint32_t end = per_frame ? sample_idx + 1 : n_samples();
            nr_render_desc d{};
            d.sample_begin = sample_idx;
            d.sample_end = end;
            d.ray_depth = ray_depth;
            d.seed = seed;
            d.do_rr = 1;
            d.rr_start_bounce = 3;
            d.rr_survival = render_mode.load() == RenderMode::INTERACTIVE ? 0.15f : 0.95f;
            d.mode = mis.load() ? NR_MODE_FAST : NR_MODE_PARITY;
            {
                std::lock_guard<std::mutex> g(ctx->mu);
                std::shared_lock<std::shared_mutex> lay(ctx->layout_mu);
                check(nr_render(ctx->h, 0, &d));
                check(nr_sync(ctx->h));
            }
            while (sample_idx < end) {
                sample_idx++;
                if (frame_hook) call_with_gil(on_frame_finished, sample_idx);
            }
            if (requests.restart_pending()) return ReturnState::RESTARTED;
        }
        return ReturnState::FINISHED;
    }

    void handle(ReturnState rs) {                                            // engine.cpp:150-168
        switch (rs) {
            case ReturnState::WAITING:
                requests.start();
                std::this_thread::sleep_for(std::chrono::milliseconds(33));
                break;
            case ReturnState::STOPPED:
                state = EngineState::IDLE;
                call_with_gil(on_stopped);
                break;
            case ReturnState::RESTARTED:
                requests.start();
                call_with_gil(on_restarted);
                break;
            case ReturnState::FINISHED:
                state = EngineState::IDLE;
                call_with_gil(on_render_finished);
                break;
        }
    }

    void idle() {                                                            // engine.cpp:96-115
        state = EngineState::IDLE;
        while (!requests.shutdown_pending()) {
            if (requests.start_pending()) {
                reset();
                state = EngineState::RENDERING;
                call_with_gil(on_render_started);
                ReturnState rs = engine_type.load() == EngineType::VIEWPORT ? viewport() : render();
                handle(rs);
                continue;
            }
            iface.update();
            std::this_thread::sleep_for(std::chrono::milliseconds(16));
        }
        call_with_gil(on_shutdown);
    }

    void set_n_samples(uint32_t n) {                                         // engine.cpp:129-134
        { std::lock_guard<std::mutex> g(ctx->mu); check(nr_sync(ctx->h)); }
        cam.p.samples_per_pixel = n;
        reset();
    }
    void set_max_depth(uint32_t v) {
        { std::lock_guard<std::mutex> g(ctx->mu); check(nr_sync(ctx->h)); }
        ray_depth = v;
        reset();
    }

    // closest hits of world-space rays (SceneGraph::hit), origin[3] + direction[3] per row
    std::vector<HitRecord> trace(py::array_t<float, py::array::c_style | py::array::forcecast> rays) {
        if (!has_scene) throw std::runtime_error("trace_closest() called before set_scene()");
        if (rays.ndim() != 2 || rays.shape(1) != 6) throw std::runtime_error("rays must have shape (n, 6)");
        const int32_t n = (int32_t)rays.shape(0);
        std::vector<nr_hit> hits(n);
        {
            std::lock_guard<std::mutex> g(ctx->mu);
            check(nr_trace_closest(ctx->h, rays.data(), n, hits.data()));
        }
        std::vector<HitRecord> out(n);
        for (int i = 0; i < n; i++) {
            const nr_hit& h = hits[i];
            HitRecord& r = out[i];
            r.object_id = h.object_id; r.material_id = h.material_id; r.prim = h.prim;
            r.front_face = h.front_face != 0; r.t = h.t;
            r.p = v3(h.p[0], h.p[1], h.p[2]);
            r.n = v3(h.n[0], h.n[1], h.n[2]);
            r.tangent = v3(h.tangent[0], h.tangent[1], h.tangent[2]);
            r.uv = v2(h.uv[0], h.uv[1]);
        }
        return out;
    }
};

// --- SceneInterface (interface.h) -------------------------------------------
void SceneInterface::update() {                                              // interface.h:51-56
    if (!enabled.load()) return;
    RenderEngine& e = *eng;
    TransferStream& st = *e.stream;
    if (teardown_pending.load()) {                                           // interface.h:202-207
        st.overlay_mask = nullptr;
        std::lock_guard<std::mutex> g(mu);
        rec = HitRecord{};
        rec.object_id = 0;
        teardown_pending = false;
        enabled = false;
        return;
    }
    if (!e.has_scene) return;
    int obj;
    { std::lock_guard<std::mutex> g(mu); obj = rec.object_id; }
    std::lock_guard<std::mutex> g(e.ctx->mu);                                // build_selection_mask, interface.h:193-199
    std::shared_lock<std::shared_mutex> lay(e.ctx->layout_mu);
    uint8_t* m = nr_view_mask_buffer(e.ctx->h, 0);
    if (!m) throw std::runtime_error(std::string("nectar_hip: ") + nr_last_error());
    check(nr_selection_mask(e.ctx->h, 0, obj, outline_radius, m));
    st.overlay_color = mask_color;
    st.overlay_mask = m;
}

// query_scene (interface.h:59-70): the picked record; a miss leaves the
// default record (object 0), as the reference's hit_test_ray does
void SceneInterface::query_scene(float u, float v) {
    RenderEngine& e = *eng;
    if (!e.has_scene) throw std::runtime_error("query_scene() called before set_scene()");
    nr_hit h{};
    {
        std::lock_guard<std::mutex> g(e.ctx->mu);
        check(nr_pick(e.ctx->h, 0, u, v, &h));
    }
    std::lock_guard<std::mutex> g(mu);
    rec = HitRecord{};
    rec.object_id = 0;
    if (h.object_id >= 0) {
        rec.object_id = h.object_id; rec.material_id = h.material_id; rec.prim = h.prim;
        rec.front_face = h.front_face != 0; rec.t = h.t;
        rec.p = v3(h.p[0], h.p[1], h.p[2]);
        rec.n = v3(h.n[0], h.n[1], h.n[2]);
        rec.tangent = v3(h.tangent[0], h.tangent[1], h.tangent[2]);
        rec.uv = v2(h.uv[0], h.uv[1]);
    }
    enabled = true;
}

SceneOutline SceneInterface::outline() {                                     // outliner.h:48-66
    RenderEngine& e = *eng;
    SceneOutline o;
    if (!e.has_scene) return o;
    auto L = lower_scene(*e.scene);
    for (size_t i = 0; i < e.scene->objects.size(); i++) {
        SceneNode n;
        n.object_id = i;
        n.material_id = (size_t)L->objs[i].material;
        n.is_light = e.scene->objects[i]->type == NR_HIT_OBJECT_LIGHT;
        n.name = n.type_name = hittable_type_name(e.scene->objects[i]->type);
        o.nodes.push_back(n);
    }
    return o;
}

HitP SceneInterface::picked() {                                              // hit_object(), interface.h:187-189
    RenderEngine& e = *eng;
    if (!e.has_scene || rec.object_id < 0 || (size_t)rec.object_id >= e.scene->objects.size())
        throw std::runtime_error("Object index not valid.");
    return e.scene->objects[(size_t)rec.object_id];
}

// interface.h:127-133: only the object records and the top-level tree are
// rebuilt on the device (nr_object_set_transform)
void SceneInterface::set_transform(const Transform& x) {
    RenderEngine& e = *eng;
    int32_t idx;
    {
        std::lock_guard<std::mutex> g(mu);
        picked()->xform = x;
        idx = rec.object_id;
    }
    float p[3], r[3], sc[3];
    for (int k = 0; k < 3; k++) { p[k] = x.position.e[k]; r[k] = x.rotation_deg.e[k]; sc[k] = x.scale.e[k]; }
    {
        std::lock_guard<std::mutex> g(e.ctx->mu);
        std::unique_lock<std::shared_mutex> lay(e.ctx->layout_mu);
        check(nr_object_set_transform(e.ctx->h, idx, p, r, sc));
    }
    e.requests.restart();
}

void SceneInterface::add_object(const HitP& h) {                             // interface.h:112-116
    RenderEngine& e = *eng;
    if (!e.has_scene) throw std::runtime_error("add_object() called before set_scene()");
    {
        std::lock_guard<std::mutex> g(mu);
        e.add_object(h);
    }
    e.requests.restart();
}

std::shared_ptr<Skylight> SceneInterface::get_skylight() {
    std::lock_guard<std::mutex> g(mu);
    if (!eng->has_scene) throw std::runtime_error("no scene");
    return eng->scene->sky;
}

void SceneInterface::swap_skylight(const Skylight& sky) {                     // interface.h:145-151
    {
        std::lock_guard<std::mutex> g(mu);
        if (!eng->has_scene) throw std::runtime_error("no scene");
        eng->scene->sky = std::make_shared<Skylight>(sky);
    }
    eng->upload_sky();
    eng->requests.restart();
}

void SceneInterface::request_skylight_update() {                             // interface.h:140-143
    if (!eng->has_scene) throw std::runtime_error("no scene");
    eng->upload_sky();
    eng->requests.restart();
}

// set_material (interface.h:155-160) only restarts the render in the reference
void SceneInterface::set_material(const Material&) { eng->requests.restart(); }

template <class T>
std::string vec_repr(const char* name, const T& v, int n) {
    std::string s = std::string(name) + "(";
    for (int i = 0; i < n; i++) s += (i ? ", " : "") + std::to_string(v.e[i]);
    return s + ")";
}

Vector3 to_v3(py::handle h) {
    if (py::isinstance<Vector3>(h)) return h.cast<Vector3>();
    if (py::isinstance<Color>(h)) { Color c = h.cast<Color>(); return v3(c.e[0], c.e[1], c.e[2]); }
    auto s = h.cast<std::vector<float>>();
    if (s.size() != 3) throw py::value_error("expected 3 components");
    return v3(s[0], s[1], s[2]);
}

}  // namespace

PYBIND11_MODULE(_pathtracer, m) {
    m.doc() = "NectarRender host module on MI355X (libnectar_hip).";

    // ---------------------------------------------------------------- core
    auto m_core = m.def_submodule("core", "Core module.");
    auto m_random = m_core.def_submodule("random", "Core RNG submodule.");
    struct Generator {};
    py::class_<Generator>(m_random, "Generator");

    auto m_vec = m_core.def_submodule("vector", "Vector module.");
    py::class_<Vector2>(m_vec, "Vector2")
        .def(py::init([](float e) { return Vector2(e); }), py::arg("e") = 0.0f)
        .def(py::init([](float x, float y) { return v2(x, y); }), py::arg("x") = 0.0f, py::arg("y") = 0.0f)
        .def(py::init([](const std::array<float, 2>& a) { return v2(a[0], a[1]); }))
        .def("__repr__", [](const Vector2& v) { return vec_repr("Vector2", v, 2); })
        .def("__getitem__", [](const Vector2& v, int i) { if (i < 0 || i > 1) throw py::index_error(); return v.e[i]; })
        .def("x", [](const Vector2& v) { return v.e[0]; })
        .def("y", [](const Vector2& v) { return v.e[1]; })
        .def("u", [](const Vector2& v) { return v.e[0]; })
        .def("v", [](const Vector2& v) { return v.e[1]; });
    py::implicitly_convertible<py::tuple, Vector2>();
    py::implicitly_convertible<py::list, Vector2>();

    py::class_<Vector3>(m_vec, "Vector3")
        .def(py::init([](float e) { return Vector3(e); }), py::arg("e") = 0.0f)
        .def(py::init([](float x, float y, float z) { return v3(x, y, z); }), py::arg("x") = 0.0f,
             py::arg("y") = 0.0f, py::arg("z") = 0.0f)
        .def(py::init([](const std::array<float, 3>& a) { return v3(a[0], a[1], a[2]); }))
        .def("__repr__", [](const Vector3& v) { return vec_repr("Vector3", v, 3); })
        .def("__getitem__", [](const Vector3& v, int i) { if (i < 0 || i > 2) throw py::index_error(); return v.e[i]; })
        .def("length", &Vector3::length)
        .def("as_array", [](const Vector3& v) { return std::array<float, 3>{v.e[0], v.e[1], v.e[2]}; })
        .def("x", [](const Vector3& v) { return v.e[0]; })
        .def("y", [](const Vector3& v) { return v.e[1]; })
        .def("z", [](const Vector3& v) { return v.e[2]; })
        .def("u", [](const Vector3& v) { return v.e[0]; })
        .def("v", [](const Vector3& v) { return v.e[1]; })
        .def("w", [](const Vector3& v) { return v.e[2]; })
        .def("__add__", [](const Vector3& a, const Vector3& b) { return vadd(a, b); })
        .def("__add__", [](const Vector3& a, float t) { return sadd(a, t); })
        .def("__radd__", [](const Vector3& a, float t) { return sadd(a, t); })
        .def("__sub__", [](const Vector3& a, const Vector3& b) { return vsub(a, b); })
        .def("__sub__", [](const Vector3& a, float t) { return ssub(a, t); })
        .def("__mul__", [](const Vector3& a, const Vector3& b) { return vmul(a, b); })
        .def("__mul__", [](const Vector3& a, float t) { return smul(a, t); })
        .def("__rmul__", [](const Vector3& a, float t) { return smul(a, t); })
        .def("__truediv__", [](const Vector3& a, const Vector3& b) { return vdiv(a, b); })
        .def("__truediv__", [](const Vector3& a, float t) { return sdiv(a, t); })
        .def("__iadd__", [](Vector3& a, const Vector3& b) { a = vadd(a, b); return a; })
        .def("__iadd__", [](Vector3& a, float t) { a = sadd(a, t); return a; })
        .def("__isub__", [](Vector3& a, const Vector3& b) { a = vsub(a, b); return a; })
        .def("__isub__", [](Vector3& a, float t) { a = ssub(a, t); return a; })
        .def("__imul__", [](Vector3& a, const Vector3& b) { a = vmul(a, b); return a; })
        .def("__imul__", [](Vector3& a, float t) { a = smul(a, t); return a; })
        .def("__itruediv__", [](Vector3& a, const Vector3& b) { a = vdiv(a, b); return a; })
        .def("__itruediv__", [](Vector3& a, float t) { a = sdiv_assign(a, t); return a; });
    py::implicitly_convertible<py::tuple, Vector3>();
    py::implicitly_convertible<py::list, Vector3>();

    py::class_<Color>(m_vec, "Color")
        .def(py::init([](float r, float g, float b) { return c3(r, g, b); }), py::arg("r") = 0.0f,
             py::arg("g") = 0.0f, py::arg("b") = 0.0f)
        .def(py::init([](const std::array<float, 3>& a) { return c3(a[0], a[1], a[2]); }))
        .def("__repr__", [](const Color& c) { return vec_repr("Color", c, 3); })
        .def("__getitem__", [](const Color& v, int i) { if (i < 0 || i > 2) throw py::index_error(); return v.e[i]; })
        .def("r", [](const Color& c) { return c.e[0]; })
        .def("g", [](const Color& c) { return c.e[1]; })
        .def("b", [](const Color& c) { return c.e[2]; })
        .def_static("black", [] { return c3(0, 0, 0); })
        .def_static("white", [] { return c3(1, 1, 1); })
        .def_static("red", [] { return c3(1, 0, 0); })
        .def_static("green", [] { return c3(0, 1, 0); })
        .def_static("blue", [] { return c3(0, 0, 1); })
        .def_static("purple", [] { return c3(1, 0, 1); })
        .def_static("yellow", [] { return c3(1, 1, 0); })
        .def_static("teal", [] { return c3(0, 1, 1); })
        .def("__add__", [](const Color& a, const Color& b) { return vadd(a, b); })
        .def("__add__", [](const Color& a, float t) { return sadd(a, t); })
        .def("__sub__", [](const Color& a, const Color& b) { return vsub(a, b); })
        .def("__sub__", [](const Color& a, float t) { return ssub(a, t); })
        .def("__mul__", [](const Color& a, const Color& b) { return vmul(a, b); })
        .def("__mul__", [](const Color& a, float t) { return smul(a, t); })
        .def("__rmul__", [](const Color& a, float t) { return smul(a, t); })
        .def("__truediv__", [](const Color& a, const Color& b) { return vdiv(a, b); })
        .def("__truediv__", [](const Color& a, float t) { return sdiv(a, t); });
    py::implicitly_convertible<py::tuple, Color>();
    py::implicitly_convertible<py::list, Color>();

    m_vec.def("dot", &dot3);
    m_vec.def("cross", &cross3);
    m_vec.def("reflect", &reflect3);
    m_vec.def("refract", &refract3);

    auto m_mat = m_core.def_submodule("matrix", "Matrix module.");
    py::class_<Matrix3>(m_mat, "Matrix3")
        .def(py::init<>())
        .def("transpose", &Matrix3::T)
        .def("T", &Matrix3::T)
        .def("right", [](const Matrix3& a) { return v3(a.m[0][0], a.m[1][0], a.m[2][0]); })
        .def("up", [](const Matrix3& a) { return v3(a.m[0][1], a.m[1][1], a.m[2][1]); })
        .def("forward", [](const Matrix3& a) { return v3(a.m[0][2], a.m[1][2], a.m[2][2]); })
        .def("numpy", [](const Matrix3& a) {
            std::array<std::array<float, 3>, 3> o;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) o[i][j] = a.m[i][j];
            return o;
        })
        .def("__mul__", [](const Matrix3& a, const Matrix3& b) { return a * b; })
        .def("__mul__", [](const Matrix3& a, const Vector3& v) { return a * v; });
    m_mat.def("rotation_from_euler", &rotation_from_euler);

    py::class_<Transform>(m_core, "Transform")
        .def(py::init<>())
        .def(py::init([](py::object p, py::object r, py::object s) { return Transform(to_v3(p), to_v3(r), to_v3(s)); }),
             py::arg("position") = v3(0, 0, 0), py::arg("rotation") = v3(0, 0, 0), py::arg("scale") = v3(1, 1, 1))
        .def("position", [](const Transform& t) { return t.position; })
        .def("rotation", [](const Transform& t) { return t.R; })
        .def("scale", [](const Transform& t) { return t.scale; })
        .def("set_position", [](Transform& t, const Vector3& p) { t.position = p; })
        .def("set_rotation", [](Transform& t, const Matrix3& R) { t.R = R; })
        .def("set_scale", [](Transform& t, const Vector3& s) { t.scale = s; })
        .def("p", [](const Transform& t) { return t.position; })
        .def("pos", [](const Transform& t) { return t.position; })
        .def("R", [](const Transform& t) { return t.R; });

    // ---------------------------------------------------------------- cuda
    auto m_cuda = m.def_submodule("cuda", "Device utility module.");
    m_cuda.def("destroy_cublas_handle", [] {}, "No BLAS handle is held; kept for atexit registration.");
    m_cuda.def("cudaDeviceSynchronize", [] {
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) py::print("Encountered the following error while attempting to synchronize device:\n",
                                       hipGetErrorString(e));
    });

    // ---------------------------------------------------------------- material
    auto m_material = m.def_submodule("material", "Material module.");
    auto m_texture = m_material.def_submodule("texture", "Texture submodule.");
    py::enum_<TextureType>(m_texture, "TextureType")
        .value("CONSTANT", TextureType::CONSTANT)
        .value("IMAGE", TextureType::IMAGE);
    py::class_<Texture, std::shared_ptr<Texture>>(m_texture, "Texture")
        .def_static("from_color", [](const Color& c) { return Texture::from_color(c); }, py::arg("color"))
        .def_static("from_color", [](float r, float g, float b) { return Texture::from_color(c3(r, g, b)); },
                    py::arg("r"), py::arg("g"), py::arg("b"))
        .def_static("from_image", &Texture::from_image, py::arg("filepath"), py::arg("scale") = 1.0f)
        .def_static("from_array", [](py::array_t<uint8_t, py::array::c_style | py::array::forcecast> a, float scale) {
            if (a.ndim() != 3 || a.shape(0) < 3) throw std::runtime_error("image array must be C x H x W with C >= 3");
            auto t = std::make_shared<Texture>();
            t->type = TextureType::IMAGE;
            t->scale = scale;
            t->C = (size_t)a.shape(0); t->H = (size_t)a.shape(1); t->W = (size_t)a.shape(2);
            t->chw = std::make_shared<std::vector<uint8_t>>(a.data(), a.data() + a.size());
            return t;
        }, py::arg("chw"), py::arg("scale") = 1.0f, "Image texture from a planar C x H x W uint8 array.")
        .def_readonly("type", &Texture::type)
        .def_readonly("filepath", &Texture::filepath)
        .def_readonly("constant_color", &Texture::constant_color)
        .def_readonly("C", &Texture::C)
        .def_readonly("H", &Texture::H)
        .def_readonly("W", &Texture::W)
        .def("__repr__", [](const Texture& t) {
            return std::string("Texture(type = ") + (t.type == TextureType::CONSTANT ? "CONSTANT" : "IMAGE");
        });

    py::enum_<MaterialType>(m_material, "MaterialType")
        .value("Lambertian", MaterialType::Lambertian)
        .value("PBR", MaterialType::PBR)
        .value("Dielectric", MaterialType::Dielectric)
        .value("Emissive", MaterialType::Emissive)
        .value("Isotropic", MaterialType::Isotropic);
    struct MaterialCore {};
    struct Lambertian : MaterialCore {};
    struct PBR : MaterialCore {};
    struct Dielectric : MaterialCore {};
    struct Emissive : MaterialCore {};
    struct Isotropic : MaterialCore {};
    py::class_<MaterialCore>(m_material, "MaterialCore");
    py::class_<Lambertian, MaterialCore>(m_material, "Lambertian");
    py::class_<PBR, MaterialCore>(m_material, "PBR");
    py::class_<Dielectric, MaterialCore>(m_material, "Dielectric");
    py::class_<Emissive, MaterialCore>(m_material, "Emissive");
    py::class_<Isotropic, MaterialCore>(m_material, "Isotropic");

    static const char* kMatNames[] = {"UNKNOWN", "Lambertian", "PBR", "Dielectric", "Emissive", "Isotropic"};
    py::class_<Material, std::shared_ptr<Material>>(m_material, "Material")
        .def(py::init<>())
        .def("material_type", [](const Material& mt) { return mt.type; })
        .def("texture_count", [](const Material& mt) { return (uint8_t)mt.tex.size(); })
        .def("get_tracked_texture", [](const Material& mt, size_t i) {
            if (i >= mt.tex.size()) throw py::index_error("texture index out of range");
            return mt.tex[i];
        })
        .def("__repr__", [](const Material& mt) { return std::string("Material(type = ") + kMatNames[(int)mt.type] + ")"; })
        .def_static("lambertian", [](const Color& c) { return Material::lambertian(Texture::from_color(c)); },
                    py::arg("color") = c3(1, 1, 1))
        .def_static("lambertian", [](TexP t) { return Material::lambertian(std::move(t)); },
                    py::arg("texture") = Texture::from_color(c3(1, 1, 1)))
        .def_static("pbr", &Material::pbr,
                    py::arg("albedo") = Texture::from_color(c3(1, 1, 1)),
                    py::arg("roughness") = Texture::from_color(Color(0.8f)),
                    py::arg("metallic") = Texture::from_color(Color(0.0f)),
                    py::arg("emission") = Texture::from_color(c3(0, 0, 0)),
                    py::arg("normal") = Texture::from_color(c3(0.5f, 0.5f, 1.0f)))
        .def_static("dielectric", &Material::dielectric, py::arg("tint") = c3(1, 1, 1), py::arg("ior") = 1.5f)
        .def_static("emissive", [](const Color& c, float b) { return Material::emissive(Texture::from_color(c), b); },
                    py::arg("albedo") = c3(1, 1, 1), py::arg("brightness") = 35.0f)
        .def_static("emissive", [](TexP t, float b) { return Material::emissive(std::move(t), b); },
                    py::arg("texture") = Texture::from_color(c3(1, 1, 1)), py::arg("brightness") = 35.0f)
        .def_static("isotropic", [](const Color& c) { return Material::isotropic(Texture::from_color(c)); },
                    py::arg("albedo") = c3(1, 1, 1))
        .def_static("isotropic", [](TexP t) { return Material::isotropic(std::move(t)); },
                    py::arg("texture") = Texture::from_color(c3(1, 1, 1)));

    // ---------------------------------------------------------------- hittable
    auto m_hittable = m.def_submodule("hittable", "Hittable module.");
    py::class_<HitRecord>(m_hittable, "HitRecord")
        .def(py::init<>())
        .def_readonly("object_id", &HitRecord::object_id)
        .def_readonly("material_id", &HitRecord::material_id)
        .def_readonly("prim", &HitRecord::prim)
        .def_readonly("p", &HitRecord::p)
        .def_readonly("n", &HitRecord::n)
        .def_readonly("tangent", &HitRecord::tangent)
        .def_readonly("uv", &HitRecord::uv)
        .def_readonly("t", &HitRecord::t)
        .def_readonly("front_face", &HitRecord::front_face)
        .def_property_readonly("mat", [](const HitRecord& r) { return r.material_id; });
    struct Quad {}; struct Sphere {}; struct Cube {}; struct Mesh {}; struct ObjectLight {}; struct ConstantMedium {};
    py::class_<Quad>(m_hittable, "Quad");
    py::class_<Sphere>(m_hittable, "Sphere");
    py::class_<Cube>(m_hittable, "Cube");
    py::class_<Mesh>(m_hittable, "Mesh");
    py::class_<ObjectLight>(m_hittable, "ObjectLight");
    py::class_<ConstantMedium>(m_hittable, "ConstantMedium");

    py::class_<Hittable, std::shared_ptr<Hittable>>(m_hittable, "Hittable")
        .def(py::init<>())
        .def_static("quad", [](py::object p, py::object r, py::object s, const Material& mt) {
            return make_obj(NR_HIT_QUAD, to_v3(p), to_v3(r), to_v3(s), mt);
        }, py::arg("position") = v3(0, 0, 0), py::arg("rotation") = v3(0, 0, 0), py::arg("scale") = v3(0, 0, 0),
           py::arg("material") = default_material())
        .def_static("cube", [](py::object p, py::object r, py::object s, const Material& mt) {
            return make_obj(NR_HIT_CUBE, to_v3(p), to_v3(r), to_v3(s), mt);
        }, py::arg("position") = v3(0, 0, 0), py::arg("rotation") = v3(0, 0, 0), py::arg("scale") = v3(0, 0, 0),
           py::arg("material") = default_material())
        .def_static("sphere", [](py::object p, float radius, const Material& mt) {
            // Transform(position): no rotation, unit scale (transform.h:17-23)
            HitP h = make_obj(NR_HIT_SPHERE, to_v3(p), v3(0, 0, 0), v3(1, 1, 1), mt);
            h->radius = radius;
            return h;
        }, py::arg("position") = v3(0, 0, 0), py::arg("radius") = 1.0f, py::arg("material") = default_material())
        .def_static("mesh", [](const std::string& path, py::object p, py::object r, py::object s, const Material& mt) {
            auto g = std::make_shared<MeshGeom>();
            g->data = nrio::load_obj(path);
            g->path = path;
            HitP h = make_obj(NR_HIT_MESH, to_v3(p), to_v3(r), to_v3(s), mt);
            h->mesh = g;
            return h;
        }, py::arg("path"), py::arg("position") = v3(0, 0, 0), py::arg("rotation") = v3(0, 0, 0),
           py::arg("scale") = v3(0, 0, 0), py::arg("material") = default_material())
        .def_static("mesh_from_arrays", [](py::array_t<float, py::array::c_style | py::array::forcecast> verts,
                                           py::array_t<uint32_t, py::array::c_style | py::array::forcecast> tris,
                                           py::object p, py::object r, py::object s, const Material& mt) {
            if (verts.ndim() != 2 || verts.shape(1) != 11) throw std::runtime_error("vertices must be (n, 11)");
            if (tris.ndim() != 2 || tris.shape(1) != 3) throw std::runtime_error("triangles must be (m, 3)");
            auto g = std::make_shared<MeshGeom>();
            g->data.vertices.assign(verts.data(), verts.data() + verts.size());
            g->data.triangles.assign(tris.data(), tris.data() + tris.size());
            HitP h = make_obj(NR_HIT_MESH, to_v3(p), to_v3(r), to_v3(s), mt);
            h->mesh = g;
            return h;
        }, py::arg("vertices"), py::arg("triangles"), py::arg("position") = v3(0, 0, 0),
           py::arg("rotation") = v3(0, 0, 0), py::arg("scale") = v3(1, 1, 1), py::arg("material") = default_material(),
           "Mesh from MeshVertex rows (position, normal, tangent, uv: 11 floats) and triangle indices.")
        .def_static("constant_medium", [](const Hittable& b, float density, TexP t) {
            return wrap(b, NR_HIT_CONSTANT_MEDIUM, density, Material::isotropic(std::move(t)));
        }, py::arg("bound_obj"), py::arg("density") = 1.0f, py::arg("texture") = Texture::from_color(c3(1, 1, 1)))
        .def_static("constant_medium", [](const Hittable& b, float density, const Color& albedo) {
            return wrap(b, NR_HIT_CONSTANT_MEDIUM, density, Material::isotropic(Texture::from_color(albedo)));
        }, py::arg("bound_obj"), py::arg("density") = 1.0f, py::arg("albedo") = c3(1, 1, 1))
        .def_static("object_light", [](const Hittable& b, float brightness, TexP t) {
            return wrap(b, NR_HIT_OBJECT_LIGHT, 0.0f, Material::emissive(std::move(t), brightness));
        }, py::arg("bound_obj"), py::arg("brightness") = 1.0f, py::arg("texture") = Texture::from_color(c3(1, 1, 1)))
        .def_static("object_light", [](const Hittable& b, float brightness, const Color& albedo) {
            return wrap(b, NR_HIT_OBJECT_LIGHT, 0.0f, Material::emissive(Texture::from_color(albedo), brightness));
        }, py::arg("bound_obj"), py::arg("brightness") = 1.0f, py::arg("albedo") = c3(1, 1, 1));

    // ---------------------------------------------------------------- light
    auto m_light = m.def_submodule("light", "Lights module.");
    py::enum_<SkylightType>(m_light, "SkylightType")
        .value("Null", SkylightType::Null)
        .value("Simple", SkylightType::Simple)
        .value("HDRI", SkylightType::HDRI);
    py::class_<SimpleSkylightConfig>(m_light, "SimpleSkylightConfig")
        .def_readwrite("start", &SimpleSkylightConfig::start)
        .def_readwrite("end", &SimpleSkylightConfig::end);
    py::class_<HDRISkylightConfig>(m_light, "HDRISkylightConfig")
        .def_readwrite("rotation", &HDRISkylightConfig::rotation)
        .def_readwrite("intensity", &HDRISkylightConfig::intensity);
    struct SimpleSky {}; struct HdriSky {};
    py::class_<SimpleSky>(m_light, "Simple");
    py::class_<HdriSky>(m_light, "HDRI");
    py::class_<Skylight, std::shared_ptr<Skylight>>(m_light, "Skylight")
        .def(py::init<>())
        .def_static("simple", [](const Color& s, const Color& e) {
            auto k = std::make_shared<Skylight>();
            k->type = SkylightType::Simple;
            k->cfg_simple.start = s;
            k->cfg_simple.end = e;
            return k;
        }, py::arg("start_color") = c3(1, 1, 1), py::arg("end_color") = c3(0.5f, 0.7f, 1.0f))
        .def_static("hdri", [] { auto k = std::make_shared<Skylight>(); k->type = SkylightType::HDRI; return k; })
        .def_static("hdri", [](const std::string& path) {
            auto k = std::make_shared<Skylight>();
            k->type = SkylightType::HDRI;
            k->load_hdri_file(path);
            return k;
        }, py::arg("filepath"))
        .def_static("hdri_from_array", [](py::array_t<float, py::array::c_style | py::array::forcecast> a,
                                          float rotation, float intensity) {
            if (a.ndim() != 3 || a.shape(0) < 3) throw std::runtime_error("HDRI array must be C x H x W with C >= 3");
            auto k = std::make_shared<Skylight>();
            k->type = SkylightType::HDRI;
            k->cfg_hdri.C = (int)a.shape(0); k->cfg_hdri.H = (int)a.shape(1); k->cfg_hdri.W = (int)a.shape(2);
            k->cfg_hdri.rotation = rotation;
            k->cfg_hdri.intensity = intensity;
            k->cfg_hdri.data = std::make_shared<std::vector<float>>(a.data(), a.data() + a.size());
            return k;
        }, py::arg("chw"), py::arg("rotation") = 0.0f, py::arg("intensity") = 1.0f,
           "HDRI sky from a planar C x H x W float32 array.")
        .def("load_hdri_file", &Skylight::load_hdri_file)
        .def("config_simple", [](Skylight& k) -> SimpleSkylightConfig& { return k.cfg_simple; }, rvp::reference_internal)
        .def("config_hdri", [](Skylight& k) -> HDRISkylightConfig& { return k.cfg_hdri; }, rvp::reference_internal);

    // ---------------------------------------------------------------- data
    auto m_data = m.def_submodule("data", "Data module.");
    py::enum_<StreamState>(m_data, "StreamState")
        .value("ACTIVE", StreamState::ACTIVE)
        .value("INACTIVE", StreamState::INACTIVE)
        .value("FROZEN", StreamState::FROZEN);
    py::enum_<TonemapMethod>(m_data, "TonemapMethod")
        .value("REINHARD", TonemapMethod::REINHARD)
        .value("REINHARD_EXTENDED", TonemapMethod::REINHARD_EXTENDED)
        .value("ACES", TonemapMethod::ACES);
    py::enum_<DenoiseFilterQuality>(m_data, "DenoiseFilterQuality")
        .value("DEFAULT", DenoiseFilterQuality::DEFAULT)
        .value("HIGH", DenoiseFilterQuality::HIGH)
        .value("BALANCED", DenoiseFilterQuality::BALANCED)
        .value("FAST", DenoiseFilterQuality::FAST);
    py::class_<StreamConfig>(m_data, "StreamConfig")
        .def(py::init<>())
        .def_readwrite("apply_denoising", &StreamConfig::apply_denoising)
        .def_readwrite("denoise_clean_auxiliaries", &StreamConfig::denoise_clean_auxiliaries)
        .def_readwrite("denoise_quality", &StreamConfig::denoise_quality)
        .def_readwrite("denoise_input_scale", &StreamConfig::denoise_input_scale)
        .def_readwrite("linear_to_gamma", &StreamConfig::linear_to_gamma)
        .def_readwrite("apply_white_balance", &StreamConfig::apply_white_balance)
        .def_readwrite("wb_temperature", &StreamConfig::wb_temperature)
        .def_readwrite("wb_tint", &StreamConfig::wb_tint)
        .def_readwrite("apply_tonemapping", &StreamConfig::apply_tonemapping)
        .def_readwrite("tm_method", &StreamConfig::tm_method)
        .def_readwrite("tm_white_point", &StreamConfig::tm_white_point)
        .def_readwrite("tm_alpha", &StreamConfig::tm_alpha);

    py::class_<TransferStream, std::shared_ptr<TransferStream>>(m_data, "TransferStream")
        .def("readback", [](TransferStream& s) {
            s.prepare();
            py::gil_scoped_release r;
            return s.readback();
        })
        .def("readback_numpy", &TransferStream::readback_numpy, "Post-processed uint8 C x H x W copy.")
        .def("readback_float", &TransferStream::readback_float,
             "Display-referred float image (white balance, tonemap, gamma, clamp) before the x255.")
        .def("update_config", [](TransferStream& s, const StreamConfig& c) {
            std::lock_guard<std::mutex> g(s.cfg_mu);
            s.cfg = c;
        })
        .def("get_config", [](TransferStream& s) { std::lock_guard<std::mutex> g(s.cfg_mu); return s.cfg; })
        .def("buffer_ptr", &TransferStream::buffer_ptr)
        .def("get_state", [](const TransferStream& s) { return s.state.load(); })
        .def("is_active", [](const TransferStream& s) { return s.state.load() == StreamState::ACTIVE; })
        .def("is_inactive", [](const TransferStream& s) { return s.state.load() == StreamState::INACTIVE; })
        .def("is_frozen", [](const TransferStream& s) { return s.state.load() == StreamState::FROZEN; })
        .def("freeze", [](TransferStream& s) { if (s.state.load() == StreamState::ACTIVE) s.state = StreamState::FROZEN; })
        .def("unfreeze", [](TransferStream& s) { if (s.state.load() == StreamState::FROZEN) s.state = StreamState::ACTIVE; })
        .def("is_linked", [](const TransferStream& s) { return (bool)s.ctx; })
        .def("has_overlay", [](const TransferStream& s) { return s.overlay_mask.load() != nullptr; })
        .def("remove_overlay", [](TransferStream& s) { s.overlay_mask = nullptr; })
        .def("shape", [](const TransferStream& s) { return std::array<size_t, 3>{s.C, s.H, s.W}; })
        .def("n_pixels", [](const TransferStream& s) { return s.H * s.W; })
        .def("n_elements", [](const TransferStream& s) { return s.C * s.H * s.W; })
        .def("n_bytes", [](const TransferStream& s) { return s.C * s.H * s.W; })
        .def("enable_denoising", [](TransferStream& s) { std::lock_guard<std::mutex> g(s.cfg_mu); s.cfg.apply_denoising = true; })
        .def("disable_denoising", [](TransferStream& s) { std::lock_guard<std::mutex> g(s.cfg_mu); s.cfg.apply_denoising = false; })
        .def("rebuild_denoiser", [](TransferStream& s) {
            std::lock_guard<std::mutex> g(s.readback_mu);
            s.denoiser.reset();                  // recreated with the shipped weights on the next readback
        });

    py::class_<DataObject>(m_data, "DataObject")
        .def("n_pixels", &DataObject::n_pixels)
        .def("n_elements", &DataObject::n_elements)
        .def("n_bytes", &DataObject::n_bytes)
        .def("shape", &DataObject::shape)
        .def("numpy", &DataObject::numpy)
        .def("numpy_float", &DataObject::numpy_float, "Float C x H x W copy of the layer.")
        .def("is_enabled", [](const DataObject& d) { return d.enabled; })
        .def("device_ptr", &DataObject::device_ptr)
        .def_readonly("C", &DataObject::C)
        .def_readonly("H", &DataObject::H)
        .def_readonly("W", &DataObject::W);

    py::enum_<LayerType>(m_data, "LayerType")
        .value("BEAUTY", LayerType::BEAUTY)
        .value("DIFFUSE", LayerType::DIFFUSE)
        .value("WORLD_NORMAL", LayerType::WORLD_NORMAL)
        .value("OBJECT_NORMAL", LayerType::OBJECT_NORMAL)
        .value("SPECULAR", LayerType::SPECULAR)
        .value("SHADOW", LayerType::SHADOW)
        .value("DEPTH", LayerType::DEPTH)
        .value("EMISSION", LayerType::EMISSION)
        .value("OBJECT_ID", LayerType::OBJECT_ID);

    py::class_<RenderLayersConfig>(m_data, "RenderLayersConfig")
        .def(py::init<>())
        .def(py::init([](bool beauty, bool diffuse, bool world_normal, bool object_normal, bool specular, bool shadow,
                         bool depth, bool emission, bool object_id) {
            RenderLayersConfig c;
            c.beauty = beauty; c.diffuse = diffuse; c.world_normal = world_normal; c.object_normal = object_normal;
            c.specular = specular; c.shadow = shadow; c.depth = depth; c.emission = emission; c.object_id = object_id;
            return c;
        }), py::arg("beauty") = true, py::arg("diffuse") = true, py::arg("world_normal") = true,
            py::arg("object_normal") = false, py::arg("specular") = false, py::arg("shadow") = false,
            py::arg("depth") = false, py::arg("emission") = false, py::arg("object_id") = false)
        .def_readwrite("beauty", &RenderLayersConfig::beauty)
        .def_readwrite("diffuse", &RenderLayersConfig::diffuse)
        .def_readwrite("world_normal", &RenderLayersConfig::world_normal)
        .def_readwrite("object_normal", &RenderLayersConfig::object_normal)
        .def_readwrite("specular", &RenderLayersConfig::specular)
        .def_readwrite("shadow", &RenderLayersConfig::shadow)
        .def_readwrite("depth", &RenderLayersConfig::depth)
        .def_readwrite("emission", &RenderLayersConfig::emission)
        .def_readwrite("object_id", &RenderLayersConfig::object_id);

    auto layer_prop = [](int L) {
        return [L](RenderLayers& r) -> DataObject& { return r.layers[L]; };
    };
    py::class_<RenderLayers, std::shared_ptr<RenderLayers>>(m_data, "RenderLayers")
        .def(py::init([](size_t h, size_t w, const RenderLayersConfig& cfg) {
            return std::make_shared<RenderLayers>(h, w, cfg);
        }), py::arg("h"), py::arg("w"), py::arg("cfg") = RenderLayersConfig())
        .def("get_layer", [](RenderLayers& r, LayerType t) -> DataObject& { return r.get(t); }, rvp::reference_internal)
        .def_readonly("H", &RenderLayers::H)
        .def_readonly("W", &RenderLayers::W)
        .def_property_readonly("beauty", layer_prop(NR_LAYER_BEAUTY), rvp::reference_internal)
        .def_property_readonly("diffuse", layer_prop(NR_LAYER_DIFFUSE), rvp::reference_internal)
        .def_property_readonly("world_normal", layer_prop(NR_LAYER_WORLD_NORMAL), rvp::reference_internal)
        .def_property_readonly("object_normal", layer_prop(NR_LAYER_OBJECT_NORMAL), rvp::reference_internal)
        .def_property_readonly("specular", layer_prop(NR_LAYER_SPECULAR), rvp::reference_internal)
        .def_property_readonly("shadow", layer_prop(NR_LAYER_SHADOW), rvp::reference_internal)
        .def_property_readonly("depth", layer_prop(NR_LAYER_DEPTH), rvp::reference_internal)
        .def_property_readonly("emission", layer_prop(NR_LAYER_EMISSION), rvp::reference_internal)
        .def_property_readonly("object_id", layer_prop(NR_LAYER_OBJECT_ID), rvp::reference_internal);

    // ---------------------------------------------------------------- scene
    auto m_scene = m.def_submodule("scene", "Scene module.");
    py::class_<SceneNode>(m_scene, "SceneNode")
        .def_readonly("object_id", &SceneNode::object_id)
        .def_readonly("parent_id", &SceneNode::parent_id)
        .def_readonly("material_id", &SceneNode::material_id)
        .def_readonly("is_light", &SceneNode::is_light)
        .def_readonly("name", &SceneNode::name)
        .def_readonly("type_name", &SceneNode::type_name);
    py::class_<SceneOutline>(m_scene, "SceneOutline")
        .def_readonly("nodes", &SceneOutline::nodes);
    py::class_<SceneInterface>(m_scene, "SceneInterface")
        .def("set_material", [](SceneInterface& si, const Material& m) { si.set_material(m); }, py::arg("mat"))
        .def("add_object", [](SceneInterface& si, const HitP& h) { si.add_object(h); }, py::arg("obj"))
        .def("query_scene", [](SceneInterface& si, float u, float v) { py::gil_scoped_release r; si.query_scene(u, v); })
        .def("select_scene_node", &SceneInterface::select_scene_node)
        .def("enable", &SceneInterface::enable)
        .def("disable", &SceneInterface::disable)
        .def("is_enabled", &SceneInterface::is_enabled)
        .def("is_disabled", [](const SceneInterface& si) { return !si.is_enabled(); })
        .def("get_scene_outline", &SceneInterface::outline)
        .def("get_transform", &SceneInterface::get_transform)
        .def("set_transform", [](SceneInterface& si, const Transform& x) { py::gil_scoped_release r; si.set_transform(x); })
        .def("get_material", &SceneInterface::get_material)
        .def("get_hit_record", [](SceneInterface& si) { std::lock_guard<std::mutex> g(si.mu); return si.rec; })
        .def("request_skylight_update", &SceneInterface::request_skylight_update)
        .def("get_skylight", &SceneInterface::get_skylight)
        .def("swap_skylight", [](SceneInterface& si, const Skylight& sky) { si.swap_skylight(sky); },
             py::arg("new_skylight"))
        .def("update", [](SceneInterface& si) { py::gil_scoped_release r; si.update(); },
             "Rebuild the selection outline overlay (the engine does this every pass and while idle).");
    py::class_<Scene, std::shared_ptr<Scene>>(m_scene, "Scene")
        .def(py::init([](py::list hittables, py::list lights, std::shared_ptr<Skylight> sky) {
            auto s = std::make_shared<Scene>();
            for (auto h : hittables) s->objects.push_back(h.cast<HitP>());
            s->n_hittables = s->objects.size();
            for (auto h : lights) s->objects.push_back(h.cast<HitP>());
            s->sky = sky ? std::make_shared<Skylight>(*sky) : std::make_shared<Skylight>();
            return s;
        }), py::arg("hittables"), py::arg("lights"), py::arg("skylight"))
        .def("n_objects", [](const Scene& s) { return s.objects.size(); })
        .def("lowered", [](const Scene& s) {
            // the C ABI descriptors this scene uploads as (objects, materials, images, meshes, sky)
            auto L = lower_scene(s);
            py::list objs, mats;
            for (const nr_object_desc& o : L->objs) {
                py::dict d;
                d["type"] = o.type; d["bound_type"] = o.bound_type; d["material"] = o.material; d["mesh"] = o.mesh;
                d["position"] = std::array<float, 3>{o.position[0], o.position[1], o.position[2]};
                d["rotation"] = std::array<float, 3>{o.rotation[0], o.rotation[1], o.rotation[2]};
                d["scale"] = std::array<float, 3>{o.scale[0], o.scale[1], o.scale[2]};
                d["radius"] = o.radius; d["density"] = o.density;
                objs.append(d);
            }
            for (const nr_material_desc& m : L->mats) {
                py::dict d;
                d["type"] = m.type;
                py::list tex;
                for (int k = 0; k < 5; k++)
                    tex.append(py::make_tuple(m.tex[k].image,
                                              std::array<float, 3>{m.tex[k].color[0], m.tex[k].color[1], m.tex[k].color[2]},
                                              m.tex[k].scale));
                d["tex"] = tex;
                d["tint"] = std::array<float, 3>{m.tint[0], m.tint[1], m.tint[2]};
                d["ior"] = m.ior; d["brightness"] = m.brightness;
                mats.append(d);
            }
            py::dict sky;
            sky["type"] = L->desc.sky.type;
            sky["start"] = std::array<float, 3>{L->desc.sky.start[0], L->desc.sky.start[1], L->desc.sky.start[2]};
            sky["end"] = std::array<float, 3>{L->desc.sky.end[0], L->desc.sky.end[1], L->desc.sky.end[2]};
            sky["rotation"] = L->desc.sky.rotation; sky["intensity"] = L->desc.sky.intensity;
            sky["shape"] = std::array<int, 3>{L->desc.sky.C, L->desc.sky.H, L->desc.sky.W};
            py::dict out;
            out["objects"] = objs; out["materials"] = mats; out["sky"] = sky;
            out["n_images"] = L->imgs.size(); out["n_meshes"] = L->meshes.size();
            return out;
        }, "The C ABI descriptors (nr_scene_desc) this scene is uploaded as.");

    // ---------------------------------------------------------------- engine
    auto m_engine = m.def_submodule("engine", "Engine module.");
    auto m_camera = m_engine.def_submodule("camera", "Camera submodule.");
    py::class_<CameraParams>(m_camera, "CameraParams")
        .def(py::init([](const Vector2& res, py::object pos, py::object rot, uint32_t spp, float focal, float focus,
                         float aperture, float sensor, float shutter) {
            CameraParams p;
            p.resolution = res; p.position = to_v3(pos); p.rotation = to_v3(rot); p.samples_per_pixel = spp;
            p.focal_length = focal; p.focus_distance = focus; p.aperture = aperture; p.sensor_width = sensor;
            p.shutter_speed = shutter;
            p.build();
            return p;
        }), py::arg("resolution"), py::arg("position"), py::arg("rotation"), py::arg("samples_per_pixel"),
            py::arg("focal_length"), py::arg("focus_distance"), py::arg("aperture"), py::arg("sensor_width"),
            py::arg("shutter_speed"))
        .def_readonly("R", &CameraParams::R)
        .def_readonly("shutter_time", &CameraParams::shutter_time)
        .def_readonly("aspect_ratio", &CameraParams::aspect_ratio)
        .def_readwrite("resolution", &CameraParams::resolution)
        .def_readwrite("position", &CameraParams::position)
        .def_readwrite("rotation", &CameraParams::rotation)
        .def_readwrite("samples_per_pixel", &CameraParams::samples_per_pixel)
        .def_readwrite("focal_length", &CameraParams::focal_length)
        .def_readwrite("focus_distance", &CameraParams::focus_distance)
        .def_readwrite("aperture", &CameraParams::aperture)
        .def_readwrite("sensor_width", &CameraParams::sensor_width)
        .def_readwrite("shutter_speed", &CameraParams::shutter_speed);

    py::class_<Camera>(m_camera, "Camera")
        .def(py::init([](const CameraParams& p) { return Camera(p); }), py::arg("params"))
        .def_readwrite("on_updated", &Camera::on_updated)
        .def("parameters", [](const Camera& c) { return c.p; })
        .def("project_to_screen", &Camera::project_to_screen, py::arg("world_point"));

    py::enum_<EngineState>(m_engine, "EngineState")
        .value("IDLE", EngineState::IDLE)
        .value("RENDERING", EngineState::RENDERING);
    py::enum_<EngineType>(m_engine, "EngineType")
        .value("PATHTRACER", EngineType::PATHTRACER)
        .value("VIEWPORT", EngineType::VIEWPORT);
    py::class_<EnginePollResponse>(m_engine, "EnginePollResponse")
        .def(py::init<>())
        .def_readwrite("should_update_camera", &EnginePollResponse::should_update_camera)
        .def_readwrite("camera_params", &EnginePollResponse::camera_params);
    py::class_<EngineRequests>(m_engine, "EngineRequests")
        .def("start", &EngineRequests::start)
        .def("stop", &EngineRequests::stop)
        .def("restart", &EngineRequests::restart)
        .def("shutdown", &EngineRequests::shutdown);

#define NR_HOOK(name, flag)                                                                    \
    def_property(#name, [](const RenderEngine& e) { return e.name; },                          \
                 [](RenderEngine& e, decltype(RenderEngine::name) f) { e.name = std::move(f); flag; })
    py::class_<RenderEngine>(m_engine, "RenderEngine")
        .def(py::init([](const Camera& c, uint32_t depth, uint32_t seed) {
            return std::make_unique<RenderEngine>(c, depth, seed);
        }), py::arg("camera"), py::arg("ray_depth") = 8u, py::arg("seed") = 54321u)
        .def("set_scene", [](RenderEngine& e, const Scene& s) { e.set_scene(s); })
        .def("idle", [](RenderEngine& e) { py::gil_scoped_release r; e.idle(); })
        .def("render", [](RenderEngine& e) { py::gil_scoped_release r; (void)e.render(); })
        .NR_HOOK(on_render_started, (void)0)
        .NR_HOOK(on_frame_finished, e.frame_hook = true)
        .NR_HOOK(on_render_finished, (void)0)
        .NR_HOOK(on_stopped, (void)0)
        .NR_HOOK(on_restarted, (void)0)
        .NR_HOOK(on_reset, (void)0)
        .NR_HOOK(on_shutdown, (void)0)
        .NR_HOOK(poll_updates, e.poll_hook = true)
        .def("camera", [](RenderEngine& e) -> Camera& { return e.cam; }, rvp::reference_internal)
        .def("layers", [](RenderEngine& e) { return e.layers; })
        .def("scene", [](RenderEngine& e) { return e.scene; })
        .def("stream", [](RenderEngine& e) { return e.stream; })
        .def("requests", [](RenderEngine& e) -> EngineRequests& { return e.requests; }, rvp::reference_internal)
        .def("update_streaming_layer", [](RenderEngine& e, LayerType t) {
            if (!nr_aov_device_ptr(e.ctx->h, 0, (int)t)) throw std::runtime_error("layer is not enabled");
            if (kLayerChannels[(int)t] != 3) throw std::runtime_error("only 3-channel layers can be streamed");
            e.stream->linked = (int)t;
        })
        .def("get_state", [](const RenderEngine& e) { return e.state.load(); })
        .def("is_idle", [](const RenderEngine& e) { return e.state.load() == EngineState::IDLE; })
        .def("is_rendering", [](const RenderEngine& e) { return e.state.load() == EngineState::RENDERING; })
        .def("get_engine_type", [](const RenderEngine& e) { return e.engine_type.load(); })
        .def("set_engine_type", [](RenderEngine& e, EngineType t) { e.engine_type = t; })
        .def("n_samples", &RenderEngine::n_samples)
        .def("set_n_samples", &RenderEngine::set_n_samples)
        .def("max_depth", [](const RenderEngine& e) { return e.ray_depth; })
        .def("set_max_depth", &RenderEngine::set_max_depth)
        .def("set_axis_grid_visible", [](RenderEngine& e, bool v) { e.show_axis_grid = v; })
        .def("set_mis", [](RenderEngine& e, bool v) { e.mis = v; e.reset(); }, py::arg("enabled"),
             "Next-event estimation to an HDRI sky with power-heuristic MIS (NR_MODE_FAST) from the next "
             "pass on; resets the accumulation (fast mode starts at pass 0).  Off (the reference's "
             "estimator) by default.  Not in the reference's API.")
        .def("mis", [](const RenderEngine& e) { return e.mis.load(); })
        .def("sample_index", [](const RenderEngine& e) { return e.sample_idx; },
             "Index of the next sample pass (1 after a reset).")
        .def("viewport", [](RenderEngine& e) { py::gil_scoped_release r; (void)e.viewport(); },
             "One viewport preview frame into the layers (what a VIEWPORT engine renders while idle).")
        .def("get_scene_interface", [](RenderEngine& e) -> SceneInterface& { return e.iface; },
             rvp::reference_internal)
        .def("trace_closest", &RenderEngine::trace,
             "Closest hits of world-space rays, an (n, 6) array of origin and direction.");
#undef NR_HOOK

    // ---------------------------------------------------------------- host
    auto m_host = m.def_submodule("host", "Host module.");
    auto m_memory = m_host.def_submodule("memory", "Memory utils submodule.");
    struct MemInfo { size_t used = 0, free = 0, total = 0; };
    auto meminfo = [] {
        MemInfo i;
        if (hipMemGetInfo(&i.free, &i.total) != hipSuccess) throw std::runtime_error("hipMemGetInfo failed");
        i.used = i.total - i.free;
        return i;
    };
    py::class_<MemInfo>(m_memory, "CUDAMemInfo")
        .def(py::init<>())
        .def_readonly("used", &MemInfo::used)
        .def_readonly("free", &MemInfo::free)
        .def_readonly("total", &MemInfo::total);
    m_memory.def("get_cuda_meminfo", meminfo);
    m_memory.def("get_cuda_memory_used", [meminfo] { return meminfo().used; });
    m_memory.def("get_cuda_memory_free", [meminfo] { return meminfo().free; });
    m_memory.def("get_cuda_memory_total", [meminfo] { return meminfo().total; });

    // ---------------------------------------------------------------- io (file ingest, no device work)
    auto m_io = m.def_submodule("io", "Scene file readers (OBJ, PNG/PNM/HDR).");
    m_io.def("load_obj", [](const std::string& path) {
        nrio::MeshData d = nrio::load_obj(path);
        py::array_t<float> v({(py::ssize_t)d.n_vertices(), (py::ssize_t)11});
        py::array_t<uint32_t> t({(py::ssize_t)d.n_triangles(), (py::ssize_t)3});
        std::memcpy(v.mutable_data(), d.vertices.data(), d.vertices.size() * sizeof(float));
        std::memcpy(t.mutable_data(), d.triangles.data(), d.triangles.size() * sizeof(uint32_t));
        return py::make_tuple(v, t);
    }, py::arg("path"), "MeshVertex rows (n, 11) and triangles (m, 3) as Hittable.mesh builds them.");
    m_io.def("load_image_uint8", [](const std::string& path) {
        nrio::Image8 im = nrio::load_image_u8(path);
        const py::ssize_t planes = (py::ssize_t)(im.chw.size() / ((size_t)im.H * im.W));
        py::array_t<uint8_t> a({planes, (py::ssize_t)im.H, (py::ssize_t)im.W});
        std::memcpy(a.mutable_data(), im.chw.data(), im.chw.size());
        return py::make_tuple(a, im.C);
    }, py::arg("path"), "Planar texture data and the channel count reported for the file.");
    m_io.def("load_image_float", [](const std::string& path) {
        nrio::ImageF im = nrio::load_image_f32(path);
        py::array_t<float> a({(py::ssize_t)im.C, (py::ssize_t)im.H, (py::ssize_t)im.W});
        std::memcpy(a.mutable_data(), im.chw.data(), im.chw.size() * sizeof(float));
        return a;
    }, py::arg("path"), "Planar float image as Skylight.hdri loads it.");
}
