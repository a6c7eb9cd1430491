// nr_device.h — device arithmetic of the MI355X path tracer.
//
// Every routine keeps the reference's operation order; compiled with
// -ffp-contract=off (IEEE +,-,*,/,sqrt on gfx950) and with the transcendentals
// of nr_mathfn.h, results are bit-identical to the same arithmetic on the
// host.  File:line citations are relative to nectar_render/csrc/ of the
// reference.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "nr_mathfn.h"
#include "nr_types.h"

#define NR_EPS 1e-4f
#define NR_FMIN 1e-8f
#define NR_FMAX 1e32f
#define NR_PI 3.1415926535897932385f
#define NR_PI2 (NR_PI * 2.0f)
#define NR_PI4 (NR_PI2 * NR_PI2)

#define NR_HD __host__ __device__ __forceinline__
#define NR_D __device__ __forceinline__

// Transcendentals of the estimators.  The reference's estimator takes the
// deterministic double-evaluated functions of nr_mathfn.h, which the oracle
// shares (bit-identical AOVs).  Fast mode, checked by convergence rather than
// per pixel, takes the device library's float functions and x^5 by products.
template <bool F>
struct Tm {
    NR_D static float cos(float x) { return nrm::cosf_(x); }
    NR_D static float sin(float x) { return nrm::sinf_(x); }
    NR_D static float atan2(float y, float x) { return nrm::atan2f_(y, x); }
    NR_D static float acos(float x) { return nrm::acosf_(x); }
    NR_D static float pow5(float x) { return nrm::powf_(x, 5.0f); }
};
template <>
struct Tm<true> {
    NR_D static float cos(float x) { return cosf(x); }
    NR_D static float sin(float x) { return sinf(x); }
    NR_D static float atan2(float y, float x) { return atan2f(y, x); }
    NR_D static float acos(float x) { return acosf(x); }
    NR_D static float pow5(float x) { const float x2 = x * x; return x2 * x2 * x; }
};

struct f3 { float x, y, z; };
struct f2 { float x, y; };

NR_HD f3 mk3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
NR_HD f3 s3(float a) { return mk3(a, a, a); }
NR_HD f3 ld3(const float* p) { return mk3(p[0], p[1], p[2]); }
NR_HD f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
NR_HD f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
NR_HD f3 operator*(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
NR_HD f3 operator*(f3 a, float t) { return mk3(a.x * t, a.y * t, a.z * t); }
NR_HD f3 operator-(f3 a) { return mk3(-a.x, -a.y, -a.z); }
NR_HD f3 adds(f3 a, float t) { return mk3(a.x + t, a.y + t, a.z + t); }
// Vector / float multiplies by the reciprocal (core/vector.h:251-254)
NR_HD f3 divs(f3 a, float t) { return a * (1.0f / t); }
// Vector / Vector adds FMIN to the divisor (core/vector.h:82-88)
NR_HD f3 divv(f3 a, f3 b) { return mk3(a.x / (b.x + NR_FMIN), a.y / (b.y + NR_FMIN), a.z / (b.z + NR_FMIN)); }
NR_HD float dot(f3 u, f3 v) { return u.x * v.x + u.y * v.y + u.z * v.z; }
NR_HD f3 cross(f3 u, f3 v) { return mk3(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x); }
NR_HD float len2(f3 v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
NR_HD float vlength(f3 v) { return sqrtf(len2(v)); }
NR_HD f3 normalize(f3 v) { return divs(v, vlength(v)); }
NR_HD float comp(f3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
NR_HD f3 reflect(f3 v, f3 n) { return v - n * (2.0f * dot(v, n)); }            // vector.h:469-474
NR_HD f3 refract(f3 uv, f3 n, float ior) {                                     // vector.h:476-485
    float cos_theta = fminf(dot(-uv, n), 1.0f);
    f3 perp = (uv + n * cos_theta) * ior;
    float k = (float)(1.0 - (double)len2(perp));
    return perp + n * (-sqrtf(fabsf(k)));
}
NR_HD f3 mv(const float* R, f3 v) {                                            // matrix.h:16-22
    return mk3(R[0] * v.x + R[1] * v.y + R[2] * v.z,
               R[3] * v.x + R[4] * v.y + R[5] * v.z,
               R[6] * v.x + R[7] * v.y + R[8] * v.z);
}
NR_HD f3 mtv(const float* R, f3 v) {                                           // R^T * v
    return mk3(R[0] * v.x + R[3] * v.y + R[6] * v.z,
               R[1] * v.x + R[4] * v.y + R[7] * v.z,
               R[2] * v.x + R[5] * v.y + R[8] * v.z);
}

struct Ray { f3 o, d, inv; };
NR_HD Ray mkray(f3 o, f3 d) { Ray r; r.o = o; r.d = d; r.inv = mk3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z); return r; }
NR_HD f3 at(const Ray& r, float t) { return r.o + r.d * t; }
NR_HD Ray to_object(const Ray& r, const GXform& x) {                           // ray.h:33-37
    f3 o = mtv(x.R, r.o - ld3(x.pos)) * ld3(x.inv_s);
    f3 d = mtv(x.R, r.d) * ld3(x.inv_s);
    return mkray(o, d);
}

// AABB::hit_t (bvh/aabb.h:135-157): entry distance or FMAX
NR_HD float box_t(const float* mn, const float* mx, const Ray& r, float tmin, float tmax) {
    // The reference leaves the axis loop as soon as tmax <= tmin.  tmin only
    // grows and tmax only shrinks, so finishing all three axes with selects
    // and testing once gives the same answer (NaN slab distances never update
    // either bound, as in the branchy form) without divergent control flow.
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float o = comp(r.o, a), iv = comp(r.inv, a);
        float t0 = (mn[a] - o) * iv;
        float t1 = (mx[a] - o) * iv;
        bool c = t0 < t1;
        float lo = c ? t0 : t1, hi = c ? t1 : t0;
        tmin = lo > tmin ? lo : tmin;
        tmax = hi < tmax ? hi : tmax;
    }
    return tmax <= tmin ? NR_FMAX : tmin;
}

// box_t for a ray whose inverse direction components are all finite and nonzero:
// then no slab distance is NaN ((P - o) * inv is finite or +-inf), so the selects
// are min / max and the result is box_t's bit for bit.
NR_HD float box_t_fin(const float* mn, const float* mx, const Ray& r, float tmin, float tmax) {
    float lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float o = comp(r.o, a), iv = comp(r.inv, a);
        const float t0 = (mn[a] - o) * iv, t1 = (mx[a] - o) * iv;
        lo[a] = fminf(t0, t1);
        hi[a] = fmaxf(t0, t1);
    }
    const float t0 = fmaxf(fmaxf(lo[0], lo[1]), fmaxf(lo[2], tmin));
    const float t1 = fminf(fminf(hi[0], hi[1]), fminf(hi[2], tmax));
    return t1 <= t0 ? NR_FMAX : t0;
}
NR_HD bool inv_finite(const Ray& r) {
    bool f = true;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float m = fabsf(comp(r.inv, a));
        f = f && m < INFINITY && m > 0.0f;
    }
    return f;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 per draw (core/random.h:50-55 with rocRAND's init semantics):
// key = seed, counter = (draw/4, 0, subsequence, 0), word draw%4,
// u = 2^-32 + word * 2^-32 in (0, 1].
// ---------------------------------------------------------------------------
NR_D uint32_t philox_word(uint32_t seed, uint32_t subseq, uint32_t draw, uint32_t subhi = 0u) {
    uint32_t c0 = draw >> 2, c1 = 0u, c2 = subseq, c3 = subhi;
    uint32_t k0 = seed, k1 = 0u;
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    uint32_t w = draw & 3u;
    return w == 0 ? c0 : (w == 1 ? c1 : (w == 2 ? c2 : c3));
}

// subhi: high word of a 64-bit subsequence (fast mode); the parity estimator
// keeps the reference's 32-bit subsequence (subhi = 0)
struct Gen {
    uint32_t seed, subseq, draw, subhi = 0u;
    NR_D float uniform() {
        uint32_t w = philox_word(seed, subseq, draw, subhi);
        draw += 1u;
        const float inv = 2.3283064e-10f;
        return inv + ((float)w * inv);
    }
    NR_D float rf(float mn, float mx) { return uniform() * (mx - mn) + mn; }
    NR_D float rf() { return uniform() * (1.0f - 0.0f) + 0.0f; }
};

// ---------------------------------------------------------------------------
// PCG hash (random/include/hash.h:16-191)
// ---------------------------------------------------------------------------
__constant__ static uint32_t NR_HASH_TABLE[64] = {
    0x9E3779B9u, 0x85EBCA6Bu, 0xC2B2AE35u, 0x27D4EB2Fu, 0x165667B1u, 0x68D2A5B5u, 0xBF58476Du, 0x94D049BBu,
    0xA3B195B5u, 0x1B873593u, 0xCC9E2D51u, 0x6C62272Eu, 0xE654025Bu, 0x2B5AD5B5u, 0x6C696E65u, 0x736F6D65u,
    0x646F6D65u, 0x70736575u, 0x6E646F6Du, 0xFF51AFD7u, 0xC4CEA4D6u, 0x45D9F3Bu,  0x3335B369u, 0xA341316Cu,
    0x52784632u, 0x68656C6Cu, 0x6F20776Fu, 0x726C6421u, 0xB5297A4Du, 0x68E31DA4u, 0x1B56C4E9u, 0xD2A98B26u,
    0x637B2B53u, 0x4F6E4F35u, 0x62DC5AB7u, 0x5A17A2C3u, 0x6EF305DAu, 0x2C9F410Du, 0x3771A6A7u, 0x15C3A7C4u,
    0x8F43CD0Bu, 0xA9DE8FA3u, 0xEBF7A401u, 0x58399E71u, 0x42B91A0Du, 0x4B608671u, 0xB41B21C7u, 0x2D5A73B7u,
    0xE50B4C1Fu, 0x95C48A21u, 0x37A07BA9u, 0xC0F4A537u, 0xA8FA763Bu, 0xD08BE5EDu, 0x4A9C5EB3u, 0x7B14EEA7u,
    0xCC3A6547u, 0x9B89D66Du, 0xE15E5E3Bu, 0x4F3D8A5Fu, 0xD4E0A97Bu, 0x3C6EF372u, 0xA54FF53Au, 0x510E527Fu,
};
#define NR_HASH_SCALAR (1.0f / 4294967296.0f)   // 1.0f / (float)0xFFFFFFFF, hash.h:16

NR_D uint32_t hmix(uint32_t v, uint32_t ci) { v ^= (v >> 16u) * NR_HASH_TABLE[ci & 63u]; return v ^ (v >> 13u); }
NR_D uint32_t pcg_hash(uint32_t i, uint32_t seed, uint32_t stream) {          // hash.h:47-61
    uint32_t s = (seed * 1759274928u) ^ (stream * 2246822519u);
    s = i ^ s;
    s = hmix(s, stream * 3u);
    s = hmix(s, stream * 3u + 1u);
    s = hmix(s, stream * 3u + 2u);
    uint32_t state = s * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
NR_D float pcg_f(float f, uint32_t seed) { return (float)pcg_hash(__float_as_uint(f), seed, 0u) * NR_HASH_SCALAR; }
NR_D float pcg_f3(f3 a, uint32_t seed) {                                       // hash.h:92-105
    uint32_t s = seed * 1759274928u;
    s ^= pcg_hash(__float_as_uint(a.x), (seed + 0u) * 2654435761u, 0u) ^ 668265263u;
    s ^= pcg_hash(__float_as_uint(a.y), (seed + 1u) * 2246822519u, 0u) ^ 374761393u;
    s ^= pcg_hash(__float_as_uint(a.z), (seed + 2u) * 3266489917u, 0u) ^ 2654435761u;
    uint32_t state = s * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (float)((word >> 22u) ^ word) * NR_HASH_SCALAR;
}
NR_D float pcg_ray(const Ray& r) {                                             // hash.h:107-124
    float a = pcg_f3(r.o, 0u);
    float b = pcg_f3(r.d, 3266489917u);
    return (a + b) * 0.5f;
}
NR_D uint32_t f2u_sat(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}
NR_D f2 pcg_vec2(float f, uint32_t seed) {                                     // hash.h:182-191
    f2 r;
    r.x = pcg_f(f, seed);
    uint32_t s = seed ^ f2u_sat(r.x * 4294967296.0f);
    r.y = pcg_f((float)(s * 747796405u + 2891336453u), 0u);
    return r;
}

// ---------------------------------------------------------------------------
// ONB (core/onb.h:5-26) and direction sampling (core/vector.h:489-519)
// ---------------------------------------------------------------------------
struct ONB {
    f3 u, v, w;
    NR_D void build(f3 n) {
        w = normalize(n);
        f3 a = (fabsf(w.x) > 0.9f) ? mk3(0.0f, 1.0f, 0.0f) : mk3(1.0f, 0.0f, 0.0f);
        v = normalize(cross(w, a));
        u = cross(w, v);
    }
    NR_D f3 xf(f3 p) const { return (u * p.x + v * p.y) + w * p.z; }
};
template <bool F = false>
NR_D f3 cosine_dir(Gen& g) {
    float r1 = g.rf();
    float r2 = g.rf();
    float phi = NR_PI2 * r1;
    return mk3(Tm<F>::cos(phi) * sqrtf(r2), Tm<F>::sin(phi) * sqrtf(r2), sqrtf(1.0f - r2));
}
NR_D f3 unit_vector(Gen& g) {
    while (true) {
        float a = g.rf(-1.0f, 1.0f);
        float b = g.rf(-1.0f, 1.0f);
        float c = g.rf(-1.0f, 1.0f);
        f3 p = mk3(a, b, c);
        float l = len2(p);
        if (1e-160 < (double)l && l <= 1.0f) return divs(p, sqrtf(l));
    }
}

// ---------------------------------------------------------------------------
// hit record in flight (HitRecord, hittable/hit_record.h)
// ---------------------------------------------------------------------------
struct HRec {
    f3 p, n, tg;
    f2 uv;
    float t;
    bool front;
    int32_t prim;
};
NR_D void hrec_clear(HRec& r) {
    r.p = s3(0.0f); r.n = s3(0.0f); r.tg = s3(0.0f); r.uv.x = 0.0f; r.uv.y = 0.0f;
    r.t = 0.0f; r.front = false; r.prim = -1;
}

// HitRecord::to_world_space (hit_record.h:24-51)
NR_D void to_world(HRec& rec, const GXform& x, const Ray& obj_ray, const Ray& world_ray, bool bias) {
    rec.p = mv(x.R, rec.p * ld3(x.scale)) + ld3(x.pos);
    rec.tg = normalize(mv(x.R, rec.tg * ld3(x.scale)));
    rec.front = dot(obj_ray.d, rec.n) < 0.0f;
    rec.n = rec.front ? rec.n : -rec.n;
    rec.n = normalize(mv(x.R, rec.n * ld3(x.inv_s)));
    rec.front = dot(world_ray.d, rec.n) < 0.0f;
    rec.n = rec.front ? rec.n : -rec.n;
    rec.tg = normalize(rec.tg - rec.n * dot(rec.tg, rec.n));
    if (vlength(rec.tg) < 1e-8f)
        rec.tg = normalize(cross(rec.n, fabsf(rec.n.x) > 0.9f ? mk3(0.0f, 1.0f, 0.0f) : mk3(1.0f, 0.0f, 0.0f)));
    if (bias) rec.p = rec.p + rec.n * NR_EPS;
    rec.t = dot(rec.p - world_ray.o, world_ray.d) / len2(world_ray.d);
}

// Quad constants (hittable/primitives.h:8-16): u = (1,0,0), v = (0,0,-1),
// n = cross(u, v) = (-0, 1, 0), normal = n, w = n / dot(n, n) = n, area 1.
NR_D f3 quad_u() { return mk3(1.0f, 0.0f, 0.0f); }
NR_D f3 quad_v() { return mk3(0.0f, 0.0f, -1.0f); }
NR_D f3 quad_n() { return cross(quad_u(), quad_v()); }

// Quad::hit (primitives.h:19-32, 73-88)
NR_D bool quad_hit(const Ray& r, float tmin, float tmax, HRec& rec) {
    f3 nrm = normalize(quad_n());
    float denom = dot(nrm, r.d);
    if (fabsf(denom) < NR_FMIN) return false;
    float t = -dot(nrm, r.o) / denom;
    if (!(tmin <= t && t <= tmax)) return false;
    rec.t = t;
    rec.p = at(r, t);
    rec.n = nrm;
    rec.tg = quad_u();
    f3 n = quad_n();
    f3 w = divs(n, dot(n, n));
    float a = dot(w, cross(rec.p, quad_v()));
    float b = dot(w, cross(quad_u(), rec.p));
    rec.uv.x = a + 0.5f;
    rec.uv.y = b + 0.5f;
    return ((-0.5f < a) && (a < 0.5f)) & ((-0.5f < b) && (b < 0.5f));
}

// Sphere::hit (hittable/shapes.h:25-58)
NR_D bool sphere_hit(const Ray& r, float tmin, float tmax, HRec& rec, float radius, const GXform& x) {
    f3 oc = -r.o;
    float a = len2(r.d);
    float h = dot(r.d, oc);
    float c = len2(oc) - radius * radius;
    float disc = h * h - a * c;
    if (disc < 0.0f) return false;
    float sq = sqrtf(disc);
    float root = (h - sq) / a;
    if (!(tmin < root && root < tmax)) {
        root = (h + sq) / a;
        if (!(tmin < root && root < tmax)) return false;
    }
    rec.t = root;
    rec.p = at(r, root);
    f3 nn = divv(rec.p, adds(ld3(x.scale) * radius, NR_FMIN));
    rec.n = nn;
    rec.tg = normalize(mk3(-rec.p.z, 0.0f, rec.p.x));
    float theta = nrm::acosf_(-nn.y);
    float phi = nrm::atan2f_(-nn.z, nn.x) + NR_PI;
    rec.uv.x = phi / NR_PI2;
    rec.uv.y = theta / NR_PI;
    return true;
}

// Cube::hit (hittable/shapes.h:142-173)
NR_D bool cube_hit(const GScene& S, const Ray& r, float tmin, float tmax, HRec& rec) {
    bool any = false;
    int face = -1;
    Ray face_ray = r;
    for (int i = 0; i < 6; i++) {
        const GXform& fx = S.cube_faces[i];
        Ray fr = to_object(r, fx);
        HRec fre;
        hrec_clear(fre);
        if (quad_hit(fr, tmin, tmax, fre)) {
            any = true; face = i; face_ray = fr; rec = fre; tmax = fre.t;
        }
    }
    if (any) to_world(rec, S.cube_faces[face], face_ray, r, false);
    return any;
}

// Möller–Trumbore (hittable/mesh.h:338-365)
NR_D bool tri_hit(const Ray& r, f3 v0, f3 v1, f3 v2, float tmin, float tmax, float& ot, float& ou, float& ov) {
    // Every test of the reference's early-out sequence, evaluated without
    // branches (a lane's later terms cannot change an earlier rejection).
    f3 e1 = v1 - v0, e2 = v2 - v0;
    f3 h = cross(r.d, e2);
    float a = dot(e1, h);
    float f = 1.0f / a;
    f3 s = r.o - v0;
    float u = f * dot(s, h);
    f3 q = cross(s, e1);
    float v = f * dot(r.d, q);
    float t = f * dot(e2, q);
    bool ok = !(fabsf(a) < NR_FMIN) && !(u < 0.0f || u > 1.0f) && !(v < 0.0f || u + v > 1.0f) &&
              (tmin <= t && t <= tmax);
    if (ok) { ot = t; ou = u; ov = v; }
    return ok;
}

// Address spaces of the traversal's memory.  Scene records hold generic
// pointers, and generic (flat) accesses wait on both the vector-memory and the
// LDS counters; typed pointers make the compiler emit ds_* for LDS and global_*
// for device memory, so a stack pop does not wait for node fetches in flight and
// vice versa.  The traversal loops also copy the table pointers they use into
// locals first: stack stores could otherwise alias the scene records and force
// a reload of each pointer on every step.
#define NR_LDS __attribute__((address_space(3)))
#define NR_GLB __attribute__((address_space(1)))
template <typename T>
NR_D const NR_GLB T* glb(const T* p) { return (const NR_GLB T*)p; }
// Clang vector types (HIP's float4 / int2 classes have no address-space-qualified
// member operators)
typedef float nr_f4 __attribute__((ext_vector_type(4)));
typedef int nr_i2 __attribute__((ext_vector_type(2)));
NR_D const NR_GLB nr_f4* glb_f4(const NR_GLB void* p) { return (const NR_GLB nr_f4*)p; }

// Traversal stack: the first lds_cap levels live in LDS (one column per lane,
// conflict-free), deeper levels spill to a per-thread column of global memory.
// Entries are (ref, entry t).  cap bounds the total depth.
struct Stack {
    NR_LDS nr_i2* lds;  // lds[level * 64 + lane]
    NR_GLB nr_i2* spill;// spill[(level - lds_cap) * spill_stride + slot]
    int lane;
    int lds_cap;
    int cap;
    size_t slot;
    size_t spill_stride;
    NR_D void push(int& sp, int32_t ref, float t) {
        nr_i2 e = {ref, __float_as_int(t)};
        if (sp < lds_cap) lds[sp * 64 + lane] = e;
        else if (sp < cap) spill[(size_t)(sp - lds_cap) * spill_stride + slot] = e;
        sp++;
    }
    NR_D void pop(int& sp, int32_t& ref, float& t) {
        sp--;
        nr_i2 e;
        if (sp < lds_cap) e = lds[sp * 64 + lane];
        else e = spill[(size_t)(sp - lds_cap) * spill_stride + slot];
        ref = e.x;
        t = __int_as_float(e.y);
    }
};

NR_D void load_node(const NR_GLB GNode* nodes, int32_t i, float* lmn, float* lmx, float* rmn, float* rmx, int32_t& l,
                    int32_t& r) {
    const NR_GLB nr_f4* p = glb_f4(nodes + i);
    const nr_f4 a = p[0];
    const nr_f4 b = p[1];
    const nr_f4 c = p[2];
    const nr_f4 d = p[3];
    lmn[0] = a.x; lmn[1] = a.y; lmn[2] = a.z; lmx[0] = a.w;
    lmx[1] = b.x; lmx[2] = b.y; rmn[0] = b.z; rmn[1] = b.w;
    rmn[2] = c.x; rmx[0] = c.y; rmx[1] = c.z; rmx[2] = c.w;
    l = __float_as_int(d.x); r = __float_as_int(d.y);
}

// Mesh::hit traversal (hittable/mesh.h:239-292) in near-first order with the
// child entry distance kept on the stack: popping (ref, t) and testing
// t < tmax equals the reference's re-test of the popped node's box, because
// tmin never changes and tmax only shrinks during a traversal.  The near child
// stays in registers (the reference pushes it and pops it straight back), only
// the far child goes to the stack.
// Wave scheduling: a lane that reaches a leaf holds it until NR_LEAF_BATCH
// lanes of the wave (or all lanes still working) hold one, then the wave runs
// the triangle test once for all of them.  Each lane still does its own steps
// in exactly the reference's order; only the interleaving across lanes changes.
// Returns the BVH slot of the closest triangle (or -1), its t, u, v; with ANY
// a lane stops at its first hit (occlusion queries).
// `overflow` is set if the stack would exceed its capacity.
#ifndef NR_LEAF_BATCH
#define NR_LEAF_BATCH 1
#endif
template <bool ANY, bool FIN>
NR_D int32_t mesh_closest_t(const GMesh& M, const Ray& r, float tmin, float tmax, Stack& st, int sp0,
                            float& bt, float& bu, float& bv, bool& overflow) {
    int32_t best = -1;
    int sp = sp0;
    if (sp0 + 1 > st.cap) { overflow = true; return -1; }
    int32_t ref = M.root_ref;
    float te = box_t(M.root_min, M.root_max, r, tmin, tmax);
    const NR_GLB GNode* const nodes = glb(M.nodes);
    const NR_GLB GTri* const tris = glb(M.tris);
    bool active = true;     // the lane still has work
    bool leaf = false;      // the lane holds a leaf to test
    while (true) {
        if (active && !leaf) {
            if (!(te < tmax)) {                        // culled: take the next entry
                if (sp > sp0) st.pop(sp, ref, te);
                else active = false;
            } else if (ref < 0) {
                leaf = true;
            } else {
                float lmn[3], lmx[3], rmn[3], rmx[3]; int32_t l, rr;
                load_node(nodes, ref, lmn, lmx, rmn, rmx, l, rr);
                float tl = FIN ? box_t_fin(lmn, lmx, r, tmin, tmax) : box_t(lmn, lmx, r, tmin, tmax);
                float tr = FIN ? box_t_fin(rmn, rmx, r, tmin, tmax) : box_t(rmn, rmx, r, tmin, tmax);
                bool left_near = tl < tr;
                int32_t near_i = left_near ? l : rr, far_i = left_near ? rr : l;
                float near_t = left_near ? tl : tr, far_t = left_near ? tr : tl;
                if (sp + 1 > st.cap) {
                    overflow = true;
                    active = false;
                } else {
                    if (far_t < tmax) st.push(sp, far_i, far_t);
                    ref = near_i;
                    te = near_t;
                }
            }
        }
        unsigned long long held = __ballot(leaf);
        unsigned long long walking = __ballot(active && !leaf);
        if (held != 0ull && (__popcll(held) >= NR_LEAF_BATCH || walking == 0ull)) {
            if (leaf) {
                int32_t slot = ~ref;
                const NR_GLB nr_f4* tp = glb_f4(tris + slot);
                const nr_f4 a = tp[0], b = tp[1], c = tp[2];
                float t, u, v;
                bool h = tri_hit(r, mk3(a.x, a.y, a.z), mk3(a.w, b.x, b.y), mk3(b.z, b.w, c.x), tmin, tmax, t, u, v);
                if (h) {
                    best = slot; tmax = t; bt = t; bu = u; bv = v;
                }
                leaf = false;
                if (ANY && h) active = false;
                else if (sp > sp0) st.pop(sp, ref, te);
                else active = false;
            }
        }
        if (__ballot(active) == 0ull) break;
    }
    return best;
}

// The binary traversal; a wave whose rays all have finite, nonzero inverse
// directions (every ray but axis-parallel ones) runs the min / max slab test.
template <bool ANY = false>
NR_D int32_t mesh_closest(const GMesh& M, const Ray& r, float tmin, float tmax, Stack& st, int sp0,
                          float& bt, float& bu, float& bv, bool& overflow) {
    if (__ballot(!inv_finite(r)) == 0ull) return mesh_closest_t<ANY, true>(M, r, tmin, tmax, st, sp0, bt, bu, bv, overflow);
    return mesh_closest_t<ANY, false>(M, r, tmin, tmax, st, sp0, bt, bu, bv, overflow);
}

// The same closest-hit query over the 4-wide layout (fast mode): one 128 B
// node per two binary levels, children entered nearest first (sorted by entry
// distance), the nearer ones pushed behind the farther.  The closest hit is the
// binary traversal's; only which of two triangles hit at exactly the same t
// is reported may differ.  Leaf batching as in mesh_closest.
template <bool ANY = false>
NR_D int32_t mesh_closest_w(const GMesh& M, const Ray& r, float tmin, float tmax, Stack& st, int sp0,
                            float& bt, float& bu, float& bv, bool& overflow) {
    int32_t best = -1;
    int sp = sp0;
    if (sp0 + 1 > st.cap) { overflow = true; return -1; }
    int32_t ref = M.root_ref;
    float te = box_t(M.root_min, M.root_max, r, tmin, tmax);
    const NR_GLB GWNode* const wnodes = glb(M.wnodes);
    const NR_GLB GTri* const tris = glb(M.tris);
    bool active = true;
    bool leaf = false;
    while (true) {
        if (active && !leaf) {
            if (!(te < tmax)) {
                if (sp > sp0) st.pop(sp, ref, te);
                else active = false;
            } else if (ref < 0) {
                leaf = true;
            } else {
                const NR_GLB nr_f4* p = glb_f4(wnodes + ref);
                const nr_f4 q0 = p[0], q1 = p[1], q2 = p[2], q3 = p[3], q4 = p[4], q5 = p[5], q6 = p[6];
                const float lx[4] = {q0.x, q0.y, q0.z, q0.w}, ly[4] = {q1.x, q1.y, q1.z, q1.w};
                const float lz[4] = {q2.x, q2.y, q2.z, q2.w}, hx[4] = {q3.x, q3.y, q3.z, q3.w};
                const float hy[4] = {q4.x, q4.y, q4.z, q4.w}, hz[4] = {q5.x, q5.y, q5.z, q5.w};
                int32_t c[4] = {__float_as_int(q6.x), __float_as_int(q6.y), __float_as_int(q6.z),
                                __float_as_int(q6.w)};
                float t[4];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const float mn[3] = {lx[k], ly[k], lz[k]}, mx[3] = {hx[k], hy[k], hz[k]};
                    t[k] = c[k] != NR_WIDE_EMPTY ? box_t(mn, mx, r, tmin, tmax) : NR_FMAX;
                }
                // sorting network over (t, ref), ascending t
#define NR_CSWAP(a, b)                                                         \
    if (t[b] < t[a]) {                                                         \
        float tt = t[a]; t[a] = t[b]; t[b] = tt;                               \
        int32_t cc = c[a]; c[a] = c[b]; c[b] = cc;                             \
    }
                NR_CSWAP(0, 1) NR_CSWAP(2, 3) NR_CSWAP(0, 2) NR_CSWAP(1, 3) NR_CSWAP(1, 2)
#undef NR_CSWAP
                const int nh = (t[0] < tmax) + (t[1] < tmax) + (t[2] < tmax) + (t[3] < tmax);
                if (nh == 0) {
                    if (sp > sp0) st.pop(sp, ref, te);
                    else active = false;
                } else if (sp + nh - 1 > st.cap) {
                    overflow = true;
                    active = false;
                } else {
                    if (nh > 3) st.push(sp, c[3], t[3]);
                    if (nh > 2) st.push(sp, c[2], t[2]);
                    if (nh > 1) st.push(sp, c[1], t[1]);
                    ref = c[0];
                    te = t[0];
                }
            }
        }
        unsigned long long held = __ballot(leaf);
        unsigned long long walking = __ballot(active && !leaf);
        if (held != 0ull && (__popcll(held) >= NR_LEAF_BATCH || walking == 0ull)) {
            if (leaf) {
                int32_t slot = ~ref;
                const NR_GLB nr_f4* tp = glb_f4(tris + slot);
                const nr_f4 a = tp[0], b = tp[1], cc = tp[2];
                float tt, u, v;
                bool h = tri_hit(r, mk3(a.x, a.y, a.z), mk3(a.w, b.x, b.y), mk3(b.z, b.w, cc.x), tmin, tmax, tt, u, v);
                if (h) {
                    best = slot; tmax = tt; bt = tt; bu = u; bv = v;
                }
                leaf = false;
                if (ANY && h) active = false;
                else if (sp > sp0) st.pop(sp, ref, te);
                else active = false;
            }
        }
        if (__ballot(active) == 0ull) break;
    }
    return best;
}

// The 4-wide traversal over quantised nodes (GQNode): one 64 B fetch per two
// binary levels instead of three 64 B nodes.  Leaf batching as in mesh_closest.
//
// Box test: the slab distance of plane code q on axis a is (origin + q s - o) inv
// = fma(q, s inv, (origin - o) inv), one FMA per plane after a few per-node terms;
// the entry plane is the lower one where inv >= 0, the upper one otherwise.  Each
// slab is widened by m = 2^-18 (|(origin - o) inv| + |origin inv| + 255 |s inv|),
// at least eight times the rounding of this form, of the decoded planes (which
// contain the exact child boxes) and of box_t's own (P - o) inv, so every box
// that box_t accepts on the exact child box is accepted here with an entry
// distance no larger: the closest hit is the binary traversal's (of two
// triangles hit at the same t either may be reported).  An axis with |inv| > 2^40
// or |o| > 2^80 does not constrain the box: its terms are NaN, which fmaxf and
// fminf (IEEE maxNum / minNum, v_max_f32 / v_min_f32) pass over, and every other
// term stays finite.  The one difference from the binary traversal: a ray lying
// exactly in a box's plane, with a zero direction component, makes box_t compute
// (P - o) * inf = NaN and, for +0, reject a box the ray touches; here that axis
// is left out, so such a ray may find a hit the binary traversal misses.
NR_D float qbyte(uint32_t w, int k) { return (float)((w >> (8 * k)) & 255u); }
template <bool ANY = false>
NR_D int32_t mesh_closest_q(const GMesh& M, const Ray& r, float tmin, float tmax, Stack& st, int sp0,
                            float& bt, float& bu, float& bv, bool& overflow) {
    int32_t best = -1;
    int sp = sp0;
    if (sp0 + 1 > st.cap) { overflow = true; return -1; }
    int32_t ref = M.root_ref;
    float te = box_t(M.root_min, M.root_max, r, tmin, tmax);
    const NR_GLB GQNode* const qnodes = glb(M.qnodes);
    const NR_GLB GTri* const tris = glb(M.tris);
    // per-ray terms of the box test
    float ic[3];
    bool neg[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float iv = comp(r.inv, a), o = comp(r.o, a);
        const bool ok = fabsf(iv) <= 0x1p40f && fabsf(o) <= 0x1p80f;
        ic[a] = ok ? iv : __builtin_nanf("");
        neg[a] = iv < 0.0f;
    }
    bool active = true;
    bool leaf = false;
    while (true) {
        if (active && !leaf) {
            if (!(te < tmax)) {
                if (sp > sp0) st.pop(sp, ref, te);
                else active = false;
            } else if (ref < 0) {
                leaf = true;
            } else {
                const NR_GLB nr_f4* p = glb_f4(qnodes + ref);
                const nr_f4 q0 = p[0], q1 = p[1], q2 = p[2], q3 = p[3];
                const uint32_t ex = __float_as_uint(q0.w);
                const float org[3] = {q0.x, q0.y, q0.z};
                const uint32_t wlo[3] = {__float_as_uint(q1.x), __float_as_uint(q1.y), __float_as_uint(q1.z)};
                const uint32_t whi[3] = {__float_as_uint(q1.w), __float_as_uint(q2.x), __float_as_uint(q2.y)};
                int32_t c[4] = {__float_as_int(q2.z), __float_as_int(q2.w), __float_as_int(q3.x), __float_as_int(q3.y)};
                float t[4];
                float Ae[3], Ax[3], B[3];
                uint32_t we[3], wx[3];
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    const float sc = __uint_as_float(((ex >> (8 * a)) & 255u) << 23);
                    const float A = (org[a] - comp(r.o, a)) * ic[a];
                    B[a] = sc * ic[a];
                    const float m =
                        __builtin_fmaf(fabsf(B[a]), 255.0f, fabsf(A) + fabsf(org[a] * ic[a])) * 0x1p-18f;
                    Ae[a] = A - m;
                    Ax[a] = A + m;
                    we[a] = neg[a] ? whi[a] : wlo[a];
                    wx[a] = neg[a] ? wlo[a] : whi[a];
                }
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const float e0 = __builtin_fmaf(qbyte(we[0], k), B[0], Ae[0]);
                    const float e1 = __builtin_fmaf(qbyte(we[1], k), B[1], Ae[1]);
                    const float e2 = __builtin_fmaf(qbyte(we[2], k), B[2], Ae[2]);
                    const float x0 = __builtin_fmaf(qbyte(wx[0], k), B[0], Ax[0]);
                    const float x1 = __builtin_fmaf(qbyte(wx[1], k), B[1], Ax[1]);
                    const float x2 = __builtin_fmaf(qbyte(wx[2], k), B[2], Ax[2]);
                    const float t0 = fmaxf(fmaxf(e0, e1), fmaxf(e2, tmin));
                    const float t1 = fminf(fminf(x0, x1), fminf(x2, tmax));
                    t[k] = (t0 <= t1 && c[k] != NR_WIDE_EMPTY) ? t0 : NR_FMAX;
                }
#define NR_CSWAP(a, b)                                                         \
    if (t[b] < t[a]) {                                                         \
        float tt = t[a]; t[a] = t[b]; t[b] = tt;                               \
        int32_t cc = c[a]; c[a] = c[b]; c[b] = cc;                             \
    }
                NR_CSWAP(0, 1) NR_CSWAP(2, 3) NR_CSWAP(0, 2) NR_CSWAP(1, 3) NR_CSWAP(1, 2)
#undef NR_CSWAP
                const int nh = (t[0] < tmax) + (t[1] < tmax) + (t[2] < tmax) + (t[3] < tmax);
                if (nh == 0) {
                    if (sp > sp0) st.pop(sp, ref, te);
                    else active = false;
                } else if (sp + nh - 1 > st.cap) {
                    overflow = true;
                    active = false;
                } else {
                    if (nh > 3) st.push(sp, c[3], t[3]);
                    if (nh > 2) st.push(sp, c[2], t[2]);
                    if (nh > 1) st.push(sp, c[1], t[1]);
                    ref = c[0];
                    te = t[0];
                }
            }
        }
        unsigned long long held = __ballot(leaf);
        unsigned long long walking = __ballot(active && !leaf);
        if (held != 0ull && (__popcll(held) >= NR_LEAF_BATCH || walking == 0ull)) {
            if (leaf) {
                int32_t slot = ~ref;
                const NR_GLB nr_f4* tp = glb_f4(tris + slot);
                const nr_f4 a = tp[0], b = tp[1], cc = tp[2];
                float tt, u, v;
                bool h = tri_hit(r, mk3(a.x, a.y, a.z), mk3(a.w, b.x, b.y), mk3(b.z, b.w, cc.x), tmin, tmax, tt, u, v);
                if (h) {
                    best = slot; tmax = tt; bt = tt; bu = u; bv = v;
                }
                leaf = false;
                if (ANY && h) active = false;
                else if (sp > sp0) st.pop(sp, ref, te);
                else active = false;
            }
        }
        if (__ballot(active) == 0ull) break;
    }
    return best;
}

// interpolate_hit (hittable/mesh.h:367-383)
NR_D void mesh_interp(const GMesh& M, int32_t slot, const Ray& r, float t, float u, float v, HRec& rec) {
    const uint32_t* ix = M.idx + 3 * (size_t)slot;
    const GVertex& a = M.verts[ix[0]];
    const GVertex& b = M.verts[ix[1]];
    const GVertex& c = M.verts[ix[2]];
    float w = 1.0f - u - v;
    rec.t = t;
    rec.p = at(r, t);
    rec.uv.x = a.uv[0] * w + b.uv[0] * u + c.uv[0] * v;
    rec.uv.y = a.uv[1] * w + b.uv[1] * u + c.uv[1] * v;
    rec.n = normalize((ld3(a.normal) * w + ld3(b.normal) * u) + ld3(c.normal) * v);
    rec.tg = normalize((ld3(a.tangent) * w + ld3(b.tangent) * u) + ld3(c.tangent) * v);
    rec.prim = M.tris[slot].orig;
}

NR_D bool mesh_hit(const GScene& S, int mesh, const Ray& r, float tmin, float tmax, HRec& rec,
                   Stack& st, int sp0, bool& overflow) {
    const GMesh& M = S.meshes[mesh];
    float t, u, v;
    int32_t slot = mesh_closest(M, r, tmin, tmax, st, sp0, t, u, v, overflow);
    if (slot < 0) return false;
    mesh_interp(M, slot, r, t, u, v, rec);
    return true;
}

// Hittable::hit of a primitive type (hittable/hittable.h:332-352)
NR_D bool prim_hit(const GScene& S, int type, const GObject& o, const Ray& r, float tmin, float tmax, HRec& rec,
                   Stack& st, int sp0, bool& overflow) {
    switch (type) {
        case 1: return quad_hit(r, tmin, tmax, rec);
        case 2: return sphere_hit(r, tmin, tmax, rec, o.radius, o.xf);
        case 3: return cube_hit(S, r, tmin, tmax, rec);
        case 4: return mesh_hit(S, o.mesh, r, tmin, tmax, rec, st, sp0, overflow);
    }
    return false;
}

// ConstantMedium::hit (hittable/volumes.h:14-45)
NR_D bool medium_hit(const GScene& S, const GObject& o, const Ray& r, float tmin, float tmax, HRec& rec,
                     Stack& st, int sp0, bool& overflow) {
    HRec r1, r2;
    hrec_clear(r1);
    hrec_clear(r2);
    if (!prim_hit(S, o.bound_type, o, r, -NR_FMAX, NR_FMAX, r1, st, sp0, overflow)) return false;
    if (!prim_hit(S, o.bound_type, o, r, r1.t + NR_EPS, NR_FMAX, r2, st, sp0, overflow)) return false;
    if (r1.t < tmin) r1.t = tmin;
    if (r2.t > tmax) r2.t = tmax;
    if (r1.t >= r2.t) return false;
    if (r1.t < 0.0f) r1.t = 0.0f;
    float ray_length = vlength(r.d);
    float inside = (r2.t - r1.t) * ray_length;
    float hit_distance = o.neg_inv_density * nrm::logf_(pcg_ray(r));
    if (hit_distance > inside) return false;
    rec.p = at(r, r1.t + hit_distance / ray_length);
    rec.n = mk3(1.0f, 0.0f, 0.0f);
    rec.front = true;
    return true;
}

// Object test of the top-level traversal (Hittable::hit_test, hittable.h:324-330):
// the world-space distance of the
// object's closest hit (Hittable::hit_test + HitRecord::to_world_space), and
// for plain meshes the triangle slot, barycentrics and object-space t so the
// full record can be rebuilt without a second traversal.
// Analytic objects, media and lights: out of line so their uniform loads (cube
// face transforms, medium parameters) are not hoisted into the registers of the
// mesh traversal loop.  Fills rec in world space.
#ifndef NR_GENERIC_INLINE
#define NR_GENERIC_INLINE 1
#endif
#if NR_GENERIC_INLINE
#define NR_GENERIC_ATTR __forceinline__
#else
#define NR_GENERIC_ATTR __attribute__((noinline))
#endif
__device__ NR_GENERIC_ATTR bool object_hit_generic(const GScene& S, const GObject& o, const Ray& r,
                                                             HRec& rec, Stack& st, int sp0, bool& overflow) {
    Ray orr = to_object(r, o.xf);
    hrec_clear(rec);
    bool h;
    if (o.type == 6)
        h = prim_hit(S, o.bound_type, o, orr, -NR_FMAX, NR_FMAX, rec, st, sp0, overflow);
    else if (o.type == 5)
        h = medium_hit(S, o, orr, NR_EPS, NR_FMAX, rec, st, sp0, overflow);
    else
        h = prim_hit(S, o.type, o, orr, NR_EPS, NR_FMAX, rec, st, sp0, overflow);
    if (h) to_world(rec, o.xf, orr, r, true);
    return h;
}

// Object types a traversal kernel is compiled for (a scene picks the smallest
// set that covers it; the code of absent types, and its registers, drop out).
enum : int { NR_TY_QUAD = 1, NR_TY_SPHERE = 2, NR_TY_MESH = 4, NR_TY_OTHER = 8, NR_TY_ALL = 15 };

// Mesh-tree layouts a traversal can use: the reference's binary tree (its visit
// order, what the reference's estimator needs), the 4-wide tree, the quantised
// 4-wide tree (fast mode).
enum : int { NR_LAY_BINARY = 0, NR_LAY_WIDE = 1, NR_LAY_QUANT = 2 };

template <int TY, bool ANY = false, int LAY = NR_LAY_BINARY>
NR_D bool object_t(const GScene& S, const GObject& o, const Ray& r, Stack& st, int sp0, bool& overflow,
                   float& tw, int32_t& slot, float& u, float& v, float& tobj) {
    HRec rec;
    bool h = false;
    slot = -1;
    const bool inl = ((TY & NR_TY_MESH) && o.type == 4) || ((TY & NR_TY_QUAD) && o.type == 1) ||
                     ((TY & NR_TY_SPHERE) && o.type == 2);
    if (inl) {   // mesh, quad, sphere inline
        Ray orr = to_object(r, o.xf);
        hrec_clear(rec);
        if ((TY & NR_TY_MESH) && o.type == 4) {
            const GMesh& M = S.meshes[o.mesh];
            slot = (LAY == NR_LAY_QUANT && M.qnodes) ? mesh_closest_q<ANY>(M, orr, NR_EPS, NR_FMAX, st, sp0, tobj, u, v, overflow)
                   : (LAY == NR_LAY_WIDE && M.wnodes) ? mesh_closest_w<ANY>(M, orr, NR_EPS, NR_FMAX, st, sp0, tobj, u, v, overflow)
                                                      : LAY == NR_LAY_BINARY ? mesh_closest<ANY>(M, orr, NR_EPS, NR_FMAX, st, sp0, tobj, u, v, overflow)
                   : mesh_closest_t<ANY, false>(M, orr, NR_EPS, NR_FMAX, st, sp0, tobj, u, v, overflow);
            h = slot >= 0;
            if (h) mesh_interp(M, slot, orr, tobj, u, v, rec);
        } else if ((TY & NR_TY_QUAD) && o.type == 1) {
            h = quad_hit(orr, NR_EPS, NR_FMAX, rec);
        } else if (TY & NR_TY_SPHERE) {
            h = sphere_hit(orr, NR_EPS, NR_FMAX, rec, o.radius, o.xf);
        }
        if (h) to_world(rec, o.xf, orr, r, true);
    } else if (TY & NR_TY_OTHER) {
        h = object_hit_generic(S, o, r, rec, st, sp0, overflow);
    }
    if (!h) return false;
    tw = rec.t;
    return true;
}

// SceneGraph::hit (scene/include/graph.h:16-66).  Returns the object table
// index of the closest hit or -1; rec receives the world-space record.  Only
// (object, t, triangle slot, u, v) stay live during the traversal; the record
// of the winner is rebuilt at the end by the same deterministic arithmetic.
// LAY: the mesh-tree layout traversed (NR_LAY_*)
template <int TY, int LAY = NR_LAY_BINARY>
NR_D int32_t scene_closest(const GScene& S, const Ray& r, HRec& rec, Stack& st, bool& overflow) {
    float tmin = NR_EPS, tmax = NR_FMAX;
    int32_t best = -1, bslot = -1;
    float bu = 0.0f, bv = 0.0f, btobj = 0.0f;
    int sp = 0;
    if (st.cap < 1 || S.n_objects == 0) return -1;
    const NR_GLB GNode* const top = glb(S.nodes);
    st.push(sp, S.root_ref, box_t(S.root_min, S.root_max, r, tmin, tmax));
    while (sp > 0) {
        int32_t ref; float te;
        st.pop(sp, ref, te);
        if (!(te < tmax)) continue;
        if (ref < 0) {
            int32_t oi = ~ref;
            float tw, u, v, tobj;
            int32_t slot;
            if (object_t<TY, false, LAY>(S, S.objects[oi], r, st, sp, overflow, tw, slot, u, v, tobj)) {
                if (tmin < tw && tw < tmax) {
                    best = oi; bslot = slot; bu = u; bv = v; btobj = tobj;
                    tmax = tw;
                }
            }
            if (overflow) break;
        } else {
            float lmn[3], lmx[3], rmn[3], rmx[3]; int32_t l, rr;
            load_node(top, ref, lmn, lmx, rmn, rmx, l, rr);
            float tl = box_t(lmn, lmx, r, tmin, tmax);
            float tr = box_t(rmn, rmx, r, tmin, tmax);
            bool left_near = tl < tr;
            int32_t near_i = left_near ? l : rr, far_i = left_near ? rr : l;
            float near_t = left_near ? tl : tr, far_t = left_near ? tr : tl;
            if (sp + 2 > st.cap) { overflow = true; break; }
            if (far_t < tmax) st.push(sp, far_i, far_t);
            st.push(sp, near_i, near_t);
        }
    }
    if (best >= 0 && !overflow) {
        const GObject& o = S.objects[best];
        if ((TY & NR_TY_MESH) && o.type == 4) {
            Ray orr = to_object(r, o.xf);
            hrec_clear(rec);
            mesh_interp(S.meshes[o.mesh], bslot, orr, btobj, bu, bv, rec);
            to_world(rec, o.xf, orr, r, true);
        } else if (((TY & NR_TY_QUAD) && o.type == 1) || ((TY & NR_TY_SPHERE) && o.type == 2)) {
            Ray orr = to_object(r, o.xf);
            hrec_clear(rec);
            bool h = o.type == 1 ? quad_hit(orr, NR_EPS, NR_FMAX, rec)
                                 : sphere_hit(orr, NR_EPS, NR_FMAX, rec, o.radius, o.xf);
            if (h) to_world(rec, o.xf, orr, r, true);
        } else if (TY & NR_TY_OTHER) {
            object_hit_generic(S, o, r, rec, st, 0, overflow);
        }
    }
    return best;
}

// Occlusion query with the visibility of scene_closest: true iff scene_closest
// would return an object for r (some object's hit lies in (NR_EPS, inf)).  A
// lane stops at the first such hit; mesh leaves stop at their first triangle.
template <int TY, int LAY = NR_LAY_WIDE>
NR_D bool scene_occluded(const GScene& S, const Ray& r, Stack& st, bool& overflow) {
    const float tmin = NR_EPS, tmax = NR_FMAX;
    int sp = 0;
    if (st.cap < 1 || S.n_objects == 0) return false;
    const NR_GLB GNode* const top = glb(S.nodes);
    st.push(sp, S.root_ref, box_t(S.root_min, S.root_max, r, tmin, tmax));
    while (sp > 0) {
        int32_t ref; float te;
        st.pop(sp, ref, te);
        if (!(te < tmax)) continue;
        if (ref < 0) {
            float tw, u, v, tobj;
            int32_t slot;
            if (object_t<TY, true, LAY>(S, S.objects[~ref], r, st, sp, overflow, tw, slot, u, v, tobj) && tmin < tw &&
                tw < tmax)
                return true;
            if (overflow) return false;
        } else {
            float lmn[3], lmx[3], rmn[3], rmx[3]; int32_t l, rr;
            load_node(top, ref, lmn, lmx, rmn, rmx, l, rr);
            float tl = box_t(lmn, lmx, r, tmin, tmax);
            float tr = box_t(rmn, rmx, r, tmin, tmax);
            bool left_near = tl < tr;
            int32_t near_i = left_near ? l : rr, far_i = left_near ? rr : l;
            float near_t = left_near ? tl : tr, far_t = left_near ? tr : tl;
            if (sp + 2 > st.cap) { overflow = true; return false; }
            if (far_t < tmax) st.push(sp, far_i, far_t);
            st.push(sp, near_i, near_t);
        }
    }
    return false;
}

// ---------------------------------------------------------------------------
// textures (material/texture.h:49-91)
// ---------------------------------------------------------------------------
NR_D f3 tex_sample(const GScene& S, const GTex& t, f2 uv) {
    if (t.image < 0) return ld3(t.color);
    const GImage& im = S.images[t.image];
    float sc = 1.0f / fmaxf(NR_FMIN, t.scale);
    float su = fmodf(uv.x * sc, 1.0f), sv = fmodf(uv.y * sc, 1.0f);
    float u = su < 0.0f ? 0.0f : (su > 1.0f ? 1.0f : su);
    float v = 1.0f - (sv < 0.0f ? 0.0f : (sv > 1.0f ? 1.0f : sv));
    int i = (int)(u * (float)im.W); if (i >= im.W) i = im.W - 1;
    int j = (int)(v * (float)im.H); if (j >= im.H) j = im.H - 1;
    size_t px = (size_t)j * im.W + i, hw = (size_t)im.H * im.W;
    f3 c = mk3((float)im.data[px], (float)im.data[hw + px], (float)im.data[2 * hw + px]);
    return divs(c, 255.0f);
}

// ---------------------------------------------------------------------------
// sky (light/include/skylight.h:33-122)
// ---------------------------------------------------------------------------
template <bool F = false>
NR_D f3 sky_sample(const GSky& sk, f3 d) {
    if (sk.type == 1) {
        float a = (float)(0.5 * ((double)d.y + 1.0));
        float oma = (float)(1.0 - (double)a);
        return ld3(sk.start) * oma + ld3(sk.end) * a;
    }
    if (sk.type == 2) {
        if (!sk.data) return s3(0.0f);
        f3 dir = mv(sk.rot, normalize(d));
        float u = Tm<F>::atan2(dir.z, dir.x) / NR_PI2 + 0.5f;
        float v = Tm<F>::acos(fminf(fmaxf(dir.y, -1.0f), 1.0f)) / NR_PI;
        int W = sk.W, H = sk.H;
        float fx = u * (float)W - 0.5f;
        float fy = v * (float)H - 0.5f;
        int x0 = (int)floorf(fx), y0 = (int)floorf(fy);
        int x1 = x0 + 1, y1 = y0 + 1;
        float tx = fx - (float)x0, ty = fy - (float)y0;
        x0 = ((x0 % W) + W) % W;
        x1 = ((x1 % W) + W) % W;
        y0 = (int)fmaxf(0.0f, fminf((float)(H - 1), (float)y0));
        y1 = (int)fmaxf(0.0f, fminf((float)(H - 1), (float)y1));
        size_t hw = (size_t)H * W;
        size_t i00 = (size_t)y0 * W + x0, i10 = (size_t)y0 * W + x1, i01 = (size_t)y1 * W + x0, i11 = (size_t)y1 * W + x1;
        f3 c00, c10, c01, c11;
        if (sk.rgbx) {                                  // the same texel values, interleaved
            const float4* q = reinterpret_cast<const float4*>(sk.rgbx);
            float4 a = q[i00], b = q[i10], c = q[i01], d = q[i11];
            c00 = mk3(a.x, a.y, a.z); c10 = mk3(b.x, b.y, b.z); c01 = mk3(c.x, c.y, c.z); c11 = mk3(d.x, d.y, d.z);
        } else {
            c00 = mk3(sk.data[i00], sk.data[hw + i00], sk.data[2 * hw + i00]);
            c10 = mk3(sk.data[i10], sk.data[hw + i10], sk.data[2 * hw + i10]);
            c01 = mk3(sk.data[i01], sk.data[hw + i01], sk.data[2 * hw + i01]);
            c11 = mk3(sk.data[i11], sk.data[hw + i11], sk.data[2 * hw + i11]);
        }
        f3 top = c00 * (1.0f - tx) + c10 * tx;
        f3 bot = c01 * (1.0f - tx) + c11 * tx;
        return (top * (1.0f - ty) + bot * ty) * sk.intensity;
    }
    return s3(0.0f);
}

// ---------------------------------------------------------------------------
// PDFs (engine/include/engine/pdf.h)
// ---------------------------------------------------------------------------
NR_D float ggx_D(float n_dot_h, float alpha) {                                // pdf.h:100-104
    float a2 = alpha * alpha;
    float d = n_dot_h * n_dot_h * (a2 - 1.0f) + 1.0f;
    return a2 / fmaxf(NR_PI * d * d, 1e-8f);
}
NR_D float smith_g1(float nx, float alpha) {                                   // pdf.h:106-111
    float a2 = alpha * alpha;
    return 2.0f * nx / (nx + sqrtf(a2 + (1.0f - a2) * nx * nx));
}
NR_D float pbr_pdf_value(const ONB& uvw, f3 view, float alpha, float sp, f3 d) {   // pdf.h:135-148
    float diffuse_pdf = fmaxf(0.0f, dot(normalize(d), uvw.w) / NR_PI);
    f3 h = normalize(view + d);
    float n_dot_h = fmaxf(dot(uvw.w, h), 0.0f);
    float v_dot_h = fmaxf(dot(view, h), 1e-4f);
    float spec_pdf = (ggx_D(n_dot_h, alpha) * n_dot_h) / (4.0f * v_dot_h);
    return (1.0f - sp) * diffuse_pdf + sp * spec_pdf;
}
template <bool F = false>
NR_D f3 pbr_pdf_generate(const ONB& uvw, f3 view, float alpha, float sp, Gen& g) { // pdf.h:150-166
    if (g.rf() < sp) {
        float u1 = g.rf();
        float u2 = g.rf();
        float cos_theta = sqrtf((1.0f - u1) / (1.0f + (alpha * alpha - 1.0f) * u1));
        float sin_theta = sqrtf(fmaxf(0.0f, 1.0f - cos_theta * cos_theta));
        float phi = NR_PI2 * u2;
        f3 h = uvw.xf(mk3(sin_theta * Tm<F>::cos(phi), sin_theta * Tm<F>::sin(phi), cos_theta));
        return normalize(h * (2.0f * dot(view, h)) - view);
    }
    return uvw.xf(cosine_dir<F>(g));
}
NR_D float cosine_pdf_value(const ONB& uvw, f3 d) {                           // pdf.h:26-29
    float c = dot(normalize(d), uvw.w);
    return fmaxf(0.0f, c / NR_PI);
}

// ---------------------------------------------------------------------------
// light sampling of the first ObjectLight (hittable.h:356-386)
// ---------------------------------------------------------------------------
NR_D float quad_light_pdf(const GObject& o, f3 origin, f3 dir) {              // primitives.h:34-49
    HRec rec;
    hrec_clear(rec);
    Ray wr = mkray(origin, dir);
    Ray orr = to_object(wr, o.xf);
    bool h = quad_hit(orr, NR_EPS, NR_FMAX, rec);
    to_world(rec, o.xf, orr, wr, true);
    if (!h) return 0.0f;
    float dist2 = rec.t * rec.t * len2(dir);
    float cosine = fabsf(dot(dir, rec.n) / vlength(dir));
    if (cosine < 1e-8f) return 0.0f;
    f3 n = quad_n();
    float world_area = vlength(n) * o.xf.scale[0] * o.xf.scale[2];
    return dist2 / (cosine * world_area);
}
NR_D f3 quad_light_random(const GObject& o, f3 origin, Gen& g) {              // primitives.h:52-61
    f3 corner = quad_u() * -0.5f - quad_v() * 0.5f;
    float r1 = g.rf();
    float r2 = g.rf();
    f3 p = (corner + quad_u() * r1) + quad_v() * r2;
    p = mv(o.xf.R, p * ld3(o.xf.scale)) + ld3(o.xf.pos);
    return p - origin;
}
NR_D float sphere_light_pdf(const GObject& o, f3 origin, f3 dir) {            // shapes.h:60-76
    HRec rec;
    hrec_clear(rec);
    if (!sphere_hit(mkray(origin, dir), NR_EPS, NR_FMAX, rec, o.radius, o.xf)) return 0.0f;
    float rad_sq = o.radius * o.radius;
    float dist2 = len2(ld3(o.xf.pos) - origin);
    float cos_theta_max = sqrtf(1.0f - rad_sq / dist2);
    float solid_angle = NR_PI2 * (1.0f - cos_theta_max);
    return 1.0f / solid_angle;
}
template <bool F = false>
NR_D f3 sphere_light_random(const GObject& o, f3 origin, Gen& g) {            // shapes.h:78-108
    f3 direction = ld3(o.xf.pos) - origin;
    float dist2 = len2(direction);
    ONB uvw;
    uvw.build(direction);
    float r1 = g.rf();
    float r2 = g.rf();
    float rad_sq = o.radius * o.radius;
    float z = 1.0f + r2 * (sqrtf(1.0f - rad_sq / dist2) - 1.0f);
    float phi = NR_PI2 * r1;
    float x = Tm<F>::cos(phi) * sqrtf(1.0f - z * z);
    float y = Tm<F>::sin(phi) * sqrtf(1.0f - z * z);
    return uvw.xf(mk3(x, y, z));
}
NR_D int light_geom(const GObject& L) {
    return L.type == 6 ? L.bound_type : (L.type == 5 ? -1 : L.type);
}
NR_D float light_pdf(const GObject& L, f3 origin, f3 dir) {
    int t = light_geom(L);
    if (t == 1) return quad_light_pdf(L, origin, dir);
    if (t == 2) return sphere_light_pdf(L, origin, dir);
    return 0.0f;
}
template <bool F = false>
NR_D f3 light_random(const GObject& L, f3 origin, Gen& g) {
    int t = light_geom(L);
    if (t == 1) return quad_light_random(L, origin, g);
    if (t == 2) return sphere_light_random<F>(L, origin, g);
    if (t == 4) return s3(0.0f);
    return mk3(1.0f, 0.0f, 0.0f);
}

// PBR::get_shading_normal (material/material.h:113-125)
NR_D f3 pbr_shading_normal(const GScene& S, const GMaterial& m, f3 n, f3 tg, f2 uv) {
    f3 tn = tex_sample(S, m.tex[4], uv);
    f3 map_n = normalize(mk3(tn.x * 2.0f - 1.0f, tn.y * 2.0f - 1.0f, tn.z * 2.0f - 1.0f));
    f3 t = normalize(tg - n * dot(tg, n));
    f3 b = cross(n, t);
    return normalize((t * map_n.x + b * map_n.y) + n * map_n.z);
}

// ---------------------------------------------------------------------------
// viewport preview (engine/src/trace.cu:99-143, engine/src/axis_grid.cu) and
// the screen-space rays of picking and selection (engine/camera.h:181-184, 215-226)
// ---------------------------------------------------------------------------
NR_HD float smoothstep_f(float a, float b, float x) {                          // core/cg_std.h:97-103
    float t = fmaxf(0.0f, fminf(1.0f, (x - a) / (b - a)));
    return t * t * (3.0f - (2.0f * t));
}
NR_HD f3 lerp3(f3 x, f3 y, float w) { return x + (y - x) * w; }               // core/cg_std.h:90-93

NR_HD float grid_mask(float u, float v, float spacing, float width) {          // axis_grid.cu:4-10
    float du = fabsf(fmodf(u + spacing * 0.5f, spacing) - spacing * 0.5f);
    float dv = fabsf(fmodf(v + spacing * 0.5f, spacing) - spacing * 0.5f);
    float d = fminf(du, dv);
    return 1.0f - smoothstep_f(0.0f, width, d);
}

// axis_grid (axis_grid.cu:12-51) with max_t = FMAX and line width 1e-3; the
// grid level's powf(10, level0) is 10^level0 (level0 is a whole number >= 0)
NR_HD bool axis_grid(const Ray& r, int axis, f3 major, f3& grid_color, float& grid_alpha) {
    float dir_axis = comp(r.d, axis);
    if (fabsf(dir_axis) < 1e-6f) return false;
    float t = -comp(r.o, axis) / dir_axis;
    if (t <= 0.0f || t >= NR_FMAX) return false;
    f3 p = at(r, t);
    p = mk3(fabsf(p.x), fabsf(p.y), fabsf(p.z));
    float u = comp(p, (axis + 1) % 3), v = comp(p, (axis + 2) % 3);
    float pixel_size = 1e-3f * t;
    float level = fmaxf(nrm::log10f_pos(fmaxf(pixel_size, 1e-5f)) + 1.0f, 0.0f);
    float level0 = floorf(level);
    float blend = level - level0;
    float spacing0 = nrm::pow10f_int((int)level0);
    float spacing1 = spacing0 * 10.0f;
    float width = pixel_size * 1.5f;
    float g0 = grid_mask(u, v, spacing0, width);
    float g1 = grid_mask(u, v, spacing1, width);
    float mask = fmaxf(g0 * (1.0f - blend), g1);
    if (mask < 0.01f) return false;
    float fade = fmaxf(0.0f, 1.0f - t / 100.0f);
    grid_color = lerp3(s3(0.35f), major, smoothstep_f(0.0f, 1.0f, g1));
    grid_alpha = mask * fade;
    return true;
}

NR_HD bool triplanar_grid(const Ray& r, f3& grid_color, float& grid_alpha) {  // axis_grid.cu:53-70
    return axis_grid(r, 0, mk3(1.0f, 0.3f, 0.3f), grid_color, grid_alpha) ||
           axis_grid(r, 1, mk3(0.3f, 1.0f, 0.3f), grid_color, grid_alpha) ||
           axis_grid(r, 2, mk3(0.3f, 0.3f, 1.0f), grid_color, grid_alpha);
}

// Material::viewport_color (material/material.h:43-45, 109-111, 218, 289-291, 335-337, 588-596)
NR_D f3 viewport_color(const GScene& S, const GMaterial& m, f2 uv) {
    switch (m.type) {
        case 1: case 2: case 4: case 5: return tex_sample(S, m.tex[0], uv);
        case 3: return ld3(m.tint);
        default: return s3(0.0f);
    }
}

// DeviceCamera::screen_space_ray (camera.h:215-226)
NR_HD Ray screen_space_ray(const GCamera& c, float su, float sv) {
    float x = su * c.res[0];
    float y = sv * c.res[1];
    float u = (x - (c.res[0] - 1.0f) * 0.5f) * (c.uvw[0] / c.res[0]);
    float v = (y - (c.res[1] - 1.0f) * 0.5f) * (c.uvw[1] / c.res[1]);
    f3 d = normalize(mv(c.R, mk3(u, v, -c.focal_length)));
    return mkray(ld3(c.pos), d);
}

// ---------------------------------------------------------------------------
// HDRI importance sampling (fast mode, with fast mode's float transcendentals).
// The sky lookup maps a direction d to
// u = atan2(z, x) / 2pi + 0.5, v = acos(y) / pi after the sky rotation R; the
// density is piecewise constant over texels in (u, v), so the solid-angle pdf is
// p(u, v) / (2 pi^2 sin(pi v)).
// ---------------------------------------------------------------------------
NR_D int cdf_find(const float* cdf, int n, float x) {          // last k with cdf[k] <= x, k in [0, n)
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (cdf[mid] <= x) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// direction (world), its solid-angle pdf and radiance; marg may be an LDS copy
NR_D f3 sky_cdf_sample(const GSky& sk, const float* marg, float u1, float u2, float& pdf) {
    const int H = sk.H, W = sk.W;
    const int j = cdf_find(marg, H, u1);
    const float m0 = marg[j], m1 = marg[j + 1];
    const float dv = m1 > m0 ? (u1 - m0) / (m1 - m0) : 0.5f;
    // texel of the row by its alias table: u2 * W picks a column and a fraction,
    // the fraction decides between the column and its alias and, rescaled,
    // places the point inside the chosen texel
    const float4* row = reinterpret_cast<const float4*>(sk.alias) + (size_t)j * W;
    const float x = u2 * (float)W;
    int i = min((int)x, W - 1);
    const float f = fminf(x - (float)i, 1.0f);
    const float4 e = row[i];
    float du, pt;
    if (f < e.x || e.x >= 1.0f) {
        du = e.x > 0.0f ? f / e.x : 0.5f;
        pt = e.z;
    } else {
        i = min(max(__float_as_int(e.y), 0), W - 1);
        du = (f - e.x) / (1.0f - e.x);
        pt = row[i].z;
    }
    const float u = ((float)i + fminf(fmaxf(du, 0.0f), 1.0f)) / (float)W;
    const float v = ((float)j + fminf(fmaxf(dv, 0.0f), 1.0f)) / (float)H;
    const float theta = v * NR_PI, phi = (u - 0.5f) * NR_PI2;
    const float st = Tm<true>::sin(theta), ct = Tm<true>::cos(theta);
    const f3 local = mk3(st * Tm<true>::cos(phi), ct, st * Tm<true>::sin(phi));
    const float p_uv = (m1 - m0) * pt * (float)(W * H);
    pdf = st > 1e-7f ? p_uv / (2.0f * NR_PI * NR_PI * st) : 0.0f;
    return mtv(sk.rot, local);
}

NR_D float sky_cdf_pdf(const GSky& sk, f3 d) {
    const int H = sk.H, W = sk.W;
    f3 dir = mv(sk.rot, normalize(d));
    float u = Tm<true>::atan2(dir.z, dir.x) / NR_PI2 + 0.5f;
    float v = Tm<true>::acos(fminf(fmaxf(dir.y, -1.0f), 1.0f)) / NR_PI;
    int i = min(max((int)(u * (float)W), 0), W - 1);
    int j = min(max((int)(v * (float)H), 0), H - 1);
    const float pt = reinterpret_cast<const float4*>(sk.alias)[(size_t)j * W + i].z;
    const float p_uv = (sk.marg[j + 1] - sk.marg[j]) * pt * (float)(W * H);
    const float st = Tm<true>::sin(v * NR_PI);
    return st > 1e-7f ? p_uv / (2.0f * NR_PI * NR_PI * st) : 0.0f;
}

NR_HD float power_heuristic(float a, float b) { return a * a / (a * a + b * b); }
