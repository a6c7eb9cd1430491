/*
 * nr_oracle.cpp — CPU oracle for the NectarRender path-tracing hot path.
 *
 * TEST INFRASTRUCTURE (see nr_oracle.h): a C++17 restatement of the reference
 * routines listed below, in IEEE fp32 with fp contraction disabled
 * (-ffp-contract=off), used by tests as the parity checker and by bench.py as
 * the timed CPU baseline.  Every function cites the reference file:line it
 * follows; paths are relative to nectar_render/csrc/ of the reference.
 *
 * Restated deliberately as the reference has it (parity mode), including its
 * defects: the FMIN-biased vector division (core/vector.h:82-92), the
 * n-1 rendered samples (engine/src/engine.cpp:43), the 1/(4*pi^2) "sphere" PDF
 * (engine/pdf.h:11-13 with constants.h:13), the unnormalised light direction
 * of Quad::random (hittable/primitives.h:62-71), Sphere::pdf_value ignoring the
 * transform (hittable/shapes.h:60-76), the Reinhard no-op (data/src/stream.cu:7-11).
 *
 * Two assumptions the reference's sources leave open (no run of the reference
 * is possible here): device-side function arguments are evaluated left to
 * right (core/vector.h:303-311, 361-371; hittable/primitives.h:62-71), and the
 * Philox stream is rocRAND's rocrand_init(seed, subsequence, offset) +
 * rocrand_uniform, the open-source equivalent of the cuRAND calls at
 * core/random.h:50-62.  Host code follows g++ (the reference's host compiler):
 * constructor arguments right to left (camera.h:55-58).
 *
 * Parity unpinned: the float bits of the reference's CUDA transcendentals under
 * --use_fast_math (CMakeLists.txt:110-112); this file uses glibc libm.
 */
#include "nr_oracle.h"
#include "../nectarrender_amd/csrc/nr_mathfn.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <random>
#include <stdexcept>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace oro {

// core/include/core/constants.h:7-13
constexpr float EPS = 1e-4f;
constexpr float FMIN = 1e-8f;
constexpr float FMAX = 1e32f;
constexpr float PI = 3.1415926535897932385f;
constexpr float PI2 = PI * 2.0f;
constexpr float PI4 = PI2 * PI2;

// ---------------------------------------------------------------------------
// vectors (core/include/core/vector.h)
// ---------------------------------------------------------------------------
struct V3 { float x, y, z; };
struct V2 { float x, y; };

static inline V3 v3(float a) { return {a, a, a}; }
static inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }      // vector.h:58-64
static inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }      // vector.h:66-72
static inline V3 mul(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }      // vector.h:74-80
static inline V3 muls(V3 a, float t) { return {a.x * t, a.y * t, a.z * t}; }        // vector.h:233-242
static inline V3 adds(V3 a, float t) { return {a.x + t, a.y + t, a.z + t}; }        // vector.h:197-204
static inline V3 subs(V3 a, float t) { return {a.x - t, a.y - t, a.z - t}; }        // vector.h:215-222
static inline V3 neg(V3 a) { return {-a.x, -a.y, -a.z}; }                           // vector.h:50-56
// operator/(T, float) multiplies by the reciprocal (vector.h:251-254)
static inline V3 divs(V3 a, float t) { return muls(a, 1.0f / t); }
// operator/(T, T) adds FMIN to the divisor (vector.h:82-88, 245-250)
static inline V3 divv(V3 a, V3 b) { return {a.x / (b.x + FMIN), a.y / (b.y + FMIN), a.z / (b.z + FMIN)}; }
static inline float dot(V3 u, V3 v) { return u.x * v.x + u.y * v.y + u.z * v.z; }   // vector.h:455-459
static inline V3 cross(V3 u, V3 v) {                                                // vector.h:461-467
    return {u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x};
}
static inline float len2(V3 v) { return v.x * v.x + v.y * v.y + v.z * v.z; }       // vector.h:100-108
static inline float length(V3 v) { return sqrtf(len2(v)); }
static inline V3 normalize(V3 v) { return divs(v, length(v)); }                     // vector.h:268-269
static inline bool near_zero(V3 v) { return length(v) < 1e-8f; }                    // vector.h:110-112
static inline float get(V3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
static inline V3 reflect(V3 v, V3 n) { return sub(v, muls(n, 2.0f * dot(v, n))); }  // vector.h:469-474
static inline V3 refract(V3 uv, V3 n, float ior) {                                  // vector.h:476-485
    float cos_theta = fminf(dot(neg(uv), n), 1.0f);
    V3 perp = muls(add(uv, muls(n, cos_theta)), ior);
    float k = (float)(1.0 - (double)len2(perp));
    V3 parallel = muls(n, -sqrtf(fabsf(k)));
    return add(perp, parallel);
}

// ---------------------------------------------------------------------------
// Matrix3 (core/include/core/matrix.h)
// ---------------------------------------------------------------------------
struct M3 { float m[3][3]; };
static inline M3 ident() { M3 r{}; r.m[0][0] = r.m[1][1] = r.m[2][2] = 1.0f; return r; }
static inline V3 mv(const M3& a, V3 v) {                                            // matrix.h:16-22
    return {a.m[0][0] * v.x + a.m[0][1] * v.y + a.m[0][2] * v.z,
            a.m[1][0] * v.x + a.m[1][1] * v.y + a.m[1][2] * v.z,
            a.m[2][0] * v.x + a.m[2][1] * v.y + a.m[2][2] * v.z};
}
static inline M3 mm(const M3& a, const M3& b) {                                     // matrix.h:24-33
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            float acc = 0.0f;
            for (int k = 0; k < 3; k++) acc += a.m[i][k] * b.m[k][j];
            r.m[i][j] = acc;
        }
    return r;
}
static inline M3 transpose(const M3& a) {
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.m[i][j] = a.m[j][i];
    return r;
}
static M3 rot_x(float a) { float c = cosf(a), s = sinf(a); M3 r = ident();          // matrix.h:64-72
    r.m[1][1] = c; r.m[1][2] = -s; r.m[2][1] = s; r.m[2][2] = c; return r; }
static M3 rot_y(float a) { float c = cosf(a), s = sinf(a); M3 r = ident();          // matrix.h:74-82
    r.m[0][0] = c; r.m[0][2] = s; r.m[2][0] = -s; r.m[2][2] = c; return r; }
static M3 rot_z(float a) { float c = cosf(a), s = sinf(a); M3 r = ident();          // matrix.h:84-92
    r.m[0][0] = c; r.m[0][1] = -s; r.m[1][0] = s; r.m[1][1] = c; return r; }
static M3 rotation_from_euler(V3 r) { return mm(mm(rot_z(r.z), rot_y(r.y)), rot_x(r.x)); } // matrix.h:94-96
static inline float deg2rad(float d) { return (float)((double)d * 0.01745329); }    // core/utils.h:11-13
static inline V3 right(const M3& a) { return {a.m[0][0], a.m[1][0], a.m[2][0]}; }   // matrix.h:45-47
static inline V3 up(const M3& a) { return {a.m[0][1], a.m[1][1], a.m[2][1]}; }      // matrix.h:48-50

// ---------------------------------------------------------------------------
// Transform (core/include/core/transform.h:10-78)
// ---------------------------------------------------------------------------
struct Xform {
    V3 pos, scale, inv_s;
    M3 R, inv_R;
};
static Xform make_xform(V3 p, V3 rot_deg, V3 scale) {                               // transform.h:25-34
    Xform x;
    x.pos = p;
    x.R = rotation_from_euler({deg2rad(rot_deg.x), deg2rad(rot_deg.y), deg2rad(rot_deg.z)});
    x.inv_R = transpose(x.R);
    x.scale = scale;
    x.inv_s = divv(v3(1.0f), adds(scale, FMIN));
    return x;
}
static Xform xform_identity_at(V3 p) {                                              // transform.h:17-23
    Xform x;
    x.pos = p; x.R = rotation_from_euler(v3(0.0f)); x.inv_R = transpose(x.R);
    x.scale = v3(1.0f); x.inv_s = v3(1.0f);
    return x;
}

// ---------------------------------------------------------------------------
// Ray (engine/include/engine/ray.h:6-45)
// ---------------------------------------------------------------------------
struct Ray {
    V3 o, d, inv;
    float tm;
};
static inline Ray make_ray(V3 o, V3 d, float tm = 0.0f) {
    return {o, d, {1.0f / d.x, 1.0f / d.y, 1.0f / d.z}, tm};
}
static inline V3 at(const Ray& r, float t) { return add(r.o, muls(r.d, t)); }        // ray.h:31
static inline Ray to_object_space(const Ray& r, const Xform& x) {                    // ray.h:33-37
    V3 o = mul(mv(x.inv_R, sub(r.o, x.pos)), x.inv_s);
    V3 d = mul(mv(x.inv_R, r.d), x.inv_s);
    return make_ray(o, d, r.tm);
}

// ---------------------------------------------------------------------------
// Interval (core/include/core/interval.h) and AABB (hittable/include/bvh/aabb.h)
// ---------------------------------------------------------------------------
struct Interval { float min, max; };
static inline bool contains(Interval i, float x) { return i.min <= x && x <= i.max; }   // interval.h:28-30
static inline bool surrounds(Interval i, float x) { return i.min < x && x < i.max; }    // interval.h:32-34

struct AABB { Interval a[3]; };
static inline AABB aabb_empty() { AABB b; for (auto& i : b.a) i = {FMAX, -FMAX}; return b; } // aabb.h:11
static inline AABB aabb_pts(V3 p, V3 q) {                                               // aabb.h:19-23
    AABB b;
    for (int k = 0; k < 3; k++) {
        float a = get(p, k), c = get(q, k);
        b.a[k] = (a <= c) ? Interval{a, c} : Interval{c, a};
    }
    return b;
}
static inline AABB aabb_union(const AABB& x, const AABB& y) {                           // aabb.h:25-29, interval.h:14-17
    AABB b;
    for (int k = 0; k < 3; k++) {
        b.a[k].min = x.a[k].min <= y.a[k].min ? x.a[k].min : y.a[k].min;
        b.a[k].max = x.a[k].max >= y.a[k].max ? x.a[k].max : y.a[k].max;
    }
    return b;
}
static inline float surface_area(const AABB& b) {                                       // aabb.h:159-162
    float dx = b.a[0].max - b.a[0].min, dy = b.a[1].max - b.a[1].min, dz = b.a[2].max - b.a[2].min;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}
// AABB::hit_t (aabb.h:135-157); hit() is hit_t(...) != FMAX (aabb.h:111-133)
static inline float aabb_hit_t(const AABB& b, const Ray& r, Interval t) {
    for (int axis = 0; axis < 3; axis++) {
        const Interval& ax = b.a[axis];
        float t0 = (ax.min - get(r.o, axis)) * get(r.inv, axis);
        float t1 = (ax.max - get(r.o, axis)) * get(r.inv, axis);
        if (t0 < t1) {
            if (t0 > t.min) t.min = t0;
            if (t1 < t.max) t.max = t1;
        } else {
            if (t1 > t.min) t.min = t1;
            if (t0 < t.max) t.max = t0;
        }
        if (t.max <= t.min) return FMAX;
    }
    return t.min;
}
static inline bool aabb_hit(const AABB& b, const Ray& r, Interval t) {
    for (int axis = 0; axis < 3; axis++) {
        const Interval& ax = b.a[axis];
        float t0 = (ax.min - get(r.o, axis)) * get(r.inv, axis);
        float t1 = (ax.max - get(r.o, axis)) * get(r.inv, axis);
        if (t0 < t1) {
            if (t0 > t.min) t.min = t0;
            if (t1 < t.max) t.max = t1;
        } else {
            if (t1 > t.min) t.min = t1;
            if (t0 < t.max) t.max = t0;
        }
        if (t.max <= t.min) return false;
    }
    return true;
}
static AABB aabb_buffer(AABB b, float e = EPS) {                                        // aabb.h:98-103, interval.h:65-68
    for (auto& i : b.a)
        if (i.max - i.min < e) { float pad = e / 2.0f; i = {i.min - pad, i.max + pad}; }
    return b;
}
static AABB aabb_transformed(const AABB& b, const Xform& x) {                           // aabb.h:73-96
    V3 c[8] = {{b.a[0].min, b.a[1].min, b.a[2].min}, {b.a[0].max, b.a[1].min, b.a[2].min},
               {b.a[0].min, b.a[1].max, b.a[2].min}, {b.a[0].max, b.a[1].max, b.a[2].min},
               {b.a[0].min, b.a[1].min, b.a[2].max}, {b.a[0].max, b.a[1].min, b.a[2].max},
               {b.a[0].min, b.a[1].max, b.a[2].max}, {b.a[0].max, b.a[1].max, b.a[2].max}};
    V3 mn = v3(FMAX), mx = v3(-FMAX);
    for (const V3& cc : c) {
        V3 w = add(mv(x.R, mul(cc, x.scale)), x.pos);
        mn = {fminf(mn.x, w.x), fminf(mn.y, w.y), fminf(mn.z, w.z)};
        mx = {fmaxf(mx.x, w.x), fmaxf(mx.y, w.y), fmaxf(mx.z, w.z)};
    }
    return aabb_buffer(aabb_pts(mn, mx));
}

// ---------------------------------------------------------------------------
// Binned SAH BVH (hittable/include/bvh/bvh.h:10-320)
// ---------------------------------------------------------------------------
struct BVHNode { AABB box; int32_t left, right, object; };

struct BVHBuilder {
    static constexpr int32_t BINS = 16;                                                 // bvh.h:33
    std::vector<BVHNode> nodes;
    std::vector<int32_t> order;
    std::vector<AABB> boxes;

    void build(const std::vector<AABB>& in) {                                           // bvh.h:36-55
        boxes = in;
        nodes.clear(); order.clear();
        if (boxes.empty()) return;
        nodes.reserve(boxes.size() * 2);
        order.resize(boxes.size());
        std::iota(order.begin(), order.end(), 0);
        build_recursive(0, (int32_t)order.size());
    }
    static float centroid(const AABB& b, int axis) { return (b.a[axis].min + b.a[axis].max) * 0.5f; }
    static int32_t bin_of(float c, float c_min, float extent) {                          // bvh.h:141-144
        int32_t bin = (int32_t)(BINS * (c - c_min) / extent);
        return bin < 0 ? 0 : (bin >= BINS ? BINS - 1 : bin);
    }
    int32_t build_recursive(int32_t start, int32_t end) {                               // bvh.h:266-320
        int32_t span = end - start;
        size_t idx = nodes.size();
        nodes.push_back({});
        if (span == 1) {
            nodes[idx].object = start; nodes[idx].left = -1; nodes[idx].right = -1;
            nodes[idx].box = boxes[order[start]];
            return (int32_t)idx;
        }
        V3 cmin = v3(FMAX), cmax = v3(-FMAX);                                           // bvh.h:102-126
        for (int32_t i = start; i < end; i++) {
            const AABB& b = boxes[order[i]];
            V3 c = {(b.a[0].min + b.a[0].max) * 0.5f, (b.a[1].min + b.a[1].max) * 0.5f,
                    (b.a[2].min + b.a[2].max) * 0.5f};
            cmin = {fminf(cmin.x, c.x), fminf(cmin.y, c.y), fminf(cmin.z, c.z)};
            cmax = {fmaxf(cmax.x, c.x), fmaxf(cmax.y, c.y), fmaxf(cmax.z, c.z)};
        }
        float best_cost = FMAX;
        int32_t best_axis = -1, best_bin = -1;
        for (int32_t axis = 0; axis < 3; axis++) {
            float c_min = get(cmin, axis), c_max = get(cmax, axis);
            float extent = c_max - c_min;
            if (extent < EPS) continue;
            AABB bin_box[BINS]; int32_t bin_cnt[BINS];
            for (int k = 0; k < BINS; k++) { bin_box[k] = aabb_empty(); bin_cnt[k] = 0; }
            for (int32_t i = start; i < end; i++) {                                     // bvh.h:128-147
                const AABB& b = boxes[order[i]];
                int32_t bin = bin_of(centroid(b, axis), c_min, extent);
                bin_cnt[bin]++;
                bin_box[bin] = bin_cnt[bin] == 1 ? b : aabb_union(bin_box[bin], b);
            }
            AABB lbox[BINS], rbox[BINS]; int32_t lcnt[BINS], rcnt[BINS];                // bvh.h:149-199
            AABB running = aabb_empty(); int32_t rc = 0; bool started = false;
            for (int32_t i = 0; i < BINS; i++) {
                if (bin_cnt[i] > 0) { running = started ? aabb_union(running, bin_box[i]) : bin_box[i]; started = true; }
                rc += bin_cnt[i]; lbox[i] = running; lcnt[i] = rc;
            }
            rc = 0; started = false;
            for (int32_t i = BINS - 1; i >= 0; i--) {
                if (bin_cnt[i] > 0) { running = started ? aabb_union(running, bin_box[i]) : bin_box[i]; started = true; }
                rc += bin_cnt[i]; rbox[i] = running; rcnt[i] = rc;
            }
            for (int32_t i = 0; i < BINS - 1; i++) {
                int32_t nl = lcnt[i], nr = rcnt[i + 1];
                if (nl == 0 || nr == 0) continue;
                float cost = surface_area(lbox[i]) * nl + surface_area(rbox[i + 1]) * nr;
                if (cost < best_cost) { best_cost = cost; best_axis = axis; best_bin = i; }
            }
        }
        if (best_axis == -1) {                                                          // bvh.h:307-315
            AABB comb = boxes[order[start]];
            for (int32_t i = start + 1; i < end; i++) comb = aabb_union(comb, boxes[order[i]]);
            float lx = comb.a[0].max - comb.a[0].min, ly = comb.a[1].max - comb.a[1].min,
                  lz = comb.a[2].max - comb.a[2].min;
            best_axis = (lx > ly && lx > lz) ? 0 : (ly > lz ? 1 : 2);
        }
        int32_t mid;
        if (best_bin == -1) {                                                           // bvh.h:201-222
            int32_t axis = best_axis;
            std::sort(order.begin() + start, order.begin() + end, [&](int32_t a, int32_t b) {
                return boxes[a].a[axis].min < boxes[b].a[axis].min;
            });
            mid = start + (end - start) / 2;
        } else {                                                                        // bvh.h:224-260
            int32_t axis = best_axis;
            float c_min = get(cmin, axis);
            float extent = get(cmax, axis) - c_min;
            int32_t bb = best_bin;
            auto it = std::partition(order.begin() + start, order.begin() + end, [&](int32_t a) {
                const AABB& b = boxes[a];
                float c = (b.a[axis].min + b.a[axis].max) * 0.5f;
                return bin_of(c, c_min, extent) <= bb;
            });
            mid = (int32_t)(it - order.begin());
        }
        nodes[idx].object = -1;
        int32_t l = build_recursive(start, mid);
        nodes[idx].left = l;
        int32_t r = build_recursive(mid, end);
        nodes[idx].right = r;
        nodes[idx].box = aabb_union(nodes[l].box, nodes[r].box);
        return (int32_t)idx;
    }
};

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11; rocRAND rocrand_philox4x32_10.h) and
// Generator (core/include/core/random.h:11-75)
// ---------------------------------------------------------------------------
static inline void philox10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
// rocrand_init(seed, subsequence, offset) then rocrand_uniform: key = seed,
// counter = (offset/4, 0, subsequence, 0), output word offset%4, u = 2^-32 + v*2^-32.
static inline float philox_uniform(uint32_t seed, uint32_t subseq, uint32_t offset) {
    uint32_t ctr[4] = {offset >> 2, 0u, subseq, 0u};
    uint32_t key[2] = {seed, 0u};
    uint32_t out[4];
    philox10(ctr, key, out);
    const float inv = 2.3283064e-10f;
    return inv + ((float)out[offset & 3u] * inv);
}
struct Gen {
    uint32_t seed, subseq, draw;
    float uniform() { float u = philox_uniform(seed, subseq, draw); draw += 1; return u; }   // random.h:50-55
    float random_float(float mn = 0.0f, float mx = 1.0f) { return uniform() * (mx - mn) + mn; } // random.h:64-66
};

// ---------------------------------------------------------------------------
// PCG hash (random/include/hash.h:16-191)
// ---------------------------------------------------------------------------
static const uint32_t HASH_CONSTANTS[64] = {                                            // hash.h:17-34
    0x9E3779B9u, 0x85EBCA6Bu, 0xC2B2AE35u, 0x27D4EB2Fu, 0x165667B1u, 0x68D2A5B5u, 0xBF58476Du, 0x94D049BBu,
    0xA3B195B5u, 0x1B873593u, 0xCC9E2D51u, 0x6C62272Eu, 0xE654025Bu, 0x2B5AD5B5u, 0x6C696E65u, 0x736F6D65u,
    0x646F6D65u, 0x70736575u, 0x6E646F6Du, 0xFF51AFD7u, 0xC4CEA4D6u, 0x45D9F3Bu,  0x3335B369u, 0xA341316Cu,
    0x52784632u, 0x68656C6Cu, 0x6F20776Fu, 0x726C6421u, 0xB5297A4Du, 0x68E31DA4u, 0x1B56C4E9u, 0xD2A98B26u,
    0x637B2B53u, 0x4F6E4F35u, 0x62DC5AB7u, 0x5A17A2C3u, 0x6EF305DAu, 0x2C9F410Du, 0x3771A6A7u, 0x15C3A7C4u,
    0x8F43CD0Bu, 0xA9DE8FA3u, 0xEBF7A401u, 0x58399E71u, 0x42B91A0Du, 0x4B608671u, 0xB41B21C7u, 0x2D5A73B7u,
    0xE50B4C1Fu, 0x95C48A21u, 0x37A07BA9u, 0xC0F4A537u, 0xA8FA763Bu, 0xD08BE5EDu, 0x4A9C5EB3u, 0x7B14EEA7u,
    0xCC3A6547u, 0x9B89D66Du, 0xE15E5E3Bu, 0x4F3D8A5Fu, 0xD4E0A97Bu, 0x3C6EF372u, 0xA54FF53Au, 0x510E527Fu,
};
static const float HASH_SCALAR = 1.0f / (float)0xFFFFFFFFu;                              // hash.h:16
static inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }   // core/utils.h:31-35
static inline uint32_t hash_mix(uint32_t v, uint32_t ci) {                               // hash.h:39-45
    v ^= (v >> 16u) * HASH_CONSTANTS[ci & 63u];
    return v ^ (v >> 13u);
}
static inline uint32_t pcg_hash_int(uint32_t i, uint32_t seed, uint32_t stream) {       // hash.h:47-61
    uint32_t s = (seed * 1759274928u) ^ (stream * 2246822519u);
    s = i ^ s;
    s = hash_mix(s, stream * 3u);
    s = hash_mix(s, stream * 3u + 1u);
    s = hash_mix(s, stream * 3u + 2u);
    uint32_t state = s * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
static inline float pcg_float(float f, uint32_t seed = 0u, uint32_t stream = 0u) {      // hash.h:75-82
    return (float)pcg_hash_int(f2u(f), seed, stream) * HASH_SCALAR;
}
static inline float pcg_float3(V3 a, uint32_t seed = 0u, uint32_t stream = 0u) {        // hash.h:92-105
    uint32_t s = seed * 1759274928u;
    s ^= pcg_hash_int(f2u(a.x), (seed + 0u) * 2654435761u, stream) ^ 668265263u;
    s ^= pcg_hash_int(f2u(a.y), (seed + 1u) * 2246822519u, stream) ^ 374761393u;
    s ^= pcg_hash_int(f2u(a.z), (seed + 2u) * 3266489917u, stream) ^ 2654435761u;
    uint32_t state = s * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (float)((word >> 22u) ^ word) * HASH_SCALAR;
}
static inline float pcg_float_ray(const Ray& r) {                                        // hash.h:107-124
    float a = pcg_float3(r.o, 0u * 2654435761u, 0u);
    float b = pcg_float3(r.d, 1u * 3266489917u, 0u);
    return (a + b) * 0.5f;
}
// float -> uint32 as the GPU converts (saturating); x*2^32 reaches 2^32 when x == 1
static inline uint32_t f2u_sat(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}
static inline V2 pcg_vector2(float f, uint32_t seed) {                                   // hash.h:182-191
    float x = pcg_float(f, seed, 0u);
    uint32_t s = seed ^ f2u_sat(x * 4294967296.0f);
    float y = pcg_float((float)(s * 747796405u + 2891336453u), 0u, 0u);
    return {x, y};
}

// ---------------------------------------------------------------------------
// ONB (core/include/core/onb.h:5-26)
// ---------------------------------------------------------------------------
struct ONB {
    V3 u, v, w;
    explicit ONB(V3 n) {
        w = normalize(n);
        V3 a = (fabsf(w.x) > 0.9f) ? V3{0.0f, 1.0f, 0.0f} : V3{1.0f, 0.0f, 0.0f};
        v = normalize(cross(w, a));
        u = cross(w, v);
    }
    V3 transform(V3 p) const { return add(add(muls(u, p.x), muls(v, p.y)), muls(w, p.z)); }
};

static inline V3 random_cosine_direction(Gen& g) {                                       // vector.h:509-519
    float r1 = g.random_float();
    float r2 = g.random_float();
    float phi = PI2 * r1;
    float x = nrm::cosf_(phi) * sqrtf(r2);
    float y = nrm::sinf_(phi) * sqrtf(r2);
    float z = sqrtf(1.0f - r2);
    return {x, y, z};
}
static inline V3 random_unit_vector(Gen& g) {                                            // vector.h:489-496
    while (true) {
        float a = g.random_float(-1.0f, 1.0f);
        float b = g.random_float(-1.0f, 1.0f);
        float c = g.random_float(-1.0f, 1.0f);
        V3 p = {a, b, c};
        float l = len2(p);
        if (1e-160 < l && l <= 1) return divs(p, sqrtf(l));
    }
}

// ---------------------------------------------------------------------------
// scene data
// ---------------------------------------------------------------------------
struct Texture {                                                                         // material/texture.h:49-91
    int image = -1; V3 color{0, 0, 0}; float scale = 1.0f;
};
struct Image { int C = 0, H = 0, W = 0; std::vector<uint8_t> px; };
struct Material {
    int type = 0; Texture tex[5]; V3 tint{1, 1, 1}; float ior = 1.5f; float brightness = 35.0f;
};
struct MeshVertex { V3 pos, normal, tangent; V2 uv; };
struct Mesh {
    std::vector<MeshVertex> verts;
    std::vector<uint32_t> tris;      // 3 per triangle, BVH order
    std::vector<int32_t> tri_orig;   // original triangle index per BVH slot
    std::vector<BVHNode> nodes;
    int depth = 0;
};
struct Object {
    int type = 0, bound_type = 0, material = 0, mesh = -1, object_id = 0;
    Xform xf;
    float radius = 0.0f;             // radius + FMIN (shapes.h:9)
    float neg_inv_density = 0.0f;    // volumes.h:13-14
    AABB local_box;
};
struct HitRec {                                                                          // hit_record.h:9-22
    int object_id = 0, material_id = 0, prim = -1;
    V3 p{0, 0, 0}, n{0, 0, 0}, tangent{0, 0, 0};
    V2 uv{0, 0};
    float t = 0.0f;
    bool front_face = false;
};
struct Counters { int64_t inner = 0, tri = 0, obj = 0; };

struct Scene {
    std::vector<Object> objects;     // BVH order (registry reorders entries, registry.h:216-234)
    std::vector<Material> materials;
    std::vector<Image> images;
    std::vector<Mesh> meshes;
    std::vector<BVHNode> nodes;
    int first_light = -1;
    nr_sky_desc sky{};
    std::vector<float> hdri;
    Xform cube_faces[6];
};

// Quad constants (primitives.h:8-16)
struct QuadConst {
    V3 u, v, w, normal; float area;
    QuadConst() {
        u = {1.0f, 0.0f, 0.0f};
        v = {0.0f, 0.0f, -1.0f};
        V3 n = cross(u, v);
        normal = normalize(n);
        w = divs(n, dot(n, n));
        area = length(n);
    }
};
static const QuadConst QC;

// HitRecord::to_world_space (hit_record.h:24-51)
static void to_world_space(HitRec& rec, const Xform& x, const Ray& obj_ray, const Ray& world_ray, bool bias) {
    rec.p = add(mv(x.R, mul(rec.p, x.scale)), x.pos);
    rec.tangent = normalize(mv(x.R, mul(rec.tangent, x.scale)));
    rec.front_face = dot(obj_ray.d, rec.n) < 0.0f;
    rec.n = rec.front_face ? rec.n : neg(rec.n);
    rec.n = normalize(mv(x.R, mul(rec.n, x.inv_s)));
    rec.front_face = dot(world_ray.d, rec.n) < 0.0f;
    rec.n = rec.front_face ? rec.n : neg(rec.n);
    rec.tangent = normalize(sub(rec.tangent, muls(rec.n, dot(rec.tangent, rec.n))));
    if (near_zero(rec.tangent))
        rec.tangent = normalize(cross(rec.n, fabsf(rec.n.x) > 0.9f ? V3{0.0f, 1.0f, 0.0f} : V3{1.0f, 0.0f, 0.0f}));
    if (bias) rec.p = add(rec.p, muls(rec.n, EPS));
    rec.t = dot(sub(rec.p, world_ray.o), world_ray.d) / len2(world_ray.d);
}

// Quad::hit (primitives.h:19-32, 73-88)
static bool quad_hit(const Ray& r, Interval rt, HitRec& rec) {
    float denom = dot(QC.normal, r.d);
    if (fabs(denom) < FMIN) return false;
    float t = -dot(QC.normal, r.o) / denom;
    if (!contains(rt, t)) return false;
    rec.t = t; rec.p = at(r, t); rec.n = QC.normal; rec.tangent = QC.u;
    float a = dot(QC.w, cross(rec.p, QC.v));
    float b = dot(QC.w, cross(QC.u, rec.p));
    rec.uv = {a + 0.5f, b + 0.5f};
    Interval unit{-0.5f, 0.5f};
    return surrounds(unit, a) & surrounds(unit, b);
}

// Sphere::hit (shapes.h:25-58)
static bool sphere_hit(const Ray& r, Interval rt, HitRec& rec, float radius, const Xform& x) {
    V3 oc = neg(r.o);
    float a = len2(r.d);
    float h = dot(r.d, oc);
    float c = len2(oc) - radius * radius;
    float disc = h * h - a * c;
    if (disc < 0) return false;
    float sqrtd = sqrtf(disc);
    float root = (h - sqrtd) / a;
    if (!surrounds(rt, root)) {
        root = (h + sqrtd) / a;
        if (!surrounds(rt, root)) return false;
    }
    rec.t = root;
    rec.p = at(r, root);
    V3 norm = divv(rec.p, adds(muls(x.scale, radius), FMIN));
    rec.n = norm;
    rec.tangent = normalize(V3{-rec.p.z, 0.0f, rec.p.x});
    float theta = nrm::acosf_(-norm.y);                                                        // shapes.h:19-23
    float phi = nrm::atan2f_(-norm.z, norm.x) + PI;
    rec.uv = {phi / PI2, theta / PI};
    return true;
}

// Cube::hit (shapes.h:142-173)
static bool cube_hit(const Scene& S, const Ray& r, Interval rt, HitRec& rec) {
    bool any = false; int face = -1; Ray face_ray{};
    for (int i = 0; i < 6; i++) {
        const Xform& fx = S.cube_faces[i];
        Ray fr = to_object_space(r, fx);
        HitRec fre;
        if (quad_hit(fr, rt, fre)) {
            any = true; face = i; face_ray = fr; rec = fre; rt.max = fre.t;
        }
    }
    if (any) to_world_space(rec, S.cube_faces[face], face_ray, r, false);
    return any;
}

// Mesh::hit (mesh.h:239-292), intersect_triangle (:338-365), interpolate_hit (:367-383)
static bool intersect_triangle(const Ray& r, V3 v0, V3 v1, V3 v2, Interval rt, float& ot, float& ou, float& ov) {
    V3 e1 = sub(v1, v0), e2 = sub(v2, v0);
    V3 h = cross(r.d, e2);
    float a = dot(e1, h);
    if (fabsf(a) < FMIN) return false;
    float f = 1.0f / a;
    V3 s = sub(r.o, v0);
    float u = f * dot(s, h);
    if (u < 0.0f || u > 1.0f) return false;
    V3 q = cross(s, e1);
    float v = f * dot(r.d, q);
    if (v < 0.0f || u + v > 1.0f) return false;
    float t = f * dot(e2, q);
    if (!contains(rt, t)) return false;
    ot = t; ou = u; ov = v;
    return true;
}
static bool mesh_hit(const Mesh& M, const Ray& r, Interval rt, HitRec& rec, Counters* cnt) {
    uint32_t stack[256];
    int sp = 0;
    stack[sp++] = 0u;
    bool any = false;
    while (sp > 0) {
        uint32_t idx = stack[--sp];
        const BVHNode& node = M.nodes[idx];
        if (!aabb_hit(node.box, r, rt)) continue;
        if (node.object != -1) {
            if (cnt) cnt->tri++;
            const uint32_t* tri = &M.tris[3 * (size_t)node.object];
            const MeshVertex& v0 = M.verts[tri[0]];
            const MeshVertex& v1 = M.verts[tri[1]];
            const MeshVertex& v2 = M.verts[tri[2]];
            float t, u, v;
            if (intersect_triangle(r, v0.pos, v1.pos, v2.pos, rt, t, u, v)) {
                any = true;
                rt.max = t;
                rec.t = t;
                rec.p = at(r, t);
                float w = 1.0f - u - v;
                rec.uv = {v0.uv.x * w + v1.uv.x * u + v2.uv.x * v, v0.uv.y * w + v1.uv.y * u + v2.uv.y * v};
                rec.n = normalize(add(add(muls(v0.normal, w), muls(v1.normal, u)), muls(v2.normal, v)));
                rec.tangent = normalize(add(add(muls(v0.tangent, w), muls(v1.tangent, u)), muls(v2.tangent, v)));
                rec.prim = M.tri_orig[node.object];
            }
        } else {
            if (cnt) cnt->inner++;
            const BVHNode& L = M.nodes[node.left];
            const BVHNode& R = M.nodes[node.right];
            float tl = aabb_hit_t(L.box, r, rt);
            float tr = aabb_hit_t(R.box, r, rt);
            bool left_near = tl < tr;
            uint32_t near_i = left_near ? node.left : node.right;
            uint32_t far_i = left_near ? node.right : node.left;
            float far_t = left_near ? tr : tl;
            if (sp + 2 > 256) throw std::runtime_error("oracle mesh stack overflow");
            if (far_t < rt.max) stack[sp++] = far_i;
            stack[sp++] = near_i;
        }
    }
    return any;
}

// Hittable::hit for the four primitive types (hittable.h:332-352)
static bool prim_hit(const Scene& S, int type, const Object& o, const Ray& r, Interval rt, HitRec& rec, Counters* cnt) {
    switch (type) {
        case NR_HIT_QUAD: return quad_hit(r, rt, rec);
        case NR_HIT_SPHERE: return sphere_hit(r, rt, rec, o.radius, o.xf);
        case NR_HIT_CUBE: return cube_hit(S, r, rt, rec);
        case NR_HIT_MESH: return mesh_hit(S.meshes[o.mesh], r, rt, rec, cnt);
    }
    return false;
}

// ConstantMedium::hit (volumes.h:14-45)
static bool medium_hit(const Scene& S, const Object& o, const Ray& r, Interval rt, HitRec& rec, Counters* cnt) {
    HitRec r1, r2;
    if (!prim_hit(S, o.bound_type, o, r, Interval{-FMAX, FMAX}, r1, cnt)) return false;
    if (!prim_hit(S, o.bound_type, o, r, Interval{r1.t + EPS, FMAX}, r2, cnt)) return false;
    if (r1.t < rt.min) r1.t = rt.min;
    if (r2.t > rt.max) r2.t = rt.max;
    if (r1.t >= r2.t) return false;
    if (r1.t < 0.0f) r1.t = 0.0f;
    float ray_length = length(r.d);
    float inside = (r2.t - r1.t) * ray_length;
    float hit_distance = o.neg_inv_density * nrm::logf_(pcg_float_ray(r));
    if (hit_distance > inside) return false;
    rec.p = at(r, r1.t + hit_distance / ray_length);
    rec.n = {1.0f, 0.0f, 0.0f};
    rec.front_face = true;
    return true;
}

static bool object_hit(const Scene& S, const Object& o, const Ray& r, Interval rt, HitRec& rec, Counters* cnt) {
    switch (o.type) {
        case NR_HIT_OBJECT_LIGHT:                                                        // lights.h:11-20
            return prim_hit(S, o.bound_type, o, r, Interval{-FMAX, FMAX}, rec, cnt);
        case NR_HIT_CONSTANT_MEDIUM:
            return medium_hit(S, o, r, rt, rec, cnt);
        default:
            return prim_hit(S, o.type, o, r, rt, rec, cnt);
    }
}

// Hittable::hit_test (hittable.h:324-330)
static bool hit_test(const Scene& S, const Object& o, const Ray& r, HitRec& rec, Counters* cnt) {
    Ray orr = to_object_space(r, o.xf);
    bool h = object_hit(S, o, orr, Interval{EPS, FMAX}, rec, cnt);
    to_world_space(rec, o.xf, orr, r, true);
    return h;
}

// SceneGraph::hit (scene/include/graph.h:16-66)
static bool scene_hit(const Scene& S, const Ray& r, Interval rt, HitRec& rec, Counters* cnt) {
    if (S.nodes.empty()) return false;
    uint32_t stack[256];
    int sp = 0;
    stack[sp++] = 0u;
    bool any = false;
    while (sp > 0) {
        uint32_t idx = stack[--sp];
        const BVHNode& node = S.nodes[idx];
        if (!aabb_hit(node.box, r, rt)) continue;
        if (node.object != -1) {
            if (cnt) cnt->obj++;
            HitRec tmp;
            const Object& o = S.objects[node.object];
            if (hit_test(S, o, r, tmp, cnt)) {
                if (surrounds(rt, tmp.t)) {
                    any = true;
                    rec = tmp;
                    rec.object_id = o.object_id;
                    rec.material_id = o.material;
                    rt.max = tmp.t;
                }
            }
        } else {
            if (cnt) cnt->inner++;
            const BVHNode& L = S.nodes[node.left];
            const BVHNode& R = S.nodes[node.right];
            float tl = aabb_hit_t(L.box, r, rt);
            float tr = aabb_hit_t(R.box, r, rt);
            bool left_near = tl < tr;
            uint32_t near_i = left_near ? node.left : node.right;
            uint32_t far_i = left_near ? node.right : node.left;
            float far_t = left_near ? tr : tl;
            if (sp + 2 > 256) throw std::runtime_error("oracle scene stack overflow");
            if (far_t < rt.max) stack[sp++] = far_i;
            stack[sp++] = near_i;
        }
    }
    return any;
}

// ---------------------------------------------------------------------------
// light PDFs of the first ObjectLight (hittable.h:356-386, primitives.h:34-62,
// shapes.h:60-108, lights.h:22-34)
// ---------------------------------------------------------------------------
static float quad_pdf_value(const Object& o, V3 origin, V3 dir) {                       // primitives.h:34-49
    HitRec rec;
    Ray wr = make_ray(origin, dir);
    Ray orr = to_object_space(wr, o.xf);
    bool h = quad_hit(orr, Interval{EPS, FMAX}, rec);
    to_world_space(rec, o.xf, orr, wr, true);
    if (!h) return 0.0f;
    float dist2 = rec.t * rec.t * len2(dir);
    float cosine = fabsf(dot(dir, rec.n) / length(dir));
    if (cosine < 1e-8f) return 0.0f;
    float world_area = QC.area * o.xf.scale.x * o.xf.scale.z;
    return dist2 / (cosine * world_area);
}
static V3 quad_random(const Object& o, V3 origin, Gen& g) {                             // primitives.h:52-61
    V3 corner = sub(muls(QC.u, -0.5f), muls(QC.v, 0.5f));
    float r1 = g.random_float();
    float r2 = g.random_float();
    V3 p = add(add(corner, muls(QC.u, r1)), muls(QC.v, r2));
    p = add(mv(o.xf.R, mul(p, o.xf.scale)), o.xf.pos);
    return sub(p, origin);
}
static float sphere_pdf_value(const Object& o, V3 origin, V3 dir) {                     // shapes.h:60-76
    HitRec rec;
    if (!sphere_hit(make_ray(origin, dir), Interval{EPS, FMAX}, rec, o.radius, o.xf)) return 0.0f;
    float rad_sq = o.radius * o.radius;
    float dist2 = len2(sub(o.xf.pos, origin));
    float cos_theta_max = sqrtf(1.0f - rad_sq / dist2);
    float solid_angle = PI2 * (1.0f - cos_theta_max);
    return 1.0f / solid_angle;
}
static V3 sphere_random(const Object& o, V3 origin, Gen& g) {                           // shapes.h:78-108
    V3 direction = sub(o.xf.pos, origin);
    float dist2 = len2(direction);
    ONB uvw(direction);
    float r1 = g.random_float();
    float r2 = g.random_float();
    float rad_sq = o.radius * o.radius;
    float z = 1.0f + r2 * (sqrtf(1.0f - rad_sq / dist2) - 1.0f);
    float phi = PI2 * r1;
    float x = nrm::cosf_(phi) * sqrtf(1.0f - z * z);
    float y = nrm::sinf_(phi) * sqrtf(1.0f - z * z);
    return uvw.transform({x, y, z});
}
static float light_pdf_value(const Object& L, V3 origin, V3 dir) {
    int t = L.type == NR_HIT_OBJECT_LIGHT ? L.bound_type : (L.type == NR_HIT_CONSTANT_MEDIUM ? -1 : L.type);
    switch (t) {
        case NR_HIT_QUAD: return quad_pdf_value(L, origin, dir);
        case NR_HIT_SPHERE: return sphere_pdf_value(L, origin, dir);
        default: return 0.0f;   // Cube / Mesh (shapes.h:175-179, mesh.h:294-298); medium not dispatched
    }
}
static V3 light_random(const Object& L, V3 origin, Gen& g) {
    int t = L.type == NR_HIT_OBJECT_LIGHT ? L.bound_type : (L.type == NR_HIT_CONSTANT_MEDIUM ? -1 : L.type);
    switch (t) {
        case NR_HIT_QUAD: return quad_random(L, origin, g);
        case NR_HIT_SPHERE: return sphere_random(L, origin, g);
        case NR_HIT_MESH: return {0.0f, 0.0f, 0.0f};
        default: return {1.0f, 0.0f, 0.0f};
    }
}

// ---------------------------------------------------------------------------
// textures (material/texture.h:49-91)
// ---------------------------------------------------------------------------
static V3 tex_sample(const Scene& S, const Texture& t, V2 uv) {
    if (t.image < 0) return t.color;
    const Image& im = S.images[t.image];
    float sc = 1.0f / fmaxf(FMIN, t.scale);
    float su = fmodf(uv.x * sc, 1.0f), sv = fmodf(uv.y * sc, 1.0f);
    float u = su < 0.0f ? 0.0f : (su > 1.0f ? 1.0f : su);
    float v = 1.0f - (sv < 0.0f ? 0.0f : (sv > 1.0f ? 1.0f : sv));
    int i = (int)(u * (float)im.W); if (i >= im.W) i = im.W - 1;
    int j = (int)(v * (float)im.H); if (j >= im.H) j = im.H - 1;
    size_t px = (size_t)j * im.W + i, hw = (size_t)im.H * im.W;
    V3 c = {(float)im.px[px], (float)im.px[hw + px], (float)im.px[2 * hw + px]};
    return divs(c, 255.0f);
}

// ---------------------------------------------------------------------------
// PDFs (engine/include/engine/pdf.h)
// ---------------------------------------------------------------------------
static inline float ggx_D(float n_dot_h, float alpha) {                                  // pdf.h:100-104
    float a2 = alpha * alpha;
    float d = n_dot_h * n_dot_h * (a2 - 1.0f) + 1.0f;
    return a2 / fmaxf(PI * d * d, 1e-8f);
}
static inline float smith_g1(float nx, float alpha) {                                    // pdf.h:106-111
    float a2 = alpha * alpha;
    return 2.0f * nx / (nx + sqrtf(a2 + (1.0f - a2) * nx * nx));
}

enum PdfType { PDF_COSINE, PDF_SPHERE, PDF_PBR };
struct Pdf {
    PdfType type = PDF_SPHERE;
    ONB uvw{V3{0, 0, 1}};
    V3 view{0, 0, 0};
    float alpha = 0.0f, spec_prob = 0.0f;

    float value(V3 d) const {
        switch (type) {
            case PDF_COSINE: {                                                           // pdf.h:26-29
                float c = dot(normalize(d), uvw.w);
                return fmaxf(0.0f, c / PI);
            }
            case PDF_SPHERE: return 1.0f / PI4;                                          // pdf.h:11-13
            case PDF_PBR: {                                                              // pdf.h:135-148
                float diffuse_pdf = fmaxf(0.0f, dot(normalize(d), uvw.w) / PI);
                V3 h = normalize(add(view, d));
                float n_dot_h = fmaxf(dot(uvw.w, h), 0.0f);
                float v_dot_h = fmaxf(dot(view, h), 1e-4f);
                float spec_pdf = (ggx_D(n_dot_h, alpha) * n_dot_h) / (4.0f * v_dot_h);
                return (1.0f - spec_prob) * diffuse_pdf + spec_prob * spec_pdf;
            }
        }
        return 0.0f;
    }
    V3 generate(Gen& g) const {
        switch (type) {
            case PDF_COSINE: return uvw.transform(random_cosine_direction(g));           // pdf.h:31-33
            case PDF_SPHERE: return random_unit_vector(g);                               // pdf.h:15-17
            case PDF_PBR: {                                                              // pdf.h:150-166
                if (g.random_float() < spec_prob) {
                    float u1 = g.random_float();
                    float u2 = g.random_float();
                    float cos_theta = sqrtf((1.0f - u1) / (1.0f + (alpha * alpha - 1.0f) * u1));
                    float sin_theta = sqrtf(fmaxf(0.0f, 1.0f - cos_theta * cos_theta));
                    float phi = PI2 * u2;
                    V3 h = uvw.transform({sin_theta * nrm::cosf_(phi), sin_theta * nrm::sinf_(phi), cos_theta});
                    return normalize(sub(muls(h, 2.0f * dot(view, h)), view));
                }
                return uvw.transform(random_cosine_direction(g));
            }
        }
        return {0.0f, 0.0f, 1.0f};
    }
};
static Pdf pdf_cosine(V3 n) { Pdf p; p.type = PDF_COSINE; p.uvw = ONB(n); return p; }
static Pdf pdf_sphere() { Pdf p; p.type = PDF_SPHERE; return p; }
static Pdf pdf_pbr(V3 n, V3 view, float rough, float sp) {                               // pdf.h:124-133
    Pdf p; p.type = PDF_PBR; p.uvw = ONB(n); p.view = view;
    p.alpha = fmaxf(rough * rough, 1e-3f); p.spec_prob = sp;
    return p;
}

// ---------------------------------------------------------------------------
// materials (material/include/material/material.h:33-368)
// ---------------------------------------------------------------------------
struct Scatter { V3 atten{1, 1, 1}; Pdf pdf; bool skip_pdf = false; Ray skip_ray{}; };

static V3 pbr_shading_normal(const Scene& S, const Material& m, const HitRec& rec) {     // material.h:113-125
    V3 tn = tex_sample(S, m.tex[4], rec.uv);
    V3 map_n = normalize(V3{tn.x * 2.0f - 1.0f, tn.y * 2.0f - 1.0f, tn.z * 2.0f - 1.0f});
    V3 t = normalize(sub(rec.tangent, muls(rec.n, dot(rec.tangent, rec.n))));
    V3 b = cross(rec.n, t);
    return normalize(add(add(muls(t, map_n.x), muls(b, map_n.y)), muls(rec.n, map_n.z)));
}

static bool mat_scatter(const Scene& S, const Material& m, const HitRec& rec, const Ray& ray, Scatter& sr, Gen& g) {
    switch (m.type) {
        case NR_MAT_LAMBERTIAN:                                                          // material.h:47-57
            sr.atten = tex_sample(S, m.tex[0], rec.uv);
            sr.pdf = pdf_cosine(rec.n);
            sr.skip_pdf = false;
            return true;
        case NR_MAT_PBR: {                                                               // material.h:127-150
            V3 n = pbr_shading_normal(S, m, rec);
            V3 albedo = tex_sample(S, m.tex[0], rec.uv);
            float metallic = tex_sample(S, m.tex[2], rec.uv).x;
            V3 F0 = add(muls(v3(0.04f), 1.0f - metallic), muls(albedo, metallic));
            float spec_prob = fminf(0.9f, fmaxf(0.1f, fmaxf(F0.x, fmaxf(F0.y, F0.z))));
            sr.atten = v3(1.0f);
            sr.skip_pdf = false;
            sr.pdf = pdf_pbr(n, normalize(neg(ray.d)), tex_sample(S, m.tex[1], rec.uv).x, spec_prob);
            return true;
        }
        case NR_MAT_DIELECTRIC: {                                                        // material.h:220-244
            sr.atten = m.tint;
            sr.skip_pdf = true;
            float ri = rec.front_face ? (1.0f / m.ior) : m.ior;
            V3 unit = normalize(ray.d);
            float cos_theta = fminf(dot(neg(unit), rec.n), 1.0f);
            float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
            bool cannot_refract = ri * sin_theta > 1.0;
            V3 dir;
            bool refl = cannot_refract;
            if (!refl) {
                float r0 = (1.0f - ri) / (1.0f + ri);                                    // material.h:262-269
                r0 = r0 * r0;
                float refl_p = r0 + (1.0f - r0) * nrm::powf_((1.0f - cos_theta), 5.0f);
                refl = refl_p > g.uniform();
            }
            dir = refl ? reflect(unit, rec.n) : refract(unit, rec.n, ri);
            sr.skip_ray = make_ray(rec.p, dir, ray.tm);
            return true;
        }
        case NR_MAT_EMISSIVE: return false;                                              // material.h:293-298
        case NR_MAT_ISOTROPIC:                                                           // material.h:339-349
            sr.atten = tex_sample(S, m.tex[0], rec.uv);
            sr.pdf = pdf_sphere();
            sr.skip_pdf = false;
            return true;
    }
    return false;
}

static V3 mat_evaluate(const Scene& S, const Material& m, const HitRec& rec, V3 view, V3 light) {
    switch (m.type) {
        case NR_MAT_LAMBERTIAN: {                                                        // material.h:59-67
            float c = dot(rec.n, normalize(light));
            if (c <= 0.0f) return v3(0.0f);
            return muls(tex_sample(S, m.tex[0], rec.uv), c / PI);
        }
        case NR_MAT_PBR: {                                                               // material.h:152-180
            V3 n = pbr_shading_normal(S, m, rec);
            float n_dot_l = dot(n, light);
            float n_dot_v = dot(n, view);
            if (n_dot_l <= 0.0f || n_dot_v <= 0.0f) return v3(0.0f);
            V3 albedo = tex_sample(S, m.tex[0], rec.uv);
            float roughness = tex_sample(S, m.tex[1], rec.uv).x;
            float metallic = tex_sample(S, m.tex[2], rec.uv).x;
            float alpha = fmaxf(roughness * roughness, 1e-3f);
            V3 h = normalize(add(view, light));
            float n_dot_h = fmaxf(dot(n, h), 0.0f);
            float v_dot_h = fmaxf(dot(view, h), 0.0f);
            V3 F0 = add(muls(v3(0.04f), 1.0f - metallic), muls(albedo, metallic));
            V3 F = add(F0, muls(sub(v3(1.0f), F0), nrm::powf_(1.0f - v_dot_h, 5.0f)));
            float D = ggx_D(n_dot_h, alpha);
            float G = smith_g1(n_dot_l, alpha) * smith_g1(n_dot_v, alpha);
            V3 spec = muls(F, D * G / fmaxf(4.0f * n_dot_v * n_dot_l, 1e-4f));
            V3 diff = muls(muls(albedo, 1.0f - metallic), 1.0f / PI);
            return muls(add(diff, spec), n_dot_l);
        }
        case NR_MAT_ISOTROPIC:                                                           // material.h:351-357
            return muls(tex_sample(S, m.tex[0], rec.uv), 1.0f / PI4);
    }
    return v3(0.0f);
}

static V3 mat_emitted(const Scene& S, const Material& m, const HitRec& rec) {
    switch (m.type) {
        case NR_MAT_PBR:                                                                 // material.h:182-188
            if (!rec.front_face) return v3(0.0f);
            return tex_sample(S, m.tex[3], rec.uv);
        case NR_MAT_EMISSIVE:                                                            // material.h:306-312
            if (!rec.front_face) return v3(0.0f);
            return muls(tex_sample(S, m.tex[0], rec.uv), m.brightness);
    }
    return v3(0.0f);
}

// ---------------------------------------------------------------------------
// sky (light/include/skylight.h:33-122, 243-254)
// ---------------------------------------------------------------------------
static V3 sky_sample(const Scene& S, const Ray& r) {
    const nr_sky_desc& sk = S.sky;
    if (sk.type == NR_SKY_SIMPLE) {
        float a = (float)(0.5 * ((double)r.d.y + 1.0));
        float oma = (float)(1.0 - (double)a);
        V3 st = {sk.start[0], sk.start[1], sk.start[2]}, en = {sk.end[0], sk.end[1], sk.end[2]};
        return add(muls(st, oma), muls(en, a));
    }
    if (sk.type == NR_SKY_HDRI) {
        if (S.hdri.empty()) return v3(0.0f);
        M3 rot = rot_y(deg2rad(sk.rotation));
        V3 dir = mv(rot, normalize(r.d));
        float u = nrm::atan2f_(dir.z, dir.x) / PI2 + 0.5f;
        float v = nrm::acosf_(fminf(fmaxf(dir.y, -1.0f), 1.0f)) / PI;
        int W = sk.W, H = sk.H;
        float fx = u * (float)W - 0.5f;
        float fy = v * (float)H - 0.5f;
        int x0 = (int)floorf(fx), y0 = (int)floorf(fy);
        int x1 = x0 + 1, y1 = y0 + 1;
        float tx = fx - (float)x0, ty = fy - (float)y0;
        x0 = ((x0 % W) + W) % W;
        x1 = ((x1 % W) + W) % W;
        y0 = (int)fmaxf(0, fminf((float)(H - 1), (float)y0));
        y1 = (int)fmaxf(0, fminf((float)(H - 1), (float)y1));
        size_t hw = (size_t)H * W;
        auto texel = [&](int x, int y) {
            size_t i = (size_t)y * W + (size_t)x;
            return V3{S.hdri[i], S.hdri[hw + i], S.hdri[2 * hw + i]};
        };
        V3 c00 = texel(x0, y0), c10 = texel(x1, y0), c01 = texel(x0, y1), c11 = texel(x1, y1);
        V3 top = add(muls(c00, 1.0f - tx), muls(c10, tx));
        V3 bot = add(muls(c01, 1.0f - tx), muls(c11, tx));
        return muls(add(muls(top, 1.0f - ty), muls(bot, ty)), sk.intensity);
    }
    return v3(0.0f);
}

// ---------------------------------------------------------------------------
// camera (engine/include/engine/camera.h:19-280)
// ---------------------------------------------------------------------------
struct Camera {
    nr_camera_desc p{};
    M3 R;
    V3 uvw;
    float shutter_time, defocus_angle, defocus_radius;
    size_t H, W;
    std::vector<V2> pattern;
};
static std::vector<V2> n_rooks(uint32_t n, uint32_t seed) {                              // camera.h:45-61
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937 rng(seed);
    std::shuffle(perm.begin(), perm.end(), rng);
    std::uniform_real_distribution<float> jitter(0.0f, 1.0f);
    std::vector<V2> out(n);
    for (uint32_t i = 0; i < n; i++) {
        float jy = jitter(rng);     // g++ evaluates the Vector2 arguments right to left
        float jx = jitter(rng);
        out[i] = {(i + jx) / (float)n, (perm[i] + jy) / (float)n};
    }
    return out;
}
static void camera_build(Camera& c, const nr_camera_desc& d, uint32_t seed) {            // camera.h:124-136
    c.p = d;
    c.R = rotation_from_euler({deg2rad(d.rotation[0]), deg2rad(d.rotation[1]), deg2rad(d.rotation[2])});
    c.shutter_time = 1.0f / (d.shutter_speed + FMIN);
    float aperture_radius = d.aperture / 2.0f;
    c.defocus_angle = 2.0f * atanf(aperture_radius / (d.focus_distance + FMIN));
    c.defocus_radius = d.focus_distance * tanf(c.defocus_angle / 2.0f);
    float aspect = d.resolution[1] / d.resolution[0];
    c.uvw = {d.sensor_width, -d.sensor_width * aspect, 0.0f};
    c.H = (size_t)d.resolution[0];
    c.W = (size_t)d.resolution[1];
    c.pattern = n_rooks(d.samples_per_pixel, seed);
}
static Ray camera_ray(const Camera& c, uint32_t x, uint32_t y, uint32_t pixel, uint32_t s, uint32_t seed, Gen& g) {
    // sample_offset (camera.h:247-259)
    V2 base = c.pattern[s];
    V2 shift = pcg_vector2((float)pixel, seed);
    V2 off = {fmodf(base.x + shift.x, 1.0f) - 0.5f, fmodf(base.y + shift.y, 1.0f) - 0.5f};
    // get_focus_point (camera.h:261-278)
    float rx = c.p.resolution[0], ry = c.p.resolution[1];
    float u = ((float)x + off.x - (rx - 1.0f) * 0.5f) * (c.uvw.x / rx);
    float v = ((float)y + off.y - (ry - 1.0f) * 0.5f) * (c.uvw.y / ry);
    float w = -c.p.focal_length;
    V3 center = mv(c.R, {u, v, w});
    float focus_scale = c.p.focus_distance / c.p.focal_length;
    V3 pos = {c.p.position[0], c.p.position[1], c.p.position[2]};
    V3 focus = add(pos, muls(center, focus_scale));
    // get_ray (camera.h:186-213)
    V3 origin;
    if (c.defocus_angle <= 0.0f) origin = pos;
    else {
        V3 du = muls(right(c.R), c.defocus_radius);
        V3 dv = muls(up(c.R), c.defocus_radius);
        V2 loc;
        while (true) {                                                                   // vector.h:303-311
            float a = g.random_float(-1.0f, 1.0f);
            float b = g.random_float(-1.0f, 1.0f);
            loc = {a, b};
            if (loc.x * loc.x + loc.y * loc.y < 1.0f) break;
        }
        V3 defocus = add(muls(du, loc.x), muls(dv, loc.y));
        origin = add(pos, defocus);
    }
    float time = c.shutter_time > 0.0f ? g.uniform() * c.shutter_time : 0.0f;
    return make_ray(origin, normalize(sub(focus, origin)), time);
}

// ---------------------------------------------------------------------------
// path (engine/src/trace.cu:11-79, engine/src/brdf.cu:4-35)
// ---------------------------------------------------------------------------
struct PathOut { V3 beauty{0, 0, 0}, diffuse{0, 0, 0}, normal{0, 0, 0}; float depth = 0.0f; int64_t segments = 0; };

// sample_cosine_brdf (brdf.cu:4-35): with an ObjectLight in the scene, the 50/50
// MixturePDF of the material's pdf and lights[0] (pdf.h:262-283), else the material's
static void sample_brdf(const Scene& S, const Scatter& sr, V3 p, Gen& g, V3& dir, float& pdf_val) {
    if (S.first_light >= 0) {
        const Object& L = S.objects[S.first_light];
        if (g.random_float() < 0.5f) dir = sr.pdf.generate(g);                          // pdf.h:274-277
        else dir = light_random(L, p, g);
        pdf_val = 0.5f * sr.pdf.value(dir) + 0.5f * light_pdf_value(L, p, dir);
    } else {
        dir = sr.pdf.generate(g);
        pdf_val = sr.pdf.value(dir);
    }
}

static void trace_path(const Scene& S, const Camera& c, const nr_render_desc& d, uint32_t s,
                       uint32_t x, uint32_t y, PathOut& out, Counters* cnt) {
    uint32_t pixel = y * (uint32_t)c.W + x;
    V3 atten = v3(1.0f);
    Gen g{d.seed, (uint32_t)((size_t)pixel + (size_t)s * (c.W * c.H)), 0u};
    Ray ray = camera_ray(c, x, y, pixel, s, d.seed, g);
    for (uint32_t bounce = 0; bounce < d.ray_depth; bounce++) {
        HitRec rec;
        out.segments++;
        // trace_ray (trace.cu:11-41)
        bool alive;
        bool hit = scene_hit(S, ray, Interval{EPS, FMAX}, rec, cnt);
        if (!hit) {
            out.beauty = add(out.beauty, mul(atten, sky_sample(S, ray)));
            alive = false;
        } else {
            if (bounce == 0) {      // DEPTH: distance from the camera to the first hit (no reference writer)
                const V3 e = sub(rec.p, ray.o);
                out.depth = std::sqrt(((e.x * e.x) + (e.y * e.y)) + (e.z * e.z));
            }
            const Material& m = S.materials[rec.material_id];
            Ray r_in = ray;
            out.beauty = add(out.beauty, mul(atten, mat_emitted(S, m, rec)));
            Scatter sr;
            if (!mat_scatter(S, m, rec, ray, sr, g)) {
                alive = false;
            } else if (sr.skip_pdf) {
                atten = mul(atten, sr.atten);
                ray = sr.skip_ray;
                alive = true;
            } else {
                V3 dir; float pdf_val;
                sample_brdf(S, sr, rec.p, g, dir, pdf_val);
                if (pdf_val <= 0.0f) {
                    alive = false;
                } else {
                    ray = make_ray(rec.p, dir, r_in.tm);
                    V3 brdf = mat_evaluate(S, m, rec, normalize(neg(r_in.d)), dir);
                    atten = mul(atten, divs(brdf, pdf_val));
                    alive = true;
                }
            }
        }
        if (!alive) break;
        if (bounce == 0) {                                                               // trace.cu:67-70
            out.diffuse = add(out.diffuse, atten);
            out.normal = add(out.normal, rec.n);
        }
        if (d.do_rr && bounce >= d.rr_start_bounce) {                                    // trace.cu:72-77
            float survival = fmaxf(atten.x, fmaxf(atten.y, atten.z));
            survival = fminf(survival, d.rr_survival);
            if (g.uniform() > survival) break;
            atten = mul(atten, v3(1.0f / survival));
        }
    }
}

}  // namespace oro

using namespace oro;

struct nro_scene { Scene s; };
struct nro_camera { Camera c; };

extern "C" {

nro_scene* nro_scene_build(const nr_scene_desc* d) {
    try {
        auto* out = new nro_scene();
        Scene& S = out->s;
        // cube face transforms (shapes.h:192-223)
        const float fp[6][3] = {{0, 0, 0.5f}, {0, 0, -0.5f}, {0.5f, 0, 0}, {-0.5f, 0, 0}, {0, 0.5f, 0}, {0, -0.5f, 0}};
        const float fr[6][3] = {{90, 0, 0}, {-90, 0, 0}, {0, 0, -90}, {0, 0, 90}, {0, 0, 0}, {180, 0, 0}};
        for (int i = 0; i < 6; i++)
            S.cube_faces[i] = make_xform({fp[i][0], fp[i][1], fp[i][2]}, {fr[i][0], fr[i][1], fr[i][2]}, v3(1.0f));
        for (int i = 0; i < d->n_images; i++) {
            Image im; im.C = d->images[i].C; im.H = d->images[i].H; im.W = d->images[i].W;
            im.px.assign(d->images[i].data, d->images[i].data + (size_t)im.C * im.H * im.W);
            S.images.push_back(std::move(im));
        }
        for (int i = 0; i < d->n_materials; i++) {
            const nr_material_desc& md = d->materials[i];
            Material m; m.type = md.type;
            for (int k = 0; k < 5; k++) {
                m.tex[k].image = md.tex[k].image;
                m.tex[k].color = {md.tex[k].color[0], md.tex[k].color[1], md.tex[k].color[2]};
                m.tex[k].scale = md.tex[k].scale;
            }
            m.tint = {md.tint[0], md.tint[1], md.tint[2]}; m.ior = md.ior; m.brightness = md.brightness;
            S.materials.push_back(m);
        }
        // meshes: triangle BVH over object-space triangle boxes (mesh.h:177-216, 320-336)
        for (int i = 0; i < d->n_meshes; i++) {
            const nr_mesh_desc& md = d->meshes[i];
            Mesh M;
            M.verts.resize(md.n_vertices);
            for (int v = 0; v < md.n_vertices; v++) {
                const float* f = md.vertices + 11 * (size_t)v;
                M.verts[v] = {{f[0], f[1], f[2]}, {f[3], f[4], f[5]}, {f[6], f[7], f[8]}, {f[9], f[10]}};
            }
            std::vector<AABB> boxes(md.n_triangles);
            for (int t = 0; t < md.n_triangles; t++) {
                V3 a = M.verts[md.triangles[3 * t]].pos, b = M.verts[md.triangles[3 * t + 1]].pos,
                   c = M.verts[md.triangles[3 * t + 2]].pos;
                V3 mn = {fminf(a.x, fminf(b.x, c.x)), fminf(a.y, fminf(b.y, c.y)), fminf(a.z, fminf(b.z, c.z))};
                V3 mx = {fmaxf(a.x, fmaxf(b.x, c.x)), fmaxf(a.y, fmaxf(b.y, c.y)), fmaxf(a.z, fmaxf(b.z, c.z))};
                boxes[t] = aabb_pts(mn, mx);
            }
            BVHBuilder bb; bb.build(boxes);
            M.nodes = bb.nodes;
            M.tris.resize(3 * (size_t)md.n_triangles);
            M.tri_orig.resize(md.n_triangles);
            for (int t = 0; t < md.n_triangles; t++) {
                int src = bb.order[t];
                M.tri_orig[t] = src;
                for (int k = 0; k < 3; k++) M.tris[3 * t + k] = md.triangles[3 * src + k];
            }
            // depth
            std::vector<std::pair<int, int>> st{{0, 1}};
            while (!st.empty()) {
                auto [n, dd] = st.back(); st.pop_back();
                M.depth = std::max(M.depth, dd);
                if (M.nodes[n].object == -1) { st.push_back({M.nodes[n].left, dd + 1}); st.push_back({M.nodes[n].right, dd + 1}); }
            }
            S.meshes.push_back(std::move(M));
        }
        // objects in registration order, then the top-level BVH reorders them
        std::vector<Object> reg(d->n_objects);
        std::vector<AABB> wboxes(d->n_objects);
        for (int i = 0; i < d->n_objects; i++) {
            const nr_object_desc& od = d->objects[i];
            Object& o = reg[i];
            o.type = od.type; o.bound_type = od.bound_type; o.material = od.material; o.mesh = od.mesh;
            o.object_id = i;
            int geom = (od.type == NR_HIT_CONSTANT_MEDIUM || od.type == NR_HIT_OBJECT_LIGHT) ? od.bound_type : od.type;
            if (geom == NR_HIT_SPHERE) o.xf = xform_identity_at({od.position[0], od.position[1], od.position[2]});
            else o.xf = make_xform({od.position[0], od.position[1], od.position[2]},
                                   {od.rotation[0], od.rotation[1], od.rotation[2]},
                                   {od.scale[0], od.scale[1], od.scale[2]});
            o.radius = od.radius + FMIN;
            o.neg_inv_density = -1.0f / (od.density + FMIN);
            switch (geom) {                                                              // build_bbox per type
                case NR_HIT_QUAD: case NR_HIT_CUBE: o.local_box = aabb_pts(v3(-0.5f), v3(0.5f)); break;
                case NR_HIT_SPHERE: o.local_box = aabb_pts(v3(-o.radius), v3(o.radius)); break;
                case NR_HIT_MESH: {                                                      // mesh.h:218-237
                    const Mesh& M = S.meshes[o.mesh];
                    V3 mn = v3(FMAX), mx = v3(-FMAX);
                    for (const auto& v : M.verts) {
                        mn = {fminf(mn.x, v.pos.x), fminf(mn.y, v.pos.y), fminf(mn.z, v.pos.z)};
                        mx = {fmaxf(mx.x, v.pos.x), fmaxf(mx.y, v.pos.y), fmaxf(mx.z, v.pos.z)};
                    }
                    o.local_box = M.verts.empty() ? aabb_empty() : aabb_pts(mn, mx);
                    break;
                }
                default: o.local_box = aabb_empty();
            }
            wboxes[i] = aabb_transformed(o.local_box, o.xf);                             // hittable.h:294-296
        }
        BVHBuilder tb; tb.build(wboxes);
        S.nodes = tb.nodes;
        for (size_t i = 0; i < tb.order.size(); i++) S.objects.push_back(reg[tb.order[i]]);
        for (size_t i = 0; i < S.objects.size(); i++)                                    // registry.h:240-244
            if (S.objects[i].type == NR_HIT_OBJECT_LIGHT) { S.first_light = (int)i; break; }
        S.sky = d->sky;
        if (d->sky.type == NR_SKY_HDRI && d->sky.data) {
            S.hdri.assign(d->sky.data, d->sky.data + (size_t)d->sky.C * d->sky.H * d->sky.W);
        }
        S.sky.data = nullptr;
        return out;
    } catch (...) {
        return nullptr;
    }
}

void nro_scene_free(nro_scene* s) { delete s; }
int nro_scene_first_light(const nro_scene* s) { return s->s.first_light; }
int nro_scene_mesh_depth(const nro_scene* s, int mesh) { return s->s.meshes.at(mesh).depth; }

static void fill_hit(const HitRec& r, bool hit, nr_hit* h) {
    if (!hit) { std::memset(h, 0, sizeof(*h)); h->object_id = -1; h->prim = -1; return; }
    h->object_id = r.object_id; h->material_id = r.material_id; h->prim = r.prim;
    h->front_face = r.front_face ? 1 : 0; h->t = r.t;
    h->p[0] = r.p.x; h->p[1] = r.p.y; h->p[2] = r.p.z;
    h->n[0] = r.n.x; h->n[1] = r.n.y; h->n[2] = r.n.z;
    h->tangent[0] = r.tangent.x; h->tangent[1] = r.tangent.y; h->tangent[2] = r.tangent.z;
    h->uv[0] = r.uv.x; h->uv[1] = r.uv.y;
}

int nro_trace_closest(const nro_scene* s, const float* rays, int n, nr_hit* hits, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    int err = 0;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; i++) {
        try {
            const float* r = rays + 6 * (size_t)i;
            Ray ray = make_ray({r[0], r[1], r[2]}, {r[3], r[4], r[5]});
            HitRec rec;
            bool h = scene_hit(s->s, ray, Interval{EPS, FMAX}, rec, nullptr);
            fill_hit(rec, h, &hits[i]);
        } catch (...) {
#pragma omp atomic write
            err = 1;
        }
    }
    return err;
}

int nro_trace_counts(const nro_scene* s, const float* rays, int n, int64_t* counts) {
    Counters c;
    for (int i = 0; i < n; i++) {
        const float* r = rays + 6 * (size_t)i;
        Ray ray = make_ray({r[0], r[1], r[2]}, {r[3], r[4], r[5]});
        HitRec rec;
        scene_hit(s->s, ray, Interval{EPS, FMAX}, rec, &c);
    }
    counts[0] = c.inner; counts[1] = c.tri; counts[2] = c.obj;
    return 0;
}

nro_camera* nro_camera_build(const nr_camera_desc* cam, uint32_t seed) {
    auto* c = new nro_camera();
    camera_build(c->c, *cam, seed);
    return c;
}
void nro_camera_free(nro_camera* c) { delete c; }

int nro_camera_ray(const nro_camera* c, uint32_t x, uint32_t y, uint32_t s, uint32_t seed, float* out7) {
    if (s >= c->c.pattern.size()) return 1;
    uint32_t pixel = y * (uint32_t)c->c.W + x;
    Gen g{seed, (uint32_t)((size_t)pixel + (size_t)s * (c->c.W * c->c.H)), 0u};
    Ray r = camera_ray(c->c, x, y, pixel, s, seed, g);
    out7[0] = r.o.x; out7[1] = r.o.y; out7[2] = r.o.z;
    out7[3] = r.d.x; out7[4] = r.d.y; out7[5] = r.d.z; out7[6] = r.tm;
    return 0;
}

int nro_nrooks(uint32_t n, uint32_t seed, float* out) {
    auto p = n_rooks(n, seed);
    for (uint32_t i = 0; i < n; i++) { out[2 * i] = p[i].x; out[2 * i + 1] = p[i].y; }
    return 0;
}

int nro_render_pass(const nro_scene* s, const nro_camera* c, const nr_render_desc* d, uint32_t sample_idx,
                    const int32_t* pixels, int n_pixels, float* beauty, float* diffuse, float* normal, float* depth,
                    int64_t* counts, int threads) {
    const Camera& cam = c->c;
    if (sample_idx >= cam.pattern.size()) return 1;
    size_t HW = cam.H * cam.W;
    int n = pixels ? n_pixels : (int)HW;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    int64_t seg = 0, inner = 0, tri = 0, obj = 0;
    int err = 0;
    // 16x16 tiles when rendering the full frame (schedule(dynamic) over tiles)
    const int TS = 16;
    int tiles_x = (int)((cam.W + TS - 1) / TS), tiles_y = (int)((cam.H + TS - 1) / TS);
    int n_items = pixels ? n : tiles_x * tiles_y;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : seg, inner, tri, obj)
    for (int it = 0; it < n_items; it++) {
        try {
            auto run = [&](uint32_t x, uint32_t y) {
                PathOut po;
                Counters cc;
                trace_path(s->s, cam, *d, sample_idx, x, y, po, counts ? &cc : nullptr);
                size_t px = (size_t)y * cam.W + x;
                if (beauty) { beauty[px] = po.beauty.x; beauty[HW + px] = po.beauty.y; beauty[2 * HW + px] = po.beauty.z; }
                if (diffuse) { diffuse[px] = po.diffuse.x; diffuse[HW + px] = po.diffuse.y; diffuse[2 * HW + px] = po.diffuse.z; }
                if (normal) { normal[px] = po.normal.x; normal[HW + px] = po.normal.y; normal[2 * HW + px] = po.normal.z; }
                if (depth) depth[px] = po.depth;
                seg += po.segments; inner += cc.inner; tri += cc.tri; obj += cc.obj;
            };
            if (pixels) {
                uint32_t px = (uint32_t)pixels[it];
                run(px % (uint32_t)cam.W, px / (uint32_t)cam.W);
            } else {
                int tx = it % tiles_x, ty = it / tiles_x;
                for (int y = ty * TS; y < std::min((int)cam.H, ty * TS + TS); y++)
                    for (int x = tx * TS; x < std::min((int)cam.W, tx * TS + TS); x++) run(x, y);
            }
        } catch (...) {
#pragma omp atomic write
            err = 1;
        }
    }
    if (counts) { counts[0] = seg; counts[1] = inner; counts[2] = tri; counts[3] = obj; counts[4] = n; }
    return err;
}

void nro_accumulate(float* avg, const float* sample, size_t n, uint32_t current_sample) {
    float fn = (float)current_sample;                                                    // data.cu:28-40
    float a = (fn - 1.0f) / fn, b = 1.0f / fn;
    for (size_t i = 0; i < n; i++) avg[i] = avg[i] * a + sample[i] * b;
}

int nro_render(const nro_scene* s, const nro_camera* c, const nr_render_desc* d,
               float* beauty, float* diffuse, float* normal, float* depth, int threads) {
    size_t n = 3 * c->c.H * c->c.W;
    std::vector<float> sb(n), sd(n), sn(n), sz(n / 3);
    for (uint32_t si = d->sample_begin; si < d->sample_end; si++) {                      // engine.cpp:42-74
        int e = nro_render_pass(s, c, d, si, nullptr, 0, sb.data(), sd.data(), sn.data(), sz.data(), nullptr, threads);
        if (e) return e;
        if (beauty) nro_accumulate(beauty, sb.data(), n, si);
        if (diffuse) nro_accumulate(diffuse, sd.data(), n, si);
        if (normal) nro_accumulate(normal, sn.data(), n, si);
        if (depth) nro_accumulate(depth, sz.data(), n / 3, si);
    }
    return 0;
}

// post: data/src/stream.cu:7-159
static V3 kelvin_to_rgb(float kelvin) {                                                  // stream.cu:65-86
    float temp = kelvin / 100.0f;
    float r, g, b;
    if (temp <= 66.0f) { r = 255.0f; g = 99.4708025861f * logf(temp) - 161.1195681661f; }
    else { r = 329.698727446f * powf(temp - 60.0f, -0.1332047592f); g = 288.1221695283f * powf(temp - 60.0f, -0.0755148492f); }
    if (temp >= 66.0f) b = 255.0f;
    else if (temp <= 19.0f) b = 0.0f;
    else b = 138.5177312231f * logf(temp - 10.0f) - 305.0447927307f;
    return {fminf(1.0f, fmaxf(0.0f, r / 255.0f)), fminf(1.0f, fmaxf(0.0f, g / 255.0f)), fminf(1.0f, fmaxf(0.0f, b / 255.0f))};
}
static inline float clamp01(float x) { return fminf(1.0f, fmaxf(0.0f, x)); }

// ---------------------------------------------------------------------------
// viewport preview (engine/src/trace.cu:99-143, engine/src/axis_grid.cu),
// picking (trace.cu:166-177) and selection (scene/src/scene_interface.cu:3-50),
// overlay (data/src/stream.cu:177-202)
// ---------------------------------------------------------------------------
static inline float cg_smoothstep(float a, float b, float x) {                             // cg_std.h:97-103
    float t = fmaxf(0.0f, fminf(1.0f, (x - a) / (b - a)));                                 // saturate, cg_std.h:82-85
    return t * t * (3.0f - (2.0f * t));
}
static inline V3 cg_lerp(V3 x, V3 y, float w) { return add(x, muls(sub(y, x), w)); }      // cg_std.h:90-93

static float grid_mask(float u, float v, float spacing, float width) {                    // axis_grid.cu:4-10
    float du = fabsf(fmodf(u + spacing * 0.5f, spacing) - spacing * 0.5f);
    float dv = fabsf(fmodf(v + spacing * 0.5f, spacing) - spacing * 0.5f);
    float d = fminf(du, dv);
    return 1.0f - cg_smoothstep(0.0f, width, d);
}

static bool axis_grid(const Ray& ray, float max_t, int axis, V3 major, V3& gc, float& ga) { // axis_grid.cu:12-51
    const float line_width = 1e-3f;
    float dir_axis = get(ray.d, axis);
    if (fabsf(dir_axis) < 1e-6f) return false;
    float t = -get(ray.o, axis) / dir_axis;
    if (t <= 0.0f || t >= max_t) return false;
    V3 p = at(ray, t);
    p = {fabsf(p.x), fabsf(p.y), fabsf(p.z)};
    int u_axis = (axis + 1) % 3, v_axis = (axis + 2) % 3;
    float u = get(p, u_axis), v = get(p, v_axis);
    float pixel_size = line_width * t;
    // log10f and powf(10, level0) through the shared transcendentals (level0 is whole)
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
    gc = cg_lerp(v3(0.35f), major, cg_smoothstep(0.0f, 1.0f, g1));
    ga = mask * fade;
    return true;
}

static bool triplanar_grid(const Ray& ray, float max_t, V3& gc, float& ga) {             // axis_grid.cu:53-70
    bool hit_grid = false;
    hit_grid = hit_grid || axis_grid(ray, max_t, 0, {1.0f, 0.3f, 0.3f}, gc, ga);
    hit_grid = hit_grid || axis_grid(ray, max_t, 1, {0.3f, 1.0f, 0.3f}, gc, ga);
    hit_grid = hit_grid || axis_grid(ray, max_t, 2, {0.3f, 0.3f, 1.0f}, gc, ga);
    return hit_grid;
}

static V3 viewport_color(const Scene& S, const Material& m, const HitRec& rec) {          // material.h:588-596
    switch (m.type) {
        case NR_MAT_LAMBERTIAN: return tex_sample(S, m.tex[0], rec.uv);                   // material.h:43-45
        case NR_MAT_PBR: return tex_sample(S, m.tex[0], rec.uv);                          // material.h:109-111
        case NR_MAT_DIELECTRIC: return m.tint;                                            // material.h:218
        case NR_MAT_EMISSIVE: return tex_sample(S, m.tex[0], rec.uv);                     // material.h:289-291
        case NR_MAT_ISOTROPIC: return tex_sample(S, m.tex[0], rec.uv);                    // material.h:335-337
        default: return v3(0.0f);
    }
}

// DeviceCamera::get_ray_center (camera.h:181-184): get_focus_point(x, y, 0, 0, 0)
static Ray center_ray(const Camera& c, uint32_t x, uint32_t y) {
    V2 base = c.pattern[0];
    V2 shift = pcg_vector2(0.0f, 0u);
    V2 off = {fmodf(base.x + shift.x, 1.0f) - 0.5f, fmodf(base.y + shift.y, 1.0f) - 0.5f};
    float rx = c.p.resolution[0], ry = c.p.resolution[1];
    float u = ((float)x + off.x - (rx - 1.0f) * 0.5f) * (c.uvw.x / rx);
    float v = ((float)y + off.y - (ry - 1.0f) * 0.5f) * (c.uvw.y / ry);
    V3 center = mv(c.R, {u, v, -c.p.focal_length});
    V3 pos = {c.p.position[0], c.p.position[1], c.p.position[2]};
    V3 focus = add(pos, muls(center, c.p.focus_distance / c.p.focal_length));
    return make_ray(pos, normalize(sub(focus, pos)));
}

// DeviceCamera::screen_space_ray (camera.h:215-226)
static Ray screen_space_ray(const Camera& c, float su, float sv) {
    float x = su * c.p.resolution[0];
    float y = sv * c.p.resolution[1];
    float u = (x - (c.p.resolution[0] - 1.0f) * 0.5f) * (c.uvw.x / c.p.resolution[0]);
    float v = (y - (c.p.resolution[1] - 1.0f) * 0.5f) * (c.uvw.y / c.p.resolution[1]);
    V3 direction = normalize(mv(c.R, {u, v, -c.p.focal_length}));
    return make_ray({c.p.position[0], c.p.position[1], c.p.position[2]}, direction);
}

int nro_viewport(const nro_scene* sc, const nro_camera* cc, int show_axis_grid, float* beauty, float* diffuse,
                 float* normal, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    const Scene& S = sc->s;
    const Camera& c = cc->c;
    const size_t H = c.H, W = c.W, HW = H * W;
    int err = 0;
#pragma omp parallel for schedule(dynamic, 4)
    for (long y = 0; y < (long)H; y++) {
        for (size_t x = 0; x < W; x++) {
            try {
                Ray ray = center_ray(c, (uint32_t)x, (uint32_t)y);
                HitRec rec;
                bool hit = scene_hit(S, ray, Interval{EPS, FMAX}, rec, nullptr);
                V3 b, dcol = v3(0.0f), n = v3(0.0f);    // sample AOVs start cleared
                if (!hit) {
                    V3 gc; float ga;
                    V3 sky = sky_sample(S, ray);
                    b = (show_axis_grid && triplanar_grid(ray, FMAX, gc, ga)) ? cg_lerp(sky, gc, ga) : sky;
                } else {
                    V3 light_vector = sub({c.p.position[0], c.p.position[1], c.p.position[2]}, rec.p);
                    float light_dist = length(light_vector);
                    V3 headlight_dir = normalize(light_vector);
                    float headlight = dot(headlight_dir, rec.n) * 0.5f + 0.5f;
                    headlight *= 1.0f / (light_dist * light_dist);
                    headlight = fminf(0.75f, headlight);
                    V3 directional_light_dir = {1.0f, 1.0f, 1.0f};
                    float directional = (dot(directional_light_dir, rec.n) + 1.0f) * 0.5f;
                    directional = fmaxf(0.1f, directional);
                    V3 col = viewport_color(S, S.materials[rec.material_id], rec);
                    b = muls(col, fmaxf(headlight, directional));
                    dcol = col;
                    n = rec.n;
                }
                const size_t i = (size_t)y * W + x;
                beauty[i] = 0.0f + b.x; beauty[HW + i] = 0.0f + b.y; beauty[2 * HW + i] = 0.0f + b.z;
                diffuse[i] = 0.0f + dcol.x; diffuse[HW + i] = 0.0f + dcol.y; diffuse[2 * HW + i] = 0.0f + dcol.z;
                normal[i] = 0.0f + n.x; normal[HW + i] = 0.0f + n.y; normal[2 * HW + i] = 0.0f + n.z;
            } catch (...) {
#pragma omp atomic write
                err = 1;
            }
        }
    }
    return err;
}

int nro_triplanar_grid(const float* ray6, float* out4) {
    Ray r = make_ray({ray6[0], ray6[1], ray6[2]}, {ray6[3], ray6[4], ray6[5]});
    V3 gc{0, 0, 0};
    float ga = 0.0f;
    bool h = triplanar_grid(r, FMAX, gc, ga);
    out4[0] = gc.x; out4[1] = gc.y; out4[2] = gc.z; out4[3] = ga;
    return h ? 1 : 0;
}

int nro_center_ray(const nro_camera* cc, uint32_t x, uint32_t y, float* out6) {
    Ray r = center_ray(cc->c, x, y);
    out6[0] = r.o.x; out6[1] = r.o.y; out6[2] = r.o.z;
    out6[3] = r.d.x; out6[4] = r.d.y; out6[5] = r.d.z;
    return 0;
}

int nro_screen_ray(const nro_camera* cc, float su, float sv, float* out6) {
    Ray r = screen_space_ray(cc->c, su, sv);
    out6[0] = r.o.x; out6[1] = r.o.y; out6[2] = r.o.z;
    out6[3] = r.d.x; out6[4] = r.d.y; out6[5] = r.d.z;
    return 0;
}

int nro_pick(const nro_scene* sc, const nro_camera* cc, float su, float sv, nr_hit* out) {
    Ray r = screen_space_ray(cc->c, su, sv);
    HitRec rec;
    bool h = scene_hit(sc->s, r, Interval{EPS, FMAX}, rec, nullptr);
    fill_hit(rec, h, out);
    return 0;
}

int nro_selection_mask(const nro_scene* sc, const nro_camera* cc, int32_t object_id, int radius, uint8_t* out) {
    const Camera& c = cc->c;
    const size_t H = c.H, W = c.W;
    std::vector<uint8_t> base(H * W);
#pragma omp parallel for schedule(dynamic, 4)
    for (long y = 0; y < (long)H; y++)
        for (size_t x = 0; x < W; x++) {                                                  // mask_selected_kernel
            float su = (x + 0.5f) / (float)W;
            float sv = (y + 0.5f) / (float)H;
            Ray r = screen_space_ray(c, su, sv);
            HitRec rec;
            bool hit = scene_hit(sc->s, r, Interval{EPS, FMAX}, rec, nullptr);
            base[(size_t)y * W + x] = (hit && rec.object_id == object_id) ? 255u : 0u;
        }
    for (size_t y = 0; y < H; y++)
        for (size_t x = 0; x < W; x++) {                                                  // build_outline_kernel
            size_t idx = y * W + x;
            out[idx] = 0u;
            if (base[idx] != 0) continue;
            bool found = false;
            for (int dy = -radius; dy <= radius && !found; dy++)
                for (int dx = -radius; dx <= radius && !found; dx++) {
                    long nx = (long)x + dx, ny = (long)y + dy;
                    if (nx < 0 || ny < 0 || nx >= (long)W || ny >= (long)H) continue;
                    if (base[(size_t)ny * W + (size_t)nx] != 0) found = true;
                }
            if (found) out[idx] = 255u;
        }
    return 0;
}

int nro_overlay(uint8_t* img, const uint8_t* mask, const float* color, int H, int W) {
    uint32_t rgb[3];
    for (int k = 0; k < 3; k++) rgb[k] = (uint8_t)(fminf(1.0f, fmaxf(0.0f, color[k])) * 255.0f);  // stream.cu:214-222
    const size_t HW = (size_t)H * W;
    for (size_t i = 0; i < HW; i++) {
        uint32_t alpha = mask[i];
        if (alpha == 0u) continue;
        uint32_t inv = 255u - alpha;
        for (int k = 0; k < 3; k++) img[k * HW + i] = (uint8_t)((rgb[k] * alpha + img[k * HW + i] * inv + 127u) / 255u);
    }
    return 0;
}

int nro_post(const float* chw, int H, int W, const nr_post_desc* d, float* out_f32, uint8_t* out_u8) {
    size_t HW = (size_t)H * W;
    V3 ref = kelvin_to_rgb(6500.0f);
    V3 tgt = kelvin_to_rgb(fmaxf(1000.0f, fminf(d->wb_temperature, 40000.0f)));
    V3 ratio = {tgt.x / fmaxf(ref.x, 1e-4f), tgt.y / fmaxf(ref.y, 1e-4f), tgt.z / fmaxf(ref.z, 1e-4f)};
    float tint = d->wb_tint * 0.0025f;
    for (size_t i = 0; i < HW; i++) {
        V3 c = {chw[i], chw[HW + i], chw[2 * HW + i]};
        if (d->apply_white_balance) {                                                    // stream.cu:88-109
            c = mul(c, ratio);
            c = {c.x * (1.0f + tint * 0.5f), c.y * (1.0f - tint), c.z * (1.0f + tint * 0.5f)};
        }
        if (d->apply_tonemapping) {                                                      // stream.cu:34-59
            V3 col = c;
            if (d->tm_method == 0) {
                if (d->fix_reinhard) col = divv(col, adds(col, 1.0f));   // stream.cu:7-11 discards this
            } else if (d->tm_method == 1) {                                              // stream.cu:13-20
                float inv_w2 = 1.0f / (d->tm_white_point * d->tm_white_point);
                V3 num = mul(col, add(v3(1.0f), muls(col, inv_w2)));
                col = divv(num, add(v3(1.0f), col));
            } else {                                                                     // stream.cu:22-32
                const float a = 2.51f, b = 0.03f, cc = 2.43f, dd = 0.59f, e = 0.14f;
                V3 num = mul(col, adds(muls(col, a), b));
                V3 den = adds(mul(col, adds(muls(col, cc), dd)), e);
                col = {clamp01(num.x / den.x), clamp01(num.y / den.y), clamp01(num.z / den.z)};
            }
            float alpha = d->tm_alpha;
            c = add(muls(c, 1.0f - alpha), muls(col, alpha));
            c = {clamp01(c.x), clamp01(c.y), clamp01(c.z)};
        }
        if (d->linear_to_gamma)                                                          // stream.cu:115-121
            c = {c.x > 0.0f ? sqrtf(c.x) : 0.0f, c.y > 0.0f ? sqrtf(c.y) : 0.0f, c.z > 0.0f ? sqrtf(c.z) : 0.0f};
        float v[3] = {c.x, c.y, c.z};
        for (int k = 0; k < 3; k++) {                                                    // stream.cu:123-140
            float x = v[k];
            if (std::isnan(x)) x = 0.0f;
            x = fmaxf(0.0f, fminf(1.0f, x));
            if (out_f32) out_f32[k * HW + i] = x;
            if (out_u8) out_u8[k * HW + i] = (uint8_t)(x * 255.0f);
        }
    }
    return 0;
}

// ---- single-vertex probes (closed-form tests of the pieces a path is made of) ----
static HitRec rec_from(const nr_hit* h) {
    HitRec r;
    r.object_id = h->object_id; r.material_id = h->material_id; r.prim = h->prim;
    r.p = {h->p[0], h->p[1], h->p[2]}; r.n = {h->n[0], h->n[1], h->n[2]};
    r.tangent = {h->tangent[0], h->tangent[1], h->tangent[2]};
    r.uv = {h->uv[0], h->uv[1]}; r.t = h->t; r.front_face = h->front_face != 0;
    return r;
}
static const Object* object_by_id(const Scene& S, int object_id) {
    for (const Object& o : S.objects)
        if (o.object_id == object_id) return &o;
    return nullptr;
}

int nro_material_evaluate(const nro_scene* s, int material, const nr_hit* rec, const float* view3, const float* light3,
                          float* out3) {
    const Scene& S = s->s;
    if (material < 0 || material >= (int)S.materials.size()) return -1;
    V3 f = mat_evaluate(S, S.materials[material], rec_from(rec), {view3[0], view3[1], view3[2]},
                        {light3[0], light3[1], light3[2]});
    out3[0] = f.x; out3[1] = f.y; out3[2] = f.z;
    return 0;
}

int nro_material_sample(const nro_scene* s, int material, const nr_hit* rec, const float* ray6, uint32_t seed,
                        uint32_t subseq, uint32_t offset, float* dir3, float* pdf, float* atten3, int* kind) {
    const Scene& S = s->s;
    if (material < 0 || material >= (int)S.materials.size()) return -1;
    const HitRec r = rec_from(rec);
    const Ray ray = make_ray({ray6[0], ray6[1], ray6[2]}, {ray6[3], ray6[4], ray6[5]});
    Gen g{seed, subseq, offset};
    Scatter sr;
    V3 dir{0, 0, 0};
    float pv = 0.0f;
    if (!mat_scatter(S, S.materials[material], r, ray, sr, g)) {
        *kind = 0;
    } else if (sr.skip_pdf) {
        *kind = 1;
        dir = sr.skip_ray.d;
        pv = -1.0f;
    } else {
        *kind = 2;
        sample_brdf(S, sr, r.p, g, dir, pv);
    }
    dir3[0] = dir.x; dir3[1] = dir.y; dir3[2] = dir.z;
    *pdf = pv;
    atten3[0] = sr.atten.x; atten3[1] = sr.atten.y; atten3[2] = sr.atten.z;
    return (int)g.draw;
}

float nro_light_pdf(const nro_scene* s, int object_id, const float* origin3, const float* dir3) {
    const Object* o = object_by_id(s->s, object_id);
    if (!o) return -1.0f;
    return light_pdf_value(*o, {origin3[0], origin3[1], origin3[2]}, {dir3[0], dir3[1], dir3[2]});
}

int nro_light_random(const nro_scene* s, int object_id, const float* origin3, uint32_t seed, uint32_t subseq,
                     uint32_t offset, float* out3) {
    const Object* o = object_by_id(s->s, object_id);
    if (!o) return -1;
    Gen g{seed, subseq, offset};
    V3 d = light_random(*o, {origin3[0], origin3[1], origin3[2]}, g);
    out3[0] = d.x; out3[1] = d.y; out3[2] = d.z;
    return 0;
}

int nro_texture_sample(const nro_scene* s, int material, int slot, float u, float v, float* out3) {
    const Scene& S = s->s;
    if (material < 0 || material >= (int)S.materials.size() || slot < 0 || slot >= 5) return -1;
    V3 c = tex_sample(S, S.materials[material].tex[slot], {u, v});
    out3[0] = c.x; out3[1] = c.y; out3[2] = c.z;
    return 0;
}

void nro_hash_constants(uint32_t* out64) { std::memcpy(out64, HASH_CONSTANTS, sizeof(HASH_CONSTANTS)); }

void nro_philox_block(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox10(ctr, key, out); }
float nro_philox_uniform(uint32_t seed, uint32_t subseq, uint32_t offset) { return philox_uniform(seed, subseq, offset); }
uint32_t nro_pcg_hash_int(uint32_t i, uint32_t seed, uint32_t stream) { return pcg_hash_int(i, seed, stream); }
float nro_pcg_float(float f, uint32_t seed, uint32_t stream) { return pcg_float(f, seed, stream); }
float nro_pcg_float3(const float* a, uint32_t seed, uint32_t stream) { return pcg_float3({a[0], a[1], a[2]}, seed, stream); }
float nro_pcg_float_ray(const float* o, const float* d) { return pcg_float_ray(make_ray({o[0], o[1], o[2]}, {d[0], d[1], d[2]})); }
void nro_pcg_vector2(float f, uint32_t seed, float* out2) { V2 v = pcg_vector2(f, seed); out2[0] = v.x; out2[1] = v.y; }

int nro_bvh_build(const float* boxes, int n, int32_t* nodes_out, float* node_boxes, int32_t* order_out) {
    std::vector<AABB> b(n);
    for (int i = 0; i < n; i++)
        b[i] = aabb_pts({boxes[6 * i], boxes[6 * i + 1], boxes[6 * i + 2]}, {boxes[6 * i + 3], boxes[6 * i + 4], boxes[6 * i + 5]});
    BVHBuilder bb; bb.build(b);
    for (size_t k = 0; k < bb.nodes.size(); k++) {
        nodes_out[3 * k] = bb.nodes[k].left; nodes_out[3 * k + 1] = bb.nodes[k].right; nodes_out[3 * k + 2] = bb.nodes[k].object;
        for (int a = 0; a < 3; a++) { node_boxes[6 * k + a] = bb.nodes[k].box.a[a].min; node_boxes[6 * k + 3 + a] = bb.nodes[k].box.a[a].max; }
    }
    for (int i = 0; i < n; i++) order_out[i] = bb.order[i];
    return (int)bb.nodes.size();
}

int nro_sky_sample(const nro_scene* s, const float* dir3, float* out3) {
    V3 c = sky_sample(s->s, make_ray({0, 0, 0}, {dir3[0], dir3[1], dir3[2]}));
    out3[0] = c.x; out3[1] = c.y; out3[2] = c.z;
    return 0;
}

float nro_ggx_distribution(float n_dot_h, float alpha) { return ggx_D(n_dot_h, alpha); }
float nro_pdf_pbr_value(const float* n, const float* v, float rough, float sp, const float* d) {
    return pdf_pbr({n[0], n[1], n[2]}, {v[0], v[1], v[2]}, rough, sp).value({d[0], d[1], d[2]});
}
int nro_pdf_pbr_generate(const float* n, const float* v, float rough, float sp, uint32_t seed, uint32_t subseq,
                         uint32_t offset, float* out3) {
    Gen g{seed, subseq, offset};
    V3 d = pdf_pbr({n[0], n[1], n[2]}, {v[0], v[1], v[2]}, rough, sp).generate(g);
    out3[0] = d.x; out3[1] = d.y; out3[2] = d.z;
    return (int)(g.draw - offset);
}
float nro_pdf_cosine_value(const float* n, const float* d) { return pdf_cosine({n[0], n[1], n[2]}).value({d[0], d[1], d[2]}); }
int nro_pdf_cosine_generate(const float* n, uint32_t seed, uint32_t subseq, uint32_t offset, float* out3) {
    Gen g{seed, subseq, offset};
    V3 d = pdf_cosine({n[0], n[1], n[2]}).generate(g);
    out3[0] = d.x; out3[1] = d.y; out3[2] = d.z;
    return (int)(g.draw - offset);
}

// deterministic transcendentals shared with the HIP kernels (nr_mathfn.h):
// which = 0 sin, 1 cos, 2 acos, 3 atan2(x, y), 4 log, 5 pow(x, y)
void nro_mathfn(int which, const float* x, const float* y, float* out, int n) {
    for (int i = 0; i < n; i++) {
        switch (which) {
            case 0: out[i] = nrm::sinf_(x[i]); break;
            case 1: out[i] = nrm::cosf_(x[i]); break;
            case 2: out[i] = nrm::acosf_(x[i]); break;
            case 3: out[i] = nrm::atan2f_(x[i], y[i]); break;
            case 4: out[i] = nrm::logf_(x[i]); break;
            case 5: out[i] = nrm::powf_(x[i], y[i]); break;
        }
    }
}

}  // extern "C"
