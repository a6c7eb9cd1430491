// nr_types.h — device-resident scene tables of the MI355X path tracer.
//
// Layout in HBM (all tables are flat arrays, one allocation each):
//   GNode    64 B   binary BVH node holding BOTH child boxes + child refs, so one
//                   interior visit is one 64 B line (4 x dwordx4 loads)
//   GWNode  128 B   4-wide mesh node: the grandchildren of binary node i at
//                   index i, boxes SoA, one 128 B line (8 x dwordx4)
//   GQNode   64 B   the same 4-wide node with its child boxes quantised to 8 bits
//                   per plane on a per-node power-of-two grid (fast mode), one
//                   64 B half line (4 x dwordx4)
//   GTri     48 B   triangle positions v0,v1,v2 in BVH leaf order + original index
//   GVertex  48 B   shading attributes (normal, tangent, uv), read once per closest hit
//   GObject 128 B   object record (transform, type, material, mesh)
//   GMaterial       5 texture slots + dielectric/emissive parameters
// Child refs: >= 0 interior node index, < 0 leaf item ~ref.
#pragma once
#include <stdint.h>

#define NR_MAX_VIEWS 16
#define NR_MAT_TYPES 6   // index 0 unused, 1..5 = nr_material_type

struct alignas(16) GXform {
    float R[9];       // world = R * (p * scale) + pos; inverse rotation is R^T
    float pos[3];
    float scale[3];
    float inv_s[3];   // 1 / ((scale + FMIN) + FMIN), transform.h:33
};

struct alignas(64) GNode {
    float lmin[3], lmax[3];
    float rmin[3], rmax[3];
    int32_t left, right;
    int32_t pad0, pad1;
};
static_assert(sizeof(GNode) == 64, "GNode must be one 64 B line");

// Wide node i holds the up to four grandchildren of binary node i (a child that
// is a leaf stands for itself), so every binary interior ref is also a wide
// node index and the binary root is the wide root.  Unused slots have ref
// NR_WIDE_EMPTY and are never box-tested.
#define NR_WIDE_EMPTY ((int32_t)0x80000000)
struct alignas(128) GWNode {
    float lo[3][4];         // [axis][child]
    float hi[3][4];
    int32_t ref[4];         // binary refs (>= 0 node, < 0 leaf ~slot)
    int32_t pad[4];
};
static_assert(sizeof(GWNode) == 128, "GWNode is one 128 B line");

// Quantised 4-wide node (fast mode): child k's box on axis a is
//   [origin[a] + qlo_k * 2^e_a, origin[a] + qhi_k * 2^e_a]
// with the 8-bit grid indices of child k in byte k of qlo[a] / qhi[a] and the
// biased exponent e_a + 127 in byte a of `exps` (so 2^e_a is a float with that
// exponent field: exact, and q * 2^e_a is exact).  The indices are rounded
// outward and checked with the same float additions the traversal performs, so
// every decoded box contains the exact child box: traversals over it visit a
// superset of the nodes and find the same closest hit.
struct alignas(64) GQNode {
    float origin[3];
    uint32_t exps;
    uint32_t qlo[3];
    uint32_t qhi[3];
    int32_t ref[4];         // binary refs as in GWNode
    int32_t pad[2];
};
static_assert(sizeof(GQNode) == 64, "GQNode is one 64 B half line");

struct alignas(16) GTri {
    float v0[3], v1[3], v2[3];
    int32_t orig;
    int32_t pad0, pad1;
};
static_assert(sizeof(GTri) == 48, "GTri is 48 B");

struct alignas(16) GVertex {
    float normal[3], tangent[3], uv[2];
    float pad[4];
};

struct alignas(16) GMesh {
    const GNode* nodes;
    const GWNode* wnodes;   // same tree, 4 wide (null for a one-triangle mesh)
    const GQNode* qnodes;   // same tree, 4 wide, quantised boxes (null for a one-triangle mesh)
    const GTri* tris;       // BVH slot order
    const uint32_t* idx;    // 3 vertex indices per BVH slot
    const GVertex* verts;
    float root_min[3], root_max[3];
    int32_t root_ref;
    int32_t depth;
};

struct alignas(16) GObject {
    GXform xf;              // 80 B (72 + alignment)
    int32_t type;
    int32_t bound_type;
    int32_t material;
    int32_t object_id;
    int32_t mesh;
    float radius;           // radius + FMIN (shapes.h:9)
    float neg_inv_density;  // -1 / (density + FMIN) (volumes.h:13-14)
    int32_t pad[3];         // pad[0] = material type of `material`
};
static_assert(sizeof(GObject) == 128, "GObject is 128 B");

struct GTex {
    int32_t image;          // -1 constant
    float color[3];
    float scale;
};

struct alignas(16) GMaterial {
    int32_t type;
    GTex tex[5];
    float tint[3];
    float ior;
    float brightness;
    int32_t pad[2];
};

struct GImage {
    const uint8_t* data;
    int32_t C, H, W;
    int32_t pad;
};

struct GSky {
    int32_t type;
    float start[3], end[3];
    float rot[9];           // rotation_y(deg2rad(rotation)), computed on the host
    float intensity;
    int32_t H, W;
    const float* data;      // CHW
    // importance sampling of the HDRI (fast mode): piecewise-constant density over
    // texels (weights of sky_cdfs); marginal CDF over rows (H + 1 floats, searched
    // in LDS) and a per-row alias table; null when not built
    const float* marg;
    // per texel, float4: Walker/Vose alias table of its row (probability of
    // keeping the texel, alias texel as int bits) and the texel's probability
    // within its row; H x W entries
    const float* alias;
    // the map as one float4 (r, g, b, 0) per texel: the bilinear lookup's four
    // texels in four 16-byte loads instead of twelve scattered planar ones
    const float* rgbx;
};

// everything a traversal/shading kernel needs, passed by value
struct GScene {
    const GNode* nodes;
    const GObject* objects;     // top-level BVH item order
    const GMesh* meshes;
    const GMaterial* materials;
    const GImage* images;
    GSky sky;
    float root_min[3], root_max[3];
    int32_t root_ref;
    int32_t n_objects;
    int32_t first_light;        // index into objects, -1 none
    int32_t stack_depth;        // LDS stack entries per lane
    GXform cube_faces[6];
};

struct GCamera {
    float R[9];
    float pos[3];
    float uvw[3];
    float res[2];               // (H, W) as the reference stores it
    float focal_length, focus_distance;
    float shutter_time, defocus_angle, defocus_radius;
    uint32_t H, W;
    const float* pattern;       // 2 floats per sample
    uint32_t n_pattern;
};

// closest-hit record written by the traversal stage for the shading stage (48 B:
// what shading reads; the front-face flag in bit 31 of `matf`, the material index
// in the rest)
struct alignas(16) GHit {
    float p[3];
    float n[3];
    float tg[3];
    float uv[2];
    uint32_t matf;
};
static_assert(sizeof(GHit) == 48, "GHit is three 16-byte words");
