/*
 * nr_oracle.h — C API of the CPU oracle (oracle/nr_oracle.cpp).
 *
 * TEST INFRASTRUCTURE.  The oracle is a plain C++17 restatement of the
 * reference's path-tracing hot path (IEEE fp32, no fp contraction, OpenMP over
 * pixel tiles).  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg load it, and only as the checker or the timed CPU baseline.
 * The product (nectarrender_amd/) never links or calls it.
 */
#ifndef NR_ORACLE_H
#define NR_ORACLE_H

#include "../include/nectar_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nro_scene nro_scene;
typedef struct nro_camera nro_camera;

/* scene */
nro_scene* nro_scene_build(const nr_scene_desc* desc);
void nro_scene_free(nro_scene* s);
int nro_scene_first_light(const nro_scene* s);   /* object index of lights[0], -1 if none */
int nro_scene_mesh_depth(const nro_scene* s, int mesh);  /* max depth of a mesh BVH */

/* closest hit of n rays (origin[3], direction[3]) — SceneGraph::hit */
int nro_trace_closest(const nro_scene* s, const float* rays, int n, nr_hit* hits, int threads);
/* traversal work of the same query: counts[0] inner nodes, [1] triangle tests,
 * [2] top-level leaf objects, summed over the rays */
int nro_trace_counts(const nro_scene* s, const float* rays, int n, int64_t* counts);

/* camera */
nro_camera* nro_camera_build(const nr_camera_desc* cam, uint32_t seed);
void nro_camera_free(nro_camera* c);
int nro_camera_ray(const nro_camera* c, uint32_t x, uint32_t y, uint32_t sample_idx,
                   uint32_t seed, float* out7 /* origin[3] dir[3] time */);
int nro_nrooks(uint32_t n, uint32_t seed, float* out2n);

/* One sample pass: writes the per-pass sample values (what trace_full_kernel
 * adds into the cleared sample AOVs) for the listed pixels (NULL = all) into
 * full-size planar CHW buffers (3 x H x W each, depth 1 x H x W: the distance
 * from the camera to the first hit, 0 on a miss; NULL buffers are skipped).
 * counts (optional, 5 values): segments, inner nodes, triangle tests,
 * top-level objects, paths. */
int nro_render_pass(const nro_scene* s, const nro_camera* c, const nr_render_desc* d,
                    uint32_t sample_idx, const int32_t* pixels, int n_pixels,
                    float* beauty, float* diffuse, float* normal, float* depth,
                    int64_t* counts, int threads);

/* Full RenderEngine::render loop: passes d->sample_begin .. d->sample_end-1,
 * each folded into the running means (accumulate_samples_kernel). */
int nro_render(const nro_scene* s, const nro_camera* c, const nr_render_desc* d,
               float* beauty, float* diffuse, float* normal, float* depth, int threads);

/* accumulate_samples_kernel on n floats */
void nro_accumulate(float* avg, const float* sample, size_t n, uint32_t current_sample);

/* post: process_stream_kernel on a 3 x H x W image */
int nro_post(const float* chw, int H, int W, const nr_post_desc* d, float* out_f32, uint8_t* out_u8);

/* viewport preview (trace_viewport_kernel): overwrites 3 x H x W beauty,
 * diffuse and normal; screen-space ray, picking, selection mask + outline,
 * overlay composite (composite_overlay_kernel) */
int nro_viewport(const nro_scene* s, const nro_camera* c, int show_axis_grid, float* beauty, float* diffuse,
                 float* normal, int threads);
int nro_triplanar_grid(const float* ray6, float* out4 /* grid colour rgb, alpha */);
int nro_center_ray(const nro_camera* c, uint32_t x, uint32_t y, float* out6);
int nro_screen_ray(const nro_camera* c, float su, float sv, float* out6);
int nro_pick(const nro_scene* s, const nro_camera* c, float su, float sv, nr_hit* out);
int nro_selection_mask(const nro_scene* s, const nro_camera* c, int32_t object_id, int radius, uint8_t* out);
int nro_overlay(uint8_t* chw_u8, const uint8_t* mask, const float* color3, int H, int W);

/* Single-vertex probes.  rec: a hit record as nr_trace_closest returns it.
 * material_evaluate: Material::evaluate(rec, view, light) (material.h:59-67, 152-180, 351-357).
 * material_sample: Material::scatter, then for non-specular materials
 * sample_cosine_brdf (brdf.cu:4-35, the 50/50 mixture with lights[0] when the scene
 * has an ObjectLight) with Philox stream (seed, subseq) from draw `offset`; kind =
 * 0 absorbed, 1 specular (dir = the scattered ray, pdf = -1), 2 sampled (dir, pdf);
 * returns the next draw index.  light_pdf / light_random: pdf_value / random of the
 * object registered as object_id (primitives.h:34-71, shapes.h:60-108).
 * texture_sample: TextureView::sample (texture.h:49-91) of a material's texture slot. */
int nro_material_evaluate(const nro_scene* s, int material, const nr_hit* rec, const float* view3, const float* light3,
                          float* out3);
int nro_material_sample(const nro_scene* s, int material, const nr_hit* rec, const float* ray6, uint32_t seed,
                        uint32_t subseq, uint32_t offset, float* dir3, float* pdf, float* atten3, int* kind);
float nro_light_pdf(const nro_scene* s, int object_id, const float* origin3, const float* dir3);
int nro_light_random(const nro_scene* s, int object_id, const float* origin3, uint32_t seed, uint32_t subseq,
                     uint32_t offset, float* out3);
int nro_texture_sample(const nro_scene* s, int material, int slot, float u, float v, float* out3);
/* the 64 mixing constants of hash.h:17-34 as the oracle holds them */
void nro_hash_constants(uint32_t* out64);

/* RNG and hashing */
void nro_philox_block(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float nro_philox_uniform(uint32_t seed, uint32_t subsequence, uint32_t offset);
uint32_t nro_pcg_hash_int(uint32_t i, uint32_t seed, uint32_t stream);
float nro_pcg_float(float f, uint32_t seed, uint32_t stream);
float nro_pcg_float3(const float* a3, uint32_t seed, uint32_t stream);
float nro_pcg_float_ray(const float* o3, const float* d3);
void nro_pcg_vector2(float f, uint32_t seed, float* out2);

/* BVH build over n boxes (min xyz, max xyz): nodes_out = (left, right, object)
 * per node, node_boxes = (min xyz, max xyz) per node, order_out = item order.
 * Returns the node count (2n-1), or -1. */
int nro_bvh_build(const float* boxes, int n, int32_t* nodes_out, float* node_boxes, int32_t* order_out);

/* sky lookup of a world direction (Skylight::sample) */
int nro_sky_sample(const nro_scene* s, const float* dir3, float* out3);

/* PDFs (pdf.h) for closed-form tests */
float nro_ggx_distribution(float n_dot_h, float alpha);
float nro_pdf_pbr_value(const float* n3, const float* view3, float roughness, float spec_prob, const float* dir3);
int nro_pdf_pbr_generate(const float* n3, const float* view3, float roughness, float spec_prob,
                         uint32_t seed, uint32_t subseq, uint32_t offset, float* out3);
float nro_pdf_cosine_value(const float* n3, const float* dir3);
int nro_pdf_cosine_generate(const float* n3, uint32_t seed, uint32_t subseq, uint32_t offset, float* out3);

/* deterministic transcendentals (nr_mathfn.h): 0 sin 1 cos 2 acos 3 atan2(x,y) 4 log 5 pow(x,y) */
void nro_mathfn(int which, const float* x, const float* y, float* out, int n);

#ifdef __cplusplus
}
#endif

#endif
