/*
 * nectar_hip.h — C ABI of the MI355X path-tracing library (libnectar_hip.so).
 *
 * The library replaces the device half of NectarRender's render pass.  The
 * pybind11 module `_pathtracer` (nectarrender_amd/csrc/bindings.cpp) lowers the
 * reference's Scene/Hittable/Material/Camera objects into the plain structs
 * below and calls these entry points; tests call them through ctypes.
 *
 * Conventions
 *   - every entry point returns NR_OK (0) or an NR_ERR_* code; the message of
 *     the last failure on the calling thread is available from nr_last_error().
 *     Nothing aborts the process (the reference aborts on CUDA errors,
 *     csrc/core/include/core/cuda.h:25-35).
 *   - all pointers in descriptors are host pointers; the library copies what it
 *     needs during the call, so descriptors may be freed on return.
 *   - images and AOV layers are planar CHW; AOV pixel index = y*W + x, channel
 *     plane offset = c*H*W (csrc/data/include/data/utils.h:18-20).
 *   - H = camera resolution[0], W = camera resolution[1]
 *     (csrc/engine/include/engine/camera.h:176-177, 357-358).
 */
#ifndef NECTAR_HIP_H
#define NECTAR_HIP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NR_ABI_VERSION 7

enum nr_status {
    NR_OK = 0,
    NR_ERR_INVALID = 1,   /* bad argument / descriptor */
    NR_ERR_HIP = 2,       /* HIP runtime failure */
    NR_ERR_STATE = 3,     /* call order (e.g. render before scene upload) */
    NR_ERR_LIMIT = 4      /* a size limit of this library was exceeded */
};

/* HittableType, csrc/hittable/include/hittable/hittable.h:13-15 */
enum nr_hittable_type {
    NR_HIT_NULL = 0,
    NR_HIT_QUAD = 1,
    NR_HIT_SPHERE = 2,
    NR_HIT_CUBE = 3,
    NR_HIT_MESH = 4,
    NR_HIT_CONSTANT_MEDIUM = 5,
    NR_HIT_OBJECT_LIGHT = 6
};

/* MaterialType, csrc/material/include/material/material.h:374-376 */
enum nr_material_type {
    NR_MAT_NULL = 0,
    NR_MAT_LAMBERTIAN = 1,
    NR_MAT_PBR = 2,
    NR_MAT_DIELECTRIC = 3,
    NR_MAT_EMISSIVE = 4,
    NR_MAT_ISOTROPIC = 5
};

/* SkylightType, csrc/light/include/skylight.h:131 */
enum nr_sky_type { NR_SKY_NULL = 0, NR_SKY_SIMPLE = 1, NR_SKY_HDRI = 2 };

/* LayerType, csrc/data/include/data/render_layers.h:11-14.  A view holds BEAUTY,
 * DIFFUSE and WORLD_NORMAL (3 x H x W, the reference's default layers) and
 * DEPTH (1 x H x W): the running mean of the distance from the camera to the
 * first hit of each sample (0 where the camera ray misses).  The reference
 * allocates DEPTH but no kernel writes it; it is written here for the
 * denoiser's input. */
enum nr_layer {
    NR_LAYER_BEAUTY = 0,
    NR_LAYER_DIFFUSE = 1,
    NR_LAYER_SPECULAR = 2,
    NR_LAYER_OBJECT_NORMAL = 3,
    NR_LAYER_WORLD_NORMAL = 4,
    NR_LAYER_SHADOW = 5,
    NR_LAYER_DEPTH = 6,
    NR_LAYER_EMISSION = 7,
    NR_LAYER_OBJECT_ID = 8
};

/* Estimator. PARITY reproduces trace_full_kernel draw for draw
 * (csrc/engine/src/trace.cu:47-79).  FAST (no counterpart in the reference)
 * adds next-event estimation to an HDRI sky: at every non-specular vertex a
 * direction drawn from the map's radiance x sin(theta) (marginal CDF over rows,
 * per-row alias tables), weighted against
 * the scattering technique (the reference's mixture with lights[0] included)
 * by the power heuristic, with a shadow ray; sample s of pixel i draws from
 * the 64-bit Philox subsequence (s << 32) | i, and passes start at 0 (pass s
 * is folded as the (s + 1)-th sample, so samples_per_pixel = n renders all n
 * pattern entries).  ObjectLights keep the reference's one-sample mixture.
 * Same expectation as PARITY; checked by convergence (white furnace, z-test
 * against PARITY), not per pixel. */
enum nr_mode { NR_MODE_PARITY = 0, NR_MODE_FAST = 1 };

/* Where nr_scene_upload builds mesh trees (BVH<T>::build, csrc/hittable/include/bvh/bvh.h:36-55,
 * for Mesh::build, csrc/hittable/include/hittable/mesh.h:177-216).  Both builders make the same
 * binned-SAH decisions and produce the same node array; DEVICE runs them on the GPU.  AUTO
 * (the default) uses the device for meshes of at least 4096 triangles. */
enum nr_build_mode { NR_BUILD_HOST = 0, NR_BUILD_DEVICE = 1, NR_BUILD_AUTO = 2 };

/* One hittable.  Transform(position, rotation_deg, scale) semantics of
 * csrc/core/include/core/transform.h:25-34; spheres use rotation 0, scale 1
 * (Transform(position), transform.h:17-23). */
typedef struct nr_object_desc {
    int32_t type;          /* nr_hittable_type */
    int32_t bound_type;    /* CONSTANT_MEDIUM / OBJECT_LIGHT: type of the bound object */
    int32_t material;      /* index into nr_scene_desc.materials */
    int32_t mesh;          /* index into nr_scene_desc.meshes when (bound_)type is MESH, else -1 */
    float position[3];
    float rotation[3];     /* Euler degrees, R = Rz * Ry * Rx (matrix.h:94-96) */
    float scale[3];
    float radius;          /* sphere radius as given to Hittable.sphere (FMIN is added inside) */
    float density;         /* constant medium density */
} nr_object_desc;

/* TextureView (csrc/material/include/material/texture.h:49-91). */
typedef struct nr_texture_desc {
    int32_t image;         /* index into nr_scene_desc.images; -1 = constant colour */
    float color[3];
    float scale;           /* uv scale of image textures */
} nr_texture_desc;

typedef struct nr_image_desc {
    int32_t C, H, W;       /* C >= 3 */
    const uint8_t* data;   /* planar CHW uint8 */
} nr_image_desc;

/* One material.  tex[] slots: LAMBERTIAN {albedo}; PBR {albedo, roughness,
 * metallic, emission, normal}; EMISSIVE {texture}; ISOTROPIC {texture}. */
typedef struct nr_material_desc {
    int32_t type;          /* nr_material_type */
    nr_texture_desc tex[5];
    float tint[3];         /* dielectric */
    float ior;             /* dielectric */
    float brightness;      /* emissive */
} nr_material_desc;

/* MeshVertex (csrc/hittable/include/hittable/mesh.h:10-16). */
typedef struct nr_mesh_desc {
    int32_t n_vertices;
    int32_t n_triangles;
    const float* vertices;     /* n_vertices x 11: position[3] normal[3] tangent[3] uv[2] */
    const uint32_t* triangles; /* n_triangles x 3 vertex indices */
} nr_mesh_desc;

typedef struct nr_sky_desc {
    int32_t type;          /* nr_sky_type */
    float start[3], end[3];/* SIMPLE */
    float rotation;        /* HDRI, degrees about +y */
    float intensity;       /* HDRI */
    int32_t C, H, W;       /* HDRI, C >= 3 */
    const float* data;     /* HDRI planar CHW fp32 */
} nr_sky_desc;

/* Objects are listed in registration order: hittables then lights
 * (csrc/scene/include/scene.h:31-39); object_id = position in this list.
 * materials[0] is the registry's default material (registry.h:23-26). */
typedef struct nr_scene_desc {
    int32_t n_objects;
    const nr_object_desc* objects;
    int32_t n_materials;
    const nr_material_desc* materials;
    int32_t n_images;
    const nr_image_desc* images;
    int32_t n_meshes;
    const nr_mesh_desc* meshes;
    nr_sky_desc sky;
} nr_scene_desc;

/* CameraParams (csrc/engine/include/engine/camera.h:68-136). */
typedef struct nr_camera_desc {
    float resolution[2];   /* (H, W) as the reference stores it */
    float position[3];
    float rotation[3];     /* Euler degrees */
    uint32_t samples_per_pixel;
    float focal_length;
    float focus_distance;
    float aperture;
    float sensor_width;
    float shutter_speed;
} nr_camera_desc;

/* One call renders sample passes [sample_begin, sample_end) of a view and
 * folds each into the view's running-mean AOVs exactly like
 * RenderEngine::render (csrc/engine/src/engine.cpp:42-74).  TraceConfig
 * fields: csrc/engine/include/engine/trace.h:10-19. */
typedef struct nr_render_desc {
    uint32_t sample_begin;
    uint32_t sample_end;
    uint32_t ray_depth;
    uint32_t seed;
    int32_t do_rr;
    uint32_t rr_start_bounce;
    float rr_survival;
    int32_t mode;          /* nr_mode */
} nr_render_desc;

/* Closest-hit record (HitRecord, csrc/hittable/include/hittable/hit_record.h).
 * prim = original triangle index for meshes, -1 otherwise. 64 bytes. */
typedef struct nr_hit {
    int32_t object_id;     /* -1 = miss */
    int32_t material_id;
    int32_t prim;
    int32_t front_face;
    float t;
    float p[3];
    float n[3];
    float tangent[3];
    float uv[2];
} nr_hit;

/* StreamConfig (csrc/data/include/data/stream.h:13-30), minus the OIDN knobs. */
typedef struct nr_post_desc {
    int32_t linear_to_gamma;
    int32_t apply_white_balance;
    float wb_temperature;
    float wb_tint;
    int32_t apply_tonemapping;
    int32_t tm_method;     /* 0 REINHARD, 1 REINHARD_EXTENDED, 2 ACES */
    float tm_white_point;
    float tm_alpha;
    int32_t fix_reinhard;  /* 0 keeps the reference's Reinhard no-op (stream.cu:7-11) */
    int32_t layer;         /* nr_layer streamed (TransferStream::link, stream.h:84-93); 3-channel layers */
} nr_post_desc;

typedef struct nr_ctx nr_ctx;

/* Version of this ABI (NR_ABI_VERSION). */
int nr_abi_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* nr_last_error(void);

/* Create a context on HIP device `device`.  Replaces the allocations of
 * RenderEngine::RenderEngine (csrc/engine/src/engine.cpp:7-34). NULL on error. */
nr_ctx* nr_create(int device);
void nr_destroy(nr_ctx* ctx);

/* Lower and upload a scene: top-level and per-mesh binned-SAH BVHs, object,
 * material and texture tables, sky.  Replaces Scene::build
 * (csrc/scene/src/scene.cpp:68-76) and the registries it drives
 * (csrc/hittable/include/hittable/registry.h:149-263,
 *  csrc/material/include/material/registry.h:85-95). */
int nr_scene_upload(nr_ctx* ctx, const nr_scene_desc* scene);

/* Set the camera of a view (0 <= view < 16) and build its sample pattern.
 * Replaces Camera::__construct (csrc/engine/include/engine/camera.h:296-305).
 * (Re)allocates and zeroes the view's AOV layers when the resolution changes. */
int nr_camera_set(nr_ctx* ctx, int view, const nr_camera_desc* cam, uint32_t seed);

/* Zero the running-mean AOVs of a view (RenderEngine::reset, engine.cpp:216-220). */
int nr_aov_reset(nr_ctx* ctx, int view);

/* Render sample passes of a view.  Replaces the loop body of
 * RenderEngine::render (engine.cpp:42-74): trace_full (trace.cu:81-94) +
 * RenderLayers::accumulate (render_layers.h:131-140).  Asynchronous on the
 * context's stream; nr_sync waits. */
int nr_render(nr_ctx* ctx, int view, const nr_render_desc* desc);

/* The same sample passes for several views (1 to NR_MAX_VIEWS, distinct) in one
 * streaming pipeline: every launch serves the paths of all of them, so small
 * and uneven views share the GPU instead of running one after another.  Each
 * view's AOVs equal what nr_render gives it with the same desc.  The batch form
 * of the reference's one-engine-per-camera loop (engine.cpp:42-74); the
 * reference has no multi-view call.  Waits for the context's previous work
 * before it starts. */
int nr_render_views(nr_ctx* ctx, const int32_t* views, int n_views, const nr_render_desc* desc);

/* nr_render restricted per tile: pixel (x, y) renders only the passes of
 * [sample_begin, sample_end) below tile_end[(y / tile_size) * cols + x / tile_size]
 * (host array of exactly rows x cols values, cols = ceil(W / tile_size), rows =
 * ceil(H / tile_size), which the call reads in full; tile_size a power of two in
 * [8, 4096]).  A tile whose limit is at most sample_begin keeps
 * its AOVs; rendering from the first pass (0 in NR_MODE_FAST, 1 otherwise)
 * restarts a tile's running mean.  The per-tile fallback of a denoised frame to
 * more samples; no counterpart in the reference. */
int nr_render_tiles(nr_ctx* ctx, int view, const nr_render_desc* desc, const uint32_t* tile_end, int tile_size);

/* Image size of a view (H = resolution[0], W = resolution[1]). */
int nr_view_size(nr_ctx* ctx, int view, int32_t* H, int32_t* W);

/* Copy one AOV layer (C x H x W fp32; C = 1 for DEPTH, 3 otherwise) of a view to host memory. */
int nr_aov_read(nr_ctx* ctx, int view, int layer, float* host_chw, size_t n_floats);

/* Device pointer of a view's AOV layer (C x H x W fp32), NULL if absent. */
float* nr_aov_device_ptr(nr_ctx* ctx, int view, int layer);

/* Closest hit of n world-space rays (origin[3], direction[3] per ray) against
 * the uploaded scene: SceneGraph::hit (csrc/scene/include/graph.h:16-66). */
int nr_trace_closest(nr_ctx* ctx, const float* rays, int32_t n, nr_hit* hits);

/* Same with device-resident rays/hits, enqueued on the context stream. */
int nr_trace_closest_device(nr_ctx* ctx, const float* d_rays, int32_t n, nr_hit* d_hits);

/* nr_trace_closest_device without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The traversal-stack spill and
 * the counters live in a file-scope device array of 16 MiB per GPU (held until the
 * process ends; its address is looked up once per device by nr_create), and rays
 * are traced in chunks of as many rays as it holds spill columns for, so any n is
 * accepted; a scene whose stack would not fit one 128-ray workgroup is refused
 * with NR_ERR_LIMIT.  The array keeps the last call's contents; the call clears
 * what must start at zero with a kernel (no memset node).  Calls that overlap on
 * one GPU (other streams, other contexts on the same device) race on the array.
 * Same hits as nr_trace_closest_device; asynchronous on the context stream. */
int nr_trace_closest_device_noalloc(nr_ctx* ctx, const float* d_rays, int32_t n, nr_hit* d_hits);

/* White balance -> tonemap -> gamma -> uint8 CHW of a view's AOV layer
 * desc->layer (BEAUTY unless relinked) (process_stream_kernel,
 * csrc/data/src/stream.cu:146-159).  If out_f32 is
 * non-NULL the display-referred float image (before x255) is written too. */
int nr_post(nr_ctx* ctx, int view, const nr_post_desc* desc, uint8_t* out_u8, float* out_f32);

/* Per-tile error estimate of a denoised frame (all device pointers, planar
 * 3 x H x W display-referred fp32, e.g. from nr_post_device): for each
 * tile_size^2 tile, the mean over its pixels and channels of
 * (D - J1)^2 - (J1 - J2)^2 / 2, where J1 and J2 are two plain renders of the
 * same view independent of D and of each other.  Its expectation is D's mean
 * squared error plus the judges' own display-space bias (tonemapping a
 * few-sample render is biased low), so it errs towards more fallback; it can
 * come out negative.  d_out receives rows x cols floats, row major.
 * Synchronous.  No counterpart in the reference: the plain path tracer judging
 * the denoiser. */
int nr_tile_error(nr_ctx* ctx, const float* d_denoised, const float* d_judge1, const float* d_judge2, int32_t H,
                  int32_t W, int32_t tile_size, float* d_out);

/* The same post-processing of any planar 3 x H x W fp32 device image (a
 * denoised frame, an accumulated reference); outputs are device pointers,
 * either may be NULL.  Synchronous. */
int nr_post_device(nr_ctx* ctx, const float* d_chw, int32_t H, int32_t W, const nr_post_desc* desc,
                   uint8_t* d_u8, float* d_f32);

/* ---- interactive mode: viewport preview, picking, selection, readback ---- */

/* RenderEngine::viewport (csrc/engine/src/engine.cpp:76-96) -> trace_viewport
 * (csrc/engine/src/trace.cu:99-161): one centre ray per pixel (get_ray_center,
 * camera.h:181-184); a hit is shaded with the material's viewport colour under a
 * headlight and a directional light, a miss shows the sky and, if show_axis_grid,
 * the triplanar axis grid (csrc/engine/src/axis_grid.cu).  Overwrites the view's
 * BEAUTY, DIFFUSE and WORLD_NORMAL layers (aovs.overwrite, engine.cpp:89-91). */
int nr_viewport(nr_ctx* ctx, int view, int show_axis_grid);

/* hit_test_ray (trace.cu:166-188) as SceneInterface::query_scene calls it
 * (csrc/scene/include/interface.h:59-70): closest hit of the view camera's
 * screen-space ray at (su, sv) in [0,1]^2 (camera.h:215-226).  Host output;
 * object_id = -1 on a miss. */
int nr_pick(nr_ctx* ctx, int view, float su, float sv, nr_hit* out);

/* selection_mask (csrc/scene/src/scene_interface.cu:3-83): d_mask (H x W device
 * bytes) = 255 on the pixels within outline_radius of, but not on, object
 * object_id as seen through the view's screen-space rays; 0 elsewhere. */
int nr_selection_mask(nr_ctx* ctx, int view, int32_t object_id, int32_t outline_radius, uint8_t* d_mask);

/* An H x W byte device buffer owned by the view (allocated on first use, freed by
 * nr_destroy; reallocated when the view grows), for callers without a device allocator of
 * their own: SceneInterface keeps its selection mask here (interface.h:193-199). */
uint8_t* nr_view_mask_buffer(nr_ctx* ctx, int view);

/* composite_overlay (csrc/data/src/stream.cu:177-227) on a planar 3 x H x W
 * device uint8 image: blend colour (clamped to [0,1], x255) with the mask as alpha. */
int nr_overlay_device(nr_ctx* ctx, uint8_t* d_img, int32_t H, int32_t W, const uint8_t* d_mask,
                      const float color[3]);

/* TransferStream::readback (csrc/data/include/data/stream.h:154-182). */
typedef struct nr_readback_desc {
    nr_post_desc post;           /* post->layer is the linked layer */
    void* denoiser;              /* nr_unet* (nectar_denoise.h): denoise the layer with the
                                    view's DIFFUSE and WORLD_NORMAL first; NULL = off */
    const uint8_t* d_overlay;    /* H x W device mask (nr_selection_mask); NULL = none */
    float overlay_color[3];
} nr_readback_desc;

/* Linked layer -> [denoise] -> post -> [overlay] -> host_u8 (3 x H x W), on the
 * view's own readback stream through pinned staging.  It may run on another
 * thread while nr_render / nr_viewport run on the same context: it reads the
 * layers as the last completed call left them, and a call that rewrites the
 * layers waits for a readback in flight.  Not concurrently with nr_camera_set,
 * nr_scene_upload or nr_destroy. */
int nr_readback(nr_ctx* ctx, int view, const nr_readback_desc* desc, uint8_t* host_u8);

/* ---- scene trees and live edits ---- */

/* Number of stream groups nr_render splits a view into (1..4, default 2): bands of
 * whole 8-pixel tile rows, each a complete wavefront pipeline on its own HIP stream,
 * so one group's traversal tail overlaps the others' work.  The result does not
 * depend on it (every pixel's samples are folded in order).  Timing mode
 * (nr_set_timing) always renders with one group. */
int nr_set_render_streams(nr_ctx* ctx, int n);

/* Launch tunings of nr_render (none changes the result: every pixel's samples
 * are folded in order with the same random numbers).  No counterpart in the
 * reference (one CUDA thread per pixel, launch shape fixed). */
typedef struct nr_tuning {
    int32_t slots;            /* path slots per pixel, 1..64 (default 4): samples of a pixel in flight;
                                 the path state and queues hold slots x pixels entries */
    int32_t stream_groups;    /* 1..4 (default 2), as nr_set_render_streams */
    int32_t shade_wg_per_cu;  /* k_shade grid: 256-thread workgroups per CU, 1..32 (default 8) */
    int32_t tail_paths;       /* tail mode: once a stream group's live paths are at most this many
                                 (default 262144; 0 = never), one kernel finishes each remaining
                                 pixel with one thread, instead of wavefront iterations that each
                                 last as long as their slowest ray */
    int32_t graph_iterations; /* 0 (default): the host launches every kernel of every streaming
                                 iteration, sizing the traversal grids from the live-path count;
                                 2..16, even: each stream group captures that many consecutive
                                 iterations once per render call as a HIP graph (grids of the
                                 group's full path capacity, surplus workgroups exit at once) and
                                 replays it, so the host issues one launch per that many
                                 iterations.  Timing mode (nr_set_timing) always uses the host loop. */
} nr_tuning;
int nr_set_tuning(nr_ctx* ctx, const nr_tuning* tuning);
int nr_get_tuning(nr_ctx* ctx, nr_tuning* tuning);

/* The project's autotuner: renders `desc` on `view` under candidate tunings
 * (coordinate descent over slots {1,2,4,8,16,32,48,64}, skipping slot counts whose
 * path state would take more than 60% of the device's memory, stream groups {1..4}, shade
 * workgroups per CU {4,8,16}, tail paths {0, 64k, 256k, 1M, 2M, 4M}, graph
 * iterations {0, 2, 8}), each timed as the faster of two renders after a
 * one-pass warmup, keeps the fastest (a candidate replaces the current one only
 * when it is at least 1% faster, so timing noise does not swap settings that
 * render equally fast), writes it to *tuning (may be NULL) and clears the view's
 * AOVs.  Give it the passes of the real renders: path slots above the desc's
 * pass count cannot show their effect. */
int nr_autotune(nr_ctx* ctx, int view, const nr_render_desc* desc, nr_tuning* tuning);

/* Mesh tree builder for later nr_scene_upload calls (enum nr_build_mode). */
int nr_set_build_mode(nr_ctx* ctx, int mode);

/* The uploaded tree of mesh `mesh`: n_nodes child-pair nodes of 64 bytes (both child boxes
 * as min[3] max[3] left then right, then the left and right refs: >= 0 node index, < 0 leaf
 * ~slot; two zero pads) and the original triangle index of every leaf slot.  nodes64 /
 * slot_orig may be NULL (sizes only); depth = nodes on the longest root-to-leaf path. */
int nr_mesh_tree(nr_ctx* ctx, int mesh, void* nodes64, int32_t* slot_orig, int32_t* n_nodes, int32_t* n_tris,
                 int32_t* depth);

/* SceneInterface::set_transform (csrc/scene/include/interface.h:127-133) as a partial rebuild
 * (the registry's mark_dirty + build_bvh, csrc/hittable/include/hittable/registry.h:216-234):
 * object `index` (registration order, = nr_hit.object_id) gets Transform(position,
 * rotation in degrees, scale); only the object records and the top-level tree are rebuilt and
 * re-uploaded; mesh trees, materials, textures and the sky stay on the device. */
int nr_object_set_transform(nr_ctx* ctx, int32_t index, const float position[3], const float rotation[3],
                            const float scale[3]);

/* SceneInterface::add_object (csrc/scene/include/interface.h:112-116) -> Scene::add_hittable
 * (csrc/scene/src/scene.cpp:101-104) and the rebuild it requests, as a partial update: the
 * object is registered last (object_id = the previous object count, registry.h:113-126).
 * `add` (may be NULL) lists the images, materials and meshes the object brings, appended to the
 * uploaded tables (its n_objects must be 0, its sky is ignored): indices continue after the
 * uploaded entries, and obj->material, obj->mesh and the new materials' texture images index the
 * combined tables.  Only a new mesh gets a tree; the top-level tree and the object table are
 * rebuilt (the registry's build_bvh, registry.h:216-234); uploaded meshes, images, materials and
 * the sky stay on the device.  Hits and renders equal those of nr_scene_upload of the combined
 * scene.  Invalid input leaves the scene as it was; a device failure leaves no scene. */
int nr_scene_add_object(nr_ctx* ctx, const nr_object_desc* obj, const nr_scene_desc* add);

/* SceneInterface::swap_skylight / request_skylight_update (interface.h:137-151 ->
 * Skylight::build in Scene::update, scene.cpp:126-127): replaces the sky (and, for an HDRI,
 * its interleaved copy and fast mode's sampling tables) and nothing else. */
int nr_scene_set_sky(nr_ctx* ctx, const nr_sky_desc* sky);

/* Wait for all work of the context.  NR_ERR_LIMIT if the last nr_render* call's
 * traversal overflowed its stack (the flag covers that call only; nr_aov_read
 * reports it too). */
int nr_sync(nr_ctx* ctx);

/* Mesh-tree layout of later nr_trace_closest(_device) calls: 0 = the
 * reference's binary tree in its near-first order (default, what the
 * reference's estimator traverses), 1 = the 4-wide layout, 2 = the 4-wide
 * layout with child boxes quantised to 8 bits per plane (rounded outward),
 * which NR_MODE_FAST's extension and shadow rays traverse.  All three find the
 * same closest hit; of two triangles hit at exactly the same t either may be
 * reported, and a ray lying exactly in a tree box's plane with a zero direction
 * component may find a hit with the quantised layout that the reference's box
 * test (0 * inf = NaN) misses.  No counterpart in the reference: the test hook of the wide
 * traversals. */
int nr_set_trace_layout(nr_ctx* ctx, int layout);

/* Per-kernel GPU time (ms) of the last nr_render / nr_trace_closest call,
 * measured with HIP events on the context stream: [0] whole call,
 * [1] traversal kernels (k_extend), [2] shading kernels, [3] generate+accumulate,
 * [4] streaming iterations (= k_extend launches) of nr_render, [5] shadow-ray
 * kernels (NR_MODE_FAST with an HDRI sky; 0 otherwise), [6] tail mode (k_tail:
 * traversal, shading and shadow rays of the last paths), [7] rays k_extend
 * traced (path segments; the rest of the render's segments ran in k_tail).
 * Returns the number of values written. */
int nr_last_timings(nr_ctx* ctx, float* ms, int n);

/* Enable/disable event timing (off by default; it adds event records). */
int nr_set_timing(nr_ctx* ctx, int enabled);

#ifdef __cplusplus
}
#endif

#endif /* NECTAR_HIP_H */
