// nectar_hip.hip — MI355X wavefront path tracer behind the C ABI of
// include/nectar_hip.h.
//
// The sample passes of a view run as a streaming wavefront pipeline with one
// path slot per pixel:
//   k_start         first camera ray (camera.h:186-278) of every pixel
//   per iteration:
//     k_counts      clears the queue counters the iteration appends to
//     k_extend      closest hit (SceneGraph::hit, graph.h:16-66) of every live
//                   path; misses add the sky and end the path; hits are
//                   compacted into one queue per material type (wave ballot +
//                   one atomic per wave and type)
//     k_shade<M>    one kernel per material type present: emission, scatter,
//                   BRDF sampling / mixture light sampling (brdf.cu:4-35),
//                   bounce-0 AOVs, Russian roulette (trace.cu:63-78); survivors
//                   are compacted into the next live queue
//     k_shadow      (fast mode) shadow rays of next-event estimation to the
//                   HDRI: an any-hit traversal, the MIS-weighted radiance is
//                   added if the ray reaches the sky
//     k_regen       folds every ended path's sample into the running-mean AOVs
//                   (accumulate_samples_kernel, data.cu:28-40) and starts the
//                   pixel's next sample in the same iteration
// The host reads the live count once per iteration to size the traversal grid.
// Path state lives in HBM as SoA float4 arrays indexed by pixel; queues hold
// pixel indices.  The per-path Philox stream is (seed, pixel + s*W*H, draw id)
// exactly as trace.cu:58 keys it, so every path makes the reference's draws,
// and each pixel's samples are folded in order.  Fast mode (no counterpart in
// the reference) keys the stream by the 64-bit (s << 32) | pixel, renders
// passes from 0 and adds next-event estimation; it has the same expectation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/nectar_denoise.h"
#include "../../include/nectar_hip.h"
#include "nr_bvh.h"
#include "nr_bvh_gpu.h"
#include "nr_device.h"
#include "nr_error.h"
#include "nr_types.h"

// ---------------------------------------------------------------------------
// error plumbing
// ---------------------------------------------------------------------------
// message of the last failure on this thread, shared by every translation unit
// of the library (nr_error.h)
std::string& nr_error_slot() {
    static thread_local std::string slot;
    return slot;
}
#define g_err nr_error_slot()
static int fail(int code, const std::string& msg) { g_err = msg; return code; }
#define HIPCK(expr)                                                                      \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(NR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));  \
    } while (0)

// counters (one int each): live queues [0,2), material queues [2,8), error flag,
// shadow-ray queue (fast mode), occupied path slots and pixels with samples not
// yet started after k_regen, rays k_extend traced during the render (64 bits
// at [12, 14))
enum { CNT_ACT = 0, CNT_MAT = 2, CNT_ERR = 8, CNT_SHD = 9, CNT_OCC = 10, CNT_PEND = 11, CNT_TRACED = 12, CNT_N = 16 };
// Path slots per pixel (a tuning of the context, 1..NR_SLOTS_MAX, default 4):
// sample s of pixel i runs in slot s % slots (path id i + (s % slots) * cap),
// so up to `slots` consecutive samples of a pixel are in flight and pixels with
// long paths do not hold back the whole view (1080p workload, 64 samples: 266
// iterations with one slot, 146 with 2, 86 with 4).  With 64, every sample of a
// 64-sample render is live at once (about 70 GB of path state and traversal stack at
// 1080p), so each traversal launch holds more rays and its end-of-launch drain is paid
// less often: 1080p MIS 1284 (8 slots) -> 1368 (16) -> 1460 (32) -> 1611 Mpaths/s (64).
#ifndef NR_SLOTS_MAX
#define NR_SLOTS_MAX 64
#endif
#define NR_SLOTS_DEFAULT 4
// Live paths (upper bound) of a stream group below which the group switches
// to tail mode (k_tail)
#ifndef NR_TAIL_PATHS
#define NR_TAIL_PATHS 262144
#endif
// rows of the HDRI marginal CDF that k_shade stages in LDS (larger maps read it from global memory)
constexpr int SKY_MARG_LDS = 4097;
#ifndef NR_LDS_LEVELS
#define NR_LDS_LEVELS 8
#endif
constexpr int LDS_STACK_LEVELS = NR_LDS_LEVELS;   // traversal stack levels kept in LDS
#define NR_MAX_GROUPS 4
#ifndef NR_RENDER_STREAMS
#define NR_RENDER_STREAMS 2
#endif
// register budget of the traversal kernels: minimum waves per SIMD (0 = compiler's
// choice).  6 caps them at 80 VGPRs: the kernels for scenes with every object
// type would otherwise take 81 and run 5 waves per SIMD; the quad / sphere /
// mesh kernels take 71-74 either way.
#ifndef NR_TRAV_OCC
#define NR_TRAV_OCC 6
#endif
// waves per SIMD of k_extend for quad / sphere / mesh scenes: at 7 (72 VGPRs) the
// binary-tree traversal loop keeps its state in registers; built for 8 it spilled
// 120 bytes per lane inside the loop and ran slower on the 1080p workload.  The
// fast-mode kernel (quantised nodes, folded box test) at 6: 80 VGPRs and 12 bytes
// of spill outside the loop; at 7 it spilled 64 bytes and reloaded per-ray terms
// inside the loop (1080p MIS workload 1185 -> 1055 Mpaths/s), at 5 1163.  The
// every-object-type kernels keep the budget of NR_TRAV_OCC.
#ifndef NR_EXT_WAVES
#define NR_EXT_WAVES 7
#endif
#ifndef NR_EXT_FAST_WAVES
#define NR_EXT_FAST_WAVES 6
#endif
// waves per SIMD of k_shadow for quad / sphere / mesh scenes: 7 (72 VGPRs, 16
// bytes of spill outside the loop) against 6 (75 VGPRs): 1080p MIS workload
// 1183 -> 1187 Mpaths/s.
#ifndef NR_SHADOW_WAVES
#define NR_SHADOW_WAVES 7
#endif
#if NR_TRAV_OCC > 0
#define NR_TRAV_ATTR __attribute__((amdgpu_waves_per_eu(NR_TRAV_OCC, 8)))
#else
#define NR_TRAV_ATTR
#endif

// Mesh-tree layouts of the fast-mode traversals (NR_LAY_*): the reference's
// estimator always walks the binary tree in the reference's order.  Shadow rays
// over the quantised nodes: k_shadow 24.87 -> 24.64 ms per 64 passes of the 1080p
// workload against the full-precision 4-wide nodes.
#ifndef NR_EXT_FAST_LAY
#define NR_EXT_FAST_LAY NR_LAY_QUANT
#endif
#ifndef NR_SHADOW_LAY
#define NR_SHADOW_LAY NR_LAY_QUANT
#endif

struct StackCfg {
    int2* spill;            // levels >= LDS_STACK_LEVELS, one column per ray of the launch
    size_t stride;          // rays in the launch
};

// Path state, `slots` slots per pixel (path id = pixel + slot * cap).  o.w
// carries the path's bounce index and d.w the sample index it is tracing (as
// integer bits).  Queues hold path ids; their stride is qcap = slots * padded pixels.
// float4 slots per path of the pending shadow ray's record (origin, direction,
// contribution, one unused: 64-byte aligned; a 48-byte record measured no faster)
#define NR_SH_STRIDE 4
struct PathPtrs {
    float4* o;          // o[2 p] origin and d[2 p] = o[2 p + 1] direction of path p
    float4* d;
    float4* atten;      // one 64-byte record per path p: atten[4 p] throughput, L[4 p] = atten[4 p + 1]
                        // radiance, diff[4 p] = atten[4 p + 2] albedo, nrm[4 p] = atten[4 p + 3] normal
    float4* L;
    float4* diff;
    float4* nrm;
    GHit* hit;
    int32_t* qact;      // 2 * qcap
    int32_t* qmat;      // NR_MAT_TYPES * qcap
    int32_t* counts;    // CNT_N
    uint8_t* finished;  // per path slot: 0 running, 1 ended (fold when its turn comes), 2 idle
    uint32_t* nf;       // per pixel: next sample to fold
    uint32_t* ns;       // per pixel: next sample to start
    float4* acc;        // running means during a render: beauty [0, cap), diffuse [cap, 2 cap), normal [2 cap, 3 cap)
    // fast mode: one pending shadow ray per path (origin, direction, MIS-weighted
    // contribution added to L if the ray reaches the sky), queued in qshd
    float4* sho;        // one record per path: sho[K p] origin, shd[K p] = sho[K p + 1] direction,
    float4* shd;        // shc[K p] = sho[K p + 2] contribution, K = NR_SH_STRIDE
    float4* shc;
    int32_t* qshd;
    int32_t* err;       // traversal-overflow flag (group 0's CNT_ERR; every group sets it)
    uint32_t cap;       // pixels
    uint32_t qcap;      // queue entries per queue (slots * padded pixels)
    uint32_t slots;     // path slots per pixel, 1..NR_SLOTS_MAX
};

// A view of a render call: its camera, where its 8x8 tiles start in the call's
// tile enumeration and where its pixels start in the path arrays.  All views of
// one call share every launch of the streaming pipeline.
struct GView {
    GCamera cam;
    uint32_t tile0;     // first tile of the view (tiles left to right within bands of 8 rows)
    uint32_t pix0;      // first path-array pixel of the view (its pixel y*W + x is pix0 + y*W + x)
    uint32_t tiles_x;   // tiles per band
    uint32_t npix;      // W * H: the subsequence of sample s is pixel + s*W*H (mod 2^32)
};

struct TraceArgs {
    uint32_t seed;
    uint32_t in;        // live queue read by this iteration (the other one is filled)
    uint32_t depth;
    int32_t do_rr;
    uint32_t rr_start;
    float rr_survival;
    uint32_t s_end;     // passes [s_begin, s_end) are rendered
    uint32_t t0, t1;    // this stream group's range of the tiled pixel enumeration (tile_pixel)
    int32_t fast;       // NR_MODE_FAST: 64-bit sample subsequence (pixel, sample), NEE + MIS
    uint32_t n_views;
    const GView* views; // device table, n_views entries
    // optional per-tile sample limit (one view): pixel (x, y) renders passes below
    // min(s_end, tile_end[(y >> tile_shift) * tile_cols + (x >> tile_shift)])
    const uint32_t* tile_end;
    uint32_t tile_shift, tile_cols;
};

__device__ __forceinline__ uint32_t pixel_end(const TraceArgs& A, uint32_t x, uint32_t y) {
    if (!A.tile_end) return A.s_end;
    return min(A.s_end, A.tile_end[(y >> A.tile_shift) * A.tile_cols + (x >> A.tile_shift)]);
}

// ---------------------------------------------------------------------------
// wave-level queue append: one atomic per wave, lanes get consecutive slots
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wave_append(bool pred, int32_t* counter) {
    unsigned long long mask = __ballot(pred);
    if (mask == 0ull) return -1;
    int lane = __lane_id();
    int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(mask));
    base = __shfl(base, leader);
    unsigned long long below = lane == 0 ? 0ull : (mask & ((~0ull) >> (64 - lane)));
    return pred ? base + __popcll(below) : -1;
}

// workgroup-level queue append: one atomic per workgroup (slots are dense, in
// wave order within the workgroup).  Every thread of the block must call it.
template <int BLOCK>
__device__ __forceinline__ int block_append(bool pred, int32_t* counter) {
    constexpr int NW = BLOCK / 64;
    __shared__ int wcount[NW];
    __shared__ int wbase[NW];
    const int wave = threadIdx.x >> 6, lane = __lane_id();
    unsigned long long mask = __ballot(pred);
    if (lane == 0) wcount[wave] = __popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        int total = 0;
        for (int w = 0; w < NW; w++) { wbase[w] = total; total += wcount[w]; }
        int base = total ? atomicAdd(counter, total) : 0;
        for (int w = 0; w < NW; w++) wbase[w] += base;
    }
    __syncthreads();
    unsigned long long below = lane == 0 ? 0ull : (mask & ((~0ull) >> (64 - lane)));
    return pred ? wbase[wave] + __popcll(below) : -1;
}

// workgroup-level append of `cnt` (< NR_SLOTS_MAX + 1) items per thread with
// one atomic per workgroup: item k of a thread lands at the position passed to
// emit(k, pos); the workgroup's items are contiguous, ordered by rank k, then by
// thread.  When `tally` is given, the workgroup's sum of `t` is added to it (one
// more atomic).  Same-address device atomics serialise, so one per workgroup
// instead of one per rank matters for launches of thousands of workgroups.  The
// per-rank ballots are taken again when the positions are handed out, so no
// per-rank state is held in registers (the rank loops end at the wave's largest
// count).
template <int BLOCK, class Emit>
__device__ __forceinline__ void block_append_ranks(uint32_t cnt, int32_t* counter, uint32_t t, int32_t* tally,
                                                   Emit&& emit) {
    constexpr int NW = BLOCK / 64;
    constexpr int NE = NR_SLOTS_MAX * NW;              // (rank, wave) entries, rank-major
    constexpr int PER = (NE + 63) / 64;
    static_assert(NW <= 64, "one wave scans the workgroup");
    __shared__ int rbase[NE];
    __shared__ int tsum[NW];
    const int wave = threadIdx.x >> 6, lane = __lane_id();
    const unsigned long long below = lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
    int tw = 0;
    for (int k = 0; k < NR_SLOTS_MAX; k++) {
        const unsigned long long mk = __ballot(cnt > (uint32_t)k), tk = __ballot(t > (uint32_t)k);
        if (lane == 0) rbase[k * NW + wave] = __popcll(mk);
        tw += __popcll(tk);
        if ((mk | tk) == 0ull) {                       // wave-uniform: the later ranks are empty
            if (lane == 0)
                for (int r = k + 1; r < NR_SLOTS_MAX; r++) rbase[r * NW + wave] = 0;
            break;
        }
    }
    if (lane == 0) tsum[wave] = tw;
    __syncthreads();
    if (wave == 0) {
        int v[PER], sum = 0;
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const int e = lane * PER + j;
            v[j] = e < NE ? rbase[e] : 0;
            sum += v[j];
        }
        int inc = sum;                                  // inclusive scan over the lanes
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(inc, d);
            if (lane >= d) inc += y;
        }
        const int total = __shfl(inc, 63);
        int ts = lane < NW ? tsum[lane] : 0;
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) ts += __shfl_xor(ts, d);
        int base = 0;
        if (lane == 0) {
            if (total) base = atomicAdd(counter, total);
            if (tally && ts) atomicAdd(tally, ts);
        }
        int run = __shfl(base, 0) + inc - sum;
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const int e = lane * PER + j;
            if (e < NE) rbase[e] = run;
            run += v[j];
        }
    }
    __syncthreads();
    for (int k = 0; k < NR_SLOTS_MAX; k++) {
        const unsigned long long mk = __ballot(cnt > (uint32_t)k);
        if (mk == 0ull) break;                          // wave-uniform
        if (cnt > (uint32_t)k) emit(k, rbase[k * NW + wave] + __popcll(mk & below));
    }
}

// workgroup-level append of each thread's item to queue `key` (1 <= key < NQ,
// 0 = none): one atomic per workgroup and non-empty queue.
template <int BLOCK, int NQ>
__device__ __forceinline__ int block_append_keyed(int key, int32_t* counters) {
    constexpr int NW = BLOCK / 64;
    __shared__ int wcount[NQ][NW];
    __shared__ int wbase[NQ][NW];
    const int wave = threadIdx.x >> 6, lane = __lane_id();
    unsigned long long mine = 0ull;
    for (int q = 1; q < NQ; q++) {
        unsigned long long m = __ballot(key == q);
        if (key == q) mine = m;
        if (lane == 0) wcount[q][wave] = __popcll(m);
    }
    __syncthreads();
    if (threadIdx.x < NQ && threadIdx.x > 0) {
        const int q = threadIdx.x;
        int total = 0;
        for (int w = 0; w < NW; w++) { wbase[q][w] = total; total += wcount[q][w]; }
        int base = total ? atomicAdd(&counters[q], total) : 0;
        for (int w = 0; w < NW; w++) wbase[q][w] += base;
    }
    __syncthreads();
    if (key <= 0) return -1;
    unsigned long long below = lane == 0 ? 0ull : (mine & ((~0ull) >> (64 - lane)));
    return wbase[key][wave] + __popcll(below);
}

__device__ __forceinline__ Stack make_stack(int2* smem, const GScene& S, const StackCfg& SC) {
    int tid = threadIdx.x;
    Stack st;
    st.lds = (NR_LDS nr_i2*)(smem + (tid >> 6) * LDS_STACK_LEVELS * 64);
    st.spill = (NR_GLB nr_i2*)SC.spill;
    st.lane = tid & 63;
    st.lds_cap = LDS_STACK_LEVELS;
    st.cap = S.stack_depth;
    st.slot = (size_t)blockIdx.x * blockDim.x + tid;
    st.spill_stride = SC.stride;
    return st;
}

// ---------------------------------------------------------------------------
// camera (engine/include/engine/camera.h:186-278)
// ---------------------------------------------------------------------------
__device__ __forceinline__ Ray camera_ray(const GCamera& c, uint32_t x, uint32_t y, uint32_t pixel, uint32_t s,
                                          uint32_t seed, Gen& g) {
    float bx = c.pattern[2 * s], by = c.pattern[2 * s + 1];
    f2 shift = pcg_vec2((float)pixel, seed);
    float ox = fmodf(bx + shift.x, 1.0f) - 0.5f;
    float oy = fmodf(by + shift.y, 1.0f) - 0.5f;
    float u = ((float)x + ox - (c.res[0] - 1.0f) * 0.5f) * (c.uvw[0] / c.res[0]);
    float v = ((float)y + oy - (c.res[1] - 1.0f) * 0.5f) * (c.uvw[1] / c.res[1]);
    float w = -c.focal_length;
    f3 center = mv(c.R, mk3(u, v, w));
    float focus_scale = c.focus_distance / c.focal_length;
    f3 pos = ld3(c.pos);
    f3 focus = pos + center * focus_scale;
    f3 origin = pos;
    if (!(c.defocus_angle <= 0.0f)) {
        f3 du = mk3(c.R[0], c.R[3], c.R[6]) * c.defocus_radius;
        f3 dv = mk3(c.R[1], c.R[4], c.R[7]) * c.defocus_radius;
        float lx, ly;
        while (true) {                                                           // vector.h:303-311
            lx = g.rf(-1.0f, 1.0f);
            ly = g.rf(-1.0f, 1.0f);
            if (lx * lx + ly * ly < 1.0f) break;
        }
        origin = pos + (du * lx + dv * ly);
    }
    if (c.shutter_time > 0.0f) (void)g.uniform();                                // ray time draw
    return mkray(origin, normalize(focus - origin));
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
template <int TY, int LAY>
__global__ void __launch_bounds__(128) NR_TRAV_ATTR k_trace_closest(const GScene* __restrict__ Sp,
                                                                    const float* __restrict__ rays, int n,
                                                                    nr_hit* __restrict__ hits,
                                                                    int32_t* __restrict__ counts, StackCfg SC) {
    extern __shared__ int2 smem[];
    const GScene& S = *Sp;
    Stack st = make_stack(smem, S, SC);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* q = rays + 6 * (size_t)i;
    Ray r = mkray(mk3(q[0], q[1], q[2]), mk3(q[3], q[4], q[5]));
    HRec rec;
    hrec_clear(rec);
    bool of = false;
    int oi = scene_closest<TY, LAY>(S, r, rec, st, of);
    if (of) atomicOr(&counts[CNT_ERR], 1);
    nr_hit h;
    if (oi < 0) {
        memset(&h, 0, sizeof(h));
        h.object_id = -1;
        h.prim = -1;
    } else {
        const GObject& o = S.objects[oi];
        h.object_id = o.object_id;
        h.material_id = o.material;
        h.prim = rec.prim;
        h.front_face = rec.front ? 1 : 0;
        h.t = rec.t;
        h.p[0] = rec.p.x; h.p[1] = rec.p.y; h.p[2] = rec.p.z;
        h.n[0] = rec.n.x; h.n[1] = rec.n.y; h.n[2] = rec.n.z;
        h.tangent[0] = rec.tg.x; h.tangent[1] = rec.tg.y; h.tangent[2] = rec.tg.z;
        h.uv[0] = rec.uv.x; h.uv[1] = rec.uv.y;
    }
    hits[i] = h;
}

// ---------------------------------------------------------------------------
// viewport preview, picking and selection (engine/src/trace.cu:99-188,
// scene/src/scene_interface.cu:3-50)
// ---------------------------------------------------------------------------
// 128-thread workgroup b of a 1-D grid covers the 16 x 8 pixel tile b (tiles in
// row-major order); a 1-D grid keeps blockIdx.x * 128 + tid a unique stack-spill slot
__device__ __forceinline__ bool tile16x8(uint32_t W, uint32_t H, uint32_t& x, uint32_t& y) {
    const uint32_t tiles_x = (W + 15) >> 4;
    x = (blockIdx.x % tiles_x) * 16 + (threadIdx.x & 15);
    y = (blockIdx.x / tiles_x) * 8 + (threadIdx.x >> 4);
    return x < W && y < H;
}

// DeviceCamera::get_ray_center (camera.h:181-184) = get_focus_point(x, y, 0, 0, 0):
// the offset of sample 0 of pixel 0 under seed 0, the same for every pixel
__device__ __forceinline__ Ray camera_center_ray(const GCamera& c, uint32_t x, uint32_t y) {
    float bx = c.pattern[0], by = c.pattern[1];
    f2 shift = pcg_vec2(0.0f, 0u);
    float ox = fmodf(bx + shift.x, 1.0f) - 0.5f;
    float oy = fmodf(by + shift.y, 1.0f) - 0.5f;
    float u = ((float)x + ox - (c.res[0] - 1.0f) * 0.5f) * (c.uvw[0] / c.res[0]);
    float v = ((float)y + oy - (c.res[1] - 1.0f) * 0.5f) * (c.uvw[1] / c.res[1]);
    f3 center = mv(c.R, mk3(u, v, -c.focal_length));
    f3 pos = ld3(c.pos);
    f3 focus = pos + center * (c.focus_distance / c.focal_length);
    return mkray(pos, normalize(focus - pos));
}

// trace_viewport_kernel: closest hit of the centre ray; a hit is shaded with a
// headlight and a fixed directional light, a miss shows the sky (and the axis
// grid).  The AOVs are cleared and then added to, so each value is 0 + v.
template <int TY>
__global__ void __launch_bounds__(128) NR_TRAV_ATTR k_viewport(const GScene* __restrict__ Sp, GCamera cam, int show_grid,
                                                               float* __restrict__ beauty, float* __restrict__ diffuse,
                                                               float* __restrict__ normal, int32_t* __restrict__ counts,
                                                               StackCfg SC) {
    extern __shared__ int2 smem[];
    const GScene& S = *Sp;
    Stack st = make_stack(smem, S, SC);
    uint32_t x, y;
    if (!tile16x8(cam.W, cam.H, x, y)) return;
    Ray r = camera_center_ray(cam, x, y);
    HRec rec;
    hrec_clear(rec);
    bool of = false;
    int oi = scene_closest<TY>(S, r, rec, st, of);
    if (of) atomicOr(&counts[CNT_ERR], 1);
    f3 b, dcol = s3(0.0f), n = s3(0.0f);
    if (oi < 0) {
        f3 sky = sky_sample(S.sky, r.d);
        f3 gc;
        float ga;
        b = (show_grid && triplanar_grid(r, gc, ga)) ? lerp3(sky, gc, ga) : sky;
    } else {
        f3 light_vector = ld3(cam.pos) - rec.p;
        float light_dist = vlength(light_vector);
        f3 headlight_dir = normalize(light_vector);
        float headlight = dot(headlight_dir, rec.n) * 0.5f + 0.5f;
        headlight *= 1.0f / (light_dist * light_dist);
        headlight = fminf(0.75f, headlight);
        float directional = (dot(mk3(1.0f, 1.0f, 1.0f), rec.n) + 1.0f) * 0.5f;
        directional = fmaxf(0.1f, directional);
        f3 col = viewport_color(S, S.materials[S.objects[oi].material], rec.uv);
        b = col * fmaxf(headlight, directional);
        dcol = col;
        n = rec.n;
    }
    const size_t HW = (size_t)cam.H * cam.W, i = (size_t)y * cam.W + x;
    beauty[i] = 0.0f + b.x; beauty[HW + i] = 0.0f + b.y; beauty[2 * HW + i] = 0.0f + b.z;
    diffuse[i] = 0.0f + dcol.x; diffuse[HW + i] = 0.0f + dcol.y; diffuse[2 * HW + i] = 0.0f + dcol.z;
    normal[i] = 0.0f + n.x; normal[HW + i] = 0.0f + n.y; normal[2 * HW + i] = 0.0f + n.z;
}

// mask_selected_kernel: 255 where the screen-space ray of the pixel centre hits
// object `object_id`
template <int TY>
__global__ void __launch_bounds__(128) NR_TRAV_ATTR k_select_mask(const GScene* __restrict__ Sp, GCamera cam,
                                                                  int32_t object_id, uint8_t* __restrict__ mask,
                                                                  int32_t* __restrict__ counts, StackCfg SC) {
    extern __shared__ int2 smem[];
    const GScene& S = *Sp;
    Stack st = make_stack(smem, S, SC);
    uint32_t x, y;
    if (!tile16x8(cam.W, cam.H, x, y)) return;
    float su = ((float)x + 0.5f) / (float)cam.W;
    float sv = ((float)y + 0.5f) / (float)cam.H;
    Ray r = screen_space_ray(cam, su, sv);
    HRec rec;
    hrec_clear(rec);
    bool of = false;
    int oi = scene_closest<TY>(S, r, rec, st, of);
    if (of) atomicOr(&counts[CNT_ERR], 1);
    mask[(size_t)y * cam.W + x] = (oi >= 0 && S.objects[oi].object_id == object_id) ? 255u : 0u;
}

// build_outline_kernel: 255 on unmasked pixels with a masked pixel within the
// (2r+1)^2 square around them
__global__ void __launch_bounds__(128) k_outline(const uint8_t* __restrict__ base, uint8_t* __restrict__ out,
                                                 uint32_t H, uint32_t W, int radius) {
    uint32_t x, y;
    if (!tile16x8(W, H, x, y)) return;
    const size_t i = (size_t)y * W + x;
    uint8_t o = 0u;
    if (base[i] == 0) {
        for (int dy = -radius; dy <= radius && !o; dy++) {
            const int ny = (int)y + dy;
            if (ny < 0 || ny >= (int)H) continue;
            for (int dx = -radius; dx <= radius; dx++) {
                const int nx = (int)x + dx;
                if (nx < 0 || nx >= (int)W) continue;
                if (base[(size_t)ny * W + nx] != 0) { o = 255u; break; }
            }
        }
    }
    out[i] = o;
}

// hit_test_ray_kernel: one screen-space ray
template <int TY>
__global__ void __launch_bounds__(64) NR_TRAV_ATTR k_pick(const GScene* __restrict__ Sp, GCamera cam, float su, float sv,
                                                          nr_hit* __restrict__ out, int32_t* __restrict__ counts,
                                                          StackCfg SC) {
    extern __shared__ int2 smem[];
    const GScene& S = *Sp;
    Stack st = make_stack(smem, S, SC);
    if (threadIdx.x != 0) return;
    Ray r = screen_space_ray(cam, su, sv);
    HRec rec;
    hrec_clear(rec);
    bool of = false;
    int oi = scene_closest<TY>(S, r, rec, st, of);
    if (of) atomicOr(&counts[CNT_ERR], 1);
    nr_hit h;
    memset(&h, 0, sizeof(h));
    h.object_id = -1;
    h.prim = -1;
    if (oi >= 0) {
        const GObject& o = S.objects[oi];
        h.object_id = o.object_id;
        h.material_id = o.material;
        h.prim = rec.prim;
        h.front_face = rec.front ? 1 : 0;
        h.t = rec.t;
        h.p[0] = rec.p.x; h.p[1] = rec.p.y; h.p[2] = rec.p.z;
        h.n[0] = rec.n.x; h.n[1] = rec.n.y; h.n[2] = rec.n.z;
        h.tangent[0] = rec.tg.x; h.tangent[1] = rec.tg.y; h.tangent[2] = rec.tg.z;
        h.uv[0] = rec.uv.x; h.uv[1] = rec.uv.y;
    }
    *out = h;
}

// composite_overlay_kernel (data/src/stream.cu:177-202): blend `rgb` into a
// planar uint8 image with the mask as alpha
__global__ void __launch_bounds__(256) k_overlay(uint8_t* __restrict__ img, const uint8_t* __restrict__ mask, size_t HW,
                                                 uint32_t r, uint32_t g, uint32_t b) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= HW) return;
    const uint32_t alpha = mask[i];
    if (alpha == 0u) return;
    const uint32_t inv = 255u - alpha;
    img[i] = (uint8_t)((r * alpha + img[i] * inv + 127u) / 255u);
    img[HW + i] = (uint8_t)((g * alpha + img[HW + i] * inv + 127u) / 255u);
    img[2 * HW + i] = (uint8_t)((b * alpha + img[2 * HW + i] * inv + 127u) / 255u);
}

// Streaming sample passes.  Every pixel owns one path slot.  When a pixel's
// path ends (sky, no scatter, zero pdf, roulette or max depth) it is flagged;
// k_regen folds the finished sample into the running-mean AOVs exactly as
// accumulate_samples_kernel does (data.cu:28-40) and starts the pixel's next
// sample at once, so traversal launches stay full.  A pixel's samples are
// folded strictly in order, so the AOVs equal the reference's pass-by-pass
// result bit for bit.

// 8x8 pixel tiles, tiles left to right within bands of 8 rows, the views one
// after another: thread t of the tiled grid maps to view v, pixel (x, y);
// partial tiles at the edges are skipped.
__device__ __forceinline__ bool tile_pixel(const TraceArgs& A, uint32_t t, uint32_t& v, uint32_t& x, uint32_t& y) {
    const uint32_t tile = t >> 6, within = t & 63;
    v = 0;
    while (v + 1 < A.n_views && tile >= A.views[v + 1].tile0) v++;
    const GView& V = A.views[v];
    const uint32_t lt = tile - V.tile0;
    x = (lt % V.tiles_x) * 8 + (within & 7);
    y = (lt / V.tiles_x) * 8 + (within >> 3);
    return x < V.cam.W && y < V.cam.H;
}

// view of a path-array pixel
__device__ __forceinline__ uint32_t pixel_view(const TraceArgs& A, uint32_t g) {
    uint32_t v = 0;
    while (v + 1 < A.n_views && g >= A.views[v + 1].pix0) v++;
    return v;
}

// Generator of pixel i's sample s: the reference's 32-bit subsequence i + s*W*H
// (trace.cu:58), or in fast mode the 64-bit subsequence (s << 32) | i, which
// never repeats across samples.
__device__ __forceinline__ Gen path_gen(const TraceArgs& A, uint32_t i, uint32_t s, uint32_t draw, uint32_t WH) {
    Gen g{A.seed, i + s * WH, draw};
    if (A.fast) { g.subseq = i; g.subhi = s; }
    return g;
}

// atten.w of a path: the solid-angle pdf of the direction it was scattered
// into (fast mode's MIS weight at a sky hit), or -1 after the camera or a
// specular bounce
__device__ __forceinline__ uint32_t path_pixel(const PathPtrs& P, uint32_t path) { return path % P.cap; }

// Starts sample s of pixel (x, y) of view V in its slot; returns the path id.
__device__ __forceinline__ uint32_t start_path(const GView& V, PathPtrs& P, uint32_t x, uint32_t y, uint32_t s,
                                               const TraceArgs& A) {
    const GCamera& cam = V.cam;
    uint32_t i = y * cam.W + x;
    uint32_t p = V.pix0 + i + (s % P.slots) * P.cap;
    Gen g = path_gen(A, i, s, 0u, V.npix);
    Ray r = camera_ray(cam, x, y, i, s, A.seed, g);
    P.o[2 * (size_t)p] = make_float4(r.o.x, r.o.y, r.o.z, __int_as_float(0));
    P.d[2 * (size_t)p] = make_float4(r.d.x, r.d.y, r.d.z, __int_as_float((int)s));
    P.atten[4 * (size_t)p] = make_float4(1.0f, 1.0f, 1.0f, -1.0f);
    // L's radiance, diff and nrm are not cleared: the first segment of the path
    // writes all three (radiance counted from zero at bounce 0).  L.w carries the
    // path's Philox draw counter.
    P.L[4 * (size_t)p] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(g.draw));
    P.finished[p] = 0;
    return p;
}

// running-mean fold of one finished sample (accumulate_samples_kernel): the
// same per-channel arithmetic on a pixel-interleaved copy of the layer
__device__ __forceinline__ void fold(float4* acc, float4 v, float a, float b) {
    float4 m = *acc;
    m.x = m.x * a + v.x * b;
    m.y = m.y * a + v.y * b;
    m.z = m.z * a + v.z * b;
    m.w = m.w * a + v.w * b;          // DEPTH (normal layer); 0 for the others
    *acc = m;
}

// Per-tile error estimate of a denoised display image D against two independent
// plain renders J1, J2 of the same pixels (display images, planar 3 x H x W):
// the mean over the tile's pixels and channels of (D - J1)^2 - (J1 - J2)^2 / 2.
// With J1 independent of D, E[(D - J1)^2] = MSE(D) + Var(J1) + bias(J1)^2 +
// 2 E[D - T] bias(J1) (T the converged display image) and E[(J1 - J2)^2] / 2 =
// Var(J1): an unbiased MSE estimate where the judges are unbiased, and above
// it by the judges' display-space bias (the display transform is concave, so a
// few-sample render displays darker than its expectation).  One workgroup per
// tile; fixed-order tree reduction (deterministic).
__global__ void __launch_bounds__(256) k_tile_error(const float* __restrict__ D, const float* __restrict__ J1,
                                                    const float* __restrict__ J2, uint32_t H, uint32_t W,
                                                    uint32_t ts, uint32_t cols, float* __restrict__ out) {
    __shared__ float red[256];
    const uint32_t tile = blockIdx.x, tx = tile % cols, ty = tile / cols;
    const uint32_t x0 = tx * ts, y0 = ty * ts;
    const uint32_t w = min(ts, W - x0), h = min(ts, H - y0);
    const size_t HW = (size_t)H * W;
    float acc = 0.0f;
    for (uint32_t k = threadIdx.x; k < w * h; k += blockDim.x) {
        const size_t i = (size_t)(y0 + k / w) * W + x0 + k % w;
        for (int c = 0; c < 3; c++) {
            const float d = D[c * HW + i], a = J1[c * HW + i], b = J2[c * HW + i];
            acc += (d - a) * (d - a) - 0.5f * (a - b) * (a - b);
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[tile] = red[0] / (float)(3 * w * h);
}

// 4-wide layout of a mesh tree (fast mode): wide node i = the grandchildren of
// binary node i, in left-to-right order.
__global__ void __launch_bounds__(256) k_wide_nodes(const GNode* __restrict__ nodes, int32_t n, GWNode* __restrict__ out) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const GNode b = nodes[i];
    GWNode w;
    int k = 0;
    auto put = [&](const float* mn, const float* mx, int32_t ref) {
        for (int a = 0; a < 3; a++) { w.lo[a][k] = mn[a]; w.hi[a][k] = mx[a]; }
        w.ref[k] = ref;
        k++;
    };
    for (int side = 0; side < 2; side++) {
        const int32_t ch = side ? b.right : b.left;
        if (ch >= 0) {
            const GNode c = nodes[ch];
            put(c.lmin, c.lmax, c.left);
            put(c.rmin, c.rmax, c.right);
        } else {
            put(side ? b.rmin : b.lmin, side ? b.rmax : b.lmax, ch);
        }
    }
    for (; k < 4; k++) {
        for (int a = 0; a < 3; a++) { w.lo[a][k] = 0.0f; w.hi[a][k] = 0.0f; }
        w.ref[k] = NR_WIDE_EMPTY;
    }
    w.pad[0] = w.pad[1] = w.pad[2] = w.pad[3] = 0;
    out[i] = w;
}

// Quantised 4-wide layout (fast mode) from the 4-wide one: per axis the grid
// starts at the union box's lower bound and steps by the smallest power of two
// 2^e (e >= -126) with origin + 255 * 2^e >= the union's upper bound; each child
// plane is rounded outward to the grid and checked with the float additions
// the traversal decodes with (origin + q * 2^e, q * 2^e exact), stepping q
// outward until the decoded box contains the exact one.  A child box of zero
// thickness on some axis is left out (ref NR_WIDE_EMPTY): its triangles all lie
// in that axis plane, and the reference's slab test (AABB::hit_t, empty when
// tmax <= tmin) never enters such a box for a ray crossing the plane, while a ray
// in the plane is parallel to every one of its triangles, which the triangle
// test rejects.
__device__ __forceinline__ float qdecode(float origin, uint32_t q, float s) { return origin + (float)q * s; }
__global__ void __launch_bounds__(256) k_quant_nodes(const GWNode* __restrict__ wn, int32_t n, GQNode* __restrict__ out) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    GWNode w = wn[i];
    for (int k = 0; k < 4; k++)
        if (w.ref[k] != NR_WIDE_EMPTY &&
            (w.lo[0][k] == w.hi[0][k] || w.lo[1][k] == w.hi[1][k] || w.lo[2][k] == w.hi[2][k]))
            w.ref[k] = NR_WIDE_EMPTY;
    GQNode q;
    uint32_t exps = 0;
    for (int a = 0; a < 3; a++) {
        float lo = NR_FMAX, hi = -NR_FMAX;
        for (int k = 0; k < 4; k++)
            if (w.ref[k] != NR_WIDE_EMPTY) { lo = fminf(lo, w.lo[a][k]); hi = fmaxf(hi, w.hi[a][k]); }
        if (!(lo <= hi)) { lo = 0.0f; hi = 0.0f; }
        int e = -126;
        const float ext = hi - lo;
        if (ext > 0.0f) {
            int k2 = 0;
            (void)frexpf(ext / 255.0f, &k2);                 // ext / 255 < 2^k2
            e = max(k2 - 1, -126);
        }
        auto scale = [](int ee) { return __uint_as_float((uint32_t)(ee + 127) << 23); };
        while (e < 127 && qdecode(lo, 255u, scale(e)) < hi) e++;
        const float sc = scale(e);
        uint32_t ql = 0, qh = 0;
        for (int k = 0; k < 4; k++) {
            uint32_t a0 = 0, b0 = 0;
            if (w.ref[k] != NR_WIDE_EMPTY) {
                float fl = floorf((w.lo[a][k] - lo) / sc), fh = ceilf((w.hi[a][k] - lo) / sc);
                int il = (int)fminf(fmaxf(fl, 0.0f), 255.0f), ih = (int)fminf(fmaxf(fh, 0.0f), 255.0f);
                while (il > 0 && qdecode(lo, (uint32_t)il, sc) > w.lo[a][k]) il--;
                while (ih < 255 && qdecode(lo, (uint32_t)ih, sc) < w.hi[a][k]) ih++;
                a0 = (uint32_t)il;
                b0 = (uint32_t)ih;
            }
            ql |= a0 << (8 * k);
            qh |= b0 << (8 * k);
        }
        q.origin[a] = lo;
        q.qlo[a] = ql;
        q.qhi[a] = qh;
        exps |= (uint32_t)(e + 127) << (8 * a);
    }
    q.exps = exps;
    for (int k = 0; k < 4; k++) q.ref[k] = w.ref[k];
    q.pad[0] = q.pad[1] = 0;
    out[i] = q;
}

// The planar C x H x W layers stay the state between calls; a render folds into
// pixel-interleaved copies (one 16-byte read-modify-write per layer and pixel)
// and writes them back at the end.
__global__ void __launch_bounds__(256) k_pack_acc(const float* __restrict__ b, const float* __restrict__ d,
                                                  const float* __restrict__ n, const float* __restrict__ z,
                                                  float4* __restrict__ acc, uint32_t HW, uint32_t cap) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= HW) return;
    acc[i] = make_float4(b[i], b[HW + i], b[2 * (size_t)HW + i], 0.0f);
    acc[cap + i] = make_float4(d[i], d[HW + i], d[2 * (size_t)HW + i], 0.0f);
    acc[2 * (size_t)cap + i] = make_float4(n[i], n[HW + i], n[2 * (size_t)HW + i], z[i]);
}

__global__ void __launch_bounds__(256) k_unpack_acc(const float4* __restrict__ acc, float* __restrict__ b,
                                                    float* __restrict__ d, float* __restrict__ n,
                                                    float* __restrict__ dep, uint32_t HW, uint32_t cap) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= HW) return;
    float4 x = acc[i], y = acc[cap + i], z = acc[2 * (size_t)cap + i];
    b[i] = x.x; b[HW + i] = x.y; b[2 * (size_t)HW + i] = x.z;
    d[i] = y.x; d[HW + i] = y.y; d[2 * (size_t)HW + i] = y.z;
    n[i] = z.x; n[HW + i] = z.y; n[2 * (size_t)HW + i] = z.z;
    dep[i] = z.w;
}

// First `slots` samples of every pixel (fewer if the range is shorter),
// appended to the live queue in tile order, one workgroup-aggregated append per
// sample rank.  The group's counters were cleared by k_clear_counts.
__global__ void __launch_bounds__(256) k_start(PathPtrs P, uint32_t s, TraceArgs A) {
    uint32_t t = A.t0 + blockIdx.x * blockDim.x + threadIdx.x, v = 0, x = 0, y = 0;
    const bool valid = t < A.t1 && tile_pixel(A, t, v, x, y);
    uint32_t g = 0, m = 0;                                          // samples started for the pixel
    if (valid) {
        const uint32_t end = pixel_end(A, x, y);
        m = end > s ? min(P.slots, end - s) : 0u;
        const GView& V = A.views[v];
        g = V.pix0 + y * V.cam.W + x;
        P.nf[g] = s;
        P.ns[g] = s + m;
    }
    block_append_ranks<256>(m, &P.counts[CNT_ACT + 0], 0u, nullptr, [&](int k, int pos) {
        P.qact[pos] = (int32_t)start_path(A.views[v], P, x, y, s + (uint32_t)k, A);
    });
    if (valid)
        for (uint32_t k = m; k < P.slots; k++) P.finished[g + ((s + k) % P.slots) * P.cap] = 2;
}

// every counter of a group but the overflow flag (shared by the groups; cleared
// once per render call by k_clear_err)
__global__ void k_clear_counts(int32_t* counts) {
    if (threadIdx.x < CNT_N && threadIdx.x != CNT_ERR) counts[threadIdx.x] = 0;
}

__global__ void k_clear_err(int32_t* err) {
    if (threadIdx.x == 0) *err = 0;
}

// Clears the counters one iteration appends to: material queues and the next live queue.
__global__ void k_counts(int32_t* counts, int32_t in) {
    int t = threadIdx.x;
    if (t < NR_MAT_TYPES) counts[CNT_MAT + t] = 0;
    if (t == 0) counts[CNT_ACT + (1 - in)] = 0;
    if (t == 1) counts[CNT_SHD] = 0;
    if (t == 2) counts[CNT_OCC] = 0;
    if (t == 3) counts[CNT_PEND] = 0;
}

// One extension segment of a path (trace.cu:19-22 and the closest hit of
// trace_ray): a miss adds the sky and ends the path, a hit stores its record.
// Returns the material type of the hit, 0 on a miss.
template <int TY, bool FAST>
__device__ __forceinline__ int extend_one(const GScene& S, const PathPtrs& P, int path, Stack& st) {
    float4 o4 = P.o[2 * (size_t)path], d4 = P.d[2 * (size_t)path];
    Ray r = mkray(mk3(o4.x, o4.y, o4.z), mk3(d4.x, d4.y, d4.z));
    HRec rec;
    hrec_clear(rec);
    bool of = false;
    int oi = scene_closest<TY, FAST ? NR_EXT_FAST_LAY : NR_LAY_BINARY>(S, r, rec, st, of);
    if (of) atomicOr(P.err, 1);
    if (oi < 0) {                                                               // trace.cu:19-22
        const bool first = __float_as_int(o4.w) == 0;
        float4 a4 = P.atten[4 * (size_t)path], l4 = first ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : P.L[4 * (size_t)path];
        if (first) {                                                            // no hit: zero AOVs
            P.diff[4 * (size_t)path] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            P.nrm[4 * (size_t)path] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        f3 Le = sky_sample<FAST>(S.sky, r.d);
        // fast mode: a scattered ray that finds the sky shares it with the
        // shadow rays of next-event estimation (power heuristic)
        if (FAST && S.sky.marg && a4.w > 0.0f) Le = Le * power_heuristic(a4.w, sky_cdf_pdf(S.sky, r.d));
        f3 L = mk3(l4.x, l4.y, l4.z) + mk3(a4.x, a4.y, a4.z) * Le;
        P.L[4 * (size_t)path] = make_float4(L.x, L.y, L.z, 0.0f);
        P.finished[path] = 1;
        return 0;
    }
    const GObject& ob = S.objects[oi];
    GHit h;
    h.p[0] = rec.p.x; h.p[1] = rec.p.y; h.p[2] = rec.p.z;
    h.n[0] = rec.n.x; h.n[1] = rec.n.y; h.n[2] = rec.n.z;
    h.tg[0] = rec.tg.x; h.tg[1] = rec.tg.y; h.tg[2] = rec.tg.z;
    h.uv[0] = rec.uv.x; h.uv[1] = rec.uv.y;
    h.matf = (uint32_t)ob.material | (rec.front ? 0x80000000u : 0u);
    P.hit[path] = h;
    return ob.pad[0];
}

template <int TY, bool FAST>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(TY == NR_TY_ALL ? (NR_TRAV_OCC > 0 ? NR_TRAV_OCC : 1) : FAST ? NR_EXT_FAST_WAVES : NR_EXT_WAVES, 8)))
k_extend(const GScene* __restrict__ Sp, PathPtrs P, TraceArgs A, StackCfg SC) {
    extern __shared__ int2 smem[];
    const GScene& S = *Sp;
    Stack st = make_stack(smem, S, SC);
    const int n = P.counts[CNT_ACT + A.in];
    if ((int)(blockIdx.x * blockDim.x) >= n) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) atomicAdd(reinterpret_cast<unsigned long long*>(P.counts + CNT_TRACED), (unsigned long long)n);
    int path = -1;
    int mtype = 0;
    if (i < n) {
        path = P.qact[(size_t)A.in * P.qcap + i];
        mtype = extend_one<TY, FAST>(S, P, path, st);
    }
    const int slot = block_append_keyed<128, NR_MAT_TYPES>(mtype, &P.counts[CNT_MAT]);
    if (mtype > 0) P.qmat[(size_t)mtype * P.qcap + slot] = path;
}

// material evaluate (material.h:59-67, 152-180, 351-357)
template <int M, bool F = false>
__device__ __forceinline__ f3 mat_eval(const GScene& S, const GMaterial& m, f3 n, f3 sn, f2 uv, f3 view, f3 l) {
    if (M == NR_MAT_LAMBERTIAN) {
        float c = dot(n, normalize(l));
        if (c <= 0.0f) return s3(0.0f);
        return tex_sample(S, m.tex[0], uv) * (c / NR_PI);
    } else if (M == NR_MAT_PBR) {
        float n_dot_l = dot(sn, l);
        float n_dot_v = dot(sn, view);
        if (n_dot_l <= 0.0f || n_dot_v <= 0.0f) return s3(0.0f);
        f3 albedo = tex_sample(S, m.tex[0], uv);
        float roughness = tex_sample(S, m.tex[1], uv).x;
        float metallic = tex_sample(S, m.tex[2], uv).x;
        float alpha = fmaxf(roughness * roughness, 1e-3f);
        f3 h = normalize(view + l);
        float n_dot_h = fmaxf(dot(sn, h), 0.0f);
        float v_dot_h = fmaxf(dot(view, h), 0.0f);
        f3 F0 = s3(0.04f) * (1.0f - metallic) + albedo * metallic;
        f3 Fr = F0 + (s3(1.0f) - F0) * Tm<F>::pow5(1.0f - v_dot_h);
        float D = ggx_D(n_dot_h, alpha);
        float G = smith_g1(n_dot_l, alpha) * smith_g1(n_dot_v, alpha);
        f3 spec = Fr * (D * G / fmaxf(4.0f * n_dot_v * n_dot_l, 1e-4f));
        f3 diff = (albedo * (1.0f - metallic)) * (1.0f / NR_PI);
        return (diff + spec) * n_dot_l;
    } else {
        return tex_sample(S, m.tex[0], uv) * (1.0f / NR_PI4);
    }
}

// Fast mode adds next-event estimation to the HDRI sky at every non-specular
// vertex whose scattered ray would still be traced (bounce + 1 < depth): a
// direction from the sky's tables (marginal CDF staged in LDS, per-row alias
// tables), weighted against the
// scattering technique with the power heuristic, queued as a shadow ray whose
// contribution k_shadow adds if it reaches the sky.
// One shading step of a path at its stored hit (trace.cu:23-40, 63-78):
// emission, scatter, mixture PDF, fast mode's next-event sample, bounce-0
// AOVs, Russian roulette.  cont: the path continues; shadow: a shadow ray is
// pending (fast mode).  marg: the HDRI's marginal CDF (LDS or global).
template <int M, bool FAST>
__device__ __forceinline__ void shade_one(const GScene& S, const PathPtrs& P, const TraceArgs& A, int path,
                                          const float* marg, bool& cont, bool& shadow) {
    constexpr bool NEE_M = FAST && (M == NR_MAT_LAMBERTIAN || M == NR_MAT_PBR || M == NR_MAT_ISOTROPIC);
    GHit h = P.hit[path];
    float4 o4 = P.o[2 * (size_t)path], d4 = P.d[2 * (size_t)path], a4 = P.atten[4 * (size_t)path];
    const uint32_t bounce = (uint32_t)__float_as_int(o4.w);
    const float4 lr = P.L[4 * (size_t)path];       // w: the draw counter
    const float4 l4 = bounce == 0 ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : lr;
    const uint32_t sample = (uint32_t)__float_as_int(d4.w);
    f3 din = mk3(d4.x, d4.y, d4.z);
    f3 atten = mk3(a4.x, a4.y, a4.z);
    f3 L = mk3(l4.x, l4.y, l4.z);
    f3 p = ld3(h.p), nn = ld3(h.n), tg = ld3(h.tg);
    f2 uv; uv.x = h.uv[0]; uv.y = h.uv[1];
    bool front = (h.matf >> 31) != 0;
    const uint32_t gp = path_pixel(P, (uint32_t)path);
    const GView& V = A.views[pixel_view(A, gp)];
    Gen g = path_gen(A, gp - V.pix0, sample, __float_as_uint(lr.w), V.npix);
    const GMaterial& m = S.materials[h.matf & 0x7FFFFFFFu];
    // emission (material.h:182-188, 306-312)
    f3 em = s3(0.0f);
    if (M == NR_MAT_PBR && front) em = tex_sample(S, m.tex[3], uv);
    if (M == NR_MAT_EMISSIVE && front) em = tex_sample(S, m.tex[0], uv) * m.brightness;
    L = L + atten * em;
    bool alive = false;
    f3 nd = din;
    float sc_pdf = -1.0f;                                                   // pdf of the scattered direction (-1: specular)
    if (M == NR_MAT_DIELECTRIC) {                                           // material.h:220-244
        float ri = front ? (1.0f / m.ior) : m.ior;
        f3 unit = normalize(din);
        float cos_theta = fminf(dot(-unit, nn), 1.0f);
        float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
        bool refl = (double)(ri * sin_theta) > 1.0;
        if (!refl) {
            float r0 = (1.0f - ri) / (1.0f + ri);
            r0 = r0 * r0;
            float refl_p = r0 + (1.0f - r0) * Tm<FAST>::pow5(1.0f - cos_theta);
            refl = refl_p > g.uniform();
        }
        nd = refl ? reflect(unit, nn) : refract(unit, nn, ri);
        atten = atten * ld3(m.tint);
        alive = true;
    } else if (M != NR_MAT_EMISSIVE) {
        ONB uvw;
        f3 view = s3(0.0f), sn = nn;
        float alpha = 0.0f, sp = 0.0f;
        if (M == NR_MAT_LAMBERTIAN) {
            uvw.build(nn);
        } else if (M == NR_MAT_PBR) {                                       // material.h:127-150
            sn = pbr_shading_normal(S, m, nn, tg, uv);
            f3 albedo = tex_sample(S, m.tex[0], uv);
            float metallic = tex_sample(S, m.tex[2], uv).x;
            f3 F0 = s3(0.04f) * (1.0f - metallic) + albedo * metallic;
            sp = fminf(0.9f, fmaxf(0.1f, fmaxf(F0.x, fmaxf(F0.y, F0.z))));
            view = normalize(-din);
            float rough = tex_sample(S, m.tex[1], uv).x;
            alpha = fmaxf(rough * rough, 1e-3f);
            uvw.build(sn);
        }
        f3 dir;
        float pdf;
        if (S.first_light >= 0) {                                          // brdf.cu:16-23
            const GObject& Lo = S.objects[S.first_light];
            if (g.rf() < 0.5f) {
                if (M == NR_MAT_LAMBERTIAN) dir = uvw.xf(cosine_dir<FAST>(g));
                else if (M == NR_MAT_PBR) dir = pbr_pdf_generate<FAST>(uvw, view, alpha, sp, g);
                else dir = unit_vector(g);
            } else {
                dir = light_random<FAST>(Lo, p, g);
            }
            float mp;
            if (M == NR_MAT_LAMBERTIAN) mp = cosine_pdf_value(uvw, dir);
            else if (M == NR_MAT_PBR) mp = pbr_pdf_value(uvw, view, alpha, sp, dir);
            else mp = 1.0f / NR_PI4;
            pdf = 0.5f * mp + 0.5f * light_pdf(Lo, p, dir);
        } else {
            if (M == NR_MAT_LAMBERTIAN) { dir = uvw.xf(cosine_dir<FAST>(g)); pdf = cosine_pdf_value(uvw, dir); }
            else if (M == NR_MAT_PBR) { dir = pbr_pdf_generate<FAST>(uvw, view, alpha, sp, g); pdf = pbr_pdf_value(uvw, view, alpha, sp, dir); }
            else { dir = unit_vector(g); pdf = 1.0f / NR_PI4; }
        }
        if (NEE_M && marg && bounce + 1 < A.depth) {
            // drawn after the scattering sample: the scattered direction does
            // not depend on whether the sky has sampling tables
            float u1 = g.uniform(), u2 = g.uniform(), ps;
            f3 ld = sky_cdf_sample(S.sky, marg, u1, u2, ps);
            if (ps > 0.0f) {
                f3 f = mat_eval<M, FAST>(S, m, nn, sn, uv, normalize(-din), ld);
                if (f.x > 0.0f || f.y > 0.0f || f.z > 0.0f) {
                    // pdf of the scattering technique in direction ld
                    float mp;
                    if (M == NR_MAT_LAMBERTIAN) mp = cosine_pdf_value(uvw, ld);
                    else if (M == NR_MAT_PBR) mp = pbr_pdf_value(uvw, view, alpha, sp, ld);
                    else mp = 1.0f / NR_PI4;
                    if (S.first_light >= 0) mp = 0.5f * mp + 0.5f * light_pdf(S.objects[S.first_light], p, ld);
                    f3 c = atten * f * sky_sample<true>(S.sky, ld) * (power_heuristic(ps, mp) / ps);
                    if (c.x > 0.0f || c.y > 0.0f || c.z > 0.0f) {
                        P.sho[NR_SH_STRIDE * (size_t)path] = make_float4(p.x, p.y, p.z, 0.0f);
                        P.shd[NR_SH_STRIDE * (size_t)path] = make_float4(ld.x, ld.y, ld.z, 0.0f);
                        P.shc[NR_SH_STRIDE * (size_t)path] = make_float4(c.x, c.y, c.z, 0.0f);
                        shadow = true;
                    }
                }
            }
        }
        if (!(pdf <= 0.0f)) {
            f3 brdf = mat_eval<M, FAST>(S, m, nn, sn, uv, normalize(-din), dir);
            atten = atten * divs(brdf, pdf);
            nd = dir;
            alive = true;
            sc_pdf = pdf;
        }
    }
    if (bounce == 0) {                                                      // trace.cu:67-70 (zero when
        f3 dv = s3(0.0f), nv = s3(0.0f);                                    // the path ends here)
        if (alive) { dv = dv + atten; nv = nv + nn; }
        P.diff[4 * (size_t)path] = make_float4(dv.x, dv.y, dv.z, 0.0f);
        // the DEPTH AOV rides in .w: distance from the camera to the first hit
        P.nrm[4 * (size_t)path] = make_float4(nv.x, nv.y, nv.z, vlength(p - mk3(o4.x, o4.y, o4.z)));
    }
    if (alive) {
        if (A.do_rr && bounce >= A.rr_start && bounce + 1 < A.depth) {      // trace.cu:72-77
            float survival = fmaxf(atten.x, fmaxf(atten.y, atten.z));
            survival = fminf(survival, A.rr_survival);
            if (g.uniform() > survival) alive = false;
            else atten = atten * s3(1.0f / survival);
        }
    }
    P.L[4 * (size_t)path] = make_float4(L.x, L.y, L.z, __uint_as_float(g.draw));
    if (alive && bounce + 1 < A.depth) {
        P.atten[4 * (size_t)path] = make_float4(atten.x, atten.y, atten.z, FAST ? sc_pdf : 0.0f);
        P.o[2 * (size_t)path] = make_float4(p.x, p.y, p.z, __int_as_float((int)(bounce + 1)));
        P.d[2 * (size_t)path] = make_float4(nd.x, nd.y, nd.z, d4.w);
        cont = true;
    } else {
        P.finished[path] = 1;
    }
}

template <int M, bool FAST>
__global__ void __launch_bounds__(256) k_shade(const GScene* __restrict__ Sp, PathPtrs P, TraceArgs A) {
    const GScene& S = *Sp;
    const int n = P.counts[CNT_MAT + M];
    const int out = 1 - A.in;
    const int stride = gridDim.x * blockDim.x;
    constexpr bool NEE_M = FAST && (M == NR_MAT_LAMBERTIAN || M == NR_MAT_PBR || M == NR_MAT_ISOTROPIC);
    __shared__ float smarg[NEE_M ? SKY_MARG_LDS : 1];
    const float* marg = S.sky.marg;
    if (NEE_M && marg && S.sky.H < SKY_MARG_LDS && (int)(blockIdx.x * blockDim.x) < n) {
        for (int k = threadIdx.x; k <= S.sky.H; k += blockDim.x) smarg[k] = marg[k];
        __syncthreads();
        marg = smarg;
    }
    for (int base = blockIdx.x * blockDim.x; base < n; base += stride) {     // workgroup-uniform trip count
    int i = base + threadIdx.x;
    bool cont = false, shadow = false;
    int path = -1;
    if (i < n) {
        path = P.qmat[(size_t)M * P.qcap + i];
        shade_one<M, FAST>(S, P, A, path, marg, cont, shadow);
    }
    int slot = block_append<256>(cont, &P.counts[CNT_ACT + out]);
    if (cont) P.qact[(size_t)out * P.qcap + slot] = path;
    if (NEE_M) {
        int ss = block_append<256>(shadow, &P.counts[CNT_SHD]);
        if (shadow) P.qshd[ss] = path;
    }
    }
}

// Shadow rays of next-event estimation (fast mode): a ray that reaches the sky
// (the visibility of scene_closest) adds its contribution to the path's
// radiance.  Runs after every k_shade of the iteration and before k_regen.
template <int TY>
__device__ __forceinline__ void shadow_one(const GScene& S, const PathPtrs& P, int path, Stack& st) {
    float4 o4 = P.sho[NR_SH_STRIDE * (size_t)path], d4 = P.shd[NR_SH_STRIDE * (size_t)path];
    bool of = false;
    bool blocked = scene_occluded<TY, NR_SHADOW_LAY>(S, mkray(mk3(o4.x, o4.y, o4.z), mk3(d4.x, d4.y, d4.z)), st, of);
    if (of) atomicOr(P.err, 1);
    if (!blocked && !of) {
        float4 c = P.shc[NR_SH_STRIDE * (size_t)path], l = P.L[4 * (size_t)path];
        P.L[4 * (size_t)path] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, l.w);
    }
}

template <int TY>
__global__ void __launch_bounds__(128)
__attribute__((amdgpu_waves_per_eu(TY == NR_TY_ALL ? (NR_TRAV_OCC > 0 ? NR_TRAV_OCC : 1) : NR_SHADOW_WAVES, 8)))
k_shadow(const GScene* __restrict__ Sp, PathPtrs P, StackCfg SC) {
    extern __shared__ int2 smem[];
    const GScene& S = *Sp;
    Stack st = make_stack(smem, S, SC);
    const int n = P.counts[CNT_SHD];
    if ((int)(blockIdx.x * blockDim.x) >= n) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) shadow_one<TY>(S, P, P.qshd[i], st);
}

// Per pixel, scanning all pixels in tile order so the new primary rays of a
// wave are coherent: fold the ended samples strictly in sample order (sample
// nf lives in slot nf % slots; a sample that ends before its predecessor
// waits), then start the next samples in the slots that became idle.  The
// occupied slots (samples started and not yet folded, ns - nf) are counted:
// that number never grows, so the host sizes later traversal grids from it.
// Folds the ended sample nf of path-array pixel i (in path slot p) into the
// running means (accumulate_samples_kernel, data.cu:28-40) and frees the slot.
__device__ __forceinline__ void fold_sample(const PathPtrs& P, const TraceArgs& A, uint32_t i, uint32_t p, uint32_t nf) {
    float fn = (float)(A.fast ? nf + 1 : nf);                   // fast mode counts pass 0 as the first
    float a = (fn - 1.0f) / fn, b = 1.0f / fn;
    const float4 lp = P.L[4 * (size_t)p];
    fold(P.acc + i, make_float4(lp.x, lp.y, lp.z, 0.0f), a, b);
    fold(P.acc + P.cap + i, P.diff[4 * (size_t)p], a, b);
    fold(P.acc + 2 * (size_t)P.cap + i, P.nrm[4 * (size_t)p], a, b);
    P.finished[p] = 2;
}

constexpr int REGEN_BLOCK = 512;     // 1.2% faster renders than 1024 or 256 on the 1080p workload
__global__ void __launch_bounds__(REGEN_BLOCK) k_regen(PathPtrs P, TraceArgs A) {
    uint32_t t = A.t0 + blockIdx.x * blockDim.x + threadIdx.x, v = 0, x = 0, y = 0;
    uint32_t n_start = 0, occ = 0, i = 0, ns0 = 0;
    bool pending = false;
    if (t < A.t1 && tile_pixel(A, t, v, x, y)) {
        const GView& V = A.views[v];
        i = V.pix0 + y * V.cam.W + x;
        const uint32_t end = pixel_end(A, x, y);
        uint32_t nf = P.nf[i], ns = P.ns[i];
#pragma unroll
        for (int k = 0; k < NR_SLOTS_MAX; k++) {
            const uint32_t p = i + (nf % P.slots) * P.cap;
            if (nf >= ns || P.finished[p] != 1) break;
            fold_sample(P, A, i, p, nf);
            nf++;
        }
        ns0 = ns;
        for (int k = 0; k < NR_SLOTS_MAX; k++) {
            if (ns >= end || ns - nf >= P.slots) break;         // slot ns % slots is idle when ns - nf < slots
            start_path(V, P, x, y, ns, A);
            n_start = k + 1;
            ns++;
        }
        P.nf[i] = nf;
        P.ns[i] = ns;
        occ = ns - nf;
        pending = ns < end;
    }
    // queue entries: the path of sample ns0 + k is pixel i's slot (ns0 + k) % slots
    block_append_ranks<REGEN_BLOCK>(n_start, &P.counts[CNT_ACT + (1 - A.in)], occ, &P.counts[CNT_OCC],
                                    [&](int k, int pos) {
        P.qact[(size_t)(1 - A.in) * P.qcap + pos] = (int32_t)(i + ((ns0 + (uint32_t)k) % P.slots) * P.cap);
    });
    // pixels that still have samples to start (tail mode waits until there are none)
    const int pend = __syncthreads_count(pending);
    if (threadIdx.x == 0 && pend) atomicAdd(&P.counts[CNT_PEND], pend);
}

// Tail mode.  Once few paths are live, a wavefront iteration costs about the
// traversal time of its slowest ray plus four launches, however few rays it
// holds, and the last paths drain over many such iterations.  When every
// pixel's samples have all been started and few paths are live, k_tail gives
// each live path one thread that runs it to its end, segment by segment, with
// the per-path steps the wavefront kernels run (extend_one, shade_one,
// shadow_one) in their order; the k_regen that follows folds every pixel's
// samples in sample order, so the AOVs are the same bit for bit.
template <bool FAST>
__device__ __forceinline__ void shade_dispatch(const GScene& S, const PathPtrs& P, const TraceArgs& A, int m, int path,
                                               bool& cont, bool& shadow) {
    const float* marg = S.sky.marg;
    switch (m) {
        case NR_MAT_LAMBERTIAN: shade_one<NR_MAT_LAMBERTIAN, FAST>(S, P, A, path, marg, cont, shadow); break;
        case NR_MAT_PBR: shade_one<NR_MAT_PBR, FAST>(S, P, A, path, marg, cont, shadow); break;
        case NR_MAT_DIELECTRIC: shade_one<NR_MAT_DIELECTRIC, FAST>(S, P, A, path, marg, cont, shadow); break;
        case NR_MAT_EMISSIVE: shade_one<NR_MAT_EMISSIVE, FAST>(S, P, A, path, marg, cont, shadow); break;
        default: shade_one<NR_MAT_ISOTROPIC, FAST>(S, P, A, path, marg, cont, shadow); break;
    }
}

template <int TY, bool FAST>
__global__ void __launch_bounds__(128) k_tail(const GScene* __restrict__ Sp, PathPtrs P, TraceArgs A, StackCfg SC) {
    extern __shared__ int2 smem[];
    const GScene& S = *Sp;
    Stack st = make_stack(smem, S, SC);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.counts[CNT_ACT + A.in]) return;
    const int p = P.qact[(size_t)A.in * P.qcap + i];
    while (true) {
        const int m = extend_one<TY, FAST>(S, P, p, st);
        if (m == 0) break;                                  // the sky ended the path
        bool cont = false, shadow = false;
        shade_dispatch<FAST>(S, P, A, m, p, cont, shadow);
        if (shadow) shadow_one<TY>(S, P, p, st);
        if (!cont) break;                                   // shade_one marked it ended
    }
}

// ray_depth 0: every pass adds a zero sample (trace.cu:63 runs no bounce).  Pass s
// is folded as sample number fs; with per-tile limits (tile_end non-null, a view
// W pixels wide) a pixel whose tile ends at or before pass s keeps its AOVs.
__global__ void __launch_bounds__(256) k_fold_zero(float* beauty, float* diffuse, float* normal, float* depth, uint32_t n,
                                                   uint32_t s, uint32_t fs, uint32_t W, const uint32_t* __restrict__ tile_end,
                                                   uint32_t tile_shift, uint32_t tile_cols) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (tile_end && tile_end[((i / W) >> tile_shift) * tile_cols + ((i % W) >> tile_shift)] <= s) return;
    float fn = (float)fs;
    float a = (fn - 1.0f) / fn, b = 1.0f / fn;
    float* layers[3] = {beauty, diffuse, normal};
    for (float* l : layers)
        for (int c = 0; c < 3; c++) l[(size_t)c * n + i] = l[(size_t)c * n + i] * a + 0.0f * b;
    depth[i] = depth[i] * a + 0.0f * b;
}

struct PostArgs {
    float ratio[3];
    float tint;
    int32_t wb, tm, tm_method, gamma, fix_reinhard;
    float white_point, alpha;
};

__device__ __forceinline__ float clamp01(float x) { return fminf(1.0f, fmaxf(0.0f, x)); }

// process_stream_kernel (data/src/stream.cu:146-159)
__global__ void __launch_bounds__(256) k_post(const float* __restrict__ img, uint32_t HW, PostArgs A,
                                              uint8_t* __restrict__ out8, float* __restrict__ outf) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= HW) return;
    f3 c = mk3(img[i], img[HW + i], img[2 * (size_t)HW + i]);
    if (A.wb) {
        c = c * ld3(A.ratio);
        c = mk3(c.x * (1.0f + A.tint * 0.5f), c.y * (1.0f - A.tint), c.z * (1.0f + A.tint * 0.5f));
    }
    if (A.tm) {
        f3 col = c;
        if (A.tm_method == 0) {
            if (A.fix_reinhard) col = divv(col, adds(col, 1.0f));
        } else if (A.tm_method == 1) {
            float inv_w2 = 1.0f / (A.white_point * A.white_point);
            f3 num = col * (s3(1.0f) + col * inv_w2);
            col = divv(num, s3(1.0f) + col);
        } else {
            const float a = 2.51f, b = 0.03f, cc = 2.43f, dd = 0.59f, e = 0.14f;
            f3 num = col * adds(col * a, b);
            f3 den = adds(col * adds(col * cc, dd), e);
            col = mk3(clamp01(num.x / den.x), clamp01(num.y / den.y), clamp01(num.z / den.z));
        }
        c = c * (1.0f - A.alpha) + col * A.alpha;
        c = mk3(clamp01(c.x), clamp01(c.y), clamp01(c.z));
    }
    if (A.gamma) c = mk3(c.x > 0.0f ? sqrtf(c.x) : 0.0f, c.y > 0.0f ? sqrtf(c.y) : 0.0f, c.z > 0.0f ? sqrtf(c.z) : 0.0f);
    float v[3] = {c.x, c.y, c.z};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float x = v[k];
        if (x != x) x = 0.0f;
        x = fmaxf(0.0f, fminf(1.0f, x));
        if (outf) outf[k * (size_t)HW + i] = x;
        if (out8) out8[k * (size_t)HW + i] = (uint8_t)(x * 255.0f);
    }
}

// ---------------------------------------------------------------------------
// host math for scene lowering (same operation order as the device side and
// the reference: core/include/core/matrix.h, transform.h, camera.h:124-136)
// ---------------------------------------------------------------------------
namespace host {
struct M3 { float m[9]; };
static M3 ident() { M3 r{}; r.m[0] = r.m[4] = r.m[8] = 1.0f; return r; }
static M3 mul(const M3& a, const M3& b) {
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            float acc = 0.0f;
            for (int k = 0; k < 3; k++) acc += a.m[3 * i + k] * b.m[3 * k + j];
            r.m[3 * i + j] = acc;
        }
    return r;
}
static M3 rx(float a) { float c = cosf(a), s = sinf(a); M3 r = ident(); r.m[4] = c; r.m[5] = -s; r.m[7] = s; r.m[8] = c; return r; }
static M3 ry(float a) { float c = cosf(a), s = sinf(a); M3 r = ident(); r.m[0] = c; r.m[2] = s; r.m[6] = -s; r.m[8] = c; return r; }
static M3 rz(float a) { float c = cosf(a), s = sinf(a); M3 r = ident(); r.m[0] = c; r.m[1] = -s; r.m[3] = s; r.m[4] = c; return r; }
static float deg2rad(float d) { return (float)((double)d * 0.01745329); }
static M3 euler(float x, float y, float z) { return mul(mul(rz(z), ry(y)), rx(x)); }

static GXform xform(const float* p, const float* rot_deg, const float* sc, bool position_only) {
    GXform x{};
    M3 R = position_only ? euler(0.0f, 0.0f, 0.0f) : euler(deg2rad(rot_deg[0]), deg2rad(rot_deg[1]), deg2rad(rot_deg[2]));
    std::memcpy(x.R, R.m, sizeof(x.R));
    for (int k = 0; k < 3; k++) {
        x.pos[k] = p[k];
        x.scale[k] = position_only ? 1.0f : sc[k];
        x.inv_s[k] = position_only ? 1.0f : 1.0f / ((sc[k] + 1e-8f) + 1e-8f);
    }
    return x;
}

static nrb::Box transformed(const nrb::Box& b, const GXform& x) {               // aabb.h:73-96
    float mn[3] = {1e32f, 1e32f, 1e32f}, mx[3] = {-1e32f, -1e32f, -1e32f};
    for (int c = 0; c < 8; c++) {
        f3 p = mk3((c & 1) ? b.hi[0] : b.lo[0], (c & 2) ? b.hi[1] : b.lo[1], (c & 4) ? b.hi[2] : b.lo[2]);
        f3 w = mv(x.R, p * ld3(x.scale)) + ld3(x.pos);
        mn[0] = fminf(mn[0], w.x); mn[1] = fminf(mn[1], w.y); mn[2] = fminf(mn[2], w.z);
        mx[0] = fmaxf(mx[0], w.x); mx[1] = fmaxf(mx[1], w.y); mx[2] = fmaxf(mx[2], w.z);
    }
    nrb::Box r = nrb::box_points(mn, mx);
    for (int k = 0; k < 3; k++)                                                  // AABB::buffer(EPS)
        if (r.hi[k] - r.lo[k] < 1e-4f) { float pad = 1e-4f / 2.0f; r.lo[k] -= pad; r.hi[k] += pad; }
    return r;
}

// SampleGenerator::generate_n_rooks (camera.h:45-61); the reference's host
// compiler evaluates the Vector2 constructor arguments right to left, so the
// second coordinate's jitter is drawn first.
static std::vector<float> n_rooks(uint32_t n, uint32_t seed) {
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    std::mt19937 rng(seed);
    std::shuffle(perm.begin(), perm.end(), rng);
    std::uniform_real_distribution<float> jitter(0.0f, 1.0f);
    std::vector<float> out(2 * (size_t)n);
    for (uint32_t i = 0; i < n; i++) {
        float jy = jitter(rng);
        float jx = jitter(rng);
        out[2 * i] = (i + jx) / (float)n;
        out[2 * i + 1] = (perm[i] + jy) / (float)n;
    }
    return out;
}
}  // namespace host

// ---------------------------------------------------------------------------
// scratch of nr_trace_closest_device_noalloc
// ---------------------------------------------------------------------------
// One file-scope array per GPU the library runs on, held until the process
// ends: bytes [0, 256) hold the counters k_trace_closest writes (the overflow
// flag at CNT_ERR), the rest traversal-stack spill columns for one chunk of rays.
// Host code never uses the array's name as an address (on the host it is only a
// stand-in for the symbol): nr_create looks up each device's copy with
// hipGetSymbolAddress once, with that device current, and every launch takes the
// pointer from that table.
constexpr size_t NR_NOALLOC_SCRATCH_BYTES = (size_t)16 << 20;
constexpr size_t NR_NOALLOC_COUNTER_BYTES = 256;
static __device__ __align__(256) int2 scratch[NR_NOALLOC_SCRATCH_BYTES / sizeof(int2)];
constexpr int NR_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[NR_MAX_DEVICES] = {};

// The device copy of `scratch` on the current device `device` (looked up once).
static int scratch_of_device(int device, void** out) {
    if (device < 0 || device >= NR_MAX_DEVICES) return fail(NR_ERR_LIMIT, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) return fail(NR_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return NR_OK;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct View {
    bool valid = false;
    nr_camera_desc desc{};
    uint32_t seed = 0;
    GCamera cam{};
    float* d_pattern = nullptr;
    uint32_t H = 0, W = 0;
    float* layer[9] = {nullptr};
    // interactive mode
    std::mutex layers_mu;                 // layer writes vs. readbacks (nr_readback)
    std::mutex rb_mu;                     // one readback of this view at a time
    uint8_t* d_sel_base = nullptr;        // selection mask before the outline pass (H x W)
    uint8_t* d_mask = nullptr;            // nr_view_mask_buffer (H x W)
    size_t mask_cap = 0;
    uint8_t* d_rb_u8 = nullptr;           // readback image (3 x H x W)
    float* d_rb_den = nullptr;            // denoised layer (3 x H x W)
    uint8_t* h_rb = nullptr;              // pinned staging (3 x H x W)
    size_t sel_cap = 0, rb_cap = 0, den_cap = 0;
};

// planes of an AOV layer (render_layers.h:178-188: SHADOW and DEPTH have one)
static inline int layer_channels(int L) { return (L == NR_LAYER_DEPTH || L == NR_LAYER_SHADOW) ? 1 : 3; }

struct nr_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    void* scratch = nullptr;            // this device's copy of `scratch` (nr_trace_closest_device_noalloc)
    bool has_scene = false;
    GScene scene{};
    GScene* d_scene = nullptr;          // device copy of `scene`
    std::vector<void*> scene_mem;
    bool mat_present[NR_MAT_TYPES] = {false};
    int obj_types = NR_TY_ALL;      // object types of the scene (NR_TY_* bits): picks the traversal kernel
    int build_mode = NR_BUILD_AUTO;
    // host copies kept for introspection and live edits
    std::vector<GImage> h_imgs;
    std::vector<GMaterial> h_mats;
    std::vector<GMesh> h_meshes;
    std::vector<int32_t> h_mesh_nodes;      // child-pair nodes per mesh
    std::vector<nrb::Box> h_mesh_local;     // object-space box per mesh
    std::vector<GObject> reg_objects;       // registration order
    std::vector<nrb::Box> reg_local;        // local (object-space) boxes
    int max_mesh_depth = 0;
    View views[NR_MAX_VIEWS];
    // path state
    PathPtrs P{};
    std::vector<void*> path_mem;
    uint32_t path_cap = 0;
    int32_t* d_err = nullptr;           // counters of nr_trace_closest (CNT_N ints)
    GView* d_views = nullptr;           // view table of the current render call (NR_MAX_VIEWS)
    uint32_t* d_tile_end = nullptr;     // per-tile sample limits of nr_render_tiles
    size_t tile_end_cap = 0;
    uint32_t q_cap = 0;                 // padded pixels the path queues hold
    // persistent launch shapes
    int n_cu = 0;
    int shade_grid = 0;                 // workgroups of 256 for k_shade (grid-stride)
    int2* spill = nullptr;              // traversal stack levels beyond LDS
    size_t spill_elems = 0;
    int32_t* h_live = nullptr;          // pinned: occupied slots and pending pixels of the last two iterations, per group
    hipEvent_t ev_count[2] = {nullptr, nullptr};
    // stream groups: disjoint bands of pixels, each a complete wavefront pipeline on
    // its own stream, so one group's traversal tail overlaps another's work
    int render_streams = NR_RENDER_STREAMS;
    int slots = NR_SLOTS_DEFAULT;       // path slots per pixel (nr_set_tuning)
    int tail_paths = NR_TAIL_PATHS;     // tail mode once a group's live-path bound is at most this (nr_set_tuning)
    int graph_iters = 0;                // streaming iterations per captured HIP graph (0: host-launched; nr_set_tuning)
    hipStream_t gstream[NR_MAX_GROUPS] = {};
    hipEvent_t gev[NR_MAX_GROUPS][2] = {};
    hipEvent_t gev_done[NR_MAX_GROUPS] = {};
    // per stream group: the last captured streaming-iteration graph and the launch
    // arguments it was captured with (reused while they do not change)
    hipGraphExec_t gexec[NR_MAX_GROUPS] = {};
    std::vector<uint8_t> gkey[NR_MAX_GROUPS];
    hipEvent_t ev_packed = nullptr;
    int64_t last_iterations = 0;
    // timing
    bool timing = false;
    int trace_layout = 0;       // mesh-tree layout of nr_trace_closest (NR_LAY_*)
    std::vector<hipEvent_t> ev;             // per-launch marks
    hipEvent_t ev_begin = nullptr, ev_end = nullptr;
    float last_ms[8] = {0};
    int n_last = 0;
    // readback (nr_readback): its own stream; ev_layers marks the last enqueued
    // layer write, ev_rb_done the last enqueued readback of the layers
    hipStream_t rb_stream = nullptr;
    hipEvent_t ev_layers = nullptr, ev_rb_done = nullptr;
};

// Enqueue-side guard of every call that rewrites a view's layers: the writes
// wait for readbacks in flight and are published to later readbacks.
struct LayerWrite {
    nr_ctx* c;
    std::lock_guard<std::mutex> lk;
    LayerWrite(nr_ctx* c_, View& v) : c(c_), lk(v.layers_mu) { (void)hipStreamWaitEvent(c->stream, c->ev_rb_done, 0); }
    ~LayerWrite() { (void)hipEventRecord(c->ev_layers, c->stream); }
};

static void free_list(std::vector<void*>& v) {
    for (void* p : v) if (p) (void)hipFree(p);
    v.clear();
}

template <typename T>
static int upload(nr_ctx* c, const T* host, size_t n, T** dev) {
    *dev = nullptr;
    if (n == 0) return NR_OK;
    HIPCK(hipMalloc((void**)dev, n * sizeof(T)));
    c->scene_mem.push_back(*dev);
    HIPCK(hipMemcpyAsync(*dev, host, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return NR_OK;
}

// Importance-sampling tables of an equirect HDRI for fast mode.  The sky
// lookup filters bilinearly, so the radiance anywhere inside texel (i, j)
// mixes its 3 x 3 neighbourhood (u wraps, v clamps): the texel weighs the
// largest channel mean (negative and non-finite values count as 0) of that
// neighbourhood times sin(theta) at the row centre, theta = pi (j + 0.5) / H.
// The density then bounds the radiance from above everywhere, including the
// half-texel rim a bright feature bleeds into, which would otherwise be found
// only by rare scattered rays.  marg is the CDF over rows (H + 1 entries, sums
// in double, ends exactly 0 and 1); alias holds per texel a float4 (keep
// probability, alias texel as int bits, the texel's probability within its
// row, 0): a Walker alias table of each row built with Vose's method (columns
// whose scaled weight is below 1 take the excess of columns above 1, in the
// order of two work stacks).  Returns false for a map with no positive texel
// (no tables, no sky sampling).
static bool sky_cdfs(const float* data, int H, int W, std::vector<float>& marg, std::vector<float>& alias) {
    const size_t hw = (size_t)H * W;
    marg.assign((size_t)H + 1, 0.0f);
    alias.assign(4 * hw, 0.0f);
    std::vector<double> mean(hw), rowmax(hw);
    for (size_t k = 0; k < hw; k++) {
        double v = 0.0;
        for (int ch = 0; ch < 3; ch++) {
            const double x = (double)data[ch * hw + k];
            if (std::isfinite(x) && x > 0.0) v += x;
        }
        mean[k] = v / 3.0;
    }
    for (int j = 0; j < H; j++)                                  // max over i - 1 .. i + 1 (wrapping)
        for (int i = 0; i < W; i++) {
            const double* r = mean.data() + (size_t)j * W;
            rowmax[(size_t)j * W + i] = std::max(r[i], std::max(r[(i + W - 1) % W], r[(i + 1) % W]));
        }
    std::vector<double> row_sum(H), w((size_t)W), q((size_t)W);
    std::vector<int> small, large;
    double total = 0.0;
    for (int j = 0; j < H; j++) {
        const double st = std::sin(M_PI * ((double)j + 0.5) / (double)H);
        const double* a = rowmax.data() + (size_t)std::max(j - 1, 0) * W;   // max over j - 1 .. j + 1 (clamped)
        const double* b = rowmax.data() + (size_t)j * W;
        const double* c = rowmax.data() + (size_t)std::min(j + 1, H - 1) * W;
        double sum = 0.0;
        for (int i = 0; i < W; i++) {
            w[i] = std::max(a[i], std::max(b[i], c[i]));
            sum += w[i];
        }
        float* ar = alias.data() + 4 * (size_t)j * W;
        small.clear();
        large.clear();
        for (int i = 0; i < W; i++) {
            q[i] = sum > 0.0 ? w[i] * (double)W / sum : 1.0;
            ar[4 * i + 2] = sum > 0.0 ? (float)(w[i] / sum) : 1.0f / (float)W;
            (q[i] < 1.0 ? small : large).push_back(i);
        }
        while (!small.empty() && !large.empty()) {
            const int sm = small.back(), lg = large.back();
            small.pop_back();
            large.pop_back();
            ar[4 * sm] = (float)q[sm];
            int32_t al = lg;
            std::memcpy(&ar[4 * sm + 1], &al, sizeof(al));
            q[lg] = (q[lg] + q[sm]) - 1.0;
            (q[lg] < 1.0 ? small : large).push_back(lg);
        }
        for (int rest : large) { ar[4 * rest] = 1.0f; std::memcpy(&ar[4 * rest + 1], &rest, sizeof(rest)); }
        for (int rest : small) { ar[4 * rest] = 1.0f; std::memcpy(&ar[4 * rest + 1], &rest, sizeof(rest)); }
        row_sum[j] = sum * st;
        total += row_sum[j];
    }
    if (!(total > 0.0) || !std::isfinite(total)) return false;
    double acc = 0.0;
    for (int j = 0; j < H; j++) {
        marg[j] = (float)(acc / total);
        acc += row_sum[j];
    }
    marg[H] = 1.0f;
    return true;
}

// Stack levels a mesh traversal of a tree of the given depth can hold: depth for
// the binary near-first order (one far child per level), 3 per wide node and
// ceil(depth / 2) + 1 wide levels for the 4-wide order.
static int mesh_stack_levels(int depth) { return std::max(depth, 3 * ((depth + 1) / 2 + 1)); }

static size_t trav_lds_bytes() { return (size_t)LDS_STACK_LEVELS * 64 * sizeof(int2) * (128 / 64); }

// Spill columns for the traversal stack levels beyond LDS, one per ray of a
// launch of n rays (level-major, so a wave's spills are contiguous).
static int stack_spill(nr_ctx* c, uint32_t n, StackCfg& SC) {
    size_t lv = c->scene.stack_depth > LDS_STACK_LEVELS ? (size_t)(c->scene.stack_depth - LDS_STACK_LEVELS) : 0;
    size_t need = lv * (size_t)n;
    if (need > c->spill_elems) {
        if (c->spill) HIPCK(hipFree(c->spill));
        c->spill = nullptr;
        c->spill_elems = 0;
        HIPCK(hipMalloc((void**)&c->spill, need * sizeof(int2)));
        c->spill_elems = need;
    }
    SC.spill = c->spill;
    SC.stride = n;
    return NR_OK;
}

// Mesh tree on the GPU (nr_bvh_gpu.h): the same node array as the host build.
// Node, triangle and index arrays are left in the scene allocations.
static int build_mesh_on_device(nr_ctx* c, const nr_mesh_desc& md, GMesh& gm, GNode*& dn, GTri*& dt, uint32_t*& di,
                                int32_t& n_nodes) {
    const int32_t nt = md.n_triangles;
    float* dv = nullptr;
    uint32_t* dtri = nullptr;
    float* bx = nullptr;
    int32_t* dord = nullptr;
    auto release = [&] {
        if (dv) (void)hipFree(dv);
        if (dtri) (void)hipFree(dtri);
        if (bx) (void)hipFree(bx);
        if (dord) (void)hipFree(dord);
    };
    auto hip = [&](hipError_t e, const char* what) -> int {
        if (e == hipSuccess) return NR_OK;
        release();
        return fail(NR_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    if (int e = hip(hipMalloc((void**)&dv, sizeof(float) * 11 * (size_t)md.n_vertices), "hipMalloc vertices")) return e;
    if (int e = hip(hipMalloc((void**)&dtri, sizeof(uint32_t) * 3 * (size_t)nt), "hipMalloc triangles")) return e;
    if (int e = hip(hipMalloc((void**)&bx, sizeof(float) * 6 * (size_t)nt), "hipMalloc boxes")) return e;
    if (int e = hip(hipMalloc((void**)&dord, sizeof(int32_t) * (size_t)nt), "hipMalloc order")) return e;
    if (int e = hip(hipMemcpyAsync(dv, md.vertices, sizeof(float) * 11 * (size_t)md.n_vertices, hipMemcpyHostToDevice,
                                   c->stream), "copy vertices")) return e;
    if (int e = hip(hipMemcpyAsync(dtri, md.triangles, sizeof(uint32_t) * 3 * (size_t)nt, hipMemcpyHostToDevice,
                                   c->stream), "copy triangles")) return e;
    float* const boxes[6] = {bx, bx + nt, bx + 2 * (size_t)nt, bx + 3 * (size_t)nt, bx + 4 * (size_t)nt,
                             bx + 5 * (size_t)nt};
    nrg::triangle_boxes(c->stream, dv, dtri, nt, boxes);
    dn = nullptr; dt = nullptr; di = nullptr;
    if (int e = hip(hipMalloc((void**)&dn, sizeof(GNode) * (size_t)(nt - 1)), "hipMalloc nodes")) return e;
    c->scene_mem.push_back(dn);
    if (int e = hip(hipMalloc((void**)&dt, sizeof(GTri) * (size_t)nt), "hipMalloc tris")) return e;
    c->scene_mem.push_back(dt);
    if (int e = hip(hipMalloc((void**)&di, sizeof(uint32_t) * 3 * (size_t)nt), "hipMalloc idx")) return e;
    c->scene_mem.push_back(di);
    nrg::Result r;
    std::string err;
    const float* const cboxes[6] = {boxes[0], boxes[1], boxes[2], boxes[3], boxes[4], boxes[5]};
    int e = nrg::build(c->stream, nt, cboxes, dn, dord, r, err);
    if (e != NR_OK) {
        release();
        return fail(e, "device tree build: " + err);
    }
    nrg::gather_triangles(c->stream, dv, dtri, dord, nt, dt, di);
    if (int e2 = hip(hipGetLastError(), "gather triangles")) return e2;
    if (int e2 = hip(hipStreamSynchronize(c->stream), "device tree build")) return e2;
    release();
    gm.root_ref = r.root_ref;
    for (int a = 0; a < 3; a++) { gm.root_min[a] = r.root_lo[a]; gm.root_max[a] = r.root_hi[a]; }
    gm.depth = r.depth;
    n_nodes = nt - 1;
    return NR_OK;
}

// one k_shade per material type present in the scene
template <bool FAST>
static void launch_shade(nr_ctx* c, dim3 gS, hipStream_t st, const PathPtrs& P, const TraceArgs& A) {
    if (c->mat_present[NR_MAT_LAMBERTIAN]) hipLaunchKernelGGL((k_shade<NR_MAT_LAMBERTIAN, FAST>), gS, dim3(256), 0, st, c->d_scene, P, A);
    if (c->mat_present[NR_MAT_PBR]) hipLaunchKernelGGL((k_shade<NR_MAT_PBR, FAST>), gS, dim3(256), 0, st, c->d_scene, P, A);
    if (c->mat_present[NR_MAT_DIELECTRIC]) hipLaunchKernelGGL((k_shade<NR_MAT_DIELECTRIC, FAST>), gS, dim3(256), 0, st, c->d_scene, P, A);
    if (c->mat_present[NR_MAT_EMISSIVE]) hipLaunchKernelGGL((k_shade<NR_MAT_EMISSIVE, FAST>), gS, dim3(256), 0, st, c->d_scene, P, A);
    if (c->mat_present[NR_MAT_ISOTROPIC]) hipLaunchKernelGGL((k_shade<NR_MAT_ISOTROPIC, FAST>), gS, dim3(256), 0, st, c->d_scene, P, A);
}

// ---------------------------------------------------------------------------
// Scene tables.  nr_scene_upload assembles them all; the live edits
// (nr_object_set_transform, nr_scene_add_object, nr_scene_set_sky) rebuild
// only the tables an edit changes and re-publish the GScene record.
// ---------------------------------------------------------------------------

// Frees one allocation of the scene's list: a table that an edit replaced.
static void scene_release(nr_ctx* c, const void* p) {
    if (!p) return;
    auto it = std::find(c->scene_mem.begin(), c->scene_mem.end(), const_cast<void*>(p));
    if (it == c->scene_mem.end()) return;
    (void)hipFree(*it);
    c->scene_mem.erase(it);
}

static int check_image(const nr_image_desc& im) {
    if (im.C < 3 || im.H <= 0 || im.W <= 0 || !im.data) return fail(NR_ERR_INVALID, "image needs C >= 3 and data");
    return NR_OK;
}

static int check_material(const nr_material_desc& md, int n_images) {
    if (md.type < 1 || md.type > 5) return fail(NR_ERR_INVALID, "bad material type");
    for (int k = 0; k < 5; k++)
        if (md.tex[k].image >= n_images) return fail(NR_ERR_INVALID, "texture image index out of range");
    return NR_OK;
}

static GMaterial make_material(const nr_material_desc& md) {
    GMaterial g;
    std::memset(&g, 0, sizeof(g));
    g.type = md.type;
    for (int k = 0; k < 5; k++) {
        g.tex[k].image = md.tex[k].image;
        for (int a = 0; a < 3; a++) g.tex[k].color[a] = md.tex[k].color[a];
        g.tex[k].scale = md.tex[k].scale;
    }
    for (int a = 0; a < 3; a++) g.tint[a] = md.tint[a];
    g.ior = md.ior;
    g.brightness = md.brightness;
    return g;
}

static int check_mesh(const nr_mesh_desc& md) {
    if (md.n_vertices < 0 || md.n_triangles < 0) return fail(NR_ERR_INVALID, "bad mesh sizes");
    if ((md.n_vertices > 0 && !md.vertices) || (md.n_triangles > 0 && !md.triangles))
        return fail(NR_ERR_INVALID, "mesh arrays missing");
    for (int64_t k = 0; k < 3 * (int64_t)md.n_triangles; k++)
        if (md.triangles[k] >= (uint32_t)md.n_vertices) return fail(NR_ERR_INVALID, "triangle index out of range");
    return NR_OK;
}

// One mesh on the device: its tree (device builder for large meshes, else the
// host binned-SAH build, the same node array either way), triangles in leaf
// order, vertex attributes, and the 4-wide and quantised 4-wide node layouts.
// `local` receives the mesh's object-space box (mesh.h:218-237).
static int build_mesh(nr_ctx* c, const nr_mesh_desc& md, GMesh& gm, int32_t& n_nodes, nrb::Box& local) {
    const bool on_device = md.n_triangles > 1 &&
                           (c->build_mode == NR_BUILD_DEVICE ||
                            (c->build_mode == NR_BUILD_AUTO && md.n_triangles >= 4096));
    gm = GMesh{};
    n_nodes = 0;
    GNode* dn = nullptr;
    GTri* dt = nullptr;
    uint32_t* di = nullptr;
    bool built = false;
    if (on_device) {
        int e = build_mesh_on_device(c, md, gm, dn, dt, di, n_nodes);
        if (e == NR_OK) built = true;
        else if (e != NR_ERR_LIMIT) return e;        // a degenerate cluster the device path refuses: host build
    }
    if (!built) {
        std::vector<nrb::Box> tb(md.n_triangles);
        for (int t = 0; t < md.n_triangles; t++) {                             // mesh.h:320-336
            const float* a = md.vertices + 11 * (size_t)md.triangles[3 * t];
            const float* b = md.vertices + 11 * (size_t)md.triangles[3 * t + 1];
            const float* cc = md.vertices + 11 * (size_t)md.triangles[3 * t + 2];
            float mn[3], mx[3];
            for (int k = 0; k < 3; k++) {
                mn[k] = fminf(a[k], fminf(b[k], cc[k]));
                mx[k] = fmaxf(a[k], fmaxf(b[k], cc[k]));
            }
            tb[t] = nrb::box_points(mn, mx);
        }
        nrb::Tree T;
        nrb::Builder().build(tb, T);
        std::vector<GNode> nodes;
        nrb::to_pairs(T, nodes, gm.root_ref, gm.root_min, gm.root_max);
        gm.depth = T.depth;
        std::vector<GTri> tris(md.n_triangles);
        std::vector<uint32_t> idx(3 * (size_t)md.n_triangles);
        for (int s = 0; s < md.n_triangles; s++) {
            int src = T.order[s];
            GTri& g = tris[s];
            for (int k = 0; k < 3; k++) {
                uint32_t vi = md.triangles[3 * src + k];
                idx[3 * (size_t)s + k] = vi;
                float* dst = k == 0 ? g.v0 : (k == 1 ? g.v1 : g.v2);
                for (int a = 0; a < 3; a++) dst[a] = md.vertices[11 * (size_t)vi + a];
            }
            g.orig = src;
            g.pad0 = g.pad1 = 0;
        }
        int e = upload(c, nodes.data(), nodes.size(), &dn);
        if (!e) e = upload(c, tris.data(), tris.size(), &dt);
        if (!e) e = upload(c, idx.data(), idx.size(), &di);
        if (e) return e;
        n_nodes = (int32_t)nodes.size();
        HIPCK(hipStreamSynchronize(c->stream));          // the host arrays end with this scope
    }
    std::vector<GVertex> verts(md.n_vertices);
    float mn[3] = {1e32f, 1e32f, 1e32f}, mx[3] = {-1e32f, -1e32f, -1e32f};
    for (int v = 0; v < md.n_vertices; v++) {
        const float* f = md.vertices + 11 * (size_t)v;
        GVertex& g = verts[v];
        for (int a = 0; a < 3; a++) { g.normal[a] = f[3 + a]; g.tangent[a] = f[6 + a]; }
        g.uv[0] = f[9]; g.uv[1] = f[10];
        g.pad[0] = g.pad[1] = g.pad[2] = g.pad[3] = 0.0f;
        for (int a = 0; a < 3; a++) { mn[a] = fminf(mn[a], f[a]); mx[a] = fmaxf(mx[a], f[a]); }
    }
    local = md.n_vertices > 0 ? nrb::box_points(mn, mx) : nrb::box_none();     // mesh.h:218-237
    GVertex* dv;
    if (int e = upload(c, verts.data(), verts.size(), &dv)) return e;
    HIPCK(hipStreamSynchronize(c->stream));
    gm.nodes = dn; gm.tris = dt; gm.idx = di; gm.verts = dv;
    gm.wnodes = nullptr;
    gm.qnodes = nullptr;
    if (n_nodes > 0) {
        GWNode* dw = nullptr;
        HIPCK(hipMalloc((void**)&dw, sizeof(GWNode) * (size_t)n_nodes));
        c->scene_mem.push_back(dw);
        hipLaunchKernelGGL(k_wide_nodes, dim3((n_nodes + 255) / 256), dim3(256), 0, c->stream, dn, n_nodes, dw);
        HIPCK(hipGetLastError());
        gm.wnodes = dw;
        GQNode* dq = nullptr;
        HIPCK(hipMalloc((void**)&dq, sizeof(GQNode) * (size_t)n_nodes));
        c->scene_mem.push_back(dq);
        hipLaunchKernelGGL(k_quant_nodes, dim3((n_nodes + 255) / 256), dim3(256), 0, c->stream, dw, n_nodes, dq);
        HIPCK(hipGetLastError());
        gm.qnodes = dq;
    }
    return NR_OK;
}

static int check_object(const nr_object_desc& od, int n_materials, int n_meshes) {
    if (od.type < 1 || od.type > 6) return fail(NR_ERR_INVALID, "bad hittable type");
    int geom = (od.type == NR_HIT_CONSTANT_MEDIUM || od.type == NR_HIT_OBJECT_LIGHT) ? od.bound_type : od.type;
    if (geom < 1 || geom > 4) return fail(NR_ERR_INVALID, "bound object must be quad, sphere, cube or mesh");
    if (geom == NR_HIT_MESH && (od.mesh < 0 || od.mesh >= n_meshes)) return fail(NR_ERR_INVALID, "mesh index out of range");
    if (od.material < 0 || od.material >= n_materials) return fail(NR_ERR_INVALID, "material index out of range");
    return NR_OK;
}

// The record of the object registered as `id` (Hittable + its registry entry)
// and its object-space box `lb`; registers its material type and object type
// with the context's kernel selection.
static GObject make_object(nr_ctx* c, const nr_object_desc& od, int32_t id, nrb::Box& lb) {
    const int geom = (od.type == NR_HIT_CONSTANT_MEDIUM || od.type == NR_HIT_OBJECT_LIGHT) ? od.bound_type : od.type;
    GObject o;
    std::memset(&o, 0, sizeof(o));
    o.xf = host::xform(od.position, od.rotation, od.scale, geom == NR_HIT_SPHERE);
    o.type = od.type;
    o.bound_type = od.bound_type;
    o.material = od.material;
    o.object_id = id;
    o.mesh = geom == NR_HIT_MESH ? od.mesh : -1;
    o.radius = od.radius + 1e-8f;
    o.neg_inv_density = -1.0f / (od.density + 1e-8f);
    o.pad[0] = c->h_mats[od.material].type;
    c->mat_present[c->h_mats[od.material].type] = true;
    c->obj_types |= od.type == NR_HIT_QUAD ? NR_TY_QUAD : od.type == NR_HIT_SPHERE ? NR_TY_SPHERE
                  : od.type == NR_HIT_MESH ? NR_TY_MESH : NR_TY_OTHER;
    if (geom == NR_HIT_QUAD || geom == NR_HIT_CUBE) {
        float a[3] = {-0.5f, -0.5f, -0.5f}, b[3] = {0.5f, 0.5f, 0.5f};
        lb = nrb::box_points(a, b);
    } else if (geom == NR_HIT_SPHERE) {
        float a[3] = {-o.radius, -o.radius, -o.radius}, b[3] = {o.radius, o.radius, o.radius};
        lb = nrb::box_points(a, b);
    } else {
        lb = c->h_mesh_local[od.mesh];
    }
    return o;
}

// Top level (the registry's build_bvh, registry.h:216-234): the binned-SAH tree
// over the world boxes of the registered objects in registration order, and
// the object records in the tree's item order.  in_place: the object count is
// the one the tables were allocated for, so they are rewritten where they are.
static int upload_top(nr_ctx* c, GScene& S, bool in_place) {
    const int32_t n = (int32_t)c->reg_objects.size();
    std::vector<nrb::Box> wb(n);
    for (int32_t i = 0; i < n; i++) wb[i] = host::transformed(c->reg_local[i], c->reg_objects[i].xf);
    nrb::Tree T;
    nrb::Builder().build(wb, T);
    std::vector<GNode> tnodes;
    nrb::to_pairs(T, tnodes, S.root_ref, S.root_min, S.root_max);
    std::vector<GObject> objs(n);
    S.first_light = -1;
    for (int32_t k = 0; k < n; k++) {
        objs[k] = c->reg_objects[T.order[k]];
        if (S.first_light < 0 && objs[k].type == NR_HIT_OBJECT_LIGHT) S.first_light = k;   // registry.h:240-244
    }
    S.n_objects = n;
    S.stack_depth = T.depth + mesh_stack_levels(c->max_mesh_depth) + 4;
    if (in_place) {
        if (!tnodes.empty())
            HIPCK(hipMemcpyAsync((void*)S.nodes, tnodes.data(), sizeof(GNode) * tnodes.size(), hipMemcpyHostToDevice,
                                 c->stream));
        if (n > 0)
            HIPCK(hipMemcpyAsync((void*)S.objects, objs.data(), sizeof(GObject) * objs.size(), hipMemcpyHostToDevice,
                                 c->stream));
    } else {
        scene_release(c, S.nodes);
        scene_release(c, S.objects);
        GNode* dn = nullptr;
        GObject* dob = nullptr;
        int e = upload(c, tnodes.data(), tnodes.size(), &dn);
        if (!e) e = upload(c, objs.data(), objs.size(), &dob);
        S.nodes = dn;
        S.objects = dob;
        if (e) {
            (void)hipStreamSynchronize(c->stream);    // copies in flight read the host arrays
            return e;
        }
    }
    HIPCK(hipStreamSynchronize(c->stream));
    return NR_OK;
}

static int check_sky(const nr_sky_desc& sky) {
    if (sky.type < NR_SKY_NULL || sky.type > NR_SKY_HDRI) return fail(NR_ERR_INVALID, "bad sky type");
    if (sky.type == NR_SKY_HDRI && sky.data && (sky.C < 3 || sky.H <= 0 || sky.W <= 0))
        return fail(NR_ERR_INVALID, "HDRI needs C >= 3");
    return NR_OK;
}

// The skylight (skylight.h:71-122) and, for an HDRI, the map, its interleaved
// copy and the importance-sampling tables of fast mode.
static int upload_sky(nr_ctx* c, const nr_sky_desc& sky, GSky& g) {
    g = GSky{};
    g.type = sky.type;
    for (int a = 0; a < 3; a++) { g.start[a] = sky.start[a]; g.end[a] = sky.end[a]; }
    {
        host::M3 R = host::ry(host::deg2rad(sky.rotation));
        std::memcpy(g.rot, R.m, sizeof(g.rot));
    }
    g.intensity = sky.intensity;
    if (sky.type != NR_SKY_HDRI || !sky.data) return NR_OK;
    g.H = sky.H; g.W = sky.W;
    float* p;
    // only the first three planes are read (skylight.h:106-114)
    if (int e = upload(c, sky.data, (size_t)3 * sky.H * sky.W, &p)) return e;
    g.data = p;
    const size_t hw = (size_t)sky.H * sky.W;
    std::vector<float> rgbx(4 * hw, 0.0f), marg, alias;
    for (size_t k = 0; k < hw; k++)
        for (int ch = 0; ch < 3; ch++) rgbx[4 * k + ch] = sky.data[ch * hw + k];
    float* px = nullptr;
    int e = upload(c, rgbx.data(), rgbx.size(), &px);
    g.rgbx = px;
    if (!e && sky_cdfs(sky.data, sky.H, sky.W, marg, alias)) {
        float *pm = nullptr, *pa = nullptr;
        e = upload(c, marg.data(), marg.size(), &pm);
        if (!e) e = upload(c, alias.data(), alias.size(), &pa);
        g.marg = pm;
        g.alias = pa;
    }
    (void)hipStreamSynchronize(c->stream);               // the host tables end with this scope
    return e;
}

// Writes the GScene record the kernels read (allocated by the first upload).
static int publish_scene(nr_ctx* c, const GScene& S) {
    if (!c->d_scene) {
        GScene* ds;
        if (int e = upload(c, &S, 1, &ds)) return e;
        c->d_scene = ds;
    } else {
        HIPCK(hipMemcpyAsync(c->d_scene, &S, sizeof(GScene), hipMemcpyHostToDevice, c->stream));
    }
    HIPCK(hipStreamSynchronize(c->stream));
    c->scene = S;
    return NR_OK;
}

template <typename T>
static int replace_table(nr_ctx* c, const std::vector<T>& host_table, const T*& dev) {
    scene_release(c, dev);
    T* p = nullptr;
    if (int e = upload(c, host_table.data(), host_table.size(), &p)) return e;
    dev = p;
    return NR_OK;
}

extern "C" {

int nr_abi_version(void) { return NR_ABI_VERSION; }
const char* nr_last_error(void) { return g_err.c_str(); }

nr_ctx* nr_create(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) { g_err = "no HIP device visible"; return nullptr; }
    if (device < 0 || device >= n) { g_err = "device index out of range"; return nullptr; }
    if (hipSetDevice(device) != hipSuccess) { g_err = "hipSetDevice failed"; return nullptr; }
    auto* c = new nr_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        g_err = "hipStreamCreate failed";
        delete c;
        return nullptr;
    }
    if (hipEventCreate(&c->ev_begin) != hipSuccess || hipEventCreate(&c->ev_end) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_count[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_count[1], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_layers, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_rb_done, hipEventDisableTiming) != hipSuccess ||
        hipStreamCreateWithFlags(&c->rb_stream, hipStreamNonBlocking) != hipSuccess) {
        g_err = "hipEventCreate failed";
        (void)hipStreamDestroy(c->stream);
        delete c;
        return nullptr;
    }
    if (hipMalloc((void**)&c->d_err, sizeof(int32_t) * CNT_N) != hipSuccess) {
        g_err = "hipMalloc failed";
        (void)hipStreamDestroy(c->stream);
        delete c;
        return nullptr;
    }
    (void)hipMemsetAsync(c->d_err, 0, sizeof(int32_t) * CNT_N, c->stream);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        g_err = "device query failed";
        nr_destroy(c);
        return nullptr;
    }
    for (int g = 0; g < NR_MAX_GROUPS; g++) {
        bool ok = hipStreamCreateWithFlags(&c->gstream[g], hipStreamNonBlocking) == hipSuccess &&
                  hipEventCreateWithFlags(&c->gev[g][0], hipEventDisableTiming) == hipSuccess &&
                  hipEventCreateWithFlags(&c->gev[g][1], hipEventDisableTiming) == hipSuccess &&
                  hipEventCreateWithFlags(&c->gev_done[g], hipEventDisableTiming) == hipSuccess;
        if (!ok) { g_err = "stream group setup failed"; nr_destroy(c); return nullptr; }
    }
    if (hipEventCreateWithFlags(&c->ev_packed, hipEventDisableTiming) != hipSuccess) {
        g_err = "hipEventCreate failed";
        nr_destroy(c);
        return nullptr;
    }
    if (hipHostMalloc((void**)&c->h_live, sizeof(int32_t) * 4 * NR_MAX_GROUPS, hipHostMallocDefault) != hipSuccess) {
        g_err = "hipHostMalloc failed";
        nr_destroy(c);
        return nullptr;
    }
    c->n_cu = prop.multiProcessorCount;
    c->shade_grid = c->n_cu * 8;
    if (scratch_of_device(device, &c->scratch) != NR_OK) {
        const std::string msg = g_err;
        nr_destroy(c);
        g_err = msg;
        return nullptr;
    }
    return c;
}

void nr_destroy(nr_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto st : c->gstream) if (st) (void)hipStreamSynchronize(st);
    free_list(c->scene_mem);
    free_list(c->path_mem);
    if (c->d_views) (void)hipFree(c->d_views);
    if (c->d_tile_end) (void)hipFree(c->d_tile_end);
    if (c->rb_stream) (void)hipStreamSynchronize(c->rb_stream);
    for (auto& v : c->views) {
        if (v.d_pattern) (void)hipFree(v.d_pattern);
        for (auto* l : v.layer) if (l) (void)hipFree(l);
        if (v.d_sel_base) (void)hipFree(v.d_sel_base);
        if (v.d_mask) (void)hipFree(v.d_mask);
        if (v.d_rb_u8) (void)hipFree(v.d_rb_u8);
        if (v.d_rb_den) (void)hipFree(v.d_rb_den);
        if (v.h_rb) (void)hipHostFree(v.h_rb);
    }
    for (auto e : c->ev) (void)hipEventDestroy(e);
    if (c->ev_begin) (void)hipEventDestroy(c->ev_begin);
    for (auto e : c->ev_count) if (e) (void)hipEventDestroy(e);
    if (c->ev_end) (void)hipEventDestroy(c->ev_end);
    if (c->d_err) (void)hipFree(c->d_err);
    if (c->spill) (void)hipFree(c->spill);
    if (c->h_live) (void)hipHostFree(c->h_live);
    if (c->ev_layers) (void)hipEventDestroy(c->ev_layers);
    for (int g = 0; g < NR_MAX_GROUPS; g++) {
        if (c->gstream[g]) { (void)hipStreamSynchronize(c->gstream[g]); (void)hipStreamDestroy(c->gstream[g]); }
        if (c->gexec[g]) (void)hipGraphExecDestroy(c->gexec[g]);
        for (auto& ev : c->gev[g]) if (ev) (void)hipEventDestroy(ev);
        if (c->gev_done[g]) (void)hipEventDestroy(c->gev_done[g]);
    }
    if (c->ev_packed) (void)hipEventDestroy(c->ev_packed);
    if (c->ev_rb_done) (void)hipEventDestroy(c->ev_rb_done);
    if (c->rb_stream) (void)hipStreamDestroy(c->rb_stream);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

int nr_scene_upload(nr_ctx* c, const nr_scene_desc* d) {
    if (!c || !d) return fail(NR_ERR_INVALID, "null context or scene");
    HIPCK(hipSetDevice(c->device));
    if (d->n_materials <= 0) return fail(NR_ERR_INVALID, "scene needs at least the default material");
    if (d->n_objects < 0 || d->n_images < 0 || d->n_meshes < 0) return fail(NR_ERR_INVALID, "negative table size");
    for (int i = 0; i < d->n_images; i++)
        if (int e = check_image(d->images[i])) return e;
    for (int i = 0; i < d->n_materials; i++)
        if (int e = check_material(d->materials[i], d->n_images)) return e;
    for (int i = 0; i < d->n_meshes; i++)
        if (int e = check_mesh(d->meshes[i])) return e;
    for (int i = 0; i < d->n_objects; i++)
        if (int e = check_object(d->objects[i], d->n_materials, d->n_meshes)) return e;
    if (int e = check_sky(d->sky)) return e;
    HIPCK(hipStreamSynchronize(c->stream));
    free_list(c->scene_mem);
    c->d_scene = nullptr;
    c->has_scene = false;
    for (auto& b : c->mat_present) b = false;
    c->obj_types = 0;
    c->h_imgs.clear();
    c->h_mats.clear();
    c->h_meshes.clear();
    c->h_mesh_nodes.clear();
    c->h_mesh_local.clear();
    c->reg_objects.clear();
    c->reg_local.clear();
    c->max_mesh_depth = 0;
    GScene S{};
    for (int i = 0; i < d->n_images; i++) {
        const nr_image_desc& im = d->images[i];
        uint8_t* p = nullptr;
        if (int e = upload(c, im.data, (size_t)im.C * im.H * im.W, &p)) return e;
        c->h_imgs.push_back(GImage{p, im.C, im.H, im.W, 0});
    }
    for (int i = 0; i < d->n_materials; i++) c->h_mats.push_back(make_material(d->materials[i]));
    for (int mi = 0; mi < d->n_meshes; mi++) {
        GMesh gm;
        int32_t n_nodes;
        nrb::Box local;
        if (int e = build_mesh(c, d->meshes[mi], gm, n_nodes, local)) return e;
        c->h_meshes.push_back(gm);
        c->h_mesh_nodes.push_back(n_nodes);
        c->h_mesh_local.push_back(local);
        c->max_mesh_depth = std::max(c->max_mesh_depth, gm.depth);
    }
    // objects in registration order (hittables then lights, scene.h:31-39)
    for (int i = 0; i < d->n_objects; i++) {
        nrb::Box lb;
        c->reg_objects.push_back(make_object(c, d->objects[i], i, lb));
        c->reg_local.push_back(lb);
    }
    if (int e = replace_table(c, c->h_imgs, S.images)) return e;
    if (int e = replace_table(c, c->h_mats, S.materials)) return e;
    if (int e = replace_table(c, c->h_meshes, S.meshes)) return e;
    if (int e = upload_top(c, S, false)) return e;
    if (int e = upload_sky(c, d->sky, S.sky)) return e;
    // cube face transforms (shapes.h:192-223)
    const float fp[6][3] = {{0, 0, 0.5f}, {0, 0, -0.5f}, {0.5f, 0, 0}, {-0.5f, 0, 0}, {0, 0.5f, 0}, {0, -0.5f, 0}};
    const float fr[6][3] = {{90, 0, 0}, {-90, 0, 0}, {0, 0, -90}, {0, 0, 90}, {0, 0, 0}, {180, 0, 0}};
    const float one[3] = {1.0f, 1.0f, 1.0f};
    for (int i = 0; i < 6; i++) S.cube_faces[i] = host::xform(fp[i], fr[i], one, false);
    if (int e = publish_scene(c, S)) return e;
    c->has_scene = true;
    return NR_OK;
}

int nr_scene_add_object(nr_ctx* c, const nr_object_desc* obj, const nr_scene_desc* add) {
    if (!c || !obj) return fail(NR_ERR_INVALID, "null argument");
    if (!c->has_scene) return fail(NR_ERR_STATE, "no scene");
    const int ni = add ? add->n_images : 0, nm = add ? add->n_materials : 0, nme = add ? add->n_meshes : 0;
    if (add && add->n_objects != 0) return fail(NR_ERR_INVALID, "the additions list no objects");
    if (ni < 0 || nm < 0 || nme < 0) return fail(NR_ERR_INVALID, "negative table size");
    const int n_img = (int)c->h_imgs.size() + ni, n_mat = (int)c->h_mats.size() + nm,
              n_mesh = (int)c->h_meshes.size() + nme;
    for (int i = 0; i < ni; i++)
        if (int e = check_image(add->images[i])) return e;
    for (int i = 0; i < nm; i++)
        if (int e = check_material(add->materials[i], n_img)) return e;
    for (int i = 0; i < nme; i++)
        if (int e = check_mesh(add->meshes[i])) return e;
    if (int e = check_object(*obj, n_mat, n_mesh)) return e;
    HIPCK(hipSetDevice(c->device));
    HIPCK(hipStreamSynchronize(c->stream));
    // a device failure below leaves no usable scene (the next call must upload one)
    c->has_scene = false;
    GScene S = c->scene;
    if (ni > 0) {
        for (int i = 0; i < ni; i++) {
            const nr_image_desc& im = add->images[i];
            uint8_t* p = nullptr;
            if (int e = upload(c, im.data, (size_t)im.C * im.H * im.W, &p)) return e;
            c->h_imgs.push_back(GImage{p, im.C, im.H, im.W, 0});
        }
        if (int e = replace_table(c, c->h_imgs, S.images)) return e;
    }
    if (nm > 0) {
        for (int i = 0; i < nm; i++) c->h_mats.push_back(make_material(add->materials[i]));
        if (int e = replace_table(c, c->h_mats, S.materials)) return e;
    }
    if (nme > 0) {
        for (int i = 0; i < nme; i++) {
            GMesh gm;
            int32_t n_nodes;
            nrb::Box local;
            if (int e = build_mesh(c, add->meshes[i], gm, n_nodes, local)) return e;
            c->h_meshes.push_back(gm);
            c->h_mesh_nodes.push_back(n_nodes);
            c->h_mesh_local.push_back(local);
            c->max_mesh_depth = std::max(c->max_mesh_depth, gm.depth);
        }
        if (int e = replace_table(c, c->h_meshes, S.meshes)) return e;
    }
    nrb::Box lb;
    c->reg_objects.push_back(make_object(c, *obj, (int32_t)c->reg_objects.size(), lb));
    c->reg_local.push_back(lb);
    if (int e = upload_top(c, S, false)) return e;
    if (int e = publish_scene(c, S)) return e;
    c->has_scene = true;
    return NR_OK;
}

int nr_scene_set_sky(nr_ctx* c, const nr_sky_desc* sky) {
    if (!c || !sky) return fail(NR_ERR_INVALID, "null argument");
    if (!c->has_scene) return fail(NR_ERR_STATE, "no scene");
    if (int e = check_sky(*sky)) return e;
    HIPCK(hipSetDevice(c->device));
    HIPCK(hipStreamSynchronize(c->stream));
    c->has_scene = false;
    GScene S = c->scene;
    const GSky old = S.sky;
    if (int e = upload_sky(c, *sky, S.sky)) return e;
    for (const void* p : {(const void*)old.data, (const void*)old.rgbx, (const void*)old.marg, (const void*)old.alias})
        scene_release(c, p);
    if (int e = publish_scene(c, S)) return e;
    c->has_scene = true;
    return NR_OK;
}

int nr_camera_set(nr_ctx* c, int view, const nr_camera_desc* cd, uint32_t seed) {
    if (!c || !cd) return fail(NR_ERR_INVALID, "null argument");
    if (view < 0 || view >= NR_MAX_VIEWS) return fail(NR_ERR_INVALID, "view index out of range");
    if (cd->resolution[0] < 1.0f || cd->resolution[1] < 1.0f) return fail(NR_ERR_INVALID, "resolution must be positive");
    if (cd->samples_per_pixel < 1) return fail(NR_ERR_INVALID, "samples_per_pixel must be >= 1");
    HIPCK(hipSetDevice(c->device));
    View& v = c->views[view];
    uint32_t H = (uint32_t)cd->resolution[0], W = (uint32_t)cd->resolution[1];
    if ((uint64_t)H * W > (1ull << 31)) return fail(NR_ERR_LIMIT, "view larger than 2^31 pixels");
    HIPCK(hipStreamSynchronize(c->stream));
    GCamera& g = v.cam;                                                          // camera.h:124-136
    host::M3 R = host::euler(host::deg2rad(cd->rotation[0]), host::deg2rad(cd->rotation[1]), host::deg2rad(cd->rotation[2]));
    std::memcpy(g.R, R.m, sizeof(g.R));
    for (int a = 0; a < 3; a++) g.pos[a] = cd->position[a];
    g.shutter_time = 1.0f / (cd->shutter_speed + 1e-8f);
    float aperture_radius = cd->aperture / 2.0f;
    g.defocus_angle = 2.0f * atanf(aperture_radius / (cd->focus_distance + 1e-8f));
    g.defocus_radius = cd->focus_distance * tanf(g.defocus_angle / 2.0f);
    float aspect = cd->resolution[1] / cd->resolution[0];
    g.uvw[0] = cd->sensor_width; g.uvw[1] = -cd->sensor_width * aspect; g.uvw[2] = 0.0f;
    g.res[0] = cd->resolution[0]; g.res[1] = cd->resolution[1];
    g.focal_length = cd->focal_length; g.focus_distance = cd->focus_distance;
    g.H = H; g.W = W;
    bool pattern_dirty = !v.valid || v.seed != seed || v.desc.samples_per_pixel != cd->samples_per_pixel;
    if (pattern_dirty) {
        std::vector<float> pat = host::n_rooks(cd->samples_per_pixel, seed);
        if (v.d_pattern) HIPCK(hipFree(v.d_pattern));
        v.d_pattern = nullptr;
        HIPCK(hipMalloc((void**)&v.d_pattern, pat.size() * sizeof(float)));
        HIPCK(hipMemcpy(v.d_pattern, pat.data(), pat.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    g.pattern = v.d_pattern;
    g.n_pattern = cd->samples_per_pixel;
    if (!v.valid || v.H != H || v.W != W) {
        for (auto& l : v.layer) { if (l) HIPCK(hipFree(l)); l = nullptr; }
        const int layers[4] = {NR_LAYER_BEAUTY, NR_LAYER_DIFFUSE, NR_LAYER_WORLD_NORMAL, NR_LAYER_DEPTH};
        for (int L : layers) {
            const size_t bytes = sizeof(float) * layer_channels(L) * (size_t)H * W;
            HIPCK(hipMalloc((void**)&v.layer[L], bytes));
            HIPCK(hipMemset(v.layer[L], 0, bytes));
        }
    }
    v.H = H; v.W = W;
    v.desc = *cd;
    v.seed = seed;
    v.valid = true;
    return NR_OK;
}

int nr_aov_reset(nr_ctx* c, int view) {
    if (!c || view < 0 || view >= NR_MAX_VIEWS || !c->views[view].valid) return fail(NR_ERR_INVALID, "bad view");
    HIPCK(hipSetDevice(c->device));
    View& v = c->views[view];
    LayerWrite w(c, v);
    for (int L = 0; L < 9; L++)
        if (v.layer[L]) HIPCK(hipMemsetAsync(v.layer[L], 0, sizeof(float) * layer_channels(L) * (size_t)v.H * v.W, c->stream));
    return NR_OK;
}

// Path state for n pixels and queues for npad padded pixels (whole 8x8 tiles).
static int ensure_paths(nr_ctx* c, uint32_t n, uint32_t npad) {
    if (n <= c->path_cap && npad <= c->q_cap && c->P.slots == (uint32_t)c->slots) return NR_OK;
    n = std::max(n, c->path_cap);
    npad = std::max(npad, c->q_cap);
    HIPCK(hipStreamSynchronize(c->stream));
    free_list(c->path_mem);
    c->path_cap = 0;
    c->q_cap = 0;
    PathPtrs& P = c->P;
    auto al = [&](void** p, size_t bytes) -> int {
        HIPCK(hipMalloc(p, bytes));
        c->path_mem.push_back(*p);
        return NR_OK;
    };
    const uint32_t slots = (uint32_t)c->slots;
    const size_t np = (size_t)slots * n;         // path slots
    // origin and direction of a path side by side (one 32-byte record, read together)
    int e = al((void**)&P.o, sizeof(float4) * 2 * np);
    P.d = P.o + 1;
    if (!e) e = al((void**)&P.atten, sizeof(float4) * 4 * np);   // throughput, radiance, albedo, normal
    if (!e) { P.L = P.atten + 1; P.diff = P.atten + 2; P.nrm = P.atten + 3; }
    if (!e) e = al((void**)&P.hit, sizeof(GHit) * np);
    const size_t nq = (size_t)slots * npad;      // entries per queue
    if (!e) e = al((void**)&P.qact, sizeof(int32_t) * 2 * nq);
    if (!e) e = al((void**)&P.qmat, sizeof(int32_t) * NR_MAT_TYPES * nq);
    if (!e) e = al((void**)&P.counts, sizeof(int32_t) * CNT_N * NR_MAX_GROUPS);
    if (!e) e = al((void**)&P.finished, sizeof(uint8_t) * np);
    if (!e) e = al((void**)&P.nf, sizeof(uint32_t) * (size_t)n);
    if (!e) e = al((void**)&P.ns, sizeof(uint32_t) * (size_t)n);
    if (!e) e = al((void**)&P.acc, sizeof(float4) * 3 * (size_t)n);
    if (!e) e = al((void**)&P.sho, sizeof(float4) * NR_SH_STRIDE * np);   // shadow ray record
    if (!e) { P.shd = P.sho + 1; P.shc = P.sho + 2; }
    if (!e) e = al((void**)&P.qshd, sizeof(int32_t) * nq);
    if (e) { free_list(c->path_mem); return e; }
    HIPCK(hipMemsetAsync(P.finished, 2, np, c->stream));
    HIPCK(hipMemsetAsync(P.counts, 0, sizeof(int32_t) * CNT_N * NR_MAX_GROUPS, c->stream));
    P.cap = n;
    P.qcap = (uint32_t)nq;
    P.err = P.counts + CNT_ERR;
    P.slots = slots;
    c->path_cap = n;
    c->q_cap = npad;
    return NR_OK;
}

static hipEvent_t ev_at(nr_ctx* c, size_t i) {
    while (c->ev.size() <= i) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev.push_back(e);
    }
    return c->ev[i];
}

// per-tile sample limits of nr_render_tiles (host table, copied to c->d_tile_end)
struct TileLimitHost {
    const uint32_t* end;
    size_t n;
    uint32_t shift, cols;
};
struct TileLimit {
    const uint32_t* d_end;
    uint32_t shift, cols;
};

static int render_impl(nr_ctx* c, const int32_t* view_ids, int n_views, const nr_render_desc* d,
                       const TileLimitHost* tlh);

int nr_render(nr_ctx* c, int view, const nr_render_desc* d) {
    const int32_t v = view;
    return render_impl(c, &v, 1, d, nullptr);
}

int nr_render_views(nr_ctx* c, const int32_t* view_ids, int n_views, const nr_render_desc* d) {
    return render_impl(c, view_ids, n_views, d, nullptr);
}

int nr_render_tiles(nr_ctx* c, int view, const nr_render_desc* d, const uint32_t* tile_end, int tile_size) {
    if (!c || !d || !tile_end) return fail(NR_ERR_INVALID, "null argument");
    if (view < 0 || view >= NR_MAX_VIEWS || !c->views[view].valid) return fail(NR_ERR_STATE, "view has no camera");
    if (tile_size < 8 || tile_size > 4096 || (tile_size & (tile_size - 1)))
        return fail(NR_ERR_INVALID, "tile_size must be a power of two in [8, 4096]");
    uint32_t shift = 0;
    while ((1 << shift) < tile_size) shift++;
    const View& v = c->views[view];
    TileLimitHost t{tile_end, (size_t)((v.H + tile_size - 1) >> shift) * ((v.W + tile_size - 1) >> shift), shift,
                    (v.W + tile_size - 1) >> shift};
    const int32_t id = view;
    return render_impl(c, &id, 1, d, &t);
}

static int render_impl(nr_ctx* c, const int32_t* view_ids, int n_views, const nr_render_desc* d,
                       const TileLimitHost* tlh) {
    if (!c || !d || !view_ids) return fail(NR_ERR_INVALID, "null argument");
    if (!c->has_scene) return fail(NR_ERR_STATE, "nr_render before nr_scene_upload");
    if (n_views < 1 || n_views > NR_MAX_VIEWS) return fail(NR_ERR_INVALID, "1 to NR_MAX_VIEWS views per call");
    std::vector<int32_t> ids(view_ids, view_ids + n_views);
    for (int32_t id : ids)
        if (id < 0 || id >= NR_MAX_VIEWS || !c->views[id].valid) return fail(NR_ERR_STATE, "view has no camera");
    {
        std::vector<int32_t> u(ids);
        std::sort(u.begin(), u.end());
        if (std::adjacent_find(u.begin(), u.end()) != u.end()) return fail(NR_ERR_INVALID, "a view appears twice");
    }
    for (int32_t id : ids)
        if (d->sample_end > c->views[id].cam.n_pattern) return fail(NR_ERR_INVALID, "sample index beyond samples_per_pixel");
    if (d->mode != NR_MODE_PARITY && d->mode != NR_MODE_FAST) return fail(NR_ERR_INVALID, "unknown render mode");
    const bool fast = d->mode == NR_MODE_FAST;
    if (!fast && d->sample_begin < 1 && d->sample_end > d->sample_begin)
        return fail(NR_ERR_INVALID, "sample passes start at 1 (running mean of passes 1..n-1)");
    HIPCK(hipSetDevice(c->device));
    // the views one after another in the tile enumeration and in the path arrays
    GView hv[NR_MAX_VIEWS];
    uint64_t tiles = 0, npix = 0;
    for (int k = 0; k < n_views; k++) {
        const View& v = c->views[ids[k]];
        hv[k].cam = v.cam;
        hv[k].tile0 = (uint32_t)tiles;
        hv[k].pix0 = (uint32_t)npix;
        hv[k].tiles_x = (v.W + 7) / 8;
        hv[k].npix = v.W * v.H;
        tiles += (uint64_t)hv[k].tiles_x * ((v.H + 7) / 8);
        npix += (uint64_t)v.W * v.H;
    }
    if (tiles * 64 * (uint64_t)c->slots > (uint64_t)INT32_MAX) return fail(NR_ERR_LIMIT, "views too large for 32-bit path ids");
    const uint32_t N = (uint32_t)npix, T = (uint32_t)tiles;
    int e = ensure_paths(c, N, T * 64);
    if (e) return e;
    // the traversal-overflow flag reports this call only (read by nr_sync, nr_aov_read
    // and timing mode)
    hipLaunchKernelGGL(k_clear_err, dim3(1), dim3(64), 0, c->stream, c->P.err);
    if (!c->d_views) HIPCK(hipMalloc((void**)&c->d_views, sizeof(GView) * NR_MAX_VIEWS));
    HIPCK(hipStreamSynchronize(c->stream));              // kernels of the previous call read the tables
    HIPCK(hipMemcpy(c->d_views, hv, sizeof(GView) * n_views, hipMemcpyHostToDevice));
    TileLimit tl_dev{nullptr, 0u, 0u};
    const TileLimit* tl = nullptr;
    if (tlh) {
        if (tlh->n > c->tile_end_cap) {
            if (c->d_tile_end) HIPCK(hipFree(c->d_tile_end));
            c->d_tile_end = nullptr;
            c->tile_end_cap = 0;
            HIPCK(hipMalloc((void**)&c->d_tile_end, sizeof(uint32_t) * tlh->n));
            c->tile_end_cap = tlh->n;
        }
        HIPCK(hipMemcpy(c->d_tile_end, tlh->end, sizeof(uint32_t) * tlh->n, hipMemcpyHostToDevice));
        tl_dev = TileLimit{c->d_tile_end, tlh->shift, tlh->cols};
        tl = &tl_dev;
    }
    const size_t lds = trav_lds_bytes();
    dim3 gS(c->shade_grid);
    size_t ei = 0;
    float t_ext = 0.0f, t_shade = 0.0f, t_other = 0.0f, t_shadow = 0.0f, t_tail = 0.0f;
    std::vector<std::pair<size_t, int>> marks;   // (event index of start, category)
    auto mark = [&](int cat) {
        if (!c->timing) return;
        hipEvent_t a = ev_at(c, ei);
        (void)hipEventRecord(a, c->stream);
        marks.push_back({ei, cat});
        ei++;
    };
    auto layer_of = [&](int k, int l) { return c->views[ids[k]].layer[l]; };
    // views whose layers this call writes, locked in ascending view order
    auto write_guards = [&]() {
        std::vector<int32_t> order(ids);
        std::sort(order.begin(), order.end());
        std::vector<std::unique_ptr<LayerWrite>> w;
        for (int32_t id : order) w.emplace_back(new LayerWrite(c, c->views[id]));
        return w;
    };
    hipEvent_t e0 = c->ev_begin, e1 = c->ev_end;
    if (c->timing) (void)hipEventRecord(e0, c->stream);
    c->last_iterations = 0;
    if (d->sample_end > d->sample_begin && d->ray_depth == 0) {
        auto w = write_guards();
        for (int k = 0; k < n_views; k++) {
            const uint32_t n = hv[k].npix;
            for (uint32_t s = d->sample_begin; s < d->sample_end; s++)
                hipLaunchKernelGGL(k_fold_zero, dim3((n + 255) / 256), dim3(256), 0, c->stream, layer_of(k, NR_LAYER_BEAUTY),
                                   layer_of(k, NR_LAYER_DIFFUSE), layer_of(k, NR_LAYER_WORLD_NORMAL),
                                   layer_of(k, NR_LAYER_DEPTH), n, s,
                                   fast ? s + 1 : s, c->views[ids[k]].W, tl ? tl->d_end : nullptr, tl ? tl->shift : 0u,
                                   tl ? tl->cols : 0u);
        }
    } else if (d->sample_end > d->sample_begin) {
        // Stream groups: G contiguous ranges of whole tiles, each with its own
        // queues, counters, stack-spill columns and stream.  Timing mode runs one
        // group on the context stream so per-kernel times are not shared with
        // another group.
        const int G = c->timing ? 1 : (int)std::max(1u, std::min<uint32_t>((uint32_t)c->render_streams, T));
        const uint32_t S = c->P.slots;
        const uint32_t m = std::min<uint32_t>(S, d->sample_end - d->sample_begin);
        struct Group {
            PathPtrs P;
            TraceArgs A;
            StackCfg SC;
            hipStream_t st;
            int64_t it = 0, bound = 0, pending = 1;
            int64_t iters = 0;      // streaming iterations enqueued (a graph replay enqueues several)
            int64_t cap = 0;        // path slots of the group: the graphs' traversal grids
            bool done = false;
        } grp[NR_MAX_GROUPS];
        const size_t lv = c->scene.stack_depth > LDS_STACK_LEVELS ? (size_t)(c->scene.stack_depth - LDS_STACK_LEVELS) : 0;
        {
            StackCfg all;
            e = stack_spill(c, S * T * 64 + 128u * NR_MAX_GROUPS, all);
            if (e) return e;
            size_t spill_off = 0;
            for (int g = 0; g < G; g++) {
                const uint32_t T0 = (uint32_t)((uint64_t)T * g / G), T1 = (uint32_t)((uint64_t)T * (g + 1) / G);
                const size_t q0 = (size_t)S * T0 * 64;              // queue entries of the groups before
                Group& gr = grp[g];
                gr.P = c->P;
                gr.P.qact = c->P.qact + q0;
                gr.P.qmat = c->P.qmat + q0;
                gr.P.qshd = c->P.qshd + q0;
                gr.P.counts = c->P.counts + CNT_N * g;
                gr.A = TraceArgs{d->seed, 0u, d->ray_depth, d->do_rr, d->rr_start_bounce, d->rr_survival,
                                 d->sample_end, T0 * 64, T1 * 64, fast ? 1 : 0, (uint32_t)n_views, c->d_views,
                                 tl ? tl->d_end : nullptr, tl ? tl->shift : 0u, tl ? tl->cols : 0u};
                const uint32_t cap_g = (S * (T1 - T0) * 64 + 127) / 128 * 128;
                gr.SC.spill = all.spill + spill_off;
                gr.SC.stride = cap_g;
                spill_off += lv * cap_g;
                gr.st = G == 1 ? c->stream : c->gstream[g];
                gr.bound = (int64_t)(T1 - T0) * 64 * m;
                gr.cap = cap_g;
            }
        }
        mark(2);
        for (int k = 0; k < n_views; k++)
            hipLaunchKernelGGL(k_pack_acc, dim3((hv[k].npix + 255) / 256), dim3(256), 0, c->stream,
                               layer_of(k, NR_LAYER_BEAUTY), layer_of(k, NR_LAYER_DIFFUSE),
                               layer_of(k, NR_LAYER_WORLD_NORMAL), layer_of(k, NR_LAYER_DEPTH), c->P.acc + hv[k].pix0,
                               hv[k].npix, c->P.cap);
        if (G > 1) {
            HIPCK(hipEventRecord(c->ev_packed, c->stream));
            for (int g = 0; g < G; g++) HIPCK(hipStreamWaitEvent(grp[g].st, c->ev_packed, 0));
        }
        for (int g = 0; g < G; g++) {
            const uint32_t nt = grp[g].A.t1 - grp[g].A.t0;
            hipLaunchKernelGGL(k_clear_counts, dim3(1), dim3(64), 0, grp[g].st, grp[g].P.counts);
            hipLaunchKernelGGL(k_start, dim3((nt + 255) / 256), dim3(256), 0, grp[g].st, grp[g].P, d->sample_begin,
                               grp[g].A);
        }
        // The occupied path slots (samples started and not folded) never grow
        // from one iteration to the next and bound the live paths, so a count
        // read back two iterations late is an upper bound for the traversal
        // grid and each group always has its next iteration queued.
        // With several groups the host takes whichever group's count is back
        // (event query) and only spins when none is: waiting on one group
        // while another one's queue runs dry leaves that group's stream idle.
        // Every exit from here on, error returns included, makes the context stream
        // wait for the group streams, so a later synchronisation of the context
        // stream (before path or scene memory is freed) covers their kernels.
        struct GroupJoin {
            nr_ctx* c;
            int G;
            const Group* grp;
            ~GroupJoin() {
                if (G <= 1) return;
                for (int g = 0; g < G; g++) {
                    (void)hipEventRecord(c->gev_done[g], grp[g].st);
                    (void)hipStreamWaitEvent(c->stream, c->gev_done[g], 0);
                }
            }
        };
        const bool other = (c->obj_types & NR_TY_OTHER) != 0;
        constexpr int TYI = NR_TY_QUAD | NR_TY_SPHERE | NR_TY_MESH;
        // One streaming iteration of a group: clear the counters it appends to, trace,
        // shade each material present, trace the shadow rays (fast mode), fold and start
        // samples.  `bound` sizes the traversal grids (workgroups past the live count
        // exit at once).
        auto enqueue_iteration = [&](Group& gr, const TraceArgs& A, int64_t bound) {
            const PathPtrs& P = gr.P;
            hipStream_t st = gr.st;
            hipLaunchKernelGGL(k_counts, dim3(1), dim3(64), 0, st, P.counts, (int32_t)A.in);
            mark(0);
            const dim3 gx((uint32_t)((bound + 127) / 128));
            if (fast) {
                if (other) hipLaunchKernelGGL((k_extend<NR_TY_ALL, true>), gx, dim3(128), lds, st, c->d_scene, P, A, gr.SC);
                else hipLaunchKernelGGL((k_extend<TYI, true>), gx, dim3(128), lds, st, c->d_scene, P, A, gr.SC);
            } else {
                if (other) hipLaunchKernelGGL((k_extend<NR_TY_ALL, false>), gx, dim3(128), lds, st, c->d_scene, P, A, gr.SC);
                else hipLaunchKernelGGL((k_extend<TYI, false>), gx, dim3(128), lds, st, c->d_scene, P, A, gr.SC);
            }
            mark(1);
            // the shading grid-stride loops need no more workgroups than live paths
            const dim3 gSi((uint32_t)std::max<int64_t>(1, std::min<int64_t>(gS.x, (bound + 255) / 256)));
            if (fast) launch_shade<true>(c, gSi, st, P, A);
            else launch_shade<false>(c, gSi, st, P, A);
            if (fast && c->scene.sky.marg) {
                mark(4);
                if (other) hipLaunchKernelGGL(k_shadow<NR_TY_ALL>, gx, dim3(128), lds, st, c->d_scene, P, gr.SC);
                else hipLaunchKernelGGL(k_shadow<TYI>, gx, dim3(128), lds, st, c->d_scene, P, gr.SC);
            }
            mark(2);
            const uint32_t nt = A.t1 - A.t0;
            hipLaunchKernelGGL(k_regen, dim3((nt + REGEN_BLOCK - 1) / REGEN_BLOCK), dim3(REGEN_BLOCK), 0, st, P, A);
        };
        // Graph mode: group g's graph of graph_iters consecutive iterations with grids of
        // the group's capacity, captured on its stream (its launch arguments are the
        // cache key: a later call with the same arguments replays the same graph).
        const bool use_graph = c->graph_iters > 0 && !c->timing;
        auto group_graph = [&](int g, Group& gr) -> int {
            struct Key {
                PathPtrs P;
                TraceArgs A;
                StackCfg SC;
                int64_t cap;
                const void* scene;
                hipStream_t st;
                int32_t iters, shade, fast, other, marg;
                uint8_t mats[NR_MAT_TYPES];
            } k;
            std::memset(&k, 0, sizeof k);         // padding bytes take part in the comparison
            k.P = gr.P; k.A = gr.A; k.SC = gr.SC; k.cap = gr.cap; k.scene = c->d_scene; k.st = gr.st;
            k.iters = c->graph_iters; k.shade = (int32_t)gS.x; k.fast = fast; k.other = other;
            k.marg = c->scene.sky.marg != nullptr;
            for (int t = 0; t < NR_MAT_TYPES; t++) k.mats[t] = c->mat_present[t] ? 1 : 0;
            const uint8_t* kb = reinterpret_cast<const uint8_t*>(&k);
            if (c->gexec[g] && c->gkey[g].size() == sizeof k && std::memcmp(c->gkey[g].data(), kb, sizeof k) == 0)
                return NR_OK;
            if (c->gexec[g]) {                    // not running: the call began by waiting for the last one
                (void)hipGraphExecDestroy(c->gexec[g]);
                c->gexec[g] = nullptr;
                c->gkey[g].clear();
            }
            HIPCK(hipStreamBeginCapture(gr.st, hipStreamCaptureModeThreadLocal));
            TraceArgs Ak = gr.A;
            for (int i = 0; i < c->graph_iters; i++) {
                enqueue_iteration(gr, Ak, gr.cap);
                Ak.in = 1 - Ak.in;
            }
            const hipError_t le = hipGetLastError();
            hipGraph_t graph = nullptr;
            const hipError_t ce = hipStreamEndCapture(gr.st, &graph);
            if (le != hipSuccess || ce != hipSuccess || !graph) {
                if (graph) (void)hipGraphDestroy(graph);
                return fail(NR_ERR_HIP, std::string("capturing the streaming iterations: ") +
                                            hipGetErrorString(le != hipSuccess ? le : ce));
            }
            hipGraphExec_t ex = nullptr;
            const hipError_t ie = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) return fail(NR_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
            c->gexec[g] = ex;
            c->gkey[g].assign(kb, kb + sizeof k);
            return NR_OK;
        };
        const int64_t max_iters = (int64_t)(d->sample_end - d->sample_begin) * d->ray_depth + 2;
        int active = G;
        {
        GroupJoin join{c, G, grp};
        while (active > 0) {
            bool progressed = false;
            for (int g = 0; g < G; g++) {
                Group& gr = grp[g];
                if (gr.done) continue;
                const int64_t it = gr.it;
                int32_t* live = c->h_live + 4 * g;        // [occupied, pending] of the last two iterations
                if (it >= 2) {
                    if (G > 1) {
                        const hipError_t q = hipEventQuery(c->gev[g][it & 1]);
                        if (q == hipErrorNotReady) continue;
                        HIPCK(q);
                    } else {
                        HIPCK(hipEventSynchronize(c->gev[g][it & 1]));
                    }
                    progressed = true;
                    gr.bound = live[2 * (it & 1)];
                    gr.pending = live[2 * (it & 1) + 1];
                    if (gr.bound == 0) { gr.done = true; active--; continue; }
                }
                progressed = true;
                // (a graph replay runs graph_iters iterations, and the drain is seen two replays late)
                if (gr.iters > max_iters + (use_graph ? 3 * c->graph_iters : 0))
                    return fail(NR_ERR_STATE, "streaming loop did not drain");
                PathPtrs& P = gr.P;
                TraceArgs& A = gr.A;
                hipStream_t st = gr.st;
                if (it >= 2 && gr.pending == 0 && gr.bound <= (int64_t)c->tail_paths) {
                    // tail mode: every sample is started; one thread per live path runs it to its end,
                    // then k_regen folds them all
                    mark(5);
                    const dim3 gt((uint32_t)((gr.bound + 127) / 128));
                    if (fast) {
                        if (other) hipLaunchKernelGGL((k_tail<NR_TY_ALL, true>), gt, dim3(128), lds, st, c->d_scene, P, A, gr.SC);
                        else hipLaunchKernelGGL((k_tail<TYI, true>), gt, dim3(128), lds, st, c->d_scene, P, A, gr.SC);
                    } else {
                        if (other) hipLaunchKernelGGL((k_tail<NR_TY_ALL, false>), gt, dim3(128), lds, st, c->d_scene, P, A, gr.SC);
                        else hipLaunchKernelGGL((k_tail<TYI, false>), gt, dim3(128), lds, st, c->d_scene, P, A, gr.SC);
                    }
                    mark(2);
                    const uint32_t nt = A.t1 - A.t0;
                    hipLaunchKernelGGL(k_regen, dim3((nt + REGEN_BLOCK - 1) / REGEN_BLOCK), dim3(REGEN_BLOCK), 0, st, P, A);
                    HIPCK(hipGetLastError());
                    gr.done = true;
                    active--;
                    continue;
                }
                if (use_graph) {
                    if (int e = group_graph(g, gr)) return e;
                    HIPCK(hipGraphLaunch(c->gexec[g], st));
                    gr.iters += c->graph_iters;     // an even count: the replay leaves A.in as it was
                } else {
                    enqueue_iteration(gr, A, gr.bound);
                    HIPCK(hipGetLastError());
                    A.in = 1 - A.in;
                    gr.iters++;
                }
                HIPCK(hipMemcpyAsync(live + 2 * (it & 1), P.counts + CNT_OCC, 2 * sizeof(int32_t), hipMemcpyDeviceToHost,
                                     st));
                HIPCK(hipEventRecord(c->gev[g][it & 1], st));
                gr.it++;
            }
            if (!progressed) std::this_thread::yield();
        }
        c->last_iterations = 0;
        for (int g = 0; g < G; g++) c->last_iterations += grp[g].iters;
        }   // GroupJoin: the context stream waits for every group
        mark(2);
        auto w = write_guards();
        for (int k = 0; k < n_views; k++)
            hipLaunchKernelGGL(k_unpack_acc, dim3((hv[k].npix + 255) / 256), dim3(256), 0, c->stream,
                               c->P.acc + hv[k].pix0, layer_of(k, NR_LAYER_BEAUTY), layer_of(k, NR_LAYER_DIFFUSE),
                               layer_of(k, NR_LAYER_WORLD_NORMAL), layer_of(k, NR_LAYER_DEPTH), hv[k].npix, c->P.cap);
    }
    HIPCK(hipGetLastError());
    if (c->timing) {
        mark(3);
        (void)hipEventRecord(e1, c->stream);
        HIPCK(hipEventSynchronize(e1));
        for (size_t k = 0; k + 1 < marks.size(); k++) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, c->ev[marks[k].first], c->ev[marks[k + 1].first]);
            if (marks[k].second == 0) t_ext += ms;
            else if (marks[k].second == 1) t_shade += ms;
            else if (marks[k].second == 4) t_shadow += ms;
            else if (marks[k].second == 5) t_tail += ms;
            else t_other += ms;
        }
        float tot = 0.0f;
        (void)hipEventElapsedTime(&tot, e0, e1);
        c->last_ms[0] = tot; c->last_ms[1] = t_ext; c->last_ms[2] = t_shade; c->last_ms[3] = t_other;
        c->last_ms[4] = (float)c->last_iterations;
        c->last_ms[5] = t_shadow;
        c->last_ms[6] = t_tail;
        unsigned long long traced = 0;                   // timing mode renders with one stream group
        HIPCK(hipMemcpy(&traced, c->P.counts + CNT_TRACED, sizeof(traced), hipMemcpyDeviceToHost));
        c->last_ms[7] = (float)traced;
        c->n_last = 8;
        int32_t err = 0;
        HIPCK(hipMemcpy(&err, c->P.counts + CNT_ERR, sizeof(int32_t), hipMemcpyDeviceToHost));
        if (err) return fail(NR_ERR_LIMIT, "traversal stack overflow");
    }
    return NR_OK;
}

int nr_view_size(nr_ctx* c, int view, int32_t* H, int32_t* W) {
    if (!c || !H || !W || view < 0 || view >= NR_MAX_VIEWS || !c->views[view].valid) return fail(NR_ERR_INVALID, "bad view");
    *H = (int32_t)c->views[view].H;
    *W = (int32_t)c->views[view].W;
    return NR_OK;
}

int nr_aov_read(nr_ctx* c, int view, int layer, float* host, size_t n_floats) {
    if (!c || !host || view < 0 || view >= NR_MAX_VIEWS || !c->views[view].valid) return fail(NR_ERR_INVALID, "bad view");
    if (layer < 0 || layer > 8 || !c->views[view].layer[layer]) return fail(NR_ERR_INVALID, "layer not allocated");
    View& v = c->views[view];
    size_t need = (size_t)layer_channels(layer) * v.H * v.W;
    if (n_floats < need) return fail(NR_ERR_INVALID, "host buffer too small");
    HIPCK(hipSetDevice(c->device));
    HIPCK(hipMemcpyAsync(host, v.layer[layer], need * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    int32_t err = 0;
    if (c->P.err) {                                     // the last render's traversal-overflow flag
        HIPCK(hipMemcpy(&err, c->P.err, sizeof(int32_t), hipMemcpyDeviceToHost));
        if (err) return fail(NR_ERR_LIMIT, "traversal stack overflow during render");
    }
    return NR_OK;
}

float* nr_aov_device_ptr(nr_ctx* c, int view, int layer) {
    if (!c || view < 0 || view >= NR_MAX_VIEWS || layer < 0 || layer > 8) return nullptr;
    return c->views[view].layer[layer];
}

int nr_set_trace_layout(nr_ctx* c, int layout) {
    if (!c) return fail(NR_ERR_INVALID, "null context");
    if (layout < NR_LAY_BINARY || layout > NR_LAY_QUANT)
        return fail(NR_ERR_INVALID, "layout must be 0 (binary), 1 (4-wide) or 2 (quantised 4-wide)");
    c->trace_layout = layout;
    return NR_OK;
}

// k_trace_closest for the scene's object types and the context's layout
static void launch_trace(nr_ctx* c, dim3 g, size_t lds, const float* rays, int n, nr_hit* hits, int32_t* counts,
                         const StackCfg& SC) {
    constexpr int TYI = NR_TY_QUAD | NR_TY_SPHERE | NR_TY_MESH;
    const dim3 b(128);
    const bool all = (c->obj_types & NR_TY_OTHER) != 0;
#define NR_TRACE_LAUNCH(TY_, LAY_) \
    hipLaunchKernelGGL((k_trace_closest<TY_, LAY_>), g, b, lds, c->stream, c->d_scene, rays, n, hits, counts, SC)
    if (c->trace_layout == NR_LAY_QUANT) { if (all) NR_TRACE_LAUNCH(NR_TY_ALL, NR_LAY_QUANT); else NR_TRACE_LAUNCH(TYI, NR_LAY_QUANT); }
    else if (c->trace_layout == NR_LAY_WIDE) { if (all) NR_TRACE_LAUNCH(NR_TY_ALL, NR_LAY_WIDE); else NR_TRACE_LAUNCH(TYI, NR_LAY_WIDE); }
    else { if (all) NR_TRACE_LAUNCH(NR_TY_ALL, NR_LAY_BINARY); else NR_TRACE_LAUNCH(TYI, NR_LAY_BINARY); }
#undef NR_TRACE_LAUNCH
}

int nr_trace_closest_device(nr_ctx* c, const float* d_rays, int32_t n, nr_hit* d_hits) {
    if (!c) return fail(NR_ERR_INVALID, "null context");
    if (!c->has_scene) return fail(NR_ERR_STATE, "trace before scene upload");
    if (n < 0) return fail(NR_ERR_INVALID, "negative ray count");
    if (n == 0) return NR_OK;
    HIPCK(hipSetDevice(c->device));
    StackCfg SC;
    int e = stack_spill(c, (uint32_t)n, SC);
    if (e) return e;
    HIPCK(hipMemsetAsync(c->d_err + CNT_ERR, 0, sizeof(int32_t), c->stream));
    hipEvent_t e0 = c->ev_begin, e1 = c->ev_end;
    if (c->timing) (void)hipEventRecord(e0, c->stream);
    launch_trace(c, dim3((n + 127) / 128), trav_lds_bytes(), d_rays, n, d_hits, c->d_err, SC);
    HIPCK(hipGetLastError());
    if (c->timing) {
        (void)hipEventRecord(e1, c->stream);
        HIPCK(hipEventSynchronize(e1));
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        c->last_ms[0] = ms; c->last_ms[1] = ms; c->last_ms[2] = 0.0f; c->last_ms[3] = 0.0f;
        c->n_last = 4;
    }
    return NR_OK;
}

int nr_trace_closest_device_noalloc(nr_ctx* c, const float* d_rays, int32_t n, nr_hit* d_hits) {
    if (!c) return fail(NR_ERR_INVALID, "null context");
    if (!c->has_scene) return fail(NR_ERR_STATE, "trace before scene upload");
    if (n < 0) return fail(NR_ERR_INVALID, "negative ray count");
    if (n == 0) return NR_OK;
    if (!d_rays || !d_hits) return fail(NR_ERR_INVALID, "null ray or hit buffer");
    if (!c->scratch) return fail(NR_ERR_STATE, "no scratch on this device");
    HIPCK(hipSetDevice(c->device));
    // rays per launch: as many spill columns as the scratch holds, whole workgroups
    const size_t lv = c->scene.stack_depth > LDS_STACK_LEVELS ? (size_t)(c->scene.stack_depth - LDS_STACK_LEVELS) : 0;
    const size_t col_bytes = lv * sizeof(int2);
    const size_t room = NR_NOALLOC_SCRATCH_BYTES - NR_NOALLOC_COUNTER_BYTES;
    size_t chunk = col_bytes ? room / col_bytes / 128 * 128 : (size_t)n;
    if (chunk == 0)
        return fail(NR_ERR_LIMIT, "the scene's traversal stack (" + std::to_string(c->scene.stack_depth) +
                                      " levels) needs more than the " + std::to_string(NR_NOALLOC_SCRATCH_BYTES >> 20) +
                                      " MiB scratch holds for one workgroup of 128 rays");
    chunk = std::min(chunk, (size_t)n);
    int32_t* counters = (int32_t*)c->scratch;
    int2* spill = (int2*)((char*)c->scratch + NR_NOALLOC_COUNTER_BYTES);
    hipLaunchKernelGGL(k_clear_err, dim3(1), dim3(64), 0, c->stream, counters + CNT_ERR);
    const size_t lds = trav_lds_bytes();
    for (size_t off = 0; off < (size_t)n; off += chunk) {
        const int m = (int)std::min(chunk, (size_t)n - off);
        const StackCfg SC{spill, (size_t)m};
        launch_trace(c, dim3((m + 127) / 128), lds, d_rays + 6 * off, m, d_hits + off, counters, SC);
        HIPCK(hipGetLastError());
    }
    return NR_OK;
}

int nr_trace_closest(nr_ctx* c, const float* rays, int32_t n, nr_hit* hits) {
    if (!c || (!rays && n > 0) || (!hits && n > 0)) return fail(NR_ERR_INVALID, "null argument");
    if (!c->has_scene) return fail(NR_ERR_STATE, "trace before scene upload");
    if (n <= 0) return n == 0 ? NR_OK : fail(NR_ERR_INVALID, "negative ray count");
    HIPCK(hipSetDevice(c->device));
    float* dr = nullptr;
    nr_hit* dh = nullptr;
    HIPCK(hipMalloc((void**)&dr, sizeof(float) * 6 * (size_t)n));
    if (hipMalloc((void**)&dh, sizeof(nr_hit) * (size_t)n) != hipSuccess) { (void)hipFree(dr); return fail(NR_ERR_HIP, "hipMalloc hits"); }
    int rc = NR_OK;
    int32_t zero = 0, err = 0;
    if (hipMemcpyAsync(dr, rays, sizeof(float) * 6 * (size_t)n, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(c->d_err + CNT_ERR, &zero, sizeof(int32_t), hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = fail(NR_ERR_HIP, "copy rays");
    if (rc == NR_OK) rc = nr_trace_closest_device(c, dr, n, dh);
    if (rc == NR_OK && (hipMemcpyAsync(hits, dh, sizeof(nr_hit) * (size_t)n, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                        hipMemcpyAsync(&err, c->d_err + CNT_ERR, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                        hipStreamSynchronize(c->stream) != hipSuccess))
        rc = fail(NR_ERR_HIP, "copy hits");
    (void)hipFree(dr);
    (void)hipFree(dh);
    if (rc == NR_OK && err) rc = fail(NR_ERR_LIMIT, "traversal stack overflow");
    return rc;
}

static float host_kelvin_channel(float kelvin, int ch) {                         // stream.cu:65-86
    float temp = kelvin / 100.0f;
    float r, g, b;
    if (temp <= 66.0f) { r = 255.0f; g = 99.4708025861f * logf(temp) - 161.1195681661f; }
    else { r = 329.698727446f * powf(temp - 60.0f, -0.1332047592f); g = 288.1221695283f * powf(temp - 60.0f, -0.0755148492f); }
    if (temp >= 66.0f) b = 255.0f;
    else if (temp <= 19.0f) b = 0.0f;
    else b = 138.5177312231f * logf(temp - 10.0f) - 305.0447927307f;
    float v = ch == 0 ? r : (ch == 1 ? g : b);
    return fminf(1.0f, fmaxf(0.0f, v / 255.0f));
}

static PostArgs post_args(const nr_post_desc* d) {
    PostArgs A{};
    float tgt_k = fmaxf(1000.0f, fminf(d->wb_temperature, 40000.0f));
    for (int k = 0; k < 3; k++)
        A.ratio[k] = host_kelvin_channel(tgt_k, k) / fmaxf(host_kelvin_channel(6500.0f, k), 1e-4f);
    A.tint = d->wb_tint * 0.0025f;
    A.wb = d->apply_white_balance; A.tm = d->apply_tonemapping; A.tm_method = d->tm_method;
    A.gamma = d->linear_to_gamma; A.fix_reinhard = d->fix_reinhard;
    A.white_point = d->tm_white_point; A.alpha = d->tm_alpha;
    return A;
}

int nr_post(nr_ctx* c, int view, const nr_post_desc* d, uint8_t* out8, float* outf) {
    if (!c || !d || view < 0 || view >= NR_MAX_VIEWS || !c->views[view].valid) return fail(NR_ERR_INVALID, "bad view");
    View& v = c->views[view];
    if (d->layer < 0 || d->layer > 8 || !v.layer[d->layer]) return fail(NR_ERR_INVALID, "post: layer not allocated");
    if (layer_channels(d->layer) != 3) return fail(NR_ERR_INVALID, "post: the layer has one channel");
    HIPCK(hipSetDevice(c->device));
    uint32_t HW = v.H * v.W;
    PostArgs A = post_args(d);
    uint8_t* d8 = nullptr;
    float* df = nullptr;
    if (out8) HIPCK(hipMalloc((void**)&d8, 3 * (size_t)HW));
    if (outf && hipMalloc((void**)&df, 3 * sizeof(float) * (size_t)HW) != hipSuccess) {
        if (d8) (void)hipFree(d8);
        return fail(NR_ERR_HIP, "hipMalloc post");
    }
    hipLaunchKernelGGL(k_post, dim3((HW + 255) / 256), dim3(256), 0, c->stream, v.layer[d->layer], HW, A, d8, df);
    int rc = NR_OK;
    if (hipGetLastError() != hipSuccess) rc = fail(NR_ERR_HIP, "k_post launch");
    if (rc == NR_OK && out8 && hipMemcpyAsync(out8, d8, 3 * (size_t)HW, hipMemcpyDeviceToHost, c->stream) != hipSuccess) rc = fail(NR_ERR_HIP, "copy post u8");
    if (rc == NR_OK && outf && hipMemcpyAsync(outf, df, 3 * sizeof(float) * (size_t)HW, hipMemcpyDeviceToHost, c->stream) != hipSuccess) rc = fail(NR_ERR_HIP, "copy post f32");
    if (hipStreamSynchronize(c->stream) != hipSuccess && rc == NR_OK) rc = fail(NR_ERR_HIP, "post sync");
    if (d8) (void)hipFree(d8);
    if (df) (void)hipFree(df);
    return rc;
}

int nr_post_device(nr_ctx* c, const float* img, int32_t H, int32_t W, const nr_post_desc* d, uint8_t* d8, float* df) {
    if (!c || !d || !img) return fail(NR_ERR_INVALID, "null argument");
    if (H <= 0 || W <= 0 || (uint64_t)H * W > (1ull << 31)) return fail(NR_ERR_INVALID, "bad image size");
    HIPCK(hipSetDevice(c->device));
    const uint32_t HW = (uint32_t)H * (uint32_t)W;
    hipLaunchKernelGGL(k_post, dim3((HW + 255) / 256), dim3(256), 0, c->stream, img, HW, post_args(d), d8, df);
    HIPCK(hipGetLastError());
    HIPCK(hipStreamSynchronize(c->stream));
    return NR_OK;
}

int nr_tile_error(nr_ctx* c, const float* d_denoised, const float* d_judge1, const float* d_judge2, int32_t H,
                  int32_t W, int32_t tile_size, float* d_out) {
    if (!c || !d_denoised || !d_judge1 || !d_judge2 || !d_out) return fail(NR_ERR_INVALID, "null argument");
    if (H <= 0 || W <= 0 || (uint64_t)H * W > (1ull << 31)) return fail(NR_ERR_INVALID, "bad image size");
    if (tile_size < 1 || tile_size > 4096) return fail(NR_ERR_INVALID, "tile_size must be in [1, 4096]");
    HIPCK(hipSetDevice(c->device));
    const uint32_t cols = (W + tile_size - 1) / tile_size, rows = (H + tile_size - 1) / tile_size;
    hipLaunchKernelGGL(k_tile_error, dim3(cols * rows), dim3(256), 0, c->stream, d_denoised, d_judge1, d_judge2,
                       (uint32_t)H, (uint32_t)W, (uint32_t)tile_size, cols, d_out);
    HIPCK(hipGetLastError());
    HIPCK(hipStreamSynchronize(c->stream));
    return NR_OK;
}

// ---------------------------------------------------------------------------
// interactive mode
// ---------------------------------------------------------------------------
static int view_of(nr_ctx* c, int view, View** out) {
    if (!c) return fail(NR_ERR_INVALID, "null context");
    if (view < 0 || view >= NR_MAX_VIEWS || !c->views[view].valid) return fail(NR_ERR_INVALID, "bad view");
    *out = &c->views[view];
    return NR_OK;
}

// 1-D grid of 128-thread workgroups over the view's 16 x 8 pixel tiles
static uint32_t tile_blocks(const View& v) { return ((v.W + 15) / 16) * ((v.H + 7) / 8); }

static int check_overflow(nr_ctx* c) {
    int32_t err = 0;
    HIPCK(hipMemcpyAsync(&err, c->d_err + CNT_ERR, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    if (err) return fail(NR_ERR_LIMIT, "traversal stack overflow");
    return NR_OK;
}

int nr_viewport(nr_ctx* c, int view, int show_axis_grid) {
    View* v = nullptr;
    if (int e = view_of(c, view, &v)) return e;
    if (!c->has_scene) return fail(NR_ERR_STATE, "nr_viewport before nr_scene_upload");
    float* lb = v->layer[NR_LAYER_BEAUTY];
    float* ld = v->layer[NR_LAYER_DIFFUSE];
    float* ln = v->layer[NR_LAYER_WORLD_NORMAL];
    if (!lb || !ld || !ln) return fail(NR_ERR_STATE, "view has no beauty/diffuse/world-normal layers");
    HIPCK(hipSetDevice(c->device));
    const uint32_t nb = tile_blocks(*v);
    StackCfg SC;
    if (int e = stack_spill(c, nb * 128u, SC)) return e;
    HIPCK(hipMemsetAsync(c->d_err + CNT_ERR, 0, sizeof(int32_t), c->stream));
    {
        LayerWrite w(c, *v);
        // aovs.overwrite (engine.cpp:89-91) copies every layer of the viewport's
        // sample AOVs; trace_viewport writes no depth, so DEPTH becomes 0
        if (v->layer[NR_LAYER_DEPTH])
            HIPCK(hipMemsetAsync(v->layer[NR_LAYER_DEPTH], 0, sizeof(float) * (size_t)v->H * v->W, c->stream));
        if (c->obj_types & NR_TY_OTHER)
            hipLaunchKernelGGL(k_viewport<NR_TY_ALL>, dim3(nb), dim3(128), trav_lds_bytes(), c->stream, c->d_scene,
                               v->cam, show_axis_grid != 0, lb, ld, ln, c->d_err, SC);
        else
            hipLaunchKernelGGL(k_viewport<NR_TY_QUAD | NR_TY_SPHERE | NR_TY_MESH>, dim3(nb), dim3(128),
                               trav_lds_bytes(), c->stream, c->d_scene, v->cam, show_axis_grid != 0, lb, ld, ln,
                               c->d_err, SC);
        HIPCK(hipGetLastError());
    }
    return check_overflow(c);
}

int nr_pick(nr_ctx* c, int view, float su, float sv, nr_hit* out) {
    View* v = nullptr;
    if (int e = view_of(c, view, &v)) return e;
    if (!out) return fail(NR_ERR_INVALID, "null output");
    if (!c->has_scene) return fail(NR_ERR_STATE, "nr_pick before nr_scene_upload");
    HIPCK(hipSetDevice(c->device));
    StackCfg SC;
    if (int e = stack_spill(c, 64u, SC)) return e;
    nr_hit* dh = nullptr;
    HIPCK(hipMalloc((void**)&dh, sizeof(nr_hit)));
    int rc = NR_OK;
    if (hipMemsetAsync(c->d_err + CNT_ERR, 0, sizeof(int32_t), c->stream) != hipSuccess) rc = fail(NR_ERR_HIP, "memset");
    if (rc == NR_OK) {
        if (c->obj_types & NR_TY_OTHER)
            hipLaunchKernelGGL(k_pick<NR_TY_ALL>, dim3(1), dim3(64), trav_lds_bytes(), c->stream, c->d_scene, v->cam,
                               su, sv, dh, c->d_err, SC);
        else
            hipLaunchKernelGGL(k_pick<NR_TY_QUAD | NR_TY_SPHERE | NR_TY_MESH>, dim3(1), dim3(64), trav_lds_bytes(),
                               c->stream, c->d_scene, v->cam, su, sv, dh, c->d_err, SC);
        if (hipGetLastError() != hipSuccess) rc = fail(NR_ERR_HIP, "k_pick launch");
    }
    if (rc == NR_OK && hipMemcpyAsync(out, dh, sizeof(nr_hit), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        rc = fail(NR_ERR_HIP, "copy hit");
    if (rc == NR_OK) rc = check_overflow(c);
    (void)hipFree(dh);
    return rc;
}

int nr_selection_mask(nr_ctx* c, int view, int32_t object_id, int32_t outline_radius, uint8_t* d_mask) {
    View* v = nullptr;
    if (int e = view_of(c, view, &v)) return e;
    if (!d_mask) return fail(NR_ERR_INVALID, "null mask");
    if (outline_radius < 0 || outline_radius > 64) return fail(NR_ERR_INVALID, "outline radius outside 0..64");
    if (!c->has_scene) return fail(NR_ERR_STATE, "nr_selection_mask before nr_scene_upload");
    HIPCK(hipSetDevice(c->device));
    const size_t HW = (size_t)v->H * v->W;
    if (v->sel_cap < HW) {
        if (v->d_sel_base) HIPCK(hipFree(v->d_sel_base));
        v->d_sel_base = nullptr;
        v->sel_cap = 0;
        HIPCK(hipMalloc((void**)&v->d_sel_base, HW));
        v->sel_cap = HW;
    }
    const uint32_t nb = tile_blocks(*v);
    StackCfg SC;
    if (int e = stack_spill(c, nb * 128u, SC)) return e;
    HIPCK(hipMemsetAsync(c->d_err + CNT_ERR, 0, sizeof(int32_t), c->stream));
    if (c->obj_types & NR_TY_OTHER)
        hipLaunchKernelGGL(k_select_mask<NR_TY_ALL>, dim3(nb), dim3(128), trav_lds_bytes(), c->stream, c->d_scene,
                           v->cam, object_id, v->d_sel_base, c->d_err, SC);
    else
        hipLaunchKernelGGL(k_select_mask<NR_TY_QUAD | NR_TY_SPHERE | NR_TY_MESH>, dim3(nb), dim3(128),
                           trav_lds_bytes(), c->stream, c->d_scene, v->cam, object_id, v->d_sel_base, c->d_err, SC);
    hipLaunchKernelGGL(k_outline, dim3(nb), dim3(128), 0, c->stream, v->d_sel_base, d_mask, v->H, v->W,
                       (int)outline_radius);
    HIPCK(hipGetLastError());
    return check_overflow(c);
}

uint8_t* nr_view_mask_buffer(nr_ctx* c, int view) {
    View* v = nullptr;
    if (view_of(c, view, &v) != NR_OK) return nullptr;
    const size_t HW = (size_t)v->H * v->W;
    if (v->mask_cap < HW) {
        if (hipSetDevice(c->device) != hipSuccess) { fail(NR_ERR_HIP, "hipSetDevice"); return nullptr; }
        if (v->d_mask) (void)hipFree(v->d_mask);
        v->d_mask = nullptr;
        v->mask_cap = 0;
        if (hipMalloc((void**)&v->d_mask, HW) != hipSuccess) { fail(NR_ERR_HIP, "hipMalloc mask"); return nullptr; }
        if (hipMemset(v->d_mask, 0, HW) != hipSuccess) { fail(NR_ERR_HIP, "hipMemset mask"); return nullptr; }
        v->mask_cap = HW;
    }
    return v->d_mask;
}

// composite_overlay (stream.cu:214-222): colour channel -> uint8 by truncation
static void overlay_rgb(const float* color, uint32_t* rgb) {
    for (int k = 0; k < 3; k++) rgb[k] = (uint32_t)(uint8_t)(fminf(1.0f, fmaxf(0.0f, color[k])) * 255.0f);
}

int nr_overlay_device(nr_ctx* c, uint8_t* d_img, int32_t H, int32_t W, const uint8_t* d_mask, const float color[3]) {
    if (!c || !d_img || !d_mask || !color) return fail(NR_ERR_INVALID, "null argument");
    if (H <= 0 || W <= 0 || (uint64_t)H * W > (1ull << 31)) return fail(NR_ERR_INVALID, "bad image size");
    HIPCK(hipSetDevice(c->device));
    uint32_t rgb[3];
    overlay_rgb(color, rgb);
    const size_t HW = (size_t)H * W;
    hipLaunchKernelGGL(k_overlay, dim3((uint32_t)((HW + 255) / 256)), dim3(256), 0, c->stream, d_img, d_mask, HW,
                       rgb[0], rgb[1], rgb[2]);
    HIPCK(hipGetLastError());
    HIPCK(hipStreamSynchronize(c->stream));
    return NR_OK;
}

int nr_readback(nr_ctx* c, int view, const nr_readback_desc* d, uint8_t* host) {
    View* v = nullptr;
    if (int e = view_of(c, view, &v)) return e;
    if (!d || !host) return fail(NR_ERR_INVALID, "null argument");
    const int layer = d->post.layer;
    if (layer < 0 || layer > 8 || !v->layer[layer]) return fail(NR_ERR_INVALID, "readback: layer not allocated");
    if (layer_channels(layer) != 3) return fail(NR_ERR_INVALID, "readback: the layer has one channel");
    if (d->denoiser && (!v->layer[NR_LAYER_DIFFUSE] || !v->layer[NR_LAYER_WORLD_NORMAL]))
        return fail(NR_ERR_STATE, "denoising needs the DIFFUSE and WORLD_NORMAL layers");
    std::lock_guard<std::mutex> one(v->rb_mu);
    HIPCK(hipSetDevice(c->device));
    const size_t HW = (size_t)v->H * v->W;
    if (v->rb_cap < HW) {
        if (v->d_rb_u8) HIPCK(hipFree(v->d_rb_u8));
        if (v->h_rb) HIPCK(hipHostFree(v->h_rb));
        v->d_rb_u8 = nullptr; v->h_rb = nullptr; v->rb_cap = 0;
        HIPCK(hipMalloc((void**)&v->d_rb_u8, 3 * HW));
        HIPCK(hipHostMalloc((void**)&v->h_rb, 3 * HW, hipHostMallocDefault));
        v->rb_cap = HW;
    }
    if (d->denoiser && v->den_cap < HW) {
        if (v->d_rb_den) HIPCK(hipFree(v->d_rb_den));
        v->d_rb_den = nullptr; v->den_cap = 0;
        HIPCK(hipMalloc((void**)&v->d_rb_den, 3 * HW * sizeof(float)));
        v->den_cap = HW;
    }
    const PostArgs A = post_args(&d->post);
    {
        std::lock_guard<std::mutex> lk(v->layers_mu);   // no layer write is enqueued until the reads are
        const float* src = v->layer[layer];
        if (d->denoiser) {
            // the U-Net runs on its own stream and returns when done
            HIPCK(hipEventSynchronize(c->ev_layers));
            int rc = nr_unet_forward(static_cast<nr_unet*>(d->denoiser), src, v->layer[NR_LAYER_DIFFUSE],
                                     v->layer[NR_LAYER_WORLD_NORMAL], v->layer[NR_LAYER_DEPTH], 1, (int32_t)v->H,
                                     (int32_t)v->W, v->d_rb_den);
            if (rc != NR_OK) return rc;
            (void)hipSetDevice(c->device);
            src = v->d_rb_den;
        } else {
            HIPCK(hipStreamWaitEvent(c->rb_stream, c->ev_layers, 0));
        }
        hipLaunchKernelGGL(k_post, dim3((uint32_t)((HW + 255) / 256)), dim3(256), 0, c->rb_stream, src, (uint32_t)HW, A,
                           v->d_rb_u8, (float*)nullptr);
        HIPCK(hipGetLastError());
        HIPCK(hipEventRecord(c->ev_rb_done, c->rb_stream));
    }
    if (d->d_overlay) {
        uint32_t rgb[3];
        overlay_rgb(d->overlay_color, rgb);
        hipLaunchKernelGGL(k_overlay, dim3((uint32_t)((HW + 255) / 256)), dim3(256), 0, c->rb_stream, v->d_rb_u8,
                           d->d_overlay, HW, rgb[0], rgb[1], rgb[2]);
        HIPCK(hipGetLastError());
    }
    HIPCK(hipMemcpyAsync(v->h_rb, v->d_rb_u8, 3 * HW, hipMemcpyDeviceToHost, c->rb_stream));
    HIPCK(hipStreamSynchronize(c->rb_stream));
    std::memcpy(host, v->h_rb, 3 * HW);
    return NR_OK;
}

// ---------------------------------------------------------------------------
// scene trees and live edits
// ---------------------------------------------------------------------------
int nr_set_render_streams(nr_ctx* c, int n) {
    if (!c) return fail(NR_ERR_INVALID, "null context");
    if (n < 1 || n > NR_MAX_GROUPS) return fail(NR_ERR_INVALID, "render streams outside 1..4");
    c->render_streams = n;
    return NR_OK;
}

int nr_set_tuning(nr_ctx* c, const nr_tuning* t) {
    if (!c || !t) return fail(NR_ERR_INVALID, "null argument");
    if (t->slots < 1 || t->slots > NR_SLOTS_MAX) return fail(NR_ERR_INVALID, "slots outside 1..64");
    if (t->stream_groups < 1 || t->stream_groups > NR_MAX_GROUPS) return fail(NR_ERR_INVALID, "stream groups outside 1..4");
    if (t->shade_wg_per_cu < 1 || t->shade_wg_per_cu > 32) return fail(NR_ERR_INVALID, "shade workgroups per CU outside 1..32");
    if (t->tail_paths < 0) return fail(NR_ERR_INVALID, "negative tail_paths");
    if (t->graph_iterations != 0 && (t->graph_iterations < 2 || t->graph_iterations > 16 || (t->graph_iterations & 1)))
        return fail(NR_ERR_INVALID, "graph_iterations must be 0 or even in 2..16");
    HIPCK(hipSetDevice(c->device));
    HIPCK(hipStreamSynchronize(c->stream));                 // path state may be reallocated by the next render
    c->slots = t->slots;
    c->render_streams = t->stream_groups;
    c->shade_grid = c->n_cu * t->shade_wg_per_cu;
    c->tail_paths = t->tail_paths;
    c->graph_iters = t->graph_iterations;
    return NR_OK;
}

int nr_get_tuning(nr_ctx* c, nr_tuning* t) {
    if (!c || !t) return fail(NR_ERR_INVALID, "null argument");
    t->slots = c->slots;
    t->stream_groups = c->render_streams;
    t->shade_wg_per_cu = c->shade_grid / std::max(1, c->n_cu);
    t->tail_paths = c->tail_paths;
    t->graph_iterations = c->graph_iters;
    return NR_OK;
}

// Coordinate descent over the tunings, timing `desc` on `view` (wall clock
// around the render and its completion; the first pass is rendered once
// untimed after each change so allocations stay outside the timing).
int nr_autotune(nr_ctx* c, int view, const nr_render_desc* d, nr_tuning* out) {
    if (!c || !d) return fail(NR_ERR_INVALID, "null argument");
    if (d->sample_end <= d->sample_begin) return fail(NR_ERR_INVALID, "autotune needs at least one pass");
    nr_tuning best;
    if (int e = nr_get_tuning(c, &best)) return e;
    auto measure = [&](const nr_tuning& t, double& ms) -> int {
        if (int e = nr_set_tuning(c, &t)) return e;
        nr_render_desc w = *d;
        w.sample_end = w.sample_begin + 1;
        if (int e = nr_aov_reset(c, view)) return e;
        if (int e = nr_render(c, view, &w)) return e;
        HIPCK(hipStreamSynchronize(c->stream));
        ms = 1e30;
        for (int rep = 0; rep < 2; rep++) {      // the faster of two timings
            if (int e = nr_aov_reset(c, view)) return e;
            HIPCK(hipStreamSynchronize(c->stream));
            auto t0 = std::chrono::steady_clock::now();
            if (int e = nr_render(c, view, d)) return e;
            HIPCK(hipStreamSynchronize(c->stream));
            ms = std::min(ms, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        }
        return NR_OK;
    };
    double best_ms = 0.0;
    if (int e = measure(best, best_ms)) return e;
    // device memory one path slot per pixel of the view takes: path state (origin /
    // direction, throughput / radiance / albedo / normal, hit, shadow ray, state byte),
    // its queue entries and its traversal-stack spill columns
    size_t dev_free = 0, dev_total = 0;
    HIPCK(hipMemGetInfo(&dev_free, &dev_total));
    const View& vw = c->views[view];
    const double npad = (double)((vw.W + 7) / 8) * (double)((vw.H + 7) / 8) * 64.0;
    const int lv = std::max(0, c->scene.stack_depth - LDS_STACK_LEVELS);
    const double slot_bytes =
        npad * (double)(2 * 16 + 4 * 16 + sizeof(GHit) + NR_SH_STRIDE * 16 + 1 + (2 + NR_MAT_TYPES + 1) * 4 + 8 * lv);
    const int slot_c[] = {1, 2, 4, 8, 16, 32, 48, 64}, group_c[] = {1, 2, 3, 4}, shade_c[] = {4, 8, 16},
              tail_c[] = {0, 65536, 262144, 1048576, 2097152, 4194304}, graph_c[] = {0, 2, 8};
    const int* cands[5] = {slot_c, group_c, shade_c, tail_c, graph_c};
    const int n_cand[5] = {8, 4, 3, 6, 3};
    for (int axis = 0; axis < 5; axis++) {
        const int* cand = cands[axis];
        const int n = n_cand[axis];
        nr_tuning base = best;
        for (int k = 0; k < n; k++) {
            nr_tuning t = base;
            int32_t* fields[5] = {&t.slots, &t.stream_groups, &t.shade_wg_per_cu, &t.tail_paths, &t.graph_iterations};
            int32_t& field = *fields[axis];
            if (field == cand[k]) continue;
            field = cand[k];
            if (axis == 0 && (double)t.slots * slot_bytes > 0.6 * (double)dev_total) continue;   // too big a view
            double ms = 0.0;
            if (int e = measure(t, ms)) return e;
            if (ms < 0.99 * best_ms) { best_ms = ms; best = t; }   // 1%: above the timings' noise
        }
    }
    if (int e = nr_set_tuning(c, &best)) return e;
    if (int e = nr_aov_reset(c, view)) return e;
    HIPCK(hipStreamSynchronize(c->stream));
    if (out) *out = best;
    return NR_OK;
}

int nr_set_build_mode(nr_ctx* c, int mode) {
    if (!c) return fail(NR_ERR_INVALID, "null context");
    if (mode < NR_BUILD_HOST || mode > NR_BUILD_AUTO) return fail(NR_ERR_INVALID, "bad build mode");
    c->build_mode = mode;
    return NR_OK;
}

int nr_mesh_tree(nr_ctx* c, int mesh, void* nodes64, int32_t* slot_orig, int32_t* n_nodes, int32_t* n_tris,
                 int32_t* depth) {
    if (!c) return fail(NR_ERR_INVALID, "null context");
    if (!c->has_scene) return fail(NR_ERR_STATE, "no scene");
    if (mesh < 0 || mesh >= (int)c->h_meshes.size()) return fail(NR_ERR_INVALID, "mesh index out of range");
    const GMesh& m = c->h_meshes[mesh];
    const int32_t nn = c->h_mesh_nodes[mesh];
    const int32_t nt = nn + 1;
    if (n_nodes) *n_nodes = nn;
    if (n_tris) *n_tris = nt;
    if (depth) *depth = m.depth;
    HIPCK(hipSetDevice(c->device));
    HIPCK(hipStreamSynchronize(c->stream));
    if (nodes64 && nn > 0) HIPCK(hipMemcpy(nodes64, m.nodes, sizeof(GNode) * (size_t)nn, hipMemcpyDeviceToHost));
    if (slot_orig) {
        std::vector<GTri> t(nt);
        HIPCK(hipMemcpy(t.data(), m.tris, sizeof(GTri) * (size_t)nt, hipMemcpyDeviceToHost));
        for (int32_t k = 0; k < nt; k++) slot_orig[k] = t[k].orig;
    }
    return NR_OK;
}

int nr_object_set_transform(nr_ctx* c, int32_t index, const float position[3], const float rotation[3],
                            const float scale[3]) {
    if (!c || !position || !rotation || !scale) return fail(NR_ERR_INVALID, "null argument");
    if (!c->has_scene) return fail(NR_ERR_STATE, "no scene");
    const int32_t n = (int32_t)c->reg_objects.size();
    if (index < 0 || index >= n) return fail(NR_ERR_INVALID, "object index out of range");
    HIPCK(hipSetDevice(c->device));
    GObject& o = c->reg_objects[index];
    const int geom = (o.type == NR_HIT_CONSTANT_MEDIUM || o.type == NR_HIT_OBJECT_LIGHT) ? o.bound_type : o.type;
    o.xf = host::xform(position, rotation, scale, geom == NR_HIT_SPHERE);
    // same object count as the tables hold: rewritten in place once the queued work is done
    HIPCK(hipStreamSynchronize(c->stream));
    GScene S = c->scene;
    if (int e = upload_top(c, S, true)) return e;
    return publish_scene(c, S);
}

int nr_sync(nr_ctx* c) {
    if (!c) return fail(NR_ERR_INVALID, "null context");
    HIPCK(hipSetDevice(c->device));
    HIPCK(hipStreamSynchronize(c->stream));
    if (c->P.err) {                                     // the last render's traversal-overflow flag
        int32_t err = 0;
        HIPCK(hipMemcpy(&err, c->P.err, sizeof(int32_t), hipMemcpyDeviceToHost));
        if (err) return fail(NR_ERR_LIMIT, "traversal stack overflow during render");
    }
    return NR_OK;
}

int nr_last_timings(nr_ctx* c, float* ms, int n) {
    if (!c || !ms) return 0;
    int k = std::min(n, c->n_last);
    for (int i = 0; i < k; i++) ms[i] = c->last_ms[i];
    return k;
}

int nr_set_timing(nr_ctx* c, int enabled) {
    if (!c) return fail(NR_ERR_INVALID, "null context");
    c->timing = enabled != 0;
    return NR_OK;
}

}  // extern "C"
