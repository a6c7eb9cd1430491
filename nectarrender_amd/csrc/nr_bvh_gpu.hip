// nr_bvh_gpu.hip — binned-SAH mesh trees built on gfx950 (see nr_bvh_gpu.h).
//
// Large nodes (more than SMALL items) are split a level at a time: one pass
// over the items reduces each node's centroid bounds and box, one pass fills
// its 3 x 16 bins (counts + boxes), one thread per node sweeps the bins exactly
// as the host builder does, and a flag scan scatters every item to the left or
// right child range.  Min/max/count reductions use integer atomics on
// order-preserving float keys, so they are exact and order-independent.  A
// workgroup whose items all belong to one node reduces in LDS first, so the
// top levels (one node, millions of items) issue one global atomic per value
// per workgroup.  Nodes of 2..SMALL items go to a list that one-wave workgroups
// finish depth-first in LDS.
//
// Item order: the host partitions with std::partition, which for vector
// iterators swaps the j-th misplaced item from the left with the j-th misplaced
// item from the right; the scatter reproduces that permutation (a misplaced
// item finds its partner by binary search on the flag prefix sums), so a node's
// items arrive in the host's order and median-split ties (std::sort, stable
// insertion sort for up to 16 items) resolve the same way.
//
// Output placement: a node over item range [s, e) has preorder index k (root 0,
// left child k + 1, right child k + 2 (m - s)) and child-pair index k - s (the
// s leaves before it in preorder are the items left of s); a leaf's slot is
// its position.  This is exactly where nr_bvh.h's recursion puts them.
#include "nr_bvh_gpu.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "../../include/nectar_hip.h"

namespace nrg {
namespace {

constexpr int BINS = 16;
constexpr int SMALL = 512;               // nodes up to this many items: one wave in LDS
constexpr int FALLBACK_CAP = 65536;      // median-split clusters larger than this: refuse
constexpr int ACC = 12;                  // centroid min[3] max[3], box min[3] max[3]
constexpr int BIN_VALS = 7;              // count, box min[3], box max[3]
constexpr int BIN_TOTAL = 3 * BINS * BIN_VALS;

// order-preserving float <-> uint32 keys (every float ordered, -0 below +0)
__host__ __device__ inline uint32_t fkey(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float fdec(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// the host builder's bin of a centroid (nr_bvh.h Builder::bin)
__device__ inline int32_t bin_of(float c, float c0, float ext) {
    int32_t k = (int32_t)(BINS * (c - c0) / ext);
    return k < 0 ? 0 : (k >= BINS ? BINS - 1 : k);
}

// Destination of position i in std::partition's swap order over [s, e) with
// m = s + (true count); pre(x) = number of true flags in [s, x) for x in [s, e].
template <typename Pre>
__device__ inline int32_t partition_dest(int32_t i, bool f, int32_t s, int32_t m, int32_t e, const Pre& pre) {
    const int32_t k = pre(e) - pre(m);              // misplaced trues = misplaced falses
    if (i < m) {
        if (f) return i;
        const int32_t j = (i - s) - pre(i);          // falses in [s, i)
        const int32_t want = k - 1 - j;              // partner: the want-th true in [m, e)
        int32_t lo = m, hi = e - 1;
        while (lo < hi) {
            const int32_t mid = (lo + hi) >> 1;
            if (pre(mid + 1) - pre(m) >= want + 1) hi = mid; else lo = mid + 1;
        }
        return lo;
    }
    if (!f) return i;
    const int32_t r = pre(i) - pre(m);               // trues in [m, i)
    const int32_t want = k - 1 - r;                  // partner: the want-th false in [s, m)
    int32_t lo = s, hi = m - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if ((mid + 1 - s) - pre(mid + 1) >= want + 1) hi = mid; else lo = mid + 1;
    }
    return lo;
}

struct Box3 { float lo[3], hi[3]; };
__device__ inline Box3 box_none() {
    Box3 r;
    for (int a = 0; a < 3; a++) { r.lo[a] = 1e32f; r.hi[a] = -1e32f; }
    return r;
}
__device__ inline Box3 box_join(const Box3& x, const Box3& y) {        // aabb.h:25-29
    Box3 r;
    for (int a = 0; a < 3; a++) {
        r.lo[a] = x.lo[a] <= y.lo[a] ? x.lo[a] : y.lo[a];
        r.hi[a] = x.hi[a] >= y.hi[a] ? x.hi[a] : y.hi[a];
    }
    return r;
}
__device__ inline float box_area(const Box3& b) {
    float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// positions (SoA, double-buffered): item box and item index
struct Pos {
    float* b[6];
    int32_t* item;
    int32_t* task;           // current large task at this position, -1 = settled
};

struct Tasks {               // one level's large nodes
    int32_t *s, *e, *k, *depth, *parent, *side;
};

struct Split {               // per large node of the level
    uint32_t* acc;           // [T][ACC] keys
    uint32_t* bins;          // [T][BIN_TOTAL]
    int32_t* axis;           // split axis
    int32_t* kb;             // last left bin, -1 = median fallback
    float* c0;               // [T][3]
    float* ext;              // [T][3]
    int32_t* m;              // split position
    int32_t* child;          // [T][2] next-level task index or -1
};

struct Out {
    GNode* nodes;
    float* root;             // [6] root box
    int32_t* depth;
    int32_t* counts;         // [0] next-level large nodes, [1] small nodes, [2] refusals
};

struct SmallList {
    int32_t *s, *e, *k, *depth, *parent, *side;
};

__device__ inline void write_child_box(GNode* nodes, float* root, int32_t parent, int32_t side, const Box3& b) {
    if (parent < 0) {
        for (int a = 0; a < 3; a++) { root[a] = b.lo[a]; root[3 + a] = b.hi[a]; }
        return;
    }
    GNode& g = nodes[parent];
    for (int a = 0; a < 3; a++) {
        if (side == 0) { g.lmin[a] = b.lo[a]; g.lmax[a] = b.hi[a]; }
        else { g.rmin[a] = b.lo[a]; g.rmax[a] = b.hi[a]; }
    }
}

__device__ inline Box3 pos_box(const Pos& P, int32_t i) {
    Box3 b;
    for (int a = 0; a < 3; a++) { b.lo[a] = P.b[a][i]; b.hi[a] = P.b[3 + a][i]; }
    return b;
}

// ---------------------------------------------------------------------------
// level kernels
// ---------------------------------------------------------------------------
__global__ void k_acc_init(int32_t T, uint32_t* acc, uint32_t* bins) {
    const int32_t t = blockIdx.x;
    if (t >= T) return;
    for (int v = threadIdx.x; v < ACC; v += blockDim.x) acc[(size_t)t * ACC + v] = (v % 6) < 3 ? 0xFFFFFFFFu : 0u;
    for (int v = threadIdx.x; v < BIN_TOTAL; v += blockDim.x) {
        const int w = v % BIN_VALS;
        bins[(size_t)t * BIN_TOTAL + v] = w == 0 ? 0u : (w <= 3 ? 0xFFFFFFFFu : 0u);
    }
}

// centroid bounds and box of every large node
__global__ void __launch_bounds__(256) k_bounds(int32_t n, Pos P, uint32_t* acc) {
    __shared__ uint32_t red[ACC];
    __shared__ int32_t blk_task;
    const int32_t i = blockIdx.x * 256 + threadIdx.x;
    const int32_t t = i < n ? P.task[i] : -1;
    if (threadIdx.x == 0) blk_task = t;
    if (threadIdx.x < ACC) red[threadIdx.x] = (threadIdx.x % 6) < 3 ? 0xFFFFFFFFu : 0u;
    __syncthreads();
    const bool uniform = __syncthreads_and(t == blk_task && t >= 0);
    uint32_t v[ACC];
    if (t >= 0) {
        for (int a = 0; a < 3; a++) {
            const float lo = P.b[a][i], hi = P.b[3 + a][i];
            const uint32_t c = fkey((lo + hi) * 0.5f);
            v[a] = c; v[3 + a] = c;
            v[6 + a] = fkey(lo); v[9 + a] = fkey(hi);
        }
    }
    if (uniform) {
        for (int q = 0; q < ACC; q++) {
            if ((q % 6) < 3) atomicMin(&red[q], v[q]);
            else atomicMax(&red[q], v[q]);
        }
        __syncthreads();
        if (threadIdx.x < ACC) {
            uint32_t* dst = &acc[(size_t)blk_task * ACC + threadIdx.x];
            if ((threadIdx.x % 6) < 3) atomicMin(dst, red[threadIdx.x]);
            else atomicMax(dst, red[threadIdx.x]);
        }
    } else if (t >= 0) {
        for (int q = 0; q < ACC; q++) {
            uint32_t* dst = &acc[(size_t)t * ACC + q];
            if ((q % 6) < 3) atomicMin(dst, v[q]);
            else atomicMax(dst, v[q]);
        }
    }
}

// decode the centroid bounds; hand the node box to the parent (or the root)
__global__ void k_node_box(int32_t T, Tasks L, const uint32_t* acc, float* c0, float* ext, Out O) {
    const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    const uint32_t* A = acc + (size_t)t * ACC;
    Box3 b;
    for (int a = 0; a < 3; a++) {
        const float cl = fdec(A[a]), ch = fdec(A[3 + a]);
        c0[t * 3 + a] = cl;
        ext[t * 3 + a] = ch - cl;
        b.lo[a] = fdec(A[6 + a]);
        b.hi[a] = fdec(A[9 + a]);
    }
    write_child_box(O.nodes, O.root, L.parent[t], L.side[t], b);
}

// one item into bin k of axis a (LDS or global bins)
__device__ inline void bin_item(uint32_t* bins, int a, int32_t k, const float* lo, const float* hi) {
    uint32_t* B = bins + (a * BINS + k) * BIN_VALS;
    atomicAdd(&B[0], 1u);
    for (int q = 0; q < 3; q++) { atomicMin(&B[1 + q], fkey(lo[q])); atomicMax(&B[4 + q], fkey(hi[q])); }
}

__global__ void __launch_bounds__(256) k_bins(int32_t n, Pos P, const float* c0, const float* ext, uint32_t* bins) {
    __shared__ uint32_t sb[BIN_TOTAL];
    __shared__ int32_t blk_task;
    const int32_t i = blockIdx.x * 256 + threadIdx.x;
    const int32_t t = i < n ? P.task[i] : -1;
    if (threadIdx.x == 0) blk_task = t;
    for (int v = threadIdx.x; v < BIN_TOTAL; v += 256) {
        const int w = v % BIN_VALS;
        sb[v] = w == 0 ? 0u : (w <= 3 ? 0xFFFFFFFFu : 0u);
    }
    __syncthreads();
    const bool uniform = __syncthreads_and(t == blk_task && t >= 0);
    if (t >= 0) {
        float lo[3], hi[3];
        for (int a = 0; a < 3; a++) { lo[a] = P.b[a][i]; hi[a] = P.b[3 + a][i]; }
        for (int a = 0; a < 3; a++) {
            const float e = ext[t * 3 + a];
            if (e < 1e-4f) continue;
            const int32_t k = bin_of((lo[a] + hi[a]) * 0.5f, c0[t * 3 + a], e);
            bin_item(uniform ? sb : bins + (size_t)t * BIN_TOTAL, a, k, lo, hi);
        }
    }
    if (uniform) {
        __syncthreads();
        uint32_t* G = bins + (size_t)blk_task * BIN_TOTAL;
        for (int v = threadIdx.x; v < BIN_TOTAL; v += 256) {
            const int w = v % BIN_VALS;
            const uint32_t x = sb[v];
            if (w == 0) { if (x) atomicAdd(&G[v], x); }
            else if (w <= 3) { if (x != 0xFFFFFFFFu) atomicMin(&G[v], x); }
            else if (x != 0u) atomicMax(&G[v], x);
        }
    }
}

// the host builder's split decision from a node's bins (nr_bvh.h Builder::split)
__device__ inline void choose(const uint32_t* bins, const float* ext, const Box3& all, int32_t s, int32_t e,
                              int32_t& axis, int32_t& kb, int32_t& m) {
    float best = 1e32f;
    int best_axis = -1, best_bin = -1;
    int32_t best_left = 0;
    for (int a = 0; a < 3; a++) {
        if (ext[a] < 1e-4f) continue;
        Box3 bb[BINS];
        int32_t bn[BINS];
        for (int k = 0; k < BINS; k++) {
            const uint32_t* B = bins + (a * BINS + k) * BIN_VALS;
            bn[k] = (int32_t)B[0];
            for (int q = 0; q < 3; q++) { bb[k].lo[q] = fdec(B[1 + q]); bb[k].hi[q] = fdec(B[4 + q]); }
        }
        Box3 lb[BINS], rb[BINS];
        int32_t lc[BINS], rc[BINS];
        Box3 run = box_none();
        int32_t cnt = 0;
        bool on = false;
        for (int k = 0; k < BINS; k++) {
            if (bn[k] > 0) { run = on ? box_join(run, bb[k]) : bb[k]; on = true; }
            cnt += bn[k]; lb[k] = run; lc[k] = cnt;
        }
        cnt = 0; on = false;
        for (int k = BINS - 1; k >= 0; k--) {
            if (bn[k] > 0) { run = on ? box_join(run, bb[k]) : bb[k]; on = true; }
            cnt += bn[k]; rb[k] = run; rc[k] = cnt;
        }
        for (int k = 0; k < BINS - 1; k++) {
            if (lc[k] == 0 || rc[k + 1] == 0) continue;
            float cost = box_area(lb[k]) * lc[k] + box_area(rb[k + 1]) * rc[k + 1];
            if (cost < best) { best = cost; best_axis = a; best_bin = k; best_left = lc[k]; }
        }
    }
    if (best_axis == -1) {
        float lx = all.hi[0] - all.lo[0], ly = all.hi[1] - all.lo[1], lz = all.hi[2] - all.lo[2];
        best_axis = (lx > ly && lx > lz) ? 0 : (ly > lz ? 1 : 2);
    }
    axis = best_axis;
    kb = best_bin;
    m = best_bin == -1 ? s + (e - s) / 2 : s + best_left;
}

__global__ void k_split(int32_t T, Tasks L, Split S, Tasks N, SmallList SL, Out O) {
    const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    const int32_t s = L.s[t], e = L.e[t], k = L.k[t], d = L.depth[t];
    const uint32_t* A = S.acc + (size_t)t * ACC;
    Box3 all;
    for (int a = 0; a < 3; a++) { all.lo[a] = fdec(A[6 + a]); all.hi[a] = fdec(A[9 + a]); }
    int32_t axis, kb, m;
    choose(S.bins + (size_t)t * BIN_TOTAL, S.ext + t * 3, all, s, e, axis, kb, m);
    if (kb == -1 && e - s > FALLBACK_CAP) atomicAdd(&O.counts[2], 1);
    S.axis[t] = axis; S.kb[t] = kb; S.m[t] = m;
    const int32_t p = k - s;
    const int32_t cs[2] = {s, m}, ce[2] = {m, e}, ck[2] = {k + 1, k + 2 * (m - s)};
    GNode& g = O.nodes[p];
    for (int side = 0; side < 2; side++) {
        const int32_t size = ce[side] - cs[side];
        int32_t ref;
        S.child[t * 2 + side] = -1;
        if (size == 1) {
            ref = ~cs[side];
            atomicMax(O.depth, d + 1);
        } else {
            ref = ck[side] - cs[side];
            if (size <= SMALL) {
                const int32_t j = atomicAdd(&O.counts[1], 1);
                SL.s[j] = cs[side]; SL.e[j] = ce[side]; SL.k[j] = ck[side]; SL.depth[j] = d + 1;
                SL.parent[j] = p; SL.side[j] = side;
            } else {
                const int32_t j = atomicAdd(&O.counts[0], 1);
                N.s[j] = cs[side]; N.e[j] = ce[side]; N.k[j] = ck[side]; N.depth[j] = d + 1;
                N.parent[j] = p; N.side[j] = side;
                S.child[t * 2 + side] = j;
            }
        }
        if (side == 0) g.left = ref; else g.right = ref;
    }
    g.pad0 = 0;
    g.pad1 = 0;
}

// left flags (binned split) or the final position (median fallback: rank by
// (box low bound on the axis, item index))
__global__ void __launch_bounds__(256) k_flags(int32_t n, Pos P, Tasks L, Split S, int32_t* flag, int32_t* fbpos) {
    const int32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int32_t t = P.task[i];
    int32_t f = 0;
    if (t >= 0) {
        const int32_t a = S.axis[t];
        if (S.kb[t] >= 0) {
            const float c = (P.b[a][i] + P.b[3 + a][i]) * 0.5f;
            f = bin_of(c, S.c0[t * 3 + a], S.ext[t * 3 + a]) <= S.kb[t] ? 1 : 0;
        } else if (L.e[t] - L.s[t] > FALLBACK_CAP) {
            fbpos[i] = i;                      // refused (the build stops after this level): stay put
        } else {
            const int32_t s = L.s[t], e = L.e[t];
            const float key = P.b[a][i];
            int32_t rank = 0;                  // stable sort by the low bound (ties keep their order)
            for (int32_t j = s; j < e; j++) {
                const float kj = P.b[a][j];
                rank += (kj < key || (kj == key && j < i)) ? 1 : 0;
            }
            fbpos[i] = s + rank;
            f = s + rank < S.m[t] ? 1 : 0;
        }
    }
    flag[i] = f;
}

__global__ void __launch_bounds__(256) k_scatter(int32_t n, Pos P, Pos Q, Tasks L, Split S, const int32_t* flag,
                                                 const int32_t* excl, const int32_t* fbpos, Out O) {
    const int32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int32_t t = P.task[i];
    int32_t np = i, nt = -1;
    if (t >= 0) {
        const int32_t s = L.s[t], e = L.e[t], m = S.m[t];
        bool left;
        if (S.kb[t] >= 0) {
            const int32_t base = excl[s];
            left = flag[i] != 0;
            np = partition_dest(i, left, s, m, e, [&](int32_t x) { return excl[x] - base; });
        } else {
            np = fbpos[i];
            left = np < m;
        }
        const int side = left ? 0 : 1;
        nt = S.child[t * 2 + side];
        if ((left ? m - s : e - m) == 1) write_child_box(O.nodes, O.root, L.k[t] - s, side, pos_box(P, i));
    }
    for (int q = 0; q < 6; q++) Q.b[q][np] = P.b[q][i];
    Q.item[np] = P.item[i];
    Q.task[np] = nt;
}

// ---------------------------------------------------------------------------
// small nodes: one wave builds a whole subtree of <= SMALL items in LDS
// ---------------------------------------------------------------------------
struct SmallLds {
    float b[6][SMALL];
    int32_t item[SMALL];
    float tb[6][SMALL];
    int32_t titem[SMALL];
    int32_t stack[SMALL][6];
    int32_t pre[SMALL + 1];      // true flags before a position (local, from the node start)
    uint8_t flg[SMALL];
    uint32_t bins[BIN_TOTAL];
    float c0[3], ext[3];
    int32_t axis, kb, m;
};

__device__ inline float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ inline float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

__global__ void __launch_bounds__(64) k_small(SmallList SL, const int32_t* count, Pos P, int32_t* order, Out O) {
    __shared__ SmallLds sm;
    if ((int32_t)blockIdx.x >= *count) return;
    const int lane = threadIdx.x;
    const int32_t S0 = SL.s[blockIdx.x], E0 = SL.e[blockIdx.x];
    const int32_t n = E0 - S0;
    for (int32_t j = lane; j < n; j += 64) {
        for (int q = 0; q < 6; q++) sm.b[q][j] = P.b[q][S0 + j];
        sm.item[j] = P.item[S0 + j];
    }
    int sp = 0;
    if (lane == 0) {
        sm.stack[0][0] = 0; sm.stack[0][1] = n; sm.stack[0][2] = SL.k[blockIdx.x];
        sm.stack[0][3] = SL.depth[blockIdx.x]; sm.stack[0][4] = SL.parent[blockIdx.x]; sm.stack[0][5] = SL.side[blockIdx.x];
    }
    sp = 1;
    __syncthreads();
    while (sp > 0) {
        sp--;
        const int32_t ls = sm.stack[sp][0], le = sm.stack[sp][1], k = sm.stack[sp][2], d = sm.stack[sp][3];
        const int32_t par = sm.stack[sp][4], side = sm.stack[sp][5];
        __syncthreads();
        // bounds
        float cl[3], ch[3];
        Box3 nb;
        for (int a = 0; a < 3; a++) { cl[a] = 1e32f; ch[a] = -1e32f; nb.lo[a] = 1e32f; nb.hi[a] = -1e32f; }
        for (int32_t j = ls + lane; j < le; j += 64)
            for (int a = 0; a < 3; a++) {
                const float lo = sm.b[a][j], hi = sm.b[3 + a][j], c = (lo + hi) * 0.5f;
                cl[a] = fminf(cl[a], c); ch[a] = fmaxf(ch[a], c);
                nb.lo[a] = fminf(nb.lo[a], lo); nb.hi[a] = fmaxf(nb.hi[a], hi);
            }
        for (int a = 0; a < 3; a++) {
            cl[a] = wave_min(cl[a]); ch[a] = wave_max(ch[a]);
            nb.lo[a] = wave_min(nb.lo[a]); nb.hi[a] = wave_max(nb.hi[a]);
        }
        if (lane == 0) write_child_box(O.nodes, O.root, par, side, nb);
        // bins
        for (int v = lane; v < BIN_TOTAL; v += 64) {
            const int w = v % BIN_VALS;
            sm.bins[v] = w == 0 ? 0u : (w <= 3 ? 0xFFFFFFFFu : 0u);
        }
        __syncthreads();
        float ext[3];
        for (int a = 0; a < 3; a++) ext[a] = ch[a] - cl[a];
        for (int32_t j = ls + lane; j < le; j += 64) {
            float lo[3], hi[3];
            for (int a = 0; a < 3; a++) { lo[a] = sm.b[a][j]; hi[a] = sm.b[3 + a][j]; }
            for (int a = 0; a < 3; a++) {
                if (ext[a] < 1e-4f) continue;
                bin_item(sm.bins, a, bin_of((lo[a] + hi[a]) * 0.5f, cl[a], ext[a]), lo, hi);
            }
        }
        __syncthreads();
        if (lane == 0) {
            int32_t axis, kb, m;
            choose(sm.bins, ext, nb, ls, le, axis, kb, m);
            sm.axis = axis; sm.kb = kb; sm.m = m;
            for (int a = 0; a < 3; a++) { sm.c0[a] = cl[a]; sm.ext[a] = ext[a]; }
        }
        __syncthreads();
        const int32_t axis = sm.axis, kb = sm.kb, m = sm.m;
        // partition into the temporary arrays, in std::partition's order
        if (kb >= 0) {
            for (int32_t j = ls + lane; j < le; j += 64)
                sm.flg[j] = bin_of((sm.b[axis][j] + sm.b[3 + axis][j]) * 0.5f, sm.c0[axis], sm.ext[axis]) <= kb ? 1 : 0;
            __syncthreads();
            int32_t run = 0;
            for (int32_t base = ls; base < le; base += 64) {
                const int32_t j = base + lane;
                const bool f = j < le && sm.flg[j];
                const uint64_t mask = __ballot(f);
                const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
                if (j < le) sm.pre[j - ls] = run + __popcll(mask & below);
                run += __popcll(mask);
            }
            if (lane == 0) sm.pre[le - ls] = run;
            __syncthreads();
            for (int32_t j = ls + lane; j < le; j += 64) {
                const int32_t np = partition_dest(j, sm.flg[j] != 0, ls, m, le, [&](int32_t x) { return sm.pre[x - ls]; });
                for (int q = 0; q < 6; q++) sm.tb[q][np] = sm.b[q][j];
                sm.titem[np] = sm.item[j];
            }
        } else {
            for (int32_t j = ls + lane; j < le; j += 64) {
                const float key = sm.b[axis][j];
                int32_t rank = 0;              // stable sort by the low bound (ties keep their order)
                for (int32_t q = ls; q < le; q++) {
                    const float kq = sm.b[axis][q];
                    rank += (kq < key || (kq == key && q < j)) ? 1 : 0;
                }
                for (int q = 0; q < 6; q++) sm.tb[q][ls + rank] = sm.b[q][j];
                sm.titem[ls + rank] = sm.item[j];
            }
        }
        __syncthreads();
        for (int32_t j = ls + lane; j < le; j += 64) {
            for (int q = 0; q < 6; q++) sm.b[q][j] = sm.tb[q][j];
            sm.item[j] = sm.titem[j];
        }
        __syncthreads();
        // node record, leaf children, children to visit
        const int32_t gs = S0 + ls, gm = S0 + m;
        const int32_t p = k - gs;
        const int32_t cs[2] = {ls, m}, ce[2] = {m, le}, ck[2] = {k + 1, k + 2 * (m - ls)};
        if (lane == 0) {
            GNode& g = O.nodes[p];
            g.left = (m - ls == 1) ? ~gs : ck[0] - gs;
            g.right = (le - m == 1) ? ~gm : ck[1] - gm;
            g.pad0 = 0;
            g.pad1 = 0;
            for (int sd = 0; sd < 2; sd++) {
                if (ce[sd] - cs[sd] != 1) continue;
                Box3 b;
                for (int a = 0; a < 3; a++) { b.lo[a] = sm.b[a][cs[sd]]; b.hi[a] = sm.b[3 + a][cs[sd]]; }
                write_child_box(O.nodes, O.root, p, sd, b);
                atomicMax(O.depth, d + 1);
            }
            for (int sd = 1; sd >= 0; sd--) {
                if (ce[sd] - cs[sd] < 2) continue;
                sm.stack[sp][0] = cs[sd]; sm.stack[sp][1] = ce[sd]; sm.stack[sp][2] = ck[sd];
                sm.stack[sp][3] = d + 1; sm.stack[sp][4] = p; sm.stack[sp][5] = sd;
                sp++;
            }
        }
        sp = __shfl(sp, 0);
        __syncthreads();
    }
    for (int32_t j = lane; j < n; j += 64) order[S0 + j] = sm.item[j];
}

__global__ void k_iota_boxes(int32_t n, Pos P, const float* b0, const float* b1, const float* b2, const float* b3,
                             const float* b4, const float* b5) {
    const int32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    P.b[0][i] = b0[i]; P.b[1][i] = b1[i]; P.b[2][i] = b2[i];
    P.b[3][i] = b3[i]; P.b[4][i] = b4[i]; P.b[5][i] = b5[i];
    P.item[i] = i;
    P.task[i] = 0;
}

__global__ void k_copy_order(int32_t n, const int32_t* item, int32_t* order) {
    const int32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) order[i] = item[i];
}

__global__ void k_tri_boxes(const float* __restrict__ v, const uint32_t* __restrict__ tris, int32_t n, float* b0,
                            float* b1, float* b2, float* b3, float* b4, float* b5) {
    const int32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const float* a = v + 11 * (size_t)tris[3 * (size_t)t];
    const float* b = v + 11 * (size_t)tris[3 * (size_t)t + 1];
    const float* c = v + 11 * (size_t)tris[3 * (size_t)t + 2];
    float* lo[3] = {b0, b1, b2};
    float* hi[3] = {b3, b4, b5};
    for (int k = 0; k < 3; k++) {                                             // mesh.h:320-336, aabb.h:19-23
        const float mn = fminf(a[k], fminf(b[k], c[k]));
        const float mx = fmaxf(a[k], fmaxf(b[k], c[k]));
        lo[k][t] = mn <= mx ? mn : mx;
        hi[k][t] = mn <= mx ? mx : mn;
    }
}

__global__ void k_gather_tris(const float* __restrict__ v, const uint32_t* __restrict__ tris,
                              const int32_t* __restrict__ order, int32_t n, GTri* out, uint32_t* idx) {
    const int32_t s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    const int32_t src = order[s];
    GTri g;
    for (int k = 0; k < 3; k++) {
        const uint32_t vi = tris[3 * (size_t)src + k];
        idx[3 * (size_t)s + k] = vi;
        float* dst = k == 0 ? g.v0 : (k == 1 ? g.v1 : g.v2);
        for (int a = 0; a < 3; a++) dst[a] = v[11 * (size_t)vi + a];
    }
    g.orig = src;
    g.pad0 = 0;
    g.pad1 = 0;
    out[s] = g;
}

template <typename T>
struct DevBuf {
    T* p = nullptr;
    hipError_t alloc(size_t n) { return hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)); }
    ~DevBuf() { if (p) (void)hipFree(p); }
};

}  // namespace

#define NRG_CK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) { err = std::string(#expr) + ": " + hipGetErrorString(e_); return NR_ERR_HIP; } \
    } while (0)

void triangle_boxes(hipStream_t st, const float* d_vertices, const uint32_t* d_tris, int32_t n, float* const boxes[6]) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_tri_boxes, dim3((n + 255) / 256), dim3(256), 0, st, d_vertices, d_tris, n, boxes[0], boxes[1],
                       boxes[2], boxes[3], boxes[4], boxes[5]);
}

void gather_triangles(hipStream_t st, const float* d_vertices, const uint32_t* d_tris, const int32_t* d_order,
                      int32_t n, GTri* d_out, uint32_t* d_idx) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather_tris, dim3((n + 255) / 256), dim3(256), 0, st, d_vertices, d_tris, d_order, n, d_out,
                       d_idx);
}

int build(hipStream_t st, int32_t n, const float* const boxes[6], GNode* d_nodes, int32_t* d_order, Result& r,
          std::string& err) {
    if (n <= 0) { err = "empty tree"; return NR_ERR_INVALID; }
    // positions, double-buffered
    DevBuf<float> pb[2][6];
    DevBuf<int32_t> pitem[2], ptask[2];
    for (int b = 0; b < 2; b++) {
        for (int q = 0; q < 6; q++) NRG_CK(pb[b][q].alloc(n));
        NRG_CK(pitem[b].alloc(n));
        NRG_CK(ptask[b].alloc(n));
    }
    Pos P[2];
    for (int b = 0; b < 2; b++) {
        for (int q = 0; q < 6; q++) P[b].b[q] = pb[b][q].p;
        P[b].item = pitem[b].p;
        P[b].task = ptask[b].p;
    }
    hipLaunchKernelGGL(k_iota_boxes, dim3((n + 255) / 256), dim3(256), 0, st, n, P[0], boxes[0], boxes[1], boxes[2],
                       boxes[3], boxes[4], boxes[5]);
    DevBuf<float> root;
    DevBuf<int32_t> depth, counts;
    NRG_CK(root.alloc(6));
    NRG_CK(depth.alloc(1));
    NRG_CK(counts.alloc(4));
    // small host -> device values go through synchronous copies (the host
    // sources are locals); the stream is drained first
    NRG_CK(hipStreamSynchronize(st));
    const int32_t one = 1, zeros[4] = {0, 0, 0, 0};
    NRG_CK(hipMemcpy(depth.p, &one, sizeof(int32_t), hipMemcpyHostToDevice));
    NRG_CK(hipMemcpy(counts.p, zeros, sizeof(zeros), hipMemcpyHostToDevice));
    Out O{d_nodes, root.p, depth.p, counts.p};
    // small-node list (each entry covers >= 2 items)
    const int32_t small_cap = n / 2 + 1;
    DevBuf<int32_t> sl[6];
    for (auto& b : sl) NRG_CK(b.alloc(small_cap));
    SmallList SL{sl[0].p, sl[1].p, sl[2].p, sl[3].p, sl[4].p, sl[5].p};
    int cur = 0;
    int32_t n_small = 0;
    if (n == 1) {
        for (int q = 0; q < 6; q++) NRG_CK(hipMemcpyAsync(root.p + q, boxes[q], 4, hipMemcpyDeviceToDevice, st));
        NRG_CK(hipMemcpy(d_order, zeros, sizeof(int32_t), hipMemcpyHostToDevice));
        r.root_ref = ~0;
    } else if (n <= SMALL) {
        const int32_t rec[6] = {0, n, 0, 1, -1, 0};
        for (int q = 0; q < 6; q++) NRG_CK(hipMemcpy(sl[q].p, &rec[q], sizeof(int32_t), hipMemcpyHostToDevice));
        n_small = 1;
        r.root_ref = 0;
    } else {
        r.root_ref = 0;
        // large-node levels
        const int32_t task_cap = n / SMALL + 2;
        DevBuf<int32_t> tl[2][6];
        for (int b = 0; b < 2; b++) for (auto& x : tl[b]) NRG_CK(x.alloc(task_cap));
        Tasks L[2];
        for (int b = 0; b < 2; b++) L[b] = Tasks{tl[b][0].p, tl[b][1].p, tl[b][2].p, tl[b][3].p, tl[b][4].p, tl[b][5].p};
        DevBuf<uint32_t> acc, bins;
        DevBuf<int32_t> axis, kb, m, child;
        DevBuf<float> c0, ext;
        NRG_CK(acc.alloc((size_t)task_cap * ACC));
        NRG_CK(bins.alloc((size_t)task_cap * BIN_TOTAL));
        NRG_CK(axis.alloc(task_cap)); NRG_CK(kb.alloc(task_cap)); NRG_CK(m.alloc(task_cap));
        NRG_CK(child.alloc((size_t)task_cap * 2));
        NRG_CK(c0.alloc((size_t)task_cap * 3)); NRG_CK(ext.alloc((size_t)task_cap * 3));
        Split S{acc.p, bins.p, axis.p, kb.p, c0.p, ext.p, m.p, child.p};
        DevBuf<int32_t> flag, excl, fbpos;
        NRG_CK(flag.alloc(n + 1)); NRG_CK(excl.alloc(n + 1)); NRG_CK(fbpos.alloc(n));
        NRG_CK(hipMemset(flag.p + n, 0, sizeof(int32_t)));    // excl[n] = total
        size_t scan_bytes = 0;
        NRG_CK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, flag.p, excl.p, n + 1, st));
        DevBuf<unsigned char> scan_tmp;
        NRG_CK(scan_tmp.alloc(scan_bytes));
        const int32_t root_rec[6] = {0, n, 0, 1, -1, 0};
        int lt = 0;
        for (int q = 0; q < 6; q++) NRG_CK(hipMemcpy(tl[0][q].p, &root_rec[q], sizeof(int32_t), hipMemcpyHostToDevice));
        int32_t T = 1;
        int32_t h_counts[4] = {0, 0, 0, 0};
        const dim3 gn((n + 255) / 256);
        while (T > 0) {
            const int32_t zero2[2] = {0, h_counts[1]};
            NRG_CK(hipMemcpy(counts.p, zero2, 2 * sizeof(int32_t), hipMemcpyHostToDevice));
            hipLaunchKernelGGL(k_acc_init, dim3(T), dim3(256), 0, st, T, acc.p, bins.p);
            hipLaunchKernelGGL(k_bounds, gn, dim3(256), 0, st, n, P[cur], acc.p);
            hipLaunchKernelGGL(k_node_box, dim3((T + 255) / 256), dim3(256), 0, st, T, L[lt], acc.p, c0.p, ext.p, O);
            hipLaunchKernelGGL(k_bins, gn, dim3(256), 0, st, n, P[cur], c0.p, ext.p, bins.p);
            hipLaunchKernelGGL(k_split, dim3((T + 255) / 256), dim3(256), 0, st, T, L[lt], S, L[1 - lt], SL, O);
            hipLaunchKernelGGL(k_flags, gn, dim3(256), 0, st, n, P[cur], L[lt], S, flag.p, fbpos.p);
            NRG_CK(hipcub::DeviceScan::ExclusiveSum(scan_tmp.p, scan_bytes, flag.p, excl.p, n + 1, st));
            hipLaunchKernelGGL(k_scatter, gn, dim3(256), 0, st, n, P[cur], P[1 - cur], L[lt], S, flag.p, excl.p,
                               fbpos.p, O);
            NRG_CK(hipGetLastError());
            NRG_CK(hipMemcpyAsync(h_counts, counts.p, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
            NRG_CK(hipStreamSynchronize(st));
            if (h_counts[2]) { err = "median-split cluster too large for the device build"; return NR_ERR_LIMIT; }
            cur = 1 - cur;
            lt = 1 - lt;
            T = h_counts[0];
        }
        n_small = h_counts[1];
    }
    // leaves of large nodes sit at their final positions; the small subtrees then
    // overwrite their ranges with their own final order
    if (n > 1) hipLaunchKernelGGL(k_copy_order, dim3((n + 255) / 256), dim3(256), 0, st, n, P[cur].item, d_order);
    if (n_small > 0) {
        NRG_CK(hipStreamSynchronize(st));
        NRG_CK(hipMemcpy(counts.p + 1, &n_small, sizeof(int32_t), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_small, dim3(n_small), dim3(64), 0, st, SL, counts.p + 1, P[cur], d_order, O);
    }
    NRG_CK(hipGetLastError());
    float h_root[6];
    int32_t h_depth = 1;
    NRG_CK(hipMemcpyAsync(h_root, root.p, sizeof(h_root), hipMemcpyDeviceToHost, st));
    NRG_CK(hipMemcpyAsync(&h_depth, depth.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    NRG_CK(hipStreamSynchronize(st));
    for (int a = 0; a < 3; a++) { r.root_lo[a] = h_root[a]; r.root_hi[a] = h_root[3 + a]; }
    r.depth = h_depth;
    return NR_OK;
}

}  // namespace nrg
