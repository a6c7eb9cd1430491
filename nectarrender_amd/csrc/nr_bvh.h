// nr_bvh.h — host BVH build for the MI355X path tracer.
//
// The tree is the reference's binned-SAH tree (hittable/include/bvh/bvh.h:10-320:
// 16 centroid bins, cost A_L*N_L + A_R*N_R, axes with extent < EPS skipped,
// longest-axis median split when no axis qualifies, one item per leaf), built
// with the same std::partition / std::sort calls so the node structure — and
// therefore the near-first visit order on the device — is the reference's.
// The output is converted to the 64 B child-pair layout of nr_types.h.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

#include "nr_types.h"

namespace nrb {

struct Box {
    float lo[3], hi[3];
};

inline Box box_points(const float* a, const float* b) {       // aabb.h:19-23
    Box r;
    for (int k = 0; k < 3; k++) {
        if (a[k] <= b[k]) { r.lo[k] = a[k]; r.hi[k] = b[k]; }
        else { r.lo[k] = b[k]; r.hi[k] = a[k]; }
    }
    return r;
}
inline Box box_join(const Box& x, const Box& y) {              // aabb.h:25-29
    Box r;
    for (int k = 0; k < 3; k++) {
        r.lo[k] = x.lo[k] <= y.lo[k] ? x.lo[k] : y.lo[k];
        r.hi[k] = x.hi[k] >= y.hi[k] ? x.hi[k] : y.hi[k];
    }
    return r;
}
inline Box box_none() {
    Box r;
    for (int k = 0; k < 3; k++) { r.lo[k] = 1e32f; r.hi[k] = -1e32f; }
    return r;
}
inline float box_area(const Box& b) {
    float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

struct Tree {
    // reference-shaped nodes: left/right node indices, item >= 0 for leaves
    std::vector<Box> box;
    std::vector<int32_t> left, right, item;
    std::vector<int32_t> order;   // slot -> input index
    int depth = 0;
};

class Builder {
public:
    static constexpr int kBins = 16;

    void build(const std::vector<Box>& items, Tree& out) {
        boxes_ = &items;
        t_ = &out;
        out.box.clear(); out.left.clear(); out.right.clear(); out.item.clear();
        out.order.resize(items.size());
        std::iota(out.order.begin(), out.order.end(), 0);
        out.depth = 0;
        if (items.empty()) return;
        size_t cap = 2 * items.size();
        out.box.reserve(cap); out.left.reserve(cap); out.right.reserve(cap); out.item.reserve(cap);
        split(0, (int32_t)items.size(), 1);
    }

private:
    const std::vector<Box>* boxes_ = nullptr;
    Tree* t_ = nullptr;

    static float mid(const Box& b, int a) { return (b.lo[a] + b.hi[a]) * 0.5f; }
    static int32_t bin(float c, float c0, float ext) {
        int32_t k = (int32_t)(kBins * (c - c0) / ext);
        return k < 0 ? 0 : (k >= kBins ? kBins - 1 : k);
    }

    int32_t split(int32_t s, int32_t e, int depth) {
        Tree& T = *t_;
        const std::vector<Box>& B = *boxes_;
        std::vector<int32_t>& ord = T.order;
        if (depth > T.depth) T.depth = depth;
        int32_t me = (int32_t)T.box.size();
        T.box.push_back(Box{}); T.left.push_back(-1); T.right.push_back(-1); T.item.push_back(-1);
        if (e - s == 1) {
            T.item[me] = s;
            T.box[me] = B[ord[s]];
            return me;
        }
        float cl[3] = {1e32f, 1e32f, 1e32f}, ch[3] = {-1e32f, -1e32f, -1e32f};
        for (int32_t i = s; i < e; i++) {
            const Box& b = B[ord[i]];
            for (int a = 0; a < 3; a++) {
                float c = mid(b, a);
                cl[a] = fminf(cl[a], c);
                ch[a] = fmaxf(ch[a], c);
            }
        }
        float best = 1e32f;
        int best_axis = -1, best_bin = -1;
        for (int a = 0; a < 3; a++) {
            float ext = ch[a] - cl[a];
            if (ext < 1e-4f) continue;
            Box bb[kBins]; int32_t bn[kBins] = {0};
            for (int k = 0; k < kBins; k++) bb[k] = box_none();
            for (int32_t i = s; i < e; i++) {
                const Box& b = B[ord[i]];
                int32_t k = bin(mid(b, a), cl[a], ext);
                bn[k]++;
                bb[k] = bn[k] == 1 ? b : box_join(bb[k], b);
            }
            // prefix (left) and suffix (right) sweeps; the running box is
            // deliberately carried from the left sweep into the right one
            Box lb[kBins], rb[kBins]; int32_t lc[kBins], rc[kBins];
            Box run = box_none(); int32_t cnt = 0; bool on = false;
            for (int k = 0; k < kBins; k++) {
                if (bn[k] > 0) { run = on ? box_join(run, bb[k]) : bb[k]; on = true; }
                cnt += bn[k]; lb[k] = run; lc[k] = cnt;
            }
            cnt = 0; on = false;
            for (int k = kBins - 1; k >= 0; k--) {
                if (bn[k] > 0) { run = on ? box_join(run, bb[k]) : bb[k]; on = true; }
                cnt += bn[k]; rb[k] = run; rc[k] = cnt;
            }
            for (int k = 0; k < kBins - 1; k++) {
                if (lc[k] == 0 || rc[k + 1] == 0) continue;
                float cost = box_area(lb[k]) * lc[k] + box_area(rb[k + 1]) * rc[k + 1];
                if (cost < best) { best = cost; best_axis = a; best_bin = k; }
            }
        }
        if (best_axis == -1) {
            Box all = B[ord[s]];
            for (int32_t i = s + 1; i < e; i++) all = box_join(all, B[ord[i]]);
            float lx = all.hi[0] - all.lo[0], ly = all.hi[1] - all.lo[1], lz = all.hi[2] - all.lo[2];
            best_axis = (lx > ly && lx > lz) ? 0 : (ly > lz ? 1 : 2);
        }
        int32_t m;
        const int a = best_axis;
        if (best_bin == -1) {
            std::sort(ord.begin() + s, ord.begin() + e,
                      [&](int32_t x, int32_t y) { return B[x].lo[a] < B[y].lo[a]; });
            m = s + (e - s) / 2;
        } else {
            const float c0 = cl[a], ext = ch[a] - cl[a];
            const int32_t kb = best_bin;
            auto it = std::partition(ord.begin() + s, ord.begin() + e,
                                     [&](int32_t x) { return bin(mid(B[x], a), c0, ext) <= kb; });
            m = (int32_t)(it - ord.begin());
        }
        int32_t l = split(s, m, depth + 1);
        int32_t r = split(m, e, depth + 1);
        T.left[me] = l;
        T.right[me] = r;
        T.box[me] = box_join(T.box[l], T.box[r]);
        return me;
    }
};

// Convert to the child-pair layout: every interior reference node becomes one
// GNode holding both child boxes; leaves become negative refs ~slot.
inline void to_pairs(const Tree& T, std::vector<GNode>& nodes, int32_t& root_ref, float* root_lo, float* root_hi) {
    nodes.clear();
    if (T.box.empty()) { root_ref = 0; return; }
    std::vector<int32_t> pidx(T.box.size(), -1);
    int32_t n_inner = 0;
    for (size_t k = 0; k < T.box.size(); k++)
        if (T.item[k] < 0) pidx[k] = n_inner++;
    nodes.resize(n_inner);
    auto ref_of = [&](int32_t k) -> int32_t { return T.item[k] >= 0 ? ~T.item[k] : pidx[k]; };
    for (size_t k = 0; k < T.box.size(); k++) {
        if (T.item[k] >= 0) continue;
        GNode& g = nodes[pidx[k]];
        const Box& L = T.box[T.left[k]];
        const Box& R = T.box[T.right[k]];
        for (int a = 0; a < 3; a++) { g.lmin[a] = L.lo[a]; g.lmax[a] = L.hi[a]; g.rmin[a] = R.lo[a]; g.rmax[a] = R.hi[a]; }
        g.left = ref_of(T.left[k]);
        g.right = ref_of(T.right[k]);
        g.pad0 = g.pad1 = 0;
    }
    root_ref = ref_of(0);
    for (int a = 0; a < 3; a++) { root_lo[a] = T.box[0].lo[a]; root_hi[a] = T.box[0].hi[a]; }
}

}  // namespace nrb
