// build.cpp -- host construction of the 4-wide BVH.
//
// Replaces the reference's single-threaded build (src/bvh.cu:515-620: 4-ary median
// split on the max-variance axis, leaves <= 8).  This builder:
//   1. splits a binary tree top-down, by default with a binned surface-area
//      heuristic (up to 256 centroid bins per axis), optionally at the object median of the
//      axis with the largest centroid extent (std::nth_element), down to leaves of
//      <= leaf_size triangles, building independent subtrees on worker threads;
//   2. collapses it into 4-wide nodes by repeatedly opening the child with the
//      largest surface area, so every wide node has up to 4 children;
//      the areas are those of the boxes grown by kDilate of the scene diagonal on every
//      side: a distance query visits a box when its search ball reaches it, so the
//      chance of a visit follows the area of the box widened by a typical ball radius
//      (measured on the 256^3 grid: +1.9% unsigned, +1.5% signed distance, rays equal;
//      0.5% to 4% of the diagonal all within 1% of each other with leaves of 16);
//   3. writes nodes in depth-first pre-order and triangles in leaf order.
// Trees deeper than the traversal stack allows (3 * depth + 1 > 64 entries) are
// refused by the caller; median splits always stay far below that.
//
// Child boxes are inflated outward by a few ulps of the scene scale so that the
// fp32 box tests stay conservative with respect to the fp32 triangle tests: a box
// is never rejected while a triangle inside it could still win.  Query results
// therefore do not depend on the tree shape.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "bvh.hpp"

namespace cubvh_amd {
namespace {

struct Aabb {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Aabb& o) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], o.lo[k]);
            hi[k] = std::max(hi[k], o.hi[k]);
        }
    }
    // surface area of the box grown by `grow` (total, over both sides) on every axis
    float area(float grow) const {
        float d0 = hi[0] - lo[0], d1 = hi[1] - lo[1], d2 = hi[2] - lo[2];
        if (d0 < 0) return 0;
        d0 += grow;
        d1 += grow;
        d2 += grow;
        return 2.0f * (d0 * d1 + d1 * d2 + d2 * d0);
    }
};

struct BNode {
    Aabb box;
    int32_t left = -1, right = -1;  // binary children (-1 for a leaf)
    int64_t first = 0;
    int32_t count = 0;
};

// Growth of the boxes in the area terms, as a fraction of the scene diagonal.
constexpr float kDilate = 0.01f;
// Centroid bins per axis of the SAH split (256 against 16: +2% distance-query rate on
// the 256^3 grid, +10% build time).
constexpr int kMaxBins = 256;

struct Builder {
    const float* tri9;
    std::vector<Aabb> tbox;
    std::vector<float> cen;  // 3 per triangle
    std::vector<int64_t> idx;
    int leaf_size;

    // Each subtree task writes into its own node vector; merged afterwards.
    struct Part {
        std::vector<BNode> nodes;
    };

    Aabb range_box(int64_t b, int64_t e) const {
        Aabb r;
        for (int64_t i = b; i < e; ++i) r.grow(tbox[idx[i]]);
        return r;
    }

    // Builds the subtree over idx[b, e) into part; returns its root index in part.
    int32_t build_rec(Part& part, int64_t b, int64_t e) {
        int32_t me = (int32_t)part.nodes.size();
        part.nodes.emplace_back();
        part.nodes[me].box = range_box(b, e);
        int64_t n = e - b;
        if (n <= leaf_size) {
            part.nodes[me].first = b;
            part.nodes[me].count = (int32_t)n;
            return me;
        }
        int64_t m = split(b, e);
        int32_t l = build_rec(part, b, m);
        int32_t r = build_rec(part, m, e);
        part.nodes[me].left = l;
        part.nodes[me].right = r;
        return me;
    }

    int split_mode = 0;  // 0 object median, 1 binned SAH
    float grow = 0.0f;   // box growth in the area terms

    int64_t split(int64_t b, int64_t e) { return split_mode == 1 ? split_sah(b, e) : split_median(b, e); }

    // Binned surface-area heuristic over up to kMaxBins centroid bins per axis; falls back to the
    // object median when every candidate leaves one side empty.
    int64_t split_sah(int64_t b, int64_t e) {
        // as many bins as faces, from 16 up to kMaxBins
        const int NB = (int)std::min<int64_t>(kMaxBins, std::max<int64_t>(16, e - b));
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int64_t i = b; i < e; ++i) {
            const float* c = &cen[3 * idx[i]];
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], c[k]);
                hi[k] = std::max(hi[k], c[k]);
            }
        }
        float best_cost = INFINITY;
        int best_axis = -1, best_bin = -1;
        for (int axis = 0; axis < 3; ++axis) {
            const float ext = hi[axis] - lo[axis];
            if (!(ext > 0.0f)) continue;
            Aabb bb[kMaxBins];
            int64_t cnt[kMaxBins] = {0};
            const float sc = NB / ext;
            for (int64_t i = b; i < e; ++i) {
                int k = (int)((cen[3 * idx[i] + axis] - lo[axis]) * sc);
                k = std::min(std::max(k, 0), NB - 1);
                cnt[k]++;
                bb[k].grow(tbox[idx[i]]);
            }
            Aabb right[kMaxBins];
            int64_t rc[kMaxBins];
            Aabb acc;
            int64_t ac = 0;
            for (int k = NB - 1; k >= 1; --k) {
                acc.grow(bb[k]);
                ac += cnt[k];
                right[k] = acc;
                rc[k] = ac;
            }
            Aabb left;
            int64_t lc = 0;
            for (int k = 0; k < NB - 1; ++k) {
                left.grow(bb[k]);
                lc += cnt[k];
                if (lc == 0 || rc[k + 1] == 0) continue;
                float cost = left.area(grow) * (float)lc + right[k + 1].area(grow) * (float)rc[k + 1];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = axis;
                    best_bin = k;
                }
            }
        }
        if (best_axis < 0) return split_median(b, e);
        const float ext = hi[best_axis] - lo[best_axis];
        const float sc = NB / ext;
        const float* cp = cen.data();
        const float l0 = lo[best_axis];
        const int ax = best_axis, bin = best_bin;
        auto mid = std::partition(idx.begin() + b, idx.begin() + e, [cp, ax, l0, sc, bin, NB](int64_t x) {
            int k = (int)((cp[3 * x + ax] - l0) * sc);
            k = std::min(std::max(k, 0), NB - 1);
            return k <= bin;
        });
        int64_t m = mid - idx.begin();
        if (m == b || m == e) return split_median(b, e);
        return m;
    }

    // Object median on the axis of largest centroid extent.
    int64_t split_median(int64_t b, int64_t e) {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int64_t i = b; i < e; ++i) {
            const float* c = &cen[3 * idx[i]];
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], c[k]);
                hi[k] = std::max(hi[k], c[k]);
            }
        }
        int axis = 0;
        float ext = hi[0] - lo[0];
        for (int k = 1; k < 3; ++k)
            if (hi[k] - lo[k] > ext) { ext = hi[k] - lo[k]; axis = k; }
        int64_t m = b + (e - b) / 2;
        const float* cp = cen.data();
        std::nth_element(idx.begin() + b, idx.begin() + m, idx.begin() + e, [cp, axis](int64_t x, int64_t y) {
            float cx = cp[3 * x + axis], cy = cp[3 * y + axis];
            return cx < cy || (cx == cy && x < y);
        });
        return m;
    }
};

// Outward rounding of a bound by a relative and an absolute margin.
inline float pad_down(float v, float abs_pad) {
    double d = (double)v - std::fabs((double)v) * 1.0e-6 - abs_pad;
    float f = (float)d;
    if ((double)f > d) f = std::nextafter(f, -INFINITY);
    return f;
}
inline float pad_up(float v, float abs_pad) {
    double d = (double)v + std::fabs((double)v) * 1.0e-6 + abs_pad;
    float f = (float)d;
    if ((double)f < d) f = std::nextafter(f, INFINITY);
    return f;
}

}  // namespace

std::string build_bvh(const float* tri9, const int64_t* ids, int64_t n_faces, int leaf_size, HostBvh& out,
                      int split_mode) {
    if (n_faces <= 0) return "mesh has no faces";
    if (n_faces > MAX_LEAF_FIRST) return "too many faces for the 27-bit leaf encoding";
    if (leaf_size < 1 || leaf_size > MAX_LEAF_SIZE) return "leaf_size must be in 1..16";

    Builder B;
    B.tri9 = tri9;
    B.leaf_size = leaf_size;
    B.split_mode = split_mode;
    B.tbox.resize(n_faces);
    B.cen.resize(3 * n_faces);
    B.idx.resize(n_faces);
    float max_abs = 0.0f;
    for (int64_t i = 0; i < n_faces; ++i) {
        const float* t = tri9 + 9 * i;
        Aabb& bb = B.tbox[i];
        for (int v = 0; v < 3; ++v)
            for (int k = 0; k < 3; ++k) {
                float x = t[3 * v + k];
                if (!std::isfinite(x)) return "vertex coordinates must be finite";
                bb.lo[k] = std::min(bb.lo[k], x);
                bb.hi[k] = std::max(bb.hi[k], x);
                max_abs = std::max(max_abs, std::fabs(x));
            }
        for (int k = 0; k < 3; ++k) B.cen[3 * i + k] = (t[k] + t[3 + k] + t[6 + k]) * (1.0f / 3.0f);
        B.idx[i] = i;
    }

    {
        const Aabb all = B.range_box(0, n_faces);
        float diag2 = 0.0f;
        for (int k = 0; k < 3; ++k) diag2 += (all.hi[k] - all.lo[k]) * (all.hi[k] - all.lo[k]);
        B.grow = kDilate * std::sqrt(diag2);
    }
    // Top levels sequentially until there are enough independent ranges, then one
    // worker per range.
    unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    struct Range {
        int64_t b, e;
        int parent;  // index into `top`, -1 for root
        bool is_left;
    };
    // Sequential top: split ranges breadth-first until count >= 4*hw or ranges small.
    struct TopNode {
        Aabb box;
        int64_t b, e;
        int l = -1, r = -1;  // indices into top (internal), or -1
        int part = -1;       // subtree index if delegated
    };
    std::vector<TopNode> top;
    top.push_back({B.range_box(0, n_faces), 0, n_faces});
    std::vector<int> frontier = {0};
    const int64_t par_min = 1 << 14;
    while (frontier.size() < 4 * hw) {
        std::vector<int> next;
        bool any = false;
        for (int t : frontier) {
            int64_t b = top[t].b, e = top[t].e;
            if (e - b <= std::max<int64_t>(par_min, leaf_size)) { next.push_back(t); continue; }
            int64_t m = B.split(b, e);
            int l = (int)top.size();
            top.push_back({B.range_box(b, m), b, m});
            int r = (int)top.size();
            top.push_back({B.range_box(m, e), m, e});
            top[t].l = l;
            top[t].r = r;
            next.push_back(l);
            next.push_back(r);
            any = true;
        }
        frontier.swap(next);
        if (!any) break;
    }
    std::vector<Builder::Part> parts(frontier.size());
    std::vector<int32_t> part_root(frontier.size());
    {
        std::vector<std::thread> th;
        std::atomic<size_t> next{0};
        auto work = [&]() {
            for (;;) {
                size_t k = next.fetch_add(1);
                if (k >= frontier.size()) return;
                int t = frontier[k];
                top[t].part = (int)k;
                part_root[k] = B.build_rec(parts[k], top[t].b, top[t].e);
            }
        };
        unsigned nth = std::min<unsigned>(hw, (unsigned)frontier.size());
        for (unsigned i = 1; i < nth; ++i) th.emplace_back(work);
        work();
        for (auto& t : th) t.join();
    }
    for (size_t k = 0; k < frontier.size(); ++k) top[frontier[k]].part = (int)k;

    // Merge everything into one binary node array.
    std::vector<BNode> bn;
    bn.reserve(2 * (size_t)n_faces / std::max(1, leaf_size) + 16);
    std::vector<int32_t> top_map(top.size(), -1);
    // Recursive merge from the top tree.
    std::function<int32_t(int)> merge_top = [&](int t) -> int32_t {
        if (top[t].part >= 0) {
            const auto& pn = parts[top[t].part].nodes;
            int32_t base = (int32_t)bn.size();
            for (const BNode& x : pn) {
                BNode y = x;
                if (y.left >= 0) { y.left += base; y.right += base; }
                bn.push_back(y);
            }
            return base + part_root[top[t].part];
        }
        int32_t me = (int32_t)bn.size();
        bn.emplace_back();
        bn[me].box = top[t].box;
        int32_t l = merge_top(top[t].l);
        int32_t r = merge_top(top[t].r);
        bn[me].left = l;
        bn[me].right = r;
        return me;
    };
    int32_t broot = merge_top(0);

    // Collapse into 4-wide nodes (pre-order).
    const float abs_pad = std::max(max_abs, 1e-30f) * 4.0e-6f;
    out.nodes.clear();
    out.nodes.reserve(bn.size() / 2 + 2);
    int32_t max_depth = 0;
    std::function<int32_t(std::vector<int32_t>, int32_t)> emit = [&](std::vector<int32_t> kids, int32_t depth) -> int32_t {
        // open the largest-area internal child until there are 4 children
        while (kids.size() < 4) {
            int best = -1;
            float ba = -1.0f;
            for (size_t k = 0; k < kids.size(); ++k) {
                const BNode& x = bn[kids[k]];
                if (x.left >= 0 && x.box.area(B.grow) > ba) { ba = x.box.area(B.grow); best = (int)k; }
            }
            if (best < 0) break;
            int32_t o = kids[best];
            kids[best] = bn[o].left;
            kids.insert(kids.begin() + best + 1, bn[o].right);
        }
        max_depth = std::max(max_depth, depth);
        int32_t me = (int32_t)out.nodes.size();
        out.nodes.emplace_back();
        {
            WideNode& w = out.nodes[me];
            std::memset(&w, 0, sizeof w);
            for (int k = 0; k < 4; ++k) {
                w.lo_x[k] = w.lo_y[k] = w.lo_z[k] = INFINITY;
                w.hi_x[k] = w.hi_y[k] = w.hi_z[k] = -INFINITY;
                w.child[k] = CHILD_EMPTY;
            }
            w.n_children = (int32_t)kids.size();
        }
        for (size_t k = 0; k < kids.size(); ++k) {
            const BNode& x = bn[kids[k]];
            int32_t ref;
            if (x.left < 0) {
                ref = make_leaf_ref(x.first, x.count);
            } else {
                ref = emit({x.left, x.right}, depth + 1);
            }
            WideNode& w = out.nodes[me];  // re-fetch: emit may reallocate
            w.child[k] = ref;
            w.lo_x[k] = pad_down(x.box.lo[0], abs_pad);
            w.lo_y[k] = pad_down(x.box.lo[1], abs_pad);
            w.lo_z[k] = pad_down(x.box.lo[2], abs_pad);
            w.hi_x[k] = pad_up(x.box.hi[0], abs_pad);
            w.hi_y[k] = pad_up(x.box.hi[1], abs_pad);
            w.hi_z[k] = pad_up(x.box.hi[2], abs_pad);
        }
        return me;
    };
    const BNode& r0 = bn[broot];
    if (r0.left < 0) {
        emit({broot}, 1);  // the whole mesh is one leaf: root with a single leaf child
    } else {
        emit({r0.left, r0.right}, 1);
    }

    out.n_faces = n_faces;
    out.leaf_size = leaf_size;
    out.depth = max_depth;
    out.stack_need = 3 * max_depth + 1;
    out.tris.assign(12 * (size_t)n_faces, 0.0f);
    for (int64_t s = 0; s < n_faces; ++s) {
        int64_t f = B.idx[s];
        const float* t = tri9 + 9 * f;
        float* o = &out.tris[12 * s];
        o[0] = t[0]; o[1] = t[1]; o[2] = t[2];
        int32_t id32 = (int32_t)ids[f];
        std::memcpy(&o[3], &id32, 4);
        o[4] = t[3]; o[5] = t[4]; o[6] = t[5];
        o[8] = t[6]; o[9] = t[7]; o[10] = t[8];
        if (ids[f] == 0) out.face0_slot = s;
    }
    compute_face_bounds(out);
    return "";
}

namespace {

// Bounding disk of a point set: centre and direction given in double, stored as
// fp32 (centre, radius, unit normal, thickness).  Radius and thickness are measured
// from the rounded centre along the rounded normal in double and rounded outward, so
// every point lies inside the stored cylinder.  A zero direction gives a sphere.
void store_disk(const double (*v)[3], int nv, const double centre[3], const double dir[3], float* o) {
    float cf[3], nf[3] = {0.0f, 0.0f, 0.0f};
    for (int k = 0; k < 3; ++k) cf[k] = (float)centre[k];
    const double nl = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    if (nl > 0.0 && std::isfinite(nl))
        for (int k = 0; k < 3; ++k) nf[k] = (float)(dir[k] / nl);
    double r = 0.0, th = 0.0;
    for (int i = 0; i < nv; ++i) {
        const double d[3] = {v[i][0] - cf[0], v[i][1] - cf[1], v[i][2] - cf[2]};
        const double h = d[0] * nf[0] + d[1] * nf[1] + d[2] * nf[2];
        const double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        r = std::max(r, std::sqrt(std::max(dd - h * h, 0.0)));
        th = std::max(th, std::fabs(h));
    }
    r = r * (1.0 + 1e-6) + 1e-30;
    th = th * (1.0 + 1e-6) + 1e-30;
    float rf = (float)r, tf = (float)th;
    if ((double)rf < r) rf = std::nextafter(rf, INFINITY);
    if ((double)tf < th) tf = std::nextafter(tf, INFINITY);
    o[0] = cf[0]; o[1] = cf[1]; o[2] = cf[2]; o[3] = rf;
    o[4] = nf[0]; o[5] = nf[1]; o[6] = nf[2]; o[7] = tf;
}

void face_vertices(const HostBvh& bvh, int64_t s, double (*v)[3]) {
    const float* t = &bvh.tris[12 * s];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) v[i][k] = t[4 * i + k];
}

void face_normal(const double (*v)[3], double* n) {
    double e1[3], e2[3];
    for (int k = 0; k < 3; ++k) {
        e1[k] = v[1][k] - v[0][k];
        e2[k] = v[2][k] - v[0][k];
    }
    n[0] = e1[1] * e2[2] - e1[2] * e2[1];
    n[1] = e1[2] * e2[0] - e1[0] * e2[2];
    n[2] = e1[0] * e2[1] - e1[1] * e2[0];
}

// fp32 vector operations of the distance formula, with its fused operations and no
// others (this file is compiled with -ffp-contract=off): bit-identical to the device.
struct V3 {
    float x, y, z;
};
V3 vsub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
float vdot(V3 a, V3 b) { return std::fmaf(a.z, b.z, std::fmaf(a.y, b.y, a.x * b.x)); }
V3 vcross(V3 a, V3 b) {
    return V3{std::fmaf(a.y, b.z, -(a.z * b.y)), std::fmaf(a.z, b.x, -(a.x * b.z)), std::fmaf(a.x, b.y, -(a.y * b.x))};
}

void store_face_constants(const float* t, float* o) {
    const V3 a{t[0], t[1], t[2]}, b{t[4], t[5], t[6]}, c{t[8], t[9], t[10]};
    const V3 v21 = vsub(b, a), v32 = vsub(c, b), v13 = vsub(a, c);
    const V3 n = vcross(v21, v13);
    const V3 g1 = vcross(v21, n), g2 = vcross(v32, n), g3 = vcross(v13, n);
    const float rows[9][4] = {
        {a.x, a.y, a.z, t[3]},          {b.x, b.y, b.z, vdot(v21, v21)},    {c.x, c.y, c.z, vdot(v32, v32)},
        {v21.x, v21.y, v21.z, vdot(v13, v13)}, {v32.x, v32.y, v32.z, vdot(n, n)}, {v13.x, v13.y, v13.z, n.x},
        {g1.x, g1.y, g1.z, n.y},        {g2.x, g2.y, g2.z, n.z},            {g3.x, g3.y, g3.z, 0.0f}};
    std::memcpy(o, rows, sizeof(rows));
}

}  // namespace

void compute_face_bounds(HostBvh& bvh) {
    const int64_t m = bvh.n_faces;
    bvh.tbnd.assign(8 * (size_t)m, 0.0f);
    bvh.lbnd.assign(8 * (size_t)m, 0.0f);
    bvh.tpre.assign(36 * (size_t)m, 0.0f);
    for (int64_t s = 0; s < m; ++s) store_face_constants(&bvh.tris[12 * s], &bvh.tpre[36 * s]);
    // faces: centroid and plane normal
    for (int64_t s = 0; s < m; ++s) {
        double v[3][3], n[3], c[3];
        face_vertices(bvh, s, v);
        face_normal(v, n);
        for (int k = 0; k < 3; ++k) c[k] = (v[0][k] + v[1][k] + v[2][k]) / 3.0;
        store_disk(v, 3, c, n, &bvh.tbnd[8 * s]);
    }
    // leaves: centre of the vertex box and the area-weighted mean normal
    for (const WideNode& w : bvh.nodes) {
        for (int k = 0; k < 4; ++k) {
            const int32_t ref = w.child[k];
            if (ref == CHILD_EMPTY || ref >= 0) continue;
            const int64_t first = leaf_ref_first(ref);
            const int count = leaf_ref_count(ref);
            if (first + count > m) continue;  // rejected by validate_bvh
            double v[3 * MAX_LEAF_SIZE][3], n[3] = {0.0, 0.0, 0.0};
            double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int j = 0; j < count; ++j) {
                double fn[3];
                face_vertices(bvh, first + j, &v[3 * j]);
                face_normal(&v[3 * j], fn);
                for (int a = 0; a < 3; ++a) {
                    n[a] += fn[a];
                    for (int i = 0; i < 3; ++i) {
                        lo[a] = std::min(lo[a], v[3 * j + i][a]);
                        hi[a] = std::max(hi[a], v[3 * j + i][a]);
                    }
                }
            }
            const double c[3] = {0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.5 * (lo[2] + hi[2])};
            store_disk(v, 3 * count, c, n, &bvh.lbnd[8 * first]);
        }
    }
    // internal subtrees: the same rule over every vertex below the child; the nodes are
    // shared out to threads (each vertex is visited once per ancestor level)
    const int64_t nn = (int64_t)bvh.nodes.size();
    bvh.cbnd.assign(32 * (size_t)nn, 0.0f);
    std::atomic<int64_t> next{0};
    auto work = [&]() {
        std::vector<int32_t> st;
        std::vector<double> pts;  // 3 per vertex
        for (;;) {
            const int64_t ni = next.fetch_add(1);
            if (ni >= nn) return;
            for (int k = 0; k < 4; ++k) {
                const int32_t top = bvh.nodes[ni].child[k];
                if (top < 0) continue;  // leaf or empty slot
                pts.clear();
                double n[3] = {0.0, 0.0, 0.0};
                double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
                st.assign(1, top);
                while (!st.empty()) {
                    const int32_t ref = st.back();
                    st.pop_back();
                    if (ref >= 0) {
                        for (int c = 0; c < 4; ++c)
                            if (bvh.nodes[ref].child[c] != CHILD_EMPTY) st.push_back(bvh.nodes[ref].child[c]);
                        continue;
                    }
                    const int64_t first = leaf_ref_first(ref);
                    const int count = leaf_ref_count(ref);
                    for (int j = 0; j < count && first + j < m; ++j) {
                        double v[3][3], fn[3];
                        face_vertices(bvh, first + j, v);
                        face_normal(v, fn);
                        for (int a = 0; a < 3; ++a) {
                            n[a] += fn[a];
                            for (int i = 0; i < 3; ++i) {
                                lo[a] = std::min(lo[a], v[i][a]);
                                hi[a] = std::max(hi[a], v[i][a]);
                            }
                        }
                        for (int i = 0; i < 3; ++i) pts.insert(pts.end(), v[i], v[i] + 3);
                    }
                }
                const double c[3] = {0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.5 * (lo[2] + hi[2])};
                store_disk(reinterpret_cast<const double(*)[3]>(pts.data()), (int)(pts.size() / 3), c, n,
                           &bvh.cbnd[32 * ni + 8 * k]);
            }
        }
    };
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const unsigned nt = nn < 64 ? 1u : hw;
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; ++t) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
}

std::string validate_bvh(HostBvh& bvh) {
    const int64_t nn = (int64_t)bvh.nodes.size();
    if (nn <= 0) return "no nodes";
    if ((int64_t)bvh.tris.size() != 12 * bvh.n_faces) return "triangle array size mismatch";
    std::vector<uint8_t> seen(nn, 0);
    std::vector<uint8_t> covered(bvh.n_faces, 0);
    int32_t max_depth = 0;
    // iterative DFS with explicit depth; each node must be reached exactly once
    std::vector<std::pair<int32_t, int32_t>> st = {{0, 1}};
    while (!st.empty()) {
        auto [ni, d] = st.back();
        st.pop_back();
        if (ni < 0 || ni >= nn) return "node reference out of range";
        if (seen[ni]) return "node reached twice";
        seen[ni] = 1;
        max_depth = std::max(max_depth, d);
        WideNode& w = bvh.nodes[ni];
        for (int k = 0; k < 4; ++k) {
            int32_t ref = w.child[k];
            if (ref == CHILD_EMPTY) {
                // the walks rely on an empty slot's box being empty (+inf .. -inf): no
                // point with a finite or infinite coordinate reaches it
                w.lo_x[k] = w.lo_y[k] = w.lo_z[k] = INFINITY;
                w.hi_x[k] = w.hi_y[k] = w.hi_z[k] = -INFINITY;
                continue;
            }
            if (ref >= 0) {
                st.push_back({ref, d + 1});
            } else {
                int64_t first = leaf_ref_first(ref);
                int cnt = leaf_ref_count(ref);
                if (first + cnt > bvh.n_faces) return "leaf range out of bounds";
                for (int j = 0; j < cnt; ++j) {
                    if (covered[first + j]) return "triangle in two leaves";
                    covered[first + j] = 1;
                }
            }
        }
    }
    for (int64_t i = 0; i < bvh.n_faces; ++i)
        if (!covered[i]) return "triangle not referenced by any leaf";
    bvh.depth = max_depth;
    bvh.stack_need = 3 * max_depth + 1;
    bvh.face0_slot = 0;
    for (int64_t s = 0; s < bvh.n_faces; ++s) {
        int32_t id32;
        std::memcpy(&id32, &bvh.tris[12 * s + 3], 4);
        if (id32 == 0) { bvh.face0_slot = s; break; }
    }
    return "";
}

}  // namespace cubvh_amd
