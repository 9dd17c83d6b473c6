// cubvh_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the reference (ashawkey/cubvh) BVH query path, used as the
// checker for the HIP product path and as the timed CPU baseline.  Nothing in the
// product (cubvh_amd/) links, loads or calls this file; only tests/, the smoke()
// check and bench.py's cpu_baseline leg do.
//
// Contents and what pins each part (the pins live in tests/test_oracle_*.py):
//
//  (1) fp32 per-triangle primitives, restated from the reference:
//        tri_dist2        include/gpu/triangle.cuh:46-68 (Inigo Quilez unsigned distance)
//        tri_ray          include/gpu/triangle.cuh:27-44 (IQ ray/triangle intersector)
//        tri_closest_pt   include/gpu/triangle.cuh:74-139
//        tri_barycentric  include/gpu/triangle.cuh:141-158 (u evaluated in double)
//        safe_divide      include/gpu/common.h:98-106
//        box_dist2        include/gpu/bounding_box.cuh:216-218
//        box_ray          include/gpu/bounding_box.cuh:151-198
//      Pinned against the FP64 brute force below (closed-form Voronoi-region
//      closest point, Moller-Trumbore) on the golden meshes, and against special
//      cases with known answers (points above a face interior, on vertices, etc.).
//      Arithmetic convention: every dot product is fmaf(z,z', fmaf(y,y', x*x')) and
//      every cross-product component is fmaf(a,b, -(c*d)); everything else is plain
//      IEEE fp32 with round-to-nearest.  The reference is compiled by nvcc with
//      default FMA contraction, whose exact contraction choices cannot be observed
//      here: parity at the ulp level against the reference's own GPU output is
//      therefore UNPINNED; parity to 1e-5 is pinned by the FP64 brute force.
//
//  (2) brute-force queries over all faces in fp32 with the canonical tie rule
//      (smallest squared distance / depth, ties to the lowest face index).  This is
//      the bit-exact parity target of the GPU kernels.
//
//  (3) FP64 brute force (independent mathematics): closest point by Voronoi regions
//      (Ericson, Real-Time Collision Detection, 5.1.5), Moller-Trumbore ray hits,
//      generalized winding number (Van Oosterom & Strackee solid angle).
//
//  (4) the reference BVH4: host build (src/bvh.cu:515-620, median split on the
//      max-variance centroid axis with std::nth_element, leaves <= 8, 4 children per
//      node) and the stack traversals closest_triangle (src/bvh.cu:300-353),
//      ray_intersect (250-298), avg_normal_around_point (356-396), the watertight
//      and raystab sign rules (398-427) and the kernel epilogues (631-740).  It
//      counts the box and triangle tests for the algorithmic-bytes figures, and is
//      the CPU baseline.  Its face choice follows the reference tie rules
//      (last visited wins for distance, first visited wins for rays).
//
//  (5) PCG32 (include/gpu/pcg32.h:33-170; O'Neill's published PCG-XSH-RR 64/32 with
//      Brown's jump-ahead) and the Fibonacci stab directions (common.h:63-96).
//      Pinned by known-answer tests against stepping and against the published
//      algorithm restated independently in Python.
//
//  (6) floodfill: the reference's iterated hook / pointer-jump loop
//      (include/gpu/floodfill.cuh:14-272, cubic grids) and a BFS over tensor-order
//      6-connectivity.  Pinned against scipy.ndimage.label and hand-built grids.
//
//  (7) the sparse-voxel hash table (include/cpu/hashtable.h:20-147, the sequential
//      form of include/gpu/hashtable.cuh:24-217): (marker, row) slots, the
//      CityHash-like slot function of the reference (kept alike so that the table can
//      be compared slot for slot with the reference's own CPU table), linear probing,
//      rows inserted in order.  Pinned against the known answers of the reference's
//      test/hashtable.py, against a brute-force dictionary (the lowest row index of
//      each key), and slot for slot against the reference's CPU table compiled from
//      its header (oracle/_ref, when the reference sources are present).
//
//  (8) sparse marching cubes: the output contract of the reference's GPU extraction
//      (include/gpu/spcumc.cuh:379-530, 631-868) restated sequentially -- edge vertices
//      keyed by (lowest lattice point, axis), ordered by a stable sort, placed from
//      the first cell holding the edge; triangles per cell in table order with the
//      second and third vertex exchanged -- and the corner averaging of its CPU path
//      (include/cpu/spmc.h:408-433, double sums in cell order).  The case table is the
//      product's (cubvh_amd/csrc/mc_table.hpp), pinned entry by entry against the
//      reference's own; the rest is pinned against the reference's CPU extraction
//      compiled from its header (same triangles and vertex positions, vertex order
//      aside) and against closed-surface invariants.
//
//  (9) backward of the unsigned distance (no counterpart in the reference): the
//      envelope-theorem gradient (grad_p = g n, grad_v_k -= w_k g n with
//      n = (p - c) / |p - c|, c rebuilt from face + uvw) in fp32 with the product's
//      operations and summation order -- per face its queries in ascending index, in
//      chunks of 32 summed separately, then the chunk sums reduced in runs of 32 (each
//      run added in order from zero) until one value is left; per vertex its corners
//      3 f + k in ascending order -- so that the product's
//      vertex gradients can be compared bit for bit.  Pinned against FP64 central
//      finite differences of the FP64 brute-force distance (tests/test_oracle_backward.py).
//
// sin/cos of the stab directions are evaluated as (float)sin((double)x): a libm
// result, not the CUDA math library.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <stack>
#include <tuple>
#include <vector>

#include "../cubvh_amd/csrc/mc_table.hpp"  // the case table (pinned against the reference's)

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ----------------------------------------------------------------------------
// fp32 vector helpers (canonical arithmetic convention, see header)
// ----------------------------------------------------------------------------
struct V3 {
    float x, y, z;
};
inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }
inline float dot(V3 a, V3 b) { return std::fmaf(a.z, b.z, std::fmaf(a.y, b.y, a.x * b.x)); }
inline V3 cross(V3 a, V3 b) {
    return v3(std::fmaf(a.y, b.z, -(a.z * b.y)), std::fmaf(a.z, b.x, -(a.x * b.z)),
              std::fmaf(a.x, b.y, -(a.y * b.x)));
}
inline float sgn(float x) { return std::copysign(1.0f, x); }  // common.h:49-51
inline float clampf(float v, float lo, float hi) {           // common.h:53-56
    return v < lo ? lo : (hi < v ? hi : v);
}
inline float stdmin(float a, float b) { return (b < a) ? b : a; }  // std::min
inline float stdmax(float a, float b) { return (a < b) ? b : a; }  // std::max

// common.h:98-106
inline float safe_divide(float num, float den, float eps = 1e-6f) {
    if (std::fabs(den) < eps) {
        if (den <= 0) return -(num / eps);
        return num / eps;
    }
    return num / den;
}

struct Tri {
    V3 a, b, c;
    int64_t id;
};

// triangle.cuh:46-68
inline float tri_dist2(const Tri& t, V3 p) {
    V3 v21 = sub(t.b, t.a), p1 = sub(p, t.a);
    V3 v32 = sub(t.c, t.b), p2 = sub(p, t.b);
    V3 v13 = sub(t.a, t.c), p3 = sub(p, t.c);
    V3 nor = cross(v21, v13);
    float s = sgn(dot(cross(v21, nor), p1)) + sgn(dot(cross(v32, nor), p2)) +
              sgn(dot(cross(v13, nor), p3));
    if (s < 2.0f) {
        float t1 = clampf(dot(v21, p1) / dot(v21, v21), 0.0f, 1.0f);
        float t2 = clampf(dot(v32, p2) / dot(v32, v32), 0.0f, 1.0f);
        float t3 = clampf(dot(v13, p3) / dot(v13, v13), 0.0f, 1.0f);
        V3 e1 = v3(std::fmaf(v21.x, t1, -p1.x), std::fmaf(v21.y, t1, -p1.y), std::fmaf(v21.z, t1, -p1.z));
        V3 e2 = v3(std::fmaf(v32.x, t2, -p2.x), std::fmaf(v32.y, t2, -p2.y), std::fmaf(v32.z, t2, -p2.z));
        V3 e3 = v3(std::fmaf(v13.x, t3, -p3.x), std::fmaf(v13.y, t3, -p3.y), std::fmaf(v13.z, t3, -p3.z));
        return stdmin(stdmin(dot(e1, e1), dot(e2, e2)), dot(e3, e3));
    }
    float q = dot(nor, p1);
    return q * q / dot(nor, nor);
}

// triangle.cuh:27-44; a miss is 1e6.
// exact_div = false: the reference, 1/(d.n) through safe_divide (clamped at 1e-6),
//   which scales u, v and t by |d.n|/1e-6 on grazing rays and can report hits far
//   from the triangle.
// exact_div = true: the canonical form used by the product, plain IEEE division
//   (d.n == 0 is a miss); identical to the reference whenever |d.n| >= 1e-6.
inline float tri_ray(const Tri& t, V3 ro, V3 rd, bool exact_div = true) {
    V3 v1v0 = sub(t.b, t.a), v2v0 = sub(t.c, t.a), rov0 = sub(ro, t.a);
    V3 n = cross(v1v0, v2v0);
    V3 q = cross(rov0, rd);
    float den = dot(rd, n);
    if (exact_div && den == 0.0f) return 1e6f;
    float d = exact_div ? 1.0f / den : safe_divide(1.0f, den);
    float u = d * -dot(q, v2v0);
    float v = d * dot(q, v1v0);
    float tt = d * -dot(n, rov0);
    if (exact_div) {
        if (!(u >= 0.0f && u <= 1.0f && v >= 0.0f && (u + v) <= 1.0f && tt >= 0.0f)) tt = 1e6f;
    } else if (u < 0.0f || u > 1.0f || v < 0.0f || (u + v) > 1.0f || tt < 0.0f) {
        tt = 1e6f;
    }
    return tt;
}

// triangle.cuh:23-25 (Eigen normalized(): divide by sqrt of squaredNorm if > 0)
inline V3 tri_normal(const Tri& t) {
    V3 n = cross(sub(t.b, t.a), sub(t.c, t.a));
    float z = dot(n, n);
    if (z > 0.0f) {
        float s = std::sqrt(z);
        return v3(n.x / s, n.y / s, n.z / s);
    }
    return n;
}

// triangle.cuh:74-106
inline bool point_in_triangle(const Tri& t, V3 p) {
    V3 la = sub(t.a, p), lb = sub(t.b, p), lc = sub(t.c, p);
    V3 u = cross(lb, lc), v = cross(lc, la), w = cross(la, lb);
    if (dot(u, v) < 0.0f) return false;
    if (dot(u, w) < 0.0f) return false;
    return true;
}

// triangle.cuh:108-112
inline V3 closest_point_to_line(V3 a, V3 b, V3 c) {
    V3 ba = sub(b, a);
    float t = dot(sub(c, a), ba) / dot(ba, ba);
    t = stdmax(stdmin(t, 1.0f), 0.0f);
    return v3(std::fmaf(t, ba.x, a.x), std::fmaf(t, ba.y, a.y), std::fmaf(t, ba.z, a.z));
}

// triangle.cuh:114-139
inline V3 tri_closest_pt(const Tri& t, V3 point) {
    V3 n = tri_normal(t);
    float s = dot(n, sub(point, t.a));
    point = v3(std::fmaf(-s, n.x, point.x), std::fmaf(-s, n.y, point.y), std::fmaf(-s, n.z, point.z));
    if (point_in_triangle(t, point)) return point;
    V3 c1 = closest_point_to_line(t.a, t.b, point);
    V3 c2 = closest_point_to_line(t.b, t.c, point);
    V3 c3 = closest_point_to_line(t.c, t.a, point);
    V3 d1 = sub(point, c1), d2 = sub(point, c2), d3 = sub(point, c3);
    float m1 = dot(d1, d1), m2 = dot(d2, d2), m3 = dot(d3, d3);
    float m = stdmin(stdmin(m1, m2), m3);
    if (m == m1) return c1;
    if (m == m2) return c2;
    return c3;
}

// triangle.cuh:141-158
inline V3 tri_barycentric(const Tri& t, V3 p) {
    V3 v0 = sub(t.b, t.a), v1 = sub(t.c, t.a), v2 = sub(p, t.a);
    float d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1);
    float d20 = dot(v2, v0), d21 = dot(v2, v1);
    float denom = std::fmaf(d00, d11, -(d01 * d01));
    float v = std::fmaf(d11, d20, -(d01 * d21)) / denom;
    float w = std::fmaf(d00, d21, -(d01 * d20)) / denom;
    float u = (float)(1.0 - (double)v - (double)w);
    return v3(u, v, w);
}

struct Box {
    V3 mn, mx;
};

// bounding_box.cuh:216-218
inline float box_dist2(const Box& b, V3 p) {
    V3 d = v3(stdmax(stdmax(b.mn.x - p.x, p.x - b.mx.x), 0.0f), stdmax(stdmax(b.mn.y - p.y, p.y - b.mx.y), 0.0f),
              stdmax(stdmax(b.mn.z - p.z, p.z - b.mx.z), 0.0f));
    return dot(d, d);
}

// Squared distance to the box widened by a relative 1e-5 (plus 1e-7) per axis: a lower
// bound that stays below every face's fp32 squared distance despite rounding; used by
// the canonical mode of the reference-BVH walk.
inline float box_dist2_padded(const Box& b, V3 p) {
    auto axis = [](float lo, float hi, float x) {
        float pad = 1e-5f * (std::fabs(lo) + std::fabs(hi)) + 1e-7f;
        return stdmax(stdmax((lo - pad) - x, x - (hi + pad)), 0.0f);
    };
    V3 d = v3(axis(b.mn.x, b.mx.x, p.x), axis(b.mn.y, b.mx.y, p.y), axis(b.mn.z, b.mx.z, p.z));
    return dot(d, d);
}

// bounding_box.cuh:151-198, returns tmin (FLT_MAX on a miss)
inline float box_ray(const Box& b, V3 pos, V3 dir) {
    const float FMAX = std::numeric_limits<float>::max();
    float tmin = safe_divide(b.mn.x - pos.x, dir.x);
    float tmax = safe_divide(b.mx.x - pos.x, dir.x);
    if (tmin > tmax) std::swap(tmin, tmax);
    float tymin = safe_divide(b.mn.y - pos.y, dir.y);
    float tymax = safe_divide(b.mx.y - pos.y, dir.y);
    if (tymin > tymax) std::swap(tymin, tymax);
    if (tmin > tymax || tymin > tmax) return FMAX;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    float tzmin = safe_divide(b.mn.z - pos.z, dir.z);
    float tzmax = safe_divide(b.mx.z - pos.z, dir.z);
    if (tzmin > tzmax) std::swap(tzmin, tzmax);
    if (tmin > tzmax || tzmin > tmax) return FMAX;
    if (tzmin > tmin) tmin = tzmin;
    return tmin;
}

// Conservative slab test used by the canonical mode of the reference-BVH walk:
// exact reciprocals (zero components replaced by +-1e-30), boxes widened by a
// relative 1e-5; returns the entry parameter, FLT_MAX on a miss.
inline float box_ray_conservative(const Box& b, V3 pos, V3 dir) {
    const float FMAX = std::numeric_limits<float>::max();
    float lo[3] = {b.mn.x, b.mn.y, b.mn.z}, hi[3] = {b.mx.x, b.mx.y, b.mx.z};
    float o[3] = {pos.x, pos.y, pos.z}, d[3] = {dir.x, dir.y, dir.z};
    float tn = -FMAX, tf = FMAX;
    for (int k = 0; k < 3; ++k) {
        float pad = 1e-5f * (std::fabs(lo[k]) + std::fabs(hi[k])) + 1e-7f;
        float dk = std::fabs(d[k]) < 1e-30f ? std::copysign(1e-30f, d[k]) : d[k];
        float inv = 1.0f / dk;
        float t0 = (lo[k] - pad - o[k]) * inv, t1 = (hi[k] + pad - o[k]) * inv;
        if (t0 > t1) std::swap(t0, t1);
        tn = std::max(tn, t0);
        tf = std::min(tf, t1);
    }
    if (tn > tf || tf < 0.0f) return FMAX;
    return tn;
}

// ----------------------------------------------------------------------------
// PCG32 (pcg32.h:33-170)
// ----------------------------------------------------------------------------
struct Pcg32 {
    uint64_t state = 0x853c49e6748fea9bULL;
    uint64_t inc = 0xda3e39cb94b95bdbULL;
    static constexpr uint64_t MULT = 0x5851f42d4c957f2dULL;
    uint32_t next_uint() {
        uint64_t old = state;
        state = old * MULT + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31));
    }
    float next_float() {
        uint32_t u = (next_uint() >> 9) | 0x3f800000u;
        float f;
        std::memcpy(&f, &u, 4);
        return f - 1.0f;
    }
    void advance(int64_t delta_) {
        uint64_t cur_mult = MULT, cur_plus = inc, acc_mult = 1u, acc_plus = 0u;
        uint64_t delta = (uint64_t)delta_;
        while (delta > 0) {
            if (delta & 1) {
                acc_mult *= cur_mult;
                acc_plus = acc_plus * cur_mult + cur_plus;
            }
            cur_plus = (cur_mult + 1) * cur_plus;
            cur_mult *= cur_mult;
            delta /= 2;
        }
        state = acc_mult * state + acc_plus;
    }
};

// common.h:63-96 with N_DIRS = 32 (epsilon 1.33)
inline float fractf(float x) { return x - std::floor(x); }
inline V3 fibonacci_dir32(uint32_t i, float ox, float oy) {
    const float eps = 1.33f;
    const float GOLDEN = 1.6180339887498948482045868343656f;
    float px = fractf(((float)i + eps) / ((float)31u + 2 * eps) + ox);
    float py = fractf((float)i / GOLDEN + oy);
    const float PI = 3.14159265358979323846f;
    float cos_theta = std::fmaf(-2.0f, px, 1.0f);
    float phi = (2.0f * PI) * (py - 0.5f);
    float sin_theta = std::sqrt(std::fmax(std::fmaf(-cos_theta, cos_theta, 1.0f), 0.0f));
    float sp = (float)std::sin((double)phi);
    float cp = (float)std::cos((double)phi);
    return v3(sin_theta * cp, sin_theta * sp, cos_theta);
}

// ----------------------------------------------------------------------------
// mesh gather (src/api_gpu.cu:28-30)
// ----------------------------------------------------------------------------
std::vector<Tri> gather(const float* V, int64_t nV, const int32_t* F, int64_t nF) {
    std::vector<Tri> t((size_t)nF);
    for (int64_t i = 0; i < nF; ++i) {
        const int32_t* f = F + 3 * i;
        for (int k = 0; k < 3; ++k)
            if (f[k] < 0 || f[k] >= nV) return {};
        t[i].a = v3(V[3 * f[0]], V[3 * f[0] + 1], V[3 * f[0] + 2]);
        t[i].b = v3(V[3 * f[1]], V[3 * f[1] + 1], V[3 * f[1] + 2]);
        t[i].c = v3(V[3 * f[2]], V[3 * f[2] + 1], V[3 * f[2] + 2]);
        t[i].id = i;
    }
    return t;
}

inline V3 load3(const float* p, int64_t i) { return v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
inline void store3(float* p, int64_t i, V3 v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}

int set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) return nthreads;
    return omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}

constexpr float MAX_DIST = 1000.0f;            // src/bvh.cu:17
constexpr float MAX_DIST_SQ = MAX_DIST * MAX_DIST;  // src/bvh.cu:18

// canonical (layout-independent) nearest face: smallest fp32 squared distance,
// ties to the lowest face index.  -1 when nothing is within MAX_DIST.
inline int64_t brute_nearest(const std::vector<Tri>& tris, V3 p, float* out_d2) {
    float best = MAX_DIST_SQ;
    int64_t bi = -1;
    for (size_t i = 0; i < tris.size(); ++i) {
        float d2 = tri_dist2(tris[i], p);
        if (d2 < best || (d2 == best && bi >= 0 && (int64_t)i < bi)) {
            best = d2;
            bi = (int64_t)i;
        }
    }
    *out_d2 = best;
    return bi;
}

// canonical watertight normal sum: all faces with tri_dist2(c) < 1e-6, unit normals
// accumulated in double (order independent), rounded to float, divided by the count.
inline V3 brute_avg_normal(const std::vector<Tri>& tris, V3 c) {
    double sx = 0, sy = 0, sz = 0;
    float cnt = 0;
    for (size_t i = 0; i < tris.size(); ++i) {
        if (tri_dist2(tris[i], c) < 1e-6f) {
            V3 n = tri_normal(tris[i]);
            sx += n.x;
            sy += n.y;
            sz += n.z;
            cnt += 1.0f;
        }
    }
    return v3((float)sx / cnt, (float)sy / cnt, (float)sz / cnt);
}

inline bool brute_any_hit(const std::vector<Tri>& tris, V3 o, V3 d) {
    for (size_t i = 0; i < tris.size(); ++i)
        if (tri_ray(tris[i], o, d) < MAX_DIST) return true;
    return false;
}

// ----------------------------------------------------------------------------
// reference BVH4 (src/bvh.cu:141-620)
// ----------------------------------------------------------------------------
struct RefNode {
    Box bb;
    int left_idx;
    int right_idx;
    int escape_idx;
};
static_assert(sizeof(RefNode) == 36, "reference node is 9 x 4 bytes");

// Triangle record as laid out by the reference (3 x Vector3f, 4 pad bytes, int64 id).
struct RefTriRecord {
    float a[3], b[3], c[3];
    int32_t pad;
    int64_t id;
};
static_assert(sizeof(RefTriRecord) == 48, "reference triangle is 48 bytes");

struct RefBvh {
    std::vector<Tri> tris;  // permuted in place by the build
    std::vector<RefNode> nodes;
};

// host-side (gcc, no FMA contraction) centroid, triangle.cuh:160-166
inline float centroid_axis(const Tri& t, int axis) {
    const float* a = &t.a.x;
    const float* b = &t.b.x;
    const float* c = &t.c.x;
    return (a[axis] + b[axis] + c[axis]) / 3.0f;
}
inline V3 centroid(const Tri& t) {
    return v3((t.a.x + t.b.x + t.c.x) / 3.0f, (t.a.y + t.b.y + t.c.y) / 3.0f, (t.a.z + t.b.z + t.c.z) / 3.0f);
}

Box box_of(std::vector<Tri>::iterator b, std::vector<Tri>::iterator e) {
    Box bb;
    bb.mn = bb.mx = b->a;
    for (auto it = b; it != e; ++it) {
        for (const V3* p : {&it->a, &it->b, &it->c}) {
            bb.mn = v3(stdmin(bb.mn.x, p->x), stdmin(bb.mn.y, p->y), stdmin(bb.mn.z, p->z));
            bb.mx = v3(stdmax(bb.mx.x, p->x), stdmax(bb.mx.y, p->y), stdmax(bb.mx.z, p->z));
        }
    }
    return bb;
}

// src/bvh.cu:515-620
void ref_build(RefBvh& bvh, uint32_t n_per_leaf) {
    constexpr int BF = 4;
    auto& tris = bvh.tris;
    auto& nodes = bvh.nodes;
    nodes.clear();
    nodes.emplace_back();
    nodes.front().bb = box_of(tris.begin(), tris.end());
    nodes.front().left_idx = nodes.front().right_idx = 0;
    struct BN {
        int node_idx;
        std::vector<Tri>::iterator begin, end;
    };
    std::stack<BN> st;
    st.push({0, tris.begin(), tris.end()});
    while (!st.empty()) {
        BN curr = st.top();
        st.pop();
        std::array<BN, BF> ch;
        ch[0].begin = curr.begin;
        ch[0].end = curr.end;
        int n_children = 1;
        while (n_children < BF) {
            for (int i = n_children - 1; i >= 0; --i) {
                auto& c = ch[i];
                float n = (float)std::distance(c.begin, c.end);
                V3 mean = v3(0, 0, 0);
                for (auto it = c.begin; it != c.end; ++it) mean = add(mean, centroid(*it));
                mean = v3(mean.x / n, mean.y / n, mean.z / n);
                V3 var = v3(0, 0, 0);
                for (auto it = c.begin; it != c.end; ++it) {
                    V3 d = sub(centroid(*it), mean);
                    var = v3(var.x + d.x * d.x, var.y + d.y * d.y, var.z + d.z * d.z);
                }
                var = v3(var.x / n, var.y / n, var.z / n);
                int axis = 0;
                float vm = var.x;
                if (var.y > vm) { axis = 1; vm = var.y; }
                if (var.z > vm) { axis = 2; }
                auto m = c.begin + std::distance(c.begin, c.end) / 2;
                std::nth_element(c.begin, m, c.end, [&](const Tri& t1, const Tri& t2) {
                    return centroid_axis(t1, axis) < centroid_axis(t2, axis);
                });
                ch[i * 2].begin = ch[i].begin;
                ch[i * 2 + 1].end = ch[i].end;
                ch[i * 2].end = ch[i * 2 + 1].begin = m;
            }
            n_children *= 2;
        }
        nodes[curr.node_idx].left_idx = (int)nodes.size();
        for (int i = 0; i < BF; ++i) {
            auto& c = ch[i];
            c.node_idx = (int)nodes.size();
            nodes.emplace_back();
            nodes.back().bb = box_of(c.begin, c.end);
            if ((uint32_t)std::distance(c.begin, c.end) <= n_per_leaf) {
                nodes.back().left_idx = -(int)std::distance(tris.begin(), c.begin) - 1;
                nodes.back().right_idx = -(int)std::distance(tris.begin(), c.end) - 1;
            } else {
                st.push(c);
            }
        }
        nodes[curr.node_idx].right_idx = (int)nodes.size();
    }
    for (auto& n : nodes) n.escape_idx = -1;
    std::function<void(int, int)> thread_bvh = [&](int ni, int esc) {
        RefNode& nd = nodes[ni];
        nd.escape_idx = esc;
        if (nd.left_idx < 0) return;
        int first = nd.left_idx, end = nd.right_idx;
        for (int c = first; c < end; ++c) thread_bvh(c, (c + 1 < end) ? (c + 1) : esc);
    };
    if (!nodes.empty()) thread_bvh(0, -1);
}

struct DistIdx {
    float dist;
    int idx;
};
// sorting network of src/bvh.cu:36-59 (descending order)
inline void cas(DistIdx& a, DistIdx& b) {
    if (a.dist < b.dist) std::swap(a, b);
}
inline void sort4_desc(DistIdx v[4]) {
    cas(v[0], v[2]);
    cas(v[1], v[3]);
    cas(v[0], v[1]);
    cas(v[2], v[3]);
    cas(v[1], v[2]);
}

struct Counts {
    int64_t box = 0;
    int64_t tri = 0;
};

// Fixed stack of 32 (include/gpu/bvh.cuh:24-57); overflow is reported by flag.
struct Stack32 {
    int e[32];
    int n = 0;
    bool over = false;
    void push(int v) {
        if (n >= 32) { over = true; return; }
        e[n++] = v;
    }
    int pop() { return e[--n]; }
    bool empty() const { return n <= 0; }
};

// Stackless escape-link traversal (src/bvh.cu:178-213), used on stack overflow.
std::pair<int, float> ref_closest_stackless(const RefBvh& b, V3 p, float maxd2, Counts* c, int tie_mode,
                                            bool* miss = nullptr) {
    float best = maxd2;
    int bi = -1;
    int idx = 0;
    while (idx != -1) {
        const RefNode& nd = b.nodes[idx];
        if (c) c->box++;
        if (tie_mode == 0 ? box_dist2(nd.bb, p) > best : box_dist2_padded(nd.bb, p) > best * (1.0f + 1.0f / 4096.0f)) {
            idx = nd.escape_idx;
            continue;
        }
        if (nd.left_idx < 0) {
            int end = -nd.right_idx - 1;
            for (int i = -nd.left_idx - 1; i < end; ++i) {
                float d2 = tri_dist2(b.tris[i], p);
                if (c) c->tri++;
                bool take = tie_mode == 0 ? (d2 <= best)
                                          : (d2 < best || (d2 == best && bi >= 0 && b.tris[i].id < b.tris[bi].id));
                if (take) { best = d2; bi = i; }
            }
            idx = nd.escape_idx;
        } else {
            idx = nd.left_idx;
        }
    }
    if (miss) *miss = (bi == -1);
    if (bi == -1) { bi = 0; best = 0.0f; }
    return {bi, std::sqrt(best)};
}

// src/bvh.cu:300-353.  tie_mode 0: reference rule (d <= best: last visited wins);
// tie_mode 1: canonical rule (ties to the lowest original face id).
std::pair<int, float> ref_closest(const RefBvh& b, V3 p, float maxd2, Counts* c, int tie_mode, bool* miss) {
    Stack32 st;
    st.push(0);
    float best = maxd2;
    int bi = -1;
    while (!st.empty()) {
        int idx = st.pop();
        const RefNode& nd = b.nodes[idx];
        if (nd.left_idx < 0) {
            int end = -nd.right_idx - 1;
            for (int i = -nd.left_idx - 1; i < end; ++i) {
                float d2 = tri_dist2(b.tris[i], p);
                if (c) c->tri++;
                bool take = tie_mode == 0 ? (d2 <= best)
                                          : (d2 < best || (d2 == best && bi >= 0 && b.tris[i].id < b.tris[bi].id));
                if (take) { best = d2; bi = i; }
            }
        } else {
            DistIdx ch[4];
            int first = nd.left_idx;
            for (int i = 0; i < 4; ++i)
                ch[i] = {tie_mode == 0 ? box_dist2(b.nodes[first + i].bb, p) : box_dist2_padded(b.nodes[first + i].bb, p),
                         first + i};
            if (c) c->box += 4;
            sort4_desc(ch);
            // canonical mode: boxes within a relative 2^-12 of the best are visited too, so
            // that a face tied with the best is never skipped for a rounding of its box
            const float lim = tie_mode == 0 ? best : best * (1.0f + 1.0f / 4096.0f);
            for (int i = 0; i < 4; ++i)
                if (ch[i].dist <= lim) st.push(ch[i].idx);
        }
        if (st.over) return ref_closest_stackless(b, p, best, c, tie_mode);
    }
    if (miss) *miss = (bi == -1);
    if (bi == -1) { bi = 0; best = 0.0f; }
    return {bi, std::sqrt(best)};
}

// src/bvh.cu:145-176
std::pair<int, float> ref_ray_stackless(const RefBvh& b, V3 ro, V3 rd, Counts* c, int tie_mode) {
    float mint = MAX_DIST;
    int bi = -1;
    int idx = 0;
    while (idx != -1) {
        const RefNode& nd = b.nodes[idx];
        if (c) c->box++;
        float tb = tie_mode == 0 ? box_ray(nd.bb, ro, rd) : box_ray_conservative(nd.bb, ro, rd);
        if (tie_mode == 0 ? (tb >= mint) : (tb > mint)) { idx = nd.escape_idx; continue; }
        if (nd.left_idx < 0) {
            int end = -nd.right_idx - 1;
            for (int i = -nd.left_idx - 1; i < end; ++i) {
                float t = tri_ray(b.tris[i], ro, rd, tie_mode != 0);
                if (c) c->tri++;
                bool take = tie_mode == 0 ? (t < mint)
                                          : (t < mint || (t == mint && bi >= 0 && b.tris[i].id < b.tris[bi].id));
                if (take) { mint = t; bi = i; }
            }
            idx = nd.escape_idx;
        } else {
            idx = nd.left_idx;
        }
    }
    return {bi, mint};
}

// src/bvh.cu:250-298.  tie_mode 0: reference (t < mint: first visited wins);
// tie_mode 1: canonical (ties to the lowest original face id; boxes visited on <=).
std::pair<int, float> ref_ray(const RefBvh& b, V3 ro, V3 rd, Counts* c, int tie_mode) {
    Stack32 st;
    st.push(0);
    float mint = MAX_DIST;
    int bi = -1;
    while (!st.empty()) {
        int idx = st.pop();
        const RefNode& nd = b.nodes[idx];
        if (nd.left_idx < 0) {
            int end = -nd.right_idx - 1;
            for (int i = -nd.left_idx - 1; i < end; ++i) {
                float t = tri_ray(b.tris[i], ro, rd, tie_mode != 0);
                if (c) c->tri++;
                bool take = tie_mode == 0 ? (t < mint)
                                          : (t < mint || (t == mint && bi >= 0 && b.tris[i].id < b.tris[bi].id));
                if (take) { mint = t; bi = i; }
            }
        } else {
            DistIdx ch[4];
            int first = nd.left_idx;
            for (int i = 0; i < 4; ++i)
                ch[i] = {tie_mode == 0 ? box_ray(b.nodes[first + i].bb, ro, rd)
                                       : box_ray_conservative(b.nodes[first + i].bb, ro, rd),
                         first + i};
            if (c) c->box += 4;
            sort4_desc(ch);
            for (int i = 0; i < 4; ++i)
                if (tie_mode == 0 ? (ch[i].dist < mint) : (ch[i].dist <= mint)) st.push(ch[i].idx);
        }
        if (st.over) return ref_ray_stackless(b, ro, rd, c, tie_mode);
    }
    return {bi, mint};
}

// src/bvh.cu:216-248
V3 ref_avg_normal_stackless(const RefBvh& b, V3 p, Counts* c) {
    const float EPS = 1e-6f;
    float tw = 0;
    V3 r = v3(0, 0, 0);
    int idx = 0;
    while (idx != -1) {
        const RefNode& nd = b.nodes[idx];
        if (c) c->box++;
        if (box_dist2(nd.bb, p) >= EPS) { idx = nd.escape_idx; continue; }
        if (nd.left_idx < 0) {
            int end = -nd.right_idx - 1;
            for (int i = -nd.left_idx - 1; i < end; ++i) {
                if (c) c->tri++;
                if (tri_dist2(b.tris[i], p) < EPS) {
                    r = add(r, tri_normal(b.tris[i]));
                    tw += 1;
                }
            }
            idx = nd.escape_idx;
        } else {
            idx = nd.left_idx;
        }
    }
    return v3(r.x / tw, r.y / tw, r.z / tw);
}

// src/bvh.cu:356-396 (fp32 sum in visit order, as the reference does)
V3 ref_avg_normal(const RefBvh& b, V3 p, Counts* c) {
    const float EPS = 1e-6f;
    Stack32 st;
    st.push(0);
    float tw = 0;
    V3 r = v3(0, 0, 0);
    while (!st.empty()) {
        int idx = st.pop();
        const RefNode& nd = b.nodes[idx];
        if (nd.left_idx < 0) {
            int end = -nd.right_idx - 1;
            for (int i = -nd.left_idx - 1; i < end; ++i) {
                if (c) c->tri++;
                if (tri_dist2(b.tris[i], p) < EPS) {
                    r = add(r, tri_normal(b.tris[i]));
                    tw += 1;
                }
            }
        } else {
            int first = nd.left_idx;
            if (c) c->box += 4;
            for (int i = 0; i < 4; ++i)
                if (box_dist2(b.nodes[first + i].bb, p) < EPS) st.push(first + i);
        }
    }
    if (st.over) return ref_avg_normal_stackless(b, p, c);
    return v3(r.x / tw, r.y / tw, r.z / tw);
}

inline void epilogue_uvw(const Tri& t, V3 p, float* uvw, int64_t i) {
    if (!uvw) return;
    V3 cp = tri_closest_pt(t, p);
    store3(uvw, i, tri_barycentric(t, cp));
}

}  // namespace

// ============================================================================
// C ABI (loaded with ctypes by the tests and bench.py's cpu_baseline leg)
// ============================================================================
extern "C" {

int orc_num_threads(void) { return set_threads(0); }

// ---- single primitives (for known-answer tests) ----------------------------
void orc_tri_dist2(const float* tri9, const float* pts, int64_t n, float* out) {
    Tri t{v3(tri9[0], tri9[1], tri9[2]), v3(tri9[3], tri9[4], tri9[5]), v3(tri9[6], tri9[7], tri9[8]), 0};
    for (int64_t i = 0; i < n; ++i) out[i] = tri_dist2(t, load3(pts, i));
}
void orc_tri_closest(const float* tri9, const float* pts, int64_t n, float* cp, float* uvw) {
    Tri t{v3(tri9[0], tri9[1], tri9[2]), v3(tri9[3], tri9[4], tri9[5]), v3(tri9[6], tri9[7], tri9[8]), 0};
    for (int64_t i = 0; i < n; ++i) {
        V3 c = tri_closest_pt(t, load3(pts, i));
        store3(cp, i, c);
        store3(uvw, i, tri_barycentric(t, c));
    }
}
void orc_tri_ray(const float* tri9, const float* ro, const float* rd, int64_t n, float* out, int exact_div) {
    Tri t{v3(tri9[0], tri9[1], tri9[2]), v3(tri9[3], tri9[4], tri9[5]), v3(tri9[6], tri9[7], tri9[8]), 0};
    for (int64_t i = 0; i < n; ++i) out[i] = tri_ray(t, load3(ro, i), load3(rd, i), exact_div != 0);
}
float orc_safe_divide(float a, float b) { return safe_divide(a, b); }

// ---- PCG32 / stab directions ------------------------------------------------
void orc_pcg32_uints(int64_t advance, int64_t n, uint32_t* out) {
    Pcg32 r;
    r.advance(advance);
    for (int64_t i = 0; i < n; ++i) out[i] = r.next_uint();
}
void orc_pcg32_offsets(int64_t n, float* out) {  // per query i: advance(2i), 2 floats
    for (int64_t i = 0; i < n; ++i) {
        Pcg32 r;
        r.advance((int64_t)(uint32_t)((uint32_t)i * 2u));
        out[2 * i] = r.next_float();
        out[2 * i + 1] = r.next_float();
    }
}
void orc_fibonacci_dirs32(float ox, float oy, float* out) {
    for (uint32_t i = 0; i < 32; ++i) store3(out, i, fibonacci_dir32(i, ox, oy));
}

// ---- fp32 brute force (canonical: ties to the lowest face index) ------------
// Returns 0 on success, -1 on a bad face index.
int orc_udf_brute(const float* V, int64_t nV, const int32_t* F, int64_t nF, const float* pts, int64_t n,
                  float* dist, int64_t* face, float* uvw, int nthreads) {
    auto tris = gather(V, nV, F, nF);
    if (tris.empty()) return -1;
    int nt = set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
    for (int64_t i = 0; i < n; ++i) {
        V3 p = load3(pts, i);
        float d2;
        int64_t bi = brute_nearest(tris, p, &d2);
        if (bi < 0) { bi = 0; d2 = 0.0f; }
        dist[i] = std::sqrt(d2);
        face[i] = bi;
        epilogue_uvw(tris[bi], p, uvw, i);
    }
    return 0;
}

int orc_ray_brute(const float* V, int64_t nV, const int32_t* F, int64_t nF, const float* ro, const float* rd,
                  int64_t n, float* pos, int64_t* face, float* depth, int nthreads) {
    auto tris = gather(V, nV, F, nF);
    if (tris.empty()) return -1;
    int nt = set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
    for (int64_t i = 0; i < n; ++i) {
        V3 o = load3(ro, i), d = load3(rd, i);
        float mint = MAX_DIST;
        int64_t bi = -1;
        for (size_t k = 0; k < tris.size(); ++k) {
            float t = tri_ray(tris[k], o, d);
            if (t < mint || (t == mint && bi >= 0 && (int64_t)k < bi)) { mint = t; bi = (int64_t)k; }
        }
        depth[i] = mint;
        face[i] = bi;
        store3(pos, i, v3(std::fmaf(mint, d.x, o.x), std::fmaf(mint, d.y, o.y), std::fmaf(mint, d.z, o.z)));
    }
    return 0;
}

// mode 0 = watertight, otherwise raystab (src/bvh.cu:430-466).  The raystab RNG is
// keyed by the query index i within this call (src/bvh.cu:663-668).
int orc_sdf_brute(const float* V, int64_t nV, const int32_t* F, int64_t nF, const float* pts, int64_t n,
                  uint32_t mode, float* dist, int64_t* face, float* uvw, int nthreads) {
    auto tris = gather(V, nV, F, nF);
    if (tris.empty()) return -1;
    int nt = set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (int64_t i = 0; i < n; ++i) {
        V3 p = load3(pts, i);
        float d2;
        int64_t bi = brute_nearest(tris, p, &d2);
        if (bi < 0) { bi = 0; d2 = 0.0f; }
        float d = std::sqrt(d2);
        const Tri& t = tris[bi];
        if (mode == 0) {
            V3 cp = tri_closest_pt(t, p);
            V3 nrm = brute_avg_normal(tris, cp);
            d = std::copysign(d, dot(nrm, sub(p, cp)));
        } else {
            Pcg32 r;
            r.advance((int64_t)(uint32_t)((uint32_t)i * 2u));
            float ox = r.next_float(), oy = r.next_float();
            bool inside = true;
            for (uint32_t k = 0; k < 32 && inside; ++k) {
                V3 dd = fibonacci_dir32(k, ox, oy);
                if (!brute_any_hit(tris, p, neg(dd)) || !brute_any_hit(tris, p, dd)) inside = false;
            }
            if (inside) d = -d;
        }
        dist[i] = d;
        face[i] = bi;
        epilogue_uvw(t, p, uvw, i);
    }
    return 0;
}

// ---- FP64 brute force ---------------------------------------------------------
// Closest point on triangle abc to p (Ericson 2005, 5.1.5), returns barycentrics.
static void closest_f64(const double* a, const double* b, const double* c, const double* p, double* q,
                        double* bary) {
    double ab[3], ac[3], ap[3];
    for (int k = 0; k < 3; ++k) { ab[k] = b[k] - a[k]; ac[k] = c[k] - a[k]; ap[k] = p[k] - a[k]; }
    auto d3 = [](const double* x, const double* y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
    double d1 = d3(ab, ap), d2 = d3(ac, ap);
    double u, v, w;
    if (d1 <= 0 && d2 <= 0) { u = 1; v = 0; w = 0; goto done; }
    {
        double bp[3];
        for (int k = 0; k < 3; ++k) bp[k] = p[k] - b[k];
        double d3v = d3(ab, bp), d4 = d3(ac, bp);
        if (d3v >= 0 && d4 <= d3v) { u = 0; v = 1; w = 0; goto done; }
        double vc = d1 * d4 - d3v * d2;
        if (vc <= 0 && d1 >= 0 && d3v <= 0) {
            double t = d1 / (d1 - d3v);
            u = 1 - t; v = t; w = 0; goto done;
        }
        double cp[3];
        for (int k = 0; k < 3; ++k) cp[k] = p[k] - c[k];
        double d5 = d3(ab, cp), d6 = d3(ac, cp);
        if (d6 >= 0 && d5 <= d6) { u = 0; v = 0; w = 1; goto done; }
        double vb = d5 * d2 - d1 * d6;
        if (vb <= 0 && d2 >= 0 && d6 <= 0) {
            double t = d2 / (d2 - d6);
            u = 1 - t; v = 0; w = t; goto done;
        }
        double va = d3v * d6 - d5 * d4;
        if (va <= 0 && (d4 - d3v) >= 0 && (d5 - d6) >= 0) {
            double t = (d4 - d3v) / ((d4 - d3v) + (d5 - d6));
            u = 0; v = 1 - t; w = t; goto done;
        }
        double den = 1.0 / (va + vb + vc);
        v = vb * den; w = vc * den; u = 1 - v - w;
    }
done:
    for (int k = 0; k < 3; ++k) q[k] = u * a[k] + v * b[k] + w * c[k];
    bary[0] = u; bary[1] = v; bary[2] = w;
}

// dist: FP64 distance to the nearest face; face: its index (lowest on exact ties);
// cpt: closest point; gap: distance of the nearest *other* face minus dist.
int orc_udf_f64(const float* V, int64_t nV, const int32_t* F, int64_t nF, const float* pts, int64_t n,
                double* dist, int64_t* face, double* cpt, double* gap, int nthreads) {
    for (int64_t i = 0; i < 3 * nF; ++i)
        if (F[i] < 0 || F[i] >= nV) return -1;
    int nt = set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (int64_t i = 0; i < n; ++i) {
        double p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
        double best = 1e300, second = 1e300, bq[3] = {0, 0, 0};
        int64_t bi = -1;
        for (int64_t f = 0; f < nF; ++f) {
            double a[3], b[3], c[3], q[3], bc[3];
            for (int k = 0; k < 3; ++k) {
                a[k] = V[3 * F[3 * f] + k];
                b[k] = V[3 * F[3 * f + 1] + k];
                c[k] = V[3 * F[3 * f + 2] + k];
            }
            closest_f64(a, b, c, p, q, bc);
            double d = std::sqrt((p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) +
                                 (p[2] - q[2]) * (p[2] - q[2]));
            if (d < best) {
                second = best;
                best = d;
                bi = f;
                bq[0] = q[0]; bq[1] = q[1]; bq[2] = q[2];
            } else if (d < second) {
                second = d;
            }
        }
        dist[i] = best;
        face[i] = bi;
        if (cpt) { cpt[3 * i] = bq[0]; cpt[3 * i + 1] = bq[1]; cpt[3 * i + 2] = bq[2]; }
        if (gap) gap[i] = second - best;
    }
    return 0;
}

// Moller-Trumbore in FP64.  t: nearest hit (1000 on a miss), face (-1 on a miss),
// margin: for the hit face, min(u, v, 1-u-v, t); ambiguous: 1 if another face has
// a near-boundary decision (barycentric within tol of an edge) at t <= hit t + tol.
int orc_ray_f64(const float* V, int64_t nV, const int32_t* F, int64_t nF, const float* ro, const float* rd,
                int64_t n, double tol, double* tout, int64_t* face, uint8_t* ambiguous, int nthreads) {
    for (int64_t i = 0; i < 3 * nF; ++i)
        if (F[i] < 0 || F[i] >= nV) return -1;
    int nt = set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (int64_t i = 0; i < n; ++i) {
        double o[3] = {ro[3 * i], ro[3 * i + 1], ro[3 * i + 2]};
        double d[3] = {rd[3 * i], rd[3 * i + 1], rd[3 * i + 2]};
        double best = 1000.0;
        int64_t bi = -1;
        double marg_t = 1e300;  // smallest t of a marginal candidate
        for (int64_t f = 0; f < nF; ++f) {
            double a[3], e1[3], e2[3], s[3];
            for (int k = 0; k < 3; ++k) {
                a[k] = V[3 * F[3 * f] + k];
                e1[k] = V[3 * F[3 * f + 1] + k] - a[k];
                e2[k] = V[3 * F[3 * f + 2] + k] - a[k];
                s[k] = o[k] - a[k];
            }
            double pv[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
            double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
            if (det == 0.0) continue;
            double inv = 1.0 / det;
            double u = (s[0] * pv[0] + s[1] * pv[1] + s[2] * pv[2]) * inv;
            double qv[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
            double v = (d[0] * qv[0] + d[1] * qv[1] + d[2] * qv[2]) * inv;
            double t = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * inv;
            double w = 1 - u - v;
            double m = std::min(std::min(u, v), std::min(w, t));
            bool hit = (m >= 0) && t < 1000.0;
            if (std::fabs(m) <= tol && t < 1000.0 + tol) marg_t = std::min(marg_t, t);
            if (hit && t < best) { best = t; bi = f; }
        }
        tout[i] = best;
        face[i] = bi;
        ambiguous[i] = (marg_t <= best + tol) ? 1 : 0;
    }
    return 0;
}

// Generalized winding number (sum of signed solid angles / 4 pi), FP64.
int orc_winding_f64(const float* V, int64_t nV, const int32_t* F, int64_t nF, const float* pts, int64_t n,
                    double* w, int nthreads) {
    for (int64_t i = 0; i < 3 * nF; ++i)
        if (F[i] < 0 || F[i] >= nV) return -1;
    int nt = set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (int64_t i = 0; i < n; ++i) {
        double acc = 0;
        for (int64_t f = 0; f < nF; ++f) {
            double a[3], b[3], c[3];
            for (int k = 0; k < 3; ++k) {
                a[k] = V[3 * F[3 * f] + k] - pts[3 * i + k];
                b[k] = V[3 * F[3 * f + 1] + k] - pts[3 * i + k];
                c[k] = V[3 * F[3 * f + 2] + k] - pts[3 * i + k];
            }
            double la = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
            double lb = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
            double lc = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            double det = a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                         a[2] * (b[0] * c[1] - b[1] * c[0]);
            double ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
            double bc = b[0] * c[0] + b[1] * c[1] + b[2] * c[2];
            double ca = c[0] * a[0] + c[1] * a[1] + c[2] * a[2];
            double den = la * lb * lc + ab * lc + bc * la + ca * lb;
            acc += 2.0 * std::atan2(det, den);
        }
        w[i] = acc / (4.0 * M_PI);
    }
    return 0;
}

// ---- reference BVH4 -------------------------------------------------------------
void* orc_refbvh_build(const float* V, int64_t nV, const int32_t* F, int64_t nF) {
    auto tris = gather(V, nV, F, nF);
    if (tris.empty()) return nullptr;
    RefBvh* b = new RefBvh;
    b->tris = std::move(tris);
    ref_build(*b, 8);  // src/api_gpu.cu:36
    return b;
}
void orc_refbvh_free(void* h) { delete (RefBvh*)h; }
int64_t orc_refbvh_num_nodes(void* h) { return (int64_t)((RefBvh*)h)->nodes.size(); }

// Export in the reference state_dict layout: nodes int32 [n, 9], triangles int32 [M, 12]
// (include/gpu/bvh.cuh:84-94, src/api_gpu.cu:110-121).  The 4 padding bytes are zeroed.
void orc_refbvh_export(void* h, int32_t* nodes9, int32_t* tris12) {
    RefBvh* b = (RefBvh*)h;
    std::memcpy(nodes9, b->nodes.data(), b->nodes.size() * sizeof(RefNode));
    for (size_t i = 0; i < b->tris.size(); ++i) {
        RefTriRecord r;
        std::memset(&r, 0, sizeof r);
        const Tri& t = b->tris[i];
        r.a[0] = t.a.x; r.a[1] = t.a.y; r.a[2] = t.a.z;
        r.b[0] = t.b.x; r.b[1] = t.b.y; r.b[2] = t.b.z;
        r.c[0] = t.c.x; r.c[1] = t.c.y; r.c[2] = t.c.z;
        r.id = t.id;
        std::memcpy(tris12 + 12 * i, &r, sizeof r);
    }
}

// A reference-layout tree from state_dict arrays (src/api_gpu.cu:47-68): nodes int32
// [n, 9], triangles int32 [M, 12].  No validation beyond the sizes: test input only.
void* orc_refbvh_import(const int32_t* nodes9, int64_t n_nodes, const int32_t* tris12, int64_t n_tris) {
    if (n_nodes <= 0 || n_tris <= 0) return nullptr;
    RefBvh* b = new RefBvh;
    b->nodes.resize((size_t)n_nodes);
    std::memcpy(b->nodes.data(), nodes9, (size_t)n_nodes * sizeof(RefNode));
    b->tris.resize((size_t)n_tris);
    for (int64_t i = 0; i < n_tris; ++i) {
        RefTriRecord r;
        std::memcpy(&r, tris12 + 12 * i, sizeof r);
        Tri& t = b->tris[(size_t)i];
        t.a = v3(r.a[0], r.a[1], r.a[2]);
        t.b = v3(r.b[0], r.b[1], r.b[2]);
        t.c = v3(r.c[0], r.c[1], r.c[2]);
        t.id = r.id;
    }
    return b;
}

// Queries through the reference BVH.  tie_mode 0 = reference rule, 1 = canonical;
// plus 2 = walk with the escape links only.
// counts (nullable): [box tests, triangle tests] summed over all queries.
int orc_refbvh_udf(void* h, const float* pts, int64_t n, float* dist, int64_t* face, float* uvw, int tie_mode,
                   int64_t* counts, int nthreads) {
    const RefBvh& b = *(RefBvh*)h;
    int nt = set_threads(nthreads);
    int64_t cb = 0, ctr = 0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt) reduction(+ : cb, ctr)
    for (int64_t i = 0; i < n; ++i) {
        V3 p = load3(pts, i);
        Counts c;
        bool miss = false;
        // tie_mode bit 1: walk with the escape links only (the reference's overflow path)
        const int tm = tie_mode & 1;
        auto r = (tie_mode & 2) ? ref_closest_stackless(b, p, MAX_DIST_SQ, counts ? &c : nullptr, tm, &miss)
                                : ref_closest(b, p, MAX_DIST_SQ, counts ? &c : nullptr, tm, &miss);
        int ti = r.first;
        if (miss && tm == 1) {  // canonical: the miss path reports face 0
            for (size_t k = 0; k < b.tris.size(); ++k)
                if (b.tris[k].id == 0) { ti = (int)k; break; }
        }
        dist[i] = r.second;
        face[i] = b.tris[ti].id;
        epilogue_uvw(b.tris[ti], p, uvw, i);
        cb += c.box;
        ctr += c.tri;
    }
    if (counts) { counts[0] = cb; counts[1] = ctr; }
    return 0;
}

int orc_refbvh_ray(void* h, const float* ro, const float* rd, int64_t n, float* pos, int64_t* face, float* depth,
                   int tie_mode, int64_t* counts, int nthreads) {
    const RefBvh& b = *(RefBvh*)h;
    int nt = set_threads(nthreads);
    int64_t cb = 0, ctr = 0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt) reduction(+ : cb, ctr)
    for (int64_t i = 0; i < n; ++i) {
        V3 o = load3(ro, i), d = load3(rd, i);
        Counts c;
        auto r = (tie_mode & 2) ? ref_ray_stackless(b, o, d, counts ? &c : nullptr, tie_mode & 1)
                                : ref_ray(b, o, d, counts ? &c : nullptr, tie_mode & 1);
        depth[i] = r.second;
        store3(pos, i, v3(std::fmaf(r.second, d.x, o.x), std::fmaf(r.second, d.y, o.y), std::fmaf(r.second, d.z, o.z)));
        face[i] = r.first >= 0 ? b.tris[r.first].id : -1;
        cb += c.box;
        ctr += c.tri;
    }
    if (counts) { counts[0] = cb; counts[1] = ctr; }
    return 0;
}

// src/bvh.cu:398-427 and kernels 631-685.  counts: [box, tri] of the distance
// traversal, [box, tri] of the sign work (normal averaging or stab rays).
int orc_refbvh_sdf(void* h, const float* pts, int64_t n, uint32_t mode, float* dist, int64_t* face, float* uvw,
                   int64_t* counts, int nthreads) {
    const RefBvh& b = *(RefBvh*)h;
    int nt = set_threads(nthreads);
    int64_t cb = 0, ctr = 0, sb = 0, st = 0;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt) reduction(+ : cb, ctr, sb, st)
    for (int64_t i = 0; i < n; ++i) {
        V3 p = load3(pts, i);
        Counts c, s;
        auto r = ref_closest(b, p, MAX_DIST_SQ, counts ? &c : nullptr, 0, nullptr);
        const Tri& t = b.tris[r.first];
        float d = r.second;
        if (mode == 0) {
            V3 cp = tri_closest_pt(t, p);
            V3 nrm = ref_avg_normal(b, cp, counts ? &s : nullptr);
            d = std::copysign(d, dot(nrm, sub(p, cp)));
        } else {
            Pcg32 rng;
            rng.advance((int64_t)(uint32_t)((uint32_t)i * 2u));
            float ox = rng.next_float(), oy = rng.next_float();
            bool inside = true;
            for (uint32_t k = 0; k < 32; ++k) {
                V3 dd = fibonacci_dir32(k, ox, oy);
                if (ref_ray(b, p, neg(dd), counts ? &s : nullptr, 0).first < 0 ||
                    ref_ray(b, p, dd, counts ? &s : nullptr, 0).first < 0) {
                    inside = false;
                    break;
                }
            }
            if (inside) d = -d;
        }
        dist[i] = d;
        face[i] = t.id;
        epilogue_uvw(t, p, uvw, i);
        cb += c.box; ctr += c.tri; sb += s.box; st += s.tri;
    }
    if (counts) { counts[0] = cb; counts[1] = ctr; counts[2] = sb; counts[3] = st; }
    return 0;
}

}  // extern "C"

namespace {
// Canonical watertight normal at c through the reference tree: every face with
// tri_dist2(c) < 1e-6 (children pruned with the padded box, which bounds every face's
// fp32 squared distance from below), unit normals summed in double, then rounded and
// divided by the count -- the same sum as brute_avg_normal over all faces.
V3 ref_avg_normal_canon(const RefBvh& b, V3 c) {
    const float EPS = 1e-6f;
    std::vector<int> st{0};
    double sx = 0, sy = 0, sz = 0;
    float cnt = 0;
    while (!st.empty()) {
        const RefNode& nd = b.nodes[st.back()];
        st.pop_back();
        if (nd.left_idx < 0) {
            int end = -nd.right_idx - 1;
            for (int i = -nd.left_idx - 1; i < end; ++i)
                if (tri_dist2(b.tris[i], c) < EPS) {
                    V3 n = tri_normal(b.tris[i]);
                    sx += n.x;
                    sy += n.y;
                    sz += n.z;
                    cnt += 1.0f;
                }
        } else {
            for (int k = 0; k < 4; ++k)
                if (box_dist2_padded(b.nodes[nd.left_idx + k].bb, c) < EPS) st.push_back(nd.left_idx + k);
        }
    }
    return v3((float)sx / cnt, (float)sy / cnt, (float)sz / cnt);
}
}  // namespace

extern "C" {

// Signed distance through the reference tree for queries that carry their raystab RNG
// index explicitly: qidx[i] is the index query i has within the caller's whole call
// (src/bvh.cu:663-668 keys the generator by it), so a sample of a large call can be
// checked on its own.  qidx NULL means 0..n-1.
// rule 0: the reference's own rules (last-visited distance ties, fp32 normal sum in
//         visit order, safe_divide intersector and slab test), as orc_refbvh_sdf;
// rule 1: canonical, the product's bit-exact target (distance ties to the lowest face
//         id, face 0 when nothing is within 1000, normal sum in double over every face
//         within squared distance 1e-6, exact-division intersector, conservative slabs).
int orc_refbvh_sdf_idx(void* h, const float* pts, const int64_t* qidx, int64_t n, uint32_t mode, int rule,
                       float* dist, int64_t* face, float* uvw, int nthreads) {
    const RefBvh& b = *(RefBvh*)h;
    const int tm = rule ? 1 : 0;
    int face0 = 0;
    if (tm == 1)
        for (size_t k = 0; k < b.tris.size(); ++k)
            if (b.tris[k].id == 0) { face0 = (int)k; break; }
    int nt = set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (int64_t i = 0; i < n; ++i) {
        V3 p = load3(pts, i);
        bool miss = false;
        auto r = ref_closest(b, p, MAX_DIST_SQ, nullptr, tm, &miss);
        int ti = (miss && tm == 1) ? face0 : r.first;
        const Tri& t = b.tris[ti];
        float d = r.second;
        if (mode == 0) {
            V3 cp = tri_closest_pt(t, p);
            V3 nrm = tm ? ref_avg_normal_canon(b, cp) : ref_avg_normal(b, cp, nullptr);
            d = std::copysign(d, dot(nrm, sub(p, cp)));
        } else {
            const int64_t q = qidx ? qidx[i] : i;
            Pcg32 rng;
            rng.advance((int64_t)(uint32_t)((uint32_t)q * 2u));
            float ox = rng.next_float(), oy = rng.next_float();
            bool inside = true;
            for (uint32_t k = 0; k < 32; ++k) {
                V3 dd = fibonacci_dir32(k, ox, oy);
                if (ref_ray(b, p, neg(dd), nullptr, tm).first < 0 || ref_ray(b, p, dd, nullptr, tm).first < 0) {
                    inside = false;
                    break;
                }
            }
            if (inside) d = -d;
        }
        dist[i] = d;
        face[i] = t.id;
        epilogue_uvw(t, p, uvw, i);
    }
    return 0;
}

// ---- floodfill ----------------------------------------------------------------
// Restatement of include/gpu/floodfill.cuh:157-272 for one batch (labels global
// across the batch).  Neighbour decoding follows the reference (x = local % W,
// y = (local / W) % H, z = local / (W*H)), which is tensor-order 6-connectivity
// only when H == W == D.  Hook passes are Jacobi sweeps (read old labels, write new);
// the compress schedule and MAX_ITERATIONS are the reference's.  Returns the number
// of iterations run, or -1 if MAX_ITERATIONS was reached without convergence.
int orc_floodfill_ref(const uint8_t* grid, int B, int H, int W, int D, int32_t* labels) {
    const int64_t Nvol = (int64_t)H * W * D, Ntot = (int64_t)B * Nvol;
    for (int64_t i = 0; i < Ntot; ++i) labels[i] = grid[i] ? -1 : (int32_t)i;
    const int diag = (int)std::ceil(std::sqrt((double)(H * H + W * W + D * D)));
    const int base = (H <= 4 && W <= 4 && D <= 4) ? 20 : diag + 50;
    const int MAXIT = std::min(base, 10 * std::max(H, std::max(W, D)));
    std::vector<int32_t> nl(labels, labels + Ntot);
    auto compress = [&]() {
        for (int64_t idx = 0; idx < Ntot; ++idx) {
            if (labels[idx] < 0) continue;
            int64_t cur = idx;
            int32_t root = labels[cur];
            int steps = 0;
            while (root != cur && root >= 0 && root < Ntot && steps < 1000) {
                cur = root;
                root = labels[cur];
                steps++;
            }
            if (steps >= 1000) { labels[idx] = (int32_t)idx; continue; }
            cur = idx;
            steps = 0;
            while (cur != root && cur >= 0 && cur < Ntot && steps < 1000) {
                int32_t nx = labels[cur];
                labels[cur] = root;
                cur = nx;
                steps++;
            }
        }
    };
    int it = 0;
    bool changed = true;
    while (changed && it < MAXIT) {
        changed = false;
        std::memcpy(nl.data(), labels, Ntot * sizeof(int32_t));
        for (int64_t idx = 0; idx < Ntot; ++idx) {
            if (grid[idx]) continue;
            int64_t b = idx / Nvol, local = idx % Nvol, bb = b * Nvol;
            int x = (int)(local % W), y = (int)((local / W) % H), z = (int)(local / ((int64_t)W * H));
            int32_t best = labels[idx];
            auto nb = [&](int64_t nidx) {
                if (nidx >= bb && nidx < bb + Nvol && !grid[nidx] && labels[nidx] >= 0) best = std::min(best, labels[nidx]);
            };
            if (x > 0) nb(idx - 1);
            if (x < W - 1) nb(idx + 1);
            if (y > 0) nb(idx - W);
            if (y < H - 1) nb(idx + W);
            if (z > 0) nb(idx - (int64_t)W * H);
            if (z < D - 1) nb(idx + (int64_t)W * H);
            if (labels[idx] >= 0 && best < labels[idx]) { nl[idx] = best; changed = true; }
        }
        std::memcpy(labels, nl.data(), Ntot * sizeof(int32_t));
        if (it % 5 == 4 || it < 10) compress();
        it++;
    }
    compress();
    return changed ? -1 : it;
}

// BFS over tensor-order 6-connectivity of a [B,H,W,D] grid (last axis fastest).
// Free voxels get the minimum linear index of their component, occupied voxels -1.
int orc_floodfill_bfs(const uint8_t* grid, int B, int H, int W, int D, int32_t* labels) {
    const int64_t Nvol = (int64_t)H * W * D, Ntot = (int64_t)B * Nvol;
    for (int64_t i = 0; i < Ntot; ++i) labels[i] = grid[i] ? -1 : -2;
    std::vector<int64_t> q;
    for (int64_t s = 0; s < Ntot; ++s) {
        if (labels[s] != -2) continue;
        q.clear();
        q.push_back(s);
        labels[s] = (int32_t)s;
        for (size_t h = 0; h < q.size(); ++h) {
            int64_t idx = q[h];
            int64_t local = idx % Nvol;
            int d = (int)(local % D), w = (int)((local / D) % W), hh = (int)(local / ((int64_t)D * W));
            int64_t nbr[6];
            int k = 0;
            if (d > 0) nbr[k++] = idx - 1;
            if (d < D - 1) nbr[k++] = idx + 1;
            if (w > 0) nbr[k++] = idx - D;
            if (w < W - 1) nbr[k++] = idx + D;
            if (hh > 0) nbr[k++] = idx - (int64_t)D * W;
            if (hh < H - 1) nbr[k++] = idx + (int64_t)D * W;
            for (int j = 0; j < k; ++j)
                if (labels[nbr[j]] == -2) { labels[nbr[j]] = (int32_t)s; q.push_back(nbr[j]); }
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// (7) hash table.  Slot function: include/cpu/hashtable.h:21-42.
static uint32_t hash_step(uint32_t h, uint32_t k) {
    h += k * 0x9E3779B9u;
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}
static int hash_home(const int32_t* key, int d, int capacity) {
    uint32_t h = 0;
    for (int i = 0; i < d; ++i) h = hash_step(h, (uint32_t)key[i]);
    int slot = (int)h % capacity;  // signed remainder, as the reference
    return slot < 0 ? slot + capacity : slot;
}

// Clears table[2 * capacity] to -1 and inserts rows 0..n-1 of coords [n, d] in order
// (include/cpu/hashtable.h:97-124, build: 126-132 with capacity max(16, 2n)).
// Returns the number of rows that found no slot.
int64_t orc_hash_build(const int32_t* coords, int64_t n, int d, int capacity, int32_t* table) {
    std::fill(table, table + 2 * (int64_t)capacity, -1);
    int64_t dropped = 0;
    for (int64_t i = 0; i < n; ++i) {
        int slot = hash_home(coords + i * d, d, capacity);
        const int begin = slot;
        bool placed = false;
        for (int a = 0; a < capacity; ++a) {
            if (table[2 * slot] == -1) {
                table[2 * slot] = slot;
                table[2 * slot + 1] = (int32_t)i;
                placed = true;
                break;
            }
            slot = (slot + 1) % capacity;
            if (slot == begin) break;
        }
        dropped += placed ? 0 : 1;
    }
    return dropped;
}

// out[i] = row index of the first stored key equal to queries[i] along its probe run,
// -1 at the first empty slot or after a full cycle (include/cpu/hashtable.h:58-86).
void orc_hash_search(const int32_t* table, int capacity, const int32_t* coords, int d, const int32_t* queries,
                     int64_t m, int32_t* out) {
    for (int64_t i = 0; i < m; ++i) {
        const int32_t* q = queries + i * d;
        int slot = hash_home(q, d, capacity);
        const int begin = slot;
        int32_t r = -1;
        for (int a = 0; a < capacity; ++a) {
            if (table[2 * slot] == -1) break;
            const int32_t row = table[2 * slot + 1];
            if (row != -1 && std::equal(q, q + d, coords + (int64_t)row * d)) {
                r = row;
                break;
            }
            slot = (slot + 1) % capacity;
            if (slot == begin) break;
        }
        out[i] = r;
    }
}

// ---------------------------------------------------------------------------
// (8) sparse marching cubes.
// Cell corner offsets and edge end corners: include/gpu/spcumc.cuh:322-333.
static const int kMcCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                    {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
static const int kMcEdge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                   {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

// Vertices (at most 12 n) and triangles (at most 5 n) of the extraction; returns 0.
int orc_smc(const int32_t* coords, const float* corners, int64_t n, float iso, int consistent, float* verts,
            int32_t* tris, int64_t* n_verts, int64_t* n_tris) {
    std::vector<float> val(corners, corners + 8 * n);
    if (consistent) {
        std::map<std::array<int64_t, 3>, std::pair<double, int>> acc;
        for (int64_t i = 0; i < n; ++i)
            for (int k = 0; k < 8; ++k) {
                auto& a = acc[{(int64_t)coords[3 * i] + kMcCorner[k][0], (int64_t)coords[3 * i + 1] + kMcCorner[k][1],
                               (int64_t)coords[3 * i + 2] + kMcCorner[k][2]}];
                a.first += (double)corners[8 * i + k];
                a.second += 1;
            }
        for (int64_t i = 0; i < n; ++i)
            for (int k = 0; k < 8; ++k) {
                const auto& a = acc[{(int64_t)coords[3 * i] + kMcCorner[k][0], (int64_t)coords[3 * i + 1] + kMcCorner[k][1],
                                     (int64_t)coords[3 * i + 2] + kMcCorner[k][2]}];
                val[8 * i + k] = (float)(a.first / a.second);
            }
    }
    struct Raw {
        std::array<int64_t, 4> key;
        float p[3];
    };
    std::vector<Raw> raw;
    std::vector<int64_t> first(n + 1, 0);
    std::vector<int> cases(n);
    for (int64_t i = 0; i < n; ++i) {
        const float* v = &val[8 * i];
        int c = 0;
        for (int k = 0; k < 8; ++k)
            if (v[k] < iso) c |= 1 << k;
        cases[i] = c;
        first[i] = (int64_t)raw.size();
        for (int e = 0; e < 12; ++e) {
            const int a = kMcEdge[e][0], b = kMcEdge[e][1];
            if (((c >> a) & 1) == ((c >> b) & 1)) continue;
            Raw r;
            for (int ax = 0; ax < 3; ++ax)
                r.key[ax] = (int64_t)coords[3 * i + ax] + std::min(kMcCorner[a][ax], kMcCorner[b][ax]);
            r.key[3] = kMcCorner[a][0] != kMcCorner[b][0] ? 0 : (kMcCorner[a][1] != kMcCorner[b][1] ? 1 : 2);
            const float va = v[a], vb = v[b], denom = vb - va;
            float t;
            if (std::fabs(denom) < 1e-30f) {
                t = 0.5f;
            } else {
                t = (iso - va) / denom;
                t = std::fmin(std::fmax(t, 0.0f), 1.0f);
            }
            for (int ax = 0; ax < 3; ++ax) {
                const float pa = (float)((int64_t)coords[3 * i + ax] + kMcCorner[a][ax]);
                const float pb = (float)((int64_t)coords[3 * i + ax] + kMcCorner[b][ax]);
                r.p[ax] = pa + t * (pb - pa);
            }
            raw.push_back(r);
        }
    }
    first[n] = (int64_t)raw.size();
    std::vector<int64_t> order(raw.size());
    for (size_t j = 0; j < order.size(); ++j) order[j] = (int64_t)j;
    std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return raw[x].key < raw[y].key; });
    std::vector<int32_t> id(raw.size());
    int64_t m = 0;
    for (size_t j = 0; j < order.size(); ++j) {
        if (j == 0 || raw[order[j - 1]].key != raw[order[j]].key) {
            std::memcpy(verts + 3 * m, raw[order[j]].p, sizeof(float) * 3);
            ++m;
        }
        id[order[j]] = (int32_t)(m - 1);
    }
    int64_t t_out = 0;
    for (int64_t i = 0; i < n; ++i) {
        const uint64_t packed = cubvh_amd::mc::kTriangles[cases[i]];
        // the slot of edge e among the cell's crossing edges, in edge order
        int slot[12], k = 0;
        for (int e = 0; e < 12; ++e) {
            const int a = kMcEdge[e][0], b = kMcEdge[e][1];
            slot[e] = (((cases[i] >> a) ^ (cases[i] >> b)) & 1) ? k++ : -1;
        }
        for (int t = 0; t < cubvh_amd::mc::triangle_count(packed); ++t) {
            const int e0 = cubvh_amd::mc::triangle_edge(packed, 3 * t), e1 = cubvh_amd::mc::triangle_edge(packed, 3 * t + 1),
                      e2 = cubvh_amd::mc::triangle_edge(packed, 3 * t + 2);
            if (slot[e0] < 0 || slot[e1] < 0 || slot[e2] < 0) return -1;  // table edge without a crossing
            tris[3 * t_out] = id[first[i] + slot[e0]];
            tris[3 * t_out + 1] = id[first[i] + slot[e2]];
            tris[3 * t_out + 2] = id[first[i] + slot[e1]];
            ++t_out;
        }
    }
    *n_verts = m;
    *n_tris = t_out;
    return 0;
}

// Backward of sum_i g_i d_i (see (9) above).  gp [n,3] is written; gv [nV,3] is
// added to.  Queries with an out-of-range face or vertex, or with d = 0 or g = 0,
// add nothing to gv (gp is g n, zero for the out-of-range ones).
int orc_udf_backward(const float* V, int64_t nV, const int32_t* F, int64_t nF, const float* p, const float* g,
                     const int64_t* face, const float* uvw, int64_t n, float* gp, float* gv) {
    std::vector<float> gn(3 * (size_t)n, 0.0f);
    std::vector<std::vector<int64_t>> by_face((size_t)nF);
    for (int64_t i = 0; i < n; ++i) {
        const int64_t f = face[i];
        bool ok = f >= 0 && f < nF;
        const int32_t ia = ok ? F[3 * f] : 0, ib = ok ? F[3 * f + 1] : 0, ic = ok ? F[3 * f + 2] : 0;
        ok = ok && ia >= 0 && ia < nV && ib >= 0 && ib < nV && ic >= 0 && ic < nV;
        float s = 0.0f;
        if (ok) {
            const float u = uvw[3 * i], v = uvw[3 * i + 1], w = uvw[3 * i + 2];
            float d[3];
            for (int k = 0; k < 3; ++k) {
                const float c = std::fmaf(w, V[3 * ic + k], std::fmaf(v, V[3 * ib + k], u * V[3 * ia + k]));
                d[k] = p[3 * i + k] - c;
            }
            const float len = std::sqrt(std::fmaf(d[2], d[2], std::fmaf(d[1], d[1], d[0] * d[0])));
            s = len > 0.0f ? g[i] / len : 0.0f;
            for (int k = 0; k < 3; ++k) gn[3 * i + k] = d[k] * s;
        }
        for (int k = 0; k < 3; ++k) gp[3 * i + k] = gn[3 * i + k];
        if (ok && s != 0.0f) by_face[(size_t)f].push_back(i);
    }
    // per face: its queries in chunks of 32 from the first, each chunk summed in order;
    // then runs of 32 consecutive sums, each added in order from zero, replace the sums
    // until one is left (a single chunk sum goes through one such run too)
    const size_t kChunk = 32;
    std::vector<float> acc(9 * (size_t)nF, 0.0f);
    std::vector<float> sums, next;
    for (int64_t f = 0; f < nF; ++f) {
        const std::vector<int64_t>& qs = by_face[(size_t)f];
        if (qs.empty()) continue;
        sums.clear();
        for (size_t c0 = 0; c0 < qs.size(); c0 += kChunk) {
            float part[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (size_t j = c0; j < std::min(qs.size(), c0 + kChunk); ++j) {
                const int64_t q = qs[j];
                for (int k = 0; k < 3; ++k) {
                    const float wk = uvw[3 * q + k];
                    for (int c = 0; c < 3; ++c) part[3 * k + c] += -wk * gn[3 * q + c];
                }
            }
            sums.insert(sums.end(), part, part + 9);
        }
        size_t m = sums.size() / 9;
        do {
            const size_t mo = (m + kChunk - 1) / kChunk;
            next.assign(9 * mo, 0.0f);
            for (size_t u = 0; u < mo; ++u)
                for (size_t j = u * kChunk; j < std::min(m, u * kChunk + kChunk); ++j)
                    for (int k = 0; k < 9; ++k) next[9 * u + k] += sums[9 * j + k];
            sums.swap(next);
            m = mo;
        } while (m > 1);
        for (int k = 0; k < 9; ++k) acc[9 * f + k] = sums[k];
    }
    std::vector<std::vector<int64_t>> by_vertex((size_t)nV);
    for (int64_t s = 0; s < 3 * nF; ++s)
        if (F[s] >= 0 && F[s] < nV) by_vertex[(size_t)F[s]].push_back(s);
    for (int64_t v = 0; v < nV; ++v) {
        float x = 0.0f, y = 0.0f, z = 0.0f;
        for (int64_t s : by_vertex[(size_t)v]) {
            x += acc[3 * s];
            y += acc[3 * s + 1];
            z += acc[3 * s + 2];
        }
        gv[3 * v] += x;
        gv[3 * v + 1] += y;
        gv[3 * v + 2] += z;
    }
    return 0;
}

}  // extern "C"
