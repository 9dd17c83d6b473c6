// query.hip -- BVH traversal kernels for gfx950 (wave64).
//
// One lane per query.  Each lane walks the 4-wide tree nearest-child-first: the
// four child bounds of a node arrive as six 16-byte loads (structure-of-arrays
// inside the 128-byte node), the four box tests run as independent lane-local
// arithmetic, the surviving children are ordered with a 5-comparator network, the
// nearest one is descended into directly and the others are pushed on a per-lane
// stack that lives in LDS (entry = child reference + its box key, laid out
// [depth][lane] so that a wave's pushes and pops hit distinct banks).  Popped
// entries are re-checked against the current bound, so work found stale after a
// better hit is skipped (the reference prunes only at push, src/bvh.cu:333-337).
//
// Parity rules (independent of the tree shape, see build.cpp for the conservative
// boxes):
//   distance: smallest fp32 squared distance, ties to the lowest original face id
//   ray:      smallest hit parameter t < 1000, ties to the lowest original face id
//   any-hit:  exists a face with t < 1000
#include <hip/hip_runtime.h>
#include <stdint.h>




#include "bvh.hpp"
#include "geom.hpp"
#include "query.hpp"

namespace cubvh_amd {
namespace dev {

constexpr float MAX_DIST = 1000.0f;                  // src/bvh.cu:17
constexpr float MAX_DIST_SQ = MAX_DIST * MAX_DIST;   // src/bvh.cu:18
constexpr float NORMAL_EPS = 1e-6f;                  // src/bvh.cu:360
// Children whose key is within this relative factor of the current best are still
// visited, which absorbs fp32 rounding differences between box and triangle tests.
constexpr float BOUND_SLACK = 1.0f + 1.0f / 4096.0f;

struct Tree : TreeArrays {
    int64_t face0_slot;
};

__host__ __device__ inline Tree make_tree(const TreeArrays& a, int64_t face0_slot) {
    Tree t;
    static_cast<TreeArrays&>(t) = a;
    t.face0_slot = face0_slot;
    return t;
}

// True when no face inside a bounding disk (centre c, radius r, unit normal n,
// thickness t; one face's or one leaf's) can come within squared distance `bound` of
// p: every point x inside the disk's cylinder has
// |p - x|^2 >= max(|h| - t, 0)^2 + max(rho - r, 0)^2 with h = n.(p - c) and rho the
// in-plane distance from c.  The margin 2^-15 (|p-c|^2 + r^2) exceeds the fp32
// rounding of this bound and of the exact face test, so a rejected face could not
// have produced a squared distance <= bound (nor tied the best).
__device__ __forceinline__ bool disk_rejects(f3 p, float4 cb, float4 nb, float bound) {
    f3 d = mk3(p.x - cb.x, p.y - cb.y, p.z - cb.z);
    float dd = dot3(d, d);
    float h = dot3(mk3(nb.x, nb.y, nb.z), d);
    float ah = fmaxf(fabsf(h) - nb.w, 0.0f);
    float rho = __builtin_amdgcn_sqrtf(fmaxf(dd - h * h, 0.0f));
    float t = fmaxf(rho - cb.w, 0.0f);
    float lb = ah * ah + t * t;
    float margin = 3.0517578125e-05f * (dd + cb.w * cb.w);
    return lb - margin > bound;
}

__device__ __forceinline__ FaceConst load_face(const float4* __restrict__ tpre, int32_t slot) {
    const float4* r = tpre + 9 * slot;
    return face_const(r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8]);
}

__device__ __forceinline__ void load_tri(const float4* __restrict__ tris, int32_t slot, f3& a, f3& b, f3& c,
                                         int32_t& id) {
    float4 t0 = tris[3 * slot + 0];
    float4 t1 = tris[3 * slot + 1];
    float4 t2 = tris[3 * slot + 2];
    a = xyz(t0);
    b = xyz(t1);
    c = xyz(t2);
    id = __float_as_int(t0.w);
}

__device__ __forceinline__ void leaf_range(int32_t ref, int32_t& first, int32_t& count) {
    first = leaf_ref_first(ref);
    count = leaf_ref_count(ref);
}

// Ascending 4-element sort network on (key, ref).
__device__ __forceinline__ void cswap(float& ka, int32_t& ra, float& kb, int32_t& rb) {
    if (kb < ka) {
        float tk = ka; ka = kb; kb = tk;
        int32_t tr = ra; ra = rb; rb = tr;
    }
}
__device__ __forceinline__ void sort4(float k[4], int32_t r[4]) {
    cswap(k[0], r[0], k[2], r[2]);
    cswap(k[1], r[1], k[3], r[3]);
    cswap(k[0], r[0], k[1], r[1]);
    cswap(k[2], r[2], k[3], r[3]);
    cswap(k[1], r[1], k[2], r[2]);
}

// Per-lane traversal stack.  The first LDS_N entries live in LDS, laid out
// [entry][lane] (entry e of lane t at e * BLOCK + t) so that a wave's accesses at
// equal depth hit distinct banks; deeper entries, which are rare, spill to a
// private array (scratch).  STACK_TOTAL bounds the depth of the whole stack: a
// nearest-first walk of a 4-wide tree of depth D holds at most 3 * D + 1 entries,
// and the library refuses trees that would need more.
constexpr int STACK_TOTAL = 64;

template <int BLOCK, int LDS_N>
struct LaneStack {
    int32_t* ref;
    float* key;
    int tid;
    int sp;
    int32_t spill_ref[STACK_TOTAL - LDS_N];
    float spill_key[STACK_TOTAL - LDS_N];
    __device__ __forceinline__ void push(int32_t r, float k) {
        if (sp < LDS_N) {
            ref[sp * BLOCK + tid] = r;
            key[sp * BLOCK + tid] = k;
        } else if (sp < STACK_TOTAL) {
            spill_ref[sp - LDS_N] = r;
            spill_key[sp - LDS_N] = k;
        } else {
            return;  // unreachable for accepted trees
        }
        ++sp;
    }
    // top entry after a decrement of sp by the caller
    __device__ __forceinline__ float key_at(int i) const {
        return i < LDS_N ? key[i * BLOCK + tid] : spill_key[i - LDS_N];
    }
    __device__ __forceinline__ int32_t ref_at(int i) const {
        return i < LDS_N ? ref[i * BLOCK + tid] : spill_ref[i - LDS_N];
    }
};

__device__ __forceinline__ void box_dist2_4(const WideNode& nd, f3 p, float d[4]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float dx = fmaxf(fmaxf(nd.lo_x[k] - p.x, p.x - nd.hi_x[k]), 0.0f);
        float dy = fmaxf(fmaxf(nd.lo_y[k] - p.y, p.y - nd.hi_y[k]), 0.0f);
        float dz = fmaxf(fmaxf(nd.lo_z[k] - p.z, p.z - nd.hi_z[k]), 0.0f);
        d[k] = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
    }
}

// A point with a NaN coordinate: every squared distance to a face is NaN, so no face
// can win (the reference's `<=` tests are false for it too) and the walks skip it.  It
// must not walk: fmaxf drops NaN operands, so its box distances are finite, and with
// three NaN coordinates zero even for an empty child slot (whose box is +inf .. -inf).
__device__ __forceinline__ bool has_nan(f3 p) { return p.x != p.x || p.y != p.y || p.z != p.z; }

__device__ __forceinline__ WideNode load_node(const WideNode* __restrict__ nodes, int32_t i) {
    const float4* src = reinterpret_cast<const float4*>(nodes + i);
    WideNode nd;
    float4* dst = reinterpret_cast<float4*>(&nd);
#pragma unroll
    for (int k = 0; k < 8; ++k) dst[k] = src[k];
    return nd;
}

// Nearest face to p.  Returns the leaf slot (-1 if nothing within MAX_DIST) and its
// squared distance in best_d2.
template <int BLOCK, int LDS_N>
__device__ int32_t closest_face(const Tree& T, f3 p, LaneStack<BLOCK, LDS_N>& st, float& best_d2, int32_t& best_id) {
    float best = MAX_DIST_SQ;
    float bound = MAX_DIST_SQ * BOUND_SLACK;
    int32_t bslot = -1, bid = 0x7fffffff;
    int32_t ref = has_nan(p) ? CHILD_EMPTY : 0;
    st.sp = 0;
    while (ref != CHILD_EMPTY) {
        if (ref >= 0) {
            WideNode nd = load_node(T.nodes, ref);
            float k[4];
            int32_t r[4] = {nd.child[0], nd.child[1], nd.child[2], nd.child[3]};
            box_dist2_4(nd, p, k);
            for (int j = 0; j < 4; ++j)  // empty slots, and internal subtrees whose disk is out of reach
                if (r[j] == CHILD_EMPTY ||
                    (r[j] >= 0 && k[j] <= bound &&
                     disk_rejects(p, T.cbnd[8 * ref + 2 * j], T.cbnd[8 * ref + 2 * j + 1], bound)))
                    k[j] = INFINITY;
            sort4(k, r);
            if (k[0] <= bound) {
                if (k[3] <= bound) st.push(r[3], k[3]);
                if (k[2] <= bound) st.push(r[2], k[2]);
                if (k[1] <= bound) st.push(r[1], k[1]);
                ref = r[0];
                continue;
            }
        } else {
            int32_t first, count;
            leaf_range(ref, first, count);
            if (disk_rejects(p, T.lbnd[2 * first], T.lbnd[2 * first + 1], bound)) count = 0;  // whole leaf out of reach
            for (int j = 0; j < count; ++j) {
                if (disk_rejects(p, T.tbnd[2 * (first + j)], T.tbnd[2 * (first + j) + 1], bound)) continue;
                const FaceConst fc = load_face(T.tpre, first + j);
                const float d2 = tri_dist2_pre(fc, p);
                const int32_t id = fc.id;
                if (d2 < best || (d2 == best && bslot >= 0 && id < bid)) {
                    best = d2;
                    bslot = first + j;
                    bid = id;
                    bound = best * BOUND_SLACK;
                }
            }
        }
        // pop the next entry that can still hold a closer face
        ref = CHILD_EMPTY;
        while (st.sp > 0) {
            --st.sp;
            if (st.key_at(st.sp) <= bound) {
                ref = st.ref_at(st.sp);
                break;
            }
        }
    }
    best_d2 = best;
    best_id = bid;
    return bslot;
}

// Entry/exit parameters of the ray against the 4 child boxes; key = entry when the
// box is hit within [0, tmax_limit], +inf otherwise.
__device__ __forceinline__ void box_ray_4(const WideNode& nd, f3 o, f3 inv, float tlimit, float k[4]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        float tx0 = (nd.lo_x[j] - o.x) * inv.x, tx1 = (nd.hi_x[j] - o.x) * inv.x;
        float ty0 = (nd.lo_y[j] - o.y) * inv.y, ty1 = (nd.hi_y[j] - o.y) * inv.y;
        float tz0 = (nd.lo_z[j] - o.z) * inv.z, tz1 = (nd.hi_z[j] - o.z) * inv.z;
        float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
        float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
        bool hit = nd.child[j] != CHILD_EMPTY && tn <= tf && tf >= 0.0f && tn <= tlimit;
        k[j] = hit ? tn : INFINITY;
    }
}

__device__ __forceinline__ f3 ray_inverse(f3 d) {
    // Exact reciprocal; zero components are replaced by a tiny value of the same
    // sign so that no slab product becomes 0 * inf.
    float x = fabsf(d.x) < 1e-30f ? copysignf(1e-30f, d.x) : d.x;
    float y = fabsf(d.y) < 1e-30f ? copysignf(1e-30f, d.y) : d.y;
    float z = fabsf(d.z) < 1e-30f ? copysignf(1e-30f, d.z) : d.z;
    return mk3(1.0f / x, 1.0f / y, 1.0f / z);
}

// Closest hit.  Returns the leaf slot (-1 on a miss) and t in best_t.
template <int BLOCK, int LDS_N>
__device__ int32_t closest_hit(const Tree& T, f3 o, f3 d, LaneStack<BLOCK, LDS_N>& st, float& best_t) {
    f3 inv = ray_inverse(d);
    float best = MAX_DIST;
    float bound = MAX_DIST * BOUND_SLACK;
    int32_t bslot = -1, bid = 0x7fffffff;
    int32_t ref = 0;
    st.sp = 0;
    for (;;) {
        if (ref >= 0) {
            WideNode nd = load_node(T.nodes, ref);
            float k[4];
            int32_t r[4] = {nd.child[0], nd.child[1], nd.child[2], nd.child[3]};
            box_ray_4(nd, o, inv, bound, k);
            sort4(k, r);
            if (k[0] <= bound) {
                if (k[3] <= bound) st.push(r[3], k[3]);
                if (k[2] <= bound) st.push(r[2], k[2]);
                if (k[1] <= bound) st.push(r[1], k[1]);
                ref = r[0];
                continue;
            }
        } else {
            int32_t first, count;
            leaf_range(ref, first, count);
            for (int j = 0; j < count; ++j) {
                f3 a, b, c;
                int32_t id;
                load_tri(T.tris, first + j, a, b, c, id);
                float t = tri_ray(a, b, c, o, d);
                if (t < best || (t == best && bslot >= 0 && id < bid)) {
                    best = t;
                    bslot = first + j;
                    bid = id;
                    bound = best * BOUND_SLACK + 1e-30f;
                }
            }
        }
        ref = CHILD_EMPTY;
        while (st.sp > 0) {
            --st.sp;
            if (st.key_at(st.sp) <= bound) {
                ref = st.ref_at(st.sp);
                break;
            }
        }
        if (ref == CHILD_EMPTY) break;
    }
    best_t = best;
    return bslot;
}

// Whether both rays o + t d and o - t d (t in [0, MAX_DIST)) hit a face: returns 3 when
// both do (the reference's stab pair, src/bvh.cu:414-424: a face hit with t < MAX_DIST
// along each), otherwise a value other than 3 -- the walk stops as soon as one
// direction is answered "no hit", since the pair is then decided.  One slab test gives
// the line's interval [tn, tf] in a box: the +d ray meets it when tf >= 0 and
// tn <= limit, the -d ray when tn <= 0 and tf >= -limit (the two rays' own slab tests,
// negated exactly).  Children are walked nearest-first by the entry distance of the
// directions still open; a stack entry carries those directions in the two low bits of
// its key and is dropped at pop when both are answered.  open1 / open2 count the stack
// entries still open for +d / -d: a direction not yet hit with no open entry and not
// open in the subtree at hand can no longer be hit.
// The walk as a resumable state: init, then step until it returns a result (>= 0), so
// that a lane can start the next line as soon as its own ends (persistent stab kernels).
struct LineWalk {
    f3 o, d, inv;
    int found, dirs, open1, open2;
    int32_t ref;

    template <int BLOCK, int LDS_N>
    __device__ __forceinline__ void init(f3 o_, f3 d_, LaneStack<BLOCK, LDS_N>& st) {
        o = o_;
        d = d_;
        inv = ray_inverse(d_);
        found = 0;
        ref = 0;
        dirs = 3;
        open1 = open2 = 0;
        st.sp = 0;
    }

    // One node or leaf visit (and the pop after it): -1 while the pair is undecided,
    // else the walk's result as any_hit_line returns it.
    template <int BLOCK, int LDS_N>
    __device__ __forceinline__ int step(const Tree& T, LaneStack<BLOCK, LDS_N>& st) {
        const float limit = MAX_DIST * BOUND_SLACK;
        if ((dirs & ~found) != 0) {
            if (ref >= 0) {
                WideNode nd = load_node(T.nodes, ref);
                float k[4];
                int32_t r[4] = {nd.child[0], nd.child[1], nd.child[2], nd.child[3]};
                const int open = 3 & ~found;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    float tx0 = (nd.lo_x[j] - o.x) * inv.x, tx1 = (nd.hi_x[j] - o.x) * inv.x;
                    float ty0 = (nd.lo_y[j] - o.y) * inv.y, ty1 = (nd.hi_y[j] - o.y) * inv.y;
                    float tz0 = (nd.lo_z[j] - o.z) * inv.z, tz1 = (nd.hi_z[j] - o.z) * inv.z;
                    float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
                    float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
                    const bool box = nd.child[j] != CHILD_EMPTY && tn <= tf;
                    const int m = box ? (((tf >= 0.0f && tn <= limit) ? 1 : 0) | ((tn <= 0.0f && tf >= -limit) ? 2 : 0)) & open
                                      : 0;
                    float key = INFINITY;
                    if (m & 1) key = fmaxf(tn, 0.0f);
                    if (m & 2) key = fminf(key, fmaxf(-tf, 0.0f));
                    k[j] = m ? __uint_as_float((__float_as_uint(key) & ~3u) | (uint32_t)m) : INFINITY;
                }
                sort4(k, r);
                if (k[0] != INFINITY) {
#pragma unroll
                    for (int j = 3; j >= 1; --j)
                        if (k[j] != INFINITY) {
                            st.push(r[j], k[j]);
                            const uint32_t m = __float_as_uint(k[j]) & 3u;
                            open1 += (int)(m & 1u);
                            open2 += (int)(m >> 1);
                        }
                    ref = r[0];
                    dirs = (int)(__float_as_uint(k[0]) & 3u);
                    const int pending = dirs | (open1 ? 1 : 0) | (open2 ? 2 : 0);
                    if ((3 & ~found & ~pending) != 0) return found;  // a direction without a hit left
                    return -1;
                }
            } else {
                int32_t first, count;
                leaf_range(ref, first, count);
                for (int j = 0; j < count && (dirs & ~found) != 0; ++j) {
                    f3 a, b, c;
                    int32_t id;
                    load_tri(T.tris, first + j, a, b, c, id);
                    found |= tri_line(a, b, c, o, d, MAX_DIST) & dirs;
                }
                if (found == 3) return 3;
            }
        }
        if ((3 & ~found & ~((open1 ? 1 : 0) | (open2 ? 2 : 0))) != 0) return found;
        if (st.sp == 0) return found;
        --st.sp;
        ref = st.ref_at(st.sp);
        dirs = (int)(__float_as_uint(st.key_at(st.sp)) & 3u);
        open1 -= dirs & 1;
        open2 -= dirs >> 1;
        return -1;
    }
};

template <int BLOCK, int LDS_N>
__device__ int any_hit_line(const Tree& T, f3 o, f3 d, LaneStack<BLOCK, LDS_N>& st) {
    LineWalk w;
    w.init(o, d, st);
    for (;;) {
        const int r = w.step(T, st);
        if (r >= 0) return r;
    }
}

// Mean unit normal of every face within squared distance NORMAL_EPS of point c
// (src/bvh.cu:356-396).  The sum is accumulated in double so that it does not
// depend on the visiting order, then rounded to float and divided by the count.
template <int BLOCK, int LDS_N>
__device__ f3 avg_normal(const Tree& T, f3 c, LaneStack<BLOCK, LDS_N>& st) {
    double sx = 0.0, sy = 0.0, sz = 0.0;
    float cnt = 0.0f;
    int32_t ref = 0;
    st.sp = 0;
    for (;;) {
        if (has_nan(c)) break;  // no face is within NORMAL_EPS of it: the mean is 0 / 0
        if (ref >= 0) {
            WideNode nd = load_node(T.nodes, ref);
            float k[4];
            box_dist2_4(nd, c, k);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (nd.child[j] != CHILD_EMPTY && k[j] < NORMAL_EPS) st.push(nd.child[j], k[j]);
        } else {
            int32_t first, count;
            leaf_range(ref, first, count);
            // disks that put a leaf or face beyond NORMAL_EPS rule out d2 < NORMAL_EPS
            if (disk_rejects(c, T.lbnd[2 * first], T.lbnd[2 * first + 1], NORMAL_EPS)) count = 0;
            for (int j = 0; j < count; ++j) {
                if (disk_rejects(c, T.tbnd[2 * (first + j)], T.tbnd[2 * (first + j) + 1], NORMAL_EPS)) continue;
                f3 a, b, cc;
                int32_t id;
                load_tri(T.tris, first + j, a, b, cc, id);
                if (tri_dist2(a, b, cc, c) < NORMAL_EPS) {
                    f3 n = tri_normal(a, b, cc);
                    sx += (double)n.x;
                    sy += (double)n.y;
                    sz += (double)n.z;
                    cnt += 1.0f;
                }
            }
        }
        if (st.sp == 0) break;
        --st.sp;
        ref = st.ref_at(st.sp);
    }
    return mk3((float)sx / cnt, (float)sy / cnt, (float)sz / cnt);
}

// ---------------------------------------------------------------------------
// Packet traversal: one wave walks the tree for its 64 queries together.
//
// The current node and the stack are wave-uniform; each stack entry carries the
// 64-bit mask of lanes that still want that subtree.  Node and face records are
// read through the constant address space with a wave-uniform address, which the
// compiler turns into scalar loads (one fetch per wave, results in SGPRs), and
// every lane tests them against its own query.  The stack lives in three VGPRs
// spread across the wave's lanes (entry e in lane e, written by a lane-id select and
// read with v_readlane), so it needs no LDS; a tree of depth D needs at most
// 3 * D + 1 <= 64 entries.  With queries in Morton order the lanes of a wave are
// close together and the union of their walks is close to a single walk.
// ---------------------------------------------------------------------------
#if defined(__HIP_DEVICE_COMPILE__)
typedef const __attribute__((address_space(4))) float* cfloat_ptr;  // constant address space
#else
typedef const float* cfloat_ptr;  // host pass: never executed
#endif

__device__ __forceinline__ float4 ld4(cfloat_ptr p) { return make_float4(p[0], p[1], p[2], p[3]); }

struct UNode {  // wave-uniform copy of a node (scalar registers)
    float4 lx, hx, ly, hy, lz, hz;
    int4 child;
};

__device__ __forceinline__ UNode load_unode(const WideNode* nodes, int32_t ref) {
    cfloat_ptr src = (cfloat_ptr)(nodes + ref);
    UNode u;
    u.lx = ld4(src + 0);
    u.hx = ld4(src + 4);
    u.ly = ld4(src + 8);
    u.hy = ld4(src + 12);
    u.lz = ld4(src + 16);
    u.hz = ld4(src + 20);
    float4 c = ld4(src + 24);
    u.child = make_int4(__float_as_int(c.x), __float_as_int(c.y), __float_as_int(c.z), __float_as_int(c.w));
    return u;
}

__device__ __forceinline__ void load_utri(const float4* tris, int32_t slot, f3& a, f3& b, f3& c, int32_t& id) {
    cfloat_ptr t = (cfloat_ptr)(tris + 3 * slot);
    float4 t0 = ld4(t), t1 = ld4(t + 4), t2 = ld4(t + 8);
    a = xyz(t0);
    b = xyz(t1);
    c = xyz(t2);
    id = __float_as_int(t0.w);
}

__device__ __forceinline__ FaceConst load_uface(const float4* tpre, int32_t slot) {
    cfloat_ptr r = (cfloat_ptr)(tpre + 9 * slot);
    return face_const(ld4(r), ld4(r + 4), ld4(r + 8), ld4(r + 12), ld4(r + 16), ld4(r + 20), ld4(r + 24),
                      ld4(r + 28), ld4(r + 32));
}

__device__ __forceinline__ float box_d2(float lx, float hx, float ly, float hy, float lz, float hz, f3 p) {
    float dx = fmaxf(fmaxf(lx - p.x, p.x - hx), 0.0f);
    float dy = fmaxf(fmaxf(ly - p.y, p.y - hy), 0.0f);
    float dz = fmaxf(fmaxf(lz - p.z, p.z - hz), 0.0f);
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

__device__ __forceinline__ int32_t uniform(int32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint64_t ballot(bool b) { return __builtin_amdgcn_ballot_w64(b); }
__device__ __forceinline__ float lane_value(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

struct WaveStack {
    int32_t ref = 0, mlo = 0, mhi = 0;  // entry e held by lane e
    int32_t sp = 0;                     // wave-uniform
    __device__ __forceinline__ void push(int32_t r, uint64_t m) {
        const bool here = (int32_t)(threadIdx.x & 63) == sp;
        ref = here ? r : ref;
        mlo = here ? (int32_t)(uint32_t)m : mlo;
        mhi = here ? (int32_t)(uint32_t)(m >> 32) : mhi;
        ++sp;
    }
    __device__ __forceinline__ bool pop(int32_t& r, uint64_t& m) {
        if (sp == 0) return false;
        --sp;
        r = __builtin_amdgcn_readlane(ref, sp);
        m = (uint64_t)(uint32_t)__builtin_amdgcn_readlane(mlo, sp) |
            ((uint64_t)(uint32_t)__builtin_amdgcn_readlane(mhi, sp) << 32);
        return true;
    }
};

// Uniform sort of 4 (key, ref, mask) triples, ascending by key.  Keys are
// non-negative floats (or +inf) passed as their bit patterns, which order the same
// way as unsigned integers, so the network runs on the scalar unit.
__device__ __forceinline__ void ucswap(uint32_t& ka, int32_t& ra, uint64_t& ma, uint32_t& kb, int32_t& rb,
                                       uint64_t& mb) {
    if (kb < ka) {
        uint32_t tk = ka; ka = kb; kb = tk;
        int32_t tr = ra; ra = rb; rb = tr;
        uint64_t tm = ma; ma = mb; mb = tm;
    }
}

// Key of the lowest lane in m (+inf bits when m is empty).
__device__ __forceinline__ uint32_t rep_key(float k, uint64_t m) {
    const uint32_t v = (uint32_t)__builtin_amdgcn_readlane(__float_as_int(k), (int)__builtin_ctzll(m | 1ull));
    return m ? v : 0x7f800000u;
}

// Exact face tests of a wave, queued in LDS as (face slot, query lane) pairs and run
// 64 at a time with one pair per lane, so that a test keeps every lane busy however
// few lanes wanted the face.  QCAP holds a full round plus one face's worth of pairs.
constexpr int QCAP = 128;
struct ExactQueue {
    uint32_t slot[QCAP];
    uint32_t lane[QCAP];
    unsigned long long key[64];  // per query lane: (squared distance bits, face id), min wins
    int32_t kslot[64];           // per query lane: slot of the face holding key
    uint32_t lanes[64];          // the lanes that reach the current leaf, in lane order
};

__device__ __forceinline__ void wave_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// Lanes below `lane` in m.
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Runs the newest min(qn, 64) queued pairs and folds their results into each query
// lane's best.  Squared distances are >= 0 (or NaN / +inf, which never win), so their
// bit patterns order like the values, and a key (d2 bits, face id) orders by distance
// and then by the lower face id: the walk's tie rule.  An unset best (bslot < 0) has
// the key (MAX_DIST_SQ, 0), so a face must come strictly closer than MAX_DIST_SQ.
__device__ __forceinline__ void run_exact(const Tree& T, ExactQueue& q, int32_t& qn, f3 p, float& best,
                                          float& bound, int32_t& bslot, int32_t& bid) {
    const int lane = (int)(threadIdx.x & 63);
    const int32_t take = qn < 64 ? qn : 64;
    const int32_t base = qn - take;
    const unsigned long long mine =
        ((unsigned long long)__float_as_uint(best) << 32) | (uint32_t)(bslot >= 0 ? bid : 0);
    q.key[lane] = mine;
    wave_fence();
    const bool has = lane < take;
    uint32_t slot = 0, src = 0;
    if (has) {
        slot = q.slot[base + lane];
        src = q.lane[base + lane];
    }
    const f3 ps = mk3(__shfl(p.x, (int)src), __shfl(p.y, (int)src), __shfl(p.z, (int)src));
    unsigned long long k = 0;
    if (has) {
        const FaceConst fc = load_face(T.tpre, (int32_t)slot);
        k = ((unsigned long long)__float_as_uint(tri_dist2_pre(fc, ps)) << 32) | (uint32_t)fc.id;
        atomicMin(&q.key[src], k);
    }
    wave_fence();
    if (has && q.key[src] == k) q.kslot[src] = (int32_t)slot;
    wave_fence();
    const unsigned long long now = q.key[lane];
    if (now != mine) {
        best = __uint_as_float((uint32_t)(now >> 32));
        bid = (int32_t)(uint32_t)now;
        bslot = q.kslot[lane];
        bound = best * BOUND_SLACK;
    }
    qn = base;
    wave_fence();
}

// A leaf wanted by more lanes than this is tested one face at a time (a round of 64
// pairs per face would cost more than the step); fewer lanes are packed into rounds.
constexpr int kStepLanes = 32;

// Nearest face for every lane of the wave (lanes with active = false only ride
// along).  Results in the lane's best / bslot / bid, as closest_face.  q is this
// wave's own LDS queue.
__device__ void closest_face_packet(const Tree& T, int32_t root, f3 p, bool active, float& best, int32_t& bslot,
                                    int32_t& bid, ExactQueue& q) {
    const int lane = (int)(threadIdx.x & 63);
    best = MAX_DIST_SQ;
    float bound = MAX_DIST_SQ * BOUND_SLACK;
    bslot = -1;
    bid = 0x7fffffff;
    uint64_t mask = ballot(active && !has_nan(p));
    if (mask == 0) return;
    WaveStack st;
    int32_t qn = 0;  // queued pairs (wave-uniform)
    int32_t ref = uniform(root);
    for (;;) {
        if (ref >= 0) {
            UNode nd = load_unode(T.nodes, ref);
            // every lane computes; the uniform lane mask selects the interested ones
            uint64_t m0 = ballot(box_d2(nd.lx.x, nd.hx.x, nd.ly.x, nd.hy.x, nd.lz.x, nd.hz.x, p) <= bound) & mask;
            uint64_t m1 = ballot(box_d2(nd.lx.y, nd.hx.y, nd.ly.y, nd.hy.y, nd.lz.y, nd.hz.y, p) <= bound) & mask;
            uint64_t m2 = ballot(box_d2(nd.lx.z, nd.hx.z, nd.ly.z, nd.hy.z, nd.lz.z, nd.hz.z, p) <= bound) & mask;
            uint64_t m3 = ballot(box_d2(nd.lx.w, nd.hx.w, nd.ly.w, nd.hy.w, nd.lz.w, nd.hz.w, p) <= bound) & mask;
            // internal subtrees: drop the lanes that the subtree's disk puts out of reach
            cfloat_ptr cp = (cfloat_ptr)(T.cbnd + 8 * ref);
            if (m0 != 0 && nd.child.x >= 0) m0 = ballot(!disk_rejects(p, ld4(cp), ld4(cp + 4), bound)) & m0;
            if (m1 != 0 && nd.child.y >= 0) m1 = ballot(!disk_rejects(p, ld4(cp + 8), ld4(cp + 12), bound)) & m1;
            if (m2 != 0 && nd.child.z >= 0) m2 = ballot(!disk_rejects(p, ld4(cp + 16), ld4(cp + 20), bound)) & m2;
            if (m3 != 0 && nd.child.w >= 0) m3 = ballot(!disk_rejects(p, ld4(cp + 24), ld4(cp + 28), bound)) & m3;
            if ((m0 | m1 | m2 | m3) != 0) {
                // order the children by the key of their lowest interested lane
                uint32_t r0 = rep_key(box_d2(nd.lx.x, nd.hx.x, nd.ly.x, nd.hy.x, nd.lz.x, nd.hz.x, p), m0);
                uint32_t r1 = rep_key(box_d2(nd.lx.y, nd.hx.y, nd.ly.y, nd.hy.y, nd.lz.y, nd.hz.y, p), m1);
                uint32_t r2 = rep_key(box_d2(nd.lx.z, nd.hx.z, nd.ly.z, nd.hy.z, nd.lz.z, nd.hz.z, p), m2);
                uint32_t r3 = rep_key(box_d2(nd.lx.w, nd.hx.w, nd.ly.w, nd.hy.w, nd.lz.w, nd.hz.w, p), m3);
                int32_t c0 = nd.child.x, c1 = nd.child.y, c2 = nd.child.z, c3 = nd.child.w;
                ucswap(r0, c0, m0, r2, c2, m2);
                ucswap(r1, c1, m1, r3, c3, m3);
                ucswap(r0, c0, m0, r1, c1, m1);
                ucswap(r2, c2, m2, r3, c3, m3);
                ucswap(r1, c1, m1, r2, c2, m2);
                if (m3) st.push(c3, m3);
                if (m2) st.push(c2, m2);
                if (m1) st.push(c1, m1);
                ref = uniform(c0);
                mask = m0;
                continue;
            }
        } else {
            int32_t first, count;
            leaf_range(ref, first, count);
            // lanes whose point may still reach a face of this leaf
            cfloat_ptr lp = (cfloat_ptr)(T.lbnd + 2 * first);
            const uint64_t lm = ballot(!disk_rejects(p, ld4(lp), ld4(lp + 4), bound)) & mask;
            if (__builtin_popcountll(lm) > kStepLanes) {
                // most lanes want the leaf: one step per face, face records uniform
                for (int j = 0; j < count; ++j) {
                    cfloat_ptr bp = (cfloat_ptr)(T.tbnd + 2 * (first + j));
                    const uint64_t cm = ballot(!disk_rejects(p, ld4(bp), ld4(bp + 4), bound)) & lm;
                    if (cm == 0) continue;  // no lane can reach this face
                    if ((cm >> lane) & 1) {
                        const uint32_t at = (uint32_t)qn + lanes_below(cm);
                        q.slot[at] = (uint32_t)(first + j);
                        q.lane[at] = (uint32_t)lane;
                    }
                    qn += __builtin_popcountll(cm);
                    if (qn >= 64) {
                        wave_fence();
                        run_exact(T, q, qn, p, best, bound, bslot, bid);
                    }
                }
            } else if (lm != 0) {
                // The (lane, face) pairs of the leaf, 64 per round: pair i is face
                // i / nl of the (i % nl)-th lane of lm.  Pairs whose face disk does
                // not reject join the exact-test queue.
                const int32_t nl = __builtin_popcountll(lm);
                if ((lm >> lane) & 1) q.lanes[lanes_below(lm)] = (uint32_t)lane;
                wave_fence();
                const float inv_nl = 1.0f / (float)nl;
                const int32_t total = nl * count;
                for (int32_t base = 0; base < total; base += 64) {
                    const int32_t i = base + lane;
                    const bool has = i < total;
                    // (i + 1/2) / nl is at least 1/(2 nl) away from an integer: exact floor
                    const int32_t j = has ? (int32_t)(((float)i + 0.5f) * inv_nl) : 0;
                    const uint32_t src = has ? q.lanes[i - j * nl] : 0u;
                    const f3 ps = mk3(__shfl(p.x, (int)src), __shfl(p.y, (int)src), __shfl(p.z, (int)src));
                    const float bs = __shfl(bound, (int)src);
                    bool pass = false;
                    if (has) {
                        const float4* bp = T.tbnd + 2 * (first + j);
                        pass = !disk_rejects(ps, bp[0], bp[1], bs);
                    }
                    const uint64_t pm = ballot(pass);
                    if (pass) {
                        const uint32_t at = (uint32_t)qn + lanes_below(pm);
                        q.slot[at] = (uint32_t)(first + j);
                        q.lane[at] = src;
                    }
                    qn += __builtin_popcountll(pm);
                    if (qn >= 64) {
                        wave_fence();
                        run_exact(T, q, qn, p, best, bound, bslot, bid);
                    }
                }
                wave_fence();
            }
        }
        if (!st.pop(ref, mask)) break;
        ref = uniform(ref);
    }
    if (qn > 0) {
        wave_fence();
        run_exact(T, q, qn, p, best, bound, bslot, bid);
    }
}

// Packet form of closest_hit: every lane traces its own ray; node and face records
// are wave-uniform.  Children are ordered by the entry distance of their lowest
// interested lane.  Results (t, slot, id) as closest_hit.
__device__ __forceinline__ float slab_entry(float lx, float hx, float ly, float hy, float lz, float hz, f3 o, f3 inv) {
    const float tx0 = (lx - o.x) * inv.x, tx1 = (hx - o.x) * inv.x;
    const float ty0 = (ly - o.y) * inv.y, ty1 = (hy - o.y) * inv.y;
    const float tz0 = (lz - o.z) * inv.z, tz1 = (hz - o.z) * inv.z;
    const float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
    const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
    return (tn <= tf && tf >= 0.0f) ? tn : INFINITY;
}

__device__ void closest_hit_packet(const Tree& T, f3 o, f3 d, bool active, float& best, int32_t& bslot,
                                   int32_t& bid) {
    const int lane = (int)(threadIdx.x & 63);
    const f3 inv = ray_inverse(d);
    best = MAX_DIST;
    float bound = MAX_DIST * BOUND_SLACK;
    bslot = -1;
    bid = 0x7fffffff;
    uint64_t mask = ballot(active);
    if (mask == 0) return;
    WaveStack st;
    int32_t ref = 0;
    for (;;) {
        if (ref >= 0) {
            UNode nd = load_unode(T.nodes, ref);
            float k0 = nd.child.x != CHILD_EMPTY ? slab_entry(nd.lx.x, nd.hx.x, nd.ly.x, nd.hy.x, nd.lz.x, nd.hz.x, o, inv) : INFINITY;
            float k1 = nd.child.y != CHILD_EMPTY ? slab_entry(nd.lx.y, nd.hx.y, nd.ly.y, nd.hy.y, nd.lz.y, nd.hz.y, o, inv) : INFINITY;
            float k2 = nd.child.z != CHILD_EMPTY ? slab_entry(nd.lx.z, nd.hx.z, nd.ly.z, nd.hy.z, nd.lz.z, nd.hz.z, o, inv) : INFINITY;
            float k3 = nd.child.w != CHILD_EMPTY ? slab_entry(nd.lx.w, nd.hx.w, nd.ly.w, nd.hy.w, nd.lz.w, nd.hz.w, o, inv) : INFINITY;
            uint64_t m0 = ballot(k0 <= bound) & mask;
            uint64_t m1 = ballot(k1 <= bound) & mask;
            uint64_t m2 = ballot(k2 <= bound) & mask;
            uint64_t m3 = ballot(k3 <= bound) & mask;
            if ((m0 | m1 | m2 | m3) != 0) {
                // order by entry distance clamped at 0 (a box around the origin enters at
                // a negative t); non-negative floats compare as their bit patterns
                uint32_t r0 = rep_key(fmaxf(k0, 0.0f), m0), r1 = rep_key(fmaxf(k1, 0.0f), m1);
                uint32_t r2 = rep_key(fmaxf(k2, 0.0f), m2), r3 = rep_key(fmaxf(k3, 0.0f), m3);
                int32_t c0 = nd.child.x, c1 = nd.child.y, c2 = nd.child.z, c3 = nd.child.w;
                ucswap(r0, c0, m0, r2, c2, m2);
                ucswap(r1, c1, m1, r3, c3, m3);
                ucswap(r0, c0, m0, r1, c1, m1);
                ucswap(r2, c2, m2, r3, c3, m3);
                ucswap(r1, c1, m1, r2, c2, m2);
                if (m3) st.push(c3, m3);
                if (m2) st.push(c2, m2);
                if (m1) st.push(c1, m1);
                ref = uniform(c0);
                mask = m0;
                continue;
            }
        } else {
            int32_t first, count;
            leaf_range(ref, first, count);
            for (int j = 0; j < count; ++j) {
                f3 a, b, c;
                int32_t id;
                load_utri(T.tris, first + j, a, b, c, id);
                if ((mask >> lane) & 1) {
                    const float t = tri_ray(a, b, c, o, d);
                    if (t < best || (t == best && bslot >= 0 && id < bid)) {
                        best = t;
                        bslot = first + j;
                        bid = id;
                        bound = best * BOUND_SLACK + 1e-30f;
                    }
                }
            }
        }
        if (!st.pop(ref, mask)) break;
        ref = uniform(ref);
    }
}

__device__ __forceinline__ f3 load_pt(const float* __restrict__ p, int64_t i) {
    return mk3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
}
__device__ __forceinline__ void store_pt(float* __restrict__ p, int64_t i, f3 v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}

// mode: 0 unsigned, 1 signed watertight, 2 signed raystab
template <int BLOCK, int LDS_N, int MODE>
__global__ __launch_bounds__(BLOCK) void distance_kernel(int64_t n, const float* __restrict__ positions,
                                                         const uint32_t* __restrict__ order,
                                                         float* __restrict__ distances, int64_t* __restrict__ face_id,
                                                         float* __restrict__ uvw, Tree T) {
    __shared__ int32_t s_ref[LDS_N * BLOCK];
    __shared__ float s_key[LDS_N * BLOCK];
    const int64_t gi = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (gi >= n) return;
    const int64_t i = order ? (int64_t)order[gi] : gi;  // query index
    LaneStack<BLOCK, LDS_N> st;
    st.ref = s_ref;
    st.key = s_key;
    st.tid = (int)threadIdx.x;
    st.sp = 0;
    const f3 p = load_pt(positions, i);
    float d2;
    int32_t bid;
    int32_t slot = closest_face(T, p, st, d2, bid);
    if (slot < 0) {  // nothing within MAX_DIST: distance 0 and face 0
        slot = (int32_t)T.face0_slot;
        d2 = 0.0f;
    }
    f3 a, b, c;
    int32_t id;
    load_tri(T.tris, slot, a, b, c, id);
    float dist = sqrtf(d2);
    if (MODE == 1) {
        f3 cp = tri_closest_pt(a, b, c, p);
        f3 nrm = avg_normal(T, cp, st);
        dist = copysignf(dist, dot3(nrm, sub3(p, cp)));
    } else if (MODE == 2) {
        Pcg32 rng;
        rng.advance((uint64_t)(uint32_t)((uint32_t)i * 2u));
        float ox = rng.next_float();
        float oy = rng.next_float();
        bool inside = true;
        for (uint32_t k = 0; k < 32; ++k) {
            f3 dir = fibonacci_dir32(k, ox, oy);
            if (any_hit_line(T, p, dir, st) != 3) {
                inside = false;
                break;
            }
        }
        if (inside) dist = -dist;
    }
    distances[i] = dist;
    face_id[i] = (int64_t)id;
    if (uvw) {
        f3 cp = tri_closest_pt(a, b, c, p);
        store_pt(uvw, i, tri_barycentric(a, b, c, cp));
    }
}

// Workgroup b of a launch runs on XCD b % 8: the dispatcher deals workgroups round-robin
// over the 8 XCDs, each with its own L2.  xcd_remap gives workgroup b the logical block
// it works on so that each XCD takes runs of `chunk` consecutive logical blocks (within
// every group of 8 chunk workgroups; a last partial group keeps its order): neighbouring
// tiles, which walk the same part of a tree, then share one L2 instead of eight.  A
// bijection of [0, nb), so every tile is still walked once.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb, uint32_t chunk) {
    if (chunk == 0) return b;
    const uint32_t group = 8u * chunk;
    if (b >= nb / group * group) return b;
    const uint32_t r = b % group;
    return b - r + (r % 8u) * chunk + r / 8u;
}

// Raystab sign of distances already computed (src/bvh.cu:408-427): a query is inside,
// and its distance negated, when all 64 stab rays from its point hit a face (the 32
// stab lines, each hit both ways).  The stab directions are keyed by the query index,
// as in distance_kernel.  Most queries are decided by their first line, a few need all
// 32.  A line walk is a chain of dependent node loads of very uneven length: a launch
// lasts at least as long as its longest walk, so the lines after the first run in a
// few batches of many lines each, in persistent waves whose idle lanes take the next
// line from a counter (a wave no longer waits for its longest walk):
//   stab_first : line 0 of every query (in the Morton order), one lane per query, which
//                also stores the query's two random offsets; the queries whose line
//                hits both ways make list 0;
//   stab_lines : lines k0 .. k0 + L - 1 of the listed queries, one item per (line,
//                entry), line-major; a line that misses marks its entry (bit 31);
//   stab_keep  : the unmarked entries make the next list;
//   stab_negate: the queries of the last list are inside.
// Lines after a query's first miss may be walked too; the decisions are the ones of
// walking a query's lines in order up to the first miss.  Queries number below 2^31,
// so bit 31 of an entry is free.  counts[0] = n, counts[1 + b] = length of list b;
// the item counter of line batch b is the 64-bit word at counts + 32 + 2 b.
constexpr uint32_t kStabMiss = 0x80000000u;
// batches of lines after line 0: [1, 4), [4, 12), [12, 32), for the roughly geometric
// decay of the undecided queries (measured on the soup's 2^22 queries: 19.6 ms per call
// against 20.8 with [1, 3), [3, 7), [7, 15), [15, 32), 21.7 with five batches and 22.0
// with [1, 4), [4, 32))
constexpr int kStabBatches = 3;
constexpr uint32_t kStabBatchFirst[kStabBatches + 1] = {1u, 4u, 12u, 32u};
// idle lanes of a wave take new items once this many of the 64 are idle (or all are)
constexpr int kRefillLanes = 16;
// waves per SIMD the persistent stab kernels are compiled for (72 VGPRs, a few spilled:
// 18.9 ms per call against 19.3 with 6 waves and 80 VGPRs; 8 waves do not fit the LDS
// and the compiler falls back to 5); the stab kernels run one-wave workgroups (19.3 ms
// against 19.6 with two) with 6 KB of LDS each, 4 x kStabWaves of them per CU (26 fit
// the LDS; the others start as the first ones leave)
constexpr int kStabWaves = 7;
constexpr int kStabBlock = 64;
constexpr int kStabBlocksPerCu = 4 * kStabWaves;

__global__ void stab_init(uint32_t* counts, int64_t n) {
    counts[threadIdx.x] = threadIdx.x == 0 ? (uint32_t)n : 0u;  // 64 threads: the 256-byte header
}

// Appends q of every lane with keep set to list (count words used), one atomic per
// wave; every lane of the wave must reach it.
__device__ __forceinline__ void append_kept(bool keep, uint32_t q, uint32_t* __restrict__ list, uint32_t* count) {
    const uint64_t m = ballot(keep);
    if (m == 0) return;
    const int lane = (int)(threadIdx.x & 63);
    const int leader = (int)__builtin_ctzll(m);
    uint32_t b = 0;
    if (lane == leader) b = atomicAdd(count, (uint32_t)__builtin_popcountll(m));
    b = (uint32_t)__builtin_amdgcn_readlane((int32_t)b, leader);
    if (keep) list[b + lanes_below(m)] = q;
}

// Persistent line walks: every lane walks items job.start(t) for t taken from
// *counter until the counter passes job.total; job.finish gets each result.  Every
// lane of the wave returns at the same time.
template <int BLOCK, int LDS_N, class Job>
__device__ void walk_lines(const Tree& T, const Job& job, unsigned long long* counter, LaneStack<BLOCK, LDS_N>& st) {
    const int lane = (int)(threadIdx.x & 63);
    LineWalk w;
    bool live = false, more = true;
    uint32_t tag = 0;
    for (;;) {
        const uint64_t idle = ballot(!live && more);
        if (idle != 0) {
            const uint64_t busy = ballot(live);
            if (busy == 0 || __builtin_popcountll(idle) >= kRefillLanes) {
                const int leader = (int)__builtin_ctzll(idle);
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(counter, (unsigned long long)__builtin_popcountll(idle));
                base = (unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)base, leader) |
                       ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)(base >> 32), leader) << 32);
                if (!live && more) {
                    const unsigned long long t = base + (unsigned long long)lanes_below(idle);
                    if (t >= job.total) {
                        more = false;
                    } else {
                        f3 o, d;
                        if (job.start(t, o, d, tag)) {
                            w.init(o, d, st);
                            live = true;
                        }
                    }
                }
            }
        }
        if (ballot(live) == 0) {
            if (ballot(more) == 0) return;
            continue;
        }
        if (live) {
            const int r = w.step(T, st);
            if (r >= 0) {
                job.finish(tag, r);
                live = false;
            }
        }
    }
}

struct LineBatchJob {  // item t: line k0 + t / m of entry t % m
    unsigned long long total;
    uint32_t m, k0;
    uint32_t* list;
    const float* positions;
    const float2* offs;
    __device__ __forceinline__ bool start(unsigned long long t, f3& o, f3& d, uint32_t& tag) const {
        const uint32_t line = (uint32_t)(t / m);
        const uint32_t j = (uint32_t)(t - (unsigned long long)line * m);
        const uint32_t e = __hip_atomic_load(&list[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (e & kStabMiss) return false;  // an earlier line of the batch already missed
        const float2 r = offs[e];
        o = load_pt(positions, e);
        d = fibonacci_dir32(k0 + line, r.x, r.y);
        tag = j;
        return true;
    }
    __device__ __forceinline__ void finish(uint32_t tag, int found) const {
        if (found != 3) __hip_atomic_fetch_or(&list[tag], kStabMiss, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
};

template <int BLOCK, int LDS_N>
__global__ __launch_bounds__(BLOCK) void stab_first(const uint32_t* __restrict__ order, uint32_t* __restrict__ list,
                                                    uint32_t* __restrict__ counts, const float* __restrict__ positions,
                                                    float2* __restrict__ offs, Tree T, uint32_t xchunk) {
    __shared__ int32_t s_ref[LDS_N * BLOCK];
    __shared__ float s_key[LDS_N * BLOCK];
    const int64_t i = (int64_t)xcd_remap(blockIdx.x, gridDim.x, xchunk) * BLOCK + threadIdx.x;
    LaneStack<BLOCK, LDS_N> st;
    st.ref = s_ref;
    st.key = s_key;
    st.tid = (int)threadIdx.x;
    st.sp = 0;
    bool keep = false;
    uint32_t q = 0;
    if (i < (int64_t)counts[0]) {
        q = order[i];
        Pcg32 rng;
        rng.advance((uint64_t)(uint32_t)(q * 2u));
        float2 o;
        o.x = rng.next_float();
        o.y = rng.next_float();
        offs[q] = o;
        keep = any_hit_line(T, load_pt(positions, q), fibonacci_dir32(0u, o.x, o.y), st) == 3;
    }
    append_kept(keep, q, list, &counts[1]);
}

template <int BLOCK, int LDS_N>
__global__ __launch_bounds__(BLOCK, kStabWaves) void stab_lines(uint32_t b, uint32_t* __restrict__ list, uint32_t* __restrict__ counts,
                                                    const float* __restrict__ positions,
                                                    const float2* __restrict__ offs, Tree T) {
    __shared__ int32_t s_ref[LDS_N * BLOCK];
    __shared__ float s_key[LDS_N * BLOCK];
    LaneStack<BLOCK, LDS_N> st;
    st.ref = s_ref;
    st.key = s_key;
    st.tid = (int)threadIdx.x;
    st.sp = 0;
    LineBatchJob job;
    job.m = counts[1 + b];
    job.k0 = kStabBatchFirst[b];
    job.total = (unsigned long long)job.m * (kStabBatchFirst[b + 1] - kStabBatchFirst[b]);
    job.list = list;
    job.positions = positions;
    job.offs = offs;
    walk_lines(T, job, (unsigned long long*)(counts + 32) + b, st);
}

// The unmarked entries of list_in (count_in of them) make list_out (count_out).
__global__ __launch_bounds__(256) void stab_keep(const uint32_t* __restrict__ list_in, const uint32_t* __restrict__ count_in,
                                                 uint32_t* __restrict__ list_out, uint32_t* __restrict__ count_out) {
    const uint32_t m = *count_in;
    const uint32_t wave0 = blockIdx.x * 256u + (threadIdx.x & ~63u);
    if (wave0 >= m) return;  // whole waves past the end; the others all reach append_kept
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    bool keep = false;
    uint32_t q = 0;
    if (j < m) {
        q = list_in[j];
        keep = (q & kStabMiss) == 0u;
    }
    append_kept(keep, q, list_out, count_out);
}

__global__ __launch_bounds__(256) void stab_negate(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count,
                                                   float* __restrict__ distances) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)*count) return;
    const uint32_t q = list[i];
    distances[q] = -distances[q];
}


constexpr int kNormalLds = 4;  // LDS entries per lane of the per-lane normal search

// Packet kernel for the unsigned (MODE 0) and watertight signed (MODE 1) distance, WAVES
// waves per workgroup.  Workgroup b walks tile xcd_remap(b) (runs of xchunk consecutive
// tiles of the Morton order per XCD); its lane i handles query order[64 tile + i] (or
// that index without a permutation).
template <int MODE, int WAVES>
__global__ __launch_bounds__(64 * WAVES, 8) void distance_packet_kernel(int64_t n, const float* __restrict__ positions,
                                                              const uint32_t* __restrict__ order,
                                                              float* __restrict__ distances,
                                                              int64_t* __restrict__ face_id, float* __restrict__ uvw,
                                                              Tree T, uint32_t xchunk) {
    const int64_t i = (int64_t)xcd_remap(blockIdx.x, gridDim.x, xchunk) * (64 * WAVES) + threadIdx.x;
    int64_t q = i;
    if (i < n && order) q = (int64_t)order[i];
    const bool valid = i < n && q < n;  // order entries >= n are skipped
    f3 p = mk3(0.0f, 0.0f, 0.0f);
    if (valid) p = load_pt(positions, q);
    float d2;
    int32_t slot, bid;
    __shared__ ExactQueue queues[WAVES];
    closest_face_packet(T, 0, p, valid, d2, slot, bid, queues[threadIdx.x >> 6]);
    if (slot < 0) {  // nothing within MAX_DIST: distance 0 and face 0
        slot = (int32_t)T.face0_slot;
        d2 = 0.0f;
    }
    f3 a, b, c;
    int32_t id = 0;
    a = b = c = mk3(0.0f, 0.0f, 0.0f);
    if (valid) load_tri(T.tris, slot, a, b, c, id);
    float dist = sqrtf(d2);
    f3 cp = mk3(0.0f, 0.0f, 0.0f);
    if (MODE == 1 || uvw) {
        if (valid) cp = tri_closest_pt(a, b, c, p);
    }
    if constexpr (MODE == 1) {
        // The normal search has a tiny radius around each lane's own closest point, so
        // lanes walk it one by one (a packet would walk the union of 64 paths).
        __shared__ int32_t s_ref[kNormalLds * 64 * WAVES];
        __shared__ float s_key[kNormalLds * 64 * WAVES];
        LaneStack<64 * WAVES, kNormalLds> st;
        st.ref = s_ref;
        st.key = s_key;
        st.tid = (int)threadIdx.x;
        st.sp = 0;
        if (valid) {
            f3 nrm = avg_normal(T, cp, st);
            dist = copysignf(dist, dot3(nrm, sub3(p, cp)));
        }
    }
    if (!valid) return;
    distances[q] = dist;
    face_id[q] = (int64_t)id;
    if (uvw) store_pt(uvw, q, tri_barycentric(a, b, c, cp));
}

// Unsigned distance over a forest: one wave per 64-slot tile of the padded query
// order; every tile belongs to one mesh, whose root node and face-0 slot come from
// mesh_info.  PERSISTENT waves take tiles from an atomic counter until none is left.
template <int PERSISTENT, int WAVES>
__global__ __launch_bounds__(64 * WAVES, PERSISTENT ? 7 : 8) void forest_distance_kernel(const float* __restrict__ positions,
                                                              const uint32_t* __restrict__ order,
                                                              const int64_t* __restrict__ pbase, int32_t n_meshes,
                                                              const int4* __restrict__ mesh_info,
                                                              float* __restrict__ distances,
                                                              int64_t* __restrict__ face_id, float* __restrict__ uvw,
                                                              TreeArrays arrays, uint32_t* counter,
                                                              int64_t max_tiles, uint32_t xcd_chunk) {
    __shared__ ExactQueue queues[WAVES];
    const int lane = (int)(threadIdx.x & 63);
    // max_tiles: the tiles the order array holds (bounds the walk even when the offsets
    // are not a CSR of the queries)
    const int64_t n_tiles = min(pbase[n_meshes] / 64, max_tiles);
    int64_t tile = PERSISTENT ? 0 : ((int64_t)xcd_remap(blockIdx.x, gridDim.x, xcd_chunk) * WAVES + (threadIdx.x >> 6));
    for (;;) {
        if (PERSISTENT) {
            uint32_t t = 0;
            if (lane == 0) t = atomicAdd(counter, 1u);
            tile = (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int32_t)t);
        }
        if (tile >= n_tiles) return;
        // mesh of this tile: largest m with pbase[m] <= 64 * tile (wave-uniform search)
        int32_t lo = 0, hi = n_meshes - 1;
        while (lo < hi) {
            int32_t mid = (lo + hi + 1) >> 1;
            if (pbase[mid] <= 64 * tile) lo = mid;
            else hi = mid - 1;
        }
        const int32_t mesh = uniform(lo);
        const int4 mi = mesh_info[mesh];
        const uint32_t qi = order[64 * tile + lane];
        const bool valid = qi != 0xffffffffu;
        f3 p = mk3(0.0f, 0.0f, 0.0f);
        if (valid) p = load_pt(positions, qi);
        const Tree T = make_tree(arrays, (int64_t)mi.y);
        float d2;
        int32_t slot, bid;
        closest_face_packet(T, uniform(mi.x), p, valid, d2, slot, bid, queues[threadIdx.x >> 6]);
        if (valid) {
            if (slot < 0) {
                slot = mi.y;
                d2 = 0.0f;
            }
            f3 a, b, c;
            int32_t id;
            load_tri(T.tris, slot, a, b, c, id);
            distances[qi] = sqrtf(d2);
            face_id[qi] = (int64_t)id;
            if (uvw) store_pt(uvw, qi, tri_barycentric(a, b, c, tri_closest_pt(a, b, c, p)));
        }
        if (!PERSISTENT) return;
    }
}

template <int BLOCK, int LDS_N>
__global__ __launch_bounds__(BLOCK) void ray_kernel(int64_t n, const float* __restrict__ rays_o,
                                                    const float* __restrict__ rays_d, float* __restrict__ positions,
                                                    int64_t* __restrict__ face_id, float* __restrict__ depth, Tree T) {
    __shared__ int32_t s_ref[LDS_N * BLOCK];
    __shared__ float s_key[LDS_N * BLOCK];
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    LaneStack<BLOCK, LDS_N> st;
    st.ref = s_ref;
    st.key = s_key;
    st.tid = (int)threadIdx.x;
    st.sp = 0;
    const f3 o = load_pt(rays_o, i);
    const f3 d = load_pt(rays_d, i);
    float t;
    int32_t slot = closest_hit(T, o, d, st, t);
    depth[i] = t;
    store_pt(positions, i, mk3(__builtin_fmaf(t, d.x, o.x), __builtin_fmaf(t, d.y, o.y), __builtin_fmaf(t, d.z, o.z)));
    int64_t fid = -1;
    if (slot >= 0) fid = (int64_t)__float_as_int(T.tris[3 * slot].w);
    face_id[i] = fid;
}

// Rays in the caller's order, one wave per 64.  A wave whose rays are nearly parallel
// and start close together (camera rays in pixel order: directions within ~5.7 degrees
// of its first ray, origins within 1% of the scene's diagonal) walks as a packet;
// any other wave walks lane by lane.  Both give closest_hit's results.
constexpr int kRayLds = 8;
template <int WAVES>
__global__ __launch_bounds__(64 * WAVES) void ray_packet_kernel(int64_t n, const float* __restrict__ rays_o,
                                                         const float* __restrict__ rays_d,
                                                         const uint32_t* __restrict__ order,
                                                         float* __restrict__ positions, int64_t* __restrict__ face_id,
                                                         float* __restrict__ depth, Tree T) {
    __shared__ int32_t s_ref[kRayLds * 64 * WAVES];
    __shared__ float s_key[kRayLds * 64 * WAVES];
    const int64_t g = (int64_t)blockIdx.x * (64 * WAVES) + threadIdx.x;
    int64_t i = g;  // ray index (lane g takes ray order[g] when an order is given)
    if (g < n && order) i = (int64_t)order[g];
    const bool valid = g < n && i < n;
    f3 o = mk3(0.0f, 0.0f, 0.0f), d = mk3(1.0f, 0.0f, 0.0f);
    if (valid) {
        o = load_pt(rays_o, i);
        d = load_pt(rays_d, i);
    }
    // coherence of the wave, judged against its first ray
    const f3 o0 = mk3(__shfl(o.x, 0), __shfl(o.y, 0), __shfl(o.z, 0));
    const f3 d0 = mk3(__shfl(d.x, 0), __shfl(d.y, 0), __shfl(d.z, 0));
    const UNode root = load_unode(T.nodes, 0);
    float diag2 = 0.0f;
    {
        const float lx = fminf(fminf(root.lx.x, root.lx.y), fminf(root.lx.z, root.lx.w));
        const float hx = fmaxf(fmaxf(root.hx.x, root.hx.y), fmaxf(root.hx.z, root.hx.w));
        const float ly = fminf(fminf(root.ly.x, root.ly.y), fminf(root.ly.z, root.ly.w));
        const float hy = fmaxf(fmaxf(root.hy.x, root.hy.y), fmaxf(root.hy.z, root.hy.w));
        const float lz = fminf(fminf(root.lz.x, root.lz.y), fminf(root.lz.z, root.lz.w));
        const float hz = fmaxf(fmaxf(root.hz.x, root.hz.y), fmaxf(root.hz.z, root.hz.w));
        // empty slots hold +inf / -inf bounds and drop out of the min / max
        diag2 = (hx - lx) * (hx - lx) + (hy - ly) * (hy - ly) + (hz - lz) * (hz - lz);
    }
    const f3 od = sub3(o, o0);
    const float dd = dot3(d, d0), n2 = dot3(d, d) * dot3(d0, d0);
    const bool near = !valid || (dd > 0.0f && dd * dd >= 0.99f * n2 && dot3(od, od) <= 1e-4f * diag2);
    float t;
    int32_t slot, bid;
    if (ballot(near) == ~0ull) {
        closest_hit_packet(T, o, d, valid, t, slot, bid);
    } else {
        if (!valid) return;
        LaneStack<64 * WAVES, kRayLds> st;
        st.ref = s_ref;
        st.key = s_key;
        st.tid = (int)threadIdx.x;
        st.sp = 0;
        slot = closest_hit(T, o, d, st, t);
        bid = slot >= 0 ? __float_as_int(T.tris[3 * slot].w) : 0;
    }
    if (!valid) return;
    depth[i] = t;
    store_pt(positions, i, mk3(__builtin_fmaf(t, d.x, o.x), __builtin_fmaf(t, d.y, o.y), __builtin_fmaf(t, d.z, o.z)));
    face_id[i] = slot >= 0 ? (int64_t)bid : -1;
}

}  // namespace dev

namespace {

constexpr int kBlock = 128;
constexpr int kLdsEntries = 12;
// Waves per workgroup of the packet distance and ray kernels: with one, a workgroup ends
// with its wave and the next starts at once (the 256^3 UDF walk 11.68 ms against 11.90
// with 4 and 13.47 with 8; the watertight SDF 889 against 868 Mq/s; camera rays 3119
// against 2899 Mrays/s).
constexpr int kDistanceWaves = 1;

// Runs of consecutive workgroups per XCD for a grid of nb one-wave workgroups over a
// Morton order: the largest power of two <= nb / 128, so that each XCD's L2 serves one
// aligned block of the order at a time (the 256^3 UDF walk: 10.80 ms with 2048 =
// 2^24 / 64 / 128, 11.38 ms without, 11.6 ms with 1536 or 3072, 11.68 ms with 8192).
uint32_t xcd_chunk_for(unsigned nb) {
    uint32_t c = 1;
    while ((uint64_t)c * 2 * 128 <= nb) c *= 2;
    return nb >= 128 * 8 ? c : 0u;
}

hipError_t launch_distance_impl(const QueryArgs& q, int mode) {
    const dev::Tree T = dev::make_tree(q.tree, q.face0_slot);
    hipStream_t s = (hipStream_t)q.stream;
    if ((mode == 0 || mode == 1) && q.packet) {
        constexpr int W = kDistanceWaves;
        const unsigned grid = (unsigned)((q.n + 64 * W - 1) / (64 * W));
        const uint32_t xchunk = xcd_chunk_for(grid);
        if (mode == 0)
            hipLaunchKernelGGL((dev::distance_packet_kernel<0, W>), dim3(grid), dim3(64 * W), 0, s, q.n, q.in0, q.order,
                               q.out_f, q.out_id, q.out_v, T, xchunk);
        else
            hipLaunchKernelGGL((dev::distance_packet_kernel<1, W>), dim3(grid), dim3(64 * W), 0, s, q.n, q.in0, q.order,
                               q.out_f, q.out_id, q.out_v, T, xchunk);
        return hipGetLastError();
    }
    const unsigned grid = (unsigned)((q.n + kBlock - 1) / kBlock);
    if (mode == 0)
        hipLaunchKernelGGL((dev::distance_kernel<kBlock, kLdsEntries, 0>), dim3(grid), dim3(kBlock), 0, s, q.n, q.in0,
                           q.order, q.out_f, q.out_id, q.out_v, T);
    else if (mode == 1)
        hipLaunchKernelGGL((dev::distance_kernel<kBlock, kLdsEntries, 1>), dim3(grid), dim3(kBlock), 0, s, q.n, q.in0,
                           q.order, q.out_f, q.out_id, q.out_v, T);
    else
        hipLaunchKernelGGL((dev::distance_kernel<kBlock, kLdsEntries, 2>), dim3(grid), dim3(kBlock), 0, s, q.n, q.in0,
                           q.order, q.out_f, q.out_id, q.out_v, T);
    return hipGetLastError();
}

}  // namespace

}  // namespace cubvh_amd

#include "forest.hpp"

namespace cubvh_amd {

namespace {

template <int WAVES>
hipError_t launch_forest_shape(const ForestArgs& a, const ForestWorkspace& ws, hipStream_t s) {
    const int64_t max_tiles = ws.npad_max / 64;
    if (a.persistent) {
        const unsigned blocks = (unsigned)std::max(1, a.n_cu * a.waves_per_cu / WAVES);
        hipLaunchKernelGGL((dev::forest_distance_kernel<1, WAVES>), dim3(blocks), dim3(64 * WAVES), 0, s, a.positions,
                           ws.order, ws.pbase, a.n_meshes, a.mesh_info, a.distances, a.face_id, a.uvw, a.tree,
                           ws.counter, max_tiles, 0u);
    } else {
        const unsigned blocks = (unsigned)((max_tiles + WAVES - 1) / WAVES);
        hipLaunchKernelGGL((dev::forest_distance_kernel<0, WAVES>), dim3(blocks), dim3(64 * WAVES), 0, s, a.positions,
                           ws.order, ws.pbase, a.n_meshes, a.mesh_info, a.distances, a.face_id, a.uvw, a.tree,
                           ws.counter, max_tiles, (uint32_t)a.xcd_chunk);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_forest_distance(const ForestArgs& a, const ForestWorkspace& ws, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    switch (a.waves_per_block) {
        case 1: return launch_forest_shape<1>(a, ws, s);
        case 2: return launch_forest_shape<2>(a, ws, s);
        case 4: return launch_forest_shape<4>(a, ws, s);
        case 8: return launch_forest_shape<8>(a, ws, s);
        default: return hipErrorInvalidValue;
    }
}

int stack_capacity_for(int need) { return need <= dev::STACK_TOTAL ? dev::STACK_TOTAL : -1; }

hipError_t launch_distance(const QueryArgs& q, int mode) {
    if (q.n <= 0) return hipSuccess;
    if (stack_capacity_for(q.stack_need) < 0) return hipErrorInvalidValue;
    return launch_distance_impl(q, mode);
}

size_t stab_workspace_bytes(int64_t n) {
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    return al(64) + 2 * al(4 * (size_t)n) + al(8 * (size_t)n);
}

hipError_t launch_stab_sign(const QueryArgs& q, void* ws) {
    if (q.n <= 0) return hipSuccess;
    if (stack_capacity_for(q.stack_need) < 0) return hipErrorInvalidValue;
    if (!q.order || !ws || q.n >= ((int64_t)1 << 31)) return hipErrorInvalidValue;  // bit 31 marks list entries
    const dev::Tree T = dev::make_tree(q.tree, q.face0_slot);
    hipStream_t s = (hipStream_t)q.stream;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    char* w = (char*)ws;
    uint32_t* counts = (uint32_t*)w;
    w += al(64);
    uint32_t* lists[2];
    lists[0] = (uint32_t*)w;
    w += al(4 * (size_t)q.n);
    lists[1] = (uint32_t*)w;
    w += al(4 * (size_t)q.n);
    float2* offs = (float2*)w;
    int dev_id = 0, n_cu = 256;
    (void)hipGetDevice(&dev_id);
    (void)hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev_id);
    // persistent waves: as many workgroups as the CUs hold at once
    constexpr int SB = dev::kStabBlock;
    const unsigned grid = (unsigned)std::min<int64_t>((q.n + SB - 1) / SB, (int64_t)n_cu * dev::kStabBlocksPerCu);
    const unsigned keep_grid = (unsigned)((q.n + 255) / 256);
    hipLaunchKernelGGL(dev::stab_init, dim3(1), dim3(64), 0, s, counts, q.n);
    const unsigned first_grid = (unsigned)((q.n + SB - 1) / SB);
    hipLaunchKernelGGL((dev::stab_first<SB, kLdsEntries>), dim3(first_grid), dim3(SB), 0, s, q.order, lists[0], counts,
                       q.in0, offs, T, xcd_chunk_for(first_grid));
    for (int b = 0; b < dev::kStabBatches; ++b) {
        hipLaunchKernelGGL((dev::stab_lines<SB, kLdsEntries>), dim3(grid), dim3(SB), 0, s, (uint32_t)b, lists[b & 1],
                           counts, q.in0, offs, T);
        hipLaunchKernelGGL(dev::stab_keep, dim3(keep_grid), dim3(256), 0, s, lists[b & 1], counts + 1 + b,
                           lists[(b + 1) & 1], counts + 2 + b);
    }
    hipLaunchKernelGGL(dev::stab_negate, dim3(keep_grid), dim3(256), 0, s, lists[dev::kStabBatches & 1],
                       counts + 1 + dev::kStabBatches, q.out_f);
    return hipGetLastError();
}

hipError_t launch_ray(const QueryArgs& q) {
    if (q.n <= 0) return hipSuccess;
    if (stack_capacity_for(q.stack_need) < 0) return hipErrorInvalidValue;
    const dev::Tree T = dev::make_tree(q.tree, q.face0_slot);
    if (q.packet) {
        constexpr int W = kDistanceWaves;  // one wave per workgroup, as the distance kernels (3119 vs 2899 Mrays/s with 4)
        // no runs per XCD here: ray costs differ between image regions, and runs put the
        // costly ones on one XCD (camera rays 2475 Mrays/s with the distance kernels' runs)
        const unsigned grid = (unsigned)((q.n + 64 * W - 1) / (64 * W));
        hipLaunchKernelGGL(dev::ray_packet_kernel<W>, dim3(grid), dim3(64 * W), 0, (hipStream_t)q.stream, q.n, q.in0,
                           q.in1, q.order, q.out_v, q.out_id, q.out_f, T);
        return hipGetLastError();
    }
    const unsigned grid = (unsigned)((q.n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL((dev::ray_kernel<kBlock, kLdsEntries>), dim3(grid), dim3(kBlock), 0, (hipStream_t)q.stream, q.n,
                       q.in0, q.in1, q.out_v, q.out_id, q.out_f, T);
    return hipGetLastError();
}

}  // namespace cubvh_amd
