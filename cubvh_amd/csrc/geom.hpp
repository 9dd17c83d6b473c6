// geom.hpp -- fp32 geometry on the device.
//
// Formulas follow the reference's primitives so that results agree with it:
//   tri_dist2        include/gpu/triangle.cuh:46-68  (Inigo Quilez unsigned distance)
//   tri_ray          include/gpu/triangle.cuh:27-44  (IQ intersector, 1e6 on a miss,
//                    exact reciprocal instead of safe_divide: see below)
//   tri_normal       include/gpu/triangle.cuh:23-25
//   tri_closest_pt   include/gpu/triangle.cuh:74-139
//   tri_barycentric  include/gpu/triangle.cuh:141-158 (u in double)
//   safe_divide      include/gpu/common.h:98-106
//   fibonacci_dir32  include/gpu/common.h:63-96 (N_DIRS = 32)
//   Pcg32            include/gpu/pcg32.h:33-170 (PCG-XSH-RR 64/32, Brown jump-ahead)
//
// Arithmetic is pinned so that it is reproducible on any IEEE machine: the library
// is compiled with -ffp-contract=off and the only fused operations are the explicit
// ones here -- a dot product is fma(z,z', fma(y,y', x*x')), a cross-product
// component is fma(a,b, -(c*d)).  Division and sqrt are IEEE correctly rounded
// (the HIP default).  sin/cos of the stab directions are evaluated in double and
// rounded to float.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cubvh_amd {
namespace dev {

struct f3 {
    float x, y, z;
};

__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 neg3(f3 a) { return mk3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot3(f3 a, f3 b) {
    return __builtin_fmaf(a.z, b.z, __builtin_fmaf(a.y, b.y, a.x * b.x));
}
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return mk3(__builtin_fmaf(a.y, b.z, -(a.z * b.y)), __builtin_fmaf(a.z, b.x, -(a.x * b.z)),
               __builtin_fmaf(a.x, b.y, -(a.y * b.x)));
}
__device__ __forceinline__ f3 xyz(float4 v) { return mk3(v.x, v.y, v.z); }
// std::min / std::max semantics (second operand only on a strict comparison)
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float clamp01(float v) { return v < 0.0f ? 0.0f : (1.0f < v ? 1.0f : v); }
__device__ __forceinline__ float sgn(float x) { return copysignf(1.0f, x); }

__device__ __forceinline__ float safe_divide(float num, float den) {
    const float eps = 1e-6f;
    if (fabsf(den) < eps) return den <= 0.0f ? -(num / eps) : num / eps;
    return num / den;
}

// Squared distance from p to triangle abc.
__device__ __forceinline__ float tri_dist2(f3 a, f3 b, f3 c, f3 p) {
    f3 v21 = sub3(b, a), p1 = sub3(p, a);
    f3 v32 = sub3(c, b), p2 = sub3(p, b);
    f3 v13 = sub3(a, c), p3 = sub3(p, c);
    f3 nor = cross3(v21, v13);
    float s = sgn(dot3(cross3(v21, nor), p1)) + sgn(dot3(cross3(v32, nor), p2)) + sgn(dot3(cross3(v13, nor), p3));
    if (s < 2.0f) {
        float t1 = clamp01(dot3(v21, p1) / dot3(v21, v21));
        float t2 = clamp01(dot3(v32, p2) / dot3(v32, v32));
        float t3 = clamp01(dot3(v13, p3) / dot3(v13, v13));
        f3 e1 = mk3(__builtin_fmaf(v21.x, t1, -p1.x), __builtin_fmaf(v21.y, t1, -p1.y), __builtin_fmaf(v21.z, t1, -p1.z));
        f3 e2 = mk3(__builtin_fmaf(v32.x, t2, -p2.x), __builtin_fmaf(v32.y, t2, -p2.y), __builtin_fmaf(v32.z, t2, -p2.z));
        f3 e3 = mk3(__builtin_fmaf(v13.x, t3, -p3.x), __builtin_fmaf(v13.y, t3, -p3.y), __builtin_fmaf(v13.z, t3, -p3.z));
        return smin(smin(dot3(e1, e1), dot3(e2, e2)), dot3(e3, e3));
    }
    float q = dot3(nor, p1);
    return q * q / dot3(nor, nor);
}

// The face-only terms of tri_dist2, computed once per face with the same operations
// (edges v21 = b - a, v32 = c - b, v13 = a - c; nor = v21 x v13; g_k = v_k x nor; the
// squared lengths are dot3 of each vector with itself).
struct FaceConst {
    f3 a, b, c, v21, v32, v13, nor, g1, g2, g3;
    float l21, l32, l13, lnn;
    int32_t id;
};

// tri_dist2 over precomputed face terms: the same operations on the same values, so
// the same result bit for bit.
__device__ __forceinline__ float tri_dist2_pre(const FaceConst& f, f3 p) {
    f3 p1 = sub3(p, f.a), p2 = sub3(p, f.b), p3 = sub3(p, f.c);
    float s = sgn(dot3(f.g1, p1)) + sgn(dot3(f.g2, p2)) + sgn(dot3(f.g3, p3));
    if (s < 2.0f) {
        float t1 = clamp01(dot3(f.v21, p1) / f.l21);
        float t2 = clamp01(dot3(f.v32, p2) / f.l32);
        float t3 = clamp01(dot3(f.v13, p3) / f.l13);
        f3 e1 = mk3(__builtin_fmaf(f.v21.x, t1, -p1.x), __builtin_fmaf(f.v21.y, t1, -p1.y),
                    __builtin_fmaf(f.v21.z, t1, -p1.z));
        f3 e2 = mk3(__builtin_fmaf(f.v32.x, t2, -p2.x), __builtin_fmaf(f.v32.y, t2, -p2.y),
                    __builtin_fmaf(f.v32.z, t2, -p2.z));
        f3 e3 = mk3(__builtin_fmaf(f.v13.x, t3, -p3.x), __builtin_fmaf(f.v13.y, t3, -p3.y),
                    __builtin_fmaf(f.v13.z, t3, -p3.z));
        return smin(smin(dot3(e1, e1), dot3(e2, e2)), dot3(e3, e3));
    }
    float q = dot3(f.nor, p1);
    return q * q / f.lnn;
}

// Unpacks a 36-float face record (bvh.hpp: tpre) given as nine float4 rows.
__device__ __forceinline__ FaceConst face_const(float4 r0, float4 r1, float4 r2, float4 r3, float4 r4, float4 r5,
                                                float4 r6, float4 r7, float4 r8) {
    FaceConst f;
    f.a = xyz(r0);
    f.id = __float_as_int(r0.w);
    f.b = xyz(r1);
    f.l21 = r1.w;
    f.c = xyz(r2);
    f.l32 = r2.w;
    f.v21 = xyz(r3);
    f.l13 = r3.w;
    f.v32 = xyz(r4);
    f.lnn = r4.w;
    f.v13 = xyz(r5);
    f.g1 = xyz(r6);
    f.g2 = xyz(r7);
    f.g3 = xyz(r8);
    f.nor = mk3(r5.w, r6.w, r7.w);
    return f;
}

// Ray parameter of the hit with triangle abc, 1e6 on a miss.  The reference takes
// 1/(d.n) through safe_divide, which clamps |d.n| at 1e-6 and on grazing rays scales
// u, v and t by |d.n|/1e-6 (reporting hits far from the triangle, or at the wrong
// depth).  Here the reciprocal is exact (d.n == 0 is a miss), which is the same
// computation whenever |d.n| >= 1e-6 and the correct one below it.
__device__ __forceinline__ float tri_ray(f3 a, f3 b, f3 c, f3 ro, f3 rd) {
    f3 v1v0 = sub3(b, a), v2v0 = sub3(c, a), rov0 = sub3(ro, a);
    f3 n = cross3(v1v0, v2v0);
    f3 q = cross3(rov0, rd);
    float den = dot3(rd, n);
    if (den == 0.0f) return 1e6f;
    float d = 1.0f / den;
    float u = d * -dot3(q, v2v0);
    float v = d * dot3(q, v1v0);
    float t = d * -dot3(n, rov0);
    if (!(u >= 0.0f && u <= 1.0f && v >= 0.0f && (u + v) <= 1.0f && t >= 0.0f)) t = 1e6f;
    return t;
}

// Both rays o + t rd and o - t rd (t >= 0) against triangle abc in one evaluation:
// bit 0 set when tri_ray(a, b, c, ro, rd) < tmax, bit 1 when tri_ray(a, b, c, ro, -rd)
// < tmax.  Negating rd negates q, den, the reciprocal and t exactly and leaves u and v
// unchanged, so these are the two calls' results bit for bit.
__device__ __forceinline__ int tri_line(f3 a, f3 b, f3 c, f3 ro, f3 rd, float tmax) {
    f3 v1v0 = sub3(b, a), v2v0 = sub3(c, a), rov0 = sub3(ro, a);
    f3 n = cross3(v1v0, v2v0);
    f3 q = cross3(rov0, rd);
    float den = dot3(rd, n);
    if (den == 0.0f) return 0;
    float d = 1.0f / den;
    float u = d * -dot3(q, v2v0);
    float v = d * dot3(q, v1v0);
    float t = d * -dot3(n, rov0);
    if (!(u >= 0.0f && u <= 1.0f && v >= 0.0f && (u + v) <= 1.0f)) return 0;
    const float tn = -t;
    return ((t >= 0.0f && t < tmax) ? 1 : 0) | ((tn >= 0.0f && tn < tmax) ? 2 : 0);
}

__device__ __forceinline__ f3 tri_normal(f3 a, f3 b, f3 c) {
    f3 n = cross3(sub3(b, a), sub3(c, a));
    float z = dot3(n, n);
    if (z > 0.0f) {
        float s = sqrtf(z);
        return mk3(n.x / s, n.y / s, n.z / s);
    }
    return n;
}

__device__ __forceinline__ f3 closest_on_segment(f3 a, f3 b, f3 p) {
    f3 ba = sub3(b, a);
    float t = dot3(sub3(p, a), ba) / dot3(ba, ba);
    t = smax(smin(t, 1.0f), 0.0f);
    return mk3(__builtin_fmaf(t, ba.x, a.x), __builtin_fmaf(t, ba.y, a.y), __builtin_fmaf(t, ba.z, a.z));
}

__device__ __forceinline__ f3 tri_closest_pt(f3 a, f3 b, f3 c, f3 point) {
    f3 n = tri_normal(a, b, c);
    float s = dot3(n, sub3(point, a));
    point = mk3(__builtin_fmaf(-s, n.x, point.x), __builtin_fmaf(-s, n.y, point.y), __builtin_fmaf(-s, n.z, point.z));
    f3 la = sub3(a, point), lb = sub3(b, point), lc = sub3(c, point);
    f3 u = cross3(lb, lc), v = cross3(lc, la), w = cross3(la, lb);
    if (!(dot3(u, v) < 0.0f) && !(dot3(u, w) < 0.0f)) return point;
    f3 c1 = closest_on_segment(a, b, point);
    f3 c2 = closest_on_segment(b, c, point);
    f3 c3 = closest_on_segment(c, a, point);
    f3 d1 = sub3(point, c1), d2 = sub3(point, c2), d3 = sub3(point, c3);
    float m1 = dot3(d1, d1), m2 = dot3(d2, d2), m3 = dot3(d3, d3);
    float m = smin(smin(m1, m2), m3);
    if (m == m1) return c1;
    if (m == m2) return c2;
    return c3;
}

__device__ __forceinline__ f3 tri_barycentric(f3 a, f3 b, f3 c, f3 p) {
    f3 v0 = sub3(b, a), v1 = sub3(c, a), v2 = sub3(p, a);
    float d00 = dot3(v0, v0), d01 = dot3(v0, v1), d11 = dot3(v1, v1);
    float d20 = dot3(v2, v0), d21 = dot3(v2, v1);
    float denom = __builtin_fmaf(d00, d11, -(d01 * d01));
    float v = __builtin_fmaf(d11, d20, -(d01 * d21)) / denom;
    float w = __builtin_fmaf(d00, d21, -(d01 * d20)) / denom;
    float u = (float)(1.0 - (double)v - (double)w);
    return mk3(u, v, w);
}

struct Pcg32 {
    uint64_t state = 0x853c49e6748fea9bULL;
    uint64_t inc = 0xda3e39cb94b95bdbULL;
    __device__ __forceinline__ uint32_t next_uint() {
        uint64_t old = state;
        state = old * 0x5851f42d4c957f2dULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31));
    }
    __device__ __forceinline__ float next_float() {
        return __uint_as_float((next_uint() >> 9) | 0x3f800000u) - 1.0f;
    }
    __device__ __forceinline__ void advance(uint64_t delta) {
        uint64_t cur_mult = 0x5851f42d4c957f2dULL, cur_plus = inc, acc_mult = 1u, acc_plus = 0u;
        while (delta > 0) {
            if (delta & 1) {
                acc_mult *= cur_mult;
                acc_plus = acc_plus * cur_mult + cur_plus;
            }
            cur_plus = (cur_mult + 1) * cur_plus;
            cur_mult *= cur_mult;
            delta >>= 1;
        }
        state = acc_mult * state + acc_plus;
    }
};

__device__ __forceinline__ float fractf(float x) { return x - floorf(x); }

__device__ __forceinline__ f3 fibonacci_dir32(uint32_t i, float ox, float oy) {
    const float eps = 1.33f;
    const float GOLDEN = 1.6180339887498948482045868343656f;
    const float PI = 3.14159265358979323846f;
    float px = fractf(((float)i + eps) / ((float)31u + 2.0f * eps) + ox);
    float py = fractf((float)i / GOLDEN + oy);
    float cos_theta = __builtin_fmaf(-2.0f, px, 1.0f);
    float phi = (2.0f * PI) * (py - 0.5f);
    float sin_theta = sqrtf(fmaxf(__builtin_fmaf(-cos_theta, cos_theta, 1.0f), 0.0f));
    double sp, cp;
    sincos((double)phi, &sp, &cp);
    return mk3(sin_theta * (float)cp, sin_theta * (float)sp, cos_theta);
}

}  // namespace dev
}  // namespace cubvh_amd
