// Device helpers shared by the particle-stage kernels (gfx950, FP64).
//
// * the Lorentz factor, graded-axis location and patch search of the reference tracer
//   (kernels/trace_trajectories_2d/trace_trajectories_2d_pyccel.py:19-39, 165-171, 215-222);
// * the clamped staggered bilinear gather (staggered_field_2d_pyccel.py:36-55);
// * 128-bit fixed-point accumulation: a value (a deposit, or one trajectory's FP64 sum of
//   its deposits into a node) is rounded once to a multiple of 2^e and added as an exact
//   integer, so a node's total does not depend on the order of the atomic adds (the
//   reference adds FP64 in an order set by its threads).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bt {

typedef int64_t i64;
typedef unsigned long long u64;

// gamma of the proper velocity u: sqrt(1 + u.u / c^2); exactly 1 for inv_c2 == 0.
__device__ __forceinline__ double lorentz_factor(double u0, double u1, double u2, double inv_c2) {
    return sqrt(1.0 + (u0 * u0 + u1 * u1 + u2 * u2) * inv_c2);
}

// The edge [nodes[c], nodes[c+1]] holding x (pyccel.py:19-33).
__device__ __forceinline__ i64 node_cell(double x, const double* __restrict__ nodes, i64 n) {
    i64 lo = 0, hi = n - 1;
    if (x >= nodes[n - 1]) {
        lo = n - 2;
    } else if (x > nodes[0]) {
        while (hi - lo > 1) {
            const i64 mid = (lo + hi) / 2;
            if (nodes[mid] <= x) lo = mid; else hi = mid;
        }
    }
    return lo;
}

// The uniform coordinate x0 + h * index of x on a graded axis (pyccel.py:36-39).
__device__ __forceinline__ double uniform_x(double x, const double* __restrict__ nodes, i64 n,
                                            double x0, double h) {
    const i64 c = node_cell(x, nodes, n);
    return x0 + h * ((double)c + (x - nodes[c]) / (nodes[c + 1] - nodes[c]));
}

// floor(s) as an index clamped to [lo, hi].  Clamping the double first keeps the
// conversion in range for points far off the mesh; the clamped index is the reference's.
__device__ __forceinline__ int clamped_floor(double s, int lo, int hi) {
    double f = floor(s);
    f = fmax(f, (double)lo);
    f = fmin(f, (double)hi);
    return (int)f;
}

// Clamped bilinear of the staggered grid f (rows of `cols` values) at fractional index
// (s, t): i in [0, imax], j in [0, jmax] (staggered_field_2d_pyccel.py:38-55).
__device__ __forceinline__ double clamped_bilinear(const double* __restrict__ f, i64 cols,
                                                   double s, double t, int imax, int jmax) {
    const int i = clamped_floor(s, 0, imax);
    const int j = clamped_floor(t, 0, jmax);
    const double u = s - (double)i;
    const double w = t - (double)j;
    // 32-bit offsets (the arrays hold fewer than 2^29 values): base + offset addressing
    const unsigned c = (unsigned)i * (unsigned)cols + (unsigned)j;
    const unsigned d = c + (unsigned)cols;
    const double f00 = f[c], f01 = f[c + 1], f10 = f[d], f11 = f[d + 1];
    const double low = f00 + w * (f01 - f00);
    const double high = f10 + w * (f11 - f10);
    return low + u * (high - low);
}

// ---- square roots through the hardware reciprocal square root -------------------------
// The reference's divisions by |v|, r and sqrt(|a|) and its hypot become products with
// 1 / sqrt(x): the hardware estimate y (v_rsq_f64, relative error below 2^-23) and one
// third-order correction.  With e = 1 - x y^2 the exact value is y (1 - e)^(-1/2) =
// y (1 + e/2 + 3e^2/8 + ...); the truncation leaves about 5/16 e^3 < 2^-66, and the rounding
// of the five operations about one unit in the last place -- far inside the 1e-10 relative bar
// of the trace; bt_arith_selftest measures it on the GPU.

// 1 / sqrt(x) for a positive normal x of at least 2^-1000.
__device__ __forceinline__ double rsqrt_nr(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double e = __builtin_fma(-(x * y), y, 1.0);
    const double p = __builtin_fma(e, 0.375, 0.5);
    return __builtin_fma(y * e, p, y);
}

// (x - o) / h on a uniform axis as one fused multiply-add with inv_h = RN(1 / h):
// RN(x inv_h - RN(o inv_h)); for o = 0 the same bits as RN(RN(x - o) inv_h), otherwise within
// an ulp of the axis' extent in cells (the product is formed once per launch).
__device__ __forceinline__ double axis_coord(double x, double o, double inv_h) {
    return __builtin_fma(x, inv_h, -(o * inv_h));
}

// 1 / d for a positive normal d: the hardware estimate r (v_rcp_f64) and one third-order
// correction, r (1 + e + e^2) with e = 1 - d r (truncation about e^3 < 2^-66).
__device__ __forceinline__ double recip_nr(double d) {
    const double r = __builtin_amdgcn_rcp(d);
    const double e = __builtin_fma(-d, r, 1.0);
    return __builtin_fma(r, __builtin_fma(e, e, e), r);
}

// 1 / sqrt(x) for any positive finite x (tiny ones scaled by an exact power of two).
__device__ __forceinline__ double rsqrt_pos(double x) {
    if (x >= 0x1p-1000) return rsqrt_nr(x);
    return rsqrt_nr(x * 0x1p600) * 0x1p300;
}

// The distance r = |(y, z)| from the axis and the direction (y / r, z / r); (0, 0) on the
// axis, where r = 0 (pyccel.py:196-205: r = hypot(y, z), then y / r, z / r when r > 0).
struct Radial {
    double r, cy, cz;
    double ir;  // 1 / r off the axis (0 on it)
};

// The common case is computed without a branch (the GPU does not trap on the reciprocal
// square root of 0 or a NaN), and one rarely taken branch replaces it near the axis.
__device__ __forceinline__ Radial radial(double y, double z) {
    const double s2 = __builtin_fma(y, y, z * z);
    Radial o;
    const double ir = rsqrt_nr(s2);
    o.r = s2 * ir;
    o.cy = y * ir;
    o.cz = z * ir;
    o.ir = ir;
    if (__builtin_expect(!(s2 >= 0x1p-1000), 0)) {
        if (s2 > 0.0) {  // within 1e-150 of the axis: the same with y, z scaled by 2^600
            const double ys = y * 0x1p600, zs = z * 0x1p600;
            const double s2s = __builtin_fma(ys, ys, zs * zs);
            const double irs = rsqrt_nr(s2s);
            o.r = (s2s * irs) * 0x1p-600;
            o.cy = ys * irs;
            o.cz = zs * irs;
            o.ir = irs * 0x1p600;
        } else {  // on the axis (or a NaN coordinate, which stays NaN in r)
            o.r = s2;
            o.cy = 0.0;
            o.cz = 0.0;
            o.ir = 0.0;
        }
    }
    return o;
}

// |u|^2 as fused multiply-adds (within an ulp of the sum of the three products).
__device__ __forceinline__ double norm2(double u0, double u1, double u2) {
    return __builtin_fma(u0, u0, __builtin_fma(u1, u1, u2 * u2));
}

// The step-size rule (pyccel.py:282-295): dt = f min(h / |v|, sqrt(2 h / |a|)) over the
// terms that exist, h when neither does; |v| = gv sqrt(v2), |a| = sqrt(a2) (gv = 1 / gamma,
// exactly 1 for classical motion).  `branch`: 1 = h / |v|, 2 = sqrt(2 h / |a|), 0 = h;
// `speed`, `accel` = |v|, |a| where they exist (0 otherwise).
struct StepRule {
    double dt, speed, accel;
    int branch;
};

__device__ __forceinline__ StepRule step_rule(double v2, double a2, double h, double f, double gv) {
    StepRule o;
    // the common case (|v| and |a| of normal size) without a branch: the same operations as
    // the general rule below on these inputs, so the same bits.  d2 = sqrt(q) can only be the
    // smaller term when q < d1^2; when q exceeds d1^2 by a relative 2^-40 (far more than the
    // few units in the last place of the computed d2 and d1^2) the speed term is chosen without
    // the third square root, with the bits of the full comparison.
    {
        const double iv = rsqrt_nr(v2);
        const double ia = rsqrt_nr(a2);
        const double q = (2.0 * h) * ia;
        const double d1 = (h * iv) / gv;
        if (__builtin_expect(v2 >= 0x1p-1000 && a2 >= 0x1p-1000 && q >= 0x1p-1000, 1)) {
            o.speed = gv * (v2 * iv);
            o.accel = a2 * ia;
            if (__builtin_expect(q > (d1 * d1) * (1.0 + 0x1p-40), 1)) {
                o.branch = 1;
                o.dt = f * d1;
                return o;
            }
            const double d2 = q * rsqrt_nr(q);
            const bool second = d2 < d1;
            o.branch = second ? 2 : 1;
            o.dt = f * (second ? d2 : d1);
            return o;
        }
    }
    double dt = -1.0;
    o.branch = 0;
    o.speed = 0.0;
    o.accel = 0.0;
    if (v2 > 0.0) {
        const double iv = rsqrt_pos(v2);
        o.speed = gv * (v2 * iv);
        dt = (h * iv) / gv;
        o.branch = 1;
    }
    if (a2 > 0.0) {
        const double ia = rsqrt_pos(a2);
        o.accel = a2 * ia;
        const double q = (2.0 * h) * ia;
        const double d2 = q * rsqrt_pos(q);
        if (dt < 0.0 || d2 < dt) {
            dt = d2;
            o.branch = 2;
        }
    }
    if (dt < 0.0) dt = h;
    o.dt = f * dt;
    return o;
}

// ---- buffer gathers -------------------------------------------------------------------
// A field array read through a buffer resource: 32-bit byte offsets (no 64-bit address
// arithmetic per load) and both values of a row pair f[c], f[c + 1] in one 16-byte load.
typedef __amdgpu_buffer_rsrc_t Rsrc;

// The raw buffer of `count` doubles at `base` (count < 2^28; gfx9 resource word 3).
__device__ __forceinline__ Rsrc field_rsrc(const double* base, i64 count) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(base), (short)0, (int)(count * 8), 0x00020000);
}

struct Pair {
    double a, b;
};

__device__ __forceinline__ Pair load_pair(Rsrc r, unsigned elem) {
    return __builtin_bit_cast(Pair, __builtin_amdgcn_raw_buffer_load_b128(r, elem * 8u, 0, 0));
}

// The same at element `elem` + `base`, with `base` the same in every lane: it goes into the
// load's scalar offset, so the lanes' offsets need no add for it.
__device__ __forceinline__ Pair load_pair_at(Rsrc r, unsigned elem, unsigned base) {
    return __builtin_bit_cast(Pair, __builtin_amdgcn_raw_buffer_load_b128(r, elem * 8u, base * 8u, 0));
}

// trunc(s) saturated to the 32-bit range, NaN to 0: the hardware conversion v_cvt_i32_f64
// (a C++ conversion of an out-of-range double is undefined, so the instruction is named).
__device__ __forceinline__ int trunc_sat_i32(double s) {
    int i;
    asm("v_cvt_i32_f64 %0, %1" : "=v"(i) : "v"(s));
    return i;
}

// floor(s) clamped to [0, hi] (hi >= 0), and the same for a NaN s (0): trunc equals floor
// from 0 up, and everything below 0 clamps to 0 either way -- one conversion and one
// integer median instead of floor, two double clamps and a conversion.  UNIFORM: `hi` is the
// same in every lane (a scalar register), and the median is the single v_med3_i32 (the
// compiler forms it only for constant bounds).
template <bool UNIFORM = false>
__device__ __forceinline__ int clamped_floor_index(double s, int hi) {
    const int t = trunc_sat_i32(s);
    if (UNIFORM) {
        int r;
        asm("v_med3_i32 %0, %1, 0, %2" : "=v"(r) : "v"(t), "s"(hi));
        return r;
    }
    return min(max(t, 0), hi);
}

// a * b + c for a, b below 2^24 (v_mad_u32_u24, full rate; a 32-bit multiply is quarter rate)
__device__ __forceinline__ unsigned mad24(unsigned a, unsigned b, unsigned c) {
    return __umul24(a, b) + c;
}

// clamped_bilinear of the grid that starts at element `base` of buffer r, for the tracer:
// the cell clamped as an integer (the same cell and fractions as clamping floor(s) in
// double), the three interpolations fused (fma: one rounding each instead of two).  Rows and
// columns stay below 2^24 (the entry points check it).  UNIFORM: the bounds are the same in
// every lane.
// The gather in two halves: bilinear_issue finds the cell and its fractions and issues the
// two row-pair loads; bilinear_value interpolates once they are in (first use of the loaded
// values), so independent work can run between the two.
struct Bilinear {
    Pair lo, hi;
    double u, w;
};

template <bool UNIFORM = false>
__device__ __forceinline__ Bilinear bilinear_issue(Rsrc r, unsigned base, i64 cols, double s, double t, int imax,
                                                   int jmax) {
    Bilinear b;
    const int i = clamped_floor_index<UNIFORM>(s, imax);
    const int j = clamped_floor_index<UNIFORM>(t, jmax);
    b.u = s - (double)i;
    b.w = t - (double)j;
    const unsigned c = mad24((unsigned)i, (unsigned)cols, (unsigned)j);
    if (UNIFORM) {  // base and the row length are the same in every lane: scalar offsets
        b.lo = load_pair_at(r, c, base);
        b.hi = load_pair_at(r, c, base + (unsigned)cols);
    } else {
        b.lo = load_pair(r, c + base);
        b.hi = load_pair(r, c + base + (unsigned)cols);
    }
    return b;
}

__device__ __forceinline__ double bilinear_value(const Bilinear& b) {
    const double low = __builtin_fma(b.w, b.lo.b - b.lo.a, b.lo.a);
    const double high = __builtin_fma(b.w, b.hi.b - b.hi.a, b.hi.a);
    return __builtin_fma(b.u, high - low, low);
}

template <bool UNIFORM = false>
__device__ __forceinline__ double clamped_bilinear_buf(Rsrc r, unsigned base, i64 cols, double s, double t,
                                                       int imax, int jmax) {
    return bilinear_value(bilinear_issue<UNIFORM>(r, base, cols, s, t, imax, jmax));
}

// ---- fixed point ---------------------------------------------------------------------
// A deposit is rounded once to an integer multiple q of 2^e (|q| < 2^94).  A node holds two
// u64 words: w0 sums the low 32 bits of each q (as an unsigned number below 2^32) and w1
// sums q >> 32 (arithmetic); the node's value is (w1 2^32 + w0) 2^e, exact in any order of
// the additions, which are no-return integer atomics.  w0 stays exact for up to 2^31
// deposits into one node.
constexpr int kWords = 2;

// The two words of q = round(value 2^-e) (to nearest, ties to even); false for a zero,
// non-finite or out-of-range deposit (dropped).  Below 2^51 the rounding is that of adding
// 1.5 2^52, whose bit pattern then holds q exactly.
__device__ __forceinline__ bool to_fixed(double value, int e, u64& w0, u64& w1) {
    const double t = ldexp(value, -e);
    if (fabs(t) < 0x1p51) {
        const i64 q = __double_as_longlong(t + 0x1.8p52) - 0x4338000000000000ll;
        w0 = (u64)q & 0xffffffffull;
        w1 = (u64)(q >> 32);
        return q != 0;
    }
    const double s = rint(t);
    if (!(fabs(s) < 0x1p94)) return false;
    if (fabs(s) < 0x1p63) {
        const i64 q = (i64)s;
        if (q == 0) return false;
        w0 = (u64)q & 0xffffffffull;
        w1 = (u64)(q >> 32);
        return true;
    }
    const double hd = floor(s * 0x1p-32);  // exact: s is a multiple of 2^11 here
    w1 = (u64)(i64)hd;
    w0 = (u64)(s - hd * 0x1p32);           // exact, in [0, 2^32)
    return true;
}

// Adds the words to a node's words in global memory (no-return atomics).
__device__ __forceinline__ void add_words(u64* __restrict__ w, u64 w0, u64 w1) {
    if (w0) __hip_atomic_fetch_add(w, w0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (w1) __hip_atomic_fetch_add(w + 1, w1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One deposit of `value` into `node` of the word array `acc`.
__device__ __forceinline__ void fixed_add(u64* __restrict__ acc, i64 node, double value, int e) {
    u64 w0, w1;
    if (to_fixed(value, e, w0, w1)) add_words(acc + kWords * node, w0, w1);
}

// The node value as a double, rounded once (correctly) from the 128-bit integer.
__device__ __forceinline__ double fixed_value(u64 lo, i64 hi_signed, int e) {
    u64 hi = (u64)hi_signed;
    const bool neg = hi_signed < 0;
    if (neg) {
        lo = ~lo + 1ull;
        hi = ~hi + (lo == 0 ? 1ull : 0ull);
    }
    double m;
    if (hi == 0) {
        m = (double)lo;
    } else {
        const int lz = __clzll((long long)hi);
        const int bits = 64 - lz;  // significant bits of hi, 1..64
        u64 top, sticky;
        if (lz == 0) {
            top = hi;
            sticky = lo != 0;
        } else {
            top = (hi << lz) | (lo >> bits);
            sticky = (lo << lz) != 0;
        }
        m = ldexp((double)(top | sticky), bits);
    }
    return ldexp(neg ? -m : m, e);
}

// The double of a node's two words: (w1 2^32 + w0) 2^e, rounded once.
__device__ __forceinline__ double words_value(u64 w0, u64 w1, int e) {
    const u64 lo = w0 + (w1 << 32);
    const u64 carry = lo < w0 ? 1ull : 0ull;
    const i64 hi = ((i64)w1 >> 32) + (i64)carry;
    return fixed_value(lo, hi, e);
}

}  // namespace bt
