// The discrete adjoint of the trace: dJ/dfx, dJ/dfy for an objective J of the plane
// crossings, for the electrostatic, classical trace on one uniform mesh (RK4 or Boris).
//
// The forward map is the trace kernel's step (kernels/trace_trajectories_2d/
// trace_trajectories_2d_pyccel.py:178-370 with gamma = 1): stage points, the field gather at
// each (pyccel.py:196-269), the step-size rule dt = f min(h / |v|, sqrt(2 h / |a|))
// (pyccel.py:282-295) and the crossing record (x_p, y, z, v) interpolated at
// s = before / (before - after) (pyccel.py:426-442).  The adjoint differentiates exactly
// that arithmetic: the chosen branch of the step rule, the gather's cell (the derivative of
// the bilinear form inside it, including the clamped cells' extrapolation), the
// interpolation fraction.  Given dJ/d(record) for every crossing (the seeds), one lane
// per trajectory
//   1. steps forward from the initial state to its last seeded step, storing the state
//      every `interval` steps (checkpoints);
//   2. for each segment between checkpoints, last first: steps forward again from the
//      checkpoint storing every step's start state, then walks the segment backward,
//      recomputing each step's stages from its start state and applying the transposed
//      step to the adjoint state (lambda_x, lambda_v); the transposed gather scatters
//      lambda_E times the bilinear weights into dJ/dfx and dJ/dfy, through a window of
//      node sums in LDS that each wave moves with its lanes (the segments run in step
//      across the wave).
// When the groups of 64 trajectories outnumber the waves the grid holds at once, the walk of
// each group runs as several tasks (pieces) of its segments, upper segments first, from the
// forward trace's checkpoints: the adjoint state at a piece's lower end is handed to the
// group's next piece through stores written through to the coherent level and a per-group
// count of finished pieces (the next piece waits for it, boundedly).  A last round of whole
// walks would hold few waves on some SIMDs and two on others; pieces make it short.
// The forward steps here repeat the trace kernel's operations in the same order, so they
// reproduce its states bit for bit.
#include <hip/hip_runtime.h>

#include "bt_device.h"
#include "bt_internal.h"

namespace bt {

struct Vec3 {
    double a, b, c;
};

__device__ __forceinline__ Vec3 vadd(Vec3 x, Vec3 y) { return {x.a + y.a, x.b + y.b, x.c + y.c}; }
// x + s y, fused (the trace kernel's stage updates)
__device__ __forceinline__ Vec3 vaxpy(Vec3 x, double s, Vec3 y) {
    return {__builtin_fma(s, y.a, x.a), __builtin_fma(s, y.b, x.b), __builtin_fma(s, y.c, x.c)};
}
__device__ __forceinline__ Vec3 vscale(double s, Vec3 y) { return {s * y.a, s * y.b, s * y.c}; }
__device__ __forceinline__ double vdot(Vec3 x, Vec3 y) { return x.a * y.a + x.b * y.b + x.c * y.c; }

// The two edge-field buffers (uniform).
struct AdjFields {
    Rsrc fx, fy;
};

// The launch's parameters in the kernel-argument segment (the kernel's only argument), read
// through a pointer that the compiler cannot follow across an empty asm: the fields that only
// a refill, a seeded step, a flush to global memory or the end of a trajectory use are loaded
// where they are used instead of holding scalar registers through the backward walk.
typedef const __attribute__((address_space(4))) AdjointParams* KAdj;
__device__ __forceinline__ KAdj adjoint_params() { return (KAdj)__builtin_amdgcn_kernarg_segment_ptr(); }
__device__ __forceinline__ KAdj cold(KAdj k) {
    asm volatile("" : "+s"(k));
    return k;
}

// The bilinear gather of the staggered grid f (rows of `cols`) at (s, t) with its partial
// derivatives in s and t and its cell and weights (the same clamped cell as the forward).
struct Gathered {
    double value, ds, dt;
    unsigned c;
    int i, j;
    double u, w;
};

// The first half of gather_d: the cell, its weights and the loads of its two row pairs
// (issued here; their values are first used by gather_finish).
struct GatherLoads {
    Gathered g;
    Pair lo, hi;
};

__device__ __forceinline__ GatherLoads gather_issue(Rsrc f, i64 cols, double s, double t, int imax, int jmax) {
    GatherLoads L;
    Gathered& g = L.g;
    g.i = clamped_floor_index<true>(s, imax);
    g.j = clamped_floor_index<true>(t, jmax);
    g.u = s - (double)g.i;
    g.w = t - (double)g.j;
    g.c = mad24((unsigned)g.i, (unsigned)cols, (unsigned)g.j);
    L.lo = load_pair(f, g.c);
    L.hi = load_pair(f, g.c + (unsigned)cols);
    return L;
}

__device__ __forceinline__ Gathered gather_finish(const GatherLoads& L) {
    Gathered g = L.g;
    const double f00 = L.lo.a, f01 = L.lo.b, f10 = L.hi.a, f11 = L.hi.b;
    // (clamped_bilinear_buf's arithmetic)
    const double low = __builtin_fma(g.w, f01 - f00, f00);
    const double high = __builtin_fma(g.w, f11 - f10, f10);
    g.value = __builtin_fma(g.u, high - low, low);
    g.ds = high - low;
    g.dt = (1.0 - g.u) * (f01 - f00) + g.u * (f11 - f10);
    return g;
}

__device__ __forceinline__ Gathered gather_d(Rsrc f, i64 cols, double s, double t, int imax, int jmax) {
    return gather_finish(gather_issue(f, cols, s, t, imax, jmax));
}

// The start state of the next backward step, read from the segment scratch while the step
// before it runs: `src` is its first component (the others `lanes` apart).
// Its loads are issued after the last field loads of the current step, so that no wait for
// a field value also waits for them (vector loads complete in the order they are issued).
struct StatePrefetch {
    const double* src;
    i64 lanes;
    Vec3 x, v;
};

__device__ __forceinline__ void prefetch_state(StatePrefetch& pf) {
    const double* s = pf.src;
    const i64 n = pf.lanes;
    pf.x = {s[0], s[n], s[2 * n]};
    pf.v = {s[3 * n], s[4 * n], s[5 * n]};
}

// No memory operation between a field gather's loads and their use.
struct NoHook {
    __device__ __forceinline__ void operator()() const {}
};

// q/m E at point q: the trace kernel's edge_field for one patch whose origin and spacing are
// the mesh's, then the radial projection (bt_trace.hip, same operations).  `hook` issues
// memory operations right after the two field loads (an empty hook: none), where no wait for
// the fields waits for them.
template <int GEOM, typename Hook = NoHook>
__device__ __forceinline__ Vec3 accel(const AdjointParams& p, const AdjFields& F, double qm, Vec3 q,
                                      Hook hook = Hook(), const Radial* known = nullptr) {
    double cy = 0.0, cz = 0.0, qy = q.b;
    if (GEOM == 1) {
        const Radial rd = known != nullptr ? *known : radial(q.b, q.c);
        qy = rd.r; cy = rd.cy; cz = rd.cz;
    } else if (GEOM == 2) {
        cy = 1.0;
        if (q.b < p.origin_y) cy = -1.0;
        qy = p.origin_y + fabs(q.b - p.origin_y);
    }
    const double sx = axis_coord(q.a, p.origin_x, p.inv_h);
    const double sy = axis_coord(qy, p.origin_y, p.inv_h);
    // (gather_issue / gather_finish: clamped_bilinear_buf's operations)
    const GatherLoads lx = gather_issue(F.fx, p.ny, sx + 0.5, sy, (int)p.nx - 1, (int)p.ny - 2);
    const GatherLoads ly = gather_issue(F.fy, p.ny + 1, sx, sy + 0.5, (int)p.nx - 2, (int)p.ny - 1);
    if (!__is_same(Hook, NoHook)) {
        __builtin_amdgcn_sched_barrier(0);
        hook();
        __builtin_amdgcn_sched_barrier(0);
    }
    const double ex = gather_finish(lx).value;
    const double ey = gather_finish(ly).value;
    if (GEOM != 0) return {qm * ex, (qm * ey) * cy, (qm * ey) * cz};
    return {qm * ex, qm * ey, qm * 0.0};
}

// ---- the wave's gradient window ------------------------------------------------------
// Each wave sums its lanes' dJ/dfx and dJ/dfy contributions in LDS for kWinCols columns x
// kWinRows rows of both grids (column i in slot i mod kWinCols) that follow its lanes'
// cells; a column leaving the window is added to global memory once.  The lanes of a wave
// trace neighbouring trajectories backward in step, so most contributions stay in LDS.
constexpr int kWinCols = 16;
constexpr int kWinRows = 64;

// per wave of a block of at most 256 threads: [0] dJ/dfx, [1] dJ/dfy sums
__shared__ double adjoint_windows[4][2][kWinCols * kWinRows];

struct GradWindow {
    int wave;    // this wave's windows in adjoint_windows
    int lo;      // first column held, -1: none
    int base;    // first row held
};

// Wave-wide min / max with every lane of the wave active: DPP row operations (quad swaps,
// half-row and row mirrors, then row broadcasts into rows 1, 3 and 2, 3) leave the result
// in lane 63, read back as a uniform value.
template <bool MAX>
__device__ __forceinline__ int wave_reduce(int v) {
    const auto op = [](int a, int b) { return MAX ? max(a, b) : min(a, b); };
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));   // quad_perm [1, 0, 3, 2]
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));   // quad_perm [2, 3, 0, 1]
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xF, 0xF, false));  // row_half_mirror
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xF, 0xF, false));  // row_mirror
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x142, 0xA, 0xF, false));  // row_bcast:15
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x143, 0xC, 0xF, false));  // row_bcast:31
    return __builtin_amdgcn_readlane(v, 63);
}

__device__ __forceinline__ int wave_min(int v) { return wave_reduce<false>(v); }

__device__ __forceinline__ int wave_max(int v) { return wave_reduce<true>(v); }

// The same over two 16-bit halves at once (v_pk_min_i16 / v_pk_max_i16): the minima or the
// maxima of two quantities below 2^15 in one reduction.
typedef short Short2 __attribute__((ext_vector_type(2)));

template <bool MAX>
__device__ __forceinline__ int wave_reduce_pk(int v) {
    const auto op = [](int a, int b) {
        const Short2 x = __builtin_bit_cast(Short2, a), y = __builtin_bit_cast(Short2, b);
        return __builtin_bit_cast(int, MAX ? __builtin_elementwise_max(x, y) : __builtin_elementwise_min(x, y));
    };
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xF, 0xF, false));
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xF, 0xF, false));
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x142, 0xA, 0xF, false));
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x143, 0xC, 0xF, false));
    return __builtin_amdgcn_readlane(v, 63);
}

template <int A>
__device__ __forceinline__ void add_node(KAdj kp, i64 cols, const GradWindow& W, int ii, int jj, double v) {
    if (W.lo >= 0 && ii >= W.lo && ii < W.lo + kWinCols && jj >= W.base && jj < W.base + kWinRows) {
        atomicAdd(&adjoint_windows[W.wave][A][(ii & (kWinCols - 1)) * kWinRows + (jj - W.base)], v);
        return;
    }
    const KAdj c = cold(kp);
    unsafeAtomicAdd((A == 0 ? c->dfx : c->dfy) + (i64)ii * cols + jj, v);
}

// A lane's running sums of lambda_E times the bilinear weights of the gathered points in one
// cell of grid A (0: fx, 1: fy; rows of `cols`), kept as the four moments sum(lam),
// sum(lam u), sum(lam w), sum(lam u w) of the contributions: a contribution in the same cell
// is four fused multiply-adds, and the four node values -- (1-u)(1-w), (1-u)w, u(1-w), uw
// weighted -- are formed once, when the lane's points leave the cell or its trajectory ends,
// and go to the wave's window or global memory.  Consecutive stages and steps mostly stay in
// one cell.
constexpr int kNoCell = -(1 << 30);

// Polls of a group's finished-piece count before a piece gives up waiting (each a round trip
// to the coherent level and a short sleep: far longer than any piece takes).
constexpr int kHandoffSpins = 1 << 19;

struct GradCache {
    int c, r;  // the cell (kNoCell: none)
    double s, su, sw, suw;
};

__device__ __forceinline__ void cache_reset(GradCache& g) {
    g.c = kNoCell;
    g.r = kNoCell;
    g.s = g.su = g.sw = g.suw = 0.0;
}

// The cell's four node values to the window or global memory; the cache empties.
template <int A>
__device__ __forceinline__ void cache_flush(KAdj kp, i64 cols, const GradWindow& W, GradCache& g) {
    const int c = g.c, r = g.r;
    if (c != kNoCell) {
        const double n11 = g.suw, n10 = g.su - g.suw, n01 = g.sw - g.suw;
        const double n00 = (g.s - g.su) - n01;
        if (n00 != 0.0) add_node<A>(kp, cols, W, c, r, n00);
        if (n01 != 0.0) add_node<A>(kp, cols, W, c, r + 1, n01);
        if (n10 != 0.0) add_node<A>(kp, cols, W, c + 1, r, n10);
        if (n11 != 0.0) add_node<A>(kp, cols, W, c + 1, r + 1, n11);
    }
    cache_reset(g);
}

// lambda_E times the bilinear weights of the gathered cell into the lane's moments of grid A.
template <int A>
__device__ __forceinline__ void scatter(KAdj kp, i64 cols, const GradWindow& W, GradCache& gc,
                                        const Gathered& g, double lam) {
    if (lam == 0.0) return;
    if (__builtin_expect(g.i != gc.c || g.j != gc.r, 0)) {
        cache_flush<A>(kp, cols, W, gc);
        gc.c = g.i;
        gc.r = g.j;
    }
    const double lu = lam * g.u;
    gc.s += lam;
    gc.su += lu;
    gc.sw = __builtin_fma(lam, g.w, gc.sw);
    gc.suw = __builtin_fma(lu, g.w, gc.suw);
}

// Column `col` of the window (slot s) to global memory; lane r takes row base + r.
__device__ __forceinline__ void flush_column(KAdj kp, const GradWindow& W, int s, int col) {
    const KAdj p = cold(kp);
    const int r = (int)__lane_id();
    const int row = W.base + r;
    const double vx = __longlong_as_double((long long)atomicExch(
        reinterpret_cast<unsigned long long*>(&adjoint_windows[W.wave][0][s * kWinRows + r]), 0ull));
    const double vy = __longlong_as_double((long long)atomicExch(
        reinterpret_cast<unsigned long long*>(&adjoint_windows[W.wave][1][s * kWinRows + r]), 0ull));
    const i64 nx = p->nx, ny = p->ny;
    if (vx != 0.0 && col <= nx && row < ny) unsafeAtomicAdd(p->dfx + (i64)col * ny + row, vx);
    if (vy != 0.0 && col < nx && row <= ny) unsafeAtomicAdd(p->dfy + (i64)col * (ny + 1) + row, vy);
}

__device__ __forceinline__ void window_flush_all(KAdj kp, GradWindow& W) {
    if (W.lo < 0) return;
    for (int s = 0; s < kWinCols; ++s) flush_column(kp, W, s, W.lo + ((s - W.lo) & (kWinCols - 1)));
    W.lo = -1;
}

// Moves the window (wave-uniform; every lane of the wave calls it) to hold the cells
// (ci, cj) of the active lanes with room below them, the direction of the backward walk.
__device__ __forceinline__ void window_follow(const AdjointParams& p, KAdj kp, GradWindow& W, bool act, int ci,
                                              int cj) {
    // the window stays when every active lane's cell lies inside it with the margins tested
    // below (two columns on both sides, a row below and two above): one compare per lane
    // instead of the wave-wide minima and maxima, which only a move needs
    const bool inside = !act || ((unsigned)(ci - 2 - W.lo) <= (unsigned)(kWinCols - 5) &&
                                 (unsigned)(cj - 1 - W.base) <= (unsigned)(kWinRows - 4));
    if (W.lo >= 0 && __ballot(!inside) == 0) return;
    int imin, imax, jmin, jmax;
    if (p.nx < 32768 && p.ny < 32768) {  // cells below 2^15 on both axes: (ci, cj) in one word
        const int both = (ci & 0xffff) | (cj << 16);
        const int mn = wave_reduce_pk<false>(act ? both : 0x7fff7fff);
        const int mx = wave_reduce_pk<true>(act ? both : -1);  // -1: (-1, -1)
        imin = (short)(mn & 0xffff); jmin = mn >> 16;
        imax = (short)(mx & 0xffff); jmax = mx >> 16;
    } else {
        imin = wave_min(act ? ci : 0x7fffffff); imax = wave_max(act ? ci : -1);
        jmin = wave_min(act ? cj : 0x7fffffff); jmax = wave_max(act ? cj : -1);
    }
    if (imax < 0) return;
    const bool rows_ok = W.lo >= 0 && jmin - 1 >= W.base && jmax + 2 < W.base + kWinRows;
    const bool cols_ok = W.lo >= 0 && imin - 2 >= W.lo && imax + 2 <= W.lo + kWinCols - 1;
    if (rows_ok && cols_ok) return;
    int nlo = imax + 2 - (kWinCols - 1);
    if (nlo > imin - 2) nlo = imin - 2;
    if (nlo < 0) nlo = 0;
    const int nbase = rows_ok ? W.base : (jmin - 2 > 0 ? jmin - 2 : 0);
    // (lanes spread over more rows than the margins allow leave the rows where they are: then
    // the columns that stay keep their sums, and a window already in place moves nothing)
    const bool same_rows = nbase == W.base;
    if (nlo == W.lo && same_rows) return;
    if (W.lo >= 0) {
        for (int s = 0; s < kWinCols; ++s) {
            const int col = W.lo + ((s - W.lo) & (kWinCols - 1));
            if (!(same_rows && col >= nlo && col < nlo + kWinCols)) flush_column(kp, W, s, col);
        }
    }
    // (wave-uniform: kept in scalar registers)
    W.lo = __builtin_amdgcn_readfirstlane(nlo);
    W.base = __builtin_amdgcn_readfirstlane(nbase);
}

// The transposed field evaluation: given lambda_g for g = q/m E(q), scatter into dJ/dfx,
// dJ/dfy and return lambda_q; `g` receives g itself (accel's value, bit for bit).  With `pf`
// (the last evaluation of a backward step), the next step's start state is requested right
// after this evaluation's field loads.
template <int GEOM>
__device__ __forceinline__ Vec3 accel_adjoint(const AdjointParams& p, const AdjFields& F, const GradWindow& W,
                                              GradCache& fx_sums, GradCache& fy_sums, double qm, Vec3 q,
                                              Vec3 lg, Vec3& g, StatePrefetch* pf = nullptr,
                                              const Radial* known = nullptr) {
    const KAdj kp = adjoint_params();
    double cy = 0.0, cz = 0.0, qy = q.b, r = 0.0, ir = 0.0;
    if (GEOM == 1) {
        const Radial rd = known != nullptr ? *known : radial(q.b, q.c);
        r = rd.r; qy = r; cy = rd.cy; cz = rd.cz; ir = rd.ir;
    } else if (GEOM == 2) {
        cy = 1.0;
        if (q.b < p.origin_y) cy = -1.0;
        qy = p.origin_y + fabs(q.b - p.origin_y);
    }
    const double sx = axis_coord(q.a, p.origin_x, p.inv_h);
    const double sy = axis_coord(qy, p.origin_y, p.inv_h);
    const GatherLoads lx = gather_issue(F.fx, p.ny, sx + 0.5, sy, (int)p.nx - 1, (int)p.ny - 2);
    const GatherLoads ly = gather_issue(F.fy, p.ny + 1, sx, sy + 0.5, (int)p.nx - 2, (int)p.ny - 1);
    if (pf != nullptr) {
        // (the barriers keep the scheduler from moving these loads above the field loads)
        __builtin_amdgcn_sched_barrier(0);
        prefetch_state(*pf);
        __builtin_amdgcn_sched_barrier(0);
    }
    const Gathered gx = gather_finish(lx);
    const Gathered gy = gather_finish(ly);
    if (GEOM != 0) g = {qm * gx.value, (qm * gy.value) * cy, (qm * gy.value) * cz};
    else g = {qm * gx.value, qm * gy.value, qm * 0.0};
    const double lex = qm * lg.a;
    double ley, lcy = 0.0, lcz = 0.0;
    if (GEOM == 1) {
        ley = qm * (lg.b * cy + lg.c * cz);
        lcy = qm * lg.b * gy.value;
        lcz = qm * lg.c * gy.value;
    } else if (GEOM == 2) {
        ley = qm * lg.b * cy;
    } else {
        ley = qm * lg.b;
    }
    scatter<0>(kp, p.ny, W, fx_sums, gx, lex);
    scatter<1>(kp, p.ny + 1, W, fy_sums, gy, ley);
    const double lsx = lex * gx.ds + ley * gy.ds;
    const double lsy = lex * gx.dt + ley * gy.dt;
    Vec3 lq = {lsx * p.inv_h, 0.0, 0.0};
    const double lqy = lsy * p.inv_h;
    if (GEOM == 1) {
        if (r > 0.0) {
            // r = |(y, z)|, cy = y / r, cz = z / r
            lq.b = lqy * cy + (lcy * cz * cz - lcz * cy * cz) * ir;
            lq.c = lqy * cz + (lcz * cy * cy - lcy * cy * cz) * ir;
        }
    } else if (GEOM == 2) {
        lq.b = lqy * cy;
    } else {
        lq.b = lqy;
    }
    return lq;
}

// dt of the step rule and its branch: 1 = h / |v|, 2 = sqrt(2 h / |a|), 0 = h (the trace
// kernel's step_rule on the same operands).
__device__ __forceinline__ double step_size(const AdjointParams& p, Vec3 v, Vec3 a, int& branch, double& speed,
                                            double& acc) {
    const StepRule r = step_rule(norm2(v.a, v.b, v.c), norm2(a.a, a.b, a.c), p.h,
                                 p.step_fraction, 1.0);
    branch = r.branch;
    speed = r.speed;
    acc = r.accel;
    return r.dt;
}

// One forward step (the trace kernel's arithmetic); `hook` runs after the loads of the step's
// last field gather (accel's hook).
template <int INTEG, int GEOM, typename Hook = NoHook>
__device__ __forceinline__ void forward_step(const AdjointParams& p, const AdjFields& F, double qm, Vec3& x,
                                             Vec3& v, Hook hook = Hook()) {
    const Vec3 a = accel<GEOM>(p, F, qm, x);
    int branch;
    double speed, acc;
    const double dt = step_size(p, v, a, branch, speed, acc);
    const double hdt = 0.5 * dt;
    if (INTEG == 0) {
        const double dt6 = dt * (1.0 / 6.0);
        Vec3 sx = v, sv = a;
        Vec3 q = vaxpy(x, hdt, v), k = vaxpy(v, hdt, a);
        Vec3 g = accel<GEOM>(p, F, qm, q);
        sx = vaxpy(sx, 2.0, k); sv = vaxpy(sv, 2.0, g);
        q = vaxpy(x, hdt, k); k = vaxpy(v, hdt, g);
        g = accel<GEOM>(p, F, qm, q);
        sx = vaxpy(sx, 2.0, k); sv = vaxpy(sv, 2.0, g);
        q = vaxpy(x, dt, k); k = vaxpy(v, dt, g);
        g = accel<GEOM>(p, F, qm, q, hook);
        sx = vadd(sx, k); sv = vadd(sv, g);
        x = vaxpy(x, dt6, sx);
        v = vaxpy(v, dt6, sv);
    } else {
        const Vec3 k = vaxpy(v, hdt, a);
        const Vec3 q = vaxpy(x, dt, k);
        const Vec3 g = accel<GEOM>(p, F, qm, q, hook);
        x = q;
        v = vaxpy(k, hdt, g);
    }
}

// The seeds of the records of step m (1-based) of trajectory k: the crossing's adjoint
// split into the step's start (ls_x, ls_v) and end (le_x, le_v) terms.
__device__ __forceinline__ void crossing_seeds(const AdjointParams& p, i64 k, i64 m, Vec3 x, Vec3 v, Vec3 y, Vec3 w,
                                               Vec3& lsx, Vec3& lsv, Vec3& lex, Vec3& lev) {
    const KAdj kc = cold(adjoint_params());
    for (int pl = 0; pl < kc->n_planes; ++pl) {
        const i64 idx = k * kc->n_planes + pl;
        i64 cnt = kc->counts[idx];
        if (cnt > kc->capacity) cnt = kc->capacity;
        for (i64 c = 0; c < cnt; ++c) {
            const double* rec = kc->hits + (idx * kc->capacity + c) * 8;
            if ((i64)rec[7] != m) continue;
            const double* sd = kc->seeds + (idx * kc->capacity + c) * 5;
            const double xp = kc->planes[pl];
            const double before = x.a - xp;
            const double after = y.a - xp;
            const double sfrac = before / (before - after);
            const double ls = sd[0] * (y.b - x.b) + sd[1] * (y.c - x.c) + sd[2] * (w.a - v.a) + sd[3] * (w.b - v.b) +
                              sd[4] * (w.c - v.c);
            const double den = before - after;
            lsx.a += ls * (-after) / (den * den);
            lex.a += ls * before / (den * den);
            lsx.b += sd[0] * (1.0 - sfrac); lsx.c += sd[1] * (1.0 - sfrac);
            lex.b += sd[0] * sfrac; lex.c += sd[1] * sfrac;
            lsv.a += sd[2] * (1.0 - sfrac); lsv.b += sd[3] * (1.0 - sfrac); lsv.c += sd[4] * (1.0 - sfrac);
            lev.a += sd[2] * sfrac; lev.b += sd[3] * sfrac; lev.c += sd[4] * sfrac;
        }
    }
}

// The largest seeded step below `below` (0 if none).
__device__ __forceinline__ i64 seeded_step(const AdjointParams& p, i64 k, i64 below) {
    const KAdj kc = cold(adjoint_params());
    i64 best = 0;
    for (int pl = 0; pl < kc->n_planes; ++pl) {
        const i64 idx = k * kc->n_planes + pl;
        i64 cnt = kc->counts[idx];
        if (cnt > kc->capacity) cnt = kc->capacity;
        for (i64 c = 0; c < cnt; ++c) {
            const double* sd = kc->seeds + (idx * kc->capacity + c) * 5;
            if (sd[0] == 0.0 && sd[1] == 0.0 && sd[2] == 0.0 && sd[3] == 0.0 && sd[4] == 0.0) continue;
            const i64 m = (i64)kc->hits[(idx * kc->capacity + c) * 8 + 7];
            if (m < below && m > best) best = m;
        }
    }
    return best;
}

// The transposed step m (start state x, v; `rx`: radial(x.b, x.c) when cylindrical): lambda of
// the end (lx, lv) -> of the start.
template <int INTEG, int GEOM>
__device__ __forceinline__ void reverse_step(const AdjointParams& p, const AdjFields& F, const GradWindow& W,
                                             GradCache& fx_sums, GradCache& fy_sums, i64 k,
                                             i64 m, bool seeded, double qm, Vec3 x, Vec3 v, const Radial& rx,
                                             Vec3& lx, Vec3& lv, StatePrefetch& pf) {
    const Vec3 a = accel<GEOM>(p, F, qm, x, NoHook(), &rx);
    int branch;
    double speed, acc;
    const double dt = step_size(p, v, a, branch, speed, acc);
    const double hdt = 0.5 * dt;
    Vec3 La = {0.0, 0.0, 0.0}, Lx, Lv;
    double Ldt = 0.0, Lhdt = 0.0;
    if (INTEG == 0) {
        const double dt6 = dt * (1.0 / 6.0);
        // The stages' points and the first stage's field; the fields at q2 and q3 come from their
        // transposed evaluations below (bit for bit accel's), and the terms that need them are
        // added once they are known.  s12 = v + 2 k1 + 2 k2 and a12 = a + 2 g1 are the stage
        // sums before k3 and g2, g3 (the same operations as the full sums).
        const Vec3 q1 = vaxpy(x, hdt, v), k1 = vaxpy(v, hdt, a);
        const Vec3 g1 = accel<GEOM>(p, F, qm, q1);
        const Vec3 q2 = vaxpy(x, hdt, k1), k2 = vaxpy(v, hdt, g1);
        const Vec3 q3 = vaxpy(x, dt, k2);
        const Vec3 s12 = vaxpy(vaxpy(v, 2.0, k1), 2.0, k2);
        const Vec3 a12 = vaxpy(a, 2.0, g1);
        Vec3 xs = {0.0, 0.0, 0.0}, vs = {0.0, 0.0, 0.0};
        if (__builtin_expect(seeded, 0)) {  // the crossing's seeds need the step's end first
            const Vec3 g2 = accel<GEOM>(p, F, qm, q2), g3 = accel<GEOM>(p, F, qm, q3);
            const Vec3 sx = vadd(s12, vaxpy(v, dt, g2)), sv = vaxpy(a12, 2.0, g2);
            const Vec3 y = vaxpy(x, dt6, sx), w = vaxpy(v, dt6, vadd(sv, g3));
            crossing_seeds(p, k, m, x, v, y, w, xs, vs, lx, lv);
        }
        Lx = lx;
        Lv = lv;
        // the step's derivative in dt6, lx.sx + lv.sv: lx.k3 + 2 lv.g2 and lv.g3 are added once
        // g2 and g3 are known
        double Ldt6 = vdot(lx, s12) + vdot(lv, a12);
        const Vec3 Lsx = vscale(dt6, lx), Lsv = vscale(dt6, lv);
        Lv = vadd(Lv, Lsx);
        Vec3 Lk1 = vscale(2.0, Lsx), Lk2 = vscale(2.0, Lsx);
        const Vec3 Lk3 = Lsx;
        La = Lsv;
        Vec3 Lg1 = vscale(2.0, Lsv), Lg2 = vscale(2.0, Lsv);
        const Vec3 Lg3 = Lsv;
        Vec3 g;
        const Vec3 Lq3 = accel_adjoint<GEOM>(p, F, W, fx_sums, fy_sums, qm, q3, Lg3, g);
        Ldt6 += vdot(lv, g);  // g3
        Lv = vadd(Lv, Lk3); Lg2 = vaxpy(Lg2, dt, Lk3);
        Lx = vadd(Lx, Lq3); Ldt += vdot(Lq3, k2); Lk2 = vaxpy(Lk2, dt, Lq3);
        const Vec3 Lq2 = accel_adjoint<GEOM>(p, F, W, fx_sums, fy_sums, qm, q2, Lg2, g);
        Ldt6 += vdot(lx, vaxpy(v, dt, g)) + 2.0 * vdot(lv, g);  // k3 = v + dt g2, and g2
        Ldt += vdot(Lk3, g);  // g2
        Lv = vadd(Lv, Lk2); Lg1 = vaxpy(Lg1, hdt, Lk2);
        Lx = vadd(Lx, Lq2); Lhdt += vdot(Lq2, k1); Lk1 = vaxpy(Lk1, hdt, Lq2);
        const Vec3 Lq1 = accel_adjoint<GEOM>(p, F, W, fx_sums, fy_sums, qm, q1, Lg1, g);
        Lhdt += vdot(Lk2, g1);
        Lv = vadd(Lv, Lk1); Lhdt += vdot(Lk1, a); La = vaxpy(La, hdt, Lk1);
        Lx = vadd(Lx, Lq1); Lhdt += vdot(Lq1, v); Lv = vaxpy(Lv, hdt, Lq1);
        Ldt += Ldt6 * (1.0 / 6.0);
        Lx = vadd(Lx, xs);
        Lv = vadd(Lv, vs);
    } else {
        const Vec3 kk = vaxpy(v, hdt, a);
        const Vec3 q = vaxpy(x, dt, kk);
        const Vec3 g1 = accel<GEOM>(p, F, qm, q);
        Vec3 xs = {0.0, 0.0, 0.0}, vs = {0.0, 0.0, 0.0};
        if (seeded) {
            const Vec3 w = vaxpy(kk, hdt, g1);
            crossing_seeds(p, k, m, x, v, q, w, xs, vs, lx, lv);
        }
        Vec3 Lk = lv;
        Lhdt += vdot(lv, g1);
        const Vec3 Lg1 = vscale(hdt, lv);
        Vec3 gq;
        const Vec3 Lq = vadd(lx, accel_adjoint<GEOM>(p, F, W, fx_sums, fy_sums, qm, q, Lg1, gq));
        Lx = Lq;
        Ldt += vdot(Lq, kk);
        Lk = vaxpy(Lk, dt, Lq);
        Lv = Lk;
        Lhdt += vdot(Lk, a);
        La = vscale(hdt, Lk);
        Lx = vadd(Lx, xs);
        Lv = vadd(Lv, vs);
    }
    Ldt += 0.5 * Lhdt;
    if (branch == 1) {
        Lv = vaxpy(Lv, -Ldt * dt / (speed * speed), v);
    } else if (branch == 2) {
        La = vaxpy(La, -0.5 * Ldt * dt / (acc * acc), a);
    }
    Vec3 ga;
    Lx = vadd(Lx, accel_adjoint<GEOM>(p, F, W, fx_sums, fy_sums, qm, x, La, ga, &pf, &rx));
    lx = Lx;
    lv = Lv;
}

// scratch: checkpoints [n_ckpt][6][lanes], then segment states [interval][6][lanes]
__device__ __forceinline__ void put_state(double* __restrict__ base, i64 lanes, i64 slot, i64 lane, Vec3 x, Vec3 v) {
    double* s = base + slot * 6 * lanes + lane;
    s[0] = x.a; s[lanes] = x.b; s[2 * lanes] = x.c;
    s[3 * lanes] = v.a; s[4 * lanes] = v.b; s[5 * lanes] = v.c;
}

__device__ __forceinline__ void get_state(const double* __restrict__ base, i64 lanes, i64 slot, i64 lane, Vec3& x,
                                          Vec3& v) {
    const double* s = base + slot * 6 * lanes + lane;
    x = {s[0], s[lanes], s[2 * lanes]};
    v = {s[3 * lanes], s[4 * lanes], s[5 * lanes]};
}

// The mesh cell of point x (the window's key), clamped to the mesh (rx: radial(x.b, x.c) when
// cylindrical).
template <int GEOM>
__device__ __forceinline__ void cell_of(const AdjointParams& p, Vec3 x, const Radial& rx, int& ci, int& cj) {
    double qy = x.b;
    if (GEOM == 1) qy = rx.r;
    else if (GEOM == 2) qy = p.origin_y + fabs(x.b - p.origin_y);
    // (floor clamped to [0, n - 1] as the saturated conversion and one median: the same cell)
    ci = clamped_floor_index<true>(axis_coord(x.a, p.origin_x, p.inv_h), (int)p.nx - 1);
    cj = clamped_floor_index<true>(axis_coord(qy, p.origin_y, p.inv_h), (int)p.ny - 1);
}

template <int INTEG, int GEOM>
__global__ __launch_bounds__(256, 2) void adjoint_kernel(AdjointParams p) {
    const unsigned lane = __lane_id();
    const u64 below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const AdjFields F = {field_rsrc(p.fxs, (p.nx + 1) * p.ny), field_rsrc(p.fys, p.nx * (p.ny + 1))};
    GradWindow W;
    W.wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    W.lo = -1;
    W.base = 0;
    GradCache fx_sums, fy_sums;  // the lane's node sums of dJ/dfx and dJ/dfy
    cache_reset(fx_sums);
    cache_reset(fy_sums);
    for (int t = (int)lane; t < kWinCols * kWinRows; t += 64) {
        adjoint_windows[W.wave][0][t] = 0.0;
        adjoint_windows[W.wave][1][t] = 0.0;
    }
    // (per-lane indices and step counts are 32-bit: the entry point bounds the trajectories,
    // the lanes and the steps below 2^31, and the loop keeps fewer registers)
    const int gl = (int)(blockIdx.x * blockDim.x + threadIdx.x);  // this lane's scratch column
    double* ck = p.scratch;
    double* seg = p.scratch + (i64)p.n_ckpt * 6 * p.lanes;
    const i64 K = p.interval;
    const int P = p.pieces;
    bool exhausted = false;
    for (;;) {
        const KAdj c = cold(adjoint_params());
        int k = -1;
        int piece = 0;  // of the task: segments of piece 0 lie above those of piece 1, ...
        i64 g = 0;      // the task's group (pieces > 1)
        if (P == 1) {
            // one task per trajectory: each lane that needs one claims the next queue entry
            const bool want = !exhausted;
            const u64 need = __ballot(want);
            if (need != 0) {
                const int count = __popcll(need);
                const int leader = __ffsll((long long)need) - 1;
                u64 base = 0;
                if ((int)lane == leader) base = atomicAdd(c->queue, (u64)count);
                base = __shfl(base, leader);
                if (want) {
                    const u64 mm = base + (u64)__popcll(need & below);
                    if (mm < (u64)c->n_active) {
                        const i64 kk = c->active[mm];
                        if (kk >= 0 && kk < c->n) k = (int)kk;
                    } else {
                        exhausted = true;
                    }
                }
            }
            if (__ballot(k >= 0 || !exhausted) == 0) break;
        } else {
            // one task per piece of a group of 64 consecutive entries, every group's first piece
            // before any second piece, and so on (wave-uniform)
            u64 t = 0;
            if (lane == 0) t = atomicAdd(c->queue, 1ull);
            t = __shfl(t, 0);
            if (t >= (u64)P * (u64)c->groups) break;
            piece = (int)(t / (u64)c->groups);
            g = (i64)(t - (u64)piece * (u64)c->groups);
            const i64 mm = g * 64 + (i64)lane;
            if (mm < c->n_active) {
                const i64 kk = c->active[mm];
                if (kk >= 0 && kk < c->n) k = (int)kk;
            }
        }
        const i64 last_seeded = k >= 0 ? seeded_step(p, k, 0x7fffffffffffffffll) : 0;
        // a trajectory with no seed, or seeded beyond the checkpoints, adds nothing; the
        // latter (a max_step below the seeded steps) is counted for the caller (once)
        bool valid = last_seeded > 0 && (last_seeded - 1) / K < p.n_ckpt;
        int last = valid ? (int)last_seeded : 0;
        if (__builtin_expect(last_seeded > 0 && !valid, 0) && piece == 0 && c->dropped != nullptr)
            __hip_atomic_fetch_add(c->dropped, (i64)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const i64 N = c->n;
        // the forward trace's checkpoints when they reach the last seeded step (pieces need them:
        // another lane continues the walk; the launch uses pieces only when they do)
        const bool stored = valid && c->fwd_ckpt != nullptr && (last - 1) / K < c->fwd_n_ckpt;
        if (__builtin_expect(P > 1 && valid && !stored, 0)) {
            if (piece == 0 && c->dropped != nullptr)
                __hip_atomic_fetch_add(c->dropped, (i64)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            valid = false;
        }
        if (!valid) last = 0;
        const double qm = valid ? c->qm[k] : 0.0;
        Vec3 x = {0.0, 0.0, 0.0}, v = {0.0, 0.0, 0.0};
        if (valid && !stored) {
            x = {c->x[k], c->x[N + k], c->x[2 * N + k]};
            v = {c->v[k], c->v[N + k], c->v[2 * N + k]};
            // 1. forward to the last seeded step, a checkpoint every K steps
            for (int m = 0; m < last; ++m) {
                if (m % K == 0) put_state(ck, p.lanes, m / K, gl, x, v);
                forward_step<INTEG, GEOM>(p, F, qm, x, v);
            }
        }
        // 2. the segments backward, in step across the wave (the window moves with them); a
        // piece takes its share of the group's segments (all of them with one piece)
        Vec3 lx = {0.0, 0.0, 0.0}, lv = {0.0, 0.0, 0.0};
        const int top = wave_max(valid ? (int)((last - 1) / K) : -1);
        const int nseg = top + 1;
        const int c_hi = top - (int)(((i64)piece * nseg) / P);
        const int c_lo = top + 1 - (int)(((i64)(piece + 1) * nseg) / P);
        // the largest seeded record step within or below this piece's segments
        int next = piece == 0 ? last : (valid ? (int)seeded_step(p, k, (i64)(c_hi + 1) * K + 1) : 0);
        if (piece > 0) {
            // the adjoint state at the end of this piece, from the group's piece above: wait
            // (bounded) until that piece is done, then read it through the coherent path
            const KAdj h = cold(adjoint_params());
            int* const flag = h->done + g;
            int seen = 0;
            for (int spin = 0; spin < kHandoffSpins; ++spin) {
                seen = __builtin_amdgcn_readfirstlane(
                    __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                if (seen >= piece) break;
                __builtin_amdgcn_s_sleep(8);
            }
            asm volatile("" ::: "memory");
            if (__builtin_expect(seen >= piece, 1)) {
                const i64 HL = h->groups * 64, at = g * 64 + (i64)lane;
                double* const hd = h->hand;
                lx = {__hip_atomic_load(hd + at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                      __hip_atomic_load(hd + HL + at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                      __hip_atomic_load(hd + 2 * HL + at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)};
                lv = {__hip_atomic_load(hd + 3 * HL + at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                      __hip_atomic_load(hd + 4 * HL + at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                      __hip_atomic_load(hd + 5 * HL + at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)};
            } else if (valid) {
                // (never in a working launch) the trajectory's gradient would miss the pieces
                // above: counted as dropped, and this piece adds nothing
                if (h->dropped != nullptr)
                    __hip_atomic_fetch_add(h->dropped, (i64)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                valid = false;
                last = 0;
                next = 0;
            }
        }
        for (int c = c_hi; c >= c_lo; --c) {
            const int m0 = c * (int)K;
            const bool has = valid && m0 < last;
            const int m1 = has ? (m0 + (int)K < last ? m0 + (int)K : last) : m0;
            if (has) {
                if (stored) {
                    const double* cp = cold(adjoint_params())->fwd_ckpt + (i64)c * 6 * N + k;
                    x = {cp[0], cp[N], cp[2 * N]};
                    v = {cp[3 * N], cp[4 * N], cp[5 * N]};
                } else {
                    get_state(ck, p.lanes, c, gl, x, v);
                }
                // each step's start state goes to the scratch once the step's last field loads
                // are issued (stores count with the loads: a store issued before a gather would
                // hold up the wait for its fields); the segment's last state stays in x, v, where
                // the backward walk starts
                const i64 lanes = p.lanes;
                for (int m = m0; m + 1 < m1; ++m) {
                    double* const dst = seg + (m - m0) * 6 * lanes + gl;
                    const Vec3 xs = x, vs = v;
                    forward_step<INTEG, GEOM>(p, F, qm, x, v, [=] {
                        dst[0] = xs.a; dst[lanes] = xs.b; dst[2 * lanes] = xs.c;
                        dst[3 * lanes] = vs.a; dst[4 * lanes] = vs.b; dst[5 * lanes] = vs.c;
                    });
                }
            }
            // x, v hold the start state of step m (the segment's last, then each step's
            // prefetched predecessor)
            for (i64 r = K - 1; r >= 0; --r) {
                const int m = m0 + (int)r;
                const bool act = has && m < m1;
                // (the step works on copies taken here, where the wait for the prefetch falls;
                // the prefetch then loads into x, v themselves, with no copy behind it)
                const Vec3 xc = x, vc = v;
                // (the start point's distance from the axis, once for the cell and the step's two
                // evaluations there)
                Radial rx = {0.0, 0.0, 0.0, 0.0};
                if (GEOM == 1) rx = radial(xc.b, xc.c);
                int ci = 0, cj = 0;
                if (act) cell_of<GEOM>(p, xc, rx, ci, cj);
                window_follow(p, adjoint_params(), W, act, ci, cj);
                if (act) {
                    const bool seeded = m + 1 == next;
                    // the prefetch always loads (slot 0 again at r = 0, unused): a load under a
                    // branch would make the following waits count it as outstanding
                    StatePrefetch pf;
                    pf.src = seg + (r > 0 ? r - 1 : 0) * 6 * p.lanes + gl;
                    pf.lanes = p.lanes;
                    reverse_step<INTEG, GEOM>(p, F, W, fx_sums, fy_sums, k, m + 1, seeded, qm, xc, vc, rx, lx, lv, pf);
                    x = pf.x;
                    v = pf.v;
                    if (seeded) next = (int)seeded_step(p, k, m + 1);
                }
            }
        }
        cache_flush<0>(adjoint_params(), p.ny, W, fx_sums);
        cache_flush<1>(adjoint_params(), p.ny + 1, W, fy_sums);
        if (piece + 1 < P) {
            // hand the adjoint state to the group's next piece: written through to the coherent
            // level, complete, and only then counted in done[g]
            const KAdj h = cold(adjoint_params());
            const i64 HL = h->groups * 64, at = g * 64 + (i64)lane;
            double* const hd = h->hand;
            __hip_atomic_store(hd + at, lx.a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(hd + HL + at, lx.b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(hd + 2 * HL + at, lx.c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(hd + 3 * HL + at, lv.a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(hd + 4 * HL + at, lv.b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(hd + 5 * HL + at, lv.c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            if (lane == 0) __hip_atomic_fetch_add(h->done + g, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            continue;
        }
        {
            const KAdj e = cold(adjoint_params());
            double* const dx = e->dx;
            double* const dv = e->dv;
            if (valid && dx != nullptr) {
                dx[k] += lx.a; dx[N + k] += lx.b; dx[2 * N + k] += lx.c;
            }
            if (valid && dv != nullptr) {
                dv[k] += lv.a; dv[N + k] += lv.b; dv[2 * N + k] += lv.c;
            }
        }
    }
    window_flush_all(adjoint_params(), W);
}

// The queue head and the groups' finished-piece counts to 0 (a kernel: no memset node in a
// captured graph).
__global__ void adjoint_prepare_kernel(u64* __restrict__ queue, int* __restrict__ done, i64 groups) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *queue = 0ull;
    const i64 stride = (i64)gridDim.x * blockDim.x;
    for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; done != nullptr && t < groups; t += stride) done[t] = 0;
}

typedef void (*AdjointKernel)(AdjointParams);

AdjointKernel select_adjoint_kernel(int integ, int geom) {
#define BT_SELA(I, G) \
    if (integ == I && geom == G) return adjoint_kernel<I, G>;
    BT_SELA(0, 0) BT_SELA(0, 1) BT_SELA(0, 2) BT_SELA(1, 0) BT_SELA(1, 1) BT_SELA(1, 2)
#undef BT_SELA
    return nullptr;
}

int adjoint_occupancy(int integ, int geom, int block) {
    AdjointKernel kern = select_adjoint_kernel(integ, geom);
    if (kern == nullptr) return 0;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, reinterpret_cast<const void*>(kern), block, 0) !=
        hipSuccess)
        return 0;
    return blocks;
}

int launch_adjoint(const AdjointParams& p, int block, int grid, hipStream_t stream, int integ, int geom) {
    AdjointKernel kern = select_adjoint_kernel(integ, geom);
    if (kern == nullptr) return -1;
    int pgrid = p.pieces > 1 ? (int)((p.groups + 255) / 256) : 1;
    if (pgrid > 1024) pgrid = 1024;
    hipLaunchKernelGGL(adjoint_prepare_kernel, dim3(pgrid), dim3(256), 0, stream, p.queue,
                       p.pieces > 1 ? p.done : nullptr, p.groups);
    if (p.n_active > 0 && grid > 0) hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, stream, p);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // namespace bt
