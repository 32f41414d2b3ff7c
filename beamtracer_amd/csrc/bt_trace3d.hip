// The 3D trajectory tracer for gfx950 (kernels/trace_trajectories_3d/trace_trajectories_3d_pyccel.py:18-657).
//
// Each lane traces one trajectory of the active list through the staggered edge fields of
// the finest patch containing each stage point (trilinear; E_x on (nx + 1, ny, nz), E_y on
// (nx, ny + 1, nz), E_z on (nx, ny, nz + 1)), an optional node magnetic field of the base
// mesh, symmetry planes at the lower y and z faces (the field of the mirror image, the
// normal E and the two other B components changing sign), RK4 or Boris's kick-rotate-kick
// (relativistic with inv_c2 != 0), until it leaves the mesh, meets an electrode sub-cell or
// reaches max_steps.  Every step deposits dq dt at its Hermite midpoint on the 8 nodes of
// its patch cell and, with has_j, the current element dq (x' - x) on the base mesh's nodes;
// both as 128-bit fixed-point words (bt_device.h: to_fixed / add_words), so the node sums do
// not depend on the order of the lanes.  A lane keeps its cell's node sums in registers and
// moves a node's sum to memory only when the node leaves its cell.  The step's arithmetic is the reference's in its
// order, with no fused multiply-add, so a trajectory's states, records and paths are
// bit-identical to the reference arithmetic; indices are clamped in floating point before
// they become integers, so a non-finite point addresses the mesh's first cell.
#include <hip/hip_runtime.h>

#include "bt_device.h"
#include "bt_internal.h"

namespace bt {

namespace {

// floor(s) clamped to [lo, hi] as an integer (NaN: lo).
__device__ __forceinline__ int floor_in(double s, int lo, int hi) {
    return (int)fmin(fmax(floor(s), (double)lo), (double)hi);
}

// (pyccel.py:237-241): f at ci and its +1, +dj, +di neighbours, trilinear.
__device__ __forceinline__ double trilinear(const double* __restrict__ f, i64 ci, i64 di, i64 dj, double a, double b,
                                           double c) {
    return (1.0 - a) * ((1.0 - b) * ((1.0 - c) * f[ci] + c * f[ci + 1])
                        + b * ((1.0 - c) * f[ci + dj] + c * f[ci + dj + 1]))
         + a * ((1.0 - b) * ((1.0 - c) * f[ci + di] + c * f[ci + di + 1])
                + b * ((1.0 - c) * f[ci + di + dj] + c * f[ci + di + dj + 1]));
}

// The finest patch containing the point, else the base mesh (pyccel.py:196-206).
__device__ __forceinline__ int patch_at(const Trace3DParams& p, double p0, double p1, double p2) {
    for (int qq = 0; qq < p.n_patches - 1; ++qq) {
        const double* r = p.patch_real + qq * 10;
        if (p0 >= r[4] && p0 <= r[7] && p1 >= r[5] && p1 <= r[8] && p2 >= r[6] && p2 <= r[9]) return qq;
    }
    return p.n_patches - 1;
}

// E and B at a stage point (pyccel.py:182-318).
__device__ void fields_at(const Trace3DParams& p, double q0, double q1, double q2, double e[3], double b[3]) {
    double p1 = q1, p2 = q2, s1 = 1.0, s2 = 1.0;
    if (p.mirror_y != 0 && p1 < p.oy) { p1 = 2.0 * p.oy - p1; s1 = -1.0; }
    if (p.mirror_z != 0 && p2 < p.oz) { p2 = 2.0 * p.oz - p2; s2 = -1.0; }
    const double p0 = q0;
    const int pp = patch_at(p, p0, p1, p2);
    const i64* pi = p.patch_int + pp * 8;
    const double* pr = p.patch_real + pp * 10;
    const int pnx = (int)pi[0], pny = (int)pi[1], pnz = (int)pi[2];
    const double phh = pr[3];
    const double sx = (p0 - pr[0]) / phh, sy = (p1 - pr[1]) / phh, sz = (p2 - pr[2]) / phh;
    // E_x at (sx + 1/2, sy, sz)
    double s = sx + 0.5;
    int i = floor_in(s, 0, pnx - 1), j = floor_in(sy, 0, pny - 2), kz = floor_in(sz, 0, pnz - 2);
    double wa = s - (double)i, wb = sy - (double)j;
    const double wc_x = sz - (double)kz;
    i64 dj = pnz, di = (i64)pny * pnz;
    e[0] = trilinear(p.fxs, pi[3] + i * di + j * dj + kz, di, dj, wa, wb, wc_x);
    // E_y at (sx, sy + 1/2, sz): the z index and weight of E_x
    i = floor_in(sx, 0, pnx - 2);
    s = sy + 0.5;
    j = floor_in(s, 0, pny - 1);
    wa = sx - (double)i;
    wb = s - (double)j;
    di = (i64)(pny + 1) * pnz;
    e[1] = s1 * trilinear(p.fys, pi[4] + i * di + j * dj + kz, di, dj, wa, wb, wc_x);
    // E_z at (sx, sy, sz + 1/2): the x index and weight of E_y
    j = floor_in(sy, 0, pny - 2);
    s = sz + 0.5;
    kz = floor_in(s, 0, pnz - 1);
    wb = sy - (double)j;
    const double wc = s - (double)kz;
    dj = pnz + 1;
    di = (i64)pny * (pnz + 1);
    e[2] = s2 * trilinear(p.fzs, pi[5] + i * di + j * dj + kz, di, dj, wa, wb, wc);
    b[0] = b[1] = b[2] = 0.0;
    if (p.has_b != 0) {  // B on the base mesh's nodes (3, nx, ny, nz)
        const double bx = (p0 - p.ox) / p.h, by = (p1 - p.oy) / p.h, bz = (p2 - p.oz) / p.h;
        const int bi = floor_in(bx, 0, p.nx - 2), bj = floor_in(by, 0, p.ny - 2), bk = floor_in(bz, 0, p.nz - 2);
        const double a = bx - (double)bi, bb = by - (double)bj, c = bz - (double)bk;
        const double w000 = (1.0 - a) * (1.0 - bb) * (1.0 - c), w001 = (1.0 - a) * (1.0 - bb) * c;
        const double w010 = (1.0 - a) * bb * (1.0 - c), w011 = (1.0 - a) * bb * c;
        const double w100 = a * (1.0 - bb) * (1.0 - c), w101 = a * (1.0 - bb) * c;
        const double w110 = a * bb * (1.0 - c), w111 = a * bb * c;
        const i64 nodes = (i64)p.nx * p.ny * p.nz, si = (i64)p.ny * p.nz, sj = p.nz;
        double f[3];
#pragma unroll
        for (int n0 = 0; n0 < 3; ++n0) {
            const double* g = p.bn + n0 * nodes + bi * si + bj * sj + bk;
            f[n0] = w000 * g[0] + w001 * g[1] + w010 * g[sj] + w011 * g[sj + 1] + w100 * g[si] + w101 * g[si + 1]
                  + w110 * g[si + sj] + w111 * g[si + sj + 1];
        }
        b[0] = s1 * s2 * f[0];
        b[1] = s2 * f[1];
        b[2] = s1 * f[2];
    }
}

// Boris's velocity update over a half step (tau = hdt / 2 * q/m): half kick, rotation with
// t scaled by 1 / gamma of v-minus, half kick (mpiparticles' relativistic_boris_update as
// trace_trajectories_3d_pyccel.py:344-352 calls it).
__device__ __forceinline__ void boris_update(double& u0, double& u1, double& u2, const double e[3], const double b[3],
                                             double hdt, double qm, double inv_c2) {
    const double hq = 0.5 * hdt * qm;
    const double m0 = u0 + hq * e[0], m1 = u1 + hq * e[1], m2 = u2 + hq * e[2];
    const double g = 1.0 / lorentz_factor(m0, m1, m2, inv_c2);
    const double t0 = hq * b[0] * g, t1 = hq * b[1] * g, t2 = hq * b[2] * g;
    const double f = 2.0 / (1.0 + (t0 * t0 + t1 * t1 + t2 * t2));
    const double r0 = m0 + (m1 * t2 - m2 * t1);
    const double r1 = m1 + (m2 * t0 - m0 * t2);
    const double r2 = m2 + (m0 * t1 - m1 * t0);
    const double c0 = f * t0, c1 = f * t1, c2 = f * t2;
    u0 = m0 + (r1 * c2 - r2 * c1) + hq * e[0];
    u1 = m1 + (r2 * c0 - r0 * c2) + hq * e[1];
    u2 = m2 + (r0 * c1 - r1 * c0) + hq * e[2];
}

// A deposit's node index and fraction along one axis of n nodes (pyccel.py:453-468); a
// non-finite coordinate gives the first node (its deposit is then dropped, not finite).
__device__ __forceinline__ void deposit_axis(double s, int n, int& i, double& w) {
    const double f = floor(s);
    w = s - f;
    if (f < 0.0) { w = 0.0; i = 0; }
    else if (f >= (double)(n - 1)) { w = 1.0; i = n - 2; }
    else if (f >= 0.0) i = (int)f;
    else i = 0;
}

// The node sums of a lane's current cell, NQ quantities: slot q holds the node whose x, y, z
// parities are the bits of q (so the nodes that two neighbouring cells share keep their
// slots), as integer-valued doubles in units of 2^e (exact below 2^53).  A deposit's shares
// are rounded once to that grid, as to_fixed rounds them, so the node totals equal a
// deposit of every share straight into the words; the cache only moves a node's sum to
// memory when the node leaves the lane's cell (or the trajectory ends).
template <int NQ>
struct NodeCache {
    int i, j, k;         // the cell (i < 0: empty)
    i64 origin, di, dj;  // its mesh's first node and strides in the word array
    i64 qstride;         // the node count between two quantities' grids
    double s[NQ][8];
};

template <int NQ>
__device__ __forceinline__ void cache_init(NodeCache<NQ>& c, i64 qstride) {
    c.i = c.j = c.k = -1;
    c.origin = c.di = c.dj = 0;
    c.qstride = qstride;
#pragma unroll
    for (int m = 0; m < NQ; ++m)
#pragma unroll
        for (int q = 0; q < 8; ++q) c.s[m][q] = 0.0;
}

// The node of parity `bit` among cell index a's two nodes a, a + 1.
__device__ __forceinline__ int node_of(int a, int bit) { return a + ((a ^ bit) & 1); }

__device__ __forceinline__ void flush_units(u64* __restrict__ acc, i64 node, double units) {
    if (units != 0.0) {
        const i64 q = (i64)units;
        add_words(acc + kWords * node, (u64)q & 0xffffffffull, (u64)(q >> 32));
    }
}

// Moves the cache to cell (i, j, k) of the mesh at `origin` with strides di, dj: the sums of
// the nodes that the new cell does not hold go to memory (all of them on another mesh).
template <int NQ>
__device__ __forceinline__ void cache_move(NodeCache<NQ>& c, u64* __restrict__ acc, int i, int j, int k, i64 origin,
                                          i64 di, i64 dj) {
    if (c.i == i && c.j == j && c.k == k && c.origin == origin) return;
    if (c.i >= 0) {
        const bool same = c.origin == origin;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int ni = node_of(c.i, q >> 2), nj = node_of(c.j, (q >> 1) & 1), nk = node_of(c.k, q & 1);
            const bool kept = same && (unsigned)(ni - i) <= 1u && (unsigned)(nj - j) <= 1u && (unsigned)(nk - k) <= 1u;
            if (!kept) {
                const i64 node = c.origin + ni * c.di + nj * c.dj + nk;
#pragma unroll
                for (int m = 0; m < NQ; ++m) {
                    flush_units(acc, node + m * c.qstride, c.s[m][q]);
                    c.s[m][q] = 0.0;
                }
            }
        }
    }
    c.i = i; c.j = j; c.k = k;
    c.origin = origin; c.di = di; c.dj = dj;
}

template <int NQ>
__device__ __forceinline__ void cache_flush(NodeCache<NQ>& c, u64* __restrict__ acc) {
    cache_move(c, acc, -1, -1, -1, -1, 0, 0);
    c.i = -1;
}

// The trilinear shares of value[m] (pyccel.py:472-480: ((1 - a)(1 - b))(1 - c) value for
// node (i, j, k), and alike) into the cache's cell, which must be the deposit's cell.
template <int NQ>
__device__ __forceinline__ void cache_add(NodeCache<NQ>& c, u64* __restrict__ acc, double a, double b, double cz,
                                         const double (&value)[NQ], int e) {
    const double a0 = 1.0 - a, b0 = 1.0 - b, c0 = 1.0 - cz;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        // slot q holds the node (i + da, j + db, k + dc) of parities q
        const int da = (c.i ^ (q >> 2)) & 1, db = (c.j ^ (q >> 1)) & 1, dc = (c.k ^ q) & 1;
        const double w = ((da ? a : a0) * (db ? b : b0)) * (dc ? cz : c0);
#pragma unroll
        for (int m = 0; m < NQ; ++m) {
            const double share = w * value[m];
            const double t = ldexp(share, -e);
            if (fabs(t) < 0x1p50) {
                const double rounded = (t + 0x1.8p52) - 0x1.8p52;  // to nearest, ties to even
                const double sum = c.s[m][q] + rounded;               // exact: below 2^51
                if (fabs(sum) < 0x1p51) {
                    c.s[m][q] = sum;
                    continue;
                }
                flush_units(acc, c.origin + m * c.qstride + (c.i + da) * c.di + (c.j + db) * c.dj + (c.k + dc),
                            c.s[m][q]);
                c.s[m][q] = rounded;
                continue;
            }
            fixed_add(acc, c.origin + m * c.qstride + (c.i + da) * c.di + (c.j + db) * c.dj + (c.k + dc), share, e);
        }
    }
}

template <int INTEG, bool HAS_B>
__device__ void trace_one(const Trace3DParams& p, i64 k) {
    const i64 N = p.n;
    double x0 = p.x[k], x1 = p.x[N + k], x2 = p.x[2 * N + k];
    double v0 = p.v[k], v1 = p.v[N + k], v2 = p.v[2 * N + k];
    const double qmk = p.qm[k], dqk = p.dq[k], inv_c2 = p.inv_c2;
    i64 slot = p.path_slot != nullptr ? p.path_slot[k] : -1;
    if (slot >= p.n_paths) slot = -1;
    i64 ns = p.steps[k];
    int st = -1;
    int have = 0;
    double ce[3] = {0.0, 0.0, 0.0}, cb[3] = {0.0, 0.0, 0.0};
    NodeCache<1> charge;
    NodeCache<3> current;
    cache_init(charge, 0);
    cache_init(current, (i64)p.nx * p.ny * p.nz);
    constexpr int NSTAGE = INTEG == 0 ? 4 : 2;
    while (ns < p.max_steps) {
        // the spacing of the finest patch at the step's start (pyccel.py:149-163)
        double c1 = x1, c2 = x2;
        if (p.mirror_y != 0 && c1 < p.oy) c1 = 2.0 * p.oy - c1;
        if (p.mirror_z != 0 && c2 < p.oz) c2 = 2.0 * p.oz - c2;
        const int start_patch = patch_at(p, x0, c1, c2);
        const double hloc = start_patch == p.n_patches - 1 ? p.h : p.patch_real[start_patch * 10 + 3];
        const double gv = 1.0 / lorentz_factor(v0, v1, v2, inv_c2);
        double q0 = x0, q1 = x1, q2 = x2, kx0 = v0, kx1 = v1, kx2 = v2;
        double sx0 = 0.0, sx1 = 0.0, sx2 = 0.0, sv0 = 0.0, sv1 = 0.0, sv2 = 0.0;
        double w0 = 0.0, w1 = 0.0, w2 = 0.0, y0, y1, y2;
        double dt = 0.0, hdt = 0.0, dt6 = 0.0;
#pragma unroll
        for (int stage = 0; stage < NSTAGE; ++stage) {
            double e[3], b[3];
            if (INTEG != 0 && stage == 0 && have != 0) {
                e[0] = ce[0]; e[1] = ce[1]; e[2] = ce[2];
                b[0] = cb[0]; b[1] = cb[1]; b[2] = cb[2];
            } else {
                fields_at(p, q0, q1, q2, e, b);
                if (!HAS_B) b[0] = b[1] = b[2] = 0.0;
            }
            double gk = 1.0 / lorentz_factor(kx0, kx1, kx2, inv_c2);  // pyccel.py:320-324
            const double g0 = qmk * (e[0] + gk * kx1 * b[2] - gk * kx2 * b[1]);
            const double g1 = qmk * (e[1] + gk * kx2 * b[0] - gk * kx0 * b[2]);
            const double g2 = qmk * (e[2] + gk * kx0 * b[1] - gk * kx1 * b[0]);
            if (stage == 0) {
                const double speed = gv * sqrt(v0 * v0 + v1 * v1 + v2 * v2);
                const double accel = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
                dt = -1.0;
                if (speed > 0.0) dt = hloc / speed;
                if (accel > 0.0) {
                    const double d2 = sqrt(2.0 * hloc / accel);
                    if (dt < 0.0 || d2 < dt) dt = d2;
                }
                if (dt < 0.0) dt = hloc;
                dt = p.step_fraction * dt;
                hdt = 0.5 * dt;
                dt6 = dt / 6.0;
                if (INTEG != 0) {  // Boris: a half step's kick-rotate-kick, then the drift
                    kx0 = v0; kx1 = v1; kx2 = v2;
                    boris_update(kx0, kx1, kx2, e, b, hdt, qmk, inv_c2);
                    gk = 1.0 / lorentz_factor(kx0, kx1, kx2, inv_c2);
                    q0 = x0 + dt * gk * kx0;
                    q1 = x1 + dt * gk * kx1;
                    q2 = x2 + dt * gk * kx2;
                    continue;
                }
                sx0 = gv * v0; sx1 = gv * v1; sx2 = gv * v2;
                sv0 = g0; sv1 = g1; sv2 = g2;
                q0 = x0 + hdt * gv * v0;
                q1 = x1 + hdt * gv * v1;
                q2 = x2 + hdt * gv * v2;
                kx0 = v0 + hdt * g0;
                kx1 = v1 + hdt * g1;
                kx2 = v2 + hdt * g2;
            } else if (INTEG != 0) {  // the second half step's kick at the end's fields
                w0 = kx0; w1 = kx1; w2 = kx2;
                boris_update(w0, w1, w2, e, b, hdt, qmk, inv_c2);
                ce[0] = e[0]; ce[1] = e[1]; ce[2] = e[2];
                cb[0] = b[0]; cb[1] = b[1]; cb[2] = b[2];
                have = 1;
            } else if (stage == 1 || stage == 2) {
                const double c = stage == 1 ? hdt : dt;
                sx0 = sx0 + 2.0 * gk * kx0; sx1 = sx1 + 2.0 * gk * kx1; sx2 = sx2 + 2.0 * gk * kx2;
                sv0 = sv0 + 2.0 * g0; sv1 = sv1 + 2.0 * g1; sv2 = sv2 + 2.0 * g2;
                q0 = x0 + c * gk * kx0; q1 = x1 + c * gk * kx1; q2 = x2 + c * gk * kx2;
                kx0 = v0 + c * g0; kx1 = v1 + c * g1; kx2 = v2 + c * g2;
            } else {
                sx0 = sx0 + gk * kx0; sx1 = sx1 + gk * kx1; sx2 = sx2 + gk * kx2;
                sv0 = sv0 + g0; sv1 = sv1 + g1; sv2 = sv2 + g2;
            }
        }
        if (INTEG == 0) {
            y0 = x0 + dt6 * sx0; y1 = x1 + dt6 * sx1; y2 = x2 + dt6 * sx2;
            w0 = v0 + dt6 * sv0; w1 = v1 + dt6 * sv1; w2 = v2 + dt6 * sv2;
        } else {
            y0 = q0; y1 = q1; y2 = q2;
        }
        // space charge: dq dt at the Hermite midpoint, trilinear (pyccel.py:432-480)
        const double gw = 1.0 / lorentz_factor(w0, w1, w2, inv_c2);
        const double m0 = 0.5 * (x0 + y0) + dt * (gv * v0 - gw * w0) / 8.0;
        double m1 = 0.5 * (x1 + y1) + dt * (gv * v1 - gw * w1) / 8.0;
        double m2 = 0.5 * (x2 + y2) + dt * (gv * v2 - gw * w2) / 8.0;
        if (p.mirror_y != 0 && m1 < p.oy) m1 = 2.0 * p.oy - m1;
        if (p.mirror_z != 0 && m2 < p.oz) m2 = 2.0 * p.oz - m2;
        {
            const int pp = patch_at(p, m0, m1, m2);
            const i64* pi = p.patch_int + pp * 8;
            const double* pr = p.patch_real + pp * 10;
            const double phh = pr[3];
            int i, j, kz;
            double a, b, c;
            deposit_axis((m0 - pr[0]) / phh, (int)pi[0], i, a);
            deposit_axis((m1 - pr[1]) / phh, (int)pi[1], j, b);
            deposit_axis((m2 - pr[2]) / phh, (int)pi[2], kz, c);
            cache_move(charge, p.acc_q, i, j, kz, pi[7], pi[1] * pi[2], pi[2]);
            const double value[1] = {dqk * dt};
            cache_add(charge, p.acc_q, a, b, c, value, p.e_q);
        }
        if (p.has_j != 0) {  // the current element dq (x' - x) on the base mesh (pyccel.py:481-540)
            double jd[3] = {dqk * (y0 - x0), dqk * (y1 - x1), dqk * (y2 - x2)};
            if (p.mirror_y != 0 && 0.5 * (x1 + y1) < p.oy) jd[1] = -jd[1];
            if (p.mirror_z != 0 && 0.5 * (x2 + y2) < p.oz) jd[2] = -jd[2];
            int bi, bj, bk;
            double ba, bb, bc;
            deposit_axis((m0 - p.ox) / p.h, p.nx, bi, ba);
            deposit_axis((m1 - p.oy) / p.h, p.ny, bj, bb);
            deposit_axis((m2 - p.oz) / p.h, p.nz, bk, bc);
            cache_move(current, p.acc_j, bi, bj, bk, 0, (i64)p.ny * p.nz, p.nz);
            cache_add(current, p.acc_j, ba, bb, bc, jd, p.e_j);
        }
        ns = ns + 1;
        for (int pl = 0; pl < p.n_planes; ++pl) {  // plane crossings (pyccel.py:543-559)
            const double xp = p.planes[pl];
            const double before = x0 - xp, after = y0 - xp;
            if ((before < 0.0) != (after < 0.0)) {
                const double s = before / (before - after);
                const i64 idx = k * p.n_planes + pl;
                const i64 cc = p.counts[idx];
                if (cc < p.capacity) {
                    double* rec = p.hits + (idx * p.capacity + cc) * 8;
                    rec[0] = xp;
                    rec[1] = x1 + s * (y1 - x1);
                    rec[2] = x2 + s * (y2 - x2);
                    rec[3] = gv * v0 + s * (gw * w0 - gv * v0);
                    rec[4] = gv * v1 + s * (gw * w1 - gv * v1);
                    rec[5] = gv * v2 + s * (gw * w2 - gv * v2);
                    rec[6] = after >= 0.0 ? 1.0 : -1.0;
                    rec[7] = 1.0 * (double)ns;
                }
                p.counts[idx] = cc + 1;
            }
        }
        x0 = y0; x1 = y1; x2 = y2;
        v0 = w0; v1 = w1; v2 = w2;
        // the end of the step: off the mesh or in an electrode (pyccel.py:567-624)
        const double px = x0;
        double py = x1, pz = x2;
        if (p.mirror_y != 0 && py < p.oy) py = 2.0 * p.oy - py;
        if (p.mirror_z != 0 && pz < p.oz) pz = 2.0 * p.oz - pz;
        if (slot >= 0) {
            const i64 cc = p.path_count[slot];
            if (cc < p.path_capacity) {
                double* rec = p.paths + (slot * p.path_capacity + cc) * 3;
                rec[0] = px; rec[1] = py; rec[2] = pz;
            }
            p.path_count[slot] = cc + 1;
        }
        if (px >= p.ux) {
            st = 0;
        } else if (px < p.ox) {
            st = 1;
        } else if (py > p.uy || py < p.oy || pz > p.uz || pz < p.oz) {
            st = 2;
        } else {
            // int((p - o) / h), at most n - 2 (the point is inside the box: not below 0)
            const double fi = (px - p.ox) / p.h, fj = (py - p.oy) / p.h, fk = (pz - p.oz) / p.h;
            const int i = (int)fmin(fmax(trunc(fi), 0.0), (double)(p.nx - 2));
            const int j = (int)fmin(fmax(trunc(fj), 0.0), (double)(p.ny - 2));
            const int kz = (int)fmin(fmax(trunc(fk), 0.0), (double)(p.nz - 2));
            const i64 cell = p.cells[((i64)i * p.cells_n1 + j) * p.cells_n2 + kz];
            if (cell >= 0) {
                const int R = p.refine;
                const double Rd = (double)R;
                const int si = (int)fmin(fmax(trunc((fi - (double)i) * Rd), 0.0), Rd - 1.0);
                const int sj = (int)fmin(fmax(trunc((fj - (double)j) * Rd), 0.0), Rd - 1.0);
                const int sk = (int)fmin(fmax(trunc((fk - (double)kz) * Rd), 0.0), Rd - 1.0);
                const i64 at = ((cell * R + si) * R + sj) * R + sk;
                const i64 e = at < p.n_subcells ? p.subcells[at] : -1;
                if (e >= 0) {
                    st = 3;
                    p.electrode[k] = e;
                } else if (e == -2) {
                    st = -2;
                }
            }
        }
        if (st != -1) break;
    }
    if (st == -1) st = 4;
    cache_flush(charge, p.acc_q);
    if (p.has_j != 0) cache_flush(current, p.acc_j);
    p.x[k] = x0; p.x[N + k] = x1; p.x[2 * N + k] = x2;
    p.v[k] = v0; p.v[N + k] = v1; p.v[2 * N + k] = v2;
    p.steps[k] = ns;
    p.status[k] = st;
}

}  // namespace

// Two waves per SIMD without a magnetic field (the node caches fit 256 registers); with one,
// the registers the B gather needs.
template <int INTEG, bool HAS_B>
__global__ __launch_bounds__(256, HAS_B ? 1 : 2) void trace3d_kernel(Trace3DParams p) {
    const i64 stride = (i64)gridDim.x * blockDim.x;
    for (i64 m = (i64)blockIdx.x * blockDim.x + threadIdx.x; m < p.n_active; m += stride) {
        const i64 k = p.active[m];
        if (k >= 0 && k < p.n) trace_one<INTEG, HAS_B>(p, k);  // an index that is no trajectory is skipped
    }
}

__global__ void zero_words_kernel(u64* __restrict__ w, i64 count) {
    const i64 stride = (i64)gridDim.x * blockDim.x;
    for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; t < count; t += stride) w[t] = 0ull;
}

// out[c] += the fixed-point total of node c
__global__ void add_words_kernel(const u64* __restrict__ acc, i64 nodes, int e, double* __restrict__ out) {
    const i64 stride = (i64)gridDim.x * blockDim.x;
    for (i64 c = (i64)blockIdx.x * blockDim.x + threadIdx.x; c < nodes; c += stride) {
        const u64 w0 = acc[kWords * c], w1 = acc[kWords * c + 1];
        if (w0 != 0 || w1 != 0) out[c] += words_value(w0, w1, e);
    }
}

static unsigned grid_for(i64 items) {
    i64 g = (items + 255) / 256;
    if (g > 4096) g = 4096;
    return (unsigned)(g < 1 ? 1 : g);
}

int launch_trace3d(const Trace3DParams& p, int integrator, i64 charge_nodes, double* charge, i64 current_nodes,
                   double* current, hipStream_t stream) {
    const i64 wq = kWords * charge_nodes, wj = p.has_j != 0 ? kWords * current_nodes : 0;
    hipLaunchKernelGGL(zero_words_kernel, dim3(grid_for(wq)), dim3(256), 0, stream, p.acc_q, wq);
    if (wj > 0) hipLaunchKernelGGL(zero_words_kernel, dim3(grid_for(wj)), dim3(256), 0, stream, p.acc_j, wj);
    if (p.n_active > 0) {
        i64 grid = (p.n_active + 255) / 256;
        if (grid > 65535) grid = 65535;
        const dim3 g((unsigned)grid), b(256);
        if (integrator == 0) {
            if (p.has_b) hipLaunchKernelGGL((trace3d_kernel<0, true>), g, b, 0, stream, p);
            else hipLaunchKernelGGL((trace3d_kernel<0, false>), g, b, 0, stream, p);
        } else {
            if (p.has_b) hipLaunchKernelGGL((trace3d_kernel<1, true>), g, b, 0, stream, p);
            else hipLaunchKernelGGL((trace3d_kernel<1, false>), g, b, 0, stream, p);
        }
    }
    hipLaunchKernelGGL(add_words_kernel, dim3(grid_for(charge_nodes)), dim3(256), 0, stream, p.acc_q, charge_nodes,
                       p.e_q, charge);
    if (wj > 0)
        hipLaunchKernelGGL(add_words_kernel, dim3(grid_for(current_nodes)), dim3(256), 0, stream, p.acc_j,
                           current_nodes, p.e_j, current);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // namespace bt
