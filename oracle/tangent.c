/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * Forward-mode tangent of the C restatement of the trace (oracle.c, trace_one): every
 * value of the classical, electrostatic trace on one uniform mesh carries its derivative
 * along a direction (dfx, dfy) of the edge fields.  The values are computed with exactly
 * the operations of trace_one (same order, -ffp-contract=off), so the trajectories, steps
 * and crossing records are the oracle's; the tangents of the records (y, z, vx, vy, vz) are
 * the directional derivatives of the discrete map: the chosen branch of the step rule
 * (trace_trajectories_2d_pyccel.py:282-295), the gather's cell (pyccel.py:196-269) and the
 * crossing fraction (pyccel.py:426-442).  tests/ check it against central differences of
 * the oracle itself and use it as the CPU tangent that the GPU adjoint must transpose.
 * Used by tests/ only.
 */
#include <math.h>
#include <stdint.h>

typedef int64_t i64;

/* oracle.c's argument struct (kept in step with it) */
typedef struct {
    const i64* active; i64 n_active;
    double* x; double* v; i64 n;
    const double* qm; const double* dq;
    i64* steps; i64* status; i64* electrode;
    const double* fxs; const double* fys;
    const i64* patch_int; const double* patch_real; i64 n_patches;
    double origin_x, origin_y, h, upper_x, upper_y;
    i64 cylindrical, integrator;
    double step_fraction;
    i64 max_steps;
    const i64* cells; i64 cells_cols;
    const i64* subcells; i64 refine;
    const double* planes; i64 n_planes;
    double* hits; i64 capacity;
    i64* counts;
    const i64* path_slot; double* paths; i64 path_capacity; i64* path_count;
    i64 num_threads; double* charge; i64 charge_stride;
    i64 graded; const double* x_nodes; i64 n_x_nodes;
    double inv_c2;
    i64 has_b; const double* b0; const double* b1; const double* b2;
} tangent_trace_args;

typedef struct {
    double v, d;  /* a value and its tangent */
} D;

static D cst(double v) { D r = {v, 0.0}; return r; }
static D dadd(D a, D b) { D r = {a.v + b.v, a.d + b.d}; return r; }
static D dsub(D a, D b) { D r = {a.v - b.v, a.d - b.d}; return r; }
static D dmul(D a, D b) { D r = {a.v * b.v, a.d * b.v + a.v * b.d}; return r; }
static D ddiv(D a, D b) { double q = a.v / b.v; D r = {q, (a.d - q * b.d) / b.v}; return r; }
static D dsqrt(D a) { double s = sqrt(a.v); D r = {s, s > 0.0 ? a.d / (2.0 * s) : 0.0}; return r; }
static D dhypot(D a, D b) {
    double s = hypot(a.v, b.v);
    D r = {s, s > 0.0 ? (a.v * a.d + b.v * b.d) / s : 0.0};
    return r;
}

/* the clamped bilinear of f at (s, t) with the tangent of f along df */
static D gather(const double* f, const double* df, i64 cols, D s, D t, i64 imax, i64 jmax, i64 off) {
    i64 i = (i64)floor(s.v);
    if (i < 0) i = 0; else if (i > imax) i = imax;
    i64 j = (i64)floor(t.v);
    if (j < 0) j = 0; else if (j > jmax) j = jmax;
    D u = dsub(s, cst((double)i)), w = dsub(t, cst((double)j));
    i64 c = off + i * cols + j;
    D f00 = {f[c], df[c]}, f01 = {f[c + 1], df[c + 1]};
    D f10 = {f[c + cols], df[c + cols]}, f11 = {f[c + cols + 1], df[c + cols + 1]};
    D low = dadd(f00, dmul(w, dsub(f01, f00)));
    D high = dadd(f10, dmul(w, dsub(f11, f10)));
    return dadd(low, dmul(u, dsub(high, low)));
}

static void accel(const tangent_trace_args* a, const double* dfx, const double* dfy, double qm, D q0, D q1, D q2,
                  D* g0, D* g1, D* g2) {
    D qy = q1, cy = cst(0.0), cz = cst(0.0);
    if (a->cylindrical == 1) {
        D r = dhypot(q1, q2);
        if (r.v > 0.0) { cy = ddiv(q1, r); cz = ddiv(q2, r); }
        qy = r;
    } else if (a->cylindrical == 2) {
        cy = cst(1.0);
        double sgn = 1.0;
        if (q1.v < a->origin_y) { cy = cst(-1.0); sgn = -1.0; }
        qy.v = a->origin_y + fabs(q1.v - a->origin_y);
        qy.d = sgn * q1.d;
    }
    const i64 pnx = a->patch_int[0], pny = a->patch_int[1], ofx = a->patch_int[2], ofy = a->patch_int[3];
    const double pox = a->patch_real[0], poy = a->patch_real[1], phh = a->patch_real[2];
    D sx = ddiv(dsub(q0, cst(pox)), cst(phh));
    D sy = ddiv(dsub(qy, cst(poy)), cst(phh));
    D ex = gather(a->fxs, dfx, pny, dadd(sx, cst(0.5)), sy, pnx - 1, pny - 2, ofx);
    D ey = gather(a->fys, dfy, pny + 1, sx, dadd(sy, cst(0.5)), pnx - 2, pny - 1, ofy);
    D q = cst(qm);
    if (a->cylindrical != 0) {
        *g0 = dmul(q, ex);
        *g1 = dmul(q, dmul(ey, cy));
        *g2 = dmul(q, dmul(ey, cz));
    } else {
        *g0 = dmul(q, ex);
        *g1 = dmul(q, ey);
        *g2 = dmul(q, cst(0.0));
    }
}

/* One trajectory with its tangent; rec_d (n, n_planes, capacity, 5) receives the tangents
 * of its crossing records. */
static void tangent_one(const tangent_trace_args* a, const double* dfx, const double* dfy, i64 k, double* rec_d) {
    const i64 nx = a->patch_int[0], ny = a->patch_int[1];
    const i64 N = a->n;
    D x0 = cst(a->x[k]), x1 = cst(a->x[N + k]), x2 = cst(a->x[2 * N + k]);
    D v0 = cst(a->v[k]), v1 = cst(a->v[N + k]), v2 = cst(a->v[2 * N + k]);
    const double qmk = a->qm[k];
    i64 ns = a->steps[k];
    i64 st = -1;
    int have = 0;
    D ca0 = cst(0.0), ca1 = cst(0.0), ca2 = cst(0.0);
    const double hloc = a->h;
    while (ns < a->max_steps) {
        D a0, a1, a2;
        if (a->integrator != 0 && have) {
            a0 = ca0; a1 = ca1; a2 = ca2;
        } else {
            accel(a, dfx, dfy, qmk, x0, x1, x2, &a0, &a1, &a2);
        }
        D speed = dmul(cst(1.0), dsqrt(dadd(dadd(dmul(v0, v0), dmul(v1, v1)), dmul(v2, v2))));
        D acc = dsqrt(dadd(dadd(dmul(a0, a0), dmul(a1, a1)), dmul(a2, a2)));
        D dt = cst(-1.0);
        if (speed.v > 0.0) dt = ddiv(cst(hloc), speed);
        if (acc.v > 0.0) {
            D d2 = dsqrt(ddiv(cst(2.0 * hloc), acc));
            if (dt.v < 0.0 || d2.v < dt.v) dt = d2;
        }
        if (dt.v < 0.0) dt = cst(hloc);
        dt = dmul(cst(a->step_fraction), dt);
        D hdt = dmul(cst(0.5), dt);
        D y0, y1, y2, w0, w1, w2;
        if (a->integrator == 0) {
            D dt6 = ddiv(dt, cst(6.0));
            D sx0 = v0, sx1 = v1, sx2 = v2, sv0 = a0, sv1 = a1, sv2 = a2;
            D q0 = dadd(x0, dmul(hdt, v0)), q1 = dadd(x1, dmul(hdt, v1)), q2 = dadd(x2, dmul(hdt, v2));
            D k0 = dadd(v0, dmul(hdt, a0)), k1 = dadd(v1, dmul(hdt, a1)), k2 = dadd(v2, dmul(hdt, a2));
            D g0, g1, g2;
            accel(a, dfx, dfy, qmk, q0, q1, q2, &g0, &g1, &g2);
            sx0 = dadd(sx0, dmul(cst(2.0), k0)); sx1 = dadd(sx1, dmul(cst(2.0), k1)); sx2 = dadd(sx2, dmul(cst(2.0), k2));
            sv0 = dadd(sv0, dmul(cst(2.0), g0)); sv1 = dadd(sv1, dmul(cst(2.0), g1)); sv2 = dadd(sv2, dmul(cst(2.0), g2));
            q0 = dadd(x0, dmul(hdt, k0)); q1 = dadd(x1, dmul(hdt, k1)); q2 = dadd(x2, dmul(hdt, k2));
            k0 = dadd(v0, dmul(hdt, g0)); k1 = dadd(v1, dmul(hdt, g1)); k2 = dadd(v2, dmul(hdt, g2));
            accel(a, dfx, dfy, qmk, q0, q1, q2, &g0, &g1, &g2);
            sx0 = dadd(sx0, dmul(cst(2.0), k0)); sx1 = dadd(sx1, dmul(cst(2.0), k1)); sx2 = dadd(sx2, dmul(cst(2.0), k2));
            sv0 = dadd(sv0, dmul(cst(2.0), g0)); sv1 = dadd(sv1, dmul(cst(2.0), g1)); sv2 = dadd(sv2, dmul(cst(2.0), g2));
            q0 = dadd(x0, dmul(dt, k0)); q1 = dadd(x1, dmul(dt, k1)); q2 = dadd(x2, dmul(dt, k2));
            k0 = dadd(v0, dmul(dt, g0)); k1 = dadd(v1, dmul(dt, g1)); k2 = dadd(v2, dmul(dt, g2));
            accel(a, dfx, dfy, qmk, q0, q1, q2, &g0, &g1, &g2);
            sx0 = dadd(sx0, k0); sx1 = dadd(sx1, k1); sx2 = dadd(sx2, k2);
            sv0 = dadd(sv0, g0); sv1 = dadd(sv1, g1); sv2 = dadd(sv2, g2);
            y0 = dadd(x0, dmul(dt6, sx0)); y1 = dadd(x1, dmul(dt6, sx1)); y2 = dadd(x2, dmul(dt6, sx2));
            w0 = dadd(v0, dmul(dt6, sv0)); w1 = dadd(v1, dmul(dt6, sv1)); w2 = dadd(v2, dmul(dt6, sv2));
        } else {
            D k0 = dadd(v0, dmul(hdt, a0)), k1 = dadd(v1, dmul(hdt, a1)), k2 = dadd(v2, dmul(hdt, a2));
            y0 = dadd(x0, dmul(dt, k0)); y1 = dadd(x1, dmul(dt, k1)); y2 = dadd(x2, dmul(dt, k2));
            D g0, g1, g2;
            accel(a, dfx, dfy, qmk, y0, y1, y2, &g0, &g1, &g2);
            w0 = dadd(k0, dmul(hdt, g0)); w1 = dadd(k1, dmul(hdt, g1)); w2 = dadd(k2, dmul(hdt, g2));
            ca0 = g0; ca1 = g1; ca2 = g2;
            have = 1;
        }
        ns = ns + 1;
        for (i64 p = 0; p < a->n_planes; ++p) {
            const double xp = a->planes[p];
            D before = dsub(x0, cst(xp)), after = dsub(y0, cst(xp));
            if ((before.v < 0.0) != (after.v < 0.0)) {
                D s = ddiv(before, dsub(before, after));
                i64 idx = k * a->n_planes + p;
                i64 c = a->counts[idx];
                if (c < a->capacity) {
                    D r1 = dadd(x1, dmul(s, dsub(y1, x1)));
                    D r2 = dadd(x2, dmul(s, dsub(y2, x2)));
                    D r3 = dadd(dmul(cst(1.0), v0), dmul(s, dsub(dmul(cst(1.0), w0), dmul(cst(1.0), v0))));
                    D r4 = dadd(dmul(cst(1.0), v1), dmul(s, dsub(dmul(cst(1.0), w1), dmul(cst(1.0), v1))));
                    D r5 = dadd(dmul(cst(1.0), v2), dmul(s, dsub(dmul(cst(1.0), w2), dmul(cst(1.0), v2))));
                    double* h = a->hits + (idx * a->capacity + c) * 8;
                    h[0] = xp; h[1] = r1.v; h[2] = r2.v; h[3] = r3.v; h[4] = r4.v; h[5] = r5.v;
                    h[6] = after.v >= 0.0 ? 1.0 : -1.0;
                    h[7] = 1.0 * (double)ns;
                    double* t = rec_d + (idx * a->capacity + c) * 5;
                    t[0] = r1.d; t[1] = r2.d; t[2] = r3.d; t[3] = r4.d; t[4] = r5.d;
                }
                a->counts[idx] = c + 1;
            }
        }
        x0 = y0; x1 = y1; x2 = y2;
        v0 = w0; v1 = w1; v2 = w2;
        /* the end tests of trace_one (values only) */
        double px = x0.v, py = x1.v;
        if (a->cylindrical == 1) py = hypot(x1.v, x2.v);
        else if (a->cylindrical == 2) py = a->origin_y + fabs(x1.v - a->origin_y);
        if (px >= a->upper_x) {
            st = 0;
        } else if (px < a->origin_x) {
            st = 1;
        } else if (py > a->upper_y || (a->cylindrical == 0 && py < a->origin_y)) {
            st = 2;
        } else {
            i64 i = (i64)((px - a->origin_x) / a->h);
            if (i > nx - 2) i = nx - 2;
            i64 j = (i64)((py - a->origin_y) / a->h);
            if (j > ny - 2) j = ny - 2;
            i64 cell = a->cells[i * a->cells_cols + j];
            if (cell >= 0) {
                i64 R = a->refine;
                i64 si = (i64)(((px - a->origin_x) / a->h - (double)i) * (double)R);
                if (si < 0) si = 0; else if (si > R - 1) si = R - 1;
                i64 sj = (i64)(((py - a->origin_y) / a->h - (double)j) * (double)R);
                if (sj < 0) sj = 0; else if (sj > R - 1) sj = R - 1;
                i64 e = a->subcells[(cell * R + si) * R + sj];
                if (e >= 0) {
                    st = 3;
                    a->electrode[k] = e;
                } else if (e == -2) {
                    st = -2;
                }
            }
        }
        if (st != -1) break;
    }
    if (st == -1) st = 4;
    a->x[k] = x0.v; a->x[N + k] = x1.v; a->x[2 * N + k] = x2.v;
    a->v[k] = v0.v; a->v[N + k] = v1.v; a->v[2 * N + k] = v2.v;
    a->steps[k] = ns;
    a->status[k] = st;
}

/* Returns -1 (nothing done) unless the trace is classical and electrostatic on one
 * uniform patch. */
int oracle_trace_tangent(const tangent_trace_args* a, const double* dfx, const double* dfy, double* rec_d) {
    if (a->n_patches != 1 || a->graded != 0 || a->inv_c2 != 0.0 || a->has_b != 0) return -1;
#pragma omp parallel for schedule(dynamic, 16)
    for (i64 m = 0; m < a->n_active; ++m) tangent_one(a, dfx, dfy, a->active[m], rec_d);
    return 0;
}
