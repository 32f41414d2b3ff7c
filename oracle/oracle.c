/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of beamtracer's particle-stage kernels, used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker and the
 * timed CPU baseline.  Nothing in beamtracer_amd/ links, loads or calls this file.
 *
 * Followed statement by statement, with the same operation order (build with
 * -ffp-contract=off so that no multiply-add is fused):
 *   staggered_field_2d       kernels/staggered_field_2d/staggered_field_2d_pyccel.py:13-55
 *   deposit_trajectories_2d  kernels/deposit_trajectories_2d/deposit_trajectories_2d_pyccel.py:13-59
 *   trace_trajectories_2d    kernels/trace_trajectories_2d/trace_trajectories_2d_pyccel.py:19-507
 *                            (second witness: trace_trajectories_2d_cuda.cu:12-362)
 *   edge_field               solver/fields.py:18-67 (_edge_field: the staggered edge fields of a
 *                            potential, the input of the trace)
 *
 * Extension beyond the reference's compiled kernel (which refuses a magnetic field,
 * solver/tracer.py:1060-1090): a node magnetic field (three components on the base mesh's
 * nodes, gathered like self_field.py:86-113) acting through the force of
 * solver/tracer.py:519-529, with Boris's rotation (solver/tracer.py:618-650) for
 * integrator 1 and v x B in every Runge--Kutta stage for integrator 0.  With has_b == 0
 * every operation is the reference's.
 *
 * Parity pins (tests/test_oracle_*.py): closed forms (bilinear fields reproduced exactly,
 * constant fields, charge conservation, uniform-field energy gain and flight time, RK4
 * order 4 and Boris order 2 on the harmonic well, |v| kept by Boris in a pure B,
 * Busch's theorem through a solenoid, the force of a uniform radial and azimuthal B,
 * exact mirror images below a symmetry plane, the Hermite midpoint of the deposit
 * under uniform acceleration), brute force on tiny grids, and an independent
 * step-synchronous NumPy restatement of solver/tracer.py:851-1054 (oracle/numpy_ref.py).
 *
 * Mutation switches (tests only): ORACLE_MUTANT_BR = -1 flips the radial B, _BTHETA = -1
 * the azimuthal B, _MIRROR_B = 1 drops the sign of B_x below a symmetry plane and
 * _MIDPOINT = -1 flips the dt (v - w) / 8 term of the deposit point, _LORENTZ = 1 takes
 * the first-order expansion 1 + u.u / 2c^2 for the Lorentz factor.  tests/ builds such
 * mutants to show that the pins above fail on each of these mistakes; the default build
 * is the reference's arithmetic.
 * The relativistic Lorentz factor (mpiparticles, un-vendored) is written from its
 * equation gamma = sqrt(1 + u.u / c^2), pinned by energy conservation in a uniform electric
 * field (gamma c^2 grows by q/m E dx; the speed at a plane follows from gamma alone): its
 * bit-level parity with mpiparticles is unpinned; for inv_c2 = 0 it is exactly 1.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef int64_t i64;

#ifndef ORACLE_MUTANT_BR
#define ORACLE_MUTANT_BR 1.0
#endif
#ifndef ORACLE_MUTANT_BTHETA
#define ORACLE_MUTANT_BTHETA 1.0
#endif
#ifndef ORACLE_MUTANT_MIRROR_B
#define ORACLE_MUTANT_MIRROR_B 0
#endif
#ifndef ORACLE_MUTANT_MIDPOINT
#define ORACLE_MUTANT_MIDPOINT 1.0
#endif
#ifndef ORACLE_MUTANT_LORENTZ
#define ORACLE_MUTANT_LORENTZ 0
#endif

/* gamma of the proper velocity u (mpiparticles lorentz_factor; equation of
 * solver/tracer.py:464-466): exactly 1 when inv_c2 == 0. */
static double lorentz_factor(double u0, double u1, double u2, double inv_c2) {
    const double s = (u0 * u0 + u1 * u1 + u2 * u2) * inv_c2;
    if (ORACLE_MUTANT_LORENTZ) return 1.0 + 0.5 * s;
    return sqrt(1.0 + s);
}

/* trace_trajectories_2d_pyccel.py:19-33 */
static i64 node_cell(double x, const double* nodes, i64 n) {
    i64 lo = 0, hi = n - 1;
    if (x >= nodes[n - 1]) {
        lo = n - 2;
    } else if (x > nodes[0]) {
        while (hi - lo > 1) {
            i64 mid = (lo + hi) / 2;
            if (nodes[mid] <= x) lo = mid; else hi = mid;
        }
    }
    return lo;
}

/* trace_trajectories_2d_pyccel.py:36-39 */
static double uniform_x(double x, const double* nodes, i64 n, double x0, double h) {
    i64 c = node_cell(x, nodes, n);
    return x0 + h * ((double)c + (x - nodes[c]) / (nodes[c + 1] - nodes[c]));
}

/* staggered_field_2d_pyccel.py:13-55 */
void oracle_staggered_field_2d(const double* px, const double* py, i64 n,
                               const double* fx, const double* fy, i64 nx, i64 ny,
                               double origin_x, double origin_y, double h,
                               double* ex, double* ey) {
    for (i64 k = 0; k < n; ++k) {
        double sx = (px[k] - origin_x) / h;
        double sy = (py[k] - origin_y) / h;
        double s = sx + 0.5;
        i64 i = (i64)floor(s);
        if (i < 0) i = 0; if (i > nx - 1) i = nx - 1;
        i64 j = (i64)floor(sy);
        if (j < 0) j = 0; if (j > ny - 2) j = ny - 2;
        double u = s - (double)i;
        double w = sy - (double)j;
        /* fx (nx + 1, ny) */
        double low = fx[i * ny + j] + w * (fx[i * ny + j + 1] - fx[i * ny + j]);
        double high = fx[(i + 1) * ny + j] + w * (fx[(i + 1) * ny + j + 1] - fx[(i + 1) * ny + j]);
        ex[k] = low + u * (high - low);
        double t = sy + 0.5;
        i = (i64)floor(sx);
        if (i < 0) i = 0; if (i > nx - 2) i = nx - 2;
        j = (i64)floor(t);
        if (j < 0) j = 0; if (j > ny - 1) j = ny - 1;
        u = sx - (double)i;
        w = t - (double)j;
        /* fy (nx, ny + 1) */
        i64 m = ny + 1;
        low = fy[i * m + j] + w * (fy[i * m + j + 1] - fy[i * m + j]);
        high = fy[(i + 1) * m + j] + w * (fy[(i + 1) * m + j + 1] - fy[(i + 1) * m + j]);
        ey[k] = low + u * (high - low);
    }
}

/* deposit_trajectories_2d_pyccel.py:13-59 (serial, in index order) */
void oracle_deposit_trajectories_2d(const double* px, const double* py, const double* q,
                                    i64 n, double origin_x, double origin_y, double h,
                                    i64 cylindrical, double* charge, i64 nx, i64 ny) {
    for (i64 k = 0; k < n; ++k) {
        double x = px[k], y = py[k];
        i64 i = (i64)floor((x - origin_x) / h);
        i64 j = (i64)floor((y - origin_y) / h);
        double t = (x - (origin_x + h * (double)i)) / h;
        double u;
        if (cylindrical != 0) {
            double r1 = origin_y + h * (double)j;
            double r2 = r1 + h;
            u = (y * y - r1 * r1) / (r2 * r2 - r1 * r1);
        } else {
            u = (y - (origin_y + h * (double)j)) / h;
        }
        if (i < 0) { t = 0.0; i = 0; }
        else if (i >= nx - 1) { t = 1.0; i = nx - 2; }
        if (j < 0) { u = 0.0; j = 0; }
        else if (j >= ny - 1) { u = 1.0; j = ny - 2; }
        double value = q[k];
        charge[i * ny + j] += (1.0 - t) * (1.0 - u) * value;
        charge[i * ny + j + 1] += (1.0 - t) * u * value;
        charge[(i + 1) * ny + j] += t * (1.0 - u) * value;
        charge[(i + 1) * ny + j + 1] += t * u * value;
    }
}

/* Every argument of trace_trajectories_2d_pyccel.py:42-78, flattened to pointers and
 * extents, plus the node magnetic field of the extension. */
typedef struct {
    const i64* active; i64 n_active;
    double* x; double* v; i64 n;             /* (3, n) component-major */
    const double* qm; const double* dq;
    i64* steps; i64* status; i64* electrode;
    const double* fxs; const double* fys;
    const i64* patch_int; const double* patch_real; i64 n_patches;  /* (P, 6), (P, 7) */
    double origin_x, origin_y, h, upper_x, upper_y;
    i64 cylindrical, integrator;
    double step_fraction;
    i64 max_steps;
    const i64* cells; i64 cells_cols;        /* (nx - 1, ny - 1) */
    const i64* subcells; i64 refine;
    const double* planes; i64 n_planes;
    double* hits; i64 capacity;              /* (n, n_planes, capacity, 8) */
    i64* counts;                             /* (n, n_planes) */
    const i64* path_slot; double* paths; i64 path_capacity; i64* path_count;
    i64 num_threads; double* charge; i64 charge_stride;   /* (num_threads, nodes) */
    i64 graded; const double* x_nodes; i64 n_x_nodes;
    double inv_c2;
    /* extension: (B_x, B_r, B_theta) cylindrical or (B_x, B_y, B_z) planar on the base
     * mesh's nodes (nx, ny); a NULL component is zero */
    i64 has_b; const double* b0; const double* b1; const double* b2;
} oracle_trace_args;

/* self_field.py:100-113: bilinear on nodes, constant beyond the mesh */
static double node_bilinear(const double* f, double sx, double sy, i64 nx, i64 ny) {
    if (f == NULL) return 0.0;
    if (sx < 0.0) sx = 0.0; if (sx > (double)(nx - 1)) sx = (double)(nx - 1);
    if (sy < 0.0) sy = 0.0; if (sy > (double)(ny - 1)) sy = (double)(ny - 1);
    i64 i = (i64)sx; if (i > nx - 2) i = nx - 2;
    i64 j = (i64)sy; if (j > ny - 2) j = ny - 2;
    double a = sx - (double)i, b = sy - (double)j;
    return (1.0 - a) * ((1.0 - b) * f[i * ny + j] + b * f[i * ny + j + 1])
         + a * ((1.0 - b) * f[(i + 1) * ny + j] + b * f[(i + 1) * ny + j + 1]);
}

/* The Cartesian B at mesh point (lq, qy) (lq uniform x), rotated like
 * solver/tracer.py:70-83 (cylindrical) or mirrored like tracer.py:547-570. */
static void magnetic_at(const oracle_trace_args* a, double lq, double qy, double cy, double cz,
                        i64 nx, i64 ny, double* B0, double* B1, double* B2) {
    double sx = (lq - a->origin_x) / a->h;
    double sy = (qy - a->origin_y) / a->h;
    double c0 = node_bilinear(a->b0, sx, sy, nx, ny);
    double c1 = node_bilinear(a->b1, sx, sy, nx, ny);
    double c2 = node_bilinear(a->b2, sx, sy, nx, ny);
    if (a->cylindrical == 1) {
        *B0 = c0;
        *B1 = ORACLE_MUTANT_BR * c1 * cy - ORACLE_MUTANT_BTHETA * c2 * cz;
        *B2 = ORACLE_MUTANT_BR * c1 * cz + ORACLE_MUTANT_BTHETA * c2 * cy;
    } else if (a->cylindrical == 2) {
        *B0 = ORACLE_MUTANT_MIRROR_B ? c0 : cy * c0;   /* cy is the mirror sign s */
        *B1 = c1;
        *B2 = cy * c2;
    } else {
        *B0 = c0; *B1 = c1; *B2 = c2;
    }
}

/* Boris's velocity update over tau (solver/tracer.py:618-632): half kick, rotation, half
 * kick; relativistic: t scaled by 1 / gamma of v-minus. */
static void boris_kick(double* u0, double* u1, double* u2, double e0, double e1, double e2,
                       double b0, double b1, double b2, double tau, double qm, double inv_c2) {
    double hq = 0.5 * tau * qm;
    double m0 = *u0 + hq * e0, m1 = *u1 + hq * e1, m2 = *u2 + hq * e2;
    double g = 1.0 / lorentz_factor(m0, m1, m2, inv_c2);
    double t0 = hq * b0 * g, t1 = hq * b1 * g, t2 = hq * b2 * g;
    double f = 2.0 / (1.0 + (t0 * t0 + t1 * t1 + t2 * t2));
    double p0 = m0 + (m1 * t2 - m2 * t1);
    double p1 = m1 + (m2 * t0 - m0 * t2);
    double p2 = m2 + (m0 * t1 - m1 * t0);
    double s0 = f * t0, s1 = f * t1, s2 = f * t2;
    double n0 = m0 + (p1 * s2 - p2 * s1);
    double n1 = m1 + (p2 * s0 - p0 * s2);
    double n2 = m2 + (p0 * s1 - p1 * s0);
    *u0 = n0 + hq * e0; *u1 = n1 + hq * e1; *u2 = n2 + hq * e2;
}

/* The edge field and mesh coordinates at Cartesian point q (trace_trajectories_2d_pyccel.py
 * 196-269): returns E_x, E_y of the finest patch and the radial direction (cy, cz). */
static void edge_field(const oracle_trace_args* a, double q0, double q1, double q2,
                       double* ex, double* ey, double* cy_out, double* cz_out,
                       double* lq_out, double* qy_out) {
    double qy = q1, cy = 0.0, cz = 0.0;
    if (a->cylindrical == 1) {
        double r = hypot(q1, q2);
        if (r > 0.0) { cy = q1 / r; cz = q2 / r; }
        qy = r;
    } else if (a->cylindrical == 2) {
        cy = 1.0;
        if (q1 < a->origin_y) cy = -1.0;
        qy = a->origin_y + fabs(q1 - a->origin_y);
    }
    double lq = q0;
    if (a->graded != 0) lq = uniform_x(q0, a->x_nodes, a->n_x_nodes, a->origin_x, a->h);
    i64 pp = a->n_patches - 1;
    for (i64 qq = 0; qq < a->n_patches - 1; ++qq) {
        const double* pr = a->patch_real + qq * 7;
        if (lq >= pr[3] && lq <= pr[5] && qy >= pr[4] && qy <= pr[6]) { pp = qq; break; }
    }
    const i64* pi = a->patch_int + pp * 6;
    const double* pr = a->patch_real + pp * 7;
    i64 pnx = pi[0], pny = pi[1], ofx = pi[2], ofy = pi[3];
    double pox = pr[0], poy = pr[1], phh = pr[2];
    double sx = (lq - pox) / phh;
    double sy = (qy - poy) / phh;
    double s = sx + 0.5;
    i64 i = (i64)floor(s);
    if (i < 0) i = 0; else if (i > pnx - 1) i = pnx - 1;
    i64 j = (i64)floor(sy);
    if (j < 0) j = 0; else if (j > pny - 2) j = pny - 2;
    double u = s - (double)i, w = sy - (double)j;
    i64 c = ofx + i * pny + j;
    const double* f = a->fxs;
    double low = f[c] + w * (f[c + 1] - f[c]);
    double high = f[c + pny] + w * (f[c + pny + 1] - f[c + pny]);
    *ex = low + u * (high - low);
    s = sy + 0.5;
    i = (i64)floor(sx);
    if (i < 0) i = 0; else if (i > pnx - 2) i = pnx - 2;
    j = (i64)floor(s);
    if (j < 0) j = 0; else if (j > pny - 1) j = pny - 1;
    u = sx - (double)i; w = s - (double)j;
    c = ofy + i * (pny + 1) + j;
    f = a->fys;
    low = f[c] + w * (f[c + 1] - f[c]);
    high = f[c + pny + 1] + w * (f[c + pny + 2] - f[c + pny + 1]);
    *ey = low + u * (high - low);
    *cy_out = cy; *cz_out = cz; *lq_out = lq; *qy_out = qy;
}

/* One trajectory from its state until it ends (trace_trajectories_2d_pyccel.py:136-507). */
static void trace_one(const oracle_trace_args* a, i64 k, double* charge_row) {
    const i64 P = a->n_patches;
    const i64 nx = a->patch_int[(P - 1) * 6 + 0];
    const i64 ny = a->patch_int[(P - 1) * 6 + 1];
    const i64 N = a->n;
    double x0 = a->x[k], x1 = a->x[N + k], x2 = a->x[2 * N + k];
    double v0 = a->v[k], v1 = a->v[N + k], v2 = a->v[2 * N + k];
    const double qmk = a->qm[k];
    const i64 slot = a->path_slot[k];
    i64 ns = a->steps[k];
    i64 st = -1;
    double dt = 0.0, hdt = 0.0, dt6 = 0.0;
    const int nstage = a->integrator == 0 ? 4 : 2;
    int have = 0;
    double ca0 = 0.0, ca1 = 0.0, ca2 = 0.0;        /* Boris: the last end's acceleration */
    double ce0 = 0.0, ce1 = 0.0, ce2 = 0.0;        /* with B: the last end's E and B */
    double cb0 = 0.0, cb1 = 0.0, cb2 = 0.0;
    const double inv_c2 = a->inv_c2;
    while (ns < a->max_steps) {
        double rr = x1;
        if (a->cylindrical == 1) rr = hypot(x1, x2);
        else if (a->cylindrical == 2) rr = a->origin_y + fabs(x1 - a->origin_y);
        double lx = x0;
        if (a->graded != 0) lx = uniform_x(x0, a->x_nodes, a->n_x_nodes, a->origin_x, a->h);
        double hloc = a->h;
        for (i64 qq = 0; qq < P - 1; ++qq) {
            const double* pr = a->patch_real + qq * 7;
            if (lx >= pr[3] && lx <= pr[5] && rr >= pr[4] && rr <= pr[6]) { hloc = pr[2]; break; }
        }
        if (a->graded != 0) {
            i64 cg = node_cell(x0, a->x_nodes, a->n_x_nodes);
            double d2 = (a->x_nodes[cg + 1] - a->x_nodes[cg]) / a->h;
            if (d2 < 1.0) hloc = hloc * d2;
        }
        double gv = 1.0 / lorentz_factor(v0, v1, v2, inv_c2);
        double q0 = x0, q1 = x1, q2 = x2;
        double kx0 = v0, kx1 = v1, kx2 = v2;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        double sx0 = 0.0, sx1 = 0.0, sx2 = 0.0, sv0 = 0.0, sv1 = 0.0, sv2 = 0.0;
        double w0 = 0.0, w1 = 0.0, w2 = 0.0, y0, y1, y2;
        double g0, g1, g2, gk;
        double e0 = 0.0, e1 = 0.0, e2 = 0.0, b0 = 0.0, b1 = 0.0, b2 = 0.0;
        for (int stage = 0; stage < nstage; ++stage) {
            gk = 1.0 / lorentz_factor(kx0, kx1, kx2, inv_c2);
            if (stage == 0 && have != 0) {
                if (a->has_b) {
                    /* the end fields of the last step with this step's velocity */
                    e0 = ce0; e1 = ce1; e2 = ce2; b0 = cb0; b1 = cb1; b2 = cb2;
                    double c0 = gk * kx0, c1 = gk * kx1, c2 = gk * kx2;
                    g0 = qmk * (e0 + (c1 * b2 - c2 * b1));
                    g1 = qmk * (e1 + (c2 * b0 - c0 * b2));
                    g2 = qmk * (e2 + (c0 * b1 - c1 * b0));
                } else {
                    g0 = ca0; g1 = ca1; g2 = ca2;
                }
            } else {
                double ex, ey, cy, cz, lq, qy;
                edge_field(a, q0, q1, q2, &ex, &ey, &cy, &cz, &lq, &qy);
                if (a->cylindrical != 0) {
                    g0 = qmk * ex;
                    g1 = qmk * (ey * cy);
                    g2 = qmk * (ey * cz);
                } else {
                    g0 = qmk * ex;
                    g1 = qmk * ey;
                    g2 = qmk * 0.0;
                }
                if (a->has_b) {
                    e0 = ex;
                    e1 = a->cylindrical != 0 ? ey * cy : ey;
                    e2 = a->cylindrical == 1 ? ey * cz : 0.0;
                    magnetic_at(a, lq, qy, cy, cz, nx, ny, &b0, &b1, &b2);
                    double c0 = gk * kx0, c1 = gk * kx1, c2 = gk * kx2;
                    g0 = qmk * (e0 + (c1 * b2 - c2 * b1));
                    g1 = qmk * (e1 + (c2 * b0 - c0 * b2));
                    g2 = qmk * (e2 + (c0 * b1 - c1 * b0));
                }
            }
            if (stage == 0) {
                a0 = g0; a1 = g1; a2 = g2;
                double speed = gv * sqrt(v0 * v0 + v1 * v1 + v2 * v2);
                double accel = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
                dt = -1.0;
                if (speed > 0.0) dt = hloc / speed;
                if (accel > 0.0) {
                    double d2 = sqrt(2.0 * hloc / accel);
                    if (dt < 0.0 || d2 < dt) dt = d2;
                }
                if (dt < 0.0) dt = hloc;
                dt = a->step_fraction * dt;
                hdt = 0.5 * dt;
                dt6 = dt / 6.0;
                if (a->integrator != 0) {
                    if (a->has_b) {
                        kx0 = v0; kx1 = v1; kx2 = v2;
                        boris_kick(&kx0, &kx1, &kx2, e0, e1, e2, b0, b1, b2, hdt, qmk, inv_c2);
                    } else {
                        kx0 = v0 + hdt * a0;
                        kx1 = v1 + hdt * a1;
                        kx2 = v2 + hdt * a2;
                    }
                    gk = 1.0 / lorentz_factor(kx0, kx1, kx2, inv_c2);
                    q0 = x0 + dt * gk * kx0;
                    q1 = x1 + dt * gk * kx1;
                    q2 = x2 + dt * gk * kx2;
                    continue;
                }
                sx0 = gv * v0; sx1 = gv * v1; sx2 = gv * v2;
                sv0 = a0; sv1 = a1; sv2 = a2;
                q0 = x0 + hdt * gv * v0;
                q1 = x1 + hdt * gv * v1;
                q2 = x2 + hdt * gv * v2;
                kx0 = v0 + hdt * a0;
                kx1 = v1 + hdt * a1;
                kx2 = v2 + hdt * a2;
            } else if (a->integrator != 0) {
                if (a->has_b) {
                    w0 = kx0; w1 = kx1; w2 = kx2;
                    boris_kick(&w0, &w1, &w2, e0, e1, e2, b0, b1, b2, hdt, qmk, inv_c2);
                    ce0 = e0; ce1 = e1; ce2 = e2; cb0 = b0; cb1 = b1; cb2 = b2;
                } else {
                    w0 = kx0 + hdt * g0;
                    w1 = kx1 + hdt * g1;
                    w2 = kx2 + hdt * g2;
                    ca0 = g0; ca1 = g1; ca2 = g2;
                }
                have = 1;
            } else if (stage == 1) {
                sx0 = sx0 + 2.0 * gk * kx0; sx1 = sx1 + 2.0 * gk * kx1; sx2 = sx2 + 2.0 * gk * kx2;
                sv0 = sv0 + 2.0 * g0; sv1 = sv1 + 2.0 * g1; sv2 = sv2 + 2.0 * g2;
                q0 = x0 + hdt * gk * kx0; q1 = x1 + hdt * gk * kx1; q2 = x2 + hdt * gk * kx2;
                kx0 = v0 + hdt * g0; kx1 = v1 + hdt * g1; kx2 = v2 + hdt * g2;
            } else if (stage == 2) {
                sx0 = sx0 + 2.0 * gk * kx0; sx1 = sx1 + 2.0 * gk * kx1; sx2 = sx2 + 2.0 * gk * kx2;
                sv0 = sv0 + 2.0 * g0; sv1 = sv1 + 2.0 * g1; sv2 = sv2 + 2.0 * g2;
                q0 = x0 + dt * gk * kx0; q1 = x1 + dt * gk * kx1; q2 = x2 + dt * gk * kx2;
                kx0 = v0 + dt * g0; kx1 = v1 + dt * g1; kx2 = v2 + dt * g2;
            } else {
                sx0 = sx0 + gk * kx0; sx1 = sx1 + gk * kx1; sx2 = sx2 + gk * kx2;
                sv0 = sv0 + g0; sv1 = sv1 + g1; sv2 = sv2 + g2;
            }
        }
        if (a->integrator == 0) {
            y0 = x0 + dt6 * sx0; y1 = x1 + dt6 * sx1; y2 = x2 + dt6 * sx2;
            w0 = v0 + dt6 * sv0; w1 = v1 + dt6 * sv1; w2 = v2 + dt6 * sv2;
        } else {
            y0 = q0; y1 = q1; y2 = q2;
        }

        /* space charge at the Hermite midpoint (pyccel.py:372-423) */
        double gw = 1.0 / lorentz_factor(w0, w1, w2, inv_c2);
        double m0 = 0.5 * (x0 + y0) + ORACLE_MUTANT_MIDPOINT * dt * (gv * v0 - gw * w0) / 8.0;
        double m1 = 0.5 * (x1 + y1) + ORACLE_MUTANT_MIDPOINT * dt * (gv * v1 - gw * w1) / 8.0;
        double m2 = 0.5 * (x2 + y2) + ORACLE_MUTANT_MIDPOINT * dt * (gv * v2 - gw * w2) / 8.0;
        if (a->cylindrical == 1) m1 = hypot(m1, m2);
        else if (a->cylindrical == 2) m1 = a->origin_y + fabs(m1 - a->origin_y);
        if (a->graded != 0) m0 = uniform_x(m0, a->x_nodes, a->n_x_nodes, a->origin_x, a->h);
        i64 pp = P - 1;
        for (i64 qq = 0; qq < P - 1; ++qq) {
            const double* pr = a->patch_real + qq * 7;
            if (m0 >= pr[3] && m0 <= pr[5] && m1 >= pr[4] && m1 <= pr[6]) { pp = qq; break; }
        }
        {
            const i64* pi = a->patch_int + pp * 6;
            const double* pr = a->patch_real + pp * 7;
            i64 pnx = pi[0], pny = pi[1], ofn = pi[5];
            double pox = pr[0], poy = pr[1], phh = pr[2];
            i64 i = (i64)floor((m0 - pox) / phh);
            i64 j = (i64)floor((m1 - poy) / phh);
            double u = (m0 - (pox + phh * (double)i)) / phh;
            double w;
            if (a->cylindrical == 1) {
                double low = poy + phh * (double)j;
                double high = low + phh;
                w = (m1 * m1 - low * low) / (high * high - low * low);
            } else {
                w = (m1 - (poy + phh * (double)j)) / phh;
            }
            if (i < 0) { u = 0.0; i = 0; }
            else if (i >= pnx - 1) { u = 1.0; i = pnx - 2; }
            if (j < 0) { w = 0.0; j = 0; }
            else if (j >= pny - 1) { w = 1.0; j = pny - 2; }
            double deposit = a->dq[k] * dt;
            i64 c = ofn + i * pny + j;
            charge_row[c] += (1.0 - u) * (1.0 - w) * deposit;
            charge_row[c + 1] += (1.0 - u) * w * deposit;
            charge_row[c + pny] += u * (1.0 - w) * deposit;
            charge_row[c + pny + 1] += u * w * deposit;
        }
        ns = ns + 1;

        /* plane crossings (pyccel.py:426-442) */
        for (i64 p = 0; p < a->n_planes; ++p) {
            double xp = a->planes[p];
            double before = x0 - xp;
            double after = y0 - xp;
            if ((before < 0.0) != (after < 0.0)) {
                double s = before / (before - after);
                i64 idx = k * a->n_planes + p;
                i64 c = a->counts[idx];
                if (c < a->capacity) {
                    double* hrec = a->hits + (idx * a->capacity + c) * 8;
                    hrec[0] = xp;
                    hrec[1] = x1 + s * (y1 - x1);
                    hrec[2] = x2 + s * (y2 - x2);
                    hrec[3] = gv * v0 + s * (gw * w0 - gv * v0);
                    hrec[4] = gv * v1 + s * (gw * w1 - gv * v1);
                    hrec[5] = gv * v2 + s * (gw * w2 - gv * v2);
                    hrec[6] = after >= 0.0 ? 1.0 : -1.0;
                    hrec[7] = 1.0 * (double)ns;
                }
                a->counts[idx] = c + 1;
            }
        }
        x0 = y0; x1 = y1; x2 = y2;
        v0 = w0; v1 = w1; v2 = w2;

        /* the end of the step: off the mesh or in an electrode (pyccel.py:450-497) */
        double px = x0, py = x1;
        if (a->cylindrical == 1) py = hypot(x1, x2);
        else if (a->cylindrical == 2) py = a->origin_y + fabs(x1 - a->origin_y);
        if (slot >= 0) {
            i64 c = a->path_count[slot];
            if (c < a->path_capacity) {
                a->paths[(slot * a->path_capacity + c) * 2 + 0] = px;
                a->paths[(slot * a->path_capacity + c) * 2 + 1] = py;
            }
            a->path_count[slot] = c + 1;
        }
        if (a->graded != 0) px = uniform_x(px, a->x_nodes, a->n_x_nodes, a->origin_x, a->h);
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
    a->x[k] = x0; a->x[N + k] = x1; a->x[2 * N + k] = x2;
    a->v[k] = v0; a->v[N + k] = v1; a->v[2 * N + k] = v2;
    a->steps[k] = ns;
    a->status[k] = st;
}

/* trace_trajectories_2d_pyccel.py:129-135: thread t traces every num_threads-th active
 * trajectory and deposits into its own slab charge[t]. */
void oracle_trace_trajectories_2d(const oracle_trace_args* a) {
    const i64 T = a->num_threads > 0 ? a->num_threads : 1;
#pragma omp parallel for schedule(static, 1) num_threads(T)
    for (i64 t = 0; t < T; ++t) {
        for (i64 m = t; m < a->n_active; m += T) {
            trace_one(a, a->active[m], a->charge + t * a->charge_stride);
        }
    }
}

/* The size of the argument struct, checked by the Python side against its ctypes mirror. */
i64 oracle_trace_args_size(void) { return (i64)sizeof(oracle_trace_args); }

/* solver/fields.py:18-67 (_edge_field), one edge of one line at a time.  Along an axis of
 * n nodes (node t of line c at phi[t * st + c * sc]) the n + 1 edges of line c go to
 * out[e * ot + c * oc]; lengths[t] is the spacing from node t to node t + 1.  The selection
 * order is the reference's: the plain difference, replaced by the surface-distance form
 * for a near node before an electrode node, then for an electrode node before a near node,
 * then, on an edge with both nodes in an electrode, by the surface edge of the one side
 * that has one (else 0). */
static double oracle_edge_interior(const double* phi, const int32_t* solid, const uint8_t* near,
                                   const double* dist_lo, const double* dist_hi, const double* lengths,
                                   i64 n, i64 st, i64 c, i64 e) {
    const i64 ta = e - 1, tb = e;                     /* the edge's nodes */
    const i64 ia = ta * st + c, ib = tb * st + c;
    const double a = phi[ia], b = phi[ib];
    const double len = lengths[ta];
    const double len_before = ta >= 1 ? lengths[ta - 1] : 1.0;
    const double len_after = ta + 1 <= n - 2 ? lengths[ta + 1] : 1.0;
    const int sa = solid[ia] >= 0, sb = solid[ib] >= 0;
    double field = (a - b) / len;
    if (near[ia] && sb) field = (a - b) / (dist_hi[ia] * len);
    if (near[ib] && sa) field = (a - b) / (dist_lo[ib] * len);
    if (sa && sb) {
        const int near0 = ta >= 1 ? near[ia - st] != 0 : 0;
        const int near3 = tb + 1 <= n - 1 ? near[ib + st] != 0 : 0;
        const double phi0 = ta >= 1 ? phi[ia - st] : 0.0;
        const double phi3 = tb + 1 <= n - 1 ? phi[ib + st] : 0.0;
        const double d0 = ta >= 1 ? dist_hi[ia - st] : 1.0;
        const double d3 = tb + 1 <= n - 1 ? dist_lo[ib + st] : 1.0;
        double inside = 0.0;
        if (near0 && !near3) inside = (phi0 - a) / (d0 * len_before);
        if (near3 && !near0) inside = (b - phi3) / (d3 * len_after);
        field = inside;
    }
    return field;
}

void oracle_edge_field(const double* phi, const int32_t* solid, const uint8_t* near,
                       const double* dist_lo, const double* dist_hi, const double* lengths,
                       i64 n, i64 m, i64 st, i64 sc, i64 symmetric_lo, i64 symmetric_hi,
                       double* out, i64 ot, i64 oc) {
    for (i64 c = 0; c < m; ++c) {
        const i64 node0 = c * sc, edge0 = c * oc;
        for (i64 e = 0; e <= n; ++e) out[e * ot + edge0] = 0.0;
        for (i64 e = 1; e <= n - 1; ++e)
            out[e * ot + edge0] = oracle_edge_interior(phi + node0, solid + node0, near + node0, dist_lo + node0,
                                                       dist_hi + node0, lengths, n, st, 0, e);
        /* the two end faces, in this order (the second may read the first when n == 2) */
        out[edge0] = symmetric_lo ? -out[ot + edge0] : 2.0 * out[ot + edge0] - out[2 * ot + edge0];
        out[n * ot + edge0] = symmetric_hi ? -out[(n - 1) * ot + edge0]
                                           : 2.0 * out[(n - 1) * ot + edge0] - out[(n - 2) * ot + edge0];
    }
}

/* ---- the 3D tracer (trace_trajectories_3d_pyccel.py:18-657) -------------------------- */

/* Every argument of trace_trajectories_3d_pyccel.py:18-60, flattened to pointers and
 * extents. */
typedef struct {
    const i64* active; i64 n_active;
    double* x; double* v; i64 n;
    const double* qm; const double* dq;
    i64* steps; i64* status; i64* electrode;
    const double* fxs; const double* fys; const double* fzs;
    const i64* patch_int; const double* patch_real; i64 n_patches;
    const double* bn; i64 has_b;
    double origin_x, origin_y, origin_z, h, upper_x, upper_y, upper_z;
    i64 mirror_y, mirror_z, integrator;
    double step_fraction;
    i64 max_steps;
    const i64* cells; i64 cells_n1, cells_n2;
    const i64* subcells; i64 refine;
    const double* planes; i64 n_planes;
    double* hits; i64 capacity; i64* counts;
    const i64* path_slot; double* paths; i64 path_capacity; i64* path_count;
    i64 num_threads; double* charge; i64 charge_stride;
    i64 has_j; double* current; i64 current_stride;
    double inv_c2;
} oracle_trace3d_args;

i64 oracle_trace3d_args_size(void) { return (i64)sizeof(oracle_trace3d_args); }

/* the index floor(s) clamped to [lo, hi] (pyccel.py:225-268) */
static i64 clamp_floor3(double s, i64 lo, i64 hi) {
    i64 i = (i64)floor(s);
    if (i < lo) i = lo; else if (i > hi) i = hi;
    return i;
}

/* the trilinear sum of pyccel.py:237-241: f at ci and its +1, +dj, +di neighbours */
static double trilinear3(const double* f, i64 ci, i64 di, i64 dj, double a, double b, double c) {
    return (1.0 - a) * ((1.0 - b) * ((1.0 - c) * f[ci] + c * f[ci + 1])
                        + b * ((1.0 - c) * f[ci + dj] + c * f[ci + dj + 1]))
         + a * ((1.0 - b) * ((1.0 - c) * f[ci + di] + c * f[ci + di + 1])
                + b * ((1.0 - c) * f[ci + di + dj] + c * f[ci + di + dj + 1]));
}

/* the finest patch containing (p0, p1, p2), else the base mesh (pyccel.py:196-206) */
static i64 patch3(const oracle_trace3d_args* a, double p0, double p1, double p2) {
    for (i64 qq = 0; qq < a->n_patches - 1; ++qq) {
        const double* r = a->patch_real + qq * 10;
        if (p0 >= r[4] && p0 <= r[7] && p1 >= r[5] && p1 <= r[8] && p2 >= r[6] && p2 <= r[9]) return qq;
    }
    return a->n_patches - 1;
}

/* E (and B) at stage point q, mirrored below a symmetry plane (pyccel.py:182-318) */
static void fields3(const oracle_trace3d_args* a, double q0, double q1, double q2, double* e, double* b) {
    double p0 = q0, p1 = q1, p2 = q2, s1 = 1.0, s2 = 1.0;
    if (a->mirror_y != 0 && p1 < a->origin_y) { p1 = 2.0 * a->origin_y - p1; s1 = -1.0; }
    if (a->mirror_z != 0 && p2 < a->origin_z) { p2 = 2.0 * a->origin_z - p2; s2 = -1.0; }
    const i64 pp = patch3(a, p0, p1, p2);
    const i64* pi = a->patch_int + pp * 8;
    const double* pr = a->patch_real + pp * 10;
    const i64 pnx = pi[0], pny = pi[1], pnz = pi[2];
    const double phh = pr[3];
    const double sx = (p0 - pr[0]) / phh, sy = (p1 - pr[1]) / phh, sz = (p2 - pr[2]) / phh;
    /* E_x on (pnx + 1, pny, pnz) at (sx + 1/2, sy, sz) */
    double s = sx + 0.5;
    i64 i = clamp_floor3(s, 0, pnx - 1), j = clamp_floor3(sy, 0, pny - 2), kz = clamp_floor3(sz, 0, pnz - 2);
    double wa = s - (double)i, wb = sy - (double)j, wc = sz - (double)kz;
    i64 dj = pnz, di = pny * pnz;
    e[0] = trilinear3(a->fxs, pi[3] + i * di + j * dj + kz, di, dj, wa, wb, wc);
    /* E_y on (pnx, pny + 1, pnz) at (sx, sy + 1/2, sz): kz and c as for E_x */
    i = clamp_floor3(sx, 0, pnx - 2);
    s = sy + 0.5;
    j = clamp_floor3(s, 0, pny - 1);
    wa = sx - (double)i;
    wb = s - (double)j;
    dj = pnz;
    di = (pny + 1) * pnz;
    e[1] = s1 * trilinear3(a->fys, pi[4] + i * di + j * dj + kz, di, dj, wa, wb, wc);
    /* E_z on (pnx, pny, pnz + 1) at (sx, sy, sz + 1/2): i and a as for E_y */
    j = clamp_floor3(sy, 0, pny - 2);
    s = sz + 0.5;
    kz = clamp_floor3(s, 0, pnz - 1);
    wb = sy - (double)j;
    wc = s - (double)kz;
    dj = pnz + 1;
    di = pny * (pnz + 1);
    e[2] = s2 * trilinear3(a->fzs, pi[5] + i * di + j * dj + kz, di, dj, wa, wb, wc);
    b[0] = b[1] = b[2] = 0.0;
    if (a->has_b != 0) {  /* B on the base mesh's nodes (3, nx, ny, nz) */
        const i64* bi = a->patch_int + (a->n_patches - 1) * 8;
        const i64 nx = bi[0], ny = bi[1], nz = bi[2];
        const double bx = (p0 - a->origin_x) / a->h, by = (p1 - a->origin_y) / a->h, bz = (p2 - a->origin_z) / a->h;
        i = clamp_floor3(bx, 0, nx - 2);
        j = clamp_floor3(by, 0, ny - 2);
        kz = clamp_floor3(bz, 0, nz - 2);
        wa = bx - (double)i;
        wb = by - (double)j;
        wc = bz - (double)kz;
        const double w000 = (1.0 - wa) * (1.0 - wb) * (1.0 - wc), w001 = (1.0 - wa) * (1.0 - wb) * wc;
        const double w010 = (1.0 - wa) * wb * (1.0 - wc), w011 = (1.0 - wa) * wb * wc;
        const double w100 = wa * (1.0 - wb) * (1.0 - wc), w101 = wa * (1.0 - wb) * wc;
        const double w110 = wa * wb * (1.0 - wc), w111 = wa * wb * wc;
        const i64 nodes = nx * ny * nz, si = ny * nz, sj = nz;
        for (int n0 = 0; n0 < 3; ++n0) {
            const double* f = a->bn + n0 * nodes + i * si + j * sj + kz;
            const double v = w000 * f[0] + w001 * f[1] + w010 * f[sj] + w011 * f[sj + 1]
                           + w100 * f[si] + w101 * f[si + 1] + w110 * f[si + sj] + w111 * f[si + sj + 1];
            if (n0 == 0) b[0] = s1 * s2 * v;
            else if (n0 == 1) b[1] = s2 * v;
            else b[2] = s1 * v;
        }
    }
}

/* a node index floor(s) with its fraction, clamped for a deposit (pyccel.py:453-468) */
static void deposit_cell3(double s, i64 n, i64* i, double* w) {
    i64 k = (i64)floor(s);
    double f = s - (double)k;
    if (k < 0) { f = 0.0; k = 0; }
    else if (k >= n - 1) { f = 1.0; k = n - 2; }
    *i = k;
    *w = f;
}

/* the 8 trilinear node shares of value at node cc (pyccel.py:472-480) */
static void deposit8(double* row, i64 cc, i64 di, i64 dj, double a, double b, double c, double value) {
    row[cc] += (1.0 - a) * (1.0 - b) * (1.0 - c) * value;
    row[cc + 1] += (1.0 - a) * (1.0 - b) * c * value;
    row[cc + dj] += (1.0 - a) * b * (1.0 - c) * value;
    row[cc + dj + 1] += (1.0 - a) * b * c * value;
    row[cc + di] += a * (1.0 - b) * (1.0 - c) * value;
    row[cc + di + 1] += a * (1.0 - b) * c * value;
    row[cc + di + dj] += a * b * (1.0 - c) * value;
    row[cc + di + dj + 1] += a * b * c * value;
}

static void trace3d_one(const oracle_trace3d_args* a, i64 k, double* charge_row, double* current_row) {
    const i64 N = a->n;
    const i64* base = a->patch_int + (a->n_patches - 1) * 8;
    const i64 nx = base[0], ny = base[1], nz = base[2];
    double x0 = a->x[k], x1 = a->x[N + k], x2 = a->x[2 * N + k];
    double v0 = a->v[k], v1 = a->v[N + k], v2 = a->v[2 * N + k];
    const double qmk = a->qm[k], inv_c2 = a->inv_c2;
    const i64 slot = a->path_slot[k];
    i64 ns = a->steps[k], st = -1;
    double dt = 0.0, hdt = 0.0, dt6 = 0.0, tau = 0.0;
    const int nstage = a->integrator == 0 ? 4 : 2;
    int have = 0;
    double ce[3] = {0.0, 0.0, 0.0}, cb[3] = {0.0, 0.0, 0.0};
    while (ns < a->max_steps) {
        /* the spacing of the finest patch at the step's start (pyccel.py:149-163) */
        double p1 = x1, p2 = x2;
        if (a->mirror_y != 0 && p1 < a->origin_y) p1 = 2.0 * a->origin_y - p1;
        if (a->mirror_z != 0 && p2 < a->origin_z) p2 = 2.0 * a->origin_z - p2;
        double hloc = a->h;
        for (i64 qq = 0; qq < a->n_patches - 1; ++qq) {
            const double* r = a->patch_real + qq * 10;
            if (x0 >= r[4] && x0 <= r[7] && p1 >= r[5] && p1 <= r[8] && p2 >= r[6] && p2 <= r[9]) {
                hloc = r[3];
                break;
            }
        }
        const double gv = 1.0 / lorentz_factor(v0, v1, v2, inv_c2);
        double q0 = x0, q1 = x1, q2 = x2, kx0 = v0, kx1 = v1, kx2 = v2;
        double sx0 = 0.0, sx1 = 0.0, sx2 = 0.0, sv0 = 0.0, sv1 = 0.0, sv2 = 0.0;
        double w0 = 0.0, w1 = 0.0, w2 = 0.0, y0, y1, y2;
        for (int stage = 0; stage < nstage; ++stage) {
            double e[3], b[3];
            if (stage == 0 && have != 0) {
                e[0] = ce[0]; e[1] = ce[1]; e[2] = ce[2];
                b[0] = cb[0]; b[1] = cb[1]; b[2] = cb[2];
            } else {
                fields3(a, q0, q1, q2, e, b);
            }
            /* pyccel.py:320-324 */
            double gk = 1.0 / lorentz_factor(kx0, kx1, kx2, inv_c2);
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
                dt = a->step_fraction * dt;
                hdt = 0.5 * dt;
                dt6 = dt / 6.0;
                if (a->integrator != 0) {  /* Boris: a half step's kick-rotate-kick, the drift */
                    tau = 0.5 * hdt * qmk;
                    kx0 = v0; kx1 = v1; kx2 = v2;
                    boris_kick(&kx0, &kx1, &kx2, e[0], e[1], e[2], b[0], b[1], b[2], hdt, qmk, inv_c2);
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
            } else if (a->integrator != 0) {  /* the second half step's kick at the end */
                w0 = kx0; w1 = kx1; w2 = kx2;
                boris_kick(&w0, &w1, &w2, e[0], e[1], e[2], b[0], b[1], b[2], hdt, qmk, inv_c2);
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
        (void)tau;
        if (a->integrator == 0) {
            y0 = x0 + dt6 * sx0; y1 = x1 + dt6 * sx1; y2 = x2 + dt6 * sx2;
            w0 = v0 + dt6 * sv0; w1 = v1 + dt6 * sv1; w2 = v2 + dt6 * sv2;
        } else {
            y0 = q0; y1 = q1; y2 = q2;
        }
        /* space charge: dq dt at the Hermite midpoint, trilinear (pyccel.py:432-480) */
        const double gw = 1.0 / lorentz_factor(w0, w1, w2, inv_c2);
        const double m0 = 0.5 * (x0 + y0) + dt * (gv * v0 - gw * w0) / 8.0;
        double m1 = 0.5 * (x1 + y1) + dt * (gv * v1 - gw * w1) / 8.0;
        double m2 = 0.5 * (x2 + y2) + dt * (gv * v2 - gw * w2) / 8.0;
        if (a->mirror_y != 0 && m1 < a->origin_y) m1 = 2.0 * a->origin_y - m1;
        if (a->mirror_z != 0 && m2 < a->origin_z) m2 = 2.0 * a->origin_z - m2;
        {
            const i64 pp = patch3(a, m0, m1, m2);
            const i64* pi = a->patch_int + pp * 8;
            const double* pr = a->patch_real + pp * 10;
            const double phh = pr[3];
            i64 i, j, kz;
            double wa, wb, wc;
            deposit_cell3((m0 - pr[0]) / phh, pi[0], &i, &wa);
            deposit_cell3((m1 - pr[1]) / phh, pi[1], &j, &wb);
            deposit_cell3((m2 - pr[2]) / phh, pi[2], &kz, &wc);
            const i64 dj = pi[2], di = pi[1] * pi[2];
            deposit8(charge_row, pi[7] + i * di + j * dj + kz, di, dj, wa, wb, wc, a->dq[k] * dt);
        }
        if (a->has_j != 0) {  /* the current element dq (x' - x) on the base mesh (pyccel.py:481-540) */
            double jd[3] = {a->dq[k] * (y0 - x0), a->dq[k] * (y1 - x1), a->dq[k] * (y2 - x2)};
            if (a->mirror_y != 0 && 0.5 * (x1 + y1) < a->origin_y) jd[1] = -jd[1];
            if (a->mirror_z != 0 && 0.5 * (x2 + y2) < a->origin_z) jd[2] = -jd[2];
            i64 bi, bj, bk;
            double ba, bb, bc;
            deposit_cell3((m0 - a->origin_x) / a->h, nx, &bi, &ba);
            deposit_cell3((m1 - a->origin_y) / a->h, ny, &bj, &bb);
            deposit_cell3((m2 - a->origin_z) / a->h, nz, &bk, &bc);
            const i64 nb = nx * ny * nz, dj = nz, di = ny * nz;
            for (int jc = 0; jc < 3; ++jc)
                deposit8(current_row, jc * nb + bi * di + bj * dj + bk, di, dj, ba, bb, bc, jd[jc]);
        }
        ns = ns + 1;
        /* plane crossings (pyccel.py:543-559) */
        for (i64 p = 0; p < a->n_planes; ++p) {
            const double xp = a->planes[p];
            const double before = x0 - xp, after = y0 - xp;
            if ((before < 0.0) != (after < 0.0)) {
                const double s = before / (before - after);
                const i64 idx = k * a->n_planes + p;
                const i64 cc = a->counts[idx];
                if (cc < a->capacity) {
                    double* rec = a->hits + (idx * a->capacity + cc) * 8;
                    rec[0] = xp;
                    rec[1] = x1 + s * (y1 - x1);
                    rec[2] = x2 + s * (y2 - x2);
                    rec[3] = gv * v0 + s * (gw * w0 - gv * v0);
                    rec[4] = gv * v1 + s * (gw * w1 - gv * v1);
                    rec[5] = gv * v2 + s * (gw * w2 - gv * v2);
                    rec[6] = after >= 0.0 ? 1.0 : -1.0;
                    rec[7] = 1.0 * (double)ns;
                }
                a->counts[idx] = cc + 1;
            }
        }
        x0 = y0; x1 = y1; x2 = y2;
        v0 = w0; v1 = w1; v2 = w2;
        /* the end of the step: off the mesh or in an electrode (pyccel.py:567-624) */
        const double px = x0;
        double py = x1, pz = x2;
        if (a->mirror_y != 0 && py < a->origin_y) py = 2.0 * a->origin_y - py;
        if (a->mirror_z != 0 && pz < a->origin_z) pz = 2.0 * a->origin_z - pz;
        if (slot >= 0) {
            const i64 cc = a->path_count[slot];
            if (cc < a->path_capacity) {
                double* rec = a->paths + (slot * a->path_capacity + cc) * 3;
                rec[0] = px; rec[1] = py; rec[2] = pz;
            }
            a->path_count[slot] = cc + 1;
        }
        if (px >= a->upper_x) {
            st = 0;
        } else if (px < a->origin_x) {
            st = 1;
        } else if (py > a->upper_y || py < a->origin_y || pz > a->upper_z || pz < a->origin_z) {
            st = 2;
        } else {
            i64 i = (i64)((px - a->origin_x) / a->h);
            if (i > nx - 2) i = nx - 2;
            i64 j = (i64)((py - a->origin_y) / a->h);
            if (j > ny - 2) j = ny - 2;
            i64 kz = (i64)((pz - a->origin_z) / a->h);
            if (kz > nz - 2) kz = nz - 2;
            const i64 cell = a->cells[(i * a->cells_n1 + j) * a->cells_n2 + kz];
            if (cell >= 0) {
                const i64 R = a->refine;
                i64 si = (i64)(((px - a->origin_x) / a->h - (double)i) * (double)R);
                if (si < 0) si = 0; else if (si > R - 1) si = R - 1;
                i64 sj = (i64)(((py - a->origin_y) / a->h - (double)j) * (double)R);
                if (sj < 0) sj = 0; else if (sj > R - 1) sj = R - 1;
                i64 sk = (i64)(((pz - a->origin_z) / a->h - (double)kz) * (double)R);
                if (sk < 0) sk = 0; else if (sk > R - 1) sk = R - 1;
                const i64 e = a->subcells[((cell * R + si) * R + sj) * R + sk];
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
    a->x[k] = x0; a->x[N + k] = x1; a->x[2 * N + k] = x2;
    a->v[k] = v0; a->v[N + k] = v1; a->v[2 * N + k] = v2;
    a->steps[k] = ns;
    a->status[k] = st;
}

/* pyccel.py:128-135: thread t traces every num_threads-th active trajectory into its own
 * slabs charge[t] and current[t]. */
void oracle_trace_trajectories_3d(const oracle_trace3d_args* a) {
    const i64 T = a->num_threads > 0 ? a->num_threads : 1;
#pragma omp parallel for schedule(static, 1) num_threads(T)
    for (i64 t = 0; t < T; ++t) {
        for (i64 m = t; m < a->n_active; m += T) {
            trace3d_one(a, a->active[m], a->charge + t * a->charge_stride,
                        a->current + (a->has_j != 0 ? t * a->current_stride : 0));
        }
    }
}
