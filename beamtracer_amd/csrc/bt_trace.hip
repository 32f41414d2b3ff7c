// The fused particle tracer for gfx950: trace_trajectories_2d.
//
// Behaviour: kernels/trace_trajectories_2d/trace_trajectories_2d_pyccel.py:42-507 (and its
// CUDA twin trace_trajectories_2d_cuda.cu:165-362): every active trajectory is stepped
// until it leaves the mesh, enters an electrode sub-cell (status 3), stops in a sub-cell
// that an electrode surface cuts (status -2, the caller decides), or reaches max_steps
// (status 4).  Each step takes the classical Runge--Kutta step (integrator 0, four field
// evaluations) or Boris's kick--drift--kick (1, one evaluation: the end field is the next
// start's), deposits dq dt at the Hermite midpoint, records plane crossings and path
// points, and tests the end point.  The operations are the reference's, in its order (the
// library is built with -ffp-contract=off, so only the fusions written out below are fused),
// with these changes of rounding, each a few units in the last place against the CPU
// restatement's 1e-10 bar:
// * hypot, the step-size rule's square roots and the divisions by r, |v| and sqrt|a| go
//   through one hardware reciprocal square root with a third-order correction (rsqrt_nr,
//   radial, step_rule in bt_device.h); the cylindrical deposit weight's division through a
//   corrected reciprocal (recip_nr); divisions by the mesh spacing and by 6 are products
//   with their reciprocals;
// * fused multiply-adds (__builtin_fma) in the RK4 stage points and velocities
//   (x + c k, v + c g), the weighted stage sums, the bilinear interpolation of the gather
//   and the Hermite midpoint; the adjoint's forward recompute fuses the same operations,
//   so the two stay bit-identical.
//
// MI355X design:
// * persistent grid, one trajectory per lane, refilled from a device work queue: a
//   wave whose lanes finish claims that many new trajectories with one atomic, so lanes
//   freed by short trajectories pick up the remaining ones;
// * state in registers (x, v component rows are coalesced loads / stores); the launch's
//   rarely used fields are read from the kernel-argument segment where they are used, so
//   the step loop's scalar registers hold only what every step needs;
// * the FP64 edge fields (2.6 MB for an 801 x 201 mesh) stay resident in each XCD's 4 MB
//   L2 and are read through buffer resources, one 16-byte load per row pair;
// * the deposit is exact fixed point (deterministic): each lane sums its trajectory's
//   deposits into the four nodes of its cell in registers, each wave sums its lanes' node
//   sums in an LDS window of 16 (or 32) columns that follows its lanes (integer LDS
//   atomics, no barriers: the window is the wave's own), and a window column that the wave
//   leaves is added to global memory once.  Global integer atomics on this GPU are executed
//   at the memory side, where same-node adds from every wave serialize; the window turns
//   one atomic per deposit into one per wave and node.  A finalize kernel converts the
//   node words and adds them to charge row 0;
// * compile-time variants on integrator, geometry, magnetic field and the general
//   features (refined patches, graded axis, relativistic motion), so the common case
//   carries no dead branches.
//
// Extension (not in the reference's compiled kernel, which refuses B): a node magnetic
// field acting through q/m (E + u/gamma x B) (solver/tracer.py:519-529) in every RK4
// stage, and through Boris's rotation (tracer.py:618-650) for integrator 1.
#include <hip/hip_runtime.h>

#include "bt_device.h"
#include "bt_internal.h"

namespace bt {

struct LaneState {
    double x0, x1, x2, v0, v1, v2;
    double qmk, dqk;
    double ca0, ca1, ca2;  // Boris (E only): the last end's acceleration
    double ce0, ce1, ce2, cb0, cb1, cb2;  // with B: the last end's E and B
    i64 k, slot, ns;
    int have;
    // the mesh coordinate across the axis, the radial direction and the cell coordinates of
    // the current point, as the last step's end test computed them (valid when `point` is set)
    double pt_y, pt_cy, pt_cz, pt_sx, pt_sy;
    int point;
    // the lane's deposit cache: the FP64 sums of the trajectory's deposits into the four nodes
    // of base-mesh cell (cci, ccj) since it reached each of them, added in step order;
    // slot (a << 1) | b holds the cell's node whose row and column have parities a and b, so a
    // node the next cell shares keeps its slot.  A sum leaves the lane -- rounded once to the
    // fixed-point units 2^e -- when the trajectory moves to a cell without that node or ends.
    // Which deposits a sum holds and their order depend on the trajectory alone, so the node
    // totals do not depend on the launch shape or on the order of the lanes.
    int cci, ccj;
    double cq[4];
    i64 ck_next, ck_index;  // the step count of the next checkpoint and its slot
    // RK4 on one patch: the first-stage gathers of the next step at the current point (E_x, E_y),
    // issued by the last step's end so that their loads run under its deposit (valid with
    // `point`)
    Bilinear pgx, pgy;
};

constexpr int kNoCell = -(1 << 30);
constexpr i64 kWindowLimit = (i64)1 << 46;  // a sum added to a window slot (the slots' bound)

// Node-field bilinear, constant beyond the mesh (solver/self_field.py:100-113): the point is
// clamped to the nodes, the cell to the last one, and the four nodes weighted as
// (1 - a)((1 - b) f00 + b f01) + a((1 - b) f10 + b f11).  The cell and weights are the same for
// the three components of a node field, so they are formed once (in 32 bits: a mesh has fewer
// than 2^24 nodes per axis), and each component reads its two node pairs (f00, f01) and
// (f10, f11) as two 16-byte loads of consecutive nodes.
struct NodeCell {
    unsigned c;  // the cell's first node, i ny + j
    double a, b;
};

__device__ __forceinline__ NodeCell node_cell(double sx, double sy, int nx, int ny) {
    sx = fmin(fmax(sx, 0.0), (double)(nx - 1));
    sy = fmin(fmax(sy, 0.0), (double)(ny - 1));
    int i = (int)sx;
    if (i > nx - 2) i = nx - 2;
    int j = (int)sy;
    if (j > ny - 2) j = ny - 2;
    NodeCell n;
    n.a = sx - (double)i;
    n.b = sy - (double)j;
    n.c = mad24((unsigned)i, (unsigned)ny, (unsigned)j);
    return n;
}

__device__ __forceinline__ double node_bilinear(const double* __restrict__ f, const NodeCell& n, int nx, int ny) {
    if (f == nullptr) return 0.0;
    const Rsrc r = field_rsrc(f, (i64)nx * ny);
    const Pair lo = load_pair(r, n.c), hi = load_pair_at(r, n.c, (unsigned)ny);
    const double a = n.a, b = n.b;
    return (1.0 - a) * ((1.0 - b) * lo.a + b * lo.b) + a * ((1.0 - b) * hi.a + b * hi.b);
}

template <int GEOM>
__device__ __forceinline__ void magnetic_at(const TraceParams& p, double lq, double qy, double cy,
                                            double cz, double& B0, double& B1, double& B2) {
    // (the node coordinates as products with RN(1 / h), as the edge field's on one patch)
    const double sx = axis_coord(lq, p.origin_x, p.inv_h);
    const double sy = axis_coord(qy, p.origin_y, p.inv_h);
    const NodeCell n = node_cell(sx, sy, (int)p.nx, (int)p.ny);
    const double c0 = node_bilinear(p.b0, n, (int)p.nx, (int)p.ny);
    const double c1 = node_bilinear(p.b1, n, (int)p.nx, (int)p.ny);
    const double c2 = node_bilinear(p.b2, n, (int)p.nx, (int)p.ny);
    if (GEOM == 1) {
        B0 = c0; B1 = c1 * cy - c2 * cz; B2 = c1 * cz + c2 * cy;
    } else if (GEOM == 2) {
        B0 = cy * c0; B1 = c1; B2 = cy * c2;
    } else {
        B0 = c0; B1 = c1; B2 = c2;
    }
}

// Boris's velocity update over tau: half kick, rotation, half kick (tracer.py:618-632).
__device__ __forceinline__ void boris_kick(double& u0, double& u1, double& u2, double e0, double e1,
                                           double e2, double b0, double b1, double b2, double tau,
                                           double qm, double inv_c2) {
    const double hq = 0.5 * tau * qm;
    const double m0 = u0 + hq * e0, m1 = u1 + hq * e1, m2 = u2 + hq * e2;
    // 1 / gamma; classical (inv_c2 = 0): sqrt(1 + |m|^2 0) is exactly 1 for a finite |m|^2 and
    // NaN otherwise, which the compare gives without the square root and the division
    double g;
    if (inv_c2 == 0.0) g = (m0 * m0 + m1 * m1 + m2 * m2) < __builtin_inf() ? 1.0 : __builtin_nan("");
    else g = 1.0 / lorentz_factor(m0, m1, m2, inv_c2);
    const double t0 = hq * b0 * g, t1 = hq * b1 * g, t2 = hq * b2 * g;
    // (2 / (1 + |t|^2) as a product with the corrected reciprocal: 1 + |t|^2 >= 1 is normal)
    const double f = 2.0 * recip_nr(1.0 + (t0 * t0 + t1 * t1 + t2 * t2));
    const double p0 = m0 + (m1 * t2 - m2 * t1);
    const double p1 = m1 + (m2 * t0 - m0 * t2);
    const double p2 = m2 + (m0 * t1 - m1 * t0);
    const double s0 = f * t0, s1 = f * t1, s2 = f * t2;
    const double n0 = m0 + (p1 * s2 - p2 * s1);
    const double n1 = m1 + (p2 * s0 - p0 * s2);
    const double n2 = m2 + (p0 * s1 - p1 * s0);
    u0 = n0 + hq * e0; u1 = n1 + hq * e1; u2 = n2 + hq * e2;
}

// The finest patch containing (px, py); the base mesh (last) if none (pyccel.py:215-222).
template <bool GENERAL>
__device__ __forceinline__ int patch_of(const TraceParams& p, double px, double py) {
    if (!GENERAL) return p.n_patches - 1;
    for (int q = 0; q < p.n_patches - 1; ++q) {
        const double* pr = p.patch_real + q * 7;
        if (px >= pr[3] && px <= pr[5] && py >= pr[4] && py <= pr[6]) return q;
    }
    return p.n_patches - 1;
}

struct PatchView {
    i64 nx, ny, ofx, ofy, ofn;
    double ox, oy, h;
};

template <bool GENERAL>
__device__ __forceinline__ PatchView patch_view(const TraceParams& p, int pp) {
    PatchView v;
    if (!GENERAL) {  // the one patch is the mesh (same origin and spacing)
        v.nx = p.nx; v.ny = p.ny; v.ofx = p.bofx; v.ofy = p.bofy; v.ofn = p.bofn;
        v.ox = p.origin_x; v.oy = p.origin_y; v.h = p.h;
        return v;
    }
    const i64* pi = p.patch_int + (i64)pp * 6;
    const double* pr = p.patch_real + (i64)pp * 7;
    v.nx = pi[0]; v.ny = pi[1]; v.ofx = pi[2]; v.ofy = pi[3]; v.ofn = pi[5];
    v.ox = pr[0]; v.oy = pr[1]; v.h = pr[2];
    return v;
}

template <bool GENERAL>
__device__ __forceinline__ double locate_x(const TraceParams& p, double x) {
    if (GENERAL && p.graded != 0) return uniform_x(x, p.x_nodes, p.n_x_nodes, p.origin_x, p.h);
    return x;
}

template <bool GENERAL>
__device__ __forceinline__ double lorentz(double u0, double u1, double u2, double inv_c2) {
    if (!GENERAL) return 1.0;
    return lorentz_factor(u0, u1, u2, inv_c2);
}

// The two edge-field buffers of a launch (uniform).
struct Fields {
    Rsrc fx, fy;
};

// The launch's parameters in the kernel-argument segment (the kernels' first argument), read
// through a pointer that the compiler cannot follow across an empty asm: fields that only rare
// paths use (a refill, the end of a trajectory, a plane record, an electrode sub-cell) are
// then loaded where they are used (scalar loads) instead of holding scalar registers through
// the whole step loop, where they were spilled.
typedef const __attribute__((address_space(4))) TraceParams* KParams;
__device__ __forceinline__ KParams kernel_params() { return (KParams)__builtin_amdgcn_kernarg_segment_ptr(); }
__device__ __forceinline__ KParams cold(KParams k) {
    asm volatile("" : "+s"(k));
    return k;
}

// E_x, E_y of the finest patch at the Cartesian point q, with the radial direction and
// the mesh coordinates (pyccel.py:196-269).
// With `known` (one patch that is the mesh), qy, cy, cz, sx and sy are the values the end
// test of the last step computed for this same point: the same operations on the same
// operands, so the same bits.  Otherwise they are computed and returned.  With `pre` as
// well, the two gathers were issued then too (the same operations: `pgx`, `pgy`).
// The value of `v` in the other lane of a pair (lanes 2t and 2t + 1; both active).
__device__ __forceinline__ double pair_swap(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0xB1, 0xF, 0xF, false);  // quad_perm [1, 0, 3, 2]
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0xB1, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Two lanes per trajectory: this lane's half of edge_field's two gathers on the one patch that
// is the mesh -- E_x in even lanes, E_y in odd ones, each with the operations of that gather
// alone -- issued; pair_fields then gives both lanes E_x and E_y.
__device__ __forceinline__ Bilinear issue_pair_gather(const TraceParams& p, const Fields& F, double sx, double sy) {
    const bool odd = (__lane_id() & 1u) != 0;
    const PatchView v = patch_view<false>(p, p.n_patches - 1);
    const double s = odd ? sx : sx + 0.5;
    const double t = odd ? sy + 0.5 : sy;
    const int imax = odd ? (int)v.nx - 2 : (int)v.nx - 1;
    const int jmax = odd ? (int)v.ny - 1 : (int)v.ny - 2;
    const unsigned cols = (unsigned)(odd ? v.ny + 1 : v.ny);
    const unsigned base = odd ? (unsigned)(p.fy_base + v.ofy) : (unsigned)(p.fx_base + v.ofx);
    Bilinear b;
    const int i = clamped_floor_index<false>(s, imax);
    const int j = clamped_floor_index<false>(t, jmax);
    b.u = s - (double)i;
    b.w = t - (double)j;
    const unsigned c = mad24((unsigned)i, cols, base + (unsigned)j);
    if (odd) {
        b.lo = load_pair(F.fy, c);
        b.hi = load_pair(F.fy, c + cols);
    } else {
        b.lo = load_pair(F.fx, c);
        b.hi = load_pair(F.fx, c + cols);
    }
    return b;
}

__device__ __forceinline__ void pair_fields(const Bilinear& b, double& ex, double& ey) {
    const bool odd = (__lane_id() & 1u) != 0;
    const double mine = bilinear_value(b), other = pair_swap(mine);
    ex = odd ? other : mine;
    ey = odd ? mine : other;
}

template <int GEOM, bool GENERAL, int LPT = 1>
__device__ __forceinline__ void edge_field(const TraceParams& p, const Fields& F, double q0, double q1, double q2,
                                           double& ex, double& ey, double& cy, double& cz,
                                           double& lq, double& qy, bool known, double& sx, double& sy,
                                           bool pre = false, const Bilinear* pgx = nullptr,
                                           const Bilinear* pgy = nullptr) {
    static_assert(LPT == 1 || !GENERAL, "two lanes per trajectory only on one uniform patch");
    if (!known) {
        cy = 0.0; cz = 0.0; qy = q1;
        if (GEOM == 1) {
            const Radial rd = radial(q1, q2);
            qy = rd.r; cy = rd.cy; cz = rd.cz;
        }
    }
    if (GEOM == 2) {
        cy = 1.0; cz = 0.0;
        if (q1 < p.origin_y) cy = -1.0;
        if (!known) qy = p.origin_y + fabs(q1 - p.origin_y);
    }
    lq = locate_x<GENERAL>(p, q0);
    const PatchView v = patch_view<GENERAL>(p, patch_of<GENERAL>(p, lq, qy));
    if (!known) {
        if (GENERAL) {
            sx = (lq - v.ox) / v.h;
            sy = (qy - v.oy) / v.h;
        } else {
            sx = axis_coord(lq, v.ox, p.inv_h);
            sy = axis_coord(qy, v.oy, p.inv_h);
        }
    }
    if (LPT == 2) {
        pair_fields(pre ? *pgx : issue_pair_gather(p, F, sx, sy), ex, ey);
        return;
    }
    if (!GENERAL && pre) {
        ex = bilinear_value(*pgx);
        ey = bilinear_value(*pgy);
        return;
    }
    ex = clamped_bilinear_buf<!GENERAL>(F.fx, (unsigned)(p.fx_base + v.ofx), v.ny, sx + 0.5, sy, (int)v.nx - 1,
                                        (int)v.ny - 2);
    ey = clamped_bilinear_buf<!GENERAL>(F.fy, (unsigned)(p.fy_base + v.ofy), v.ny + 1, sx, sy + 0.5, (int)v.nx - 2,
                              (int)v.ny - 1);
}

// The first-stage gathers of the next step at the end point, with its cell coordinates (sx, sy)
// on the one patch that is the mesh (edge_field's gathers, issued now: RK4 only; a pair's
// lane issues its half into gx).
template <int LPT = 1>
__device__ __forceinline__ void issue_first_stage(const TraceParams& p, const Fields& F, double sx, double sy,
                                                  Bilinear& gx, Bilinear& gy) {
    if (LPT == 2) {
        gx = issue_pair_gather(p, F, sx, sy);
        return;
    }
    const PatchView v = patch_view<false>(p, p.n_patches - 1);
    gx = bilinear_issue<true>(F.fx, (unsigned)(p.fx_base + v.ofx), v.ny, sx + 0.5, sy, (int)v.nx - 1, (int)v.ny - 2);
    gy = bilinear_issue<true>(F.fy, (unsigned)(p.fy_base + v.ofy), v.ny + 1, sx, sy + 0.5, (int)v.nx - 2,
                              (int)v.ny - 1);
}

// ---- the deposit window ---------------------------------------------------------------
// Each wave sums its lanes' deposits in its own window of node sums in LDS that follows it:
// columns lo .. lo + WC - 1 (column c in slot c mod WC) and rows base .. base + 63 of one
// design's mesh, one signed 64-bit word per node.  The placement is wave-uniform (scalar
// registers), so a flush from a lane tests the node with two unsigned compares and adds with
// a no-return LDS atomic; a node outside goes to global memory directly.  There is no barrier
// anywhere in the step loop: only the wave itself reads, adds to and moves its window, and a
// wave's LDS operations execute in program order.  After each round of kRoundSteps steps the
// wave reads the cache cells of its first and last lanes that hold one (its lanes took
// consecutive queue entries: neighbours at the emitter) and, when they leave the window with
// room for a round of travel, moves it: the columns that leave are flushed (lane l adds row
// l), and the whole window when the rows or the design change (rare: the rows sit centred on
// the lanes) and at least every kFlushRounds rounds.

constexpr int kWinCols = 16;   // the columns of a window (a wide window: kWideCols)
constexpr int kWideCols = 32;
constexpr int kWinRows = 64;   // the wave size: in a flush lane l takes row l
constexpr int kRoundSteps = 2;        // steps between window checks
constexpr int kFlushRounds = 512;     // every slot is emptied at least this often
// A slot sums a node's deposits in one signed 64-bit word.  What enters the window are the
// lanes' cached sums below 2^46 units in magnitude (larger ones go to the node's words in
// global memory); a lane adds at most one of them to a node per step (when its trajectory
// leaves the node or ends), so a slot takes at most 64 per step, and it is flushed at least
// every kFlushRounds rounds: |sum| < kFlushRounds * kRoundSteps * 64 * 2^46 = 2^62.
static_assert(kFlushRounds * kRoundSteps * 64 <= (1 << 16), "window slots must stay below 2^62");

// The LDS of the windows: wave w of a workgroup owns words [w WC 64, (w + 1) WC 64).
extern __shared__ u64 bt_window_words[];

// A wave's window: its placement (wave-uniform) and the mesh of its design (flush addresses).
struct WaveWin {
    int slot0;             // this wave's first word in bt_window_words
    int lo, base, design;  // columns lo .. lo + WC - 1, rows base .. base + 63 of `design`
    int ny;                // rows of that design's mesh
    i64 ofn;               // its first node in the accumulator
};

// Splits a slot's signed sum q into the two node words (low 32 bits, q >> 32).
__device__ __forceinline__ void add_sum(u64* __restrict__ w, u64 q) {
    add_words(w, q & 0xffffffffull, (u64)((i64)q >> 32));
}

// A double below 2^51 in magnitude rounded to the nearest integer (ties to even), as a signed
// 64-bit integer: adding 1.5 2^52 rounds it, and the sum's bit pattern then holds it.
__device__ __forceinline__ i64 exact_i64(double q) {
    return __double_as_longlong(q + 0x1.8p52) - 0x4338000000000000ll;
}

// A lane's FP64 sum of a node's deposits leaves the lane, in units of 2^e (every deposit is
// scaled by 2^-e, exactly, before it is summed): rounded once to an integer (to nearest, ties
// to even) and added to node (ii, jj).  The common case -- a sum below 2^46 units whose node
// the wave's window holds -- is one no-return LDS atomic; any other sum goes to the node's
// words in global memory through to_fixed, which takes sums of any size and drops a non-finite
// or out-of-range one as it drops such a single deposit.
template <bool USE_WIN, int WC>
__device__ __forceinline__ void sum_leaves(const WaveWin& w, KParams kp, i64 ofn, i64 ny, int design,
                                           int ii, int jj, double units) {
    if (__builtin_expect(USE_WIN && fabs(units) < 0x1p46 && design == w.design &&
                         (unsigned)(ii - w.lo) < (unsigned)WC && (unsigned)(jj - w.base) < (unsigned)kWinRows, 1)) {
        atomicAdd(&bt_window_words[w.slot0 + (ii & (WC - 1)) * kWinRows + (jj - w.base)],
                  (u64)exact_i64(units));  // two's complement
        return;
    }
    u64 lo, hi;
    if (to_fixed(units, 0, lo, hi)) add_words(cold(kp)->acc + kWords * (ofn + (i64)ii * ny + jj), lo, hi);
}

// Moves the lane's cache to cell (i, j): the sums of nodes that both cells share stay in
// their slots, the others leave the lane.  The cache's cell (c, r) holds columns c, c + 1 and
// rows r, r + 1; column c stays iff i is c - 1 or c, column c + 1 iff i is c or c + 1 (rows
// alike); slot q holds column c when q's column parity is c's, else c + 1.
//
// With two lanes per trajectory (LPT 2) each lane of the pair keeps the two nodes whose column
// parity is its own (the lane's parity): its slot b holds the node of row parity b, and the
// pair together holds what one lane holds alone -- each node's sum is the same sum, added and
// rounded the same way.
template <int LPT>
__device__ __forceinline__ int lane_column_parity() {
    return LPT == 2 ? (int)(__lane_id() & 1u) : 0;
}

template <bool USE_WIN, int WC, int LPT = 1>
__device__ __forceinline__ void cache_move(LaneState& s, const WaveWin& w, KParams kp, i64 ofn, i64 ny,
                                           int design, int i, int j) {
    if (i == s.cci && j == s.ccj) return;
    const int c = s.cci, r = s.ccj;
    const bool keep_c = (unsigned)(c - i) <= 1u, keep_c1 = (unsigned)(c + 1 - i) <= 1u;
    const bool keep_r = (unsigned)(r - j) <= 1u, keep_r1 = (unsigned)(r + 1 - j) <= 1u;
    const int a = lane_column_parity<LPT>();
#pragma unroll
    for (int q = 0; q < 4 / LPT; ++q) {
        const int qa = LPT == 2 ? a : (q >> 1), qb = LPT == 2 ? q : (q & 1);  // the slot's parities
        const bool at_c = ((c ^ qa) & 1) == 0;  // slot q's column is c (else c + 1)
        const bool at_r = ((r ^ qb) & 1) == 0;
        const bool kept = (at_c ? keep_c : keep_c1) && (at_r ? keep_r : keep_r1);
        if (!kept && s.cq[q] != 0.0)
            sum_leaves<USE_WIN, WC>(w, kp, ofn, ny, design, at_c ? c : c + 1, at_r ? r : r + 1, s.cq[q]);
        s.cq[q] = kept ? s.cq[q] : 0.0;
    }
    s.cci = i;
    s.ccj = j;
}

// Empties the lane's cache (the trajectory ended).
template <bool USE_WIN, int WC, int LPT = 1>
__device__ __forceinline__ void cache_flush(LaneState& s, const WaveWin& w, KParams kp, i64 ofn, i64 ny,
                                            int design) {
    const int c = s.cci, r = s.ccj;
    const int a = lane_column_parity<LPT>();
#pragma unroll
    for (int q = 0; q < 4 / LPT; ++q) {
        const int qa = LPT == 2 ? a : (q >> 1), qb = LPT == 2 ? q : (q & 1);
        if (s.cq[q] != 0.0)
            sum_leaves<USE_WIN, WC>(w, kp, ofn, ny, design, c + ((c ^ qa) & 1), r + ((r ^ qb) & 1), s.cq[q]);
        s.cq[q] = 0.0;
    }
    s.cci = kNoCell;
    s.ccj = kNoCell;
}

// The 4-node deposit of one step into cell (i, j) of a patch whose nodes start at `ofn`:
// node (i + a, j + b) receives ((a ? u : 1 - u) (b ? w : 1 - w)) deposit, evaluated in that
// order (pyccel.py:407-423), `deposit` in units of 2^e.  On the base mesh (in_base) the values
// add to the lane's FP64 sums (rounded to integers when a sum leaves the lane); on a refined
// patch each value is rounded once to an integer (ties to even) and goes to the node words in
// global memory.  The node totals are exact integer sums of rounded values, so the order of
// the lanes does not change them; a non-finite deposit is dropped.
template <bool USE_WIN, int WC, int LPT = 1>
__device__ __forceinline__ void deposit4(LaneState& s, const WaveWin& ww, KParams kp, i64 ofn, i64 ny,
                                         bool in_base, int design, int i, int j, double u, double w,
                                         double deposit) {
    if (__builtin_expect(!(fabs(deposit) < __builtin_inf()), 0)) return;  // (the four values would not be finite)
    if (in_base) cache_move<USE_WIN, WC, LPT>(s, ww, kp, ofn, ny, design, i, j);
    const double u0 = 1.0 - u, w0 = 1.0 - w;
    const bool io = (i & 1) != 0, jo = (j & 1) != 0;
    const int a = lane_column_parity<LPT>();
#pragma unroll
    for (int q = 0; q < 4 / LPT; ++q) {  // slot q: the node of parities (q >> 1, q & 1); (a, q) for a pair
        const int qa = LPT == 2 ? a : (q >> 1), qb = LPT == 2 ? q : (q & 1);
        const bool ia = io != (qa != 0);  // the node is i + 1
        const bool jb = jo != (qb != 0);  // the node is j + 1
        const double value = ((ia ? u : u0) * (jb ? w : w0)) * deposit;
        if (in_base) {
            s.cq[q] += value;
            continue;
        }
        u64 lo, hi;
        if (to_fixed(value, 0, lo, hi)) add_words(cold(kp)->acc + kWords * (ofn + (i64)(i + ia) * ny + (j + jb)), lo, hi);
    }
}

// Columns first .. last of the wave's window (a range it holds) to global memory, slots
// emptied: lane l takes row l of each column.  Called by the whole wave.
template <int WC>
__device__ __forceinline__ void window_flush_columns(const WaveWin& w, KParams kp, int first, int last) {
    const int lane = (int)__lane_id();
    for (int c = first; c <= last; ++c) {
        const int idx = w.slot0 + (c & (WC - 1)) * kWinRows + lane;
        const u64 q = bt_window_words[idx];
        if (q != 0) {
            add_sum(cold(kp)->acc + kWords * (w.ofn + (i64)c * w.ny + w.base + lane), q);
            bt_window_words[idx] = 0ull;
        }
    }
}

// The whole window to global memory (the wave ends, or re-bases).
template <int WC>
__device__ __forceinline__ void window_flush_all(const WaveWin& w, KParams kp) {
    if (w.design >= 0) window_flush_columns<WC>(w, kp, w.lo, w.lo + WC - 1);
}

// The wave's window at the start: empty, placed nowhere (no design).
template <int WC>
__device__ __forceinline__ WaveWin window_start() {
    WaveWin w;
    w.slot0 = (int)(threadIdx.x / 64u) * WC * kWinRows;
    w.lo = 0; w.base = 0; w.design = -1; w.ny = 0; w.ofn = 0;
    const int lane = (int)__lane_id();
    for (int c = 0; c < WC; ++c) bt_window_words[w.slot0 + c * kWinRows + lane] = 0ull;
    return w;
}

// After a round: follow the lanes.  `q` is the mesh of the lanes being stepped (for a batch,
// that of the wave's current design) and `counts` marks them.  The first and last such lanes
// holding a cell give the span the window must hold for another round of travel (a trajectory
// moves at most one cell per step); when it does not, or `force`, the window moves: columns
// placed with the trailing lane kRoundSteps inside the low end, rows centred on the lanes'
// rows.  Wave-uniform; called by the whole wave.
template <int WC>
__device__ __forceinline__ void window_round(WaveWin& w, const TraceParams& q, KParams kp,
                                             const LaneState& s, bool counts, bool force) {
    const u64 has = __ballot(counts && s.cci != kNoCell);
    if (has == 0) {
        if (force) {
            window_flush_all<WC>(w, kp);
            w.design = -1;
        }
        return;
    }
    const int f = __ffsll((long long)has) - 1, l = 63 - __clzll((long long)has);
    const int c0 = __builtin_amdgcn_readlane(s.cci, f), c1 = __builtin_amdgcn_readlane(s.cci, l);
    const int r0 = __builtin_amdgcn_readlane(s.ccj, f), r1 = __builtin_amdgcn_readlane(s.ccj, l);
    const int cmin = min(c0, c1), cmax = max(c0, c1), rmin = min(r0, r1), rmax = max(r0, r1);
    const bool same = q.design == w.design;
    const bool rows_ok = same && rmin - kRoundSteps >= w.base && rmax + 1 + kRoundSteps < w.base + kWinRows;
    const bool cols_ok = cmin - kRoundSteps >= w.lo && cmax + 1 + kRoundSteps < w.lo + WC;
    if (__builtin_expect(rows_ok && cols_ok && !force, 1)) return;
    int nlo = cmin - kRoundSteps;
    if (nlo < 0) nlo = 0;
    int nbase = ((rmin + rmax + 1) >> 1) - kWinRows / 2;
    if (nbase > q.ny - kWinRows) nbase = q.ny - kWinRows;
    if (nbase < 0) nbase = 0;
    // rows that stay (in room, or lanes spread over more rows than the room allows, centred
    // where the window already is): only columns leave, those of [lo, lo + WC) outside
    // [nlo, nlo + WC) -- none when the columns stay too
    if (same && !force && (rows_ok || nbase == w.base)) {
        if (nlo == w.lo) return;
        if (nlo > w.lo) window_flush_columns<WC>(w, kp, w.lo, min(nlo, w.lo + WC) - 1);
        else window_flush_columns<WC>(w, kp, max(nlo + WC, w.lo), w.lo + WC - 1);
        w.lo = nlo;
        return;
    }
    window_flush_all<WC>(w, kp);
    w.lo = nlo; w.base = nbase; w.design = q.design; w.ny = q.ny; w.ofn = q.bofn;
}

// The record of a crossing of plane pl at xp during the step from (x, u) to (y, w)
// (pyccel.py:426-442): the point and velocity interpolated at the crossing, the direction and
// the step; the count goes on past the capacity (the caller then retraces).
__device__ __forceinline__ void record_crossing(KParams kp, const LaneState& s, int pl, double xp,
                                                double x0, double x1, double x2, double y0, double y1, double y2,
                                                double gv, double v0, double v1, double v2, double gw, double w0,
                                                double w1, double w2, i64 ns) {
    const KParams p = cold(kp);
    const double before = x0 - xp;
    const double after = y0 - xp;
    const double sfrac = before / (before - after);
    const i64 idx = s.k * p->n_planes + pl;
    const i64 c = p->counts[idx];
    if (c < p->capacity) {
        double* rec = p->hits + (idx * p->capacity + c) * 8;
        rec[0] = xp;
        rec[1] = x1 + sfrac * (y1 - x1);
        rec[2] = x2 + sfrac * (y2 - x2);
        rec[3] = gv * v0 + sfrac * (gw * w0 - gv * v0);
        rec[4] = gv * v1 + sfrac * (gw * w1 - gv * v1);
        rec[5] = gv * v2 + sfrac * (gw * w2 - gv * v2);
        rec[6] = after >= 0.0 ? 1.0 : -1.0;
        rec[7] = 1.0 * (double)ns;
    }
    p->counts[idx] = c + 1;
}

// One step of the trajectory in `s`; returns the status set by the end tests (-1: go on).
// LPT 2: both lanes of a pair step the same trajectory with the same operations (the same
// bits), each gathering one of the two fields and keeping the deposits of its two nodes; the
// even lane writes the trajectory's records.
template <int INTEG, int GEOM, bool HAS_B, bool GENERAL, bool USE_WIN, int WC, int LPT = 1>
__device__ __forceinline__ int step(const TraceParams& p, KParams kp, const Fields& F, LaneState& s,
                                    const WaveWin& ww, double plane0) {
    const bool writer = LPT == 1 || (__lane_id() & 1u) == 0;
    const double inv_c2 = p.inv_c2;
    double x0 = s.x0, x1 = s.x1, x2 = s.x2, v0 = s.v0, v1 = s.v1, v2 = s.v2;
    const double qmk = s.qmk;
    // the spacing of the finest patch at the step's start (pyccel.py:155-177)
    double hloc = p.h;
    if (GENERAL) {
        double rr = x1;
        if (GEOM == 1) rr = radial(x1, x2).r;
        else if (GEOM == 2) rr = p.origin_y + fabs(x1 - p.origin_y);
        const double lx = locate_x<GENERAL>(p, x0);
        for (int qq = 0; qq < p.n_patches - 1; ++qq) {
            const double* pr = p.patch_real + qq * 7;
            if (lx >= pr[3] && lx <= pr[5] && rr >= pr[4] && rr <= pr[6]) { hloc = pr[2]; break; }
        }
        if (p.graded != 0) {
            const i64 cg = node_cell(x0, p.x_nodes, p.n_x_nodes);
            const double d2 = (p.x_nodes[cg + 1] - p.x_nodes[cg]) / p.h;
            if (d2 < 1.0) hloc = hloc * d2;
        }
    }
    const double gv = 1.0 / lorentz<GENERAL>(v0, v1, v2, inv_c2);
    double q0 = x0, q1 = x1, q2 = x2;
    double kx0 = v0, kx1 = v1, kx2 = v2;
    double sx0 = 0.0, sx1 = 0.0, sx2 = 0.0, sv0 = 0.0, sv1 = 0.0, sv2 = 0.0;
    double w0 = 0.0, w1 = 0.0, w2 = 0.0, y0, y1, y2;
    double dt = 0.0, hdt = 0.0, dt6 = 0.0;
    double e0 = 0.0, e1 = 0.0, e2 = 0.0, b0 = 0.0, b1 = 0.0, b2 = 0.0;
    double end_y = 0.0, end_cy = 0.0, end_cz = 0.0, end_sx = 0.0, end_sy = 0.0;  // Boris: the end point's
    constexpr int NSTAGE = INTEG == 0 ? 4 : 2;
#pragma unroll
    for (int stage = 0; stage < NSTAGE; ++stage) {
        double gk = 1.0 / lorentz<GENERAL>(kx0, kx1, kx2, inv_c2);
        double g0, g1, g2;
        if (INTEG != 0 && stage == 0 && s.have != 0) {
            if (HAS_B) {
                e0 = s.ce0; e1 = s.ce1; e2 = s.ce2; b0 = s.cb0; b1 = s.cb1; b2 = s.cb2;
                const double c0 = gk * kx0, c1 = gk * kx1, c2 = gk * kx2;
                g0 = qmk * (e0 + (c1 * b2 - c2 * b1));
                g1 = qmk * (e1 + (c2 * b0 - c0 * b2));
                g2 = qmk * (e2 + (c0 * b1 - c1 * b0));
            } else {
                g0 = s.ca0; g1 = s.ca1; g2 = s.ca2;
            }
        } else {
            double ex, ey, cy, cz, lq, qy, psx, psy;
            const bool known = !GENERAL && stage == 0 && s.point != 0;
            if (known) { qy = s.pt_y; cy = s.pt_cy; cz = s.pt_cz; psx = s.pt_sx; psy = s.pt_sy; }
            edge_field<GEOM, GENERAL, LPT>(p, F, q0, q1, q2, ex, ey, cy, cz, lq, qy, known, psx, psy,
                                           INTEG == 0 && known, &s.pgx, &s.pgy);
            if (INTEG != 0 && stage == 1) { end_y = qy; end_cy = cy; end_cz = cz; end_sx = psx; end_sy = psy; }
            if (GEOM != 0) {
                const double qe = qmk * ey;
                g0 = qmk * ex;
                g1 = qe * cy;
                g2 = qe * cz;
            } else {
                g0 = qmk * ex;
                g1 = qmk * ey;
                g2 = qmk * 0.0;
            }
            if (HAS_B) {
                e0 = ex;
                e1 = GEOM != 0 ? ey * cy : ey;
                e2 = GEOM == 1 ? ey * cz : 0.0;
                magnetic_at<GEOM>(p, lq, qy, cy, cz, b0, b1, b2);
                const double c0 = gk * kx0, c1 = gk * kx1, c2 = gk * kx2;
                g0 = qmk * (e0 + (c1 * b2 - c2 * b1));
                g1 = qmk * (e1 + (c2 * b0 - c0 * b2));
                g2 = qmk * (e2 + (c0 * b1 - c1 * b0));
            }
        }
        if (stage == 0) {
            const double a0 = g0, a1 = g1, a2 = g2;
            dt = step_rule(norm2(v0, v1, v2), norm2(a0, a1, a2), hloc, p.step_fraction, gv).dt;
            hdt = 0.5 * dt;
            dt6 = dt * (1.0 / 6.0);
            if (INTEG != 0) {
                if (HAS_B) {
                    kx0 = v0; kx1 = v1; kx2 = v2;
                    boris_kick(kx0, kx1, kx2, e0, e1, e2, b0, b1, b2, hdt, qmk, GENERAL ? inv_c2 : 0.0);
                } else {
                    kx0 = __builtin_fma(hdt, a0, v0);
                    kx1 = __builtin_fma(hdt, a1, v1);
                    kx2 = __builtin_fma(hdt, a2, v2);
                }
                gk = 1.0 / lorentz<GENERAL>(kx0, kx1, kx2, inv_c2);
                q0 = __builtin_fma(dt * gk, kx0, x0);
                q1 = __builtin_fma(dt * gk, kx1, x1);
                q2 = __builtin_fma(dt * gk, kx2, x2);
                continue;
            }
            // (the stage updates are fused multiply-adds, as the adjoint's recompute)
            sx0 = gv * v0; sx1 = gv * v1; sx2 = gv * v2;
            sv0 = a0; sv1 = a1; sv2 = a2;
            q0 = __builtin_fma(hdt * gv, v0, x0);
            q1 = __builtin_fma(hdt * gv, v1, x1);
            q2 = __builtin_fma(hdt * gv, v2, x2);
            kx0 = __builtin_fma(hdt, a0, v0);
            kx1 = __builtin_fma(hdt, a1, v1);
            kx2 = __builtin_fma(hdt, a2, v2);
        } else if (INTEG != 0) {
            if (HAS_B) {
                w0 = kx0; w1 = kx1; w2 = kx2;
                boris_kick(w0, w1, w2, e0, e1, e2, b0, b1, b2, hdt, qmk, GENERAL ? inv_c2 : 0.0);
                s.ce0 = e0; s.ce1 = e1; s.ce2 = e2; s.cb0 = b0; s.cb1 = b1; s.cb2 = b2;
            } else {
                w0 = __builtin_fma(hdt, g0, kx0);
                w1 = __builtin_fma(hdt, g1, kx1);
                w2 = __builtin_fma(hdt, g2, kx2);
                s.ca0 = g0; s.ca1 = g1; s.ca2 = g2;
            }
            s.have = 1;
        } else if (stage == 1 || stage == 2) {
            const double c = stage == 1 ? hdt : dt;
            sx0 = __builtin_fma(2.0 * gk, kx0, sx0); sx1 = __builtin_fma(2.0 * gk, kx1, sx1);
            sx2 = __builtin_fma(2.0 * gk, kx2, sx2);
            sv0 = __builtin_fma(2.0, g0, sv0); sv1 = __builtin_fma(2.0, g1, sv1); sv2 = __builtin_fma(2.0, g2, sv2);
            q0 = __builtin_fma(c * gk, kx0, x0); q1 = __builtin_fma(c * gk, kx1, x1); q2 = __builtin_fma(c * gk, kx2, x2);
            kx0 = __builtin_fma(c, g0, v0); kx1 = __builtin_fma(c, g1, v1); kx2 = __builtin_fma(c, g2, v2);
        } else {
            sx0 = __builtin_fma(gk, kx0, sx0); sx1 = __builtin_fma(gk, kx1, sx1); sx2 = __builtin_fma(gk, kx2, sx2);
            sv0 = sv0 + g0; sv1 = sv1 + g1; sv2 = sv2 + g2;
        }
    }
    if (INTEG == 0) {
        y0 = __builtin_fma(dt6, sx0, x0); y1 = __builtin_fma(dt6, sx1, x1); y2 = __builtin_fma(dt6, sx2, x2);
        w0 = __builtin_fma(dt6, sv0, v0); w1 = __builtin_fma(dt6, sv1, v1); w2 = __builtin_fma(dt6, sv2, v2);
    } else {
        y0 = q0; y1 = q1; y2 = q2;
    }

    // the end point (the next step's start): its distance from the axis, radial direction and
    // cell coordinates, and the electrode entry of its cell, requested here so that the load
    // overlaps the deposit below (pyccel.py:450-497 test them after the deposit; the order of
    // these reads does not change a result).  Cell indices are clamped, so the load stays in
    // bounds for a point off the mesh too.
    const bool end_known = !GENERAL && INTEG != 0 && s.have != 0;
    double py = y1, pcy = 0.0, pcz = 0.0;
    if (end_known) {
        py = end_y; pcy = end_cy; pcz = end_cz;
    } else if (GEOM == 1) {
        const Radial rd = radial(y1, y2);
        py = rd.r; pcy = rd.cy; pcz = rd.cz;
    } else if (GEOM == 2) {
        py = p.origin_y + fabs(y1 - p.origin_y);
    }
    const double lx = locate_x<GENERAL>(p, y0);
    const double ci = end_known ? end_sx : (GENERAL ? (lx - p.origin_x) / p.h : axis_coord(lx, p.origin_x, p.inv_h));
    const double cj = end_known ? end_sy : (GENERAL ? (py - p.origin_y) / p.h : axis_coord(py, p.origin_y, p.inv_h));
    // clamped on both sides as trace_trajectories_2d_cuda.cu:340-341: min(trunc(max(c, 0)),
    // n - 2, cells - 1), taken as the saturated trunc(c) clamped to [0, that bound] (the
    // same integer for an integer bound)
    const int ei = clamped_floor_index<true>(ci, p.cell_imax);
    const int ej = clamped_floor_index<true>(cj, p.cell_jmax);
    const i64 cell = p.cells[mad24((unsigned)ei, (unsigned)p.cells_cols, (unsigned)p.cells_base + (unsigned)ej)];
    // RK4 on one patch: the next step's first-stage gathers at the end point, issued before the
    // deposit so that their loads run under it (used if the trajectory goes on: `point`)
    if (INTEG == 0 && !GENERAL) issue_first_stage<LPT>(p, F, ci, cj, s.pgx, s.pgy);

    // space charge: dq dt at the Hermite midpoint, area-weighted (pyccel.py:372-423)
    const double gw = 1.0 / lorentz<GENERAL>(w0, w1, w2, inv_c2);
    const double dt8 = dt * 0.125;
    double m0 = __builtin_fma(dt8, gv * v0 - gw * w0, 0.5 * (x0 + y0));
    double m1 = __builtin_fma(dt8, gv * v1 - gw * w1, 0.5 * (x1 + y1));
    const double m2 = __builtin_fma(dt8, gv * v2 - gw * w2, 0.5 * (x2 + y2));
    if (GEOM == 1) m1 = radial(m1, m2).r;
    else if (GEOM == 2) m1 = p.origin_y + fabs(m1 - p.origin_y);
    m0 = locate_x<GENERAL>(p, m0);
    {
        // the cell and the area weights of the midpoint, clamped to the patch as the reference
        // clamps them (pyccel.py:406-417: a point below the first node deposits on it with
        // weight 1, one beyond the last on the last; a NaN point on the first): the cell index
        // by the saturated conversion and a median, the fractions clamped to [0, 1]
        const PatchView pv = patch_view<GENERAL>(p, patch_of<GENERAL>(p, m0, m1));
        const double sxm = GENERAL ? (m0 - pv.ox) / pv.h : axis_coord(m0, pv.ox, p.inv_h);
        const double sym = GENERAL ? (m1 - pv.oy) / pv.h : axis_coord(m1, pv.oy, p.inv_h);
        const int i = clamped_floor_index<!GENERAL>(sxm, (int)pv.nx - 2);  // (a patch has fewer than 2^29 nodes)
        const int j = clamped_floor_index<!GENERAL>(sym, (int)pv.ny - 2);
        const double u = fmin(fmax(sxm - (double)i, 0.0), 1.0);
        double w;
        if (GEOM == 1) {
            const double low = pv.oy + pv.h * (double)j;
            const double high = low + pv.h;
            w = __builtin_fma(m1, m1, -(low * low)) * recip_nr(high * high - low * low);
        } else {
            w = sym - (double)j;
        }
        w = fmin(fmax(w, 0.0), 1.0);
        deposit4<USE_WIN, WC, LPT>(s, ww, kp, pv.ofn, pv.ny, !GENERAL || pv.ofn == p.bofn, p.design, i, j, u, w,
                                   s.dqk * dt);
    }
    const i64 ns = s.ns + 1;
    s.ns = ns;

    // plane crossings (pyccel.py:426-442); the first plane's position is in a register, and a
    // record's address is formed only for a crossing (rare)
    if (p.n_planes > 0) {
        if (__builtin_expect((x0 - plane0 < 0.0) != (y0 - plane0 < 0.0), 0) && writer)
            record_crossing(kp, s, 0, plane0, x0, x1, x2, y0, y1, y2, gv, v0, v1, v2, gw, w0, w1, w2, ns);
        if (__builtin_expect(p.n_planes > 1, 0)) {
            for (int pl = 1; pl < p.n_planes; ++pl) {
                const double xp = p.planes[p.planes_base + pl];
                if ((x0 - xp < 0.0) != (y0 - xp < 0.0) && writer)
                    record_crossing(kp, s, pl, xp, x0, x1, x2, y0, y1, y2, gv, v0, v1, v2, gw, w0, w1, w2, ns);
            }
        }
    }
    s.x0 = y0; s.x1 = y1; s.x2 = y2; s.v0 = w0; s.v1 = w1; s.v2 = w2;

    // the end of the step: off the mesh or in an electrode (pyccel.py:450-497)
    s.point = 0;
    if (s.slot >= 0) {
        const i64 c = p.path_count[s.slot];
        if (c < p.path_capacity) {
            p.paths[(s.slot * p.path_capacity + c) * 2 + 0] = y0;
            p.paths[(s.slot * p.path_capacity + c) * 2 + 1] = py;
        }
        p.path_count[s.slot] = c + 1;
    }
    int st = -1;
    if (lx >= p.upper_x) {
        st = 0;
    } else if (lx < p.origin_x) {
        st = 1;
    } else if (py > p.upper_y || (GEOM == 0 && py < p.origin_y)) {
        st = 2;
    } else {
        s.pt_y = py; s.pt_cy = pcy; s.pt_cz = pcz; s.pt_sx = ci; s.pt_sy = cj; s.point = 1;
        if (cell >= 0) {
            const KParams c = cold(kp);
            const int refine = c->refine;
            const double R = (double)refine;
            const int si = (int)fmin(fmax((ci - (double)ei) * R, 0.0), R - 1.0);
            const int sj = (int)fmin(fmax((cj - (double)ej) * R, 0.0), R - 1.0);
            const i64 at = (cell * refine + si) * refine + sj;
            const i64 e = at < p.n_subcells ? c->subcells[p.subcells_base + at] : -1;
            if (e >= 0) {
                st = 3;
                if (writer) c->electrode[s.k] = e;
            } else if (e == -2) {
                st = -2;
            }
        }
    }
    return st;
}

__device__ __forceinline__ void store_end(KParams kp, const LaneState& s, int st) {
    const KParams p = cold(kp);
    const i64 N = p->n, k = s.k;
    double* const x = p->x;
    double* const v = p->v;
    x[k] = s.x0; x[N + k] = s.x1; x[2 * N + k] = s.x2;
    v[k] = s.v0; v[N + k] = s.v1; v[2 * N + k] = s.v2;
    p->steps[k] = s.ns;
    p->status[k] = st;
}

// ---- batched form -------------------------------------------------------------------
// A design's mesh: its table entry over the batch's shared fields (the arenas, the
// accumulator and the trajectory arrays stay those of the launch; the design's nodes start
// at its node offset in the accumulator).  Called with a wave-uniform design, so that the
// mesh parameters stay in scalar registers.
__device__ __forceinline__ void load_design(TraceParams& lp, const TraceParams& p, int d) {
    // the design table does not change during a launch: read through the constant address
    // space, a uniform entry is a few scalar loads and stays in scalar registers
    typedef const __attribute__((address_space(4))) BatchDesign* Table;
    const auto& b = ((Table)p.designs)[d];
    lp.nx = (int)b.nx; lp.ny = (int)b.ny;
    lp.origin_x = b.origin_x; lp.origin_y = b.origin_y; lp.h = b.h;
    lp.upper_x = b.upper_x; lp.upper_y = b.upper_y;
    lp.box = b.origin_x; lp.boy = b.origin_y; lp.bh = b.h;
    lp.inv_bh = 1.0 / b.h; lp.inv_h = lp.inv_bh;
    lp.bofx = 0; lp.bofy = 0; lp.bofn = (int)b.node_offset;
    lp.fx_base = (int)b.fx_offset; lp.fy_base = (int)b.fy_offset;
    lp.cells_base = (int)b.cells_offset; lp.cells_rows = (int)b.cells_rows; lp.cells_cols = (int)b.cells_cols;
    lp.cell_imax = (int)(b.nx - 2 < b.cells_rows - 1 ? b.nx - 2 : b.cells_rows - 1);
    lp.cell_jmax = (int)(b.ny - 2 < b.cells_cols - 1 ? b.ny - 2 : b.cells_cols - 1);
    lp.subcells_base = (int)b.subcells_offset; lp.n_subcells = (int)b.n_subcells;
    lp.planes_base = d * p.n_planes;
    lp.deposit_exponent = b.deposit_exponent;
    lp.design = d;
    // a design's node magnetic field is its run of the launch's node arenas
    lp.b0 = p.b0 != nullptr ? p.b0 + b.node_offset : nullptr;
    lp.b1 = p.b1 != nullptr ? p.b1 + b.node_offset : nullptr;
    lp.b2 = p.b2 != nullptr ? p.b2 + b.node_offset : nullptr;
}

// The persistent trace loop of one wave.  Waves run independently (no barrier): each claims
// trajectories for its free lanes from the launch's queue, steps them in rounds of
// kRoundSteps steps and keeps its own deposit window.  One mesh (BATCHED false): every lane
// uses the launch's parameters.  BATCHED: every design's trajectories share the queue; the
// wave steps the lanes of one design at a time, with that design's mesh in wave-uniform
// registers, and turns to the design of its first busy lane when none of the current one is
// left (lanes of other designs wait: a queue laid out in runs of whole waves per design
// never makes a wave wait).
//
// RECORDS: the launch may record paths and checkpoints (else that code is compiled out; a
// batched launch records neither).
//
// LPT 2: two lanes per trajectory (lanes 2t, 2t + 1 of a wave take the same trajectory and
// step it together; RECORDS false).
template <int INTEG, int GEOM, bool HAS_B, bool GENERAL, bool BATCHED, bool RECORDS, int WC = kWinCols, int LPT = 1>
__device__ __forceinline__ void trace_loop(const TraceParams& p) {
    static_assert(LPT == 1 || (!GENERAL && !RECORDS), "two lanes per trajectory: one uniform patch, no records");
    const KParams kp = kernel_params();  // p itself, for the rarely used fields
    WaveWin ww = window_start<WC>();
    const unsigned lane = __lane_id();
    const u64 below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const bool writer = LPT == 1 || (lane & 1u) == 0;  // the lane that writes its trajectory's end
    TraceParams dp = p;  // BATCHED: the mesh of design wd
    int wd = -1;
    const TraceParams& q = BATCHED ? dp : p;  // the mesh of the lanes being stepped
    const Fields F = {field_rsrc(p.fxs, p.n_fxs), field_rsrc(p.fys, p.n_fys)};
    // the first plane's position, read once (for a batch, once per design of the wave)
    double plane0 = (!BATCHED && p.n_planes > 0) ? p.planes[p.planes_base] : 0.0;
    LaneState s;
    s.k = -1;
    s.cci = s.ccj = kNoCell;
    s.cq[0] = s.cq[1] = s.cq[2] = s.cq[3] = 0.0;
    int my_design = 0;      // BATCHED: the design of the lane's trajectory
    bool exhausted = false;  // wave-uniform: the queue has no entry left for this wave
    bool refill = true;      // with refill_block: only when every lane of the wave is free
    int rounds = 0;
    for (;;) {
        if (!exhausted && refill) {
            const KParams c = cold(kp);
            const bool want = s.k < 0;
            const u64 need = __ballot(want);
            bool end_seen = false;
            if (need != 0) {
                // (a pair's two lanes are free together and take one entry: the pair's rank)
                const int count = __popcll(need) / LPT;
                const int leader = __ffsll((long long)need) - 1;
                u64 base = 0;
                if ((int)lane == leader) base = atomicAdd(c->queue, (u64)count);
                base = __shfl(base, leader);
                if (want) {
                    const u64 m = base + (u64)(__popcll(need & below) / LPT);
                    if (m < (u64)c->n_active) {
                        const i64 k = c->active[m];
                        const i64 N = c->n;
                        bool valid = k >= 0 && k < N;
                        if (BATCHED && valid) {
                            my_design = c->design_of[k];
                            valid = my_design >= 0 && my_design < c->n_designs;
                        }
                        // an index that is not a trajectory (or of no design) is skipped: the
                        // lane asks again
                        if (valid) {
                            s.k = k;
                            s.x0 = c->x[k]; s.x1 = c->x[N + k]; s.x2 = c->x[2 * N + k];
                            s.v0 = c->v[k]; s.v1 = c->v[N + k]; s.v2 = c->v[2 * N + k];
                            s.qmk = c->qm[k];
                            // the trajectory's deposits in units of its design's fixed-point
                            // grid 2^e (an exact scaling): its lane sums are then ready to round
                            int e = c->deposit_exponent;
                            if (BATCHED) {
                                typedef const __attribute__((address_space(4))) BatchDesign* Table;
                                e = ((Table)c->designs)[my_design].deposit_exponent;
                            }
                            s.dqk = ldexp(c->dq[k], -e);
                            s.slot = (RECORDS && c->path_slot != nullptr) ? c->path_slot[k] : -1;
                            if (s.slot >= c->n_paths) s.slot = -1;
                            s.ns = c->steps[k];
                            s.have = 0;
                            s.point = 0;
                            s.ca0 = s.ca1 = s.ca2 = 0.0;
                            s.ce0 = s.ce1 = s.ce2 = s.cb0 = s.cb1 = s.cb2 = 0.0;
                            if (RECORDS && c->ckpt != nullptr) {  // the first multiple of the interval >= ns
                                s.ck_index = (s.ns + c->ckpt_interval - 1) / c->ckpt_interval;
                                s.ck_next = s.ck_index * c->ckpt_interval;
                            }
                            if (s.ns >= c->max_steps) {  // no step: the loop of the reference ends at once
                                if (writer) store_end(kp, s, 4);
                                s.k = -1;
                            }
                        }
                    } else {
                        end_seen = true;
                    }
                }
            }
            // claims are handed out in order: once one lane sees the end, later claims would too
            if (__ballot(end_seen) != 0) exhausted = true;
        }
        const u64 busy = __ballot(s.k >= 0);
        if (busy == 0) {
            if (exhausted) break;
            continue;  // every claim was skipped: ask again
        }
        if (BATCHED && __ballot(s.k >= 0 && my_design == wd) == 0) {
            wd = __builtin_amdgcn_readlane(my_design, __ffsll((long long)busy) - 1);
            load_design(dp, p, wd);
            typedef const __attribute__((address_space(4))) double* Planes;  // read-only during the launch
            plane0 = p.n_planes > 0 ? ((Planes)p.planes)[dp.planes_base] : 0.0;
        }
        // a round of steps; a lane whose trajectory ends waits for the next round's refill
        for (int r = 0; r < kRoundSteps; ++r) {
            if (s.k >= 0 && (!BATCHED || my_design == wd)) {
                if (RECORDS && p.ckpt != nullptr && s.ns == s.ck_next) {
                    if (s.ck_index < p.n_ckpt) {
                        double* c = p.ckpt + s.ck_index * 6 * p.n + s.k;
                        c[0] = s.x0; c[p.n] = s.x1; c[2 * p.n] = s.x2;
                        c[3 * p.n] = s.v0; c[4 * p.n] = s.v1; c[5 * p.n] = s.v2;
                    }
                    s.ck_index += 1;
                    s.ck_next += p.ckpt_interval;
                }
                const int st = step<INTEG, GEOM, HAS_B, GENERAL, true, WC, LPT>(q, kp, F, s, ww, plane0);
                if (st != -1 || s.ns >= p.max_steps) {
                    if (writer) store_end(kp, s, st != -1 ? st : 4);
                    s.k = -1;
                    cache_flush<true, WC, LPT>(s, ww, kp, q.bofn, q.ny, q.design);
                }
            }
        }
        window_round<WC>(ww, q, kp, s, !BATCHED || my_design == wd, ++rounds % kFlushRounds == 0);
        refill = !p.refill_block || __ballot(s.k >= 0) == 0;
    }
    window_flush_all<WC>(ww, kp);
}

template <int INTEG, int GEOM, bool HAS_B, bool GENERAL, bool RECORDS, int WC, int LPT = 1>
__global__ __launch_bounds__(256) void trace_kernel(TraceParams p) {
    trace_loop<INTEG, GEOM, HAS_B, GENERAL, false, RECORDS, WC, LPT>(p);
}

// Every design's trajectories in one persistent launch.  Held to three waves per SIMD (168
// VGPRs): a batch always fills the GPU, so the third wave's latency hiding pays.  WC: the
// columns of each wave's deposit window (a wide window of 32 holds lanes spread further
// along x, at twice the LDS); LPT: lanes per trajectory (1 or 2).
template <int INTEG, int GEOM, bool HAS_B, int WC, int LPT = 1>
__global__ __launch_bounds__(256, 3) void batched_kernel(TraceParams p) {
    trace_loop<INTEG, GEOM, HAS_B, false, true, false, WC, LPT>(p);
}

// The start of a batched trace from the emitter: x, v from x0, v0; steps 0; status and
// electrode -1; plane counts 0; the node charges 0 (a kernel, so that a captured graph
// holds no memset node).
__global__ void batched_reset_kernel(TraceParams p, const double* __restrict__ x0, const double* __restrict__ v0,
                                     i64 total_nodes) {
    const i64 stride = (i64)gridDim.x * blockDim.x;
    const i64 first = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    for (i64 t = first; t < 3 * p.n; t += stride) {
        p.x[t] = x0[t];
        p.v[t] = v0[t];
    }
    for (i64 t = first; t < p.n; t += stride) {
        p.steps[t] = 0;
        p.status[t] = -1;
        p.electrode[t] = -1;
    }
    for (i64 t = first; t < p.n * p.n_planes; t += stride) p.counts[t] = 0;
    for (i64 t = first; t < total_nodes; t += stride) p.charge[t] = 0.0;
}

// charge[design d's node c] += its fixed-point total, with d's exponent (blockIdx.y = d)
__global__ void batched_finalize_kernel(const u64* __restrict__ acc, const BatchDesign* __restrict__ designs,
                                        double* __restrict__ charge) {
    const BatchDesign& b = designs[blockIdx.y];
    const i64 stride = (i64)gridDim.x * blockDim.x;
    for (i64 c = (i64)blockIdx.x * blockDim.x + threadIdx.x; c < b.n_nodes; c += stride) {
        const i64 node = b.node_offset + c;
        const u64 w0 = acc[kWords * node], w1 = acc[kWords * node + 1];
        if (w0 != 0 || w1 != 0) charge[node] += words_value(w0, w1, b.deposit_exponent);
    }
}

// zero the fixed-point accumulator and the queue head
__global__ void prepare_kernel(u64* __restrict__ acc, i64 words, u64* __restrict__ queue) {
    const i64 stride = (i64)gridDim.x * blockDim.x;
    for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; t < words; t += stride) acc[t] = 0ull;
    if (blockIdx.x == 0 && threadIdx.x == 0) *queue = 0ull;
}

// charge[c] += the fixed-point total of node c
__global__ void finalize_kernel(const u64* __restrict__ acc, i64 nodes, int e, double* __restrict__ charge) {
    const i64 stride = (i64)gridDim.x * blockDim.x;
    for (i64 c = (i64)blockIdx.x * blockDim.x + threadIdx.x; c < nodes; c += stride) {
        const u64 w0 = acc[kWords * c], w1 = acc[kWords * c + 1];
        if (w0 != 0 || w1 != 0) charge[c] += words_value(w0, w1, e);
    }
}

// The LDS of a workgroup's deposit windows: one window per wave.
static size_t window_bytes(int block, bool wide) {
    return (size_t)(block / 64) * (wide ? kWideCols : kWinCols) * kWinRows * sizeof(u64);
}

template <int INTEG, int GEOM, bool HAS_B, bool GENERAL, bool RECORDS, int WC, int LPT = 1>
static TraceKernel kernel_ptr() {
    return trace_kernel<INTEG, GEOM, HAS_B, GENERAL, RECORDS, WC, LPT>;
}

// The general variants (refined patches, graded axis, relativistic motion) always record and
// use the narrow window; the others come with either window width, and without records also
// with two lanes per trajectory (`pairs`, ignored otherwise).
TraceKernel select_trace_kernel(int integ, int geom, bool has_b, bool general, bool records, bool wide,
                                bool pairs) {
    if (pairs && !general && !records) {
#define BT_SELP(I, G, B, W) \
        if (integ == I && geom == G && has_b == B && wide == W) return kernel_ptr<I, G, B, false, false, W ? kWideCols : kWinCols, 2>();
#define BT_SELP_G(I, G) BT_SELP(I, G, false, false) BT_SELP(I, G, false, true) BT_SELP(I, G, true, false) BT_SELP(I, G, true, true)
        BT_SELP_G(0, 0) BT_SELP_G(0, 1) BT_SELP_G(0, 2) BT_SELP_G(1, 0) BT_SELP_G(1, 1) BT_SELP_G(1, 2)
#undef BT_SELP_G
#undef BT_SELP
        return nullptr;
    }
#define BT_SEL(I, G, B, X, R, W) \
    if (integ == I && geom == G && has_b == B && general == X && (X || records == R) && (X || wide == W)) \
        return kernel_ptr<I, G, B, X, R, W ? kWideCols : kWinCols>();
#define BT_SEL_B(I, G, B) BT_SEL(I, G, B, false, false, false) BT_SEL(I, G, B, false, true, false) \
                          BT_SEL(I, G, B, false, false, true) BT_SEL(I, G, B, false, true, true) \
                          BT_SEL(I, G, B, true, true, false)
#define BT_SEL_G(I, G) BT_SEL_B(I, G, false) BT_SEL_B(I, G, true)
    BT_SEL_G(0, 0) BT_SEL_G(0, 1) BT_SEL_G(0, 2)
    BT_SEL_G(1, 0) BT_SEL_G(1, 1) BT_SEL_G(1, 2)
#undef BT_SEL_G
#undef BT_SEL_B
#undef BT_SEL
    return nullptr;
}

int launch_trace(const TraceParams& p, int block, int grid, hipStream_t stream, int integ, int geom,
                 bool has_b, bool general, bool records, bool wide, bool pairs, i64 nodes) {
    TraceKernel kern = select_trace_kernel(integ, geom, has_b, general, records, wide, pairs);
    if (kern == nullptr) return -1;
    const i64 words = kWords * nodes;
    int zgrid = (int)((words + 255) / 256);
    if (zgrid > 4096) zgrid = 4096;
    if (zgrid < 1) zgrid = 1;
    hipLaunchKernelGGL(prepare_kernel, dim3(zgrid), dim3(256), 0, stream, p.acc, words, p.queue);
    if (p.n_active > 0 && grid > 0)
        hipLaunchKernelGGL(kern, dim3(grid), dim3(block), window_bytes(block, !general && wide), stream, p);
    int fgrid = (int)((nodes + 255) / 256);
    if (fgrid > 4096) fgrid = 4096;
    if (fgrid < 1) fgrid = 1;
    hipLaunchKernelGGL(finalize_kernel, dim3(fgrid), dim3(256), 0, stream, p.acc, nodes,
                       p.deposit_exponent, p.charge);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

TraceKernel select_batched_kernel(int integ, int geom, bool has_b, bool wide, bool pairs) {
#define BT_SELB(I, G, B) \
    if (integ == I && geom == G && has_b == B) \
        return pairs ? (wide ? batched_kernel<I, G, B, kWideCols, 2> : batched_kernel<I, G, B, kWinCols, 2>) \
                     : (wide ? batched_kernel<I, G, B, kWideCols> : batched_kernel<I, G, B, kWinCols>);
    BT_SELB(0, 0, false) BT_SELB(0, 1, false) BT_SELB(0, 2, false)
    BT_SELB(1, 0, false) BT_SELB(1, 1, false) BT_SELB(1, 2, false)
    BT_SELB(0, 0, true) BT_SELB(0, 1, true) BT_SELB(0, 2, true)
    BT_SELB(1, 0, true) BT_SELB(1, 1, true) BT_SELB(1, 2, true)
#undef BT_SELB
    return nullptr;
}

int launch_batched(const TraceParams& p, const double* x0, const double* v0, bool reset, int block, int grid,
                   hipStream_t stream, int integ, int geom, bool has_b, bool wide, bool pairs, i64 total_nodes,
                   i64 max_design_nodes) {
    TraceKernel kern = select_batched_kernel(integ, geom, has_b, wide, pairs);
    if (kern == nullptr) return -1;
    if (reset) {
        i64 work = 3 * p.n > total_nodes ? 3 * p.n : total_nodes;
        int rgrid = (int)((work + 255) / 256);
        if (rgrid > 8192) rgrid = 8192;
        if (rgrid < 1) rgrid = 1;
        hipLaunchKernelGGL(batched_reset_kernel, dim3(rgrid), dim3(256), 0, stream, p, x0, v0, total_nodes);
    }
    const i64 words = kWords * total_nodes;
    int zgrid = (int)((words + 255) / 256);
    if (zgrid > 8192) zgrid = 8192;
    if (zgrid < 1) zgrid = 1;
    hipLaunchKernelGGL(prepare_kernel, dim3(zgrid), dim3(256), 0, stream, p.acc, words, p.queue);
    if (p.n_active > 0 && grid > 0)
        hipLaunchKernelGGL(kern, dim3(grid), dim3(block), window_bytes(block, wide), stream, p);
    int fgrid = (int)((max_design_nodes + 255) / 256);
    if (fgrid > 1024) fgrid = 1024;
    if (fgrid < 1) fgrid = 1;
    if (p.n_designs > 0)
        hipLaunchKernelGGL(batched_finalize_kernel, dim3(fgrid, (unsigned)p.n_designs), dim3(256), 0, stream,
                           p.acc, p.designs, p.charge);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int batched_occupancy(int integ, int geom, bool has_b, bool wide, bool pairs, int block) {
    TraceKernel kern = select_batched_kernel(integ, geom, has_b, wide, pairs);
    if (kern == nullptr) return 0;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, reinterpret_cast<const void*>(kern),
                                                     block, window_bytes(block, wide)) != hipSuccess)
        return 0;
    return blocks;
}

// The kernels' square-root arithmetic on arrays: radial(y[k], z[k]) -> r, cy, cz and
// rs[k] = 1 / sqrt(y[k]) for y[k] > 0 (0 otherwise).
__global__ void arith_selftest_kernel(const double* __restrict__ y, const double* __restrict__ z, i64 n,
                                      double* __restrict__ r, double* __restrict__ cy, double* __restrict__ cz,
                                      double* __restrict__ rs) {
    const i64 stride = (i64)gridDim.x * blockDim.x;
    for (i64 k = (i64)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const Radial rd = radial(y[k], z[k]);
        r[k] = rd.r;
        cy[k] = rd.cy;
        cz[k] = rd.cz;
        rs[k] = y[k] > 0.0 ? rsqrt_pos(y[k]) : 0.0;
    }
}

int launch_arith_selftest(const double* y, const double* z, i64 n, double* r, double* cy, double* cz, double* rs,
                          hipStream_t stream) {
    int grid = (int)((n + 255) / 256);
    if (grid > 8192) grid = 8192;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(arith_selftest_kernel, dim3(grid), dim3(256), 0, stream, y, z, n, r, cy, cz, rs);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int trace_occupancy(int integ, int geom, bool has_b, bool general, bool records, bool wide, bool pairs, int block) {
    TraceKernel kern = select_trace_kernel(integ, geom, has_b, general, records, wide, pairs);
    if (kern == nullptr) return 0;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, reinterpret_cast<const void*>(kern),
                                                     block, window_bytes(block, !general && wide)) != hipSuccess)
        return 0;
    return blocks;
}

}  // namespace bt
