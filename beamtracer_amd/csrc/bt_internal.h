// Internal launch parameters shared by the C ABI (bt_api.hip) and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bt {

typedef int64_t i64;
typedef unsigned long long u64;

// One design of a batched trace: its mesh and the offsets of its arrays in the batch's
// arenas (same layout as bt_batch_design in include/beamtracer_hip.h).
struct BatchDesign {
    i64 nx, ny;
    double origin_x, origin_y, h, upper_x, upper_y;
    i64 fx_offset, fy_offset;
    i64 cells_offset, cells_rows, cells_cols;
    i64 subcells_offset, n_subcells;
    i64 node_offset, n_nodes;
    int32_t deposit_exponent, reserved;
};

// Device pointers and scalars of one trace launch (see bt_trace2d_args in
// include/beamtracer_hip.h for the meaning of each).  Counts, sizes and offsets that the
// entry points bound below 2^31 are 32-bit: the step loop holds them in scalar registers.
struct TraceParams {
    const i64* active; i64 n_active;
    double* x; double* v; int n;
    const double* qm; const double* dq;
    i64* steps; i64* status; i64* electrode;
    const double* fxs; const double* fys;
    i64 n_fxs, n_fys;      // doubles in fxs, fys (read through 32-bit buffer offsets)
    int fx_base, fy_base;  // added to every patch offset (a batched lane's design arenas)
    const i64* patch_int; const double* patch_real; int n_patches;
    double origin_x, origin_y, h, upper_x, upper_y;
    double step_fraction; i64 max_steps;
    const i64* cells; int cells_rows, cells_cols;
    int cell_imax, cell_jmax;  // min(nx - 2, cells_rows - 1), min(ny - 2, cells_cols - 1)
    const i64* subcells; int n_subcells; int refine;
    const double* planes; int n_planes;
    // element offsets of the lane's design in cells, subcells and planes (a batched launch;
    // 0 for one mesh): 32-bit offsets from the launch's arrays instead of a 64-bit pointer
    // per lane keep the batched loop's registers down
    int cells_base, subcells_base, planes_base;
    double* hits; int capacity; i64* counts;
    const i64* path_slot; double* paths; int n_paths; int path_capacity; i64* path_count;
    int graded; const double* x_nodes; i64 n_x_nodes;
    int reuse;   // one patch with the mesh's origin and spacing: end-test values carry over
    int refill_block;  // refill a workgroup's lanes together (BT_REFILL_BLOCK)
    double* ckpt; int n_ckpt, ckpt_interval;  // checkpoints [c][6][n] of the states, or null
    double inv_c2;
    const double* b0; const double* b1; const double* b2;
    // base (last) patch = the mesh; filled on the host
    int nx, ny;
    int bofx, bofy, bofn;
    double box, boy, bh;
    double inv_bh, inv_h;  // RN(1 / bh), RN(1 / h): divisions by the spacings as products
    // deposit accumulator and work queue (workspace)
    u64* acc; int deposit_exponent; u64* queue;   // node words, fixed-point exponent, work queue
    double* charge;
    // batched form: the design table and every trajectory's design (device), else null;
    // `design` is the design whose mesh these parameters describe (0 for one mesh)
    const BatchDesign* designs; const int32_t* design_of; int n_designs;
    int design;
};

typedef void (*TraceKernel)(TraceParams);

// `records`: paths or checkpoints may be written (else a variant without that code runs)
// `wide`: each wave's deposit window spans 32 columns instead of 16 (beams whose lanes spread along x)
// `pairs`: two lanes per trajectory (without records, on one uniform patch; ignored otherwise).
TraceKernel select_trace_kernel(int integ, int geom, bool has_b, bool general, bool records, bool wide,
                                bool pairs);
int launch_trace(const TraceParams& p, int block, int grid, hipStream_t stream, int integ, int geom,
                 bool has_b, bool general, bool records, bool wide, bool pairs, i64 nodes);
int trace_occupancy(int integ, int geom, bool has_b, bool general, bool records, bool wide, bool pairs, int block);

// Batched form: one persistent launch over the trajectories of every design (single-patch,
// uniform, classical meshes); `reset` first restores x, v from x0, v0 and clears steps,
// status, electrode, counts and charge.
TraceKernel select_batched_kernel(int integ, int geom, bool has_b, bool wide, bool pairs);
int launch_batched(const TraceParams& p, const double* x0, const double* v0, bool reset, int block, int grid,
                   hipStream_t stream, int integ, int geom, bool has_b, bool wide, bool pairs, i64 total_nodes,
                   i64 max_design_nodes);
int batched_occupancy(int integ, int geom, bool has_b, bool wide, bool pairs, int block);

// The device-array helpers (bt_memory.hip); `bad` and `count` are device counters.
int launch_fill_i64(i64* dst, i64 value, i64 n, hipStream_t stream);
int launch_fill_f64(double* dst, double value, i64 n, hipStream_t stream);
int launch_abs_max(const double* a, i64 n, u64* out, hipStream_t stream);
int launch_max_i64(const i64* a, i64 n, u64* out, hipStream_t stream);
int launch_select_equal(const i64* a, i64 n, i64 value, i64* idx, u64* count, hipStream_t stream);
int launch_gather_columns(const double* src, i64 rows, i64 n, const i64* idx, i64 m, double* dst, u64* bad,
                          hipStream_t stream);
int launch_scatter(i64* dst, i64 n, const i64* idx, const i64* values, i64 m, u64* bad, hipStream_t stream);
int launch_gather_records(const double* hits, i64 n, i64 n_planes, i64 capacity, i64 plane, const i64* trajectory,
                          const i64* slot, i64 m, double* dst, u64* bad, hipStream_t stream);

// The adjoint of a trace (bt_adjoint.hip): seeds dJ/d(record) of the plane crossings ->
// dJ/dfx, dJ/dfy (and the initial states' adjoints), one uniform mesh, classical, no B.
struct AdjointParams {
    const i64* active; i64 n_active; i64 n;
    const double* x; const double* v; const double* qm;
    const double* fxs; const double* fys; i64 nx, ny;
    double origin_x, origin_y, h, step_fraction;
    double inv_h;  // RN(1 / h): divisions by the spacing as products
    const double* planes; int n_planes;
    // element offsets of the lane's design in cells, subcells and planes (a batched launch;
    // 0 for one mesh): 32-bit offsets from the launch's arrays instead of a 64-bit pointer
    // per lane keep the batched loop's registers down
    int cells_base, subcells_base, planes_base;
    const double* hits; i64 capacity; const i64* counts; const double* seeds;
    i64 interval, n_ckpt, lanes;
    const double* fwd_ckpt; i64 fwd_n_ckpt;  // the forward trace's checkpoints [c][6][n], or null
    double* scratch;
    double* dfx; double* dfy; double* dx; double* dv;
    u64* queue;
    i64* dropped;  // += seeded trajectories beyond the checkpoints (their gradient is missing), or null
    // the backward walk of each group of 64 queue entries in `pieces` tasks (1: one task per
    // group, claimed per lane): the adjoint state passes from one piece to the next through
    // `hand` [6][groups * 64], `done[g]` counting group g's finished pieces
    int pieces;
    i64 groups;
    int* done;
    double* hand;
};

int launch_adjoint(const AdjointParams& p, int block, int grid, hipStream_t stream, int integ, int geom);
int adjoint_occupancy(int integ, int geom, int block);

int launch_arith_selftest(const double* y, const double* z, i64 n, double* r, double* cy, double* cz, double* rs,
                          hipStream_t stream);

int launch_edge_field(const double* phi, const int32_t* solid, const uint8_t* near, const double* distance,
                      const double* dx, double h, i64 nx, i64 ny, int faces, double* fx, double* fy,
                      hipStream_t stream);
int launch_gather(const double* px, const double* py, i64 n, const double* fx, const double* fy,
                  i64 nx, i64 ny, double ox, double oy, double h, double* ex, double* ey,
                  hipStream_t stream);
int launch_l2_gather(const double* f, i64 rows, i64 cols, int lookups, int blocks, unsigned seed, double* sink,
                     hipStream_t stream);
int launch_l2_line(const double* f, i64 n_lines, int lines, int blocks, unsigned seed, double* sink,
                   hipStream_t stream);
int launch_deposit(const double* px, const double* py, const double* q, i64 n, double ox,
                   double oy, double h, i64 cylindrical, double* charge, i64 nx, i64 ny, int e,
                   u64* acc, hipStream_t stream);

// Device pointers and scalars of one 3D trace launch (bt_trace3d_args in
// include/beamtracer_hip.h gives the meaning of each; the tables are device copies).
struct Trace3DParams {
    const i64* active; i64 n_active, n;
    double* x; double* v; const double* qm; const double* dq;
    i64* steps; i64* status; i64* electrode;
    const double* fxs; const double* fys; const double* fzs;
    const i64* patch_int; const double* patch_real; int n_patches;
    const double* bn; int has_b;
    double ox, oy, oz, h, ux, uy, uz;
    int mirror_y, mirror_z;
    double step_fraction; i64 max_steps;
    const i64* cells; int cells_n1, cells_n2;
    const i64* subcells; i64 n_subcells; int refine;
    const double* planes; int n_planes;
    // element offsets of the lane's design in cells, subcells and planes (a batched launch;
    // 0 for one mesh): 32-bit offsets from the launch's arrays instead of a 64-bit pointer
    // per lane keep the batched loop's registers down
    int cells_base, subcells_base, planes_base;
    double* hits; int capacity; i64* counts;
    const i64* path_slot; double* paths; i64 n_paths; int path_capacity; i64* path_count;
    int has_j; double inv_c2;
    int nx, ny, nz;          // the base mesh
    u64* acc_q; u64* acc_j;  // fixed-point words of the charge and current nodes
    int e_q, e_j;            // their exponents
};

// The 3D trace: zero the words, trace, add the words to charge and current (row 0).
int launch_trace3d(const Trace3DParams& p, int integrator, i64 charge_nodes, double* charge, i64 current_nodes,
                   double* current, hipStream_t stream);

}  // namespace bt
