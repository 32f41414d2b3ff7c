/*
 * beamtracer_hip.h -- C ABI of the MI355X (gfx950) particle-stage kernels.
 *
 * The three kernels of beamtracer's particle stage, with the argument meaning of the
 * reference kernels they replace (paths relative to the reference's src/beamtracer/):
 *
 *   bt_staggered_field_2d      kernels/staggered_field_2d/staggered_field_2d_pyccel.py:13-23
 *   bt_deposit_trajectories_2d kernels/deposit_trajectories_2d/deposit_trajectories_2d_pyccel.py:13-22
 *   bt_trace_trajectories_2d   kernels/trace_trajectories_2d/trace_trajectories_2d_pyccel.py:42-78
 *                              (called from solver/tracer.py:1237-1247)
 *
 * and, on the field side of the trace, the edge fields of a potential:
 *
 *   bt_edge_field_2d           solver/fields.py:18-67 (StaggeredElectricField, fields.py:116-152)
 *
 * and the 3D tracer of BeamTracer3D:
 *
 *   bt_trace_trajectories_3d   kernels/trace_trajectories_3d/trace_trajectories_3d_pyccel.py:18-60
 *                              (called from solver/tracer3d.py:405)
 *
 * Every entry point exists in two forms:
 *   bt_<name>          host arrays (NumPy memory), synchronous: the drop-in for the Pyccel
 *                      kernel.  Arrays are copied to the device and the outputs back.
 *   bt_<name>_device   device arrays (hipMalloc / torch memory), asynchronous on `stream`
 *                      (a hipStream_t, NULL for the null stream).
 *
 * Arrays are C-contiguous; int64_t where the reference has Pyccel `int`, double where it
 * has `float`.  Outputs are mutated in place exactly as the reference kernels do, except
 * that `charge` receives the whole deposit in its row 0 (as the reference's CUDA kernel,
 * kernels/trace_trajectories_2d/trace_trajectories_2d_cuda.cu:1-4), summed exactly in
 * fixed point with least significant bit 2^deposit_exponent (a trajectory's deposits into a
 * node during one stay in its cells are added in FP64 in step order and rounded once to that
 * grid), so the node charges do not depend on the order of the trajectories or the launch
 * shape.
 *
 * Return value: 0 on success, a negative BT_E_* code otherwise (never aborts);
 * bt_last_error() describes the last failure of the calling thread.
 */
#ifndef BEAMTRACER_HIP_H
#define BEAMTRACER_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BT_OK 0
#define BT_E_ARG (-1)        /* an argument is inconsistent (shape, count, enum)       */
#define BT_E_NOMEM (-2)      /* a device or host allocation failed                     */
#define BT_E_LAUNCH (-3)     /* a kernel launch or a HIP runtime call failed           */
#define BT_E_DEVICE (-4)     /* no usable HIP device                                   */
#define BT_E_UNSUPPORTED (-5)

/* flags of a trace: refill a wave's lanes only when all of them are free, so that the lanes
 * of a wave start their trajectories together and stay in the same cells (their deposits
 * then add up in the wave's LDS window; the lanes of trajectories that end early wait).
 * Results do not depend on it. */
#define BT_REFILL_BLOCK 1
/* flags of a trace: no trajectory records a path (every path_slot is negative); the library
 * then runs a variant without the path code (the host form finds this itself).  A path slot
 * that is not negative is ignored under this flag. */
#define BT_NO_PATHS 2
/* flags of a trace: each wave's deposit window in LDS spans 32 mesh columns instead of 16
 * (faster for beams whose lanes spread along x; results unchanged; ignored with refined
 * patches, a graded axis or relativistic motion). */
#define BT_WIDE_WINDOW 4

/* flag of bt_batch_occupancy only: the variant with a node magnetic field (a batch whose
 * bt_batch_args carry magnetic arenas runs it). */
#define BT_MAGNETIC 8
/* flags of a trace: two lanes per trajectory -- the lanes of a pair step the same trajectory
 * with the same operations, one gathering E_x and the other E_y, each keeping the deposits of
 * its two nodes of the cell (the launch holds twice the lanes; results unchanged, bit for
 * bit; ignored with path records, checkpoints, refined patches, a graded axis or relativistic
 * motion).  The launch-shape tuners time it against one lane per trajectory. */
#define BT_LANE_PAIRS 16

/* Every argument of trace_trajectories_2d (trace_trajectories_2d_pyccel.py:42-78) as a
 * pointer with its extents, then the node magnetic field that the reference's compiled
 * kernel lacks (its NumPy path takes B, solver/tracer.py:519-529, 605-650), then launch
 * and workspace controls. */
typedef struct bt_trace2d_args {
    const int64_t* active; int64_t n_active;          /* active (n_active,)          */
    double* x; double* v; int64_t n;                  /* x, v (3, n)                 */
    const double* qm; const double* dq;               /* (n,)                        */
    int64_t* steps; int64_t* status; int64_t* electrode;  /* (n,)                    */
    const double* fxs; int64_t n_fxs;                 /* flat fx of every patch      */
    const double* fys; int64_t n_fys;                 /* flat fy of every patch      */
    const int64_t* patch_int; const double* patch_real; int64_t n_patches; /* (P,6),(P,7) */
    double origin_x, origin_y, h, upper_x, upper_y;
    int64_t cylindrical;                              /* 0 planar, 1 (x,r), 2 mirror */
    int64_t integrator;                               /* 0 RK4, 1 Boris              */
    double step_fraction;
    int64_t max_steps;
    const int64_t* cells; int64_t cells_rows, cells_cols;   /* (nx-1, ny-1)          */
    const int64_t* subcells; int64_t n_subcells; int64_t refine;
    const double* planes; int64_t n_planes;
    double* hits; int64_t capacity;                   /* (n, n_planes, capacity, 8)  */
    int64_t* counts;                                  /* (n, n_planes)               */
    const int64_t* path_slot;                         /* (n,)                        */
    double* paths; int64_t n_paths; int64_t path_capacity;   /* (n_paths, cap, 2)    */
    int64_t* path_count;                              /* (n_paths,)                  */
    int64_t num_threads;                              /* ignored (one slab)          */
    double* charge; int64_t charge_rows, charge_cols; /* row 0 receives the deposit  */
    int64_t graded; const double* x_nodes; int64_t n_x_nodes;
    double inv_c2;
    /* extension: a node magnetic field on the base (last) patch's nodes (nx, ny),
     * (B_x, B_r, B_theta) cylindrical or (B_x, B_y, B_z) planar; NULL components are 0 */
    int64_t has_b; const double* b0; const double* b1; const double* b2;
    /* controls */
    int32_t deposit_exponent;   /* fixed-point LSB 2^e of the deposit; see bt_deposit_exponent */
    int32_t block_size;         /* threads per block (64..256, a multiple of 64); 0 = default 256 */
    int32_t max_blocks;         /* persistent grid size; 0 = fill the device            */
    int32_t flags;              /* BT_REFILL_BLOCK | BT_NO_PATHS | BT_WIDE_WINDOW, or 0 */
    void* workspace;            /* device scratch of bt_trace_workspace_bytes(); NULL = library-owned,
                                   shared by every call on the device: one call at a time */
    int64_t workspace_bytes;
    /* extension: the state (x, v) of every trajectory before each step whose count is a
     * multiple of checkpoint_interval, into checkpoints[c][6][n] for c < n_checkpoints
     * (device form only; NULL: none) -- the adjoint's checkpoints */
    double* checkpoints; int64_t n_checkpoints; int64_t checkpoint_interval;
} bt_trace2d_args;

/* Library version string. */
const char* bt_version(void);
/* Description of the last error on the calling thread ("" if none). */
const char* bt_last_error(void);
/* sizeof(bt_trace2d_args), for binding checks. */
int64_t bt_trace_args_size(void);

/* The fixed-point exponent e (LSB = 2^e) for deposits whose typical magnitude is
 * `scale` (e.g. max|dq| times a typical step, or max|q|): 40 bits below it. */
int32_t bt_deposit_exponent(double scale);

/* Device scratch needed by a trace or a deposit onto `n_nodes` nodes. */
int64_t bt_trace_workspace_bytes(int64_t n_nodes);

/* trace_trajectories_2d: replaces the kernel call of solver/tracer.py:1237-1247. */
int bt_trace_trajectories_2d(const bt_trace2d_args* args);
int bt_trace_trajectories_2d_device(const bt_trace2d_args* args, void* stream);

/* The device form without any allocation.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments and results as bt_trace_trajectories_2d_device, with
 * args->workspace NULL: the node words, queue and tables live in a fixed device array of
 * BT_NOALLOC_SCRATCH_BYTES that the library holds on each GPU until the process ends, so the
 * mesh (all patches) may hold at most BT_NOALLOC_MAX_NODES nodes (larger: BT_E_ARG, nothing
 * launched).  Every call on one GPU shares that array: calls that overlap on one GPU race. */
#define BT_NOALLOC_MAX_NODES 1048304
#define BT_NOALLOC_SCRATCH_BYTES (16 * 1048304 + 256 + 4096)
int bt_trace_trajectories_2d_noalloc(const bt_trace2d_args* args, void* stream);

/* staggered_field_2d(px, py, fx, fy, origin_x, origin_y, h, ex, ey):
 * fx (nx + 1, ny), fy (nx, ny + 1); replaces staggered_field_2d_pyccel.py:13-55 as called
 * from solver/fields.py:162-165. */
int bt_staggered_field_2d(const double* px, const double* py, int64_t n,
                          const double* fx, const double* fy, int64_t nx, int64_t ny,
                          double origin_x, double origin_y, double h,
                          double* ex, double* ey);
int bt_staggered_field_2d_device(const double* px, const double* py, int64_t n,
                                 const double* fx, const double* fy, int64_t nx, int64_t ny,
                                 double origin_x, double origin_y, double h,
                                 double* ex, double* ey, void* stream);

/* deposit_trajectories_2d(px, py, q, origin_x, origin_y, h, cylindrical, charge):
 * charge (nx, ny) += the area-weighted deposit; replaces
 * deposit_trajectories_2d_pyccel.py:13-59 as called from solver/tracer.py:812-820. */
int bt_deposit_trajectories_2d(const double* px, const double* py, const double* q,
                               int64_t n, double origin_x, double origin_y, double h,
                               int64_t cylindrical, double* charge, int64_t nx, int64_t ny,
                               int32_t deposit_exponent);
int bt_deposit_trajectories_2d_device(const double* px, const double* py, const double* q,
                                      int64_t n, double origin_x, double origin_y, double h,
                                      int64_t cylindrical, double* charge, int64_t nx,
                                      int64_t ny, int32_t deposit_exponent,
                                      void* workspace, int64_t workspace_bytes,
                                      void* stream);

/* ---- the 3D tracer ----------------------------------------------------------------------
 * Every argument of trace_trajectories_3d (trace_trajectories_3d_pyccel.py:18-60, called
 * from solver/tracer3d.py:405) as a pointer with its extents, in the reference's order and
 * meaning: states x, v (3, n); edge fields of the patches flattened one after the other in
 * fxs, fys, fzs; patch_int (P, 8) = node counts, the three field offsets, node count, node
 * offset; patch_real (P, 10) = origin, spacing, box (finest first, the mesh last); bn
 * (3, nx, ny, nz) the base mesh's node B (has_b); cells (nx - 1, ny - 1, nz - 1) and
 * subcells (refine^3 per cell); hits (n, n_planes, capacity, 8); paths (n_paths,
 * path_capacity, 3); charge (rows, nodes) and, with has_j, current (rows, 3 nx ny nz).
 * patch_int, patch_real and planes are read on the host (the library uploads them).  The
 * charge and the current go to row 0, each summed exactly in fixed point with the least
 * significant bits 2^deposit_exponent and 2^current_exponent. */
typedef struct bt_trace3d_args {
    const int64_t* active; int64_t n_active;
    double* x; double* v; int64_t n;
    const double* qm; const double* dq;
    int64_t* steps; int64_t* status; int64_t* electrode;
    const double* fxs; int64_t n_fxs;
    const double* fys; int64_t n_fys;
    const double* fzs; int64_t n_fzs;
    const int64_t* patch_int; const double* patch_real; int64_t n_patches;
    const double* bn; int64_t has_b;
    double origin_x, origin_y, origin_z, h, upper_x, upper_y, upper_z;
    int64_t mirror_y, mirror_z, integrator;
    double step_fraction; int64_t max_steps;
    const int64_t* cells; int64_t cells_n0, cells_n1, cells_n2;
    const int64_t* subcells; int64_t n_subcells; int64_t refine;
    const double* planes; int64_t n_planes;
    double* hits; int64_t capacity; int64_t* counts;
    const int64_t* path_slot; double* paths; int64_t n_paths, path_capacity; int64_t* path_count;
    int64_t num_threads;
    double* charge; int64_t charge_rows, charge_cols;
    int64_t has_j; double* current; int64_t current_rows, current_cols;
    double inv_c2;
    int32_t deposit_exponent, current_exponent;
    void* workspace; int64_t workspace_bytes;  /* device scratch (NULL: the library's own, shared by
                                                  every call on the device: one call at a time) */
} bt_trace3d_args;

/* sizeof(bt_trace3d_args), for binding checks. */
int64_t bt_trace3d_args_size(void);
/* Device scratch of a 3D trace (the node words of charge and current, the tables); -1 for
 * arguments that fail the checks. */
int64_t bt_trace3d_workspace_bytes(const bt_trace3d_args* args);
/* trace_trajectories_3d: replaces the kernel call of solver/tracer3d.py:405-430. */
int bt_trace_trajectories_3d(const bt_trace3d_args* args);
int bt_trace_trajectories_3d_device(const bt_trace3d_args* args, void* stream);

/* ---- the trace's input: edge fields of a potential -----------------------------------
 * The staggered edge fields of the node potential phi (nx, ny): fx (nx + 1, ny) along x and
 * fy (nx, ny + 1) along r.  Replaces the host computation of solver/fields.py:18-67
 * (_edge_field) as StaggeredElectricField builds them (solver/fields.py:116-152): a
 * Vlasov iteration can then keep the fields in HBM from the potential to the trace.
 *   solid     (nx, ny) electrode index of each node, or -1 (IonOpticsMesh.solid);
 *   near      (nx, ny) 1 for a free node with an electrode neighbour, else 0 (never 1 on an
 *             electrode node: IonOpticsMesh.near_solid);
 *   distance  (4, nx, ny) the bracketed surface distance in units of the edge towards -x,
 *             +x, -r, +r (1 where no surface; IonOpticsMesh.distance);
 *   dx        the nx - 1 spacings along x (a graded axis), or NULL for h;
 *   faces     bit k set: face k (x_min, x_max, y_min, y_max) mirrors the field (the axis or
 *             a symmetry plane), else the end edge extrapolates linearly.
 * Bit-identical to the reference's NumPy arithmetic (IEEE quotients). */
int bt_edge_field_2d(const double* phi, const int32_t* solid, const uint8_t* near, const double* distance,
                     const double* dx, double h, int64_t nx, int64_t ny, int32_t faces, double* fx, double* fy);
int bt_edge_field_2d_device(const double* phi, const int32_t* solid, const uint8_t* near,
                            const double* distance, const double* dx, double h, int64_t nx, int64_t ny,
                            int32_t faces, double* fx, double* fy, void* stream);

/* ---- batched form ---------------------------------------------------------------------
 * Many designs (meshes, fields, beams) traced in one persistent launch: the reference
 * evaluates each candidate of beamtracer.optimization in a separate process
 * (optimization/search.py:200 optimize, through the ProcessPoolExecutor of
 * optimization/pool.py:110-160); here every design's trajectories share one work queue.
 * Each design is one uniform patch (no refinement, no graded axis), classical
 * (inv_c2 = 0) and electrostatic; its arrays sit at offsets in shared arenas.  Per design,
 * the results equal those of bt_trace_trajectories_2d on the same inputs (same exponent),
 * bit for bit. */
typedef struct bt_batch_design {
    int64_t nx, ny;                       /* nodes of the design's mesh                    */
    double origin_x, origin_y, h, upper_x, upper_y;
    int64_t fx_offset, fy_offset;         /* its fx (nx+1, ny) in fxs, fy (nx, ny+1) in fys */
    int64_t cells_offset, cells_rows, cells_cols;   /* its cells (>= nx-1, >= ny-1)        */
    int64_t subcells_offset, n_subcells;  /* its sub-cell electrode table                  */
    int64_t node_offset, n_nodes;         /* its nx * ny node charges in charge            */
    int32_t deposit_exponent;             /* its fixed-point exponent                      */
    int32_t reserved;
} bt_batch_design;

/* The arrays of a batched trace (device memory), as bt_trace2d_args for one design, over
 * the concatenated trajectories of every design; design_of[k] is trajectory k's design. */
typedef struct bt_batch_args {
    const int64_t* active; int64_t n_active;
    double* x; double* v; int64_t n;              /* (3, n)                                 */
    const double* x0; const double* v0;           /* non-NULL: restore x, v from them and clear
                                                     steps, status, electrode, counts and
                                                     charge before the trace                */
    const double* qm; const double* dq;
    int64_t* steps; int64_t* status; int64_t* electrode;
    const int32_t* design_of;                     /* (n,)                                   */
    const double* fxs; const double* fys; const int64_t* cells; const int64_t* subcells;
    const double* planes; int64_t n_planes;       /* (n_designs, n_planes)                  */
    double* hits; int64_t capacity; int64_t* counts;   /* (n, n_planes, capacity, 8), (n, n_planes) */
    double* charge;                               /* (bt_batch_plan_nodes(),) += deposits   */
    int64_t cylindrical, integrator;
    double step_fraction; int64_t max_steps; int64_t refine;
    int32_t block_size, max_blocks;               /* 0: 256 threads, all resident blocks    */
    int32_t flags, reserved;                      /* BT_REFILL_BLOCK: a workgroup takes new
                                                     trajectories together (consecutive
                                                     queue entries: mostly one design);
                                                     BT_WIDE_WINDOW: 32-column deposit
                                                     windows                                */
    const double* b0; const double* b1; const double* b2;  /* node magnetic field arenas of
                                                     bt_batch_plan_nodes() doubles, a design's
                                                     nodes at its node_offset ((B_x, B_r,
                                                     B_theta) cylindrical, (B_x, B_y, B_z)
                                                     planar); all NULL: no field, the
                                                     electrostatic variant; a NULL component
                                                     is zero */
} bt_batch_args;

int64_t bt_batch_args_size(void);
int64_t bt_batch_design_size(void);
/* Validates the design table (host memory) against the arena sizes, copies it to the
 * device and allocates the fixed-point workspace of every node. */
int bt_batch_plan_create(const bt_batch_design* designs, int64_t n_designs, int64_t n_fxs,
                         int64_t n_fys, int64_t n_cells, int64_t n_subcells, void** plan);
int bt_batch_plan_destroy(void* plan);
/* Replaces the plan's design table (same count, arenas and node total; e.g. new deposit
 * exponents after a field update); waits for the device first. */
int bt_batch_plan_update(void* plan, const bt_batch_design* designs, int64_t n_designs);
/* Total nodes of the plan's designs (the length of `charge`). */
int64_t bt_batch_plan_nodes(void* plan);
/* Resident blocks per CU of the batched kernel (for launch-shape tuning); `flags` are the
   trace's flags, of which BT_WIDE_WINDOW (the 32-column deposit windows) and BT_MAGNETIC
   (the variant with a magnetic field) count here. */
int bt_batch_occupancy(int64_t integrator, int64_t cylindrical, int32_t block_size, int32_t flags);
/* One batched trace, asynchronous on `stream`; allocates nothing. */
int bt_batch_trace_device(void* plan, const bt_batch_args* args, void* stream);
/* Captures bt_batch_trace_device(plan, args) as a HIP graph (kernel nodes only). */
int bt_batch_graph_create(void* plan, const bt_batch_args* args, void** graph);
int bt_graph_launch(void* graph, void* stream);
int bt_graph_destroy(void* graph);

/* ---- adjoint ---------------------------------------------------------------------------
 * The discrete adjoint of bt_trace_trajectories_2d for an objective J of the plane
 * crossings (no counterpart in the reference, which has no gradients): given the forward
 * trace's records (hits, counts) and seeds dJ/d(y, z, vx, vy, vz) of every record, adds
 * dJ/dfx and dJ/dfy of the edge fields (and dJ/dx, dJ/dv of the initial states) for the
 * electrostatic, classical trace on one uniform mesh.  Each trajectory is walked backward
 * segment by segment, each segment recomputed from its checkpoint: the forward trace's
 * (`checkpoints`), or its own, stepped again from the initial state.  With the forward
 * trace's checkpoints and more groups of 64 trajectories than the grid holds waves, each
 * group's walk runs in pieces handed between waves.  Device memory, asynchronous on
 * `stream`; the scratch is (max_step / interval + 1 + interval) * 48 bytes per lane of the
 * resident grid, plus 48 bytes per trajectory and 4 per group of 64 for the pieces.
 * n and max_step + interval stay below 2^31. */
typedef struct bt_adjoint_args {
    const int64_t* active; int64_t n_active;      /* the trajectories with seeded records  */
    const double* x; const double* v; int64_t n;  /* initial states (3, n)                 */
    const double* qm;                             /* (n,)                                  */
    const double* fx; const double* fy;           /* (nx + 1, ny), (nx, ny + 1)             */
    int64_t nx, ny;
    double origin_x, origin_y, h;
    int64_t cylindrical, integrator;
    double step_fraction;
    const double* planes; int64_t n_planes;
    const double* hits; int64_t capacity; const int64_t* counts;  /* the forward trace's   */
    const double* seeds;                          /* (n, n_planes, capacity, 5)             */
    int64_t max_step;                             /* largest seeded record step            */
    const double* checkpoints;                    /* the forward trace's [c][6][n] at this  */
    int64_t n_checkpoints;                        /* interval, or NULL: stepped again here  */
    int32_t interval;                             /* steps between checkpoints (>= 1)      */
    int32_t block_size;                           /* 0: 256                                 */
    double* dfx; double* dfy;                     /* += dJ/dfx, dJ/dfy                      */
    double* dx; double* dv;                       /* += dJ/dx, dJ/dv (3, n), or NULL        */
    void* workspace; int64_t workspace_bytes;     /* device scratch of bt_adjoint_workspace_bytes()
                                                     (calls on several streams need their own);
                                                     NULL: library-owned, one call at a time */
    int64_t* dropped;                             /* device, or NULL: += the number of seeded
                                                     trajectories whose last seeded step is
                                                     beyond max_step (their gradients are
                                                     missing: max_step was too small) */
} bt_adjoint_args;

/* Measurement helper (no counterpart in the reference): `blocks` x 256 lanes each make
 * `lookups` random bilinear lookups (two 16-byte row-pair loads, 32 B) in the device field of
 * rows x cols doubles (< 2^28 values) and write one sum each to sink[blocks * 256]; the
 * caller times the launch on `stream` for the L2 gather bandwidth of the tracer's access
 * pattern. */
int bt_l2_gather_bandwidth(const double* field, int64_t rows, int64_t cols, int32_t lookups, int32_t blocks,
                           double* sink, void* stream);
/* Measurement helper (no counterpart in the reference): `blocks` x 256 lanes, in groups of 8,
 * read `lines` random whole 128-byte lines each of the device field of n_lines lines
 * (n_lines * 16 doubles, < 2^24 lines), 16 bytes a lane, and write one sum per lane to
 * sink[blocks * 256]: timed on `stream`, the L2 bandwidth in useful bytes of reads that find
 * nothing in L1 (the L2 term of the tracer's roofline). */
int bt_l2_line_bandwidth(const double* field, int64_t n_lines, int32_t lines, int32_t blocks, double* sink,
                         void* stream);

/* Self-test of the kernels' square-root arithmetic (no counterpart in the reference): for
 * device arrays of n doubles, the radial distance r[k] = |(y[k], z[k])| and direction
 * cy[k] = y[k] / r[k], cz[k] = z[k] / r[k] (0 on the axis) as the tracer computes them from
 * one reciprocal square root, and rs[k] = 1 / sqrt(y[k]) for y[k] > 0 (0 otherwise).  The
 * tests bound their distance from the correctly rounded values in units in the last place. */
int bt_arith_selftest(const double* y, const double* z, int64_t n, double* r, double* cy, double* cz, double* rs,
                      void* stream);

int64_t bt_adjoint_args_size(void);
/* Scratch of an adjoint call: its queue head and checkpoint/segment states per lane. */
int64_t bt_adjoint_workspace_bytes(int64_t n_active, int64_t max_step, int32_t interval, int64_t integrator,
                                   int64_t cylindrical, int32_t block_size);
int bt_trace_adjoint_device(const bt_adjoint_args* args, void* stream);

/* ---- device arrays owned by the library ------------------------------------------------
   No counterpart in the reference (its caller keeps host arrays): the host code's
   device-resident state (resident beams, batched designs, edge fields on the GPU) lives in
   memory allocated here, and these are the operations it needs between traces.  `stream`
   NULL is the null stream.  Copies to and from the host return once complete; the helpers
   that report a count or a value return once their kernel has run.  An index out of range is
   skipped and reported as BT_E_ARG.  The helpers share one per-device counter: calls on one
   device must not overlap. */
int bt_compute_units(void);                      /* of the current device (or an error code) */
int bt_device_alloc(int64_t bytes, void** out);  /* *out NULL for 0 bytes */
int bt_device_free(void* ptr);
int bt_copy_to_device(void* dst, const void* src, int64_t bytes, void* stream);
int bt_copy_to_host(void* dst, const void* src, int64_t bytes, void* stream);
int bt_copy_on_device(void* dst, const void* src, int64_t bytes, void* stream);  /* asynchronous */
int bt_fill_i64(int64_t* dst, int64_t value, int64_t n, void* stream);           /* asynchronous */
int bt_fill_f64(double* dst, double value, int64_t n, void* stream);             /* asynchronous */
int bt_synchronize(void* stream);
/* *out = max |a[i]| (NaN if any is NaN; 0 for n = 0) */
int bt_abs_max_f64(const double* a, int64_t n, double* out, void* stream);
/* *out = max a[i] (n >= 1) */
int bt_max_i64(const int64_t* a, int64_t n, int64_t* out, void* stream);
/* idx[0 .. *count) = the i with a[i] == value, in no particular order (idx holds n) */
int bt_select_equal_i64(const int64_t* a, int64_t n, int64_t value, int64_t* idx, int64_t* count, void* stream);
/* dst (rows, m) = src (rows, n) at the columns idx[0 .. m) */
int bt_gather_columns_f64(const double* src, int64_t rows, int64_t n, const int64_t* idx, int64_t m, double* dst,
                          void* stream);
/* dst (n)[idx[k]] = values[k] for k < m */
int bt_scatter_i64(int64_t* dst, int64_t n, const int64_t* idx, const int64_t* values, int64_t m, void* stream);
/* dst (m, 8) = hits (n, n_planes, capacity, 8)[trajectory[r], plane, slot[r], :] */
int bt_gather_records(const double* hits, int64_t n, int64_t n_planes, int64_t capacity, int64_t plane,
                      const int64_t* trajectory, const int64_t* slot, int64_t m, double* dst, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* BEAMTRACER_HIP_H */
