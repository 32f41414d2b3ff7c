// The C ABI of libbeamtracer_hip.so (declared in include/beamtracer_hip.h).
//
// Host forms (bt_<name>): the drop-in for the reference's host kernels.  They validate
// the shapes, copy the arrays to the device (grow-only buffers kept per device), launch,
// copy the outputs back and return when done.
// Device forms (bt_<name>_device): take device pointers and launch asynchronously on the
// caller's stream; only the small patch and plane tables of the trace are read on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/beamtracer_hip.h"
#include "bt_internal.h"

using bt::i64;
using bt::u64;

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& message) {
    g_error = message;
    return code;
}

int hip_fail(hipError_t err, const char* what) {
    return fail(BT_E_LAUNCH, std::string(what) + ": " + hipGetErrorString(err));
}

constexpr int kMaxPatches = 16;
constexpr int kMaxPlanes = 256;
constexpr i64 kMax32 = (i64)1 << 31;  // bound of the 32-bit launch parameters
// bound of a mesh's nodes and cells per axis: row offsets are formed with 24-bit multiplies
constexpr i64 kMaxAxis = ((i64)1 << 24) - 1;
constexpr size_t kTableBytes = 4096;

size_t round256(size_t b) { return (b + 255) & ~size_t(255); }

struct Buffer {
    void* ptr = nullptr;
    size_t bytes = 0;
};

struct DeviceState {
    std::vector<Buffer> buffers;   // host-form staging, by slot
    Buffer workspace;              // library-owned workspace
    Buffer adjoint;                // library-owned adjoint scratch
    Buffer counters;               // the device-array helpers' counters
    hipStream_t stream = nullptr;  // host-form stream
    int cus = 0;
};

std::mutex g_mutex;
std::map<int, DeviceState> g_devices;

int current_device(int* dev) {
    hipError_t err = hipGetDevice(dev);
    if (err != hipSuccess) return hip_fail(err, "hipGetDevice");
    return BT_OK;
}

// per-device state, created on first use (caller holds g_mutex)
int device_state(DeviceState** out) {
    int dev = 0;
    int rc = current_device(&dev);
    if (rc) return rc;
    DeviceState& s = g_devices[dev];
    if (s.cus == 0) {
        hipDeviceProp_t prop;
        hipError_t err = hipGetDeviceProperties(&prop, dev);
        if (err != hipSuccess) return hip_fail(err, "hipGetDeviceProperties");
        s.cus = prop.multiProcessorCount;
        s.buffers.resize(32);
    }
    *out = &s;
    return BT_OK;
}

int ensure(Buffer& b, size_t bytes) {
    if (bytes == 0) bytes = 8;
    if (b.bytes >= bytes) return BT_OK;
    if (b.ptr) (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
    hipError_t err = hipMalloc(&b.ptr, bytes);
    if (err != hipSuccess) return fail(BT_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(err));
    b.bytes = bytes;
    return BT_OK;
}

// The node words of a deposit: 2 words of 8 bytes per node.
size_t node_bytes(i64 nodes) { return round256((size_t)(16 * nodes)); }
size_t workspace_bytes(i64 nodes) { return node_bytes(nodes) + 256 + kTableBytes; }

// ---- the allocation-free trace (bt_trace_trajectories_2d_noalloc) -------------------------
// Its workspace (node words, queue, tables) is this file-scope device array: every GPU the
// library runs on holds one copy until the process ends, and every call on that GPU shares it
// (calls that overlap on one GPU race; the trace's own prepare kernel zeroes the words and the
// queue at the start of each call).  In host code `scratch` names a host-side stand-in, so the
// device address of each GPU's copy is taken once per device with hipGetSymbolAddress and
// used everywhere instead.
constexpr size_t kScratchBytes = (size_t)BT_NOALLOC_SCRATCH_BYTES;
static_assert(kScratchBytes == 16 * (size_t)BT_NOALLOC_MAX_NODES + 256 + kTableBytes,
              "the header's node bound must match the scratch layout");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct NoAllocState {
    void* scratch = nullptr;  // this device's copy of `scratch`
    int cus = 0;
};
constexpr int kMaxDevices = 64;
NoAllocState g_noalloc[kMaxDevices];  // fixed: setting a device up allocates nothing

// The current device's scratch address and CU count, looked up on its first call (caller
// holds g_mutex).
int noalloc_state(NoAllocState** out) {
    int dev = 0;
    int rc = current_device(&dev);
    if (rc) return rc;
    if (dev < 0 || dev >= kMaxDevices) return fail(BT_E_DEVICE, "device index beyond the allocation-free table");
    NoAllocState& s = g_noalloc[dev];
    if (s.scratch == nullptr) {
        void* ptr = nullptr;
        hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (err != hipSuccess) return hip_fail(err, "hipGetSymbolAddress(scratch)");
        int cus = 0;
        err = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (err != hipSuccess) return hip_fail(err, "hipDeviceGetAttribute");
        s.cus = cus;
        s.scratch = ptr;
    }
    *out = &s;
    return BT_OK;
}

// The fixed launch-independent checks of a trace (host-readable tables only).
int check_trace(const bt_trace2d_args* a) {
    if (a == nullptr) return fail(BT_E_ARG, "args is NULL");
    if (a->n < 0 || a->n_active < 0) return fail(BT_E_ARG, "negative trajectory count");
    if (a->n_patches < 1 || a->n_patches > kMaxPatches)
        return fail(BT_E_ARG, "n_patches must be 1..16");
    if (a->n_planes < 0 || a->n_planes > kMaxPlanes) return fail(BT_E_ARG, "n_planes must be 0..256");
    if (a->cylindrical < 0 || a->cylindrical > 2) return fail(BT_E_ARG, "cylindrical must be 0, 1 or 2");
    if (a->integrator < 0 || a->integrator > 1) return fail(BT_E_ARG, "integrator must be 0 or 1");
    if (!(a->h > 0.0)) return fail(BT_E_ARG, "h must be positive");
    if (a->refine < 1) return fail(BT_E_ARG, "refine must be positive");
    if (a->capacity < 0 || a->path_capacity < 0 || a->n_paths < 0)
        return fail(BT_E_ARG, "negative buffer size");
    if (a->n >= kMax32 || a->capacity >= kMax32 || a->path_capacity >= kMax32 || a->n_paths >= kMax32 ||
        a->refine >= (1 << 15) || a->n_subcells >= kMax32 || a->charge_cols >= kMax32)
        return fail(BT_E_ARG, "trajectory, buffer, sub-cell and node counts must stay below 2^31");
    if (a->checkpoints != nullptr && (a->n_checkpoints >= kMax32 || a->checkpoint_interval >= kMax32))
        return fail(BT_E_ARG, "checkpoint counts must stay below 2^31");
    if (a->patch_int == nullptr || a->patch_real == nullptr)
        return fail(BT_E_ARG, "patch tables are NULL");
    if (a->n_planes > 0 && a->planes == nullptr) return fail(BT_E_ARG, "planes is NULL");
    if (a->graded != 0 && a->n_x_nodes < 2) return fail(BT_E_ARG, "graded axis needs 2+ nodes");
    if (a->n_fxs < 1 || a->n_fys < 1 || a->n_fxs >= ((i64)1 << 28) || a->n_fys >= ((i64)1 << 28))
        return fail(BT_E_ARG, "fxs and fys must hold 1 to 2^28 - 1 values");
    i64 nodes = 0;
    for (i64 q = 0; q < a->n_patches; ++q) {
        const int64_t* pi = a->patch_int + q * 6;
        const i64 nx = pi[0], ny = pi[1];
        if (nx < 2 || ny < 2) return fail(BT_E_ARG, "a patch has fewer than 2 nodes per axis");
        if (pi[2] < 0 || pi[2] + (nx + 1) * ny > a->n_fxs) return fail(BT_E_ARG, "fx offset beyond fxs");
        if (pi[3] < 0 || pi[3] + nx * (ny + 1) > a->n_fys) return fail(BT_E_ARG, "fy offset beyond fys");
        if ((nx + 1) * (ny + 1) >= ((i64)1 << 29)) return fail(BT_E_ARG, "a patch holds 2^29 nodes or more");
        if (nx >= kMaxAxis || ny >= kMaxAxis) return fail(BT_E_ARG, "a patch has 2^24 - 1 nodes per axis or more");
        if (pi[5] < 0 || pi[5] + nx * ny > a->charge_cols)
            return fail(BT_E_ARG, "node offset beyond charge");
        if (pi[5] + nx * ny > nodes) nodes = pi[5] + nx * ny;
        if (!(a->patch_real[q * 7 + 2] > 0.0)) return fail(BT_E_ARG, "a patch spacing is not positive");
    }
    const int64_t* base = a->patch_int + (a->n_patches - 1) * 6;
    if (a->cells_rows < base[0] - 1 || a->cells_cols < base[1] - 1)
        return fail(BT_E_ARG, "cells must be at least (nx - 1, ny - 1)");
    if (a->cells_rows * a->cells_cols >= ((i64)1 << 31)) return fail(BT_E_ARG, "cells holds 2^31 entries or more");
    if (a->cells_cols >= kMaxAxis) return fail(BT_E_ARG, "cells has 2^24 - 1 columns or more");
    if (a->n_subcells < 1) return fail(BT_E_ARG, "subcells is empty");
    if (a->charge_rows < 1) return fail(BT_E_ARG, "charge has no rows");
    return BT_OK;
}

bt::TraceParams make_params(const bt_trace2d_args* a) {
    bt::TraceParams p;
    std::memset(&p, 0, sizeof(p));
    p.active = a->active; p.n_active = a->n_active;
    p.x = a->x; p.v = a->v; p.n = (int)a->n;
    p.qm = a->qm; p.dq = a->dq;
    p.steps = (i64*)a->steps; p.status = (i64*)a->status; p.electrode = (i64*)a->electrode;
    p.fxs = a->fxs; p.fys = a->fys;
    p.n_fxs = a->n_fxs; p.n_fys = a->n_fys;
    p.n_patches = (int)a->n_patches;
    p.origin_x = a->origin_x; p.origin_y = a->origin_y; p.h = a->h;
    p.upper_x = a->upper_x; p.upper_y = a->upper_y;
    p.step_fraction = a->step_fraction; p.max_steps = a->max_steps;
    p.cells = (const i64*)a->cells; p.cells_rows = (int)a->cells_rows; p.cells_cols = (int)a->cells_cols;
    p.subcells = (const i64*)a->subcells; p.n_subcells = (int)a->n_subcells; p.refine = (int)a->refine;
    p.n_planes = (int)a->n_planes;
    p.hits = a->hits; p.capacity = (int)a->capacity; p.counts = (i64*)a->counts;
    p.path_slot = (const i64*)a->path_slot; p.paths = a->paths; p.n_paths = (int)a->n_paths;
    p.path_capacity = (int)a->path_capacity; p.path_count = (i64*)a->path_count;
    p.graded = (int)a->graded; p.x_nodes = a->x_nodes; p.n_x_nodes = a->n_x_nodes;
    p.inv_c2 = a->inv_c2;
    if (a->has_b) { p.b0 = a->b0; p.b1 = a->b1; p.b2 = a->b2; }
    const int64_t* base = a->patch_int + (a->n_patches - 1) * 6;
    const double* basr = a->patch_real + (a->n_patches - 1) * 7;
    p.nx = (int)base[0]; p.ny = (int)base[1];
    p.bofx = (int)base[2]; p.bofy = (int)base[3]; p.bofn = (int)base[5];
    p.box = basr[0]; p.boy = basr[1]; p.bh = basr[2];
    p.inv_bh = 1.0 / basr[2];
    p.inv_h = 1.0 / a->h;
    p.deposit_exponent = a->deposit_exponent;
    p.cell_imax = std::min(p.nx - 2, p.cells_rows - 1);
    p.cell_jmax = std::min(p.ny - 2, p.cells_cols - 1);
    p.reuse = a->n_patches == 1 && basr[0] == a->origin_x && basr[1] == a->origin_y && basr[2] == a->h;
    p.refill_block = (a->flags & BT_REFILL_BLOCK) != 0;
    if (a->checkpoints != nullptr && a->n_checkpoints > 0 && a->checkpoint_interval > 0) {
        p.ckpt = a->checkpoints; p.n_ckpt = (int)a->n_checkpoints; p.ckpt_interval = (int)a->checkpoint_interval;
    }
    return p;
}

i64 trace_nodes(const bt_trace2d_args* a) {
    i64 nodes = 0;
    for (i64 q = 0; q < a->n_patches; ++q) {
        const int64_t* pi = a->patch_int + q * 6;
        if (pi[5] + pi[0] * pi[1] > nodes) nodes = pi[5] + pi[0] * pi[1];
    }
    return nodes;
}

// Place the accumulator, queue and tables in the workspace, upload the tables, launch.
int run_trace(const bt_trace2d_args* a, bt::TraceParams p, void* ws, hipStream_t stream, int cus) {
    const i64 nodes = trace_nodes(a);
    char* base = static_cast<char*>(ws);
    p.acc = reinterpret_cast<u64*>(base);
    p.queue = reinterpret_cast<u64*>(base + node_bytes(nodes));
    char* tables = base + node_bytes(nodes) + 256;
    const size_t pi_bytes = (size_t)a->n_patches * 6 * 8, pr_bytes = (size_t)a->n_patches * 7 * 8;
    i64* d_pi = reinterpret_cast<i64*>(tables);
    double* d_pr = reinterpret_cast<double*>(tables + pi_bytes);
    double* d_pl = reinterpret_cast<double*>(tables + pi_bytes + pr_bytes);
    hipError_t err = hipMemcpyAsync(d_pi, a->patch_int, pi_bytes, hipMemcpyHostToDevice, stream);
    if (err == hipSuccess)
        err = hipMemcpyAsync(d_pr, a->patch_real, pr_bytes, hipMemcpyHostToDevice, stream);
    if (err == hipSuccess && a->n_planes > 0)
        err = hipMemcpyAsync(d_pl, a->planes, (size_t)a->n_planes * 8, hipMemcpyHostToDevice, stream);
    if (err != hipSuccess) return hip_fail(err, "table upload");
    p.patch_int = d_pi; p.patch_real = d_pr; p.planes = d_pl;

    const int integ = (int)a->integrator, geom = (int)a->cylindrical;
    const bool has_b = a->has_b != 0;
    // the uniform variant takes one patch that is the mesh itself (same origin and spacing)
    const bool general = a->n_patches > 1 || a->graded != 0 || a->inv_c2 != 0.0 || p.reuse == 0;
    const bool records = (a->flags & BT_NO_PATHS) == 0 || p.ckpt != nullptr;
    const bool wide = (a->flags & BT_WIDE_WINDOW) != 0;
    // two lanes per trajectory: without records, on one uniform patch (ignored otherwise)
    const bool pairs = (a->flags & BT_LANE_PAIRS) != 0 && !general && !records;
    const i64 lpt = pairs ? 2 : 1;
    int block = a->block_size > 0 ? a->block_size : 256;
    if (block % 64 != 0 || block > 256) return fail(BT_E_ARG, "block_size must be a multiple of 64, at most 256");
    int grid = 0;
    if (a->n_active > 0) {
        int per_cu = bt::trace_occupancy(integ, geom, has_b, general, records, wide, pairs, block);
        if (per_cu < 1) per_cu = 1;
        i64 full = (i64)per_cu * cus;
        i64 need = (a->n_active * lpt + block - 1) / block;
        grid = (int)(need < full ? need : full);
        if (a->max_blocks > 0 && grid > a->max_blocks) grid = a->max_blocks;
        if (grid < 1) grid = 1;
    }
    p.charge = a->charge;  // row 0
    int rc = bt::launch_trace(p, block, grid, stream, integ, geom, has_b, general, records, wide, pairs, nodes);
    if (rc) return fail(BT_E_LAUNCH, std::string("trace launch: ") + hipGetErrorString(hipGetLastError()));
    return BT_OK;
}

template <typename T>
int upload(DeviceState& s, int slot, const T* host, size_t count, T** dev, hipStream_t st) {
    int rc = ensure(s.buffers[slot], count * sizeof(T));
    if (rc) return rc;
    *dev = static_cast<T*>(s.buffers[slot].ptr);
    if (count > 0 && host != nullptr) {
        hipError_t err = hipMemcpyAsync(*dev, host, count * sizeof(T), hipMemcpyHostToDevice, st);
        if (err != hipSuccess) return hip_fail(err, "hipMemcpyAsync H2D");
    }
    return BT_OK;
}

template <typename T>
int download(T* host, const T* dev, size_t count, hipStream_t st) {
    if (count == 0 || host == nullptr) return BT_OK;
    hipError_t err = hipMemcpyAsync(host, dev, count * sizeof(T), hipMemcpyDeviceToHost, st);
    if (err != hipSuccess) return hip_fail(err, "hipMemcpyAsync D2H");
    return BT_OK;
}

int host_stream(DeviceState& s) {
    if (s.stream) return BT_OK;
    hipError_t err = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
    if (err != hipSuccess) return hip_fail(err, "hipStreamCreate");
    return BT_OK;
}

// A batch of designs: the validated design table on the device and the fixed-point
// workspace of all their nodes (allocated once, so that launches allocate nothing and can
// be captured in a graph).
struct BatchPlan {
    int device = 0;
    int cus = 0;
    std::vector<bt_batch_design> designs;
    bt::BatchDesign* d_designs = nullptr;
    void* workspace = nullptr;
    i64 total_nodes = 0, max_nodes = 0;
    i64 n_fxs = 0, n_fys = 0, n_cells = 0, n_subcells = 0;
};

struct GraphHandle {
    hipGraphExec_t exec = nullptr;
};

static_assert(sizeof(bt_batch_design) == sizeof(bt::BatchDesign), "bt_batch_design layout");

// Every design against the arena sizes; the node total and the largest design.
int check_designs(const bt_batch_design* designs, i64 n_designs, i64 n_fxs, i64 n_fys, i64 n_cells,
                  i64 n_subcells, i64* total_out, i64* most_out) {
    if (n_designs < 1 || designs == nullptr) return fail(BT_E_ARG, "no designs");
    if (n_fxs >= ((i64)1 << 28) || n_fys >= ((i64)1 << 28)) return fail(BT_E_ARG, "field arenas of 2^28 values or more");
    if (n_designs > 65535) return fail(BT_E_ARG, "at most 65535 designs");
    i64 total = 0, most = 0;
    for (i64 d = 0; d < n_designs; ++d) {
        const bt_batch_design& b = designs[d];
        const std::string at = "design " + std::to_string(d) + ": ";
        if (b.nx < 2 || b.ny < 2) return fail(BT_E_ARG, at + "fewer than 2 nodes per axis");
        if ((b.nx + 1) * (b.ny + 1) >= ((i64)1 << 29)) return fail(BT_E_ARG, at + "2^29 nodes or more");
        if (b.nx >= kMaxAxis || b.ny >= kMaxAxis || b.cells_cols >= kMaxAxis)
            return fail(BT_E_ARG, at + "2^24 - 1 nodes or cells per axis or more");
        if (!(b.h > 0.0)) return fail(BT_E_ARG, at + "h must be positive");
        if (b.fx_offset < 0 || b.fx_offset + (b.nx + 1) * b.ny > n_fxs) return fail(BT_E_ARG, at + "fx beyond fxs");
        if (b.fy_offset < 0 || b.fy_offset + b.nx * (b.ny + 1) > n_fys) return fail(BT_E_ARG, at + "fy beyond fys");
        if (b.cells_rows < b.nx - 1 || b.cells_cols < b.ny - 1)
            return fail(BT_E_ARG, at + "cells must be at least (nx - 1, ny - 1)");
        if (b.cells_offset < 0 || b.cells_offset + b.cells_rows * b.cells_cols > n_cells)
            return fail(BT_E_ARG, at + "cells beyond the cell arena");
        if (b.cells_rows * b.cells_cols >= ((i64)1 << 31)) return fail(BT_E_ARG, at + "2^31 cells or more");
        if (b.n_subcells < 1 || b.subcells_offset < 0 || b.subcells_offset + b.n_subcells > n_subcells)
            return fail(BT_E_ARG, at + "subcells beyond the sub-cell arena");
        if (b.n_nodes != b.nx * b.ny || b.node_offset < 0) return fail(BT_E_ARG, at + "n_nodes must be nx * ny");
        if (b.node_offset + b.n_nodes >= kMax32 || b.n_subcells >= kMax32)
            return fail(BT_E_ARG, at + "node and sub-cell counts must stay below 2^31");
        if (b.cells_offset + b.cells_rows * b.cells_cols >= kMax32 || b.subcells_offset + b.n_subcells >= kMax32)
            return fail(BT_E_ARG, at + "cell and sub-cell arena offsets must stay below 2^31");
        if (b.node_offset + b.n_nodes > total) total = b.node_offset + b.n_nodes;
        if (b.n_nodes > most) most = b.n_nodes;
    }
    *total_out = total;
    *most_out = most;
    return BT_OK;
}

int check_batch(const BatchPlan* plan, const bt_batch_args* a) {
    if (plan == nullptr) return fail(BT_E_ARG, "plan is NULL");
    if (a == nullptr) return fail(BT_E_ARG, "args is NULL");
    if (a->n < 0 || a->n_active < 0) return fail(BT_E_ARG, "negative trajectory count");
    if (a->n_planes < 0 || a->n_planes > kMaxPlanes) return fail(BT_E_ARG, "n_planes must be 0..256");
    if (a->cylindrical < 0 || a->cylindrical > 2) return fail(BT_E_ARG, "cylindrical must be 0, 1 or 2");
    if (a->integrator < 0 || a->integrator > 1) return fail(BT_E_ARG, "integrator must be 0 or 1");
    if (a->refine < 1) return fail(BT_E_ARG, "refine must be positive");
    if (a->capacity < 0) return fail(BT_E_ARG, "negative capacity");
    if (a->n >= kMax32 || a->capacity >= kMax32 || a->refine >= (1 << 15))
        return fail(BT_E_ARG, "trajectory and buffer counts must stay below 2^31");
    if (a->block_size != 0 && (a->block_size % 64 != 0 || a->block_size < 64 || a->block_size > 256))
        return fail(BT_E_ARG, "block_size must be a multiple of 64 in 64..256");
    if ((a->x0 == nullptr) != (a->v0 == nullptr)) return fail(BT_E_ARG, "x0 and v0 go together");
    if (a->n > 0 && (a->x == nullptr || a->v == nullptr || a->qm == nullptr || a->dq == nullptr ||
                     a->steps == nullptr || a->status == nullptr || a->electrode == nullptr ||
                     a->design_of == nullptr))
        return fail(BT_E_ARG, "a trajectory array is NULL");
    if (a->n_active > 0 && a->active == nullptr) return fail(BT_E_ARG, "active is NULL");
    if (a->n_planes > 0 && (a->planes == nullptr || a->counts == nullptr || (a->capacity > 0 && a->hits == nullptr)))
        return fail(BT_E_ARG, "plane arrays are NULL");
    if (a->fxs == nullptr || a->fys == nullptr || a->cells == nullptr || a->subcells == nullptr ||
        a->charge == nullptr)
        return fail(BT_E_ARG, "a field, lookup or charge array is NULL");
    return BT_OK;
}

bt::TraceParams batch_params(const BatchPlan* plan, const bt_batch_args* a) {
    bt::TraceParams p;
    std::memset(&p, 0, sizeof(p));
    p.active = a->active; p.n_active = a->n_active;
    p.x = a->x; p.v = a->v; p.n = (int)a->n;
    p.qm = a->qm; p.dq = a->dq;
    p.steps = (i64*)a->steps; p.status = (i64*)a->status; p.electrode = (i64*)a->electrode;
    p.fxs = a->fxs; p.fys = a->fys;
    p.n_fxs = plan->n_fxs; p.n_fys = plan->n_fys;
    p.n_patches = 1;
    p.step_fraction = a->step_fraction; p.max_steps = a->max_steps;
    p.cells = (const i64*)a->cells; p.subcells = (const i64*)a->subcells; p.refine = (int)a->refine;
    p.planes = a->planes; p.n_planes = (int)a->n_planes;
    p.hits = a->hits; p.capacity = (int)a->capacity; p.counts = (i64*)a->counts;
    p.reuse = 1;  // every design is one uniform patch: the mesh itself
    char* base = static_cast<char*>(plan->workspace);
    p.acc = reinterpret_cast<u64*>(base);
    p.queue = reinterpret_cast<u64*>(base + node_bytes(plan->total_nodes));
    p.charge = a->charge;
    p.designs = plan->d_designs; p.design_of = a->design_of; p.n_designs = (int)plan->designs.size();
    p.refill_block = (a->flags & BT_REFILL_BLOCK) != 0;
    // the magnetic arenas: each design's lanes add its node offset (load_design)
    p.b0 = a->b0; p.b1 = a->b1; p.b2 = a->b2;
    return p;
}

bool batch_has_b(const bt_batch_args* a) { return a->b0 != nullptr || a->b1 != nullptr || a->b2 != nullptr; }

int run_batch(const BatchPlan* plan, const bt_batch_args* a, hipStream_t stream) {
    const bt::TraceParams p = batch_params(plan, a);
    const int integ = (int)a->integrator, geom = (int)a->cylindrical;
    const int block = a->block_size > 0 ? a->block_size : 256;
    const bool wide = (a->flags & BT_WIDE_WINDOW) != 0;
    const bool pairs = (a->flags & BT_LANE_PAIRS) != 0;
    const bool has_b = batch_has_b(a);
    int grid = 0;
    if (a->n_active > 0) {
        int per_cu = bt::batched_occupancy(integ, geom, has_b, wide, pairs, block);
        if (per_cu < 1) per_cu = 1;
        const i64 full = (i64)per_cu * plan->cus;
        const i64 need = (a->n_active * (pairs ? 2 : 1) + block - 1) / block;
        grid = (int)(need < full ? need : full);
        if (a->max_blocks > 0 && grid > a->max_blocks) grid = a->max_blocks;
        if (grid < 1) grid = 1;
    }
    const bool reset = a->x0 != nullptr;
    if (bt::launch_batched(p, a->x0, a->v0, reset, block, grid, stream, integ, geom, has_b, wide, pairs, plan->total_nodes,
                           plan->max_nodes))
        return fail(BT_E_LAUNCH, std::string("batched launch: ") + hipGetErrorString(hipGetLastError()));
    return BT_OK;
}

// The pieces each group of 64 trajectories' backward walk is split into, when the groups
// outnumber the waves the grid holds at once: whole groups would leave the last round of
// waves partly filled (2e5 trajectories on 2048 waves: a second round of 1077 waves that
// costs nearly a full one); pieces of 1/P of a walk make the last round's share 1/P as large.
// Twelve pieces (fewer when the walks have fewer segments) measured best on the einzel lens's
// 24-segment walks (2e5 trajectories): 8 pieces 16.59 ms, 12 16.29, 16 16.60 (uneven pieces of
// one and two segments); on an earlier build 6 18.52, 10 18.50, 24 18.61 against 12's 18.30,
// and one piece 24.5 (2 pieces 22.0, 4 20.9).
int adjoint_pieces(i64 groups, i64 waves, i64 segments) {
    if (waves < 1 || groups <= waves || segments < 2) return 1;
    return (int)(segments < 12 ? segments : 12);
}

// The grid, checkpoint slots, groups and scratch bytes of an adjoint call: the queue head,
// the per-lane checkpoint and segment states, and (when the groups outnumber the resident
// waves) the groups' finished-piece counts and the adjoint states handed between pieces.
size_t adjoint_shape(i64 n_active, i64 max_step, i64 interval, int integ, int geom, int block, int cus, int* grid_out,
                     i64* n_ckpt_out, i64* groups_out = nullptr, size_t* hand_at = nullptr) {
    int grid = 0;
    if (n_active > 0) {
        int per_cu = bt::adjoint_occupancy(integ, geom, block);
        if (per_cu < 1) per_cu = 1;
        const i64 full = (i64)per_cu * cus;
        const i64 need = (n_active + block - 1) / block;
        grid = (int)(need < full ? need : full);
        if (grid < 1) grid = 1;
    }
    const i64 lanes = (i64)grid * block;
    const i64 n_ckpt = max_step > 0 ? (max_step - 1) / interval + 1 : 1;
    *grid_out = grid;
    *n_ckpt_out = n_ckpt;
    size_t bytes = 256 + (size_t)(n_ckpt + interval) * 6 * 8 * (size_t)(lanes > 0 ? lanes : 1);
    const i64 groups = (n_active + 63) / 64;
    if (groups_out != nullptr) *groups_out = groups;
    if (groups > lanes / 64) {
        bytes = (bytes + 255) / 256 * 256;
        if (hand_at != nullptr) *hand_at = bytes;  // done[groups], then hand[6][groups * 64]
        bytes += ((size_t)groups * 4 + 255) / 256 * 256 + (size_t)groups * 64 * 6 * 8;
    } else if (hand_at != nullptr) {
        *hand_at = 0;
    }
    return bytes;
}

}  // namespace

extern "C" {

const char* bt_version(void) { return "beamtracer_hip 0.1.0 (gfx950, fp64, fixed-point deposit)"; }

const char* bt_last_error(void) { return g_error.c_str(); }

int64_t bt_trace_args_size(void) { return (int64_t)sizeof(bt_trace2d_args); }

int32_t bt_deposit_exponent(double scale) {
    if (!(scale > 0.0) || !std::isfinite(scale)) return -1074 + 60;
    return (int32_t)std::ilogb(scale) - 40;
}

int64_t bt_trace_workspace_bytes(int64_t n_nodes) { return (int64_t)workspace_bytes(n_nodes); }

int bt_trace_trajectories_2d_device(const bt_trace2d_args* a, void* stream) {
    int rc = check_trace(a);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g_mutex);
    DeviceState* s = nullptr;
    if ((rc = device_state(&s))) return rc;
    const i64 nodes = trace_nodes(a);
    void* ws = a->workspace;
    if (ws == nullptr) {
        if ((rc = ensure(s->workspace, workspace_bytes(nodes)))) return rc;
        ws = s->workspace.ptr;
    } else if (a->workspace_bytes < (int64_t)workspace_bytes(nodes)) {
        return fail(BT_E_ARG, "workspace smaller than bt_trace_workspace_bytes()");
    }
    return run_trace(a, make_params(a), ws, static_cast<hipStream_t>(stream), s->cus);
}

int bt_trace_trajectories_2d_noalloc(const bt_trace2d_args* a, void* stream) {
    int rc = check_trace(a);
    if (rc) return rc;
    if (a->workspace != nullptr) return fail(BT_E_ARG, "the allocation-free trace uses its own scratch: workspace must be NULL");
    const size_t need = workspace_bytes(trace_nodes(a));
    if (need > kScratchBytes)
        return fail(BT_E_ARG, "the allocation-free trace holds at most " + std::to_string(BT_NOALLOC_MAX_NODES) +
                                  " mesh nodes (its scratch is " + std::to_string(kScratchBytes) + " bytes); " +
                                  std::to_string(trace_nodes(a)) + " given");
    std::lock_guard<std::mutex> lock(g_mutex);
    NoAllocState* s = nullptr;
    if ((rc = noalloc_state(&s))) return rc;
    return run_trace(a, make_params(a), s->scratch, static_cast<hipStream_t>(stream), s->cus);
}

int bt_trace_trajectories_2d(const bt_trace2d_args* a) {
    int rc = check_trace(a);
    if (rc) return rc;
    if (a->checkpoints != nullptr) return fail(BT_E_UNSUPPORTED, "checkpoints need the device form");
    std::lock_guard<std::mutex> lock(g_mutex);
    DeviceState* s = nullptr;
    if ((rc = device_state(&s))) return rc;
    if ((rc = host_stream(*s))) return rc;
    hipStream_t st = s->stream;
    const i64 n = a->n, nodes = trace_nodes(a);
    bt_trace2d_args d = *a;
    if (a->path_slot != nullptr) {  // no trajectory records a path: the variant without that code
        bool any = false;
        for (i64 k = 0; k < n && !any; ++k) any = a->path_slot[k] >= 0;
        if (!any) d.flags |= BT_NO_PATHS;
    }
    double *x, *v, *hits, *paths, *charge, *fxs, *fys, *xn, *b0 = nullptr, *b1 = nullptr, *b2 = nullptr;
    const double *qm, *dq;
    int64_t *active, *steps, *status, *electrode, *counts, *slot, *pcount, *cells, *subcells;
    const size_t nh = (size_t)n * a->n_planes * a->capacity * 8;
    const size_t np = (size_t)a->n_paths * a->path_capacity * 2;
    const int64_t* bpi = a->patch_int + (a->n_patches - 1) * 6;
    const size_t nb = (size_t)bpi[0] * bpi[1];
    double* qm_d; double* dq_d;
    if ((rc = upload(*s, 0, a->active, (size_t)a->n_active, &active, st))) return rc;
    if ((rc = upload(*s, 1, a->x, (size_t)(3 * n), &x, st))) return rc;
    if ((rc = upload(*s, 2, a->v, (size_t)(3 * n), &v, st))) return rc;
    if ((rc = upload(*s, 3, a->qm, (size_t)n, &qm_d, st))) return rc;
    if ((rc = upload(*s, 4, a->dq, (size_t)n, &dq_d, st))) return rc;
    qm = qm_d; dq = dq_d;
    if ((rc = upload(*s, 5, a->steps, (size_t)n, &steps, st))) return rc;
    if ((rc = upload(*s, 6, a->status, (size_t)n, &status, st))) return rc;
    if ((rc = upload(*s, 7, a->electrode, (size_t)n, &electrode, st))) return rc;
    if ((rc = upload(*s, 8, a->fxs, (size_t)a->n_fxs, &fxs, st))) return rc;
    if ((rc = upload(*s, 9, a->fys, (size_t)a->n_fys, &fys, st))) return rc;
    if ((rc = upload(*s, 10, a->cells, (size_t)(a->cells_rows * a->cells_cols), &cells, st))) return rc;
    if ((rc = upload(*s, 11, a->subcells, (size_t)a->n_subcells, &subcells, st))) return rc;
    if ((rc = upload(*s, 12, a->hits, nh, &hits, st))) return rc;
    if ((rc = upload(*s, 13, a->counts, (size_t)n * a->n_planes, &counts, st))) return rc;
    if ((rc = upload(*s, 14, a->path_slot, (size_t)n, &slot, st))) return rc;
    if ((rc = upload(*s, 15, a->paths, np, &paths, st))) return rc;
    if ((rc = upload(*s, 16, a->path_count, (size_t)a->n_paths, &pcount, st))) return rc;
    if ((rc = upload(*s, 17, (const double*)nullptr, (size_t)nodes, &charge, st))) return rc;
    if ((rc = upload(*s, 18, a->x_nodes, (size_t)a->n_x_nodes, &xn, st))) return rc;
    if (a->has_b) {
        if (a->b0 && (rc = upload(*s, 19, a->b0, nb, &b0, st))) return rc;
        if (a->b1 && (rc = upload(*s, 20, a->b1, nb, &b1, st))) return rc;
        if (a->b2 && (rc = upload(*s, 21, a->b2, nb, &b2, st))) return rc;
    }
    hipError_t err = hipMemsetAsync(charge, 0, (size_t)nodes * 8, st);
    if (err != hipSuccess) return hip_fail(err, "hipMemsetAsync");
    d.active = active; d.x = x; d.v = v; d.qm = qm; d.dq = dq; d.steps = steps; d.status = status;
    d.electrode = electrode; d.fxs = fxs; d.fys = fys; d.cells = cells; d.subcells = subcells;
    d.hits = hits; d.counts = counts; d.path_slot = slot; d.paths = paths; d.path_count = pcount;
    d.charge = charge; d.x_nodes = xn; d.b0 = b0; d.b1 = b1; d.b2 = b2;
    if ((rc = ensure(s->workspace, workspace_bytes(nodes)))) return rc;
    if ((rc = run_trace(&d, make_params(&d), s->workspace.ptr, st, s->cus))) return rc;
    std::vector<double> added((size_t)nodes);
    if ((rc = download(a->x, x, (size_t)(3 * n), st))) return rc;
    if ((rc = download(a->v, v, (size_t)(3 * n), st))) return rc;
    if ((rc = download(a->steps, steps, (size_t)n, st))) return rc;
    if ((rc = download(a->status, status, (size_t)n, st))) return rc;
    if ((rc = download(a->electrode, electrode, (size_t)n, st))) return rc;
    if ((rc = download(a->hits, hits, nh, st))) return rc;
    if ((rc = download(a->counts, counts, (size_t)n * a->n_planes, st))) return rc;
    if ((rc = download(a->paths, paths, np, st))) return rc;
    if ((rc = download(a->path_count, pcount, (size_t)a->n_paths, st))) return rc;
    if ((rc = download(added.data(), charge, (size_t)nodes, st))) return rc;
    err = hipStreamSynchronize(st);
    if (err != hipSuccess) return hip_fail(err, "trace");
    for (i64 c = 0; c < nodes; ++c) a->charge[c] += added[(size_t)c];
    return BT_OK;
}

int bt_staggered_field_2d_device(const double* px, const double* py, int64_t n, const double* fx,
                                 const double* fy, int64_t nx, int64_t ny, double origin_x,
                                 double origin_y, double h, double* ex, double* ey, void* stream) {
    if (n < 0) return fail(BT_E_ARG, "negative point count");
    if (nx < 2 || ny < 2) return fail(BT_E_ARG, "the mesh needs 2+ nodes per axis");
    if (!(h > 0.0)) return fail(BT_E_ARG, "h must be positive");
    int rc = bt::launch_gather(px, py, n, fx, fy, nx, ny, origin_x, origin_y, h, ex, ey,
                               static_cast<hipStream_t>(stream));
    if (rc) return fail(BT_E_LAUNCH, std::string("gather launch: ") + hipGetErrorString(hipGetLastError()));
    return BT_OK;
}

int bt_staggered_field_2d(const double* px, const double* py, int64_t n, const double* fx,
                          const double* fy, int64_t nx, int64_t ny, double origin_x, double origin_y,
                          double h, double* ex, double* ey) {
    if (n < 0) return fail(BT_E_ARG, "negative point count");
    if (nx < 2 || ny < 2) return fail(BT_E_ARG, "the mesh needs 2+ nodes per axis");
    if (!(h > 0.0)) return fail(BT_E_ARG, "h must be positive");
    std::lock_guard<std::mutex> lock(g_mutex);
    DeviceState* s = nullptr;
    int rc;
    if ((rc = device_state(&s))) return rc;
    if ((rc = host_stream(*s))) return rc;
    hipStream_t st = s->stream;
    double *dpx, *dpy, *dfx, *dfy, *dex, *dey;
    if ((rc = upload(*s, 0, px, (size_t)n, &dpx, st))) return rc;
    if ((rc = upload(*s, 1, py, (size_t)n, &dpy, st))) return rc;
    if ((rc = upload(*s, 2, fx, (size_t)((nx + 1) * ny), &dfx, st))) return rc;
    if ((rc = upload(*s, 3, fy, (size_t)(nx * (ny + 1)), &dfy, st))) return rc;
    if ((rc = upload(*s, 4, (const double*)nullptr, (size_t)n, &dex, st))) return rc;
    if ((rc = upload(*s, 5, (const double*)nullptr, (size_t)n, &dey, st))) return rc;
    if (bt::launch_gather(dpx, dpy, n, dfx, dfy, nx, ny, origin_x, origin_y, h, dex, dey, st))
        return fail(BT_E_LAUNCH, std::string("gather launch: ") + hipGetErrorString(hipGetLastError()));
    if ((rc = download(ex, dex, (size_t)n, st))) return rc;
    if ((rc = download(ey, dey, (size_t)n, st))) return rc;
    hipError_t err = hipStreamSynchronize(st);
    if (err != hipSuccess) return hip_fail(err, "gather");
    return BT_OK;
}

// ---- the 3D tracer ---------------------------------------------------------------------

int64_t bt_trace3d_args_size(void) { return (int64_t)sizeof(bt_trace3d_args); }

// The launch-independent checks of a 3D trace (host-readable tables only): every patch's
// fields and nodes within their arrays, the lookup the base mesh's shape.
static int check_trace3d(const bt_trace3d_args* a) {
    if (a == nullptr) return fail(BT_E_ARG, "args is NULL");
    if (a->n < 0 || a->n_active < 0) return fail(BT_E_ARG, "negative trajectory count");
    if (a->n >= kMax32 || a->n_active >= kMax32) return fail(BT_E_ARG, "2^31 or more trajectories");
    if (a->n_patches < 1 || a->n_patches > kMaxPatches) return fail(BT_E_ARG, "n_patches must be 1..16");
    if (a->patch_int == nullptr || a->patch_real == nullptr) return fail(BT_E_ARG, "patch tables are NULL");
    if (a->n_planes < 0 || a->n_planes > kMaxPlanes) return fail(BT_E_ARG, "n_planes must be 0..256");
    if (a->n_planes > 0 && a->planes == nullptr) return fail(BT_E_ARG, "planes is NULL");
    if (a->integrator < 0 || a->integrator > 1) return fail(BT_E_ARG, "integrator must be 0 or 1");
    if (!(a->h > 0.0)) return fail(BT_E_ARG, "h must be positive");
    if (a->refine < 1 || a->refine > 1024) return fail(BT_E_ARG, "refine must be 1..1024");
    if (a->capacity < 0 || a->path_capacity < 0 || a->n_paths < 0) return fail(BT_E_ARG, "negative buffer size");
    if (a->charge == nullptr || a->charge_rows < 1) return fail(BT_E_ARG, "charge needs a row");
    for (i64 q = 0; q < a->n_patches; ++q) {
        const int64_t* pi = a->patch_int + q * 8;
        const double ph = a->patch_real[q * 10 + 3];
        if (pi[0] < 2 || pi[1] < 2 || pi[2] < 2 || pi[0] > 65536 || pi[1] > 65536 || pi[2] > 65536)
            return fail(BT_E_ARG, "a patch needs 2..65536 nodes per axis");
        if (!(ph > 0.0)) return fail(BT_E_ARG, "a patch spacing must be positive");
        const i64 cells = pi[0] * pi[1] * pi[2];
        if (cells >= kMax32) return fail(BT_E_ARG, "a patch of 2^31 or more nodes");
        if (pi[3] < 0 || pi[3] + (pi[0] + 1) * pi[1] * pi[2] > a->n_fxs || pi[4] < 0 ||
            pi[4] + pi[0] * (pi[1] + 1) * pi[2] > a->n_fys || pi[5] < 0 || pi[5] + pi[0] * pi[1] * (pi[2] + 1) > a->n_fzs)
            return fail(BT_E_ARG, "a patch's edge fields exceed fxs, fys or fzs");
        if (pi[7] < 0 || pi[7] + cells > a->charge_cols) return fail(BT_E_ARG, "a patch's nodes exceed charge");
    }
    const int64_t* base = a->patch_int + (a->n_patches - 1) * 8;
    if (a->cells == nullptr || a->cells_n0 != base[0] - 1 || a->cells_n1 != base[1] - 1 || a->cells_n2 != base[2] - 1)
        return fail(BT_E_ARG, "cells must be (nx - 1, ny - 1, nz - 1) of the base mesh");
    if (a->subcells == nullptr || a->n_subcells < 1) return fail(BT_E_ARG, "subcells is empty");
    if (a->has_b != 0 && a->bn == nullptr) return fail(BT_E_ARG, "has_b without bn");
    const i64 nb = base[0] * base[1] * base[2];
    if (a->has_j != 0 && (a->current == nullptr || a->current_rows < 1 || a->current_cols < 3 * nb))
        return fail(BT_E_ARG, "has_j needs current of 3 nx ny nz columns");
    if (a->n > 0 && (a->x == nullptr || a->v == nullptr || a->qm == nullptr || a->dq == nullptr || a->steps == nullptr ||
                     a->status == nullptr || a->electrode == nullptr || a->path_slot == nullptr))
        return fail(BT_E_ARG, "null trajectory array");
    if (a->n_active > 0 && a->active == nullptr) return fail(BT_E_ARG, "active is NULL");
    return BT_OK;
}

static size_t trace3d_bytes(const bt_trace3d_args* a) {
    const i64 current_nodes = a->has_j != 0 ? a->current_cols : 0;
    return node_bytes(a->charge_cols) + node_bytes(current_nodes) + round256((size_t)a->n_patches * 8 * 8) +
           round256((size_t)a->n_patches * 10 * 8) + round256((size_t)(a->n_planes + 1) * 8);
}

int64_t bt_trace3d_workspace_bytes(const bt_trace3d_args* a) {
    return check_trace3d(a) ? -1 : (int64_t)trace3d_bytes(a);
}

// Place the words and tables in the workspace, upload the tables, launch.
static int run_trace3d(const bt_trace3d_args* a, void* ws, hipStream_t st) {
    char* w = static_cast<char*>(ws);
    const i64 current_nodes = a->has_j != 0 ? a->current_cols : 0;
    u64* acc_q = reinterpret_cast<u64*>(w);
    w += node_bytes(a->charge_cols);
    u64* acc_j = reinterpret_cast<u64*>(w);
    w += node_bytes(current_nodes);
    i64* d_pi = reinterpret_cast<i64*>(w);
    w += round256((size_t)a->n_patches * 8 * 8);
    double* d_pr = reinterpret_cast<double*>(w);
    w += round256((size_t)a->n_patches * 10 * 8);
    double* d_pl = reinterpret_cast<double*>(w);
    hipError_t err = hipMemcpyAsync(d_pi, a->patch_int, (size_t)a->n_patches * 64, hipMemcpyHostToDevice, st);
    if (err == hipSuccess) err = hipMemcpyAsync(d_pr, a->patch_real, (size_t)a->n_patches * 80, hipMemcpyHostToDevice, st);
    if (err == hipSuccess && a->n_planes > 0)
        err = hipMemcpyAsync(d_pl, a->planes, (size_t)a->n_planes * 8, hipMemcpyHostToDevice, st);
    if (err != hipSuccess) return hip_fail(err, "table upload");
    const int64_t* base = a->patch_int + (a->n_patches - 1) * 8;
    bt::Trace3DParams p;
    p.active = a->active; p.n_active = a->n_active; p.n = a->n;
    p.x = a->x; p.v = a->v; p.qm = a->qm; p.dq = a->dq;
    p.steps = a->steps; p.status = a->status; p.electrode = a->electrode;
    p.fxs = a->fxs; p.fys = a->fys; p.fzs = a->fzs;
    p.patch_int = d_pi; p.patch_real = d_pr; p.n_patches = (int)a->n_patches;
    p.bn = a->bn; p.has_b = a->has_b != 0;
    p.ox = a->origin_x; p.oy = a->origin_y; p.oz = a->origin_z; p.h = a->h;
    p.ux = a->upper_x; p.uy = a->upper_y; p.uz = a->upper_z;
    p.mirror_y = a->mirror_y != 0; p.mirror_z = a->mirror_z != 0;
    p.step_fraction = a->step_fraction; p.max_steps = a->max_steps;
    p.cells = a->cells; p.cells_n1 = (int)a->cells_n1; p.cells_n2 = (int)a->cells_n2;
    p.subcells = a->subcells; p.n_subcells = a->n_subcells; p.refine = (int)a->refine;
    p.planes = d_pl; p.n_planes = (int)a->n_planes;
    p.hits = a->hits; p.capacity = (int)(a->capacity < kMax32 ? a->capacity : kMax32 - 1); p.counts = a->counts;
    p.path_slot = a->path_slot; p.paths = a->paths; p.n_paths = a->n_paths;
    p.path_capacity = (int)(a->path_capacity < kMax32 ? a->path_capacity : kMax32 - 1); p.path_count = a->path_count;
    p.has_j = a->has_j != 0; p.inv_c2 = a->inv_c2;
    p.nx = (int)base[0]; p.ny = (int)base[1]; p.nz = (int)base[2];
    p.acc_q = acc_q; p.acc_j = acc_j; p.e_q = a->deposit_exponent; p.e_j = a->current_exponent;
    if (bt::launch_trace3d(p, (int)a->integrator, a->charge_cols, a->charge, current_nodes, a->current, st))
        return fail(BT_E_LAUNCH, std::string("3D trace launch: ") + hipGetErrorString(hipGetLastError()));
    return BT_OK;
}

int bt_trace_trajectories_3d_device(const bt_trace3d_args* a, void* stream) {
    int rc = check_trace3d(a);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g_mutex);
    DeviceState* s = nullptr;
    if ((rc = device_state(&s))) return rc;
    void* ws = a->workspace;
    if (ws == nullptr) {
        if ((rc = ensure(s->workspace, trace3d_bytes(a)))) return rc;
        ws = s->workspace.ptr;
    } else if (a->workspace_bytes < (int64_t)trace3d_bytes(a)) {
        return fail(BT_E_ARG, "workspace smaller than bt_trace3d_workspace_bytes()");
    }
    return run_trace3d(a, ws, static_cast<hipStream_t>(stream));
}

int bt_trace_trajectories_3d(const bt_trace3d_args* a) {
    int rc = check_trace3d(a);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g_mutex);
    DeviceState* s = nullptr;
    if ((rc = device_state(&s))) return rc;
    if ((rc = host_stream(*s))) return rc;
    hipStream_t st = s->stream;
    const i64 n = a->n;
    const int64_t* base = a->patch_int + (a->n_patches - 1) * 8;
    const size_t nb = (size_t)(base[0] * base[1] * base[2]);
    const size_t nh = (size_t)n * a->n_planes * a->capacity * 8, np = (size_t)a->n_paths * a->path_capacity * 3;
    const i64 current_nodes = a->has_j != 0 ? a->current_cols : 0;
    bt_trace3d_args d = *a;
    int64_t *active, *steps, *status, *electrode, *cells, *subcells, *counts, *slot, *pcount;
    double *x, *v, *qm, *dq, *fxs, *fys, *fzs, *bn = nullptr, *hits, *paths, *charge, *current = nullptr;
    if ((rc = upload(*s, 0, a->active, (size_t)a->n_active, &active, st))) return rc;
    if ((rc = upload(*s, 1, a->x, (size_t)(3 * n), &x, st))) return rc;
    if ((rc = upload(*s, 2, a->v, (size_t)(3 * n), &v, st))) return rc;
    if ((rc = upload(*s, 3, a->qm, (size_t)n, &qm, st))) return rc;
    if ((rc = upload(*s, 4, a->dq, (size_t)n, &dq, st))) return rc;
    if ((rc = upload(*s, 5, a->steps, (size_t)n, &steps, st))) return rc;
    if ((rc = upload(*s, 6, a->status, (size_t)n, &status, st))) return rc;
    if ((rc = upload(*s, 7, a->electrode, (size_t)n, &electrode, st))) return rc;
    if ((rc = upload(*s, 8, a->fxs, (size_t)a->n_fxs, &fxs, st))) return rc;
    if ((rc = upload(*s, 9, a->fys, (size_t)a->n_fys, &fys, st))) return rc;
    if ((rc = upload(*s, 10, a->fzs, (size_t)a->n_fzs, &fzs, st))) return rc;
    if (a->has_b != 0 && (rc = upload(*s, 11, a->bn, 3 * nb, &bn, st))) return rc;
    if ((rc = upload(*s, 12, a->cells, (size_t)(a->cells_n0 * a->cells_n1 * a->cells_n2), &cells, st))) return rc;
    if ((rc = upload(*s, 13, a->subcells, (size_t)a->n_subcells, &subcells, st))) return rc;
    if ((rc = upload(*s, 14, a->hits, nh, &hits, st))) return rc;
    if ((rc = upload(*s, 15, a->counts, (size_t)n * a->n_planes, &counts, st))) return rc;
    if ((rc = upload(*s, 16, a->path_slot, (size_t)n, &slot, st))) return rc;
    if ((rc = upload(*s, 17, a->paths, np, &paths, st))) return rc;
    if ((rc = upload(*s, 18, a->path_count, (size_t)a->n_paths, &pcount, st))) return rc;
    if ((rc = upload(*s, 19, (const double*)nullptr, (size_t)a->charge_cols, &charge, st))) return rc;
    if (current_nodes > 0 && (rc = upload(*s, 20, (const double*)nullptr, (size_t)current_nodes, &current, st)))
        return rc;
    hipError_t err = hipMemsetAsync(charge, 0, (size_t)a->charge_cols * 8, st);
    if (err == hipSuccess && current_nodes > 0) err = hipMemsetAsync(current, 0, (size_t)current_nodes * 8, st);
    if (err != hipSuccess) return hip_fail(err, "hipMemsetAsync");
    d.active = active; d.x = x; d.v = v; d.qm = qm; d.dq = dq; d.steps = steps; d.status = status;
    d.electrode = electrode; d.fxs = fxs; d.fys = fys; d.fzs = fzs; d.bn = bn; d.cells = cells;
    d.subcells = subcells; d.hits = hits; d.counts = counts; d.path_slot = slot; d.paths = paths;
    d.path_count = pcount; d.charge = charge; d.current = current;
    if ((rc = ensure(s->workspace, trace3d_bytes(&d)))) return rc;
    if ((rc = run_trace3d(&d, s->workspace.ptr, st))) return rc;
    std::vector<double> added_q((size_t)a->charge_cols), added_j((size_t)current_nodes);
    if ((rc = download(a->x, x, (size_t)(3 * n), st))) return rc;
    if ((rc = download(a->v, v, (size_t)(3 * n), st))) return rc;
    if ((rc = download(a->steps, steps, (size_t)n, st))) return rc;
    if ((rc = download(a->status, status, (size_t)n, st))) return rc;
    if ((rc = download(a->electrode, electrode, (size_t)n, st))) return rc;
    if ((rc = download(a->hits, hits, nh, st))) return rc;
    if ((rc = download(a->counts, counts, (size_t)n * a->n_planes, st))) return rc;
    if ((rc = download(a->paths, paths, np, st))) return rc;
    if ((rc = download(a->path_count, pcount, (size_t)a->n_paths, st))) return rc;
    if ((rc = download(added_q.data(), charge, (size_t)a->charge_cols, st))) return rc;
    if (current_nodes > 0 && (rc = download(added_j.data(), current, (size_t)current_nodes, st))) return rc;
    err = hipStreamSynchronize(st);
    if (err != hipSuccess) return hip_fail(err, "3D trace");
    for (i64 c = 0; c < a->charge_cols; ++c) a->charge[c] += added_q[(size_t)c];
    for (i64 c = 0; c < current_nodes; ++c) a->current[c] += added_j[(size_t)c];
    return BT_OK;
}

static int check_edge_field(const double* phi, const int32_t* solid, const uint8_t* near, const double* distance,
                            double h, int64_t nx, int64_t ny, const double* fx, const double* fy) {
    if (nx < 2 || ny < 2) return fail(BT_E_ARG, "the mesh needs 2+ nodes per axis");
    if (nx > (1ll << 28) || ny > (1ll << 28) || nx * ny > (1ll << 40)) return fail(BT_E_ARG, "mesh too large");
    if (!(h > 0.0)) return fail(BT_E_ARG, "h must be positive");
    if (!phi || !solid || !near || !distance || !fx || !fy) return fail(BT_E_ARG, "null array");
    return BT_OK;
}

int bt_edge_field_2d_device(const double* phi, const int32_t* solid, const uint8_t* near,
                            const double* distance, const double* dx, double h, int64_t nx, int64_t ny,
                            int32_t faces, double* fx, double* fy, void* stream) {
    int rc = check_edge_field(phi, solid, near, distance, h, nx, ny, fx, fy);
    if (rc) return rc;
    if (bt::launch_edge_field(phi, solid, near, distance, dx, h, nx, ny, faces, fx, fy,
                              static_cast<hipStream_t>(stream)))
        return fail(BT_E_LAUNCH, std::string("edge field launch: ") + hipGetErrorString(hipGetLastError()));
    return BT_OK;
}

int bt_edge_field_2d(const double* phi, const int32_t* solid, const uint8_t* near, const double* distance,
                     const double* dx, double h, int64_t nx, int64_t ny, int32_t faces, double* fx, double* fy) {
    int rc = check_edge_field(phi, solid, near, distance, h, nx, ny, fx, fy);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g_mutex);
    DeviceState* s = nullptr;
    if ((rc = device_state(&s))) return rc;
    if ((rc = host_stream(*s))) return rc;
    hipStream_t st = s->stream;
    const size_t nodes = (size_t)(nx * ny);
    double *dphi, *ddist, *ddx = nullptr, *dfx, *dfy;
    int32_t* dsolid;
    uint8_t* dnear;
    if ((rc = upload(*s, 0, phi, nodes, &dphi, st))) return rc;
    if ((rc = upload(*s, 1, solid, nodes, &dsolid, st))) return rc;
    if ((rc = upload(*s, 2, near, nodes, &dnear, st))) return rc;
    if ((rc = upload(*s, 3, distance, 4 * nodes, &ddist, st))) return rc;
    if (dx != nullptr && (rc = upload(*s, 4, dx, (size_t)(nx - 1), &ddx, st))) return rc;
    if ((rc = upload(*s, 5, (const double*)nullptr, (size_t)((nx + 1) * ny), &dfx, st))) return rc;
    if ((rc = upload(*s, 6, (const double*)nullptr, (size_t)(nx * (ny + 1)), &dfy, st))) return rc;
    if (bt::launch_edge_field(dphi, dsolid, dnear, ddist, ddx, h, nx, ny, faces, dfx, dfy, st))
        return fail(BT_E_LAUNCH, std::string("edge field launch: ") + hipGetErrorString(hipGetLastError()));
    if ((rc = download(fx, dfx, (size_t)((nx + 1) * ny), st))) return rc;
    if ((rc = download(fy, dfy, (size_t)(nx * (ny + 1)), st))) return rc;
    hipError_t err = hipStreamSynchronize(st);
    if (err != hipSuccess) return hip_fail(err, "edge field");
    return BT_OK;
}

int bt_deposit_trajectories_2d_device(const double* px, const double* py, const double* q, int64_t n,
                                      double origin_x, double origin_y, double h,
                                      int64_t cylindrical, double* charge, int64_t nx, int64_t ny,
                                      int32_t deposit_exponent, void* workspace,
                                      int64_t workspace_bytes_, void* stream) {
    if (n < 0) return fail(BT_E_ARG, "negative point count");
    if (nx < 2 || ny < 2) return fail(BT_E_ARG, "the mesh needs 2+ nodes per axis");
    if (!(h > 0.0)) return fail(BT_E_ARG, "h must be positive");
    void* ws = workspace;
    if (ws == nullptr) {
        std::lock_guard<std::mutex> lock(g_mutex);
        DeviceState* s = nullptr;
        int rc;
        if ((rc = device_state(&s))) return rc;
        if ((rc = ensure(s->workspace, workspace_bytes(nx * ny)))) return rc;
        ws = s->workspace.ptr;
    } else if (workspace_bytes_ < (int64_t)workspace_bytes(nx * ny)) {
        return fail(BT_E_ARG, "workspace smaller than bt_trace_workspace_bytes()");
    }
    if (bt::launch_deposit(px, py, q, n, origin_x, origin_y, h, cylindrical, charge, nx, ny,
                           deposit_exponent, static_cast<u64*>(ws), static_cast<hipStream_t>(stream)))
        return fail(BT_E_LAUNCH, std::string("deposit launch: ") + hipGetErrorString(hipGetLastError()));
    return BT_OK;
}

int bt_deposit_trajectories_2d(const double* px, const double* py, const double* q, int64_t n,
                               double origin_x, double origin_y, double h, int64_t cylindrical,
                               double* charge, int64_t nx, int64_t ny, int32_t deposit_exponent) {
    if (n < 0) return fail(BT_E_ARG, "negative point count");
    if (nx < 2 || ny < 2) return fail(BT_E_ARG, "the mesh needs 2+ nodes per axis");
    if (!(h > 0.0)) return fail(BT_E_ARG, "h must be positive");
    std::lock_guard<std::mutex> lock(g_mutex);
    DeviceState* s = nullptr;
    int rc;
    if ((rc = device_state(&s))) return rc;
    if ((rc = host_stream(*s))) return rc;
    hipStream_t st = s->stream;
    const i64 nodes = nx * ny;
    double *dpx, *dpy, *dq, *dc;
    if ((rc = upload(*s, 0, px, (size_t)n, &dpx, st))) return rc;
    if ((rc = upload(*s, 1, py, (size_t)n, &dpy, st))) return rc;
    if ((rc = upload(*s, 2, q, (size_t)n, &dq, st))) return rc;
    if ((rc = upload(*s, 3, (const double*)nullptr, (size_t)nodes, &dc, st))) return rc;
    hipError_t err = hipMemsetAsync(dc, 0, (size_t)nodes * 8, st);
    if (err != hipSuccess) return hip_fail(err, "hipMemsetAsync");
    if ((rc = ensure(s->workspace, workspace_bytes(nodes)))) return rc;
    if (bt::launch_deposit(dpx, dpy, dq, n, origin_x, origin_y, h, cylindrical, dc, nx, ny,
                           deposit_exponent, static_cast<u64*>(s->workspace.ptr), st))
        return fail(BT_E_LAUNCH, std::string("deposit launch: ") + hipGetErrorString(hipGetLastError()));
    std::vector<double> added((size_t)nodes);
    if ((rc = download(added.data(), dc, (size_t)nodes, st))) return rc;
    err = hipStreamSynchronize(st);
    if (err != hipSuccess) return hip_fail(err, "deposit");
    for (i64 c = 0; c < nodes; ++c) charge[c] += added[(size_t)c];
    return BT_OK;
}


int bt_l2_gather_bandwidth(const double* field, int64_t rows, int64_t cols, int32_t lookups, int32_t blocks,
                           double* sink, void* stream) {
    if (field == nullptr || sink == nullptr || rows < 2 || cols < 2 || lookups < 1 || blocks < 1)
        return fail(BT_E_ARG, "bad arguments");
    if (rows * cols >= ((i64)1 << 28)) return fail(BT_E_ARG, "the field must hold fewer than 2^28 values");
    if (bt::launch_l2_gather(field, rows, cols, lookups, blocks, 0x5eedu, sink, static_cast<hipStream_t>(stream)))
        return fail(BT_E_LAUNCH, std::string("gather launch: ") + hipGetErrorString(hipGetLastError()));
    return BT_OK;
}

int bt_l2_line_bandwidth(const double* field, int64_t n_lines, int32_t lines, int32_t blocks, double* sink,
                         void* stream) {
    if (field == nullptr || sink == nullptr || n_lines < 1 || lines < 1 || blocks < 1)
        return fail(BT_E_ARG, "bad arguments");
    if (n_lines >= ((i64)1 << 24)) return fail(BT_E_ARG, "the field must hold fewer than 2^24 lines");
    if (bt::launch_l2_line(field, n_lines, lines, blocks, 0x1234u, sink, static_cast<hipStream_t>(stream)))
        return fail(BT_E_LAUNCH, std::string("line launch: ") + hipGetErrorString(hipGetLastError()));
    return BT_OK;
}

int bt_arith_selftest(const double* y, const double* z, int64_t n, double* r, double* cy, double* cz, double* rs,
                      void* stream) {
    if (n < 0 || (n > 0 && (y == nullptr || z == nullptr || r == nullptr || cy == nullptr || cz == nullptr ||
                            rs == nullptr)))
        return fail(BT_E_ARG, "bad arguments");
    if (n == 0) return BT_OK;
    if (bt::launch_arith_selftest(y, z, n, r, cy, cz, rs, static_cast<hipStream_t>(stream)))
        return fail(BT_E_LAUNCH, std::string("selftest launch: ") + hipGetErrorString(hipGetLastError()));
    return BT_OK;
}

int64_t bt_adjoint_args_size(void) { return (int64_t)sizeof(bt_adjoint_args); }

int64_t bt_adjoint_workspace_bytes(int64_t n_active, int64_t max_step, int32_t interval, int64_t integrator,
                                   int64_t cylindrical, int32_t block_size) {
    if (n_active < 0 || max_step < 0 || interval < 1 || integrator < 0 || integrator > 1 || cylindrical < 0 ||
        cylindrical > 2)
        return fail(BT_E_ARG, "bad arguments");
    const int block = block_size > 0 ? block_size : 256;
    std::lock_guard<std::mutex> lock(g_mutex);
    DeviceState* s = nullptr;
    int rc = device_state(&s);
    if (rc) return rc;
    int grid = 0;
    i64 n_ckpt = 1;
    return (int64_t)adjoint_shape(n_active, max_step, interval, (int)integrator, (int)cylindrical, block, s->cus,
                                  &grid, &n_ckpt);
}

int bt_trace_adjoint_device(const bt_adjoint_args* a, void* stream) {
    if (a == nullptr) return fail(BT_E_ARG, "args is NULL");
    if (a->n < 0 || a->n_active < 0) return fail(BT_E_ARG, "negative trajectory count");
    if (a->nx < 2 || a->ny < 2) return fail(BT_E_ARG, "fewer than 2 nodes per axis");
    if ((a->nx + 1) * (a->ny + 1) >= ((i64)1 << 29)) return fail(BT_E_ARG, "2^29 nodes or more");
    if (a->nx >= kMaxAxis || a->ny >= kMaxAxis) return fail(BT_E_ARG, "2^24 - 1 nodes per axis or more");
    if (!(a->h > 0.0)) return fail(BT_E_ARG, "h must be positive");
    if (a->cylindrical < 0 || a->cylindrical > 2) return fail(BT_E_ARG, "cylindrical must be 0, 1 or 2");
    if (a->integrator < 0 || a->integrator > 1) return fail(BT_E_ARG, "integrator must be 0 or 1");
    if (a->n_planes < 1 || a->n_planes > kMaxPlanes) return fail(BT_E_ARG, "n_planes must be 1..256");
    if (a->capacity < 1) return fail(BT_E_ARG, "capacity must be positive");
    if (a->interval < 1) return fail(BT_E_ARG, "interval must be positive");
    if (a->max_step < 0) return fail(BT_E_ARG, "negative max_step");
    if (a->n >= ((i64)1 << 31) || a->max_step + a->interval >= ((i64)1 << 31))
        return fail(BT_E_ARG, "2^31 trajectories or steps or more");
    if (a->block_size != 0 && (a->block_size % 64 != 0 || a->block_size < 64 || a->block_size > 256))
        return fail(BT_E_ARG, "block_size must be a multiple of 64 in 64..256");
    if (a->n_active > 0 && (a->active == nullptr || a->x == nullptr || a->v == nullptr || a->qm == nullptr ||
                            a->fx == nullptr || a->fy == nullptr || a->planes == nullptr || a->hits == nullptr ||
                            a->counts == nullptr || a->seeds == nullptr || a->dfx == nullptr || a->dfy == nullptr))
        return fail(BT_E_ARG, "an array is NULL");
    std::lock_guard<std::mutex> lock(g_mutex);
    DeviceState* s = nullptr;
    int rc = device_state(&s);
    if (rc) return rc;
    const int integ = (int)a->integrator, geom = (int)a->cylindrical;
    const int block = a->block_size > 0 ? a->block_size : 256;
    int grid = 0;
    i64 n_ckpt = 1, groups = 0;
    size_t hand_at = 0;
    const size_t bytes = adjoint_shape(a->n_active, a->max_step, a->interval, integ, geom, block, s->cus, &grid,
                                       &n_ckpt, &groups, &hand_at);
    const i64 lanes = (i64)grid * block;
    void* ws = a->workspace;
    if (ws == nullptr) {
        if ((rc = ensure(s->adjoint, bytes))) return rc;
        ws = s->adjoint.ptr;
    } else if (a->workspace_bytes < (int64_t)bytes) {
        return fail(BT_E_ARG, "workspace smaller than bt_adjoint_workspace_bytes()");
    }
    bt::AdjointParams p;
    std::memset(&p, 0, sizeof(p));
    p.active = (const i64*)a->active; p.n_active = a->n_active; p.n = a->n;
    p.x = a->x; p.v = a->v; p.qm = a->qm;
    p.fxs = a->fx; p.fys = a->fy; p.nx = a->nx; p.ny = a->ny;
    p.origin_x = a->origin_x; p.origin_y = a->origin_y; p.h = a->h; p.step_fraction = a->step_fraction;
    p.inv_h = 1.0 / a->h;
    p.planes = a->planes; p.n_planes = (int)a->n_planes;
    p.hits = a->hits; p.capacity = a->capacity; p.counts = (const i64*)a->counts; p.seeds = a->seeds;
    p.interval = a->interval; p.n_ckpt = n_ckpt; p.lanes = lanes;
    if (a->checkpoints != nullptr && a->n_checkpoints > 0) {
        p.fwd_ckpt = a->checkpoints;
        p.fwd_n_ckpt = a->n_checkpoints;
    }
    p.queue = static_cast<u64*>(ws);
    p.scratch = reinterpret_cast<double*>(static_cast<char*>(ws) + 256);
    p.dfx = a->dfx; p.dfy = a->dfy; p.dx = a->dx; p.dv = a->dv;
    p.dropped = a->dropped;
    // pieces need the forward trace's checkpoints up to every seeded step (another lane
    // continues a walk, so no lane's own checkpoints can serve)
    p.pieces = 1;
    p.groups = groups;
    if (hand_at != 0 && p.fwd_ckpt != nullptr && p.fwd_n_ckpt >= n_ckpt) {
        p.pieces = adjoint_pieces(groups, lanes / 64, n_ckpt);
        p.done = reinterpret_cast<int*>(static_cast<char*>(ws) + hand_at);
        p.hand = reinterpret_cast<double*>(static_cast<char*>(ws) + hand_at + ((size_t)groups * 4 + 255) / 256 * 256);
    }
    if (bt::launch_adjoint(p, block, grid, static_cast<hipStream_t>(stream), integ, geom))
        return fail(BT_E_LAUNCH, std::string("adjoint launch: ") + hipGetErrorString(hipGetLastError()));
    return BT_OK;
}

int64_t bt_batch_args_size(void) { return (int64_t)sizeof(bt_batch_args); }

int64_t bt_batch_design_size(void) { return (int64_t)sizeof(bt_batch_design); }

int bt_batch_plan_create(const bt_batch_design* designs, int64_t n_designs, int64_t n_fxs, int64_t n_fys,
                         int64_t n_cells, int64_t n_subcells, void** plan_out) {
    if (plan_out == nullptr) return fail(BT_E_ARG, "plan is NULL");
    *plan_out = nullptr;
    i64 total = 0, most = 0;
    int rc = check_designs(designs, n_designs, n_fxs, n_fys, n_cells, n_subcells, &total, &most);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g_mutex);
    DeviceState* s = nullptr;
    if ((rc = device_state(&s))) return rc;
    BatchPlan* plan = new BatchPlan();
    plan->designs.assign(designs, designs + n_designs);
    plan->total_nodes = total;
    plan->max_nodes = most;
    plan->n_fxs = n_fxs; plan->n_fys = n_fys; plan->n_cells = n_cells; plan->n_subcells = n_subcells;
    plan->cus = s->cus;
    (void)hipGetDevice(&plan->device);
    hipError_t err = hipMalloc(reinterpret_cast<void**>(&plan->d_designs), sizeof(bt::BatchDesign) * (size_t)n_designs);
    if (err == hipSuccess) err = hipMalloc(&plan->workspace, node_bytes(total) + 256);
    if (err == hipSuccess)
        err = hipMemcpy(plan->d_designs, designs, sizeof(bt::BatchDesign) * (size_t)n_designs, hipMemcpyHostToDevice);
    if (err != hipSuccess) {
        if (plan->d_designs) (void)hipFree(plan->d_designs);
        if (plan->workspace) (void)hipFree(plan->workspace);
        delete plan;
        return fail(BT_E_NOMEM, std::string("batch plan: ") + hipGetErrorString(err));
    }
    *plan_out = plan;
    return BT_OK;
}

int bt_batch_plan_destroy(void* handle) {
    BatchPlan* plan = static_cast<BatchPlan*>(handle);
    if (plan == nullptr) return BT_OK;
    (void)hipDeviceSynchronize();
    if (plan->d_designs) (void)hipFree(plan->d_designs);
    if (plan->workspace) (void)hipFree(plan->workspace);
    delete plan;
    return BT_OK;
}

int bt_batch_plan_update(void* handle, const bt_batch_design* designs, int64_t n_designs) {
    BatchPlan* plan = static_cast<BatchPlan*>(handle);
    if (plan == nullptr) return fail(BT_E_ARG, "plan is NULL");
    if (n_designs != (i64)plan->designs.size()) return fail(BT_E_ARG, "the design count differs from the plan's");
    i64 total = 0, most = 0;
    int rc = check_designs(designs, n_designs, plan->n_fxs, plan->n_fys, plan->n_cells, plan->n_subcells,
                           &total, &most);
    if (rc) return rc;
    if (total > plan->total_nodes) return fail(BT_E_ARG, "the designs hold more nodes than the plan");
    hipError_t err = hipDeviceSynchronize();
    if (err == hipSuccess)
        err = hipMemcpy(plan->d_designs, designs, sizeof(bt::BatchDesign) * (size_t)n_designs, hipMemcpyHostToDevice);
    if (err != hipSuccess) return hip_fail(err, "bt_batch_plan_update");
    plan->designs.assign(designs, designs + n_designs);
    plan->max_nodes = most > plan->max_nodes ? most : plan->max_nodes;
    return BT_OK;
}

int64_t bt_batch_plan_nodes(void* handle) {
    const BatchPlan* plan = static_cast<const BatchPlan*>(handle);
    return plan ? plan->total_nodes : -1;
}

int bt_batch_occupancy(int64_t integrator, int64_t cylindrical, int32_t block_size, int32_t flags) {
    if (integrator < 0 || integrator > 1 || cylindrical < 0 || cylindrical > 2) return fail(BT_E_ARG, "bad variant");
    if (block_size % 64 != 0 || block_size < 64 || block_size > 256) return fail(BT_E_ARG, "bad block size");
    return bt::batched_occupancy((int)integrator, (int)cylindrical, (flags & BT_MAGNETIC) != 0,
                                 (flags & BT_WIDE_WINDOW) != 0, (flags & BT_LANE_PAIRS) != 0, block_size);
}

int bt_batch_trace_device(void* handle, const bt_batch_args* a, void* stream) {
    const BatchPlan* plan = static_cast<const BatchPlan*>(handle);
    int rc = check_batch(plan, a);
    if (rc) return rc;
    return run_batch(plan, a, static_cast<hipStream_t>(stream));
}

int bt_batch_graph_create(void* handle, const bt_batch_args* a, void** graph_out) {
    const BatchPlan* plan = static_cast<const BatchPlan*>(handle);
    if (graph_out == nullptr) return fail(BT_E_ARG, "graph is NULL");
    *graph_out = nullptr;
    int rc = check_batch(plan, a);
    if (rc) return rc;
    hipStream_t st = nullptr;
    hipError_t err = hipStreamCreate(&st);
    if (err != hipSuccess) return hip_fail(err, "hipStreamCreate");
    hipGraph_t graph = nullptr;
    err = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (err != hipSuccess) {
        (void)hipStreamDestroy(st);
        return hip_fail(err, "hipStreamBeginCapture");
    }
    rc = run_batch(plan, a, st);
    err = hipStreamEndCapture(st, &graph);
    (void)hipStreamDestroy(st);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (err != hipSuccess) return hip_fail(err, "hipStreamEndCapture");
    GraphHandle* g = new GraphHandle();
    err = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (err != hipSuccess) {
        delete g;
        return hip_fail(err, "hipGraphInstantiate");
    }
    *graph_out = g;
    return BT_OK;
}

int bt_graph_launch(void* handle, void* stream) {
    GraphHandle* g = static_cast<GraphHandle*>(handle);
    if (g == nullptr || g->exec == nullptr) return fail(BT_E_ARG, "graph is NULL");
    hipError_t err = hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream));
    if (err != hipSuccess) return hip_fail(err, "hipGraphLaunch");
    return BT_OK;
}

int bt_graph_destroy(void* handle) {
    GraphHandle* g = static_cast<GraphHandle*>(handle);
    if (g == nullptr) return BT_OK;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    delete g;
    return BT_OK;
}

}  // extern "C"

// ---- device arrays owned by the library -------------------------------------------------
// The host code's device-resident state lives in memory allocated here; the helpers below
// are the few operations it needs between traces.  Copies to and from the host wait for
// `stream`; the helpers that report a count or a value wait too.

namespace {

// the per-device counter words of the helpers (zeroed before each use), g_mutex held
int helper_counters(u64** out) {
    DeviceState* s = nullptr;
    int rc = device_state(&s);
    if (rc) return rc;
    rc = ensure(s->counters, 256);
    if (rc) return rc;
    *out = static_cast<u64*>(s->counters.ptr);
    return BT_OK;
}

int read_counter(const u64* dev, u64* host, hipStream_t stream, const char* what) {
    hipError_t err = hipMemcpyAsync(host, dev, sizeof(u64), hipMemcpyDeviceToHost, stream);
    if (err == hipSuccess) err = hipStreamSynchronize(stream);
    return err == hipSuccess ? BT_OK : hip_fail(err, what);
}

int bad_indices(u64 bad) {
    return bad ? fail(BT_E_ARG, std::to_string(bad) + " index(es) out of range") : BT_OK;
}

}  // namespace

extern "C" {

int bt_compute_units(void) {
    std::lock_guard<std::mutex> lock(g_mutex);
    DeviceState* s = nullptr;
    int rc = device_state(&s);
    return rc ? rc : s->cus;
}

int bt_device_alloc(int64_t bytes, void** out) {
    if (out == nullptr) return fail(BT_E_ARG, "out is NULL");
    *out = nullptr;
    if (bytes < 0) return fail(BT_E_ARG, "negative size");
    if (bytes == 0) return BT_OK;
    hipError_t err = hipMalloc(out, (size_t)bytes);
    if (err != hipSuccess) {
        *out = nullptr;
        return fail(BT_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(err));
    }
    return BT_OK;
}

int bt_device_free(void* ptr) {
    if (ptr == nullptr) return BT_OK;
    hipError_t err = hipFree(ptr);
    return err == hipSuccess ? BT_OK : hip_fail(err, "hipFree");
}

int bt_copy_to_device(void* dst, const void* src, int64_t bytes, void* stream) {
    if (bytes < 0) return fail(BT_E_ARG, "negative size");
    if (bytes == 0) return BT_OK;
    if (dst == nullptr || src == nullptr) return fail(BT_E_ARG, "NULL pointer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t err = hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, st);
    if (err == hipSuccess) err = hipStreamSynchronize(st);
    return err == hipSuccess ? BT_OK : hip_fail(err, "copy to device");
}

int bt_copy_to_host(void* dst, const void* src, int64_t bytes, void* stream) {
    if (bytes < 0) return fail(BT_E_ARG, "negative size");
    if (bytes == 0) return BT_OK;
    if (dst == nullptr || src == nullptr) return fail(BT_E_ARG, "NULL pointer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t err = hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, st);
    if (err == hipSuccess) err = hipStreamSynchronize(st);
    return err == hipSuccess ? BT_OK : hip_fail(err, "copy to host");
}

int bt_copy_on_device(void* dst, const void* src, int64_t bytes, void* stream) {
    if (bytes < 0) return fail(BT_E_ARG, "negative size");
    if (bytes == 0) return BT_OK;
    if (dst == nullptr || src == nullptr) return fail(BT_E_ARG, "NULL pointer");
    hipError_t err = hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream));
    return err == hipSuccess ? BT_OK : hip_fail(err, "copy on device");
}

int bt_fill_i64(int64_t* dst, int64_t value, int64_t n, void* stream) {
    if (n < 0 || (n > 0 && dst == nullptr)) return fail(BT_E_ARG, "bad array");
    return bt::launch_fill_i64(dst, value, n, static_cast<hipStream_t>(stream)) ? hip_fail(hipGetLastError(), "fill")
                                                                                : BT_OK;
}

int bt_fill_f64(double* dst, double value, int64_t n, void* stream) {
    if (n < 0 || (n > 0 && dst == nullptr)) return fail(BT_E_ARG, "bad array");
    return bt::launch_fill_f64(dst, value, n, static_cast<hipStream_t>(stream)) ? hip_fail(hipGetLastError(), "fill")
                                                                                : BT_OK;
}

int bt_synchronize(void* stream) {
    hipError_t err = hipStreamSynchronize(static_cast<hipStream_t>(stream));
    return err == hipSuccess ? BT_OK : hip_fail(err, "hipStreamSynchronize");
}

int bt_abs_max_f64(const double* a, int64_t n, double* out, void* stream) {
    if (out == nullptr || n < 0 || (n > 0 && a == nullptr)) return fail(BT_E_ARG, "bad array");
    std::lock_guard<std::mutex> lock(g_mutex);
    u64* counters = nullptr;
    int rc = helper_counters(&counters);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t err = hipMemsetAsync(counters, 0, sizeof(u64), st);
    if (err != hipSuccess) return hip_fail(err, "hipMemsetAsync");
    if (bt::launch_abs_max(a, n, counters, st)) return hip_fail(hipGetLastError(), "abs max");
    u64 bits = 0;
    rc = read_counter(counters, &bits, st, "abs max");
    if (rc) return rc;
    std::memcpy(out, &bits, sizeof(double));
    return BT_OK;
}

int bt_max_i64(const int64_t* a, int64_t n, int64_t* out, void* stream) {
    if (out == nullptr || n < 1 || a == nullptr) return fail(BT_E_ARG, "bad array");
    std::lock_guard<std::mutex> lock(g_mutex);
    u64* counters = nullptr;
    int rc = helper_counters(&counters);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t err = hipMemsetAsync(counters, 0, sizeof(u64), st);
    if (err != hipSuccess) return hip_fail(err, "hipMemsetAsync");
    if (bt::launch_max_i64(a, n, counters, st)) return hip_fail(hipGetLastError(), "max");
    u64 key = 0;
    rc = read_counter(counters, &key, st, "max");
    *out = (int64_t)(key ^ 0x8000000000000000ull);
    return rc;
}

int bt_select_equal_i64(const int64_t* a, int64_t n, int64_t value, int64_t* idx, int64_t* count, void* stream) {
    if (count == nullptr || n < 0 || (n > 0 && (a == nullptr || idx == nullptr))) return fail(BT_E_ARG, "bad array");
    std::lock_guard<std::mutex> lock(g_mutex);
    u64* counters = nullptr;
    int rc = helper_counters(&counters);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t err = hipMemsetAsync(counters, 0, sizeof(u64), st);
    if (err != hipSuccess) return hip_fail(err, "hipMemsetAsync");
    if (bt::launch_select_equal(a, n, value, idx, counters, st)) return hip_fail(hipGetLastError(), "select");
    u64 found = 0;
    rc = read_counter(counters, &found, st, "select");
    *count = (int64_t)found;
    return rc;
}

int bt_gather_columns_f64(const double* src, int64_t rows, int64_t n, const int64_t* idx, int64_t m, double* dst,
                          void* stream) {
    if (rows < 0 || n < 0 || m < 0) return fail(BT_E_ARG, "negative size");
    if (rows == 0 || m == 0) return BT_OK;
    if (src == nullptr || idx == nullptr || dst == nullptr) return fail(BT_E_ARG, "NULL pointer");
    std::lock_guard<std::mutex> lock(g_mutex);
    u64* counters = nullptr;
    int rc = helper_counters(&counters);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t err = hipMemsetAsync(counters, 0, sizeof(u64), st);
    if (err != hipSuccess) return hip_fail(err, "hipMemsetAsync");
    if (bt::launch_gather_columns(src, rows, n, idx, m, dst, counters, st)) return hip_fail(hipGetLastError(), "gather");
    u64 bad = 0;
    rc = read_counter(counters, &bad, st, "gather");
    return rc ? rc : bad_indices(bad);
}

int bt_scatter_i64(int64_t* dst, int64_t n, const int64_t* idx, const int64_t* values, int64_t m, void* stream) {
    if (n < 0 || m < 0) return fail(BT_E_ARG, "negative size");
    if (m == 0) return BT_OK;
    if (dst == nullptr || idx == nullptr || values == nullptr) return fail(BT_E_ARG, "NULL pointer");
    std::lock_guard<std::mutex> lock(g_mutex);
    u64* counters = nullptr;
    int rc = helper_counters(&counters);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t err = hipMemsetAsync(counters, 0, sizeof(u64), st);
    if (err != hipSuccess) return hip_fail(err, "hipMemsetAsync");
    if (bt::launch_scatter(dst, n, idx, values, m, counters, st)) return hip_fail(hipGetLastError(), "scatter");
    u64 bad = 0;
    rc = read_counter(counters, &bad, st, "scatter");
    return rc ? rc : bad_indices(bad);
}

int bt_gather_records(const double* hits, int64_t n, int64_t n_planes, int64_t capacity, int64_t plane,
                      const int64_t* trajectory, const int64_t* slot, int64_t m, double* dst, void* stream) {
    if (n < 0 || n_planes < 1 || capacity < 1 || m < 0) return fail(BT_E_ARG, "bad record shape");
    if (plane < 0 || plane >= n_planes) return fail(BT_E_ARG, "plane out of range");
    if (m == 0) return BT_OK;
    if (hits == nullptr || trajectory == nullptr || slot == nullptr || dst == nullptr)
        return fail(BT_E_ARG, "NULL pointer");
    std::lock_guard<std::mutex> lock(g_mutex);
    u64* counters = nullptr;
    int rc = helper_counters(&counters);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t err = hipMemsetAsync(counters, 0, sizeof(u64), st);
    if (err != hipSuccess) return hip_fail(err, "hipMemsetAsync");
    if (bt::launch_gather_records(hits, n, n_planes, capacity, plane, trajectory, slot, m, dst, counters, st))
        return hip_fail(hipGetLastError(), "gather records");
    u64 bad = 0;
    rc = read_counter(counters, &bad, st, "gather records");
    return rc ? rc : bad_indices(bad);
}

}  // extern "C"
