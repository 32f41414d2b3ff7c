// The staggered edge fields of a potential for gfx950: the input of the trace.
//
// E_x on the nx + 1 edges along x of each of the ny lines (fx, (nx + 1, ny)) and E_r on the
// ny + 1 edges along r of each of the nx lines (fy, (nx, ny + 1)), from the node potential
// (solver/fields.py:18-67, called by StaggeredElectricField, fields.py:116-152):
//
//   * an interior edge between nodes a and b takes (phi_a - phi_b) / len, with len the
//     spacing between them;
//   * between a free node next to an electrode ("near") and the electrode's node the
//     difference is taken over the bracketed surface distance d len instead (d of the
//     near node towards the electrode; the electrode-to-near side wins if both apply);
//   * an edge with both nodes inside an electrode takes the field of the one adjacent
//     edge that crosses a surface, if exactly one side has such an edge, else 0 (the
//     reference's "extrapolate from the one adjacent surface": that edge's value is the
//     same quotient, since a near node is never an electrode node);
//   * the two end edges of a line mirror (-first) or extrapolate (2 first - second) the
//     edges next to them, the lower end first (a line of two nodes reads it at the upper).
//
// One thread per edge, fx's edges first, both arrays in row order so that neighbouring
// lanes read neighbouring nodes: an HBM-streaming kernel of ~30 bytes per edge.  The
// divisions are IEEE (correctly rounded), so the fields are bit-identical to the host's.
#include <hip/hip_runtime.h>

#include "bt_device.h"
#include "bt_internal.h"

namespace bt {

namespace {

// One line of nodes along the axis of the edges: node t at t * step from `base`.
struct Line {
    const double* __restrict__ phi;
    const int32_t* __restrict__ solid;
    const uint8_t* __restrict__ near;
    const double* __restrict__ dist_lo;  // surface distance towards the lower neighbour
    const double* __restrict__ dist_hi;  // towards the upper neighbour
    const double* __restrict__ len;      // len[t]: node t to t + 1 (nullptr: uniform h)
    double h;
    i64 n, step, base;

    __device__ double spacing(i64 t) const { return len != nullptr ? len[t] : h; }
    __device__ i64 at(i64 t) const { return base + t * step; }
    __device__ bool is_solid(i64 t) const { return solid[at(t)] >= 0; }
    __device__ bool is_near(i64 t) const { return near[at(t)] != 0; }
};

// Interior edge e (nodes e - 1, e) by the plain or surface-distance rule.
__device__ double crossing_edge(const Line& L, i64 e) {
    const i64 a = e - 1, b = e;
    const double drop = L.phi[L.at(a)] - L.phi[L.at(b)];
    const double len = L.spacing(a);
    if (L.is_near(b) && L.is_solid(a)) return drop / (L.dist_lo[L.at(b)] * len);
    if (L.is_near(a) && L.is_solid(b)) return drop / (L.dist_hi[L.at(a)] * len);
    return drop / len;
}

// Interior edge e, 1 <= e <= n - 1.
__device__ double interior_edge(const Line& L, i64 e) {
    if (!(L.is_solid(e - 1) && L.is_solid(e))) return crossing_edge(L, e);
    const bool below = e >= 2 && L.is_near(e - 2);     // a surface edge on the lower side
    const bool above = e + 1 <= L.n - 1 && L.is_near(e + 1);
    if (below && !above) return crossing_edge(L, e - 1);
    if (above && !below) return crossing_edge(L, e + 1);
    return 0.0;
}

__device__ double lower_end(const Line& L, bool mirror) {
    const double first = interior_edge(L, 1);
    const double second = L.n >= 3 ? interior_edge(L, 2) : 0.0;  // (the upper end, still 0)
    return mirror ? -first : 2.0 * first - second;
}

__device__ double edge_value(const Line& L, i64 e, bool mirror_lo, bool mirror_hi) {
    if (e == 0) return lower_end(L, mirror_lo);
    if (e < L.n) return interior_edge(L, e);
    const double first = interior_edge(L, L.n - 1);
    const double second = L.n >= 3 ? interior_edge(L, L.n - 2) : lower_end(L, mirror_lo);
    return mirror_hi ? -first : 2.0 * first - second;
}

}  // namespace

__global__ __launch_bounds__(256) void edge_field_kernel(const double* __restrict__ phi,
                                                         const int32_t* __restrict__ solid,
                                                         const uint8_t* __restrict__ near,
                                                         const double* __restrict__ distance,
                                                         const double* __restrict__ dx, double h, i64 nx,
                                                         i64 ny, int faces, double* __restrict__ fx,
                                                         double* __restrict__ fy) {
    const i64 nodes = nx * ny;
    const i64 n_fx = (nx + 1) * ny, total = n_fx + nx * (ny + 1);
    const i64 stride = (i64)gridDim.x * blockDim.x;
    for (i64 g = (i64)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) {
        if (g < n_fx) {  // fx[e][j]: edge e of line j along x
            const i64 e = g / ny, j = g - e * ny;
            const Line L = {phi, solid, near, distance, distance + nodes, dx, h, nx, ny, j};
            fx[g] = edge_value(L, e, (faces & 1) != 0, (faces & 2) != 0);
        } else {  // fy[i][e]: edge e of line i along r
            const i64 q = g - n_fx;
            const i64 i = q / (ny + 1), e = q - i * (ny + 1);
            const Line L = {phi, solid, near, distance + 2 * nodes, distance + 3 * nodes, nullptr, h, ny, 1, i * ny};
            fy[q] = edge_value(L, e, (faces & 4) != 0, (faces & 8) != 0);
        }
    }
}

int launch_edge_field(const double* phi, const int32_t* solid, const uint8_t* near, const double* distance,
                      const double* dx, double h, i64 nx, i64 ny, int faces, double* fx, double* fy,
                      hipStream_t stream) {
    const i64 total = (nx + 1) * ny + nx * (ny + 1);
    if (total == 0) return 0;
    i64 blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(edge_field_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, phi, solid, near, distance,
                       dx, h, nx, ny, faces, fx, fy);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace bt
