// The standalone gather and deposit kernels for gfx950.
//
// staggered_field_2d: (E_x, E_y) at points from the staggered edge fields, clamped bilinear
// (kernels/staggered_field_2d/staggered_field_2d_pyccel.py:13-55).  One point per lane,
// grid-stride; px/py/ex/ey are streamed (coalesced, HBM-bound) while the 8 field values
// of a point come from L2.
//
// deposit_trajectories_2d: the area-weighted deposit of charges at points, bilinear in
// planar geometry and linear in r^2 in cylindrical geometry, charge beyond the mesh on
// the boundary nodes (deposit_trajectories_2d_pyccel.py:13-59).  Each weight is rounded
// once to the fixed-point grid 2^e and added with integer atomics into a 128-bit node
// accumulator, then converted and added to `charge`: the result does not depend on the
// order in which lanes arrive.
#include <hip/hip_runtime.h>

#include "bt_device.h"
#include "bt_internal.h"

namespace bt {

__global__ __launch_bounds__(256) void gather_kernel(const double* __restrict__ px,
                                                     const double* __restrict__ py, i64 n,
                                                     const double* __restrict__ fx,
                                                     const double* __restrict__ fy, i64 nx, i64 ny,
                                                     double ox, double oy, double h,
                                                     double* __restrict__ ex,
                                                     double* __restrict__ ey) {
    const i64 stride = (i64)gridDim.x * blockDim.x;
    for (i64 k = (i64)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const double sx = (px[k] - ox) / h;
        const double sy = (py[k] - oy) / h;
        ex[k] = clamped_bilinear(fx, ny, sx + 0.5, sy, (int)nx - 1, (int)ny - 2);
        ey[k] = clamped_bilinear(fy, ny + 1, sx, sy + 0.5, (int)nx - 2, (int)ny - 1);
    }
}

__global__ __launch_bounds__(256) void deposit_kernel(const double* __restrict__ px,
                                                      const double* __restrict__ py,
                                                      const double* __restrict__ q, i64 n,
                                                      double ox, double oy, double h,
                                                      i64 cylindrical, i64 nx, i64 ny, int e,
                                                      u64* __restrict__ acc) {
    const i64 stride = (i64)gridDim.x * blockDim.x;
    for (i64 k = (i64)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const double x = px[k], y = py[k];
        const double fi = floor((x - ox) / h);
        const double fj = floor((y - oy) / h);
        double t = (x - (ox + h * fi)) / h;
        double u;
        if (cylindrical != 0) {
            const double r1 = oy + h * fj;
            const double r2 = r1 + h;
            u = (y * y - r1 * r1) / (r2 * r2 - r1 * r1);
        } else {
            u = (y - (oy + h * fj)) / h;
        }
        i64 i, j;
        // !(f >= 0) also sends a NaN point to the first node (its deposit is dropped)
        if (!(fi >= 0.0)) { t = 0.0; i = 0; }
        else if (fi >= (double)(nx - 1)) { t = 1.0; i = nx - 2; }
        else i = (i64)fi;
        if (!(fj >= 0.0)) { u = 0.0; j = 0; }
        else if (fj >= (double)(ny - 1)) { u = 1.0; j = ny - 2; }
        else j = (i64)fj;
        const double value = q[k];
        const i64 c = i * ny + j;
        fixed_add(acc, c, (1.0 - t) * (1.0 - u) * value, e);
        fixed_add(acc, c + 1, (1.0 - t) * u * value, e);
        fixed_add(acc, c + ny, t * (1.0 - u) * value, e);
        fixed_add(acc, c + ny + 1, t * u * value, e);
    }
}

__global__ void zero_words(u64* __restrict__ acc, i64 words) {
    const i64 stride = (i64)gridDim.x * blockDim.x;
    for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; t < words; t += stride) acc[t] = 0ull;
}

__global__ void add_fixed(const u64* __restrict__ acc, i64 nodes, int e, double* __restrict__ charge) {
    const i64 stride = (i64)gridDim.x * blockDim.x;
    for (i64 c = (i64)blockIdx.x * blockDim.x + threadIdx.x; c < nodes; c += stride) {
        const u64 w0 = acc[kWords * c], w1 = acc[kWords * c + 1];
        if (w0 != 0 || w1 != 0) charge[c] += words_value(w0, w1, e);
    }
}

static int grid_for(i64 n, int block, int cap) {
    i64 g = (n + block - 1) / block;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
}

int launch_gather(const double* px, const double* py, i64 n, const double* fx, const double* fy,
                  i64 nx, i64 ny, double ox, double oy, double h, double* ex, double* ey,
                  hipStream_t stream) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(gather_kernel, dim3(grid_for(n, 256, 65536)), dim3(256), 0, stream, px, py,
                       n, fx, fy, nx, ny, ox, oy, h, ex, ey);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_deposit(const double* px, const double* py, const double* q, i64 n, double ox,
                   double oy, double h, i64 cylindrical, double* charge, i64 nx, i64 ny, int e,
                   u64* acc, hipStream_t stream) {
    const i64 nodes = nx * ny;
    hipLaunchKernelGGL(zero_words, dim3(grid_for(kWords * nodes, 256, 4096)), dim3(256), 0, stream, acc,
                       kWords * nodes);
    if (n > 0)
        hipLaunchKernelGGL(deposit_kernel, dim3(grid_for(n, 256, 65536)), dim3(256), 0, stream, px,
                           py, q, n, ox, oy, h, cylindrical, nx, ny, e, acc);
    hipLaunchKernelGGL(add_fixed, dim3(grid_for(nodes, 256, 4096)), dim3(256), 0, stream, acc, nodes,
                       e, charge);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

// ---- L2 gather bandwidth (measurement) ---------------------------------------------------
// Random 16-byte row-pair gathers, two per lookup as the tracer's bilinear does (f[c], f[c+1]
// and f[c+cols], f[c+cols+1]), from a field of `rows` x `cols` doubles read through a buffer
// resource: each lane draws its cells from a 32-bit hash, so a field larger than the CU's
// L1 serves them from L2.  The sum of what was read goes to sink[lane] (so nothing is
// dropped); the bytes moved are 32 per lookup.
__device__ __forceinline__ unsigned hash32(unsigned x) {
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__global__ __launch_bounds__(256) void l2_gather_kernel(const double* field, unsigned rows, unsigned cols,
                                                        int lookups, unsigned seed, double* __restrict__ sink) {
    const Rsrc f = field_rsrc(field, (i64)rows * cols);
    const unsigned lane = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned state = hash32(lane ^ seed) | 1u;  // xorshift32 needs a nonzero state
    double acc = 0.0;
#pragma unroll 4
    for (int k = 0; k < lookups; ++k) {
        state ^= state << 13; state ^= state >> 17; state ^= state << 5;
        // row in [0, rows - 2], col in [0, cols - 2] by multiply-high range reduction
        const unsigned row = __umulhi(state, rows - 1), col = __umulhi(state * 0x9e3779b9u, cols - 1);
        const unsigned c = row * cols + col;
        const Pair lo = load_pair(f, c), hi = load_pair(f, c + cols);
        acc += (lo.a + lo.b) + (hi.a + hi.b);
    }
    sink[lane] = acc;
}

// Whole-line reads of a field resident in L2: each group of 8 lanes reads one 128-byte line
// (16 bytes a lane) at a random line of the field, `lines` times, so every byte that L2 sends
// is used: the L2 bandwidth in useful bytes for reads that reuse nothing in L1.
__global__ __launch_bounds__(256) void l2_line_kernel(const double* field, unsigned n_lines, int lines, unsigned seed,
                                                      double* __restrict__ sink) {
    const Rsrc f = field_rsrc(field, (i64)n_lines * 16);
    const unsigned lane = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned state = hash32((lane >> 3) ^ seed) | 1u;  // one stream of lines per group of 8 lanes
    const unsigned part = (lane & 7u) * 2u;               // this lane's two doubles of the line
    double acc = 0.0;
#pragma unroll 4
    for (int k = 0; k < lines; ++k) {
        state ^= state << 13; state ^= state >> 17; state ^= state << 5;
        const unsigned line = __umulhi(state, n_lines);
        const Pair q = load_pair(f, line * 16u + part);
        acc += q.a + q.b;
    }
    sink[lane] = acc;
}

int launch_l2_line(const double* f, i64 n_lines, int lines, int blocks, unsigned seed, double* sink,
                   hipStream_t stream) {
    hipLaunchKernelGGL(l2_line_kernel, dim3(blocks), dim3(256), 0, stream, f, (unsigned)n_lines, lines, seed, sink);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_l2_gather(const double* f, i64 rows, i64 cols, int lookups, int blocks, unsigned seed, double* sink,
                     hipStream_t stream) {
    hipLaunchKernelGGL(l2_gather_kernel, dim3(blocks), dim3(256), 0, stream, f, (unsigned)rows, (unsigned)cols,
                       lookups, seed, sink);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // namespace bt
