// Small device kernels behind the library's own device arrays: fills, the largest magnitude
// of an array, the indices where an array equals a value, and index gathers and scatters.
// The host code keeps its device-resident state (resident beams, batched designs, edge
// fields) in memory the library allocates; these kernels are the few operations that state
// needs between traces.  Every index is checked against its array: an index out of range
// is skipped and counted, and the entry point reports it.
#include "bt_internal.h"

namespace bt {

namespace {

template <typename T>
__global__ void fill_kernel(T* __restrict__ dst, T value, i64 n) {
    const i64 stride = (i64)gridDim.x * blockDim.x;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = value;
}

// the largest |a[i]| as the bits of a non-negative double: their order is the numbers'
// (a NaN's bits exceed infinity's, so a NaN anywhere gives a NaN)
__global__ void abs_max_kernel(const double* __restrict__ a, i64 n, u64* __restrict__ out) {
    const i64 stride = (i64)gridDim.x * blockDim.x;
    u64 best = 0;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const u64 bits = (u64)__double_as_longlong(a[i]) & 0x7fffffffffffffffull;
        best = bits > best ? bits : best;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const u64 other = __shfl_down(best, off);
        best = other > best ? other : best;
    }
    if ((threadIdx.x & 63) == 0 && best != 0) atomicMax(out, best);
}

// the largest a[i] through the order-keeping map of int64 onto uint64 (the sign bit flipped)
__global__ void max_i64_kernel(const i64* __restrict__ a, i64 n, u64* __restrict__ out) {
    const i64 stride = (i64)gridDim.x * blockDim.x;
    u64 best = 0;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const u64 key = (u64)a[i] ^ 0x8000000000000000ull;
        best = key > best ? key : best;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const u64 other = __shfl_down(best, off);
        best = other > best ? other : best;
    }
    if ((threadIdx.x & 63) == 0 && best != 0) atomicMax(out, best);
}

__global__ void select_equal_kernel(const i64* __restrict__ a, i64 n, i64 value, i64* __restrict__ idx,
                                    u64* __restrict__ count) {
    const i64 stride = (i64)gridDim.x * blockDim.x;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        if (a[i] == value) idx[atomicAdd(count, 1ull)] = i;
}

// dst[r][k] = src[r][idx[k]] for rows r < rows (src rows of n, dst rows of m)
__global__ void gather_columns_kernel(const double* __restrict__ src, i64 rows, i64 n, const i64* __restrict__ idx,
                                      i64 m, double* __restrict__ dst, u64* __restrict__ bad) {
    const i64 stride = (i64)gridDim.x * blockDim.x;
    for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; t < rows * m; t += stride) {
        const i64 r = t / m, k = t - r * m;
        const i64 j = idx[k];
        if (j < 0 || j >= n) {
            if (r == 0) atomicAdd(bad, 1ull);
            dst[t] = __longlong_as_double(0x7ff8000000000000ll);
            continue;
        }
        dst[t] = src[r * n + j];
    }
}

// dst[idx[k]] = values[k]
__global__ void scatter_kernel(i64* __restrict__ dst, i64 n, const i64* __restrict__ idx,
                               const i64* __restrict__ values, i64 m, u64* __restrict__ bad) {
    const i64 stride = (i64)gridDim.x * blockDim.x;
    for (i64 k = (i64)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += stride) {
        const i64 j = idx[k];
        if (j < 0 || j >= n) {
            atomicAdd(bad, 1ull);
            continue;
        }
        dst[j] = values[k];
    }
}

// dst[r] = hits[trajectory[r], plane, slot[r], :] (8 values each)
__global__ void gather_records_kernel(const double* __restrict__ hits, i64 n, i64 n_planes, i64 capacity,
                                      i64 plane, const i64* __restrict__ trajectory, const i64* __restrict__ slot,
                                      i64 m, double* __restrict__ dst, u64* __restrict__ bad) {
    const i64 stride = (i64)gridDim.x * blockDim.x;
    for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; t < 8 * m; t += stride) {
        const i64 r = t >> 3, f = t & 7;
        const i64 k = trajectory[r], c = slot[r];
        if (k < 0 || k >= n || c < 0 || c >= capacity) {
            if (f == 0) atomicAdd(bad, 1ull);
            dst[t] = __longlong_as_double(0x7ff8000000000000ll);
            continue;
        }
        dst[t] = hits[((k * n_planes + plane) * capacity + c) * 8 + f];
    }
}

int grid_for(i64 work) {
    i64 g = (work + 255) / 256;
    if (g > 4096) g = 4096;
    return g < 1 ? 1 : (int)g;
}

}  // namespace

int launch_fill_i64(i64* dst, i64 value, i64 n, hipStream_t stream) {
    if (n > 0) hipLaunchKernelGGL(fill_kernel<i64>, dim3(grid_for(n)), dim3(256), 0, stream, dst, value, n);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_fill_f64(double* dst, double value, i64 n, hipStream_t stream) {
    if (n > 0) hipLaunchKernelGGL(fill_kernel<double>, dim3(grid_for(n)), dim3(256), 0, stream, dst, value, n);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_abs_max(const double* a, i64 n, u64* out, hipStream_t stream) {
    if (n > 0) hipLaunchKernelGGL(abs_max_kernel, dim3(grid_for(n)), dim3(256), 0, stream, a, n, out);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_max_i64(const i64* a, i64 n, u64* out, hipStream_t stream) {
    if (n > 0) hipLaunchKernelGGL(max_i64_kernel, dim3(grid_for(n)), dim3(256), 0, stream, a, n, out);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_select_equal(const i64* a, i64 n, i64 value, i64* idx, u64* count, hipStream_t stream) {
    if (n > 0) hipLaunchKernelGGL(select_equal_kernel, dim3(grid_for(n)), dim3(256), 0, stream, a, n, value, idx, count);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_gather_columns(const double* src, i64 rows, i64 n, const i64* idx, i64 m, double* dst, u64* bad,
                          hipStream_t stream) {
    if (rows > 0 && m > 0)
        hipLaunchKernelGGL(gather_columns_kernel, dim3(grid_for(rows * m)), dim3(256), 0, stream, src, rows, n, idx, m,
                           dst, bad);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_scatter(i64* dst, i64 n, const i64* idx, const i64* values, i64 m, u64* bad, hipStream_t stream) {
    if (m > 0) hipLaunchKernelGGL(scatter_kernel, dim3(grid_for(m)), dim3(256), 0, stream, dst, n, idx, values, m, bad);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_gather_records(const double* hits, i64 n, i64 n_planes, i64 capacity, i64 plane, const i64* trajectory,
                          const i64* slot, i64 m, double* dst, u64* bad, hipStream_t stream) {
    if (m > 0)
        hipLaunchKernelGGL(gather_records_kernel, dim3(grid_for(8 * m)), dim3(256), 0, stream, hits, n, n_planes,
                           capacity, plane, trajectory, slot, m, dst, bad);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // namespace bt
