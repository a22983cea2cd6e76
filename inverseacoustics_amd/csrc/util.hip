// Small batched vector kernels: real-matrix x complex-vector products for the
// bidiagonal singular vectors, squared norms, precision conversion and the
// sound-power inner product W = 1/2 Re(v^H p) S/m (CUDA/utils.cuh:265-289).
#include "ia_internal.hpp"

namespace ia {

namespace {

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// y = Q x (T=0) with Q column-major n x n, or y = Q^T x (T=1); real Q, complex x.
// T=0: thread per row i, loop over k (Q[i + k n], coalesced across threads).
// T=1: wave per output i, lanes over k (Q[k + i n], contiguous).
__global__ __launch_bounds__(256) void k_rmv_n(int n, const double *__restrict__ Qall,
                                              const cx<double> *__restrict__ xall, size_t sx,
                                              cx<double> *__restrict__ yall, size_t sy) {
    __shared__ cx<double> xs[1024];
    const int b = blockIdx.y;
    const double *Q = Qall + (size_t)b * n * n;
    const cx<double> *x = xall + (size_t)b * sx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double ar = 0, ai = 0;
    for (int k0 = 0; k0 < n; k0 += 1024) {
        __syncthreads();
        for (int k = threadIdx.x; k < 1024 && k0 + k < n; k += blockDim.x) xs[k] = x[k0 + k];
        __syncthreads();
        if (i < n) {
            const int kn = min(1024, n - k0);
            for (int k = 0; k < kn; ++k) {
                double q = Q[(size_t)(k0 + k) * n + i];
                ar += q * xs[k].re;
                ai += q * xs[k].im;
            }
        }
    }
    if (i < n) yall[(size_t)b * sy + i] = cx<double>{ar, ai};
}

__global__ __launch_bounds__(256) void k_rmv_t(int n, const double *__restrict__ Qall,
                                              const cx<double> *__restrict__ xall, size_t sx,
                                              cx<double> *__restrict__ yall, size_t sy) {
    const int b = blockIdx.y;
    const double *Q = Qall + (size_t)b * n * n;
    const cx<double> *x = xall + (size_t)b * sx;
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const double *Qi = Q + (size_t)i * n;
    double ar = 0, ai = 0;
    for (int k = lane; k < n; k += 64) {
        double q = Qi[k];
        ar += q * x[k].re;
        ai += q * x[k].im;
    }
    ar = wsum(ar);
    ai = wsum(ai);
    if (lane == 0) yall[(size_t)b * sy + i] = cx<double>{ar, ai};
}

template <typename T>
__global__ __launch_bounds__(256) void k_norm2(int n, const cx<T> *__restrict__ x, size_t stride,
                                              double *out) {
    __shared__ double red[4];
    const int b = blockIdx.x;
    const cx<T> *xb = x + (size_t)b * stride;
    double acc = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += (double)abs2(xb[i]);
    acc = wsum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[b] = red[0] + red[1] + red[2] + red[3];
}

template <typename Td, typename Ts>
__global__ void k_convert(size_t n, const cx<Ts> *__restrict__ src, cx<Td> *__restrict__ dst) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t step = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += step) dst[i] = cvt<Td>(src[i]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_power(int m, const cx<T> *__restrict__ v, size_t sv,
                                              const cx<T> *__restrict__ p, size_t sp, double scale,
                                              double *W) {
    __shared__ double red[4];
    const int b = blockIdx.x;
    double acc = 0;
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        cx<T> a = v[(size_t)b * sv + i], c = p[(size_t)b * sp + i];
        acc += (double)a.re * (double)c.re + (double)a.im * (double)c.im;  // Re(conj(v) p)
    }
    acc = wsum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) W[b] = 0.5 * (red[0] + red[1] + red[2] + red[3]) * scale;
}

}  // namespace

void launch_real_matvec(int n, const double *Q, int transpose, const cx<double> *x, size_t stridex,
                        cx<double> *y, size_t stridey, int batch, hipStream_t s) {
    if (!transpose) {
        dim3 grid((n + 255) / 256, batch);
        hipLaunchKernelGGL(k_rmv_n, grid, dim3(256), 0, s, n, Q, x, stridex, y, stridey);
    } else {
        dim3 grid((n + 3) / 4, batch);
        hipLaunchKernelGGL(k_rmv_t, grid, dim3(256), 0, s, n, Q, x, stridex, y, stridey);
    }
}

template <typename T>
void launch_norm2(int n, const cx<T> *x, size_t stride, double *out, int batch, hipStream_t s) {
    hipLaunchKernelGGL(k_norm2<T>, dim3(batch), dim3(256), 0, s, n, x, stride, out);
}

template <typename Td, typename Ts>
void launch_convert(size_t n, const cx<Ts> *src, cx<Td> *dst, hipStream_t s) {
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) return;
    hipLaunchKernelGGL((k_convert<Td, Ts>), dim3((unsigned)blocks), dim3(256), 0, s, n, src, dst);
}

template <typename T>
void launch_power(int m, const cx<T> *v, size_t stridev, const cx<T> *p, size_t stridep,
                  double scale, double *W, int batch, hipStream_t s) {
    hipLaunchKernelGGL(k_power<T>, dim3(batch), dim3(256), 0, s, m, v, stridev, p, stridep, scale,
                       W);
}

template void launch_norm2<double>(int, const cx<double> *, size_t, double *, int, hipStream_t);
template void launch_norm2<float>(int, const cx<float> *, size_t, double *, int, hipStream_t);
template void launch_convert<double, float>(size_t, const cx<float> *, cx<double> *, hipStream_t);
template void launch_convert<float, double>(size_t, const cx<double> *, cx<float> *, hipStream_t);
template void launch_convert<double, double>(size_t, const cx<double> *, cx<double> *, hipStream_t);
template void launch_power<double>(int, const cx<double> *, size_t, const cx<double> *, size_t,
                                   double, double *, int, hipStream_t);
template void launch_power<float>(int, const cx<float> *, size_t, const cx<float> *, size_t, double,
                                  double *, int, hipStream_t);

}  // namespace ia

namespace ia {
namespace {
__global__ void k_rows_to_cols(int R, int C, const double *__restrict__ Ptall,
                               cx<double> *__restrict__ dst) {
    const int b = blockIdx.z, i = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= R) return;
    const double *Pt = Ptall + (size_t)b * C * C;
    double v = (k < C) ? Pt[(size_t)k * C + i] : 0.0;
    dst[(size_t)b * R * C + (size_t)i * R + k] = cx<double>{v, 0.0};
}
template <typename T>
__global__ void k_conj_transpose(int rows, int cols, const cx<T> *__restrict__ src, size_t ss,
                                 cx<T> *__restrict__ dst, size_t sd) {
    __shared__ cx<T> tile[32][33];
    const int b = blockIdx.z;
    const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: ty 0..7
    for (int yy = ty; yy < 32; yy += 8) {
        int r = r0 + tx, c = c0 + yy;
        if (r < rows && c < cols) tile[yy][tx] = src[(size_t)b * ss + (size_t)c * rows + r];
    }
    __syncthreads();
    for (int yy = ty; yy < 32; yy += 8) {
        int c = c0 + tx, r = r0 + yy;  // dst(c, r) = conj(src(r, c)), dst is cols x rows
        if (r < rows && c < cols) dst[(size_t)b * sd + (size_t)r * cols + c] = conj(tile[tx][yy]);
    }
}
__global__ void k_pad_copy(int n, int ny, const cx<double> *__restrict__ x, size_t sx,
                           cx<double> *__restrict__ y, size_t sy) {
    const int b = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ny) return;
    y[(size_t)b * sy + i] = (i < n) ? x[(size_t)b * sx + i] : cx<double>{0, 0};
}
}  // namespace

void launch_rows_to_cols(int R, int C, const double *Pt, cx<double> *dst, int batch, hipStream_t s) {
    dim3 grid((R + 255) / 256, C, batch);
    hipLaunchKernelGGL(k_rows_to_cols, grid, dim3(256), 0, s, R, C, Pt, dst);
}
template <typename T>
void launch_conj_transpose(int rows, int cols, const cx<T> *src, size_t ss, cx<T> *dst, size_t sd,
                           int batch, hipStream_t s) {
    dim3 grid((rows + 31) / 32, (cols + 31) / 32, batch);
    hipLaunchKernelGGL(k_conj_transpose<T>, grid, dim3(256), 0, s, rows, cols, src, ss, dst, sd);
}
void launch_pad_copy(int n, int ny, const cx<double> *x, size_t sx, cx<double> *y, size_t sy,
                     int batch, hipStream_t s) {
    dim3 grid((ny + 255) / 256, batch);
    hipLaunchKernelGGL(k_pad_copy, grid, dim3(256), 0, s, n, ny, x, sx, y, sy);
}
template void launch_conj_transpose<double>(int, int, const cx<double> *, size_t, cx<double> *,
                                            size_t, int, hipStream_t);
template void launch_conj_transpose<float>(int, int, const cx<float> *, size_t, cx<float> *,
                                           size_t, int, hipStream_t);
}  // namespace ia
