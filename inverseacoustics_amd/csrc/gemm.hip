// Batched complex GEMM on CDNA4 matrix cores.
//
//   C = alpha op(A) op(B) + beta C,   op in {'N', 'C'} (conjugate transpose)
//
// Replaces the cuBLAS Zgemm/Zgeam call sites of the reference
// (CUDA/ibem.cuh:398, CUDA/invr.cuh:402-433) and provides the trailing
// updates of the LU factorisation and the triangular solves.
//
// Complex products are four real MFMAs (re*re, -im*im, re*im, im*re) of
// v_mfma_f64_16x16x4_f64 (FP64 form) or v_mfma_f32_16x16x4_f32 (mixed
// form) with real accumulators.  Each wave computes the TRANSPOSED 32x32
// product op(B)^T op(A)^T so that the 16x16 MFMA result rows land on
// consecutive lanes along the column-major C rows: one store instruction
// writes 4 columns x 16 contiguous rows.
//
// Workgroup tile 64x64, K step 16, 4 waves (one per SIMD), operands staged
// in LDS as split real/imaginary planes in k-major order so that the MFMA
// operand reads of 16 consecutive lanes hit 16 consecutive words.
#include <rocblas/rocblas.h>

#include <cstdlib>

#include "ia_internal.hpp"

namespace ia {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T>
struct mfma_t;
// Result layout of the 16x16x4 MFMAs: lane L holds column L%16 and four rows
// r(i), i = 0..3.  FP64: r = L/16 + 4 i.  FP32: r = 4 (L/16) + i.
template <>
struct mfma_t<double> {
    typedef d4 acc;
    static __device__ __forceinline__ int row(int lk, int i) { return lk + 4 * i; }
    static __device__ __forceinline__ acc run(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <>
struct mfma_t<float> {
    typedef f4 acc;
    static __device__ __forceinline__ int row(int lk, int i) { return 4 * lk + i; }
    static __device__ __forceinline__ acc run(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

constexpr int BM = 64, BN = 64, BK = 16, PAD = 4;

template <typename T, char OPA, char OPB, bool BETA0>
__global__ __launch_bounds__(256) void k_zgemm(int M, int N, int K, cx<T> alpha,
                                              const cx<T> *__restrict__ A, int lda, size_t sA,
                                              const cx<T> *__restrict__ B, int ldb, size_t sB,
                                              cx<T> beta, cx<T> *__restrict__ C, int ldc,
                                              size_t sC) {
    using acc_t = typename mfma_t<T>::acc;
    __shared__ T As_re[BK][BM + PAD], As_im[BK][BM + PAD];
    __shared__ T Bs_re[BK][BN + PAD], Bs_im[BK][BN + PAD];

    const int bz = blockIdx.z;
    A += (size_t)bz * sA;
    B += (size_t)bz * sB;
    C += (size_t)bz * sC;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    const int wm = (w & 1) * 32, wn = (w >> 1) * 32;
    const int l16 = lane & 15, lk = lane >> 4;

    acc_t cre[2][2], cim[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            cre[a][b] = acc_t{0, 0, 0, 0};
            cim[a][b] = acc_t{0, 0, 0, 0};
        }

    for (int k0 = 0; k0 < K; k0 += BK) {
        // ---- stage op(A) tile (BM x BK) as As[k][m]
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            int r, kk;
            cx<T> v{0, 0};
            if (OPA == 'N') {
                r = t & 63;
                kk = (t >> 6) + 4 * it;
                int gr = m0 + r, gk = k0 + kk;
                if (gr < M && gk < K) v = A[(size_t)gk * lda + gr];
            } else {
                kk = t & 15;
                r = (t >> 4) + 16 * it;
                int gr = m0 + r, gk = k0 + kk;
                if (gr < M && gk < K) {
                    v = A[(size_t)gr * lda + gk];
                    v.im = -v.im;
                }
            }
            As_re[kk][r] = v.re;
            As_im[kk][r] = v.im;
        }
        // ---- stage op(B) tile (BK x BN) as Bs[k][n]
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            int c, kk;
            cx<T> v{0, 0};
            if (OPB == 'N') {
                kk = t & 15;
                c = (t >> 4) + 16 * it;
                int gc = n0 + c, gk = k0 + kk;
                if (gc < N && gk < K) v = B[(size_t)gc * ldb + gk];
            } else {
                c = t & 63;
                kk = (t >> 6) + 4 * it;
                int gc = n0 + c, gk = k0 + kk;
                if (gc < N && gk < K) {
                    v = B[(size_t)gk * ldb + gc];
                    v.im = -v.im;
                }
            }
            Bs_re[kk][c] = v.re;
            Bs_im[kk][c] = v.im;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < BK; ks += 4) {
            T are[2], aim[2], naim[2], bre[2], bim[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                are[a] = As_re[ks + lk][wm + a * 16 + l16];
                aim[a] = As_im[ks + lk][wm + a * 16 + l16];
                naim[a] = -aim[a];
                bre[a] = Bs_re[ks + lk][wn + a * 16 + l16];
                bim[a] = Bs_im[ks + lk][wn + a * 16 + l16];
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    // transposed product: MFMA-A <- op(B)^T rows (n), MFMA-B <- op(A)^T cols (m)
                    cre[a][b] = mfma_t<T>::run(bre[b], are[a], cre[a][b]);
                    cre[a][b] = mfma_t<T>::run(bim[b], naim[a], cre[a][b]);
                    cim[a][b] = mfma_t<T>::run(bim[b], are[a], cim[a][b]);
                    cim[a][b] = mfma_t<T>::run(bre[b], aim[a], cim[a][b]);
                }
        }
        __syncthreads();
    }
    // ---- epilogue: lane holds D'[n = row(lk, i)][m = l16]
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                int gm = m0 + wm + a * 16 + l16;
                int gn = n0 + wn + b * 16 + mfma_t<T>::row(lk, i);
                if (gm < M && gn < N) {
                    T xr = cre[a][b][i], xi = cim[a][b][i];
                    cx<T> r{alpha.re * xr - alpha.im * xi, alpha.re * xi + alpha.im * xr};
                    size_t o = (size_t)gn * ldc + gm;
                    if (!BETA0) {
                        cx<T> c0 = C[o];
                        r.re += beta.re * c0.re - beta.im * c0.im;
                        r.im += beta.re * c0.im + beta.im * c0.re;
                    }
                    C[o] = r;
                }
            }
}

template <typename T, char OPA, char OPB>
void dispatch_beta(int M, int N, int K, cx<T> alpha, const cx<T> *A, int lda, size_t sA,
                   const cx<T> *B, int ldb, size_t sB, cx<T> beta, cx<T> *C, int ldc, size_t sC,
                   int batch, hipStream_t s) {
    dim3 grid((M + BM - 1) / BM, (N + BN - 1) / BN, batch);
    if (beta.re == 0 && beta.im == 0)
        hipLaunchKernelGGL((k_zgemm<T, OPA, OPB, true>), grid, dim3(256), 0, s, M, N, K, alpha, A,
                           lda, sA, B, ldb, sB, beta, C, ldc, sC);
    else
        hipLaunchKernelGGL((k_zgemm<T, OPA, OPB, false>), grid, dim3(256), 0, s, M, N, K, alpha, A,
                           lda, sA, B, ldb, sB, beta, C, ldc, sC);
}

}  // namespace

namespace {

// rocBLAS handle of the current device and stream for this host thread
// (created once per pair; streams that run concurrently never share a
// handle's workspace)
rocblas_handle blas_handle(hipStream_t s) {
    struct Slot {
        int dev;
        hipStream_t s;
        rocblas_handle h;
    };
    static thread_local Slot slots[16];
    static thread_local int used = 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    for (int i = 0; i < used; ++i)
        if (slots[i].dev == dev && slots[i].s == s) return slots[i].h;
    if (used == 16) return nullptr;  // the caller falls back to the own kernel
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) return nullptr;
    if (rocblas_set_stream(h, s) != rocblas_status_success) {
        (void)rocblas_destroy_handle(h);
        return nullptr;
    }
    slots[used++] = Slot{dev, s, h};
    return h;
}

rocblas_operation blas_op(char op) {
    return op == 'C' ? rocblas_operation_conjugate_transpose : rocblas_operation_none;
}

// Plain batched products go to rocBLAS (Tensile kernels tuned for gfx950
// reach 70-95 % of the matrix-core peak on these shapes, against ~35 % for
// the single-buffered 64 x 64 tile above, which remains for tiny products
// and when IA_GEMM_OWN is set).  Returns false when not handled.
template <typename T>
bool blas_gemm(char opA, char opB, int M, int N, int K, cx<T> alpha, const cx<T> *A, int lda,
               size_t sA, const cx<T> *B, int ldb, size_t sB, cx<T> beta, cx<T> *C, int ldc,
               size_t sC, int batch, hipStream_t s) {
    if ((double)M * N * K < 32768.0 || getenv("IA_GEMM_OWN")) return false;
    rocblas_handle h = blas_handle(s);
    if (!h) return false;
    rocblas_status st;
    if (sizeof(T) == sizeof(double))
        st = rocblas_zgemm_strided_batched(
            h, blas_op(opA), blas_op(opB), M, N, K,
            reinterpret_cast<const rocblas_double_complex *>(&alpha),
            reinterpret_cast<const rocblas_double_complex *>(A), lda, (rocblas_stride)sA,
            reinterpret_cast<const rocblas_double_complex *>(B), ldb, (rocblas_stride)sB,
            reinterpret_cast<const rocblas_double_complex *>(&beta),
            reinterpret_cast<rocblas_double_complex *>(C), ldc, (rocblas_stride)sC, batch);
    else
        st = rocblas_cgemm_strided_batched(
            h, blas_op(opA), blas_op(opB), M, N, K,
            reinterpret_cast<const rocblas_float_complex *>(&alpha),
            reinterpret_cast<const rocblas_float_complex *>(A), lda, (rocblas_stride)sA,
            reinterpret_cast<const rocblas_float_complex *>(B), ldb, (rocblas_stride)sB,
            reinterpret_cast<const rocblas_float_complex *>(&beta),
            reinterpret_cast<rocblas_float_complex *>(C), ldc, (rocblas_stride)sC, batch);
    return st == rocblas_status_success;
}

}  // namespace

template <typename T>
void launch_gemm(char opA, char opB, int M, int N, int K, cx<T> alpha, const cx<T> *A, int lda,
                 size_t strideA, const cx<T> *B, int ldb, size_t strideB, cx<T> beta, cx<T> *C,
                 int ldc, size_t strideC, int batch, hipStream_t s) {
    if (M <= 0 || N <= 0 || batch <= 0) return;
    if (blas_gemm<T>(opA, opB, M, N, K, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc,
                     strideC, batch, s))
        return;
    if (opA == 'N' && opB == 'N')
        dispatch_beta<T, 'N', 'N'>(M, N, K, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc,
                                   strideC, batch, s);
    else if (opA == 'C' && opB == 'N')
        dispatch_beta<T, 'C', 'N'>(M, N, K, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc,
                                   strideC, batch, s);
    else if (opA == 'N' && opB == 'C')
        dispatch_beta<T, 'N', 'C'>(M, N, K, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc,
                                   strideC, batch, s);
    else
        dispatch_beta<T, 'C', 'C'>(M, N, K, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc,
                                   strideC, batch, s);
}

template void launch_gemm<double>(char, char, int, int, int, cx<double>, const cx<double> *, int,
                                  size_t, const cx<double> *, int, size_t, cx<double>,
                                  cx<double> *, int, size_t, int, hipStream_t);
template void launch_gemm<float>(char, char, int, int, int, cx<float>, const cx<float> *, int,
                                 size_t, const cx<float> *, int, size_t, cx<float>, cx<float> *,
                                 int, size_t, int, hipStream_t);

}  // namespace ia
