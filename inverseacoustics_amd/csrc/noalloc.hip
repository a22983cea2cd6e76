// ia_regularize_noalloc: the regularised solve from a given SVD (the
// computation of ia_regularize) for callers that cannot allocate device
// memory -- code under HIP graph capture, tight serving loops.  Inputs and
// outputs are the caller's device buffers; the intermediate vectors (the
// coordinates beta = U^H b, the filtered coefficients phi and the selection
// scalars) live in a file-scope device array of fixed size instead of
// buffers allocated per call, and the two matrix-vector products are
// dedicated kernels (no library routine that might allocate).
#include <hip/hip_runtime.h>

#include <mutex>

#include "ia_internal.hpp"

constexpr int NA_PMAX = 4096;  // largest p (singular values)
// beta and phi (p complex128 each), |b|^2, and the selection outputs:
// 128 KiB + 32 bytes
constexpr size_t NA_BETA = 0;
constexpr size_t NA_PHI = NA_BETA + 16 * (size_t)NA_PMAX;
constexpr size_t NA_SCAL = NA_PHI + 16 * (size_t)NA_PMAX;  // |b|^2, lambda
constexpr size_t NA_INTS = NA_SCAL + 2 * sizeof(double);    // grid index, TSVD k
constexpr size_t NA_BYTES = NA_INTS + 2 * sizeof(int);

// One per device (each GPU holds its own copy until the process ends).
// Every call on a device shares it: calls that overlap on one device race.
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

namespace ia {

namespace {

std::mutex g_mu;
constexpr int NA_MAXDEV = 64;
void *g_ptr[NA_MAXDEV] = {};

// Device address of this device's scratch array (looked up once per device
// with that device current; the host-side name is only a stand-in).
ia_status scratch_ptr(unsigned char **out) {
    int dev = 0;
    IA_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NA_MAXDEV) return fail(IA_ERR_DEVICE, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(IA_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_ptr[dev] = p;
    }
    *out = static_cast<unsigned char *>(g_ptr[dev]);
    return IA_OK;
}

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// beta_i = sum_r conj(U[r, i]) b[r] (i < p): one wave per i, lanes over the
// contiguous column
__global__ __launch_bounds__(256) void k_na_uhb(int m, int p, const cx<double> *__restrict__ U,
                                               const cx<double> *__restrict__ b,
                                               cx<double> *__restrict__ beta) {
    const int lane = threadIdx.x & 63, i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= p) return;
    const cx<double> *Ui = U + (size_t)i * m;
    double ar = 0, ai = 0;
    for (int r = lane; r < m; r += 64) {
        const cx<double> c = cmulc(Ui[r], b[r]);
        ar += c.re;
        ai += c.im;
    }
    ar = wsum(ar);
    ai = wsum(ai);
    if (lane == 0) beta[i] = cx<double>{ar, ai};
}

// x_r = sum_i V[r, i] phi_i: thread per row, coalesced over each column
__global__ __launch_bounds__(256) void k_na_vphi(int n, int p, const cx<double> *__restrict__ V,
                                                const cx<double> *__restrict__ phi,
                                                cx<double> *__restrict__ x) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    cx<double> acc{0, 0};
    for (int i = 0; i < p; ++i) acc = acc + V[(size_t)i * n + r] * phi[i];
    x[r] = acc;
}

// caller outputs: lambda and the index (mode 0: the grid index, 1: the
// discrete-GCV truncation index, 2: the caller's fixed truncation index)
__global__ void k_na_finish(const double *__restrict__ lam_s, const int *__restrict__ ints,
                            int mode, int kfix, double *__restrict__ lambda,
                            int *__restrict__ index) {
    if (threadIdx.x != 0) return;
    if (lambda) *lambda = *lam_s;
    if (index) *index = mode == 0 ? ints[0] : (mode == 1 ? ints[1] : kfix);
}

}  // namespace

}  // namespace ia

using namespace ia;

extern "C" ia_status ia_regularize_noalloc(int m, int n, int p, const ia_z *u, const double *s,
                                           const ia_z *v, const ia_z *b, const ia_reg_params *prm,
                                           ia_z *x, double *lambda, int *index, void *stream) {
    IA_TRY(check_device());
    if (!prm || m <= 0 || n <= 0 || p <= 0 || !u || !s || !v || !b || !x)
        return fail(IA_ERR_ARG, "ia_regularize_noalloc: bad arguments");
    if (p > NA_PMAX)
        return fail(IA_ERR_ARG, "ia_regularize_noalloc: p = %d exceeds the scratch limit %d", p,
                    NA_PMAX);
    RegDev rd{(int)prm->filter, (int)prm->select, prm->n_lambda > 0 ? prm->n_lambda : 50, prm->k,
              prm->decades > 0 ? prm->decades : 8.0, prm->lambda};
    if (rd.filter < IA_FILTER_TIKHONOV || rd.filter > IA_FILTER_DSVD)
        return fail(IA_ERR_ARG, "ia_regularize_noalloc: bad filter");
    if (rd.select < IA_SELECT_GCV || rd.select > IA_SELECT_RANK)
        return fail(IA_ERR_ARG, "ia_regularize_noalloc: bad select");
    if (rd.n_lambda > 512) return fail(IA_ERR_ARG, "n_lambda must be <= 512");
    const bool tsvd_fixed = prm->filter == IA_FILTER_TSVD && prm->select == IA_SELECT_FIXED;
    if (tsvd_fixed && (prm->k < 0 || prm->k > p)) return fail(IA_ERR_ARG, "k out of range");
    unsigned char *sp = nullptr;
    IA_TRY(scratch_ptr(&sp));
    cx<double> *beta = reinterpret_cast<cx<double> *>(sp + NA_BETA);
    cx<double> *phi = reinterpret_cast<cx<double> *>(sp + NA_PHI);
    double *bn2 = reinterpret_cast<double *>(sp + NA_SCAL);
    double *lam_s = bn2 + 1;
    int *ints = reinterpret_cast<int *>(sp + NA_INTS);
    hipStream_t st = (hipStream_t)stream;
    const cx<double> *U = reinterpret_cast<const cx<double> *>(u);
    const cx<double> *V = reinterpret_cast<const cx<double> *>(v);
    const cx<double> *B = reinterpret_cast<const cx<double> *>(b);
    hipLaunchKernelGGL(k_na_uhb, dim3((p + 3) / 4), dim3(256), 0, st, m, p, U, B, beta);
    launch_norm2<double>(m, B, m, bn2, 1, st);
    launch_select(p, m, s, beta, bn2, rd, lam_s, ints, ints + 1, phi, 1, st);
    hipLaunchKernelGGL(k_na_vphi, dim3((n + 255) / 256), dim3(256), 0, st, n, p, V, phi,
                       reinterpret_cast<cx<double> *>(x));
    const int mode = prm->filter != IA_FILTER_TSVD ? 0 : (tsvd_fixed ? 2 : 1);
    hipLaunchKernelGGL(k_na_finish, dim3(1), dim3(64), 0, st, lam_s, ints, mode, prm->k, lambda,
                       index);
    IA_HIP(hipGetLastError());
    return IA_OK;
}
