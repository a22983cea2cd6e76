// HELS transfer matrices on gfx950.
//
// Restates CUDA/hels.cuh:222-323 (the CPU/GSL path; the reference's GPU
// kernels hels.cuh:86-183 use FP32 tgammaf normalisations and doubly
// recursive upward Bessel recurrences, unstable for n > kr):
//   Y_n^m = sqrt((2n+1)(n-m)!/(4 pi (n+m)!)) P_n^m(cos th) e^{i m ph}, m >= 0,
//   Y_n^{-m} = (-1)^m conj(Y_n^m), P_n^m with the Condon-Shortley phase
//   (= (-1)^m (1-x^2)^{m/2} (2m)!/(2^m m!) C^{(m+1/2)}_{n-m}(x)),
//   M = h_n(k r) Y, Q = k h_n'(k r) Y / (i rho c k), h_n = j_n + i y_n,
// column i = n^2 + n + m (CUDA order, hels.cuh:92).
//
// One thread per point and frequency, everything in FP64:
//   y_n upward (stable), j_n by Miller's downward recurrence normalised on
//   (j_0, j_1), h_n' = (n/x) h_n - h_{n+1}; associated Legendre functions by
//   the standard three-term recurrence in n for each m.
// Each thread's j_n and y_n (n = 0..N+1) live in LDS, sized by the launch's
// order and laid out [n][thread] so that a wave's accesses to one n hit 64
// consecutive words; stores are column-major with consecutive threads on
// consecutive points.
#include "ia_internal.hpp"

namespace ia {

namespace {

constexpr int HELS_NMAX = 80;

constexpr int HELS_TPB = 64;

template <typename T, bool DERIV>
__global__ __launch_bounds__(HELS_TPB) void k_hels(int npts, int N, const double *__restrict__ kv,
                                                  double rho_c, const double *__restrict__ azelr,
                                                  cx<T> *__restrict__ M, cx<T> *__restrict__ Q,
                                                  size_t stride) {
    extern __shared__ double hsh[];  // j_n then y_n, (N + 2) values each per thread
    double *jb_ = hsh + threadIdx.x, *yb_ = hsh + (size_t)(N + 2) * HELS_TPB + threadIdx.x;
    auto jb = [&](int n) -> double & { return jb_[n * HELS_TPB]; };
    auto yb = [&](int n) -> double & { return yb_[n * HELS_TPB]; };
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (j >= npts) return;
    const double k = kv[f];
    const double az = azelr[j], el = azelr[j + npts], r = azelr[j + 2 * npts];
    const double x = k * r;
    const int n_t = (N + 1) * (N + 1);
    cx<T> *Mf = M + (size_t)f * stride;
    cx<T> *Qf = DERIV ? Q + (size_t)f * stride : nullptr;

    // ---- spherical Bessel functions j_n, y_n for n = 0..N+1
    double sx, cx_;
    sincos(x, &sx, &cx_);
    yb(0) = -cx_ / x;
    yb(1) = -cx_ / (x * x) - sx / x;
    for (int n = 1; n <= N; ++n) yb(n + 1) = (2.0 * n + 1.0) / x * yb(n) - yb(n - 1);
    {
        // Miller: run j_{n-1} = (2n+1)/x j_n - j_{n+1} down from far above N,
        // keep n <= N+1, normalise on the exact (j_0, j_1).
        const int nstart = N + 22 + (int)(1.5 * x);
        double jp1 = 0.0, jc = 1.0;  // j_{n+1}, j_n at n = nstart
        for (int n = nstart; n >= 1; --n) {
            if (n <= N + 1) jb(n) = jc;
            const double jm1 = (2.0 * n + 1.0) / x * jc - jp1;
            jp1 = jc;
            jc = jm1;
            if (fabs(jc) > 1e250) {
                const double sc = 1e-250;
                jc *= sc;
                jp1 *= sc;
                for (int q = n; q <= N + 1; ++q) jb(q) *= sc;
            }
        }
        jb(0) = jc;
        const double j0 = sx / x, j1 = sx / (x * x) - cx_ / x;
        const double tn = fmax(fabs(jb(0)), fabs(jb(1)));
        const double a0 = jb(0) / tn, a1 = jb(1) / tn;
        const double s = (j0 * a0 + j1 * a1) / (a0 * a0 + a1 * a1) / tn;
        for (int n = 0; n <= N + 1; ++n) jb(n) *= s;
    }
    // ---- associated Legendre, m outer, n inner
    const double ct = cos(el);
    const double st = sqrt(fmax(0.0, 1.0 - ct * ct));
    const double inv4pi = 0.079577471545947667884;
    double pmm = 1.0;     // P_m^m
    double fac2m = 1.0;   // (2m)!
    for (int m = 0; m <= N; ++m) {
        if (m > 0) {
            pmm *= -(2.0 * m - 1.0) * st;
            fac2m *= (2.0 * m - 1.0) * (2.0 * m);
        }
        double sphi, cphi;
        sincos((double)m * az, &sphi, &cphi);
        double pnm2 = 0.0, pnm1 = pmm;
        double ratio = 1.0 / fac2m;  // (n-m)!/(n+m)! at n = m
        for (int n = m; n <= N; ++n) {
            double pnm;
            if (n == m) {
                pnm = pmm;
            } else {
                pnm = ((2.0 * n - 1.0) * ct * pnm1 - (n + m - 1.0) * pnm2) / (double)(n - m);
                ratio *= (double)(n - m) / (double)(n + m);
                pnm2 = pnm1;
                pnm1 = pnm;
            }
            if (n == m) {
                pnm2 = 0.0;
                pnm1 = pmm;
            }
            const double nrm = sqrt((2.0 * n + 1.0) * inv4pi * ratio) * pnm;
            const double yre = nrm * cphi, yim = nrm * sphi;
            // h_n and k h_n'(kr) / (i rho c k) = h_n'(x) (-i)/(rho c)
            const double hre = jb(n), him = yb(n);
            cx<double> ypos{yre, yim};
            cx<double> yneg{(m & 1) ? -yre : yre, (m & 1) ? yim : -yim};
            cx<double> h{hre, him};
            const int ip = n * n + n + m, in = n * n + n - m;
            cx<double> mp = h * ypos;
            Mf[(size_t)ip * npts + j] = cvt<T>(mp);
            if (m > 0) Mf[(size_t)in * npts + j] = cvt<T>(h * yneg);
            if (DERIV) {
                cx<double> dh{(double)n / x * jb(n) - jb(n + 1), (double)n / x * yb(n) - yb(n + 1)};
                cx<double> g{dh.im / rho_c, -dh.re / rho_c};  // dh * (-i/(rho c))
                Qf[(size_t)ip * npts + j] = cvt<T>(g * ypos);
                if (m > 0) Qf[(size_t)in * npts + j] = cvt<T>(g * yneg);
            }
        }
    }
    (void)n_t;
}

}  // namespace

int hels_nmax() { return HELS_NMAX; }

template <typename T>
void launch_hels(int npts, int N, int nf, const double *k, double rho_c, const double *azelr,
                 cx<T> *M, cx<T> *Q, size_t stride, hipStream_t s) {
    dim3 grid((npts + HELS_TPB - 1) / HELS_TPB, nf);
    // 2 (N + 2) doubles per thread: 84 KB at the largest order (80)
    const size_t lds = sizeof(double) * 2 * (size_t)(N + 2) * HELS_TPB;
    (void)hipFuncSetAttribute((const void *)k_hels<T, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipFuncSetAttribute((const void *)k_hels<T, false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (Q)
        hipLaunchKernelGGL((k_hels<T, true>), grid, dim3(HELS_TPB), lds, s, npts, N, k, rho_c,
                           azelr, M, Q, stride);
    else
        hipLaunchKernelGGL((k_hels<T, false>), grid, dim3(HELS_TPB), lds, s, npts, N, k, rho_c,
                           azelr, M, Q, stride);
}

template void launch_hels<double>(int, int, int, const double *, double, const double *,
                                  cx<double> *, cx<double> *, size_t, hipStream_t);
template void launch_hels<float>(int, int, int, const double *, double, const double *, cx<float> *,
                                 cx<float> *, size_t, hipStream_t);

}  // namespace ia
