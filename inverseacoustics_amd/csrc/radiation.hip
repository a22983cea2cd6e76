// Radiation-mode sweep on elemental radiators.
//
// The reference's radiation-mode path (MATLAB/invrS.m:157-170,
// CUDA/invr.cuh:426-433) takes the leading singular vectors Q of a
// radiation-resistance matrix R, maps them to the microphones (H_r = H Q)
// and solves the regularised problem in mode space.  Here R is the
// elemental-radiator form for np equal patches of area A on a baffled plane,
//   R_ij = (w^2 rho A^2 / (4 pi c)) sin(k r_ij) / (k r_ij),
// so that W = v^H R v for patch velocities v (Elliott & Johnson, JASA 94
// (1993) 2194), and the microphone transfer is the Rayleigh integral
//   G_ij = -(i w rho A / (2 pi)) e^{ik R_ij} / R_ij.
// Per frequency, every frequency of a chunk in one launch per stage:
//   1. R / (w^2 rho A^2 / (4 pi c)) assembled on the device (k_rad_R);
//   2. its FP64 eigendecomposition: R is symmetric positive semi-definite,
//      so its singular values are its eigenvalues and its left singular
//      vectors its modes -- the batched bidiagonal SVD of svd.hip, with only
//      the n leading vectors pushed through the reflectors, each rotated to
//      real with its largest entry positive (k_realign); modes below the
//      eigenvalue floor dropped, the kept ones orthonormalised (k_orth_modes);
//   3. H_r = G Q (G assembled by k_rad_G, one batched GEMM);
//   4. the regularised solve of H_r y = b (regsolve_batched: the SVD of H_r
//      in complex128 or complex64, parameter choice in FP64);
//   5. W = sum_i e_i |y_i|^2 and v = Q y.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "ia_internal.hpp"

namespace ia {

template <typename T>
ia_status regsolve_batched(int rows, int cols, int batch, const cx<double> *A, size_t sA,
                           const cx<double> *b, size_t sb, const RegDev &prm, cx<double> *x,
                           size_t sx, double *lam, int *idx, int *kts, double *smax,
                           hipStream_t s, double *gap);
RegDev to_regdev(const ia_reg_params &p);
ia_status check_reg(const ia_reg_params &p, int pmax);

namespace {

constexpr int RT = 64;  // tile of the assembly kernels

// S_f = R_f / (w^2 rho A^2/(4 pi c)): sin(k r)/(k r), column-major complex
// (imaginary part zero) for the complex SVD.  Tile of 64 x 64 per
// workgroup, the tile's 64 column points staged in LDS, each thread owns a
// row and walks 16 columns: stores of 64 consecutive entries per column.
__global__ __launch_bounds__(256) void k_rad_R(int np, const double *__restrict__ kv,
                                              const double *__restrict__ p,
                                              cx<double> *__restrict__ S, size_t sS) {
    __shared__ double sc[3][RT];
    const int f = blockIdx.z, t = threadIdx.x;
    const int i0 = blockIdx.x * RT, j0 = blockIdx.y * RT;
    if (t < RT) {
        const int j = min(j0 + t, np - 1);
        sc[0][t] = p[j];
        sc[1][t] = p[j + np];
        sc[2][t] = p[j + 2 * np];
    }
    __syncthreads();
    const int i = i0 + (t & (RT - 1));
    if (i >= np) return;
    const double k = kv[f], x = p[i], y = p[i + np], z = p[i + 2 * np];
    cx<double> *Sf = S + (size_t)f * sS;
    const int cb = (t >> 6) * 16;
    for (int cc = 0; cc < 16; ++cc) {
        const int jl = cb + cc, j = j0 + jl;
        if (j >= np) break;
        const double dx = x - sc[0][jl], dy = y - sc[1][jl], dz = z - sc[2][jl];
        const double kr = k * sqrt(dx * dx + dy * dy + dz * dz);
        Sf[(size_t)j * np + i] = cx<double>{kr == 0.0 ? 1.0 : sin(kr) / kr, 0.0};
    }
}

// G_f (l x np) = -(i w rho A/(2 pi)) e^{ikR}/R, mic i, patch j
__global__ __launch_bounds__(256) void k_rad_G(int l, int np, const double *__restrict__ kv,
                                              double rho_c_A, const double *__restrict__ r_h,
                                              const double *__restrict__ p,
                                              cx<double> *__restrict__ G, size_t sG) {
    __shared__ double sc[3][RT];
    const int f = blockIdx.z, t = threadIdx.x;
    const int i0 = blockIdx.x * RT, j0 = blockIdx.y * RT;
    if (t < RT) {
        const int j = min(j0 + t, np - 1);
        sc[0][t] = p[j];
        sc[1][t] = p[j + np];
        sc[2][t] = p[j + 2 * np];
    }
    __syncthreads();
    const int i = i0 + (t & (RT - 1));
    if (i >= l) return;
    const double k = kv[f], x = r_h[i], y = r_h[i + l], z = r_h[i + 2 * l];
    // w rho A / (2 pi) = k (rho c A) / (2 pi)
    const double amp = k * rho_c_A * 0.15915494309189533577;
    cx<double> *Gf = G + (size_t)f * sG;
    const int cb = (t >> 6) * 16;
    for (int cc = 0; cc < 16; ++cc) {
        const int jl = cb + cc, j = j0 + jl;
        if (j >= np) break;
        const double dx = x - sc[0][jl], dy = y - sc[1][jl], dz = z - sc[2][jl];
        const double R = sqrt(dx * dx + dy * dy + dz * dz);
        double sn, cs;
        sincos(k * R, &sn, &cs);
        const double a = amp / R;
        // -i a (cs + i sn) = a sn - i a cs
        Gf[(size_t)j * l + i] = cx<double>{a * sn, -a * cs};
    }
}

// dst (np x n per frequency): column j = row j of Pt (np x np), the
// bidiagonal's left singular vector j
__global__ void k_first_cols(int np, int n, const double *__restrict__ Pt,
                             cx<double> *__restrict__ dst) {
    const int f = blockIdx.z, j = blockIdx.y;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= np) return;
    dst[(size_t)f * np * n + (size_t)j * np + r] =
        cx<double>{Pt[(size_t)f * np * np + (size_t)r * np + j], 0.0};
}

// Rotate column j of U (np x n per frequency) by the conjugate phase of its
// largest entry: the mode of a real symmetric matrix comes out real with
// that entry positive; the (round-off) imaginary part is dropped.
__global__ __launch_bounds__(256) void k_realign(int np, int n, cx<double> *__restrict__ U) {
    __shared__ double sv[256];
    __shared__ int si[256];
    const int f = blockIdx.y, j = blockIdx.x, t = threadIdx.x;
    cx<double> *u = U + (size_t)f * np * n + (size_t)j * np;
    double best = -1;
    int bi = 0;
    for (int r = t; r < np; r += 256) {
        const double a = abs2(u[r]);
        if (a > best) {
            best = a;
            bi = r;
        }
    }
    sv[t] = best;
    si[t] = bi;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) {
            // larger magnitude wins, ties to the lower row
            if (sv[t + h] > sv[t] || (sv[t + h] == sv[t] && si[t + h] < si[t])) {
                sv[t] = sv[t + h];
                si[t] = si[t + h];
            }
        }
        __syncthreads();
    }
    const cx<double> um = u[si[0]];
    const double nm = sqrt(abs2(um));
    const cx<double> ph = nm > 0 ? cx<double>{um.re / nm, -um.im / nm} : cx<double>{1, 0};
    __syncthreads();
    for (int r = t; r < np; r += 256) {
        const cx<double> v = u[r] * ph;
        u[r] = cx<double>{v.re, 0.0};
    }
}

__device__ __forceinline__ double block_sum1024(double v, double *red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double a = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) a += red[i];
    return a;
}

// Modes kept: e_i >= floor e_1 (below that a mode radiates a negligible
// fraction of the leading mode's power per unit amplitude and its shape is
// not resolved in FP64: the invariant subspace of the values near the
// round-off floor is defined only up to an arbitrary rotation, which would
// leave the mode-space fit basis-dependent).  The kept modes are made
// orthonormal by two passes of modified Gram-Schmidt (they already are to
// about eps e_1 / gap); the dropped ones are zeroed, so their columns of
// H_r and their amplitudes vanish.  One 1024-thread workgroup per frequency.
__global__ __launch_bounds__(1024) void k_orth_modes(int np, int n, const double *__restrict__ sig,
                                                    double floor_rel, cx<double> *__restrict__ U,
                                                    double *__restrict__ modes) {
    __shared__ double red[16];
    const int f = blockIdx.x, t = threadIdx.x;
    const double *sg = sig + (size_t)f * np;
    cx<double> *Uf = U + (size_t)f * np * n;
    int nk = 0;
    while (nk < n && sg[nk] >= floor_rel * sg[0] && sg[0] > 0) ++nk;
    for (int j = 0; j < nk; ++j) {
        cx<double> *uj = Uf + (size_t)j * np;
        for (int pass = 0; pass < 2; ++pass)
            for (int i = 0; i < j; ++i) {
                const cx<double> *ui = Uf + (size_t)i * np;
                double d = 0;
                for (int r = t; r < np; r += 1024) d += ui[r].re * uj[r].re;
                d = block_sum1024(d, red);
                for (int r = t; r < np; r += 1024) uj[r].re -= d * ui[r].re;
            }
        double nr = 0;
        for (int r = t; r < np; r += 1024) nr += uj[r].re * uj[r].re;
        nr = block_sum1024(nr, red);
        const double inv = nr > 0 ? 1.0 / sqrt(nr) : 0.0;
        for (int r = t; r < np; r += 1024) uj[r].re *= inv;
    }
    __syncthreads();
    for (size_t q = t; q < (size_t)np * n; q += 1024) {
        const int j = (int)(q / np);
        if (j >= nk) Uf[q] = cx<double>{0, 0};
        if (modes) modes[(size_t)f * np * n + q] = j < nk ? Uf[q].re : 0.0;
    }
}

// e = scale sigma (n leading values), W = sum e_i |y_i|^2
__global__ __launch_bounds__(256) void k_modal_power(int np, int n, const double *__restrict__ kv,
                                                    double scale_c, const double *__restrict__ sig,
                                                    const cx<double> *__restrict__ y,
                                                    double *__restrict__ eig,
                                                    double *__restrict__ W) {
    __shared__ double red[4];
    const int f = blockIdx.x, t = threadIdx.x;
    const double k = kv[f], scale = scale_c * k * k;
    double acc = 0;
    for (int i = t; i < n; i += 256) {
        const double e = scale * sig[(size_t)f * np + i];
        if (eig) eig[(size_t)f * n + i] = e;
        acc += e * abs2(y[(size_t)f * n + i]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((t & 63) == 0) red[t >> 6] = acc;
    __syncthreads();
    if (t == 0 && W) W[f] = red[0] + red[1] + red[2] + red[3];
}

}  // namespace

// device workspace of the radiation sweep, kept per host thread
struct RadCache {
    DevBuf S, U, G, H, y, v, lam, idx, kts, smax, W, eig, modes, k, p, rh, b;
    SvdWork<double> sw;
    void release() {
        for (DevBuf *q : {&S, &U, &G, &H, &y, &v, &lam, &idx, &kts, &smax, &W, &eig, &modes, &k, &p,
                          &rh, &b, &sw.d, &sw.e, &sw.sig, &sw.tauL, &sw.tauR, &sw.Vr, &sw.QBt,
                          &sw.PBt, &sw.work, &sw.gw, &sw.Rb, &sw.tauQ, &sw.qw, &sw.Bs, &sw.uR2,
                          &sw.uL2, &sw.tR2, &sw.tL2, &sw.Pm, &sw.tauP})
            q->reset();
    }
};
static RadCache &rad_cache() {
    static thread_local RadCache *c = new RadCache();
    return *c;
}
void release_rad_cache() { rad_cache().release(); }

}  // namespace ia

using namespace ia;

extern "C" ia_status ia_radiation_sweep(const ia_radiation_desc *desc, const double *k,
                                        const double *r_p, const double *r_h, const ia_z *b,
                                        ia_radiation_out *out) {
    IA_TRY(check_device());
    if (!desc || !k || !r_p || !r_h || !b)
        return fail(IA_ERR_ARG, "ia_radiation_sweep: NULL argument");
    const int np = desc->np, l = desc->l, F = desc->n_freq, n = desc->n_modes;
    if (np < 2 || l <= 0 || F <= 0 || n <= 0 || n > np)
        return fail(IA_ERR_ARG, "ia_radiation_sweep: bad sizes (np >= 2, 1 <= n_modes <= np)");
    if (np > 4096)
        return fail(IA_ERR_ARG, "ia_radiation_sweep: at most 4096 patches (the two-stage SVD limit)");
    if (!(desc->area > 0)) return fail(IA_ERR_ARG, "ia_radiation_sweep: patch area must be > 0");
    IA_TRY(check_reg(desc->reg, std::min(l, n)));
    const bool dk = is_device_ptr(k);
    if (!dk)
        for (int f = 0; f < F; ++f)
            if (!(k[f] > 0) || !std::isfinite(k[f]))
                return fail(IA_ERR_ARG, "ia_radiation_sweep: k[%d] must be > 0", f);
    if (!is_device_ptr(b) && !finite_all((const double *)b, 2 * (size_t)l * F))
        return fail(IA_ERR_NONFINITE, "ia_radiation_sweep: non-finite pressures");
    hipStream_t s = nullptr;
    RadCache &W = rad_cache();
    const double c = air().c, rho = air().rho, A = desc->area;
    // w^2 rho A^2/(4 pi c) = k^2 (c rho A^2 / (4 pi))
    const double scale_c = c * rho * A * A / (4 * M_PI);
    const RegDev prm = to_regdev(desc->reg);
    const double eig_floor = desc->eig_floor == 0.0 ? 1e-8 : std::max(desc->eig_floor, 0.0);
    IA_TRY(to_device(W.k, k, sizeof(double) * F, s));
    IA_TRY(to_device(W.p, r_p, sizeof(double) * 3 * np, s));
    IA_TRY(to_device(W.rh, r_h, sizeof(double) * 3 * l, s));
    IA_TRY(to_device(W.b, b, sizeof(ia_z) * (size_t)l * F, s));
    // frequencies per chunk: the complex copy of R, the stage-one reflector
    // store and the bidiagonal vectors take ~1 GB per frequency at np = 4096
    const size_t per = (size_t)np * np * (sizeof(cx<double>) * 2 + 2 * sizeof(double));
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
        (void)hipGetLastError();
        fr = (size_t)16 << 30;
    }
    // (the chase runs one workgroup per frequency: a larger chunk keeps more
    // compute units busy)
    int chunk = desc->chunk > 0 ? desc->chunk : (int)std::max<size_t>(1, (fr / 5 * 2) / per);
    chunk = std::min(std::min(chunk, F), 128);
    ia_radiation_out o{};
    if (out) o = *out;
    const size_t nn = (size_t)np * np;
    for (int f0 = 0; f0 < F; f0 += chunk) {
        const int nf = std::min(chunk, F - f0);
        const double *kf = W.k.as<double>() + f0;
        IA_TRY(W.S.alloc(sizeof(cx<double>) * nn * nf));
        cx<double> *S = W.S.as<cx<double>>();
        hipLaunchKernelGGL(k_rad_R, dim3((np + RT - 1) / RT, (np + RT - 1) / RT, nf), dim3(256), 0,
                           s, np, kf, W.p.as<double>(), S, nn);
        IA_HIP(hipGetLastError());
        SvdWork<double> &sw = W.sw;
        IA_TRY(sw.alloc(np, np, nf));
        IA_TRY(svd_factor<double>(sw, S, s));
        // the n leading modes: U = U_b [P_B; 0] (columns 0..n-1)
        IA_TRY(W.U.alloc(sizeof(cx<double>) * (size_t)np * n * nf));
        cx<double> *U = W.U.as<cx<double>>();
        hipLaunchKernelGGL(k_first_cols, dim3((np + 255) / 256, n, nf), dim3(256), 0, s, np, n,
                           sw.PBt.as<double>(), U);
        svd_apply_u<double>(sw, S, U, np, n, s);
        double *md = nullptr;
        if (o.modes) {
            IA_TRY(W.modes.alloc(sizeof(double) * (size_t)np * n * nf));
            md = W.modes.as<double>();
        }
        hipLaunchKernelGGL(k_realign, dim3(n, nf), dim3(256), 0, s, np, n, U);
        hipLaunchKernelGGL(k_orth_modes, dim3(nf), dim3(1024), 0, s, np, n, sw.sig.as<double>(),
                           eig_floor, U, md);
        // H_r = G Q
        IA_TRY(W.G.alloc(sizeof(cx<double>) * (size_t)l * np * nf));
        IA_TRY(W.H.alloc(sizeof(cx<double>) * (size_t)l * n * nf));
        cx<double> *G = W.G.as<cx<double>>(), *H = W.H.as<cx<double>>();
        hipLaunchKernelGGL(k_rad_G, dim3((l + RT - 1) / RT, (np + RT - 1) / RT, nf), dim3(256), 0,
                           s, l, np, kf, rho * c * A, W.rh.as<double>(), W.p.as<double>(), G,
                           (size_t)l * np);
        const cx<double> one{1, 0}, zero{0, 0};
        launch_gemm<double>('N', 'N', l, n, np, one, G, l, (size_t)l * np, U, np, (size_t)np * n,
                            zero, H, l, (size_t)l * n, nf, s);
        IA_HIP(hipGetLastError());
        // regularised solve in mode space
        IA_TRY(W.y.alloc(sizeof(cx<double>) * (size_t)n * nf));
        IA_TRY(W.lam.alloc(sizeof(double) * nf));
        IA_TRY(W.idx.alloc(sizeof(int) * nf));
        IA_TRY(W.kts.alloc(sizeof(int) * nf));
        IA_TRY(W.smax.alloc(sizeof(double) * nf));
        const cx<double> *bf = W.b.as<cx<double>>() + (size_t)f0 * l;
        if (desc->precision == IA_PREC_MIXED)
            IA_TRY(regsolve_batched<float>(l, n, nf, H, (size_t)l * n, bf, l, prm,
                                           W.y.as<cx<double>>(), n, W.lam.as<double>(),
                                           W.idx.as<int>(), W.kts.as<int>(), W.smax.as<double>(), s,
                                           nullptr));
        else
            IA_TRY(regsolve_batched<double>(l, n, nf, H, (size_t)l * n, bf, l, prm,
                                            W.y.as<cx<double>>(), n, W.lam.as<double>(),
                                            W.idx.as<int>(), W.kts.as<int>(), W.smax.as<double>(),
                                            s, nullptr));
        // W = sum e |y|^2, v = Q y
        IA_TRY(W.W.alloc(sizeof(double) * nf));
        IA_TRY(W.eig.alloc(sizeof(double) * (size_t)n * nf));
        hipLaunchKernelGGL(k_modal_power, dim3(nf), dim3(256), 0, s, np, n, kf, scale_c,
                           sw.sig.as<double>(), W.y.as<cx<double>>(), W.eig.as<double>(),
                           W.W.as<double>());
        if (o.velocity) {
            IA_TRY(W.v.alloc(sizeof(cx<double>) * (size_t)np * nf));
            launch_gemm<double>('N', 'N', np, 1, n, one, U, np, (size_t)np * n,
                                W.y.as<cx<double>>(), n, n, zero, W.v.as<cx<double>>(), np, np, nf,
                                s);
        }
        IA_HIP(hipGetLastError());
        IA_HIP(hipStreamSynchronize(s));
        if (o.power) IA_HIP(hipMemcpy(o.power + f0, W.W.p, sizeof(double) * nf, hipMemcpyDefault));
        if (o.lambda) IA_HIP(hipMemcpy(o.lambda + f0, W.lam.p, sizeof(double) * nf, hipMemcpyDefault));
        if (o.lambda_index)
            IA_HIP(hipMemcpy(o.lambda_index + f0, W.idx.p, sizeof(int) * nf, hipMemcpyDefault));
        if (o.k_tsvd) IA_HIP(hipMemcpy(o.k_tsvd + f0, W.kts.p, sizeof(int) * nf, hipMemcpyDefault));
        if (o.eigval)
            IA_HIP(hipMemcpy(o.eigval + (size_t)f0 * n, W.eig.p, sizeof(double) * n * nf,
                             hipMemcpyDefault));
        if (o.modes)
            IA_HIP(hipMemcpy(o.modes + (size_t)f0 * np * n, W.modes.p,
                             sizeof(double) * (size_t)np * n * nf, hipMemcpyDefault));
        if (o.amplitude)
            IA_HIP(hipMemcpy(o.amplitude + (size_t)f0 * n, W.y.p, sizeof(ia_z) * (size_t)n * nf,
                             hipMemcpyDefault));
        if (o.velocity)
            IA_HIP(hipMemcpy(o.velocity + (size_t)f0 * np, W.v.p, sizeof(ia_z) * (size_t)np * nf,
                             hipMemcpyDefault));
        if (o.status) {
            std::vector<int> ok(nf, IA_OK);
            IA_HIP(hipMemcpy(o.status + f0, ok.data(), sizeof(int) * nf, hipMemcpyDefault));
        }
    }
    return IA_OK;
}
