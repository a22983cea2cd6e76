// Gradients of the ESM regularised solve (hand-written adjoint).
//
// The reference has no derivative code; the chain differentiated here is its
// forward path: sources q_j = r_s,j - d n_j, D and M (CUDA/esm.cuh:1-36,
// MATLAB/esmV.m:20-29), H_e = M_hq D^-1 (esm.cuh:59-92), the standard-form
// Tikhonov solution x at a fixed lambda (tikhonovG.m:29-30), the misfit
// J = ||H_e x - b||^2 and the sound power W = 1/2 Re(x^H M_sq D^-1 x) S/m
// (to_pressure / sound_power, CUDA/utils.cuh:199-289).
//
// With A = H_e = U diag(s) V^H (l < m), N = A^H A + lam^2 I, z = N^-1 x,
// r = A x - b, P = M_sq D^-1, h = (P + P^H) x, w = N^-1 h, q = D^-1 x,
// the first-order changes are (Re <G, dM> = Re sum conj(G_ij) dM_ij):
//   dJ = Re <G_A^J, dA>,  G_A^J = 2 r (x + lam^2 z)^H + 2 lam^2 (A z) x^H
//   dW = Re <G_A^W, dA> + Re <S/2m x q^H, dM_sq> - Re <S/2m (P^H x) q^H, dD>,
//        G_A^W = -(S/2m) (r w^H + (A w) x^H)
// and dA = dM_hq D^-1 - A dD D^-1 moves every cotangent onto the assembled
// matrices as rank-two outer products (A^H r = -lam^2 x and
// A^H A y = N y - lam^2 y remove every product with A^H):
//   M_hq:  J: r c1^H + (A z) c2^H,  c1 = 2 (q + lam^2 D^-1 z), c2 = 2 lam^2 q
//          W: -(S/2m) (r w_^H + (A w) q^H),  w_ = D^-1 w
//   D:     J: lam^2 x c1^H - (x - lam^2 z) c2^H
//          W: (S/2m) (-lam^2 x w_^H + (p_s - lam^2 w) q^H),  p_s = M_sq q
//   M_sq:  W: (S/2m) x q^H
//   log lam: dJ = 4 lam^4 Re(x^H z),  dW = -(S/m) lam^2 Re(h^H z).
// The matrices are never stored for the backward pass: the reduction
// kernels recompute every entry's derivative with respect to d and the
// microphone positions and contract it with the outer products on the fly.
// z, w and x come from the bidiagonal SVD of the forward pass (only the
// vectors are pushed through its reflectors); N^-1 h = h / lam^2 -
// V diag(s^2 / (lam^2 (s^2 + lam^2))) V^H h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "ia_internal.hpp"

namespace ia {

RegDev to_regdev(const ia_reg_params &p);
ia_status check_reg(const ia_reg_params &p, int pmax);

namespace {

constexpr double INV4PI = 0.079577471545947667884;
constexpr int GT = 256;  // threads of the gradient reductions

__device__ __forceinline__ double wsum64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// block sum over GT threads, result in every thread
__device__ __forceinline__ double bsum(double v, double *red) {
    v = wsum64(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0;
#pragma unroll
    for (int i = 0; i < GT / 64; ++i) s += red[i];
    return s;
}

// src = r_s - d nrm (3m, SoA)
__global__ void k_retreat(int n3, const double *__restrict__ r_s, const double *__restrict__ nrm,
                          double d, double *__restrict__ src) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) src[i] = r_s[i] - d * nrm[i];
}

// coordinates in the right singular basis of A: vec 0 = x (s beta / g, the
// Tikhonov coefficients of the sweep's selection kernel), vec 1 =
// z = N^-1 x (s beta / g^2), g = s^2 + lam^2
__global__ void k_coef_xz(int C, const double *__restrict__ sig, const cx<double> *__restrict__ beta,
                          const double *__restrict__ lam, cx<double> *__restrict__ co) {
    const int f = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    const double s = sig[(size_t)f * C + i], L = lam[f];
    const double g = s * s + L * L;
    const cx<double> b = beta[(size_t)f * C + i];
    const double a = s / g;
    co[(size_t)f * 2 * C + i] = a * b;
    co[(size_t)f * 2 * C + C + i] = (a / g) * b;
}

// coordinates of V diag(-s^2 / (lam^2 (s^2 + lam^2))) t
__global__ void k_coef_w(int C, const double *__restrict__ sig, const double *__restrict__ lam,
                         cx<double> *__restrict__ t) {
    const int f = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    const double s = sig[(size_t)f * C + i], L2 = lam[f] * lam[f];
    const double a = -(s * s) / (L2 * (s * s + L2));
    t[(size_t)f * C + i] = a * t[(size_t)f * C + i];
}

// out_i = sum_j M(p_i, q_j) in_j (CT = false, thread per point i) or
// out_j = sum_i conj(M(p_i, q_j)) in_i (CT = true, thread per source j),
// M = i rho c k e^{ikR} / (4 pi R), NV vectors at once.  in / out: vector v
// of frequency f at f * sf + v * sv (input and output alike).
template <int NV, bool CT>
__global__ __launch_bounds__(GT) void k_m_apply(int npts, int m, const double *__restrict__ kv,
                                               double rho_c, const double *__restrict__ pts,
                                               const double *__restrict__ src,
                                               const cx<double> *__restrict__ in, size_t sfi,
                                               size_t svi, cx<double> *__restrict__ out,
                                               size_t sfo, size_t svo) {
    __shared__ double sx[3][GT];
    __shared__ cx<double> sv[NV][GT];
    const int f = blockIdx.y, t = threadIdx.x;
    const int nout = CT ? m : npts, nin = CT ? npts : m;
    const double *opts = CT ? src : pts;
    const double *ipts = CT ? pts : src;
    const int o = blockIdx.x * GT + t;
    const double k = kv[f], amp = rho_c * k * INV4PI;
    double x = 0, y = 0, z = 0;
    if (o < nout) {
        x = opts[o];
        y = opts[o + nout];
        z = opts[o + 2 * nout];
    }
    double ar[NV], ai[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) ar[v] = ai[v] = 0;
    for (int j0 = 0; j0 < nin; j0 += GT) {
        __syncthreads();
        const int j = j0 + t;
        if (j < nin) {
            sx[0][t] = ipts[j];
            sx[1][t] = ipts[j + nin];
            sx[2][t] = ipts[j + 2 * nin];
#pragma unroll
            for (int v = 0; v < NV; ++v) sv[v][t] = in[(size_t)f * sfi + (size_t)v * svi + j];
        }
        __syncthreads();
        const int jn = min(GT, nin - j0);
        for (int jj = 0; jj < jn; ++jj) {
            const double dx = x - sx[0][jj], dy = y - sx[1][jj], dz = z - sx[2][jj];
            const double R = sqrt(dx * dx + dy * dy + dz * dz);
            double sn, cs;
            sincos(k * R, &sn, &cs);
            const double a = amp / R;
            const double mre = -a * sn, mim = CT ? -a * cs : a * cs;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const cx<double> q = sv[v][jj];
                ar[v] += mre * q.re - mim * q.im;
                ai[v] += mre * q.im + mim * q.re;
            }
        }
    }
    if (o < nout)
#pragma unroll
        for (int v = 0; v < NV; ++v)
            out[(size_t)f * sfo + (size_t)v * svo + o] = cx<double>{ar[v], ai[v]};
}

// elementwise helpers over nf frequencies of n entries
// r = Ax - b
__global__ void k_sub(int n, const cx<double> *__restrict__ a, size_t sa,
                      const cx<double> *__restrict__ b, size_t sb, cx<double> *__restrict__ out,
                      size_t so) {
    const int f = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[(size_t)f * so + i] = a[(size_t)f * sa + i] - b[(size_t)f * sb + i];
}
// out = a + b (mode 0) or a + b / lam_f^2 (mode 1)
__global__ void k_axpy(int n, const cx<double> *__restrict__ a, size_t sa,
                       const cx<double> *__restrict__ b, size_t sb, const double *__restrict__ lam,
                       int mode, cx<double> *__restrict__ out, size_t so) {
    const int f = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double L = lam[f];
    const double c = mode == 0 ? 1.0 : 1.0 / (L * L);
    out[(size_t)f * so + i] = a[(size_t)f * sa + i] + c * b[(size_t)f * sb + i];
}

// Per-frequency scalars: J = |r|^2, W = (S/2m) Re(x^H p_s),
// dJ/dlog lam = 4 lam^4 Re(x^H z), dW/dlog lam = -(S/m) lam^2 Re(h^H z).
__global__ __launch_bounds__(GT) void k_scalars(int l, int m, const cx<double> *__restrict__ r,
                                               const cx<double> *__restrict__ X2,
                                               const cx<double> *__restrict__ ps,
                                               const cx<double> *__restrict__ h,
                                               const double *__restrict__ lam, double SW,
                                               double *__restrict__ out) {
    __shared__ double red[GT / 64];
    const int f = blockIdx.x, t = threadIdx.x;
    double a = 0;
    for (int i = t; i < l; i += GT) a += abs2(r[(size_t)f * l + i]);
    const double J = bsum(a, red);
    double w = 0, xz = 0, hz = 0;
    const cx<double> *x = X2 + (size_t)f * 2 * m, *z = x + m;
    for (int i = t; i < m; i += GT) {
        const cx<double> xi = x[i], zi = z[i], pi = ps[(size_t)f * m + i], hi = h[(size_t)f * m + i];
        w += xi.re * pi.re + xi.im * pi.im;
        xz += xi.re * zi.re + xi.im * zi.im;
        hz += hi.re * zi.re + hi.im * zi.im;
    }
    w = bsum(w, red);
    xz = bsum(xz, red);
    hz = bsum(hz, red);
    if (t == 0) {
        const double L2 = lam[f] * lam[f];
        out[(size_t)f * 4 + 0] = J;
        out[(size_t)f * 4 + 1] = SW * w;
        out[(size_t)f * 4 + 2] = 4.0 * L2 * L2 * xz;
        out[(size_t)f * 4 + 3] = -2.0 * SW * L2 * hz;
    }
}

// Column data of the reductions, per source j: position, normal and the
// four column weights c1, c2, w_ = D^-1 w, q.
struct ColW {
    double x, y, z, nx, ny, nz;
    cx<double> c1, c2, wh, q;
};

// Microphone rows: for mic i and source j, grad_{p_i} M_ij = kap G1 Delta
// and dM_ij/dd = kap G1 (Delta . n_j), kap = i rho c k / (4 pi),
// G1 = e^{ikR} (ikR - 1) / R^3, Delta = p_i - q_j.  Thread per mic: the mic
// gradients are complete per thread; the d parts are reduced per workgroup.
__global__ __launch_bounds__(GT) void k_grad_mic(
    int l, int m, const double *__restrict__ kv, double rho_c, const double *__restrict__ r_h,
    const double *__restrict__ src, const double *__restrict__ nrm,
    const cx<double> *__restrict__ rr, const cx<double> *__restrict__ AX,
    const cx<double> *__restrict__ Aw, const cx<double> *__restrict__ QZ,
    const cx<double> *__restrict__ wh, const double *__restrict__ lam, double SW,
    double *__restrict__ gmicJ, double *__restrict__ gmicW, double *__restrict__ part,
    int nparts) {
    __shared__ ColW cw[GT];
    __shared__ double red[GT / 64];
    const int f = blockIdx.y, t = threadIdx.x;
    const int i = blockIdx.x * GT + t;
    const double k = kv[f], L2 = lam[f] * lam[f];
    const double kap = rho_c * k * INV4PI;  // kap = i * this
    double px = 0, py = 0, pz = 0;
    cx<double> ri{0, 0}, azi{0, 0}, awi{0, 0};
    if (i < l) {
        px = r_h[i];
        py = r_h[i + l];
        pz = r_h[i + 2 * l];
        ri = rr[(size_t)f * l + i];
        azi = AX[(size_t)f * 2 * l + l + i];
        awi = Aw[(size_t)f * l + i];
    }
    const cx<double> *q = QZ + (size_t)f * 2 * m, *dz = q + m;
    double gJ[3] = {0, 0, 0}, gW[3] = {0, 0, 0}, dJ = 0, dW = 0;
    for (int j0 = 0; j0 < m; j0 += GT) {
        __syncthreads();
        const int j = j0 + t;
        if (j < m) {
            ColW c;
            c.x = src[j];
            c.y = src[j + m];
            c.z = src[j + 2 * m];
            c.nx = nrm[j];
            c.ny = nrm[j + m];
            c.nz = nrm[j + 2 * m];
            const cx<double> qj = q[j], dzj = dz[j];
            c.c1 = 2.0 * (qj + L2 * dzj);
            c.c2 = (2.0 * L2) * qj;
            c.wh = wh[(size_t)f * m + j];
            c.q = qj;
            cw[t] = c;
        }
        __syncthreads();
        if (i < l) {
            const int jn = min(GT, m - j0);
            for (int jj = 0; jj < jn; ++jj) {
                const ColW &c = cw[jj];
                const double dx = px - c.x, dy = py - c.y, dzz = pz - c.z;
                const double R2 = dx * dx + dy * dy + dzz * dzz, R = sqrt(R2);
                double sn, cs;
                sincos(k * R, &sn, &cs);
                // kap G1 = i kap' (ikR - 1) e^{ikR} / R^3
                //        = kap' / R^3 * i (-cs - kR sn + i (kR cs - sn))
                //        = kap' / R^3 * (sn - kR cs + i (-cs - kR sn))
                const double kR = k * R, sc = kap / (R2 * R);
                const cx<double> g1{sc * (sn - kR * cs), sc * (-cs - kR * sn)};
                // conj(G_ij): J: conj(r_i) c1_j + conj(Az_i) c2_j
                //             W: -SW (conj(r_i) w_j + conj(Aw_i) q_j)
                const cx<double> wJ = cmulc(ri, c.c1) + cmulc(azi, c.c2);
                const cx<double> wW = (-SW) * (cmulc(ri, c.wh) + cmulc(awi, c.q));
                const cx<double> pj = wJ * g1, pw = wW * g1;
                gJ[0] += pj.re * dx;
                gJ[1] += pj.re * dy;
                gJ[2] += pj.re * dzz;
                gW[0] += pw.re * dx;
                gW[1] += pw.re * dy;
                gW[2] += pw.re * dzz;
                const double dn = dx * c.nx + dy * c.ny + dzz * c.nz;
                dJ += pj.re * dn;
                dW += pw.re * dn;
            }
        }
    }
    if (i < l) {
        for (int c = 0; c < 3; ++c) {
            gmicJ[(size_t)f * 3 * l + (size_t)c * l + i] = gJ[c];
            gmicW[(size_t)f * 3 * l + (size_t)c * l + i] = gW[c];
        }
    }
    dJ = bsum(dJ, red);
    dW = bsum(dW, red);
    if (t == 0) {
        part[((size_t)f * nparts + blockIdx.x) * 2 + 0] = dJ;
        part[((size_t)f * nparts + blockIdx.x) * 2 + 1] = dW;
    }
}

// Surface rows: D_ij = F (n_i . Delta), F = e^{ikR} (ikR - 1) / (4 pi R^3),
// Delta = r_i - q_j:  dD_ij/dd = F' (n_i . Delta)(Delta . n_j) / R + F (n_i . n_j),
// F' = e^{ikR} (3 - 3ikR - k^2 R^2) / (4 pi R^4); dM_sq,ij/dd = kap G1 (Delta . n_j).
__global__ __launch_bounds__(GT) void k_grad_surf(
    int m, const double *__restrict__ kv, double rho_c, const double *__restrict__ r_s,
    const double *__restrict__ src, const double *__restrict__ nrm,
    const cx<double> *__restrict__ X2, const cx<double> *__restrict__ QZ,
    const cx<double> *__restrict__ ps, const cx<double> *__restrict__ wv,
    const cx<double> *__restrict__ wh, const double *__restrict__ lam, double SW,
    double *__restrict__ part, int nparts, int poff) {
    __shared__ ColW cw[GT];
    __shared__ double red[GT / 64];
    const int f = blockIdx.y, t = threadIdx.x;
    const int i = blockIdx.x * GT + t;
    const double k = kv[f], L2 = lam[f] * lam[f];
    const double kap = rho_c * k * INV4PI;
    double px = 0, py = 0, pz = 0, nx = 0, ny = 0, nz = 0;
    cx<double> xi{0, 0}, ui{0, 0}, vi{0, 0};
    const cx<double> *x = X2 + (size_t)f * 2 * m, *z = x + m;
    if (i < m) {
        px = r_s[i];
        py = r_s[i + m];
        pz = r_s[i + 2 * m];
        nx = nrm[i];
        ny = nrm[i + m];
        nz = nrm[i + 2 * m];
        xi = x[i];
        ui = xi - L2 * z[i];                                   // x - lam^2 z
        vi = ps[(size_t)f * m + i] - L2 * wv[(size_t)f * m + i];  // p_s - lam^2 w
    }
    const cx<double> *q = QZ + (size_t)f * 2 * m, *dz = q + m;
    double dJ = 0, dW = 0;
    for (int j0 = 0; j0 < m; j0 += GT) {
        __syncthreads();
        const int j = j0 + t;
        if (j < m) {
            ColW c;
            c.x = src[j];
            c.y = src[j + m];
            c.z = src[j + 2 * m];
            c.nx = nrm[j];
            c.ny = nrm[j + m];
            c.nz = nrm[j + 2 * m];
            const cx<double> qj = q[j], dzj = dz[j];
            c.c1 = 2.0 * (qj + L2 * dzj);
            c.c2 = (2.0 * L2) * qj;
            c.wh = wh[(size_t)f * m + j];
            c.q = qj;
            cw[t] = c;
        }
        __syncthreads();
        if (i < m) {
            const int jn = min(GT, m - j0);
            for (int jj = 0; jj < jn; ++jj) {
                const ColW &c = cw[jj];
                const double dx = px - c.x, dy = py - c.y, dzz = pz - c.z;
                const double R2 = dx * dx + dy * dy + dzz * dzz, R = sqrt(R2);
                double sn, cs;
                sincos(k * R, &sn, &cs);
                const double kR = k * R;
                const double niD = nx * dx + ny * dy + nz * dzz;
                const double Dnj = dx * c.nx + dy * c.ny + dzz * c.nz;
                const double ninj = nx * c.nx + ny * c.ny + nz * c.nz;
                // (ikR - 1) e^{ikR} = (-cs - kR sn) + i (kR cs - sn)
                const double ir3 = INV4PI / (R2 * R);
                const cx<double> F{(-cs - kR * sn) * ir3, (kR * cs - sn) * ir3};
                // (3 - 3ikR - k^2R^2) e^{ikR}
                const double a = 3.0 - kR * kR, bq = -3.0 * kR;
                const double ir4 = ir3 / R;
                const cx<double> Fp{(a * cs - bq * sn) * ir4, (a * sn + bq * cs) * ir4};
                const cx<double> dD = (niD * Dnj / R) * Fp + ninj * F;
                // dM_sq/dd = i kap' G1 (Delta . n_j), i kap' G1 = 4 pi kap' i F
                const double sc = 4.0 * M_PI * kap * Dnj;
                const cx<double> dM{-sc * F.im, sc * F.re};
                // conj(G^J_D)_ij = lam^2 conj(x_i) c1_j - conj(u_i) c2_j
                const cx<double> wJ = L2 * cmulc(xi, c.c1) - cmulc(ui, c.c2);
                // conj(G^W_D)_ij = SW (-lam^2 conj(x_i) w_j + conj(v_i) q_j),
                // conj(G^W_S)_ij = SW conj(x_i) q_j
                const cx<double> wW = SW * ((-L2) * cmulc(xi, c.wh) + cmulc(vi, c.q));
                const cx<double> wS = SW * cmulc(xi, c.q);
                dJ += (wJ * dD).re;
                dW += (wW * dD).re + (wS * dM).re;
            }
        }
    }
    dJ = bsum(dJ, red);
    dW = bsum(dW, red);
    if (t == 0) {
        part[((size_t)f * nparts + poff + blockIdx.x) * 2 + 0] = dJ;
        part[((size_t)f * nparts + poff + blockIdx.x) * 2 + 1] = dW;
    }
}

// d-gradients: the workgroup partials of both reductions, summed in a fixed order
__global__ void k_sum_parts(int nparts, const double *__restrict__ part, double *__restrict__ out,
                            int nf) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    double a = 0, b = 0;
    for (int p = 0; p < nparts; ++p) {
        a += part[((size_t)f * nparts + p) * 2];
        b += part[((size_t)f * nparts + p) * 2 + 1];
    }
    out[(size_t)f * 2] = a;
    out[(size_t)f * 2 + 1] = b;
}

}  // namespace

// Workspace of one gradient batch (kept per host thread between calls).
// Allocated with new and never destroyed at thread exit, so no hipFree runs
// after the HIP runtime has shut down; ia_release() frees the buffers.
struct GradCache {
    DevBuf D, Y, ipiv, perm, info, c, beta, cn2, bn2, co, tmp, vR, X2, QZ, AX, r, ps, t1, h, hR,
        tco, wv, wh, Aw, part, phi, kidx, src;
    SvdWork<double> sw;
    void release() {
        for (DevBuf *b : {&D, &Y, &ipiv, &perm, &info, &c, &beta, &cn2, &bn2, &co, &tmp, &vR, &X2,
                          &QZ, &AX, &r, &ps, &t1, &h, &hR, &tco, &wv, &wh, &Aw, &part, &phi, &kidx,
                          &src, &sw.d, &sw.e, &sw.sig, &sw.tauL, &sw.tauR, &sw.Vr, &sw.QBt,
                          &sw.PBt, &sw.work, &sw.gw, &sw.Rb, &sw.tauQ, &sw.qw, &sw.Bs, &sw.uR2,
                          &sw.uL2, &sw.tR2, &sw.tL2, &sw.Pm, &sw.tauP})
            b->reset();
    }
};
static GradCache &grad_cache() {
    static thread_local GradCache *c = new GradCache();
    return *c;
}

void release_grad_cache() { grad_cache().release(); }

static ia_status gradient_batch(int m, int l, int nf, const ia_grad_desc &desc, const double *k,
                                const double *r_s, const double *nrm, const double *src,
                                const double *r_h, const cx<double> *b, const double *lam_in,
                                bool select, double *lam_dev, double *scal, double *dd,
                                double *gmicJ, double *gmicW, cx<double> *vel, hipStream_t s) {
    GradCache &W = grad_cache();
    const int R = m, C = l;
    const double rho_c = air().rho * air().c, SW = desc.area / (2.0 * m);
    const size_t mm = (size_t)m * m;
    IA_TRY(W.D.alloc(sizeof(cx<double>) * nf * mm));
    IA_TRY(W.Y.alloc(sizeof(cx<double>) * nf * (size_t)m * l));
    IA_TRY(W.ipiv.alloc(sizeof(int) * nf * (size_t)m));
    IA_TRY(W.perm.alloc(2 * sizeof(int) * nf * (size_t)m));
    IA_TRY(W.info.alloc(sizeof(int) * nf));
    cx<double> *Dm = W.D.as<cx<double>>(), *Y = W.Y.as<cx<double>>();
    int *idx = W.perm.as<int>(), *inv = idx + (size_t)m * nf;
    // ---- forward: assembly, LU, Y = P H_e^H, SVD
    launch_esm_D<double>(m, nf, k, r_s, src, nrm, Dm, mm, s, false);
    launch_esm_M<double>(l, m, nf, k, rho_c, r_h, src, Y, (size_t)m * l, true, s);
    IA_HIP(hipGetLastError());
    IA_HIP(hipMemsetAsync(W.info.p, 0, sizeof(int) * nf, s));
    IA_TRY(getrf_batched<double>(m, Dm, m, mm, W.ipiv.as<int>(), W.info.as<int>(), nf, s));
    IA_TRY(getrs_c_permuted<double>(m, l, Dm, m, mm, Y, m, (size_t)m * l, nf, s));
    IA_TRY(lu_perm_index(m, W.ipiv.as<int>(), idx, inv, nf, s));
    SvdWork<double> &sw = W.sw;
    IA_TRY(sw.alloc(R, C, nf));
    IA_TRY(svd_factor<double>(sw, Y, s));
    const double *sig = sw.sig.as<double>();
    // beta = Q_B^T V_b^H b (coordinates of b in the left singular basis of A)
    IA_TRY(W.c.alloc(sizeof(cx<double>) * nf * (size_t)C));
    IA_TRY(W.beta.alloc(sizeof(cx<double>) * nf * (size_t)C));
    IA_TRY(W.cn2.alloc(sizeof(double) * nf));
    IA_TRY(W.bn2.alloc(sizeof(double) * nf));
    cx<double> *cv = W.c.as<cx<double>>(), *beta = W.beta.as<cx<double>>();
    IA_HIP(hipMemcpyAsync(cv, b, sizeof(cx<double>) * nf * (size_t)l, hipMemcpyDeviceToDevice, s));
    launch_norm2<double>(l, cv, l, W.bn2.as<double>(), nf, s);
    svd_apply_v<double>(sw, true, cv, C, 1, s);
    launch_norm2<double>(C, cv, C, W.cn2.as<double>(), nf, s);
    launch_real_matvec(C, sw.QBt.as<double>(), 0, cv, C, beta, C, nf, s);
    if (select) {
        // the parameter of the forward sweep, then held fixed
        IA_TRY(W.phi.alloc(sizeof(cx<double>) * nf * (size_t)C));
        IA_TRY(W.kidx.alloc(2 * sizeof(int) * nf));
        const RegDev prm = to_regdev(desc.reg);
        launch_select(C, l, sig, beta, W.bn2.as<double>(), prm, lam_dev, W.kidx.as<int>(),
                      W.kidx.as<int>() + nf, W.phi.as<cx<double>>(), nf, s, nullptr,
                      W.cn2.as<double>(), svd_floor_rel<double>());
    } else {
        IA_HIP(hipMemcpyAsync(lam_dev, lam_in, sizeof(double) * nf, hipMemcpyDeviceToDevice, s));
    }
    // ---- x and z = N^-1 x in physical order: V_A c = P^T U_b [P_B c; 0]
    IA_TRY(W.co.alloc(sizeof(cx<double>) * nf * 2 * (size_t)C));
    IA_TRY(W.tmp.alloc(sizeof(cx<double>) * nf * 2 * (size_t)C));
    IA_TRY(W.vR.alloc(sizeof(cx<double>) * nf * 2 * (size_t)R));
    IA_TRY(W.X2.alloc(sizeof(cx<double>) * nf * 2 * (size_t)m));
    cx<double> *co = W.co.as<cx<double>>(), *tmp = W.tmp.as<cx<double>>(),
               *vR = W.vR.as<cx<double>>(), *X2 = W.X2.as<cx<double>>();
    const dim3 gC((C + 255) / 256, nf), gm((m + 255) / 256, nf), gl((l + 255) / 256, nf);
    hipLaunchKernelGGL(k_coef_xz, gC, dim3(256), 0, s, C, sig, beta, lam_dev, co);
    auto apply_VA = [&](const cx<double> *coef, size_t scoef, int nvec, cx<double> *out,
                        size_t sout) {
        // coef: vector v of frequency f at f scoef + v C; out likewise with m
        for (int v = 0; v < nvec; ++v) {
            launch_real_matvec(C, sw.PBt.as<double>(), 1, coef + (size_t)v * C, scoef,
                               tmp + (size_t)v * C, (size_t)nvec * C, nf, s);
            launch_pad_copy(C, R, tmp + (size_t)v * C, (size_t)nvec * C, vR + (size_t)v * R,
                            (size_t)nvec * R, nf, s);
        }
        svd_apply_u<double>(sw, Y, vR, R, nvec, s);
        launch_row_gather<double>(m, nvec, inv, vR, R, (size_t)nvec * R, out, (int)m, sout, nf, s);
    };
    apply_VA(co, 2 * (size_t)C, 2, X2, 2 * (size_t)m);
    // ---- q = D^-1 x, D^-1 z; A x, A z; r = A x - b; p_s; P^H x; h
    IA_TRY(W.QZ.alloc(sizeof(cx<double>) * nf * 2 * (size_t)m));
    cx<double> *QZ = W.QZ.as<cx<double>>();
    IA_HIP(hipMemcpyAsync(QZ, X2, sizeof(cx<double>) * nf * 2 * (size_t)m, hipMemcpyDeviceToDevice,
                          s));
    IA_TRY(getrs_batched<double>('N', m, 2, Dm, m, mm, W.ipiv.as<int>(), QZ, m, 2 * (size_t)m, nf,
                                 s));
    IA_TRY(W.AX.alloc(sizeof(cx<double>) * nf * 2 * (size_t)l));
    IA_TRY(W.r.alloc(sizeof(cx<double>) * nf * (size_t)l));
    IA_TRY(W.ps.alloc(sizeof(cx<double>) * nf * (size_t)m));
    IA_TRY(W.t1.alloc(sizeof(cx<double>) * nf * (size_t)m));
    IA_TRY(W.h.alloc(sizeof(cx<double>) * nf * (size_t)m));
    cx<double> *AX = W.AX.as<cx<double>>(), *rr = W.r.as<cx<double>>(), *ps = W.ps.as<cx<double>>(),
               *t1 = W.t1.as<cx<double>>(), *hv = W.h.as<cx<double>>();
    hipLaunchKernelGGL((k_m_apply<2, false>), gl, dim3(GT), 0, s, l, m, k, rho_c, r_h, src, QZ,
                       2 * (size_t)m, (size_t)m, AX, 2 * (size_t)l, (size_t)l);
    hipLaunchKernelGGL(k_sub, gl, dim3(256), 0, s, l, AX, 2 * (size_t)l, b, (size_t)l, rr, (size_t)l);
    hipLaunchKernelGGL((k_m_apply<1, false>), gm, dim3(GT), 0, s, m, m, k, rho_c, r_s, src, QZ,
                       2 * (size_t)m, (size_t)m, ps, (size_t)m, (size_t)m);
    hipLaunchKernelGGL((k_m_apply<1, true>), gm, dim3(GT), 0, s, m, m, k, rho_c, r_s, src, X2,
                       2 * (size_t)m, (size_t)m, t1, (size_t)m, (size_t)m);
    IA_TRY(getrs_batched<double>('C', m, 1, Dm, m, mm, W.ipiv.as<int>(), t1, m, (size_t)m, nf, s));
    hipLaunchKernelGGL(k_axpy, gm, dim3(256), 0, s, m, ps, (size_t)m, t1, (size_t)m, lam_dev, 0, hv,
                       (size_t)m);
    // ---- w = N^-1 h = h / lam^2 + V_A diag(-s^2/(lam^2 (s^2 + lam^2))) V_A^H h
    IA_TRY(W.hR.alloc(sizeof(cx<double>) * nf * (size_t)R));
    IA_TRY(W.tco.alloc(sizeof(cx<double>) * nf * (size_t)C));
    IA_TRY(W.wv.alloc(sizeof(cx<double>) * nf * (size_t)m));
    cx<double> *hR = W.hR.as<cx<double>>(), *tco = W.tco.as<cx<double>>(), *wv = W.wv.as<cx<double>>();
    launch_row_gather<double>(m, 1, idx, hv, m, (size_t)m, hR, R, (size_t)R, nf, s);
    svd_apply_uh<double>(sw, Y, hR, R, 1, s);
    launch_real_matvec(C, sw.PBt.as<double>(), 0, hR, R, tco, C, nf, s);
    hipLaunchKernelGGL(k_coef_w, gC, dim3(256), 0, s, C, sig, lam_dev, tco);
    apply_VA(tco, (size_t)C, 1, t1, (size_t)m);
    hipLaunchKernelGGL(k_axpy, gm, dim3(256), 0, s, m, t1, (size_t)m, hv, (size_t)m, lam_dev, 1, wv,
                       (size_t)m);
    // ---- w_ = D^-1 w, A w
    IA_TRY(W.wh.alloc(sizeof(cx<double>) * nf * (size_t)m));
    IA_TRY(W.Aw.alloc(sizeof(cx<double>) * nf * (size_t)l));
    cx<double> *wh = W.wh.as<cx<double>>(), *Aw = W.Aw.as<cx<double>>();
    IA_HIP(hipMemcpyAsync(wh, wv, sizeof(cx<double>) * nf * (size_t)m, hipMemcpyDeviceToDevice, s));
    IA_TRY(getrs_batched<double>('N', m, 1, Dm, m, mm, W.ipiv.as<int>(), wh, m, (size_t)m, nf, s));
    hipLaunchKernelGGL((k_m_apply<1, false>), gl, dim3(GT), 0, s, l, m, k, rho_c, r_h, src, wh,
                       (size_t)m, (size_t)m, Aw, (size_t)l, (size_t)l);
    // ---- scalars and the two reductions
    hipLaunchKernelGGL(k_scalars, dim3(nf), dim3(GT), 0, s, l, m, rr, X2, ps, hv, lam_dev, SW, scal);
    const int pmic = (l + GT - 1) / GT, psurf = (m + GT - 1) / GT, np = pmic + psurf;
    IA_TRY(W.part.alloc(sizeof(double) * 2 * (size_t)np * nf));
    hipLaunchKernelGGL(k_grad_mic, dim3(pmic, nf), dim3(GT), 0, s, l, m, k, rho_c, r_h, src, nrm, rr,
                       AX, Aw, QZ, wh, lam_dev, SW, gmicJ, gmicW, W.part.as<double>(), np);
    hipLaunchKernelGGL(k_grad_surf, dim3(psurf, nf), dim3(GT), 0, s, m, k, rho_c, r_s, src, nrm, X2,
                       QZ, ps, wv, wh, lam_dev, SW, W.part.as<double>(), np, pmic);
    hipLaunchKernelGGL(k_sum_parts, dim3((nf + 63) / 64), dim3(64), 0, s, np, W.part.as<double>(),
                       dd, nf);
    if (vel)
        IA_HIP(hipMemcpy2DAsync(vel, sizeof(cx<double>) * m, X2, sizeof(cx<double>) * 2 * m,
                                sizeof(cx<double>) * m, nf, hipMemcpyDefault, s));
    IA_HIP(hipGetLastError());
    return IA_OK;
}

}  // namespace ia

using namespace ia;

extern "C" ia_status ia_esm_gradient(const ia_grad_desc *desc, const double *k, const double *r_s,
                                     const double *nrm, double d, const double *r_h, const ia_z *b,
                                     const double *lambda, ia_grad_out *out) {
    IA_TRY(check_device());
    if (!desc || !k || !r_s || !nrm || !r_h || !b)
        return fail(IA_ERR_ARG, "ia_esm_gradient: NULL argument");
    const int m = desc->m, l = desc->l, F = desc->n_freq;
    if (m <= 0 || l <= 0 || F <= 0) return fail(IA_ERR_ARG, "ia_esm_gradient: bad sizes");
    if (l > m)
        return fail(IA_ERR_ARG, "ia_esm_gradient: needs l <= m (more sources than microphones)");
    if (!std::isfinite(d)) return fail(IA_ERR_NONFINITE, "ia_esm_gradient: non-finite d");
    const bool select = (lambda == nullptr);
    if (select) {
        IA_TRY(check_reg(desc->reg, l));
        if (desc->reg.filter != IA_FILTER_TIKHONOV)
            return fail(IA_ERR_ARG, "ia_esm_gradient: the selection must use the Tikhonov filter");
    }
    if (!is_device_ptr(k))
        for (int f = 0; f < F; ++f)
            if (!(k[f] > 0) || !std::isfinite(k[f]))
                return fail(IA_ERR_ARG, "ia_esm_gradient: k[%d] must be > 0", f);
    if (lambda && !is_device_ptr(lambda))
        for (int f = 0; f < F; ++f)
            if (!(lambda[f] > 0) || !std::isfinite(lambda[f]))
                return fail(IA_ERR_ARG, "ia_esm_gradient: lambda[%d] must be > 0", f);
    hipStream_t s = nullptr;
    GradCache &W = grad_cache();
    DevBuf dk, drs, dnrm, drh, db, dlin;
    IA_TRY(to_device(dk, k, sizeof(double) * F, s));
    IA_TRY(to_device(drs, r_s, sizeof(double) * 3 * m, s));
    IA_TRY(to_device(dnrm, nrm, sizeof(double) * 3 * m, s));
    IA_TRY(to_device(drh, r_h, sizeof(double) * 3 * l, s));
    IA_TRY(to_device(db, b, sizeof(ia_z) * (size_t)l * F, s));
    if (lambda) IA_TRY(to_device(dlin, lambda, sizeof(double) * F, s));
    IA_TRY(W.src.alloc(sizeof(double) * 3 * m));
    hipLaunchKernelGGL(k_retreat, dim3((3 * m + 255) / 256), dim3(256), 0, s, 3 * m,
                       drs.as<double>(), dnrm.as<double>(), d, W.src.as<double>());
    const int chunk = std::min(F, desc->chunk > 0 ? desc->chunk : 64);
    DevBuf dlam, dsc, ddd, dgJ, dgW, dvel;
    IA_TRY(dlam.alloc(sizeof(double) * F));
    IA_TRY(dsc.alloc(sizeof(double) * 4 * F));
    IA_TRY(ddd.alloc(sizeof(double) * 2 * F));
    IA_TRY(dgJ.alloc(sizeof(double) * 3 * (size_t)l * F));
    IA_TRY(dgW.alloc(sizeof(double) * 3 * (size_t)l * F));
    ia_grad_out o{};
    if (out) o = *out;
    if (o.velocity) IA_TRY(dvel.alloc(sizeof(ia_z) * (size_t)m * F));
    for (int f0 = 0; f0 < F; f0 += chunk) {
        const int nf = std::min(chunk, F - f0);
        IA_TRY(gradient_batch(m, l, nf, *desc, dk.as<double>() + f0, drs.as<double>(),
                              dnrm.as<double>(), W.src.as<double>(), drh.as<double>(),
                              db.as<cx<double>>() + (size_t)f0 * l,
                              lambda ? dlin.as<double>() + f0 : nullptr, select,
                              dlam.as<double>() + f0, dsc.as<double>() + (size_t)4 * f0,
                              ddd.as<double>() + (size_t)2 * f0,
                              dgJ.as<double>() + (size_t)3 * l * f0,
                              dgW.as<double>() + (size_t)3 * l * f0,
                              o.velocity ? dvel.as<cx<double>>() + (size_t)m * f0 : nullptr, s));
    }
    IA_HIP(hipStreamSynchronize(s));
    std::vector<double> sc(4 * (size_t)F), dd(2 * (size_t)F), lam(F);
    IA_HIP(hipMemcpy(sc.data(), dsc.p, sizeof(double) * 4 * F, hipMemcpyDeviceToHost));
    IA_HIP(hipMemcpy(dd.data(), ddd.p, sizeof(double) * 2 * F, hipMemcpyDeviceToHost));
    IA_HIP(hipMemcpy(lam.data(), dlam.p, sizeof(double) * F, hipMemcpyDeviceToHost));
    for (int f = 0; f < F; ++f) {
        if (o.lambda) o.lambda[f] = lam[f];
        if (o.misfit) o.misfit[f] = sc[4 * f];
        if (o.power) o.power[f] = sc[4 * f + 1];
        if (o.dmisfit_dloglam) o.dmisfit_dloglam[f] = sc[4 * f + 2];
        if (o.dpower_dloglam) o.dpower_dloglam[f] = sc[4 * f + 3];
        if (o.dmisfit_dd) o.dmisfit_dd[f] = dd[2 * f];
        if (o.dpower_dd) o.dpower_dd[f] = dd[2 * f + 1];
    }
    if (o.dmisfit_dmic)
        IA_HIP(hipMemcpy(o.dmisfit_dmic, dgJ.p, sizeof(double) * 3 * (size_t)l * F, hipMemcpyDefault));
    if (o.dpower_dmic)
        IA_HIP(hipMemcpy(o.dpower_dmic, dgW.p, sizeof(double) * 3 * (size_t)l * F, hipMemcpyDefault));
    if (o.velocity)
        IA_HIP(hipMemcpy(o.velocity, dvel.p, sizeof(ia_z) * (size_t)m * F, hipMemcpyDefault));
    return IA_OK;
}
