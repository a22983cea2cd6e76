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
//   2. its n leading FP64 eigenpairs: for np >= 1024 patches by randomized
//      subspace iteration with Rayleigh-Ritz (real FP64 GEMMs with R, the
//      small Ritz problem by the batched SVD; each pair checked by its
//      residual, the full SVD when a frequency does not converge), else by
//      the full SVD -- R is symmetric positive semi-definite, so its
//      singular values are its eigenvalues and its left singular vectors its
//      modes: the batched bidiagonal SVD of svd.hip, with only the n leading
//      vectors pushed through the reflectors.  Each mode is rotated to real
//      with its largest entry positive (k_realign); modes below the
//      eigenvalue floor dropped, the kept ones orthonormalised (k_orth_modes);
//   3. H_r = G Q (G assembled by k_rad_G, one batched GEMM);
//   4. the regularised solve of H_r y = b (regsolve_batched: the SVD of H_r
//      in complex128 or complex64, parameter choice in FP64);
//   5. W = sum_i e_i |y_i|^2 and v = Q y.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ia_internal.hpp"

namespace ia {

template <typename T>
ia_status regsolve_batched(int rows, int cols, int batch, const cx<double> *A, size_t sA,
                           const cx<double> *b, size_t sb, const RegDev &prm, cx<double> *x,
                           size_t sx, double *lam, int *idx, int *kts, double *smax,
                           hipStream_t s, double *gap);
template <typename T>
ia_status regsolve_async(int rows, int cols, int batch, const cx<double> *A, size_t sA,
                         const cx<double> *b, size_t sb, const RegDev &prm, cx<double> *x,
                         size_t sx, double *lam, int *idx, int *kts, double *smax,
                         hipStream_t s);
template <typename T>
void regsolve_addresses(std::vector<uintptr_t> &v);
RegDev to_regdev(const ia_reg_params &p);
ia_status check_reg(const ia_reg_params &p, int pmax);

namespace {

constexpr int RT = 64;  // tile of the assembly kernels

// S_f = R_f / (w^2 rho A^2/(4 pi c)): sin(k r)/(k r), column-major, real
// (E = double, the subspace iteration) or complex with imaginary part zero
// (E = cx<double>, the complex SVD).  Tile of 64 x 64 per
// workgroup, the tile's 64 column points staged in LDS, each thread owns a
// row and walks 16 columns: stores of 64 consecutive entries per column.
template <typename E>
__global__ __launch_bounds__(256) void k_rad_R(int np, const double *__restrict__ kv,
                                              const double *__restrict__ p,
                                              E *__restrict__ S, size_t sS) {
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
    E *Sf = S + (size_t)f * sS;
    const int cb = (t >> 6) * 16;
    for (int cc = 0; cc < 16; ++cc) {
        const int jl = cb + cc, j = j0 + jl;
        if (j >= np) break;
        const double dx = x - sc[0][jl], dy = y - sc[1][jl], dz = z - sc[2][jl];
        const double kr = k * sqrt(dx * dx + dy * dy + dz * dz);
        const double v = kr == 0.0 ? 1.0 : sin(kr) / kr;
        if constexpr (sizeof(E) == sizeof(double))
            Sf[(size_t)j * np + i] = v;
        else
            Sf[(size_t)j * np + i] = E{v, 0.0};
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
// leave the mode-space fit basis-dependent).  gs: the kept modes are made
// orthonormal by classical Gram-Schmidt applied twice (Giraud, Langou &
// Rozloznik, Comput. Math. Appl. 50 (2005) 1069: orthogonal to working
// precision; they already are to about eps e_1 / gap); the dropped ones are
// zeroed, so their columns of H_r and their amplitudes vanish.  One
// 1024-thread workgroup per frequency: the column being orthonormalised
// sits in LDS, the 16 waves take the dot products with the earlier columns
// in turn, so a pass costs two barriers instead of one reduction per
// earlier column.  Dynamic LDS: (np + n) doubles.
__global__ __launch_bounds__(1024) void k_orth_modes(int np, int n, const double *__restrict__ sig,
                                                    double floor_rel, bool gs,
                                                    cx<double> *__restrict__ U,
                                                    double *__restrict__ modes) {
    extern __shared__ double xs[];  // np: the column; n: its dot products
    __shared__ double red[16];
    double *dd = xs + np;
    const int f = blockIdx.x, t = threadIdx.x, lane = t & 63, w = t >> 6;
    const double *sg = sig + (size_t)f * n;
    cx<double> *Uf = U + (size_t)f * np * n;
    int nk = 0;
    while (nk < n && sg[nk] >= floor_rel * sg[0] && sg[0] > 0) ++nk;
    for (int j = 0; gs && j < nk; ++j) {
        cx<double> *uj = Uf + (size_t)j * np;
        for (int r = t; r < np; r += 1024) xs[r] = uj[r].re;
        __syncthreads();
        for (int pass = 0; pass < 2 && j > 0; ++pass) {
            for (int i = w; i < j; i += 16) {
                const cx<double> *ui = Uf + (size_t)i * np;
                double d = 0;
                for (int r = lane; r < np; r += 64) d += ui[r].re * xs[r];
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) d += __shfl_xor(d, o, 64);
                if (lane == 0) dd[i] = d;
            }
            __syncthreads();
            for (int r = t; r < np; r += 1024) {
                double v = xs[r];
                for (int i = 0; i < j; ++i) v -= dd[i] * Uf[(size_t)i * np + r].re;
                xs[r] = v;
            }
            __syncthreads();
        }
        double nr = 0;
        for (int r = t; r < np; r += 1024) nr += xs[r] * xs[r];
        nr = block_sum1024(nr, red);
        const double inv = nr > 0 ? 1.0 / sqrt(nr) : 0.0;
        for (int r = t; r < np; r += 1024) uj[r] = cx<double>{xs[r] * inv, 0.0};
        __syncthreads();
    }
    __syncthreads();
    for (size_t q = t; q < (size_t)np * n; q += 1024) {
        const int j = (int)(q / np);
        if (j >= nk) Uf[q] = cx<double>{0, 0};
        if (modes) modes[(size_t)f * np * n + q] = j < nk ? Uf[q].re : 0.0;
    }
}

void launch_orth_modes(int np, int n, int nf, const double *sig, double floor_rel, bool gs,
                       cx<double> *U, double *modes, hipStream_t s) {
    const size_t lds = sizeof(double) * ((size_t)np + n);
    (void)hipFuncSetAttribute((const void *)k_orth_modes,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_orth_modes, dim3(nf), dim3(1024), lds, s, np, n, sig, floor_rel, gs, U,
                       modes);
}

// e = scale sigma (the n leading values, stride n), W = sum e_i |y_i|^2
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
        const double e = scale * sig[(size_t)f * n + i];
        if (eig) eig[(size_t)f * n + i] = e;
        acc += e * abs2(y[(size_t)f * n + i]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((t & 63) == 0) red[t >> 6] = acc;
    __syncthreads();
    if (t == 0 && W) W[f] = red[0] + red[1] + red[2] + red[3];
}

// ------------------------------------------------------- subspace iteration
// The n leading eigenpairs of R (real symmetric PSD) by randomized subspace
// iteration with Rayleigh-Ritz (Halko, Martinsson & Tropp, SIAM Review 53
// (2011) 217, Algorithms 4.4 and 5.3): Y = R Omega for an np x kk sketch
// Omega, then Q = orth(Y) (Householder QR, Q formed from its reflectors) and
// Y = R Q; B = Q^T Y = Q^T R Q (kk x kk) is eigen-decomposed by the batched
// SVD (B is PSD: its singular vectors are its eigenvectors) and the Ritz
// pairs (theta_j, Q u_j) are accepted when each of the n leading ones has
// ||R Q u_j - theta_j Q u_j|| <= tol theta_1 (R Q u_j = Y u_j).  Frequencies
// short of that take further rounds Q = orth(Y), Y = R Q with the check.
// On the 64 x 64 plate with n = 128 the first check passes up to ~1950 Hz
// and one more round serves the rest up to 2500 Hz.  Every product with R is
// a real FP64 GEMM on the matrix cores.

// sketch entries: uniform in [-1, 1) from a counter hash (splitmix64)
__global__ void k_sketch(size_t n, double *__restrict__ Om) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    unsigned long long z = (unsigned long long)q * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    Om[q] = (double)(z >> 11) * 0x1.0p-52 - 1.0;
}

__global__ void k_real_to_cx(size_t n, const double *__restrict__ x, cx<double> *__restrict__ y) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) y[q] = cx<double>{x[q], 0.0};
}

__global__ void k_cx_to_real(size_t n, const cx<double> *__restrict__ x, double *__restrict__ y) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) y[q] = x[q].re;
}

// y = re(y): a real matrix formed by a complex product keeps only its real part
__global__ void k_keep_real(size_t n, cx<double> *__restrict__ y) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) y[q].im = 0.0;
}

// Bc = (B + B^T) / 2 as complex (kk x kk per frequency, blockIdx.y)
__global__ void k_sym_cx(int kk, const double *__restrict__ B, cx<double> *__restrict__ Bc) {
    const size_t per = (size_t)kk * kk;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= per) return;
    const int r = (int)(q % kk), c = (int)(q / kk);
    const double *Bf = B + (size_t)blockIdx.y * per;
    Bc[(size_t)blockIdx.y * per + q] = cx<double>{0.5 * (Bf[q] + Bf[(size_t)r * kk + c]), 0.0};
}

// rr[f n + j] = ||Z_j - theta_j M_j|| / theta_0 (theta: stride kk), Z = Y U_B,
// M = Q U_B (np x n per frequency); one workgroup per (j, f)
__global__ __launch_bounds__(256) void k_ritz_res(int np, int n, int kk,
                                                 const double *__restrict__ sig,
                                                 const double *__restrict__ M,
                                                 const double *__restrict__ Z,
                                                 double *__restrict__ rr) {
    __shared__ double red[4];
    const int j = blockIdx.x, f = blockIdx.y, t = threadIdx.x;
    const double th = sig[(size_t)f * kk + j], th0 = sig[(size_t)f * kk];
    const size_t o = ((size_t)f * n + j) * np;
    double acc = 0;
    for (int r = t; r < np; r += 256) {
        const double d = Z[o + r] - th * M[o + r];
        acc += d * d;
    }
#pragma unroll
    for (int w = 32; w > 0; w >>= 1) acc += __shfl_xor(acc, w, 64);
    if ((t & 63) == 0) red[t >> 6] = acc;
    __syncthreads();
    if (t == 0) {
        const double nr = sqrt(red[0] + red[1] + red[2] + red[3]);
        rr[(size_t)f * n + j] = th0 > 0 ? nr / th0 : (nr == 0 ? 0.0 : INFINITY);
    }
}

// sg[f n + j] = sig[f stride + j], j < n
__global__ void k_lead(int n, int stride, int nf, const double *__restrict__ sig,
                       double *__restrict__ sg) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n * nf) return;
    sg[q] = sig[(size_t)(q / n) * stride + q % n];
}

inline unsigned blocks256(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// The decisions of one eigen-solve of a chunk: the window [a, b) of each
// Rayleigh-Ritz round of the subspace iteration (round 0 covers the chunk),
// the frequencies sent to the full SVD and how many of those one full-SVD
// launch takes.  They depend on the geometry and the wavenumbers only (not on
// the pressures), so a graph that replays them serves every later call with
// the same ones.
struct RadPlan {
    std::vector<std::pair<int, int>> rounds;
    std::vector<int> failed;
    int fc = 0;
};

// device workspace of the radiation sweep, kept per host thread; with
// IA_GRAPH the captured sweep, its key and plan, and the stream it replays on
struct RadCache {
    DevBuf S, U, sg, G, H, y, v, lam, idx, kts, smax, W, eig, modes, k, p, rh, b;
    DevBuf fidx, kg, U2, sg2;  // frequencies the full SVD serves
    // subspace iteration
    DevBuf Rr, Om, Yr, Qr, Yc, Qc, tq, qw, Br, Bc, UBc, UBr, Mr, Zr, rr, th;
    SvdWork<double> sw, sw2;
    hipStream_t gs = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<uintptr_t> gkey;
    void drop_graph() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
        gkey.clear();
    }
    void addresses(std::vector<uintptr_t> &out) const {
        for (const DevBuf *q : {&S, &U, &sg, &G, &H, &y, &v, &lam, &idx, &kts, &smax, &W, &eig,
                                &modes, &k, &p, &rh, &b, &Rr, &Om, &Yr, &Qr, &Yc, &Qc, &tq, &qw,
                                &Br, &Bc, &UBc, &UBr, &Mr, &Zr, &rr, &th, &fidx, &kg, &U2, &sg2})
            out.push_back((uintptr_t)q->p);
        sw.addresses(out);
        sw2.addresses(out);
    }
    void release() {
        drop_graph();
        for (DevBuf *q : {&S, &U, &sg, &G, &H, &y, &v, &lam, &idx, &kts, &smax, &W, &eig, &modes,
                          &k, &p, &rh, &b, &Rr, &Om, &Yr, &Qr, &Yc, &Qc, &tq, &qw, &Br, &Bc, &UBc,
                          &UBr, &Mr, &Zr, &rr, &th, &fidx, &kg, &U2, &sg2})
            q->reset();
        sw.release();
        sw2.release();
    }
};
static RadCache &rad_cache() {
    static thread_local RadCache *c = new RadCache();
    return *c;
}
void release_rad_cache() { rad_cache().release(); }

namespace {

// sketch width: n + max(64, n/2) columns, a multiple of 32
int rad_sketch_cols(int np, int n) {
    const int kk = (n + std::max(64, n / 2) + 31) / 32 * 32;
    return std::min(kk, np);
}

// Eigen-solver: by default the subspace iteration (O(np^2 kk) per
// frequency) serves np >= 1024 with kk <= np / 4 and falls back to the full
// SVD (O(np^3)) for a chunk it does not converge on; the full SVD serves the
// rest.  IA_RAD_MODES = "full" forces the full SVD, "subspace" the subspace
// iteration (kk < np) with no fallback (IA_ERR_NOCONV instead).
enum RadSolver { RAD_FULL, RAD_AUTO, RAD_SUBSPACE };
RadSolver rad_solver(int np, int kk) {
    const char *e = std::getenv("IA_RAD_MODES");
    if (e && std::strcmp(e, "full") == 0) return RAD_FULL;
    if (e && std::strcmp(e, "subspace") == 0) return kk < np ? RAD_SUBSPACE : RAD_FULL;
    return np >= 1024 && 4 * kk <= np ? RAD_AUTO : RAD_FULL;
}

// Ritz residual tolerance relative to theta_1 (IA_RAD_TOL overrides): a
// backward-stable full eigensolver's accuracy, eigenvalues to 1e-12 theta_1
double rad_tol() {
    const char *e = std::getenv("IA_RAD_TOL");
    return (e && *e) ? std::atof(e) : 2e-13;
}
constexpr int RAD_ROUNDS = 6;  // Rayleigh-Ritz checks at most

size_t free_bytes() {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
        (void)hipGetLastError();
        fr = (size_t)16 << 30;
    }
    return fr;
}

// device bytes per frequency of the full SVD: the complex copy of R, the
// stage-one reflector store and the bidiagonal vectors
size_t full_bytes(int np) { return (size_t)np * np * (sizeof(cx<double>) * 2 + 2 * sizeof(double)); }

// U (np x n per frequency, real values in complex form) and sg (n per
// frequency): the n leading eigenpairs of nf matrices R by the full SVD,
// the kept modes orthonormal, the dropped ones zero
ia_status full_modes(RadCache &W, int np, int n, const double *kf, int nf, double floor_rel,
                     cx<double> *U, double *sg, hipStream_t s) {
    const size_t nn = (size_t)np * np;
    IA_TRY(W.S.alloc(sizeof(cx<double>) * nn * nf));
    cx<double> *S = W.S.as<cx<double>>();
    hipLaunchKernelGGL(k_rad_R<cx<double>>, dim3((np + RT - 1) / RT, (np + RT - 1) / RT, nf),
                       dim3(256), 0, s, np, kf, W.p.as<double>(), S, nn);
    IA_HIP(hipGetLastError());
    SvdWork<double> &sw = W.sw;
    IA_TRY(sw.alloc(np, np, nf));
    IA_TRY(svd_factor<double>(sw, S, s));
    // the n leading modes: U = U_b [P_B; 0] (columns 0..n-1)
    hipLaunchKernelGGL(k_first_cols, dim3((np + 255) / 256, n, nf), dim3(256), 0, s, np, n,
                       sw.PBt.as<double>(), U);
    svd_apply_u<double>(sw, S, U, np, n, s);
    hipLaunchKernelGGL(k_lead, dim3(blocks256((size_t)n * nf)), dim3(256), 0, s, n, np, nf,
                       sw.sig.as<double>(), sg);
    hipLaunchKernelGGL(k_realign, dim3(n, nf), dim3(256), 0, s, np, n, U);
    launch_orth_modes(np, n, nf, sg, floor_rel, true, U, nullptr, s);
    IA_HIP(hipGetLastError());
    return IA_OK;
}

// device bytes per frequency of the subspace iteration (R, the sketch-sized
// blocks, B, the Ritz blocks)
size_t subspace_bytes(int np, int n, int kk) {
    const size_t pk = (size_t)np * kk;
    return (size_t)np * np * sizeof(double) + pk * (2 * sizeof(double) + 2 * sizeof(cx<double>)) +
           geqrf_work_elems(np, kk) * sizeof(cx<double>) +
           (size_t)kk * kk * (sizeof(double) + sizeof(cx<double>)) +
           (size_t)np * n * 2 * sizeof(double) + (size_t)kk * n * 24;
}

// The same by subspace iteration.  After each Rayleigh-Ritz check a
// frequency whose residual misses the tolerance is given up on when the
// Ritz values' ratio rho = theta_kk / theta_n (about the contraction of one
// round) cannot close the gap in the rounds left, rho^left res > tol -- an
// acoustically large plate at high frequency has a flat spectrum past mode
// n and converges slowly.  The rounds stop when every frequency has
// converged or been given up on; failed lists the latter (chunk-relative),
// whose U and sg are to be overwritten.
ia_status subspace_modes(RadCache &W, int np, int n, int kk, const double *kf, int nf,
                         double tol, double floor_rel, cx<double> *U, double *sg, hipStream_t s,
                         std::vector<int> &failed, RadPlan *rec, const RadPlan *play) {
    const size_t nn = (size_t)np * np, pk = (size_t)np * kk, pn = (size_t)np * n,
                 kn = (size_t)kk * n, k2 = (size_t)kk * kk;
    IA_TRY(W.Rr.alloc(sizeof(double) * nn * nf));
    IA_TRY(W.Om.alloc(sizeof(double) * pk));
    IA_TRY(W.Yr.alloc(sizeof(double) * pk * nf));
    IA_TRY(W.Qr.alloc(sizeof(double) * pk * nf));
    IA_TRY(W.Yc.alloc(sizeof(cx<double>) * pk * nf));
    IA_TRY(W.Qc.alloc(sizeof(cx<double>) * pk * nf));
    IA_TRY(W.tq.alloc(sizeof(cx<double>) * kk * nf));
    IA_TRY(W.qw.alloc(sizeof(cx<double>) * geqrf_work_elems(np, kk) * nf));
    IA_TRY(W.Br.alloc(sizeof(double) * k2 * nf));
    IA_TRY(W.Bc.alloc(sizeof(cx<double>) * k2 * nf));
    IA_TRY(W.UBc.alloc(sizeof(cx<double>) * kn * nf));
    IA_TRY(W.UBr.alloc(sizeof(double) * kn * nf));
    IA_TRY(W.Mr.alloc(sizeof(double) * pn * nf));
    IA_TRY(W.Zr.alloc(sizeof(double) * pn * nf));
    IA_TRY(W.rr.alloc(sizeof(double) * n * nf));
    IA_TRY(W.th.alloc(sizeof(double) * kk * nf));
    double *Rr = W.Rr.as<double>(), *Om = W.Om.as<double>(), *Yr = W.Yr.as<double>(),
           *Qr = W.Qr.as<double>(), *Br = W.Br.as<double>(), *UBr = W.UBr.as<double>(),
           *Mr = W.Mr.as<double>(), *Zr = W.Zr.as<double>();
    cx<double> *Yc = W.Yc.as<cx<double>>(), *Qc = W.Qc.as<cx<double>>(),
               *tq = W.tq.as<cx<double>>(), *qw = W.qw.as<cx<double>>(),
               *Bc = W.Bc.as<cx<double>>(), *UBc = W.UBc.as<cx<double>>();
    hipLaunchKernelGGL(k_rad_R<double>, dim3((np + RT - 1) / RT, (np + RT - 1) / RT, nf),
                       dim3(256), 0, s, np, kf, W.p.as<double>(), Rr, nn);
    hipLaunchKernelGGL(k_sketch, dim3(blocks256(pk)), dim3(256), 0, s, pk, Om);
    // Y = R Omega (one sketch for every frequency)
    launch_dgemm('N', np, kk, np, Rr, np, nn, Om, np, 0, Yr, np, pk, nf, s);
    IA_HIP(hipGetLastError());
    double *TH = W.th.as<double>();
    SvdWork<double> &sw = W.sw2;
    // Q = orth(Y), then Y = R Q, for the nw frequencies from a
    auto step = [&](int a, int nw) -> ia_status {
        const size_t oP = (size_t)a * pk;
        hipLaunchKernelGGL(k_real_to_cx, dim3(blocks256(pk * nw)), dim3(256), 0, s, pk * nw,
                           Yr + oP, Yc + oP);
        IA_TRY(launch_geqrf<double>(np, kk, Yc + oP, np, pk, tq + (size_t)a * kk, qw, nw, s));
        IA_TRY(launch_orgqr<double>(np, kk, Yc + oP, np, pk, tq + (size_t)a * kk, Qc + oP, np, pk,
                                    qw, nw, s));
        hipLaunchKernelGGL(k_cx_to_real, dim3(blocks256(pk * nw)), dim3(256), 0, s, pk * nw,
                           Qc + oP, Qr + oP);
        launch_dgemm('N', np, kk, np, Rr + (size_t)a * nn, np, nn, Qr + oP, np, pk, Yr + oP, np,
                     pk, nw, s);
        IA_HIP(hipGetLastError());
        return IA_OK;
    };
    // Rayleigh-Ritz on span(Q) for the nw frequencies from a: B = Q^T Y =
    // U_B Theta U_B^T; Ritz vectors M = Q U_B, R M = Y U_B, their residuals
    auto ritz = [&](int a, int nw) -> ia_status {
        const size_t oP = (size_t)a * pk, oB = (size_t)a * k2, oU = (size_t)a * kn,
                     oM = (size_t)a * pn;
        launch_dgemm('T', kk, kk, np, Qr + oP, np, pk, Yr + oP, np, pk, Br + oB, kk, k2, nw, s);
        hipLaunchKernelGGL(k_sym_cx, dim3(blocks256(k2), nw), dim3(256), 0, s, kk, Br + oB,
                           Bc + oB);
        IA_TRY(sw.alloc(kk, kk, nw));
        IA_TRY(svd_factor<double>(sw, Bc + oB, s));
        hipLaunchKernelGGL(k_first_cols, dim3((kk + 255) / 256, n, nw), dim3(256), 0, s, kk, n,
                           sw.PBt.as<double>(), UBc + oU);
        svd_apply_u<double>(sw, Bc + oB, UBc + oU, kk, n, s);
        hipLaunchKernelGGL(k_realign, dim3(n, nw), dim3(256), 0, s, kk, n, UBc + oU);
        hipLaunchKernelGGL(k_cx_to_real, dim3(blocks256(kn * nw)), dim3(256), 0, s, kn * nw,
                           UBc + oU, UBr + oU);
        launch_dgemm('N', np, n, kk, Qr + oP, np, pk, UBr + oU, kk, kn, Mr + oM, np, pn, nw, s);
        launch_dgemm('N', np, n, kk, Yr + oP, np, pk, UBr + oU, kk, kn, Zr + oM, np, pn, nw, s);
        hipLaunchKernelGGL(k_lead, dim3(blocks256((size_t)kk * nw)), dim3(256), 0, s,
                           kk, kk, nw, sw.sig.as<double>(), TH + (size_t)a * kk);
        hipLaunchKernelGGL(k_ritz_res, dim3(n, nw), dim3(256), 0, s, np, n, kk,
                           TH + (size_t)a * kk, Mr + oM, Zr + oM, W.rr.as<double>() + (size_t)a * n);
        IA_HIP(hipGetLastError());
        return IA_OK;
    };
    // round 0 checks the basis of R Omega; later rounds iterate the window
    // [a, b) of the frequencies still short of the tolerance that can reach
    // it (in a sweep sorted by frequency, the highest ones)
    IA_TRY(step(0, nf));
    if (play) {
        // the recorded rounds, no host reads (graph capture)
        for (size_t round = 0; round < play->rounds.size(); ++round) {
            const int a = play->rounds[round].first, b = play->rounds[round].second;
            if (round > 0) IA_TRY(step(a, b - a));
            IA_TRY(ritz(a, b - a));
        }
        failed = play->failed;
    }
    std::vector<double> rr((size_t)n * nf), th((size_t)kk * nf), res(nf, INFINITY);
    int a = 0, b = nf;
    for (int round = 0; !play && round < RAD_ROUNDS && a < b; ++round) {
        if (rec) rec->rounds.emplace_back(a, b);
        if (round > 0) IA_TRY(step(a, b - a));
        IA_TRY(ritz(a, b - a));
        IA_HIP(hipStreamSynchronize(s));
        IA_HIP(hipMemcpy(rr.data() + (size_t)a * n, W.rr.as<double>() + (size_t)a * n,
                         sizeof(double) * n * (b - a), hipMemcpyDeviceToHost));
        IA_HIP(hipMemcpy(th.data() + (size_t)a * kk, TH + (size_t)a * kk,
                         sizeof(double) * kk * (b - a), hipMemcpyDeviceToHost));
        const int left = RAD_ROUNDS - 1 - round;
        int na = b, nb = a;
        for (int f = a; f < b; ++f) {
            double m = 0;
            for (int j = 0; j < n; ++j) {
                const double v = rr[(size_t)f * n + j];
                m = std::isfinite(v) ? std::max(m, v) : INFINITY;
                if (!std::isfinite(m)) break;
            }
            res[f] = m;
            if (m <= tol) continue;
            const double tn = th[(size_t)f * kk + n - 1], tk = th[(size_t)f * kk + kk - 1];
            const double rho = tn > 0 ? std::min(1.0, tk / tn) : 1.0;
            if (left > 0 && std::isfinite(m) && std::pow(rho, left) * m <= tol) {
                na = std::min(na, f);
                nb = std::max(nb, f + 1);
            }
        }
        a = na;
        b = nb;
    }
    if (!play) {
        failed.clear();
        for (int f = 0; f < nf; ++f)
            if (!(res[f] <= tol)) failed.push_back(f);
        if (rec) rec->failed = failed;
    }
    // the modes: U_B's kept columns orthonormalised (then so are those of
    // M = Q U_B, to Q's orthogonality), the dropped ones zero
    hipLaunchKernelGGL(k_lead, dim3(blocks256((size_t)n * nf)), dim3(256), 0, s, n, kk, nf, TH,
                       sg);
    launch_orth_modes(kk, n, nf, sg, floor_rel, true, UBc, nullptr, s);
    hipLaunchKernelGGL(k_cx_to_real, dim3(blocks256(kn * nf)), dim3(256), 0, s, kn * nf, UBc, UBr);
    launch_dgemm('N', np, n, kk, Qr, np, pk, UBr, kk, kn, Mr, np, pn, nf, s);
    hipLaunchKernelGGL(k_real_to_cx, dim3(blocks256(pn * nf)), dim3(256), 0, s, pn * nf, Mr, U);
    IA_HIP(hipGetLastError());
    return IA_OK;
}

// k of the listed frequencies, then their modes and values moved into place
__global__ void k_gather_k(int m, const int *__restrict__ idx, const double *__restrict__ kf,
                           double *__restrict__ kg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) kg[i] = kf[idx[i]];
}
__global__ void k_scatter_modes(int np, int n, const int *__restrict__ idx,
                                const cx<double> *__restrict__ U2, const double *__restrict__ sg2,
                                cx<double> *__restrict__ U, double *__restrict__ sg) {
    const size_t per = (size_t)np * n;
    const int i = blockIdx.y, f = idx[i];
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < per;
         q += (size_t)gridDim.x * blockDim.x)
        U[(size_t)f * per + q] = U2[(size_t)i * per + q];
    if (blockIdx.x == 0)
        for (int j = threadIdx.x; j < n; j += blockDim.x) sg[(size_t)f * n + j] = sg2[(size_t)i * n + j];
}

// One chunk of the sweep from the eigen-solve to W and v, enqueued on s.
// rec: record the eigen-solver's decisions; play: replay recorded ones (no
// host reads, uploads or synchronisation: graph capture).
ia_status rad_chunk(RadCache &W, const ia_radiation_desc &desc, int f0, int nf, RadSolver solver,
                    int kk, double tol, double eig_floor, const RegDev &prm, bool want_modes,
                    bool want_vel, hipStream_t s, RadPlan *rec, const RadPlan *play) {
    const int np = desc.np, l = desc.l, n = desc.n_modes;
    const double c = air().c, rho = air().rho, A = desc.area;
    // w^2 rho A^2/(4 pi c) = k^2 (c rho A^2 / (4 pi))
    const double scale_c = c * rho * A * A / (4 * M_PI);
    const bool sub = solver != RAD_FULL;
    const double *kf = W.k.as<double>() + f0;
    IA_TRY(W.U.alloc(sizeof(cx<double>) * (size_t)np * n * nf));
    IA_TRY(W.sg.alloc(sizeof(double) * (size_t)n * nf));
    cx<double> *U = W.U.as<cx<double>>();
    double *sg = W.sg.as<double>();
    std::vector<int> failed;
    if (sub) {
        IA_TRY(subspace_modes(W, np, n, kk, kf, nf, tol, eig_floor, U, sg, s, failed, rec, play));
        if (!failed.empty() && solver == RAD_SUBSPACE)
            return fail(IA_ERR_NOCONV,
                        "ia_radiation_sweep: subspace iteration missed the Ritz tolerance %g "
                        "at %zu of frequencies %d..%d (first %d)",
                        tol, failed.size(), f0, f0 + nf - 1, f0 + failed[0]);
    } else {
        for (int f = 0; f < nf; ++f) failed.push_back(f);
        if (rec) rec->failed = failed;
    }
    if (!failed.empty()) {
        // the full SVD for the listed frequencies, as many at a time as
        // memory holds, into U2 / sg2 and moved into place
        const int m = (int)failed.size();
        const int fc = play ? play->fc
                            : desc.chunk > 0
                                  ? m
                                  : (int)std::min<size_t>(
                                        std::min<size_t>(m, 128),
                                        std::max<size_t>(1, (free_bytes() / 5 * 2) / full_bytes(np)));
        if (rec) rec->fc = fc;
        IA_TRY(W.fidx.alloc(sizeof(int) * m));
        IA_TRY(W.kg.alloc(sizeof(double) * m));
        IA_TRY(W.U2.alloc(sizeof(cx<double>) * (size_t)np * n * fc));
        IA_TRY(W.sg2.alloc(sizeof(double) * (size_t)n * fc));
        if (!play)
            IA_HIP(hipMemcpy(W.fidx.p, failed.data(), sizeof(int) * m, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_gather_k, dim3(blocks256(m)), dim3(256), 0, s, m, W.fidx.as<int>(),
                           kf, W.kg.as<double>());
        for (int g0 = 0; g0 < m; g0 += fc) {
            const int ng = std::min(fc, m - g0);
            IA_TRY(full_modes(W, np, n, W.kg.as<double>() + g0, ng, eig_floor,
                              W.U2.as<cx<double>>(), W.sg2.as<double>(), s));
            hipLaunchKernelGGL(k_scatter_modes, dim3(256, ng), dim3(256), 0, s, np, n,
                               W.fidx.as<int>() + g0, W.U2.as<cx<double>>(), W.sg2.as<double>(),
                               U, sg);
            IA_HIP(hipGetLastError());
        }
    }
    double *md = nullptr;
    if (want_modes) {
        IA_TRY(W.modes.alloc(sizeof(double) * (size_t)np * n * nf));
        md = W.modes.as<double>();
    }
    // each mode real with its largest entry positive; kept ones already
    // orthonormal, dropped ones zero
    hipLaunchKernelGGL(k_realign, dim3(n, nf), dim3(256), 0, s, np, n, U);
    launch_orth_modes(np, n, nf, sg, eig_floor, false, U, md, s);
    // H_r = G Q
    IA_TRY(W.G.alloc(sizeof(cx<double>) * (size_t)l * np * nf));
    IA_TRY(W.H.alloc(sizeof(cx<double>) * (size_t)l * n * nf));
    cx<double> *G = W.G.as<cx<double>>(), *H = W.H.as<cx<double>>();
    hipLaunchKernelGGL(k_rad_G, dim3((l + RT - 1) / RT, (np + RT - 1) / RT, nf), dim3(256), 0, s,
                       l, np, kf, rho * c * A, W.rh.as<double>(), W.p.as<double>(), G,
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
    cx<double> *y = W.y.as<cx<double>>();
    if (desc.precision == IA_PREC_MIXED)
        IA_TRY(play ? regsolve_async<float>(l, n, nf, H, (size_t)l * n, bf, l, prm, y, n,
                                            W.lam.as<double>(), W.idx.as<int>(), W.kts.as<int>(),
                                            W.smax.as<double>(), s)
                    : regsolve_batched<float>(l, n, nf, H, (size_t)l * n, bf, l, prm, y, n,
                                              W.lam.as<double>(), W.idx.as<int>(),
                                              W.kts.as<int>(), W.smax.as<double>(), s, nullptr));
    else
        IA_TRY(play ? regsolve_async<double>(l, n, nf, H, (size_t)l * n, bf, l, prm, y, n,
                                             W.lam.as<double>(), W.idx.as<int>(), W.kts.as<int>(),
                                             W.smax.as<double>(), s)
                    : regsolve_batched<double>(l, n, nf, H, (size_t)l * n, bf, l, prm, y, n,
                                               W.lam.as<double>(), W.idx.as<int>(),
                                               W.kts.as<int>(), W.smax.as<double>(), s, nullptr));
    // W = sum e |y|^2, v = Q y
    IA_TRY(W.W.alloc(sizeof(double) * nf));
    IA_TRY(W.eig.alloc(sizeof(double) * (size_t)n * nf));
    hipLaunchKernelGGL(k_modal_power, dim3(nf), dim3(256), 0, s, np, n, kf, scale_c, sg, y,
                       W.eig.as<double>(), W.W.as<double>());
    if (want_vel) {
        IA_TRY(W.v.alloc(sizeof(cx<double>) * (size_t)np * nf));
        launch_gemm<double>('N', 'N', np, 1, n, one, U, np, (size_t)np * n, y, n, n, zero,
                            W.v.as<cx<double>>(), np, np, nf, s);
    }
    IA_HIP(hipGetLastError());
    return IA_OK;
}

}  // namespace

// Single-frequency helpers of ia_invr and ia_radiation_modes (ibem_api.hip):
// the conversions and the mode alignment stay on the device.
void launch_real_to_cx(size_t n, const double *x, cx<double> *y, hipStream_t s) {
    hipLaunchKernelGGL(k_real_to_cx, dim3(blocks256(n)), dim3(256), 0, s, n, x, y);
}
void launch_cx_to_real(size_t n, const cx<double> *x, double *y, hipStream_t s) {
    hipLaunchKernelGGL(k_cx_to_real, dim3(blocks256(n)), dim3(256), 0, s, n, x, y);
}
void launch_keep_real(size_t n, cx<double> *y, hipStream_t s) {
    hipLaunchKernelGGL(k_keep_real, dim3(blocks256(n)), dim3(256), 0, s, n, y);
}
void launch_realign(int np, int n, int nf, cx<double> *U, hipStream_t s) {
    hipLaunchKernelGGL(k_realign, dim3(n, nf), dim3(256), 0, s, np, n, U);
}
void launch_rad_R(int np, const double *kv, const double *p, cx<double> *S, int nf,
                  hipStream_t s) {
    hipLaunchKernelGGL(k_rad_R<cx<double>>, dim3((np + RT - 1) / RT, (np + RT - 1) / RT, nf),
                       dim3(256), 0, s, np, kv, p, S, (size_t)np * np);
}

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
    const RegDev prm = to_regdev(desc->reg);
    const double eig_floor = desc->eig_floor == 0.0 ? 1e-8 : std::max(desc->eig_floor, 0.0);
    IA_TRY(to_device(W.k, k, sizeof(double) * F, s));
    IA_TRY(to_device(W.p, r_p, sizeof(double) * 3 * np, s));
    IA_TRY(to_device(W.rh, r_h, sizeof(double) * 3 * l, s));
    IA_TRY(to_device(W.b, b, sizeof(ia_z) * (size_t)l * F, s));
    const int kk = rad_sketch_cols(np, n);
    const RadSolver solver = rad_solver(np, kk);
    const bool sub = solver != RAD_FULL;
    const double tol = rad_tol();
    // frequencies per chunk: as many as two fifths of the free memory hold
    // (the full SVD: ~1 GB per frequency at np = 4096; its chase runs one
    // workgroup per frequency, so a larger chunk keeps more compute units
    // busy), at most 128 (full SVD) or 256 (subspace iteration)
    const size_t fixed = (size_t)l * np * sizeof(cx<double>);  // G
    const size_t per = fixed + (sub ? subspace_bytes(np, n, kk) : full_bytes(np));
    int chunk = desc->chunk > 0 ? desc->chunk
                                : (int)std::max<size_t>(1, (free_bytes() / 5 * 2) / per);
    chunk = std::min(std::min(chunk, F), sub ? 256 : 128);
    ia_radiation_out o{};
    if (out) o = *out;
    const bool want_modes = o.modes != nullptr, want_vel = o.velocity != nullptr;
    auto chunk_work = [&](int f0, int nf, hipStream_t st, RadPlan *rec,
                          const RadPlan *play) -> ia_status {
        return rad_chunk(W, *desc, f0, nf, solver, kk, tol, eig_floor, prm, want_modes, want_vel,
                         st, rec, play);
    };
    auto copy_out = [&](int f0, int nf) -> ia_status {
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
        return IA_OK;
    };
    if (graph_mode() && chunk >= F) {
        // IA_GRAPH=1 (one chunk): the sweep is captured into a graph that
        // replays the eigen-solver's decisions of an eager run; the key holds
        // the sizes, parameters, a hash of the wavenumbers and patch
        // positions (the decisions depend on them) and every buffer address
        auto make_key = [&](std::vector<uintptr_t> &key) -> ia_status {
            std::vector<double> hv((size_t)F + 3 * (size_t)np);
            IA_HIP(hipMemcpy(hv.data(), W.k.p, sizeof(double) * F, hipMemcpyDeviceToHost));
            IA_HIP(hipMemcpy(hv.data() + F, W.p.p, sizeof(double) * 3 * np, hipMemcpyDeviceToHost));
            uint64_t h = 1469598103934665603ull;
            for (double v : hv) h = (h ^ (uint64_t)dbits(v)) * 1099511628211ull;
            key = {(uintptr_t)F, (uintptr_t)np, (uintptr_t)l, (uintptr_t)n, (uintptr_t)kk,
                   (uintptr_t)desc->precision, (uintptr_t)prm.filter, (uintptr_t)prm.select,
                   (uintptr_t)prm.n_lambda, (uintptr_t)prm.k_fixed, dbits(prm.decades),
                   dbits(prm.lambda_fixed), dbits(A), dbits(c), dbits(rho), dbits(eig_floor),
                   dbits(tol), (uintptr_t)solver, (uintptr_t)want_modes, (uintptr_t)want_vel,
                   (uintptr_t)h};
            W.addresses(key);
            if (desc->precision == IA_PREC_MIXED)
                regsolve_addresses<float>(key);
            else
                regsolve_addresses<double>(key);
            return IA_OK;
        };
        if (!W.gs) IA_HIP(hipStreamCreateWithFlags(&W.gs, hipStreamNonBlocking));
        static thread_local RadPlan plan;
        std::vector<uintptr_t> key;
        IA_TRY(make_key(key));
        if (W.gexec && key == W.gkey) {
            // the plan's full-SVD list is the one piece of state the graph
            // reads that another call may have overwritten
            if (!plan.failed.empty())
                IA_HIP(hipMemcpyAsync(W.fidx.p, plan.failed.data(), sizeof(int) * plan.failed.size(),
                                      hipMemcpyHostToDevice, s));
            IA_HIP(hipStreamSynchronize(s));  // inputs copied on the legacy stream
            IA_HIP(hipGraphLaunch(W.gexec, W.gs));
            IA_HIP(hipStreamSynchronize(W.gs));
            return copy_out(0, F);
        }
        W.drop_graph();
        RadPlan rec;
        IA_TRY(chunk_work(0, F, s, &rec, nullptr));
        IA_HIP(hipStreamSynchronize(s));
        IA_TRY(copy_out(0, F));
        IA_TRY(make_key(key));  // addresses after this run's allocations
        IA_HIP(hipStreamBeginCapture(W.gs, hipStreamCaptureModeThreadLocal));
        const ia_status st = chunk_work(0, F, W.gs, nullptr, &rec);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(W.gs, &g);
        if (st == IA_OK && ce == hipSuccess && g &&
            hipGraphInstantiate(&W.gexec, g, nullptr, nullptr, 0) != hipSuccess)
            W.gexec = nullptr;
        if (g) (void)hipGraphDestroy(g);
        IA_TRY(st);
        if (!W.gexec) {
            (void)hipGetLastError();
            return fail(IA_ERR_DEVICE, "ia_radiation_sweep: capture into a graph failed");
        }
        W.gkey = key;
        plan = rec;
        return IA_OK;
    }
    for (int f0 = 0; f0 < F; f0 += chunk) {
        const int nf = std::min(chunk, F - f0);
        IA_TRY(chunk_work(f0, nf, s, nullptr, nullptr));
        IA_HIP(hipStreamSynchronize(s));
        IA_TRY(copy_out(f0, nf));
    }
    return IA_OK;
}
