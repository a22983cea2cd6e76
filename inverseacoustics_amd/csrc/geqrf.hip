// Batched blocked Householder QR of a tall matrix, Y = Q R (R >= C), used to
// shrink the bidiagonalisation of tall transfer matrices: the SVD of the
// C x C factor R shares singular values and right vectors with Y, and the
// left vectors are Q times those of R.  The bidiagonalisation streams its
// trailing block twice per column, so reducing R x C to C x C first cuts
// that traffic by ~(R/C)(3 - C/R)/2, while the QR itself is GEMM work on
// the matrix cores.
//
// Per panel of NBQ = 64 columns (LAPACK geqrf/larft scheme):
//   k_qr_panel   one workgroup per matrix: unblocked Householder QR of the
//                panel (H_i = I - tau_i v_i v_i^H, LAPACK zlarfg), explicit
//                copy of the reflectors V (unit diagonal, zeros above);
//   G = V^H V    GEMM;  k_qr_tmat builds T with H_0 ... H_{nb-1} = I - V T V^H;
//   trailing     A2 <- (I - V T^H V^H) A2 = A2 - V (T^H (V^H A2)), three GEMMs.
// The reflectors stay below the diagonal of Y with tau per column, the form
// the vector transforms (k_apply_left) read.
#include "householder.hpp"

namespace ia {

namespace {

using hh::BT;
using hh::NW;
using hh::Refl;
using hh::block_sum;
using hh::larfg;
using hh::wave_sum;

constexpr int NB = 32;   // band-reduction panels (the band width)
constexpr int NBQ = 64;  // QR panels

template <typename T>
__global__ __launch_bounds__(BT) void k_qr_panel(int R, int C, int j0, int nbp,
                                                cx<T> *__restrict__ Yall, int ldy, size_t sY,
                                                cx<T> *__restrict__ tauall,
                                                cx<T> *__restrict__ Vwall, int vcols) {
    __shared__ double red[NW];
    const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    cx<T> *A = Yall + (size_t)b * sY;
    cx<T> *tau = tauall + (size_t)b * C;
    cx<T> *Vw = Vwall + (size_t)b * R * vcols;
    for (int i = 0; i < nbp; ++i) {
        const int j = j0 + i;
        cx<T> *Aj = A + (size_t)j * ldy;
        double p = 0;
        for (int r = j + 1 + t; r < R; r += BT) p += (double)abs2(Aj[r]);
        const double xn2 = block_sum(p, red);
        const Refl rf = larfg(cvt<double>(Aj[j]), xn2);
        const cx<T> sc = cvt<T>(rf.scale);
        __syncthreads();  // A[j,j] read by all before it is overwritten
        for (int r = j + 1 + t; r < R; r += BT) Aj[r] = Aj[r] * sc;
        if (t == 0) {
            Aj[j] = cx<T>{(T)rf.beta, 0};
            tau[j] = cvt<T>(rf.tau);
        }
        __syncthreads();
        // remaining panel columns: A[:,c] -= conj(tau) u (u^H A[:,c]), one wave per column
        const cx<double> ctau = conj(rf.tau);
        for (int c = j + 1 + wv; c < j0 + nbp; c += NW) {
            cx<T> *Ac = A + (size_t)c * ldy;
            double ar = 0, ai = 0;
            for (int r = j + lane; r < R; r += 64) {
                const cx<T> uu = (r == j) ? cx<T>{1, 0} : Aj[r];
                const cx<T> y = Ac[r];
                ar += (double)uu.re * y.re + (double)uu.im * y.im;
                ai += (double)uu.re * y.im - (double)uu.im * y.re;
            }
            const cx<double> f = ctau * cx<double>{wave_sum(ar), wave_sum(ai)};
            for (int r = j + lane; r < R; r += 64) {
                const cx<double> uu = (r == j) ? cx<double>{1, 0} : cvt<double>(Aj[r]);
                Ac[r] = cvt<T>(cvt<double>(Ac[r]) - uu * f);
            }
        }
        __syncthreads();
    }
    // explicit reflectors of the panel: Vw[:, i] (rows j0..R-1)
    for (int i = 0; i < nbp; ++i) {
        const int j = j0 + i;
        const cx<T> *Aj = A + (size_t)j * ldy;
        for (int r = j0 + t; r < R; r += BT) {
            cx<T> v = (r < j) ? cx<T>{0, 0} : (r == j) ? cx<T>{1, 0} : Aj[r];
            Vw[(size_t)i * R + r] = v;
        }
    }
}

// Blocked form of k_qr_panel: the panel is factored SB columns at a time
// with those columns held in LDS (rows j0..R-1), and each sub-panel's
// reflectors are applied to the panel's remaining columns as one block,
// A <- (I - V T V^H)^H A: two passes over each column (V^H a, then
// a -= V T^H (V^H a)) instead of two passes per reflector.  Same outputs
// (reflectors below the diagonal, tau, explicit V in Vw) as k_qr_panel.
//
// Inside a sub-panel one pass over the rows per column does the previous
// reflector's work (scale its column, update the later columns), gathers
// that reflector's Gram column v_j^H v_{i-1} against the earlier ones, and
// the sums the next reflector needs (its column's norm below the diagonal
// and the raw products x^H a of the later columns); one block reduction
// then lets wave 0 form the reflector (LAPACK zlarfg), the update factors
// f = conj(tau) u^H a with u^H a = a[rc] + conj(scale) x^H a, and the next
// column of T (LAPACK larft, forward).  Two barriers per column.
template <typename T, int SB, int RPL, int TB>
__global__ __launch_bounds__(BT) void k_qr_panel_lds(int R, int C, int j0, int nbp,
                                                    cx<T> *__restrict__ Yall, int ldy, size_t sY,
                                                    cx<T> *__restrict__ tauall,
                                                    cx<T> *__restrict__ Vwall, int vcols) {
    // norm, then per column slot q (re, im): the Gram entry v_q^H v_{i-1}
    // for q < i - 1, x^H a_q for q > i
    constexpr int NRED = 1 + 2 * SB;
    static_assert(NRED <= 64, "one lane per reduced value");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *Vs = reinterpret_cast<cx<T> *>(smem_raw);  // SB columns of M rows
    __shared__ double red[NW][NRED];
    __shared__ cx<T> fs[SB];       // update factors of the current reflector
    __shared__ cx<double> prm[2];  // its scale and (beta, 0)
    __shared__ cx<double> taus[SB];
    __shared__ cx<double> Gc[SB];  // Gram column of the reflector just finished
    __shared__ cx<double> Ts[SB][SB];
    __shared__ cx<double> Zs[NW][SB];
    const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    cx<T> *A = Yall + (size_t)b * sY;
    cx<T> *tau = tauall + (size_t)b * C;
    cx<T> *Vw = Vwall + (size_t)b * R * vcols;
    const int M = R - j0;
    for (int cs = j0; cs < j0 + nbp; cs += SB) {
        const int w = min(SB, j0 + nbp - cs);
        const int r0 = cs - j0;  // local row of the sub-panel's first diagonal
        for (int i = 0; i < w; ++i)
            for (int r = t; r < M; r += BT) Vs[(size_t)i * M + r] = A[(size_t)(cs + i) * ldy + j0 + r];
        __syncthreads();
        // ---- unblocked QR of the sub-panel in LDS; pass i applies
        // reflector i - 1 and gathers the sums of column i
        for (int i = 0; i <= w; ++i) {
            const int rp = r0 + i - 1;  // diagonal row of reflector i - 1
            const int rc = rp + 1;      // diagonal row of column i
            cx<T> scT{0, 0};
            T beta = 0;
            if (i > 0) {
                scT = cvt<T>(prm[0]);
                beta = (T)prm[1].re;
            }
            T p = 0;
            cx<T> sq[SB];
#pragma unroll
            for (int q = 0; q < SB; ++q) sq[q] = cx<T>{0, 0};
            for (int r = t; r < M; r += BT) {
                if (i > 0 && r >= rp) {
                    cx<T> u{1, 0};
                    if (r == rp) {
                        Vs[(size_t)(i - 1) * M + r] = cx<T>{beta, 0};
                    } else {
                        u = Vs[(size_t)(i - 1) * M + r] * scT;
                        Vs[(size_t)(i - 1) * M + r] = u;
                    }
#pragma unroll
                    for (int q = 0; q < SB; ++q) {
                        if (q >= i && q < w) {  // later columns: a -= u f
                            cx<T> &a = Vs[(size_t)q * M + r];
                            a = a - u * fs[q];
                        }
                        if (q < i - 1)  // Gram column: conj(u_q) u (u_q final below its unit)
                            sq[q] = sq[q] + cmulc(Vs[(size_t)q * M + r], u);
                    }
                }
                if (i < w && r > rc) {
                    const cx<T> x = Vs[(size_t)i * M + r];
                    p += abs2(x);
#pragma unroll
                    for (int q = 0; q < SB; ++q)
                        if (q > i && q < w) sq[q] = sq[q] + cmulc(x, Vs[(size_t)q * M + r]);
                }
            }
            const double pd = (double)hh::wave_sum_dpp(p);
#pragma unroll
            for (int q = 0; q < SB; ++q) {
                double sr = 0, si = 0;
                if (q < i - 1 || (q > i && q < w)) {
                    sr = (double)hh::wave_sum_dpp(sq[q].re);
                    si = (double)hh::wave_sum_dpp(sq[q].im);
                }
                if (lane == 0) {
                    red[wv][1 + 2 * q] = sr;
                    red[wv][2 + 2 * q] = si;
                }
            }
            if (lane == 0) red[wv][0] = pd;
            __syncthreads();
            if (wv == 0) {
                double v = 0;
                if (lane < NRED)
                    for (int q = 0; q < NW; ++q) v += red[q][lane];
                const int ql = min(lane, SB - 1);
                const double sre = __shfl(v, 1 + 2 * ql, 64), sim = __shfl(v, 2 + 2 * ql, 64);
                // T column i - 1: T[j][i-1] = -tau_{i-1} sum_{l=j}^{i-2} T[j][l] G[l][i-1]
                if (i > 0) {
                    if (lane < i - 1) Gc[lane] = cx<double>{sre, sim};
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    const cx<double> tp = taus[i - 1];
                    if (lane < i - 1) {
                        cx<double> acc{0, 0};
                        for (int l = lane; l < i - 1; ++l) acc = acc + Ts[lane][l] * Gc[l];
                        Ts[lane][i - 1] = cx<double>{0, 0} - tp * acc;
                    }
                    if (lane == i - 1) Ts[lane][lane] = tp;
                }
                if (i < w) {
                    const double xn2 = __shfl(v, 0, 64);
                    const Refl rf = larfg(cvt<double>(Vs[(size_t)i * M + rc]), xn2);
                    if (lane > i && lane < w) {
                        const cx<double> d = cvt<double>(Vs[(size_t)lane * M + rc]) +
                                             conj(rf.scale) * cx<double>{sre, sim};
                        fs[lane] = cvt<T>(conj(rf.tau) * d);
                    }
                    if (lane == 0) {
                        prm[0] = rf.scale;
                        prm[1] = cx<double>{rf.beta, 0};
                        taus[i] = rf.tau;
                    }
                }
            }
            __syncthreads();
        }
        // ---- write back: factored columns and explicit reflectors
        // (tau leaves LDS here: a global store inside the column loop would
        // make every barrier there wait for it)
        if (t < w) tau[cs + t] = cvt<T>(taus[t]);
        for (int i = 0; i < w; ++i) {
            const int rc = cs + i - j0;
            for (int r = t; r < M; r += BT) {
                const cx<T> v = Vs[(size_t)i * M + r];
                A[(size_t)(cs + i) * ldy + j0 + r] = v;
                Vw[(size_t)(cs + i - j0) * R + j0 + r] =
                    (r < rc) ? cx<T>{0, 0} : (r == rc) ? cx<T>{1, 0} : v;
            }
        }
        const int k0 = cs + w, k1 = j0 + nbp;
        if (k0 < k1) {
            // ---- remaining panel columns: a <- a - V T^H (V^H a), one wave per column.
            // The first 64 RPL rows of the column stay in registers between
            // the two passes (all their loads in flight at once); rows beyond
            // are streamed twice, TB rows per lane at a time.
            for (int k = k0 + wv; k < k1; k += NW) {
                cx<T> *Ak = A + (size_t)k * ldy + j0;
                cx<T> ar[RPL];
#pragma unroll
                for (int q = 0; q < RPL; ++q) {
                    const int r = r0 + lane + 64 * q;
                    ar[q] = (r < M) ? Ak[r] : cx<T>{0, 0};
                }
                cx<T> acc[SB];
                // rows r0 .. r0 + 63 hold the unit diagonal and zeros above it
#pragma unroll
                for (int i = 0; i < SB; ++i) {
                    const int r = r0 + lane;
                    acc[i] = cx<T>{0, 0};
                    if (i < w && r >= r0 + i && r < M) {
                        const cx<T> u = (r == r0 + i) ? cx<T>{1, 0} : Vs[(size_t)i * M + r];
                        acc[i] = cmulc(u, ar[0]);
                    }
                }
#pragma unroll
                for (int q = 1; q < RPL; ++q) {
                    const int r = r0 + lane + 64 * q;
                    if (r < M) {
#pragma unroll
                        for (int i = 0; i < SB; ++i)
                            if (i < w) acc[i] = acc[i] + cmulc(Vs[(size_t)i * M + r], ar[q]);
                    }
                    __builtin_amdgcn_sched_barrier(0);  // keep the LDS reads of one row block live at a time
                }
                // rows beyond: blocks of TB rows per lane, loads issued together
                for (int base = r0 + lane + 64 * RPL; base < M; base += 64 * TB) {
                    cx<T> tb[TB];
#pragma unroll
                    for (int q = 0; q < TB; ++q) {
                        const int r = base + 64 * q;
                        tb[q] = (r < M) ? Ak[r] : cx<T>{0, 0};
                    }
#pragma unroll
                    for (int q = 0; q < TB; ++q) {
                        const int r = base + 64 * q;
                        if (r < M) {
#pragma unroll
                            for (int i = 0; i < SB; ++i)
                                if (i < w) acc[i] = acc[i] + cmulc(Vs[(size_t)i * M + r], tb[q]);
                        }
                        __builtin_amdgcn_sched_barrier(0);
                    }
                }
#pragma unroll
                for (int i = 0; i < SB; ++i) {
                    const cx<double> v{wave_sum((double)acc[i].re), wave_sum((double)acc[i].im)};
                    if (lane == 0) Zs[wv][i] = v;
                }
                // z = T^H acc (lane 0), then a <- a - V z with z read from LDS
                if (lane == 0) {
                    // in place, last entry first: z_jj reads entries 0..jj only
                    for (int jj = w - 1; jj >= 0; --jj) {
                        cx<double> sacc{0, 0};
                        for (int ii = 0; ii <= jj; ++ii) sacc = sacc + conj(Ts[ii][jj]) * Zs[wv][ii];
                        Zs[wv][jj] = sacc;
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();
                cx<T> zt[SB];
#pragma unroll
                for (int i = 0; i < SB; ++i) zt[i] = cvt<T>(Zs[wv][i]);
                {
                    const int r = r0 + lane;
                    if (r < M) {
                        cx<T> a = ar[0];
#pragma unroll
                        for (int i = 0; i < SB; ++i)
                            if (i < w && r >= r0 + i) {
                                const cx<T> u = (r == r0 + i) ? cx<T>{1, 0} : Vs[(size_t)i * M + r];
                                a = a - u * zt[i];
                            }
                        Ak[r] = a;
                    }
                }
#pragma unroll
                for (int q = 1; q < RPL; ++q) {
                    const int r = r0 + lane + 64 * q;
                    if (r < M) {
                        cx<T> a = ar[q];
#pragma unroll
                        for (int i = 0; i < SB; ++i)
                            if (i < w) a = a - Vs[(size_t)i * M + r] * zt[i];
                        Ak[r] = a;
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
                for (int base = r0 + lane + 64 * RPL; base < M; base += 64 * TB) {
                    cx<T> tb[TB];
#pragma unroll
                    for (int q = 0; q < TB; ++q) {
                        const int r = base + 64 * q;
                        tb[q] = (r < M) ? Ak[r] : cx<T>{0, 0};
                    }
#pragma unroll
                    for (int q = 0; q < TB; ++q) {
                        const int r = base + 64 * q;
                        if (r < M) {
                            cx<T> a = tb[q];
#pragma unroll
                            for (int i = 0; i < SB; ++i)
                                if (i < w) a = a - Vs[(size_t)i * M + r] * zt[i];
                            Ak[r] = a;
                        }
                        __builtin_amdgcn_sched_barrier(0);
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();
            }
        }
        __syncthreads();
    }
}

// Panel QR: the LDS-blocked kernel when SB columns of the panel fit in
// 128 KB of LDS (SB = 8, 4 or 2), else the column-by-column kernel.
// vcols: columns per matrix of the explicit-reflector buffer Vw.
template <typename T>
void launch_qr_panel(int R, int C, int j0, int nbp, cx<T> *Y, int ldy, size_t sY, cx<T> *tau,
                     cx<T> *Vw, int vcols, int batch, hipStream_t s) {
    const size_t colb = sizeof(cx<T>) * (size_t)(R - j0);
    const size_t cap = 128 * 1024;
    auto go = [&](auto kern, int sb) {
        const size_t lds = colb * sb;
        (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds);
        hipLaunchKernelGGL(kern, dim3(batch), dim3(BT), lds, s, R, C, j0, nbp, Y, ldy, sY, tau, Vw,
                           vcols);
    };
    // RPL rows of a column per lane kept in registers, TB per streamed block
    // (sizes that keep every variant within 128 VGPRs without spills)
    constexpr bool F = sizeof(T) == 4;
    if (colb * 8 <= cap)
        go(k_qr_panel_lds<T, 8, F ? 12 : 6, F ? 3 : 2>, 8);
    else if (colb * 4 <= cap)
        go(k_qr_panel_lds<T, 4, F ? 12 : 8, F ? 3 : 2>, 4);
    else if (colb * 2 <= cap)
        go(k_qr_panel_lds<T, 2, F ? 12 : 8, F ? 3 : 2>, 2);
    else
        hipLaunchKernelGGL(k_qr_panel<T>, dim3(batch), dim3(BT), 0, s, R, C, j0, nbp, Y, ldy, sY,
                           tau, Vw, vcols);
}

// T (nbp x nbp, upper) with H_0 ... H_{nbp-1} = I - V T V^H from G = V^H V
// and tau (LAPACK larft, forward columnwise); one wave per matrix; G and T
// are P x P (leading dimension P) per matrix.
template <typename T, int P>
__global__ __launch_bounds__(64) void k_qr_tmat(int C, int j0, int nbp, const cx<T> *__restrict__ Gall,
                                               const cx<T> *__restrict__ tauall,
                                               cx<T> *__restrict__ Tall) {
    static_assert(P <= 64, "one lane per row of T");
    // upper triangle packed by columns: T[r][c] (r <= c) at c (c + 1) / 2 + r
    __shared__ cx<double> Ts[P * (P + 1) / 2];
    __shared__ cx<double> z[P];
    const int b = blockIdx.x, t = threadIdx.x;
    const cx<T> *G = Gall + (size_t)b * P * P;
    const cx<T> *tau = tauall + (size_t)b * C + j0;
    cx<T> *Tm = Tall + (size_t)b * P * P;
    const cx<double> tau_t = t < nbp ? cvt<double>(tau[t]) : cx<double>{0, 0};
    // column i of G (rows t < P), prefetched one step ahead
    cx<T> gnext = t < P ? G[t] : cx<T>{0, 0};
    for (int i = 0; i < nbp; ++i) {
        const cx<T> gcur = gnext;
        if (i + 1 < nbp && t < P) gnext = G[(size_t)(i + 1) * P + t];
        const cx<double> ti{__shfl(tau_t.re, i, 64), __shfl(tau_t.im, i, 64)};
        if (t < i) z[t] = cx<double>{0, 0} - ti * cvt<double>(gcur);  // G[t, i]
        __syncthreads();
        const int ci = i * (i + 1) / 2;
        if (t < i) {
            // two interleaved partial sums shorten the dependent FMA chain
            cx<double> a0{0, 0}, a1{0, 0};
            int l = t;
            for (; l + 1 < i; l += 2) {
                a0 = a0 + Ts[l * (l + 1) / 2 + t] * z[l];
                a1 = a1 + Ts[(l + 1) * (l + 2) / 2 + t] * z[l + 1];
            }
            if (l < i) a0 = a0 + Ts[l * (l + 1) / 2 + t] * z[l];
            Ts[ci + t] = a0 + a1;
        }
        if (t == 0) Ts[ci + i] = ti;
        __syncthreads();
    }
    for (int e = t; e < P * P; e += 64) {
        const int r = e % P, c = e / P;
        Tm[e] = (r <= c && c < nbp) ? cvt<T>(Ts[c * (c + 1) / 2 + r]) : cx<T>{0, 0};
    }
}

// Rb (C x C) = upper triangle of Y's leading C x C block, zeros below
template <typename T>
__global__ void k_copy_upper(int C, const cx<T> *__restrict__ Yall, int ldy, size_t sY,
                             cx<T> *__restrict__ Rball) {
    const int b = blockIdx.z, c = blockIdx.y;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= C) return;
    const cx<T> v = Yall[(size_t)b * sY + (size_t)c * ldy + r];
    Rball[(size_t)b * C * C + (size_t)c * C + r] = (r <= c) ? v : cx<T>{0, 0};
}

// LQ panel of the band reduction, as a QR of the conjugate transpose:
// P[(k + q) + (k + i) ldp] = conj(Z[k + i, k + NB + q]), q < N2, i < nbp
template <typename T>
__global__ void k_lq_gather(int k, int N2, int nbp, const cx<T> *__restrict__ Zall, int ldz,
                            size_t sZ, cx<T> *__restrict__ Pall, int ldp, size_t sP) {
    const int b = blockIdx.z, i = blockIdx.y;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= N2 || i >= nbp) return;
    const cx<T> z = Zall[(size_t)b * sZ + (size_t)(k + NB + q) * ldz + (k + i)];
    Pall[(size_t)b * sP + (size_t)(k + i) * ldp + (k + q)] = conj(z);
}

// rows k..k+nbp-1 of Z after the LQ panel: Z[k + i, k + NB + q] =
// conj(R_w[q, i]) for q <= i (lower triangle), zero beyond
template <typename T>
__global__ void k_lq_scatter(int k, int nbp, cx<T> *__restrict__ Zall, int ldz, size_t sZ,
                             const cx<T> *__restrict__ Pall, int ldp, size_t sP) {
    const int b = blockIdx.y;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nbp * nbp) return;
    const int i = e % nbp, q = e / nbp;
    const cx<T> r = Pall[(size_t)b * sP + (size_t)(k + i) * ldp + (k + q)];
    Zall[(size_t)b * sZ + (size_t)(k + NB + q) * ldz + (k + i)] =
        (q <= i) ? conj(r) : cx<T>{0, 0};
}

}  // namespace

size_t geqrf_work_elems(int R, int C) {
    return (size_t)R * NBQ + 2 * (size_t)NBQ * NBQ + 2 * (size_t)NBQ * C;
}

template <typename T>
ia_status launch_geqrf(int R, int C, cx<T> *Y, int ldy, size_t sY, cx<T> *tau, cx<T> *work,
                       int batch, hipStream_t s) {
    if (R < C) return fail(IA_ERR_ARG, "QR needs R >= C (%d x %d)", R, C);
    constexpr int P = NBQ;
    cx<T> *Vw = work;
    cx<T> *G = Vw + (size_t)batch * R * P;
    cx<T> *Tm = G + (size_t)batch * P * P;
    cx<T> *W = Tm + (size_t)batch * P * P;
    cx<T> *W2 = W + (size_t)batch * P * C;
    const cx<T> one{1, 0}, zero{0, 0}, mone{-1, 0};
    const size_t sV = (size_t)R * P, sT = (size_t)P * P, sW = (size_t)P * C;
    for (int j0 = 0; j0 < C; j0 += P) {
        const int nbp = std::min(P, C - j0);
        launch_qr_panel<T>(R, C, j0, nbp, Y, ldy, sY, tau, Vw, P, batch, s);
        const int N = C - j0 - nbp, M = R - j0;
        if (N <= 0) continue;
        const cx<T> *V0 = Vw + j0;  // rows j0.. of the explicit panel reflectors
        // G = V^H V
        launch_gemm<T>('C', 'N', nbp, nbp, M, one, V0, R, sV, V0, R, sV, zero, G, P, sT, batch, s);
        hipLaunchKernelGGL((k_qr_tmat<T, P>), dim3(batch), dim3(64), 0, s, C, j0, nbp, G, tau, Tm);
        cx<T> *A2 = Y + j0 + (size_t)(j0 + nbp) * ldy;
        // W = V^H A2, W2 = T^H W, A2 -= V W2
        launch_gemm<T>('C', 'N', nbp, N, M, one, V0, R, sV, A2, ldy, sY, zero, W, P, sW, batch, s);
        launch_gemm<T>('C', 'N', nbp, N, nbp, one, Tm, P, sT, W, P, sW, zero, W2, P, sW, batch, s);
        launch_gemm<T>('N', 'N', M, N, nbp, mone, V0, R, sV, W2, P, sW, one, A2, ldy, sY, batch, s);
    }
    IA_HIP(hipGetLastError());
    return IA_OK;
}

// Stage one of the two-stage bidiagonalisation: Z (R x C, R >= C) ->
// Q1^H Z P1, upper band of width NB in the leading C x C block.  Per panel
// k: QR of the NB columns k.. (reflectors below the diagonal of Z, tau in
// tauQ) and its trailing update from the left; then, while columns remain
// beyond k + NB, the LQ of rows k..k+NB-1 over columns k+NB.. computed as
// a QR of their conjugate transpose in Pm (C-NB x C-NB per matrix; the
// reflectors stay there, tau in tauP) and applied from the right to the
// rows below, Z2 <- Z2 (I - V T V^H).  All trailing updates are GEMMs.
// tauP has C - NB entries per matrix (the reflectors of Pm's columns).
size_t band_work_elems(int R, int C) {
    return (size_t)R * NB + 2 * (size_t)NB * NB + 2 * (size_t)NB * std::max(R, C);
}

template <typename T>
ia_status launch_gebrd_band(int R, int C, cx<T> *Z, int ldz, size_t sZ, cx<T> *tauQ, cx<T> *Pm,
                            cx<T> *tauP, cx<T> *work, int batch, hipStream_t s) {
    if (R < C) return fail(IA_ERR_ARG, "band reduction needs R >= C (%d x %d)", R, C);
    cx<T> *Vw = work;
    cx<T> *G = Vw + (size_t)batch * R * NB;
    cx<T> *Tm = G + (size_t)batch * NB * NB;
    const int Wl = std::max(R, C);
    cx<T> *W = Tm + (size_t)batch * NB * NB;
    cx<T> *W2 = W + (size_t)batch * NB * Wl;
    const cx<T> one{1, 0}, zero{0, 0}, mone{-1, 0};
    const size_t sT = (size_t)NB * NB, sW = (size_t)NB * Wl;
    const int Cp = C - NB;
    const size_t sP = (size_t)std::max(Cp, 1) * std::max(Cp, 1);
    for (int k = 0; k < C; k += NB) {
        const int w = std::min(NB, C - k);
        // ---- QR of columns k..k+w-1, left update of columns k+w..
        launch_qr_panel<T>(R, C, k, w, Z, ldz, sZ, tauQ, Vw, NB, batch, s);
        const size_t sV = (size_t)R * NB;
        const int N = C - k - w, M = R - k;
        if (N > 0) {
            const cx<T> *V0 = Vw + k;
            launch_gemm<T>('C', 'N', w, w, M, one, V0, R, sV, V0, R, sV, zero, G, NB, sT, batch, s);
            hipLaunchKernelGGL((k_qr_tmat<T, NB>), dim3(batch), dim3(64), 0, s, C, k, w, G, tauQ, Tm);
            cx<T> *A2 = Z + k + (size_t)(k + w) * ldz;
            launch_gemm<T>('C', 'N', w, N, M, one, V0, R, sV, A2, ldz, sZ, zero, W, NB, sW, batch, s);
            launch_gemm<T>('C', 'N', w, N, w, one, Tm, NB, sT, W, NB, sW, zero, W2, NB, sW, batch, s);
            launch_gemm<T>('N', 'N', M, N, w, mone, V0, R, sV, W2, NB, sW, one, A2, ldz, sZ, batch,
                           s);
        }
        if (k + NB >= C) continue;
        // ---- LQ of rows k..k+NB-1 over columns k+NB..C-1
        const int N2 = C - k - NB;           // columns of the row panel
        const int nbp = std::min(NB, N2);    // reflectors (rows with columns left)
        hipLaunchKernelGGL(k_lq_gather<T>, dim3((N2 + 255) / 256, nbp, batch), dim3(256), 0, s, k,
                           N2, nbp, Z, ldz, sZ, Pm, Cp, sP);
        launch_qr_panel<T>(Cp, Cp, k, nbp, Pm, Cp, sP, tauP, Vw, NB, batch, s);
        hipLaunchKernelGGL(k_lq_scatter<T>, dim3((nbp * nbp + 255) / 256, batch), dim3(256), 0, s,
                           k, nbp, Z, ldz, sZ, Pm, Cp, sP);
        // rows below the reflected ones: Z2 <- Z2 - ((Z2 V) T) V^H
        const int M2 = R - k - nbp;
        if (M2 <= 0) continue;
        const size_t sV2 = (size_t)Cp * NB;
        const cx<T> *V0 = Vw + k;
        cx<T> *Z2 = Z + (k + nbp) + (size_t)(k + NB) * ldz;
        launch_gemm<T>('C', 'N', nbp, nbp, N2, one, V0, Cp, sV2, V0, Cp, sV2, zero, G, NB, sT,
                       batch, s);
        hipLaunchKernelGGL((k_qr_tmat<T, NB>), dim3(batch), dim3(64), 0, s, Cp, k, nbp, G, tauP, Tm);
        launch_gemm<T>('N', 'N', M2, nbp, N2, one, Z2, ldz, sZ, V0, Cp, sV2, zero, W, Wl, sW, batch,
                       s);
        launch_gemm<T>('N', 'N', M2, nbp, nbp, one, W, Wl, sW, Tm, NB, sT, zero, W2, Wl, sW, batch,
                       s);
        launch_gemm<T>('N', 'C', M2, N2, nbp, mone, W2, Wl, sW, V0, Cp, sV2, one, Z2, ldz, sZ,
                       batch, s);
    }
    IA_HIP(hipGetLastError());
    return IA_OK;
}

template <typename T>
void launch_copy_upper(int C, const cx<T> *Y, int ldy, size_t sY, cx<T> *Rb, int batch,
                       hipStream_t s) {
    dim3 g((C + 255) / 256, C, batch);
    hipLaunchKernelGGL(k_copy_upper<T>, g, dim3(256), 0, s, C, Y, ldy, sY, Rb);
}

template ia_status launch_geqrf<double>(int, int, cx<double> *, int, size_t, cx<double> *,
                                        cx<double> *, int, hipStream_t);
template ia_status launch_geqrf<float>(int, int, cx<float> *, int, size_t, cx<float> *,
                                       cx<float> *, int, hipStream_t);
template ia_status launch_gebrd_band<double>(int, int, cx<double> *, int, size_t, cx<double> *,
                                             cx<double> *, cx<double> *, cx<double> *, int,
                                             hipStream_t);
template ia_status launch_gebrd_band<float>(int, int, cx<float> *, int, size_t, cx<float> *,
                                            cx<float> *, cx<float> *, cx<float> *, int,
                                            hipStream_t);
template void launch_copy_upper<double>(int, const cx<double> *, int, size_t, cx<double> *, int,
                                        hipStream_t);
template void launch_copy_upper<float>(int, const cx<float> *, int, size_t, cx<float> *, int,
                                       hipStream_t);

}  // namespace ia
