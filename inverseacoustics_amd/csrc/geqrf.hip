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
//                a tall panel whose LDS sub-blocks hold fewer than 8
//                columns is factored as two 32-column halves, the first
//                half's reflectors applied to the second by three GEMMs;
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
        {
            // all of a thread's loads in flight before its LDS stores (the
            // launcher sizes the LDS to at most 128 KB, so M <= LROWS * BT)
            constexpr int LROWS = (128 * 1024 / ((int)sizeof(cx<T>) * SB) + BT - 1) / BT;
            cx<T> v[LROWS][SB];
#pragma unroll
            for (int q = 0; q < LROWS; ++q) {
                const int r = min(t + q * BT, M - 1);
#pragma unroll
                for (int i = 0; i < SB; ++i) v[q][i] = A[(size_t)(cs + min(i, w - 1)) * ldy + j0 + r];
            }
#pragma unroll
            for (int q = 0; q < LROWS; ++q) {
                const int r = t + q * BT;
#pragma unroll
                for (int i = 0; i < SB; ++i)
                    if (i < w && r < M) Vs[(size_t)i * M + r] = v[q][i];
            }
        }
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
// and tau (the T of LAPACK larft, forward columnwise), built by doubling
// instead of column by column: for adjacent blocks of h reflectors,
// (I - V1 T11 V1^H)(I - V2 T22 V2^H) = I - [V1 V2] [T11 T12; 0 T22] [V1 V2]^H
// with T12 = -T11 (V1^H V2) T22, so log2(P) levels of small independent
// products, starting from T_ii = tau_i, replace the P dependent
// triangular products of the column recurrence.  One workgroup of 256
// threads per matrix; G, tau and T are P x P (leading dimension P).
template <typename T, int P>
__global__ __launch_bounds__(256) void k_qr_tmat(int C, int j0, int nbp, const cx<T> *__restrict__ Gall,
                                                const cx<T> *__restrict__ tauall,
                                                cx<T> *__restrict__ Tall) {
    constexpr int NT = 256;
    // strictly upper part: G until the level that replaces the block by T12;
    // diagonal: tau; zero outside nbp (so padded reflectors contribute nothing)
    __shared__ cx<double> Ms[P][P + 1];
    __shared__ cx<double> tmp[P * P / 4];  // G12 T22 of every pair of a level
    const int b = blockIdx.x, t = threadIdx.x;
    const cx<T> *G = Gall + (size_t)b * P * P;
    const cx<T> *tau = tauall + (size_t)b * C + j0;
    cx<T> *Tm = Tall + (size_t)b * P * P;
    for (int e = t; e < P * P; e += NT) {
        const int r = e % P, c = e / P;
        cx<double> v{0, 0};
        if (c < nbp) {
            if (r < c) v = cvt<double>(G[e]);
            else if (r == c) v = cvt<double>(tau[r]);
        }
        Ms[r][c] = v;
    }
    __syncthreads();
    for (int h = 1; h < P; h *= 2) {
        const int n = (P / (2 * h)) * h * h;  // outputs of the level
        // tmp[r, c] = sum_{l <= c} G12[r, l] T22[l, c]
        for (int e = t; e < n; e += NT) {
            const int pr = e / (h * h), r = e % h, c = (e / h) % h, o = pr * 2 * h;
            cx<double> acc{0, 0};
            for (int l = 0; l <= c; ++l) acc = acc + Ms[o + r][o + h + l] * Ms[o + h + l][o + h + c];
            tmp[e] = acc;
        }
        __syncthreads();
        // T12[r, c] = -sum_{l >= r} T11[r, l] tmp[l, c]
        for (int e = t; e < n; e += NT) {
            const int pr = e / (h * h), r = e % h, c = (e / h) % h, o = pr * 2 * h;
            const cx<double> *tc = tmp + (size_t)pr * h * h + (size_t)c * h;
            cx<double> acc{0, 0};
            for (int l = r; l < h; ++l) acc = acc + Ms[o + r][o + l] * tc[l];
            Ms[o + r][o + h + c] = cx<double>{0, 0} - acc;
        }
        __syncthreads();
    }
    for (int e = t; e < P * P; e += NT) {
        const int r = e % P, c = e / P;
        Tm[e] = (r <= c && c < nbp) ? cvt<T>(Ms[r][c]) : cx<T>{0, 0};
    }
}

// zero rows [0, nr) of columns [0, nc) of every matrix's slab (ld ld,
// stride sV)
template <typename T>
__global__ void k_zero_block(int nr, int nc, cx<T> *__restrict__ Vall, int ld, size_t sV) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nr * nc) return;
    Vall[(size_t)blockIdx.y * sV + (size_t)(e / nr) * ld + e % nr] = cx<T>{0, 0};
}

template <typename T>
__global__ void k_copy_upper(int C, const cx<T> *__restrict__ Yall, int ldy, size_t sY,
                             cx<T> *__restrict__ Rball) {
    const int b = blockIdx.z, c = blockIdx.y;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= C) return;
    const cx<T> v = Yall[(size_t)b * sY + (size_t)c * ldy + r];
    Rball[(size_t)b * C * C + (size_t)c * C + r] = (r <= c) ? v : cx<T>{0, 0};
}

// LQ panel of the band reduction, as a QR of the conjugate transpose, with
// the left update of the panel's rows folded in:
// P[(k + q) + (k + i) ldp] = conj(Z[k + i, k + NB + q] - sum_j V1[k + i, j] W1[j, q])
// for i < nbp, q < N2, j < w (nbp, w <= 32).  One workgroup per 32 columns
// q; the three 32 x 32 tiles are read with consecutive lanes on consecutive
// addresses and P is written the same way.
template <typename T>
__global__ __launch_bounds__(256) void k_lq_gather(int k, int N2, int nbp, int w,
                                                   const cx<T> *__restrict__ Zall, int ldz,
                                                   size_t sZ, const cx<T> *__restrict__ V1all,
                                                   int ldv, size_t sV,
                                                   const cx<T> *__restrict__ W1all, int ldw,
                                                   size_t sW, cx<T> *__restrict__ Pall, int ldp,
                                                   size_t sP) {
    __shared__ cx<T> Vt[32][33];  // V1[k + i, j]
    __shared__ cx<T> Wt[32][33];  // W1[j, q0 + c]
    __shared__ cx<T> At[32][33];  // Z[k + i, k + NB + q0 + c]
    const int b = blockIdx.y, q0 = blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const cx<T> *Z = Zall + (size_t)b * sZ;
    const cx<T> *V1 = V1all + (size_t)b * sV;
    const cx<T> *W1 = W1all + (size_t)b * sW;
    const cx<T> zero{0, 0};
    for (int c = ty; c < 32; c += 8) {
        Vt[tx][c] = (c < w && tx < nbp) ? V1[(size_t)c * ldv + k + tx] : zero;
        const bool qv = q0 + c < N2;
        Wt[tx][c] = (qv && tx < w) ? W1[(size_t)(q0 + c) * ldw + tx] : zero;
        At[tx][c] = (qv && tx < nbp) ? Z[(size_t)(k + NB + q0 + c) * ldz + k + tx] : zero;
    }
    __syncthreads();
    if (q0 + tx >= N2) return;
    cx<T> *P = Pall + (size_t)b * sP;
    for (int i = ty; i < nbp; i += 8) {
        cx<T> acc = At[i][tx];
        for (int j = 0; j < w; ++j) acc = acc - Vt[i][j] * Wt[j][tx];
        P[(size_t)(k + i) * ldp + k + q0 + tx] = conj(acc);
    }
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

// B[r0 + j, q] = conj(V[k + q, j]) for j < nbp <= 32, q < N2: the LQ
// reflectors as rows of the band update's right operand, transposed
// through a 32 x 32 LDS tile so that reads and writes are both contiguous
template <typename T>
__global__ __launch_bounds__(256) void k_conj_rows(int k, int N2, int nbp,
                                                   const cx<T> *__restrict__ Vall, int ldv,
                                                   size_t sV, cx<T> *__restrict__ Ball, int ldb,
                                                   size_t sB, int r0) {
    __shared__ cx<T> tile[32][33];
    const int b = blockIdx.y, q0 = blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const cx<T> *V = Vall + (size_t)b * sV;
    cx<T> *B = Ball + (size_t)b * sB;
    for (int j = ty; j < nbp; j += 8)
        if (q0 + tx < N2) tile[j][tx] = V[(size_t)j * ldv + k + q0 + tx];
    __syncthreads();
    for (int qq = ty; qq < 32; qq += 8)
        if (q0 + qq < N2 && tx < nbp) B[(size_t)(q0 + qq) * ldb + r0 + tx] = conj(tile[tx][qq]);
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
    constexpr int H = P / 2;
    const size_t sTh = (size_t)H * H;
    for (int j0 = 0; j0 < C; j0 += P) {
        const int nbp = std::min(P, C - j0), M = R - j0;
        const cx<T> *V0 = Vw + j0;  // rows j0.. of the explicit panel reflectors
        if (nbp > H && (size_t)M * sizeof(cx<T>) * 8 > 128 * 1024) {
            // the panel as two halves when its LDS sub-blocks are narrower
            // than 8 columns: the panel kernel streams the columns right of
            // each sub-block once per sub-block, so the first half's
            // reflectors reach the second half as one blocked update (three
            // GEMMs) instead of through H / SB sub-block passes (with 8-column
            // sub-blocks the extra launches cost what the passes save)
            launch_qr_panel<T>(R, C, j0, H, Y, ldy, sY, tau, Vw, P, batch, s);
            const int nr = nbp - H;
            cx<T> *Ar = Y + j0 + (size_t)(j0 + H) * ldy;
            launch_gemm<T>('C', 'N', H, H, M, one, V0, R, sV, V0, R, sV, zero, G, H, sTh, batch, s);
            hipLaunchKernelGGL((k_qr_tmat<T, H>), dim3(batch), dim3(256), 0, s, C, j0, H, G, tau,
                               Tm);
            launch_gemm<T>('C', 'N', H, nr, M, one, V0, R, sV, Ar, ldy, sY, zero, W, H, sW, batch, s);
            launch_gemm<T>('C', 'N', H, nr, H, one, Tm, H, sTh, W, H, sW, zero, W2, H, sW, batch, s);
            launch_gemm<T>('N', 'N', M, nr, H, mone, V0, R, sV, W2, H, sW, one, Ar, ldy, sY, batch,
                           s);
            // second half: its reflectors start at row j0 + H; rows j0 .. j0 + H - 1
            // of their explicit columns are zero
            launch_qr_panel<T>(R, C, j0 + H, nr, Y, ldy, sY, tau, Vw + (size_t)H * R, P, batch, s);
            hipLaunchKernelGGL(k_zero_block<T>, dim3((H * nr + 255) / 256, batch), dim3(256), 0, s,
                               H, nr, Vw + (size_t)H * R + j0, R, sV);
        } else {
            launch_qr_panel<T>(R, C, j0, nbp, Y, ldy, sY, tau, Vw, P, batch, s);
        }
        const int N = C - j0 - nbp;
        if (N <= 0) continue;
        // G = V^H V
        launch_gemm<T>('C', 'N', nbp, nbp, M, one, V0, R, sV, V0, R, sV, zero, G, P, sT, batch, s);
        hipLaunchKernelGGL((k_qr_tmat<T, P>), dim3(batch), dim3(256), 0, s, C, j0, nbp, G, tau, Tm);
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
// tauQ, explicit V1 in the first NB columns of Vw1); then, while columns
// remain beyond k + NB, the LQ of rows k..k+NB-1 over columns k+NB..
// computed as a QR of their conjugate transpose in Pm (C-NB x C-NB per
// matrix; the reflectors stay there, explicit V2 in Vw2, tau in tauP).
//
// The two trailing updates of the rows below the LQ rows are applied as
// one rank-2NB GEMM.  With A the trailing block before either update,
// W1 = T1^H V1^H A, A' = A - V1 W1 (left) and A'' = A' - A' V2 T2 V2^H
// (right); since A' V2 = A V2 - V1 (W1 V2),
//     A'' = A - [V1, X] [W1; V2^H],   X = (A V2 - V1 (W1 V2)) T2,
// so A is read by V1^H A and A V2 and read and written once by the update,
// instead of twice each by two rank-NB updates.  The LQ rows themselves
// take the left update inside k_lq_gather, on their way into Pm (what they
// hold beyond the band afterwards is never read).
// tauP has C - NB entries per matrix (the reflectors of Pm's columns).
size_t band_work_elems(int R, int C) {
    const size_t Wl = std::max(R, C);
    return 2 * (size_t)NB * R + (size_t)NB * C + 3 * (size_t)NB * NB + 3 * (size_t)NB * Wl;
}

template <typename T>
ia_status launch_gebrd_band(int R, int C, cx<T> *Z, int ldz, size_t sZ, cx<T> *tauQ, cx<T> *Pm,
                            cx<T> *tauP, cx<T> *work, int batch, hipStream_t s) {
    if (R < C) return fail(IA_ERR_ARG, "band reduction needs R >= C (%d x %d)", R, C);
    constexpr int NB2 = 2 * NB;
    const int Wl = std::max(R, C);
    cx<T> *Vw1 = work;                             // R x 2NB: [V1, X]
    cx<T> *Vw2 = Vw1 + (size_t)batch * R * NB2;    // C x NB: V2
    cx<T> *G = Vw2 + (size_t)batch * C * NB;       // NB x NB
    cx<T> *Tm = G + (size_t)batch * NB * NB;       // NB x NB: T1, then T2
    cx<T> *U = Tm + (size_t)batch * NB * NB;       // NB x NB: W1 V2
    cx<T> *Wt = U + (size_t)batch * NB * NB;       // NB x Wl: V1^H A, then A V2
    cx<T> *Bm = Wt + (size_t)batch * NB * Wl;      // 2NB x C: [W1; V2^H]
    const cx<T> one{1, 0}, zero{0, 0}, mone{-1, 0};
    const size_t sV1 = (size_t)R * NB2, sT = (size_t)NB * NB, sW = (size_t)NB * Wl,
                 sB = (size_t)NB2 * C;
    const int Cp = C - NB;
    const size_t sP = (size_t)std::max(Cp, 1) * std::max(Cp, 1);
    const size_t sV2 = (size_t)std::max(Cp, 1) * NB;
    // complex64: A V2 as its conjugate transpose V2^H A^H, the faster form
    // of that product in single precision (complex128 prefers A V2)
    constexpr bool F = sizeof(T) == sizeof(float);
    for (int k = 0; k < C; k += NB) {
        const int w = std::min(NB, C - k);
        // ---- QR of columns k..k+w-1
        launch_qr_panel<T>(R, C, k, w, Z, ldz, sZ, tauQ, Vw1, NB2, batch, s);
        const int N = C - k - w, M = R - k;
        if (N <= 0) continue;
        const cx<T> *V1 = Vw1 + k;
        launch_gemm<T>('C', 'N', w, w, M, one, V1, R, sV1, V1, R, sV1, zero, G, NB, sT, batch, s);
        hipLaunchKernelGGL((k_qr_tmat<T, NB>), dim3(batch), dim3(256), 0, s, C, k, w, G, tauQ, Tm);
        cx<T> *A2 = Z + k + (size_t)(k + w) * ldz;
        // W1 = T1^H V1^H A (rows 0..w-1 of Bm)
        launch_gemm<T>('C', 'N', w, N, M, one, V1, R, sV1, A2, ldz, sZ, zero, Wt, NB, sW, batch, s);
        launch_gemm<T>('C', 'N', w, N, w, one, Tm, NB, sT, Wt, NB, sW, zero, Bm, NB2, sB, batch, s);
        if (k + NB >= C) {
            // last panel: left update only
            launch_gemm<T>('N', 'N', M, N, w, mone, V1, R, sV1, Bm, NB2, sB, one, A2, ldz, sZ, batch,
                           s);
            continue;
        }
        // ---- LQ of rows k..k+NB-1 over columns k+NB..C-1 (here w == NB, N == N2)
        const int N2 = N;
        const int nbp = std::min(NB, N2);  // reflectors (rows with columns left)
        // the LQ rows take the left update on their way into Pm
        hipLaunchKernelGGL(k_lq_gather<T>, dim3((N2 + 31) / 32, batch), dim3(256), 0, s, k, N2, nbp,
                           w, Z, ldz, sZ, Vw1, R, sV1, Bm, NB2, sB, Pm, Cp, sP);
        launch_qr_panel<T>(Cp, Cp, k, nbp, Pm, Cp, sP, tauP, Vw2, NB, batch, s);
        hipLaunchKernelGGL(k_lq_scatter<T>, dim3((nbp * nbp + 255) / 256, batch), dim3(256), 0, s,
                           k, nbp, Z, ldz, sZ, Pm, Cp, sP);
        const int M2 = R - k - nbp;  // rows below the LQ rows
        if (M2 <= 0) continue;
        const cx<T> *V2 = Vw2 + k;
        launch_gemm<T>('C', 'N', nbp, nbp, N2, one, V2, Cp, sV2, V2, Cp, sV2, zero, G, NB, sT,
                       batch, s);
        hipLaunchKernelGGL((k_qr_tmat<T, NB>), dim3(batch), dim3(256), 0, s, Cp, k, nbp, G, tauP, Tm);
        // V2^H as rows w..w+nbp-1 of Bm
        hipLaunchKernelGGL(k_conj_rows<T>, dim3((N2 + 31) / 32, batch), dim3(256), 0, s, k, N2, nbp,
                           Vw2, Cp, sV2, Bm, NB2, sB, w);
        cx<T> *Z2 = A2 + nbp;
        const cx<T> *V1b = V1 + nbp;           // rows k+nbp.. of V1
        cx<T> *X = Vw1 + (size_t)NB * R + k + nbp;  // column w.. of [V1, X]
        if (F) {
            // U^H = V2^H W1^H, Wt = (A V2)^H - U^H V1^H, X = Wt^H T2
            launch_gemm<T>('C', 'C', nbp, w, N2, one, V2, Cp, sV2, Bm, NB2, sB, zero, U, NB, sT,
                           batch, s);
            launch_gemm<T>('C', 'C', nbp, M2, N2, one, V2, Cp, sV2, Z2, ldz, sZ, zero, Wt, NB, sW,
                           batch, s);
            launch_gemm<T>('N', 'C', nbp, M2, w, mone, U, NB, sT, V1b, R, sV1, one, Wt, NB, sW,
                           batch, s);
            launch_gemm<T>('C', 'N', M2, nbp, nbp, one, Wt, NB, sW, Tm, NB, sT, zero, X, R, sV1,
                           batch, s);
        } else {
            // U = W1 V2, Wt = A V2 - V1 U, X = Wt T2
            launch_gemm<T>('N', 'N', w, nbp, N2, one, Bm, NB2, sB, V2, Cp, sV2, zero, U, NB, sT,
                           batch, s);
            launch_gemm<T>('N', 'N', M2, nbp, N2, one, Z2, ldz, sZ, V2, Cp, sV2, zero, Wt, Wl, sW,
                           batch, s);
            launch_gemm<T>('N', 'N', M2, nbp, w, mone, V1b, R, sV1, U, NB, sT, one, Wt, Wl, sW,
                           batch, s);
            launch_gemm<T>('N', 'N', M2, nbp, nbp, one, Wt, Wl, sW, Tm, NB, sT, zero, X, R, sV1,
                           batch, s);
        }
        // A'' = A - [V1, X] [W1; V2^H]
        launch_gemm<T>('N', 'N', M2, N2, w + nbp, mone, V1b, R, sV1, Bm, NB2, sB, one, Z2, ldz, sZ,
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

namespace {

// Q = [I; 0] (R x C per matrix, ld ldq)
template <typename T>
__global__ void k_set_eye(int R, int C, cx<T> *__restrict__ Qall, int ldq, size_t sQ) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)R * C) return;
    const int r = (int)(q % R), c = (int)(q / R);
    Qall[(size_t)blockIdx.y * sQ + (size_t)c * ldq + r] = cx<T>{r == c ? (T)1 : (T)0, (T)0};
}

// The explicit reflectors of columns j0..j0+nbp-1 of Y (unit diagonal,
// zeros above, Y below) in the layout k_qr_panel writes: column c of Vw
// (R x P per matrix, ld R) holds reflector j0 + c at its absolute rows.
template <typename T>
__global__ void k_refl_explicit(int R, int j0, int nbp, const cx<T> *__restrict__ Yall, int ldy,
                                size_t sY, cx<T> *__restrict__ Vall, int P) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)R * nbp) return;
    const int r = (int)(q % R), c = (int)(q / R), j = j0 + c;
    const cx<T> *Y = Yall + (size_t)blockIdx.y * sY;
    cx<T> v{0, 0};
    if (r == j) v = cx<T>{1, 0};
    else if (r > j) v = Y[(size_t)j * ldy + r];
    Vall[(size_t)blockIdx.y * R * P + (size_t)c * R + r] = v;
}

}  // namespace

// Q (R x C) = H_0 ... H_{C-1} [I; 0] from the reflectors launch_geqrf left
// in Y and tau (LAPACK orgqr, blocked): the panels taken last to first, each
// applied as X2 <- (I - V T V^H) X2 = X2 - V (T (V^H X2)) to rows and
// columns j0.. of Q (the columns before j0 are still e_c then, and the
// rows before j0 of the later ones zero), three GEMMs per panel; work as
// for launch_geqrf.
template <typename T>
ia_status launch_orgqr(int R, int C, const cx<T> *Y, int ldy, size_t sY, const cx<T> *tau,
                       cx<T> *Q, int ldq, size_t sQ, cx<T> *work, int batch, hipStream_t s) {
    if (R < C) return fail(IA_ERR_ARG, "orgqr needs R >= C (%d x %d)", R, C);
    constexpr int P = NBQ;
    cx<T> *Vw = work;
    cx<T> *G = Vw + (size_t)batch * R * P;
    cx<T> *Tm = G + (size_t)batch * P * P;
    cx<T> *W = Tm + (size_t)batch * P * P;
    cx<T> *W2 = W + (size_t)batch * P * C;
    const cx<T> one{1, 0}, zero{0, 0}, mone{-1, 0};
    const size_t sV = (size_t)R * P, sT = (size_t)P * P, sW = (size_t)P * C;
    hipLaunchKernelGGL(k_set_eye<T>, dim3((unsigned)(((size_t)R * C + 255) / 256), batch),
                       dim3(256), 0, s, R, C, Q, ldq, sQ);
    for (int j0 = ((C - 1) / P) * P; j0 >= 0; j0 -= P) {
        const int nbp = std::min(P, C - j0), M = R - j0, N = C - j0;
        hipLaunchKernelGGL(k_refl_explicit<T>,
                           dim3((unsigned)(((size_t)R * nbp + 255) / 256), batch), dim3(256), 0,
                           s, R, j0, nbp, Y, ldy, sY, Vw, P);
        const cx<T> *V0 = Vw + j0;
        launch_gemm<T>('C', 'N', nbp, nbp, M, one, V0, R, sV, V0, R, sV, zero, G, P, sT, batch, s);
        hipLaunchKernelGGL((k_qr_tmat<T, P>), dim3(batch), dim3(256), 0, s, C, j0, nbp, G, tau, Tm);
        cx<T> *X2 = Q + j0 + (size_t)j0 * ldq;
        launch_gemm<T>('C', 'N', nbp, N, M, one, V0, R, sV, X2, ldq, sQ, zero, W, P, sW, batch, s);
        launch_gemm<T>('N', 'N', nbp, N, nbp, one, Tm, P, sT, W, P, sW, zero, W2, P, sW, batch, s);
        launch_gemm<T>('N', 'N', M, N, nbp, mone, V0, R, sV, W2, P, sW, one, X2, ldq, sQ, batch, s);
    }
    IA_HIP(hipGetLastError());
    return IA_OK;
}

template ia_status launch_orgqr<double>(int, int, const cx<double> *, int, size_t,
                                        const cx<double> *, cx<double> *, int, size_t,
                                        cx<double> *, int, hipStream_t);

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
