// Batched complex LU with partial pivoting and multi-right-hand-side solves.
//
// Replaces the QR least-squares solves the reference uses for the square
// ESM/IBEM systems (cusolverDnZZgels at CUDA/esm.cuh:73-85 and
// CUDA/ibem.cuh:389-391; zgels_ at esm.cuh:153-168): for a square
// nonsingular D both give D^-1, and the reference factors the same D twice
// (esm.cuh:75,85) where this code factors it once.
//
// getrf: right-looking blocked LU, panel width NB.  Per panel:
//   k_panel     one workgroup per matrix factors the NB-column panel
//               (pivot = first max of |re|+|im|, the LAPACK izamax rule);
//   k_laswp     applies the panel's row interchanges to the other columns;
//   k_trsm_ll   U12 = L11^-1 A12 (unit lower, one thread per column);
//   launch_gemm A22 -= L21 U12 on the matrix cores.
// getrs: P D = L U;  op 'N' solves D X = B, op 'C' solves D^H X = B, with
// blocked triangular solves (diagonal block per column thread, GEMM update).
#include "ia_internal.hpp"

#include <cstdlib>

namespace ia {

namespace {

constexpr int NB = 32;    // diagonal solves and panel columns
constexpr int NBO = 128;  // rank of the trailing GEMM updates

template <typename T>
__device__ __forceinline__ T cabs1(cx<T> a) { return fabs(a.re) + fabs(a.im); }

template <typename T>
__global__ __launch_bounds__(512) void k_panel(int n, int k, int jb, cx<T> *A, int lda,
                                              size_t strideA, int *ipiv, int *info) {
    __shared__ T red_v[512];
    __shared__ int red_i[512];
    __shared__ cx<T> urow[NB];
    __shared__ cx<T> piv_inv;
    const int b = blockIdx.x;
    cx<T> *Ab = A + (size_t)b * strideA;
    int *ip = ipiv + (size_t)b * n;
    const int t = threadIdx.x;
    for (int c = 0; c < jb; ++c) {
        const int col = k + c;
        cx<T> *Ac = Ab + (size_t)col * lda;
        // pivot search over rows col..n-1
        T best = (T)-1;
        int bi = n;
        for (int r = col + t; r < n; r += 512) {
            T v = cabs1(Ac[r]);
            if (v > best) {
                best = v;
                bi = r;
            }
        }
        red_v[t] = best;
        red_i[t] = bi;
        __syncthreads();
        for (int s = 256; s > 0; s >>= 1) {
            if (t < s) {
                T v2 = red_v[t + s];
                int i2 = red_i[t + s];
                if (v2 > red_v[t] || (v2 == red_v[t] && i2 < red_i[t])) {
                    red_v[t] = v2;
                    red_i[t] = i2;
                }
            }
            __syncthreads();
        }
        const int p = red_i[0];
        const T pv = red_v[0];
        if (t == 0) {
            ip[col] = p;
            if (pv == (T)0 && info && info[b] == 0) info[b] = col + 1;
        }
        // swap rows col and p within the panel columns
        if (p != col && t < jb) {
            cx<T> *a1 = Ab + (size_t)(k + t) * lda;
            cx<T> tmp = a1[col];
            a1[col] = a1[p];
            a1[p] = tmp;
        }
        __syncthreads();
        if (t < jb) urow[t] = Ab[(size_t)(k + t) * lda + col];
        if (t == 0) {
            cx<T> d = Ac[col];
            piv_inv = (pv == (T)0) ? cx<T>{0, 0} : cdiv(cx<T>{1, 0}, d);
        }
        __syncthreads();
        const cx<T> pinv = piv_inv;
        for (int r = col + 1 + t; r < n; r += 512) {
            cx<T> l = Ac[r] * pinv;
            Ac[r] = l;
            for (int cc = c + 1; cc < jb; ++cc) {
                cx<T> *a2 = Ab + (size_t)(k + cc) * lda;
                a2[r] = a2[r] - l * urow[cc];
            }
        }
        __syncthreads();
    }
}

// Blocked form of k_panel: SB columns at a time are held in LDS (rows
// k..n-1) and factored there (same pivot rule); each sub-panel's row
// interchanges are then applied to the panel's other columns, and its L
// to the panel's later columns as one rank-SB update (one pass per column
// instead of one per pivot).  Thread 0 finally lists the panel's
// interchanges as a row permutation, perm = (count, dst[], src[]), for
// k_laswp_perm.
constexpr int LBT = 1024;
constexpr size_t LDS_CAP = 128 * 1024;
constexpr int LNW = LBT / 64;
constexpr int PERM_MAX = 2 * NB;

// (value, row) of the first maximum over the wave, in every lane: the DPP
// reduction pattern of the wave sums (quad permutes, half-row and row
// mirrors, row broadcasts), masked-off lanes keeping their own pair.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ int dpp_keep(int v) {
    return __builtin_amdgcn_update_dpp(v, v, CTRL, ROWS, 0xF, false);
}
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ float dpp_keep(float v) {
    return __builtin_bit_cast(float, dpp_keep<CTRL, ROWS>(__builtin_bit_cast(int, v)));
}
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ double dpp_keep(double v) {
    const long long u = __builtin_bit_cast(long long, v);
    const int lo = dpp_keep<CTRL, ROWS>((int)u), hi = dpp_keep<CTRL, ROWS>((int)(u >> 32));
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float lane63(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
}
__device__ __forceinline__ double lane63(double v) {
    const long long u = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane((int)u, 63), hi = __builtin_amdgcn_readlane((int)(u >> 32), 63);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
template <int CTRL, int ROWS = 0xF, typename T>
__device__ __forceinline__ void argmax_step(T &best, int &bi) {
    const T v2 = dpp_keep<CTRL, ROWS>(best);
    const int i2 = dpp_keep<CTRL, ROWS>(bi);
    if (v2 > best || (v2 == best && i2 < bi)) {
        best = v2;
        bi = i2;
    }
}
template <typename T>
__device__ __forceinline__ void argmax_dpp(T &best, int &bi) {
    argmax_step<0xB1>(best, bi);
    argmax_step<0x4E>(best, bi);
    argmax_step<0x141>(best, bi);
    argmax_step<0x140>(best, bi);
    argmax_step<0x142, 0xA>(best, bi);
    argmax_step<0x143, 0xC>(best, bi);
    best = lane63(best);
    bi = __builtin_amdgcn_readlane(bi, 63);
}

template <typename T, int SB>
__global__ __launch_bounds__(LBT) void k_panel_lds(int n, int k, int jb, cx<T> *A, int lda,
                                                  size_t strideA, int *ipiv, int *info,
                                                  int *permall) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *Vs = reinterpret_cast<cx<T> *>(smem_raw);  // SB columns of M rows
    __shared__ T red_v[LNW];
    __shared__ int red_i[LNW];
    __shared__ cx<T> pinv_s;
    __shared__ int ipl[NB];  // the panel's pivot rows (to global ipiv after each sub-panel)
    __shared__ int zpiv;     // first zero pivot column + 1 of this panel, 0 if none
    const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    cx<T> *Ab = A + (size_t)b * strideA;
    int *ip = ipiv + (size_t)b * n;
    const int M = n - k;
    if (t == 0) zpiv = 0;
    for (int cs = k; cs < k + jb; cs += SB) {
        const int w = min(SB, k + jb - cs);
        {
            // all of a thread's loads in flight before its LDS stores (the
            // launcher sizes the LDS to at most LDS_CAP, so M <= LROWS * LBT)
            constexpr int LROWS = (int)((LDS_CAP / (sizeof(cx<T>) * SB) + LBT - 1) / LBT);
            cx<T> v[LROWS][SB];
#pragma unroll
            for (int q = 0; q < LROWS; ++q) {
                const int r = min(t + q * LBT, M - 1);
#pragma unroll
                for (int i = 0; i < SB; ++i) v[q][i] = Ab[(size_t)(cs + min(i, w - 1)) * lda + k + r];
            }
#pragma unroll
            for (int q = 0; q < LROWS; ++q) {
                const int r = t + q * LBT;
#pragma unroll
                for (int i = 0; i < SB; ++i)
                    if (i < w && r < M) Vs[(size_t)i * M + r] = v[q][i];
            }
        }
        __syncthreads();
        // pivot candidates of the sub-panel's first column; later columns'
        // candidates are gathered by the update pass before them
        T best = (T)-1;
        int bi = M;
        for (int r = cs - k + t; r < M; r += LBT) {
            const T v = cabs1(Vs[r]);
            if (v > best) {
                best = v;
                bi = r;
            }
        }
        for (int i = 0; i < w; ++i) {
            const int rc = cs + i - k;
            cx<T> *col = Vs + (size_t)i * M;
            // pivot: first maximum of |re| + |im| over rows rc..M-1
            argmax_dpp(best, bi);
            if (lane == 0) {
                red_v[wv] = best;
                red_i[wv] = bi;
            }
            __syncthreads();
            if (wv == 0) {
                // the waves' candidates (lanes >= LNW hold none), then lane q
                // swaps rows rc and bx of sub-panel column q
                T bv = lane < LNW ? red_v[lane] : (T)-1;
                int bx = lane < LNW ? red_i[lane] : M;
                argmax_dpp(bv, bx);
                if (bx != rc && lane < w) {
                    const cx<T> tmp = Vs[(size_t)lane * M + rc];
                    Vs[(size_t)lane * M + rc] = Vs[(size_t)lane * M + bx];
                    Vs[(size_t)lane * M + bx] = tmp;
                }
                if (lane == 0) {
                    ipl[cs - k + i] = k + bx;
                    if (bv == (T)0 && zpiv == 0) zpiv = cs + i + 1;
                }
                // lane i swapped column i (the pivot column); its new diagonal
                if (lane == i) {
                    const cx<T> d = col[rc];
                    pinv_s = (bv == (T)0) ? cx<T>{0, 0} : cdiv(cx<T>{1, 0}, d);
                }
            }
            __syncthreads();
            const cx<T> pinv = pinv_s;
            best = (T)-1;
            bi = M;
            cx<T> *nxt = col + M;  // column i + 1
            for (int r = rc + 1 + t; r < M; r += LBT) {
                const cx<T> l = col[r] * pinv;
                col[r] = l;
                for (int q = i + 1; q < w; ++q) {
                    cx<T> *cq = Vs + (size_t)q * M;
                    cq[r] = cq[r] - l * cq[rc];
                }
                if (i + 1 < w) {
                    const T v = cabs1(nxt[r]);
                    if (v > best) {
                        best = v;
                        bi = r;
                    }
                }
            }
            __syncthreads();
        }
        if (t < w) ip[cs + t] = ipl[cs - k + t];
        for (int i = 0; i < w; ++i)
            for (int r = t; r < M; r += LBT) Ab[(size_t)(cs + i) * lda + k + r] = Vs[(size_t)i * M + r];
        // this sub-panel's interchanges on the panel's other columns
        if (t < jb) {
            const int cc = k + t;
            if (cc < cs || cc >= cs + w) {
                cx<T> *Ac = Ab + (size_t)cc * lda;
                for (int i = 0; i < w; ++i) {
                    const int r1 = cs + i, r2 = ipl[cs - k + i];
                    if (r2 != r1) {
                        const cx<T> tmp = Ac[r1];
                        Ac[r1] = Ac[r2];
                        Ac[r2] = tmp;
                    }
                }
            }
        }
        __syncthreads();
        // later panel columns: U12 = L11^-1 a[cs..cs+w), a[cs+w..] -= L21 U12
        const int r0 = cs - k;
        for (int j = cs + w + wv; j < k + jb; j += LNW) {
            cx<T> *Aj = Ab + (size_t)j * lda + k;
            // tiny forward solve, redundantly in every lane (its w loads
            // issued together)
            cx<T> x[SB];
#pragma unroll
            for (int i = 0; i < SB; ++i) x[i] = i < w ? Aj[r0 + i] : cx<T>{0, 0};
#pragma unroll
            for (int i = 0; i < SB; ++i)
                if (i < w)
                    for (int l = 0; l < i; ++l) x[i] = x[i] - Vs[(size_t)l * M + r0 + i] * x[l];
            if (lane < w) {
#pragma unroll
                for (int i = 0; i < SB; ++i)
                    if (i == lane) Aj[r0 + i] = x[i];
            }
            // rank-w update, CH rows per lane at a time with their loads in flight together
            constexpr int CH = 8;
            for (int rb = r0 + w + lane; rb < M; rb += 64 * CH) {
                cx<T> a[CH];
#pragma unroll
                for (int q = 0; q < CH; ++q) {
                    const int r = rb + 64 * q;
                    a[q] = r < M ? Aj[r] : cx<T>{0, 0};
                }
#pragma unroll
                for (int q = 0; q < CH; ++q) {
                    const int r = rb + 64 * q;
                    if (r < M) {
#pragma unroll
                        for (int i = 0; i < SB; ++i)
                            if (i < w) a[q] = a[q] - Vs[(size_t)i * M + r] * x[i];
                        Aj[r] = a[q];
                    }
                }
            }
        }
        __syncthreads();
    }
    // the panel's interchanges as one row permutation (for the other
    // columns): the swaps replayed on a row map in LDS (the sub-panel buffer,
    // no longer needed), then every touched row lists itself once
    int *pmap = reinterpret_cast<int *>(smem_raw);  // pmap[r - k], rows k..n-1
    __shared__ int pcnt;
    if (t < jb) {
        pmap[t] = k + t;
        pmap[ipl[t] - k] = ipl[t];
    }
    if (t == 0) pcnt = 0;
    __syncthreads();
    if (t == 0)
        for (int c = 0; c < jb; ++c) {
            const int p = ipl[c] - k;
            if (p != c) {
                const int tmp = pmap[c];
                pmap[c] = pmap[p];
                pmap[p] = tmp;
            }
        }
    __syncthreads();
    int *pm = permall + (size_t)b * (1 + 2 * PERM_MAX);
    if (t < 2 * jb) {
        // candidate rows: k + t (t < jb), or the pivot row ipl[t - jb] when it
        // lies outside the panel rows and is its first occurrence
        int r = -1;
        if (t < jb) {
            r = k + t;
        } else {
            const int q = t - jb, p = ipl[q];
            bool first = p >= k + jb;
            for (int q2 = 0; q2 < q && first; ++q2)
                if (ipl[q2] == p) first = false;
            if (first) r = p;
        }
        if (r >= 0 && pmap[r - k] != r) {
            const int slot = atomicAdd(&pcnt, 1);
            pm[1 + slot] = r;
            pm[1 + PERM_MAX + slot] = pmap[r - k];
        }
    }
    __syncthreads();
    if (t == 0) {
        pm[0] = pcnt;
        if (zpiv && info && info[b] == 0) info[b] = zpiv;
    }
}

// new[dst_q] = old[src_q] for the listed rows, in every column of [c0, n)
// outside [k, k+jb) (the columns left of c0 take their interchanges later,
// in k_laswp_deferred); one wave per column (all loads before the stores).
template <typename T>
__global__ __launch_bounds__(256) void k_laswp_perm(int n, int k, int jb, int c0, cx<T> *A,
                                                   int lda, size_t strideA, const int *permall) {
    const int b = blockIdx.y;
    const int lane = threadIdx.x & 63;
    int col = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (col >= n - jb - c0) return;
    col += c0;
    if (col >= k) col += jb;
    const int *pm = permall + (size_t)b * (1 + 2 * PERM_MAX);
    const int m = pm[0];
    if (lane >= m) return;
    cx<T> *Ac = A + (size_t)b * strideA + (size_t)col * lda;
    const cx<T> v = Ac[pm[1 + PERM_MAX + lane]];
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    Ac[pm[1 + lane]] = v;
}

// apply interchanges ipiv[k..k+jb) to the columns of [c0, n) outside [k, k+jb)
template <typename T>
__global__ void k_laswp(int n, int k, int jb, int c0, cx<T> *A, int lda, size_t strideA,
                        const int *ipiv) {
    const int b = blockIdx.y;
    int col = blockIdx.x * blockDim.x + threadIdx.x;
    int nout = n - jb - c0;
    if (col >= nout) return;
    col += c0;
    if (col >= k) col += jb;
    cx<T> *Ac = A + (size_t)b * strideA + (size_t)col * lda;
    const int *ip = ipiv + (size_t)b * n;
    for (int c = k; c < k + jb; ++c) {
        int p = ip[c];
        if (p != c) {
            cx<T> tmp = Ac[c];
            Ac[c] = Ac[p];
            Ac[p] = tmp;
        }
    }
}

// Deferred interchanges of the columns left of each outer panel.  For the
// outer-panel boundary a = (J + 1) NBO, idx_a gathers rows a..n-1 after the
// interchanges a..n-1 in order: new[r] = old[idx_a[r]].  Going down from
// a = n (identity), idx_a[r] = sigma_a(idx_{a+1}[r]) with sigma_a the
// transposition (a, ipiv[a]), so every thread composes its own rows with
// no exchange between threads.  One workgroup per matrix; idxall holds
// (n / NBO) * n ints per matrix, boundary J at J * n.
constexpr int DEF_ROWS = 16;  // rows per thread: deferral for n <= 256 * 16
__global__ __launch_bounds__(256) void k_perm_suffix(int n, const int *ipiv, int *idxall) {
    extern __shared__ int ips[];
    const int b = blockIdx.x, t = threadIdx.x;
    const int *ip = ipiv + (size_t)b * n;
    for (int r = t; r < n; r += 256) ips[r] = ip[r];
    __syncthreads();
    int v[DEF_ROWS];
#pragma unroll
    for (int q = 0; q < DEF_ROWS; ++q) v[q] = t + 256 * q;
    int *idx = idxall + (size_t)b * (n / NBO) * n;
    const int amin = NBO;
    for (int a = n - 1; a >= amin; --a) {
        const int pa = ips[a];
#pragma unroll
        for (int q = 0; q < DEF_ROWS; ++q) v[q] = v[q] == a ? pa : v[q] == pa ? a : v[q];
        if (a % NBO == 0) {
            int *ia = idx + (size_t)(a / NBO - 1) * n;
#pragma unroll
            for (int q = 0; q < DEF_ROWS; ++q) {
                const int r = t + 256 * q;
                if (r >= a && r < n) ia[r] = v[q];
            }
        }
    }
}

// columns [J NBO, (J + 1) NBO) with a = (J + 1) NBO < n: rows a..n-1
// gathered through idx_a; every thread reads its rows' sources before the
// barrier and writes after it.  Grid (columns / 8, batch), 8 columns per
// workgroup, all in one outer panel.
template <typename T>
__global__ __launch_bounds__(256) void k_laswp_deferred(int n, int ncols, const int *idxall,
                                                        cx<T> *A, int lda, size_t strideA) {
    const int b = blockIdx.y, t = threadIdx.x;
    const int c0 = blockIdx.x * 8;
    if (c0 >= ncols) return;
    const int J = c0 / NBO, a = (J + 1) * NBO;
    const int *ia = idxall + (size_t)b * (n / NBO) * n + (size_t)J * n;
    int src[DEF_ROWS];
#pragma unroll
    for (int q = 0; q < DEF_ROWS; ++q) {
        const int r = t + 256 * q;
        src[q] = (r >= a && r < n) ? ia[r] : -1;
    }
    cx<T> *Ab = A + (size_t)b * strideA;
    for (int c = c0; c < min(c0 + 8, ncols); ++c) {
        cx<T> *col = Ab + (size_t)c * lda;
        cx<T> val[DEF_ROWS];
#pragma unroll
        for (int q = 0; q < DEF_ROWS; ++q) val[q] = src[q] >= 0 ? col[src[q]] : cx<T>{0, 0};
        __syncthreads();
#pragma unroll
        for (int q = 0; q < DEF_ROWS; ++q)
            if (src[q] >= 0) col[t + 256 * q] = val[q];
    }
}

// x <- op(block)^-1 x for one right-hand side held in registers; JB > 0
// fixes the block size at compile time, JB == 0 takes it from jb
template <typename T, int MODE, int JB>
__device__ __forceinline__ void trsm_diag_solve(cx<T> (&x)[NB], const cx<T> (&Ts)[NB][NB + 1],
                                                const cx<T> (&dinv)[NB], int jb = NB) {
    const int n = JB > 0 ? JB : jb;
    if (MODE == 0 || MODE == 1) {
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            if (i < n) {
                cx<T> sacc = x[i];
#pragma unroll
                for (int j = 0; j < NB; ++j)
                    if (j < i) sacc = sacc - Ts[i][j] * x[j];
                x[i] = (MODE == 0) ? sacc : sacc * dinv[i];
            }
        }
    } else {
#pragma unroll
        for (int i = NB - 1; i >= 0; --i) {
            if (i < n) {
                cx<T> sacc = x[i];
#pragma unroll
                for (int j = 0; j < NB; ++j)
                    if (j > i && j < n) sacc = sacc - Ts[i][j] * x[j];
                x[i] = (MODE == 3) ? sacc : sacc * dinv[i];
            }
        }
    }
}

// Diagonal-block triangular solve, one thread per right-hand-side column.
//   MODE 0: unit lower L (from A's lower part)        forward
//   MODE 1: U^H, U non-unit upper (A's upper part)     forward
//   MODE 2: non-unit upper U                           backward
//   MODE 3: L^H, L unit lower                          backward
// The block is rows/cols [i0, i0+jb) of A; B rows [i0, i0+jb).
// The right-hand-side tile is staged through
// LDS: the block's rows of NC columns are loaded and stored with coalesced
// accesses (each column's jb rows are contiguous), every thread solves its
// column from LDS, and the non-unit diagonal is inverted once per block.
template <typename T, int MODE, int NC>
__global__ __launch_bounds__(NC) void k_trsm_diag_lds(int n, int i0, int jb, const cx<T> *A,
                                                    int lda, size_t strideA, cx<T> *B, int ldb,
                                                    size_t strideB, int nrhs) {
    __shared__ cx<T> Ts[NB][NB + 1];
    __shared__ cx<T> dinv[NB];
    __shared__ cx<T> tile[NB][NC + 1];
    const int b = blockIdx.y, t = threadIdx.x;
    const int col0 = blockIdx.x * NC;
    const cx<T> *Ab = A + (size_t)b * strideA;
    cx<T> *Bb = B + (size_t)b * strideB + i0;
    // every load of the block and of the tile is issued before the first
    // LDS store (clamped addresses, out-of-range values dropped), so the
    // staging costs one memory latency rather than one per element
    static_assert((NB * NB) % NC == 0 && (NB * NC) % NC == 0, "whole rounds");
    constexpr int LA = NB * NB / NC;
    cx<T> va[LA], vb[NB];
#pragma unroll
    for (int m = 0; m < LA; ++m) {
        const int e = t + m * NC, r = min(e % NB, jb - 1), c = min(e / NB, jb - 1);
        va[m] = Ab[(size_t)(i0 + c) * lda + (i0 + r)];  // A[r][c]
    }
#pragma unroll
    for (int m = 0; m < NB; ++m) {
        const int e = t + m * NC, i = min(e % NB, jb - 1), c = min(col0 + e / NB, nrhs - 1);
        vb[m] = Bb[(size_t)c * ldb + i];
    }
#pragma unroll
    for (int m = 0; m < LA; ++m) {
        const int e = t + m * NC, r = e % NB, c = e / NB;
        if (r < jb && c < jb) {
            if (MODE == 0 || MODE == 2)
                Ts[r][c] = va[m];
            else
                Ts[c][r] = conj(va[m]);
        }
    }
#pragma unroll
    for (int m = 0; m < NB; ++m) {
        const int e = t + m * NC, i = e % NB, c = e / NB;
        if (i < jb && col0 + c < nrhs) tile[i][c] = vb[m];
    }
    __syncthreads();
    if ((MODE == 1 || MODE == 2) && t < jb) dinv[t] = cdiv(cx<T>{1, 0}, Ts[t][t]);
    __syncthreads();
    cx<T> x[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) x[i] = i < jb ? tile[i][t] : cx<T>{0, 0};
    // full blocks take a branch-free copy of the solve, whose LDS reads the
    // compiler can issue ahead of the dependent products
    if (jb == NB)
        trsm_diag_solve<T, MODE, NB>(x, Ts, dinv);
    else
        trsm_diag_solve<T, MODE, 0>(x, Ts, dinv, jb);
#pragma unroll
    for (int i = 0; i < NB; ++i)
        if (i < jb) tile[i][t] = x[i];
    __syncthreads();
    for (int e = t; e < NB * NC; e += NC) {
        const int i = e % NB, c = e / NB;
        if (i < jb && col0 + c < nrhs) Bb[(size_t)(col0 + c) * ldb + i] = tile[i][c];
    }
}

// Direct form (complex128): one thread per column reading its jb rows
// straight from memory (all loads issued before the solve), for the
// occupancy that the LDS tile would cost at 16 bytes per element.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_trsm_diag_direct(int n, int i0, int jb, const cx<T> *A,
                                                         int lda, size_t strideA, cx<T> *B,
                                                         int ldb, size_t strideB, int nrhs) {
    __shared__ cx<T> Ts[NB][NB + 1];
    __shared__ cx<T> dinv[NB];
    const int b = blockIdx.y, t = threadIdx.x;
    const cx<T> *Ab = A + (size_t)b * strideA;
    for (int e = t; e < jb * jb; e += blockDim.x) {
        const int r = e % jb, c = e / jb;
        const cx<T> v = Ab[(size_t)(i0 + c) * lda + (i0 + r)];
        if (MODE == 0 || MODE == 2)
            Ts[r][c] = v;
        else
            Ts[c][r] = conj(v);
    }
    __syncthreads();
    if ((MODE == 1 || MODE == 2) && t < jb) dinv[t] = cdiv(cx<T>{1, 0}, Ts[t][t]);
    __syncthreads();
    const int col = blockIdx.x * blockDim.x + t;
    if (col >= nrhs) return;
    cx<T> *Bc = B + (size_t)b * strideB + (size_t)col * ldb + i0;
    cx<T> x[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) x[i] = Bc[i < jb ? i : 0];
#pragma unroll
    for (int i = 0; i < NB; ++i)
        if (i >= jb) x[i] = cx<T>{0, 0};
    if (MODE == 0 || MODE == 1) {
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            if (i < jb) {
                cx<T> sacc = x[i];
#pragma unroll
                for (int j = 0; j < NB; ++j)
                    if (j < i) sacc = sacc - Ts[i][j] * x[j];
                x[i] = (MODE == 0) ? sacc : sacc * dinv[i];
            }
        }
    } else {
#pragma unroll
        for (int i = NB - 1; i >= 0; --i) {
            if (i < jb) {
                cx<T> sacc = x[i];
#pragma unroll
                for (int j = 0; j < NB; ++j)
                    if (j > i && j < jb) sacc = sacc - Ts[i][j] * x[j];
                x[i] = (MODE == 3) ? sacc : sacc * dinv[i];
            }
        }
    }
#pragma unroll
    for (int i = 0; i < NB; ++i)
        if (i < jb) Bc[i] = x[i];
}

template <typename T, int MODE>
void launch_trsm_diag(int n, int i0, int jb, const cx<T> *A, int lda, size_t sA, cx<T> *B,
                      int ldb, size_t sB, int nrhs, int batch, hipStream_t s) {
    if constexpr (sizeof(T) == sizeof(float)) {
        constexpr int NC = 256;
        dim3 grid((nrhs + NC - 1) / NC, batch);
        hipLaunchKernelGGL((k_trsm_diag_lds<T, MODE, NC>), grid, dim3(NC), 0, s, n, i0, jb, A,
                           lda, sA, B, ldb, sB, nrhs);
    } else {
        dim3 grid((nrhs + 255) / 256, batch);
        hipLaunchKernelGGL((k_trsm_diag_direct<T, MODE>), grid, dim3(256), 0, s, n, i0, jb, A,
                           lda, sA, B, ldb, sB, nrhs);
    }
}

// permutation index from interchanges: perm applied as (P B)[r] = B[idx[r]];
// one workgroup per matrix, the interchanges replayed in order in LDS
__global__ __launch_bounds__(256) void k_perm_index(int n, const int *ipiv, int *idx, int *inv) {
    extern __shared__ int pl[];  // ipiv (n), index (n)
    int *ps = pl, *xs = pl + n;
    const int b = blockIdx.x, t = threadIdx.x;
    const int *ip = ipiv + (size_t)b * n;
    int *ix = idx + (size_t)b * n;
    int *iv = inv + (size_t)b * n;
    for (int r = t; r < n; r += 256) {
        ps[r] = ip[r];
        xs[r] = r;
    }
    __syncthreads();
    if (t == 0)
        for (int c = 0; c < n; ++c) {
            const int p = ps[c];
            if (p != c) {
                const int tmp = xs[c];
                xs[c] = xs[p];
                xs[p] = tmp;
            }
        }
    __syncthreads();
    for (int r = t; r < n; r += 256) {
        const int v = xs[r];
        ix[r] = v;
        iv[v] = r;
    }
}

// the same for orders whose index lists exceed LDS: one thread per matrix
__global__ void k_perm_index_global(int n, const int *ipiv, int *idx, int *inv, int batch) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    const int *ip = ipiv + (size_t)b * n;
    int *ix = idx + (size_t)b * n;
    int *iv = inv + (size_t)b * n;
    for (int r = 0; r < n; ++r) ix[r] = r;
    for (int c = 0; c < n; ++c) {
        int p = ip[c];
        if (p != c) {
            int tmp = ix[c];
            ix[c] = ix[p];
            ix[p] = tmp;
        }
    }
    for (int r = 0; r < n; ++r) iv[ix[r]] = r;
}

// dst[r, c] = src[perm[r], c]
template <typename T>
__global__ void k_row_gather(int n, int nrhs, const int *perm, const cx<T> *src, int lds,
                             size_t strideS, cx<T> *dst, int ldd, size_t strideD) {
    const int b = blockIdx.z;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (r >= n) return;
    const int *pb = perm + (size_t)b * n;
    dst[(size_t)b * strideD + (size_t)c * ldd + r] = src[(size_t)b * strideS + (size_t)c * lds + pb[r]];
}

template <typename T>
__global__ void k_copy_cols(int n, int nrhs, const cx<T> *src, int lds, size_t strideS, cx<T> *dst,
                            int ldd, size_t strideD) {
    const int b = blockIdx.z;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (r >= n) return;
    dst[(size_t)b * strideD + (size_t)c * ldd + r] = src[(size_t)b * strideS + (size_t)c * lds + r];
}

// Two-level blocking: diagonal blocks of NBO columns are solved with NB-wide
// diagonal solves and rank-NB updates confined to the block, then the rest of
// B gets one rank-NBO GEMM update per block (four times fewer passes over B
// than rank-NB updates).
// Left-looking outer blocks (IA_TRSM_RIGHT unset): block O first takes the
// update from every solved row at once, B_O -= A[O, done] X_done, one
// product with K = |done| that reads each block of B once, instead of a
// rank-NBO update of all remaining rows after every block.
inline bool trsm_left_looking() {
    const char *e = std::getenv("IA_TRSM_RIGHT");
    return !(e && *e && *e != '0');
}

template <typename T, int MODE>
void trsm_blocked(int n, int nrhs, const cx<T> *A, int lda, size_t sA, cx<T> *B, int ldb,
                  size_t sB, int batch, hipStream_t s) {
    const cx<T> mone{-1, 0}, one{1, 0};
    const bool forward = (MODE == 0 || MODE == 1);
    const bool left = trsm_left_looking();
    const int nob = (n + NBO - 1) / NBO;
    for (int ob = 0; ob < nob; ++ob) {
        const int blkO = forward ? ob : (nob - 1 - ob);
        const int o0 = blkO * NBO, ojb = min(NBO, n - o0), o1 = o0 + ojb;
        if (left) {
            if (forward && o0 > 0) {
                if (MODE == 0)  // B_O -= L[O, :o0] X[:o0]
                    launch_gemm<T>('N', 'N', ojb, nrhs, o0, mone, A + o0, lda, sA, B, ldb, sB,
                                   one, B + o0, ldb, sB, batch, s);
                else  // B_O -= (U[:o0, O])^H X[:o0]
                    launch_gemm<T>('C', 'N', ojb, nrhs, o0, mone, A + (size_t)o0 * lda, lda, sA,
                                   B, ldb, sB, one, B + o0, ldb, sB, batch, s);
            } else if (!forward && o1 < n) {
                if (MODE == 2)  // B_O -= U[O, o1:] X[o1:]
                    launch_gemm<T>('N', 'N', ojb, nrhs, n - o1, mone, A + (size_t)o1 * lda + o0,
                                   lda, sA, B + o1, ldb, sB, one, B + o0, ldb, sB, batch, s);
                else  // B_O -= (L[o1:, O])^H X[o1:]
                    launch_gemm<T>('C', 'N', ojb, nrhs, n - o1, mone, A + (size_t)o0 * lda + o1,
                                   lda, sA, B + o1, ldb, sB, one, B + o0, ldb, sB, batch, s);
            }
        }
        const int nblk = (ojb + NB - 1) / NB;
        for (int bi = 0; bi < nblk; ++bi) {
            const int blk = forward ? bi : (nblk - 1 - bi);
            const int i0 = o0 + blk * NB, jb = min(NB, o1 - i0);
            launch_trsm_diag<T, MODE>(n, i0, jb, A, lda, sA, B, ldb, sB, nrhs, batch, s);
            if (forward) {
                const int rest = o1 - i0 - jb;
                if (rest > 0) {
                    if (MODE == 0)  // B[i0+jb:o1] -= L[i0+jb:o1, I] X_I
                        launch_gemm<T>('N', 'N', rest, nrhs, jb, mone,
                                       A + (size_t)i0 * lda + i0 + jb, lda, sA, B + i0, ldb, sB,
                                       one, B + i0 + jb, ldb, sB, batch, s);
                    else  // B[i0+jb:o1] -= (U[I, i0+jb:o1])^H X_I
                        launch_gemm<T>('C', 'N', rest, nrhs, jb, mone,
                                       A + (size_t)(i0 + jb) * lda + i0, lda, sA, B + i0, ldb, sB,
                                       one, B + i0 + jb, ldb, sB, batch, s);
                }
            } else {
                const int rest = i0 - o0;
                if (rest > 0) {
                    if (MODE == 2)  // B[o0:i0] -= U[o0:i0, I] X_I
                        launch_gemm<T>('N', 'N', rest, nrhs, jb, mone, A + (size_t)i0 * lda + o0,
                                       lda, sA, B + i0, ldb, sB, one, B + o0, ldb, sB, batch, s);
                    else  // B[o0:i0] -= (L[I, o0:i0])^H X_I
                        launch_gemm<T>('C', 'N', rest, nrhs, jb, mone, A + (size_t)o0 * lda + i0,
                                       lda, sA, B + i0, ldb, sB, one, B + o0, ldb, sB, batch, s);
                }
            }
        }
        if (left) continue;
        if (forward) {
            const int rest = n - o1;
            if (rest > 0) {
                if (MODE == 0)  // B[o1:] -= L[o1:, O] X_O
                    launch_gemm<T>('N', 'N', rest, nrhs, ojb, mone, A + (size_t)o0 * lda + o1, lda,
                                   sA, B + o0, ldb, sB, one, B + o1, ldb, sB, batch, s);
                else  // B[o1:] -= (U[O, o1:])^H X_O
                    launch_gemm<T>('C', 'N', rest, nrhs, ojb, mone, A + (size_t)o1 * lda + o0, lda,
                                   sA, B + o0, ldb, sB, one, B + o1, ldb, sB, batch, s);
            }
        } else if (o0 > 0) {
            if (MODE == 2)  // B[:o0] -= U[:o0, O] X_O
                launch_gemm<T>('N', 'N', o0, nrhs, ojb, mone, A + (size_t)o0 * lda, lda, sA,
                               B + o0, ldb, sB, one, B, ldb, sB, batch, s);
            else  // B[:o0] -= (L[O, :o0])^H X_O
                launch_gemm<T>('C', 'N', o0, nrhs, ojb, mone, A + o0, lda, sA, B + o0, ldb, sB,
                               one, B, ldb, sB, batch, s);
        }
    }
}

// wave sum (every lane active) over the DPP lane network; result in every lane
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ int dpp_zero(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xF, false);
}
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ float dppz_f(float v) {
    return __builtin_bit_cast(float, dpp_zero<CTRL, ROWS>(__builtin_bit_cast(int, v)));
}
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ double dppz_d(double v) {
    const long long u = __builtin_bit_cast(long long, v);
    const int lo = dpp_zero<CTRL, ROWS>((int)u), hi = dpp_zero<CTRL, ROWS>((int)(u >> 32));
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
template <typename R>
__device__ __forceinline__ R hh_wave_sum(R v) {
    if constexpr (sizeof(R) == 4) {
        v += dppz_f<0xB1>(v);
        v += dppz_f<0x4E>(v);
        v += dppz_f<0x141>(v);
        v += dppz_f<0x140>(v);
        v += dppz_f<0x142, 0xA>(v);
        v += dppz_f<0x143, 0xC>(v);
    } else {
        v += dppz_d<0xB1>(v);
        v += dppz_d<0x4E>(v);
        v += dppz_d<0x141>(v);
        v += dppz_d<0x140>(v);
        v += dppz_d<0x142, 0xA>(v);
        v += dppz_d<0x143, 0xC>(v);
    }
    return lane63(v);
}

// One right-hand side, op 'N' (LAPACK getrs): b <- U^-1 L^-1 P b in a single
// launch, one workgroup per matrix with b resident in LDS throughout.  The
// interchanges are applied in order by one thread in LDS; each triangle is
// walked in diagonal blocks of G1 rows: the block is staged in LDS and solved
// by one wave (lane l owns x_l, the solved entry is broadcast by a lane
// shuffle), then every thread updates its rows of the rest of b with the
// block's columns (x_r -= A[r, blk] x_blk), so each element of L and U is
// read from memory once.
constexpr int G1 = 64;
constexpr int G1T = 1024;

// broadcast from a wave-uniform lane through the scalar unit (no LDS)
__device__ __forceinline__ float bcast_r(float v, int src) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), src));
}
__device__ __forceinline__ double bcast_r(double v, int src) {
    const long long u = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane((int)u, src), hi = __builtin_amdgcn_readlane((int)(u >> 32), src);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
template <typename T>
__device__ __forceinline__ cx<T> bcast_cx(cx<T> v, int src) {
    return cx<T>{bcast_r(v.re, src), bcast_r(v.im, src)};
}

template <typename T>
__global__ __launch_bounds__(G1T) void k_getrs1(int n, const cx<T> *__restrict__ Aall, int lda,
                                               size_t sA, const int *__restrict__ ipivall,
                                               cx<T> *__restrict__ Ball, size_t sB) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *xs = reinterpret_cast<cx<T> *>(smem_raw);  // n
    cx<T> *Ds = xs + n;                                // G1 x G1 block, column-major, ld G1 + 1
    int *ps = reinterpret_cast<int *>(Ds + (size_t)G1 * (G1 + 1));  // n pivots
    constexpr int LD = G1 + 1;
    const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const cx<T> *A = Aall + (size_t)b * sA;
    cx<T> *B = Ball + (size_t)b * sB;
    const int *ip = ipivall + (size_t)b * n;
    for (int r = t; r < n; r += G1T) {
        xs[r] = B[r];
        ps[r] = ip[r];
    }
    __syncthreads();
    if (t == 0)
        for (int c = 0; c < n; ++c) {
            const int p = ps[c];
            if (p != c) {
                const cx<T> tmp = xs[c];
                xs[c] = xs[p];
                xs[p] = tmp;
            }
        }
    __syncthreads();
    // ---- L (unit lower), top to bottom
    for (int i0 = 0; i0 < n; i0 += G1) {
        const int jb = min(G1, n - i0);
        for (int e = t; e < jb * jb; e += G1T) {
            const int l = e % jb, j = e / jb;
            Ds[j * LD + l] = A[(size_t)(i0 + j) * lda + i0 + l];
        }
        __syncthreads();
        if (wv == 0) {
            cx<T> x = lane < jb ? xs[i0 + lane] : cx<T>{0, 0};
            for (int j = 0; j < jb; ++j) {
                const cx<T> xj = bcast_cx(x, j);
                if (lane > j && lane < jb) x = x - Ds[j * LD + lane] * xj;
            }
            if (lane < jb) xs[i0 + lane] = x;
        }
        __syncthreads();
        for (int r = i0 + jb + t; r < n; r += G1T) {
            cx<T> acc{0, 0};
#pragma unroll 16
            for (int j = 0; j < jb; ++j) acc = acc + A[(size_t)(i0 + j) * lda + r] * xs[i0 + j];
            xs[r] = xs[r] - acc;
        }
        __syncthreads();
    }
    // ---- U (upper, non-unit), bottom to top
    const int nblk = (n + G1 - 1) / G1;
    for (int bk = nblk - 1; bk >= 0; --bk) {
        const int i0 = bk * G1, jb = min(G1, n - i0);
        for (int e = t; e < jb * jb; e += G1T) {
            const int l = e % jb, j = e / jb;
            Ds[j * LD + l] = A[(size_t)(i0 + j) * lda + i0 + l];
        }
        __syncthreads();
        if (wv == 0) {
            cx<T> x = lane < jb ? xs[i0 + lane] : cx<T>{0, 0};
            const cx<T> dl = lane < jb ? cdiv(cx<T>{1, 0}, Ds[lane * LD + lane]) : cx<T>{0, 0};
            for (int j = jb - 1; j >= 0; --j) {
                if (lane == j) x = x * dl;
                const cx<T> xj = bcast_cx(x, j);
                if (lane < j) x = x - Ds[j * LD + lane] * xj;
            }
            if (lane < jb) xs[i0 + lane] = x;
        }
        __syncthreads();
        for (int r = t; r < i0; r += G1T) {
            cx<T> acc{0, 0};
#pragma unroll 16
            for (int j = 0; j < jb; ++j) acc = acc + A[(size_t)(i0 + j) * lda + r] * xs[i0 + j];
            xs[r] = xs[r] - acc;
        }
        __syncthreads();
    }
    for (int r = t; r < n; r += G1T) B[r] = xs[r];
}

// op 'C', one right-hand side: x <- D^-1 b for A = P D^H factored as L U,
// D = U^H L^H P: forward with U^H (lower, non-unit), backward with L^H
// (upper, unit), then the interchanges in reverse order.  Row r of U^H is
// column r of U (contiguous), so the update of the rest of b after a block
// gives each row to one wave: lanes read the block's 64 entries of that
// column in one coalesced access and a DPP sum forms x_r's correction.
template <typename T>
__global__ __launch_bounds__(G1T) void k_getrs1c(int n, const cx<T> *__restrict__ Aall, int lda,
                                                size_t sA, const int *__restrict__ ipivall,
                                                cx<T> *__restrict__ Ball, size_t sB) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *xs = reinterpret_cast<cx<T> *>(smem_raw);  // n
    cx<T> *Ds = xs + n;                                // op(A) block: Ds[j LD + l] = op[l][j]
    int *ps = reinterpret_cast<int *>(Ds + (size_t)G1 * (G1 + 1));  // n pivots
    constexpr int LD = G1 + 1, NWV = G1T / 64;
    const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const cx<T> *A = Aall + (size_t)b * sA;
    cx<T> *B = Ball + (size_t)b * sB;
    const int *ip = ipivall + (size_t)b * n;
    for (int r = t; r < n; r += G1T) {
        xs[r] = B[r];
        ps[r] = ip[r];
    }
    // op(A)[i0 + l][i0 + j] = conj(A[i0 + j][i0 + l]), read along j
    auto stage = [&](int i0, int jb) {
        for (int e = t; e < jb * jb; e += G1T) {
            const int j = e % jb, l = e / jb;
            Ds[j * LD + l] = conj(A[(size_t)(i0 + l) * lda + i0 + j]);
        }
    };
    // rows [r0, r1) of b lose op(A)[r, blk] x_blk; each wave takes RB rows
    // at a time, their loads issued together
    auto update = [&](int i0, int jb, int r0, int r1) {
        constexpr int RB = 8;
        const cx<T> xl = lane < jb ? xs[i0 + lane] : cx<T>{0, 0};
        for (int rb = r0 + wv; rb < r1; rb += NWV * RB) {
            cx<T> a[RB];
#pragma unroll
            for (int q = 0; q < RB; ++q) {
                const int r = rb + q * NWV;
                a[q] = (r < r1 && lane < jb) ? A[(size_t)r * lda + i0 + lane] : cx<T>{0, 0};
            }
#pragma unroll
            for (int q = 0; q < RB; ++q) {
                const int r = rb + q * NWV;
                const cx<T> v = cmulc(a[q], xl);
                const cx<T> d{hh_wave_sum(v.re), hh_wave_sum(v.im)};
                if (lane == 0 && r < r1) xs[r] = xs[r] - d;
            }
        }
    };
    __syncthreads();
    // ---- U^H (lower, non-unit), top to bottom
    for (int i0 = 0; i0 < n; i0 += G1) {
        const int jb = min(G1, n - i0);
        stage(i0, jb);
        __syncthreads();
        if (wv == 0) {
            cx<T> x = lane < jb ? xs[i0 + lane] : cx<T>{0, 0};
            const cx<T> dl = lane < jb ? cdiv(cx<T>{1, 0}, Ds[lane * LD + lane]) : cx<T>{0, 0};
            for (int j = 0; j < jb; ++j) {
                if (lane == j) x = x * dl;
                const cx<T> xj = bcast_cx(x, j);
                if (lane > j && lane < jb) x = x - Ds[j * LD + lane] * xj;
            }
            if (lane < jb) xs[i0 + lane] = x;
        }
        __syncthreads();
        update(i0, jb, i0 + jb, n);
        __syncthreads();
    }
    // ---- L^H (upper, unit), bottom to top
    const int nblk = (n + G1 - 1) / G1;
    for (int bk = nblk - 1; bk >= 0; --bk) {
        const int i0 = bk * G1, jb = min(G1, n - i0);
        stage(i0, jb);
        __syncthreads();
        if (wv == 0) {
            cx<T> x = lane < jb ? xs[i0 + lane] : cx<T>{0, 0};
            for (int j = jb - 1; j >= 0; --j) {
                const cx<T> xj = bcast_cx(x, j);
                if (lane < j) x = x - Ds[j * LD + lane] * xj;
            }
            if (lane < jb) xs[i0 + lane] = x;
        }
        __syncthreads();
        update(i0, jb, 0, i0);
        __syncthreads();
    }
    // x = P^T z: the interchanges undone in reverse order
    if (t == 0)
        for (int c = n - 1; c >= 0; --c) {
            const int p = ps[c];
            if (p != c) {
                const cx<T> tmp = xs[c];
                xs[c] = xs[p];
                xs[p] = tmp;
            }
        }
    __syncthreads();
    for (int r = t; r < n; r += G1T) B[r] = xs[r];
}

template <typename T>
size_t getrs1_lds(int n) {
    return sizeof(cx<T>) * ((size_t)n + (size_t)G1 * (G1 + 1)) + sizeof(int) * (size_t)n;
}

}  // namespace

size_t getrf_work_ints(int n, int batch) {
    return ((size_t)(1 + 2 * PERM_MAX) + (size_t)(n / NBO) * n) * (size_t)batch;
}

template <typename T>
ia_status getrf_batched(int n, cx<T> *A, int lda, size_t strideA, int *ipiv, int *info, int batch,
                        hipStream_t s, int *work) {
    if (n <= 0 || batch <= 0) return IA_OK;
    const cx<T> mone{-1, 0}, one{1, 0};
    int *perm = nullptr;  // per matrix: the current panel's row permutation
    if (work) {
        perm = work;
    } else if (hipMallocAsync((void **)&perm, sizeof(int) * (1 + 2 * PERM_MAX) * (size_t)batch,
                              s) != hipSuccess) {
        (void)hipGetLastError();
        perm = nullptr;
    }
    // panels of NBO columns, each factored as NB-wide sub-panels whose
    // updates stay inside the panel; then U12 = L11^-1 A12 and one rank-NBO
    // update of the trailing matrix per panel
    // interchanges of columns left of the current outer panel are deferred
    // to one gather per column at the end (n <= 256 DEF_ROWS)
    const bool defer = n > NBO && n <= 256 * DEF_ROWS;
    for (int p0 = 0; p0 < n; p0 += NBO) {
        const int pend = min(n, p0 + NBO);
        const int c0 = defer ? p0 : 0;
        for (int k = p0; k < pend; k += NB) {
            const int jb = min(NB, pend - k);
            const size_t colb = sizeof(cx<T>) * (size_t)(n - k);
            const int sb = colb * 8 <= LDS_CAP ? 8 : colb * 4 <= LDS_CAP ? 4 : colb * 2 <= LDS_CAP ? 2 : 0;
            if (sb > 0 && perm) {
                auto go = [&](auto kern) {
                    (void)hipFuncSetAttribute((const void *)kern,
                                              hipFuncAttributeMaxDynamicSharedMemorySize,
                                              (int)(colb * sb));
                    hipLaunchKernelGGL(kern, dim3(batch), dim3(LBT), colb * sb, s, n, k, jb, A, lda,
                                       strideA, ipiv, info, perm);
                };
                if (sb == 8)
                    go(k_panel_lds<T, 8>);
                else if (sb == 4)
                    go(k_panel_lds<T, 4>);
                else
                    go(k_panel_lds<T, 2>);
                if (n - jb - c0 > 0) {
                    dim3 g((n - jb - c0 + 3) / 4, batch);
                    hipLaunchKernelGGL(k_laswp_perm<T>, g, dim3(256), 0, s, n, k, jb, c0, A, lda,
                                       strideA, (const int *)perm);
                }
            } else {
                hipLaunchKernelGGL(k_panel<T>, dim3(batch), dim3(512), 0, s, n, k, jb, A, lda,
                                   strideA, ipiv, info);
                if (n - jb - c0 > 0) {
                    dim3 g((n - jb - c0 + 255) / 256, batch);
                    hipLaunchKernelGGL(k_laswp<T>, g, dim3(256), 0, s, n, k, jb, c0, A, lda,
                                       strideA, ipiv);
                }
            }
            const int rin = pend - k - jb;
            if (rin > 0) {
                launch_trsm_diag<T, 0>(n, k, jb, (const cx<T> *)A, lda, strideA,
                                       A + (size_t)(k + jb) * lda, lda, strideA, rin, batch, s);
                launch_gemm<T>('N', 'N', n - k - jb, rin, jb, mone, A + (size_t)k * lda + k + jb,
                               lda, strideA, A + (size_t)(k + jb) * lda + k, lda, strideA, one,
                               A + (size_t)(k + jb) * lda + k + jb, lda, strideA, batch, s);
            }
        }
        const int rest = n - pend, pjb = pend - p0;
        if (rest > 0) {
            trsm_blocked<T, 0>(pjb, rest, A + (size_t)p0 * lda + p0, lda, strideA,
                               A + (size_t)pend * lda + p0, lda, strideA, batch, s);
            launch_gemm<T>('N', 'N', rest, rest, pjb, mone, A + (size_t)p0 * lda + pend, lda,
                           strideA, A + (size_t)pend * lda + p0, lda, strideA, one,
                           A + (size_t)pend * lda + pend, lda, strideA, batch, s);
        }
    }
    if (defer) {
        // columns 0 .. (last outer panel start) - 1
        const int ncols = ((n - 1) / NBO) * NBO;
        int *idx = work ? work + (size_t)(1 + 2 * PERM_MAX) * batch : nullptr;
        if (!work)
            IA_HIP(hipMallocAsync((void **)&idx, sizeof(int) * (size_t)(n / NBO) * n * batch, s));
        const size_t lds = sizeof(int) * (size_t)n;
        (void)hipFuncSetAttribute((const void *)k_perm_suffix,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(k_perm_suffix, dim3(batch), dim3(256), lds, s, n, (const int *)ipiv,
                           idx);
        hipLaunchKernelGGL(k_laswp_deferred<T>, dim3((ncols + 7) / 8, batch), dim3(256), 0, s, n,
                           ncols, (const int *)idx, A, lda, strideA);
        if (!work) IA_HIP(hipFreeAsync(idx, s));
    }
    if (perm && !work) IA_HIP(hipFreeAsync(perm, s));
    IA_HIP(hipGetLastError());
    return IA_OK;
}

template <typename T>
size_t getrs_work_bytes(int n, int nrhs, int batch) {
    // the row index and its inverse, then the right-hand sides (16-byte aligned)
    const size_t ib = (sizeof(int) * 2 * (size_t)n * batch + 15) / 16 * 16;
    return ib + sizeof(cx<T>) * (size_t)n * nrhs * batch;
}

template <typename T>
ia_status getrs_batched(char op, int n, int nrhs, const cx<T> *A, int lda, size_t strideA,
                        const int *ipiv, cx<T> *B, int ldb, size_t strideB, int batch,
                        hipStream_t s, void *work) {
    if (n <= 0 || nrhs <= 0 || batch <= 0) return IA_OK;
    if (nrhs == 1 && getrs1_lds<T>(n) <= 160 * 1024) {
        const size_t lds = getrs1_lds<T>(n);
        const void *kern = op == 'N' ? (const void *)k_getrs1<T> : (const void *)k_getrs1c<T>;
        (void)hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (op == 'N')
            hipLaunchKernelGGL(k_getrs1<T>, dim3(batch), dim3(G1T), lds, s, n, A, lda, strideA,
                               ipiv, B, strideB);
        else
            hipLaunchKernelGGL(k_getrs1c<T>, dim3(batch), dim3(G1T), lds, s, n, A, lda, strideA,
                               ipiv, B, strideB);
        IA_HIP(hipGetLastError());
        return IA_OK;
    }
    int *idx = nullptr;
    cx<T> *tmp = nullptr;
    if (work) {
        idx = static_cast<int *>(work);
        tmp = reinterpret_cast<cx<T> *>(static_cast<char *>(work) +
                                        (sizeof(int) * 2 * (size_t)n * batch + 15) / 16 * 16);
    } else {
        IA_HIP(hipMallocAsync((void **)&idx, sizeof(int) * 2 * (size_t)n * batch, s));
        IA_HIP(hipMallocAsync((void **)&tmp, sizeof(cx<T>) * (size_t)n * nrhs * batch, s));
    }
    int *inv = idx + (size_t)n * batch;
    const ia_status ps = lu_perm_index(n, ipiv, idx, inv, batch, s);
    if (ps != IA_OK) {
        if (!work) {
            (void)hipFreeAsync(tmp, s);
            (void)hipFreeAsync(idx, s);
        }
        return ps;
    }
    dim3 pg((n + 255) / 256, nrhs, batch);
    const size_t sT = (size_t)n * nrhs;
    if (op == 'N') {
        // L U X = P B
        hipLaunchKernelGGL(k_row_gather<T>, pg, dim3(256), 0, s, n, nrhs, idx, (const cx<T> *)B,
                           ldb, strideB, tmp, n, sT);
        trsm_blocked<T, 0>(n, nrhs, A, lda, strideA, tmp, n, sT, batch, s);
        trsm_blocked<T, 2>(n, nrhs, A, lda, strideA, tmp, n, sT, batch, s);
        hipLaunchKernelGGL(k_copy_cols<T>, pg, dim3(256), 0, s, n, nrhs, (const cx<T> *)tmp, n, sT,
                           B, ldb, strideB);
    } else {
        // U^H L^H (P X) = B
        hipLaunchKernelGGL(k_copy_cols<T>, pg, dim3(256), 0, s, n, nrhs, (const cx<T> *)B, ldb,
                           strideB, tmp, n, sT);
        trsm_blocked<T, 1>(n, nrhs, A, lda, strideA, tmp, n, sT, batch, s);
        trsm_blocked<T, 3>(n, nrhs, A, lda, strideA, tmp, n, sT, batch, s);
        // X = P^T V  ->  X[r] = V[inv[r]]
        hipLaunchKernelGGL(k_row_gather<T>, pg, dim3(256), 0, s, n, nrhs, inv, (const cx<T> *)tmp,
                           n, sT, B, ldb, strideB);
    }
    if (!work) {
        IA_HIP(hipFreeAsync(tmp, s));
        IA_HIP(hipFreeAsync(idx, s));
    }
    IA_HIP(hipGetLastError());
    return IA_OK;
}

ia_status lu_perm_index(int n, const int *ipiv, int *idx, int *inv, int batch, hipStream_t s) {
    if (n <= 0 || batch <= 0) return IA_OK;
    const size_t plds = 2 * sizeof(int) * (size_t)n;
    if (plds <= 160 * 1024) {
        (void)hipFuncSetAttribute((const void *)k_perm_index,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)plds);
        hipLaunchKernelGGL(k_perm_index, dim3(batch), dim3(256), plds, s, n, ipiv, idx, inv);
    } else {
        hipLaunchKernelGGL(k_perm_index_global, dim3((batch + 63) / 64), dim3(64), 0, s, n, ipiv,
                           idx, inv, batch);
    }
    IA_HIP(hipGetLastError());
    return IA_OK;
}

// ---- one right-hand side for a few matrices (A = P^T L U, op 'N'): the
// single-workgroup solve above leaves most of the device idle at small
// batches, so here each 64-row block of each triangle is one launch whose
// workgroups each solve the block's diagonal triangle (redundantly, one
// wave) and update their own GM_ROWS rows of the rest.  The solved blocks
// go to a second vector (the block's unsolved entries stay readable by
// every workgroup of the launch): forward from B into S, backward from S
// into B.
constexpr int GM_ROWS = 256;

// B <- P B: the interchanges in order, in LDS (one workgroup per matrix)
template <typename T>
__global__ __launch_bounds__(256) void k_perm_rhs(int n, const int *__restrict__ ipivall,
                                                 cx<T> *__restrict__ Ball, size_t sB) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *xs = reinterpret_cast<cx<T> *>(smem_raw);
    int *ps = reinterpret_cast<int *>(xs + n);
    const int b = blockIdx.x, t = threadIdx.x;
    cx<T> *B = Ball + (size_t)b * sB;
    const int *ip = ipivall + (size_t)b * n;
    for (int r = t; r < n; r += 256) {
        xs[r] = B[r];
        ps[r] = ip[r];
    }
    __syncthreads();
    if (t == 0)
        for (int c = 0; c < n; ++c) {
            const int p = ps[c];
            if (p != c) {
                const cx<T> tmp = xs[c];
                xs[c] = xs[p];
                xs[p] = tmp;
            }
        }
    __syncthreads();
    for (int r = t; r < n; r += 256) B[r] = xs[r];
}

// LOWER: x_blk = L_blk^-1 (unit) src_blk into dst, rows below lose
// L[r, blk] x_blk in src.  Upper: x_blk = U_blk^-1 src_blk into dst, rows
// above lose U[r, blk] x_blk in src.
template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void k_trsv_block(int n, int i0, const cx<T> *__restrict__ Aall,
                                                   int lda, size_t sA, cx<T> *__restrict__ srcall,
                                                   cx<T> *__restrict__ dstall, size_t sB) {
    __shared__ cx<T> Ds[G1][G1 + 1];  // Ds[j][l] = A[i0 + l, i0 + j]
    __shared__ cx<T> xb[G1];
    const int b = blockIdx.y, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const cx<T> *A = Aall + (size_t)b * sA;
    cx<T> *src = srcall + (size_t)b * sB;
    cx<T> *dst = dstall + (size_t)b * sB;
    const int jb = min(G1, n - i0);
    for (int e = t; e < jb * jb; e += 256) {
        const int l = e % jb, j = e / jb;
        Ds[j][l] = A[(size_t)(i0 + j) * lda + i0 + l];
    }
    __syncthreads();
    if (wv == 0) {
        cx<T> x = lane < jb ? src[i0 + lane] : cx<T>{0, 0};
        if (LOWER) {
            for (int j = 0; j < jb; ++j) {
                const cx<T> xj = bcast_cx(x, j);
                if (lane > j && lane < jb) x = x - Ds[j][lane] * xj;
            }
        } else {
            const cx<T> dl = lane < jb ? cdiv(cx<T>{1, 0}, Ds[lane][lane]) : cx<T>{0, 0};
            for (int j = jb - 1; j >= 0; --j) {
                if (lane == j) x = x * dl;
                const cx<T> xj = bcast_cx(x, j);
                if (lane < j) x = x - Ds[j][lane] * xj;
            }
        }
        if (lane < jb) {
            xb[lane] = x;
            if (blockIdx.x == 0) dst[i0 + lane] = x;
        }
    }
    __syncthreads();
    const int r0 = (LOWER ? i0 + jb : 0) + blockIdx.x * GM_ROWS;
    const int r1 = min(r0 + GM_ROWS, LOWER ? n : i0);
    for (int r = r0 + t; r < r1; r += 256) {
        cx<T> acc{0, 0};
#pragma unroll 16
        for (int j = 0; j < jb; ++j) acc = acc + A[(size_t)(i0 + j) * lda + r] * xb[j];
        src[r] = src[r] - acc;
    }
}

template <typename T>
ia_status getrs1_spread(int n, const cx<T> *A, int lda, size_t strideA, const int *ipiv,
                        cx<T> *B, size_t strideB, cx<T> *S, int batch, hipStream_t s) {
    if (n <= 0 || batch <= 0) return IA_OK;
    const size_t lds = (sizeof(cx<T>) + sizeof(int)) * (size_t)n;
    if (lds > 64 * 1024) return getrs_batched<T>('N', n, 1, A, lda, strideA, ipiv, B, n, strideB, batch, s);
    (void)hipFuncSetAttribute((const void *)k_perm_rhs<T>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_perm_rhs<T>, dim3(batch), dim3(256), lds, s, n, ipiv, B, strideB);
    const int nblk = (n + G1 - 1) / G1;
    auto groups = [](int rows) { return (unsigned)std::max(1, (rows + GM_ROWS - 1) / GM_ROWS); };
    for (int bk = 0; bk < nblk; ++bk) {
        const int i0 = bk * G1, jb = std::min(G1, n - i0);
        hipLaunchKernelGGL((k_trsv_block<T, true>), dim3(groups(n - i0 - jb), batch), dim3(256), 0,
                           s, n, i0, A, lda, strideA, B, S, strideB);
    }
    for (int bk = nblk - 1; bk >= 0; --bk) {
        const int i0 = bk * G1;
        hipLaunchKernelGGL((k_trsv_block<T, false>), dim3(groups(i0), batch), dim3(256), 0, s, n,
                           i0, A, lda, strideA, S, B, strideB);
    }
    IA_HIP(hipGetLastError());
    return IA_OK;
}

template <typename T>
ia_status getrs_prepermuted(int n, int nrhs, const cx<T> *A, int lda, size_t strideA, cx<T> *B,
                            int ldb, size_t strideB, int batch, hipStream_t s) {
    if (n <= 0 || nrhs <= 0 || batch <= 0) return IA_OK;
    trsm_blocked<T, 0>(n, nrhs, A, lda, strideA, B, ldb, strideB, batch, s);
    trsm_blocked<T, 2>(n, nrhs, A, lda, strideA, B, ldb, strideB, batch, s);
    IA_HIP(hipGetLastError());
    return IA_OK;
}

template <typename T>
ia_status getrs_c_permuted(int n, int nrhs, const cx<T> *A, int lda, size_t strideA, cx<T> *B,
                           int ldb, size_t strideB, int batch, hipStream_t s) {
    if (n <= 0 || nrhs <= 0 || batch <= 0) return IA_OK;
    trsm_blocked<T, 1>(n, nrhs, A, lda, strideA, B, ldb, strideB, batch, s);
    trsm_blocked<T, 3>(n, nrhs, A, lda, strideA, B, ldb, strideB, batch, s);
    IA_HIP(hipGetLastError());
    return IA_OK;
}

template <typename T>
void launch_row_gather(int n, int nrhs, const int *perm, const cx<T> *src, int lds, size_t sS,
                       cx<T> *dst, int ldd, size_t sD, int batch, hipStream_t s) {
    if (n <= 0 || nrhs <= 0 || batch <= 0) return;
    dim3 g((n + 255) / 256, nrhs, batch);
    hipLaunchKernelGGL(k_row_gather<T>, g, dim3(256), 0, s, n, nrhs, perm, src, lds, sS, dst, ldd,
                       sD);
}

template <typename T>
ia_status trsm_upper_batched(int n, int nrhs, const cx<T> *A, int lda, size_t strideA, cx<T> *B,
                             int ldb, size_t strideB, int batch, hipStream_t s) {
    if (n <= 0 || nrhs <= 0 || batch <= 0) return IA_OK;
    trsm_blocked<T, 2>(n, nrhs, A, lda, strideA, B, ldb, strideB, batch, s);
    IA_HIP(hipGetLastError());
    return IA_OK;
}

template ia_status trsm_upper_batched<double>(int, int, const cx<double> *, int, size_t,
                                              cx<double> *, int, size_t, int, hipStream_t);
template ia_status getrs_c_permuted<double>(int, int, const cx<double> *, int, size_t,
                                            cx<double> *, int, size_t, int, hipStream_t);
template ia_status getrs_c_permuted<float>(int, int, const cx<float> *, int, size_t, cx<float> *,
                                           int, size_t, int, hipStream_t);
template void launch_row_gather<double>(int, int, const int *, const cx<double> *, int, size_t,
                                        cx<double> *, int, size_t, int, hipStream_t);
template void launch_row_gather<float>(int, int, const int *, const cx<float> *, int, size_t,
                                       cx<float> *, int, size_t, int, hipStream_t);
template ia_status getrs_prepermuted<double>(int, int, const cx<double> *, int, size_t,
                                             cx<double> *, int, size_t, int, hipStream_t);
template ia_status getrs_prepermuted<float>(int, int, const cx<float> *, int, size_t, cx<float> *,
                                            int, size_t, int, hipStream_t);
template ia_status getrf_batched<double>(int, cx<double> *, int, size_t, int *, int *, int,
                                         hipStream_t, int *);
template ia_status getrf_batched<float>(int, cx<float> *, int, size_t, int *, int *, int,
                                        hipStream_t, int *);
template ia_status getrs1_spread<double>(int, const cx<double> *, int, size_t, const int *,
                                         cx<double> *, size_t, cx<double> *, int, hipStream_t);
template ia_status getrs_batched<double>(char, int, int, const cx<double> *, int, size_t,
                                         const int *, cx<double> *, int, size_t, int, hipStream_t,
                                         void *);
template ia_status getrs_batched<float>(char, int, int, const cx<float> *, int, size_t,
                                        const int *, cx<float> *, int, size_t, int, hipStream_t,
                                        void *);
template size_t getrs_work_bytes<double>(int, int, int);
template size_t getrs_work_bytes<float>(int, int, int);

}  // namespace ia
