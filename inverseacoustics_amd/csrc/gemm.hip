// Batched complex GEMM on CDNA4 matrix cores.
//
//   C = alpha op(A) op(B) + beta C,   op in {'N', 'C'} (conjugate transpose)
//
// Replaces the cuBLAS Zgemm/Zgeam call sites of the reference
// (CUDA/ibem.cuh:398, CUDA/invr.cuh:402-433) and carries every product of
// the sweep: the trailing updates of the LU, triangular-solve, QR and band
// reductions.
//
// Complex products by the 3M (Gauss) method: three real MFMAs per complex
// multiply-add, re*re, im*im and (re+im)*(re+im), into three accumulators;
// real part = S1 - S2, imaginary part = S3 - S1 - S2 (normwise stable,
// Higham, "Stability of a method for multiplying complex matrices with
// three real matrix multiplications", SIMAX 13, 1992).  The sums re+im are
// formed once per element when the tile is written to LDS.  MFMAs are
// v_mfma_f64_16x16x4_f64 (FP64 form) or v_mfma_f32_16x16x4_f32 (mixed
// form).  Each wave computes the TRANSPOSED product op(B)^T op(A)^T so that
// the 16x16 MFMA result rows land on consecutive lanes along the
// column-major C rows: one store instruction writes 4 columns x 16
// contiguous rows.
//
// Workgroup: 4 waves in a 2 x 2 grid; operands staged in LDS as split
// planes in k-major order so that the MFMA operand reads of 16 consecutive
// lanes hit 16 consecutive words; two LDS buffers, the next stage loaded
// from global memory into registers while the current one is multiplied.
//
// Deep complex64 products (depth >= 64, both dimensions > 32) take a second
// kernel, k_cgemm_x6, on the bf16 matrix cores: the FP32 operands are split
// exactly into three bf16 parts and six partial products are accumulated in
// FP32, which keeps FP32-level products (see k_cgemm_x6 below).

#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>

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

constexpr int PAD = 4;

// Tile shape: each of the 4 waves (2 x 2) computes (16 WMB) x (16 WNB) of C,
// the workgroup (32 WMB) x (32 WNB); BK = depth of one LDS stage; OCC =
// waves per SIMD the register allocation must allow.

// Batch-major workgroup order with the batch dealt round-robin over the 8
// XCDs (workgroup h runs on XCD h mod 8): every tile of one matrix runs on
// one XCD, so the panels it shares stay in that XCD's L2.
__device__ __forceinline__ void tile_of(int h, int tiles, int batch, int &b, int &t) {
    if ((batch & 7) == 0) {
        const int x = h & 7, slot = h >> 3;
        b = (slot / tiles) * 8 + x;
        t = slot % tiles;
    } else {
        b = h / tiles;
        t = h % tiles;
    }
}

// One stage of an R x BK operand tile (R = rows of the tile) from global
// memory into registers.  KC: element (r, k) is X[r ldx + k] (contiguous
// along k), else X[k ldx + r] (contiguous along r); consecutive threads take
// consecutive addresses either way.  CONJ: the tile holds the conjugates.
template <typename T, bool KC, bool CONJ, int R, int BK>
struct TileLoader {
    static constexpr int NE = R * BK / 256;
    static_assert(NE * 256 == R * BK, "tile must split evenly over 256 threads");
    // KC with two elements per thread: the pair (r, 2j), (r, 2j + 1) is
    // adjacent in memory and loads as one vector
    static constexpr bool PAIR = KC && NE == 2 && BK % 2 == 0;
    static __device__ __forceinline__ void coords(int t, int i, int &r, int &k) {
        if (PAIR) {
            k = (t % (BK / 2)) * 2 + i;
            r = t / (BK / 2);
        } else if (KC) {
            k = t % BK;
            r = t / BK + i * (256 / BK);
        } else {
            r = t % R;
            k = t / R + i * (256 / R);
        }
    }
    static __device__ __forceinline__ void load(const cx<T> *__restrict__ X, int ldx, int r0,
                                                int nr, int k0, int K, int t, cx<T> (&v)[NE]) {
        if constexpr (PAIR) {
            int r, k;
            coords(t, 0, r, k);
            const int gr = r0 + r, gk = k0 + k;
            const cx<T> *p = X + (size_t)gr * ldx + gk;
            if (gr < nr && gk + 1 < K && ((uintptr_t)p & (2 * sizeof(cx<T>) - 1)) == 0) {
                if constexpr (sizeof(T) == sizeof(float)) {
                    const float4 f = *reinterpret_cast<const float4 *>(p);
                    v[0] = cx<T>{f.x, f.y};
                    v[1] = cx<T>{f.z, f.w};
                } else {
                    const double2 *q = reinterpret_cast<const double2 *>(p);
                    const double2 a = q[0], b = q[1];
                    v[0] = cx<T>{a.x, a.y};
                    v[1] = cx<T>{b.x, b.y};
                }
                return;
            }
        }
#pragma unroll
        for (int i = 0; i < NE; ++i) {
            int r, k;
            coords(t, i, r, k);
            const int gr = r0 + r, gk = k0 + k;
            v[i] = cx<T>{0, 0};
            if (gr < nr && gk < K) v[i] = KC ? X[(size_t)gr * ldx + gk] : X[(size_t)gk * ldx + gr];
        }
    }
    // split planes, k-major: S[0][k][r] = re, S[1][k][r] = im and, with
    // three planes, S[2][k][r] = re + im
    template <int NP>
    static __device__ __forceinline__ void store(T (*S)[BK][R + PAD], int t,
                                                 const cx<T> (&v)[NE]) {
#pragma unroll
        for (int i = 0; i < NE; ++i) {
            int r, k;
            coords(t, i, r, k);
            const T re = v[i].re, im = CONJ ? -v[i].im : v[i].im;
            S[0][k][r] = re;
            S[1][k][r] = im;
            if (NP == 3) S[2][k][r] = re + im;
        }
    }
};

// C = alpha op(A) op(B) + beta C over a batch (strides sA, sB, sC).
// Double-buffered: the next stage's tiles are loaded into registers while the
// matrix cores work on the current one, then written to the other LDS buffer;
// one barrier per stage.
template <typename T, char OPA, char OPB, int WMB, int WNB, int BK, int OCC, bool M3, bool BETA0>
__global__ __launch_bounds__(256, OCC) void k_zgemm(int M, int N, int K, cx<T> alpha,
                                              const cx<T> *__restrict__ A, int lda, size_t sA,
                                              const cx<T> *__restrict__ B, int ldb, size_t sB,
                                              cx<T> beta, cx<T> *__restrict__ C, int ldc,
                                              size_t sC, int tiles_m, int tiles_n, int batch) {
    using acc_t = typename mfma_t<T>::acc;
    constexpr int BM = 32 * WMB, BN = 32 * WNB;
    // op(A) (BM x BK): A[k lda + m] or conj(A[m lda + k]);
    // op(B)^T (BN x BK): B[n ldb + k] or conj(B[k ldb + n])
    using LA = TileLoader<T, OPA == 'C', OPA == 'C', BM, BK>;
    using LB = TileLoader<T, OPB == 'N', OPB == 'C', BN, BK>;
    constexpr int NP = M3 ? 3 : 2;  // LDS planes per operand
    __shared__ T As[2][NP][BK][BM + PAD];
    __shared__ T Bs[2][NP][BK][BN + PAD];

    int bz, tile;
    tile_of(blockIdx.x, tiles_m * tiles_n, batch, bz, tile);
    A += (size_t)bz * sA;
    B += (size_t)bz * sB;
    C += (size_t)bz * sC;
    const int m0 = (tile % tiles_m) * BM, n0 = (tile / tiles_m) * BN;
    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    const int wm = (w & 1) * 16 * WMB, wn = (w >> 1) * 16 * WNB;
    const int l16 = lane & 15, lk = lane >> 4;

    // 4M: cre = sum(re re - im im), cim = sum(im re + re im).  3M (Gauss):
    // cre = sum(re re), cim = sum(im im), c3 = sum((re + im)(re + im)); the
    // real part is cre - cim and the imaginary part c3 - cre - cim.
    acc_t cre[WMB][WNB], cim[WMB][WNB], c3[M3 ? WMB : 1][M3 ? WNB : 1];
#pragma unroll
    for (int a = 0; a < WMB; ++a)
#pragma unroll
        for (int b = 0; b < WNB; ++b) {
            cre[a][b] = acc_t{0, 0, 0, 0};
            cim[a][b] = acc_t{0, 0, 0, 0};
            if (M3) c3[M3 ? a : 0][M3 ? b : 0] = acc_t{0, 0, 0, 0};
        }

    cx<T> ra[LA::NE], rb[LB::NE];
    const int nst = (K + BK - 1) / BK;
    if (nst > 0) {
        LA::load(A, lda, m0, M, 0, K, t, ra);
        LB::load(B, ldb, n0, N, 0, K, t, rb);
        LA::template store<NP>(As[0], t, ra);
        LB::template store<NP>(Bs[0], t, rb);
    }
    __syncthreads();
    for (int st = 0; st < nst; ++st) {
        const int cur = st & 1;
        const bool more = st + 1 < nst;
        if (more) {
            LA::load(A, lda, m0, M, (st + 1) * BK, K, t, ra);
            LB::load(B, ldb, n0, N, (st + 1) * BK, K, t, rb);
        }
#pragma unroll
        for (int ks = 0; ks < BK; ks += 4) {
            T are[WMB], aim[WMB], asum[WMB], bre[WNB], bim[WNB], bsum[WNB];
#pragma unroll
            for (int a = 0; a < WMB; ++a) {
                are[a] = As[cur][0][ks + lk][wm + a * 16 + l16];
                aim[a] = As[cur][1][ks + lk][wm + a * 16 + l16];
                if (M3) asum[a] = As[cur][NP - 1][ks + lk][wm + a * 16 + l16];
            }
#pragma unroll
            for (int b = 0; b < WNB; ++b) {
                bre[b] = Bs[cur][0][ks + lk][wn + b * 16 + l16];
                bim[b] = Bs[cur][1][ks + lk][wn + b * 16 + l16];
                if (M3) bsum[b] = Bs[cur][NP - 1][ks + lk][wn + b * 16 + l16];
            }
            // transposed product: MFMA-A <- op(B)^T rows (n), MFMA-B <- op(A)^T
            // cols (m); the first products of every block before the second
            // ones, so that no MFMA waits on the one just issued
            if (M3) {
#pragma unroll
                for (int a = 0; a < WMB; ++a)
#pragma unroll
                    for (int b = 0; b < WNB; ++b) {
                        cre[a][b] = mfma_t<T>::run(bre[b], are[a], cre[a][b]);
                        cim[a][b] = mfma_t<T>::run(bim[b], aim[a], cim[a][b]);
                        c3[M3 ? a : 0][M3 ? b : 0] =
                            mfma_t<T>::run(bsum[b], asum[a], c3[M3 ? a : 0][M3 ? b : 0]);
                    }
            } else {
#pragma unroll
                for (int a = 0; a < WMB; ++a)
#pragma unroll
                    for (int b = 0; b < WNB; ++b) {
                        cre[a][b] = mfma_t<T>::run(bre[b], are[a], cre[a][b]);
                        cim[a][b] = mfma_t<T>::run(bim[b], are[a], cim[a][b]);
                    }
#pragma unroll
                for (int a = 0; a < WMB; ++a)
#pragma unroll
                    for (int b = 0; b < WNB; ++b) {
                        cre[a][b] = mfma_t<T>::run(bim[b], -aim[a], cre[a][b]);
                        cim[a][b] = mfma_t<T>::run(bre[b], aim[a], cim[a][b]);
                    }
            }
        }
        if (more) {
            LA::template store<NP>(As[cur ^ 1], t, ra);
            LB::template store<NP>(Bs[cur ^ 1], t, rb);
        }
        __syncthreads();
    }
    // ---- epilogue: lane holds D'[n = row(lk, i)][m = l16]
#pragma unroll
    for (int a = 0; a < WMB; ++a)
#pragma unroll
        for (int b = 0; b < WNB; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int gm = m0 + wm + a * 16 + l16;
                const int gn = n0 + wn + b * 16 + mfma_t<T>::row(lk, i);
                if (gm < M && gn < N) {
                    T xr = cre[a][b][i], xi = cim[a][b][i];
                    if (M3) {
                        const T s1 = xr, s2 = xi;
                        xr = s1 - s2;
                        xi = c3[M3 ? a : 0][M3 ? b : 0][i] - s1 - s2;
                    }
                    cx<T> r{alpha.re * xr - alpha.im * xi, alpha.re * xi + alpha.im * xr};
                    const size_t o = (size_t)gn * ldc + gm;
                    if (!BETA0) {
                        const cx<T> c0 = C[o];
                        r.re += beta.re * c0.re - beta.im * c0.im;
                        r.im += beta.re * c0.im + beta.im * c0.re;
                    }
                    C[o] = r;
                }
            }
}

// ---------------------------------------------------------------- complex64
// on the bf16 matrix cores.  Every FP32 operand v is split exactly into three
// bf16 parts by truncation, v = h + m + l (8 + 8 + 8 significant bits), and
// the product keeps the six terms hh, hm, mh, hl, lh, mm: the dropped ones
// are below 2^-23 |a b|, so the result carries FP32-level products with
// FP32 accumulation (v_mfma_f32_16x16x32_bf16, k = 32 per instruction).
// Complex products as S1 = sum re re, S2 = sum im im, S3 = sum (re im + im re):
// real part S1 - S2, imaginary part S3.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4 mfma_bf16(u32x4 a, u32x4 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// the three bf16 parts of v as the high halves of three words
__device__ __forceinline__ void split3(float v, unsigned &h, unsigned &m, unsigned &l) {
    h = __float_as_uint(v) & 0xffff0000u;
    const float r1 = v - __uint_as_float(h);
    m = __float_as_uint(r1) & 0xffff0000u;
    const float r2 = r1 - __uint_as_float(m);
    l = __float_as_uint(r2) & 0xffff0000u;
}

constexpr int XB = 64, XK = 32, XKP = XK / 8;  // tile, stage depth, row (16-byte chunks)
// chunk c of row r is stored at chunk c ^ (r & 3): the 64 rows a wave writes
// at one k (64-byte row stride) then fall on different banks, and a wave's
// fragment reads (16 rows x 4 chunks) still cover 1 KB without conflict
__device__ __forceinline__ int xsw(int r, int c) { return c ^ (r & 3); }

// Operand tile (64 rows x 32 k) of one thread: one row, 8 consecutive k.
// KC: element (r, k) is X[r ldx + k] (contiguous along k; four threads per
// row), else X[k ldx + r] (contiguous along r; 64 threads per k).
template <bool KC, bool CONJ>
struct SplitLoader {
    int r, kg;
    __device__ __forceinline__ explicit SplitLoader(int t) {
        if (KC) {
            r = t >> 2;
            kg = t & 3;
        } else {
            r = t & 63;
            kg = t >> 6;
        }
    }
    __device__ __forceinline__ void load(const cx<float> *__restrict__ X, int ldx, int r0, int nr,
                                         int k0, int K, cx<float> (&v)[8]) const {
        const int gr = r0 + r;
        if (KC) {
            // the thread's 8 consecutive k (64 bytes) as four 16-byte loads
            // when they are all in range and aligned
            const cx<float> *p = X + (size_t)gr * ldx + (k0 + kg * 8);
            if (gr < nr && k0 + kg * 8 + 8 <= K && ((uintptr_t)p & 15) == 0) {
                const float4 *q = reinterpret_cast<const float4 *>(p);
#pragma unroll
                for (int h = 0; h < 4; ++h) {
                    const float4 f = q[h];
                    v[2 * h] = cx<float>{f.x, f.y};
                    v[2 * h + 1] = cx<float>{f.z, f.w};
                }
                return;
            }
        }
        if (!KC && gr < nr && k0 + kg * 8 + 8 <= K) {
            // the thread's 8 k (one row, eight columns) all in range: eight
            // plain loads, each a 512-byte access of 64 consecutive rows
            const cx<float> *p = X + (size_t)(k0 + kg * 8) * ldx + gr;
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = p[(size_t)j * ldx];
            return;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int gk = k0 + kg * 8 + j;
            v[j] = cx<float>{0, 0};
            if (gr < nr && gk < K) v[j] = KC ? X[(size_t)gr * ldx + gk] : X[(size_t)gk * ldx + gr];
        }
    }
    // planes 0..2: re (h, m, l), 3..5: im (h, m, l)
    __device__ __forceinline__ void store(u32x4 (*S)[XB][XKP], const cx<float> (&v)[8]) const {
        unsigned p[6][8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            split3(v[j].re, p[0][j], p[1][j], p[2][j]);
            split3(CONJ ? -v[j].im : v[j].im, p[3][j], p[4][j], p[5][j]);
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            u32x4 w;
#pragma unroll
            for (int j = 0; j < 4; ++j) w[j] = (p[q][2 * j] >> 16) | p[q][2 * j + 1];
            S[q][r][xsw(r, kg)] = w;
        }
    }
};

template <char OPA, char OPB, bool BETA0>
// 48 KB of LDS and at most 168 VGPRs: three workgroups per CU
__global__ __launch_bounds__(256, 3) void k_cgemm_x6(int M, int N, int K, cx<float> alpha,
                                                    const cx<float> *__restrict__ A, int lda,
                                                    size_t sA, const cx<float> *__restrict__ B,
                                                    int ldb, size_t sB, cx<float> beta,
                                                    cx<float> *__restrict__ C, int ldc, size_t sC,
                                                    int tiles_m, int tiles_n, int batch) {
    __shared__ u32x4 As[6][XB][XKP];
    __shared__ u32x4 Bs[6][XB][XKP];
    int bz, tile;
    tile_of(blockIdx.x, tiles_m * tiles_n, batch, bz, tile);
    A += (size_t)bz * sA;
    B += (size_t)bz * sB;
    C += (size_t)bz * sC;
    const int m0 = (tile % tiles_m) * XB, n0 = (tile / tiles_m) * XB;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = (w & 1) * 32, wn = (w >> 1) * 32;
    const int l16 = lane & 15, g = lane >> 4;
    // op(A) (64 x 32): A[k lda + m] or conj(A[m lda + k]); op(B)^T (64 x 32):
    // B[n ldb + k] or conj(B[k ldb + n])
    const SplitLoader<OPA == 'C', OPA == 'C'> la(t);
    const SplitLoader<OPB == 'N', OPB == 'C'> lb(t);
    f4 s1[2][2], s2[2][2], s3[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) s1[a][b] = s2[a][b] = s3[a][b] = f4{0, 0, 0, 0};
    cx<float> va[8], vb[8];
    const int nst = (K + XK - 1) / XK;
    if (nst > 0) {
        la.load(A, lda, m0, M, 0, K, va);
        lb.load(B, ldb, n0, N, 0, K, vb);
        la.store(As, va);
        lb.store(Bs, vb);
    }
    __syncthreads();
    // one stage's products: the B fragments stay in registers, the A
    // fragments are read per row block
    auto stage = [&]() {
        constexpr int PX[6] = {0, 0, 1, 0, 2, 1}, PY[6] = {0, 1, 0, 2, 0, 1};
        u32x4 fb[2][6];
#pragma unroll
        for (int q = 0; q < 6; ++q)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int r = wn + 16 * b + l16;
                fb[b][q] = Bs[q][r][xsw(r, g)];
            }
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            u32x4 fa[6];
            const int ra = wm + 16 * a + l16;
#pragma unroll
            for (int q = 0; q < 6; ++q) fa[q] = As[q][ra][xsw(ra, g)];
            // (x, y) split pairs kept: hh, hm, mh, hl, lh, mm; MFMA-A <- op(B)^T
            // rows (n), MFMA-B <- op(A)^T columns (m)
#pragma unroll
            for (int pr = 0; pr < 6; ++pr)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    const int x = PX[pr], y = PY[pr];
                    s1[a][b] = mfma_bf16(fb[b][y], fa[x], s1[a][b]);
                    s2[a][b] = mfma_bf16(fb[b][3 + y], fa[3 + x], s2[a][b]);
                    s3[a][b] = mfma_bf16(fb[b][3 + y], fa[x], s3[a][b]);
                    s3[a][b] = mfma_bf16(fb[b][y], fa[3 + x], s3[a][b]);
                }
        }
    };
    for (int st = 0; st + 1 < nst; ++st) {
        la.load(A, lda, m0, M, (st + 1) * XK, K, va);
        lb.load(B, ldb, n0, N, (st + 1) * XK, K, vb);
        stage();
        __syncthreads();  // every wave has read the stage
        la.store(As, va);
        lb.store(Bs, vb);
        __syncthreads();
    }
    // last stage: with beta != 0 the reads of C are issued before its
    // products (into the registers the next-stage loads no longer need), so
    // their latency overlaps the MFMAs instead of following them
    cx<float> c0v[2][2][4];
    if (!BETA0) {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int gm = m0 + wm + a * 16 + l16;
                    const int gn = n0 + wn + b * 16 + 4 * g + i;
                    c0v[a][b][i] = (gm < M && gn < N) ? C[(size_t)gn * ldc + gm] : cx<float>{0, 0};
                }
    }
    if (nst > 0) stage();
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int gm = m0 + wm + a * 16 + l16;
                const int gn = n0 + wn + b * 16 + 4 * g + i;
                if (gm < M && gn < N) {
                    const float xr = s1[a][b][i] - s2[a][b][i], xi = s3[a][b][i];
                    cx<float> r{alpha.re * xr - alpha.im * xi, alpha.re * xi + alpha.im * xr};
                    if (!BETA0) {
                        const cx<float> c0 = c0v[a][b][i];
                        r.re += beta.re * c0.re - beta.im * c0.im;
                        r.im += beta.re * c0.im + beta.im * c0.re;
                    }
                    C[(size_t)gn * ldc + gm] = r;
                }
            }
}

template <char OPA, char OPB>
void launch_x6(int M, int N, int K, cx<float> alpha, const cx<float> *A, int lda, size_t sA,
               const cx<float> *B, int ldb, size_t sB, cx<float> beta, cx<float> *C, int ldc,
               size_t sC, int batch, hipStream_t s) {
    const int tm = (M + XB - 1) / XB, tn = (N + XB - 1) / XB;
    const dim3 grid((unsigned)((size_t)tm * tn * batch));
    if (beta.re == 0 && beta.im == 0)
        hipLaunchKernelGGL((k_cgemm_x6<OPA, OPB, true>), grid, dim3(256), 0, s, M, N, K, alpha, A,
                           lda, sA, B, ldb, sB, beta, C, ldc, sC, tm, tn, batch);
    else
        hipLaunchKernelGGL((k_cgemm_x6<OPA, OPB, false>), grid, dim3(256), 0, s, M, N, K, alpha, A,
                           lda, sA, B, ldb, sB, beta, C, ldc, sC, tm, tn, batch);
}

template <typename T, char OPA, char OPB, int WMB, int WNB, int BK, int OCC, bool M3>
void launch_tiles(int M, int N, int K, cx<T> alpha, const cx<T> *A, int lda, size_t sA,
                  const cx<T> *B, int ldb, size_t sB, cx<T> beta, cx<T> *C, int ldc, size_t sC,
                  int batch, hipStream_t s) {
    const int tm = (M + 32 * WMB - 1) / (32 * WMB), tn = (N + 32 * WNB - 1) / (32 * WNB);
    const dim3 grid((unsigned)((size_t)tm * tn * batch));
    if (beta.re == 0 && beta.im == 0)
        hipLaunchKernelGGL((k_zgemm<T, OPA, OPB, WMB, WNB, BK, OCC, M3, true>), grid, dim3(256), 0, s, M, N, K,
                           alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, tm, tn, batch);
    else
        hipLaunchKernelGGL((k_zgemm<T, OPA, OPB, WMB, WNB, BK, OCC, M3, false>), grid, dim3(256), 0, s, M, N,
                           K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, tm, tn, batch);
}

// Kernel candidates of one product: the bf16 split kernel (complex64 only)
// and the 3M kernel with 32 x 32, 32 x 64, 64 x 32 and 64 x 64 tiles.
enum GemmCand { GC_X6 = 0, GC_T11 = 1, GC_T12 = 2, GC_T21 = 3, GC_T22 = 4, GC_COUNT = 5 };

template <typename T, char OPA, char OPB>
void launch_cand(int cand, int M, int N, int K, cx<T> alpha, const cx<T> *A, int lda, size_t sA,
                 const cx<T> *B, int ldb, size_t sB, cx<T> beta, cx<T> *C, int ldc, size_t sC,
                 int batch, hipStream_t s) {
    constexpr int BK = 8, OCC = sizeof(T) == sizeof(float) ? 4 : 2;
#define IA_TILES(WMB, WNB)                                                                        \
    launch_tiles<T, OPA, OPB, WMB, WNB, BK, OCC, true>(M, N, K, alpha, A, lda, sA, B, ldb, sB, \
                                                           beta, C, ldc, sC, batch, s)
    switch (cand) {
    case GC_X6:
        if constexpr (sizeof(T) == sizeof(float)) {
            launch_x6<OPA, OPB>(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, s);
            return;
        }
        IA_TILES(2, 2);
        return;
    case GC_T11: IA_TILES(1, 1); return;
    case GC_T12: IA_TILES(1, 2); return;
    case GC_T21: IA_TILES(2, 1); return;
    default: IA_TILES(2, 2); return;
    }
#undef IA_TILES
}

// Built-in choice (measured on the sweep's products, batch 256): complex64
// products of depth >= 64 with both dimensions > 32 on the bf16 matrix
// cores (1.33x the FP32 3M kernel at K >= 512, ahead by 10-20 % at K = 64
// to 128; at K = 32 its one stage, split and barriers make it slower);
// otherwise the 3M kernel with 64 x 64 tiles, 32-row (column)
// tiles when op(A) has at most 32 rows (C at most 32 columns), and 32 x 32
// tiles when 64 x 64 ones would leave fewer than four workgroups per CU
// (small batches).  Four 64 x 64 workgroups share a CU: their barriers,
// epilogues and global-load waits overlap each other's MFMAs, which larger
// tiles at lower occupancy measured slower at.
int default_cand(bool f32, int M, int N, int K, int batch) {
    if (f32 && K >= 64 && M > 32 && N > 32) return GC_X6;
    const double wg64 = (double)((M + 63) / 64) * ((N + 63) / 64) * batch;
    if (M <= 32 && N <= 32) return GC_T11;
    if (M <= 32) return GC_T12;
    if (N <= 32) return GC_T21;
    if (wg64 < 4.0 * 256) return GC_T11;
    return GC_T22;
}

// ---- tuning table: the autotuner (csrc/bench/gemm_tune.hip) times every
// candidate on each product a run issued and writes the fastest per exact
// shape; the library reads the table once, from IA_GEMM_TUNE (a path; "off"
// disables it) or else gemm_tune.txt beside libia.so, and falls back to the
// built-in choice for shapes it does not list.  Line format:
//   GEMMTUNE <sizeof real> <opA> <opB> M N K batch <beta == 0> <candidate> ...
struct TuneKey {
    int bytes, opa, opb, M, N, K, batch, beta0;
    bool operator==(const TuneKey &o) const {
        return bytes == o.bytes && opa == o.opa && opb == o.opb && M == o.M && N == o.N &&
               K == o.K && batch == o.batch && beta0 == o.beta0;
    }
};
struct TuneHash {
    size_t operator()(const TuneKey &k) const {
        size_t h = 1469598103934665603ull;
        for (int v : {k.bytes, k.opa, k.opb, k.M, k.N, k.K, k.batch, k.beta0})
            h = (h ^ (size_t)(unsigned)v) * 1099511628211ull;
        return h;
    }
};
using TuneTable = std::unordered_map<TuneKey, int, TuneHash>;

std::string default_tune_path() {
    Dl_info info;
    if (dladdr((const void *)&default_cand, &info) && info.dli_fname) {
        std::string so(info.dli_fname);
        const size_t slash = so.find_last_of('/');
        return (slash == std::string::npos ? std::string(".") : so.substr(0, slash)) +
               "/gemm_tune.txt";
    }
    return "gemm_tune.txt";
}

const TuneTable &tune_table() {
    static TuneTable table;
    static std::once_flag once;
    std::call_once(once, [] {
        const char *e = std::getenv("IA_GEMM_TUNE");
        if (e && (std::strcmp(e, "off") == 0 || std::strcmp(e, "0") == 0)) return;
        const std::string path = (e && *e) ? std::string(e) : default_tune_path();
        FILE *f = std::fopen(path.c_str(), "r");
        if (!f) return;
        char line[512];
        while (std::fgets(line, sizeof line, f)) {
            TuneKey k{};
            char oa = 0, ob = 0;
            int cand = -1;
            if (std::sscanf(line, "GEMMTUNE %d %c %c %d %d %d %d %d %d", &k.bytes, &oa, &ob, &k.M,
                            &k.N, &k.K, &k.batch, &k.beta0, &cand) == 9 &&
                cand >= 0 && cand < GC_COUNT) {
                k.opa = oa;
                k.opb = ob;
                table[k] = cand;
            }
        }
        std::fclose(f);
    });
    return table;
}

template <typename T, char OPA, char OPB>
void dispatch_tiles(int M, int N, int K, cx<T> alpha, const cx<T> *A, int lda, size_t sA,
                    const cx<T> *B, int ldb, size_t sB, cx<T> beta, cx<T> *C, int ldc, size_t sC,
                    int batch, hipStream_t s) {
    constexpr bool f32 = sizeof(T) == sizeof(float);
    const bool beta0 = beta.re == 0 && beta.im == 0;
    int cand = -1;
    // IA_GEMM_FORCE=<candidate>: every product on that kernel (tests)
    if (const char *fe = std::getenv("IA_GEMM_FORCE"); fe && *fe) cand = std::atoi(fe);
    if (cand < 0 || cand >= GC_COUNT) {
        const TuneTable &tt = tune_table();
        cand = -1;
        if (!tt.empty()) {
            const auto it = tt.find(TuneKey{(int)sizeof(T), OPA, OPB, M, N, K, batch, beta0 ? 1 : 0});
            if (it != tt.end()) cand = it->second;
        }
        if (cand < 0) cand = default_cand(f32, M, N, K, batch);
    }
    if (!f32 && cand == GC_X6) cand = default_cand(false, M, N, K, batch);
    launch_cand<T, OPA, OPB>(cand, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC,
                             batch, s);
}

// ------------------------------------------------------------------ FP64 real
// C = op(A) B over a batch, op in {'N', 'T'}, real double: the products of a
// real symmetric matrix with a block of vectors (the radiation-mode subspace
// iteration).  The scheme of k_zgemm with one LDS plane per operand: 4 waves
// in a 2 x 2 grid, each computing the transposed product on
// v_mfma_f64_16x16x4f64 so that result rows land on consecutive lanes along
// the column-major C rows.
template <char OPA, int WMB, int WNB, int BK, int OCC>
__global__ __launch_bounds__(256, OCC) void k_dgemm(int M, int N, int K,
                                                    const double *__restrict__ A, int lda,
                                                    size_t sA, const double *__restrict__ B,
                                                    int ldb, size_t sB, double *__restrict__ C,
                                                    int ldc, size_t sC, int tiles_m, int tiles_n,
                                                    int batch) {
    constexpr int BM = 32 * WMB, BN = 32 * WNB;
    constexpr int NA = BM * BK / 256, NB = BN * BK / 256;
    static_assert(NA * 256 == BM * BK && NB * 256 == BN * BK, "tiles split over 256 threads");
    __shared__ double As[2][BK][BM + PAD];
    __shared__ double Bs[2][BK][BN + PAD];
    int bz, tile;
    tile_of(blockIdx.x, tiles_m * tiles_n, batch, bz, tile);
    A += (size_t)bz * sA;
    B += (size_t)bz * sB;
    C += (size_t)bz * sC;
    const int m0 = (tile % tiles_m) * BM, n0 = (tile / tiles_m) * BN;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = (w & 1) * 16 * WMB, wn = (w >> 1) * 16 * WNB;
    const int l16 = lane & 15, lk = lane >> 4;
    // element i of this thread's share of a stage: op(A) row r, depth k
    // ('N': A[k lda + m], consecutive threads along m; 'T': A[m lda + k],
    // consecutive threads along k); op(B)^T row r = column n of B, B[n ldb + k]
    auto coord_a = [&](int i, int &r, int &k) {
        if (OPA == 'N') {
            r = t % BM;
            k = t / BM + i * (256 / BM);
        } else {
            k = t % BK;
            r = t / BK + i * (256 / BK);
        }
    };
    auto coord_b = [&](int i, int &r, int &k) {
        k = t % BK;
        r = t / BK + i * (256 / BK);
    };
    double ra[NA], rb[NB];
    auto load = [&](int k0) {
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            int r, k;
            coord_a(i, r, k);
            const int gm = m0 + r, gk = k0 + k;
            ra[i] = (gm < M && gk < K)
                        ? (OPA == 'N' ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk])
                        : 0.0;
        }
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            int r, k;
            coord_b(i, r, k);
            const int gn = n0 + r, gk = k0 + k;
            rb[i] = (gn < N && gk < K) ? B[(size_t)gn * ldb + gk] : 0.0;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            int r, k;
            coord_a(i, r, k);
            As[buf][k][r] = ra[i];
        }
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            int r, k;
            coord_b(i, r, k);
            Bs[buf][k][r] = rb[i];
        }
    };
    d4 acc[WMB][WNB];
#pragma unroll
    for (int a = 0; a < WMB; ++a)
#pragma unroll
        for (int b = 0; b < WNB; ++b) acc[a][b] = d4{0, 0, 0, 0};
    const int nst = (K + BK - 1) / BK;
    if (nst > 0) {
        load(0);
        store(0);
    }
    __syncthreads();
    for (int st = 0; st < nst; ++st) {
        const int cur = st & 1;
        const bool more = st + 1 < nst;
        if (more) load((st + 1) * BK);
#pragma unroll
        for (int ks = 0; ks < BK; ks += 4) {
            double av[WMB], bv[WNB];
#pragma unroll
            for (int a = 0; a < WMB; ++a) av[a] = As[cur][ks + lk][wm + a * 16 + l16];
#pragma unroll
            for (int b = 0; b < WNB; ++b) bv[b] = Bs[cur][ks + lk][wn + b * 16 + l16];
#pragma unroll
            for (int a = 0; a < WMB; ++a)
#pragma unroll
                for (int b = 0; b < WNB; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[b], av[a], acc[a][b], 0,
                                                                     0, 0);
        }
        if (more) store(cur ^ 1);
        __syncthreads();
    }
    // lane holds C[m = l16][n = lk + 4 i] of its 16 x 16 block
#pragma unroll
    for (int a = 0; a < WMB; ++a)
#pragma unroll
        for (int b = 0; b < WNB; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int gm = m0 + wm + a * 16 + l16;
                const int gn = n0 + wn + b * 16 + mfma_t<double>::row(lk, i);
                if (gm < M && gn < N) C[(size_t)gn * ldc + gm] = acc[a][b][i];
            }
}

template <char OPA, int WMB, int WNB>
void launch_dtiles(int M, int N, int K, const double *A, int lda, size_t sA, const double *B,
                   int ldb, size_t sB, double *C, int ldc, size_t sC, int batch, hipStream_t s) {
    constexpr int BK = 16, OCC = 2;
    const int tm = (M + 32 * WMB - 1) / (32 * WMB), tn = (N + 32 * WNB - 1) / (32 * WNB);
    const dim3 grid((unsigned)((size_t)tm * tn * batch));
    hipLaunchKernelGGL((k_dgemm<OPA, WMB, WNB, BK, OCC>), grid, dim3(256), 0, s, M, N, K, A, lda,
                       sA, B, ldb, sB, C, ldc, sC, tm, tn, batch);
}

}  // namespace

void launch_dgemm(char opA, int M, int N, int K, const double *A, int lda, size_t strideA,
                  const double *B, int ldb, size_t strideB, double *C, int ldc, size_t strideC,
                  int batch, hipStream_t s) {
    if (M <= 0 || N <= 0 || batch <= 0) return;
    if (opA == 'N')
        launch_dtiles<'N', 2, 2>(M, N, K, A, lda, strideA, B, ldb, strideB, C, ldc, strideC,
                                 batch, s);
    else
        launch_dtiles<'T', 2, 2>(M, N, K, A, lda, strideA, B, ldb, strideB, C, ldc, strideC,
                                 batch, s);
}

namespace {

// IA_GEMM_LOG=<path>: every product is appended to <path> as
// "GEMMLOG <sizeof real> <opA> <opB> M N K batch lda ldb ldc <beta == 0>"
// (the input of the rate benchmark csrc/bench/gemm_rates.hip)
void log_call(int real_bytes, char opA, char opB, int M, int N, int K, int batch, int lda,
              int ldb, int ldc, bool beta0) {
    static const char *path = std::getenv("IA_GEMM_LOG");
    if (!path || !*path) return;
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    if (FILE *f = std::fopen(path, "a")) {
        std::fprintf(f, "GEMMLOG %d %c %c %d %d %d %d %d %d %d %d\n", real_bytes, opA, opB, M, N, K,
                     batch, lda, ldb, ldc, beta0 ? 1 : 0);
        std::fclose(f);
    }
}

}  // namespace

template <typename T>
void launch_gemm(char opA, char opB, int M, int N, int K, cx<T> alpha, const cx<T> *A, int lda,
                 size_t strideA, const cx<T> *B, int ldb, size_t strideB, cx<T> beta, cx<T> *C,
                 int ldc, size_t strideC, int batch, hipStream_t s) {
    if (M <= 0 || N <= 0 || batch <= 0) return;
    log_call((int)sizeof(T), opA, opB, M, N, K, batch, lda, ldb, ldc,
             beta.re == 0 && beta.im == 0);
    if (opA == 'N' && opB == 'N')
        dispatch_tiles<T, 'N', 'N'>(M, N, K, alpha, A, lda, strideA, B, ldb, strideB, beta, C,
                                    ldc, strideC, batch, s);
    else if (opA == 'C' && opB == 'N')
        dispatch_tiles<T, 'C', 'N'>(M, N, K, alpha, A, lda, strideA, B, ldb, strideB, beta, C,
                                    ldc, strideC, batch, s);
    else if (opA == 'N' && opB == 'C')
        dispatch_tiles<T, 'N', 'C'>(M, N, K, alpha, A, lda, strideA, B, ldb, strideB, beta, C,
                                    ldc, strideC, batch, s);
    else
        dispatch_tiles<T, 'C', 'C'>(M, N, K, alpha, A, lda, strideA, B, ldb, strideB, beta, C,
                                    ldc, strideC, batch, s);
}

template void launch_gemm<double>(char, char, int, int, int, cx<double>, const cx<double> *, int,
                                  size_t, const cx<double> *, int, size_t, cx<double>,
                                  cx<double> *, int, size_t, int, hipStream_t);
template void launch_gemm<float>(char, char, int, int, int, cx<float>, const cx<float> *, int,
                                 size_t, const cx<float> *, int, size_t, cx<float>, cx<float> *,
                                 int, size_t, int, hipStream_t);

}  // namespace ia
