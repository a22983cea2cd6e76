// Batched complex SVD building blocks.
//
// Replaces the reference's cusolverDnZgesvd ('S','A') in csvd
// (CUDA/tsvd.cuh:39,93) and LAPACK zgesvd_ in invr (CUDA/invr.cuh:889).
// Algorithm (all frequencies of a batch in one launch per stage):
//   1. Householder bidiagonalisation Y = U_b B V_b^H (R >= C), B real upper
//      bidiagonal (LAPACK gebrd conventions: reflectors H = I - tau u u^H
//      with u_0 = 1, beta real), after a QR pre-reduction of tall Y
//      (geqrf.hip, gebrd.hip);
//   2. singular values of B by bisection on the Golub-Kahan tridiagonal
//      T = tridiag(d_0, e_0, d_1, e_1, ...) whose eigenvalues are +-sigma
//      (Sturm counts; one thread per singular value);
//   3. singular vectors of B by inverse iteration on T - sigma I
//      (pivoted tridiagonal LU, three solves or two in the complex64 form,
//      one thread per vector); the
//      eigenvector interleaves (q_0, p_0, q_1, p_1, ...) with B q = sigma p;
//   4. back-transformation: only the vectors a solve needs (U^H b, V phi)
//      are pushed through the reflectors.
#include "householder.hpp"

namespace ia {

namespace {

using hh::BT;
using hh::NW;
using hh::Refl;
using hh::block_sum;
using hh::larfg;
using hh::wave_sum;

// ------------------------------------------------------------ bisection
// #eigenvalues of the zero-diagonal symmetric tridiagonal with squared
// off-diagonals t2[0..N-2] that are < x.  t2[i] / q is formed from the
// hardware reciprocal refined by two Newton steps (|q| >= pivmin > 0), a
// relative error of a few ulps that the count tolerates like rounding.
__device__ __forceinline__ double recip(double q) {
    double r = __builtin_amdgcn_rcp(q);
    r = __builtin_fma(r, __builtin_fma(-q, r, 1.0), r);
    r = __builtin_fma(r, __builtin_fma(-q, r, 1.0), r);
    return r;
}

__device__ __forceinline__ int sturm_count(int N, const double *t2, double x, double pivmin) {
    int cnt = 0;
    double q = -x;
    if (q < 0) cnt++;
    for (int i = 0; i < N - 1; ++i) {
        if (fabs(q) < pivmin) q = -pivmin;
        q = __builtin_fma(-t2[i], recip(q), -x);
        if (q < 0) cnt++;
    }
    return cnt;
}

// The count and S = d/dx log|det(T - x I)| = sum_i q_i' / q_i from the same
// recurrence (q_{i+1}' = -1 + t2_i q_i' / q_i^2): a Newton step for the
// root of det is 1 / S.
__device__ __forceinline__ int sturm_count_d(int N, const double *t2, double x, double pivmin,
                                             double &S) {
    int cnt = 0;
    double q = -x, dq = -1.0, sum = 0.0;
    if (q < 0) cnt++;
    for (int i = 0; i < N - 1; ++i) {
        if (fabs(q) < pivmin) q = -pivmin;
        const double r = recip(q);
        sum = __builtin_fma(dq, r, sum);
        dq = __builtin_fma(t2[i] * r * r, dq, -1.0);
        q = __builtin_fma(-t2[i], r, -x);
        if (q < 0) cnt++;
    }
    if (fabs(q) < pivmin) q = -pivmin;
    S = __builtin_fma(dq, recip(q), sum);
    return cnt;
}

// Workgroups of 1024 threads; K consecutive lanes share one value (1024 / K
// values per workgroup, blockIdx.x picks the range): each step the K lanes
// take Sturm counts at the K interior points that cut the interval into
// K + 1 equal parts, so the interval shrinks (K + 1)-fold per step instead
// of 2-fold (K = 1: plain bisection).  Small batches use K > 1 to fill the
// device with a shorter chain per value.
constexpr int GL = 10, GPTS = 1 << GL;
constexpr int GGEO = GPTS / 2, GLIN = GPTS - GGEO;  // geometric / linear grid points
// Stopping width max(rtol hi, atol bound).  NEWTON (one lane per value):
// once the bracket holds only its value and is narrower than 1e-3 of it,
// the Sturm pass also returns the Newton step for det(T - x I), and the
// next point is the Newton iterate when it falls inside the bracket (the
// midpoint otherwise); a step below rtol x / 4 ends the search at the
// iterate (quadratic convergence leaves it far inside the tolerance).
template <int K, bool NEWTON>
__global__ __launch_bounds__(1024) void k_bisect(int n, const double *__restrict__ dall,
                                                const double *__restrict__ eall,
                                                double *__restrict__ sall, double rtol,
                                                double atol) {
    extern __shared__ double t2[];
    __shared__ int cgrid[GPTS];
    __shared__ double s_bound, s_pivmin;
    const int b = blockIdx.y;
    const double *d = dall + (size_t)b * n;
    const double *e = eall + (size_t)b * n;
    const int N = 2 * n;
    const int t = threadIdx.x;
    double lmax = 0, gmax = 0;
    for (int i = t; i < N - 1; i += blockDim.x) {
        double v = (i & 1) ? e[i >> 1] : d[i >> 1];
        t2[i] = v * v;
        lmax = fmax(lmax, v * v);
    }
    for (int i = t; i < n; i += blockDim.x) {
        double g = fabs(d[i]) + (i > 0 ? fabs(e[i - 1]) : 0.0);
        double g2 = fabs(d[i]) + (i < n - 1 ? fabs(e[i]) : 0.0);
        gmax = fmax(gmax, fmax(g, g2));
    }
    __shared__ double red[2][1024];
    red[0][t] = lmax;
    red[1][t] = gmax;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (t < s) {
            red[0][t] = fmax(red[0][t], red[0][t + s]);
            red[1][t] = fmax(red[1][t], red[1][t + s]);
        }
        __syncthreads();
    }
    if (t == 0) {
        s_bound = red[1][0] * (1.0 + 8.0 * 2.220446049250313e-16) + 1e-300;
        s_pivmin = fmax(red[0][0] * 2.2250738585072014e-308, 1e-300);
    }
    __syncthreads();
    const double bound = s_bound, pivmin = s_pivmin;
    // every value's first bisection levels in one step: Sturm counts on a
    // shared grid, one point per thread.  Above bound / GLIN the grid is
    // linear (GLIN points); below it geometric, 8 points per octave over 64
    // octaves, so that small values also start from an interval of ~9 %
    // relative width instead of [0, bound / GLIN]
    auto gpt = [&](int j) -> double {
        if (j <= 0) return 0.0;
        if (j >= GPTS) return bound;
        if (j > GGEO) return bound * ((double)(j - GGEO) / GLIN);
        return (bound / GLIN) * exp2(-(double)(GGEO + 1 - j) / 8.0);
    };
    for (int j = t + 1; j < GPTS; j += blockDim.x)
        cgrid[j] = sturm_count(N, t2, gpt(j), pivmin) - n;
    __syncthreads();
    static_assert(K == 1 || K == 2 || K == 4 || K == 8, "lanes per value");
    const int jk = t % K;  // this lane's point among the value's K
    for (int i = blockIdx.x * (1024 / K) + t / K; i < n; i += gridDim.x * (1024 / K)) {
        const int target = n - 1 - i;  // sigma_i (descending) has target values below it
        // first grid point with more than target values below it (GPTS if none)
        int jl = 0, jh = GPTS;  // cgrid(jl) <= target < cgrid(jh), cgrid(0) = 0, cgrid(GPTS) = n
        while (jh - jl > 1) {
            const int jm = (jl + jh) >> 1;
            if (cgrid[jm] > target)
                jh = jm;
            else
                jl = jm;
        }
        double lo = gpt(jl), hi = gpt(jh);
        int clo = jl > 0 ? cgrid[jl] : 0, chi = jh < GPTS ? cgrid[jh] : n;
        double xn = -1.0, res = -1.0;  // Newton iterate; result once it has converged
        for (int it = 0; it < 128; ++it) {
            if (hi - lo <= fmax(rtol * hi, atol * bound)) break;
            if (K == 1) {
                const bool nt = NEWTON && clo == target && chi == target + 1 && hi - lo <= 1e-3 * hi;
                const double p = (nt && xn > lo && xn < hi) ? xn : 0.5 * (lo + hi);
                double S = 0.0;
                const int c = (nt ? sturm_count_d(N, t2, p, pivmin, S) : sturm_count(N, t2, p, pivmin)) - n;
                if (c > target) {
                    hi = p;
                    chi = c;
                } else {
                    lo = p;
                    clo = c;
                }
                if (nt) {
                    const double step = 1.0 / S;
                    xn = p - step;
                    if (fabs(step) <= 0.25 * rtol * p && xn >= lo && xn <= hi) {
                        res = xn;
                        break;
                    }
                }
            } else {
                // the K lanes of a value are consecutive and share lo, hi and
                // the loop exit, so they reach the ballot together
                const double w = (hi - lo) / (K + 1);
                const double pj = lo + w * (jk + 1);
                const bool above = sturm_count(N, t2, pj, pivmin) - n > target;
                const unsigned long long bal = __ballot(above);
                const int base = (threadIdx.x & 63) - jk;
                const unsigned grp = (unsigned)(bal >> base) & ((1u << K) - 1);
                // first point with more than target values below it
                const int f = grp ? __builtin_ctz(grp) : K;
                const double nlo = f == 0 ? lo : lo + w * f;
                const double nhi = f == K ? hi : lo + w * (f + 1);
                lo = nlo;
                hi = nhi;
            }
        }
        if (jk == 0) sall[(size_t)b * n + i] = res >= 0.0 ? res : 0.5 * (lo + hi);
    }
}

// ----------------------------------------------------- inverse iteration
// One thread per singular value: pivoted LU of T - sigma I (N = 2n), three
// solves (two with float storage) from a deterministic pseudo-random start,
// split into q (even) and p (odd).  Stored per row: U's diagonal, the
// multiplier, the pivot flag and the iterate.  Scratch is one slab per wave with the 64 lanes interleaved
// (element k of lane l at (wave N + k) 64 + l): every step is one 512-byte
// coalesced access and each wave streams through consecutive memory.
// G is the scratch column count (64 per wave).  Outputs QBt, PBt hold the
// vectors as ROWS: QBt[i + k n] = q_i[k].
__device__ __forceinline__ double hash01(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    x ^= x >> 31;
    return (double)(x >> 11) * (1.0 / 9007199254740992.0) - 0.5;
}

// Every pass walks the N steps in blocks of IU: the block's operands are
// loaded into registers first (all loads of a block in flight together),
// then the serial recurrence runs on registers; the pivot flags of the
// backward pass are fetched one block ahead so that the conditional loads of
// the second superdiagonal need no extra round trip.
constexpr int IU = 6;

// Four waves per SIMD (at most 128 VGPRs): 256 frequencies x 1024 singular
// values are 4096 single-wave workgroups, one round on 1024 SIMDs.
// S: storage type of the factorization and the iterate (the arithmetic is
// FP64 either way).  float halves the scratch traffic, which bounds the
// kernel; the mixed form uses it, its bidiagonal carrying complex64
// rounding of the same order already.
template <typename S>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_inviter(int n, int b0, const double *__restrict__ dall,
                                               const double *__restrict__ eall,
                                               const double *__restrict__ sall,
                                               double *__restrict__ QBt, double *__restrict__ PBt,
                                               void *__restrict__ workv, int G) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int bl = blockIdx.y;       // local batch index within this call
    const int b = b0 + bl;
    if (i >= n) return;
    const int lane = threadIdx.x;
    const size_t wid = (size_t)bl * gridDim.x + blockIdx.x;  // this wave's scratch slab
    const int N = 2 * n;
    const double *d = dall + (size_t)b * n;
    const double *e = eall + (size_t)b * n;
    const double sigma = sall[(size_t)b * n + i];
    S *work = reinterpret_cast<S *>(workv);
    // U's superdiagonals are not stored: the backward solve rebuilds them
    // from the pivot flags, the multipliers and T (see backward below)
    S *DD = work;                      // U diagonal
    S *DL = work + (size_t)N * G;      // multipliers
    S *X = work + (size_t)2 * N * G;   // iterate
    unsigned char *PV = reinterpret_cast<unsigned char *>(work + (size_t)3 * N * G);
#define AT(arr, k) arr[(wid * N + (k)) * 64 + lane]
    auto toff = [&](int k) { return (k & 1) ? e[k >> 1] : d[k >> 1]; };
    double tnorm = 0;
    for (int k = 0; k < N - 1; ++k) tnorm = fmax(tnorm, fabs(toff(k)));
    const double eps = 2.220446049250313e-16;
    // (never below the storage type's smallest normal: a stored pivot must
    // stay invertible)
    const double tiny = fmax(eps * tnorm, sizeof(S) == 4 ? 1e-37 : 1e-300);
    auto clampd = [&](double v) { return fabs(v) < tiny ? (v < 0 ? -tiny : tiny) : v; };
    auto start = [&](int k) {
        return hash01(((uint64_t)b << 40) ^ ((uint64_t)i << 20) ^ k) + 0.25;
    };
    const int Np = (N + IU - 1) / IU * IU;
    // ---- one pass: LU of T - sigma I with partial pivoting (LAPACK gttrf),
    // the random start vector and the first solve with L
    {
        double dd = -sigma, du = toff(0), x = start(0);
        for (int kb = 0; kb < N - 1; kb += IU) {
            double tv[IU + 1];  // off-diagonals k = kb .. kb + IU (0 from N - 1 on)
#pragma unroll
            for (int u = 0; u <= IU; ++u) tv[u] = (kb + u < N - 1) ? toff(kb + u) : 0.0;
#pragma unroll
            for (int u = 0; u < IU; ++u) {
                const int k = kb + u;
                if (k >= N - 1) break;
                const double dlk = tv[u], ddn = -sigma, dun = tv[u + 1];
                const double xn = start(k + 1);
                if (fabs(dd) >= fabs(dlk)) {
                    const double fact = dd != 0.0 ? dlk / dd : 0.0;
                    AT(PV, k) = 0;
                    AT(DL, k) = fact;
                    AT(DD, k) = clampd(dd);
                    AT(X, k) = x;
                    x = xn - fact * x;
                    dd = ddn - fact * du;
                    du = dun;
                } else {
                    const double fact = dd / dlk;
                    AT(PV, k) = 1;
                    AT(DL, k) = fact;
                    AT(DD, k) = clampd(dlk);
                    AT(X, k) = xn;
                    x = x - fact * xn;
                    dd = du - fact * ddn;
                    du = -fact * dun;
                }
            }
        }
        AT(DD, N - 1) = clampd(dd);
        AT(X, N - 1) = x;
    }
    // backward solve with U; returns 1 / max |x|.  Row k of U holds
    // (DD_k, u_k, w_k): pivoted, u_k = -sigma and w_k = t_{k+1} (the rows
    // swapped in); unpivoted, w_k = 0 and u_k = t_k, or -l_{k-1} t_k when
    // row k - 1 was pivoted (t_k: T's off-diagonal k, l: the multipliers).
    // Pivot flags are fetched two blocks ahead, so that the flag of the row
    // below each block is known before the block's conditional loads.
    auto backward = [&]() {
        double scale = 0, x2 = 0, x1 = 0;  // x[k+2], x[k+1]
        auto pvat = [&](int k) { return (k >= 0 && k < N - 1) ? (int)AT(PV, k) : 0; };
        int pvc[IU], pvn[IU];
#pragma unroll
        for (int u = 0; u < IU; ++u) {
            pvc[u] = pvat(Np - IU + u);
            pvn[u] = pvat(Np - 2 * IU + u);
        }
        for (int kb = Np - IU; kb >= 0; kb -= IU) {
            double xv[IU], dd[IU], dlm[IU], tk[IU + 1];
            int pnn[IU];
#pragma unroll
            for (int u = 0; u <= IU; ++u) tk[u] = (kb + u < N - 1) ? toff(kb + u) : 0.0;
#pragma unroll
            for (int u = 0; u < IU; ++u) {
                const int k = kb + u;
                const int pm = u > 0 ? pvc[u - 1] : pvn[IU - 1];  // row k - 1 pivoted
                xv[u] = 0.0;
                dd[u] = 1.0;
                dlm[u] = 0.0;
                if (k < N) {
                    xv[u] = AT(X, k);
                    dd[u] = AT(DD, k);
                    if (pm && k >= 1) dlm[u] = AT(DL, k - 1);
                }
                pnn[u] = pvat(kb - 2 * IU + u);
            }
#pragma unroll
            for (int u = IU - 1; u >= 0; --u) {
                const int k = kb + u;
                if (k >= N) continue;
                const int pm = u > 0 ? pvc[u - 1] : pvn[IU - 1];
                double sv = xv[u];
                if (k < N - 1) {
                    const double uk = pvc[u] ? -sigma : pm ? -dlm[u] * tk[u] : tk[u];
                    sv -= uk * x1;
                    if (pvc[u]) sv -= tk[u + 1] * x2;
                }
                sv /= dd[u];
                AT(X, k) = sv;
                x2 = x1;
                x1 = sv;
                scale = fmax(scale, fabs(sv));
            }
#pragma unroll
            for (int u = 0; u < IU; ++u) {
                pvc[u] = pvn[u];
                pvn[u] = pnn[u];
            }
        }
        return scale > 0 ? 1.0 / scale : 1.0;
    };
    // forward solve with L of the iterate scaled by inv
    auto forward = [&](double inv) {
        double x = AT(X, 0) * inv;
        for (int kb = 0; kb < N - 1; kb += IU) {
            double xn[IU], fl[IU];
            int pv[IU];
#pragma unroll
            for (int u = 0; u < IU; ++u) {
                const int k = kb + u;
                if (k < N - 1) {
                    xn[u] = AT(X, k + 1);
                    fl[u] = AT(DL, k);
                    pv[u] = AT(PV, k);
                }
            }
#pragma unroll
            for (int u = 0; u < IU; ++u) {
                const int k = kb + u;
                if (k >= N - 1) break;
                const double xk = xn[u] * inv, fk = fl[u];
                if (pv[u] == 0) {
                    AT(X, k) = x;
                    x = xk - fk * x;
                } else {
                    AT(X, k) = xk;
                    x = x - fk * xk;
                }
            }
        }
        AT(X, N - 1) = x;
    };
    // three solves with FP64 storage (LAPACK stein: the first reaches its
    // growth test, two more follow); with float storage the stored
    // factors' own rounding (~6e-8 |T| / gap) bounds the vector already
    // after the second, whose start-vector residue is that bound squared
    double inv = backward();
    forward(inv);
    inv = backward();
    if (sizeof(S) == sizeof(double)) {
        forward(inv);
        inv = backward();
    }
    // ---- split, normalise, fix the sign of p so that p^T B q >= 0
    constexpr int HU = IU / 2;  // singular-vector entries per block
    double nq = 0, np = 0, sgn = 0;
    for (int kb = 0; kb < n; kb += HU) {
        double xv[2 * HU + 1], dv[HU], ev[HU];
#pragma unroll
        for (int u = 0; u <= 2 * HU; ++u) xv[u] = (2 * kb + u < N) ? AT(X, 2 * kb + u) : 0.0;
#pragma unroll
        for (int u = 0; u < HU; ++u) {
            dv[u] = (kb + u < n) ? d[kb + u] : 0.0;
            ev[u] = (kb + u < n - 1) ? e[kb + u] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < HU; ++u) {
            const int k = kb + u;
            if (k >= n) break;
            const double q = xv[2 * u] * inv, p = xv[2 * u + 1] * inv;
            const double qn = k < n - 1 ? xv[2 * u + 2] * inv : 0.0;
            nq += q * q;
            np += p * p;
            sgn += p * (dv[u] * q + (k < n - 1 ? ev[u] * qn : 0.0));
        }
    }
    nq = sqrt(nq);
    np = sqrt(np);
    const double iq = nq > 0 ? inv / nq : 0.0;
    double ip = np > 0 ? inv / np : 0.0;
    if (sgn < 0) ip = -ip;
    double *Qb = QBt + (size_t)b * n * n;
    double *Pb = PBt + (size_t)b * n * n;
    for (int kb = 0; kb < n; kb += HU) {
        double xv[2 * HU];
#pragma unroll
        for (int u = 0; u < 2 * HU; ++u) xv[u] = (2 * kb + u < N) ? AT(X, 2 * kb + u) : 0.0;
#pragma unroll
        for (int u = 0; u < HU; ++u) {
            const int k = kb + u;
            if (k >= n) break;
            Qb[(size_t)k * n + i] = xv[2 * u] * iq;
            Pb[(size_t)k * n + i] = xv[2 * u + 1] * ip;
        }
    }
#undef AT
}

// --------------------------------------------- clusters of multiple values
// Inverse iteration resolves each vector to an angle of about eps s_1 / gap.
// For an exact multiple singular value (split only by rounding, a gap of a
// few eps s_1) every vector of the multiplet comes out as a different random
// member of its invariant subspace, so the vectors are independent but not
// orthogonal.  Members of such a cluster -- consecutive values closer than
// tau_abs s_1 and tau_rel s_i, at or above floor_rel s_1 -- are made
// orthonormal by two passes of modified Gram-Schmidt on the Golub-Kahan
// vectors (q_i, p_i) (LAPACK's stein reorthogonalises within clusters in
// the same way).  Any orthonormal basis of the multiplet's subspace serves:
// the filtered solution weights all of its vectors alike.  One workgroup
// per matrix; the cluster walk is uniform over the workgroup.
// SEP: q and p are orthonormalised each on its own (used when the full
// singular bases are returned: for tiny values the Golub-Kahan vectors also
// carry components of the -s partners, so orthogonal (q, p) pairs need not
// have orthogonal q and p parts).
template <bool SEP>
__global__ __launch_bounds__(BT) void k_cluster_orth(int n, const double *__restrict__ sall,
                                                     double *__restrict__ QBt,
                                                     double *__restrict__ PBt, double tau_abs,
                                                     double tau_rel, double floor_rel) {
    __shared__ double red[NW];
    const int b = blockIdx.x, t = threadIdx.x;
    const double *sg = sall + (size_t)b * n;
    double *Q = QBt + (size_t)b * n * n;
    double *P = PBt + (size_t)b * n * n;
    const double s1 = sg[0];
    if (!(s1 > 0)) return;
    const double thr = tau_abs * s1, lo = floor_rel * s1;
    int i0 = 0;
    while (i0 < n) {
        int i1 = i0 + 1;
        while (i1 < n && sg[i1] >= lo && sg[i1 - 1] - sg[i1] <= thr &&
               sg[i1 - 1] - sg[i1] <= tau_rel * sg[i1])
            ++i1;
        if (i1 - i0 >= 2 && sg[i0] >= lo) {
            for (int j = i0 + 1; j < i1; ++j) {
                for (int pass = 0; pass < 2; ++pass) {
                    for (int i = i0; i < j; ++i) {
                        double dq = 0, dp = 0;
                        for (int k = t; k < n; k += BT) {
                            dq += Q[i + (size_t)k * n] * Q[j + (size_t)k * n];
                            dp += P[i + (size_t)k * n] * P[j + (size_t)k * n];
                        }
                        double cq, cp;
                        if (SEP) {
                            cq = block_sum(dq, red);
                            cp = block_sum(dp, red);
                        } else {
                            // |(q_i, p_i)|^2 = 2: the projection coefficient is d / 2
                            cq = cp = 0.5 * block_sum(dq + dp, red);
                        }
                        for (int k = t; k < n; k += BT) {
                            Q[j + (size_t)k * n] -= cq * Q[i + (size_t)k * n];
                            P[j + (size_t)k * n] -= cp * P[i + (size_t)k * n];
                        }
                    }
                }
                double nq = 0, np = 0;
                for (int k = t; k < n; k += BT) {
                    const double q = Q[j + (size_t)k * n], pp = P[j + (size_t)k * n];
                    nq += q * q;
                    np += pp * pp;
                }
                nq = block_sum(nq, red);
                np = block_sum(np, red);
                const double iq = nq > 0 ? 1.0 / sqrt(nq) : 0.0, ip = np > 0 ? 1.0 / sqrt(np) : 0.0;
                for (int k = t; k < n; k += BT) {
                    Q[j + (size_t)k * n] *= iq;
                    P[j + (size_t)k * n] *= ip;
                }
            }
        }
        i0 = i1;
    }
}

// --------------------------------------------------- reflector application
// v <- U_b v (FWD=false: H_0 H_1 ... H_{C-1} v) or U_b^H v, length-R vectors;
// reflector j lives in column j of Y below the diagonal (unit at row j).
template <typename T, bool HERM>
__global__ __launch_bounds__(256) void k_apply_left(int R, int C, const cx<T> *__restrict__ Yall,
                                                   int ldy, size_t sY,
                                                   const cx<T> *__restrict__ tauall,
                                                   cx<double> *__restrict__ vall, size_t sv,
                                                   int nvec) {
    __shared__ double red[2][4];
    const int vec = blockIdx.x, b = blockIdx.y;
    const cx<T> *Y = Yall + (size_t)b * sY;
    const cx<T> *tau = tauall + (size_t)b * C;
    cx<double> *v = vall + (size_t)b * nvec * sv + (size_t)vec * sv;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int jj = 0; jj < C; ++jj) {
        const int j = HERM ? jj : (C - 1 - jj);
        cx<double> tj = cvt<double>(tau[j]);
        if (tj.re == 0.0 && tj.im == 0.0) continue;
        const cx<T> *uc = Y + (size_t)j * ldy;
        double ar = 0, ai = 0;
        for (int r = j + t; r < R; r += 256) {
            cx<double> uu = (r == j) ? cx<double>{1, 0} : cvt<double>(uc[r]);
            cx<double> x = v[r];
            ar += uu.re * x.re + uu.im * x.im;
            ai += uu.re * x.im - uu.im * x.re;
        }
        ar = wave_sum(ar);
        ai = wave_sum(ai);
        __syncthreads();
        if (lane == 0) {
            red[0][w] = ar;
            red[1][w] = ai;
        }
        __syncthreads();
        cx<double> dot{red[0][0] + red[0][1] + red[0][2] + red[0][3],
                       red[1][0] + red[1][1] + red[1][2] + red[1][3]};
        cx<double> f = (HERM ? conj(tj) : tj) * dot;
        for (int r = j + t; r < R; r += 256) {
            cx<double> uu = (r == j) ? cx<double>{1, 0} : cvt<double>(uc[r]);
            v[r] = v[r] - uu * f;
        }
        __syncthreads();  // the next reflector's dot product reads rows other waves wrote
    }
}

// Same product with the vector held in registers (row r = t + 256 q in
// x[q], R <= 256 RPT) and the reflector columns streamed through a ring of
// PD register slots, each loaded PD reflectors ahead of its use, so that the
// memory latency overlaps the preceding reflectors' arithmetic.  One barrier
// per reflector: the four waves' partial dot products alternate between two
// LDS slots.
template <typename T, bool HERM, int RPT, int PD>
__global__ __launch_bounds__(256) void k_apply_left_reg(int R, int C,
                                                       const cx<T> *__restrict__ Yall, int ldy,
                                                       size_t sY, const cx<T> *__restrict__ tauall,
                                                       cx<double> *__restrict__ vall, size_t sv,
                                                       int nvec) {
    __shared__ double red[2][2][4];
    const int vec = blockIdx.x, b = blockIdx.y;
    const cx<T> *Y = Yall + (size_t)b * sY;
    const cx<T> *tau = tauall + (size_t)b * C;
    cx<double> *v = vall + (size_t)b * nvec * sv + (size_t)vec * sv;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    cx<double> x[RPT];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int r = t + 256 * q;
        x[q] = r < R ? v[r] : cx<double>{0, 0};
    }
    auto refl = [&](int jj) { return HERM ? jj : (C - 1 - jj); };
    cx<T> ring[PD][RPT];
    cx<T> rtau[PD];
    auto fetch = [&](int jj, cx<T> *u, cx<T> &tj) {
        if (jj < C) {
            const int j = refl(jj);
            const cx<T> *uc = Y + (size_t)j * ldy;
            tj = tau[j];
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int r = t + 256 * q;
                u[q] = (r > j && r < R) ? uc[r] : cx<T>{0, 0};
            }
        }
    };
#pragma unroll
    for (int sl = 0; sl < PD; ++sl) fetch(sl, ring[sl], rtau[sl]);
    for (int jj0 = 0; jj0 < C; jj0 += PD) {
#pragma unroll
        for (int sl = 0; sl < PD; ++sl) {
            const int jj = jj0 + sl;
            if (jj >= C) break;
            const int j = refl(jj);
            cx<T> u[RPT];
#pragma unroll
            for (int q = 0; q < RPT; ++q) u[q] = ring[sl][q];
            const cx<double> tj = cvt<double>(rtau[sl]);
            fetch(jj + PD, ring[sl], rtau[sl]);
            if (tj.re == 0.0 && tj.im == 0.0) continue;  // uniform: every thread read tau[j]
            double ar = 0, ai = 0;
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int r = t + 256 * q;
                const cx<double> uu = (r == j) ? cx<double>{1, 0} : cvt<double>(u[q]);
                if (r >= j && r < R) {
                    ar += uu.re * x[q].re + uu.im * x[q].im;
                    ai += uu.re * x[q].im - uu.im * x[q].re;
                }
            }
            ar = wave_sum(ar);
            ai = wave_sum(ai);
            const int pb = jj & 1;
            if (lane == 0) {
                red[pb][0][w] = ar;
                red[pb][1][w] = ai;
            }
            __syncthreads();
            const cx<double> dot{red[pb][0][0] + red[pb][0][1] + red[pb][0][2] + red[pb][0][3],
                                 red[pb][1][0] + red[pb][1][1] + red[pb][1][2] + red[pb][1][3]};
            const cx<double> f = (HERM ? conj(tj) : tj) * dot;
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int r = t + 256 * q;
                const cx<double> uu = (r == j) ? cx<double>{1, 0} : cvt<double>(u[q]);
                if (r >= j && r < R) x[q] = x[q] - uu * f;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int r = t + 256 * q;
        if (r < R) v[r] = x[q];
    }
}

// v <- V_b v or V_b^H v, length-C vectors; reflector j lives in Vr column j,
// rows j+1..C-1 (unit at row j+1).
template <typename T, bool HERM>
__global__ __launch_bounds__(256) void k_apply_right(int C, const cx<T> *__restrict__ Vrall,
                                                    const cx<T> *__restrict__ tauall,
                                                    cx<double> *__restrict__ vall, size_t sv,
                                                    int nvec) {
    __shared__ double red[2][4];
    const int vec = blockIdx.x, b = blockIdx.y;
    const cx<T> *Vr = Vrall + (size_t)b * C * C;
    const cx<T> *tau = tauall + (size_t)b * C;
    cx<double> *v = vall + (size_t)b * nvec * sv + (size_t)vec * sv;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int jj = 0; jj < C - 1; ++jj) {
        const int j = HERM ? jj : (C - 2 - jj);
        cx<double> tj = cvt<double>(tau[j]);
        if (tj.re == 0.0 && tj.im == 0.0) continue;
        const cx<T> *vc = Vr + (size_t)j * C;
        double ar = 0, ai = 0;
        for (int r = j + 1 + t; r < C; r += 256) {
            cx<double> uu = cvt<double>(vc[r]);
            cx<double> x = v[r];
            ar += uu.re * x.re + uu.im * x.im;
            ai += uu.re * x.im - uu.im * x.re;
        }
        ar = wave_sum(ar);
        ai = wave_sum(ai);
        __syncthreads();
        if (lane == 0) {
            red[0][w] = ar;
            red[1][w] = ai;
        }
        __syncthreads();
        cx<double> dot{red[0][0] + red[0][1] + red[0][2] + red[0][3],
                       red[1][0] + red[1][1] + red[1][2] + red[1][3]};
        cx<double> f = (HERM ? conj(tj) : tj) * dot;
        for (int r = j + 1 + t; r < C; r += 256) {
            cx<double> uu = cvt<double>(vc[r]);
            v[r] = v[r] - uu * f;
        }
        __syncthreads();  // the next reflector's dot product reads rows other waves wrote
    }
}

// Unblocked Householder QR A = Q R (R >= C), one workgroup per matrix, the
// LAPACK geqr2 conventions of the bidiagonalisation's left reflectors:
// reflector j in column j below the diagonal (unit at row j), R on and above
// the diagonal.  Used for the CHIEF least-squares solves (the reference's
// cusolverDnZZgels, CUDA/ibem.cuh:389-391).
template <typename T>
__global__ __launch_bounds__(BT) void k_geqr2(int R, int C, cx<T> *__restrict__ Aall, int lda,
                                             size_t sA, cx<T> *__restrict__ tauall) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *u = reinterpret_cast<cx<T> *>(smem_raw);
    cx<T> *w = u + R;
    double *red = reinterpret_cast<double *>(w + C);
    __shared__ cx<double> bc_val;
    const int b = blockIdx.x;
    cx<T> *A = Aall + (size_t)b * sA;
    cx<T> *tau = tauall + (size_t)b * C;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    for (int j = 0; j < C; ++j) {
        cx<T> *Aj = A + (size_t)j * lda;
        double p = 0;
        for (int r = j + 1 + t; r < R; r += BT) p += (double)abs2(Aj[r]);
        double xn2 = block_sum(p, red);
        if (t == 0) bc_val = cvt<double>(Aj[j]);
        __syncthreads();
        Refl rf = larfg(bc_val, xn2);
        const cx<T> sc = cvt<T>(rf.scale);
        for (int r = j + 1 + t; r < R; r += BT) {
            cx<T> uv = Aj[r] * sc;
            u[r] = uv;
            Aj[r] = uv;
        }
        if (t == 0) {
            u[j] = cx<T>{1, 0};
            Aj[j] = cx<T>{(T)rf.beta, 0};
            tau[j] = cvt<T>(rf.tau);
        }
        __syncthreads();
        for (int c = j + 1 + wv; c < C; c += NW) {
            const cx<T> *Ac = A + (size_t)c * lda;
            double ar = 0, ai = 0;
            for (int r = j + lane; r < R; r += 64) {
                cx<T> y = Ac[r], uu = u[r];
                ar += (double)(uu.re * y.re + uu.im * y.im);
                ai += (double)(uu.re * y.im - uu.im * y.re);
            }
            ar = wave_sum(ar);
            ai = wave_sum(ai);
            if (lane == 0) w[c] = cx<T>{(T)ar, (T)ai};
        }
        __syncthreads();
        const cx<T> ctau = conj(cvt<T>(rf.tau));
        for (int r = j + t; r < R; r += BT) {
            const cx<T> cu = ctau * u[r];
            for (int c = j + 1; c < C; ++c) {
                cx<T> *yp = A + (size_t)c * lda + r;
                *yp = *yp - cu * w[c];
            }
        }
        __syncthreads();
    }
}

}  // namespace


void launch_bisect(int n, const double *d, const double *e, double *sigma, int batch,
                   hipStream_t s, bool f32) {
    size_t lds = sizeof(double) * (2 * (size_t)n);
    // FP64 form: full relative precision (down to 1e-3 eps of the bound).
    // Complex64 form: 1e-12 relative or 1e-13 of the bound, a quarter fewer
    // steps; the selection and the inverse-iteration shifts still see the
    // values of the (complex64-rounded) bidiagonal to well below its own
    // rounding (1e-9 relative measurably moved nearly tied GCV choices)
    const double eps = 2.220446049250313e-16;
    const double rtol = f32 ? 1e-12 : 2.0 * eps, atol = f32 ? 1e-13 : 1e-3 * eps;
    // lanes per value: the most that keeps n * batch * K threads within
    // about four waves per SIMD of the device (a full batch stays at plain
    // bisection, which already fills it; the recheck's few frequencies get
    // eight lanes and a third of the steps)
    const double vals = (double)n * batch;
    const int K = vals * 8 <= 262144.0 ? 8 : vals * 4 <= 262144.0 ? 4 : vals * 2 <= 262144.0 ? 2 : 1;
    const unsigned gx = (unsigned)((n + 1024 / K - 1) / (1024 / K));
    auto go = [&](auto kern) {
        (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds);
        hipLaunchKernelGGL(kern, dim3(gx, batch), dim3(1024), lds, s, n, d, e, sigma, rtol, atol);
    };
    // complex64 form at full batches: Newton steps once a value is isolated
    if (K == 8)
        go(k_bisect<8, false>);
    else if (K == 4)
        go(k_bisect<4, false>);
    else if (K == 2)
        go(k_bisect<2, false>);
    else if (f32)
        go(k_bisect<1, true>);
    else
        go(k_bisect<1, false>);
}

// scratch columns: 64 per wave, (n + 63) / 64 waves per matrix
static size_t inviter_cols(int n, int sub) { return (size_t)(n + 63) / 64 * 64 * sub; }

size_t inviter_work_bytes(int n, int sub, bool f32) {
    size_t G = inviter_cols(n, sub), N = 2 * (size_t)n;
    return 3 * N * G * (f32 ? sizeof(float) : sizeof(double)) + N * G + 256;
}

void launch_inviter(int n, const double *d, const double *e, const double *sigma, double *QBt,
                    double *PBt, void *work, int sub, int batch, hipStream_t s, bool f32) {
    for (int b0 = 0; b0 < batch; b0 += sub) {
        int nb = min(sub, batch - b0);
        dim3 grid((n + 63) / 64, nb);
        if (f32)
            hipLaunchKernelGGL(k_inviter<float>, grid, dim3(64), 0, s, n, b0, d, e, sigma, QBt, PBt,
                               work, (int)inviter_cols(n, sub));
        else
            hipLaunchKernelGGL(k_inviter<double>, grid, dim3(64), 0, s, n, b0, d, e, sigma, QBt,
                               PBt, work, (int)inviter_cols(n, sub));
    }
    // multiplets: values within 1e-12 s_1 (and 1e-3 relative) of each other,
    // at or above the pooling floor of the precision form
    hipLaunchKernelGGL(k_cluster_orth<false>, dim3(batch), dim3(BT), 0, s, n, sigma, QBt, PBt,
                       1e-12, 1e-3, f32 ? 1e-6 : 1e-12);
}

void launch_orthonormalise_tail(int n, const double *sigma, double *QBt, double *PBt, int batch,
                                hipStream_t s) {
    // every run of values closer than 1e-4 s_1 to its neighbour: beyond such
    // gaps inverse iteration leaves orthogonality errors below ~1e-12
    hipLaunchKernelGGL(k_cluster_orth<true>, dim3(batch), dim3(BT), 0, s, n, sigma, QBt, PBt, 1e-4,
                       1e300, 0.0);
}

template <typename T>
ia_status launch_geqr2(int R, int C, cx<T> *A, int lda, size_t strideA, cx<T> *tau, int batch,
                       hipStream_t s) {
    size_t lds = ((size_t)R + C) * sizeof(cx<T>) + NW * sizeof(double) + 64;
    if (lds > 160 * 1024) return fail(IA_ERR_ARG, "QR of %d x %d exceeds the on-chip limit", R, C);
    (void)hipFuncSetAttribute((const void *)k_geqr2<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    hipLaunchKernelGGL(k_geqr2<T>, dim3(batch), dim3(BT), lds, s, R, C, A, lda, strideA, tau);
    return IA_OK;
}

template <typename T>
void launch_apply_left(bool herm, int R, int C, const cx<T> *Y, int ldy, size_t strideY,
                       const cx<T> *tauL, cx<double> *v, size_t stridev, int nvec, int batch,
                       hipStream_t s) {
    dim3 grid(nvec, batch);
    constexpr int PD = sizeof(T) == 4 ? 8 : 4;
    auto reg = [&](auto k_h, auto k_n) {
        if (herm)
            hipLaunchKernelGGL(k_h, grid, dim3(256), 0, s, R, C, Y, ldy, strideY, tauL, v, stridev,
                               nvec);
        else
            hipLaunchKernelGGL(k_n, grid, dim3(256), 0, s, R, C, Y, ldy, strideY, tauL, v, stridev,
                               nvec);
    };
    if (R <= 256 * 4)
        reg(k_apply_left_reg<T, true, 4, PD>, k_apply_left_reg<T, false, 4, PD>);
    else if (R <= 256 * 8)
        reg(k_apply_left_reg<T, true, 8, PD>, k_apply_left_reg<T, false, 8, PD>);
    else if (herm)
        hipLaunchKernelGGL((k_apply_left<T, true>), grid, dim3(256), 0, s, R, C, Y, ldy, strideY,
                           tauL, v, stridev, nvec);
    else
        hipLaunchKernelGGL((k_apply_left<T, false>), grid, dim3(256), 0, s, R, C, Y, ldy, strideY,
                           tauL, v, stridev, nvec);
}

template <typename T>
void launch_apply_right(bool herm, int C, const cx<T> *Vr, const cx<T> *tauR, cx<double> *v,
                        size_t stridev, int nvec, int batch, hipStream_t s) {
    dim3 grid(nvec, batch);
    if (herm)
        hipLaunchKernelGGL((k_apply_right<T, true>), grid, dim3(256), 0, s, C, Vr, tauR, v, stridev,
                           nvec);
    else
        hipLaunchKernelGGL((k_apply_right<T, false>), grid, dim3(256), 0, s, C, Vr, tauR, v,
                           stridev, nvec);
}

template ia_status launch_geqr2<double>(int, int, cx<double> *, int, size_t, cx<double> *, int,
                                         hipStream_t);
template ia_status launch_geqr2<float>(int, int, cx<float> *, int, size_t, cx<float> *, int,
                                       hipStream_t);

#define INST(T)                                                                                  \
    template void launch_apply_left<T>(bool, int, int, const cx<T> *, int, size_t,              \
                                       const cx<T> *, cx<double> *, size_t, int, int,           \
                                       hipStream_t);                                            \
    template void launch_apply_right<T>(bool, int, const cx<T> *, const cx<T> *, cx<double> *,  \
                                        size_t, int, int, hipStream_t);
INST(double)
INST(float)
#undef INST

}  // namespace ia
