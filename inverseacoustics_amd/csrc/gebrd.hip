// Panel-blocked Householder bidiagonalisation Y = U_b B V_b^H (R >= C).
//
// Same reflectors and storage as the fused kernel of svd.hip (left
// reflector j in column j of Y below the diagonal with tauL[j], right
// reflector j in row j of Vr with tauR[j], B = bidiag(d, e)), computed with
// the deferred-update scheme of LAPACK's gebrd/labrd:  inside a panel of NB
// columns the trailing matrix is not touched; the matrix the unblocked
// algorithm would see at step i is
//     A_i = A - U Yw^H - Xw Vw^H        (U: left vectors, Vw: right vectors,
//                                        Yw = tau A^H u, Xw = sigma A' v)
// and the two matrix-vector products of a step read the trailing block
// once each.  After the panel, the trailing block is updated by two complex
// GEMMs of rank NB on the matrix cores.  Per column the HBM traffic is two
// reads of the trailing block, against two reads and two writes for the
// fused unblocked kernel.
//
// One 1024-thread workgroup per frequency for the panel kernel.  LDS holds
// the current left vector u (R), the right vector v (C), the panel's small
// dot products and the partial sums of the row-split matrix-vector product.
#include "householder.hpp"

#include <cstdint>
#include <cstdlib>

namespace ia {

namespace {

using hh::BT;
using hh::NW;
using hh::Refl;
using hh::block_sum;
using hh::larfg;
using hh::wave_sum;

constexpr int NB = 32;

// Rows are read in groups of G consecutive entries per lane: G = 2 for
// complex64 (one 16-byte load per lane, as for complex128), else 1.
template <typename T, int G>
struct RowGroup;
template <typename T>
struct RowGroup<T, 1> {
    static __device__ __forceinline__ void load(const cx<T> *col, int g, cx<double> (&o)[1]) {
        o[0] = cvt<double>(col[g]);
    }
};
template <>
struct RowGroup<float, 2> {
    static __device__ __forceinline__ void load(const cx<float> *col, int g, cx<double> (&o)[2]) {
        const float4 a = reinterpret_cast<const float4 *>(col)[g];
        o[0] = cx<double>{a.x, a.y};
        o[1] = cx<double>{a.z, a.w};
    }
};

// Unroll depth of the streaming loops: UL independent 16-byte loads per lane
// are in flight before their products are accumulated (a CU needs ~64 KB in
// flight to cover the HBM latency at its share of the bandwidth).
#ifndef IA_UL32
#define IA_UL32 4
#endif
#ifndef IA_UL64
#define IA_UL64 2
#endif
template <typename T>
constexpr int unroll_of() { return sizeof(T) == sizeof(double) ? IA_UL64 : IA_UL32; }

// sum_{r >= j} conj(u_r) col_r by one wave (rows below j of u are zero)
template <typename T, int G>
__device__ __forceinline__ cx<double> wave_dot_u(const cx<T> *col, const cx<T> *u, int j, int R,
                                                  int lane) {
    constexpr int UL = unroll_of<T>();
    double ar = 0, ai = 0;
    const int ng = R / G;
    int g = j / G + lane;
    for (; g + 64 * (UL - 1) < ng; g += 64 * UL) {
        cx<double> y[UL][G], uu[UL][G];
#pragma unroll
        for (int q = 0; q < UL; ++q) RowGroup<T, G>::load(col, g + 64 * q, y[q]);
#pragma unroll
        for (int q = 0; q < UL; ++q) RowGroup<T, G>::load(u, g + 64 * q, uu[q]);
#pragma unroll
        for (int q = 0; q < UL; ++q)
#pragma unroll
            for (int e = 0; e < G; ++e) {
                ar += uu[q][e].re * y[q][e].re + uu[q][e].im * y[q][e].im;
                ai += uu[q][e].re * y[q][e].im - uu[q][e].im * y[q][e].re;
            }
    }
    for (; g < ng; g += 64) {
        cx<double> y0[G], u0[G];
        RowGroup<T, G>::load(col, g, y0);
        RowGroup<T, G>::load(u, g, u0);
#pragma unroll
        for (int e = 0; e < G; ++e) {
            ar += u0[e].re * y0[e].re + u0[e].im * y0[e].im;
            ai += u0[e].re * y0[e].im - u0[e].im * y0[e].re;
        }
    }
    return cx<double>{wave_sum(ar), wave_sum(ai)};
}

// acc[e] += sum_{c in [c0, c1)} A[G g + e, c] v[c]
template <typename T, int G>
__device__ __forceinline__ void rows_times_v(const cx<T> *Y, size_t ldy, const cx<T> *v, int g,
                                             int c0, int c1, cx<double> (&acc)[G]) {
    constexpr int UL = unroll_of<T>();
    int c = c0;
    for (; c + UL - 1 < c1; c += UL) {
        cx<double> a[UL][G];
#pragma unroll
        for (int q = 0; q < UL; ++q) RowGroup<T, G>::load(Y + (size_t)(c + q) * ldy, g, a[q]);
#pragma unroll
        for (int q = 0; q < UL; ++q) {
            const cx<double> vq = cvt<double>(v[c + q]);
#pragma unroll
            for (int e = 0; e < G; ++e) acc[e] = acc[e] + a[q][e] * vq;
        }
    }
    for (; c < c1; ++c) {
        cx<double> a0[G];
        RowGroup<T, G>::load(Y + (size_t)c * ldy, g, a0);
        const cx<double> v0 = cvt<double>(v[c]);
#pragma unroll
        for (int e = 0; e < G; ++e) acc[e] = acc[e] + a0[e] * v0;
    }
}

// acc[q] -= U[r, k] py[k] + Xw[r, k] pv[k] (k < i) + u[r] py[i], r = G g + q
template <typename T, int G>
__device__ __forceinline__ void rows_corrections(const cx<T> *Y, size_t ldy, const cx<T> *Xw,
                                                 int R, const cx<T> *u, int j0, int i, int g,
                                                 const cx<double> *py, const cx<double> *pv,
                                                 cx<double> (&acc)[G]) {
    for (int k = 0; k < i; ++k) {
        cx<double> uk[G], xk[G];
        RowGroup<T, G>::load(Y + (size_t)(j0 + k) * ldy, g, uk);
        RowGroup<T, G>::load(Xw + (size_t)k * R, g, xk);
#pragma unroll
        for (int q = 0; q < G; ++q) acc[q] = acc[q] - uk[q] * py[k] - xk[q] * pv[k];
    }
    cx<double> ur[G];
    RowGroup<T, G>::load(u, g, ur);
#pragma unroll
    for (int q = 0; q < G; ++q) acc[q] = acc[q] - ur[q] * py[i];
}

// x = sigma (A v - U py - Xw pv - u py_i) for rows r > j of the step
template <typename T, int G>
__device__ __forceinline__ void panel_gemv2(const cx<T> *Y, int ldy, cx<T> *Xw, int R, int C,
                                            const cx<T> *u, const cx<T> *v, int j0, int i, int j,
                                            const cx<double> *py, const cx<double> *pv,
                                            cx<double> sig, cx<double> *part, int lane, int wv) {
    const int ng = R / G, g0 = (j + 1) / G;
    const int nrb = (ng - g0 + 63) / 64;
    const int nt = C - 1 - j;
    if (nrb >= NW) {
        for (int rb = wv; rb < nrb; rb += NW) {
            const int g = g0 + rb * 64 + lane;
            if (g >= ng) continue;
            cx<double> acc[G];
#pragma unroll
            for (int q = 0; q < G; ++q) acc[q] = cx<double>{0, 0};
            rows_times_v<T, G>(Y, ldy, v, g, j + 1, C, acc);
            rows_corrections<T, G>(Y, ldy, Xw, R, u, j0, i, g, py, pv, acc);
#pragma unroll
            for (int q = 0; q < G; ++q)
                if (G * g + q > j) Xw[G * g + q + (size_t)i * R] = cvt<T>(sig * acc[q]);
        }
    } else {
        // fewer row blocks than waves: split the trailing columns into ns
        // slices and reduce the partial sums through LDS
        const int ns = NW / nrb;
        const int rb = wv % nrb, sl = wv / nrb;
        const int g = g0 + rb * 64 + lane;
        cx<double> acc[G];
#pragma unroll
        for (int q = 0; q < G; ++q) acc[q] = cx<double>{0, 0};
        if (sl < ns && g < ng) {
            const int chunk = (nt + ns - 1) / ns;
            const int c0 = j + 1 + sl * chunk, c1 = min(C, c0 + chunk);
            rows_times_v<T, G>(Y, ldy, v, g, c0, c1, acc);
            if (sl == 0) rows_corrections<T, G>(Y, ldy, Xw, R, u, j0, i, g, py, pv, acc);
        }
#pragma unroll
        for (int q = 0; q < G; ++q) part[(wv * 64 + lane) * 2 + q] = acc[q];
        __syncthreads();
        if (wv < nrb && g < ng) {
#pragma unroll
            for (int q = 0; q < G; ++q) {
                cx<double> sum{0, 0};
                for (int z = 0; z < ns; ++z) sum = sum + part[((z * nrb + wv) * 64 + lane) * 2 + q];
                if (G * g + q > j) Xw[G * g + q + (size_t)i * R] = cvt<T>(sig * sum);
            }
        }
    }
}

template <typename T>
__global__ __launch_bounds__(BT) void k_labrd(int R, int C, int j0, int nbp,
                                             cx<T> *__restrict__ Yall, int ldy, size_t sY,
                                             double *__restrict__ dall, double *__restrict__ eall,
                                             cx<T> *__restrict__ tauLall,
                                             cx<T> *__restrict__ tauRall,
                                             cx<T> *__restrict__ Vrall, cx<T> *__restrict__ Xwall,
                                             cx<T> *__restrict__ Ywall,
                                             cx<T> *__restrict__ Vwall) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<double> *part = reinterpret_cast<cx<double> *>(smem_raw);  // NW * 64 * 2
    cx<double> *pu = part + NW * 64 * 2;                            // NB: u^H U_k
    cx<double> *pq = pu + NB;                                       // NB: u^H Xw_k
    cx<double> *py = pq + NB;                                       // NB+1: Yw_k^H v
    cx<double> *pv = py + NB + 1;                                   // NB: Vw_k^H v
    cx<double> *cy = pv + NB;                                       // NB: conj(Yw[j,k])
    cx<double> *cv = cy + NB;                                       // NB: conj(Vw[j,k])
    cx<double> *uj = cv + NB;                                       // NB: U[j,k]
    cx<double> *xj = uj + NB;                                       // NB: Xw[j,k]
    double *red = reinterpret_cast<double *>(xj + NB);              // NW
    cx<T> *u = reinterpret_cast<cx<T> *>(red + NW);                 // R
    cx<T> *v = u + R;                                               // C

    const int b = blockIdx.x;
    cx<T> *Y = Yall + (size_t)b * sY;
    double *d = dall + (size_t)b * C;
    double *e = eall + (size_t)b * C;
    cx<T> *tauL = tauLall + (size_t)b * C;
    cx<T> *tauR = tauRall + (size_t)b * C;
    cx<T> *Vr = Vrall + (size_t)b * C * C;
    cx<T> *Xw = Xwall + (size_t)b * R * NB;
    cx<T> *Yw = Ywall + (size_t)b * C * NB;
    cx<T> *Vw = Vwall + (size_t)b * C * NB;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const bool pairs = ((R | ldy) & 1) == 0;  // complex64 rows read two at a time

    for (int i = 0; i < nbp; ++i) {
        const int j = j0 + i;
        // ---- row-j coefficients of the deferred updates
        if (t < i) {
            cy[t] = conj(cvt<double>(Yw[j + (size_t)t * C]));
            cv[t] = conj(cvt<double>(Vw[j + (size_t)t * C]));
            uj[t] = cvt<double>(Y[j + (size_t)(j0 + t) * ldy]);
            xj[t] = cvt<double>(Xw[j + (size_t)t * R]);
        }
        __syncthreads();
        // ---- column j of A_i (rows >= j) and its left reflector
        double p = 0;
        for (int r = j + t; r < R; r += BT) {
            cx<double> a = cvt<double>(Y[r + (size_t)j * ldy]);
            for (int k = 0; k < i; ++k)
                a = a - cvt<double>(Y[r + (size_t)(j0 + k) * ldy]) * cy[k] -
                    cvt<double>(Xw[r + (size_t)k * R]) * cv[k];
            u[r] = cvt<T>(a);
            if (r > j) p += abs2(a);
        }
        const double xn2 = block_sum(p, red);  // syncs: u visible
        const Refl rf = larfg(cvt<double>(u[j]), xn2);
        const cx<T> sc = cvt<T>(rf.scale);
        __syncthreads();  // u[j] read by all before it is overwritten
        for (int r = j + 1 + t; r < R; r += BT) {
            const cx<T> uv = u[r] * sc;
            u[r] = uv;
            Y[r + (size_t)j * ldy] = uv;
        }
        if (t == 0) {
            u[j] = cx<T>{1, 0};
            if (j & 1) u[j - 1] = cx<T>{0, 0};  // row pairs start at even rows
            Y[j + (size_t)j * ldy] = cx<T>{(T)rf.beta, 0};
            d[j] = rf.beta;
            tauL[j] = cvt<T>(rf.tau);
        }
        __syncthreads();
        if (j == C - 1) {
            if (t == 0) e[j] = 0.0;
            break;
        }
        // ---- w = u^H [A trailing | U | Xw] (one wave per column)
        const int nt = C - 1 - j;
        const int nvc = nt + 2 * i;
        for (int vc = wv; vc < nvc; vc += NW) {
            const cx<T> *col;
            if (vc < nt)
                col = Y + (size_t)(j + 1 + vc) * ldy;
            else if (vc < nt + i)
                col = Y + (size_t)(j0 + vc - nt) * ldy;
            else
                col = Xw + (size_t)(vc - nt - i) * R;
            cx<double> wd;
            if constexpr (sizeof(T) == 4) {
                if (pairs)
                    wd = wave_dot_u<T, 2>(col, u, j, R, lane);
                else
                    wd = wave_dot_u<T, 1>(col, u, j, R, lane);
            } else {
                wd = wave_dot_u<T, 1>(col, u, j, R, lane);
            }
            const double ar = wd.re, ai = wd.im;
            if (lane == 0) {
                const cx<double> w{ar, ai};
                if (vc < nt)
                    v[j + 1 + vc] = cvt<T>(w);  // raw u^H A, corrected below
                else if (vc < nt + i)
                    pu[vc - nt] = w;
                else
                    pq[vc - nt - i] = w;
            }
        }
        __syncthreads();
        // ---- y = tau conj(u^H A_i) and row j of (I - conj(tau) u u^H) A_i
        const cx<double> tau = rf.tau;
        double pr = 0;
        for (int c = j + 1 + t; c < C; c += BT) {
            cx<double> w = cvt<double>(v[c]);
            cx<double> rowc = cvt<double>(Y[j + (size_t)c * ldy]);
            for (int k = 0; k < i; ++k) {
                const cx<double> cyk = conj(cvt<double>(Yw[c + (size_t)k * C]));
                const cx<double> cvk = conj(cvt<double>(Vw[c + (size_t)k * C]));
                w = w - pu[k] * cyk - pq[k] * cvk;
                rowc = rowc - uj[k] * cyk - xj[k] * cvk;
            }
            const cx<double> y = tau * conj(w);
            Yw[c + (size_t)i * C] = cvt<T>(y);
            rowc = rowc - conj(y);
            v[c] = cvt<T>(rowc);
            if (c >= j + 2) pr += abs2(rowc);
        }
        const double xr2 = block_sum(pr, red);  // syncs: v, Yw[:, i] visible
        const Refl rr = larfg(conj(cvt<double>(v[j + 1])), xr2);
        const cx<T> scr = cvt<T>(rr.scale);
        __syncthreads();  // v[j+1] read by all before it is overwritten
        for (int c = j + 1 + t; c < C; c += BT) {
            const cx<T> vv = (c == j + 1) ? cx<T>{1, 0} : conj(v[c]) * scr;
            v[c] = vv;
            Vr[(size_t)j * C + c] = vv;
            Vw[c + (size_t)i * C] = vv;
            Y[j + (size_t)c * ldy] = (c == j + 1) ? cx<T>{(T)rr.beta, 0} : vv;
        }
        if (t == 0) {
            e[j] = rr.beta;
            tauR[j] = cvt<T>(rr.tau);
        }
        __syncthreads();
        // ---- py[k] = Yw_k^H v (k <= i), pv[k] = Vw_k^H v (k < i), over c > j
        for (int q = wv; q < 2 * i + 1; q += NW) {
            const cx<T> *col = (q <= i) ? Yw + (size_t)q * C : Vw + (size_t)(q - i - 1) * C;
            double ar = 0, ai = 0;
            for (int c = j + 1 + lane; c < C; c += 64) {
                const cx<T> a = col[c], vv = v[c];
                ar += (double)a.re * vv.re + (double)a.im * vv.im;
                ai += (double)a.re * vv.im - (double)a.im * vv.re;
            }
            ar = wave_sum(ar);
            ai = wave_sum(ai);
            if (lane == 0) {
                if (q <= i)
                    py[q] = cx<double>{ar, ai};
                else
                    pv[q - i - 1] = cx<double>{ar, ai};
            }
        }
        __syncthreads();
        // ---- x = sigma (A v - U py - Xw pv - u py_i), rows r > j
        const cx<double> sig = rr.tau;
        if constexpr (sizeof(T) == 4) {
            if (pairs)
                panel_gemv2<T, 2>(Y, ldy, Xw, R, C, u, v, j0, i, j, py, pv, sig, part, lane, wv);
            else
                panel_gemv2<T, 1>(Y, ldy, Xw, R, C, u, v, j0, i, j, py, pv, sig, part, lane, wv);
        } else {
            panel_gemv2<T, 1>(Y, ldy, Xw, R, C, u, v, j0, i, j, py, pv, sig, part, lane, wv);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------
// Split form for small batches: the same step as k_labrd in four launches,
// so that the two matrix-vector products of a step spread over many
// workgroups per frequency instead of one.  Per frequency state lives in
// global memory: uvec (R), vvec (C), wbuf (C + 2 NB raw products) and
// sbuf (tau, sigma, py[NB + 1], pv[NB]).
constexpr int SB = 2 + 2 * NB + 1;  // complex scalars of sbuf per frequency

// left reflector of column j (one workgroup per frequency)
template <typename T>
__global__ __launch_bounds__(BT) void k_lb_left(int R, int C, int j0, int i, cx<T> *__restrict__ Yall,
                                               int ldy, size_t sY, double *__restrict__ dall,
                                               double *__restrict__ eall,
                                               cx<T> *__restrict__ tauLall,
                                               const cx<T> *__restrict__ Xwall,
                                               const cx<T> *__restrict__ Ywall,
                                               const cx<T> *__restrict__ Vwall,
                                               cx<T> *__restrict__ uall, cx<double> *__restrict__ sall) {
    __shared__ cx<double> cy[NB], cv[NB];
    __shared__ double red[NW];
    const int b = blockIdx.x, t = threadIdx.x;
    const int j = j0 + i;
    cx<T> *Y = Yall + (size_t)b * sY;
    const cx<T> *Xw = Xwall + (size_t)b * R * NB;
    const cx<T> *Yw = Ywall + (size_t)b * C * NB;
    const cx<T> *Vw = Vwall + (size_t)b * C * NB;
    cx<T> *u = uall + (size_t)b * R;
    if (t < i) {
        cy[t] = conj(cvt<double>(Yw[j + (size_t)t * C]));
        cv[t] = conj(cvt<double>(Vw[j + (size_t)t * C]));
    }
    __syncthreads();
    double p = 0;
    for (int r = j + t; r < R; r += BT) {
        cx<double> a = cvt<double>(Y[r + (size_t)j * ldy]);
        for (int k = 0; k < i; ++k)
            a = a - cvt<double>(Y[r + (size_t)(j0 + k) * ldy]) * cy[k] -
                cvt<double>(Xw[r + (size_t)k * R]) * cv[k];
        u[r] = cvt<T>(a);
        if (r > j) p += abs2(a);
    }
    const double xn2 = block_sum(p, red);
    const Refl rf = larfg(cvt<double>(u[j]), xn2);
    const cx<T> sc = cvt<T>(rf.scale);
    __syncthreads();
    for (int r = j + 1 + t; r < R; r += BT) {
        const cx<T> uv = u[r] * sc;
        u[r] = uv;
        Y[r + (size_t)j * ldy] = uv;
    }
    if (t == 0) {
        u[j] = cx<T>{1, 0};
        Y[j + (size_t)j * ldy] = cx<T>{(T)rf.beta, 0};
        dall[(size_t)b * C + j] = rf.beta;
        tauLall[(size_t)b * C + j] = cvt<T>(rf.tau);
        sall[(size_t)b * SB] = rf.tau;
        if (j == C - 1) eall[(size_t)b * C + j] = 0.0;
    }
}

// w = u^H [A trailing | U | Xw], one wave per column; grid (ceil(nvc/4), batch)
template <typename T>
__global__ __launch_bounds__(256) void k_lb_gemv1(int R, int C, int j0, int i,
                                                 const cx<T> *__restrict__ Yall, int ldy, size_t sY,
                                                 const cx<T> *__restrict__ Xwall,
                                                 const cx<T> *__restrict__ uall,
                                                 cx<double> *__restrict__ wall) {
    const int b = blockIdx.y, lane = threadIdx.x & 63;
    const int vc = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int j = j0 + i, nt = C - 1 - j, nvc = nt + 2 * i;
    if (vc >= nvc) return;
    const cx<T> *Y = Yall + (size_t)b * sY;
    const cx<T> *Xw = Xwall + (size_t)b * R * NB;
    const cx<T> *u = uall + (size_t)b * R;
    const cx<T> *col;
    if (vc < nt)
        col = Y + (size_t)(j + 1 + vc) * ldy;
    else if (vc < nt + i)
        col = Y + (size_t)(j0 + vc - nt) * ldy;
    else
        col = Xw + (size_t)(vc - nt - i) * R;
    double ar0 = 0, ai0 = 0, ar1 = 0, ai1 = 0;
    int r = j + lane;
    for (; r + 64 < R; r += 128) {
        const cx<T> y0 = col[r], y1 = col[r + 64];
        const cx<T> u0 = u[r], u1 = u[r + 64];
        ar0 += (double)u0.re * y0.re + (double)u0.im * y0.im;
        ai0 += (double)u0.re * y0.im - (double)u0.im * y0.re;
        ar1 += (double)u1.re * y1.re + (double)u1.im * y1.im;
        ai1 += (double)u1.re * y1.im - (double)u1.im * y1.re;
    }
    if (r < R) {
        const cx<T> y0 = col[r], u0 = u[r];
        ar0 += (double)u0.re * y0.re + (double)u0.im * y0.im;
        ai0 += (double)u0.re * y0.im - (double)u0.im * y0.re;
    }
    const double ar = wave_sum(ar0 + ar1), ai = wave_sum(ai0 + ai1);
    if (lane == 0) wall[(size_t)b * (C + 2 * NB) + vc] = cx<double>{ar, ai};
}

// y, row j, right reflector, and the small products py, pv (one workgroup
// per frequency)
template <typename T>
__global__ __launch_bounds__(BT) void k_lb_right(int R, int C, int j0, int i, cx<T> *__restrict__ Yall,
                                                int ldy, size_t sY, double *__restrict__ eall,
                                                cx<T> *__restrict__ tauRall, cx<T> *__restrict__ Vrall,
                                                const cx<T> *__restrict__ Xwall,
                                                cx<T> *__restrict__ Ywall, cx<T> *__restrict__ Vwall,
                                                cx<T> *__restrict__ vall,
                                                const cx<double> *__restrict__ wall,
                                                cx<double> *__restrict__ sall) {
    __shared__ cx<double> pu[NB], pq[NB], uj[NB], xj[NB];
    __shared__ double red[NW];
    const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int j = j0 + i, nt = C - 1 - j;
    cx<T> *Y = Yall + (size_t)b * sY;
    const cx<T> *Xw = Xwall + (size_t)b * R * NB;
    cx<T> *Yw = Ywall + (size_t)b * C * NB;
    cx<T> *Vw = Vwall + (size_t)b * C * NB;
    cx<T> *Vr = Vrall + (size_t)b * C * C;
    cx<T> *v = vall + (size_t)b * C;
    const cx<double> *w = wall + (size_t)b * (C + 2 * NB);
    cx<double> *S = sall + (size_t)b * SB;
    if (t < i) {
        pu[t] = w[nt + t];
        pq[t] = w[nt + i + t];
        uj[t] = cvt<double>(Y[j + (size_t)(j0 + t) * ldy]);
        xj[t] = cvt<double>(Xw[j + (size_t)t * R]);
    }
    __syncthreads();
    const cx<double> tau = S[0];
    double pr = 0;
    for (int c = j + 1 + t; c < C; c += BT) {
        cx<double> wc = w[c - j - 1];
        cx<double> rowc = cvt<double>(Y[j + (size_t)c * ldy]);
        for (int k = 0; k < i; ++k) {
            const cx<double> cyk = conj(cvt<double>(Yw[c + (size_t)k * C]));
            const cx<double> cvk = conj(cvt<double>(Vw[c + (size_t)k * C]));
            wc = wc - pu[k] * cyk - pq[k] * cvk;
            rowc = rowc - uj[k] * cyk - xj[k] * cvk;
        }
        const cx<double> y = tau * conj(wc);
        Yw[c + (size_t)i * C] = cvt<T>(y);
        rowc = rowc - conj(y);
        v[c] = cvt<T>(rowc);
        if (c >= j + 2) pr += abs2(rowc);
    }
    const double xr2 = block_sum(pr, red);  // v and Yw[:, i] visible in the workgroup
    const Refl rr = larfg(conj(cvt<double>(v[j + 1])), xr2);
    const cx<T> scr = cvt<T>(rr.scale);
    __syncthreads();
    for (int c = j + 1 + t; c < C; c += BT) {
        const cx<T> vv = (c == j + 1) ? cx<T>{1, 0} : conj(v[c]) * scr;
        v[c] = vv;
        Vr[(size_t)j * C + c] = vv;
        Vw[c + (size_t)i * C] = vv;
        Y[j + (size_t)c * ldy] = (c == j + 1) ? cx<T>{(T)rr.beta, 0} : vv;
    }
    if (t == 0) {
        eall[(size_t)b * C + j] = rr.beta;
        tauRall[(size_t)b * C + j] = cvt<T>(rr.tau);
        S[1] = rr.tau;
    }
    __syncthreads();
    for (int q = wv; q < 2 * i + 1; q += NW) {
        const cx<T> *col = (q <= i) ? Yw + (size_t)q * C : Vw + (size_t)(q - i - 1) * C;
        double ar = 0, ai = 0;
        for (int c = j + 1 + lane; c < C; c += 64) {
            const cx<T> a = col[c], vv = v[c];
            ar += (double)a.re * vv.re + (double)a.im * vv.im;
            ai += (double)a.re * vv.im - (double)a.im * vv.re;
        }
        ar = wave_sum(ar);
        ai = wave_sum(ai);
        if (lane == 0) S[2 + q] = cx<double>{ar, ai};  // py[0..i] then pv[0..i-1]
    }
}

// x = sigma (A v - U py - Xw pv - u py_i) on 64 rows per workgroup, the
// trailing columns split over the four waves; grid (ceil((R-j-1)/64), batch)
template <typename T>
__global__ __launch_bounds__(256) void k_lb_gemv2(int R, int C, int j0, int i,
                                                 const cx<T> *__restrict__ Yall, int ldy, size_t sY,
                                                 cx<T> *__restrict__ Xwall,
                                                 const cx<T> *__restrict__ uall,
                                                 const cx<T> *__restrict__ vall,
                                                 const cx<double> *__restrict__ sall) {
    __shared__ cx<double> part[4][64];
    const int b = blockIdx.y, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int j = j0 + i, nt = C - 1 - j;
    const int r = j + 1 + blockIdx.x * 64 + lane;
    const cx<T> *Y = Yall + (size_t)b * sY;
    cx<T> *Xw = Xwall + (size_t)b * R * NB;
    const cx<T> *v = vall + (size_t)b * C;
    const cx<double> *S = sall + (size_t)b * SB;
    cx<double> acc{0, 0}, acc1{0, 0};
    if (r < R) {
        const int chunk = (nt + 3) / 4;
        const int c0 = j + 1 + wv * chunk, c1 = min(C, c0 + chunk);
        int c = c0;
        for (; c + 1 < c1; c += 2) {
            const cx<T> a0 = Y[r + (size_t)c * ldy], a1 = Y[r + (size_t)(c + 1) * ldy];
            acc = acc + cvt<double>(a0) * cvt<double>(v[c]);
            acc1 = acc1 + cvt<double>(a1) * cvt<double>(v[c + 1]);
        }
        if (c < c1) acc = acc + cvt<double>(Y[r + (size_t)c * ldy]) * cvt<double>(v[c]);
    }
    part[wv][lane] = acc + acc1;
    __syncthreads();
    if (wv == 0 && r < R) {
        cx<double> x = part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane];
        for (int k = 0; k < i; ++k)
            x = x - cvt<double>(Y[r + (size_t)(j0 + k) * ldy]) * S[2 + k] -
                cvt<double>(Xw[r + (size_t)k * R]) * S[3 + i + k];
        x = x - cvt<double>(uall[(size_t)b * R + r]) * S[2 + i];
        Xw[r + (size_t)i * R] = cvt<T>(S[1] * x);
    }
}

}  // namespace

size_t gebrd_lds_bytes(int R, int C, size_t elem) {
    return ((size_t)NW * 64 * 2 + 8 * NB + 1) * sizeof(cx<double>) + NW * sizeof(double) +
           ((size_t)R + C) * elem;
}

size_t gebrd_work_elems(int R, int C) {
    // Xw, Yw, Vw panels + split-form state (u, v, raw products, scalars)
    return ((size_t)R + 2 * (size_t)C) * NB + (size_t)R + C + 2 * ((size_t)C + 2 * NB + SB) + 8;
}

namespace {
// batches below this run the split form (CU count / 2 on MI355X; override
// with IA_GEBRD_SPLIT_BELOW)
int split_below() {
    const char *e = getenv("IA_GEBRD_SPLIT_BELOW");
    return e ? atoi(e) : 128;
}
}  // namespace

template <typename T>
ia_status launch_gebrd(int R, int C, cx<T> *Y, int ldy, size_t sY, double *d, double *e,
                       cx<T> *tauL, cx<T> *tauR, cx<T> *Vr, cx<T> *work, int batch,
                       hipStream_t s, KernelClock *clk) {
    if (R < C) return fail(IA_ERR_ARG, "bidiagonalisation needs R >= C (%d x %d)", R, C);
    const size_t lds = gebrd_lds_bytes(R, C, sizeof(cx<T>));
    if (lds > 160 * 1024)
        return fail(IA_ERR_ARG, "bidiagonalisation of %d x %d exceeds the on-chip limit", R, C);
    IA_HIP(hipFuncSetAttribute((const void *)k_labrd<T>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    cx<T> *Xw = work;
    cx<T> *Yw = Xw + (size_t)batch * R * NB;
    cx<T> *Vw = Yw + (size_t)batch * C * NB;
    cx<T> *uv = Vw + (size_t)batch * C * NB;                 // split form: u (R), v (C)
    cx<T> *vv = uv + (size_t)batch * R;
    cx<double> *wb = reinterpret_cast<cx<double> *>(
        (reinterpret_cast<uintptr_t>(vv + (size_t)batch * C) + 15) & ~(uintptr_t)15);
    cx<double> *sb = wb + (size_t)batch * (C + 2 * NB);
    const bool split = batch < split_below();
    const cx<T> mone{-1, 0}, one{1, 0};
    for (int j0 = 0; j0 < C; j0 += NB) {
        const int nbp = std::min(NB, C - j0);
        if (clk) clk->mark(s);
        if (!split) {
            hipLaunchKernelGGL(k_labrd<T>, dim3(batch), dim3(BT), lds, s, R, C, j0, nbp, Y, ldy,
                               sY, d, e, tauL, tauR, Vr, Xw, Yw, Vw);
        } else {
            for (int i = 0; i < nbp; ++i) {
                const int j = j0 + i;
                hipLaunchKernelGGL(k_lb_left<T>, dim3(batch), dim3(BT), 0, s, R, C, j0, i, Y, ldy,
                                   sY, d, e, tauL, Xw, Yw, Vw, uv, sb);
                if (j == C - 1) break;
                const int nvc = C - 1 - j + 2 * i;
                hipLaunchKernelGGL(k_lb_gemv1<T>, dim3((nvc + 3) / 4, batch), dim3(256), 0, s, R, C,
                                   j0, i, Y, ldy, sY, Xw, uv, wb);
                hipLaunchKernelGGL(k_lb_right<T>, dim3(batch), dim3(BT), 0, s, R, C, j0, i, Y, ldy,
                                   sY, e, tauR, Vr, Xw, Yw, Vw, vv, wb, sb);
                hipLaunchKernelGGL(k_lb_gemv2<T>, dim3((R - j - 1 + 63) / 64, batch), dim3(256), 0,
                                   s, R, C, j0, i, Y, ldy, sY, Xw, uv, vv, sb);
            }
        }
        if (clk) {
            clk->mark(s);
            // algorithmic bytes: the two products of step j read the trailing
            // block (rows >= j, resp. > j; columns > j) once each
            double by = 0;
            for (int j = j0; j < j0 + nbp && j < C - 1; ++j)
                by += (double)(2 * (R - j) - 1) * (double)(C - 1 - j);
            clk->bytes += by * sizeof(cx<T>) * batch;
            clk->launches++;
        }
        const int s0 = j0 + nbp;
        if (s0 < C) {
            const int M = R - s0, N = C - s0;
            cx<T> *A = Y + s0 + (size_t)s0 * ldy;
            // A -= U Yw^H + Xw Vw^H on rows/columns s0..
            launch_gemm<T>('N', 'C', M, N, nbp, mone, Y + s0 + (size_t)j0 * ldy, ldy, sY, Yw + s0,
                           C, (size_t)C * NB, one, A, ldy, sY, batch, s);
            launch_gemm<T>('N', 'C', M, N, nbp, mone, Xw + s0, R, (size_t)R * NB, Vw + s0, C,
                           (size_t)C * NB, one, A, ldy, sY, batch, s);
        }
    }
    IA_HIP(hipGetLastError());
    return IA_OK;
}

template ia_status launch_gebrd<double>(int, int, cx<double> *, int, size_t, double *, double *,
                                        cx<double> *, cx<double> *, cx<double> *, cx<double> *,
                                        int, hipStream_t, KernelClock *);
template ia_status launch_gebrd<float>(int, int, cx<float> *, int, size_t, double *, double *,
                                       cx<float> *, cx<float> *, cx<float> *, cx<float> *, int,
                                       hipStream_t, KernelClock *);

}  // namespace ia
