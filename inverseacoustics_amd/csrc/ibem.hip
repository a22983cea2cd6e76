// IBEM (collocation, linear triangles) assembly on gfx950.
//
// Restates MATLAB/ibemS.m:18-156 and CUDA/ibem.cuh:424-770 (the CPU path;
// the reference GPU path ibem.cuh:1-261 launches ~2e6 dynamic-parallelism
// child grids per frequency, frees their scratch before they finish and
// accumulates with unordered double atomics).  For element e with nodes
// (P0, P1, P2), n = (P0 - P2) x (P1 - P2), x_q = xi1 P0 + xi2 P1 + (1-xi1-xi2) P2,
// N = (xi1, xi2, 1 - xi1 - xi2), a row point p receives
//   B_a = (1/4pi) sum_q w N_a |n| e^{ikR}/R,  A_a = (1/4pi) sum_q w N_a F (ikR-1) e^{ikR}/R,
//   F = n.(x_q - p)/R^2,  C = (1/4pi) sum_q w (-F/R)  (collocation rows),
// with the singular page of the collocated vertex when p is a node of e.
//
// Gather layout, no atomics: one thread owns matrix entry (row o, column c)
// and walks the elements adjacent to node c (CSR adjacency); a wave holds
// 64 consecutive rows of one column, so the element loop is uniform and
// the column-major stores are coalesced.  The singular elements of a surface
// row (those with the row's own node, integrated on the page rule of that
// vertex) are summed by a second kernel with the page points across a
// wave's lanes.  The diagonal term C(o) is a separate block reduction over
// all elements.
#include "ia_internal.hpp"

namespace ia {

namespace {

constexpr double INV4PI = 0.079577471545947667884;

struct Elem {
    double p[3][3];
    double n[3];
    double an;
};

__device__ __forceinline__ void load_elem(const double *X, int m, const int *E, int ne, int e,
                                          Elem &el, int nd[3]) {
    for (int a = 0; a < 3; ++a) {
        nd[a] = E[e + a * ne];
        for (int d = 0; d < 3; ++d) el.p[a][d] = X[nd[a] + d * m];
    }
    double u[3], v[3];
    for (int d = 0; d < 3; ++d) {
        u[d] = el.p[0][d] - el.p[2][d];
        v[d] = el.p[1][d] - el.p[2][d];
    }
    el.n[0] = u[1] * v[2] - u[2] * v[1];
    el.n[1] = u[2] * v[0] - u[0] * v[2];
    el.n[2] = u[0] * v[1] - u[1] * v[0];
    el.an = sqrt(el.n[0] * el.n[0] + el.n[1] * el.n[1] + el.n[2] * el.n[2]);
}

// sum over the rule of shape function `a` for row point pt: A, B (+ C)
__device__ __forceinline__ void integrate(const Elem &el, const double pt[3], double k,
                                          const double *rule, int nq, int a, double &Ar,
                                          double &Ai, double &Br, double &Bi, double &Cs) {
    for (int q = 0; q < nq; ++q) {
        const double x1 = rule[q], x2 = rule[q + nq], w = rule[q + 2 * nq];
        const double N[3] = {x1, x2, 1.0 - x1 - x2};
        double d[3];
        for (int t = 0; t < 3; ++t)
            d[t] = N[0] * el.p[0][t] + N[1] * el.p[1][t] + N[2] * el.p[2][t] - pt[t];
        const double R2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        const double R = sqrt(R2);
        const double F = (el.n[0] * d[0] + el.n[1] * d[1] + el.n[2] * d[2]) / R2;
        double sn, cs;
        sincos(k * R, &sn, &cs);
        const double gr = cs / R, gi = sn / R;  // e^{ikR}/R
        const double wn = w * N[a];
        Br += wn * el.an * gr;
        Bi += wn * el.an * gi;
        // F (ikR - 1) G = F ((-1) gr - kR gi) + i F (kR gr - gi)
        const double kR = k * R;
        Ar += wn * F * (-gr - kR * gi);
        Ai += wn * F * (kR * gr - gi);
        Cs += w * (-F / R);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_ibem_rows(int rows, const double *__restrict__ pts,
                                                  int npts_ld, int surface, int m,
                                                  const double *__restrict__ X,
                                                  const int *__restrict__ E, int ne,
                                                  const int *__restrict__ adj_ptr,
                                                  const int *__restrict__ adj,
                                                  const double *__restrict__ rule, int nq,
                                                  const double *__restrict__ kv, double rho_c,
                                                  cx<T> *__restrict__ Aout, cx<T> *__restrict__ Bout,
                                                  int ldo, size_t stride) {
    const int o = blockIdx.x * 64 + (threadIdx.x & 63);
    const int c = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int f = blockIdx.z;
    if (o >= rows || c >= m) return;
    const double k = kv[f];
    const double pt[3] = {pts[o], pts[o + npts_ld], pts[o + 2 * npts_ld]};
    double Ar = 0, Ai = 0, Br = 0, Bi = 0, Cs = 0;
    for (int q = adj_ptr[c]; q < adj_ptr[c + 1]; ++q) {
        const int e = adj[q];
        Elem el;
        int nd[3];
        load_elem(X, m, E, ne, e, el, nd);
        const int a = (nd[0] == c) ? 0 : ((nd[1] == c) ? 1 : 2);
        // an element with the row's own node is singular: its page integral
        // is added by k_ibem_sing (one wave per entry, the page points over
        // the lanes) instead of here, where the lanes of a wave would wait
        // for the few that carry it
        if (surface && (nd[0] == o || nd[1] == o || nd[2] == o)) continue;
        integrate(el, pt, k, rule, nq, a, Ar, Ai, Br, Bi, Cs);
    }
    // B scaled by i rho c k: (Br + i Bi) i s = -s Bi + i s Br
    const double s = rho_c * k;
    const size_t off = (size_t)f * stride + (size_t)c * ldo + o;
    Aout[off] = cx<T>{(T)(INV4PI * Ar), (T)(INV4PI * Ai)};
    Bout[off] = cx<T>{(T)(-s * INV4PI * Bi), (T)(s * INV4PI * Br)};
}

// Singular entries, one wave per (entry, frequency): entry s has row
// o = sing_oc[2 s], column c = sing_oc[2 s + 1] and the contributions
// sing_list[3 j .. 3 j + 2] = (element e, o's vertex in e, c's vertex in e),
// j in [sing_ptr[s], sing_ptr[s + 1]); the lanes take the page points in
// turn, the four sums are reduced across the wave in a fixed order and lane
// 0 adds them to A[o, c] and B[o, c] (each entry has one wave: no atomics).
template <typename T>
__global__ __launch_bounds__(256) void k_ibem_sing(int nse, const int *__restrict__ sing_ptr,
                                                  const int *__restrict__ sing_oc,
                                                  const int *__restrict__ sing_list, int m,
                                                  const double *__restrict__ X,
                                                  const int *__restrict__ E, int ne,
                                                  const double *__restrict__ pages, int nqs,
                                                  const double *__restrict__ kv, double rho_c,
                                                  cx<T> *__restrict__ A, cx<T> *__restrict__ B,
                                                  int ldo, size_t stride) {
    const int sidx = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int f = blockIdx.y;
    if (sidx >= nse) return;
    const int o = sing_oc[2 * sidx], c = sing_oc[2 * sidx + 1];
    const double k = kv[f];
    const double pt[3] = {X[o], X[o + m], X[o + 2 * m]};
    double Ar = 0, Ai = 0, Br = 0, Bi = 0, Cs = 0;
    for (int j = sing_ptr[sidx]; j < sing_ptr[sidx + 1]; ++j) {
        const int e = sing_list[3 * j], ikn = sing_list[3 * j + 1], a = sing_list[3 * j + 2];
        Elem el;
        int nd[3];
        load_elem(X, m, E, ne, e, el, nd);
        const double *pg = pages + (size_t)ikn * 3 * nqs;
        for (int q0 = 0; q0 < nqs; q0 += 64) {
            const int q = q0 + lane;
            if (q >= nqs) break;
            const double x1 = pg[q], x2 = pg[q + nqs], w = pg[q + 2 * nqs];
            const double N[3] = {x1, x2, 1.0 - x1 - x2};
            double d[3];
            for (int t = 0; t < 3; ++t)
                d[t] = N[0] * el.p[0][t] + N[1] * el.p[1][t] + N[2] * el.p[2][t] - pt[t];
            const double R2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            const double R = sqrt(R2);
            const double F = (el.n[0] * d[0] + el.n[1] * d[1] + el.n[2] * d[2]) / R2;
            double sn, cs;
            sincos(k * R, &sn, &cs);
            const double gr = cs / R, gi = sn / R;
            const double wn = w * N[a];
            Br += wn * el.an * gr;
            Bi += wn * el.an * gi;
            const double kR = k * R;
            Ar += wn * F * (-gr - kR * gi);
            Ai += wn * F * (kR * gr - gi);
        }
    }
    (void)Cs;
#pragma unroll
    for (int w = 32; w > 0; w >>= 1) {
        Ar += __shfl_xor(Ar, w, 64);
        Ai += __shfl_xor(Ai, w, 64);
        Br += __shfl_xor(Br, w, 64);
        Bi += __shfl_xor(Bi, w, 64);
    }
    if (lane == 0) {
        const double s = rho_c * k;
        const size_t off = (size_t)f * stride + (size_t)c * ldo + o;
        const cx<T> a0 = A[off], b0 = B[off];
        A[off] = cx<T>{(T)((double)a0.re + INV4PI * Ar), (T)((double)a0.im + INV4PI * Ai)};
        B[off] = cx<T>{(T)((double)b0.re - s * INV4PI * Bi), (T)((double)b0.im + s * INV4PI * Br)};
    }
}

// A[o,o] -= 1 + C(o), C(o) = (1/4pi) sum_e sum_q w (-F/R); one block per node
template <typename T>
__global__ __launch_bounds__(256) void k_ibem_cdiag(int m, const double *__restrict__ X,
                                                   const int *__restrict__ E, int ne,
                                                   const double *__restrict__ rule, int nq,
                                                   const double *__restrict__ pages, int nqs,
                                                   cx<T> *__restrict__ A, int lda, size_t stride,
                                                   int nf, double *__restrict__ Cout) {
    __shared__ double red[256];
    const int o = blockIdx.x;
    const double pt[3] = {X[o], X[o + m], X[o + 2 * m]};
    double acc = 0;
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
        Elem el;
        int nd[3];
        load_elem(X, m, E, ne, e, el, nd);
        int ikn = -1;
        for (int t = 0; t < 3; ++t)
            if (nd[t] == o) ikn = t;
        const double *r = (ikn >= 0) ? pages + (size_t)ikn * 3 * nqs : rule;
        const int n = (ikn >= 0) ? nqs : nq;
        for (int q = 0; q < n; ++q) {
            const double x1 = r[q], x2 = r[q + n], w = r[q + 2 * n];
            const double N[3] = {x1, x2, 1.0 - x1 - x2};
            double d[3];
            for (int t = 0; t < 3; ++t)
                d[t] = N[0] * el.p[0][t] + N[1] * el.p[1][t] + N[2] * el.p[2][t] - pt[t];
            const double R2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            const double R = sqrt(R2);
            const double F = (el.n[0] * d[0] + el.n[1] * d[1] + el.n[2] * d[2]) / R2;
            acc += w * (-F / R);
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double C = INV4PI * red[0];
        if (Cout) Cout[o] = C;
        for (int f = 0; f < nf; ++f) {
            cx<T> *p = A + (size_t)f * stride + (size_t)o * lda + o;
            p->re -= (T)(1.0 + C);
        }
    }
}

// Gram matrix G_ab = sum_e sum_q w |n| N_a N_b (invr.cuh:603-625), entry per thread
__global__ void k_gram(int m, const double *__restrict__ X, const int *__restrict__ E, int ne,
                       const int *__restrict__ adj_ptr, const int *__restrict__ adj,
                       const double *__restrict__ rule, int nq, double *__restrict__ G) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (a >= m) return;
    double acc = 0;
    for (int q = adj_ptr[b]; q < adj_ptr[b + 1]; ++q) {
        const int e = adj[q];
        Elem el;
        int nd[3];
        load_elem(X, m, E, ne, e, el, nd);
        int la = -1, lb = -1;
        for (int t = 0; t < 3; ++t) {
            if (nd[t] == a) la = t;
            if (nd[t] == b) lb = t;
        }
        if (la < 0) continue;
        for (int qq = 0; qq < nq; ++qq) {
            const double x1 = rule[qq], x2 = rule[qq + nq], w = rule[qq + 2 * nq];
            const double N[3] = {x1, x2, 1.0 - x1 - x2};
            acc += w * el.an * N[la] * N[lb];
        }
    }
    G[(size_t)b * m + a] = acc;
}

}  // namespace

template <typename T>
void launch_ibem_rows(int rows, const double *pts, int npts_ld, bool surface, int m,
                      const double *X, const int *E, int ne, const int *adj_ptr, const int *adj,
                      const double *rule, int nq, const double *k, int nf, double rho_c, cx<T> *A,
                      cx<T> *B, int ldo, size_t stride, hipStream_t s) {
    dim3 grid((rows + 63) / 64, (m + 3) / 4, nf);
    hipLaunchKernelGGL(k_ibem_rows<T>, grid, dim3(256), 0, s, rows, pts, npts_ld, surface ? 1 : 0,
                       m, X, E, ne, adj_ptr, adj, rule, nq, k, rho_c, A, B, ldo, stride);
}

// The singular entries of the surface rows (row o, column c, o and c nodes of
// one element): sum over the elements e with both nodes of the page integral
// of o's vertex with c's shape function, added to A and B.
template <typename T>
void launch_ibem_sing(int nse, const int *sing_ptr, const int *sing_oc, const int *sing_list,
                      int m, const double *X, const int *E, int ne, const double *pages, int nqs,
                      const double *k, int nf, double rho_c, cx<T> *A, cx<T> *B, int ldo,
                      size_t stride, hipStream_t s) {
    if (nse <= 0) return;
    dim3 grid((nse + 3) / 4, nf);
    hipLaunchKernelGGL(k_ibem_sing<T>, grid, dim3(256), 0, s, nse, sing_ptr, sing_oc, sing_list,
                       m, X, E, ne, pages, nqs, k, rho_c, A, B, ldo, stride);
}

template <typename T>
void launch_ibem_cdiag(int m, const double *X, const int *E, int ne, const double *rule, int nq,
                       const double *pages, int nqs, cx<T> *A, int lda, size_t stride, int nf,
                       double *C, hipStream_t s) {
    hipLaunchKernelGGL(k_ibem_cdiag<T>, dim3(m), dim3(256), 0, s, m, X, E, ne, rule, nq, pages, nqs,
                       A, lda, stride, nf, C);
}

void launch_gram(int m, const double *X, const int *E, int ne, const int *adj_ptr, const int *adj,
                 const double *rule, int nq, double *G, hipStream_t s) {
    dim3 grid((m + 255) / 256, m);
    hipLaunchKernelGGL(k_gram, grid, dim3(256), 0, s, m, X, E, ne, adj_ptr, adj, rule, nq, G);
}

template void launch_ibem_rows<double>(int, const double *, int, bool, int, const double *,
                                       const int *, int, const int *, const int *, const double *,
                                       int, const double *, int, double, cx<double> *,
                                       cx<double> *, int, size_t, hipStream_t);
template void launch_ibem_rows<float>(int, const double *, int, bool, int, const double *,
                                      const int *, int, const int *, const int *, const double *,
                                      int, const double *, int, double, cx<float> *, cx<float> *,
                                      int, size_t, hipStream_t);
template void launch_ibem_sing<double>(int, const int *, const int *, const int *, int,
                                       const double *, const int *, int, const double *, int,
                                       const double *, int, double, cx<double> *, cx<double> *,
                                       int, size_t, hipStream_t);
template void launch_ibem_sing<float>(int, const int *, const int *, const int *, int,
                                      const double *, const int *, int, const double *, int,
                                      const double *, int, double, cx<float> *, cx<float> *, int,
                                      size_t, hipStream_t);
template void launch_ibem_cdiag<double>(int, const double *, const int *, int, const double *, int,
                                        const double *, int, cx<double> *, int, size_t, int,
                                        double *, hipStream_t);
template void launch_ibem_cdiag<float>(int, const double *, const int *, int, const double *, int,
                                       const double *, int, cx<float> *, int, size_t, int, double *,
                                       hipStream_t);

}  // namespace ia
