// IBEM and radiation-mode entry points (CUDA/ibem.cuh:263-770,
// CUDA/invr.cuh:263-909): quadrature tables, batched assembly, the
// Z = A_ss \ B_ss solve (LU when square, Householder QR least squares with
// CHIEF rows), H_b = A_hs Z - B_hs, the IBEM sweep, and radiation modes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ia_internal.hpp"

namespace ia {

template <typename T>
void launch_ibem_rows(int rows, const double *pts, int npts_ld, bool surface, int m,
                      const double *X, const int *E, int ne, const int *adj_ptr, const int *adj,
                      const double *rule, int nq, const double *k, int nf, double rho_c, cx<T> *A,
                      cx<T> *B, int ldo, size_t stride, hipStream_t s);
template <typename T>
void launch_ibem_sing(int nse, const int *sing_ptr, const int *sing_oc, const int *sing_list,
                      int m, const double *X, const int *E, int ne, const double *pages, int nqs,
                      const double *k, int nf, double rho_c, cx<T> *A, cx<T> *B, int ldo,
                      size_t stride, hipStream_t s);
template <typename T>
void launch_ibem_cdiag(int m, const double *X, const int *E, int ne, const double *rule, int nq,
                       const double *pages, int nqs, cx<T> *A, int lda, size_t stride, int nf,
                       double *C, hipStream_t s);
void launch_gram(int m, const double *X, const int *E, int ne, const int *adj_ptr, const int *adj,
                 const double *rule, int nq, double *G, hipStream_t s);
template <typename T>
ia_status regsolve_batched(int rows, int cols, int batch, const cx<double> *A, size_t sA,
                           const cx<double> *b, size_t sb, const RegDev &prm, cx<double> *x,
                           size_t sx, double *lam, int *idx, int *kts, double *smax, hipStream_t s,
                           double *gap = nullptr);
template <typename T>
ia_status regsolve_async(int rows, int cols, int batch, const cx<double> *A, size_t sA,
                         const cx<double> *b, size_t sb, const RegDev &prm, cx<double> *x,
                         size_t sx, double *lam, int *idx, int *kts, double *smax,
                         hipStream_t s);
template <typename T>
void regsolve_addresses(std::vector<uintptr_t> &v);
template <typename T>
ia_status regsolve_reserve(int rows, int cols, int batch);
RegDev to_regdev(const ia_reg_params &p);
ia_status check_reg(const ia_reg_params &p, int pmax);

namespace {

// ------------------------------------------------------------ quadrature
struct Quad {
    int nq = 0, nqs = 0;
    std::vector<double> rule;   // xi1[nq] xi2[nq] w[nq]
    std::vector<double> pages;  // 3 pages of xi1[nqs] xi2[nqs] w[nqs]
};

// 7-point degree-5 rule of Radon (Stroud T2:5-1), weights summing to 1/2.
void gauss7(Quad &q) {
    const double s15 = std::sqrt(15.0);
    const double a1 = (6 - s15) / 21, w1 = (155 - s15) / 1200;
    const double a2 = (6 + s15) / 21, w2 = (155 + s15) / 1200;
    const double pts[7][3] = {{1.0 / 3, 1.0 / 3, 9.0 / 40}, {a1, a1, w1}, {1 - 2 * a1, a1, w1},
                              {a1, 1 - 2 * a1, w1},       {a2, a2, w2}, {1 - 2 * a2, a2, w2},
                              {a2, 1 - 2 * a2, w2}};
    q.nq = 7;
    q.rule.assign(21, 0.0);
    for (int i = 0; i < 7; ++i) {
        q.rule[i] = pts[i][0];
        q.rule[i + 7] = pts[i][1];
        q.rule[i + 14] = 0.5 * pts[i][2];
    }
}

// Gauss-Legendre nodes/weights on [-1, 1] by Newton on P_n
void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w) {
    x.resize(n);
    w.resize(n);
    for (int i = 0; i < n; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1, p1 = z;
            for (int j = 2; j <= n; ++j) {
                double p2 = ((2.0 * j - 1) * z * p1 - (j - 1.0) * p0) / j;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1);
            double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        double p0 = 1, p1 = z;
        for (int j = 2; j <= n; ++j) {
            double p2 = ((2.0 * j - 1) * z * p1 - (j - 1.0) * p0) / j;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (z * p1 - p0) / (z * z - 1);
        x[i] = z;
        w[i] = 2 / ((1 - z * z) * dp * dp);
    }
}

// Duffy-collapsed ng x ng Gauss-Legendre rule with the singular vertex at
// page v's node (barycentric (1,0,0) rotated), weights summing to 1/2.
void duffy(Quad &q, int ng) {
    std::vector<double> x, w;
    gauss_legendre(ng, x, w);
    q.nqs = ng * ng;
    q.pages.assign(9 * (size_t)q.nqs, 0.0);
    for (int v = 0; v < 3; ++v) {
        double *pg = q.pages.data() + (size_t)v * 3 * q.nqs;
        int t = 0;
        for (int i = 0; i < ng; ++i)
            for (int j = 0; j < ng; ++j, ++t) {
                const double u = 0.5 * (x[i] + 1), vv = 0.5 * (x[j] + 1);
                const double wt = 0.25 * w[i] * w[j] * u;
                double bary[3];
                bary[v] = 1 - u;
                bary[(v + 1) % 3] = u * vv;
                bary[(v + 2) % 3] = u * (1 - vv);
                pg[t] = bary[0];
                pg[t + q.nqs] = bary[1];
                pg[t + 2 * q.nqs] = wt;
            }
    }
}

Quad &quad() {
    static thread_local Quad q;
    if (q.nq == 0) {
        gauss7(q);
        duffy(q, 40);
    }
    return q;
}

std::vector<std::string> read_lines(const char *path, ia_status &st) {
    std::vector<std::string> out;
    FILE *fp = fopen(path, "r");
    if (!fp) {
        st = fail(IA_ERR_IO, "Can't open %s", path);
        return out;
    }
    char buf[4096];
    while (fgets(buf, sizeof(buf), fp)) {
        std::string s(buf);
        while (!s.empty() && (s.back() == '\n' || s.back() == '\r' || s.back() == ' ')) s.pop_back();
        if (!s.empty()) out.push_back(s);
    }
    fclose(fp);
    st = IA_OK;
    return out;
}

struct Mesh {
    int m = 0, ne = 0;
    std::vector<int> E, adj_ptr, adj;
    // singular entries of the surface rows: (row o, column c) with o and c
    // nodes of one element, and per entry the (element, o's vertex, c's
    // vertex) triples of the elements holding both
    std::vector<int> sing_ptr, sing_oc, sing_list;
};

ia_status build_mesh(int m, int n, const int *e_n, Mesh &M) {
    M.m = m;
    M.ne = n;
    M.E.assign(e_n, e_n + 3 * (size_t)n);
    M.adj_ptr.assign(m + 1, 0);
    for (int e = 0; e < n; ++e)
        for (int a = 0; a < 3; ++a) {
            int v = M.E[e + a * n];
            if (v < 0 || v >= m) return fail(IA_ERR_ARG, "element %d has node %d out of range", e, v);
            M.adj_ptr[v + 1]++;
        }
    for (int v = 0; v < m; ++v) M.adj_ptr[v + 1] += M.adj_ptr[v];
    M.adj.assign(M.adj_ptr[m], 0);
    std::vector<int> fill(M.adj_ptr.begin(), M.adj_ptr.end() - 1);
    for (int e = 0; e < n; ++e)
        for (int a = 0; a < 3; ++a) {
            int v = M.E[e + a * n];
            if (a > 0 && (v == M.E[e] || (a == 2 && v == M.E[e + n])))
                return fail(IA_ERR_ARG, "element %d is degenerate", e);
            M.adj[fill[v]++] = e;
        }
    M.sing_ptr.assign(1, 0);
    M.sing_oc.clear();
    M.sing_list.clear();
    std::vector<int> cols;
    for (int o = 0; o < m; ++o) {
        cols.clear();
        for (int q = M.adj_ptr[o]; q < M.adj_ptr[o + 1]; ++q)
            for (int a = 0; a < 3; ++a) cols.push_back(M.E[M.adj[q] + a * n]);
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        for (int c : cols) {
            for (int q = M.adj_ptr[o]; q < M.adj_ptr[o + 1]; ++q) {
                const int e = M.adj[q];
                int ikn = -1, a = -1;
                for (int t = 0; t < 3; ++t) {
                    if (M.E[e + t * n] == o) ikn = t;
                    if (M.E[e + t * n] == c) a = t;
                }
                if (a < 0) continue;
                M.sing_list.insert(M.sing_list.end(), {e, ikn, a});
            }
            M.sing_oc.push_back(o);
            M.sing_oc.push_back(c);
            M.sing_ptr.push_back((int)(M.sing_list.size() / 3));
        }
    }
    return IA_OK;
}

// Device-side IBEM system for nf wavenumbers: Z (m x m) and H_b (l x m), FP64,
// with the assembly and factorisation scratch (buffers only grow, so a kept
// IbemDev allocates nothing on a repeated call of the same size).
struct IbemDev {
    DevBuf X, E, adj_ptr, adj, sing_ptr, sing_oc, sing_list, rule, pages, cp, rlh, k, Z, H;
    DevBuf A, B, Ah, Bh, Ad, Bd, Ahd, Bhd, ipiv, info, tau, lw, gw;  // lw, gw: LU scratch
    void release() {
        for (DevBuf *b : {&X, &E, &adj_ptr, &adj, &sing_ptr, &sing_oc, &sing_list, &rule, &pages,
                          &cp, &rlh, &k, &Z, &H, &A, &B, &Ah, &Bh, &Ad, &Bd, &Ahd, &Bhd, &ipiv,
                          &info, &tau, &lw, &gw})
            b->reset();
    }
    void addresses(std::vector<uintptr_t> &v) const {
        for (const DevBuf *b : {&X, &E, &adj_ptr, &adj, &sing_ptr, &sing_oc, &sing_list, &rule,
                                &pages, &cp, &rlh, &k, &Z, &H, &A, &B, &Ah, &Bh, &Ad, &Bd, &Ahd,
                                &Bhd, &ipiv, &info, &tau, &lw, &gw})
            v.push_back((uintptr_t)b->p);
    }
};

// launch sizes of one IBEM build that come from the mesh and the rules
struct IbemShape {
    int nse = 0, nq = 0, nqs = 0;
};

__global__ void k_zero_ints(int *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Host side of an IBEM build for nf wavenumbers: the mesh adjacency, every
// input uploaded and every device buffer sized (nothing of it is captured
// into a graph).
template <typename T>
ia_status ibem_upload(int m, int l, int n, const double *r_ms, const int *e_n,
                      const double *r_lh, int n_ch, const double *cp, const double *k, int nf,
                      IbemDev &D, hipStream_t s, IbemShape &sh) {
    Mesh M;
    IA_TRY(build_mesh(m, n, e_n, M));
    const Quad &Q = quad();
    const int mt = m + n_ch;
    IA_TRY(to_device(D.X, r_ms, sizeof(double) * 3 * m, s));
    IA_TRY(to_device(D.E, M.E.data(), sizeof(int) * 3 * (size_t)n, s));
    IA_TRY(to_device(D.adj_ptr, M.adj_ptr.data(), sizeof(int) * (m + 1), s));
    IA_TRY(to_device(D.adj, M.adj.data(), sizeof(int) * M.adj.size(), s));
    sh.nse = (int)M.sing_oc.size() / 2;
    sh.nq = Q.nq;
    sh.nqs = Q.nqs;
    IA_TRY(to_device(D.sing_ptr, M.sing_ptr.data(), sizeof(int) * M.sing_ptr.size(), s));
    IA_TRY(to_device(D.sing_oc, M.sing_oc.data(), sizeof(int) * M.sing_oc.size(), s));
    IA_TRY(to_device(D.sing_list, M.sing_list.data(), sizeof(int) * M.sing_list.size(), s));
    IA_TRY(to_device(D.rule, Q.rule.data(), sizeof(double) * Q.rule.size(), s));
    IA_TRY(to_device(D.pages, Q.pages.data(), sizeof(double) * Q.pages.size(), s));
    if (l > 0) IA_TRY(to_device(D.rlh, r_lh, sizeof(double) * 3 * l, s));
    if (n_ch > 0) IA_TRY(to_device(D.cp, cp, sizeof(double) * 3 * n_ch, s));
    IA_TRY(to_device(D.k, k, sizeof(double) * nf, s));
    const size_t sS = (size_t)mt * m, sH = (size_t)std::max(l, 1) * m;
    IA_TRY(D.A.alloc(sizeof(cx<T>) * sS * nf));
    IA_TRY(D.B.alloc(sizeof(cx<T>) * sS * nf));
    IA_TRY(D.Ah.alloc(sizeof(cx<T>) * sH * nf));
    IA_TRY(D.Bh.alloc(sizeof(cx<T>) * sH * nf));
    if (sizeof(T) != sizeof(double)) {
        IA_TRY(D.Ad.alloc(sizeof(cx<double>) * sS * nf));
        IA_TRY(D.Bd.alloc(sizeof(cx<double>) * sS * nf));
        IA_TRY(D.Ahd.alloc(sizeof(cx<double>) * sH * nf));
        IA_TRY(D.Bhd.alloc(sizeof(cx<double>) * sH * nf));
    }
    IA_TRY(D.Z.alloc(sizeof(cx<double>) * (size_t)m * m * nf));
    if (n_ch == 0) {
        IA_TRY(D.ipiv.alloc(sizeof(int) * (size_t)m * nf));
        IA_TRY(D.info.alloc(sizeof(int) * nf));
        IA_TRY(D.lw.alloc(sizeof(int) * getrf_work_ints(m, nf)));
        IA_TRY(D.gw.alloc(getrs_work_bytes<double>(m, m, nf)));
    } else {
        IA_TRY(D.tau.alloc(sizeof(cx<double>) * (size_t)m * nf));
    }
    if (l > 0) IA_TRY(D.H.alloc(sizeof(cx<double>) * sH * nf));
    return IA_OK;
}

// Device side of the build, enqueued on s: assembly of Z and H_b, the solve
// for Z (LU, or least squares with CHIEF points) and H.  check: wait for the
// LU and fail on a singular system (without it, the caller reads D.info).
template <typename T>
ia_status ibem_enqueue(int m, int l, int n, int n_ch, int nf, IbemDev &D, const IbemShape &sh,
                       hipStream_t s, bool check) {
    const int mt = m + n_ch;
    const double rho_c = air().rho * air().c;
    DevBuf &bA = D.A, &bB = D.B, &bAh = D.Ah, &bBh = D.Bh;
    const size_t sS = (size_t)mt * m, sH = (size_t)std::max(l, 1) * m;
    const int *dE = D.E.as<int>();
    launch_ibem_rows<T>(m, D.X.as<double>(), m, true, m, D.X.as<double>(), dE, n,
                        D.adj_ptr.as<int>(), D.adj.as<int>(), D.rule.as<double>(), sh.nq,
                        D.k.as<double>(), nf, rho_c, bA.as<cx<T>>(), bB.as<cx<T>>(), mt, sS, s);
    launch_ibem_sing<T>(sh.nse, D.sing_ptr.as<int>(), D.sing_oc.as<int>(), D.sing_list.as<int>(),
                        m, D.X.as<double>(), dE, n, D.pages.as<double>(), sh.nqs,
                        D.k.as<double>(), nf, rho_c, bA.as<cx<T>>(), bB.as<cx<T>>(), mt, sS, s);
    if (n_ch > 0)
        launch_ibem_rows<T>(n_ch, D.cp.as<double>(), n_ch, false, m, D.X.as<double>(), dE, n,
                            D.adj_ptr.as<int>(), D.adj.as<int>(), D.rule.as<double>(), sh.nq,
                            D.k.as<double>(), nf, rho_c, bA.as<cx<T>>() + m, bB.as<cx<T>>() + m,
                            mt, sS, s);
    if (l > 0)
        launch_ibem_rows<T>(l, D.rlh.as<double>(), l, false, m, D.X.as<double>(), dE, n,
                            D.adj_ptr.as<int>(), D.adj.as<int>(), D.rule.as<double>(), sh.nq,
                            D.k.as<double>(), nf, rho_c, bAh.as<cx<T>>(), bBh.as<cx<T>>(), l, sH,
                            s);
    launch_ibem_cdiag<T>(m, D.X.as<double>(), dE, n, D.rule.as<double>(), sh.nq,
                         D.pages.as<double>(), sh.nqs, bA.as<cx<T>>(), mt, sS, nf, nullptr, s);
    IA_HIP(hipGetLastError());
    // solve in FP64 (the mixed form assembles in complex64)
    cx<double> *A, *B, *Ah, *Bh;
    if (sizeof(T) == sizeof(double)) {
        A = reinterpret_cast<cx<double> *>(bA.p);
        B = reinterpret_cast<cx<double> *>(bB.p);
        Ah = reinterpret_cast<cx<double> *>(bAh.p);
        Bh = reinterpret_cast<cx<double> *>(bBh.p);
    } else {
        A = D.Ad.as<cx<double>>();
        B = D.Bd.as<cx<double>>();
        Ah = D.Ahd.as<cx<double>>();
        Bh = D.Bhd.as<cx<double>>();
        launch_convert<double, float>(sS * nf, reinterpret_cast<cx<float> *>(bA.p), A, s);
        launch_convert<double, float>(sS * nf, reinterpret_cast<cx<float> *>(bB.p), B, s);
        launch_convert<double, float>(sH * nf, reinterpret_cast<cx<float> *>(bAh.p), Ah, s);
        launch_convert<double, float>(sH * nf, reinterpret_cast<cx<float> *>(bBh.p), Bh, s);
    }
    cx<double> *Z = D.Z.as<cx<double>>();
    if (n_ch == 0) {
        hipLaunchKernelGGL(k_zero_ints, dim3((nf + 255) / 256), dim3(256), 0, s, D.info.as<int>(),
                           nf);
        IA_TRY(getrf_batched<double>(m, A, m, sS, D.ipiv.as<int>(), D.info.as<int>(), nf, s,
                                     D.lw.as<int>()));
        IA_TRY(getrs_batched<double>('N', m, m, A, m, sS, D.ipiv.as<int>(), B, m, sS, nf, s,
                                     D.gw.p));
        if (check) {
            std::vector<int> hinfo(nf);
            IA_HIP(hipMemcpy(hinfo.data(), D.info.p, sizeof(int) * nf, hipMemcpyDeviceToHost));
            for (int f = 0; f < nf; ++f)
                if (hinfo[f])
                    return fail(IA_ERR_SINGULAR, "IBEM system singular at k[%d] (pivot %d)", f,
                                hinfo[f]);
        }
        IA_HIP(hipMemcpyAsync(Z, B, sizeof(cx<double>) * sS * nf, hipMemcpyDeviceToDevice, s));
    } else {
        // least squares through Householder QR (the reference's ZZgels)
        IA_TRY(launch_geqr2<double>(mt, m, A, mt, sS, D.tau.as<cx<double>>(), nf, s));
        launch_apply_left<double>(true, mt, m, A, mt, sS, D.tau.as<cx<double>>(), B, mt, m, nf, s);
        IA_TRY(trsm_upper_batched<double>(m, m, A, mt, sS, B, mt, sS, nf, s));
        IA_HIP(hipMemcpy2DAsync(Z, sizeof(cx<double>) * m, B, sizeof(cx<double>) * mt,
                                sizeof(cx<double>) * m, (size_t)m * nf, hipMemcpyDeviceToDevice, s));
    }
    if (l > 0) {
        const cx<double> one{1, 0}, mone{-1, 0};
        launch_gemm<double>('N', 'N', l, m, m, one, Ah, l, sH, Z, m, (size_t)m * m, mone, Bh, l, sH,
                            nf, s);
        IA_HIP(hipMemcpyAsync(D.H.p, Bh, sizeof(cx<double>) * sH * nf, hipMemcpyDeviceToDevice, s));
    }
    IA_HIP(hipGetLastError());
    return IA_OK;
}

template <typename T>
ia_status ibem_build(int m, int l, int n, const double *r_ms, const int *e_n, const double *r_lh,
                     int n_ch, const double *cp, const double *k, int nf, IbemDev &D,
                     hipStream_t s) {
    IbemShape sh;
    IA_TRY(ibem_upload<T>(m, l, n, r_ms, e_n, r_lh, n_ch, cp, k, nf, D, s, sh));
    IA_TRY(ibem_enqueue<T>(m, l, n, n_ch, nf, D, sh, s, true));
    IA_HIP(hipStreamSynchronize(s));
    return IA_OK;
}

ia_status ibem_args(int m, int l, int n, const double *r_ms, const int *e_n, const double *r_lh,
                    int n_ch, const double *cp) {
    if (m <= 0 || n <= 0 || l < 0 || n_ch < 0) return fail(IA_ERR_ARG, "IBEM: bad sizes");
    if (!r_ms || !e_n || (l > 0 && !r_lh) || (n_ch > 0 && !cp))
        return fail(IA_ERR_ARG, "IBEM: NULL input");
    if (!finite_all(r_ms, 3 * (size_t)m)) return fail(IA_ERR_NONFINITE, "IBEM: non-finite nodes");
    return IA_OK;
}

// The IBEM sweep's device workspace, kept per host thread between calls;
// ia_release() frees it.
// With IA_GRAPH: the captured sweep, its key and the stream it replays on.
struct IbemSweepCache {
    IbemDev D;
    DevBuf b, x, p, lam, idx, kt, sm, W;
    hipStream_t gs = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<uintptr_t> gkey;
    void drop_graph() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
        gkey.clear();
    }
};
IbemSweepCache &ibem_cache() {
    static thread_local IbemSweepCache *c = new IbemSweepCache();
    return *c;
}

}  // namespace

void release_ibem_cache() {
    IbemSweepCache &c = ibem_cache();
    c.drop_graph();
    c.D.release();
    for (DevBuf *b : {&c.b, &c.x, &c.p, &c.lam, &c.idx, &c.kt, &c.sm, &c.W}) b->reset();
}

}  // namespace ia

using namespace ia;

extern "C" {

ia_status ia_quadrature_load(const char *ipf_path, const char *ipfs_path, int *n_q, double **ipf,
                             int *n_qs, double **ipfs) {
    if (!ipf_path || !ipfs_path || !n_q || !ipf || !n_qs || !ipfs)
        return fail(IA_ERR_ARG, "ia_quadrature_load: NULL argument");
    ia_status st;
    auto L = read_lines(ipf_path, st);
    IA_TRY(st);
    if (L.empty()) return fail(IA_ERR_IO, "%s: empty", ipf_path);
    const int nq = atoi(L[0].c_str());
    if (nq <= 0 || (int)L.size() < nq + 1) return fail(IA_ERR_IO, "%s: bad point count", ipf_path);
    std::vector<double> r(3 * (size_t)nq);
    for (int i = 0; i < nq; ++i)
        if (sscanf(L[i + 1].c_str(), "%lf %lf %lf", &r[i], &r[i + nq], &r[i + 2 * nq]) != 3)
            return fail(IA_ERR_IO, "%s: malformed row %d", ipf_path, i + 2);
    auto S = read_lines(ipfs_path, st);
    IA_TRY(st);
    if (S.empty()) return fail(IA_ERR_IO, "%s: empty", ipfs_path);
    const int nqs = atoi(S[0].c_str());
    if (nqs <= 0) return fail(IA_ERR_IO, "%s: bad point count", ipfs_path);
    std::vector<double> p(9 * (size_t)nqs);
    int page = -1, i = 0;
    for (size_t t = 1; t < S.size(); ++t) {
        if (S[t].rfind("$PAGE", 0) == 0) {
            page = atoi(S[t].c_str() + 5) - 1;
            i = 0;
            if (page < 0 || page > 2) return fail(IA_ERR_IO, "%s: bad page tag", ipfs_path);
            continue;
        }
        if (page < 0 || i >= nqs) return fail(IA_ERR_IO, "%s: row outside a page", ipfs_path);
        double *pg = p.data() + (size_t)page * 3 * nqs;
        if (sscanf(S[t].c_str(), "%lf %lf %lf", &pg[i], &pg[i + nqs], &pg[i + 2 * nqs]) != 3)
            return fail(IA_ERR_IO, "%s: malformed row %zu", ipfs_path, t + 1);
        ++i;
    }
    *n_q = nq;
    *n_qs = nqs;
    *ipf = (double *)malloc(sizeof(double) * r.size());
    *ipfs = (double *)malloc(sizeof(double) * p.size());
    if (!*ipf || !*ipfs) return fail(IA_ERR_NOMEM, "ia_quadrature_load: malloc");
    memcpy(*ipf, r.data(), sizeof(double) * r.size());
    memcpy(*ipfs, p.data(), sizeof(double) * p.size());
    return IA_OK;
}

ia_status ia_set_quadrature(int n_q, const double *ipf, int n_qs, const double *ipfs) {
    Quad &q = quad();
    if (n_q == 0 && n_qs == 0) {  // restore the built-in rules
        gauss7(q);
        duffy(q, 40);
        return IA_OK;
    }
    if (n_q <= 0 || n_qs <= 0 || !ipf || !ipfs) return fail(IA_ERR_ARG, "ia_set_quadrature: bad tables");
    q.nq = n_q;
    q.rule.assign(ipf, ipf + 3 * (size_t)n_q);
    q.nqs = n_qs;
    q.pages.assign(ipfs, ipfs + 9 * (size_t)n_qs);
    return IA_OK;
}

ia_status ia_ibem(int m, int l, int n, const double *r_ms, const int *e_n, const double *r_lh,
                  double k, int n_ch, const double *cp, ia_z **z, ia_z **h_b) {
    IA_TRY(check_device());
    IA_TRY(ibem_args(m, l, n, r_ms, e_n, r_lh, n_ch, cp));
    if (!(k > 0) || !std::isfinite(k)) return fail(IA_ERR_ARG, "ia_ibem: k must be > 0");
    IbemDev D;
    IA_TRY(ibem_build<double>(m, l, n, r_ms, e_n, r_lh, n_ch, cp, &k, 1, D, nullptr));
    if (z) IA_TRY(host_out<ia_z>(z, D.Z.p, (size_t)m * m));
    if (h_b && l > 0) IA_TRY(host_out<ia_z>(h_b, D.H.p, (size_t)l * m));
    return IA_OK;
}

void ibem(int m, int l, int n, double *r_ms, int *e_n, double *r_lh, ia_z **m_e, ia_z **h, double k,
          int n_ch, double *cp) {
    if (m_e) *m_e = nullptr;
    if (h) *h = nullptr;
    ia_status st = ia_ibem(m, l, n, r_ms, e_n, r_lh, k, n_ch, cp, m_e, h);
    if (st != IA_OK) fprintf(stderr, "ibem: %s\n", ia_last_error());
}

void ibemgpu(int m, int l, int n, double *r_msg, int *e_ng, double *r_lhg, ia_z **me, ia_z **he,
             double k, int n_ch, double *cp) {
    // inputs may be device pointers (ibem.cuh:263); gather them to the host
    std::vector<double> rs(3 * (size_t)m), rh(3 * (size_t)std::max(l, 0));
    std::vector<int> en(3 * (size_t)n);
    bool ok = hipMemcpy(rs.data(), r_msg, sizeof(double) * rs.size(), hipMemcpyDefault) == hipSuccess &&
              hipMemcpy(en.data(), e_ng, sizeof(int) * en.size(), hipMemcpyDefault) == hipSuccess &&
              (l <= 0 || hipMemcpy(rh.data(), r_lhg, sizeof(double) * rh.size(), hipMemcpyDefault) == hipSuccess);
    if (!ok) {
        (void)hipGetLastError();
        if (me) *me = nullptr;
        if (he) *he = nullptr;
        fprintf(stderr, "ibemgpu: cannot read inputs\n");
        return;
    }
    ibem(m, l, n, rs.data(), en.data(), rh.data(), me, he, k, n_ch, cp);
}

ia_status ia_ibem_sweep(const ia_ibem_desc *desc, const double *k, const double *r_ms,
                        const int *e_n, const double *r_lh, const double *cp, const ia_z *b,
                        ia_sweep_out *out) {
    IA_TRY(check_device());
    if (!desc || !k || !b) return fail(IA_ERR_ARG, "ia_ibem_sweep: NULL argument");
    const int m = desc->m, l = desc->l, n = desc->n_elem, n_ch = desc->n_ch, F = desc->n_freq;
    IA_TRY(ibem_args(m, l, n, r_ms, e_n, r_lh, n_ch, cp));
    if (l <= 0 || F <= 0) return fail(IA_ERR_ARG, "ia_ibem_sweep: need microphones and frequencies");
    for (int f = 0; f < F; ++f)
        if (!(k[f] > 0) || !std::isfinite(k[f])) return fail(IA_ERR_ARG, "ia_ibem_sweep: k[%d] must be > 0", f);
    if (!finite_all((const double *)b, 2 * (size_t)l * F))
        return fail(IA_ERR_NONFINITE, "ia_ibem_sweep: non-finite pressures");
    IA_TRY(check_reg(desc->reg, std::min(l, m)));
    const RegDev prm = to_regdev(desc->reg);
    hipStream_t s = nullptr;
    ia_sweep_out o{};
    if (out) o = *out;
    const int chunk = desc->chunk > 0 ? std::min(desc->chunk, F) : std::min(F, 64);
    IbemSweepCache &C = ibem_cache();
    IbemDev &D = C.D;
    const bool mixed = desc->precision == IA_PREC_MIXED;
    DevBuf &db = C.b, &bx = C.x, &bp = C.p, &bl = C.lam, &bi = C.idx, &bk = C.kt, &bs = C.sm,
           &bW = C.W;
    // host side of a chunk: inputs uploaded, buffers sized
    auto upload = [&](int f0, int nf, IbemShape &sh) -> ia_status {
        if (mixed)
            IA_TRY(ibem_upload<float>(m, l, n, r_ms, e_n, r_lh, n_ch, cp, k + f0, nf, D, s, sh));
        else
            IA_TRY(ibem_upload<double>(m, l, n, r_ms, e_n, r_lh, n_ch, cp, k + f0, nf, D, s, sh));
        IA_TRY(to_device(db, b + (size_t)f0 * l, sizeof(ia_z) * (size_t)l * nf, s));
        IA_TRY(bx.alloc(sizeof(cx<double>) * (size_t)m * nf));
        IA_TRY(bp.alloc(sizeof(cx<double>) * (size_t)m * nf));
        IA_TRY(bl.alloc(sizeof(double) * nf));
        IA_TRY(bi.alloc(sizeof(int) * nf));
        IA_TRY(bk.alloc(sizeof(int) * nf));
        IA_TRY(bs.alloc(sizeof(double) * nf));
        return bW.alloc(sizeof(double) * nf);
    };
    // device side: the build, the regularised solve, p_s = Z v and W =
    // 1/2 Re(v^H p_s) S/m; graph: no host reads or synchronisation
    auto enqueue = [&](int nf, const IbemShape &sh, hipStream_t st, bool graph) -> ia_status {
        if (mixed)
            IA_TRY(ibem_enqueue<float>(m, l, n, n_ch, nf, D, sh, st, !graph));
        else
            IA_TRY(ibem_enqueue<double>(m, l, n, n_ch, nf, D, sh, st, !graph));
        const cx<double> *H = D.H.as<cx<double>>();
        if (mixed)
            IA_TRY(graph ? regsolve_async<float>(l, m, nf, H, (size_t)l * m, db.as<cx<double>>(), l,
                                                 prm, bx.as<cx<double>>(), m, bl.as<double>(),
                                                 bi.as<int>(), bk.as<int>(), bs.as<double>(), st)
                         : regsolve_batched<float>(l, m, nf, H, (size_t)l * m,
                                                   db.as<cx<double>>(), l, prm,
                                                   bx.as<cx<double>>(), m, bl.as<double>(),
                                                   bi.as<int>(), bk.as<int>(), bs.as<double>(), st));
        else
            IA_TRY(graph ? regsolve_async<double>(l, m, nf, H, (size_t)l * m, db.as<cx<double>>(),
                                                  l, prm, bx.as<cx<double>>(), m, bl.as<double>(),
                                                  bi.as<int>(), bk.as<int>(), bs.as<double>(), st)
                         : regsolve_batched<double>(l, m, nf, H, (size_t)l * m,
                                                    db.as<cx<double>>(), l, prm,
                                                    bx.as<cx<double>>(), m, bl.as<double>(),
                                                    bi.as<int>(), bk.as<int>(), bs.as<double>(),
                                                    st));
        const cx<double> one{1, 0}, zero{0, 0};
        launch_gemm<double>('N', 'N', m, 1, m, one, D.Z.as<cx<double>>(), m, (size_t)m * m,
                            bx.as<cx<double>>(), m, m, zero, bp.as<cx<double>>(), m, m, nf, st);
        launch_power<double>(m, bx.as<cx<double>>(), m, bp.as<cx<double>>(), m, desc->area / m,
                             bW.as<double>(), nf, st);
        IA_HIP(hipGetLastError());
        return IA_OK;
    };
    auto copy_out = [&](int f0, int nf) -> ia_status {
        if (o.lambda) IA_HIP(hipMemcpy(o.lambda + f0, bl.p, sizeof(double) * nf, hipMemcpyDeviceToHost));
        if (o.lambda_index) IA_HIP(hipMemcpy(o.lambda_index + f0, bi.p, sizeof(int) * nf, hipMemcpyDeviceToHost));
        if (o.k_tsvd) IA_HIP(hipMemcpy(o.k_tsvd + f0, bk.p, sizeof(int) * nf, hipMemcpyDeviceToHost));
        if (o.power) IA_HIP(hipMemcpy(o.power + f0, bW.p, sizeof(double) * nf, hipMemcpyDeviceToHost));
        if (o.sigma_max) IA_HIP(hipMemcpy(o.sigma_max + f0, bs.p, sizeof(double) * nf, hipMemcpyDeviceToHost));
        if (o.velocity)
            IA_HIP(hipMemcpy(o.velocity + (size_t)f0 * m, bx.p, sizeof(ia_z) * (size_t)m * nf, hipMemcpyDeviceToHost));
        if (o.pressure)
            IA_HIP(hipMemcpy(o.pressure + (size_t)f0 * m, bp.p, sizeof(ia_z) * (size_t)m * nf, hipMemcpyDeviceToHost));
        if (o.status)
            for (int f = 0; f < nf; ++f) o.status[f0 + f] = IA_OK;
        return IA_OK;
    };
    if (graph_mode() && chunk >= F) {
        // IA_GRAPH=1 (one chunk): the device side is captured once per key
        // and replayed after each call's uploads; a singular system is
        // reported from the LU's info after the replay
        IbemShape sh;
        IA_TRY(upload(0, F, sh));
        if (mixed)
            IA_TRY(regsolve_reserve<float>(l, m, F));
        else
            IA_TRY(regsolve_reserve<double>(l, m, F));
        std::vector<uintptr_t> key = {
            (uintptr_t)m, (uintptr_t)l, (uintptr_t)n, (uintptr_t)n_ch, (uintptr_t)F, (uintptr_t)mixed,
            (uintptr_t)prm.filter, (uintptr_t)prm.select, (uintptr_t)prm.n_lambda,
            (uintptr_t)prm.k_fixed, dbits(prm.decades), dbits(prm.lambda_fixed),
            dbits(desc->area), dbits(air().rho * air().c), (uintptr_t)sh.nse, (uintptr_t)sh.nq,
            (uintptr_t)sh.nqs};
        D.addresses(key);
        for (const DevBuf *q : {&db, &bx, &bp, &bl, &bi, &bk, &bs, &bW}) key.push_back((uintptr_t)q->p);
        if (mixed)
            regsolve_addresses<float>(key);
        else
            regsolve_addresses<double>(key);
        if (!C.gs) IA_HIP(hipStreamCreateWithFlags(&C.gs, hipStreamNonBlocking));
        if (!C.gexec || C.gkey != key) {
            C.drop_graph();
            IA_HIP(hipStreamSynchronize(s));
            IA_HIP(hipStreamBeginCapture(C.gs, hipStreamCaptureModeThreadLocal));
            const ia_status st = enqueue(F, sh, C.gs, true);
            hipGraph_t g = nullptr;
            const hipError_t ce = hipStreamEndCapture(C.gs, &g);
            if (st == IA_OK && ce == hipSuccess && g &&
                hipGraphInstantiate(&C.gexec, g, nullptr, nullptr, 0) != hipSuccess)
                C.gexec = nullptr;
            if (g) (void)hipGraphDestroy(g);
            IA_TRY(st);
            if (!C.gexec) {
                (void)hipGetLastError();
                return fail(IA_ERR_DEVICE, "ia_ibem_sweep: capture into a graph failed");
            }
            C.gkey = key;
        }
        IA_HIP(hipStreamSynchronize(s));  // the uploads ran on the legacy stream
        IA_HIP(hipGraphLaunch(C.gexec, C.gs));
        IA_HIP(hipStreamSynchronize(C.gs));
        if (n_ch == 0) {
            std::vector<int> hinfo(F);
            IA_HIP(hipMemcpy(hinfo.data(), D.info.p, sizeof(int) * F, hipMemcpyDeviceToHost));
            for (int f = 0; f < F; ++f)
                if (hinfo[f])
                    return fail(IA_ERR_SINGULAR, "IBEM system singular at k[%d] (pivot %d)", f,
                                hinfo[f]);
        }
        return copy_out(0, F);
    }
    for (int f0 = 0; f0 < F; f0 += chunk) {
        const int nf = std::min(chunk, F - f0);
        IbemShape sh;
        IA_TRY(upload(f0, nf, sh));
        IA_TRY(enqueue(nf, sh, s, false));
        IA_HIP(hipStreamSynchronize(s));
        IA_TRY(copy_out(f0, nf));
    }
    return IA_OK;
}

// ------------------------------------------------------------ radiation modes
ia_status ia_invr(int m, int l, int n, const double *r_ms, const int *e_n, const double *r_lh,
                  double k, int n_ch, const double *cp, int n_am, ia_z **q, ia_z **m_r,
                  ia_z **h_r, ia_z **g_mass) {
    IA_TRY(check_device());
    IA_TRY(ibem_args(m, l, n, r_ms, e_n, r_lh, n_ch, cp));
    if (!(k > 0) || !std::isfinite(k)) return fail(IA_ERR_ARG, "ia_invr: k must be > 0");
    if (n_am <= 0) return fail(IA_ERR_ARG, "ia_invr: n_am must be > 0");
    n_am = std::min(n_am, m);  // invrS.m:162-164
    hipStream_t s = nullptr;
    IbemDev D;
    IA_TRY(ibem_build<double>(m, l, n, r_ms, e_n, r_lh, n_ch, cp, &k, 1, D, s));
    // G (invr.cuh:603-625), R = real(G Z) (invrV.m:112)
    const Quad &Q = quad();
    DevBuf bG, bGz, bR;
    IA_TRY(bG.alloc(sizeof(double) * (size_t)m * m));
    launch_gram(m, D.X.as<double>(), D.E.as<int>(), n, D.adj_ptr.as<int>(), D.adj.as<int>(),
                D.rule.as<double>(), Q.nq, bG.as<double>(), s);
    IA_TRY(bGz.alloc(sizeof(cx<double>) * (size_t)m * m));
    launch_real_to_cx((size_t)m * m, bG.as<double>(), bGz.as<cx<double>>(), s);
    IA_TRY(bR.alloc(sizeof(cx<double>) * (size_t)m * m));
    const cx<double> one{1, 0}, zero{0, 0};
    launch_gemm<double>('N', 'N', m, m, m, one, bGz.as<cx<double>>(), m, 0, D.Z.as<cx<double>>(),
                        m, 0, zero, bR.as<cx<double>>(), m, 0, 1, s);
    launch_keep_real((size_t)m * m, bR.as<cx<double>>(), s);
    // left singular vectors of R (invrS.m:161): explicit U = U_b [P_B; 0]
    SvdWork<double> sw;
    IA_TRY(sw.alloc(m, m, 1));
    IA_TRY(svd_factor<double>(sw, bR.as<cx<double>>(), s));
    DevBuf bU;
    IA_TRY(bU.alloc(sizeof(cx<double>) * (size_t)m * m));
    launch_rows_to_cols(m, m, sw.PBt.as<double>(), bU.as<cx<double>>(), 1, s);
    svd_apply_u<double>(sw, bR.as<cx<double>>(), bU.as<cx<double>>(), m, n_am, s);
    // modes of a real symmetric matrix are real up to a phase: rotate each to
    // real with its largest entry positive
    launch_realign(m, n_am, 1, bU.as<cx<double>>(), s);
    DevBuf bMr, bHr;
    IA_TRY(bMr.alloc(sizeof(cx<double>) * (size_t)m * n_am));
    launch_gemm<double>('N', 'N', m, n_am, m, one, D.Z.as<cx<double>>(), m, 0, bU.as<cx<double>>(),
                        m, 0, zero, bMr.as<cx<double>>(), m, 0, 1, s);
    if (l > 0) {
        IA_TRY(bHr.alloc(sizeof(cx<double>) * (size_t)l * n_am));
        launch_gemm<double>('N', 'N', l, n_am, m, one, D.H.as<cx<double>>(), l, 0,
                            bU.as<cx<double>>(), m, 0, zero, bHr.as<cx<double>>(), l, 0, 1, s);
    }
    IA_HIP(hipGetLastError());
    IA_HIP(hipDeviceSynchronize());
    if (q) IA_TRY(host_out<ia_z>(q, bU.p, (size_t)m * n_am));
    if (m_r) IA_TRY(host_out<ia_z>(m_r, bMr.p, (size_t)m * n_am));
    if (h_r && l > 0) IA_TRY(host_out<ia_z>(h_r, bHr.p, (size_t)l * n_am));
    if (g_mass) IA_TRY(host_out<ia_z>(g_mass, bGz.p, (size_t)m * m));
    return IA_OK;
}

// argument order of CUDA/invr.cuh:457 (m_e, h, q_e)
void invr(int m, int l, int n, double *r_ms, int *e_n, double *r_lh, ia_z **m_e, ia_z **h, ia_z **q,
          double k, int n_ch, double *cp, int n_am) {
    if (q) *q = nullptr;
    if (m_e) *m_e = nullptr;
    if (h) *h = nullptr;
    ia_status st = ia_invr(m, l, n, r_ms, e_n, r_lh, k, n_ch, cp, n_am, q, m_e, h, nullptr);
    if (st != IA_OK) fprintf(stderr, "invr: %s\n", ia_last_error());
}

void invrgpu(int m, int l, int n, double *r_msg, int *e_ng, double *r_lhg, ia_z **me, ia_z **he,
             ia_z **q, double k, int n_ch, double *cp, int n_rm) {
    std::vector<double> rs(3 * (size_t)m), rh(3 * (size_t)std::max(l, 0));
    std::vector<int> en(3 * (size_t)n);
    bool ok = hipMemcpy(rs.data(), r_msg, sizeof(double) * rs.size(), hipMemcpyDefault) == hipSuccess &&
              hipMemcpy(en.data(), e_ng, sizeof(int) * en.size(), hipMemcpyDefault) == hipSuccess &&
              (l <= 0 || hipMemcpy(rh.data(), r_lhg, sizeof(double) * rh.size(), hipMemcpyDefault) == hipSuccess);
    if (!ok) {
        (void)hipGetLastError();
        if (q) *q = nullptr;
        if (me) *me = nullptr;
        if (he) *he = nullptr;
        fprintf(stderr, "invrgpu: cannot read inputs\n");
        return;
    }
    invr(m, l, n, rs.data(), en.data(), rh.data(), me, he, q, k, n_ch, cp, n_rm);
}

ia_status ia_radiation_modes(int np, const double *r_p, double a_e, double k, int n_modes,
                             double **eigval, double **modes) {
    IA_TRY(check_device());
    if (np <= 0 || !r_p || !(a_e > 0) || !(k > 0) || n_modes <= 0)
        return fail(IA_ERR_ARG, "ia_radiation_modes: bad arguments");
    n_modes = std::min(n_modes, np);
    const double c = air().c, rho = air().rho, w = k * c;
    const double scale = w * w * rho * a_e * a_e / (4 * M_PI * c);
    hipStream_t s = nullptr;
    // S = R / scale = sin(k r_ij)/(k r_ij) assembled on the device
    DevBuf bp, bk, bR, bU, bM;
    IA_TRY(to_device(bp, r_p, sizeof(double) * 3 * (size_t)np, s));
    IA_TRY(to_device(bk, &k, sizeof(double), s));
    IA_TRY(bR.alloc(sizeof(cx<double>) * (size_t)np * np));
    launch_rad_R(np, bk.as<double>(), bp.as<double>(), bR.as<cx<double>>(), 1, s);
    SvdWork<double> sw;
    IA_TRY(sw.alloc(np, np, 1));
    IA_TRY(svd_factor<double>(sw, bR.as<cx<double>>(), s));
    IA_TRY(bU.alloc(sizeof(cx<double>) * (size_t)np * np));
    launch_rows_to_cols(np, np, sw.PBt.as<double>(), bU.as<cx<double>>(), 1, s);
    svd_apply_u<double>(sw, bR.as<cx<double>>(), bU.as<cx<double>>(), np, n_modes, s);
    launch_realign(np, n_modes, 1, bU.as<cx<double>>(), s);
    IA_TRY(bM.alloc(sizeof(double) * (size_t)np * n_modes));
    launch_cx_to_real((size_t)np * n_modes, bU.as<cx<double>>(), bM.as<double>(), s);
    IA_HIP(hipGetLastError());
    IA_HIP(hipDeviceSynchronize());
    std::vector<double> sig(n_modes);
    IA_HIP(hipMemcpy(sig.data(), sw.sig.p, sizeof(double) * n_modes, hipMemcpyDeviceToHost));
    IA_TRY(host_out<double>(modes, bM.p, (size_t)np * n_modes));
    *eigval = (double *)malloc(sizeof(double) * n_modes);
    if (!*eigval) return fail(IA_ERR_NOMEM, "ia_radiation_modes: malloc");
    for (int j = 0; j < n_modes; ++j) (*eigval)[j] = sig[j] * scale;
    return IA_OK;
}

}  // extern "C"
