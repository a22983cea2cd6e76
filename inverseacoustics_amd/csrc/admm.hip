// ADMM regularised solutions (MATLAB/ADMMG.m:1-89), batched over systems.
//
// For A (m x n) and b: M = 2 A^H A + rho I is inverted once per system
// (batched LU on the matrix cores), r0 = 2 A^H b, and the iteration
//     c = M^-1 (r0 + rho x - u),  x = shrink(c + u / rho),
//     u <- u + gamma rho (c - x)
// runs in one persistent workgroup per system: each iteration streams M^-1
// once (one matrix-vector product, HBM-bound) and does the shrinkage, the
// dual update and the three convergence norms in registers; every system
// stops at its own convergence (|c - c_old|, |c - x|, |u - u_old| <= eps).
// Shrinkage: L1 complex soft threshold with the continuation
// lambda = 5 lambda_0 * 0.75^k while lambda > lambda_0 / 10 (ADMMG.m:3-30),
// L2 x = (rho c + u) / (2 lambda_0 + rho) (ADMMG.m:31-48), Lp generalised
// soft thresholding with N fixed-point steps (ADMMG.m:49-88; Zuo et al.,
// ICCV 2013).  The start x_0 = A\b as MATLAB's mldivide computes it
// (ADMMG.m:7,35,55): for a rectangular A the basic solution of Householder
// QR with column pivoting (k_qrcp_basic, LAPACK zlaqp2's pivot order and
// norm downdating), for a square A the solution of the system (TSVD at the
// numerical rank, which is A^-1 b for a nonsingular A).
#include "householder.hpp"

namespace ia {

template <typename T>
ia_status regsolve_batched(int rows, int cols, int batch, const cx<double> *A, size_t sA,
                           const cx<double> *b, size_t sb, const RegDev &prm, cx<double> *x,
                           size_t sx, double *lam, int *idx, int *kts, double *smax, hipStream_t s,
                           double *gap = nullptr);

namespace {

using hh::BT;
using hh::NW;
using hh::block_sum;

struct AdmmDev {
    double lambda0, rho, gamma, eps2, p;
    int max_iter, norm, n_gst;
};

// N fixed-point steps s <- |y| - lp p s^(p-1) from s = |y| (entries above
// the threshold only, where s stays positive)
__device__ __forceinline__ double gst_mag(double ay, double lam_p, double p, int N) {
    double s = ay;
    for (int d = 0; d < N; ++d) s = ay - lam_p * p * pow(s, p - 1.0);
    return s;
}

__global__ __launch_bounds__(BT) void k_admm(int n, const cx<double> *__restrict__ Minvall,
                                            const cx<double> *__restrict__ r0all,
                                            cx<double> *__restrict__ xall,
                                            cx<double> *__restrict__ uall,
                                            cx<double> *__restrict__ call, AdmmDev prm,
                                            int *__restrict__ iters) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<double> *r = reinterpret_cast<cx<double> *>(smem_raw);  // n
    __shared__ double red[NW];
    const int b = blockIdx.x, t = threadIdx.x;
    const cx<double> *Minv = Minvall + (size_t)b * n * n;
    const cx<double> *r0 = r0all + (size_t)b * n;
    cx<double> *x = xall + (size_t)b * n;
    cx<double> *u = uall + (size_t)b * n;
    cx<double> *c = call + (size_t)b * n;
    const double rho = prm.rho, gr = prm.gamma * prm.rho;
    for (int i = t; i < n; i += BT) {
        c[i] = cx<double>{0, 0};
        u[i] = (prm.norm == 3) ? x[i] : cx<double>{0, 0};  // ADMMG.m:56 starts the Lp dual at x_0
    }
    const double lam_e = prm.lambda0 / 10.0;
    double lam = (prm.norm == 2) ? prm.lambda0 : 5.0 * prm.lambda0;
    int it = 0;
    bool done = false;
    __syncthreads();
    for (;;) {
        // the continuation runs while lambda > lambda_0 / 10 (ADMMG.m:10,59)
        if (prm.norm != 2 && !(lam > lam_e)) break;
        const double lr = lam / rho;              // L1 threshold
        const double lam_p = 2.0 * lam / rho;     // Lp
        const double tp = 2.0 * lam_p * (1.0 - prm.p);
        const double tau = (prm.norm == 3)
                               ? pow(tp, 1.0 / (2.0 - prm.p)) +
                                     lam_p * prm.p * pow(tp, (prm.p - 1.0) / (2.0 - prm.p))
                               : 0.0;
        const double dlr = 2.0 * prm.lambda0 + rho;
        for (int k = 0; k < prm.max_iter; ++k) {
            for (int i = t; i < n; i += BT) r[i] = r0[i] + rho * x[i] - u[i];
            __syncthreads();
            double d1 = 0, d2 = 0, d3 = 0;
            for (int i = t; i < n; i += BT) {
                cx<double> a0{0, 0}, a1{0, 0};
                int j = 0;
                for (; j + 1 < n; j += 2) {
                    a0 = a0 + Minv[i + (size_t)j * n] * r[j];
                    a1 = a1 + Minv[i + (size_t)(j + 1) * n] * r[j + 1];
                }
                if (j < n) a0 = a0 + Minv[i + (size_t)j * n] * r[j];
                const cx<double> cn = a0 + a1, ci = c[i], ui = u[i];
                cx<double> xn;
                if (prm.norm == 2) {
                    xn = (1.0 / dlr) * (rho * cn + ui);
                } else {
                    const cx<double> y = cn + (1.0 / rho) * ui;
                    const double ay = sqrt(abs2(y));
                    if (prm.norm == 1) {
                        const double m = fmax(0.0, ay - lr);
                        xn = (ay > 0) ? (m / ay) * y : cx<double>{m, 0};
                    } else {
                        xn = (ay > tau && ay > 0) ? (gst_mag(ay, lam_p, prm.p, prm.n_gst) / ay) * y
                                                  : cx<double>{0, 0};
                    }
                }
                const cx<double> un = ui + gr * (cn - xn);
                d1 += abs2(cn - ci);
                d2 += abs2(cn - xn);
                d3 += abs2(un - ui);
                c[i] = cn;
                x[i] = xn;
                u[i] = un;
            }
            // block_sum syncs: every r[] read is finished before the next
            // iteration rewrites it, and x, u are visible to the workgroup
            d1 = block_sum(d1, red);
            d2 = block_sum(d2, red);
            d3 = block_sum(d3, red);
            ++it;
            if (d1 <= prm.eps2 && d2 <= prm.eps2 && d3 <= prm.eps2) {
                done = true;
                break;
            }
        }
        if (prm.norm == 2 || done) break;
        lam = fmax(0.75 * lam, lam_e);
    }
    if (t == 0 && iters) iters[b] = it;
}

// M = 2 A^H A + rho I: add rho to the diagonal of the product
// x = A \ b for a rectangular A (m x n, column-major copy W, b copy y; both
// overwritten), one workgroup per system: Householder QR with column
// pivoting as LAPACK zlaqp2 -- step j takes the column with the largest
// remaining norm (the first on ties), reflects it (zlarfg), applies H^H to
// the columns right of it and to y, and downdates their norms, recomputing
// a norm when the downdate has lost more than half its digits
// ((1 - (|a_jc| / vn1)^2) (vn1 / vn2)^2 <= sqrt(eps)); then the rank
// r = #{|R_ii| > max(m, n) eps |R_00|} and the basic solution
// x[jp[i]] = (R[:r, :r]^-1 (Q^H b)[:r])_i, zero elsewhere.  Scratch per
// system: norms vn (2n doubles) and the column order jp (n ints).
__global__ __launch_bounds__(BT) void k_qrcp_basic(int m, int n, cx<double> *__restrict__ Wall,
                                                  cx<double> *__restrict__ yall,
                                                  double *__restrict__ vall,
                                                  int *__restrict__ jall,
                                                  cx<double> *__restrict__ xall) {
    __shared__ double red[NW];
    __shared__ double rv[NW];
    __shared__ int ri[NW];
    __shared__ int s_int;
    __shared__ cx<double> s_x;
    const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    cx<double> *W = Wall + (size_t)b * m * n;
    cx<double> *y = yall + (size_t)b * m;
    double *vn1 = vall + (size_t)b * 2 * n, *vn2 = vn1 + n;
    int *jp = jall + (size_t)b * n;
    cx<double> *x = xall + (size_t)b * n;
    const double tol3z = 1.0536712127723509e-08;  // sqrt(2^-53), LAPACK's sqrt(dlamch('E'))
    for (int c = wv; c < n; c += NW) {
        const cx<double> *col = W + (size_t)c * m;
        double q = 0;
        for (int r = lane; r < m; r += 64) q += abs2(col[r]);
        q = hh::wave_sum(q);
        if (lane == 0) {
            vn1[c] = vn2[c] = sqrt(q);
            jp[c] = c;
        }
    }
    __syncthreads();
    const int k = min(m, n);
    for (int j = 0; j < k; ++j) {
        // ---- pivot: the first column of largest vn1 among j..n-1
        double best = -1.0;
        int bi = n;
        for (int c = j + t; c < n; c += BT)
            if (vn1[c] > best) {
                best = vn1[c];
                bi = c;
            }
        for (int o = 32; o > 0; o >>= 1) {
            const double ob = __shfl_xor(best, o, 64);
            const int oi = __shfl_xor(bi, o, 64);
            if (ob > best || (ob == best && oi < bi)) {
                best = ob;
                bi = oi;
            }
        }
        if (lane == 0) {
            rv[wv] = best;
            ri[wv] = bi;
        }
        __syncthreads();
        if (t == 0) {
            double bb = rv[0];
            int ii = ri[0];
            for (int w = 1; w < NW; ++w)
                if (rv[w] > bb || (rv[w] == bb && ri[w] < ii)) {
                    bb = rv[w];
                    ii = ri[w];
                }
            s_int = ii;
        }
        __syncthreads();
        const int p = s_int < n ? s_int : j;  // every remaining norm NaN: keep column j
        if (p != j) {
            cx<double> *cj = W + (size_t)j * m, *cp = W + (size_t)p * m;
            for (int r = t; r < m; r += BT) {
                const cx<double> a = cj[r];
                cj[r] = cp[r];
                cp[r] = a;
            }
            if (t == 0) {
                const int q = jp[p];
                jp[p] = jp[j];
                jp[j] = q;
                vn1[p] = vn1[j];
                vn2[p] = vn2[j];
            }
        }
        __syncthreads();
        // ---- reflector of column j (rows j..m-1)
        cx<double> *cj = W + (size_t)j * m;
        double q = 0;
        for (int r = j + 1 + t; r < m; r += BT) q += abs2(cj[r]);
        const double xn2 = block_sum(q, red);
        const hh::Refl rf = hh::larfg(cj[j], xn2);
        __syncthreads();  // cj[j] read by all before it is overwritten
        for (int r = j + 1 + t; r < m; r += BT) cj[r] = cj[r] * rf.scale;
        if (t == 0) cj[j] = cx<double>{rf.beta, 0};
        __syncthreads();
        // ---- H^H = I - conj(tau) v v^H on columns j+1..n-1 and on y (c == n),
        // one wave per column; lane 0 holds row j
        const cx<double> ctau = conj(rf.tau);
        for (int c = j + 1 + wv; c <= n; c += NW) {
            cx<double> *col = (c < n) ? W + (size_t)c * m : y;
            double sr = 0, si = 0;
            for (int r = j + lane; r < m; r += 64) {
                const cx<double> v = (r == j) ? cx<double>{1, 0} : cj[r];
                const cx<double> a = col[r];
                sr += v.re * a.re + v.im * a.im;
                si += v.re * a.im - v.im * a.re;
            }
            const cx<double> f = ctau * cx<double>{hh::wave_sum(sr), hh::wave_sum(si)};
            double below = 0;
            cx<double> ajc{0, 0};
            for (int r = j + lane; r < m; r += 64) {
                const cx<double> v = (r == j) ? cx<double>{1, 0} : cj[r];
                const cx<double> a = col[r] - v * f;
                col[r] = a;
                if (r == j)
                    ajc = a;
                else
                    below += abs2(a);
            }
            if (c < n) {
                const double are = __shfl(ajc.re, 0, 64), aim = __shfl(ajc.im, 0, 64);
                const double below2 = hh::wave_sum(below);
                if (lane == 0 && vn1[c] != 0.0) {
                    const double v1 = vn1[c], v2 = vn2[c];
                    const double rat = sqrt(are * are + aim * aim) / v1;
                    const double temp = fmax(1.0 - rat * rat, 0.0);
                    const double temp2 = temp * (v1 / v2) * (v1 / v2);
                    if (temp2 <= tol3z) {
                        const double nv = (j < m - 1) ? sqrt(below2) : 0.0;
                        vn1[c] = nv;
                        vn2[c] = nv;
                    } else {
                        vn1[c] = v1 * sqrt(temp);
                    }
                }
            }
        }
        __syncthreads();
    }
    // ---- rank and the basic solution
    if (t == 0) {
        const double tol = (double)max(m, n) * 2.220446049250313e-16 * sqrt(abs2(W[0]));
        int r = 0;
        for (int i = 0; i < k; ++i)
            if (sqrt(abs2(W[(size_t)i * m + i])) > tol) ++r;
        s_int = r;
    }
    __syncthreads();
    const int r = s_int;
    for (int i = r - 1; i >= 0; --i) {
        if (t == 0) {
            s_x = cdiv(y[i], W[(size_t)i * m + i]);
            y[i] = s_x;
        }
        __syncthreads();
        const cx<double> xi = s_x;
        const cx<double> *ci = W + (size_t)i * m;
        for (int q2 = t; q2 < i; q2 += BT) y[q2] = y[q2] - ci[q2] * xi;
        __syncthreads();
    }
    for (int c = t; c < n; c += BT) x[c] = cx<double>{0, 0};
    __syncthreads();
    for (int i = t; i < r; i += BT) x[jp[i]] = y[i];
}

__global__ void k_add_diag(int n, cx<double> *M, double rho) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
    if (i < n) M[(size_t)b * n * n + (size_t)i * n + i].re += rho;
}

__global__ void k_identity(int n, cx<double> *I) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, col = blockIdx.y, b = blockIdx.z;
    if (i < n) I[(size_t)b * n * n + (size_t)col * n + i] = cx<double>{i == col ? 1.0 : 0.0, 0.0};
}

}  // namespace

ia_status admm_batched(int rows, int cols, int batch, const cx<double> *A, const cx<double> *b,
                       const ia_admm_params &P, cx<double> *xout, int *iters, hipStream_t s) {
    const int n = cols;
    const size_t lds = sizeof(cx<double>) * (size_t)n;
    if (lds > 128 * 1024) return fail(IA_ERR_ARG, "ia_admm: %d unknowns exceed the on-chip limit", n);
    DevBuf bM, bMi, bipiv, binfo, br0, bx, bu, bc, bit;
    const size_t nn = (size_t)n * n;
    IA_TRY(bM.alloc(sizeof(cx<double>) * nn * batch));
    IA_TRY(bMi.alloc(sizeof(cx<double>) * nn * batch));
    IA_TRY(bipiv.alloc(sizeof(int) * (size_t)n * batch));
    IA_TRY(binfo.alloc(sizeof(int) * batch));
    IA_TRY(br0.alloc(sizeof(cx<double>) * (size_t)n * batch));
    IA_TRY(bx.alloc(sizeof(cx<double>) * (size_t)n * batch));
    IA_TRY(bu.alloc(sizeof(cx<double>) * (size_t)n * batch));
    IA_TRY(bc.alloc(sizeof(cx<double>) * (size_t)n * batch));
    IA_TRY(bit.alloc(sizeof(int) * batch));
    const size_t sA = (size_t)rows * cols;
    const cx<double> two{2, 0}, zero{0, 0};
    // M = 2 A^H A + rho I, M^-1 by LU and n right-hand sides
    launch_gemm<double>('C', 'N', n, n, rows, two, A, rows, sA, A, rows, sA, zero,
                        bM.as<cx<double>>(), n, nn, batch, s);
    hipLaunchKernelGGL(k_add_diag, dim3((n + 255) / 256, batch), dim3(256), 0, s, n,
                       bM.as<cx<double>>(), P.rho);
    hipLaunchKernelGGL(k_identity, dim3((n + 255) / 256, n, batch), dim3(256), 0, s, n,
                       bMi.as<cx<double>>());
    IA_HIP(hipMemsetAsync(binfo.p, 0, sizeof(int) * batch, s));
    IA_TRY(getrf_batched<double>(n, bM.as<cx<double>>(), n, nn, bipiv.as<int>(), binfo.as<int>(),
                                 batch, s));
    IA_TRY(getrs_batched<double>('N', n, n, bM.as<cx<double>>(), n, nn, bipiv.as<int>(),
                                 bMi.as<cx<double>>(), n, nn, batch, s));
    // r0 = 2 A^H b
    launch_gemm<double>('C', 'N', n, 1, rows, two, A, rows, sA, b, rows, rows, zero,
                        br0.as<cx<double>>(), n, n, batch, s);
    // x0 = A \ b as MATLAB's mldivide: the QR-with-column-pivoting basic
    // solution for a rectangular A, the solution of the system for a square
    // one (TSVD at the numerical rank: A^-1 b when A is nonsingular)
    DevBuf bW, by, bv, bj;
    if (rows != cols) {
        IA_TRY(bW.alloc(sizeof(cx<double>) * sA * batch));
        IA_TRY(by.alloc(sizeof(cx<double>) * (size_t)rows * batch));
        IA_TRY(bv.alloc(sizeof(double) * 2 * (size_t)cols * batch));
        IA_TRY(bj.alloc(sizeof(int) * (size_t)cols * batch));
        IA_HIP(hipMemcpyAsync(bW.p, A, sizeof(cx<double>) * sA * batch, hipMemcpyDeviceToDevice, s));
        IA_HIP(hipMemcpyAsync(by.p, b, sizeof(cx<double>) * (size_t)rows * batch,
                              hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_qrcp_basic, dim3(batch), dim3(BT), 0, s, rows, cols,
                           bW.as<cx<double>>(), by.as<cx<double>>(), bv.as<double>(),
                           bj.as<int>(), bx.as<cx<double>>());
        IA_HIP(hipGetLastError());
    } else {
        RegDev rk{IA_FILTER_TSVD, IA_SELECT_RANK, 50, 0, 8.0, 0.0};
        IA_TRY(regsolve_batched<double>(rows, cols, batch, A, sA, b, rows, rk,
                                        bx.as<cx<double>>(), n, nullptr, nullptr, nullptr,
                                        nullptr, s));
    }
    AdmmDev d{P.lambda0, P.rho, P.gamma, P.epsilon * P.epsilon, P.p, P.max_iter, P.norm,
              P.n_gst};
    hipLaunchKernelGGL(k_admm, dim3(batch), dim3(BT), lds, s, n, bMi.as<cx<double>>(),
                       br0.as<cx<double>>(), bx.as<cx<double>>(), bu.as<cx<double>>(),
                       bc.as<cx<double>>(), d, bit.as<int>());
    IA_HIP(hipGetLastError());
    IA_HIP(hipMemcpyAsync(xout, bc.p, sizeof(cx<double>) * (size_t)n * batch, hipMemcpyDefault, s));
    if (iters) IA_HIP(hipMemcpyAsync(iters, bit.p, sizeof(int) * batch, hipMemcpyDefault, s));
    IA_HIP(hipStreamSynchronize(s));
    std::vector<int> info(batch);
    IA_HIP(hipMemcpy(info.data(), binfo.p, sizeof(int) * batch, hipMemcpyDeviceToHost));
    for (int f = 0; f < batch; ++f)
        if (info[f]) return fail(IA_ERR_SINGULAR, "ia_admm: 2 A^H A + rho I singular in system %d", f);
    return IA_OK;
}

}  // namespace ia

using namespace ia;

extern "C" ia_status ia_admm(int rows, int cols, int batch, const ia_z *a, const ia_z *b,
                             const ia_admm_params *prm, ia_z *x, int *iterations) {
    IA_TRY(check_device());
    if (rows <= 0 || cols <= 0 || batch <= 0 || !a || !b || !prm || !x)
        return fail(IA_ERR_ARG, "ia_admm: bad arguments");
    const ia_admm_params &P = *prm;
    if (P.norm < 1 || P.norm > 3) return fail(IA_ERR_ARG, "ia_admm: norm must be 1, 2 or 3");
    if (!(P.rho > 0) || !(P.gamma > 0) || !(P.epsilon >= 0) || P.max_iter <= 0 ||
        !(P.lambda0 >= 0) || !std::isfinite(P.lambda0) || !std::isfinite(P.rho))
        return fail(IA_ERR_ARG, "ia_admm: bad parameters");
    if (P.norm == 3 && (!(P.p > 0 && P.p < 1) || P.n_gst <= 0))
        return fail(IA_ERR_ARG, "ia_admm: Lp needs 0 < p < 1 and n_gst > 0");
    hipStream_t s = nullptr;
    DevBuf dA, db, dx;
    IA_TRY(to_device(dA, a, sizeof(ia_z) * (size_t)rows * cols * batch, s));
    IA_TRY(to_device(db, b, sizeof(ia_z) * (size_t)rows * batch, s));
    IA_TRY(dx.alloc(sizeof(ia_z) * (size_t)cols * batch));
    DevBuf dit;
    IA_TRY(dit.alloc(sizeof(int) * batch));
    IA_TRY(admm_batched(rows, cols, batch, dA.as<cx<double>>(), db.as<cx<double>>(), P,
                        dx.as<cx<double>>(), dit.as<int>(), s));
    IA_HIP(hipMemcpy(x, dx.p, sizeof(ia_z) * (size_t)cols * batch, hipMemcpyDefault));
    if (iterations) IA_HIP(hipMemcpy(iterations, dit.p, sizeof(int) * batch, hipMemcpyDefault));
    return IA_OK;
}
