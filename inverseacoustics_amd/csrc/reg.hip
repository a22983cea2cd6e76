// Regularisation-parameter choice and filtered coefficients, one workgroup
// per frequency, everything in FP64.
//
// Restates MATLAB/regtools: gcvG.m:19-52 (GCV), l_curveG.m:21-53 with
// lcfunG.m:1-49 (L-curve), quasioptG.m:16-41, tikhonovG.m:26-30,
// tsvdG.m:16-23 (tsvd.cuh:122-172 with betak tsvd.cuh:1-8), dsvdG.m:20-26,
// plus:
//   * the configured grid lam_j = s_1 10^(-dec + dec j/(n-1)) with Hansen's
//     standard-form Tikhonov GCV (f = lam^2/(s^2+lam^2)), argmin with ties to
//     the lowest index;
//   * the discrete TSVD GCV G(k) = (sum_{i>k}|beta_i|^2 + delta0)/(rows-k)^2,
//     k = 1..p-1 (Hansen's regtools gcv.m 'tsvd' branch, which the reference
//     does not ship).
// Output phi satisfies x = V phi, so the caller back-transforms one vector.
#include "ia_internal.hpp"

namespace ia {

namespace {

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Forsythe-Malcolm-Moler FMIN (golden section + parabolic interpolation),
// the algorithm of MATLAB fminbnd with TolX = 1e-4 (gcvG.m:51).
template <typename F>
__device__ double fminbnd(F fun, double ax, double bx, double tol_x) {
    const double seps = 1.4901161193847656e-08;  // sqrt(eps)
    const double cg = 0.5 * (3.0 - sqrt(5.0));
    double a = ax, b = bx;
    double v = a + cg * (b - a), w = v, xf = v;
    double d = 0, e = 0;
    double fx = fun(xf), fv = fx, fw = fx;
    double xm = 0.5 * (a + b);
    double tol1 = seps * fabs(xf) + tol_x / 3.0, tol2 = 2.0 * tol1;
    int nfev = 1;
    while (fabs(xf - xm) > (tol2 - 0.5 * (b - a))) {
        bool golden = true;
        if (fabs(e) > tol1) {
            golden = false;
            double r = (xf - w) * (fx - fv);
            double q = (xf - v) * (fx - fw);
            double p = (xf - v) * q - (xf - w) * r;
            q = 2.0 * (q - r);
            if (q > 0.0) p = -p;
            q = fabs(q);
            r = e;
            e = d;
            if (fabs(p) < fabs(0.5 * q * r) && p > q * (a - xf) && p < q * (b - xf)) {
                d = p / q;
                double u = xf + d;
                if ((u - a) < tol2 || (b - u) < tol2) d = (xm - xf) >= 0 ? tol1 : -tol1;
            } else {
                golden = true;
            }
        }
        if (golden) {
            e = (xf >= xm) ? (a - xf) : (b - xf);
            d = cg * e;
        }
        double u = xf + (d >= 0 ? 1.0 : -1.0) * fmax(fabs(d), tol1);
        double fu = fun(u);
        ++nfev;
        if (fu <= fx) {
            if (u >= xf)
                a = xf;
            else
                b = xf;
            v = w;
            fv = fw;
            w = xf;
            fw = fx;
            xf = u;
            fx = fu;
        } else {
            if (u < xf)
                a = u;
            else
                b = u;
            if (fu <= fw || w == xf) {
                v = w;
                fv = fw;
                w = u;
                fw = fu;
            } else if (fu <= fv || v == xf || v == w) {
                v = u;
                fv = fu;
            }
        }
        xm = 0.5 * (a + b);
        tol1 = seps * fabs(xf) + tol_x / 3.0;
        tol2 = 2.0 * tol1;
        if (nfev >= 500) break;
    }
    return xf;
}

// lcfunG.m: negative L-curve curvature at one lambda (sequential over p);
// pool = energy of the unresolved components (filter limit s -> 0: f = 0)
__device__ double lcfun1(double la, int p, const double *s, const double *ab2, const double *axi2,
                         bool tik, bool LS, double rhoLS2, double pool) {
    double eta2 = 0, rho2 = pool, phi = 0, psi = 0, dphi = 0, dpsi = 0;
    for (int i = 0; i < p; ++i) {
        double f = tik ? (s[i] * s[i]) / (s[i] * s[i] + la * la) : s[i] / (s[i] + la);
        double cf = 1.0 - f;
        double f1 = -2.0 * f * cf / la;
        double f2 = -f1 * (3.0 - 4.0 * f) / la;
        eta2 += f * f * axi2[i];
        rho2 += cf * cf * ab2[i];
        phi += f * f1 * axi2[i];
        psi += cf * f1 * ab2[i];
        dphi += (f1 * f1 + f * f2) * axi2[i];
        dpsi += (-f1 * f1 + cf * f2) * ab2[i];
    }
    if (LS) rho2 += rhoLS2;
    double eta = sqrt(eta2), rho = sqrt(rho2);
    double deta = phi / eta, drho = -psi / rho;
    double ddeta = dphi / eta - deta * (deta / eta);
    double ddrho = -dpsi / rho - drho * (drho / rho);
    double dle = deta / eta, dlr = drho / rho;
    double ddle = ddeta / eta - dle * dle, ddlr = ddrho / rho - dlr * dlr;
    return -(dlr * ddle - ddlr * dle) / pow(dlr * dlr + dle * dle, 1.5);
}

__global__ __launch_bounds__(256) void k_select(int p, int rows, const double *__restrict__ sall,
                                               const cx<double> *__restrict__ ball,
                                               const double *__restrict__ bn2all,
                                               const double *__restrict__ cn2all, double floor_rel,
                                               RegDev prm, double *lam_out, int *idx_out,
                                               int *k_out, double *gap_out,
                                               cx<double> *__restrict__ phiall,
                                               double *__restrict__ curve_out) {
    extern __shared__ double sh[];
    double *s = sh;            // p
    double *ab2 = sh + p;      // |beta|^2, p
    double *tail = sh + 2 * p; // suffix sums, p+1
    double *axi2 = sh + 3 * p + 1;  // |beta/s|^2 (0 where s = 0), p
    __shared__ double Gv[512];
    __shared__ int Gi[512];
    __shared__ double s_total, s_pool, s_lam, s_gap, s_gapk;
    __shared__ int s_k, s_idx;
    const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const double *sg = sall + (size_t)b * p;
    const cx<double> *beta = ball + (size_t)b * p;
    cx<double> *phi = phiall + (size_t)b * p;
    // Singular values below floor_rel * s_1 sit at the round-off floor of the
    // SVD: their vectors are not resolved (a cluster's computed vectors need
    // not be orthogonal), so their coordinates are pooled.  The pool's energy
    // is |c|^2 minus the resolved part (c = the vector whose coordinates in
    // the singular basis are beta, |c| exact under the Householder
    // transforms) and it enters every sum with the filter's s -> 0 limit,
    // which differs from the per-component weight by O((floor/lambda)^2).
    // When that energy is within the orthogonality error of the computed
    // vectors (below 1e-9 |c|^2 for an FP64 SVD, 1e-5 for complex64) the
    // difference carries no information and the pooled coordinates' own sum
    // is kept; with nothing below the floor there is no pool.
    const double sfloor = floor_rel * sg[0];
    for (int i = t; i < p; i += blockDim.x) {
        s[i] = sg[i];
        const bool res = !(sg[i] < sfloor);
        ab2[i] = res ? abs2(beta[i]) : 0.0;
        axi2[i] = (res && sg[i] > 0) ? ab2[i] / (sg[i] * sg[i]) : 0.0;
    }
    __syncthreads();
    if (t == 0) {
        double acc = 0, allb = 0;
        tail[p] = 0;
        for (int i = p - 1; i >= 0; --i) {
            acc += ab2[i];
            tail[i] = acc;
        }
        for (int i = 0; i < p; ++i) allb += abs2(beta[i]);
        const double tot = cn2all ? cn2all[b] : allb;
        int npool = 0;
        for (int i = 0; i < p; ++i) npool += (sg[i] < sfloor) ? 1 : 0;
        const double direct = allb - acc;  // sum of the pooled |beta|^2
        const double comp = tot - acc;
        const double ortho = (floor_rel < 1e-9) ? 1e-9 : 1e-5;
        const double pool = (npool == 0) ? 0.0 : (comp > ortho * tot) ? comp : fmax(direct, 0.0);
        for (int i = 0; i <= p; ++i) tail[i] += pool;
        s_total = tot;
        s_pool = pool;
    }
    __syncthreads();
    const double bn2 = bn2all[b];
    const double total = s_total;
    const double pool = s_pool;
    const double delta0 = (rows > p && bn2 - total > 0) ? bn2 - total : 0.0;
    const double smax = s[0];

    // ---- discrete TSVD GCV, k = 1..p-1 (argmin, lowest k on ties)
    {
        double best = INFINITY;
        int bk = p;
        for (int k = 1 + t; k < p; k += blockDim.x) {
            double den = (double)(rows - k);
            double G = (tail[k] + delta0) / (den * den);
            if (G < best) {
                best = G;
                bk = k;
            }
        }
        Gv[t] = best;
        Gi[t] = bk;
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if (t < st) {
                if (Gv[t + st] < Gv[t] || (Gv[t + st] == Gv[t] && Gi[t + st] < Gi[t])) {
                    Gv[t] = Gv[t + st];
                    Gi[t] = Gi[t + st];
                }
            }
            __syncthreads();
        }
        if (t == 0) s_k = (p > 1) ? Gi[0] : p;
        __syncthreads();
        // relative gap to the runner-up k (how far a perturbation of the
        // spectrum must move G before the chosen index changes)
        const double gbest = Gv[0];
        __syncthreads();
        double second = INFINITY;
        for (int k = 1 + t; k < p; k += blockDim.x) {
            if (k == s_k) continue;
            double den = (double)(rows - k);
            second = fmin(second, (tail[k] + delta0) / (den * den));
        }
        Gv[t] = second;
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if (t < st) Gv[t] = fmin(Gv[t], Gv[t + st]);
            __syncthreads();
        }
        if (t == 0) s_gapk = (gbest > 0 && gbest < INFINITY) ? (Gv[0] - gbest) / gbest : INFINITY;
        if (t == 0) s_gap = INFINITY;
        __syncthreads();
    }
    const bool tik = (prm.filter == IA_FILTER_TIKHONOV);
    double lam = 0.0;
    int idx = -1;
    int kk = s_k;
    if (prm.filter != IA_FILTER_TSVD) {
        if (prm.select == IA_SELECT_GCV || prm.select == IA_SELECT_LCURVE) {
            const int nl = prm.n_lambda;
            for (int j = wv; j < nl; j += blockDim.x / 64) {
                double la = smax * pow(10.0, -prm.decades + prm.decades * j / (double)(nl - 1));
                double val;
                if (prm.select == IA_SELECT_GCV) {
                    double num = 0, sf = 0;
                    for (int i = lane; i < p; i += 64) {
                        double fc = tik ? (la * la) / (s[i] * s[i] + la * la) : la / (s[i] + la);
                        num += fc * fc * ab2[i];
                        sf += fc;
                    }
                    num = wsum(num) + pool;
                    sf = wsum(sf);
                    double den = (double)(rows - p) + sf;
                    val = (num + delta0) / (den * den);
                } else {
                    double sums[6] = {0, 0, 0, 0, 0, 0};
                    for (int i = lane; i < p; i += 64) {
                        double f = tik ? (s[i] * s[i]) / (s[i] * s[i] + la * la) : s[i] / (s[i] + la);
                        double cf = 1.0 - f, f1 = -2.0 * f * cf / la, f2 = -f1 * (3.0 - 4.0 * f) / la;
                        double axi2 = (s[i] > 0) ? ab2[i] / (s[i] * s[i]) : 0.0;
                        sums[0] += f * f * axi2;
                        sums[1] += cf * cf * ab2[i];
                        sums[2] += f * f1 * axi2;
                        sums[3] += cf * f1 * ab2[i];
                        sums[4] += (f1 * f1 + f * f2) * axi2;
                        sums[5] += (-f1 * f1 + cf * f2) * ab2[i];
                    }
#pragma unroll
                    for (int q = 0; q < 6; ++q) sums[q] = wsum(sums[q]);
                    sums[1] += pool;
                    if (rows > p) sums[1] += delta0;
                    double eta = sqrt(sums[0]), rho = sqrt(sums[1]);
                    double deta = sums[2] / eta, drho = -sums[3] / rho;
                    double ddeta = sums[4] / eta - deta * (deta / eta);
                    double ddrho = -sums[5] / rho - drho * (drho / rho);
                    double dle = deta / eta, dlr = drho / rho;
                    double ddle = ddeta / eta - dle * dle, ddlr = ddrho / rho - dlr * dlr;
                    val = -(dlr * ddle - ddlr * dle) / pow(dlr * dlr + dle * dle, 1.5);
                    if (val != val) val = INFINITY;
                }
                if (lane == 0) {
                    Gv[j] = val;
                    Gi[j] = j;
                    if (curve_out) curve_out[(size_t)b * nl + j] = val;
                }
            }
            __syncthreads();
            if (t == 0) {
                double best = INFINITY;
                int bj = 0;
                for (int j = 0; j < nl; ++j)
                    if (Gv[j] < best) {
                        best = Gv[j];
                        bj = j;
                    }
                double second = INFINITY;
                for (int j = 0; j < nl; ++j)
                    if (j != bj) second = fmin(second, Gv[j]);
                s_gap = (best < INFINITY && best != 0) ? (second - best) / fabs(best) : INFINITY;
                s_idx = bj;
                s_lam = smax * pow(10.0, -prm.decades + prm.decades * bj / (double)(nl - 1));
            }
            __syncthreads();
            lam = s_lam;
            idx = s_idx;
        } else if (prm.select == IA_SELECT_FIXED) {
            lam = prm.lambda_fixed;
        } else {
            // literal regtools modes: 200-point grid + fminbnd, one thread
            if (t == 0) {
                const double eps = 2.220446049250313e-16;
                const int np = 200;
                double rpm, ratio;
                if (prm.select == IA_SELECT_QUASIOPT) {
                    rpm = s[p - 1];
                    ratio = pow(s[0] / s[p - 1], 1.0 / (np - 1));
                } else {
                    rpm = fmax(s[p - 1], s[0] * 16.0 * eps);
                    ratio = pow(s[0] / rpm, 1.0 / (np - 1));
                }
                double best = INFINITY, second = INFINITY;
                int bj = 0;
                const bool LS = (rows > p);
                const double m_n = (double)(rows - p);
                for (int j = 0; j < np; ++j) {
                    double la = pow(ratio, (double)(np - 1 - j)) * rpm;
                    double val = 0;
                    if (prm.select == IA_SELECT_GCV_REGTOOLS) {
                        // gcvG.m:38-40: f = lam/(s+lam) on the grid for both methods
                        double num = pool, sf = 0;
                        for (int i = 0; i < p; ++i) {
                            double f = la / (s[i] + la);
                            num += f * f * ab2[i];
                            sf += f;
                        }
                        val = (num + delta0) / ((m_n + sf) * (m_n + sf));
                    } else if (prm.select == IA_SELECT_LCURVE_REGTOOLS) {
                        val = lcfun1(la, p, s, ab2, axi2, tik, LS, delta0, pool);
                    } else {
                        double q2 = 0;
                        for (int i = 0; i < p; ++i) {
                            double f = tik ? s[i] * s[i] / (s[i] * s[i] + la * la) : s[i] / (s[i] + la);
                            double xi2 = ab2[i] / (s[i] * s[i]);
                            q2 += (1 - f) * (1 - f) * f * f * xi2;
                        }
                        val = sqrt(q2);
                    }
                    if (val < best) {
                        second = best;
                        best = val;
                        bj = j;
                    } else if (val < second) {
                        second = val;
                    }
                }
                s_gap = (best < INFINITY && best != 0) ? (second - best) / fabs(best) : INFINITY;
                double lo = pow(ratio, (double)(np - 1 - min(bj + 1, np - 1))) * rpm;
                double hi = pow(ratio, (double)(np - 1 - max(bj - 1, 0))) * rpm;
                double res;
                if (prm.select == IA_SELECT_GCV_REGTOOLS) {
                    res = fminbnd(
                        [&](double la) {
                            double num = pool, sf = 0;
                            for (int i = 0; i < p; ++i) {
                                double f = tik ? (la * la) / (s[i] + la * la) : la / (s[i] + la);
                                num += f * f * ab2[i];
                                sf += f;
                            }
                            return (num + delta0) / ((m_n + sf) * (m_n + sf));
                        },
                        lo, hi, 1e-4);
                } else if (prm.select == IA_SELECT_LCURVE_REGTOOLS) {
                    res = fminbnd([&](double la) { return lcfun1(la, p, s, ab2, axi2, tik, LS, delta0, pool); },
                                  lo, hi, 1e-4);
                } else {
                    res = fminbnd(
                        [&](double la) {
                            double q2 = 0;
                            for (int i = 0; i < p; ++i) {
                                double f = tik ? s[i] * s[i] / (s[i] * s[i] + la * la) : s[i] / (s[i] + la);
                                q2 += (1 - f) * (1 - f) * f * f * ab2[i] / (s[i] * s[i]);
                            }
                            return sqrt(q2);
                        },
                        lo, hi, 1e-4);
                }
                s_lam = res;
                s_idx = bj;
            }
            __syncthreads();
            lam = s_lam;
            idx = s_idx;
        }
    } else if (prm.select == IA_SELECT_FIXED) {
        kk = prm.k_fixed;
    } else if (prm.select == IA_SELECT_RANK) {
        // numerical rank with the cutoff of LAPACK-based least squares
        // (numpy lstsq rcond = eps max(rows, cols))
        const int dim = prm.k_fixed > 0 ? prm.k_fixed : max(rows, p);  // max(rows, cols)
        const double cut = 2.220446049250313e-16 * (double)dim * smax;
        int r = 0;
        for (int i = 0; i < p; ++i) r += (s[i] > cut) ? 1 : 0;
        kk = r;
    }
    // ---- filtered coefficients
    for (int i = t; i < p; i += blockDim.x) {
        cx<double> bi = beta[i];
        double g;
        if (prm.filter == IA_FILTER_TIKHONOV)
            g = s[i] / (s[i] * s[i] + lam * lam);
        else if (prm.filter == IA_FILTER_DSVD)
            g = 1.0 / (s[i] + lam);
        else
            g = (i < kk) ? 1.0 / s[i] : 0.0;
        phi[i] = cx<double>{g * bi.re, g * bi.im};
    }
    if (t == 0) {
        if (lam_out) lam_out[b] = lam;
        if (idx_out) idx_out[b] = idx;
        if (k_out) k_out[b] = s_k;
        // gap of the selection that decides x: the lambda search, or the
        // discrete-GCV k when the filter is TSVD
        if (gap_out) gap_out[b] = (prm.filter == IA_FILTER_TSVD) ? s_gapk : s_gap;
    }
}

}  // namespace

void launch_select(int p, int rows, const double *sigma, const cx<double> *beta,
                   const double *bnorm2, RegDev prm, double *lam_out, int *idx_out, int *k_out,
                   cx<double> *phi, int batch, hipStream_t s, double *gap_out,
                   const double *cnorm2, double floor_rel, double *curve_out) {
    size_t lds = sizeof(double) * (4 * (size_t)p + 1);
    (void)hipFuncSetAttribute((const void *)k_select, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    hipLaunchKernelGGL(k_select, dim3(batch), dim3(256), lds, s, p, rows, sigma, beta, bnorm2,
                       cnorm2, floor_rel, prm, lam_out, idx_out, k_out, gap_out, phi, curve_out);
}

}  // namespace ia
