// Analytic line-source scenario (MATLAB/linesource.m:1-65), batched over
// wavenumbers.
//
// A line source of length L and width a vibrating with velocity U0 lies on
// the y axis, centred at the origin, sampled at n_ls points y_j =
// linspace(-L/2, L/2, n_ls).  With G = e^{ikr}/r, r = |x - (0, y_j, 0)|:
//
//   p(x)   = i (rho c U0 k a / 2) trapz_j G                  (linesource.m:33-34,55-56)
//   v_n(x) = (U0 a / 2) trapz_j (ikr - 1) e^{ikr}/r^2 (n.(x - y_j))/r      (:58-60)
//
// p at the field points r_lh and on the surface r_ms, v_n on the surface with
// the node normals of the triangle mesh e_n: the sum over a node's triangles
// of (x1 - x3) x (x2 - x3), normalised (linesource.m:36-53).
//
// Work: one thread per (point, wavenumber); the n_ls-term trapezoid sum runs in
// registers (one sincos, one sqrt, one reciprocal per term), so the kernel is
// bound by FP64 issue, not memory.  Field and surface points share one launch:
// points [0, l) are field points (pressure only), [l, l + m) surface points.
#include "ia_internal.hpp"

#include <cmath>
#include <vector>

namespace ia {

namespace {

// Source abscissa j of linspace(-L/2, L/2, n): the last point is L/2 exactly.
__device__ __forceinline__ double ls_y(int j, int n, double h, double half) {
    return j == n - 1 ? half : -half + j * h;
}

__global__ __launch_bounds__(256) void k_linesource(int l, int m, const double *__restrict__ r_lh,
                                                    const double *__restrict__ r_ms,
                                                    const double *__restrict__ nv, int n_ls,
                                                    double len, const double *__restrict__ kv,
                                                    double amp_p, double amp_v,
                                                    cx<double> *__restrict__ xp,
                                                    cx<double> *__restrict__ xsf,
                                                    cx<double> *__restrict__ vsf) {
    const int f = blockIdx.y;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= l + m) return;
    const bool surf = t >= l;
    const int i = surf ? t - l : t;
    const double *P = surf ? r_ms : r_lh;
    const int np = surf ? m : l;
    const double x = P[i], y = P[i + np], z = P[i + 2 * np];
    double nx = 0, ny = 0, nz = 0;
    if (surf) {
        nx = nv[i];
        ny = nv[i + m];
        nz = nv[i + 2 * m];
    }
    const double k = kv[f];
    const double half = 0.5 * len;
    const double h = n_ls > 1 ? len / (n_ls - 1) : 0.0;
    const double rho2 = x * x + z * z;
    const double ndr0 = nx * x + nz * z;
    double pr = 0, pi = 0, vr = 0, vi = 0;
    double yprev = ls_y(0, n_ls, h, half);
    for (int j = 0; j < n_ls; ++j) {
        const double yj = yprev;
        const double ynext = j + 1 < n_ls ? ls_y(j + 1, n_ls, h, half) : yj;
        // trapezoid weight: half the span of the neighbouring intervals
        const double w = 0.5 * (ynext - (j > 0 ? ls_y(j - 1, n_ls, h, half) : yj));
        yprev = ynext;
        const double dy = y - yj;
        const double r = sqrt(rho2 + dy * dy);
        const double ir = 1.0 / r;
        double sn, cs;
        sincos(k * r, &sn, &cs);
        const double wg = w * ir;
        pr += wg * cs;
        pi += wg * sn;
        if (surf) {
            // (ikr - 1) e^{ikr} / r^2 * (n.(x - y_j)) / r
            const double kr = k * r;
            const double s = w * ir * ir * ir * (ndr0 + ny * dy);
            vr += s * (-cs - kr * sn);
            vi += s * (kr * cs - sn);
        }
    }
    const size_t o = (size_t)f * np + i;
    // i * amp_p * (pr + i pi)
    cx<double> p{-amp_p * k * pi, amp_p * k * pr};
    if (surf) {
        xsf[o] = p;
        vsf[o] = cx<double>{amp_v * vr, amp_v * vi};
    } else {
        xp[o] = p;
    }
}

// Node normals of a triangle mesh (linesource.m:36-53): per triangle
// (x1 - x3) x (x2 - x3), summed over the triangles of each node, normalised.
// A node that no triangle uses gets 0/0 = NaN, as in the reference.
ia_status node_normals(int m, int n, const double *r_ms, const int *e_n, std::vector<double> &nv) {
    nv.assign(3 * (size_t)m, 0.0);
    for (int e = 0; e < n; ++e) {
        const int a = e_n[e], b = e_n[e + n], c = e_n[e + 2 * n];
        if (a < 0 || a >= m || b < 0 || b >= m || c < 0 || c >= m)
            return fail(IA_ERR_ARG, "ia_linesource: element %d has a node out of range", e);
        const double d1x = r_ms[a] - r_ms[c], d1y = r_ms[a + m] - r_ms[c + m],
                     d1z = r_ms[a + 2 * m] - r_ms[c + 2 * m];
        const double d2x = r_ms[b] - r_ms[c], d2y = r_ms[b + m] - r_ms[c + m],
                     d2z = r_ms[b + 2 * m] - r_ms[c + 2 * m];
        const double cxv = d1y * d2z - d1z * d2y, cyv = d1z * d2x - d1x * d2z,
                     czv = d1x * d2y - d1y * d2x;
        for (int v : {a, b, c}) {
            nv[v] += cxv;
            nv[v + m] += cyv;
            nv[v + 2 * (size_t)m] += czv;
        }
    }
    for (int i = 0; i < m; ++i) {
        const double s = std::sqrt(nv[i] * nv[i] + nv[i + m] * nv[i + m] +
                                   nv[i + 2 * (size_t)m] * nv[i + 2 * (size_t)m]);
        nv[i] /= s;
        nv[i + m] /= s;
        nv[i + 2 * (size_t)m] /= s;
    }
    return IA_OK;
}

}  // namespace

}  // namespace ia

using namespace ia;

extern "C" ia_status ia_linesource(int n_ls, int m, int n_elem, const double *r_ms,
                                   const int *e_n, int l, const double *r_lh, double rho,
                                   double c, double u0, const double *k, int n_freq, double a,
                                   double len, ia_z *xp, ia_z *xsf, ia_z *vsf) {
    IA_TRY(check_device());
    if (n_ls < 1 || m < 0 || l < 0 || n_elem < 0 || n_freq <= 0 || !k || m + l == 0)
        return fail(IA_ERR_ARG, "ia_linesource: bad sizes");
    if ((m > 0 && (!r_ms || !xsf || !vsf || !e_n || n_elem == 0)) || (l > 0 && (!r_lh || !xp)))
        return fail(IA_ERR_ARG, "ia_linesource: NULL argument");
    if (!std::isfinite(rho) || !std::isfinite(c) || !std::isfinite(u0) || !std::isfinite(a) ||
        !std::isfinite(len) || !(len >= 0))
        return fail(IA_ERR_ARG, "ia_linesource: bad constants");
    std::vector<double> nv;
    if (m > 0) IA_TRY(node_normals(m, n_elem, r_ms, e_n, nv));
    hipStream_t s = nullptr;
    DevBuf dk, dlh, dms, dnv, dxp, dxs, dvs;
    IA_TRY(to_device(dk, k, sizeof(double) * n_freq, s));
    if (l > 0) {
        IA_TRY(to_device(dlh, r_lh, sizeof(double) * 3 * (size_t)l, s));
        IA_TRY(dxp.alloc(sizeof(ia_z) * (size_t)l * n_freq));
    }
    if (m > 0) {
        IA_TRY(to_device(dms, r_ms, sizeof(double) * 3 * (size_t)m, s));
        IA_TRY(to_device(dnv, nv.data(), sizeof(double) * nv.size(), s));
        IA_TRY(dxs.alloc(sizeof(ia_z) * (size_t)m * n_freq));
        IA_TRY(dvs.alloc(sizeof(ia_z) * (size_t)m * n_freq));
    }
    const double amp_p = 0.5 * rho * c * u0 * a;  // times k in the kernel
    const double amp_v = 0.5 * u0 * a;
    // n_ls = 1: trapz over one sample is 0 (the kernel's weight is 0 too).
    dim3 grid((unsigned)((l + m + 255) / 256), (unsigned)n_freq);
    hipLaunchKernelGGL(k_linesource, grid, dim3(256), 0, s, l, m, dlh.as<double>(),
                       dms.as<double>(), dnv.as<double>(), n_ls, len, dk.as<double>(), amp_p, amp_v,
                       dxp.as<cx<double>>(), dxs.as<cx<double>>(), dvs.as<cx<double>>());
    IA_HIP(hipGetLastError());
    if (l > 0) IA_HIP(hipMemcpy(xp, dxp.p, sizeof(ia_z) * (size_t)l * n_freq, hipMemcpyDefault));
    if (m > 0) {
        IA_HIP(hipMemcpy(xsf, dxs.p, sizeof(ia_z) * (size_t)m * n_freq, hipMemcpyDefault));
        IA_HIP(hipMemcpy(vsf, dvs.p, sizeof(ia_z) * (size_t)m * n_freq, hipMemcpyDefault));
    }
    return IA_OK;
}
