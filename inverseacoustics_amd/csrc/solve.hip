// Batched regularised least-squares solve for an arbitrary transfer matrix,
// and the HELS entry points built on it.
//
// regsolve: for each batch item, A (rows x cols) = U diag(s) V^H by the
// bidiagonal SVD of svd.hip, beta = U^H b, parameter choice (reg.hip), and
// x = V phi -- the chain csvd -> gcvG/l_curveG -> tikhonovG/tsvd of the
// reference (tsvd.cuh:10-172, MATLAB/regtools) without forming U or V.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

#include "ia_internal.hpp"

namespace ia {

template <typename T>
void launch_hels(int npts, int N, int nf, const double *k, double rho_c, const double *azelr,
                 cx<T> *M, cx<T> *Q, size_t stride, hipStream_t s);
int hels_nmax();

// A is complex128 on the device; T selects the SVD precision.
template <typename T>
ia_status regsolve_batched(int rows, int cols, int batch, const cx<double> *A, size_t sA,
                           const cx<double> *b, size_t sb, const RegDev &prm, cx<double> *x,
                           size_t sx, double *lam, int *idx, int *kts, double *smax,
                           hipStream_t s, double *gap = nullptr) {
    const bool tall = rows >= cols;
    const int R = tall ? rows : cols, C = tall ? cols : rows;
    DevBuf bY, bYd, bc, bbeta, bphi, bn2, bcn2, bv;
    IA_TRY(bY.alloc(sizeof(cx<T>) * (size_t)R * C * batch));
    cx<T> *Y = bY.as<cx<T>>();
    cx<double> *Yd = nullptr;
    if (sizeof(T) == sizeof(double)) {
        Yd = reinterpret_cast<cx<double> *>(Y);
    } else {
        IA_TRY(bYd.alloc(sizeof(cx<double>) * (size_t)R * C * batch));
        Yd = bYd.as<cx<double>>();
    }
    if (tall) {
        IA_HIP(hipMemcpy2DAsync(Yd, sizeof(cx<double>) * (size_t)R * C, A, sizeof(cx<double>) * sA,
                                sizeof(cx<double>) * (size_t)R * C, batch, hipMemcpyDeviceToDevice,
                                s));
    } else {
        launch_conj_transpose<double>(rows, cols, A, sA, Yd, (size_t)R * C, batch, s);
    }
    if (sizeof(T) != sizeof(double))
        launch_convert<float, double>((size_t)R * C * batch, Yd, reinterpret_cast<cx<float> *>(Y), s);
    SvdWork<T> sw;
    IA_TRY(sw.alloc(R, C, batch));
    IA_TRY(svd_factor<T>(sw, Y, s));
    IA_TRY(bc.alloc(sizeof(cx<double>) * (size_t)R * batch));
    IA_TRY(bbeta.alloc(sizeof(cx<double>) * (size_t)C * batch));
    IA_TRY(bphi.alloc(sizeof(cx<double>) * (size_t)C * batch));
    IA_TRY(bv.alloc(sizeof(cx<double>) * (size_t)R * batch));
    IA_TRY(bn2.alloc(sizeof(double) * batch));
    IA_TRY(bcn2.alloc(sizeof(double) * batch));
    cx<double> *c = bc.as<cx<double>>();
    launch_norm2<double>(rows, b, sb, bn2.as<double>(), batch, s);
    if (tall) {
        // beta = P_B^T (U_b^H b)[:C]
        IA_HIP(hipMemcpy2DAsync(c, sizeof(cx<double>) * R, b, sizeof(cx<double>) * sb,
                                sizeof(cx<double>) * rows, batch, hipMemcpyDeviceToDevice, s));
        svd_apply_uh<T>(sw, Y, c, R, 1, s);
        launch_norm2<double>(C, c, R, bcn2.as<double>(), batch, s);
        launch_real_matvec(C, sw.PBt.template as<double>(), 0, c, R, bbeta.as<cx<double>>(), C,
                           batch, s);
    } else {
        // beta = Q_B^T (V_b^H b)
        IA_HIP(hipMemcpy2DAsync(c, sizeof(cx<double>) * C, b, sizeof(cx<double>) * sb,
                                sizeof(cx<double>) * rows, batch, hipMemcpyDeviceToDevice, s));
        svd_apply_v<T>(sw, true, c, C, 1, s);
        launch_norm2<double>(C, c, C, bcn2.as<double>(), batch, s);
        launch_real_matvec(C, sw.QBt.template as<double>(), 0, c, C, bbeta.as<cx<double>>(), C,
                           batch, s);
    }
    RegDev prm2 = prm;
    if (prm2.select == IA_SELECT_RANK) prm2.k_fixed = std::max(rows, cols);  // rank cutoff dimension
    launch_select(C, rows, sw.sig.template as<double>(), bbeta.as<cx<double>>(), bn2.as<double>(),
                  prm2, lam, idx, kts, bphi.as<cx<double>>(), batch, s, gap, bcn2.as<double>(),
                  svd_floor_rel<T>());
    cx<double> *v = bv.as<cx<double>>();
    if (tall) {
        // x = V_b (Q_B phi)
        launch_real_matvec(C, sw.QBt.template as<double>(), 1, bphi.as<cx<double>>(), C, v, C,
                           batch, s);
        svd_apply_v<T>(sw, false, v, C, 1, s);
        IA_HIP(hipMemcpy2DAsync(x, sizeof(cx<double>) * sx, v, sizeof(cx<double>) * C,
                                sizeof(cx<double>) * C, batch, hipMemcpyDeviceToDevice, s));
    } else {
        // x = U_b [P_B phi; 0]
        launch_real_matvec(C, sw.PBt.template as<double>(), 1, bphi.as<cx<double>>(), C,
                           bbeta.as<cx<double>>(), C, batch, s);
        launch_pad_copy(C, R, bbeta.as<cx<double>>(), C, v, R, batch, s);
        svd_apply_u<T>(sw, Y, v, R, 1, s);
        IA_HIP(hipMemcpy2DAsync(x, sizeof(cx<double>) * sx, v, sizeof(cx<double>) * R,
                                sizeof(cx<double>) * R, batch, hipMemcpyDeviceToDevice, s));
    }
    if (smax)
        IA_HIP(hipMemcpy2DAsync(smax, sizeof(double), sw.sig.p, sizeof(double) * C, sizeof(double),
                                batch, hipMemcpyDeviceToDevice, s));
    IA_HIP(hipGetLastError());
    IA_HIP(hipStreamSynchronize(s));
    return IA_OK;
}

template ia_status regsolve_batched<double>(int, int, int, const cx<double> *, size_t,
                                            const cx<double> *, size_t, const RegDev &,
                                            cx<double> *, size_t, double *, int *, int *, double *,
                                            hipStream_t, double *);
template ia_status regsolve_batched<float>(int, int, int, const cx<double> *, size_t,
                                           const cx<double> *, size_t, const RegDev &, cx<double> *,
                                           size_t, double *, int *, int *, double *, hipStream_t,
                                           double *);

RegDev to_regdev(const ia_reg_params &p) {
    return RegDev{(int)p.filter, (int)p.select, p.n_lambda > 0 ? p.n_lambda : 50, p.k,
                  p.decades > 0 ? p.decades : 8.0, p.lambda};
}

ia_status check_reg(const ia_reg_params &p, int pmax) {
    if (p.filter < IA_FILTER_TIKHONOV || p.filter > IA_FILTER_DSVD)
        return fail(IA_ERR_ARG, "bad filter");
    if (p.select < IA_SELECT_GCV || p.select > IA_SELECT_RANK)
        return fail(IA_ERR_ARG, "bad select");
    if (p.n_lambda > 512) return fail(IA_ERR_ARG, "n_lambda must be <= 512");
    if (p.filter == IA_FILTER_TSVD && p.select == IA_SELECT_FIXED && (p.k < 1 || p.k > pmax))
        return fail(IA_ERR_ARG, "k out of range");
    return IA_OK;
}

}  // namespace ia

using namespace ia;

extern "C" {

// --------------------------------------------------------------- HELS
ia_status ia_hels(int m, int l, const double *r_ms_s, const double *r_lh_s, double k, int n_sh,
                  ia_z **m_e, ia_z **h, ia_z **q_e) {
    IA_TRY(check_device());
    if (m <= 0 || l <= 0 || !r_ms_s || !r_lh_s) return fail(IA_ERR_ARG, "ia_hels: bad sizes");
    if (n_sh < 0 || n_sh > hels_nmax())
        return fail(IA_ERR_ARG, "ia_hels: expansion order must be in [0, %d]", hels_nmax());
    if (!(k > 0) || !std::isfinite(k)) return fail(IA_ERR_ARG, "ia_hels: k must be > 0");
    const int n_t = (n_sh + 1) * (n_sh + 1);
    DevBuf ds, dh, dk, bM, bQ, bH;
    IA_TRY(to_device(ds, r_ms_s, sizeof(double) * 3 * m, nullptr));
    IA_TRY(to_device(dh, r_lh_s, sizeof(double) * 3 * l, nullptr));
    IA_TRY(to_device(dk, &k, sizeof(double), nullptr));
    IA_TRY(bM.alloc(sizeof(ia_z) * (size_t)m * n_t));
    IA_TRY(bQ.alloc(sizeof(ia_z) * (size_t)m * n_t));
    IA_TRY(bH.alloc(sizeof(ia_z) * (size_t)l * n_t));
    const double rho_c = air().rho * air().c;
    launch_hels<double>(m, n_sh, 1, dk.as<double>(), rho_c, ds.as<double>(), bM.as<cx<double>>(),
                        bQ.as<cx<double>>(), 0, nullptr);
    launch_hels<double>(l, n_sh, 1, dk.as<double>(), rho_c, dh.as<double>(), bH.as<cx<double>>(),
                        nullptr, 0, nullptr);
    IA_HIP(hipGetLastError());
    IA_HIP(hipDeviceSynchronize());
    if (m_e) IA_TRY(host_out<ia_z>(m_e, bM.p, (size_t)m * n_t));
    if (q_e) IA_TRY(host_out<ia_z>(q_e, bQ.p, (size_t)m * n_t));
    if (h) IA_TRY(host_out<ia_z>(h, bH.p, (size_t)l * n_t));
    return IA_OK;
}

void hels(int m, int l, double *r_ms_s, double *r_lh_s, ia_z **m_e, ia_z **h, ia_z **q_e, double k,
          int n_sh) {
    if (m_e) *m_e = nullptr;
    if (h) *h = nullptr;
    if (q_e) *q_e = nullptr;
    ia_status st = ia_hels(m, l, r_ms_s, r_lh_s, k, n_sh, m_e, h, q_e);
    if (st != IA_OK) fprintf(stderr, "hels: %s\n", ia_last_error());
}

void helsgpu(int m, int l, double *r_ms_sg, double *r_lh_sg, ia_z **me, ia_z **he, ia_z **qe,
             double k, int n_sh) {
    hels(m, l, r_ms_sg, r_lh_sg, me, he, qe, k, n_sh);
}

ia_status ia_hels_sweep(const ia_hels_desc *desc, const int *order, const double *k,
                        const double *azelr_s, const double *azelr_h, const ia_z *b,
                        ia_sweep_out *out) {
    IA_TRY(check_device());
    if (!desc || !order || !k || !azelr_s || !azelr_h || !b)
        return fail(IA_ERR_ARG, "ia_hels_sweep: NULL argument");
    const int m = desc->m, l = desc->l, F = desc->n_freq;
    if (m <= 0 || l <= 0 || F <= 0) return fail(IA_ERR_ARG, "ia_hels_sweep: bad sizes");
    for (int f = 0; f < F; ++f) {
        if (order[f] < 0 || order[f] > hels_nmax())
            return fail(IA_ERR_ARG, "ia_hels_sweep: order[%d] out of range", f);
        if (!(k[f] > 0) || !std::isfinite(k[f]))
            return fail(IA_ERR_ARG, "ia_hels_sweep: k[%d] must be > 0", f);
    }
    if (!finite_all((const double *)b, 2 * (size_t)l * F))
        return fail(IA_ERR_NONFINITE, "ia_hels_sweep: non-finite pressures");
    int pmax = 1 << 30;
    for (int f = 0; f < F; ++f) pmax = std::min(pmax, std::min(l, (order[f] + 1) * (order[f] + 1)));
    IA_TRY(check_reg(desc->reg, pmax));
    const RegDev prm = to_regdev(desc->reg);
    const double rho_c = air().rho * air().c;
    hipStream_t s = nullptr;
    DevBuf ds, dh;
    IA_TRY(to_device(ds, azelr_s, sizeof(double) * 3 * m, s));
    IA_TRY(to_device(dh, azelr_h, sizeof(double) * 3 * l, s));
    ia_sweep_out o{};
    if (out) o = *out;
    // group frequencies of equal order: one batched launch per system size
    std::map<int, std::vector<int>> groups;
    for (int f = 0; f < F; ++f) groups[order[f]].push_back(f);
    for (auto &g : groups) {
        const int N = g.first, nb = (int)g.second.size(), n_t = (N + 1) * (N + 1);
        std::vector<double> kg(nb);
        std::vector<ia_z> bg((size_t)nb * l);
        for (int i = 0; i < nb; ++i) {
            kg[i] = k[g.second[i]];
            memcpy(&bg[(size_t)i * l], b + (size_t)g.second[i] * l, sizeof(ia_z) * l);
        }
        DevBuf dk, db, bH, bM, bQ, bx, bv, bp, blam, bidx, bkt, bsm, bW;
        IA_TRY(to_device(dk, kg.data(), sizeof(double) * nb, s));
        IA_TRY(to_device(db, bg.data(), sizeof(ia_z) * (size_t)nb * l, s));
        IA_TRY(bH.alloc(sizeof(cx<double>) * (size_t)l * n_t * nb));
        IA_TRY(bM.alloc(sizeof(cx<double>) * (size_t)m * n_t * nb));
        IA_TRY(bQ.alloc(sizeof(cx<double>) * (size_t)m * n_t * nb));
        IA_TRY(bx.alloc(sizeof(cx<double>) * (size_t)n_t * nb));
        IA_TRY(bv.alloc(sizeof(cx<double>) * (size_t)m * nb));
        IA_TRY(bp.alloc(sizeof(cx<double>) * (size_t)m * nb));
        IA_TRY(blam.alloc(sizeof(double) * nb));
        IA_TRY(bidx.alloc(sizeof(int) * nb));
        IA_TRY(bkt.alloc(sizeof(int) * nb));
        IA_TRY(bsm.alloc(sizeof(double) * nb));
        IA_TRY(bW.alloc(sizeof(double) * nb));
        launch_hels<double>(l, N, nb, dk.as<double>(), rho_c, dh.as<double>(), bH.as<cx<double>>(),
                            nullptr, (size_t)l * n_t, s);
        launch_hels<double>(m, N, nb, dk.as<double>(), rho_c, ds.as<double>(), bM.as<cx<double>>(),
                            bQ.as<cx<double>>(), (size_t)m * n_t, s);
        IA_HIP(hipGetLastError());
        ia_status st;
        if (desc->precision == IA_PREC_MIXED)
            st = regsolve_batched<float>(l, n_t, nb, bH.as<cx<double>>(), (size_t)l * n_t,
                                         db.as<cx<double>>(), l, prm, bx.as<cx<double>>(), n_t,
                                         blam.as<double>(), bidx.as<int>(), bkt.as<int>(),
                                         bsm.as<double>(), s);
        else
            st = regsolve_batched<double>(l, n_t, nb, bH.as<cx<double>>(), (size_t)l * n_t,
                                          db.as<cx<double>>(), l, prm, bx.as<cx<double>>(), n_t,
                                          blam.as<double>(), bidx.as<int>(), bkt.as<int>(),
                                          bsm.as<double>(), s);
        IA_TRY(st);
        // v = Q_e c, p = M_e c (to_velocity / to_pressure), W = 1/2 Re(v^H p) S/m
        const cx<double> one{1, 0}, zero{0, 0};
        launch_gemm<double>('N', 'N', m, 1, n_t, one, bQ.as<cx<double>>(), m, (size_t)m * n_t,
                            bx.as<cx<double>>(), n_t, n_t, zero, bv.as<cx<double>>(), m, m, nb, s);
        launch_gemm<double>('N', 'N', m, 1, n_t, one, bM.as<cx<double>>(), m, (size_t)m * n_t,
                            bx.as<cx<double>>(), n_t, n_t, zero, bp.as<cx<double>>(), m, m, nb, s);
        launch_power<double>(m, bv.as<cx<double>>(), m, bp.as<cx<double>>(), m, desc->area / m,
                             bW.as<double>(), nb, s);
        IA_HIP(hipGetLastError());
        IA_HIP(hipStreamSynchronize(s));
        std::vector<double> hl(nb), hs(nb), hW(nb);
        std::vector<int> hi(nb), hk(nb);
        std::vector<ia_z> hv((size_t)m * nb), hp((size_t)m * nb);
        IA_HIP(hipMemcpy(hl.data(), blam.p, sizeof(double) * nb, hipMemcpyDeviceToHost));
        IA_HIP(hipMemcpy(hs.data(), bsm.p, sizeof(double) * nb, hipMemcpyDeviceToHost));
        IA_HIP(hipMemcpy(hW.data(), bW.p, sizeof(double) * nb, hipMemcpyDeviceToHost));
        IA_HIP(hipMemcpy(hi.data(), bidx.p, sizeof(int) * nb, hipMemcpyDeviceToHost));
        IA_HIP(hipMemcpy(hk.data(), bkt.p, sizeof(int) * nb, hipMemcpyDeviceToHost));
        IA_HIP(hipMemcpy(hv.data(), bv.p, sizeof(ia_z) * m * nb, hipMemcpyDeviceToHost));
        IA_HIP(hipMemcpy(hp.data(), bp.p, sizeof(ia_z) * m * nb, hipMemcpyDeviceToHost));
        for (int i = 0; i < nb; ++i) {
            const int f = g.second[i];
            if (o.lambda) o.lambda[f] = hl[i];
            if (o.lambda_index) o.lambda_index[f] = hi[i];
            if (o.k_tsvd) o.k_tsvd[f] = hk[i];
            if (o.power) o.power[f] = hW[i];
            if (o.sigma_max) o.sigma_max[f] = hs[i];
            if (o.velocity) memcpy(o.velocity + (size_t)f * m, &hv[(size_t)i * m], sizeof(ia_z) * m);
            if (o.pressure) memcpy(o.pressure + (size_t)f * m, &hp[(size_t)i * m], sizeof(ia_z) * m);
            if (o.status) o.status[f] = IA_OK;
        }
    }
    return IA_OK;
}

ia_status ia_regsolve(int rows, int cols, int batch, const ia_z *a, const ia_z *b,
                      const ia_reg_params *prm, ia_precision precision, ia_z *x, double *lambda,
                      int *index, int *k_tsvd, double *sigma_max) {
    IA_TRY(check_device());
    if (rows <= 0 || cols <= 0 || batch <= 0 || !a || !b || !prm || !x)
        return fail(IA_ERR_ARG, "ia_regsolve: bad arguments");
    IA_TRY(check_reg(*prm, std::min(rows, cols)));
    hipStream_t s = nullptr;
    DevBuf dA, db, dx, dl, di, dk, dsm;
    IA_TRY(to_device(dA, a, sizeof(ia_z) * (size_t)rows * cols * batch, s));
    IA_TRY(to_device(db, b, sizeof(ia_z) * (size_t)rows * batch, s));
    IA_TRY(dx.alloc(sizeof(ia_z) * (size_t)cols * batch));
    IA_TRY(dl.alloc(sizeof(double) * batch));
    IA_TRY(di.alloc(sizeof(int) * batch));
    IA_TRY(dk.alloc(sizeof(int) * batch));
    IA_TRY(dsm.alloc(sizeof(double) * batch));
    const RegDev rd = to_regdev(*prm);
    if (precision == IA_PREC_MIXED)
        IA_TRY(regsolve_batched<float>(rows, cols, batch, dA.as<cx<double>>(), (size_t)rows * cols,
                                       db.as<cx<double>>(), rows, rd, dx.as<cx<double>>(), cols,
                                       dl.as<double>(), di.as<int>(), dk.as<int>(),
                                       dsm.as<double>(), s));
    else
        IA_TRY(regsolve_batched<double>(rows, cols, batch, dA.as<cx<double>>(), (size_t)rows * cols,
                                        db.as<cx<double>>(), rows, rd, dx.as<cx<double>>(), cols,
                                        dl.as<double>(), di.as<int>(), dk.as<int>(),
                                        dsm.as<double>(), s));
    IA_HIP(hipMemcpy(x, dx.p, sizeof(ia_z) * (size_t)cols * batch, hipMemcpyDefault));
    if (lambda) IA_HIP(hipMemcpy(lambda, dl.p, sizeof(double) * batch, hipMemcpyDefault));
    if (index) IA_HIP(hipMemcpy(index, di.p, sizeof(int) * batch, hipMemcpyDefault));
    if (k_tsvd) IA_HIP(hipMemcpy(k_tsvd, dk.p, sizeof(int) * batch, hipMemcpyDefault));
    if (sigma_max) IA_HIP(hipMemcpy(sigma_max, dsm.p, sizeof(double) * batch, hipMemcpyDefault));
    return IA_OK;
}

}  // extern "C"
