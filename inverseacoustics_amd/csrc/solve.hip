// Batched regularised least-squares solve for an arbitrary transfer matrix,
// and the HELS entry points built on it.
//
// regsolve: for each batch item, A (rows x cols) = U diag(s) V^H by the
// bidiagonal SVD of svd.hip, beta = U^H b, parameter choice (reg.hip), and
// x = V phi -- the chain csvd -> gcvG/l_curveG -> tikhonovG/tsvd of the
// reference (tsvd.cuh:10-172, MATLAB/regtools) without forming U or V.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
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

// Device workspace of one batched regularised solve.  Kept per host thread,
// precision and slot between calls (buffers only grow; ia_release() frees
// them); slot 0 serves the blocking callers, slots 1.. the HELS order
// groups that run side by side on their own streams.
template <typename T>
struct RegsolveWork {
    DevBuf bY, bYd, bc, bbeta, bphi, bn2, bcn2, bv;
    SvdWork<T> sw;
    ia_status reserve(int rows, int cols, int batch) {
        const bool tall = rows >= cols;
        const int R = tall ? rows : cols, C = tall ? cols : rows;
        IA_TRY(bY.alloc(sizeof(cx<T>) * (size_t)R * C * batch));
        if (sizeof(T) != sizeof(double))
            IA_TRY(bYd.alloc(sizeof(cx<double>) * (size_t)R * C * batch));
        IA_TRY(sw.alloc(R, C, batch));
        IA_TRY(bc.alloc(sizeof(cx<double>) * (size_t)R * batch));
        IA_TRY(bbeta.alloc(sizeof(cx<double>) * (size_t)C * batch));
        IA_TRY(bphi.alloc(sizeof(cx<double>) * (size_t)C * batch));
        IA_TRY(bv.alloc(sizeof(cx<double>) * (size_t)R * batch));
        IA_TRY(bn2.alloc(sizeof(double) * batch));
        return bcn2.alloc(sizeof(double) * batch);
    }
    void release() {
        for (DevBuf *b : {&bY, &bYd, &bc, &bbeta, &bphi, &bn2, &bcn2, &bv}) b->reset();
        sw.release();
    }
    void addresses(std::vector<uintptr_t> &v) const {
        for (const DevBuf *b : {&bY, &bYd, &bc, &bbeta, &bphi, &bn2, &bcn2, &bv})
            v.push_back((uintptr_t)b->p);
        sw.addresses(v);
    }
};

constexpr int HELS_SLOTS = 8;  // order groups in flight at once

template <typename T>
RegsolveWork<T> &regsolve_work(int slot) {
    static thread_local RegsolveWork<T> *w[HELS_SLOTS + 1] = {};
    if (!w[slot]) w[slot] = new RegsolveWork<T>();
    return *w[slot];
}

// Enqueue the solve on s with workspace W (nothing waits; W and the outputs
// stay in use until s reaches this point).  A is complex128 on the device;
// T selects the SVD precision.
template <typename T>
ia_status regsolve_enqueue(RegsolveWork<T> &W, int rows, int cols, int batch,
                           const cx<double> *A, size_t sA, const cx<double> *b, size_t sb,
                           const RegDev &prm, cx<double> *x, size_t sx, double *lam, int *idx,
                           int *kts, double *smax, hipStream_t s, double *gap) {
    const bool tall = rows >= cols;
    const int R = tall ? rows : cols, C = tall ? cols : rows;
    IA_TRY(W.reserve(rows, cols, batch));
    cx<T> *Y = W.bY.template as<cx<T>>();
    cx<double> *Yd = sizeof(T) == sizeof(double) ? reinterpret_cast<cx<double> *>(Y)
                                                 : W.bYd.template as<cx<double>>();
    if (tall) {
        IA_HIP(hipMemcpy2DAsync(Yd, sizeof(cx<double>) * (size_t)R * C, A, sizeof(cx<double>) * sA,
                                sizeof(cx<double>) * (size_t)R * C, batch, hipMemcpyDeviceToDevice,
                                s));
    } else {
        launch_conj_transpose<double>(rows, cols, A, sA, Yd, (size_t)R * C, batch, s);
    }
    if (sizeof(T) != sizeof(double))
        launch_convert<float, double>((size_t)R * C * batch, Yd, reinterpret_cast<cx<float> *>(Y), s);
    SvdWork<T> &sw = W.sw;
    IA_TRY(svd_factor<T>(sw, Y, s));
    cx<double> *c = W.bc.template as<cx<double>>();
    launch_norm2<double>(rows, b, sb, W.bn2.template as<double>(), batch, s);
    if (tall) {
        // beta = P_B^T (U_b^H b)[:C]
        IA_HIP(hipMemcpy2DAsync(c, sizeof(cx<double>) * R, b, sizeof(cx<double>) * sb,
                                sizeof(cx<double>) * rows, batch, hipMemcpyDeviceToDevice, s));
        svd_apply_uh<T>(sw, Y, c, R, 1, s);
        launch_norm2<double>(C, c, R, W.bcn2.template as<double>(), batch, s);
        launch_real_matvec(C, sw.PBt.template as<double>(), 0, c, R, W.bbeta.template as<cx<double>>(), C,
                           batch, s);
    } else {
        // beta = Q_B^T (V_b^H b)
        IA_HIP(hipMemcpy2DAsync(c, sizeof(cx<double>) * C, b, sizeof(cx<double>) * sb,
                                sizeof(cx<double>) * rows, batch, hipMemcpyDeviceToDevice, s));
        svd_apply_v<T>(sw, true, c, C, 1, s);
        launch_norm2<double>(C, c, C, W.bcn2.template as<double>(), batch, s);
        launch_real_matvec(C, sw.QBt.template as<double>(), 0, c, C, W.bbeta.template as<cx<double>>(), C,
                           batch, s);
    }
    RegDev prm2 = prm;
    if (prm2.select == IA_SELECT_RANK) prm2.k_fixed = std::max(rows, cols);  // rank cutoff dimension
    launch_select(C, rows, sw.sig.template as<double>(), W.bbeta.template as<cx<double>>(),
                  W.bn2.template as<double>(), prm2, lam, idx, kts, W.bphi.template as<cx<double>>(), batch, s, gap,
                  W.bcn2.template as<double>(), svd_floor_rel<T>());
    cx<double> *v = W.bv.template as<cx<double>>();
    if (tall) {
        // x = V_b (Q_B phi)
        launch_real_matvec(C, sw.QBt.template as<double>(), 1, W.bphi.template as<cx<double>>(), C, v, C,
                           batch, s);
        svd_apply_v<T>(sw, false, v, C, 1, s);
        IA_HIP(hipMemcpy2DAsync(x, sizeof(cx<double>) * sx, v, sizeof(cx<double>) * C,
                                sizeof(cx<double>) * C, batch, hipMemcpyDeviceToDevice, s));
    } else {
        // x = U_b [P_B phi; 0]
        launch_real_matvec(C, sw.PBt.template as<double>(), 1, W.bphi.template as<cx<double>>(), C,
                           W.bbeta.template as<cx<double>>(), C, batch, s);
        launch_pad_copy(C, R, W.bbeta.template as<cx<double>>(), C, v, R, batch, s);
        svd_apply_u<T>(sw, Y, v, R, 1, s);
        IA_HIP(hipMemcpy2DAsync(x, sizeof(cx<double>) * sx, v, sizeof(cx<double>) * R,
                                sizeof(cx<double>) * R, batch, hipMemcpyDeviceToDevice, s));
    }
    if (smax)
        IA_HIP(hipMemcpy2DAsync(smax, sizeof(double), sw.sig.p, sizeof(double) * C, sizeof(double),
                                batch, hipMemcpyDeviceToDevice, s));
    IA_HIP(hipGetLastError());
    return IA_OK;
}

// Blocking form (slot-0 workspace): returns when s has finished the solve.
template <typename T>
ia_status regsolve_batched(int rows, int cols, int batch, const cx<double> *A, size_t sA,
                           const cx<double> *b, size_t sb, const RegDev &prm, cx<double> *x,
                           size_t sx, double *lam, int *idx, int *kts, double *smax,
                           hipStream_t s, double *gap = nullptr) {
    IA_TRY(regsolve_enqueue<T>(regsolve_work<T>(0), rows, cols, batch, A, sA, b, sb, prm, x, sx,
                               lam, idx, kts, smax, s, gap));
    IA_HIP(hipStreamSynchronize(s));
    return IA_OK;
}

// Non-blocking form on the slot-0 workspace (callers that capture the solve
// into a graph), and that workspace's buffer addresses (a graph key).
template <typename T>
ia_status regsolve_async(int rows, int cols, int batch, const cx<double> *A, size_t sA,
                         const cx<double> *b, size_t sb, const RegDev &prm, cx<double> *x,
                         size_t sx, double *lam, int *idx, int *kts, double *smax,
                         hipStream_t s) {
    return regsolve_enqueue<T>(regsolve_work<T>(0), rows, cols, batch, A, sA, b, sb, prm, x, sx,
                               lam, idx, kts, smax, s, nullptr);
}
template <typename T>
void regsolve_addresses(std::vector<uintptr_t> &v) {
    regsolve_work<T>(0).addresses(v);
}
template <typename T>
ia_status regsolve_reserve(int rows, int cols, int batch) {
    return regsolve_work<T>(0).reserve(rows, cols, batch);
}
template ia_status regsolve_reserve<double>(int, int, int);
template ia_status regsolve_reserve<float>(int, int, int);
template ia_status regsolve_async<double>(int, int, int, const cx<double> *, size_t,
                                          const cx<double> *, size_t, const RegDev &, cx<double> *,
                                          size_t, double *, int *, int *, double *, hipStream_t);
template ia_status regsolve_async<float>(int, int, int, const cx<double> *, size_t,
                                         const cx<double> *, size_t, const RegDev &, cx<double> *,
                                         size_t, double *, int *, int *, double *, hipStream_t);
template void regsolve_addresses<double>(std::vector<uintptr_t> &);
template void regsolve_addresses<float>(std::vector<uintptr_t> &);

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

// HELS transfer matrices and coefficients of one order group (per slot)
struct HelsSlot {
    DevBuf bH, bM, bQ, bx;
    hipStream_t s = nullptr;
};
// inputs in group order and outputs of one HELS sweep; with IA_GRAPH the
// captured sweep (every group on its slot stream, forked from and joined
// back to the stream gs), the key it was captured for (sizes, parameters,
// the order groups, every buffer address) and its fork / join events
struct HelsIo {
    DevBuf ds, dh, dk, db, lam, idx, kt, sm, W, v, p;
    hipStream_t gs = nullptr;
    hipEvent_t ev[HELS_SLOTS + 1] = {};
    hipGraphExec_t gexec = nullptr;
    std::vector<uintptr_t> gkey;
    void drop_graph() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
        gkey.clear();
    }
};
// plain regularised solve (ia_regsolve): device copies of its arguments
struct RegsolveIo {
    DevBuf A, b, x, lam, idx, kt, sm;
};

HelsSlot &hels_slot(int i) {
    static thread_local HelsSlot *h[HELS_SLOTS] = {};
    if (!h[i]) h[i] = new HelsSlot();
    return *h[i];
}
HelsIo &hels_io() {
    static thread_local HelsIo *h = new HelsIo();
    return *h;
}
RegsolveIo &regsolve_io() {
    static thread_local RegsolveIo *r = new RegsolveIo();
    return *r;
}

// the slot's stream: created once per host thread, default priority,
// non-blocking (ordered against the legacy stream by explicit waits)
ia_status hels_stream(int i, hipStream_t *out) {
    HelsSlot &h = hels_slot(i);
    if (!h.s) IA_HIP(hipStreamCreateWithFlags(&h.s, hipStreamNonBlocking));
    *out = h.s;
    return IA_OK;
}

void release_solve_cache() {
    for (int i = 0; i <= HELS_SLOTS; ++i) {
        regsolve_work<double>(i).release();
        regsolve_work<float>(i).release();
    }
    for (int i = 0; i < HELS_SLOTS; ++i)
        for (DevBuf *b : {&hels_slot(i).bH, &hels_slot(i).bM, &hels_slot(i).bQ, &hels_slot(i).bx})
            b->reset();
    HelsIo &h = hels_io();
    h.drop_graph();
    for (DevBuf *b : {&h.ds, &h.dh, &h.dk, &h.db, &h.lam, &h.idx, &h.kt, &h.sm, &h.W, &h.v, &h.p})
        b->reset();
    RegsolveIo &r = regsolve_io();
    for (DevBuf *b : {&r.A, &r.b, &r.x, &r.lam, &r.idx, &r.kt, &r.sm}) b->reset();
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
    const bool mixed = desc->precision == IA_PREC_MIXED;
    ia_sweep_out o{};
    if (out) o = *out;

    // Frequencies of equal order form one batched system size.  The groups
    // run side by side, each on its own stream and workspace (slot), with a
    // single wait at the end; frequencies are laid out in group order.
    std::map<int, std::vector<int>> groups;
    for (int f = 0; f < F; ++f) groups[order[f]].push_back(f);
    std::vector<int> gN, goff, gnb, perm;
    for (auto &g : groups) {
        gN.push_back(g.first);
        goff.push_back((int)perm.size());
        gnb.push_back((int)g.second.size());
        perm.insert(perm.end(), g.second.begin(), g.second.end());
    }
    const int G = (int)gN.size(), NS = std::min(G, HELS_SLOTS);
    std::vector<double> kp(F);
    std::vector<ia_z> bp((size_t)F * l);
    for (int i = 0; i < F; ++i) {
        kp[i] = k[perm[i]];
        memcpy(&bp[(size_t)i * l], b + (size_t)perm[i] * l, sizeof(ia_z) * l);
    }
    HelsIo &io = hels_io();
    IA_TRY(to_device(io.ds, azelr_s, sizeof(double) * 3 * m, nullptr));
    IA_TRY(to_device(io.dh, azelr_h, sizeof(double) * 3 * l, nullptr));
    IA_TRY(to_device(io.dk, kp.data(), sizeof(double) * F, nullptr));
    IA_TRY(to_device(io.db, bp.data(), sizeof(ia_z) * (size_t)F * l, nullptr));
    IA_TRY(io.lam.alloc(sizeof(double) * F));
    IA_TRY(io.idx.alloc(sizeof(int) * F));
    IA_TRY(io.kt.alloc(sizeof(int) * F));
    IA_TRY(io.sm.alloc(sizeof(double) * F));
    IA_TRY(io.W.alloc(sizeof(double) * F));
    IA_TRY(io.v.alloc(sizeof(cx<double>) * (size_t)F * m));
    IA_TRY(io.p.alloc(sizeof(cx<double>) * (size_t)F * m));
    // every buffer sized before the first group is enqueued, so that no
    // slot reallocates under work in flight (groups beyond HELS_SLOTS share
    // a slot and follow each other on its stream)
    for (int g = 0; g < G; ++g) {
        HelsSlot &S = hels_slot(g % NS);
        const int n_t = (gN[g] + 1) * (gN[g] + 1), nb = gnb[g];
        IA_TRY(S.bH.alloc(sizeof(cx<double>) * (size_t)l * n_t * nb));
        IA_TRY(S.bM.alloc(sizeof(cx<double>) * (size_t)m * n_t * nb));
        IA_TRY(S.bQ.alloc(sizeof(cx<double>) * (size_t)m * n_t * nb));
        IA_TRY(S.bx.alloc(sizeof(cx<double>) * (size_t)n_t * nb));
        if (mixed)
            IA_TRY(regsolve_work<float>(1 + g % NS).reserve(l, n_t, nb));
        else
            IA_TRY(regsolve_work<double>(1 + g % NS).reserve(l, n_t, nb));
    }
    std::vector<hipStream_t> st(NS);
    for (int i = 0; i < NS; ++i) IA_TRY(hels_stream(i, &st[i]));
    // the inputs were copied on the legacy stream
    IA_HIP(hipStreamSynchronize(nullptr));
    // every group's work, each on its slot's stream
    auto enqueue = [&]() -> ia_status {
        for (int g = 0; g < G; ++g) {
            const int sl = g % NS, N = gN[g], nb = gnb[g], off = goff[g], n_t = (N + 1) * (N + 1);
            HelsSlot &S = hels_slot(sl);
            hipStream_t sg = st[sl];
            const double *kg = io.dk.as<double>() + off;
            const cx<double> *bg = io.db.as<cx<double>>() + (size_t)off * l;
            cx<double> *H = S.bH.as<cx<double>>(), *Me = S.bM.as<cx<double>>(),
                       *Qe = S.bQ.as<cx<double>>(), *x = S.bx.as<cx<double>>();
            launch_hels<double>(l, N, nb, kg, rho_c, io.dh.as<double>(), H, nullptr,
                                (size_t)l * n_t, sg);
            launch_hels<double>(m, N, nb, kg, rho_c, io.ds.as<double>(), Me, Qe, (size_t)m * n_t,
                                sg);
            if (mixed)
                IA_TRY(regsolve_enqueue<float>(regsolve_work<float>(1 + sl), l, n_t, nb, H,
                                               (size_t)l * n_t, bg, l, prm, x, n_t,
                                               io.lam.as<double>() + off, io.idx.as<int>() + off,
                                               io.kt.as<int>() + off, io.sm.as<double>() + off, sg,
                                               nullptr));
            else
                IA_TRY(regsolve_enqueue<double>(regsolve_work<double>(1 + sl), l, n_t, nb, H,
                                                (size_t)l * n_t, bg, l, prm, x, n_t,
                                                io.lam.as<double>() + off, io.idx.as<int>() + off,
                                                io.kt.as<int>() + off, io.sm.as<double>() + off,
                                                sg, nullptr));
            // v = Q_e c, p = M_e c (to_velocity / to_pressure), W = 1/2 Re(v^H p) S/m
            const cx<double> one{1, 0}, zero{0, 0};
            cx<double> *v = io.v.as<cx<double>>() + (size_t)off * m;
            cx<double> *pr = io.p.as<cx<double>>() + (size_t)off * m;
            launch_gemm<double>('N', 'N', m, 1, n_t, one, Qe, m, (size_t)m * n_t, x, n_t, n_t,
                                zero, v, m, m, nb, sg);
            launch_gemm<double>('N', 'N', m, 1, n_t, one, Me, m, (size_t)m * n_t, x, n_t, n_t,
                                zero, pr, m, m, nb, sg);
            launch_power<double>(m, v, m, pr, m, desc->area / m, io.W.as<double>() + off, nb, sg);
            if (hipGetLastError() != hipSuccess)
                return fail(IA_ERR_DEVICE, "ia_hels_sweep: launch failed");
        }
        return IA_OK;
    };
    ia_status rc = IA_OK;
    if (graph_mode()) {
        // IA_GRAPH=1: the sweep (the fork to the slot streams, every group,
        // the join) is captured once per key into a graph and replayed with
        // one launch; ~2,000 small launches per sweep become graph nodes
        std::vector<uintptr_t> key = {
            (uintptr_t)F, (uintptr_t)m, (uintptr_t)l, (uintptr_t)mixed, (uintptr_t)prm.filter,
            (uintptr_t)prm.select, (uintptr_t)prm.n_lambda, (uintptr_t)prm.k_fixed,
            dbits(prm.decades), dbits(prm.lambda_fixed), dbits(desc->area), dbits(rho_c),
            (uintptr_t)G, (uintptr_t)NS};
        for (int g = 0; g < G; ++g) {
            key.push_back((uintptr_t)gN[g]);
            key.push_back((uintptr_t)gnb[g]);
        }
        for (const DevBuf *q :
             {&io.ds, &io.dh, &io.dk, &io.db, &io.lam, &io.idx, &io.kt, &io.sm, &io.W, &io.v, &io.p})
            key.push_back((uintptr_t)q->p);
        for (int i = 0; i < NS; ++i) {
            const HelsSlot &S = hels_slot(i);
            for (const DevBuf *q : {&S.bH, &S.bM, &S.bQ, &S.bx}) key.push_back((uintptr_t)q->p);
            if (mixed)
                regsolve_work<float>(1 + i).addresses(key);
            else
                regsolve_work<double>(1 + i).addresses(key);
        }
        if (!io.gs) {
            IA_HIP(hipStreamCreateWithFlags(&io.gs, hipStreamNonBlocking));
            for (hipEvent_t &e : io.ev) IA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        if (!io.gexec || io.gkey != key) {
            io.drop_graph();
            IA_HIP(hipStreamBeginCapture(io.gs, hipStreamCaptureModeThreadLocal));
            hipError_t e = hipEventRecord(io.ev[0], io.gs);
            for (int i = 0; i < NS && e == hipSuccess; ++i) e = hipStreamWaitEvent(st[i], io.ev[0], 0);
            rc = e == hipSuccess ? enqueue()
                                 : fail(IA_ERR_DEVICE, "ia_hels_sweep: fork of the capture failed");
            // every slot stream joins back (also after a failure), so that the
            // capture ends
            for (int i = 0; i < NS; ++i) {
                (void)hipEventRecord(io.ev[1 + i], st[i]);
                (void)hipStreamWaitEvent(io.gs, io.ev[1 + i], 0);
            }
            hipGraph_t gr = nullptr;
            const hipError_t ce = hipStreamEndCapture(io.gs, &gr);
            if (rc == IA_OK && ce == hipSuccess && gr &&
                hipGraphInstantiate(&io.gexec, gr, nullptr, nullptr, 0) != hipSuccess)
                io.gexec = nullptr;
            if (gr) (void)hipGraphDestroy(gr);
            if (rc == IA_OK && !io.gexec) {
                (void)hipGetLastError();
                rc = fail(IA_ERR_DEVICE, "ia_hels_sweep: capture into a graph failed");
            }
            if (rc == IA_OK) io.gkey = key;
        }
        if (rc == IA_OK && hipGraphLaunch(io.gexec, io.gs) != hipSuccess)
            rc = fail(IA_ERR_DEVICE, "ia_hels_sweep: %s", hipGetErrorString(hipGetLastError()));
        if (hipStreamSynchronize(io.gs) != hipSuccess && rc == IA_OK)
            rc = fail(IA_ERR_DEVICE, "ia_hels_sweep: %s", hipGetErrorString(hipGetLastError()));
    } else {
        rc = enqueue();
        // wait for every group, also after a failed enqueue (the workspaces
        // are reused by the next call)
        for (int i = 0; i < NS; ++i)
            if (hipStreamSynchronize(st[i]) != hipSuccess && rc == IA_OK)
                rc = fail(IA_ERR_DEVICE, "ia_hels_sweep: %s",
                          hipGetErrorString(hipGetLastError()));
    }
    IA_TRY(rc);
    std::vector<double> hl(F), hs(F), hW(F);
    std::vector<int> hi(F), hk(F);
    std::vector<ia_z> hv((size_t)m * F), hp((size_t)m * F);
    IA_HIP(hipMemcpy(hl.data(), io.lam.p, sizeof(double) * F, hipMemcpyDeviceToHost));
    IA_HIP(hipMemcpy(hs.data(), io.sm.p, sizeof(double) * F, hipMemcpyDeviceToHost));
    IA_HIP(hipMemcpy(hW.data(), io.W.p, sizeof(double) * F, hipMemcpyDeviceToHost));
    IA_HIP(hipMemcpy(hi.data(), io.idx.p, sizeof(int) * F, hipMemcpyDeviceToHost));
    IA_HIP(hipMemcpy(hk.data(), io.kt.p, sizeof(int) * F, hipMemcpyDeviceToHost));
    if (o.velocity) IA_HIP(hipMemcpy(hv.data(), io.v.p, sizeof(ia_z) * (size_t)m * F, hipMemcpyDeviceToHost));
    if (o.pressure) IA_HIP(hipMemcpy(hp.data(), io.p.p, sizeof(ia_z) * (size_t)m * F, hipMemcpyDeviceToHost));
    for (int i = 0; i < F; ++i) {
        const int f = perm[i];
        if (o.lambda) o.lambda[f] = hl[i];
        if (o.lambda_index) o.lambda_index[f] = hi[i];
        if (o.k_tsvd) o.k_tsvd[f] = hk[i];
        if (o.power) o.power[f] = hW[i];
        if (o.sigma_max) o.sigma_max[f] = hs[i];
        if (o.velocity) memcpy(o.velocity + (size_t)f * m, &hv[(size_t)i * m], sizeof(ia_z) * m);
        if (o.pressure) memcpy(o.pressure + (size_t)f * m, &hp[(size_t)i * m], sizeof(ia_z) * m);
        if (o.status) o.status[f] = IA_OK;
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
    RegsolveIo &io = regsolve_io();
    DevBuf &dA = io.A, &db = io.b, &dx = io.x, &dl = io.lam, &di = io.idx, &dk = io.kt,
           &dsm = io.sm;
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
