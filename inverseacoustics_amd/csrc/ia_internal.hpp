// Internal declarations shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ia.h"

namespace ia {

// ------------------------------------------------------------ complex
template <typename T>
struct alignas(2 * sizeof(T)) cx {
    T re, im;
};
using z64 = cx<double>;
using c32 = cx<float>;

template <typename T>
__host__ __device__ __forceinline__ cx<T> mk(T r, T i) { return cx<T>{r, i}; }
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator-(cx<T> a) { return {-a.re, -a.im}; }
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator*(T s, cx<T> a) { return {s * a.re, s * a.im}; }
template <typename T>
__host__ __device__ __forceinline__ cx<T> conj(cx<T> a) { return {a.re, -a.im}; }
template <typename T>
__host__ __device__ __forceinline__ T abs2(cx<T> a) { return a.re * a.re + a.im * a.im; }
// conj(a) * b
template <typename T>
__host__ __device__ __forceinline__ cx<T> cmulc(cx<T> a, cx<T> b) {
    return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> cdiv(cx<T> a, cx<T> b) {
    // Smith's algorithm
    if (fabs(b.re) >= fabs(b.im)) {
        T r = b.im / b.re, d = b.re + b.im * r;
        return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
    } else {
        T r = b.re / b.im, d = b.im + b.re * r;
        return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
    }
}
template <typename To, typename Ti>
__host__ __device__ __forceinline__ cx<To> cvt(cx<Ti> a) { return {(To)a.re, (To)a.im}; }

// ------------------------------------------------------------- errors
void set_error(const char *fmt, ...);
ia_status fail(ia_status st, const char *fmt, ...);

#define IA_HIP(call)                                                              \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess)                                                     \
            return ::ia::fail(IA_ERR_DEVICE, "%s:%d %s: %s", __FILE__, __LINE__,  \
                              #call, hipGetErrorString(e_));                      \
    } while (0)

#define IA_TRY(expr)                         \
    do {                                     \
        ia_status s_ = (expr);               \
        if (s_ != IA_OK) return s_;          \
    } while (0)

// ------------------------------------------------------ device buffer
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { reset(); }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    ia_status alloc(size_t n) {
        if (n <= bytes && p) return IA_OK;
        reset();
        if (n == 0) return IA_OK;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            return fail(IA_ERR_NOMEM, "hipMalloc(%zu): %s", n, hipGetErrorString(e));
        }
        bytes = n;
        return IA_OK;
    }
    template <typename T>
    T *as() const { return static_cast<T *>(p); }
};

// air constants of this thread
struct Air {
    double c = 343.5;   // CUDA/constants.cuh:2
    double rho = 1.19;  // CUDA/constants.cuh:5
};
Air &air();

bool is_device_ptr(const void *p);
// api.hip: IA_GRAPH set (and not "0"): the sweeps are captured into HIP
// graphs and replayed; the bits of a double as a graph-key entry
bool graph_mode();
uintptr_t dbits(double v);
// Copies n bytes of `src` (host or device) into a fresh device buffer.
ia_status to_device(DevBuf &dst, const void *src, size_t n, hipStream_t s);

// --------------------------------------------------------- launchers
// All launchers are asynchronous on `s`; matrices are column-major,
// batched with element strides.  T is double or float.

// ESM assembly (esm.hip)
// herm: store D^H instead of D (same column-major layout)
template <typename T>
void launch_esm_D(int m, int nf, const double *k, const double *r_s, const double *src,
                  const double *nrm, cx<T> *D, size_t strideD, hipStream_t s, bool herm = false);
template <typename T>
void launch_esm_M(int rows, int m, int nf, const double *k, double rho_c, const double *r_rows,
                  const double *src, cx<T> *M, size_t strideM, bool conj_transpose, hipStream_t s,
                  const int *rowperm = nullptr);
template <typename T>
void launch_esm_msq_apply(int m, int nf, const double *k, double rho_c, const double *r_s,
                          const double *src, const cx<T> *q, size_t strideq, cx<T> *p,
                          size_t stridep, hipStream_t s);

// complex GEMM (gemm.hip): C = alpha op(A) op(B) + beta C, op in {'N','C'}
template <typename T>
void launch_gemm(char opA, char opB, int M, int N, int K, cx<T> alpha, const cx<T> *A, int lda,
                 size_t strideA, const cx<T> *B, int ldb, size_t strideB, cx<T> beta, cx<T> *C,
                 int ldc, size_t strideC, int batch, hipStream_t s);
// real FP64 GEMM (gemm.hip): C = op(A) B, op in {'N','T'}, column-major
void launch_dgemm(char opA, int M, int N, int K, const double *A, int lda, size_t strideA,
                  const double *B, int ldb, size_t strideB, double *C, int ldc, size_t strideC,
                  int batch, hipStream_t s);

// LU (lu.hip)
// work (optional): getrf_work_ints(n, batch) ints of device scratch; without
// it the scratch is taken with hipMallocAsync on s
size_t getrf_work_ints(int n, int batch);
template <typename T>
ia_status getrf_batched(int n, cx<T> *A, int lda, size_t strideA, int *ipiv, int *info, int batch,
                        hipStream_t s, int *work = nullptr);
// Solve op(A) X = B with A = P^T L U from getrf; op 'N' or 'C'.  B n x nrhs.
// x <- A^-1 b for one right-hand side per matrix (A = P^T L U from
// getrf_batched, op 'N'), for a few matrices: many workgroups per matrix,
// one launch per 64-row block; S: n entries per matrix of workspace (stride
// strideB).
template <typename T>
ia_status getrs1_spread(int n, const cx<T> *A, int lda, size_t strideA, const int *ipiv,
                        cx<T> *B, size_t strideB, cx<T> *S, int batch, hipStream_t s);
// work (optional): getrs_work_bytes<T>(n, nrhs, batch) bytes of device
// scratch for nrhs > 1; without it the scratch is taken with hipMallocAsync
template <typename T>
size_t getrs_work_bytes(int n, int nrhs, int batch);
template <typename T>
ia_status getrs_batched(char op, int n, int nrhs, const cx<T> *A, int lda, size_t strideA,
                        const int *ipiv, cx<T> *B, int ldb, size_t strideB, int batch,
                        hipStream_t s, void *work = nullptr);

// getrf's interchanges as a row index per matrix: (P B)[r] = B[idx[r]],
// inv the inverse permutation (n ints each per matrix)
ia_status lu_perm_index(int n, const int *ipiv, int *idx, int *inv, int batch, hipStream_t s);
// A X = B for B given already in pivoted row order (B = P B0): the two
// triangular solves of getrs op 'N', in place
template <typename T>
ia_status getrs_prepermuted(int n, int nrhs, const cx<T> *A, int lda, size_t strideA, cx<T> *B,
                            int ldb, size_t strideB, int batch, hipStream_t s);
// A^H V = B with A = P^T L U, leaving V = P X (X the solution) in B: the two
// triangular solves of getrs op 'C' without the final row permutation
template <typename T>
ia_status getrs_c_permuted(int n, int nrhs, const cx<T> *A, int lda, size_t strideA, cx<T> *B,
                           int ldb, size_t strideB, int batch, hipStream_t s);
// dst[r, c] = src[perm[r], c] per matrix (perm: n ints per matrix)
template <typename T>
void launch_row_gather(int n, int nrhs, const int *perm, const cx<T> *src, int lds, size_t sS,
                       cx<T> *dst, int ldd, size_t sD, int batch, hipStream_t s);
// upper-triangular solve A X = B with the upper triangle of A (non-unit)
template <typename T>
ia_status trsm_upper_batched(int n, int nrhs, const cx<T> *A, int lda, size_t strideA, cx<T> *B,
                             int ldb, size_t strideB, int batch, hipStream_t s);
// Householder QR in place (reflectors below the diagonal, R above)
template <typename T>
ia_status launch_geqr2(int R, int C, cx<T> *A, int lda, size_t strideA, cx<T> *tau, int batch,
                       hipStream_t s);

// SVD building blocks (svd.hip, gebrd.hip)
// panel-blocked bidiagonalisation Y = U_b B V_b^H (R >= C); work holds
// gebrd_work_elems(R, C) complex T per batch item
size_t gebrd_lds_bytes(int R, int C, size_t elem);
size_t gebrd_work_elems(int R, int C);
// blocked Householder QR Y = Q R in place (geqrf.hip); work holds
// geqrf_work_elems(R, C) complex T per batch item
size_t geqrf_work_elems(int R, int C);
template <typename T>
ia_status launch_geqrf(int R, int C, cx<T> *Y, int ldy, size_t sY, cx<T> *tau, cx<T> *work,
                       int batch, hipStream_t s);
// Q (R x C, ld ldq) = H_0 ... H_{C-1} [I; 0] from launch_geqrf's reflectors
// in Y and tau (blocked, GEMMs); work as for launch_geqrf
template <typename T>
ia_status launch_orgqr(int R, int C, const cx<T> *Y, int ldy, size_t sY, const cx<T> *tau,
                       cx<T> *Q, int ldq, size_t sQ, cx<T> *work, int batch, hipStream_t s);
template <typename T>
void launch_copy_upper(int C, const cx<T> *Y, int ldy, size_t sY, cx<T> *Rb, int batch,
                       hipStream_t s);
// two-stage bidiagonalisation: stage one (geqrf.hip) reduces Z (R x C) to an
// upper band of width 32 (left reflectors below Z's diagonal, tau in tauQ;
// right reflectors as the columns of Pm, (C-32) x (C-32) per matrix, tau in
// tauP); stage two (bulge.hip) chases the band to the real bidiagonal d, e,
// keeping its reflectors (uR/tR, uL/tL: bulge_refl_elems(C) slots of 32).
constexpr int BAND_NB = 32;
size_t band_work_elems(int R, int C);
template <typename T>
ia_status launch_gebrd_band(int R, int C, cx<T> *Z, int ldz, size_t sZ, cx<T> *tauQ, cx<T> *Pm,
                            cx<T> *tauP, cx<T> *work, int batch, hipStream_t s);
int bulge_tmax(int C);
size_t bulge_band_elems(int C);
size_t bulge_band_pad_elems();
size_t bulge_refl_elems(int C);
size_t bulge_apply_lds(int C);
// bytes the chase's tasks read and write (each task's two blocks once)
double bulge_alg_bytes(int C, size_t elem);
struct KernelClock;
// clk (optional) brackets the chase kernel with HIP events and adds its
// bulge_alg_bytes
template <typename T>
ia_status launch_bulge(int C, const cx<T> *Z, int ldz, size_t sZ, cx<T> *Bs, cx<T> *uR,
                       cx<T> *uL, cx<T> *tR, cx<T> *tL, double *d, double *e, int batch,
                       hipStream_t s, KernelClock *clk = nullptr);
// v <- H v over one side's stage-two reflectors: fwd = true applies
// P2^H (uR) or Q2^H (uL), fwd = false applies P2 or Q2
template <typename T>
void launch_bulge_apply(bool fwd, int C, const cx<T> *u, const cx<T> *tau, cx<double> *v,
                        size_t sv, int nvec, int batch, hipStream_t s);

// free the device workspace the gradient and radiation-mode entry points
// keep per host thread
void release_grad_cache();
void release_rad_cache();
// ... the regularised-solve / HELS workspaces (solve.hip) and the IBEM
// sweep's (ibem_api.hip)
void release_solve_cache();
void release_ibem_cache();
// ... and the certified recheck's (recheck.hip)
void release_recheck_cache();

// Certified selection of the complex64 ESM sweep (recheck.hip): for
// Tikhonov + grid GCV with l <= min(m, 4096), the nearly tied frequencies L
// (sweep indices; curve: their complex64 GCV values, n_lambda each, host)
// decided in FP64 from the Gram form of the GCV function, their outputs
// written into o; the frequencies it cannot decide are appended to full.
bool gcv_recheck_applies(const ia_sweep_desc &D);
ia_status gcv_recheck(const ia_sweep_desc &D, const std::vector<int> &L,
                      const std::vector<double> &curve, double tau, const double *k,
                      const double *r_s, const double *src, const double *nrm, const double *r_h,
                      const ia_z *b, const ia_sweep_out &o, std::vector<int> &full,
                      hipStream_t s);

// ADMM (admm.hip): A (rows x cols, batch stride rows*cols), b, c on the device
ia_status admm_batched(int rows, int cols, int batch, const cx<double> *A, const cx<double> *b,
                       const ia_admm_params &P, cx<double> *c, int *iters, hipStream_t s);

// HIP events around every launch of one kernel on its stream, with the
// algorithmic bytes those launches move (read after the stream is synced)
struct KernelClock {
    std::vector<hipEvent_t> ev;
    double bytes = 0;
    int launches = 0;
    KernelClock() = default;
    KernelClock(const KernelClock &) = delete;
    KernelClock &operator=(const KernelClock &) = delete;
    ~KernelClock() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
    void mark(hipStream_t s) {
        hipEvent_t e;
        if (hipEventCreate(&e) == hipSuccess) {
            (void)hipEventRecord(e, s);
            ev.push_back(e);
        } else {
            (void)hipGetLastError();
        }
    }
    double ms() const {
        double t = 0;
        for (size_t i = 0; i + 1 < ev.size(); i += 2) {
            float x = 0;
            if (hipEventElapsedTime(&x, ev[i], ev[i + 1]) == hipSuccess) t += x;
        }
        return t;
    }
};
template <typename T>
ia_status launch_gebrd(int R, int C, cx<T> *Y, int ldy, size_t sY, double *d, double *e,
                       cx<T> *tauL, cx<T> *tauR, cx<T> *Vr, cx<T> *work, int batch,
                       hipStream_t s, KernelClock *clk = nullptr);
// f32: the bidiagonal comes from the complex64 reduction (coarser stop)
void launch_bisect(int n, const double *d, const double *e, double *sigma, int batch,
                   hipStream_t s, bool f32 = false);
// f32: the scratch holds the factorization and iterate in float (the
// complex64 form), else in double
size_t inviter_work_bytes(int n, int sub, bool f32 = false);
// QBt/PBt: n x n per batch item, vector i stored as row i (QBt[i + k n]).
void launch_inviter(int n, const double *d, const double *e, const double *sigma, double *QBt,
                    double *PBt, void *work, int sub, int batch, hipStream_t s, bool f32 = false);
// When the full singular bases are formed (csvd): the q and p vectors of
// every run of values closer than 1e-4 s_1 orthonormalised each on its own
void launch_orthonormalise_tail(int n, const double *sigma, double *QBt, double *PBt, int batch,
                                hipStream_t s);
// v <- U_b v (herm=false) or U_b^H v, length-R vectors, nvec per batch item
template <typename T>
void launch_apply_left(bool herm, int R, int C, const cx<T> *Y, int ldy, size_t strideY,
                       const cx<T> *tauL, cx<double> *v, size_t stridev, int nvec, int batch,
                       hipStream_t s);
// v <- V_b v or V_b^H v, length-C vectors (Vr: C x C per batch item)
template <typename T>
void launch_apply_right(bool herm, int C, const cx<T> *Vr, const cx<T> *tauR, cx<double> *v,
                        size_t stridev, int nvec, int batch, hipStream_t s);

// regularisation (reg.hip)
struct RegDev {
    int filter, select, n_lambda, k_fixed;
    double decades, lambda_fixed;
};
// Given sigma (p), c = Q_B^T V_b^H b (p, complex), |b|^2, rows: selects the
// parameter and writes phi (p, complex) such that x = V phi.
// gap_out (optional): relative GCV / L-curve gap between the chosen and the
// runner-up parameter (discrete-GCV k for TSVD), INFINITY when none.
// Singular values below floor_rel * s_1 are pooled: their coordinates are
// replaced by one component of energy |c|^2 - sum(resolved |beta|^2) at s = 0
// (cnorm2 = |c|^2 of the vector with beta = Q^T c; NULL: sum |beta|^2).
// floor_rel: 1e-12 for an FP64 SVD, 1e-6 for a complex64 one.
// curve_out (optional, grid GCV / L-curve only): the n_lambda criterion
// values of each batch item.
void launch_select(int p, int rows, const double *sigma, const cx<double> *beta,
                   const double *bnorm2, RegDev prm, double *lam_out, int *idx_out, int *k_out,
                   cx<double> *phi, int batch, hipStream_t s, double *gap_out = nullptr,
                   const double *cnorm2 = nullptr, double floor_rel = 1e-12,
                   double *curve_out = nullptr);
template <typename T>
constexpr double svd_floor_rel() { return sizeof(T) == sizeof(double) ? 1e-12 : 1e-6; }

// small helpers (util.hip)
void launch_real_matvec(int n, const double *Q, int transpose, const cx<double> *x,
                        size_t stridex, cx<double> *y, size_t stridey, int batch, hipStream_t s);
template <typename T>
void launch_norm2(int n, const cx<T> *x, size_t stride, double *out, int batch, hipStream_t s);
template <typename Td, typename Ts>
void launch_convert(size_t n, const cx<Ts> *src, cx<Td> *dst, hipStream_t s);
template <typename T>
void launch_power(int m, const cx<T> *v, size_t stridev, const cx<T> *p, size_t stridep,
                  double scale, double *W, int batch, hipStream_t s);

}  // namespace ia

namespace ia {
// util.hip additions
// dst (R x C column-major, per batch) column i = row i of Pt (C x C, Pt[i + k C]),
// rows C..R-1 zero.
void launch_rows_to_cols(int R, int C, const double *Pt, cx<double> *dst, int batch, hipStream_t s);
// radiation.hip: y = (x, 0); y = re(x); y.im = 0 (n entries); column j of U
// (np x n per frequency) rotated by the conjugate phase of its largest entry
// and made real; S_f = sin(k_f r_ij)/(k_f r_ij) for points p (np x 3,
// column-major), as complex, np x np per frequency
void launch_real_to_cx(size_t n, const double *x, cx<double> *y, hipStream_t s);
void launch_cx_to_real(size_t n, const cx<double> *x, double *y, hipStream_t s);
void launch_keep_real(size_t n, cx<double> *y, hipStream_t s);
void launch_realign(int np, int n, int nf, cx<double> *U, hipStream_t s);
void launch_rad_R(int np, const double *kv, const double *p, cx<double> *S, int nf, hipStream_t s);
// dst (cols x rows) = src^H (rows x cols), batched.
template <typename T>
void launch_conj_transpose(int rows, int cols, const cx<T> *src, size_t ss, cx<T> *dst, size_t sd,
                           int batch, hipStream_t s);
// y[b][0..n) = x[b][0..n), y[b][n..ny) = 0 (cx<double>)
void launch_pad_copy(int n, int ny, const cx<double> *x, size_t sx, cx<double> *y, size_t sy,
                     int batch, hipStream_t s);
}  // namespace ia

#include <algorithm>
#include <cmath>
#include <cstdlib>

namespace ia {
// ------------------------------------------------ shared host helpers

inline bool finite_all(const double *p, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

inline ia_status check_device() {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return fail(IA_ERR_DEVICE, "no HIP device available (this library has no CPU fallback)");
    }
    return IA_OK;
}

template <typename T>
inline ia_status host_out(T **dst, const void *dsrc, size_t count) {
    if (!dst) return IA_OK;
    *dst = (T *)malloc(sizeof(T) * (count ? count : 1));
    if (!*dst) return fail(IA_ERR_NOMEM, "malloc(%zu)", sizeof(T) * count);
    if (count) IA_HIP(hipMemcpy(*dst, dsrc, sizeof(T) * count, hipMemcpyDeviceToHost));
    return IA_OK;
}

// ------------------------------------------------------ SVD driver
// Factor Y (R x C, R >= C, ld R, batch stride R*C, complex T) in place:
// bidiagonal reflectors in Y/Vr/tau, singular values sig (descending),
// bidiagonal singular vectors QBt/PBt (vector i in row i).  When R >= 3C/2
// Y is first reduced to R = Q^H Y (blocked QR on the matrix cores) and the
// C x C factor is bidiagonalised instead: the left singular basis is then
// U = Q diag(U_R, I), applied by svd_apply_u / svd_apply_uh.
// Two-stage bidiagonalisation for C in [128, 4096] unless IA_ONE_STAGE is
// set (the one-stage panel kernel then handles every size).
inline bool two_stage_for(int C) {
    const char *e = std::getenv("IA_ONE_STAGE");
    if (e && *e && *e != '0') return false;
    return C >= 128 && C <= 4096;
}

template <typename T>
struct SvdWork {
    DevBuf d, e, sig, tauL, tauR, Vr, QBt, PBt, work, gw, Rb, tauQ, qw;
    DevBuf Bs, uR2, uL2, tR2, tL2, Pm, tauP;   // two-stage form
    int R = 0, C = 0, batch = 0, sub = 1;
    bool qr = false, ts = false;
    void release() {
        for (DevBuf *b : {&d, &e, &sig, &tauL, &tauR, &Vr, &QBt, &PBt, &work, &gw, &Rb, &tauQ, &qw,
                          &Bs, &uR2, &uL2, &tR2, &tL2, &Pm, &tauP})
            b->reset();
    }
    // the address of every buffer (a reallocation invalidates a captured graph)
    void addresses(std::vector<uintptr_t> &v) const {
        for (const DevBuf *b : {&d, &e, &sig, &tauL, &tauR, &Vr, &QBt, &PBt, &work, &gw, &Rb,
                                &tauQ, &qw, &Bs, &uR2, &uL2, &tR2, &tL2, &Pm, &tauP})
            v.push_back((uintptr_t)b->p);
    }
    ia_status alloc(int R_, int C_, int batch_) {
        R = R_;
        C = C_;
        batch = batch_;
        qr = 2 * (size_t)R >= 3 * (size_t)C && C >= 64;
        ts = two_stage_for(C);
        size_t bc = (size_t)batch * C;
        IA_TRY(d.alloc(bc * sizeof(double)));
        IA_TRY(e.alloc(bc * sizeof(double)));
        IA_TRY(sig.alloc(bc * sizeof(double)));
        IA_TRY(tauL.alloc(bc * sizeof(cx<T>)));
        IA_TRY(QBt.alloc(bc * C * sizeof(double)));
        IA_TRY(PBt.alloc(bc * C * sizeof(double)));
        if (ts) {
            const size_t ns = bulge_refl_elems(C), cp = (size_t)(C - BAND_NB);
            IA_TRY(Bs.alloc((bulge_band_elems(C) * batch + bulge_band_pad_elems()) *
                            sizeof(cx<T>)));
            IA_TRY(uR2.alloc(ns * BAND_NB * batch * sizeof(cx<T>)));
            IA_TRY(uL2.alloc(ns * BAND_NB * batch * sizeof(cx<T>)));
            IA_TRY(tR2.alloc(ns * batch * sizeof(cx<T>)));
            IA_TRY(tL2.alloc(ns * batch * sizeof(cx<T>)));
            IA_TRY(Pm.alloc(cp * cp * batch * sizeof(cx<T>)));
            IA_TRY(tauP.alloc(cp * batch * sizeof(cx<T>)));
            IA_TRY(gw.alloc(band_work_elems(qr ? C : R, C) * batch * sizeof(cx<T>)));
        } else {
            IA_TRY(tauR.alloc(bc * sizeof(cx<T>)));
            IA_TRY(Vr.alloc(bc * C * sizeof(cx<T>)));
            IA_TRY(gw.alloc(gebrd_work_elems(qr ? C : R, C) * batch * sizeof(cx<T>)));
        }
        if (qr) {
            IA_TRY(Rb.alloc(bc * C * sizeof(cx<T>)));
            IA_TRY(tauQ.alloc(bc * sizeof(cx<T>)));
            IA_TRY(qw.alloc(geqrf_work_elems(R, C) * batch * sizeof(cx<T>)));
        }
        // inverse-iteration scratch: as many frequencies per launch as a
        // quarter of the free device memory holds (1..32 GiB), so one launch
        // covers a whole 256-frequency batch (one thread per singular value
        // needs >= 1024 * 256 threads to fill the device)
        constexpr bool f32 = sizeof(T) == sizeof(float);
        size_t per = inviter_work_bytes(C, 1, f32);
        if (work.bytes >= inviter_work_bytes(C, batch, f32)) {
            sub = batch;  // a kept buffer already holds the whole batch
        } else {
            size_t fr = 0, tot = 0;
            if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
                (void)hipGetLastError();
                fr = (size_t)4 << 30;
            }
            const size_t cap = std::min<size_t>((size_t)32 << 30,
                                                std::max<size_t>((size_t)1 << 30, fr / 4));
            sub = (int)std::max<size_t>(1, std::min<size_t>((size_t)batch, cap / per));
        }
        IA_TRY(work.alloc(inviter_work_bytes(C, sub, f32)));
        return IA_OK;
    }
};

// QR (when chosen) + bidiagonalisation of Y; clk times the panel kernel
template <typename T>
inline ia_status svd_reduce(SvdWork<T> &w, cx<T> *Y, hipStream_t s, KernelClock *clk = nullptr) {
    const int R = w.R, C = w.C, batch = w.batch;
    cx<T> *Z = Y;
    int Rz = R;
    if (w.qr) {
        IA_TRY(launch_geqrf<T>(R, C, Y, R, (size_t)R * C, w.tauQ.template as<cx<T>>(),
                               w.qw.template as<cx<T>>(), batch, s));
        launch_copy_upper<T>(C, Y, R, (size_t)R * C, w.Rb.template as<cx<T>>(), batch, s);
        Z = w.Rb.template as<cx<T>>();
        Rz = C;
    }
    if (w.ts) {
        IA_TRY(launch_gebrd_band<T>(Rz, C, Z, Rz, (size_t)Rz * C, w.tauL.template as<cx<T>>(),
                                    w.Pm.template as<cx<T>>(), w.tauP.template as<cx<T>>(),
                                    w.gw.template as<cx<T>>(), batch, s));
        return launch_bulge<T>(C, Z, Rz, (size_t)Rz * C, w.Bs.template as<cx<T>>(),
                               w.uR2.template as<cx<T>>(), w.uL2.template as<cx<T>>(),
                               w.tR2.template as<cx<T>>(), w.tL2.template as<cx<T>>(),
                               w.d.template as<double>(), w.e.template as<double>(), batch, s,
                               clk);
    }
    return launch_gebrd<T>(Rz, C, Z, Rz, (size_t)Rz * C, w.d.template as<double>(),
                           w.e.template as<double>(), w.tauL.template as<cx<T>>(),
                           w.tauR.template as<cx<T>>(), w.Vr.template as<cx<T>>(),
                           w.gw.template as<cx<T>>(), batch, s, clk);
}

template <typename T>
inline ia_status svd_factor(SvdWork<T> &w, cx<T> *Y, hipStream_t s) {
    const int C = w.C, batch = w.batch;
    IA_TRY(svd_reduce<T>(w, Y, s));
    launch_bisect(C, w.d.template as<double>(), w.e.template as<double>(), w.sig.template as<double>(),
                  batch, s, sizeof(T) == sizeof(float));
    IA_HIP(hipGetLastError());
    launch_inviter(C, w.d.template as<double>(), w.e.template as<double>(),
                   w.sig.template as<double>(), w.QBt.template as<double>(),
                   w.PBt.template as<double>(), w.work.p, w.sub, batch, s,
                   sizeof(T) == sizeof(float));
    IA_HIP(hipGetLastError());
    return IA_OK;
}

// v <- U_b v for nvec length-R vectors per batch item (stride sv), where
// Y = U_b B V_b^H; Y is the matrix svd_reduce was given.  Rows C..R-1 of v
// must hold the vector's tail (zero for U_b [y; 0]).
template <typename T>
inline void svd_apply_u(SvdWork<T> &w, const cx<T> *Y, cx<double> *v, size_t sv, int nvec,
                        hipStream_t s) {
    const int R = w.R, C = w.C;
    if (w.ts)
        launch_bulge_apply<T>(false, C, w.uL2.template as<cx<T>>(), w.tL2.template as<cx<T>>(), v,
                              sv, nvec, w.batch, s);
    if (w.qr) {
        launch_apply_left<T>(false, C, C, w.Rb.template as<cx<T>>(), C, (size_t)C * C,
                             w.tauL.template as<cx<T>>(), v, sv, nvec, w.batch, s);
        launch_apply_left<T>(false, R, C, Y, R, (size_t)R * C, w.tauQ.template as<cx<T>>(), v, sv,
                             nvec, w.batch, s);
    } else {
        launch_apply_left<T>(false, R, C, Y, R, (size_t)R * C, w.tauL.template as<cx<T>>(), v, sv,
                             nvec, w.batch, s);
    }
}

// v <- U_b^H v (length-R vectors); the first C entries are the coordinates
// in the left singular basis of the bidiagonal
template <typename T>
inline void svd_apply_uh(SvdWork<T> &w, const cx<T> *Y, cx<double> *v, size_t sv, int nvec,
                         hipStream_t s) {
    const int R = w.R, C = w.C;
    if (w.qr) {
        launch_apply_left<T>(true, R, C, Y, R, (size_t)R * C, w.tauQ.template as<cx<T>>(), v, sv,
                             nvec, w.batch, s);
        launch_apply_left<T>(true, C, C, w.Rb.template as<cx<T>>(), C, (size_t)C * C,
                             w.tauL.template as<cx<T>>(), v, sv, nvec, w.batch, s);
    } else {
        launch_apply_left<T>(true, R, C, Y, R, (size_t)R * C, w.tauL.template as<cx<T>>(), v, sv,
                             nvec, w.batch, s);
    }
    if (w.ts)
        launch_bulge_apply<T>(true, C, w.uL2.template as<cx<T>>(), w.tL2.template as<cx<T>>(), v,
                              sv, nvec, w.batch, s);
}

// v <- V_b v (herm = false) or V_b^H v for nvec length-C vectors per batch
// item, Y = U_b B V_b^H
template <typename T>
inline void svd_apply_v(SvdWork<T> &w, bool herm, cx<double> *v, size_t sv, int nvec,
                        hipStream_t s) {
    const int C = w.C;
    if (!w.ts) {
        launch_apply_right<T>(herm, C, w.Vr.template as<cx<T>>(), w.tauR.template as<cx<T>>(), v,
                              sv, nvec, w.batch, s);
        return;
    }
    const int cp = C - BAND_NB;
    // V_b = P1 P2: P1 from the columns of Pm acting on entries BAND_NB..C-1
    if (herm) {
        launch_apply_left<T>(true, cp, cp, w.Pm.template as<cx<T>>(), cp, (size_t)cp * cp,
                             w.tauP.template as<cx<T>>(), v + BAND_NB, sv, nvec, w.batch, s);
        launch_bulge_apply<T>(true, C, w.uR2.template as<cx<T>>(), w.tR2.template as<cx<T>>(), v,
                              sv, nvec, w.batch, s);
    } else {
        launch_bulge_apply<T>(false, C, w.uR2.template as<cx<T>>(), w.tR2.template as<cx<T>>(), v,
                              sv, nvec, w.batch, s);
        launch_apply_left<T>(false, cp, cp, w.Pm.template as<cx<T>>(), cp, (size_t)cp * cp,
                             w.tauP.template as<cx<T>>(), v + BAND_NB, sv, nvec, w.batch, s);
    }
}

}  // namespace ia
