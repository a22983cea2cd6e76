/*
 * inverseacoustics_amd C ABI
 *
 * Drop-in boundary for the frequency-domain inverse solve of
 * LuisCorralR/InverseAcoustics (CUDA/esm.cuh, hels.cuh, ibem.cuh, invr.cuh,
 * tsvd.cuh, utils.cuh), implemented with HIP kernels for MI355X (gfx950).
 *
 * Conventions (identical to the reference):
 *   - complex values are {double re, im} pairs, layout-compatible with
 *     cuDoubleComplex, hipDoubleComplex, std::complex<double> and C99
 *     double _Complex;
 *   - matrices are column-major;
 *   - point sets are structure-of-arrays: x[0..n) y[0..n) z[0..n);
 *   - connectivity is structure-of-arrays and 0-based;
 *   - e^{+ikR} outgoing waves with e^{-i w t} time dependence;
 *   - the legacy entry points return host buffers allocated with malloc();
 *     the caller frees them with free().
 *
 * Differences from the reference, all deliberate:
 *   - no C++ exception crosses this boundary: the ia_* variants return an
 *     ia_status, the legacy void wrappers print the message and leave the
 *     outputs NULL (ia_last_error() holds the text);
 *   - c and rho are runtime parameters (ia_set_air); the default is the
 *     reference's compile-time C_AIR 343.5, RHO_AIR 1.19
 *     (CUDA/constants.cuh:2,5);
 *   - esm works for l != m (the reference's CUDA layout needs l == m,
 *     CUDA/esm.cuh:34,142,165-168);
 *   - the *gpu variants accept host or device pointers for their inputs.
 */
#ifndef IA_H
#define IA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } ia_z;
typedef struct { float re, im; } ia_c;

typedef enum {
    IA_OK = 0,
    IA_ERR_ARG = 1,        /* invalid argument (size, range, NULL)          */
    IA_ERR_DEVICE = 2,     /* HIP runtime error                             */
    IA_ERR_NOMEM = 3,      /* host or device allocation failed              */
    IA_ERR_SINGULAR = 4,   /* exact zero pivot in an LU factorisation       */
    IA_ERR_NONFINITE = 5,  /* NaN/Inf in inputs or results                  */
    IA_ERR_IO = 6,         /* file missing or malformed                     */
    IA_ERR_NOCONV = 7      /* iteration cap reached                         */
} ia_status;

/* ------------------------------------------------------------ runtime */
const char *ia_last_error(void);
const char *ia_version(void);
/* Speed of sound and air density used by every entry point of this thread.
 * Replaces the compile-time C_AIR / RHO_AIR of CUDA/constants.cuh:1-6. */
ia_status ia_set_air(double c, double rho);
ia_status ia_get_air(double *c, double *rho);
/* Number of visible HIP devices (0 when no GPU). */
int ia_device_count(void);
/* Wait for outstanding device work of this thread and free the device
 * workspace the sweeps keep between calls (it only grows, per host thread
 * and precision, so repeated sweeps allocate nothing). */
void ia_release(void);

/* -------------------------------------------------- ESM (esm.cuh) */
/* replaces esm(), CUDA/esm.cuh:112.  M_e = M_sq D^-1 (m x m),
 * H_e = M_hq D^-1 (l x m); r_ms (3m) surface points, r_lh (3l) field
 * points, sources at r_i * r_ms with radial normals (esm.cuh:9-15). */
void esm(int m, int l, double *r_ms, double *r_lh, ia_z **m_e, ia_z **h,
         double k, double r_i);
/* replaces esmgpu(), CUDA/esm.cuh:38 (same outputs; inputs host or device). */
void esmgpu(int m, int l, double *r_msg, double *r_lhg, ia_z **me, ia_z **he,
            double k, double r_i);
/* Status-returning general form: explicit source points src (3m) and unit
 * normals nrm (3m, NULL = radial).  Any output pointer may be NULL.
 * d, m_sq (m x m) and m_hq (l x m) return the assembled matrices. */
ia_status ia_esm(int m, int l, const double *r_ms, const double *r_lh,
                 const double *src, const double *nrm, double k,
                 ia_z **m_e, ia_z **h_e, ia_z **d, ia_z **m_sq, ia_z **m_hq);

/* ------------------------------------------------- HELS (hels.cuh) */
/* replaces hels(), CUDA/hels.cuh:222 and helsgpu(), hels.cuh:185.
 * r_ms_s (3m), r_lh_s (3l): spherical (azimuth, elevation, radius) SoA as
 * cartesian_2_spherical returns them.  m_e (m x n_t) = h_n(k r) Y_n^m,
 * h (l x n_t), q_e (m x n_t) = k h_n'(k r) Y_n^m / (i rho c k);
 * n_t = (n_sh+1)^2, column i -> n = floor(sqrt(i)), m = i - n - n^2
 * (CUDA order, hels.cuh:92).  FP64 recurrences (the CPU/GSL semantics),
 * 0 <= n_sh <= 80. */
void hels(int m, int l, double *r_ms_s, double *r_lh_s, ia_z **m_e, ia_z **h,
          ia_z **q_e, double k, int n_sh);
void helsgpu(int m, int l, double *r_ms_sg, double *r_lh_sg, ia_z **me,
             ia_z **he, ia_z **qe, double k, int n_sh);
ia_status ia_hels(int m, int l, const double *r_ms_s, const double *r_lh_s,
                  double k, int n_sh, ia_z **m_e, ia_z **h, ia_z **q_e);

/* --------------------------------------------- SVD + filters (tsvd.cuh) */
/* replaces csvd(), CUDA/tsvd.cuh:10: compact SVD a = u diag(s) v^H.
 * m >= n: u m x n, s n, v n x n;  m < n: u m x m, s m, v n x m. */
void csvd(int m, int n, ia_z *a, ia_z **u, double **s, ia_z **v);
ia_status ia_csvd(int m, int n, const ia_z *a, ia_z **u, double **s, ia_z **v);
/* replaces tsvd(), CUDA/tsvd.cuh:122: x = V(:,1:k) diag(1/s(1:k)) U(:,1:p)^H b
 * (u m x p, v n x p); 1 <= k <= p as the reference requires (tsvd.cuh:124). */
void tsvd(int m, int n, int p, int k, ia_z *u, double *s, ia_z *v, ia_z *b,
          ia_z **x);
/* Tikhonov (tikhonovG.m:26-30) and damped SVD (dsvdG.m:23-26) solutions
 * from an SVD, lambda >= 0. */
ia_status ia_tikhonov(int m, int n, int p, const ia_z *u, const double *s,
                      const ia_z *v, const ia_z *b, double lambda, ia_z **x);
ia_status ia_dsvd(int m, int n, int p, const ia_z *u, const double *s,
                  const ia_z *v, const ia_z *b, double lambda, ia_z **x);

/* ------------------------------------------------------ dense kernels */
/* C = alpha op(A) op(B) + beta C, op in {'N','C'}, column-major complex128;
 * host or device pointers, synchronous.  Replaces the cublasZgemm call
 * sites of CUDA/ibem.cuh:398 and CUDA/invr.cuh:402-433. */
ia_status ia_zgemm(char opa, char opb, int m, int n, int k, ia_z alpha,
                   const ia_z *a, int lda, const ia_z *b, int ldb, ia_z beta,
                   ia_z *c, int ldc);
/* Same product in complex64 (FP32 matrix cores, FP32 accumulation). */
ia_status ia_cgemm(char opa, char opb, int m, int n, int k, ia_c alpha,
                   const ia_c *a, int lda, const ia_c *b, int ldb, ia_c beta,
                   ia_c *c, int ldc);
/* Solve D X = B (op 'N') or D^H X = B (op 'C') for square D (n x n) with
 * partial-pivoting LU; B (n x nrhs) is overwritten by X.  Replaces the
 * square cusolverDnZZgels / zgels_ solves of CUDA/esm.cuh:73-85,153-168. */
ia_status ia_zgesv(char op, int n, int nrhs, const ia_z *d, ia_z *b);

/* ------------------------------------------------ utils (utils.cuh) */
/* replaces to_pressure(), to_velocity(), sound_power() (utils.cuh:199,232,265) */
void to_pressure(int m, int n, ia_z *me, ia_z *x, ia_z **p);
void to_velocity(int m, int n, ia_z *qe, ia_z *x, ia_z **v);
void sound_power(int m, ia_z *p, ia_z *v, double *w, double s);
/* replaces sphere(), utils.cuh:384: pulsating sphere of radius a, velocity u. */
void sphere(double *r_ms, double *r_lh, ia_z **p_s, ia_z **p_h, ia_z **v_s,
            double k, double u, double a, int m, int l);
/* replaces geometry(), utils.cuh:291 (Gmsh MSH 2.2, triangles only,
 * degenerate triangles dropped, field points r_ms (1 + 2 pi t / k)). */
ia_status ia_geometry(const char *file, double k, double t, double **r_ms,
                      double **r_lh, int **e_n, int *m, int *l, int *n);
void geometry(double **r_ms, double **r_lh, int **e_n, double k, double t,
              int *m_out, int *l_out, int *n_out, char *file);
/* replaces cartesian_2_spherical(), utils.cuh:414. */
void cartesian_2_spherical(int m, double *r_xx, double **azelr);

/* ---------------------------------------- parameter choice (regtools) */
typedef enum {
    IA_FILTER_TIKHONOV = 0,  /* f = s^2/(s^2+lam^2)                */
    IA_FILTER_TSVD = 1,      /* f = 1 for i < k                    */
    IA_FILTER_DSVD = 2       /* f = s/(s+lam)                      */
} ia_filter;

typedef enum {
    IA_SELECT_GCV = 0,        /* argmin of GCV on the lambda grid (Tikhonov,
                                 dsvd) / discrete GCV over k (TSVD)         */
    IA_SELECT_LCURVE = 1,     /* argmin of lcfun (negative curvature)       */
    IA_SELECT_FIXED = 2,      /* use lambda / k given by the caller         */
    IA_SELECT_GCV_REGTOOLS = 3,   /* literal gcvG.m: 200-point grid + fminbnd */
    IA_SELECT_LCURVE_REGTOOLS = 4,/* literal l_curveG.m: 200 points + fminbnd */
    IA_SELECT_QUASIOPT = 5,       /* quasioptG.m: 200 points + fminbnd        */
    IA_SELECT_RANK = 6            /* TSVD: k = numerical rank (singular values
                                     > eps max(rows, cols) s_1): the minimum-
                                     norm least-squares solution            */
} ia_select;

typedef struct {
    ia_filter filter;
    ia_select select;
    int n_lambda;        /* grid points (default 50)                    */
    double decades;      /* grid spans sigma_max 10^-decades .. sigma_max */
    double lambda;       /* IA_SELECT_FIXED: lambda                     */
    int k;               /* IA_SELECT_FIXED + TSVD: truncation index    */
} ia_reg_params;

/* One regularised solve from an SVD: selects lambda (or k) and returns
 * x (n), the selected lambda, its grid index and the TSVD index. */
ia_status ia_regularize(int m, int n, int p, const ia_z *u, const double *s,
                        const ia_z *v, const ia_z *b, const ia_reg_params *prm,
                        ia_z **x, double *lambda, int *index);

/* The solve of ia_regularize for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Every pointer is a device
 * pointer: u (m x p), s (p), v (n x p), b (m) in; x (n), lambda (1 double)
 * and index (1 int, as ia_regularize returns it) out, lambda and index may
 * be NULL.  The work is enqueued on `stream` (a hipStream_t; NULL = default
 * stream) and nothing is allocated: the intermediates live in a fixed
 * 128 KiB device array per GPU, so p <= 4096 (larger p returns IA_ERR_ARG),
 * and calls that overlap on one GPU race. */
ia_status ia_regularize_noalloc(int m, int n, int p, const ia_z *u, const double *s,
                                const ia_z *v, const ia_z *b, const ia_reg_params *prm,
                                ia_z *x, double *lambda, int *index, void *stream);

/* -------------------------------------------------- batched ESM sweep */
typedef enum { IA_PREC_FP64 = 0, IA_PREC_MIXED = 1 } ia_precision;

typedef struct {
    int m, l;                 /* surface points (= sources), microphones   */
    int n_freq;               /* frequencies in the sweep                  */
    ia_precision precision;
    ia_reg_params reg;
    double area;              /* S in W = 1/2 Re(v^H p) S/m                */
    int chunk;                /* frequencies per device batch (0 = auto)   */
    /* IA_PREC_MIXED only: a frequency whose selection is nearly tied (the
     * GCV / L-curve value of the runner-up parameter within this relative
     * gap of the chosen one) is solved again in FP64 and its outputs
     * replaced.  0 = 2e-4, < 0 = never. */
    double recheck_gap;
} ia_sweep_desc;

typedef struct {
    double *lambda;           /* [n_freq] selected lambda                  */
    int *lambda_index;        /* [n_freq] grid index (Tikhonov/dsvd)       */
    int *k_tsvd;              /* [n_freq] discrete-GCV TSVD index          */
    double *power;            /* [n_freq] W                                */
    double *sigma_max;        /* [n_freq]                                  */
    ia_z *velocity;           /* [n_freq * m] surface normal velocity      */
    ia_z *pressure;           /* [n_freq * m] surface pressure             */
    int *status;              /* [n_freq] ia_status per frequency          */
    double *selection_gap;    /* [n_freq] relative GCV / L-curve gap between
                                 the chosen and the runner-up parameter (k
                                 for TSVD); small = nearly tied choice     */
} ia_sweep_out;

/* Solve an ESM sweep: for each frequency f, assemble D, M_hq at k[f],
 * H_e = M_hq D^-1, SVD, select the regularisation parameter, solve
 * H_e v = b_f, and reconstruct p_s = M_sq D^-1 v and W.  Inputs are
 * host or device pointers (r_s, src, nrm: 3m; r_h: 3l; b: n_freq*l
 * column per frequency); outputs are host pointers, any may be NULL. */
ia_status ia_esm_sweep(const ia_sweep_desc *desc, const double *k,
                       const double *r_s, const double *src, const double *nrm,
                       const double *r_h, const ia_z *b, ia_sweep_out *out);

/* Same, with every pointer (inputs and outputs) on the device and the work
 * enqueued on `stream` (a hipStream_t; NULL = default stream).  Returns
 * after the work is enqueued; scratch is owned by the library. */
ia_status ia_esm_sweep_device(const ia_sweep_desc *desc, const double *k,
                              const double *r_s, const double *src,
                              const double *nrm, const double *r_h,
                              const ia_z *b, ia_sweep_out *out, void *stream);

/* ------------------------------- gradients of the ESM solve (adjoint) */
/* No counterpart in the reference, which has no derivative code; the chain
 * differentiated is its forward path esm.cuh:1-92 (assembly, H_e = M_hq D^-1),
 * tikhonovG.m:29-30 (filtered solution) and utils.cuh:199-289 (surface
 * pressure, sound power).  For each frequency f: sources q = r_s - d nrm
 * (retreat d along the unit normals nrm), the Tikhonov solution v of
 * H_e v = b_f at lambda_f (held fixed), the misfit J = ||H_e v - b_f||^2 and
 * W = 1/2 Re(v^H M_sq D^-1 v) S/m, with their gradients with respect to d,
 * the microphone coordinates r_h and log lambda_f, by a hand-written adjoint
 * that recomputes the assembled matrices in the backward pass.  FP64,
 * l <= m.  lambda: [n_freq] values, or NULL to select each lambda_f first by
 * desc->reg (Tikhonov filter; the chosen value is then held fixed).  Inputs
 * host or device pointers, outputs host or device pointers (any NULL). */
typedef struct {
    int m, l, n_freq;
    ia_reg_params reg;        /* used when lambda is NULL                  */
    double area;              /* S in W                                    */
    int chunk;                /* frequencies per device batch (0 = auto)   */
} ia_grad_desc;

typedef struct {
    double *lambda;           /* [n_freq] lambda used                      */
    double *misfit;           /* [n_freq] J                                */
    double *power;            /* [n_freq] W                                */
    double *dmisfit_dd;       /* [n_freq] dJ/dd                            */
    double *dpower_dd;        /* [n_freq] dW/dd                            */
    double *dmisfit_dmic;     /* [n_freq * 3l] dJ/dr_h, SoA (x, y, z blocks)
                                 per frequency                             */
    double *dpower_dmic;      /* [n_freq * 3l] dW/dr_h                     */
    double *dmisfit_dloglam;  /* [n_freq] dJ/dlog(lambda)                  */
    double *dpower_dloglam;   /* [n_freq] dW/dlog(lambda)                  */
    ia_z *velocity;           /* [n_freq * m] v                            */
} ia_grad_out;

ia_status ia_esm_gradient(const ia_grad_desc *desc, const double *k, const double *r_s,
                          const double *nrm, double d, const double *r_h, const ia_z *b,
                          const double *lambda, ia_grad_out *out);

/* Batched HELS sweep with a per-frequency expansion order (uneven system
 * sizes are grouped by order): H (l x n_t) regularised solve for the
 * coefficients c, then v = Q_e c (to_velocity) and p = M_e c (to_pressure)
 * on the m surface points, W = 1/2 Re(v^H p) S/m.  Host pointers:
 * order[n_freq], k[n_freq], azelr_s (3m), azelr_h (3l), b (n_freq*l). */
typedef struct {
    int m, l, n_freq;
    ia_precision precision;
    ia_reg_params reg;
    double area;
} ia_hels_desc;
ia_status ia_hels_sweep(const ia_hels_desc *desc, const int *order,
                        const double *k, const double *azelr_s,
                        const double *azelr_h, const ia_z *b,
                        ia_sweep_out *out);

/* ------------------------------------------------- IBEM (ibem.cuh) */
/* Quadrature tables: regular rule (n_q points: xi1, xi2, w SoA) and three
 * singular pages of n_qs points each (page p: xi1, xi2, w SoA at
 * 3*n_qs*p).  ia_quadrature_load reads the reference's text format
 * (count line, then "xi1 xi2 w" rows; IPFs with $PAGE 1..3 sections,
 * CUDA/ibem.cuh:431-477); the returned buffers are malloc'd.  Without a
 * call to ia_set_quadrature the library uses Radon's 7-point degree-5
 * rule and Duffy-collapsed 40x40 Gauss-Legendre singular pages. */
ia_status ia_quadrature_load(const char *ipf_path, const char *ipfs_path,
                             int *n_q, double **ipf, int *n_qs, double **ipfs);
/* Set the tables used by the IBEM / radiation-mode entry points of this
 * thread (copied); n_q = n_qs = 0 restores the built-in rules. */
ia_status ia_set_quadrature(int n_q, const double *ipf, int n_qs,
                            const double *ipfs);
/* replaces ibem(), CUDA/ibem.cuh:424 and ibemgpu(), ibem.cuh:263.
 * m_e = Z = A_ss \ B_ss (m x m; least squares when n_ch CHIEF points are
 * given), h = H_b = A_hs Z - B_hs (l x m); r_ms (3m) and r_lh (3l) SoA,
 * e_n (3n) 0-based triangles SoA, cp (3 n_ch) CHIEF points SoA. */
void ibem(int m, int l, int n, double *r_ms, int *e_n, double *r_lh,
          ia_z **m_e, ia_z **h, double k, int n_ch, double *cp);
void ibemgpu(int m, int l, int n, double *r_msg, int *e_ng, double *r_lhg,
             ia_z **me, ia_z **he, double k, int n_ch, double *cp);
ia_status ia_ibem(int m, int l, int n, const double *r_ms, const int *e_n,
                  const double *r_lh, double k, int n_ch, const double *cp,
                  ia_z **z, ia_z **h_b);

/* Batched IBEM sweep (ibemS.m pipeline per frequency): assemble Z and H_b
 * for every k[f], regularised solve H_b v = b_f, p_s = Z v and
 * W = 1/2 Re(v^H p_s) S/m.  Host pointers; b holds n_freq columns of l. */
typedef struct {
    int m, l, n_elem, n_ch, n_freq;
    ia_precision precision;   /* MIXED: complex64 assembly and SVD        */
    ia_reg_params reg;
    double area;
    int chunk;                /* frequencies per device batch (0 = auto)   */
} ia_ibem_desc;
ia_status ia_ibem_sweep(const ia_ibem_desc *desc, const double *k,
                        const double *r_ms, const int *e_n, const double *r_lh,
                        const double *cp, const ia_z *b, ia_sweep_out *out);

/* ------------------------------------ radiation modes (invr.cuh) */
/* replaces invr(), CUDA/invr.cuh:457 and invrgpu(), invr.cuh:263, with the
 * argument order of the reference: m_e = Z Q (m x n_am), h = H_b Q
 * (l x n_am), q_e = Q (m x n_am), the leading left singular vectors of
 * R = real(G Z) (MATLAB/invrS.m:160-161), G the Gram matrix of the linear
 * shape functions. */
void invr(int m, int l, int n, double *r_ms, int *e_n, double *r_lh,
          ia_z **m_e, ia_z **h, ia_z **q_e, double k, int n_ch, double *cp,
          int n_am);
void invrgpu(int m, int l, int n, double *r_msg, int *e_ng, double *r_lhg,
             ia_z **me, ia_z **he, ia_z **qe, double k, int n_ch, double *cp,
             int n_rm);
ia_status ia_invr(int m, int l, int n, const double *r_ms, const int *e_n,
                  const double *r_lh, double k, int n_ch, const double *cp,
                  int n_am, ia_z **q, ia_z **m_r, ia_z **h_r, ia_z **g_mass);
/* Elemental-radiator radiation resistance of np patches of area a_e
 * (r_p: 3 np SoA): R_ij = (w^2 rho a_e^2/(4 pi c)) sin(k r_ij)/(k r_ij),
 * R_ii with sin(x)/x = 1; returns the n_modes leading eigenvalues
 * (descending) and real modes (np x n_modes, largest entry positive). */
ia_status ia_radiation_modes(int np, const double *r_p, double a_e, double k,
                             int n_modes, double **eigval, double **modes);

/* Batched radiation-mode sweep on elemental radiators (the reference's
 * radiation-mode path MATLAB/invrS.m:157-170, CUDA/invr.cuh:426-433, with
 * the elemental-radiator resistance instead of the BEM one): np patches of
 * equal area A on a baffled plane (r_p: 3 np SoA), l microphones (r_h: 3 l),
 * pressures b (n_freq columns of l).  Per frequency: the n_modes leading
 * FP64 eigenpairs (e, Q) of R_ij = (w^2 rho A^2/(4 pi c)) sin(k r_ij)/(k r_ij)
 * (assembled on the device), H_r = G Q with the Rayleigh transfer
 * G_ij = -(i w rho A/(2 pi)) e^{ik R_ij}/R_ij, the regularised solution y of
 * H_r y = b_f (IA_PREC_MIXED: complex64 SVD of H_r, FP64 parameter choice),
 * W = sum_i e_i |y_i|^2 (= v^H R v) and v = Q y.  Modes with
 * e_i < eig_floor e_1 are dropped (zero columns of Q and H_r, zero
 * amplitudes): their shapes are not resolved in FP64.  np <= 4096.  Inputs
 * host or device pointers; outputs host or device pointers, any NULL. */
typedef struct {
    int np, l, n_freq, n_modes;
    ia_precision precision;
    ia_reg_params reg;
    double area;              /* patch area A                              */
    int chunk;                /* frequencies per device batch (0 = auto)   */
    double eig_floor;         /* relative eigenvalue floor (0 = 1e-8,
                                 < 0 = keep all n_modes)                   */
} ia_radiation_desc;

typedef struct {
    double *power;            /* [n_freq] W                                */
    double *lambda;           /* [n_freq] selected lambda                  */
    int *lambda_index;        /* [n_freq] grid index                       */
    int *k_tsvd;              /* [n_freq] discrete-GCV TSVD index          */
    double *eigval;           /* [n_freq * n_modes] e, descending          */
    double *modes;            /* [n_freq * np * n_modes] Q, column-major per
                                 frequency, largest entry of each positive */
    ia_z *amplitude;          /* [n_freq * n_modes] y                      */
    ia_z *velocity;           /* [n_freq * np] v                           */
    int *status;              /* [n_freq]                                  */
} ia_radiation_out;

ia_status ia_radiation_sweep(const ia_radiation_desc *desc, const double *k,
                             const double *r_p, const double *r_h, const ia_z *b,
                             ia_radiation_out *out);

/* Batched regularised solve of A_f x_f = b_f for any transfer matrices
 * (rows x cols each, column-major, batch stride rows*cols; host or device
 * pointers): SVD, parameter choice, filtered solution.  lambda, index
 * (grid index), k_tsvd and sigma_max are [batch] outputs (any may be NULL). */
ia_status ia_regsolve(int rows, int cols, int batch, const ia_z *a,
                      const ia_z *b, const ia_reg_params *prm,
                      ia_precision precision, ia_z *x, double *lambda,
                      int *index, int *k_tsvd, double *sigma_max);

/* ADMM regularised solutions (MATLAB/ADMMG.m:1-89) of a batch of systems
 * A_f c_f = b_f (rows x cols each, column-major, batch stride rows*cols;
 * host or device pointers), FP64.  norm 1: L1 (complex soft threshold,
 * continuation lambda = 5 lambda0 * 0.75^k while > lambda0 / 10); norm 2:
 * L2; norm 3: Lp with 0 < p < 1 (generalised soft thresholding, n_gst
 * fixed-point steps).  Iterations stop when |c - c_old|, |c - x| and
 * |u - u_old| are all <= epsilon, at most max_iter per continuation stage.
 * The start is the minimum-norm least-squares solution (ADMMG.m uses
 * MATLAB's A\b).  c: [batch * cols]; iterations: [batch] (may be NULL). */
typedef struct {
    double lambda0, rho, gamma, epsilon;
    int max_iter, norm;
    double p;
    int n_gst;
} ia_admm_params;
ia_status ia_admm(int rows, int cols, int batch, const ia_z *a, const ia_z *b,
                  const ia_admm_params *prm, ia_z *c, int *iterations);

/* ------------------------------------- scenarios (MATLAB/linesource.m) */
/* Analytic line source (MATLAB/linesource.m:1-65), batched over n_freq
 * wavenumbers k: a source of length len and width a on the y axis, centred
 * at the origin, vibrating with velocity u0, sampled at n_ls points.
 * xp = i (rho c u0 k a / 2) trapz(e^{ikr}/r) at the l field points r_lh
 * (l x n_freq); xsf the same on the m surface points r_ms, vsf the normal
 * velocity (u0 a / 2) trapz((ikr - 1) e^{ikr}/r^2 n.(x - y)/r) there
 * (m x n_freq each), n the normalised sum of (x1 - x3) x (x2 - x3) over a
 * node's triangles of e_n (3 n_elem, 0-based SoA).  rho and c are
 * arguments, as in linesource.m.  Geometry: host pointers; outputs: host or
 * device pointers.  m = 0 or l = 0 skips that output. */
ia_status ia_linesource(int n_ls, int m, int n_elem, const double *r_ms,
                        const int *e_n, int l, const double *r_lh, double rho,
                        double c, double u0, const double *k, int n_freq,
                        double a, double len, ia_z *xp, ia_z *xsf, ia_z *vsf);

/* ------------------------------------ text output (CUDA/utils.cuh) */
/* replace write_file_cuDoubleComplex(), write_file_string(),
 * write_file_double() (utils.cuh:32-74): mode is the fopen mode ("w" or
 * "a"); complex entries "%f,%f" and real entries "%f", separated by ';',
 * one '\n' at the end.  Status variants return IA_ERR_IO when the file
 * cannot be opened (the reference writes through a NULL stream). */
ia_status ia_write_complex(const char *mode, const char *file, const ia_z *v,
                           int n);
ia_status ia_write_double(const char *mode, const char *file, const double *v,
                          int n);
ia_status ia_write_string(const char *mode, const char *file, const char *text);
void write_file_cuDoubleComplex(const char *wa, char *file, ia_z *v, int n);
void write_file_string(const char *wa, char *file, char *text);
void write_file_double(const char *wa, char *file, double *v, int n);
/* replaces write_matrix_cuDoubleComplex() (utils.cuh:116): r x c column-major
 * matrix, row by row, "%f,%f," (fmt "f") or "%e,%e," (fmt "e") per entry,
 * '\n' per row and a final '\n'.  The legacy name writes "matrix.txt". */
ia_status ia_write_matrix(const char *fmt, const char *file, const ia_z *a,
                          int r, int c);
void write_matrix_cuDoubleComplex(const char *d, ia_z *a, int r, int c);
/* replace print_matrix_cuDoubleComplex(), print_matrix_double()
 * (utils.cuh:76,159): the same matrices on stdout. */
void print_matrix_cuDoubleComplex(const char *d, ia_z *a, int r, int c);
void print_matrix_double(const char *d, double *a, int r, int c);

/* Phase timings (ms) of the last sweep on this thread, plus the summed
 * HIP-event duration, algorithmic bytes and launch count of the panel
 * bidiagonalisation kernel (the sweep's dominant kernel). */
typedef struct {
    double assemble, factor, solve_he, bidiag, bisect, vectors, select,
           post, total;
    double bidiag_kernel_ms, bidiag_kernel_bytes, bidiag_kernel_launches;
    double rechecked;         /* frequencies decided again in FP64 (mixed) */
    double recheck_ms;        /* time of those FP64 solves                 */
    double recheck_full;      /* of the rechecked, those that took the full
                                 FP64 sweep (the others were decided from the
                                 Gram form of the GCV function)            */
} ia_stats;
ia_status ia_last_stats(ia_stats *st);

#ifdef __cplusplus
}
#endif
#endif /* IA_H */
