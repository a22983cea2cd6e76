// ESM transfer-matrix assembly on gfx950.
//
// Restates CUDA/esm.cuh:1-36 (esmg_surface_kernel / esmg_field_kernel) and
// MATLAB/esmV.m:20-29:
//   D_ij   = ((ikR - 1)/(4 pi R^2)) e^{ikR} cos(theta_ij),  R = |r_i - q_j|
//   M_ij   = i rho c k e^{ikR}/(4 pi R)
// with cos(theta_ij) = n_i.(r_i - q_j)/R and n_i the unit normal (radial,
// r_i/|r_i|, when no normals are given: esm.cuh:11-15).
//
// Layout: one workgroup per 64x64 tile of one frequency (blockIdx.z).  The
// 64 source points of the tile are staged in LDS; each thread owns one row
// and walks 16 columns, so every store instruction writes 64 consecutive
// complex entries of a column (coalesced), unlike the reference's
// d[js + is*m] transposed stores.
#include "ia_internal.hpp"

namespace ia {

namespace {

constexpr int TILE = 64;
constexpr double INV4PI = 0.079577471545947667884;  // 1/(4 pi)

template <typename T>
__device__ __forceinline__ void sincos_t(T x, T *s, T *c);
template <>
__device__ __forceinline__ void sincos_t<double>(double x, double *s, double *c) { sincos(x, s, c); }
template <>
__device__ __forceinline__ void sincos_t<float>(float x, float *s, float *c) { sincosf(x, s, c); }

template <typename T, bool HERM>
__global__ __launch_bounds__(256) void k_esm_D(int m, const double *__restrict__ kv,
                                              const double *__restrict__ r_s,
                                              const double *__restrict__ src,
                                              const double *__restrict__ nrm,
                                              cx<T> *__restrict__ D, size_t strideD) {
    // D[i][j] (surface point i, source j) at D[j m + i]; HERM stores the
    // conjugate transpose, D^H[j][i] at D[i m + j].  The 64 lanes of a wave
    // walk the storage-contiguous index (i, or j for HERM) so that every
    // store is coalesced; the other index's coordinates are staged in LDS.
    __shared__ T sq[6][TILE];
    const int f = blockIdx.z;
    const int i0 = blockIdx.x * TILE, j0 = blockIdx.y * TILE;
    const int t = threadIdx.x;
    if (t < TILE) {
        if (!HERM) {  // sources of the tile
            const int j = j0 + t;
            for (int q = 0; q < 3; ++q) sq[q][t] = j < m ? (T)src[j + q * m] : (T)0;
        } else {  // surface points and normals of the tile
            const int i = i0 + t;
            T x = 0, y = 0, z = 0, nx = 0, ny = 0, nz = 0;
            if (i < m) {
                x = (T)r_s[i];
                y = (T)r_s[i + m];
                z = (T)r_s[i + 2 * m];
                if (nrm) {
                    nx = (T)nrm[i];
                    ny = (T)nrm[i + m];
                    nz = (T)nrm[i + 2 * m];
                } else {
                    const T rn = sqrt(x * x + y * y + z * z);
                    nx = x / rn;
                    ny = y / rn;
                    nz = z / rn;
                }
            }
            sq[0][t] = x;
            sq[1][t] = y;
            sq[2][t] = z;
            sq[3][t] = nx;
            sq[4][t] = ny;
            sq[5][t] = nz;
        }
    }
    __syncthreads();
    const T k = (T)kv[f];
    cx<T> *Df = D + (size_t)f * strideD;
    auto entry = [&](T x, T y, T z, T nx, T ny, T nz, T sx, T sy, T sz) {
        T dx = x - sx, dy = y - sy, dz = z - sz;
        T R2 = dx * dx + dy * dy + dz * dz;
        T R = sqrt(R2);
        T cth = (nx * dx + ny * dy + nz * dz) / R;
        T sn, cs;
        sincos_t<T>(k * R, &sn, &cs);
        // ((ikR - 1) e^{ikR}) = (-1 + i kR)(cs + i sn) = (-cs - kR sn) + i(kR cs - sn)
        T kR = k * R;
        T scale = (T)INV4PI * cth / R2;
        return cx<T>{(-cs - kR * sn) * scale, (kR * cs - sn) * scale};
    };
    const int lb = (t >> 6) * 16;
    if (!HERM) {
        const int i = i0 + (t & (TILE - 1));
        if (i >= m) return;
        const T x = (T)r_s[i], y = (T)r_s[i + m], z = (T)r_s[i + 2 * m];
        T nx, ny, nz;
        if (nrm) {
            nx = (T)nrm[i];
            ny = (T)nrm[i + m];
            nz = (T)nrm[i + 2 * m];
        } else {
            T rn = sqrt(x * x + y * y + z * z);
            nx = x / rn;
            ny = y / rn;
            nz = z / rn;
        }
        for (int jj = 0; jj < 16; ++jj) {
            const int jl = lb + jj, j = j0 + jl;
            if (j >= m) break;
            Df[(size_t)j * m + i] = entry(x, y, z, nx, ny, nz, sq[0][jl], sq[1][jl], sq[2][jl]);
        }
    } else {
        const int j = j0 + (t & (TILE - 1));
        if (j >= m) return;
        const T sx = (T)src[j], sy = (T)src[j + m], sz = (T)src[j + 2 * m];
        for (int ii = 0; ii < 16; ++ii) {
            const int il = lb + ii, i = i0 + il;
            if (i >= m) break;
            Df[(size_t)i * m + j] = conj(entry(sq[0][il], sq[1][il], sq[2][il], sq[3][il], sq[4][il],
                                               sq[5][il], sx, sy, sz));
        }
    }
}

// M (rows x m) or M^H (m x rows): i rho c k e^{ikR}/(4 pi R), R = |p_i - q_j|.
// With rowperm (CT only, m entries per frequency) output row r of M^H is
// source rowperm[r], i.e. the rows come out in a pivoted order P M^H.
template <typename T, bool CT>
__global__ __launch_bounds__(256) void k_esm_M(int rows, int m, const double *__restrict__ kv,
                                              double rho_c, const double *__restrict__ pr,
                                              const double *__restrict__ src,
                                              cx<T> *__restrict__ M, size_t strideM,
                                              const int *__restrict__ rowperm) {
    // CT=false: tile over (i rows of M, j sources); thread owns row i.
    // CT=true : output is m x rows; tile over (j rows of M^H, i columns);
    //           thread owns row j, walks columns i (mics) staged in LDS.
    __shared__ T sp[3][TILE];
    const int f = blockIdx.z;
    const int t = threadIdx.x;
    const int nr_out = CT ? m : rows;   // output rows
    const int nc_out = CT ? rows : m;   // output cols
    const int r0 = blockIdx.x * TILE, c0 = blockIdx.y * TILE;
    const double *colpts = CT ? pr : src;
    const int ncp = CT ? rows : m;
    const double *rowpts = CT ? src : pr;
    const int nrp = CT ? m : rows;
    if (t < TILE) {
        int c = c0 + t;
        if (c < nc_out) {
            sp[0][t] = (T)colpts[c];
            sp[1][t] = (T)colpts[c + ncp];
            sp[2][t] = (T)colpts[c + 2 * ncp];
        } else {
            sp[0][t] = sp[1][t] = sp[2][t] = (T)0;
        }
    }
    __syncthreads();
    const int r = r0 + (t & (TILE - 1));
    if (r >= nr_out) return;
    const T k = (T)kv[f];
    const T amp = (T)(rho_c * kv[f] * INV4PI);
    const int rp = (CT && rowperm) ? rowperm[(size_t)f * m + r] : r;
    const T x = (T)rowpts[rp], y = (T)rowpts[rp + nrp], z = (T)rowpts[rp + 2 * nrp];
    cx<T> *Mf = M + (size_t)f * strideM;
    const int cb = (t >> 6) * 16;
    for (int cc = 0; cc < 16; ++cc) {
        const int cl = cb + cc, c = c0 + cl;
        if (c >= nc_out) break;
        T dx = x - sp[0][cl], dy = y - sp[1][cl], dz = z - sp[2][cl];
        T R = sqrt(dx * dx + dy * dy + dz * dz);
        T sn, cs;
        sincos_t<T>(k * R, &sn, &cs);
        T a = amp / R;
        // i a (cs + i sn) = -a sn + i a cs
        cx<T> v{-a * sn, a * cs};
        if (CT) v.im = -v.im;
        Mf[(size_t)c * nr_out + r] = v;
    }
}

// p_i = sum_j M_sq(i,j) q_j with M_sq recomputed on the fly (no m x m store).
template <typename T>
__global__ __launch_bounds__(256) void k_esm_msq_apply(int m, const double *__restrict__ kv,
                                                      double rho_c, const double *__restrict__ r_s,
                                                      const double *__restrict__ src,
                                                      const cx<T> *__restrict__ q, size_t strideq,
                                                      cx<T> *__restrict__ p, size_t stridep) {
    __shared__ T sq[4][256];
    __shared__ cx<T> sv[256];
    const int f = blockIdx.y;
    const int t = threadIdx.x;
    const int i = blockIdx.x * 256 + t;
    const T k = (T)kv[f];
    const T amp = (T)(rho_c * kv[f] * INV4PI);
    T x = 0, y = 0, z = 0;
    if (i < m) {
        x = (T)r_s[i];
        y = (T)r_s[i + m];
        z = (T)r_s[i + 2 * m];
    }
    const cx<T> *qf = q + (size_t)f * strideq;
    T acc_re = 0, acc_im = 0;
    for (int j0 = 0; j0 < m; j0 += 256) {
        __syncthreads();
        int j = j0 + t;
        if (j < m) {
            sq[0][t] = (T)src[j];
            sq[1][t] = (T)src[j + m];
            sq[2][t] = (T)src[j + 2 * m];
            sv[t] = qf[j];
        } else {
            sq[0][t] = sq[1][t] = sq[2][t] = (T)1e30;
            sv[t] = cx<T>{0, 0};
        }
        __syncthreads();
        const int jn = min(256, m - j0);
        for (int jj = 0; jj < jn; ++jj) {
            T dx = x - sq[0][jj], dy = y - sq[1][jj], dz = z - sq[2][jj];
            T sn, cs, a;
            if constexpr (sizeof(T) == sizeof(float)) {
                // the complex64 form's surface pressure only feeds the sound
                // power: hardware reciprocal square root and sine / cosine
                // (absolute error ~1e-6 per entry, far below the 0.1 dB budget)
                const float r2 = dx * dx + dy * dy + dz * dz;
                const float ri = rsqrtf(r2);
                __sincosf(k * (r2 * ri), &sn, &cs);
                a = amp * ri;
            } else {
                const T R = sqrt(dx * dx + dy * dy + dz * dz);
                sincos_t<T>(k * R, &sn, &cs);
                a = amp / R;
            }
            T mre = -a * sn, mim = a * cs;
            cx<T> qv = sv[jj];
            acc_re += mre * qv.re - mim * qv.im;
            acc_im += mre * qv.im + mim * qv.re;
        }
    }
    if (i < m) p[(size_t)f * stridep + i] = cx<T>{acc_re, acc_im};
}

}  // namespace

template <typename T>
void launch_esm_D(int m, int nf, const double *k, const double *r_s, const double *src,
                  const double *nrm, cx<T> *D, size_t strideD, hipStream_t s, bool herm) {
    dim3 grid((m + TILE - 1) / TILE, (m + TILE - 1) / TILE, nf);
    if (herm)
        hipLaunchKernelGGL((k_esm_D<T, true>), grid, dim3(256), 0, s, m, k, r_s, src, nrm, D,
                           strideD);
    else
        hipLaunchKernelGGL((k_esm_D<T, false>), grid, dim3(256), 0, s, m, k, r_s, src, nrm, D,
                           strideD);
}

template <typename T>
void launch_esm_M(int rows, int m, int nf, const double *k, double rho_c, const double *r_rows,
                  const double *src, cx<T> *M, size_t strideM, bool conj_transpose,
                  hipStream_t s, const int *rowperm) {
    if (conj_transpose) {
        dim3 grid((m + TILE - 1) / TILE, (rows + TILE - 1) / TILE, nf);
        hipLaunchKernelGGL((k_esm_M<T, true>), grid, dim3(256), 0, s, rows, m, k, rho_c, r_rows,
                           src, M, strideM, rowperm);
    } else {
        dim3 grid((rows + TILE - 1) / TILE, (m + TILE - 1) / TILE, nf);
        hipLaunchKernelGGL((k_esm_M<T, false>), grid, dim3(256), 0, s, rows, m, k, rho_c, r_rows,
                           src, M, strideM, nullptr);
    }
}

template <typename T>
void launch_esm_msq_apply(int m, int nf, const double *k, double rho_c, const double *r_s,
                          const double *src, const cx<T> *q, size_t strideq, cx<T> *p,
                          size_t stridep, hipStream_t s) {
    dim3 grid((m + 255) / 256, nf);
    hipLaunchKernelGGL(k_esm_msq_apply<T>, grid, dim3(256), 0, s, m, k, rho_c, r_s, src, q,
                       strideq, p, stridep);
}

template void launch_esm_D<double>(int, int, const double *, const double *, const double *,
                                   const double *, cx<double> *, size_t, hipStream_t, bool);
template void launch_esm_D<float>(int, int, const double *, const double *, const double *,
                                  const double *, cx<float> *, size_t, hipStream_t, bool);
template void launch_esm_M<double>(int, int, int, const double *, double, const double *,
                                   const double *, cx<double> *, size_t, bool, hipStream_t,
                                   const int *);
template void launch_esm_M<float>(int, int, int, const double *, double, const double *,
                                  const double *, cx<float> *, size_t, bool, hipStream_t,
                                  const int *);
template void launch_esm_msq_apply<double>(int, int, const double *, double, const double *,
                                           const double *, const cx<double> *, size_t,
                                           cx<double> *, size_t, hipStream_t);
template void launch_esm_msq_apply<float>(int, int, const double *, double, const double *,
                                          const double *, const cx<float> *, size_t, cx<float> *,
                                          size_t, hipStream_t);

}  // namespace ia
