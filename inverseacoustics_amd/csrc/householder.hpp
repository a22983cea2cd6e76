// Device helpers shared by the Householder kernels (svd.hip, gebrd.hip,
// geqrf.hip, bulge.hip): wave/block reductions for 1024-thread workgroups
// and the complex reflector generator.
#pragma once
#include "ia_internal.hpp"

namespace ia {
namespace hh {

constexpr int BT = 1024;  // threads of a Householder workgroup
constexpr int NW = BT / 64;

// Wave sums on the DPP lane network (no LDS traffic): pairs and quads by
// quad permutes, 8 and 16 lanes by the half-row and row mirrors, then
// row_bcast:15 / row_bcast:31 fold rows 0..3 into row 3; lane 63 holds the
// total, which is broadcast with readlane.  Lanes whose DPP source is masked
// off add zero.
namespace dpp {
constexpr int QP_1032 = 0xB1, QP_2301 = 0x4E, HALF_MIRROR = 0x141, MIRROR = 0x140,
              BCAST15 = 0x142, BCAST31 = 0x143;
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ float mv(float v) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, ROWS, 0xF, false));
}
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ double mv(double v) {
    const long long u = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)u, CTRL, ROWS, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(u >> 32), CTRL, ROWS, 0xF, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
}  // namespace dpp

template <typename R>
__device__ __forceinline__ R wave_sum_dpp(R v) {
    v += dpp::mv<dpp::QP_1032>(v);
    v += dpp::mv<dpp::QP_2301>(v);
    v += dpp::mv<dpp::HALF_MIRROR>(v);
    v += dpp::mv<dpp::MIRROR>(v);
    v += dpp::mv<dpp::BCAST15, 0xA>(v);
    v += dpp::mv<dpp::BCAST31, 0xC>(v);
    if constexpr (sizeof(R) == 4) {
        return __builtin_bit_cast(R, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
    } else {
        const long long u = __builtin_bit_cast(long long, v);
        const int lo = __builtin_amdgcn_readlane((int)u, 63);
        const int hi = __builtin_amdgcn_readlane((int)(u >> 32), 63);
        return __builtin_bit_cast(R, ((long long)hi << 32) | (unsigned)lo);
    }
}

// sum over the 64 lanes of a wave (every lane active), in every lane
__device__ __forceinline__ double wave_sum(double v) { return wave_sum_dpp(v); }

// block sum over BT threads; red must hold NW doubles; all threads get result
__device__ __forceinline__ double block_sum(double v, double *red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) s += red[i];
    return s;
}

struct Refl {
    double beta;
    cx<double> tau, scale;
};

// LAPACK zlarfg: H^H (alpha, x) = (beta, 0), H = I - tau u u^H, u = (1, x*scale)
__device__ __forceinline__ Refl larfg(cx<double> alpha, double xn2) {
    Refl r;
    if (xn2 == 0.0 && alpha.im == 0.0) {
        r.beta = alpha.re;
        r.tau = {0, 0};
        r.scale = {0, 0};
        return r;
    }
    double nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xn2);
    double beta = (alpha.re >= 0) ? -nrm : nrm;
    r.beta = beta;
    r.tau = {(beta - alpha.re) / beta, -alpha.im / beta};
    r.scale = cdiv(cx<double>{1, 0}, cx<double>{alpha.re - beta, alpha.im});
    return r;
}

}  // namespace hh
}  // namespace ia
