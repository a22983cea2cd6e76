// Device helpers shared by the Householder kernels (svd.hip, gebrd.hip):
// wave/block reductions for 1024-thread workgroups and the complex
// reflector generator.
#pragma once
#include "ia_internal.hpp"

namespace ia {
namespace hh {

constexpr int BT = 1024;  // threads of a Householder workgroup
constexpr int NW = BT / 64;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

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
