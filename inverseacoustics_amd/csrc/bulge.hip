// Stage two of the two-stage bidiagonalisation: an upper band matrix of
// width NB (stage one's panel width) is reduced to real upper bidiagonal
// form by bulge chasing, and the vector transforms of the reflectors.
//
// Sweep s (s = 0 .. C-2) cleans row s.  Its task t touches
//   t = 0:  r = s,                 c = s + 1
//   t >= 1: r = s + 1 + (t-1) NB,  c = r + NB        (while c <= C - 2)
// with L = min(NB, C - c):
//   right step: H_R from conj(B[r, c:c+L]) (LAPACK zlarfg), B[r:c+L, c:c+L] <- B H_R,
//               so B[r, c] = beta (real), B[r, c+1:c+L] = 0;
//   left step:  H_L from B[c:c+L, c], B[c:c+L, c:min(C, c+L+NB)] <- H_L^H B,
//               so B[c, c] = beta (real), B[c+1:c+L, c] = 0.
// Band B = Q2 Bd P2^H with Q2 = H_L(1) H_L(2) ..., P2 = H_R(1) H_R(2) ... in
// sequential (sweep, task) order.  Task (s, t) runs at time step 3 s + t:
// tasks of one time step touch disjoint blocks, and any two tasks whose
// blocks meet run in sequential order, so the pipelined result is the
// sequential one bit for bit.  One workgroup per matrix; wave w runs the
// sweeps s = w mod NWB; a barrier closes every time step.
//
// Band storage: B(i, j) at j * LD + (i - j + KU), KU = 2 NB - 1, LD = 3 NB,
// which holds the fill of the chase (up to 2 NB - 1 above the diagonal and
// NB - 1 below it).
#include "householder.hpp"

namespace ia {

namespace {

using hh::Refl;
using hh::larfg;
using hh::wave_sum;

constexpr int BNB = 32;
constexpr int BLD = 3 * BNB;
constexpr int BKU = 2 * BNB - 1;

__host__ __device__ inline int ntask_of(int C, int s) {
    if (s > C - 2) return 0;
    const int rest = C - 3 - s;
    return 1 + (rest >= 0 ? rest / BNB : 0);
}

__device__ __forceinline__ size_t bidx(int i, int j) { return (size_t)j * BLD + (i - j + BKU); }

// Band of Z (upper, width BNB) into band storage, zeros elsewhere.
template <typename T>
__global__ void k_band_extract(int C, const cx<T> *__restrict__ Zall, int ldz, size_t sZ,
                               cx<T> *__restrict__ Ball) {
    const int b = blockIdx.y;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)C * BLD) return;
    const int j = (int)(e / BLD), q = (int)(e % BLD);
    const int i = j + q - BKU;
    cx<T> v{0, 0};
    if (i >= 0 && i <= j && j - i <= BNB) v = Zall[(size_t)b * sZ + (size_t)j * ldz + i];
    Ball[(size_t)b * C * BLD + e] = v;
}

// wave-level LDS handoff: lane 0's writes visible to every lane of the wave
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

template <typename T>
__device__ __forceinline__ void bulge_task(int C, cx<T> *__restrict__ B, int r, int c, int L,
                           cx<double> *__restrict__ us, cx<T> *__restrict__ uR,
                           cx<T> *__restrict__ tR, cx<T> *__restrict__ uL, cx<T> *__restrict__ tL,
                           int lane) {
    cx<T> x[BNB];
    // ---------------- right step: lane = row r + lane, columns c..c+L-1
    const int nr = c + L - r;
    if (lane < nr) {
#pragma unroll
        for (int q = 0; q < BNB; ++q)
            if (q < L) x[q] = B[bidx(r + lane, c + q)];
    }
    if (lane == 0) {
        double xn2 = 0;
#pragma unroll
        for (int q = 1; q < BNB; ++q)
            if (q < L) xn2 += (double)x[q].re * x[q].re + (double)x[q].im * x[q].im;
        const Refl rf = larfg(conj(cvt<double>(x[0])), xn2);
        us[0] = cx<double>{1, 0};
#pragma unroll
        for (int q = 1; q < BNB; ++q)
            if (q < L) us[q] = conj(cvt<double>(x[q])) * rf.scale;
        us[BNB] = rf.tau;
        us[BNB + 1] = cx<double>{rf.beta, 0};
    }
    wave_sync();
    {
        const cx<double> tau = us[BNB];
        if (lane < nr) {
            cx<double> w{0, 0};
#pragma unroll
            for (int q = 0; q < BNB; ++q)
                if (q < L) w = w + cvt<double>(x[q]) * us[q];
            const cx<double> f = tau * w;
            if (lane == 0) {
                x[0] = cx<T>{(T)us[BNB + 1].re, 0};
#pragma unroll
                for (int q = 1; q < BNB; ++q) x[q] = cx<T>{0, 0};
            } else {
#pragma unroll
                for (int q = 0; q < BNB; ++q)
                    if (q < L) x[q] = cvt<T>(cvt<double>(x[q]) - f * conj(us[q]));
            }
#pragma unroll
            for (int q = 0; q < BNB; ++q)
                if (q < L) B[bidx(r + lane, c + q)] = x[q];
        }
        if (lane < L) uR[lane] = cvt<T>(us[lane]);
        if (lane == 0) *tR = cvt<T>(tau);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    // ---------------- left step: lane = column c + lane, rows c..c+L-1
    const int nc = min(C, c + L + BNB) - c;
    if (lane < nc) {
#pragma unroll
        for (int p = 0; p < BNB; ++p)
            if (p < L) x[p] = B[bidx(c + p, c + lane)];
    }
    if (lane == 0) {
        double xn2 = 0;
#pragma unroll
        for (int p = 1; p < BNB; ++p)
            if (p < L) xn2 += (double)x[p].re * x[p].re + (double)x[p].im * x[p].im;
        const Refl rf = larfg(cvt<double>(x[0]), xn2);
        us[0] = cx<double>{1, 0};
#pragma unroll
        for (int p = 1; p < BNB; ++p)
            if (p < L) us[p] = cvt<double>(x[p]) * rf.scale;
        us[BNB] = rf.tau;
        us[BNB + 1] = cx<double>{rf.beta, 0};
    }
    wave_sync();
    {
        const cx<double> tau = us[BNB];
        if (lane < nc) {
            cx<double> w{0, 0};
#pragma unroll
            for (int p = 0; p < BNB; ++p)
                if (p < L) w = w + conj(us[p]) * cvt<double>(x[p]);
            const cx<double> f = conj(tau) * w;
            if (lane == 0) {
                x[0] = cx<T>{(T)us[BNB + 1].re, 0};
#pragma unroll
                for (int p = 1; p < BNB; ++p) x[p] = cx<T>{0, 0};
            } else {
#pragma unroll
                for (int p = 0; p < BNB; ++p)
                    if (p < L) x[p] = cvt<T>(cvt<double>(x[p]) - us[p] * f);
            }
#pragma unroll
            for (int p = 0; p < BNB; ++p)
                if (p < L) B[bidx(c + p, c + lane)] = x[p];
        }
        if (lane < L) uL[lane] = cvt<T>(us[lane]);
        if (lane == 0) *tL = cvt<T>(tau);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

template <typename T, int NWB>
__global__ __launch_bounds__(64 * NWB) void k_bulge(int C, cx<T> *__restrict__ Ball,
                                                    cx<T> *__restrict__ uRall,
                                                    cx<T> *__restrict__ uLall,
                                                    cx<T> *__restrict__ tRall,
                                                    cx<T> *__restrict__ tLall, int tmax,
                                                    int nsteps, double *__restrict__ dall,
                                                    double *__restrict__ eall) {
    __shared__ cx<double> ush[NWB][BNB + 2];
    const int b = blockIdx.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    cx<T> *B = Ball + (size_t)b * C * BLD;
    const size_t nslot = (size_t)(C - 1) * tmax;
    cx<T> *uR = uRall + (size_t)b * nslot * BNB;
    cx<T> *uL = uLall + (size_t)b * nslot * BNB;
    cx<T> *tR = tRall + (size_t)b * nslot;
    cx<T> *tL = tLall + (size_t)b * nslot;
    for (int ts = 0; ts < nsteps; ++ts) {
        const int shi = min(C - 2, ts / 3);
        const int lo = ts - tmax + 1;
        const int slo = lo <= 0 ? 0 : (lo + 2) / 3;
        const int s0 = slo + ((wv - slo % NWB) + NWB) % NWB;
        for (int s = s0; s <= shi; s += NWB) {
            const int t = ts - 3 * s;
            if (t < 0 || t >= ntask_of(C, s)) continue;
            const int r = t == 0 ? s : s + 1 + (t - 1) * BNB;
            const int c = t == 0 ? s + 1 : r + BNB;
            const int L = min(BNB, C - c);
            const size_t slot = (size_t)s * tmax + t;
            bulge_task<T>(C, B, r, c, L, ush[wv], uR + slot * BNB, tR + slot, uL + slot * BNB,
                          tL + slot, lane);
        }
        __syncthreads();
    }
    double *d = dall + (size_t)b * C, *e = eall + (size_t)b * C;
    for (int i = threadIdx.x; i < C; i += blockDim.x) {
        d[i] = (double)B[bidx(i, i)].re;
        e[i] = i + 1 < C ? (double)B[bidx(i, i + 1)].re : 0.0;
    }
}

// v <- H v for the stage-two reflectors of one side, in sequential order
// (FWD, with H^H: P2^H v, Q2^H v) or reverse order (P2 v, Q2 v).  One wave
// per (vector, matrix); the vector lives in LDS; the reflectors of CH task
// slots are staged through LDS with the next chunk's loads in flight.
constexpr int CH = 32;

template <typename T, bool FWD>
__global__ __launch_bounds__(64) void k_bulge_apply(int C, const cx<T> *__restrict__ uall,
                                                    const cx<T> *__restrict__ tall, int tmax,
                                                    cx<double> *__restrict__ vall, size_t sv,
                                                    int nvec) {
    extern __shared__ cx<double> vs[];          // C entries
    __shared__ cx<double> ub[CH][BNB];
    __shared__ cx<double> tb[CH];
    const int vec = blockIdx.x, b = blockIdx.y, lane = threadIdx.x;
    const size_t nslot = (size_t)(C - 1) * tmax;
    const cx<T> *u = uall + (size_t)b * nslot * BNB;
    const cx<T> *tau = tall + (size_t)b * nslot;
    cx<double> *v = vall + (size_t)b * nvec * sv + (size_t)vec * sv;
    for (int i = lane; i < C; i += 64) vs[i] = v[i];
    const long nchunk = (long)((nslot + CH - 1) / CH);
    // chunk loads: CH * BNB reflector entries, 64 lanes -> CH*BNB/64 each
    constexpr int PER = CH * BNB / 64;
    cx<T> pre[PER];
    cx<T> pret{0, 0};
    auto load = [&](long ch) {
        const size_t base = (size_t)ch * CH;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const size_t e = base * BNB + (size_t)k * 64 + lane;
            pre[k] = e < nslot * BNB ? u[e] : cx<T>{0, 0};
        }
        if (lane < CH) pret = base + lane < nslot ? tau[base + lane] : cx<T>{0, 0};
    };
    load(FWD ? 0 : nchunk - 1);
    for (long it = 0; it < nchunk; ++it) {
        const long ch = FWD ? it : nchunk - 1 - it;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int e = k * 64 + lane;
            ub[e / BNB][e % BNB] = cvt<double>(pre[k]);
        }
        if (lane < CH) tb[lane] = cvt<double>(pret);
        __syncthreads();
        if (it + 1 < nchunk) load(FWD ? ch + 1 : ch - 1);
        for (int kk = 0; kk < CH; ++kk) {
            const int k = FWD ? kk : CH - 1 - kk;
            const size_t slot = (size_t)ch * CH + k;
            if (slot >= nslot) continue;
            const int s = (int)(slot / tmax), t = (int)(slot % tmax);
            if (t >= ntask_of(C, s)) continue;
            const int c = t == 0 ? s + 1 : s + 1 + t * BNB;
            const int L = min(BNB, C - c);
            const cx<double> tk = FWD ? conj(tb[k]) : tb[k];
            cx<double> x{0, 0}, uu{0, 0};
            if (lane < L) {
                x = vs[c + lane];
                uu = ub[k][lane];
            }
            const cx<double> pr = conj(uu) * x;
            const cx<double> dot{wave_sum(pr.re), wave_sum(pr.im)};
            if (lane < L) vs[c + lane] = x - uu * (tk * dot);
            __syncthreads();
        }
    }
    __syncthreads();
    for (int i = lane; i < C; i += 64) v[i] = vs[i];
}

}  // namespace

int bulge_tmax(int C) { return ntask_of(C, 0); }

size_t bulge_band_elems(int C) { return (size_t)C * BLD; }
size_t bulge_refl_elems(int C) { return C > 1 ? (size_t)(C - 1) * ntask_of(C, 0) : 0; }

template <typename T>
ia_status launch_bulge(int C, const cx<T> *Z, int ldz, size_t sZ, cx<T> *Bs, cx<T> *uR,
                       cx<T> *uL, cx<T> *tR, cx<T> *tL, double *d, double *e, int batch,
                       hipStream_t s) {
    if (C < 2) return fail(IA_ERR_ARG, "bulge chasing needs C >= 2");
    const size_t nb = (size_t)C * BLD;
    hipLaunchKernelGGL(k_band_extract<T>, dim3((unsigned)((nb + 255) / 256), batch), dim3(256), 0,
                       s, C, Z, ldz, sZ, Bs);
    const int tmax = ntask_of(C, 0);
    int nsteps = 0;
    for (int sw = 0; sw <= C - 2; ++sw) nsteps = std::max(nsteps, 3 * sw + ntask_of(C, sw));
    // eight waves: a task keeps one block row (32 complex values) per lane
    // plus its double-precision work in registers (~230-250 VGPRs)
    hipLaunchKernelGGL((k_bulge<T, 8>), dim3(batch), dim3(512), 0, s, C, Bs, uR, uL, tR, tL, tmax,
                       nsteps, d, e);
    IA_HIP(hipGetLastError());
    return IA_OK;
}

size_t bulge_apply_lds(int C) { return sizeof(cx<double>) * (size_t)C; }

double bulge_alg_bytes(int C, size_t elem) {
    double n = 0;
    for (int sw = 0; sw <= C - 2; ++sw)
        for (int t = 0; t < ntask_of(C, sw); ++t) {
            const int r = t == 0 ? sw : sw + 1 + (t - 1) * BNB;
            const int c = t == 0 ? sw + 1 : r + BNB;
            const int L = std::min(BNB, C - c);
            const int nc = std::min(C, c + L + BNB) - c;
            n += (double)(c + L - r) * L + (double)L * nc;
        }
    return 2.0 * n * elem;
}

template <typename T>
void launch_bulge_apply(bool fwd, int C, const cx<T> *u, const cx<T> *tau, cx<double> *v,
                        size_t sv, int nvec, int batch, hipStream_t s) {
    if (C < 2) return;
    const int tmax = ntask_of(C, 0);
    dim3 grid(nvec, batch);
    const size_t lds = bulge_apply_lds(C);
    if (fwd)
        hipLaunchKernelGGL((k_bulge_apply<T, true>), grid, dim3(64), lds, s, C, u, tau, tmax, v,
                           sv, nvec);
    else
        hipLaunchKernelGGL((k_bulge_apply<T, false>), grid, dim3(64), lds, s, C, u, tau, tmax, v,
                           sv, nvec);
}

#define INST(T)                                                                               \
    template ia_status launch_bulge<T>(int, const cx<T> *, int, size_t, cx<T> *, cx<T> *,      \
                                       cx<T> *, cx<T> *, cx<T> *, double *, double *, int,     \
                                       hipStream_t);                                          \
    template void launch_bulge_apply<T>(bool, int, const cx<T> *, const cx<T> *, cx<double> *, \
                                        size_t, int, int, hipStream_t);
INST(double)
INST(float)
#undef INST

}  // namespace ia
