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
// sequential (sweep, task) order.  One workgroup per matrix; wave w runs the
// sweeps s = w mod NWB in order, and task (s, t) waits (LDS progress keys)
// until tasks 0..t+2 of sweep s-1 are done: that covers every earlier task
// whose block meets its own (tests/test_chase_schedule.py checks every
// pair of steps on the host), and tasks that are ready together touch
// disjoint blocks, so the result is the sequential one bit for bit.  The
// least a step needs is less: a right step only tasks 0..t+1 of sweep s-1,
// a left step also the right step of task t+2.  Starting right steps that
// early, with keys per half-step, measured slower (complex128 chase 84
// against 63 ms at 256 matrices; complex64 needs its right-hand block
// loaded with the right step's rows, which only the full wait allows).
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

// Wave-level LDS handoff: one wave's LDS operations execute in order, so
// this only has to stop the compiler from moving memory operations across
// it (a wavefront-scope fence: no wait for outstanding global stores).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // keep each phase's loads and registers inside the phase
    __builtin_amdgcn_sched_barrier(0);
}

// Per-wave LDS: the overlap block B[c:c+L, c:c+L] (row-major, padded), the
// current reflector u (in the data precision) and tau, beta.
template <typename T>
struct WaveLds {
    cx<T> tile[BNB][BNB + 1];
    cx<T> u[BNB];
    cx<double> pv[2];
};

// Reflector of the L values v_q held by lanes q < L (LAPACK zlarfg on
// (v_0, v_1..)), generated in double; u lands in us.u, tau/beta in us.pv.
template <typename T>
__device__ __forceinline__ void wave_larfg(cx<double> vq, int L, int lane, WaveLds<T> &us) {
    const double a2 = (lane >= 1 && lane < L) ? vq.re * vq.re + vq.im * vq.im : 0.0;
    const double xn2 = wave_sum(a2);
    const cx<double> alpha{__shfl(vq.re, 0, 64), __shfl(vq.im, 0, 64)};
    const Refl rf = larfg(alpha, xn2);
    // entries past L are zero so that the updates need no predicates
    if (lane < BNB) us.u[lane] = lane == 0 ? cx<T>{1, 0} : lane < L ? cvt<T>(vq * rf.scale) : cx<T>{0, 0};
    if (lane == 0) {
        us.pv[0] = rf.tau;
        us.pv[1] = cx<double>{rf.beta, 0};
    }
    wave_sync();
}

__device__ __forceinline__ float fmat(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmat(double a, double b, double c) { return __builtin_fma(a, b, c); }

// w += a * b and w += conj(a) * b with four fused multiply-adds each
template <typename T>
__device__ __forceinline__ void cmac(cx<T> &w, cx<T> a, cx<T> b) {
    w.re = fmat(a.re, b.re, fmat(-a.im, b.im, w.re));
    w.im = fmat(a.re, b.im, fmat(a.im, b.re, w.im));
}
template <typename T>
__device__ __forceinline__ void cmac_conj(cx<T> &w, cx<T> a, cx<T> b) {
    w.re = fmat(a.re, b.re, fmat(a.im, b.im, w.re));
    w.im = fmat(a.re, b.im, fmat(-a.im, b.re, w.im));
}
// x -= f * conj(u)
template <typename T>
__device__ __forceinline__ void csub_fconj(cx<T> &x, cx<T> f, cx<T> u) {
    x.re = fmat(-f.re, u.re, fmat(-f.im, u.im, x.re));
    x.im = fmat(-f.im, u.re, fmat(f.re, u.im, x.im));
}
// y -= u * f
template <typename T>
__device__ __forceinline__ void csub_uf(cx<T> &y, cx<T> u, cx<T> f) {
    y.re = fmat(-u.re, f.re, fmat(u.im, f.im, y.re));
    y.im = fmat(-u.re, f.im, fmat(-u.im, f.re, y.im));
}

// Band offset of element (i, j): j * BLD + i - j + BKU.
__device__ __forceinline__ int boff(int i, int j) { return j * (BLD - 1) + i + BKU; }

// Opaque copy of a lane offset: addresses derived from it are recomputed at
// their use instead of being kept live across the update (register budget).
__device__ __forceinline__ int fresh(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

// One task (s, t) on one wave.  Right step: lanes 0..nt-1 hold the rows
// r..c-1 above the overlap block, lanes 32..32+L-1 the overlap rows
// c..c+L-1, each over columns c..c+L-1.  Left step: lanes 0..L-1 hold the
// overlap columns (through the LDS tile, already right-updated), lanes
// 32..32+nR-1 the columns c+L.. right of it, each over rows c..c+L-1.  With
// PRE the right-hand columns are loaded together with the right step's
// rows.  FULL: L == BNB (all but the last task of a sweep).  The data is
// updated in its own precision; reflectors are generated in double.
//
// Carry: the right-hand block of task t (rows c..c+BNB-1, columns
// c+BNB..) is exactly the block above the overlap of task t+1 of the same
// sweep, which this wave runs next, and no other task touches it in
// between; so it stays in the LDS tile (to_tile) instead of going to memory
// and back (from_tile: task t+1 reads its top rows there).
template <typename T, bool PRE, bool FULL>
__device__ __forceinline__ void bulge_task(int C, cx<T> *__restrict__ B, int r, int c, int Ld,
                                           bool from_tile, bool to_tile, WaveLds<T> &us,
                                           cx<T> *__restrict__ uR, cx<T> *__restrict__ tR,
                                           cx<T> *__restrict__ uL, cx<T> *__restrict__ tL,
                                           int lane) {
    const int L = FULL ? BNB : Ld;
    const int nt = c - r;                                 // 1 (first task) or BNB
    const int nR = min(C, c + L + BNB) - (c + L);         // columns right of the overlap
    const bool top = lane < nt, ovl = lane >= 32 && lane - 32 < L;
    const int row = top ? r + lane : c + (lane - 32);
    const bool rcol = lane >= 32 && lane - 32 < nR;
    const int col = c + L + (lane - 32);
    const cx<T> *__restrict__ pr = B + boff(row, c);     // + q (BLD - 1): column c + q
    cx<T> x[BNB], y[BNB];
    // Loads are unconditional, with fixed offsets, so that all of them are
    // in flight at once: entries past L lie inside the padded band buffer
    // and are zeroed afterwards.
    const cx<T> zero{0, 0};
    if (top && from_tile) {
#pragma unroll
        for (int q = 0; q < BNB; ++q) x[q] = us.tile[lane][q];
    } else if (top || ovl) {
#pragma unroll
        for (int q = 0; q < BNB; ++q) x[q] = pr[q * (BLD - 1)];
    }
    if (PRE && rcol) {
        const cx<T> *__restrict__ pc = B + boff(c, col);  // + p: row c + p
#pragma unroll
        for (int p = 0; p < BNB; ++p) y[p] = pc[p];
    }
    if (!FULL) {
#pragma unroll
        for (int q = 0; q < BNB; ++q)
            if (q >= L) {
                x[q] = zero;
                if (PRE) y[q] = zero;
            }
    }
    cx<double> rq{0, 0};
    if (lane < L)
        rq = conj(cvt<double>(from_tile ? us.tile[0][lane] : B[boff(r, c + lane)]));
    wave_sync();  // tile reads done before the tile is rewritten
    // ---------------- right step: B[r:c+L, c:c+L] <- B H_R
    wave_larfg<T>(rq, L, lane, us);
    {
        const cx<T> tau = cvt<T>(us.pv[0]);
        if (top || ovl) {
            // u is read from LDS in chunks of 8 so that only the data block
            // and one chunk are live in registers
            cx<T> w0{0, 0}, w1{0, 0};   // two chains
#pragma unroll
            for (int q0 = 0; q0 < BNB; q0 += 8) {
#pragma unroll
                for (int q = q0; q < q0 + 8; q += 2) {
                    cmac(w0, x[q], us.u[q]);
                    cmac(w1, x[q + 1], us.u[q + 1]);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            const cx<T> f = tau * (w0 + w1);
#pragma unroll
            for (int q0 = 0; q0 < BNB; q0 += 8) {
#pragma unroll
                for (int q = q0; q < q0 + 8; ++q) csub_fconj(x[q], f, us.u[q]);
                __builtin_amdgcn_sched_barrier(0);
            }
            if (lane == 0) {
                x[0] = cx<T>{(T)us.pv[1].re, 0};
#pragma unroll
                for (int q = 1; q < BNB; ++q) x[q] = cx<T>{0, 0};
            }
        }
        if (top) {
            cx<T> *__restrict__ pw = B + fresh(boff(row, c));
#pragma unroll
            for (int q = 0; q < BNB; ++q)
                if (FULL || q < L) pw[q * (BLD - 1)] = x[q];
        }
        if (ovl) {
#pragma unroll
            for (int q = 0; q < BNB; ++q) us.tile[lane - 32][q] = x[q];
        }
        if (lane < L) uR[lane] = us.u[lane];
        if (lane == 0) *tR = cvt<T>(us.pv[0]);
    }
    wave_sync();
    // ---------------- left step: B[c:c+L, c:c+L+nR] <- H_L^H B
    if (lane < L) {
#pragma unroll
        for (int p = 0; p < BNB; ++p) {
            const cx<T> v = us.tile[p][lane];
            y[p] = (FULL || p < L) ? v : zero;
        }
    }
    if (!PRE && rcol) {
        const cx<T> *__restrict__ pc = B + boff(c, col);
#pragma unroll
        for (int p = 0; p < BNB; ++p) y[p] = pc[p];
        if (!FULL) {
#pragma unroll
            for (int p = 0; p < BNB; ++p)
                if (p >= L) y[p] = zero;
        }
    }
    const cx<double> cq = lane < L ? cvt<double>(us.tile[lane][0]) : cx<double>{0, 0};
    wave_larfg<T>(cq, L, lane, us);
    {
        const cx<T> ctau = conj(cvt<T>(us.pv[0]));
        if (lane < L || rcol) {
            cx<T> w0{0, 0}, w1{0, 0};
#pragma unroll
            for (int p0 = 0; p0 < BNB; p0 += 8) {
#pragma unroll
                for (int p = p0; p < p0 + 8; p += 2) {
                    cmac_conj(w0, us.u[p], y[p]);
                    cmac_conj(w1, us.u[p + 1], y[p + 1]);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            const cx<T> f = ctau * (w0 + w1);
#pragma unroll
            for (int p0 = 0; p0 < BNB; p0 += 8) {
#pragma unroll
                for (int p = p0; p < p0 + 8; ++p) csub_uf(y[p], us.u[p], f);
                __builtin_amdgcn_sched_barrier(0);
            }
            if (lane == 0) {
                y[0] = cx<T>{(T)us.pv[1].re, 0};
#pragma unroll
                for (int p = 1; p < BNB; ++p) y[p] = cx<T>{0, 0};
            }
            if (lane >= 32 && to_tile) {
#pragma unroll
                for (int p = 0; p < BNB; ++p) us.tile[p][lane - 32] = y[p];
            } else {
                cx<T> *__restrict__ pw = B + fresh(boff(c, lane < L ? c + lane : col));
#pragma unroll
                for (int p = 0; p < BNB; ++p)
                    if (FULL || p < L) pw[p] = y[p];
            }
        }
        if (lane < L) uL[lane] = us.u[lane];
        if (lane == 0) *tL = cvt<T>(us.pv[0]);
    }
    wave_sync();  // u / pv reads done before the next task rewrites them
}

// Progress keys: slot w holds s * KEYM + (tasks of sweep s done) for the
// sweep s = w mod NWB its wave is on; keys only grow.
constexpr int KEYM = 1 << 16;
// A wait that outlasts this many polls (seconds) marks the result NaN
// instead of hanging; it cannot happen when the dependencies are met.
constexpr long SPIN_MAX = 1L << 26;

__device__ __forceinline__ bool wait_key(const int *p, int need) {
    long n = 0;
    while (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < need) {
        if (++n > SPIN_MAX) return false;
        __builtin_amdgcn_s_sleep(2);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    return true;
}

// Dataflow chase: wave w runs sweeps s = w, w + NWB, ... in order; task
// (s, t) starts once tasks 0..t+2 of sweep s-1 are done (the tasks of
// sweep s-1 it overlaps); tasks that are ready together touch disjoint
// blocks, so the result is the sequential one.
template <typename T, int NWB, bool PRE>
__global__ __launch_bounds__(64 * NWB) void k_bulge(int C, cx<T> *__restrict__ Ball,
                                                    cx<T> *__restrict__ uRall,
                                                    cx<T> *__restrict__ uLall,
                                                    cx<T> *__restrict__ tRall,
                                                    cx<T> *__restrict__ tLall, int tmax,
                                                    double *__restrict__ dall,
                                                    double *__restrict__ eall) {
    __shared__ WaveLds<T> ush[NWB];
    __shared__ int prog[NWB];
    __shared__ int bad;
    const int b = blockIdx.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    cx<T> *B = Ball + (size_t)b * C * BLD;
    const size_t nslot = (size_t)(C - 1) * tmax;
    cx<T> *uR = uRall + (size_t)b * nslot * BNB;
    cx<T> *uL = uLall + (size_t)b * nslot * BNB;
    cx<T> *tR = tRall + (size_t)b * nslot;
    cx<T> *tL = tLall + (size_t)b * nslot;
    if (lane == 0) prog[wv] = -1;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    bool ok = true;
    for (int s = wv; s <= C - 2 && ok; s += NWB) {
        const int nts = ntask_of(C, s);
        const int ntp = s > 0 ? ntask_of(C, s - 1) : 0;
        for (int t = 0; t < nts; ++t) {
            if (s > 0 && !wait_key(&prog[(s - 1) % NWB], (s - 1) * KEYM + min(t + 3, ntp))) {
                ok = false;
                break;
            }
            const int r = t == 0 ? s : s + 1 + (t - 1) * BNB;
            const int c = t == 0 ? s + 1 : r + BNB;
            const int L = min(BNB, C - c);
            const size_t slot = (size_t)s * tmax + t;
            const bool from_tile = t > 0, to_tile = t + 1 < nts;
            if (L == BNB)
                bulge_task<T, PRE, true>(C, B, r, c, L, from_tile, to_tile, ush[wv],
                                         uR + slot * BNB, tR + slot, uL + slot * BNB, tL + slot,
                                         lane);
            else
                bulge_task<T, PRE, false>(C, B, r, c, L, from_tile, to_tile, ush[wv],
                                          uR + slot * BNB, tR + slot, uL + slot * BNB, tL + slot,
                                          lane);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            if (lane == 0)
                __hip_atomic_store(&prog[wv], s * KEYM + t + 1, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    if (!ok && lane == 0) bad = 1;
    // a wave that gave up publishes completion so no other wave waits on it
    if (lane == 0)
        __hip_atomic_store(&prog[wv], 0x7fffffff, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __syncthreads();
    double *d = dall + (size_t)b * C, *e = eall + (size_t)b * C;
    const double nan = __builtin_nan("");
    for (int i = threadIdx.x; i < C; i += blockDim.x) {
        d[i] = bad ? nan : (double)B[boff(i, i)].re;
        e[i] = bad ? nan : (i + 1 < C ? (double)B[boff(i, i + 1)].re : 0.0);
    }
}

// v <- H v for the stage-two reflectors of one side, in sequential order
// (FWD, with H^H: P2^H v, Q2^H v) or reverse order (P2 v, Q2 v).  Reflector
// (s, t) acts on entries c..c+L-1 of the vector, c = s + 1 + 32 t: the
// reflectors of one sweep touch disjoint ranges, and (s, t) overlaps only
// (s + 1, t - 1) and (s + 1, t) of the next sweep.  One workgroup (NWA
// waves) per (vector, matrix) with the vector in LDS; the sweeps are taken
// in application order p = 0, 1, ... (s = p forward, s = C - 2 - p in
// reverse), wave w taking p = w mod NWA, and reflector t of sweep p starts
// once sweep p - 1 has finished the reflectors it overlaps (t and t + 1
// forward, t - 1 and t in reverse), signalled through LDS progress keys as
// in the chase; reflectors that run together touch disjoint entries, so
// the result is the sequential product.  The next reflector of a sweep is
// loaded while the current one is applied.
constexpr int NWA = 16;
static_assert(BNB == 32, "one reflector entry per lane of each half-wave");

// (sum of v over lanes 0..31, sum over lanes 32..63), in every lane: the
// row sums of the DPP wave sum, then row_bcast:15 into rows 1 and 3 only
__device__ __forceinline__ cx<double> half_sums(double v) {
    using namespace hh::dpp;
    v += mv<QP_1032>(v);
    v += mv<QP_2301>(v);
    v += mv<HALF_MIRROR>(v);
    v += mv<MIRROR>(v);
    v += mv<BCAST15, 0xA>(v);
    const long long w = __builtin_bit_cast(long long, v);
    auto lane_val = [&](int l) {
        const int lo = __builtin_amdgcn_readlane((int)w, l);
        const int hi = __builtin_amdgcn_readlane((int)(w >> 32), l);
        return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
    };
    return {lane_val(31), lane_val(63)};
}

template <typename T, bool FWD>
__global__ __launch_bounds__(64 * NWA) void k_bulge_apply(int C, const cx<T> *__restrict__ uall,
                                                          const cx<T> *__restrict__ tall,
                                                          int tmax, cx<double> *__restrict__ vall,
                                                          size_t sv, int nvec) {
    extern __shared__ cx<double> vs[];          // C entries
    __shared__ int prog[NWA];
    __shared__ int bad;
    const int vec = blockIdx.x, b = blockIdx.y;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const size_t nslot = (size_t)(C - 1) * tmax;
    const cx<T> *u = uall + (size_t)b * nslot * BNB;
    const cx<T> *tau = tall + (size_t)b * nslot;
    cx<double> *v = vall + (size_t)b * nvec * sv + (size_t)vec * sv;
    for (int i = threadIdx.x; i < C; i += blockDim.x) vs[i] = v[i];
    if (lane == 0) prog[wv] = -1;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    const int np = C - 1;  // sweeps
    bool ok = true;
    for (int p = wv; p < np && ok; p += NWA) {
        const int s = FWD ? p : np - 1 - p;
        const int nts = ntask_of(C, s);
        // reflectors of the sweep applied before this one (p - 1)
        const int sp = FWD ? s - 1 : s + 1;
        const int ntp = p > 0 ? ntask_of(C, sp) : 0;
        const size_t slot0 = (size_t)s * tmax;
        cx<T> nu = u[slot0 * BNB + (lane & 31)];  // both halves hold the reflector
        cx<T> nt = tau[slot0];
        for (int t = 0; t < nts; ++t) {
            const cx<T> cu = nu, ct = nt;
            if (t + 1 < nts) {
                nu = u[(slot0 + t + 1) * BNB + (lane & 31)];
                nt = tau[slot0 + t + 1];
            }
            if (p > 0 && !wait_key(&prog[(p - 1) % NWA],
                                   (p - 1) * KEYM + min(FWD ? t + 2 : t + 1, ntp))) {
                ok = false;
                break;
            }
            const int c = s + 1 + t * BNB;
            const int L = min(BNB, C - c);
            const cx<double> tk = FWD ? conj(cvt<double>(ct)) : cvt<double>(ct);
            // lane q and lane 32 + q both hold entry q: the low half sums the
            // real parts of conj(u) x, the high half the imaginary parts, so
            // one 32-lane reduction serves both
            const int q = lane & 31;
            cx<double> x{0, 0}, uu{0, 0};
            if (q < L) {
                x = vs[c + q];
                uu = cvt<double>(cu);
            }
            const cx<double> pr = conj(uu) * x;
            const cx<double> dot = half_sums(lane < 32 ? pr.re : pr.im);
            if (lane < L) vs[c + lane] = x - uu * (tk * dot);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            if (lane == 0)
                __hip_atomic_store(&prog[wv], p * KEYM + t + 1, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    if (!ok && lane == 0) bad = 1;
    // a wave that gave up publishes completion so no other wave waits on it
    if (lane == 0)
        __hip_atomic_store(&prog[wv], 0x7fffffff, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __syncthreads();
    const double nan = __builtin_nan("");
    for (int i = threadIdx.x; i < C; i += blockDim.x) v[i] = bad ? cx<double>{nan, nan} : vs[i];
}

}  // namespace

int bulge_tmax(int C) { return ntask_of(C, 0); }

size_t bulge_band_elems(int C) { return (size_t)C * BLD; }
// the chase reads (and discards) up to BNB columns past a matrix's band
size_t bulge_band_pad_elems() { return (size_t)(BNB + 1) * BLD; }
size_t bulge_refl_elems(int C) { return C > 1 ? (size_t)(C - 1) * ntask_of(C, 0) : 0; }

template <typename T>
ia_status launch_bulge(int C, const cx<T> *Z, int ldz, size_t sZ, cx<T> *Bs, cx<T> *uR,
                       cx<T> *uL, cx<T> *tR, cx<T> *tL, double *d, double *e, int batch,
                       hipStream_t s, KernelClock *clk) {
    if (C < 2) return fail(IA_ERR_ARG, "bulge chasing needs C >= 2");
    const size_t nb = (size_t)C * BLD;
    hipLaunchKernelGGL(k_band_extract<T>, dim3((unsigned)((nb + 255) / 256), batch), dim3(256), 0,
                       s, C, Z, ldz, sZ, Bs);
    const int tmax = ntask_of(C, 0);
    // eight waves; complex64 also holds the right-hand block in registers
    // from the start of a task (two rows of 32 values per lane)
    constexpr bool pre = sizeof(T) == sizeof(float);
    if (clk) clk->mark(s);
    hipLaunchKernelGGL((k_bulge<T, 8, pre>), dim3(batch), dim3(512), 0, s, C, Bs, uR, uL, tR, tL,
                       tmax, d, e);
    if (clk) {
        clk->mark(s);
        clk->launches += 1;
        clk->bytes += bulge_alg_bytes(C, sizeof(cx<T>)) * batch;
    }
    IA_HIP(hipGetLastError());
    return IA_OK;
}

size_t bulge_apply_lds(int C) { return sizeof(cx<double>) * (size_t)C; }

double bulge_alg_bytes(int C, size_t elem) {
    // per task: the overlap block and the block right of it come from
    // memory; the rows above the overlap come from memory only for a
    // sweep's first task (later ones receive them in LDS); the rows above and
    // the overlap go back to memory, the right-hand block only after a
    // sweep's last task; plus the two reflectors
    double n = 0;
    for (int sw = 0; sw <= C - 2; ++sw) {
        const int nts = ntask_of(C, sw);
        for (int t = 0; t < nts; ++t) {
            const int r = t == 0 ? sw : sw + 1 + (t - 1) * BNB;
            const int c = t == 0 ? sw + 1 : r + BNB;
            const int L = std::min(BNB, C - c);
            const int nR = std::min(C, c + L + BNB) - (c + L);
            const int nt = c - r;
            n += 2.0 * L * L + (double)L * nR + (double)nt * L;   // OV in/out, RH in, TB out
            if (t == 0) n += (double)nt * L;                      // TB in
            if (t + 1 == nts) n += (double)L * nR;                // RH out
            n += 2.0 * L;                                         // reflectors
        }
    }
    return n * elem;
}

template <typename T>
void launch_bulge_apply(bool fwd, int C, const cx<T> *u, const cx<T> *tau, cx<double> *v,
                        size_t sv, int nvec, int batch, hipStream_t s) {
    if (C < 2) return;
    const int tmax = ntask_of(C, 0);
    dim3 grid(nvec, batch);
    const size_t lds = bulge_apply_lds(C);
    // C = 4096 needs the whole 64 KB of dynamic LDS
    (void)hipFuncSetAttribute((const void *)k_bulge_apply<T, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipFuncSetAttribute((const void *)k_bulge_apply<T, false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (fwd)
        hipLaunchKernelGGL((k_bulge_apply<T, true>), grid, dim3(64 * NWA), lds, s, C, u, tau, tmax,
                           v, sv, nvec);
    else
        hipLaunchKernelGGL((k_bulge_apply<T, false>), grid, dim3(64 * NWA), lds, s, C, u, tau,
                           tmax, v, sv, nvec);
}

#define INST(T)                                                                               \
    template ia_status launch_bulge<T>(int, const cx<T> *, int, size_t, cx<T> *, cx<T> *,      \
                                       cx<T> *, cx<T> *, cx<T> *, double *, double *, int,     \
                                       hipStream_t, KernelClock *);                           \
    template void launch_bulge_apply<T>(bool, int, const cx<T> *, const cx<T> *, cx<double> *, \
                                        size_t, int, int, hipStream_t);
INST(double)
INST(float)
#undef INST

}  // namespace ia
