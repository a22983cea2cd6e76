// Autotuner of the batched complex GEMM (csrc/gemm.hip): for every distinct
// product a run of the library issued, every kernel candidate is timed on
// operands of that shape and the fastest is written to the tuning table the
// library reads (inverseacoustics_amd/gemm_tune.txt by default).
//
//   IA_GEMM_TUNE=off IA_GEMM_LOG=calls.txt python bench.py --steps 1 --warmup 0 --only mixed
//   IA_GEMM_TUNE=off IA_GEMM_LOG=calls.txt python bench.py --steps 1 --warmup 0 --only fp64
//   hipcc --offload-arch=gfx950 -O3 -std=c++17 -Xclang -target-feature \
//       -Xclang -packed-fp32-ops -Iinclude inverseacoustics_amd/csrc/bench/gemm_tune.hip \
//       -o inverseacoustics_amd/_build/gemm_tune
//   inverseacoustics_amd/_build/gemm_tune inverseacoustics_amd/gemm_tune.txt calls.txt
//
// Candidates: the bf16 split kernel (complex64 only) and the 3M kernel with
// 32 x 32, 32 x 64, 64 x 32 and 64 x 64 tiles (IA_TUNE_VERBOSE=1 prints every
// candidate's time per shape).  Each is run once to warm
// up, then timed over three launches (HIP events, the minimum kept), on
// pseudo-random operands with the logged leading dimensions and batch.  A
// candidate replaces the built-in choice only when it is at least 3 % faster
// (run-to-run spread of one launch is about 1-2 %), and every candidate's
// result is checked against the built-in choice's (largest difference over
// max |C|, FP32 or FP64 level) before it may be chosen.
#include "../gemm.hip"

#include <algorithm>
#include <map>
#include <string>
#include <tuple>
#include <vector>

using namespace ia;

#define CK(x)                                                                             \
    do {                                                                                  \
        hipError_t e = (x);                                                               \
        if (e != hipSuccess) {                                                            \
            std::printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__);          \
            std::exit(1);                                                                 \
        }                                                                                 \
    } while (0)

template <typename T>
__global__ void k_fill(cx<T> *p, size_t n, unsigned seed) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        unsigned h = (unsigned)(i * 2654435761u) ^ seed;
        h ^= h >> 13;
        h *= 0x5bd1e995;
        h ^= h >> 15;
        unsigned g = h * 1664525u + 1013904223u;
        p[i] = cx<T>{(T)(((h & 0xffffff) / 8388608.0 - 1.0)), (T)(((g >> 8) / 8388608.0 - 1.0))};
    }
}

// out[0] = max |a - b|, out[1] = max |b| (1-norm of the complex entries)
template <typename T>
__global__ void k_maxdiff(const cx<T> *a, const cx<T> *b, int M, int N, int ld, size_t sC,
                          int batch, double *out) {
    double m = 0, x = 0;
    const size_t tot = (size_t)M * N * batch;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
         e += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % M);
        const size_t r = e / M;
        const int j = (int)(r % N), bb = (int)(r / N);
        const size_t o = bb * sC + (size_t)j * ld + i;
        const double dr = (double)a[o].re - (double)b[o].re, di = (double)a[o].im - (double)b[o].im;
        m = fmax(m, (isnan(dr) || isnan(di)) ? INFINITY : fabs(dr) + fabs(di));
        x = fmax(x, fabs((double)b[o].re) + fabs((double)b[o].im));
    }
    atomicMax((unsigned long long *)&out[0], __double_as_longlong(m));
    atomicMax((unsigned long long *)&out[1], __double_as_longlong(x));
}

struct Call {
    char oa, ob;
    int M, N, K, batch, lda, ldb, ldc, beta0;
};

template <typename T>
void run_cand(int cand, const Call &c, const cx<T> *A, size_t sA, const cx<T> *B, size_t sB,
              cx<T> *C, size_t sC) {
    const cx<T> al{-1, 0}, be = c.beta0 ? cx<T>{0, 0} : cx<T>{1, 0};
#define IA_RUN(OA, OB)                                                                           \
    launch_cand<T, OA, OB>(cand, c.M, c.N, c.K, al, A, c.lda, sA, B, c.ldb, sB, be, C, c.ldc, sC, \
                           c.batch, 0)
    if (c.oa == 'N' && c.ob == 'N')
        IA_RUN('N', 'N');
    else if (c.oa == 'C' && c.ob == 'N')
        IA_RUN('C', 'N');
    else if (c.oa == 'N' && c.ob == 'C')
        IA_RUN('N', 'C');
    else
        IA_RUN('C', 'C');
#undef IA_RUN
}

template <typename T>
void tune(const std::vector<Call> &all, FILE *out, double &t_def, double &t_best) {
    // distinct shapes in first-seen order, with their call counts
    std::vector<Call> calls;
    std::vector<int> count;
    for (const Call &c : all) {
        auto same = [&](const Call &d) {
            return d.oa == c.oa && d.ob == c.ob && d.M == c.M && d.N == c.N && d.K == c.K &&
                   d.batch == c.batch && d.beta0 == c.beta0;
        };
        auto it = std::find_if(calls.begin(), calls.end(), same);
        if (it == calls.end()) {
            calls.push_back(c);
            count.push_back(1);
        } else {
            ++count[it - calls.begin()];
        }
    }
    size_t mA = 0, mB = 0, mC = 0;
    for (const Call &c : calls) {
        mA = std::max(mA, (size_t)c.lda * (c.oa == 'N' ? c.K : c.M) * c.batch);
        mB = std::max(mB, (size_t)c.ldb * (c.ob == 'N' ? c.N : c.K) * c.batch);
        mC = std::max(mC, (size_t)c.ldc * c.N * c.batch);
    }
    cx<T> *A, *B, *C0, *C1, *C2;
    CK(hipMalloc(&A, mA * sizeof(cx<T>)));
    CK(hipMalloc(&B, mB * sizeof(cx<T>)));
    CK(hipMalloc(&C0, mC * sizeof(cx<T>)));
    CK(hipMalloc(&C1, mC * sizeof(cx<T>)));
    CK(hipMalloc(&C2, mC * sizeof(cx<T>)));
    hipLaunchKernelGGL(k_fill<T>, dim3(4096), dim3(256), 0, 0, A, mA, 1u);
    hipLaunchKernelGGL(k_fill<T>, dim3(4096), dim3(256), 0, 0, B, mB, 2u);
    hipLaunchKernelGGL(k_fill<T>, dim3(4096), dim3(256), 0, 0, C0, mC, 3u);
    double *d;
    CK(hipMalloc(&d, 16));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    const bool f32 = sizeof(T) == sizeof(float);
    const double tol = f32 ? 4e-6 : 1e-13;
    std::map<std::string, std::tuple<int, double, double>> changed;
    for (size_t q = 0; q < calls.size(); ++q) {
        const Call &c = calls[q];
        const size_t sA = (size_t)c.lda * (c.oa == 'N' ? c.K : c.M),
                     sB = (size_t)c.ldb * (c.ob == 'N' ? c.N : c.K), sC = (size_t)c.ldc * c.N;
        const size_t nC = sC * c.batch;
        const int def = default_cand(f32, c.M, c.N, c.K, c.batch);
        // reference result: the built-in choice
        CK(hipMemcpy(C1, C0, nC * sizeof(cx<T>), hipMemcpyDeviceToDevice));
        run_cand<T>(def, c, A, sA, B, sB, C1, sC);
        double tc[GC_COUNT];
        for (int cand = 0; cand < GC_COUNT; ++cand) {
            tc[cand] = INFINITY;
            if (cand == GC_X6 && !f32) continue;
            CK(hipMemcpy(C2, C0, nC * sizeof(cx<T>), hipMemcpyDeviceToDevice));
            run_cand<T>(cand, c, A, sA, B, sB, C2, sC);
            CK(hipMemset(d, 0, 16));
            hipLaunchKernelGGL(k_maxdiff<T>, dim3(1024), dim3(256), 0, 0, C2, C1, c.M, c.N, c.ldc,
                               sC, c.batch, d);
            double h[2];
            CK(hipMemcpy(h, d, 16, hipMemcpyDeviceToHost));
            const double rel = h[0] / (h[1] > 0 ? h[1] : 1.0);
            if (!(rel <= tol * std::sqrt((double)std::max(c.K, 1)))) {
                std::printf("  candidate %d of %c%c %d %d %d x%d differs by %.2e: not eligible\n",
                            cand, c.oa, c.ob, c.M, c.N, c.K, c.batch, rel);
                continue;
            }
            float best = INFINITY;
            for (int rep = 0; rep < 3; ++rep) {
                float ms = 0;
                CK(hipEventRecord(e0, 0));
                run_cand<T>(cand, c, A, sA, B, sB, C2, sC);
                CK(hipEventRecord(e1, 0));
                CK(hipEventSynchronize(e1));
                CK(hipEventElapsedTime(&ms, e0, e1));
                best = std::min(best, ms);
            }
            tc[cand] = best;
        }
        if (std::getenv("IA_TUNE_VERBOSE")) {
            std::printf("  %d %c%c M%d N%d K%d x%d:", (int)sizeof(T), c.oa, c.ob, c.M, c.N, c.K,
                        c.batch);
            for (int cand = 0; cand < GC_COUNT; ++cand) std::printf(" %.4f", tc[cand]);
            std::printf("\n");
        }
        int pick = def;
        for (int cand = 0; cand < GC_COUNT; ++cand)
            if (tc[cand] < tc[pick]) pick = cand;
        if (!(tc[pick] < 0.97 * tc[def])) pick = def;
        t_def += count[q] * tc[def];
        t_best += count[q] * tc[pick];
        std::fprintf(out, "GEMMTUNE %d %c %c %d %d %d %d %d %d  %.4f %.4f %d\n", (int)sizeof(T),
                     c.oa, c.ob, c.M, c.N, c.K, c.batch, c.beta0, pick, tc[pick], tc[def], count[q]);
        if (pick != def) {
            char key[96];
            std::snprintf(key, sizeof key, "%d %c%c %d %d %d x%d", (int)sizeof(T), c.oa, c.ob, c.M,
                          c.N, c.K, c.batch);
            changed[key] = {pick, tc[pick] * count[q], tc[def] * count[q]};
        }
    }
    for (auto &kv : changed)
        std::printf("  %-32s candidate %d: %.3f ms against %.3f ms (all calls)\n", kv.first.c_str(),
                    std::get<0>(kv.second), std::get<1>(kv.second), std::get<2>(kv.second));
    std::fflush(stdout);
    (void)hipFree(A);
    (void)hipFree(B);
    (void)hipFree(C0);
    (void)hipFree(C1);
    (void)hipFree(C2);
    (void)hipFree(d);
}

std::vector<Call> load(const char *path, int want) {
    std::vector<Call> v;
    FILE *f = std::fopen(path, "r");
    if (!f) {
        std::printf("no %s\n", path);
        std::exit(1);
    }
    char line[256];
    while (std::fgets(line, sizeof line, f)) {
        Call c;
        int sz;
        if (std::sscanf(line, "GEMMLOG %d %c %c %d %d %d %d %d %d %d %d", &sz, &c.oa, &c.ob, &c.M,
                        &c.N, &c.K, &c.batch, &c.lda, &c.ldb, &c.ldc, &c.beta0) == 11 &&
            sz == want)
            v.push_back(c);
    }
    std::fclose(f);
    return v;
}

int main(int argc, char **argv) {
    if (argc < 3) {
        std::printf("usage: %s table.txt calls.txt [calls.txt ...]\n", argv[0]);
        return 2;
    }
    FILE *out = std::fopen(argv[1], "w");
    if (!out) {
        std::printf("cannot write %s\n", argv[1]);
        return 1;
    }
    std::fprintf(out,
                 "# GEMM tuning table (csrc/bench/gemm_tune.hip): sizeof(real) opA opB M N K batch "
                 "beta==0 candidate, then the candidate's and the built-in choice's time (ms) "
                 "and the calls of that shape in the log; candidates 0 = bf16 split, 1..4 = 3M "
                 "with 32x32, 32x64, 64x32, 64x64 tiles\n");
    std::vector<Call> c4, c8;
    for (int i = 2; i < argc; ++i) {
        std::vector<Call> a = load(argv[i], 4), b = load(argv[i], 8);
        c4.insert(c4.end(), a.begin(), a.end());
        c8.insert(c8.end(), b.begin(), b.end());
    }
    double d4 = 0, b4 = 0, d8 = 0, b8 = 0;
    if (!c4.empty()) tune<float>(c4, out, d4, b4);
    if (!c8.empty()) tune<double>(c8, out, d8, b8);
    std::fclose(out);
    std::printf("complex64: %zu calls, built-in %.2f ms, tuned %.2f ms\n", c4.size(), d4, b4);
    std::printf("complex128: %zu calls, built-in %.2f ms, tuned %.2f ms\n", c8.size(), d8, b8);
    return 0;
}
