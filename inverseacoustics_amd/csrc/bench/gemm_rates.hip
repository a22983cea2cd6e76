// Rate table of the batched complex GEMM (csrc/gemm.hip) on the products
// a run of the library actually issued, beside rocBLAS's strided-batched
// cgemm/zgemm on the same shapes (the vendor library is only the yardstick
// here; the library itself does not link it).
//
//   IA_GEMM_LOG=calls.txt python bench.py --steps 1 --warmup 0 --only mixed
//   hipcc --offload-arch=gfx950 -O3 -std=c++17 -Xclang -target-feature \
//       -Xclang -packed-fp32-ops -Iinclude inverseacoustics_amd/csrc/bench/gemm_rates.hip \
//       -o inverseacoustics_amd/_build/gemm_rates -L/opt/rocm/lib -lrocblas
//   inverseacoustics_amd/_build/gemm_rates calls.txt
//
// Every logged call is replayed once for agreement (largest |own - rocBLAS|
// over max |C|) and once timed for each library (HIP events, one launch per
// call as in the sweep), on operands of the logged shapes filled with
// pseudo-random values; rows group the calls by operation, depth K and the
// smaller output dimension.
#include "../gemm.hip"
#include <rocblas/rocblas.h>
#include <vector>
#include <string>
#include <map>
using namespace ia;
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__); exit(1);} } while (0)
template <typename T>
__global__ void k_fill(cx<T> *p, size_t n, unsigned seed) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) {
        unsigned h = (unsigned)(i * 2654435761u) ^ seed;
        h ^= h >> 13; h *= 0x5bd1e995; h ^= h >> 15;
        unsigned g = h * 1664525u + 1013904223u;
        // irrational-looking values so that products round
        p[i] = cx<T>{(T)(((h & 0xffffff) / 8388608.0 - 1.0) * 1.0000001), (T)(((g >> 8) / 8388608.0 - 1.0) * 0.9999997)};
    }
}
template <typename T>
__global__ void k_maxdiff(const cx<T> *a, const cx<T> *b, int M, int N, int ld, size_t sC, int batch, double *out) {
    double m = 0, x = 0;
    size_t tot = (size_t)M * N * batch;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
        int i = e % M; size_t r = e / M; int j = r % N; int bb = r / N;
        size_t o = bb * sC + (size_t)j * ld + i;
        double dr = (double)a[o].re - (double)b[o].re, di = (double)a[o].im - (double)b[o].im;
        m = fmax(m, (isnan(dr) || isnan(di)) ? INFINITY : fabs(dr) + fabs(di));
        x = fmax(x, fabs((double)b[o].re) + fabs((double)b[o].im));
    }
    atomicMax((unsigned long long *)&out[0], __double_as_longlong(m));
    atomicMax((unsigned long long *)&out[1], __double_as_longlong(x));
}
static rocblas_handle H;
template <typename T>
void ref(char oa, char ob, int M, int N, int K, cx<T> al, const cx<T> *A, int lda, size_t sA,
         const cx<T> *B, int ldb, size_t sB, cx<T> be, cx<T> *C, int ldc, size_t sC, int batch) {
    auto op = [](char c) { return c == 'C' ? rocblas_operation_conjugate_transpose : rocblas_operation_none; };
    rocblas_status st;
    if (sizeof(T) == 4)
        st = rocblas_cgemm_strided_batched(H, op(oa), op(ob), M, N, K, (const rocblas_float_complex *)&al,
            (const rocblas_float_complex *)A, lda, sA, (const rocblas_float_complex *)B, ldb, sB,
            (const rocblas_float_complex *)&be, (rocblas_float_complex *)C, ldc, sC, batch);
    else
        st = rocblas_zgemm_strided_batched(H, op(oa), op(ob), M, N, K, (const rocblas_double_complex *)&al,
            (const rocblas_double_complex *)A, lda, sA, (const rocblas_double_complex *)B, ldb, sB,
            (const rocblas_double_complex *)&be, (rocblas_double_complex *)C, ldc, sC, batch);
    if (st != rocblas_status_success) { printf("rocblas status %d\n", (int)st); exit(1); }
}
struct Call { char oa, ob; int M, N, K, batch, lda, ldb, ldc, beta0; };

template <typename T>
void replay(const std::vector<Call> &calls) {
    size_t mA = 0, mB = 0, mC = 0;
    for (auto &c : calls) {
        mA = std::max(mA, (size_t)c.lda * (c.oa == 'N' ? c.K : c.M) * c.batch);
        mB = std::max(mB, (size_t)c.ldb * (c.ob == 'N' ? c.N : c.K) * c.batch);
        mC = std::max(mC, (size_t)c.ldc * c.N * c.batch);
    }
    cx<T> *A, *B, *C0, *C1, *C2;
    CK(hipMalloc(&A, mA * sizeof(cx<T>))); CK(hipMalloc(&B, mB * sizeof(cx<T>)));
    CK(hipMalloc(&C0, mC * sizeof(cx<T>))); CK(hipMalloc(&C1, mC * sizeof(cx<T>))); CK(hipMalloc(&C2, mC * sizeof(cx<T>)));
    hipLaunchKernelGGL(k_fill<T>, dim3(4096), dim3(256), 0, 0, A, mA, 1u);
    hipLaunchKernelGGL(k_fill<T>, dim3(4096), dim3(256), 0, 0, B, mB, 2u);
    hipLaunchKernelGGL(k_fill<T>, dim3(4096), dim3(256), 0, 0, C0, mC, 3u);
    double *d; CK(hipMalloc(&d, 16));
    hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
    double tr = 0, tm = 0, worst = 0; int nw = -1;
    std::map<std::string, std::pair<double, double>> bykey;
    std::map<std::string, double> flops;
    std::map<std::string, int> ncalls;
    for (size_t q = 0; q < calls.size(); ++q) {
        const Call &c = calls[q];
        size_t sA = (size_t)c.lda * (c.oa == 'N' ? c.K : c.M), sB = (size_t)c.ldb * (c.ob == 'N' ? c.N : c.K), sC = (size_t)c.ldc * c.N;
        cx<T> al{-1, 0}, be = c.beta0 ? cx<T>{0, 0} : cx<T>{1, 0};
        size_t nC = sC * c.batch;
        CK(hipMemcpy(C1, C0, nC * sizeof(cx<T>), hipMemcpyDeviceToDevice));
        CK(hipMemcpy(C2, C0, nC * sizeof(cx<T>), hipMemcpyDeviceToDevice));
        // warm + check
        ref<T>(c.oa, c.ob, c.M, c.N, c.K, al, A, c.lda, sA, B, c.ldb, sB, be, C1, c.ldc, sC, c.batch);
        launch_gemm<T>(c.oa, c.ob, c.M, c.N, c.K, al, A, c.lda, sA, B, c.ldb, sB, be, C2, c.ldc, sC, c.batch, 0);
        CK(hipDeviceSynchronize());
        CK(hipMemset(d, 0, 16));
        hipLaunchKernelGGL(k_maxdiff<T>, dim3(1024), dim3(256), 0, 0, C2, C1, c.M, c.N, c.ldc, sC, c.batch, d);
        double h[2]; CK(hipMemcpy(h, d, 16, hipMemcpyDeviceToHost));
        double rel = h[0] / (h[1] > 0 ? h[1] : 1);
        if (rel > worst) { worst = rel; nw = (int)q; }
        float a = 0, b = 0;
        CK(hipEventRecord(e0, 0));
        ref<T>(c.oa, c.ob, c.M, c.N, c.K, al, A, c.lda, sA, B, c.ldb, sB, be, C1, c.ldc, sC, c.batch);
        CK(hipEventRecord(e1, 0)); CK(hipEventSynchronize(e1)); CK(hipEventElapsedTime(&a, e0, e1));
        CK(hipEventRecord(e0, 0));
        launch_gemm<T>(c.oa, c.ob, c.M, c.N, c.K, al, A, c.lda, sA, B, c.ldb, sB, be, C2, c.ldc, sC, c.batch, 0);
        CK(hipEventRecord(e1, 0)); CK(hipEventSynchronize(e1)); CK(hipEventElapsedTime(&b, e0, e1));
        tr += a; tm += b;
        char key[64];
        snprintf(key, sizeof key, "%c%c K%s %s", c.oa, c.ob, c.K >= 512 ? ">=512" : std::to_string(c.K).c_str(),
                 std::min(c.M, c.N) <= 32 ? "min<=32" : (std::min(c.M, c.N) <= 64 ? "min<=64" : "big"));
        bykey[key].first += a; bykey[key].second += b;
        flops[key] += 8.0 * c.M * c.N * c.K * c.batch;
        ncalls[key] += 1;
    }
    printf("%s: %zu calls  rocBLAS %.2f ms  own %.2f ms  own/rocBLAS %.3f  largest |own - rocBLAS| / max|C| %.2e (call %d: %c%c %d %d %d)\n",
           sizeof(T) == 4 ? "complex64" : "complex128", calls.size(), tr, tm, tm / tr, worst, nw,
           calls[nw].oa, calls[nw].ob, calls[nw].M, calls[nw].N, calls[nw].K);
    const double pk = sizeof(T) == 4 ? 157.3 : 78.6;
    printf("   %-22s %6s %10s | %9s %8s %6s | %9s %8s %6s | %6s\n", "class (op, K, min(M,N))", "calls", "GFLOP",
           "rocBLAS ms", "TFLOP/s", "frac", "own ms", "TFLOP/s", "frac", "own/rb");
    for (auto &kv : bykey) {
        const double f = flops[kv.first], ra = f / (kv.second.first / 1e3) / 1e12, ro = f / (kv.second.second / 1e3) / 1e12;
        printf("   %-22s %6d %10.1f | %9.2f %8.1f %6.3f | %9.2f %8.1f %6.3f | %6.3f\n", kv.first.c_str(), ncalls[kv.first],
               f / 1e9, kv.second.first, ra, ra / pk, kv.second.second, ro, ro / pk, kv.second.second / kv.second.first);
    }
    fflush(stdout);
    (void)hipFree(A); (void)hipFree(B); (void)hipFree(C0); (void)hipFree(C1); (void)hipFree(C2);
}

std::vector<Call> load(const char *path, int want) {
    std::vector<Call> v; FILE *f = fopen(path, "r"); char line[256];
    if (!f) { printf("no %s\n", path); exit(1); }
    while (fgets(line, sizeof line, f)) {
        Call c; int sz;
        if (sscanf(line, "GEMMLOG %d %c %c %d %d %d %d %d %d %d %d", &sz, &c.oa, &c.ob, &c.M, &c.N, &c.K, &c.batch, &c.lda, &c.ldb, &c.ldc, &c.beta0) == 11 && sz == want)
            v.push_back(c);
    }
    fclose(f);
    return v;
}
int main(int argc, char **argv) {
    if (argc < 2) {
        printf("usage: %s calls.txt [calls.txt ...]\n", argv[0]);
        return 2;
    }
    rocblas_create_handle(&H);
    for (int i = 1; i < argc; ++i) {
        std::vector<Call> c4 = load(argv[i], 4), c8 = load(argv[i], 8);
        printf("%s\n", argv[i]);
        if (!c4.empty()) replay<float>(c4);
        if (!c8.empty()) replay<double>(c8);
    }
    return 0;
}
