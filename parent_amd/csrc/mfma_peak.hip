// mfma_peak.x -- measures the matrix-core rates of the GPU that bound the
// one-hot MFMA pair kernel (counts = A_i^T A_j over frames, 2 B_i B_j
// operations per pair-frame update):
//   int8  v_mfma_i32_32x32x32_i8             (the operand type the kernel uses)
//   fp4   v_mfma_scale_f32_32x32x64_f8f6f4   with both operands e2m1 and unit
//         scales: one-hot values 0 / 1.0 are exact in fp4, and FP32 sums of
//         0 / 1 products are exact below 2^24, so it could serve the same
//         counts with twice the K per instruction
// Each wave keeps 4 independent accumulators in flight; 8 workgroups of 256
// threads per compute unit.  Prints one JSON line: operations per second and
// per clock per CU for each form.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr unsigned kUnitScales = 0x7F7F7F7Fu;  // E8M0 127 = 2^0 in every byte

__global__ __launch_bounds__(256) void k_int8(int iters, int* out) {
    const v4i a = {(int)threadIdx.x, 1, 2, 3}, b = {3, 2, 1, (int)threadIdx.x};
    v16i c0 = {}, c1 = {}, c2 = {}, c3 = {};
    for (int i = 0; i < iters; i++) {
        c0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c1, 0, 0, 0);
        c2 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c2, 0, 0, 0);
        c3 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c3, 0, 0, 0);
    }
    int s = 0;
    for (int v = 0; v < 16; v++) s += c0[v] + c1[v] + c2[v] + c3[v];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_fp4(int iters, float* out) {
    const v8i a = {(int)threadIdx.x, 0x22, 0x2020, 3, 0, 0, 0, 0}, b = {0x02, 0x20, 0x22, (int)threadIdx.x, 0, 0, 0, 0};
    v16f c0 = {}, c1 = {}, c2 = {}, c3 = {};
    for (int i = 0; i < iters; i++) {
        c0 = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c0, 4, 4, 0, kUnitScales, 0, kUnitScales);
        c1 = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c1, 4, 4, 0, kUnitScales, 0, kUnitScales);
        c2 = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c2, 4, 4, 0, kUnitScales, 0, kUnitScales);
        c3 = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c3, 4, 4, 0, kUnitScales, 0, kUnitScales);
    }
    float s = 0;
    for (int v = 0; v < 16; v++) s += c0[v] + c1[v] + c2[v] + c3[v];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

}  // namespace

int main() {
    int dev = 0, cus = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeClockRate, dev) != hipSuccess) {
        fprintf(stderr, "no HIP device\n");
        return 1;
    }
    const int blocks = cus * 8, iters = 20000;
    int* o1 = nullptr;
    float* o2 = nullptr;
    if (hipMalloc(&o1, (size_t)blocks * 256 * 4) != hipSuccess || hipMalloc(&o2, (size_t)blocks * 256 * 4) != hipSuccess) {
        fprintf(stderr, "hipMalloc failed\n");
        return 1;
    }
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return 1;
    // MACs per instruction: 32 x 32 x 32 (int8), 32 x 32 x 64 (fp4); 2 operations per MAC
    const double per_call = (double)blocks * 4 * iters * 4;
    double best[2] = {0, 0};
    for (int rep = 0; rep < 3; rep++) {
        float ms = 0;
        (void)hipEventRecord(a);
        k_int8<<<blocks, 256>>>(iters, o1);
        (void)hipEventRecord(b);
        if (hipEventSynchronize(b) != hipSuccess) return 1;
        (void)hipEventElapsedTime(&ms, a, b);
        if (rep) best[0] = std::max(best[0], per_call * 32.0 * 32 * 32 * 2 / (ms * 1e-3));
        (void)hipEventRecord(a);
        k_fp4<<<blocks, 256>>>(iters, o2);
        (void)hipEventRecord(b);
        if (hipEventSynchronize(b) != hipSuccess) return 1;
        (void)hipEventElapsedTime(&ms, a, b);
        if (rep) best[1] = std::max(best[1], per_call * 32.0 * 32 * 64 * 2 / (ms * 1e-3));
    }
    const double clk = (double)khz * 1e3;
    printf("{\"cus\": %d, \"clock_hz\": %.0f, \"int8_32x32x32\": {\"ops_per_s\": %.4e, \"ops_per_clk_per_cu\": %.1f}, "
           "\"fp4_32x32x64_scaled\": {\"ops_per_s\": %.4e, \"ops_per_clk_per_cu\": %.1f}}\n",
           cus, clk, best[0], best[0] / clk / cus, best[1], best[1] / clk / cus);
    (void)hipFree(o1);
    (void)hipFree(o2);
    return 0;
}
