// lds_peak.x -- measures the LDS atomic-add rate of the GPU, the roofline of
// the atomic pair kernels (one ds_add_u32 per pair-frame update).
//
// Three address patterns, each a run of ds_add_u32 with no other memory
// traffic:
//   conflict_free  every wave's 64 lanes hit 64 consecutive words
//   random_2500    uniform random words of a 2,500-word table (a 50 x 50
//                  histogram), the pattern of the pair kernels at 50 bins
//   same_address   all lanes of a wave hit one word
// Prints one JSON line: atomics/s and atomics per clock per CU per pattern.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

namespace {

constexpr int kWords = 10240;
constexpr int kBlock = 1024;
constexpr int kAddr = 16;

__device__ __forceinline__ unsigned hash32(unsigned x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_lds(unsigned* out, int iters) {
    __shared__ unsigned h[kWords];
    for (int i = threadIdx.x; i < kWords; i += kBlock) h[i] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned addr[kAddr];
#pragma unroll
    for (int k = 0; k < kAddr; k++) {
        if (MODE == 0) addr[k] = ((wave * kAddr + k) * 64 + lane) % kWords;
        else if (MODE == 1) addr[k] = hash32(blockIdx.x * 7919u + threadIdx.x * 131u + k) % 2500u;
        else addr[k] = (wave * kAddr + k) % kWords;
    }
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int k = 0; k < kAddr; k++) atomicAdd(&h[addr[k]], 1u);
    }
    __syncthreads();
    unsigned s = 0;
    for (int i = threadIdx.x; i < kWords; i += kBlock) s += h[i];
    out[(size_t)blockIdx.x * kBlock + threadIdx.x] = s;
}

template <int MODE>
double run(int blocks, int iters, unsigned* dout) {
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    k_lds<MODE><<<blocks, kBlock>>>(dout, 4);  // warm-up
    (void)hipEventRecord(a);
    k_lds<MODE><<<blocks, kBlock>>>(dout, iters);
    (void)hipEventRecord(b);
    (void)hipEventSynchronize(b);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    const double n = (double)blocks * kBlock * kAddr * iters;
    return n / (ms * 1e-3);
}

}  // namespace

int main() {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, 0) != hipSuccess) {
        fprintf(stderr, "no HIP device\n");
        return 1;
    }
    const int cus = p.multiProcessorCount;
    const double clk = p.clockRate * 1e3;  // Hz
    const int blocks = cus * 4;
    unsigned* dout = nullptr;
    if (hipMalloc(&dout, (size_t)blocks * kBlock * sizeof(unsigned)) != hipSuccess) return 1;
    const double r0 = run<0>(blocks, 2048, dout);
    const double r1 = run<1>(blocks, 2048, dout);
    const double r2 = run<2>(blocks, 256, dout);
    auto per = [&](double r) { return r / (cus * clk); };
    printf("{\"device\": \"%s\", \"cus\": %d, \"clock_mhz\": %.0f, "
           "\"conflict_free\": {\"atomics_per_s\": %.4e, \"per_clk_per_cu\": %.2f}, "
           "\"random_2500\": {\"atomics_per_s\": %.4e, \"per_clk_per_cu\": %.2f}, "
           "\"same_address\": {\"atomics_per_s\": %.4e, \"per_clk_per_cu\": %.2f}}\n",
           p.gcnArchName, cus, clk / 1e6, r0, per(r0), r1, per(r1), r2, per(r2));
    (void)hipFree(dout);
    return 0;
}
