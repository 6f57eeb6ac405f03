// lds_peak.x -- measures the LDS atomic-add rate of the GPU, the roofline of
// the atomic pair kernels (one ds_add_u32 per pair-frame update).
//
// Three address patterns, each a run of ds_add_u32 with no other memory
// traffic:
//   conflict_free  every wave's 64 lanes hit 64 consecutive words
//   random_2500    uniform random words of a 2,500-word table (a 50 x 50
//                  histogram), the pattern of the pair kernels at 50 bins
//   same_address   all lanes of a wave hit one word
//   stride_N       lane l hits word l*N (bank structure: N = 32 puts every
//                  lane of a 32-bank LDS in one bank)
//   lane_sliced    lane l only touches words w with w % 32 == l % 32 (random
//                  word otherwise); lanes l and l+32 share words, adding to
//                  different bytes -- the layout of per-lane sub-histograms
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
        else if (MODE == 2) addr[k] = (wave * kAddr + k) % kWords;
        else if (MODE == 3) addr[k] = (lane * 2) % kWords;
        else if (MODE == 4) addr[k] = (lane * 16) % kWords;
        else if (MODE == 5) addr[k] = (lane * 32) % kWords;
        else if (MODE == 6) addr[k] = (lane * 64) % kWords;
        else addr[k] = ((hash32(blockIdx.x * 7919u + threadIdx.x * 131u + k) % 320u) * 32u + (lane & 31)) % kWords;
    }
    const unsigned inc = MODE == 7 ? (1u << (8 * ((lane >> 5) & 1))) : 1u;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int k = 0; k < kAddr; k++) atomicAdd(&h[addr[k]], inc);
    }
    __syncthreads();
    unsigned s = 0;
    for (int i = threadIdx.x; i < kWords; i += kBlock) s += h[i];
    out[(size_t)blockIdx.x * kBlock + threadIdx.x] = s;
}

// The sliced-histogram pattern over a dynamic table of `words` words, one
// 1024-thread workgroup per CU: word = cell * 16 + lane % 16 with random
// cells in [0, words / 16) (MODE 0), or uniform random words (MODE 1).
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_lds_big(unsigned* out, int iters, int words) {
    extern __shared__ unsigned hb[];
    for (int i = threadIdx.x; i < words; i += kBlock) hb[i] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    unsigned addr[kAddr];
#pragma unroll
    for (int k = 0; k < kAddr; k++) {
        const unsigned h = hash32(blockIdx.x * 7919u + threadIdx.x * 131u + k);
        addr[k] = MODE == 0 ? (h % (unsigned)(words / 16)) * 16u + (lane & 15) : h % (unsigned)words;
    }
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int k = 0; k < kAddr; k++) atomicAdd(&hb[addr[k]], 1u);
    }
    __syncthreads();
    unsigned s = 0;
    for (int i = threadIdx.x; i < words; i += kBlock) s += hb[i];
    out[(size_t)blockIdx.x * kBlock + threadIdx.x] = s;
}

template <int MODE>
double run_big(int blocks, int iters, int words, unsigned* dout) {
    const size_t lds = (size_t)words * 4;
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_lds_big<MODE>), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    k_lds_big<MODE><<<blocks, kBlock, lds>>>(dout, 4, words);
    (void)hipEventRecord(a);
    k_lds_big<MODE><<<blocks, kBlock, lds>>>(dout, iters, words);
    (void)hipEventRecord(b);
    (void)hipEventSynchronize(b);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    return (double)blocks * kBlock * kAddr * iters / (ms * 1e-3);
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
    const double r3 = run<3>(blocks, 1024, dout);
    const double r4 = run<4>(blocks, 512, dout);
    const double r5 = run<5>(blocks, 256, dout);
    const double r6 = run<6>(blocks, 256, dout);
    const double r7 = run<7>(blocks, 2048, dout);
    const double r8 = run_big<0>(cus, 2048, 40000, dout);  // sliced, 160 KB, 1 WG/CU
    const double r9 = run_big<0>(cus, 2048, 10000, dout);  // sliced, 40 KB, 1 WG/CU
    const double r10 = run_big<1>(cus, 2048, 40000, dout); // random, 160 KB, 1 WG/CU
    const double r11 = run_big<0>(cus * 2, 2048, 20000, dout);  // sliced, 80 KB, 2 WG/CU
    auto per = [&](double r) { return r / (cus * clk); };
    printf("{\"device\": \"%s\", \"cus\": %d, \"clock_mhz\": %.0f", p.gcnArchName, cus, clk / 1e6);
    const char* names[] = {"conflict_free", "random_2500", "same_address", "stride_2", "stride_16",
                           "stride_32", "stride_64", "lane_sliced", "sliced_160KB_1wg", "sliced_40KB_1wg",
                           "random_160KB_1wg", "sliced_80KB_2wg"};
    const double rates[] = {r0, r1, r2, r3, r4, r5, r6, r7, r8, r9, r10, r11};
    for (int k = 0; k < 12; k++)
        printf(", \"%s\": {\"atomics_per_s\": %.4e, \"per_clk_per_cu\": %.2f}", names[k], rates[k], per(rates[k]));
    printf("}\n");
    (void)hipFree(dout);
    return 0;
}
