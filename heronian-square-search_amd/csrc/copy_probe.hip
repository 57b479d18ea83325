// copy_probe.hip -- a plain device-to-device copy of a known byte count.
//
// Used only to calibrate the FETCH_SIZE / WRITE_SIZE hardware counters that
// rocprofv3 --pmc reports for the search kernels: the copy reads and writes
// exactly `bytes` bytes, so counter / bytes is the scale factor.
//
//   copy_probe [BYTES]      (default 1 GiB; prints bytes, milliseconds, GB/s)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copy_kernel(const uint4* __restrict__ in, uint4* __restrict__ out, size_t n) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = in[i];
}

#define CHECK(x)                                                                  \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) {                                                   \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));          \
            return 1;                                                             \
        }                                                                         \
    } while (0)

int main(int argc, char** argv) {
    size_t bytes = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : (size_t)1 << 30;
    bytes &= ~(size_t)15;
    size_t n = bytes / 16;
    uint4 *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMemset(a, 1, bytes));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    CHECK(hipEventRecord(e0));
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, 0, a, b, n);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    std::printf("copy bytes %zu ms %.3f GB/s %.1f (read+write)\n", bytes, ms, 2.0 * bytes / ms / 1e6);
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
