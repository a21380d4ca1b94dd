// Calibration for rocprofv3 byte counters: a plain device-to-device copy of a known size.
// FETCH_SIZE / WRITE_SIZE of `calib_copy_kernel` divided by the bytes it must move gives the factor
// that turns the counters of the engine's kernels into HBM bytes.
//   calib_copy [MiB]   (default 1024): copies MiB mebibytes 5 times, prints the expected KiB per launch.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void __launch_bounds__(256) calib_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                         size_t n) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

int main(int argc, char** argv) {
    const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1024;
    const size_t bytes = mib << 20;
    const size_t n = bytes / sizeof(uint4);
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        std::fprintf(stderr, "allocation failed\n");
        return 1;
    }
    (void)hipMemset(a, 1, bytes);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    float best = 1e30f;
    for (int r = 0; r < 5; ++r) {
        (void)hipEventRecord(e0);
        hipLaunchKernelGGL(calib_copy_kernel, dim3(8192), dim3(256), 0, 0, a, b, n);
        (void)hipEventRecord(e1);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        best = ms < best ? ms : best;
    }
    std::printf("copy %zu MiB: read %zu KiB + write %zu KiB per launch, best %.3f ms = %.1f GB/s\n", mib, bytes >> 10,
                bytes >> 10, best, 2.0 * bytes / (best * 1e-3) / 1e9);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
