// copy_probe.hip -- calibration kernel for HBM byte counters.
//
// Copies N bytes (default 256 MiB) from one device buffer to another with a
// plain float4 grid-stride kernel, 5 launches.  Under `rocprofv3 --pmc
// FETCH_SIZE` / `WRITE_SIZE` the per-dispatch values of this kernel are the
// counters' reading for exactly N bytes read and N bytes written, which
// scales the readings of the render kernels into bytes.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>

__global__ void copy_f4(const float4* __restrict__ src, float4* __restrict__ dst, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int main(int argc, char** argv)
{
    size_t bytes = argc > 1 ? strtoull(argv[1], 0, 10) : (256ull << 20);
    bytes &= ~(size_t)15;
    const size_t n = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        fprintf(stderr, "hipMalloc failed\n");
        return 1;
    }
    if (hipMemset(a, 1, bytes) != hipSuccess) return 1;
    for (int it = 0; it < 5; ++it)
        hipLaunchKernelGGL(copy_f4, dim3(4096), dim3(256), 0, 0, a, b, n);
    if (hipDeviceSynchronize() != hipSuccess) {
        fprintf(stderr, "copy failed\n");
        return 1;
    }
    printf("copy_probe: 5 x %zu bytes copied\n", bytes);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
