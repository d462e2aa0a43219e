// Utility kernels of the Coulomb library: zero fill (used instead of hipMemset so that the
// clearing step can live inside a captured HIP graph) and the per-slice dz^2 update of the
// iterative form done on the device.
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

extern "C" __global__ void kttZeroFill16(uint4* data, unsigned long long count)
{
    const unsigned long long stride = static_cast<unsigned long long>(gridDim.x) * blockDim.x;
    for (unsigned long long i = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
    {
        data[i] = make_uint4(0u, 0u, 0u, 0u);
    }
}

// Plain 16-byte copy; its HBM counters calibrate the profiler's FETCH_SIZE / WRITE_SIZE.
extern "C" __global__ void kttCopy16(uint4* destination, const uint4* source, unsigned long long count)
{
    const unsigned long long stride = static_cast<unsigned long long>(gridDim.x) * blockDim.x;
    for (unsigned long long i = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
    {
        destination[i] = source[i];
    }
}

extern "C" __global__ void kttZeroFill4(unsigned int* data, unsigned long long count)
{
    const unsigned long long stride = static_cast<unsigned long long>(gridDim.x) * blockDim.x;
    for (unsigned long long i = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
    {
        data[i] = 0u;
    }
}
