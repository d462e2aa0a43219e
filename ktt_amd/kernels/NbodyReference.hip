// Reference N-body step: one body per thread, every source read from global memory, j in
// ascending order; the counterpart of Examples/Nbody/NbodyReference.cu:40-75 (same kernel name
// and arguments, float4 positions (x, y, z, mass) and velocities (vx, vy, vz, 0)). It validates
// the tuned kernel inside the tuner; n is the launch's total thread count.

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

extern "C" __global__ void nbody_kernel_reference(float timeDelta, const float4* __restrict__ oldBodyInfo,
    float4* __restrict__ newBodyInfo, const float4* __restrict__ oldVel, float4* __restrict__ newVel, float damping,
    float softeningSqr)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = blockDim.x * gridDim.x;
    const float4 p = oldBodyInfo[i];
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    for (int j = 0; j < n; ++j)
    {
        const float4 q = oldBodyInfo[j];
        const float dx = q.x - p.x;
        const float dy = q.y - p.y;
        const float dz = q.z - p.z;
        const float invr = rsqrtf(dx * dx + dy * dy + dz * dz + softeningSqr);
        const float f = q.w * invr * invr * invr;
        ax += f * dx;
        ay += f * dy;
        az += f * dz;
    }
    const float4 v = oldVel[i];
    const float dtdt = damping * timeDelta * timeDelta;
    newBodyInfo[i] = make_float4(p.x + (timeDelta * v.x + dtdt * ax), p.y + (timeDelta * v.y + dtdt * ay),
        p.z + (timeDelta * v.z + dtdt * az), p.w);
    newVel[i] = make_float4(v.x + timeDelta * ax, v.y + timeDelta * ay, v.z + timeDelta * az, v.w);
}
