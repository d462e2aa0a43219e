// One point per thread over one slice; atomInfo[i].z holds dz^2 and the
// result is stored with "=" (Examples/CoulombSum2d/CoulombSum2dReference.cl:1-21
// semantics; the row length is the launch width).
// Launch: block (16, 16, 1), grid (Gx / 16, Gy / 16, 1).
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

extern "C" __global__ void directCoulombSumReference(const float4* atomInfo, int numberOfAtoms, float gridSpacing,
    float* energyGrid)
{
    const int xIndex = blockIdx.x * blockDim.x + threadIdx.x;
    const int yIndex = blockIdx.y * blockDim.y + threadIdx.y;
    const int rowLength = gridDim.x * blockDim.x;
    const float coordX = gridSpacing * xIndex;
    const float coordY = gridSpacing * yIndex;
    float energyValue = 0.0f;
    for (int i = 0; i < numberOfAtoms; ++i)
    {
        const float4 a = atomInfo[i];
        const float dX = coordX - a.x;
        const float dY = coordY - a.y;
        energyValue += a.w * rsqrtf(dX * dX + dY * dY + a.z);
    }
    energyGrid[static_cast<size_t>(yIndex) * rowLength + xIndex] = energyValue;
}
