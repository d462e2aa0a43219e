// One point per thread over the whole G^3 grid with the atoms' real z, "+="
// store (Examples/CoulombSum3dIterative/CoulombSum3dIterativeReference.cl:1-26
// semantics; the grid extent is the launch size).
// Launch: block (16, 16, 1), grid (G / 16, G / 16, G).
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

extern "C" __global__ void directCoulombSumReference(const float4* atomInfo, int numberOfAtoms, float gridSpacing,
    float* energyGrid)
{
    const int xIndex = blockIdx.x * blockDim.x + threadIdx.x;
    const int yIndex = blockIdx.y * blockDim.y + threadIdx.y;
    const int zIndex = blockIdx.z * blockDim.z + threadIdx.z;
    const size_t sizeX = static_cast<size_t>(gridDim.x) * blockDim.x;
    const size_t sizeY = static_cast<size_t>(gridDim.y) * blockDim.y;
    const float coordX = gridSpacing * xIndex;
    const float coordY = gridSpacing * yIndex;
    const float coordZ = gridSpacing * zIndex;
    float energyValue = 0.0f;
    for (int i = 0; i < numberOfAtoms; ++i)
    {
        const float4 a = atomInfo[i];
        const float dX = coordX - a.x;
        const float dY = coordY - a.y;
        const float dZ = coordZ - a.z;
        energyValue += a.w * rsqrtf(dX * dX + dY * dY + dZ * dZ);
    }
    energyGrid[(zIndex * sizeY + yIndex) * sizeX + xIndex] += energyValue;
}
