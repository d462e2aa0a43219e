// One grid point per thread, FP32, "+=" store: the validation kernel of the
// 3D example (Examples/CoulombSum3d/CoulombSum3dReference.cu:1-28 semantics:
// same arguments, coordinates h * index, rsqrt, accumulation into the grid).
// The reference's entry guard tests x twice (CoulombSum3dReference.cu:7); this
// kernel tests x, y and z.
// Launch: block (16, 16, 1), grid (G / 16, G / 16, G).
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

extern "C" __global__ void directCoulombSumReference(const float4* atomInfo, int numberOfAtoms, float gridSpacing,
    int gridSize, float* energyGrid)
{
    const int xIndex = blockIdx.x * blockDim.x + threadIdx.x;
    const int yIndex = blockIdx.y * blockDim.y + threadIdx.y;
    const int zIndex = blockIdx.z * blockDim.z + threadIdx.z;
    if (xIndex >= gridSize || yIndex >= gridSize || zIndex >= gridSize)
    {
        return;
    }

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
    const size_t outIndex = (static_cast<size_t>(zIndex) * gridSize + yIndex) * gridSize + xIndex;
    energyGrid[outIndex] += energyValue;
}
