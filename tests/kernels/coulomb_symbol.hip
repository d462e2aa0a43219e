// Test kernel of symbol arguments (ktt::Tuner::AddArgumentSymbol): the atoms are read from a
// __constant__ variable filled by the tuner, not from a kernel parameter. One grid point per
// thread, "+=" store, the semantics of Examples/CoulombSum3d/CoulombSum3dReference.cu:1-28.
// Launch: block (16, 16, 1), grid (G / 16, G / 16, G).
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

#ifndef SYMBOL_ATOMS
#define SYMBOL_ATOMS 256
#endif

__constant__ float4 symbolAtoms[SYMBOL_ATOMS];

extern "C" __global__ void coulombFromSymbol(int numberOfAtoms, float gridSpacing, int gridSize, float* energyGrid)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z;
    if (x >= gridSize || y >= gridSize || z >= gridSize)
    {
        return;
    }
    const float px = gridSpacing * x;
    const float py = gridSpacing * y;
    const float pz = gridSpacing * z;
    float energy = 0.0f;
    for (int i = 0; i < numberOfAtoms && i < SYMBOL_ATOMS; ++i)
    {
        const float4 a = symbolAtoms[i];
        const float dx = px - a.x;
        const float dy = py - a.y;
        const float dz = pz - a.z;
        energy += a.w * rsqrtf(dx * dx + dy * dy + dz * dz);
    }
    energyGrid[(static_cast<size_t>(z) * gridSize + y) * gridSize + x] += energy;
}
