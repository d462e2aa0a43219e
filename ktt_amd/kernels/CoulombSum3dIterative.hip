// CoulombSum3dIterative slice kernel for gfx950.
//
// Argument list and semantics of the reference slice kernel
// (Examples/CoulombSum3dIterative/CoulombSum3dIterative.cl:35): the 2D
// kernel plus zIndex, writing slice zIndex of a G^3 z-major grid
// (CoulombSum3dIterative.cl:41, offset Gx * Gy * zIndex).  atomInfo[i].z (or
// atomInfoZ[i]) holds the host-precomputed dz^2 of atom i to slice zIndex
// (CoulombSum3dIterative.cpp:160-179).  "+=" store.
//
// directCoulombSumSlice is the fused alternative: it takes the atoms' real z
// and forms dz^2 on the device, so the per-slice host precompute and upload
// disappear.  Same launch shape.
//
// Launch: block (WORK_GROUP_SIZE_X, WORK_GROUP_SIZE_Y, 1),
//         grid (G / (WORK_GROUP_SIZE_X * POINTS_X), G / (WORK_GROUP_SIZE_Y * ROWS_PER_THREAD), atom blocks).
#include "coulomb_core.h"

using E = kc::DefinedEngine;

namespace
{

template <int MODE>
__device__ __forceinline__ void sliceKernel(float* shared, const kc::AtomSource src, const int numberOfAtoms,
    const float gridSpacing, const int zIndex, float* energyGrid)
{
    const int tid = threadIdx.y * WORK_GROUP_SIZE_X + threadIdx.x;
    const int rowLength = gridDim.x * E::SpanX;
    const int rowsTotal = gridDim.y * E::Rows;
    const int rowBase = blockIdx.y * E::Rows;
    const int x0 = (blockIdx.x * WORK_GROUP_SIZE_X + threadIdx.x) * POINTS_X;
    const int row0 = rowBase + threadIdx.y * ROWS_PER_THREAD;
    const size_t outOffset = static_cast<size_t>(rowLength) * rowsTotal * zIndex;

    int a0, a1;
    kc::atomBlockRange(numberOfAtoms, blockIdx.z, gridDim.z, ATOM_TILE, a0, a1);

    float xLocal[POINTS_X];
#pragma unroll
    for (int i = 0; i < POINTS_X; ++i)
    {
        xLocal[i] = gridSpacing * static_cast<float>(x0 + i);
    }
    float part[ROWS_PER_THREAD][POINTS_X];
    float unused[ROWS_PER_THREAD][POINTS_X];
    E::clear(part);

    // Volume mode over a single plane: row = zIndex * rowsTotal + y.
    const int planeRowBase = (MODE == kc::Volume) ? zIndex * rowsTotal + rowBase : rowBase;
    const int planeRows = (MODE == kc::Volume) ? (zIndex + 1) * rowsTotal : rowsTotal;
    E::accumulate<MODE, false>(shared, src, a0, a1, gridSpacing, tid, planeRowBase, planeRows, rowsTotal, 0.0, xLocal, part,
        unused, unused);

    if (gridDim.z == 1)
    {
        E::store<false>(energyGrid, rowLength, outOffset, rowLength, rowsTotal, x0, row0, part);
    }
    else
    {
        E::store<true>(energyGrid, rowLength, outOffset, rowLength, rowsTotal, x0, row0, part);
    }
}

} // namespace

extern "C" __global__ __launch_bounds__(KC_THREADS) void directCoulombSum(const float4* atomInfo, const float* atomInfoX,
    const float* atomInfoY, const float* atomInfoZ, const float* atomInfoW, int numberOfAtoms, float gridSpacing,
    int zIndex, float* energyGrid)
{
    __shared__ __attribute__((aligned(16))) float shared[E::SharedFloats];
    const kc::AtomSource src{atomInfo, atomInfoX, atomInfoY, atomInfoZ, atomInfoW};
    sliceKernel<kc::Slice>(shared, src, numberOfAtoms, gridSpacing, zIndex, energyGrid);
}

extern "C" __global__ __launch_bounds__(KC_THREADS) void directCoulombSumSlice(const float4* atomInfo,
    const float* atomInfoX, const float* atomInfoY, const float* atomInfoZ, const float* atomInfoW, int numberOfAtoms,
    float gridSpacing, int zIndex, float* energyGrid)
{
    __shared__ __attribute__((aligned(16))) float shared[E::SharedFloats];
    const kc::AtomSource src{atomInfo, atomInfoX, atomInfoY, atomInfoZ, atomInfoW};
    sliceKernel<kc::Volume>(shared, src, numberOfAtoms, gridSpacing, zIndex, energyGrid);
}
