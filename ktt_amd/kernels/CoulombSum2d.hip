// CoulombSum2d for gfx950: one slice of the grid.
//
// Argument list and semantics of the reference 2D kernel
// (Examples/CoulombSum2d/CoulombSum2d.cl:35): atomInfo[i].z (or atomInfoZ[i])
// holds the precomputed dz^2 of atom i to the slice, not its z
// (CoulombSum2d.cl:126); the slice is row-major with rows of length
// gridDim.x * WORK_GROUP_SIZE_X * POINTS_X (CoulombSum2d.cl:38-41 derives the
// row length from the launch size the same way); results are added with
// "+=" (CoulombSum2d.cl:945).
//
// Launch: block (WORK_GROUP_SIZE_X, WORK_GROUP_SIZE_Y, 1),
//         grid (Gx / (WORK_GROUP_SIZE_X * POINTS_X), Gy / (WORK_GROUP_SIZE_Y * ROWS_PER_THREAD), atom blocks).
// Several atom blocks add their partial sums with FP32 atomics.
#include "coulomb_core.h"

using E = kc::DefinedEngine;

extern "C" __global__ __launch_bounds__(KC_THREADS) void directCoulombSum(const float4* atomInfo, const float* atomInfoX,
    const float* atomInfoY, const float* atomInfoZ, const float* atomInfoW, int numberOfAtoms, float gridSpacing,
    float* energyGrid)
{
    __shared__ __attribute__((aligned(16))) float shared[E::SharedFloats];
    const kc::AtomSource src{atomInfo, atomInfoX, atomInfoY, atomInfoZ, atomInfoW};
    const int tid = threadIdx.y * WORK_GROUP_SIZE_X + threadIdx.x;
    const int rowLength = gridDim.x * E::SpanX;
    const int rowsTotal = gridDim.y * E::Rows;
    const int rowBase = blockIdx.y * E::Rows;
    const int x0 = (blockIdx.x * WORK_GROUP_SIZE_X + threadIdx.x) * POINTS_X;
    const int row0 = rowBase + threadIdx.y * ROWS_PER_THREAD;

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

    E::accumulate<kc::Slice, false>(shared, src, a0, a1, gridSpacing, tid, rowBase, rowsTotal, rowsTotal, 0.0, xLocal, part,
        unused, unused);

    if (gridDim.z == 1)
    {
        E::store<false>(energyGrid, rowLength, 0, rowLength, rowsTotal, x0, row0, part);
    }
    else
    {
        E::store<true>(energyGrid, rowLength, 0, rowLength, rowsTotal, x0, row0, part);
    }
}
