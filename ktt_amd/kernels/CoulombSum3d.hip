// CoulombSum3d for gfx950: the whole 256^3-class grid in one launch.
//
// directCoulombSum keeps the argument list of the reference kernel
// (Examples/CoulombSum3d/CoulombSum3d.cu:1), its z-major layout
// (index z*G*G + y*G + x, CoulombSum3d.cu:10-11), its grid coordinates
// (h * index, CoulombSum3d.cu:13-15) and its "+=" store (CoulombSum3d.cu:46).
// The computation is the row/tile engine of coulomb_core.h.
//
// Launch (CUDA-style grid of workgroups):
//   block = (WORK_GROUP_SIZE_X, WORK_GROUP_SIZE_Y, 1)
//   grid  = (ceil(G / (WORK_GROUP_SIZE_X * POINTS_X)),
//            ceil(G * G / (WORK_GROUP_SIZE_Y * ROWS_PER_THREAD)),
//            atom blocks)
// with the row groups folded into x when there are more than 65,535 of them
// (kc::volumeGroup). With more than one atom block the plain kernel adds its
// partial sums with FP32 atomics; the compensated kernel merges them in FP64.
#include "coulomb_volume.h"

using E = kc::DefinedEngine;

extern "C" __global__ __launch_bounds__(KC_THREADS) void directCoulombSum(const float4* atomInfo, const float* atomInfoX,
    const float* atomInfoY, const float* atomInfoZ, const float* atomInfoW, int numberOfAtoms, float gridSpacing,
    int gridSize, float* energyGrid)
{
    __shared__ __attribute__((aligned(16))) float shared[E::SharedFloats];
    const kc::VolumeGroup group = kc::volumeGroup(gridSize, E::SpanX, E::Rows);
    if (!group.valid)
    {
        return;
    }
    const kc::AtomSource src{atomInfo, atomInfoX, atomInfoY, atomInfoZ, atomInfoW};
    const int tid = threadIdx.y * WORK_GROUP_SIZE_X + threadIdx.x;
    const int rowsTotal = gridSize * gridSize;
    const int rowBase = group.row * E::Rows;
    const int x0 = (group.x * WORK_GROUP_SIZE_X + threadIdx.x) * POINTS_X;
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

    E::accumulate<kc::Volume, false>(shared, src, a0, a1, gridSpacing, tid, rowBase, rowsTotal, gridSize, 0.0, xLocal, part,
        unused, unused);

    if (gridDim.z == 1)
    {
        E::store<false>(energyGrid, gridSize, 0, gridSize, rowsTotal, x0, row0, part);
    }
    else
    {
        E::store<true>(energyGrid, gridSize, 0, gridSize, rowsTotal, x0, row0, part);
    }
}

// Compensated form (tile-relative displacements, a two-sum per point per atom
// tile, FP64 merge of atom-block partials in this launch): the body is
// kc::volumeCompensated in coulomb_volume.h, which documents the partials and
// per-split-tile counters.
extern "C" __global__ __launch_bounds__(KC_THREADS) void directCoulombSumCompensated(const float4* atomInfo,
    const float* atomInfoX, const float* atomInfoY, const float* atomInfoZ, const float* atomInfoW, int numberOfAtoms,
    float gridSpacing, int gridSize, float* energyGrid, float2* partials, unsigned int* tileCounters, int splitFrom)
{
    const kc::AtomSource src{atomInfo, atomInfoX, atomInfoY, atomInfoZ, atomInfoW};
    kc::volumeCompensated<E>(src, numberOfAtoms, gridSpacing, gridSize, energyGrid, partials, tileCounters, splitFrom);
}
