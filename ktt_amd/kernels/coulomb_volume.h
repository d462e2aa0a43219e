// Whole-volume compensated Coulomb summation over one tile shape, shared by the
// hipRTC-compiled CoulombSum3d.hip kernel (shape from the -D defines) and the
// library's precompiled kernels (fixed shapes).
//
// Grid layout and coordinates follow Examples/CoulombSum3d/CoulombSum3d.cu:10-15
// (index z*G*G + y*G + x, point (h*x, h*y, h*z)); the store adds into the grid
// (CoulombSum3d.cu:46).
//
// Compensated form: tile-relative displacements (x relative to the
// workgroup's first x, formed in FP64), a two-sum per point per atom tile, and
// an FP64 merge of the atom-block partials inside the launch.
//
// Tiles are numbered row group * x groups + x group (kc::volumeGroup; without a
// fold of the row groups that is blockIdx.y * gridDim.x + blockIdx.x).  Tiles below
// splitFrom take the whole atom range in block 0 and store directly; their
// other blocks exit at once.  Tiles from splitFrom on split the atom range over
// the gridDim.z blocks and meet in partials:
//   partials     float2[gridDim.z * (tiles - splitFrom) * threads * points],
//                (sum, comp) per point.  Unused when gridDim.z == 1.
//   tileCounters uint[tiles - splitFrom], one per split tile, zero before the
//                launch; every launch leaves them zero again.
// The last workgroup of a split tile to finish (counted with a device-scope
// atomic) sums the partials of all its blocks in FP64 and adds them to
// energyGrid.  splitFrom = 0 splits every tile; splitting only the tiles of the
// last, partial round of workgroups keeps every slot of the GPU busy to the end
// while the other tiles need no partials.
#pragma once

#include "coulomb_core.h"

namespace kc
{

template <class E>
__device__ __forceinline__ void volumeCompensated(const AtomSource src, const int numberOfAtoms, const float gridSpacing,
    const int gridSize, float* energyGrid, float2* partials, unsigned int* tileCounters, const int splitFrom)
{
    __shared__ __attribute__((aligned(16))) float shared[E::SharedFloats];
    __shared__ unsigned int sLast;
    constexpr int PX = E::PointsX;
    constexpr int ROWS = E::RowsPerThread;
    const VolumeGroup group = volumeGroup(gridSize, E::SpanX, E::Rows);
    if (!group.valid)
    {
        return;
    }
    const int tid = threadIdx.y * E::LanesX + threadIdx.x;
    const int rowsTotal = gridSize * gridSize;
    const int rowBase = group.row * E::Rows;
    const int xBase = group.x * E::SpanX;
    const int xFirst = threadIdx.x * PX;
    const int x0 = xBase + xFirst;
    const int row0 = rowBase + threadIdx.y * ROWS;
    const size_t tiles = static_cast<size_t>(group.groupsX) * static_cast<size_t>(group.groupsY);
    const size_t tile = static_cast<size_t>(group.row) * group.groupsX + group.x;
    const bool split = gridDim.z > 1 && tile >= static_cast<size_t>(splitFrom);
    if (gridDim.z > 1 && !split && blockIdx.z != 0)
    {
        return;  // a whole tile: block 0 takes every atom
    }

    int a0 = 0;
    int a1 = numberOfAtoms;
    if (split)
    {
        atomBlockRange(numberOfAtoms, blockIdx.z, gridDim.z, E::Tile, a0, a1);
    }

    const double h = static_cast<double>(gridSpacing);
    const double originX = h * xBase;
    float xLocal[PX];
#pragma unroll
    for (int i = 0; i < PX; ++i)
    {
        xLocal[i] = static_cast<float>(h * (xFirst + i));
    }

    float part[ROWS][PX];
    float sum[ROWS][PX];
    float comp[ROWS][PX];
    E::clear(part);
    E::clear(sum);
    E::clear(comp);

    E::template accumulate<Volume, true>(shared, src, a0, a1, gridSpacing, tid, rowBase, rowsTotal, gridSize, originX, xLocal,
        part, sum, comp);

    if (!split)
    {
        E::storeCompensated(energyGrid, gridSize, 0, gridSize, rowsTotal, x0, row0, sum, comp);
        return;
    }

    // Partials of the split tiles laid out [block][split tile][point-in-thread][thread]
    // so that the lanes of a wave store consecutive float2s.
    const size_t splitTiles = tiles - static_cast<size_t>(splitFrom);
    const size_t slot = tile - static_cast<size_t>(splitFrom);
    const size_t tileStride = static_cast<size_t>(E::Threads) * E::Points;
    float2* mine = partials + (static_cast<size_t>(blockIdx.z) * splitTiles + slot) * tileStride;
#pragma unroll
    for (int k = 0; k < ROWS; ++k)
    {
#pragma unroll
        for (int i = 0; i < PX; ++i)
        {
            mine[(k * PX + i) * E::Threads + tid] = make_float2(sum[k][i], comp[k][i]);
        }
    }

    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (tid == 0)
    {
        const unsigned int arrived = __hip_atomic_fetch_add(&tileCounters[slot], 1u, __ATOMIC_ACQ_REL,
            __HIP_MEMORY_SCOPE_AGENT);
        sLast = (arrived + 1 == gridDim.z) ? 1u : 0u;
    }
    __syncthreads();
    if (sLast == 0u)
    {
        return;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");

#pragma unroll
    for (int k = 0; k < ROWS; ++k)
    {
        const int row = row0 + k;
#pragma unroll
        for (int i = 0; i < PX; ++i)
        {
            double total = 0.0;
            for (unsigned int b = 0; b < gridDim.z; ++b)
            {
                const float2 sc = partials[(static_cast<size_t>(b) * splitTiles + slot) * tileStride
                    + (k * PX + i) * E::Threads + tid];
                total += addPair(-0.0, sc.x, sc.y);  // -0.0 + x == x for every x
            }
            if (row < rowsTotal && x0 + i < gridSize)
            {
                float* p = energyGrid + static_cast<size_t>(row) * gridSize + x0 + i;
                *p = static_cast<float>(static_cast<double>(*p) + total);
            }
        }
    }
    if (tid == 0)
    {
        __hip_atomic_store(&tileCounters[slot], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

} // namespace kc
