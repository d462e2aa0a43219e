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
// With more than one atom block the plain kernel adds its partial sums with
// FP32 atomics; the compensated kernel merges them in FP64 instead.
#include "coulomb_core.h"

using E = kc::DefinedEngine;

extern "C" __global__ __launch_bounds__(KC_THREADS) void directCoulombSum(const float4* atomInfo, const float* atomInfoX,
    const float* atomInfoY, const float* atomInfoZ, const float* atomInfoW, int numberOfAtoms, float gridSpacing,
    int gridSize, float* energyGrid)
{
    __shared__ __attribute__((aligned(16))) float shared[E::SharedFloats];
    const kc::AtomSource src{atomInfo, atomInfoX, atomInfoY, atomInfoZ, atomInfoW};
    const int tid = threadIdx.y * WORK_GROUP_SIZE_X + threadIdx.x;
    const int rowsTotal = gridSize * gridSize;
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

// Compensated form: tile-relative displacements (x relative to the
// workgroup's first x, formed in FP64), a two-sum per point per atom tile, and
// an FP64 merge of the atom-block partials inside this single launch.
//
// Tiles are numbered blockIdx.y * gridDim.x + blockIdx.x.  Tiles below
// splitFrom take the whole atom range in block 0 and store directly; their
// other blocks exit at once.  Tiles from splitFrom on split the atom range over
// the gridDim.z blocks and meet in partials:
//   partials     float2[gridDim.z * (tiles - splitFrom) * threads * points],
//                (sum, comp) per point.  Unused when gridDim.z == 1.
//   tileCounters uint[tiles], zero before the first launch; every launch
//                leaves them zero again.
// The last workgroup of a split tile to finish (counted with a device-scope
// atomic) sums the partials of all its blocks in FP64 and adds them to
// energyGrid.  splitFrom = 0 splits every tile; splitting only the tiles of the
// last, partial round of workgroups keeps every slot of the GPU busy to the end
// while the other tiles need no partials.
extern "C" __global__ __launch_bounds__(KC_THREADS) void directCoulombSumCompensated(const float4* atomInfo,
    const float* atomInfoX, const float* atomInfoY, const float* atomInfoZ, const float* atomInfoW, int numberOfAtoms,
    float gridSpacing, int gridSize, float* energyGrid, float2* partials, unsigned int* tileCounters, int splitFrom)
{
    __shared__ __attribute__((aligned(16))) float shared[E::SharedFloats];
    const kc::AtomSource src{atomInfo, atomInfoX, atomInfoY, atomInfoZ, atomInfoW};
    const int tid = threadIdx.y * WORK_GROUP_SIZE_X + threadIdx.x;
    const int rowsTotal = gridSize * gridSize;
    const int rowBase = blockIdx.y * E::Rows;
    const int xBase = blockIdx.x * E::SpanX;
    const int xFirst = threadIdx.x * POINTS_X;
    const int x0 = xBase + xFirst;
    const int row0 = rowBase + threadIdx.y * ROWS_PER_THREAD;
    const size_t tile = static_cast<size_t>(blockIdx.y) * gridDim.x + blockIdx.x;
    const bool split = gridDim.z > 1 && tile >= static_cast<size_t>(splitFrom);
    if (gridDim.z > 1 && !split && blockIdx.z != 0)
    {
        return;  // a whole tile: block 0 takes every atom
    }

    int a0 = 0;
    int a1 = numberOfAtoms;
    if (split)
    {
        kc::atomBlockRange(numberOfAtoms, blockIdx.z, gridDim.z, ATOM_TILE, a0, a1);
    }

    const double h = static_cast<double>(gridSpacing);
    const double originX = h * xBase;
    float xLocal[POINTS_X];
#pragma unroll
    for (int i = 0; i < POINTS_X; ++i)
    {
        xLocal[i] = static_cast<float>(h * (xFirst + i));
    }

    float part[ROWS_PER_THREAD][POINTS_X];
    float sum[ROWS_PER_THREAD][POINTS_X];
    float comp[ROWS_PER_THREAD][POINTS_X];
    E::clear(part);
    E::clear(sum);
    E::clear(comp);

    E::accumulate<kc::Volume, true>(shared, src, a0, a1, gridSpacing, tid, rowBase, rowsTotal, gridSize, originX, xLocal, part,
        sum, comp);

    if (!split)
    {
        E::storeCompensated(energyGrid, gridSize, 0, gridSize, rowsTotal, x0, row0, sum, comp);
        return;
    }

    // Partials of the split tiles laid out [block][split tile][point-in-thread][thread]
    // so that the lanes of a wave store consecutive float2s.
    const size_t splitTiles = static_cast<size_t>(gridDim.x) * gridDim.y - static_cast<size_t>(splitFrom);
    const size_t slot = tile - static_cast<size_t>(splitFrom);
    const size_t tileStride = static_cast<size_t>(E::Threads) * E::Points;
    float2* mine = partials + (static_cast<size_t>(blockIdx.z) * splitTiles + slot) * tileStride;
#pragma unroll
    for (int k = 0; k < ROWS_PER_THREAD; ++k)
    {
#pragma unroll
        for (int i = 0; i < POINTS_X; ++i)
        {
            mine[(k * POINTS_X + i) * E::Threads + tid] = make_float2(sum[k][i], comp[k][i]);
        }
    }

    __shared__ unsigned int sLast;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (tid == 0)
    {
        const unsigned int arrived = __hip_atomic_fetch_add(&tileCounters[tile], 1u, __ATOMIC_ACQ_REL,
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
    for (int k = 0; k < ROWS_PER_THREAD; ++k)
    {
        const int row = row0 + k;
#pragma unroll
        for (int i = 0; i < POINTS_X; ++i)
        {
            double total = 0.0;
            for (unsigned int b = 0; b < gridDim.z; ++b)
            {
                const float2 sc = partials[(static_cast<size_t>(b) * splitTiles + slot) * tileStride
                    + (k * POINTS_X + i) * E::Threads + tid];
                total += static_cast<double>(sc.x) + static_cast<double>(sc.y);
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
        __hip_atomic_store(&tileCounters[tile], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}
