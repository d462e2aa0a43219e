// Batched direct Coulomb summation: many independent systems (own atoms, own
// cubic grid, own spacing) in one persistent launch driven by a device-side
// work queue.  Compensated FP32 form (tile-relative x, two-sum per atom tile,
// FP64 store), as directCoulombSumCompensated with one atom block.
//
// Work items are (system, x block, row block) tiles listed by the host in
// decreasing order of cost (longest-processing-time first).  Every workgroup
// takes the next item with one device-scope atomic and exits when the queue is
// empty, so all workgroups finish once the last item has been handed out.  The
// counter is cleared by coulombBatchedReset, launched just before the
// persistent kernel (the host captures both in one HIP graph).
//
// Systems are grouped into up to five tile-shape buckets (by grid side);
// bucket b uses BUCKETb_WGX x (BATCH_THREADS / BUCKETb_WGX) threads,
// BUCKETb_PX x BUCKETb_ROWS points per thread, BUCKETb_TILE atoms per LDS
// tile and an unroll of BUCKETb_UNROLL atoms, chosen per bucket by the autotuner.
#include "coulomb_core.h"

#ifndef BATCH_THREADS
#define BATCH_THREADS 256
#endif
#ifndef BUCKET0_WGX
#define BUCKET0_WGX 32
#endif
#ifndef BUCKET0_PX
#define BUCKET0_PX 1
#endif
#ifndef BUCKET0_ROWS
#define BUCKET0_ROWS 4
#endif
#ifndef BUCKET0_TILE
#define BUCKET0_TILE 64
#endif
#ifndef BUCKET1_WGX
#define BUCKET1_WGX 16
#endif
#ifndef BUCKET1_PX
#define BUCKET1_PX 3
#endif
#ifndef BUCKET1_ROWS
#define BUCKET1_ROWS 4
#endif
#ifndef BUCKET1_TILE
#define BUCKET1_TILE 64
#endif
#ifndef BUCKET2_WGX
#define BUCKET2_WGX 32
#endif
#ifndef BUCKET2_PX
#define BUCKET2_PX 2
#endif
#ifndef BUCKET2_ROWS
#define BUCKET2_ROWS 8
#endif
#ifndef BUCKET2_TILE
#define BUCKET2_TILE 64
#endif
#ifndef BUCKET3_WGX
#define BUCKET3_WGX 16
#endif
#ifndef BUCKET3_PX
#define BUCKET3_PX 3
#endif
#ifndef BUCKET3_ROWS
#define BUCKET3_ROWS 4
#endif
#ifndef BUCKET3_TILE
#define BUCKET3_TILE 128
#endif
#ifndef BUCKET4_WGX
#define BUCKET4_WGX 64
#endif
#ifndef BUCKET4_PX
#define BUCKET4_PX 2
#endif
#ifndef BUCKET4_ROWS
#define BUCKET4_ROWS 8
#endif
#ifndef BUCKET4_TILE
#define BUCKET4_TILE 128
#endif
#ifndef BUCKET0_UNROLL
#define BUCKET0_UNROLL 8
#endif
#ifndef BUCKET1_UNROLL
#define BUCKET1_UNROLL 8
#endif
#ifndef BUCKET2_UNROLL
#define BUCKET2_UNROLL 8
#endif
#ifndef BUCKET3_UNROLL
#define BUCKET3_UNROLL 16
#endif
#ifndef BUCKET4_UNROLL
#define BUCKET4_UNROLL 8
#endif

// One system of the batch (32 bytes; the host writes the same layout).
struct BatchSystem
{
    int atomOffset;        // first atom in the packed atom array
    int atomCount;
    int gridSize;          // cubic grid side G
    int bucket;            // tile-shape bucket, 0..4
    float spacing;
    int reserved;
    long long gridOffset;  // first float of the system's grid in the packed output
};

namespace
{

using E0 = kc::Engine<BUCKET0_WGX, BATCH_THREADS / BUCKET0_WGX, BUCKET0_PX, BUCKET0_ROWS, BUCKET0_TILE, BUCKET0_UNROLL>;
using E1 = kc::Engine<BUCKET1_WGX, BATCH_THREADS / BUCKET1_WGX, BUCKET1_PX, BUCKET1_ROWS, BUCKET1_TILE, BUCKET1_UNROLL>;
using E2 = kc::Engine<BUCKET2_WGX, BATCH_THREADS / BUCKET2_WGX, BUCKET2_PX, BUCKET2_ROWS, BUCKET2_TILE, BUCKET2_UNROLL>;
using E3 = kc::Engine<BUCKET3_WGX, BATCH_THREADS / BUCKET3_WGX, BUCKET3_PX, BUCKET3_ROWS, BUCKET3_TILE, BUCKET3_UNROLL>;
using E4 = kc::Engine<BUCKET4_WGX, BATCH_THREADS / BUCKET4_WGX, BUCKET4_PX, BUCKET4_ROWS, BUCKET4_TILE, BUCKET4_UNROLL>;

static_assert(E0::Threads == BATCH_THREADS && E1::Threads == BATCH_THREADS && E2::Threads == BATCH_THREADS
    && E3::Threads == BATCH_THREADS && E4::Threads == BATCH_THREADS, "every bucket shape must use BATCH_THREADS threads");

constexpr int MaxOf(const int a, const int b)
{
    return a > b ? a : b;
}

constexpr int kSharedFloats = MaxOf(MaxOf(MaxOf(E0::SharedFloats, E1::SharedFloats), MaxOf(E2::SharedFloats, E3::SharedFloats)),
    E4::SharedFloats);

// One (x block, row block) tile of one system with engine shape E.
template <typename E>
__device__ __forceinline__ void processTile(float* shared, const float4* atoms, const BatchSystem& system, const int xBase,
    const int rowBase, float* grids)
{
    const int tid = threadIdx.x;
    const int G = system.gridSize;
    const int rowsTotal = G * G;
    const int xFirst = (tid % E::LanesX) * E::PointsX;
    const int x0 = xBase + xFirst;
    const int row0 = rowBase + (tid / E::LanesX) * E::RowsPerThread;
    const double h = static_cast<double>(system.spacing);
    const double originX = h * xBase;
    const kc::AtomSource src{atoms + system.atomOffset, nullptr, nullptr, nullptr, nullptr};

    float xLocal[E::PointsX];
#pragma unroll
    for (int i = 0; i < E::PointsX; ++i)
    {
        xLocal[i] = static_cast<float>(h * (xFirst + i));
    }
    float part[E::RowsPerThread][E::PointsX];
    float sum[E::RowsPerThread][E::PointsX];
    float comp[E::RowsPerThread][E::PointsX];
    E::clear(part);
    E::clear(sum);
    E::clear(comp);

    E::template accumulate<kc::Volume, true>(shared, src, 0, system.atomCount, system.spacing, tid, rowBase, rowsTotal, G,
        originX, xLocal, part, sum, comp);
    E::storeCompensated(grids, G, static_cast<size_t>(system.gridOffset), G, rowsTotal, x0, row0, sum, comp);
}

__device__ __forceinline__ void dispatchTile(float* shared, const float4* atoms, const BatchSystem& system, const int xBase,
    const int rowBase, float* grids)
{
    switch (system.bucket)
    {
    case 0: processTile<E0>(shared, atoms, system, xBase, rowBase, grids); break;
    case 1: processTile<E1>(shared, atoms, system, xBase, rowBase, grids); break;
    case 2: processTile<E2>(shared, atoms, system, xBase, rowBase, grids); break;
    case 3: processTile<E3>(shared, atoms, system, xBase, rowBase, grids); break;
    default: processTile<E4>(shared, atoms, system, xBase, rowBase, grids); break;
    }
}

} // namespace

extern "C" __global__ void coulombBatchedReset(unsigned int* queueCounter)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
    {
        *queueCounter = 0u;
    }
}

// items: int4 (system, xBase, rowBase, unused) in processing order.
extern "C" __global__ __launch_bounds__(BATCH_THREADS) void coulombBatchedPersistent(const float4* atoms,
    const BatchSystem* systems, const int4* items, int itemCount, float* grids, unsigned int* queueCounter)
{
    __shared__ __attribute__((aligned(16))) float shared[kSharedFloats];
    __shared__ int sItem;
    while (true)
    {
        if (threadIdx.x == 0)
        {
            sItem = static_cast<int>(atomicAdd(queueCounter, 1u));
        }
        __syncthreads();
        const int item = sItem;
        __syncthreads();
        if (item >= itemCount)
        {
            break;
        }
        const int4 work = items[item];
        const BatchSystem system = systems[work.x];
        dispatchTile(shared, atoms, system, work.y, work.z, grids);
    }
}

// One system with bucket-0 shape, one workgroup per tile: the unit the autotuner
// times for a bucket.  Launch: grid (ceil(G / span), ceil(G*G / rows)), block BATCH_THREADS.
extern "C" __global__ __launch_bounds__(BATCH_THREADS) void coulombBatchedSystem(const float4* atoms, int numberOfAtoms,
    float gridSpacing, int gridSize, float* energyGrid)
{
    __shared__ __attribute__((aligned(16))) float shared[E0::SharedFloats];
    BatchSystem system;
    system.atomOffset = 0;
    system.atomCount = numberOfAtoms;
    system.gridSize = gridSize;
    system.bucket = 0;
    system.spacing = gridSpacing;
    system.reserved = 0;
    system.gridOffset = 0;
    processTile<E0>(shared, atoms, system, blockIdx.x * E0::SpanX, blockIdx.y * E0::Rows, energyGrid);
}
