// CoulombSum3d for the reference's own tuning space, on gfx950.
//
// Kernel name, arguments, tuning parameters and launch geometry are those of
// Examples/CoulombSum3d/CoulombSum3d.cu:1-47 (identical to Examples/JSON/CoulombSum3d.cu), so
// the reference's example space (Examples/CoulombSum3d/CoulombSum3d.cpp:139-209) and its JSON
// script (Examples/JSON/CoulombSum3d.json) run on ComputeApi::HIP with this file as the kernel:
//
//   launch   block (WORK_GROUP_SIZE_X, WORK_GROUP_SIZE_Y, 1),
//            grid (ceil(G / WORK_GROUP_SIZE_X), ceil(G / WORK_GROUP_SIZE_Y), ceil(G / Z_ITERATIONS))
//   thread   grid points (x, y, z0 .. z0 + Z_ITERATIONS - 1), z0 = blockIdx.z * Z_ITERATIONS,
//            coordinates h * index, V added into the grid ("+=", index z*G*G + y*G + x)
//   WORK_GROUP_SIZE_Z    1 (the only value the reference's spaces use)
//   Z_ITERATIONS         z points per thread
//   INNER_UNROLL_FACTOR  unroll of the loops over a thread's z points (0: the compiler's choice)
//   USE_CONSTANT_MEMORY  0: atoms staged through LDS in tiles of 256, read by every lane of a wave
//                        at one address; 1: atoms read from memory at a wave-uniform index (scalar
//                        loads through the scalar data cache, gfx950's constant cache)
//   USE_SOA              0: float4 atoms {x, y, z, q}; 1: the X, Y, Z, W arrays
//   VECTOR_SIZE          atoms each thread stages per step (the reference's OpenCL space varies
//                        it with USE_SOA = 1; CUDA's holds 1)
//
// The body is this library's own. The z points of a thread share one bracket per atom,
// B = (x - X)^2 + (y - Y)^2, so an interaction costs a subtraction, an FMA, a v_rsq_f32 and an
// FMA. Per atom a wave runs its FMAs at priority 0 and then its v_rsq_f32 at priority 3 (the
// issue phases of coulomb_core.h, so that two waves' FMAs issue in one quad-cycle). The z
// displacement is h * index - Z for every point (the reference steps dZ += h, which adds a
// rounding per step). Lanes past the grid's edge stage atoms with the others and store nothing.
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

#ifndef WORK_GROUP_SIZE_X
#define WORK_GROUP_SIZE_X 32
#endif
#ifndef WORK_GROUP_SIZE_Y
#define WORK_GROUP_SIZE_Y 2
#endif
#ifndef WORK_GROUP_SIZE_Z
#define WORK_GROUP_SIZE_Z 1
#endif
#ifndef Z_ITERATIONS
#define Z_ITERATIONS 8
#endif
#ifndef INNER_UNROLL_FACTOR
#define INNER_UNROLL_FACTOR 0
#endif
#ifndef USE_CONSTANT_MEMORY
#define USE_CONSTANT_MEMORY 0
#endif
#ifndef USE_SOA
#define USE_SOA 0
#endif
#ifndef VECTOR_SIZE
#define VECTOR_SIZE 1
#endif

#define KZ_THREADS (WORK_GROUP_SIZE_X * WORK_GROUP_SIZE_Y)
#define KZ_TILE 256

static_assert(WORK_GROUP_SIZE_Z == 1, "one thread in z per workgroup");
static_assert(Z_ITERATIONS >= 1, "at least one z point per thread");
static_assert(VECTOR_SIZE >= 1 && KZ_TILE % VECTOR_SIZE == 0, "VECTOR_SIZE divides the atom tile");

#if INNER_UNROLL_FACTOR > 0
#define KZ_UNROLL_Z _Pragma("unroll INNER_UNROLL_FACTOR")
#else
#define KZ_UNROLL_Z
#endif

namespace
{

// Far away with no charge: rsqrt stays finite and the term is exactly 0.
constexpr float kPadCoord = -1.0e6f;

__device__ __forceinline__ float4 loadAtom(const float4* __restrict__ atomInfo, const float* __restrict__ atomInfoX,
    const float* __restrict__ atomInfoY, const float* __restrict__ atomInfoZ, const float* __restrict__ atomInfoW, const int i)
{
#if USE_SOA
    return make_float4(atomInfoX[i], atomInfoY[i], atomInfoZ[i], atomInfoW[i]);
#else
    return atomInfo[i];
#endif
}

// One atom against the thread's z points, in issue phases: finish the previous atom (its r^-1
// in rr, its charge qPrev), form this atom's r^2 in rr, then take every r^-1/2 at priority 3.
__device__ __forceinline__ void phasedAtom(const float4 a, const float cx, const float cy, const float (&hz)[Z_ITERATIONS],
    float (&acc)[Z_ITERATIONS], float (&rr)[Z_ITERATIONS], float& qPrev)
{
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_setprio(0);
    KZ_UNROLL_Z
    for (int j = 0; j < Z_ITERATIONS; ++j)
    {
        acc[j] = fmaf(qPrev, rr[j], acc[j]);
    }
    const float dx = cx - a.x;
    const float dy = cy - a.y;
    const float bracket = fmaf(dx, dx, dy * dy);
    KZ_UNROLL_Z
    for (int j = 0; j < Z_ITERATIONS; ++j)
    {
        const float dz = hz[j] - a.z;
        rr[j] = fmaf(dz, dz, bracket);
    }
    qPrev = a.w;
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_setprio(3);
    KZ_UNROLL_Z
    for (int j = 0; j < Z_ITERATIONS; ++j)
    {
        rr[j] = __builtin_amdgcn_rsqf(rr[j]);
    }
}

} // namespace

extern "C" __global__ __launch_bounds__(KZ_THREADS) void directCoulombSum(const float4* __restrict__ atomInfo,
    const float* __restrict__ atomInfoX, const float* __restrict__ atomInfoY, const float* __restrict__ atomInfoZ,
    const float* __restrict__ atomInfoW, int numberOfAtoms, float gridSpacing, int gridSize, float* __restrict__ energyGrid)
{
    const int x = blockIdx.x * WORK_GROUP_SIZE_X + threadIdx.x;
    const int y = blockIdx.y * WORK_GROUP_SIZE_Y + threadIdx.y;
    const int z0 = blockIdx.z * Z_ITERATIONS;
    const float cx = gridSpacing * static_cast<float>(x);
    const float cy = gridSpacing * static_cast<float>(y);
    float hz[Z_ITERATIONS];
    float acc[Z_ITERATIONS];
    float rr[Z_ITERATIONS];
    KZ_UNROLL_Z
    for (int j = 0; j < Z_ITERATIONS; ++j)
    {
        hz[j] = gridSpacing * static_cast<float>(z0 + j);
        acc[j] = 0.0f;
        rr[j] = 0.0f;
    }
    float qPrev = 0.0f;

#if USE_CONSTANT_MEMORY
    // The atom index is the same in every lane: the loads go through the scalar data cache.
    for (int i = 0; i < numberOfAtoms; ++i)
    {
        phasedAtom(loadAtom(atomInfo, atomInfoX, atomInfoY, atomInfoZ, atomInfoW, i), cx, cy, hz, acc, rr, qPrev);
    }
#else
    __shared__ float4 tile[KZ_TILE];
    const int tid = threadIdx.y * WORK_GROUP_SIZE_X + threadIdx.x;
    for (int tileStart = 0; tileStart < numberOfAtoms; tileStart += KZ_TILE)
    {
        __syncthreads();  // the previous tile is consumed
        for (int k = tid * VECTOR_SIZE; k < KZ_TILE; k += KZ_THREADS * VECTOR_SIZE)
        {
#pragma unroll
            for (int v = 0; v < VECTOR_SIZE; ++v)
            {
                const int i = tileStart + k + v;
                tile[k + v] = i < numberOfAtoms ? loadAtom(atomInfo, atomInfoX, atomInfoY, atomInfoZ, atomInfoW, i)
                                                : make_float4(kPadCoord, kPadCoord, kPadCoord, 0.0f);
            }
        }
        __syncthreads();
        const int count = numberOfAtoms - tileStart < KZ_TILE ? numberOfAtoms - tileStart : KZ_TILE;
        if (count == KZ_TILE)
        {
#pragma unroll 4
            for (int k = 0; k < KZ_TILE; ++k)
            {
                phasedAtom(tile[k], cx, cy, hz, acc, rr, qPrev);
            }
        }
        else
        {
            for (int k = 0; k < count; ++k)
            {
                phasedAtom(tile[k], cx, cy, hz, acc, rr, qPrev);
            }
        }
    }
#endif
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_setprio(0);
    KZ_UNROLL_Z
    for (int j = 0; j < Z_ITERATIONS; ++j)
    {
        acc[j] = fmaf(qPrev, rr[j], acc[j]);
    }

    if (x >= gridSize || y >= gridSize)
    {
        return;
    }
    const size_t slice = static_cast<size_t>(gridSize) * gridSize;
    float* out = energyGrid + static_cast<size_t>(z0) * slice + static_cast<size_t>(y) * gridSize + x;
    KZ_UNROLL_Z
    for (int j = 0; j < Z_ITERATIONS; ++j)
    {
        if (z0 + j < gridSize)
        {
            out[j * slice] += acc[j];
        }
    }
}
