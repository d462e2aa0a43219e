// Direct Coulomb summation engine for gfx950 (CDNA4, wave64), shared by every
// Coulomb kernel source in this directory.
//
// V(p) = sum_i q_i / |p - r_i| on a regular grid with origin 0 and spacing h
// (Examples/CoulombSum3d/CoulombSum3d.cu:13-38 defines the same quantity).
//
// Work decomposition.  The grid is a set of "rows": all x points that share one
// (y, z).  A workgroup owns WORK_GROUP_SIZE_X * POINTS_X consecutive x points of
// WORK_GROUP_SIZE_Y * ROWS_PER_THREAD consecutive rows; each thread keeps a
// POINTS_X x ROWS_PER_THREAD block of partial sums in registers.  Since
//     |p - r|^2 = (x - X)^2 + [(y - Y)^2 + (z - Z)^2]
// the bracket depends only on (row, atom).  Atoms are staged through LDS one
// tile at a time; while staging, the workgroup computes the bracket for every
// (atom, row) of the tile once and stores it in LDS.  The inner loop is then
//     per atom:   POINTS_X x (sub, mul)                      dx^2
//     per point:  add, v_rsq_f32, fma                        q * rsqrt(dx^2 + B)
// i.e. ~2 VALU + 1 transcendental per interaction, the LDS reads being
// wave-uniform broadcasts.  There is no MFMA: the sum is not a contraction.
//
// Tuning parameters arrive as preprocessor defines (hipRTC -D options):
//   WORK_GROUP_SIZE_X, WORK_GROUP_SIZE_Y  workgroup shape (threads)
//   POINTS_X                              x points per thread (register run)
//   ROWS_PER_THREAD                       rows per thread
//   ATOM_TILE                             atoms per LDS tile
//   INNER_UNROLL_FACTOR                   unroll of the per-tile atom loop
//   USE_SOA                               0: float4 atoms, 1: X/Y/Z/W arrays
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

#ifndef WORK_GROUP_SIZE_X
#define WORK_GROUP_SIZE_X 64
#endif
#ifndef WORK_GROUP_SIZE_Y
#define WORK_GROUP_SIZE_Y 4
#endif
#ifndef POINTS_X
#define POINTS_X 4
#endif
#ifndef ROWS_PER_THREAD
#define ROWS_PER_THREAD 16
#endif
#ifndef ATOM_TILE
#define ATOM_TILE 64
#endif
#ifndef INNER_UNROLL_FACTOR
#define INNER_UNROLL_FACTOR 1
#endif
#ifndef USE_SOA
#define USE_SOA 0
#endif

#define KC_THREADS (WORK_GROUP_SIZE_X * WORK_GROUP_SIZE_Y)
#define KC_ROWS (WORK_GROUP_SIZE_Y * ROWS_PER_THREAD)
#define KC_SPAN_X (WORK_GROUP_SIZE_X * POINTS_X)
#define KC_POINTS (POINTS_X * ROWS_PER_THREAD)
// Bracket precompute: each thread owns KC_PRE_ROWS rows and every
// KC_PRE_STRIDE-th atom of a tile.
#define KC_PRE_ROWS ((KC_ROWS > KC_THREADS) ? (KC_ROWS / KC_THREADS) : 1)
#define KC_PRE_STRIDE ((KC_THREADS > KC_ROWS) ? (KC_THREADS / KC_ROWS) : 1)

static_assert(KC_THREADS <= 1024, "workgroup larger than the device limit");
static_assert((KC_THREADS % KC_ROWS) == 0 || (KC_ROWS % KC_THREADS) == 0,
              "rows per workgroup and threads per workgroup must divide one another");
static_assert(ATOM_TILE % KC_PRE_STRIDE == 0, "ATOM_TILE must be a multiple of threads / rows");
static_assert(ATOM_TILE * (KC_ROWS + 4) * 4 <= 160 * 1024, "LDS tile exceeds 160 KiB");

namespace kc
{

// Padding atom used past the end of an atom range: zero charge, far away, so
// rsqrt stays finite and the term is exactly 0.
constexpr float kPadCoord = -1.0e6f;

// Row geometry: how the per-row bracket B(row, atom) is formed.
//   Volume: row = z * G + y, B = (h*y - Y)^2 + (h*z - Z)^2
//   Slice:  row = y,         B = (h*y - Y)^2 + Z   (Z already holds dz^2,
//           Examples/CoulombSum2d/CoulombSum2d.cl:126 semantics)
enum RowMode
{
    Volume = 0,
    Slice = 1
};

struct AtomSource
{
    const float4* aos;
    const float* x;
    const float* y;
    const float* z;
    const float* w;

    __device__ __forceinline__ float4 load(int i) const
    {
#if USE_SOA == 0
        return aos[i];
#else
        return make_float4(x[i], y[i], z[i], w[i]);
#endif
    }
};

// Error-free transformation s + e = a + b (Knuth two-sum); fold e into c.
__device__ __forceinline__ void twoSumInto(float& s, float& c, const float b)
{
#pragma clang fp reassociate(off) contract(off)
    const float a = s;
    const float t = a + b;
    const float bp = t - a;
    const float e = (a - (t - bp)) + (b - bp);
    s = t;
    c += e;
}

// Accumulates atoms [a0, a1) into this thread's point block.
//   part  plain FP32 partial sums (COMP == false: the result)
//   sum, comp  (COMP == true) running two-sum pair; part is folded in per tile
// Coordinates of the thread's x points are xLocal[i]; in compensated form they
// are relative to originX (subtracted from the atoms in FP64 while staging), and
// the row brackets are formed in FP64 from absolute coordinates, so every
// displacement carries one FP32 rounding at most.
template <int MODE, bool COMP>
__device__ __forceinline__ void accumulateAtoms(const AtomSource src, const int a0, const int a1, const float h,
    const int rowBase, const int rowsTotal, const int rowsPerPlane, const double originX,
    const float (&xLocal)[POINTS_X], float (&part)[ROWS_PER_THREAD][POINTS_X],
    float (&sum)[ROWS_PER_THREAD][POINTS_X], float (&comp)[ROWS_PER_THREAD][POINTS_X])
{
    __shared__ float2 sXQ[ATOM_TILE];                       // (X, q), X relative to originX when COMP
    __shared__ float2 sYZ[ATOM_TILE];                       // raw (Y, Z) for the bracket precompute
    __shared__ __attribute__((aligned(16))) float sB[ATOM_TILE * KC_ROWS];

    const int tid = threadIdx.y * WORK_GROUP_SIZE_X + threadIdx.x;
    const int rowOffset = threadIdx.y * ROWS_PER_THREAD;

    // Coordinates of the rows this thread precomputes brackets for.
    float preY[KC_PRE_ROWS];
    float preZ[KC_PRE_ROWS];
    double preYd[KC_PRE_ROWS];
    double preZd[KC_PRE_ROWS];
    const int preRow0 = tid % KC_ROWS;
    const int preAtom0 = tid / KC_ROWS;
#pragma unroll
    for (int m = 0; m < KC_PRE_ROWS; ++m)
    {
        int row = rowBase + preRow0 + m * KC_THREADS;
        row = row < rowsTotal ? row : rowsTotal - 1;
        const int y = (MODE == Volume) ? (row % rowsPerPlane) : row;
        const int z = (MODE == Volume) ? (row / rowsPerPlane) : 0;
        preY[m] = h * static_cast<float>(y);
        preZ[m] = h * static_cast<float>(z);
        preYd[m] = static_cast<double>(h) * y;
        preZd[m] = static_cast<double>(h) * z;
    }

    for (int tileStart = a0; tileStart < a1; tileStart += ATOM_TILE)
    {
        __syncthreads();   // previous tile fully consumed
        for (int a = tid; a < ATOM_TILE; a += KC_THREADS)
        {
            const int ai = tileStart + a;
            const float4 at = (ai < a1) ? src.load(ai) : make_float4(kPadCoord, kPadCoord, kPadCoord, 0.0f);
            float xs = at.x;
            if constexpr (COMP)
            {
                xs = static_cast<float>(static_cast<double>(at.x) - originX);
            }
            sXQ[a] = make_float2(xs, at.w);
            sYZ[a] = make_float2(at.y, at.z);
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < KC_PRE_ROWS; ++m)
        {
            for (int a = preAtom0; a < ATOM_TILE; a += KC_PRE_STRIDE)
            {
                const float2 yz = sYZ[a];
                float b;
                if constexpr (COMP)
                {
                    const double dy = preYd[m] - static_cast<double>(yz.x);
                    if constexpr (MODE == Volume)
                    {
                        const double dz = preZd[m] - static_cast<double>(yz.y);
                        b = static_cast<float>(dy * dy + dz * dz);
                    }
                    else
                    {
                        b = static_cast<float>(dy * dy + static_cast<double>(yz.y));
                    }
                }
                else
                {
                    const float dy = preY[m] - yz.x;
                    if constexpr (MODE == Volume)
                    {
                        const float dz = preZ[m] - yz.y;
                        b = dy * dy + dz * dz;
                    }
                    else
                    {
                        b = dy * dy + yz.y;
                    }
                }
                sB[a * KC_ROWS + preRow0 + m * KC_THREADS] = b;
            }
        }
        __syncthreads();

#pragma unroll INNER_UNROLL_FACTOR
        for (int a = 0; a < ATOM_TILE; ++a)
        {
            const float2 xq = sXQ[a];
            float dx2[POINTS_X];
#pragma unroll
            for (int i = 0; i < POINTS_X; ++i)
            {
                const float d = xLocal[i] - xq.x;
                dx2[i] = d * d;
            }
            const float* bRow = &sB[a * KC_ROWS + rowOffset];
            if constexpr ((ROWS_PER_THREAD % 4) == 0 && (KC_ROWS % 4) == 0)
            {
#pragma unroll
                for (int k = 0; k < ROWS_PER_THREAD; k += 4)
                {
                    const float4 b4 = *reinterpret_cast<const float4*>(bRow + k);
                    const float bv[4] = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                    {
#pragma unroll
                        for (int i = 0; i < POINTS_X; ++i)
                        {
                            part[k + j][i] = fmaf(xq.y, __builtin_amdgcn_rsqf(dx2[i] + bv[j]), part[k + j][i]);
                        }
                    }
                }
            }
            else
            {
#pragma unroll
                for (int k = 0; k < ROWS_PER_THREAD; ++k)
                {
                    const float bk = bRow[k];
#pragma unroll
                    for (int i = 0; i < POINTS_X; ++i)
                    {
                        part[k][i] = fmaf(xq.y, __builtin_amdgcn_rsqf(dx2[i] + bk), part[k][i]);
                    }
                }
            }
        }

        if constexpr (COMP)
        {
#pragma unroll
            for (int k = 0; k < ROWS_PER_THREAD; ++k)
            {
#pragma unroll
                for (int i = 0; i < POINTS_X; ++i)
                {
                    twoSumInto(sum[k][i], comp[k][i], part[k][i]);
                    part[k][i] = 0.0f;
                }
            }
        }
    }
}

// Atom range [a0, a1) of atom block `block` out of `blocks`; blocks are whole
// tiles so that every block but the last is full.
__device__ __forceinline__ void atomBlockRange(const int numberOfAtoms, const int block, const int blocks, int& a0, int& a1)
{
    const int perBlock = (numberOfAtoms + blocks - 1) / blocks;
    const int chunk = ((perBlock + ATOM_TILE - 1) / ATOM_TILE) * ATOM_TILE;
    a0 = block * chunk;
    a1 = a0 + chunk;
    a0 = a0 < numberOfAtoms ? a0 : numberOfAtoms;
    a1 = a1 < numberOfAtoms ? a1 : numberOfAtoms;
}

// energyGrid[row * rowLength + x] += v for this thread's block, skipping points
// outside [0, rowLength) x [0, rowsTotal).  atomic: several atom blocks add
// into the same points.
template <bool ATOMIC>
__device__ __forceinline__ void storeBlock(float* energyGrid, const size_t rowStride, const size_t outOffset,
    const int rowLength, const int rowsTotal, const int x0, const int row0, const float (&v)[ROWS_PER_THREAD][POINTS_X])
{
    const bool vectorPath = !ATOMIC && (POINTS_X % 4) == 0 && (rowLength % 4) == 0 && (rowStride % 4) == 0
        && (outOffset % 4) == 0;
#pragma unroll
    for (int k = 0; k < ROWS_PER_THREAD; ++k)
    {
        const int row = row0 + k;
        if (row >= rowsTotal)
        {
            break;
        }
        float* out = energyGrid + outOffset + static_cast<size_t>(row) * rowStride;
        if (vectorPath && x0 + POINTS_X <= rowLength)
        {
#pragma unroll
            for (int i = 0; i < POINTS_X; i += 4)
            {
                float4* p = reinterpret_cast<float4*>(out + x0 + i);
                float4 o = *p;
                o.x += v[k][i];
                o.y += v[k][i + 1];
                o.z += v[k][i + 2];
                o.w += v[k][i + 3];
                *p = o;
            }
        }
        else
        {
#pragma unroll
            for (int i = 0; i < POINTS_X; ++i)
            {
                const int x = x0 + i;
                if (x < rowLength)
                {
                    if constexpr (ATOMIC)
                    {
                        unsafeAtomicAdd(out + x, v[k][i]);
                    }
                    else
                    {
                        out[x] += v[k][i];
                    }
                }
            }
        }
    }
}

} // namespace kc
