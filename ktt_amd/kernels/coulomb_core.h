// Direct Coulomb summation engine for gfx950 (CDNA4, wave64), shared by every
// Coulomb kernel source in this directory.
//
// V(p) = sum_i q_i / |p - r_i| on a regular grid with origin 0 and spacing h
// (Examples/CoulombSum3d/CoulombSum3d.cu:13-38 defines the same quantity).
//
// Work decomposition.  The grid is a set of "rows": all x points that share one
// (y, z).  A workgroup owns WGX * PX consecutive x points of WGY * ROWS
// consecutive rows; each thread keeps a PX x ROWS block of partial sums in
// registers.  Since
//     |p - r|^2 = (x - X)^2 + [(y - Y)^2 + (z - Z)^2]
// the bracket depends only on (row, atom).  Atoms are staged through LDS one
// tile at a time; while staging, the workgroup computes the bracket for every
// (atom, row) of the tile once and stores it in LDS.  The inner loop is then
//     per atom:   PX subtractions                            dx
//     per point:  fma(dx, dx, B), v_rsq_f32, fma(q, r^-1, acc)
// i.e. 2 VALU + 1 transcendental per interaction, the LDS reads being
// wave-uniform broadcasts (no bank conflicts).  There is no MFMA: the sum is
// not a contraction.
//
// The engine is a template over the tile shape so that one kernel can hold
// several shapes (the batched kernel uses one per grid-size bucket).  The
// single-shape kernels take their shape from preprocessor defines (hipRTC -D
// options, tuned by ktt):
//   WORK_GROUP_SIZE_X, WORK_GROUP_SIZE_Y  workgroup shape (threads)
//   POINTS_X                              x points per thread (register run)
//   ROWS_PER_THREAD                       rows per thread
//   ATOM_TILE                             atoms per LDS tile
//   INNER_UNROLL_FACTOR                   unroll of the per-tile atom loop
//   USE_SOA                               0: float4 atoms, 1: X/Y/Z/W arrays
//   ISSUE_PHASES                          0: the compiler interleaves the FMAs and v_rsq_f32
//                                         of each point; 1: per atom a phase of FMAs at wave
//                                         priority 0, then a phase of v_rsq_f32 at priority 3
//                                         (Engine::phasedTile); 2: as 1, reading the next
//                                         atom's LDS data one atom ahead
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
#define POINTS_X 2
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
#ifndef ISSUE_PHASES
#define ISSUE_PHASES 1
#endif

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

// prior + (s + c) of a two-sum pair, in FP64 and in the order (prior + s) + c. An infinite term
// (an atom exactly on a grid point: r^2 = 0) makes s infinite and c NaN (the two-sum's error of
// inf - inf); s, infinite or NaN from then on, is then the IEEE result of the plain sum, as in
// the plain form. A finite s comes with a finite c.
__device__ __forceinline__ double addPair(const double prior, const float s, const float c)
{
    return __builtin_isfinite(s) ? prior + static_cast<double>(s) + static_cast<double>(c) : prior + static_cast<double>(s);
}

// The (x group, row group) of this workgroup in a whole-volume launch of a G^3
// grid: groupsX = ceil(G / spanX) groups of x points times groupsY =
// ceil(G^2 / rowsPerGroup) groups of rows. A launch's y dimension holds at most
// 65,535 workgroups, so a launch with more row groups folds them into x:
// gridDim.x = fold * groupsX and row group = blockIdx.y * fold + blockIdx.x /
// groupsX. A launch with gridDim.x == groupsX has fold 1 and maps blockIdx as it
// is. Workgroups beyond the groups (past fold * groupsX in x, or in the fold's
// last, partial run of row groups) are not valid and exit.
struct VolumeGroup
{
    int x;
    int row;
    int groupsX;
    int groupsY;
    bool valid;
};

__device__ __forceinline__ VolumeGroup volumeGroup(const int gridSize, const int spanX, const int rowsPerGroup)
{
    VolumeGroup g;
    g.groupsX = (gridSize + spanX - 1) / spanX;
    const unsigned int rows = static_cast<unsigned int>(gridSize) * static_cast<unsigned int>(gridSize);
    g.groupsY = static_cast<int>((rows + static_cast<unsigned int>(rowsPerGroup) - 1u) / static_cast<unsigned int>(rowsPerGroup));
    const int perRow = static_cast<int>(gridDim.x) / g.groupsX;
    const int fold = perRow > 1 ? perRow : 1;
    const int bx = static_cast<int>(blockIdx.x);
    g.x = bx % g.groupsX;
    g.row = static_cast<int>(blockIdx.y) * fold + bx / g.groupsX;
    g.valid = bx < fold * g.groupsX && g.row < g.groupsY;
    return g;
}

// Atom range [a0, a1) of atom block `block` out of `blocks`; blocks are whole
// tiles so that every block but the last is full.
__device__ __forceinline__ void atomBlockRange(const int numberOfAtoms, const int block, const int blocks, const int tile,
    int& a0, int& a1)
{
    const int perBlock = (numberOfAtoms + blocks - 1) / blocks;
    const int chunk = ((perBlock + tile - 1) / tile) * tile;
    a0 = block * chunk;
    a1 = a0 + chunk;
    a0 = a0 < numberOfAtoms ? a0 : numberOfAtoms;
    a1 = a1 < numberOfAtoms ? a1 : numberOfAtoms;
}

template <int WGX, int WGY, int PX, int ROWS, int TILE, int UNROLL>
struct Engine
{
    static constexpr int LanesX = WGX;            // lanes along x
    static constexpr int RowGroups = WGY;         // thread rows
    static constexpr int PointsX = PX;
    static constexpr int RowsPerThread = ROWS;
    static constexpr int Tile = TILE;
    static constexpr int Threads = WGX * WGY;
    static constexpr int Rows = WGY * ROWS;       // rows per workgroup
    static constexpr int SpanX = WGX * PX;        // x points per workgroup
    static constexpr int Points = PX * ROWS;      // points per thread
    // Bracket precompute: each thread owns PreRows rows and every PreStride-th atom of a tile.
    static constexpr int PreRows = Rows > Threads ? Rows / Threads : 1;
    static constexpr int PreStride = Threads > Rows ? Threads / Rows : 1;
    // LDS: (X, q) and (Y, Z) per atom, then TILE x Rows brackets.
    static constexpr int SharedFloats = 4 * TILE + TILE * Rows;

    static_assert(Threads <= 1024, "workgroup larger than the device limit");
    static_assert((Threads % Rows) == 0 || (Rows % Threads) == 0,
                  "rows per workgroup and threads per workgroup must divide one another");
    static_assert(TILE % PreStride == 0, "ATOM_TILE must be a multiple of threads / rows");
    static_assert(SharedFloats * 4 <= 160 * 1024, "LDS tile exceeds 160 KiB");
    static_assert(UNROLL >= 1, "unroll factor must be at least 1");

    // Accumulates atoms [a0, a1) into this thread's point block.
    //   part  plain FP32 partial sums (COMP == false: the result)
    //   sum, comp  (COMP == true) running two-sum pair; part is folded in per tile
    // tid is the flat thread index (lane x = tid % WGX, row group = tid / WGX).
    // The thread's x coordinates are xLocal[i]; in compensated form they are
    // relative to originX (subtracted from the atoms in FP64 while staging), and
    // the row brackets are formed in FP64 from absolute coordinates, so every
    // displacement carries one FP32 rounding at most.
    template <int MODE, bool COMP>
    static __device__ __forceinline__ void accumulate(float* shared, const AtomSource src, const int a0, const int a1,
        const float h, const int tid, const int rowBase, const int rowsTotal, const int rowsPerPlane, const double originX,
        const float (&xLocal)[PX], float (&part)[ROWS][PX], float (&sum)[ROWS][PX], float (&comp)[ROWS][PX])
    {
        float2* sXQ = reinterpret_cast<float2*>(shared);
        float2* sYZ = reinterpret_cast<float2*>(shared + 2 * TILE);
        float* sB = shared + 4 * TILE;

        const int rowOffset = (tid / WGX) * ROWS;

        // Coordinates of the rows this thread precomputes brackets for.
        float preY[PreRows];
        float preZ[PreRows];
        double preYd[PreRows];
        double preZd[PreRows];
        const int preRow0 = tid % Rows;
        const int preAtom0 = tid / Rows;
#pragma unroll
        for (int m = 0; m < PreRows; ++m)
        {
            int row = rowBase + preRow0 + m * Threads;
            row = row < rowsTotal ? row : rowsTotal - 1;
            const int y = (MODE == Volume) ? (row % rowsPerPlane) : row;
            const int z = (MODE == Volume) ? (row / rowsPerPlane) : 0;
            preY[m] = h * static_cast<float>(y);
            preZ[m] = h * static_cast<float>(z);
            preYd[m] = static_cast<double>(h) * y;
            preZd[m] = static_cast<double>(h) * z;
        }

        for (int tileStart = a0; tileStart < a1; tileStart += TILE)
        {
            __syncthreads();   // previous tile fully consumed
            for (int a = tid; a < TILE; a += Threads)
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
            for (int m = 0; m < PreRows; ++m)
            {
                for (int a = preAtom0; a < TILE; a += PreStride)
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
                    sB[a * Rows + preRow0 + m * Threads] = b;
                }
            }
            __syncthreads();

#if ISSUE_PHASES
            phasedTile(sXQ, sB, rowOffset, xLocal, part);
#else
#pragma unroll UNROLL
            for (int a = 0; a < TILE; ++a)
            {
                const float2 xq = sXQ[a];
                float dx[PX];
#pragma unroll
                for (int i = 0; i < PX; ++i)
                {
                    dx[i] = xLocal[i] - xq.x;
                }
                const float* bRow = &sB[a * Rows + rowOffset];
                if constexpr ((ROWS % 4) == 0 && (Rows % 4) == 0)
                {
#pragma unroll
                    for (int k = 0; k < ROWS; k += 4)
                    {
                        const float4 b4 = *reinterpret_cast<const float4*>(bRow + k);
                        const float bv[4] = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                        {
#pragma unroll
                            for (int i = 0; i < PX; ++i)
                            {
                                part[k + j][i] = fmaf(xq.y, __builtin_amdgcn_rsqf(fmaf(dx[i], dx[i], bv[j])), part[k + j][i]);
                            }
                        }
                    }
                }
                else
                {
#pragma unroll
                    for (int k = 0; k < ROWS; ++k)
                    {
                        const float bk = bRow[k];
#pragma unroll
                        for (int i = 0; i < PX; ++i)
                        {
                            part[k][i] = fmaf(xq.y, __builtin_amdgcn_rsqf(fmaf(dx[i], dx[i], bk)), part[k][i]);
                        }
                    }
                }
            }
#endif

            if constexpr (COMP)
            {
#pragma unroll
                for (int k = 0; k < ROWS; ++k)
                {
#pragma unroll
                    for (int i = 0; i < PX; ++i)
                    {
                        twoSumInto(sum[k][i], comp[k][i], part[k][i]);
                        part[k][i] = 0.0f;
                    }
                }
            }
        }
    }

    // One LDS tile of atoms in issue phases. A gfx950 SIMD can issue two VALU instructions of two
    // waves in one quad-cycle when neither is a transcendental (and neither reads an SGPR), while
    // a v_rsq_f32 issues alone for two quad-cycles. Waves that interleave the two kinds rarely
    // present two FMAs at once, so each wave alternates a phase of FMAs (accumulate the previous
    // atom's q * r^-1, form this atom's r^2) at priority 0 with a phase of v_rsq_f32 at
    // priority 3: the arbiter runs the waves' v_rsq_f32 phases one after another while the
    // others wait in their FMA phases, and those then issue in pairs. Every point still adds
    // its atoms in tile order with the same operations, so the sums are bit-identical to the
    // interleaved loop.
    // (X, q) and this thread's row brackets of atom a of the tile.
    static __device__ __forceinline__ void loadAtom(const float2* sXQ, const float* sB, const int rowOffset, const int a,
        float2& xq, float (&bv)[ROWS])
    {
        xq = sXQ[a];
        const float* bRow = &sB[a * Rows + rowOffset];
        if constexpr ((ROWS % 4) == 0 && (Rows % 4) == 0)
        {
#pragma unroll
            for (int k = 0; k < ROWS; k += 4)
            {
                const float4 b4 = *reinterpret_cast<const float4*>(bRow + k);
                bv[k] = b4.x;
                bv[k + 1] = b4.y;
                bv[k + 2] = b4.z;
                bv[k + 3] = b4.w;
            }
        }
        else
        {
#pragma unroll
            for (int k = 0; k < ROWS; ++k)
            {
                bv[k] = bRow[k];
            }
        }
    }

    static __device__ __forceinline__ void phasedTile(const float2* sXQ, const float* sB, const int rowOffset,
        const float (&xLocal)[PX], float (&part)[ROWS][PX])
    {
        float rr[ROWS][PX];  // r^2 of the atom in flight, then its r^-1
        clear(rr);
        float qPrev = 0.0f;  // charge of the atom whose r^-1 is in rr (0 before the first)
#if ISSUE_PHASES == 2
        float2 nextXq;  // the next atom's LDS data, read one atom ahead
        float nextB[ROWS];
        loadAtom(sXQ, sB, rowOffset, 0, nextXq, nextB);
#endif
#pragma unroll UNROLL
        for (int a = 0; a < TILE; ++a)
        {
            __builtin_amdgcn_sched_barrier(0);
            __builtin_amdgcn_s_setprio(0);
            float2 xq;
            float bv[ROWS];
#if ISSUE_PHASES == 2
            xq = nextXq;
#pragma unroll
            for (int k = 0; k < ROWS; ++k)
            {
                bv[k] = nextB[k];
            }
#else
            loadAtom(sXQ, sB, rowOffset, a, xq, bv);
#endif
            float dx[PX];
#pragma unroll
            for (int i = 0; i < PX; ++i)
            {
                dx[i] = xLocal[i] - xq.x;
            }
#pragma unroll
            for (int k = 0; k < ROWS; ++k)
            {
#pragma unroll
                for (int i = 0; i < PX; ++i)
                {
                    part[k][i] = fmaf(qPrev, rr[k][i], part[k][i]);
                    rr[k][i] = fmaf(dx[i], dx[i], bv[k]);
                }
            }
            qPrev = xq.y;
#if ISSUE_PHASES == 2
            loadAtom(sXQ, sB, rowOffset, a + 1 < TILE ? a + 1 : a, nextXq, nextB);
#endif
            __builtin_amdgcn_sched_barrier(0);
            __builtin_amdgcn_s_setprio(3);
#pragma unroll
            for (int k = 0; k < ROWS; ++k)
            {
#pragma unroll
                for (int i = 0; i < PX; ++i)
                {
                    rr[k][i] = __builtin_amdgcn_rsqf(rr[k][i]);
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_setprio(0);
#pragma unroll
        for (int k = 0; k < ROWS; ++k)
        {
#pragma unroll
            for (int i = 0; i < PX; ++i)
            {
                part[k][i] = fmaf(qPrev, rr[k][i], part[k][i]);
            }
        }
    }

    static __device__ __forceinline__ void clear(float (&v)[ROWS][PX])
    {
#pragma unroll
        for (int k = 0; k < ROWS; ++k)
        {
#pragma unroll
            for (int i = 0; i < PX; ++i)
            {
                v[k][i] = 0.0f;
            }
        }
    }

    // energyGrid[outOffset + row * rowStride + x] += v for this thread's block,
    // skipping points outside [0, rowLength) x [0, rowsTotal).  ATOMIC: several
    // atom blocks add into the same points.
    template <bool ATOMIC>
    static __device__ __forceinline__ void store(float* energyGrid, const size_t rowStride, const size_t outOffset,
        const int rowLength, const int rowsTotal, const int x0, const int row0, const float (&v)[ROWS][PX])
    {
        const bool vectorPath = !ATOMIC && (PX % 4) == 0 && (rowLength % 4) == 0 && (rowStride % 4) == 0 && (outOffset % 4) == 0;
#pragma unroll
        for (int k = 0; k < ROWS; ++k)
        {
            const int row = row0 + k;
            if (row >= rowsTotal)
            {
                break;
            }
            float* out = energyGrid + outOffset + static_cast<size_t>(row) * rowStride;
            if (vectorPath && x0 + PX <= rowLength)
            {
#pragma unroll
                for (int i = 0; i < PX; i += 4)
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
                for (int i = 0; i < PX; ++i)
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

    // energyGrid[...] = grid + (sum + comp), merged in FP64 (compensated form, one atom block).
    static __device__ __forceinline__ void storeCompensated(float* energyGrid, const size_t rowStride, const size_t outOffset,
        const int rowLength, const int rowsTotal, const int x0, const int row0, const float (&sum)[ROWS][PX],
        const float (&comp)[ROWS][PX])
    {
#pragma unroll
        for (int k = 0; k < ROWS; ++k)
        {
            const int row = row0 + k;
            if (row >= rowsTotal)
            {
                break;
            }
            float* out = energyGrid + outOffset + static_cast<size_t>(row) * rowStride;
#pragma unroll
            for (int i = 0; i < PX; ++i)
            {
                if (x0 + i < rowLength)
                {
                    out[x0 + i] = static_cast<float>(addPair(static_cast<double>(out[x0 + i]), sum[k][i], comp[k][i]));
                }
            }
        }
    }
};

// The shape selected by the preprocessor defines.
using DefinedEngine = Engine<WORK_GROUP_SIZE_X, WORK_GROUP_SIZE_Y, POINTS_X, ROWS_PER_THREAD, ATOM_TILE, INNER_UNROLL_FACTOR>;

} // namespace kc

#define KC_THREADS (WORK_GROUP_SIZE_X * WORK_GROUP_SIZE_Y)
