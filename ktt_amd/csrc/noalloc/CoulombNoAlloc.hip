// CoulombSum3d in the compensated form with no device allocation per call.
//
// For callers that cannot call hipMalloc: code under HIP graph capture, tight serving loops,
// harnesses that forbid allocation.  hipcc builds this file into libktt_hip_noalloc.so, which
// libktt_hip.so loads on the first call of ktt_coulomb3d_compensated_noalloc and forwards to
// ktt_noalloc_coulomb3d_compensated below.  Results are those of ktt_coulomb3d with
// KTT_COULOMB_COMPENSATED and the default configuration (same kernel body,
// kc::volumeCompensated, at the same tile shapes, with the tail split):
//   grid side a multiple of 256: 128 x 2 threads, 2 x 8 points per thread, tile 256, unroll 16
//   otherwise:                    64 x 2 threads, same per-thread block.
//
// Scratch.  The split tiles' (sum, comp) partials and their arrival counters live in the
// file-scope array `scratch` (kScratchBytes per device, held by every GPU the library runs on
// until the process ends).  Its device address is looked up once per device with
// hipGetSymbolAddress, the device current, and kept in g_Devices; kernels receive that
// address, never the array's host-side name.  Every call on one GPU shares that GPU's array:
// the counters are cleared by a fill kernel at the start of every call, and calls that
// overlap on one GPU (two streams, or two threads) race on the partials.
//
// Set-up (first call on a device): symbol lookup, compute-unit count and the kernels'
// occupancy.  Make that first call outside graph capture.  After it, a call launches one
// fill kernel and the Coulomb kernel on the caller's stream and allocates nothing.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdio>
#include <mutex>

#include "coulomb_volume.h"

namespace
{

// Scratch bound.  A split round has blocks x tail <= resident workgroups (blocks is at most
// resident / tail), so the partials take at most (resident threads) x 16 points x 8 B; with one
// counter per tail tile.  At the kernels' occupancy on MI355X (5 waves per SIMD, 1,280
// threads per CU, 256 CUs) that is 40 MiB + 1 KiB; the array covers up to 8 waves per SIMD
// (64 MiB of partials) plus the counters.  A plan that needs more is rejected.
constexpr size_t kScratchBytes = (size_t(64) << 20) + (size_t(64) << 10);
constexpr int kMaxDevices = 64;
constexpr long kMaxTailBlocks = 16;  // as ktt_coulomb3d
constexpr int kTile = 256;

} // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace
{

using WideShape = kc::Engine<128, 2, 2, 8, kTile, 16>;
using NarrowShape = kc::Engine<64, 2, 2, 8, kTile, 16>;

struct DeviceState
{
    bool ready = false;
    unsigned char* scratch = nullptr;  // device address of `scratch` on this device
    int computeUnits = 0;
    int residentWide = 0;    // workgroups per CU
    int residentNarrow = 0;
};

DeviceState g_Devices[kMaxDevices];
std::mutex g_Mutex;

} // namespace

extern "C" __global__ __launch_bounds__(WideShape::Threads) void kttNoAllocCompensatedWide(const float4* atomInfo, int numberOfAtoms,
    float gridSpacing, int gridSize, float* energyGrid, float2* partials, unsigned int* tileCounters, int splitFrom)
{
    const kc::AtomSource src{atomInfo, nullptr, nullptr, nullptr, nullptr};
    kc::volumeCompensated<WideShape>(src, numberOfAtoms, gridSpacing, gridSize, energyGrid, partials, tileCounters, splitFrom);
}

extern "C" __global__ __launch_bounds__(NarrowShape::Threads) void kttNoAllocCompensatedNarrow(const float4* atomInfo,
    int numberOfAtoms, float gridSpacing, int gridSize, float* energyGrid, float2* partials, unsigned int* tileCounters,
    int splitFrom)
{
    const kc::AtomSource src{atomInfo, nullptr, nullptr, nullptr, nullptr};
    kc::volumeCompensated<NarrowShape>(src, numberOfAtoms, gridSpacing, gridSize, energyGrid, partials, tileCounters, splitFrom);
}

extern "C" __global__ __launch_bounds__(256) void kttNoAllocClear(unsigned int* values, unsigned int count)
{
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
    {
        values[i] = 0u;
    }
}

namespace
{

int Fail(char* message, const size_t length, const char* text)
{
    if (message != nullptr && length > 0)
    {
        std::snprintf(message, length, "%s", text);
    }
    return -1;
}

int FailHip(char* message, const size_t length, const char* what, const hipError_t status)
{
    if (message != nullptr && length > 0)
    {
        std::snprintf(message, length, "%s: %s", what, hipGetErrorString(status));
    }
    return -1;
}

// Set-up of the current device, once: the device address of `scratch`, the CU count and
// each kernel's resident workgroups per CU.
int EnsureDevice(DeviceState*& state, char* message, const size_t length)
{
    int device = 0;
    hipError_t status = hipGetDevice(&device);
    if (status != hipSuccess)
    {
        return FailHip(message, length, "hipGetDevice", status);
    }
    if (device < 0 || device >= kMaxDevices)
    {
        return Fail(message, length, "device index beyond the 64 devices the no-allocation entry point tracks");
    }
    std::lock_guard<std::mutex> lock(g_Mutex);
    DeviceState& entry = g_Devices[device];
    if (!entry.ready)
    {
        void* pointer = nullptr;
        status = hipGetSymbolAddress(&pointer, HIP_SYMBOL(scratch));
        if (status != hipSuccess || pointer == nullptr)
        {
            return FailHip(message, length, "hipGetSymbolAddress(scratch)", status);
        }
        int units = 0;
        status = hipDeviceGetAttribute(&units, hipDeviceAttributeMultiprocessorCount, device);
        if (status != hipSuccess)
        {
            return FailHip(message, length, "hipDeviceGetAttribute(multiProcessorCount)", status);
        }
        int wide = 0;
        int narrow = 0;
        status = hipOccupancyMaxActiveBlocksPerMultiprocessor(&wide, reinterpret_cast<const void*>(kttNoAllocCompensatedWide),
            WideShape::Threads, 0);
        if (status != hipSuccess)
        {
            return FailHip(message, length, "hipOccupancyMaxActiveBlocksPerMultiprocessor", status);
        }
        status = hipOccupancyMaxActiveBlocksPerMultiprocessor(&narrow,
            reinterpret_cast<const void*>(kttNoAllocCompensatedNarrow), NarrowShape::Threads, 0);
        if (status != hipSuccess)
        {
            return FailHip(message, length, "hipOccupancyMaxActiveBlocksPerMultiprocessor", status);
        }
        entry.scratch = static_cast<unsigned char*>(pointer);
        entry.computeUnits = units;
        entry.residentWide = wide;
        entry.residentNarrow = narrow;
        entry.ready = true;
    }
    state = &entry;
    return 0;
}

} // namespace

// V of n_atoms float4 atoms added into the G^3 grid (see ktt_coulomb3d_compensated_noalloc in
// include/ktt_hip.h).  Returns 0, or -1 with the reason in message.
extern "C" __attribute__((visibility("default"))) int ktt_noalloc_coulomb3d_compensated(const float* atoms, int atomCount,
    float spacing, int gridSize, float* energyGrid, void* stream, char* message, size_t messageLength)
{
    if (atomCount < 0)
    {
        return Fail(message, messageLength, "Number of atoms must not be negative");
    }
    if (atomCount > 0 && atoms == nullptr)
    {
        return Fail(message, messageLength, "Atom buffer is null");
    }
    if (energyGrid == nullptr)
    {
        return Fail(message, messageLength, "Energy grid buffer is null");
    }
    if (reinterpret_cast<size_t>(atoms) % 16 != 0)
    {
        return Fail(message, messageLength, "Atom buffer must be 16-byte aligned (float4 records)");
    }
    // G^2 rows must fit an int; row groups beyond a launch's 65,535 fold into x (kc::volumeGroup).
    if (gridSize <= 0 || gridSize > 46340)
    {
        return Fail(message, messageLength, "grid size must be in [1, 46340] (G^2 rows in an int)");
    }
    DeviceState* state = nullptr;
    if (EnsureDevice(state, message, messageLength) != 0)
    {
        return -1;
    }

    // The launch plan of ktt_coulomb3d (compensated, default configuration, TAIL_SPLIT=1).
    const bool wide = gridSize % 256 == 0;
    const long threads = wide ? WideShape::Threads : NarrowShape::Threads;
    const long span = wide ? WideShape::SpanX : NarrowShape::SpanX;
    const long points = WideShape::Points;
    const long rowsPerGroup = WideShape::Rows;
    const long groupsX = (gridSize + span - 1) / span;
    const long groupsY = (static_cast<long>(gridSize) * gridSize + rowsPerGroup - 1) / rowsPerGroup;
    const long tiles = groupsX * groupsY;
    const long slots = static_cast<long>(state->computeUnits) * (wide ? state->residentWide : state->residentNarrow);
    const long tail = slots > 0 ? tiles % slots : 0;
    const long atomTiles = atomCount > 0 ? (atomCount + kTile - 1) / kTile : 1;
    const long fill = tail > 0 ? slots / tail : 1;
    long blocks = fill < atomTiles ? fill : atomTiles;
    blocks = blocks < kMaxTailBlocks ? blocks : kMaxTailBlocks;
    blocks = blocks > 1 ? blocks : 1;
    const long splitFrom = blocks > 1 ? tiles - tail : tiles;
    const long splitTiles = tiles - splitFrom;

    // Scratch layout: partials float2[blocks][split tiles][points][threads], then one counter
    // per split tile at the next 256-byte boundary.
    const size_t partialBytes = blocks > 1 ? static_cast<size_t>(blocks) * splitTiles * points * threads * sizeof(float2) : 0;
    const size_t counterOffset = (partialBytes + 255) & ~size_t(255);
    const size_t needed = counterOffset + static_cast<size_t>(splitTiles) * sizeof(unsigned int);
    if (needed > kScratchBytes)
    {
        if (message != nullptr && messageLength > 0)
        {
            std::snprintf(message, messageLength,
                "grid %d^3 needs %zu bytes of scratch; the no-allocation entry point holds at most %zu bytes per device", gridSize,
                needed, kScratchBytes);
        }
        return -1;
    }
    float2* partials = reinterpret_cast<float2*>(state->scratch);
    unsigned int* counters = reinterpret_cast<unsigned int*>(state->scratch + counterOffset);
    const hipStream_t hipStream = static_cast<hipStream_t>(stream);

    if (splitTiles > 0)
    {
        const unsigned int count = static_cast<unsigned int>(splitTiles);
        const unsigned int groups = (count + 255) / 256 < 1024 ? (count + 255) / 256 : 1024;
        hipLaunchKernelGGL(kttNoAllocClear, dim3(groups), dim3(256), 0, hipStream, counters, count);
    }
    const long fold = (groupsY + 65534) / 65535;
    const dim3 grid(static_cast<unsigned int>(groupsX * fold), static_cast<unsigned int>((groupsY + fold - 1) / fold),
        static_cast<unsigned int>(blocks));
    const float4* atomInfo = reinterpret_cast<const float4*>(atoms);
    const int split = static_cast<int>(splitFrom);
    if (wide)
    {
        hipLaunchKernelGGL(kttNoAllocCompensatedWide, grid, dim3(128, 2), 0, hipStream, atomInfo, atomCount, spacing, gridSize,
            energyGrid, partials, counters, split);
    }
    else
    {
        hipLaunchKernelGGL(kttNoAllocCompensatedNarrow, grid, dim3(64, 2), 0, hipStream, atomInfo, atomCount, spacing, gridSize,
            energyGrid, partials, counters, split);
    }
    const hipError_t status = hipGetLastError();
    if (status != hipSuccess)
    {
        return FailHip(message, messageLength, "launch of the compensated Coulomb kernel", status);
    }
    return 0;
}
