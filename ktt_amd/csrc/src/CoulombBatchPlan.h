#pragma once

// Host-side plan of a batch of Coulomb systems (include/ktt_hip.h section 5): the system
// records the persistent kernel reads, the per-bucket tile shapes, and the work list of
// (system, tile) items in longest-processing-time order. Needs no device, so the CPU tests
// check its invariants (every tile of every system listed once, the count the device list is
// sized for).

#include <cstddef>
#include <vector>

namespace ktt
{
namespace batch
{

constexpr int kBatchThreads = 256;
constexpr int kMaxBuckets = 5;

// Layout shared with struct BatchSystem in kernels/CoulombBatched.hip.
struct HostSystem
{
    int atomOffset;
    int atomCount;
    int gridSize;
    int bucket;
    float spacing;
    int reserved;
    long long gridOffset;
};
static_assert(sizeof(HostSystem) == 32, "BatchSystem layout");

struct BucketShape
{
    int wgx;
    int px;
    int rows;
    int tile;
    int unroll = 1;  // unroll of the per-tile atom loop

    int Span() const { return wgx * px; }
    int RowsPerGroup() const { return (kBatchThreads / wgx) * rows; }
};

/** Number of (system, tile) items: per system, ceil(G / span) x ceil(G^2 / rows per group) of its
 *  bucket's shape. */
size_t CountWorkItems(const std::vector<HostSystem>& systems, const std::vector<BucketShape>& shapes);

/** The work list as flattened int4 records (system, xBase, rowBase, 0), costliest first (computed
 *  points of the padded tile times staged atoms, plus the stored points); equal costs keep
 *  system and tile order. Its length is 4 x CountWorkItems. */
std::vector<int> BuildWorkItems(const std::vector<HostSystem>& systems, const std::vector<BucketShape>& shapes);

/** Throws KttException unless flat holds whole int4 records and at most capacity of them. */
void CheckItemCapacity(const std::vector<int>& flat, size_t capacity);

} // namespace batch
} // namespace ktt
