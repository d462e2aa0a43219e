// Batched Coulomb systems (include/ktt_hip.h section 5): packing of many independent systems,
// size buckets, the longest-processing-time work list, the persistent queue kernel captured in
// a HIP graph, the per-system baseline, and per-bucket autotuning through ktt::Tuner.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <set>
#include <vector>

#include <Ktt.h>

#include "../../../include/ktt_hip.h"
#include "Common.h"
#include "CoulombBatchPlan.h"
#include "CoulombRuntime.h"
#include "HipEngine.h"
#include "Json.h"

namespace ktt
{

namespace
{

using batch::BucketShape;
using batch::HostSystem;
using batch::kBatchThreads;
using batch::kMaxBuckets;

// Defaults for buckets 0..4 (the kernel's own defaults; the host always passes them explicitly):
// the fastest shapes of a bucket tuning run (MCMC, 24 configurations per bucket) on the
// 1,024-system batch of grid sides 32, 48, 64, 96 and 128 on MI355X with the phased inner loop:
// 244.4 ms against 248.0 ms for the shapes tuned with the interleaved loop.
const BucketShape kDefaultShapes[kMaxBuckets] = {{32, 1, 4, 64, 8}, {16, 3, 4, 64, 8}, {32, 2, 8, 64, 8}, {16, 3, 4, 128, 16},
    {64, 2, 8, 128, 8}};

// Single-system kernel configuration of the per-system baseline, by grid side. Every parameter
// is given, so that the baseline does not follow the defaults chosen for 256^3 grids: atom tiles
// of 128 with the loop unrolled 8 times are the fastest of those measured on the 1,024-system
// batch (unroll 16, the 256^3 default, is slower on these small grids). One atom block and, for
// the compensated form, no tail split (TAIL_SPLIT=0): each system is one launch of whole tiles,
// with no FP64 partials (a small system has fewer tiles than the GPU holds at once, so the
// library default would split every one of them).
std::string PlainConfigForSide(const int side)
{
    const std::string rest = ",ATOM_TILE=128,INNER_UNROLL_FACTOR=8,ATOM_BLOCKS=1,TAIL_SPLIT=0";
    if (side % 128 == 0) return "WORK_GROUP_SIZE_X=64,WORK_GROUP_SIZE_Y=4,POINTS_X=2,ROWS_PER_THREAD=8" + rest;
    if (side % 96 == 0) return "WORK_GROUP_SIZE_X=32,WORK_GROUP_SIZE_Y=8,POINTS_X=3,ROWS_PER_THREAD=8" + rest;
    if (side % 64 == 0) return "WORK_GROUP_SIZE_X=32,WORK_GROUP_SIZE_Y=8,POINTS_X=2,ROWS_PER_THREAD=8" + rest;
    if (side % 48 == 0) return "WORK_GROUP_SIZE_X=16,WORK_GROUP_SIZE_Y=16,POINTS_X=3,ROWS_PER_THREAD=8" + rest;
    return "WORK_GROUP_SIZE_X=32,WORK_GROUP_SIZE_Y=8,POINTS_X=1,ROWS_PER_THREAD=8" + rest;
}

void CheckShape(const BucketShape& shape, const int bucket)
{
    const std::string name = "bucket " + std::to_string(bucket);
    if (shape.wgx <= 0 || kBatchThreads % shape.wgx != 0 || shape.px <= 0 || shape.rows <= 0 || shape.tile <= 0
        || shape.unroll <= 0)
    {
        throw KttException(name + ": WGX must divide " + std::to_string(kBatchThreads) + " and PX, ROWS, TILE, UNROLL must be positive");
    }
    const int rows = shape.RowsPerGroup();
    if (kBatchThreads % rows != 0 && rows % kBatchThreads != 0)
    {
        throw KttException(name + ": rows per workgroup (" + std::to_string(rows) + ") and " + std::to_string(kBatchThreads)
            + " threads must divide one another");
    }
    const int stride = kBatchThreads > rows ? kBatchThreads / rows : 1;
    if (shape.tile % stride != 0)
    {
        throw KttException(name + ": TILE must be a multiple of " + std::to_string(stride));
    }
}

} // namespace

class CoulombBatch
{
public:
    CoulombBatch(const float* hostAtoms, const int* atomCounts, const int* gridSizes, const float* spacings, const int systemCount,
        const char* config);
    ~CoulombBatch();

    void Run(hipStream_t stream, const bool useGraph, float* kernelMs);
    void RunPerSystem(const int form, hipStream_t stream, float* kernelMs);
    void Zero(hipStream_t stream);
    void SetShapes(const std::vector<BucketShape>& shapes);

    float* Grids() const { return m_Grids; }
    size_t GridPoints() const { return m_GridPoints; }
    int SystemCount() const { return static_cast<int>(m_Systems.size()); }
    const std::vector<HostSystem>& Systems() const { return m_Systems; }
    int BucketCount() const { return static_cast<int>(m_BucketSides.size()); }
    const std::vector<int>& BucketSides() const { return m_BucketSides; }
    const std::vector<BucketShape>& Shapes() const { return m_Shapes; }
    /** Number of work items (each one int4 record in the device list). */
    int ItemCount() const { return static_cast<int>(m_Items.size() / 4); }
    long long Interactions() const;
    std::string ShapeString() const;

    // Device pointers for the tuner's user buffers.
    float* DeviceAtoms() const { return m_Atoms; }
    void* DeviceSystems() const { return m_DeviceSystems; }
    void* DeviceItems() const { return m_DeviceItems; }
    unsigned int* DeviceCounter() const { return m_Counter; }
    int PersistentGroups() const { return m_PersistentGroups; }
    std::vector<int> BuildItems(const std::vector<BucketShape>& shapes) const;
    void UploadItems(const std::vector<int>& items);
    size_t ItemCapacity() const { return m_ItemCapacity; }
    /** Grows the device work list (before a tuner takes its pointer). */
    void EnsureItemCapacity(const size_t items);
    /** Items of the whole batch when bucket uses its finest admissible tiling. */
    size_t WorstCaseItems(const int bucket, const int minimumSpan, const int minimumRows) const;

private:
    std::vector<HostSystem> m_Systems;
    std::vector<int> m_BucketSides;
    std::vector<BucketShape> m_Shapes;
    std::vector<int> m_Items;   // flattened int4 records
    size_t m_ItemCapacity;
    int m_GroupsPerCu;
    int m_PersistentGroups;
    float* m_Atoms;
    void* m_DeviceSystems;
    void* m_DeviceItems;
    float* m_Grids;
    unsigned int* m_Counter;
    size_t m_GridPoints;
    std::shared_ptr<HipKernel> m_Persistent;
    std::shared_ptr<HipKernel> m_Reset;
    hipGraph_t m_Graph;
    hipGraphExec_t m_GraphExec;
    hipStream_t m_GraphStream;

    void CompileKernels();
    void DestroyGraph();
    void LaunchKernels(hipStream_t stream);
};

CoulombBatch::CoulombBatch(const float* hostAtoms, const int* atomCounts, const int* gridSizes, const float* spacings,
    const int systemCount, const char* config) :
    m_ItemCapacity(0),
    m_GroupsPerCu(4),
    m_PersistentGroups(0),
    m_Atoms(nullptr),
    m_DeviceSystems(nullptr),
    m_DeviceItems(nullptr),
    m_Grids(nullptr),
    m_Counter(nullptr),
    m_GridPoints(0),
    m_Graph(nullptr),
    m_GraphExec(nullptr),
    m_GraphStream(nullptr)
{
    if (systemCount <= 0 || atomCounts == nullptr || gridSizes == nullptr || spacings == nullptr)
    {
        throw KttException("a batch needs at least one system and its atom counts, grid sizes and spacings");
    }
    auto& runtime = CoulombRuntime::Get();
    runtime.EnsureDevice();

    std::set<int> sides;
    long long totalAtoms = 0;
    for (int s = 0; s < systemCount; ++s)
    {
        if (atomCounts[s] < 0 || gridSizes[s] <= 0 || gridSizes[s] > 2048 || !(spacings[s] > 0.0f))
        {
            throw KttException("system " + std::to_string(s) + " has an invalid atom count, grid size or spacing");
        }
        sides.insert(gridSizes[s]);
        totalAtoms += atomCounts[s];
    }
    if (totalAtoms > 0 && hostAtoms == nullptr)
    {
        throw KttException("atom buffer is null");
    }
    if (totalAtoms > 0x7fffffffLL)
    {
        throw KttException("too many atoms in one batch");
    }
    if (sides.size() > static_cast<size_t>(kMaxBuckets))
    {
        throw KttException("a batch may hold at most " + std::to_string(kMaxBuckets) + " distinct grid sizes");
    }
    m_BucketSides.assign(sides.begin(), sides.end());

    const CoulombConfig parsed = CoulombConfig::Parse(config, CoulombConfig());
    m_GroupsPerCu = static_cast<int>(parsed.Get("GROUPS_PER_CU", 4));
    if (m_GroupsPerCu <= 0 || m_GroupsPerCu > 64)
    {
        throw KttException("GROUPS_PER_CU must be in [1, 64]");
    }
    for (int b = 0; b < kMaxBuckets; ++b)
    {
        const std::string prefix = "BUCKET" + std::to_string(b) + "_";
        BucketShape shape = kDefaultShapes[b];
        shape.wgx = static_cast<int>(parsed.Get(prefix + "WGX", shape.wgx));
        shape.px = static_cast<int>(parsed.Get(prefix + "PX", shape.px));
        shape.rows = static_cast<int>(parsed.Get(prefix + "ROWS", shape.rows));
        shape.tile = static_cast<int>(parsed.Get(prefix + "TILE", shape.tile));
        shape.unroll = static_cast<int>(parsed.Get(prefix + "UNROLL", shape.unroll));
        CheckShape(shape, b);
        m_Shapes.push_back(shape);
    }

    long long atomOffset = 0;
    long long gridOffset = 0;
    for (int s = 0; s < systemCount; ++s)
    {
        HostSystem system{};
        system.atomOffset = static_cast<int>(atomOffset);
        system.atomCount = atomCounts[s];
        system.gridSize = gridSizes[s];
        system.bucket = static_cast<int>(std::find(m_BucketSides.begin(), m_BucketSides.end(), gridSizes[s]) - m_BucketSides.begin());
        system.spacing = spacings[s];
        system.gridOffset = gridOffset;
        m_Systems.push_back(system);
        atomOffset += atomCounts[s];
        gridOffset += static_cast<long long>(gridSizes[s]) * gridSizes[s] * gridSizes[s];
    }
    m_GridPoints = static_cast<size_t>(gridOffset);

    // Item capacity for the configured shapes; the tuner grows it for its own space.
    m_ItemCapacity = std::max<size_t>(batch::CountWorkItems(m_Systems, m_Shapes), 1);

    CheckHipError(hipMalloc(reinterpret_cast<void**>(&m_Atoms), std::max<size_t>(static_cast<size_t>(atomOffset), 1) * 16), "hipMalloc(atoms)");
    CheckHipError(hipMalloc(&m_DeviceSystems, m_Systems.size() * sizeof(HostSystem)), "hipMalloc(systems)");
    CheckHipError(hipMalloc(&m_DeviceItems, m_ItemCapacity * 16), "hipMalloc(items)");
    CheckHipError(hipMalloc(reinterpret_cast<void**>(&m_Grids), m_GridPoints * sizeof(float)), "hipMalloc(grids)");
    CheckHipError(hipMalloc(reinterpret_cast<void**>(&m_Counter), 256), "hipMalloc(counter)");
    if (atomOffset > 0)
    {
        CheckHipError(hipMemcpy(m_Atoms, hostAtoms, static_cast<size_t>(atomOffset) * 16, hipMemcpyHostToDevice), "hipMemcpy(atoms)");
    }
    CheckHipError(hipMemcpy(m_DeviceSystems, m_Systems.data(), m_Systems.size() * sizeof(HostSystem), hipMemcpyHostToDevice),
        "hipMemcpy(systems)");
    CheckHipError(hipMemset(m_Counter, 0, 256), "hipMemset(counter)");
    CheckHipError(hipMemset(m_Grids, 0, m_GridPoints * sizeof(float)), "hipMemset(grids)");

    hipDeviceProp_t properties;
    int device = 0;
    CheckHipError(hipGetDevice(&device), "hipGetDevice");
    CheckHipError(hipGetDeviceProperties(&properties, device), "hipGetDeviceProperties");
    m_PersistentGroups = properties.multiProcessorCount * m_GroupsPerCu;

    UploadItems(BuildItems(m_Shapes));
    CompileKernels();
}

CoulombBatch::~CoulombBatch()
{
    DestroyGraph();
    IgnoreHipError(hipDeviceSynchronize());
    IgnoreHipError(hipFree(m_Atoms));
    IgnoreHipError(hipFree(m_DeviceSystems));
    IgnoreHipError(hipFree(m_DeviceItems));
    IgnoreHipError(hipFree(m_Grids));
    IgnoreHipError(hipFree(m_Counter));
}

std::vector<int> CoulombBatch::BuildItems(const std::vector<BucketShape>& shapes) const
{
    return batch::BuildWorkItems(m_Systems, shapes);
}

void CoulombBatch::EnsureItemCapacity(const size_t items)
{
    if (items <= m_ItemCapacity)
    {
        return;
    }
    DestroyGraph();
    CheckHipError(hipDeviceSynchronize(), "hipDeviceSynchronize");
    IgnoreHipError(hipFree(m_DeviceItems));
    m_DeviceItems = nullptr;
    CheckHipError(hipMalloc(&m_DeviceItems, items * 16), "hipMalloc(items)");
    m_ItemCapacity = items;
    UploadItems(m_Items);
}

size_t CoulombBatch::WorstCaseItems(const int bucket, const int minimumSpan, const int minimumRows) const
{
    size_t total = 0;
    for (size_t s = 0; s < m_Systems.size(); ++s)
    {
        const auto& system = m_Systems[s];
        const size_t g = static_cast<size_t>(system.gridSize);
        if (system.bucket == bucket)
        {
            total += ((g + minimumSpan - 1) / minimumSpan) * ((g * g + minimumRows - 1) / minimumRows);
        }
        else
        {
            const auto& shape = m_Shapes[static_cast<size_t>(system.bucket)];
            total += ((g + shape.Span() - 1) / shape.Span()) * ((g * g + shape.RowsPerGroup() - 1) / shape.RowsPerGroup());
        }
    }
    return total;
}

void CoulombBatch::UploadItems(const std::vector<int>& items)
{
    batch::CheckItemCapacity(items, m_ItemCapacity);
    m_Items = items;
    if (!items.empty())
    {
        CheckHipError(hipMemcpy(m_DeviceItems, items.data(), items.size() * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy(items)");
    }
}

void CoulombBatch::CompileKernels()
{
    CoulombConfig defines;
    for (int b = 0; b < kMaxBuckets; ++b)
    {
        const std::string prefix = "BUCKET" + std::to_string(b) + "_";
        defines.Set(prefix + "WGX", std::to_string(m_Shapes[b].wgx));
        defines.Set(prefix + "PX", std::to_string(m_Shapes[b].px));
        defines.Set(prefix + "ROWS", std::to_string(m_Shapes[b].rows));
        defines.Set(prefix + "TILE", std::to_string(m_Shapes[b].tile));
        defines.Set(prefix + "UNROLL", std::to_string(m_Shapes[b].unroll));
    }
    defines.Set("BATCH_THREADS", std::to_string(kBatchThreads));
    auto& runtime = CoulombRuntime::Get();
    m_Persistent = runtime.GetKernel("CoulombBatched.hip", "coulombBatchedPersistent", defines.Defines({}));
    m_Reset = runtime.GetKernel("CoulombBatched.hip", "coulombBatchedReset", {});
}

void CoulombBatch::SetShapes(const std::vector<BucketShape>& shapes)
{
    for (size_t b = 0; b < shapes.size(); ++b)
    {
        CheckShape(shapes[b], static_cast<int>(b));
    }
    DestroyGraph();
    m_Shapes = shapes;
    UploadItems(BuildItems(m_Shapes));
    CompileKernels();
}

std::string CoulombBatch::ShapeString() const
{
    std::string result;
    for (int b = 0; b < kMaxBuckets; ++b)
    {
        const std::string prefix = "BUCKET" + std::to_string(b) + "_";
        result += (result.empty() ? "" : ",") + prefix + "WGX=" + std::to_string(m_Shapes[b].wgx) + "," + prefix + "PX="
            + std::to_string(m_Shapes[b].px) + "," + prefix + "ROWS=" + std::to_string(m_Shapes[b].rows) + "," + prefix + "TILE="
            + std::to_string(m_Shapes[b].tile) + "," + prefix + "UNROLL=" + std::to_string(m_Shapes[b].unroll);
    }
    return result + ",GROUPS_PER_CU=" + std::to_string(m_GroupsPerCu);
}

long long CoulombBatch::Interactions() const
{
    long long total = 0;
    for (const auto& system : m_Systems)
    {
        total += static_cast<long long>(system.gridSize) * system.gridSize * system.gridSize * system.atomCount;
    }
    return total;
}

void CoulombBatch::DestroyGraph()
{
    if (m_GraphExec != nullptr)
    {
        IgnoreHipError(hipGraphExecDestroy(m_GraphExec));
        m_GraphExec = nullptr;
    }
    if (m_Graph != nullptr)
    {
        IgnoreHipError(hipGraphDestroy(m_Graph));
        m_Graph = nullptr;
    }
    m_GraphStream = nullptr;
}

void CoulombBatch::LaunchKernels(hipStream_t stream)
{
    unsigned int* counter = m_Counter;
    void* resetArguments[] = {&counter};
    const unsigned int one[3] = {1, 1, 1};
    const unsigned int resetBlock[3] = {64, 1, 1};
    CoulombRuntime::Launch(*m_Reset, one, resetBlock, resetArguments, stream, nullptr);

    const float* atoms = m_Atoms;
    const void* systems = m_DeviceSystems;
    const void* items = m_DeviceItems;
    int itemCount = ItemCount();
    // The kernel reads items[0, itemCount): they must all be in the uploaded list.
    if (itemCount < 0 || static_cast<size_t>(itemCount) * 4 != m_Items.size() || static_cast<size_t>(itemCount) > m_ItemCapacity)
    {
        throw KttException("batched work list is inconsistent with its device buffer");
    }
    float* grids = m_Grids;
    void* arguments[] = {&atoms, &systems, &items, &itemCount, &grids, &counter};
    const unsigned int groups = static_cast<unsigned int>(std::max(1, std::min(m_PersistentGroups, std::max(itemCount, 1))));
    const unsigned int grid[3] = {groups, 1, 1};
    const unsigned int block[3] = {static_cast<unsigned int>(kBatchThreads), 1, 1};
    CoulombRuntime::Launch(*m_Persistent, grid, block, arguments, stream, nullptr);
}

void CoulombBatch::Run(hipStream_t stream, const bool useGraph, float* kernelMs)
{
    hipEvent_t start = nullptr;
    hipEvent_t end = nullptr;
    if (kernelMs != nullptr)
    {
        CheckHipError(hipEventCreate(&start), "hipEventCreate");
        CheckHipError(hipEventCreate(&end), "hipEventCreate");
    }
    if (useGraph)
    {
        if (m_GraphExec == nullptr || m_GraphStream != stream)
        {
            DestroyGraph();
            // A private stream for capture: capturing the null stream is not allowed.
            hipStream_t capture = nullptr;
            CheckHipError(hipStreamCreateWithFlags(&capture, hipStreamNonBlocking), "hipStreamCreateWithFlags");
            CheckHipError(hipStreamBeginCapture(capture, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            try
            {
                LaunchKernels(capture);
            }
            catch (...)
            {
                hipGraph_t partial = nullptr;
                IgnoreHipError(hipStreamEndCapture(capture, &partial));
                if (partial != nullptr)
                {
                    IgnoreHipError(hipGraphDestroy(partial));
                }
                IgnoreHipError(hipStreamDestroy(capture));
                throw;
            }
            CheckHipError(hipStreamEndCapture(capture, &m_Graph), "hipStreamEndCapture");
            IgnoreHipError(hipStreamDestroy(capture));
            CheckHipError(hipGraphInstantiate(&m_GraphExec, m_Graph, nullptr, nullptr, 0), "hipGraphInstantiate");
            m_GraphStream = stream;
        }
        if (start != nullptr)
        {
            CheckHipError(hipEventRecord(start, stream), "hipEventRecord");
        }
        CheckHipError(hipGraphLaunch(m_GraphExec, stream), "hipGraphLaunch");
    }
    else
    {
        if (start != nullptr)
        {
            CheckHipError(hipEventRecord(start, stream), "hipEventRecord");
        }
        LaunchKernels(stream);
    }
    if (end != nullptr)
    {
        CheckHipError(hipEventRecord(end, stream), "hipEventRecord");
        const hipError_t sync = hipEventSynchronize(end);
        float milliseconds = 0.0f;
        if (sync == hipSuccess)
        {
            IgnoreHipError(hipEventElapsedTime(&milliseconds, start, end));
        }
        IgnoreHipError(hipEventDestroy(start));
        IgnoreHipError(hipEventDestroy(end));
        CheckHipError(sync, "hipEventSynchronize(batch)");
        *kernelMs = milliseconds;
    }
}

void CoulombBatch::RunPerSystem(const int form, hipStream_t stream, float* kernelMs)
{
    hipEvent_t start = nullptr;
    hipEvent_t end = nullptr;
    if (kernelMs != nullptr)
    {
        CheckHipError(hipEventCreate(&start), "hipEventCreate");
        CheckHipError(hipEventCreate(&end), "hipEventCreate");
        CheckHipError(hipEventRecord(start, stream), "hipEventRecord");
    }
    for (const auto& system : m_Systems)
    {
        const std::string config = PlainConfigForSide(system.gridSize);
        const int status = ktt_coulomb3d(m_Atoms + static_cast<size_t>(system.atomOffset) * 4, system.atomCount, system.spacing,
            system.gridSize, m_Grids + system.gridOffset, form, config.c_str(), stream, nullptr);
        if (status != 0)
        {
            throw KttException(std::string("per-system launch failed: ") + ktt_hip_last_error());
        }
    }
    if (end != nullptr)
    {
        CheckHipError(hipEventRecord(end, stream), "hipEventRecord");
        CheckHipError(hipEventSynchronize(end), "hipEventSynchronize");
        IgnoreHipError(hipEventElapsedTime(kernelMs, start, end));
        IgnoreHipError(hipEventDestroy(start));
        IgnoreHipError(hipEventDestroy(end));
    }
}

void CoulombBatch::Zero(hipStream_t stream)
{
    if (ktt_hip_zero(m_Grids, m_GridPoints * sizeof(float), stream) != 0)
    {
        throw KttException(std::string("zeroing the batch grids failed: ") + ktt_hip_last_error());
    }
}

// ---------------------------------------------------------------- tuning

/** Tunes the tile shape of one bucket: a ktt composite kernel (zero fill, queue reset,
 *  persistent kernel) over the systems of that bucket, with a launcher that rebuilds the
 *  work list for the configuration under test. */
static KernelResult TuneBucket(CoulombBatch& batch, const int bucket, const int searcher, const int maxConfigurations,
    const std::string& resultsFile, BucketShape& best, double& bestMs, int& tested)
{
    const std::string directory = CoulombRuntime::Get().GetKernelDirectory();
    // Finest tiling of the space below: span 16 x 1 points, (256 / 64) * 2 rows.
    batch.EnsureItemCapacity(batch.WorstCaseItems(bucket, 16, 8));
    Tuner tuner(0, 0, ComputeApi::HIP);
    tuner.SetGlobalSizeType(GlobalSizeType::CUDA);
    tuner.SetTimeUnit(TimeUnit::Microseconds);
    tuner.SetValidationMode(ValidationMode::None);

    const DimensionVector one(1, 1, 1);
    const KernelDefinitionId zero = tuner.AddKernelDefinitionFromFile("kttZeroFill4", directory + "/coulomb_util.hip",
        DimensionVector(1024), DimensionVector(256));
    const KernelDefinitionId reset = tuner.AddKernelDefinitionFromFile("coulombBatchedReset", directory + "/CoulombBatched.hip",
        one, DimensionVector(64));
    const KernelDefinitionId persistent = tuner.AddKernelDefinitionFromFile("coulombBatchedPersistent",
        directory + "/CoulombBatched.hip", DimensionVector(static_cast<size_t>(batch.PersistentGroups())),
        DimensionVector(static_cast<size_t>(kBatchThreads)));
    if (zero == InvalidKernelDefinitionId || reset == InvalidKernelDefinitionId || persistent == InvalidKernelDefinitionId)
    {
        throw KttException("unable to load the batched kernel sources from " + directory);
    }

    // Scalars first, then user buffers wrapping the batch's device memory.
    const uint64_t gridWords = batch.GridPoints();
    const ArgumentId wordsId = tuner.AddArgumentScalar(static_cast<unsigned long long>(gridWords));
    const ArgumentId itemCountId = tuner.AddArgumentScalar(batch.ItemCount());
    const ArgumentId gridsId = tuner.AddArgumentVector<float>(batch.Grids(), gridWords * sizeof(float), ArgumentAccessType::ReadWrite,
        ArgumentMemoryLocation::Device);
    const ArgumentId atomsId = tuner.AddArgumentVector<float>(batch.DeviceAtoms(), 16, ArgumentAccessType::ReadOnly,
        ArgumentMemoryLocation::Device);
    const ArgumentId systemsId = tuner.AddArgumentVector<int>(batch.DeviceSystems(), static_cast<size_t>(batch.SystemCount()) * 32,
        ArgumentAccessType::ReadOnly, ArgumentMemoryLocation::Device);
    const ArgumentId itemsId = tuner.AddArgumentVector<int>(batch.DeviceItems(), batch.ItemCapacity() * 16, ArgumentAccessType::ReadOnly,
        ArgumentMemoryLocation::Device);
    const ArgumentId counterId = tuner.AddArgumentVector<unsigned int>(batch.DeviceCounter(), 256, ArgumentAccessType::ReadWrite,
        ArgumentMemoryLocation::Device);
    tuner.SetArguments(zero, {gridsId, wordsId});
    tuner.SetArguments(reset, {counterId});
    tuner.SetArguments(persistent, {atomsId, systemsId, itemsId, itemCountId, gridsId, counterId});

    const KernelId kernel = tuner.CreateCompositeKernel("CoulombBatchedBucket" + std::to_string(bucket), {zero, reset, persistent});
    const std::string prefix = "BUCKET" + std::to_string(bucket) + "_";
    tuner.AddParameter(kernel, prefix + "WGX", std::vector<uint64_t>{16, 32, 64});
    tuner.AddParameter(kernel, prefix + "PX", std::vector<uint64_t>{1, 2, 3, 4});
    tuner.AddParameter(kernel, prefix + "ROWS", std::vector<uint64_t>{2, 4, 8});
    tuner.AddParameter(kernel, prefix + "TILE", std::vector<uint64_t>{32, 64, 128});
    tuner.AddParameter(kernel, prefix + "UNROLL", std::vector<uint64_t>{1, 8, 16});
    // Every other bucket keeps its current shape (single-valued parameters are defines too).
    for (int b = 0; b < kMaxBuckets; ++b)
    {
        if (b == bucket)
        {
            continue;
        }
        const std::string other = "BUCKET" + std::to_string(b) + "_";
        tuner.AddParameter(kernel, other + "WGX", std::vector<uint64_t>{static_cast<uint64_t>(batch.Shapes()[b].wgx)});
        tuner.AddParameter(kernel, other + "PX", std::vector<uint64_t>{static_cast<uint64_t>(batch.Shapes()[b].px)});
        tuner.AddParameter(kernel, other + "ROWS", std::vector<uint64_t>{static_cast<uint64_t>(batch.Shapes()[b].rows)});
        tuner.AddParameter(kernel, other + "TILE", std::vector<uint64_t>{static_cast<uint64_t>(batch.Shapes()[b].tile)});
        tuner.AddParameter(kernel, other + "UNROLL", std::vector<uint64_t>{static_cast<uint64_t>(batch.Shapes()[b].unroll)});
    }
    const int side = batch.BucketSides()[static_cast<size_t>(bucket)];
    tuner.AddConstraint(kernel, {prefix + "WGX", prefix + "PX", prefix + "ROWS", prefix + "TILE"},
        [side](const std::vector<uint64_t>& v)
        {
            const uint64_t span = v[0] * v[1];
            const uint64_t rows = (kBatchThreads / v[0]) * v[2];
            const uint64_t stride = kBatchThreads > rows ? kBatchThreads / rows : 1;
            const bool divides = (kBatchThreads % rows == 0 || rows % kBatchThreads == 0) && v[3] % stride == 0;
            // No more than a quarter of the lanes idle along x, at most 32 points per thread.
            const uint64_t blocks = (static_cast<uint64_t>(side) + span - 1) / span;
            return divides && blocks * span * 4 <= static_cast<uint64_t>(side) * 5 && v[1] * v[2] <= 32
                && v[3] * (rows + 4) * 4 <= 64 * 1024;
        });

    tuner.SetLauncher(kernel, [&batch, zero, reset, persistent, bucket, itemCountId](ComputeInterface& interface)
    {
        const auto& pairs = interface.GetCurrentConfiguration().GetPairs();
        std::vector<BucketShape> shapes = batch.Shapes();
        const std::string prefix = "BUCKET" + std::to_string(bucket) + "_";
        shapes[static_cast<size_t>(bucket)] = BucketShape{ParameterPair::GetParameterValue<int>(pairs, prefix + "WGX"),
            ParameterPair::GetParameterValue<int>(pairs, prefix + "PX"), ParameterPair::GetParameterValue<int>(pairs, prefix + "ROWS"),
            ParameterPair::GetParameterValue<int>(pairs, prefix + "TILE"), ParameterPair::GetParameterValue<int>(pairs, prefix + "UNROLL")};
        const auto items = batch.BuildItems(shapes);
        batch.UploadItems(items);
        int itemCount = static_cast<int>(items.size() / 4);
        interface.UpdateScalarArgument(itemCountId, &itemCount);
        const uint64_t zeroGroups = std::min<uint64_t>((batch.GridPoints() + 255) / 256, 8192);
        interface.RunKernel(zero, DimensionVector(zeroGroups), DimensionVector(256));
        interface.RunKernel(reset, DimensionVector(1), DimensionVector(64));
        const uint64_t groups = std::max<uint64_t>(1, std::min<uint64_t>(batch.PersistentGroups(), std::max(itemCount, 1)));
        interface.RunKernel(persistent, DimensionVector(groups), DimensionVector(kBatchThreads));
    });

    switch (searcher)
    {
    case KTT_SEARCHER_RANDOM:
        tuner.SetSearcher(kernel, std::make_unique<RandomSearcher>());
        break;
    case KTT_SEARCHER_MCMC:
        tuner.SetSearcher(kernel, std::make_unique<McmcSearcher>());
        break;
    default:
        tuner.SetSearcher(kernel, std::make_unique<DeterministicSearcher>());
        break;
    }
    std::unique_ptr<StopCondition> stop;
    if (maxConfigurations > 0)
    {
        stop = std::make_unique<ConfigurationCount>(static_cast<uint64_t>(maxConfigurations));
    }
    const auto results = tuner.Tune(kernel, std::move(stop));
    tested = static_cast<int>(results.size());
    if (!resultsFile.empty() && !results.empty())
    {
        tuner.SaveResults(results, resultsFile + "_bucket" + std::to_string(bucket), OutputFormat::JSON,
            {{"grid_side", std::to_string(side)}, {"systems", std::to_string(batch.SystemCount())}});
    }
    const KernelResult* winner = nullptr;
    for (const auto& result : results)
    {
        if (result.IsValid() && (winner == nullptr || result.GetKernelDuration() < winner->GetKernelDuration()))
        {
            winner = &result;
        }
    }
    // The bucket's shape before tuning competes too: a short random search may not beat it.
    const BucketShape incumbent = batch.Shapes()[static_cast<size_t>(bucket)];
    KernelResult incumbentResult;
    const ParameterInput incumbentInput = {{prefix + "WGX", static_cast<uint64_t>(incumbent.wgx)},
        {prefix + "PX", static_cast<uint64_t>(incumbent.px)}, {prefix + "ROWS", static_cast<uint64_t>(incumbent.rows)},
        {prefix + "TILE", static_cast<uint64_t>(incumbent.tile)}, {prefix + "UNROLL", static_cast<uint64_t>(incumbent.unroll)}};
    const auto inList = [](const int value, const std::initializer_list<int> list)
    {
        return std::find(list.begin(), list.end(), value) != list.end();
    };
    if (inList(incumbent.wgx, {16, 32, 64}) && inList(incumbent.px, {1, 2, 3, 4}) && inList(incumbent.rows, {2, 4, 8})
        && inList(incumbent.tile, {32, 64, 128}) && inList(incumbent.unroll, {1, 8, 16}))
    {
        incumbentResult = tuner.Run(kernel, tuner.CreateConfiguration(kernel, incumbentInput), {});
        if (incumbentResult.IsValid() && (winner == nullptr || incumbentResult.GetKernelDuration() < winner->GetKernelDuration()))
        {
            winner = &incumbentResult;
        }
    }
    if (winner == nullptr)
    {
        throw KttException("no configuration of bucket " + std::to_string(bucket) + " ran successfully");
    }
    const auto& pairs = winner->GetConfiguration().GetPairs();
    best = BucketShape{ParameterPair::GetParameterValue<int>(pairs, prefix + "WGX"), ParameterPair::GetParameterValue<int>(pairs, prefix + "PX"),
        ParameterPair::GetParameterValue<int>(pairs, prefix + "ROWS"), ParameterPair::GetParameterValue<int>(pairs, prefix + "TILE"),
        ParameterPair::GetParameterValue<int>(pairs, prefix + "UNROLL")};
    bestMs = static_cast<double>(winner->GetKernelDuration()) / 1.0e6;
    return *winner;
}


namespace batch
{

size_t CountWorkItems(const std::vector<HostSystem>& systems, const std::vector<BucketShape>& shapes)
{
    size_t total = 0;
    for (const auto& system : systems)
    {
        const auto& shape = shapes.at(static_cast<size_t>(system.bucket));
        const size_t g = static_cast<size_t>(system.gridSize);
        total += ((g + shape.Span() - 1) / shape.Span()) * ((g * g + shape.RowsPerGroup() - 1) / shape.RowsPerGroup());
    }
    return total;
}

std::vector<int> BuildWorkItems(const std::vector<HostSystem>& systems, const std::vector<BucketShape>& shapes)
{
    struct Item
    {
        long long cost;
        int system;
        int xBase;
        int rowBase;
    };
    std::vector<Item> list;
    list.reserve(CountWorkItems(systems, shapes));
    for (size_t s = 0; s < systems.size(); ++s)
    {
        const auto& system = systems[s];
        const auto& shape = shapes.at(static_cast<size_t>(system.bucket));
        const int g = system.gridSize;
        const long long rows = static_cast<long long>(g) * g;
        // Cost: computed points (padded lanes included) times staged atoms, plus the stored points.
        const long long tiles = (system.atomCount + shape.tile - 1) / shape.tile;
        for (int xBase = 0; xBase < g; xBase += shape.Span())
        {
            for (long long rowBase = 0; rowBase < rows; rowBase += shape.RowsPerGroup())
            {
                const long long width = std::min<long long>(shape.Span(), g - xBase);
                const long long height = std::min<long long>(shape.RowsPerGroup(), rows - rowBase);
                list.push_back({static_cast<long long>(shape.Span()) * shape.RowsPerGroup() * tiles * shape.tile + width * height,
                    static_cast<int>(s), xBase, static_cast<int>(rowBase)});
            }
        }
    }
    // Longest processing time first; ties keep a deterministic order.
    std::stable_sort(list.begin(), list.end(), [](const Item& a, const Item& b) { return a.cost > b.cost; });
    std::vector<int> flat;
    flat.reserve(list.size() * 4);
    for (const auto& item : list)
    {
        flat.push_back(item.system);
        flat.push_back(item.xBase);
        flat.push_back(item.rowBase);
        flat.push_back(0);
    }
    return flat;
}

void CheckItemCapacity(const std::vector<int>& flat, const size_t capacity)
{
    if (flat.size() % 4 != 0)
    {
        throw KttException("work list of " + std::to_string(flat.size()) + " ints is not a list of int4 records");
    }
    if (flat.size() / 4 > capacity)
    {
        throw KttException("work list of " + std::to_string(flat.size() / 4) + " items exceeds the batch capacity of "
            + std::to_string(capacity));
    }
}

} // namespace batch

} // namespace ktt

using namespace ktt;

#define KTT_C_TRY try {
#define KTT_C_CATCH } catch (const std::exception& exception) { return SetLastError(exception.what()); } catch (...) { return SetLastError("unknown error"); }

extern "C" {

int ktt_coulomb_batch_create(const float* hostAtoms, const int* atomCounts, const int* gridSizes, const float* spacings,
    int systemCount, const char* config, void** batch)
{
    KTT_C_TRY
    if (batch == nullptr)
    {
        return SetLastError("batch handle pointer is null");
    }
    *batch = new CoulombBatch(hostAtoms, atomCounts, gridSizes, spacings, systemCount, config);
    return 0;
    KTT_C_CATCH
}

int ktt_coulomb_batch_destroy(void* batch)
{
    KTT_C_TRY
    delete static_cast<CoulombBatch*>(batch);
    return 0;
    KTT_C_CATCH
}

int ktt_coulomb_batch_info(void* batch, float** grids, size_t* gridPoints, long long* gridOffsets, int* items,
    long long* interactions, char* shapes, size_t shapesLength)
{
    KTT_C_TRY
    if (batch == nullptr)
    {
        return SetLastError("batch is null");
    }
    const auto& value = *static_cast<CoulombBatch*>(batch);
    if (grids != nullptr)
    {
        *grids = value.Grids();
    }
    if (gridPoints != nullptr)
    {
        *gridPoints = value.GridPoints();
    }
    if (gridOffsets != nullptr)
    {
        for (int s = 0; s < value.SystemCount(); ++s)
        {
            gridOffsets[s] = value.Systems()[static_cast<size_t>(s)].gridOffset;
        }
    }
    if (items != nullptr)
    {
        *items = value.ItemCount();
    }
    if (interactions != nullptr)
    {
        *interactions = value.Interactions();
    }
    if (shapes != nullptr && shapesLength > 0)
    {
        std::snprintf(shapes, shapesLength, "%s", value.ShapeString().c_str());
    }
    return 0;
    KTT_C_CATCH
}

int ktt_coulomb_batch_zero(void* batch, void* stream)
{
    KTT_C_TRY
    static_cast<CoulombBatch*>(batch)->Zero(static_cast<hipStream_t>(stream));
    return 0;
    KTT_C_CATCH
}

int ktt_coulomb_batch_run(void* batch, int useGraph, void* stream, float* kernelMs)
{
    KTT_C_TRY
    static_cast<CoulombBatch*>(batch)->Run(static_cast<hipStream_t>(stream), useGraph != 0, kernelMs);
    return 0;
    KTT_C_CATCH
}

int ktt_coulomb_batch_run_per_system(void* batch, int form, void* stream, float* kernelMs)
{
    KTT_C_TRY
    static_cast<CoulombBatch*>(batch)->RunPerSystem(form, static_cast<hipStream_t>(stream), kernelMs);
    return 0;
    KTT_C_CATCH
}

int ktt_coulomb_batch_tune(void* batch, int searcher, int maxConfigurationsPerBucket, const char* resultsFile, char* bestConfig,
    size_t bestConfigLength)
{
    KTT_C_TRY
    if (batch == nullptr)
    {
        return SetLastError("batch is null");
    }
    auto& value = *static_cast<CoulombBatch*>(batch);
    const std::string prefix = resultsFile != nullptr ? resultsFile : "";
    std::vector<BucketShape> shapes = value.Shapes();
    JsonValue summary = JsonValue::MakeObject();
    JsonValue buckets = JsonValue::MakeArray();
    for (int bucket = 0; bucket < value.BucketCount(); ++bucket)
    {
        BucketShape best{};
        double bestMs = 0.0;
        int tested = 0;
        TuneBucket(value, bucket, searcher, maxConfigurationsPerBucket, prefix, best, bestMs, tested);
        shapes[static_cast<size_t>(bucket)] = best;
        // The remaining buckets are tuned with this bucket's winner in place.
        value.SetShapes(shapes);
        JsonValue entry = JsonValue::MakeObject();
        entry["bucket"] = bucket;
        entry["grid_side"] = value.BucketSides()[static_cast<size_t>(bucket)];
        entry["WGX"] = best.wgx;
        entry["PX"] = best.px;
        entry["ROWS"] = best.rows;
        entry["TILE"] = best.tile;
        entry["UNROLL"] = best.unroll;
        entry["kernel_ms"] = bestMs;
        entry["configurations_tested"] = tested;
        buckets.PushBack(entry);
    }
    summary["buckets"] = buckets;
    summary["config"] = value.ShapeString();
    if (!prefix.empty())
    {
        std::ofstream stream(prefix + "_winners.json");
        if (!stream.is_open())
        {
            return SetLastError("unable to write " + prefix + "_winners.json");
        }
        stream << summary.Dump(2) << "\n";
    }
    if (bestConfig != nullptr && bestConfigLength > 0)
    {
        std::snprintf(bestConfig, bestConfigLength, "%s", value.ShapeString().c_str());
    }
    return 0;
    KTT_C_CATCH
}

} // extern "C"
