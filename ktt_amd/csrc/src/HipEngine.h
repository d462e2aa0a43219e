// ComputeApi::HIP backend of the compute-engine boundary.
//
// Behaviour follows the reference's CUDA engine (Source/ComputeEngine/Cuda/CudaEngine.cpp),
// re-expressed for HIP on gfx950:
//   * kernels compile through hipRTC on a kernel-cache miss; the configuration's
//     parameters are passed as -DNAME=VALUE options; defaults are
//     --offload-arch=<device gfx arch> -O3 plus the user's options (setting options
//     clears the cache, CudaEngine.cpp:741-757); LRU capacity 10 (CudaEngine.cpp:33);
//   * code objects load with hipModuleLoadData, launch with hipModuleLaunchKernel,
//     are timed with hipEvents recorded around the launch on the queue's stream
//     (CudaKernel.cpp:43-70);
//   * buffers are hipMalloc'd (device), hipHostMalloc'd (host / host zero-copy), or
//     hipMallocManaged (unified); transfers are hipMemcpyAsync timed by events;
//   * local size above the device limit, or LDS above the per-workgroup limit, raise
//     DeviceLimitsExceeded (CudaEngine.cpp:115-135); hipRTC errors raise CompilerError.
#pragma once

#include <hip/hip_runtime_api.h>

#include <list>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include <ComputeInterface.h>

#include "ComputeEngine.h"

namespace ktt
{

void CheckHipError(const hipError_t value, const std::string& call);

/** For cleanup paths whose failure cannot be acted on. */
inline void IgnoreHipError(const hipError_t value)
{
    static_cast<void>(value);
}

/** A compiled and loaded kernel. */
class HipKernel
{
public:
    HipKernel(const std::string& name, const std::vector<char>& codeObject, const std::string& loweredName,
        const Nanoseconds compilationTime);
    ~HipKernel();
    HipKernel(const HipKernel&) = delete;
    HipKernel& operator=(const HipKernel&) = delete;

    const std::string& GetName() const { return m_Name; }
    hipFunction_t GetFunction() const { return m_Function; }
    const KernelCompilationData& GetCompilationData() const { return m_CompilationData; }
    Nanoseconds GetCompilationTime() const { return m_CompilationTime; }
    size_t GetStaticLdsSize() const { return m_CompilationData.m_LocalMemorySize; }
    /** Copies bytes of host data into the module's global variable loweredName (a symbol
     *  argument); the variable must hold at least that many bytes. */
    void FillSymbol(const std::string& loweredName, const void* data, const size_t bytes) const;

private:
    std::string m_Name;
    hipModule_t m_Module;
    hipFunction_t m_Function;
    KernelCompilationData m_CompilationData;
    Nanoseconds m_CompilationTime;
};

/** hipRTC front end shared by the engine and the direct C entry points. */
class HipCompiler
{
public:
    /** Compiles source for the given options; returns the code object and sets loweredName. Throws CompilerError.
     *  With KTT_HIP_CACHE_DIR set, code objects are cached on disk under that directory. */
    static std::vector<char> Compile(const std::string& source, const std::string& programName, const std::string& kernelName,
        const std::vector<std::string>& options, std::string& loweredName, std::string* log = nullptr);
    /** The same, also lowering the names of global variables (symbol arguments): loweredSymbols
     *  receives the name of each symbols[i] in the code object. */
    static std::vector<char> Compile(const std::string& source, const std::string& programName, const std::string& kernelName,
        const std::vector<std::string>& options, const std::vector<std::string>& symbols, std::string& loweredName,
        std::vector<std::string>& loweredSymbols, std::string* log = nullptr);
    /** The hipRTC compilation itself, without the cache. */
    static std::vector<char> CompileWithRtc(const std::string& source, const std::string& programName, const std::string& kernelName,
        const std::vector<std::string>& options, std::string& loweredName, std::string* log = nullptr,
        const std::vector<std::string>& symbols = {}, std::vector<std::string>* loweredSymbols = nullptr);
    /** Default options for the current device (or gfx950 when none is visible). */
    static std::vector<std::string> DefaultOptions(const std::string& architecture);
    /** Architecture of the device without feature suffixes, e.g. "gfx950". */
    static std::string DeviceArchitecture(const int device);
    /** Marketing name of the device; when the runtime reports none (some containers hide it),
     *  hipDeviceGetName's, else "AMD GPU (<architecture>)". */
    static std::string DeviceName(const int device);
};

class HipEngine : public ComputeEngine
{
public:
    explicit HipEngine(const PlatformIndex platformIndex, const DeviceIndex deviceIndex, const uint32_t queueCount);
    explicit HipEngine(const ComputeApiInitializer& initializer, std::vector<QueueId>& assignedQueueIds);
    ~HipEngine() override;

    ComputeActionId RunKernelAsync(const KernelComputeData& data, const QueueId queueId, const bool powerMeasurementAllowed = true) override;
    ComputationResult WaitForComputeAction(const ComputeActionId id) override;
    void ClearData(const KernelComputeId& id) override;
    void ClearKernelData(const std::string& kernelName) override;

    ComputationResult RunKernelWithProfiling(const KernelComputeData& data, const QueueId queueId) override;
    void SetProfilingCounters(const std::vector<std::string>& counters) override;
    bool IsProfilingSessionActive(const KernelComputeId& id) override;
    uint64_t GetRemainingProfilingRuns(const KernelComputeId& id) override;
    bool HasAccurateRemainingProfilingRuns() const override;
    bool SupportsMultiInstanceProfiling() const override;
    bool IsProfilingActive() const override;
    void SetProfiling(const bool profiling) override;

    TransferActionId UploadArgument(KernelArgument& kernelArgument, const QueueId queueId) override;
    TransferActionId UpdateArgument(const ArgumentId& id, const QueueId queueId, const void* data, const size_t dataSize) override;
    TransferActionId DownloadArgument(const ArgumentId& id, const QueueId queueId, void* destination, const size_t dataSize) override;
    TransferActionId CopyArgument(const ArgumentId& destination, const QueueId queueId, const ArgumentId& source,
        const size_t dataSize) override;
    TransferResult WaitForTransferAction(const TransferActionId id) override;
    void ResizeArgument(const ArgumentId& id, const size_t newSize, const bool preserveData) override;
    void GetUnifiedMemoryBufferHandle(const ArgumentId& id, UnifiedBufferMemory& handle) override;
    void AddCustomBuffer(KernelArgument& kernelArgument, ComputeBuffer buffer) override;
    void ClearBuffer(const ArgumentId& id) override;
    void ClearBuffers() override;
    bool HasBuffer(const ArgumentId& id) override;

    QueueId AddComputeQueue(ComputeQueue queue) override;
    void RemoveComputeQueue(const QueueId id) override;
    QueueId GetDefaultQueue() const override;
    std::vector<QueueId> GetAllQueues() const override;
    void SynchronizeQueue(const QueueId queueId) override;
    void SynchronizeQueues() override;
    void SynchronizeDevice() override;

    std::vector<PlatformInfo> GetPlatformInfo() const override;
    std::vector<DeviceInfo> GetDeviceInfo(const PlatformIndex platformIndex) const override;
    PlatformInfo GetCurrentPlatformInfo() const override;
    DeviceInfo GetCurrentDeviceInfo() const override;
    ComputeApi GetComputeApi() const override;
    GlobalSizeType GetGlobalSizeType() const override;

    void SetCompilerOptions(const std::string& options, const bool overrideDefault = false) override;
    void SetGlobalSizeType(const GlobalSizeType type) override;
    void SetAutomaticGlobalSizeCorrection(const bool flag) override;
    void SetKernelCacheCapacity(const uint64_t capacity) override;
    void ClearKernelCache() override;
    void EnsureThreadContext() override;

    /** Device pointer of a buffer argument (for launchers that drive kernels themselves). */
    void* GetDeviceBuffer(const ArgumentId& id);
    hipStream_t GetStream(const QueueId id) const;

private:
    struct Buffer
    {
        void* devicePointer = nullptr;
        void* hostPointer = nullptr;      // pinned allocation for host locations
        size_t size = 0;
        ArgumentMemoryLocation location = ArgumentMemoryLocation::Device;
        bool userOwned = false;
    };

    struct ComputeAction
    {
        std::shared_ptr<HipKernel> kernel;
        hipEvent_t start = nullptr;
        hipEvent_t end = nullptr;
        DimensionVector globalSize;
        DimensionVector localSize;
        Nanoseconds overhead = 0;
        Nanoseconds compilationOverhead = 0;
    };

    struct TransferAction
    {
        hipEvent_t start = nullptr;
        hipEvent_t end = nullptr;
        Nanoseconds overhead = 0;
    };

    int m_DeviceIndex;
    hipDeviceProp_t m_Properties;
    std::string m_Architecture;
    std::map<QueueId, hipStream_t> m_Streams;
    std::map<QueueId, bool> m_StreamOwned;
    QueueId m_NextQueueId;
    std::map<ArgumentId, Buffer> m_Buffers;
    std::map<ComputeActionId, ComputeAction> m_ComputeActions;
    std::map<TransferActionId, TransferAction> m_TransferActions;
    ComputeActionId m_NextComputeActionId;
    TransferActionId m_NextTransferActionId;
    std::list<std::pair<KernelComputeId, std::shared_ptr<HipKernel>>> m_KernelCache;
    uint64_t m_KernelCacheCapacity;
    std::string m_CompilerOptions;
    bool m_OverrideDefaultOptions;
    GlobalSizeType m_GlobalSizeType;
    bool m_GlobalSizeCorrection;
    bool m_Profiling;
    std::vector<std::string> m_ProfilingCounters;  // empty: the service's default counters

    // Multi-pass collection of one launch: a timing run, then one run per counter pass.
    struct ProfilingSession
    {
        std::vector<std::string> counters;
        size_t passes = 0;
        size_t runsDone = 0;           // 0 before the timing run
        Nanoseconds duration = 0;      // of the timing run
        std::vector<double> values;    // indexed like counters
    };
    std::map<KernelComputeId, ProfilingSession> m_ProfilingSessions;
    std::vector<std::string> m_PlannedCounters;  // counter list of the service's current plan
    std::vector<size_t> m_PassOf;

    // Clock warm-up after a compilation: the engine clock falls while the GPU waits for the
    // compiler, so before the first launch of a freshly compiled kernel a small arithmetic kernel
    // runs for this long (KTT_HIP_CLOCK_WARMUP_MS, default 50; 0 turns it off).
    double m_ClockWarmupMs;
    std::shared_ptr<HipKernel> m_WarmupKernel;
    void* m_WarmupOut = nullptr;

    void Initialize();
    void WarmClock(hipStream_t stream);
    ComputationResult RunWithStaticProfile(const KernelComputeData& data, const QueueId queueId);
    void AddStaticCounters(const KernelComputeData& data, const KernelCompilationData& compilation, KernelProfilingData& profiling) const;
    void EnsurePlan(const std::vector<std::string>& counters);
    std::shared_ptr<HipKernel> LoadKernel(const KernelComputeData& data, bool& compiled);
    std::vector<std::string> BuildOptions(const KernelComputeData& data) const;
    Buffer& GetBuffer(const ArgumentId& id);
    void FreeBuffer(Buffer& buffer);
    TransferActionId RecordTransfer(hipEvent_t start, hipEvent_t end, const Nanoseconds overhead);
};

} // namespace ktt
