// The compute-engine boundary (Source/ComputeEngine/ComputeEngine.h:19-81 of the
// reference): kernel runs, profiling, buffers, queues, information and settings.
// KernelComputeData carries one launch (the reference's KernelComputeData plays this role):
// name, source, the configuration's defines, global and local sizes after the thread
// modifiers, bound arguments, and the key under which the engine caches the compiled kernel.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include <DimensionVector.h>
#include <Info.h>
#include <KernelConfiguration.h>
#include <KernelResult.h>
#include <KttTypes.h>

#include "KernelArgument.h"

namespace ktt
{

class KernelDefinition;

/** Everything the engine needs for one launch of one kernel definition. */
class KernelComputeData
{
public:
    KernelComputeData(const KernelDefinition& definition, const KernelConfiguration& configuration, const DimensionVector& globalSize,
        const DimensionVector& localSize);

    void SetGlobalSize(const DimensionVector& globalSize) { m_Global = globalSize; }
    void SetLocalSize(const DimensionVector& localSize) { m_Local = localSize; }
    void SetArguments(const std::vector<KernelArgument*>& arguments) { m_Arguments = arguments; }
    /** Exchanges the positions of two arguments; throws if either is not bound. */
    void SwapArguments(const ArgumentId& first, const ArgumentId& second);

    const std::string& GetName() const { return m_Name; }
    const std::string& GetTemplatedName() const { return m_Templated; }
    /** Source without the configuration's defines (the engine passes those as -D options). */
    const std::string& GetDefaultSource() const { return m_Source; }
    const std::string& GetConfigurationPrefix() const { return m_Defines; }
    std::string GetSource() const { return m_Defines + m_Source; }
    const std::string& GetSourceDirectory() const { return m_Directory; }
    /** Kernel-cache key: the templated name and the configuration's defines. */
    KernelComputeId GetUniqueIdentifier() const { return m_Templated + "\n" + m_Defines; }
    const DimensionVector& GetGlobalSize() const { return m_Global; }
    const DimensionVector& GetLocalSize() const { return m_Local; }
    const KernelConfiguration& GetConfiguration() const { return m_Configuration; }
    const std::vector<KernelArgument*>& GetArguments() const { return m_Arguments; }

private:
    std::string m_Name;
    std::string m_Templated;
    std::string m_Source;
    std::string m_Defines;
    std::string m_Directory;
    KernelConfiguration m_Configuration;
    DimensionVector m_Global;
    DimensionVector m_Local;
    std::vector<KernelArgument*> m_Arguments;
};

class TransferResult
{
public:
    TransferResult(const Nanoseconds duration = 0, const Nanoseconds overhead = 0) :
        m_Duration(duration),
        m_Overhead(overhead)
    {}

    Nanoseconds GetDuration() const { return m_Duration; }
    Nanoseconds GetOverhead() const { return m_Overhead; }

private:
    Nanoseconds m_Duration;
    Nanoseconds m_Overhead;
};

class ComputeEngine
{
public:
    virtual ~ComputeEngine() = default;

    // Kernel methods
    virtual ComputeActionId RunKernelAsync(const KernelComputeData& data, const QueueId queueId, const bool powerMeasurementAllowed = true) = 0;
    virtual ComputationResult WaitForComputeAction(const ComputeActionId id) = 0;
    virtual void ClearData(const KernelComputeId& id) = 0;
    virtual void ClearKernelData(const std::string& kernelName) = 0;

    // Profiling methods
    virtual ComputationResult RunKernelWithProfiling(const KernelComputeData& data, const QueueId queueId) = 0;
    virtual void SetProfilingCounters(const std::vector<std::string>& counters) = 0;
    virtual bool IsProfilingSessionActive(const KernelComputeId& id) = 0;
    virtual uint64_t GetRemainingProfilingRuns(const KernelComputeId& id) = 0;
    virtual bool HasAccurateRemainingProfilingRuns() const = 0;
    virtual bool SupportsMultiInstanceProfiling() const = 0;
    virtual bool IsProfilingActive() const = 0;
    virtual void SetProfiling(const bool profiling) = 0;

    // Buffer methods
    virtual TransferActionId UploadArgument(KernelArgument& kernelArgument, const QueueId queueId) = 0;
    virtual TransferActionId UpdateArgument(const ArgumentId& id, const QueueId queueId, const void* data, const size_t dataSize) = 0;
    virtual TransferActionId DownloadArgument(const ArgumentId& id, const QueueId queueId, void* destination, const size_t dataSize) = 0;
    virtual TransferActionId CopyArgument(const ArgumentId& destination, const QueueId queueId, const ArgumentId& source,
        const size_t dataSize) = 0;
    virtual TransferResult WaitForTransferAction(const TransferActionId id) = 0;
    virtual void ResizeArgument(const ArgumentId& id, const size_t newSize, const bool preserveData) = 0;
    virtual void GetUnifiedMemoryBufferHandle(const ArgumentId& id, UnifiedBufferMemory& handle) = 0;
    virtual void AddCustomBuffer(KernelArgument& kernelArgument, ComputeBuffer buffer) = 0;
    virtual void ClearBuffer(const ArgumentId& id) = 0;
    virtual void ClearBuffers() = 0;
    virtual bool HasBuffer(const ArgumentId& id) = 0;

    // Queue methods
    virtual QueueId AddComputeQueue(ComputeQueue queue) = 0;
    virtual void RemoveComputeQueue(const QueueId id) = 0;
    virtual QueueId GetDefaultQueue() const = 0;
    virtual std::vector<QueueId> GetAllQueues() const = 0;
    virtual void SynchronizeQueue(const QueueId queueId) = 0;
    virtual void SynchronizeQueues() = 0;
    virtual void SynchronizeDevice() = 0;

    // Information retrieval methods
    virtual std::vector<PlatformInfo> GetPlatformInfo() const = 0;
    virtual std::vector<DeviceInfo> GetDeviceInfo(const PlatformIndex platformIndex) const = 0;
    virtual PlatformInfo GetCurrentPlatformInfo() const = 0;
    virtual DeviceInfo GetCurrentDeviceInfo() const = 0;
    virtual ComputeApi GetComputeApi() const = 0;
    virtual GlobalSizeType GetGlobalSizeType() const = 0;

    // Utility methods
    virtual void SetCompilerOptions(const std::string& options, const bool overrideDefault = false) = 0;
    virtual void SetGlobalSizeType(const GlobalSizeType type) = 0;
    virtual void SetAutomaticGlobalSizeCorrection(const bool flag) = 0;
    virtual void SetKernelCacheCapacity(const uint64_t capacity) = 0;
    virtual void ClearKernelCache() = 0;
    virtual void EnsureThreadContext() = 0;
};

} // namespace ktt
