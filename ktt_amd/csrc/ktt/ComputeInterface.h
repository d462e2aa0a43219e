/** @file ComputeInterface.h
  * Interface handed to kernel launchers (Source/Api/ComputeInterface.h:30-350), plus the
  * initializer for user-provided HIP streams (Source/Api/ComputeApiInitializer.h:26-38).
  */
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include <DimensionVector.h>
#include <KernelConfiguration.h>
#include <KttTypes.h>

namespace ktt
{

class ComputeInterface
{
public:
    virtual ~ComputeInterface() = default;

    virtual void RunKernel(const KernelDefinitionId id) = 0;
    virtual void RunKernel(const KernelDefinitionId id, const DimensionVector& globalSize, const DimensionVector& localSize) = 0;
    virtual ComputeActionId RunKernelAsync(const KernelDefinitionId id, const QueueId queue) = 0;
    virtual ComputeActionId RunKernelAsync(const KernelDefinitionId id, const QueueId queue, const DimensionVector& globalSize,
        const DimensionVector& localSize) = 0;
    virtual void WaitForComputeAction(const ComputeActionId id) = 0;

    virtual void RunKernelWithProfiling(const KernelDefinitionId id) = 0;
    virtual void RunKernelWithProfiling(const KernelDefinitionId id, const DimensionVector& globalSize,
        const DimensionVector& localSize) = 0;
    virtual uint64_t GetRemainingProfilingRuns(const KernelDefinitionId id) const = 0;
    virtual uint64_t GetRemainingProfilingRuns() const = 0;

    virtual QueueId GetDefaultQueue() const = 0;
    virtual std::vector<QueueId> GetAllQueues() const = 0;
    virtual void SynchronizeQueue(const QueueId queue) = 0;
    virtual void SynchronizeQueues() = 0;
    virtual void SynchronizeDevice() = 0;

    virtual const DimensionVector& GetCurrentGlobalSize(const KernelDefinitionId id) const = 0;
    virtual const DimensionVector& GetCurrentLocalSize(const KernelDefinitionId id) const = 0;
    virtual const KernelConfiguration& GetCurrentConfiguration() const = 0;
    virtual KernelRunMode GetRunMode() const = 0;

    virtual void ChangeArguments(const KernelDefinitionId id, const std::vector<ArgumentId>& arguments) = 0;
    virtual void SwapArguments(const KernelDefinitionId id, const ArgumentId& first, const ArgumentId& second) = 0;
    virtual void UpdateScalarArgument(const ArgumentId& id, const void* data) = 0;
    virtual void UpdateLocalArgument(const ArgumentId& id, const size_t dataSize) = 0;

    virtual void UploadBuffer(const ArgumentId& id) = 0;
    virtual TransferActionId UploadBufferAsync(const ArgumentId& id, const QueueId queue) = 0;
    virtual void DownloadBuffer(const ArgumentId& id, void* destination, const size_t dataSize = 0) = 0;
    virtual TransferActionId DownloadBufferAsync(const ArgumentId& id, const QueueId queue, void* destination,
        const size_t dataSize = 0) = 0;
    virtual void UpdateBuffer(const ArgumentId& id, const void* data, const size_t dataSize = 0) = 0;
    virtual TransferActionId UpdateBufferAsync(const ArgumentId& id, const QueueId queue, const void* data,
        const size_t dataSize = 0) = 0;
    virtual void CopyBuffer(const ArgumentId& destination, const ArgumentId& source, const size_t dataSize = 0) = 0;
    virtual TransferActionId CopyBufferAsync(const ArgumentId& destination, const ArgumentId& source, const QueueId queue,
        const size_t dataSize = 0) = 0;
    virtual void WaitForTransferAction(const TransferActionId id) = 0;
    virtual void ResizeBuffer(const ArgumentId& id, const size_t newDataSize, const bool preserveData) = 0;
    virtual void ClearBuffer(const ArgumentId& id) = 0;
    virtual bool HasBuffer(const ArgumentId& id) = 0;
    virtual void GetUnifiedMemoryBufferHandle(const ArgumentId& id, UnifiedBufferMemory& memoryHandle) = 0;
    virtual bool GetProfiling(const KernelId id) = 0;
};

/** User-provided context (unused by HIP: one device per process context) and queues (hipStream_t). */
class ComputeApiInitializer
{
public:
    explicit ComputeApiInitializer(ComputeContext context, const std::vector<ComputeQueue>& queues) :
        m_Context(context),
        m_Queues(queues)
    {}

    ComputeContext GetContext() const { return m_Context; }
    const std::vector<ComputeQueue>& GetQueues() const { return m_Queues; }

private:
    ComputeContext m_Context;
    std::vector<ComputeQueue> m_Queues;
};

} // namespace ktt
