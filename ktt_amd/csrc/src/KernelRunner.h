// Running one kernel configuration: buffer setup with read-only caching, the launcher,
// result generation and validation (Source/KernelRunner/KernelRunner.cpp:23-343,
// ComputeLayer.cpp:17-180, ComputeLayerData.cpp:124-156, ResultValidator.cpp:16-270,
// ResultValidator.inl:53-108, ValidationData.cpp:17-191 of the reference).
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include <ComputeInterface.h>
#include <KernelResult.h>

#include "ComputeEngine.h"
#include "Kernel.h"
#include "KernelArgument.h"

namespace ktt
{

class ComputeLayerData
{
public:
    explicit ComputeLayerData(const Kernel& kernel, const KernelConfiguration& configuration, const KernelDimensions& dimensions,
        const KernelRunMode mode);

    void AddPartialResult(const ComputationResult& result);
    void IncreaseOverhead(const Nanoseconds overhead);
    void AddArgumentOverride(const ArgumentId& id, const KernelArgument& argument);
    void ChangeArguments(const KernelDefinitionId id, std::vector<KernelArgument*> arguments);
    void SwapArguments(const KernelDefinitionId id, const ArgumentId& first, const ArgumentId& second);

    const Kernel& GetKernel() const { return m_Kernel; }
    const KernelConfiguration& GetConfiguration() const { return m_Configuration; }
    KernelRunMode GetRunMode() const { return m_RunMode; }
    const KernelComputeData& GetComputeData(const KernelDefinitionId id) const;
    KernelComputeData& GetComputeData(const KernelDefinitionId id);
    KernelResult GenerateResult(const Nanoseconds launcherDuration) const;

private:
    const Kernel& m_Kernel;
    KernelConfiguration m_Configuration;
    KernelRunMode m_RunMode;
    std::map<KernelDefinitionId, KernelComputeData> m_ComputeData;
    std::map<ArgumentId, std::unique_ptr<KernelArgument>> m_Overrides;
    std::vector<ComputationResult> m_PartialResults;
    Nanoseconds m_DataOverhead;

    Nanoseconds CalculateLauncherOverhead() const;
};

class ComputeLayer : public ComputeInterface
{
public:
    explicit ComputeLayer(ComputeEngine& engine, KernelArgumentManager& argumentManager);

    void RunKernel(const KernelDefinitionId id) override;
    void RunKernel(const KernelDefinitionId id, const DimensionVector& globalSize, const DimensionVector& localSize) override;
    ComputeActionId RunKernelAsync(const KernelDefinitionId id, const QueueId queue) override;
    ComputeActionId RunKernelAsync(const KernelDefinitionId id, const QueueId queue, const DimensionVector& globalSize,
        const DimensionVector& localSize) override;
    void WaitForComputeAction(const ComputeActionId id) override;
    void RunKernelWithProfiling(const KernelDefinitionId id) override;
    void RunKernelWithProfiling(const KernelDefinitionId id, const DimensionVector& globalSize, const DimensionVector& localSize) override;
    uint64_t GetRemainingProfilingRuns(const KernelDefinitionId id) const override;
    uint64_t GetRemainingProfilingRuns() const override;
    QueueId GetDefaultQueue() const override;
    std::vector<QueueId> GetAllQueues() const override;
    void SynchronizeQueue(const QueueId queue) override;
    void SynchronizeQueues() override;
    void SynchronizeDevice() override;
    const DimensionVector& GetCurrentGlobalSize(const KernelDefinitionId id) const override;
    const DimensionVector& GetCurrentLocalSize(const KernelDefinitionId id) const override;
    const KernelConfiguration& GetCurrentConfiguration() const override;
    KernelRunMode GetRunMode() const override;
    void ChangeArguments(const KernelDefinitionId id, const std::vector<ArgumentId>& arguments) override;
    void SwapArguments(const KernelDefinitionId id, const ArgumentId& first, const ArgumentId& second) override;
    void UpdateScalarArgument(const ArgumentId& id, const void* data) override;
    void UpdateLocalArgument(const ArgumentId& id, const size_t dataSize) override;
    void UploadBuffer(const ArgumentId& id) override;
    TransferActionId UploadBufferAsync(const ArgumentId& id, const QueueId queue) override;
    void DownloadBuffer(const ArgumentId& id, void* destination, const size_t dataSize = 0) override;
    TransferActionId DownloadBufferAsync(const ArgumentId& id, const QueueId queue, void* destination, const size_t dataSize = 0) override;
    void UpdateBuffer(const ArgumentId& id, const void* data, const size_t dataSize = 0) override;
    TransferActionId UpdateBufferAsync(const ArgumentId& id, const QueueId queue, const void* data, const size_t dataSize = 0) override;
    void CopyBuffer(const ArgumentId& destination, const ArgumentId& source, const size_t dataSize = 0) override;
    TransferActionId CopyBufferAsync(const ArgumentId& destination, const ArgumentId& source, const QueueId queue,
        const size_t dataSize = 0) override;
    void WaitForTransferAction(const TransferActionId id) override;
    void ResizeBuffer(const ArgumentId& id, const size_t newDataSize, const bool preserveData) override;
    void ClearBuffer(const ArgumentId& id) override;
    bool HasBuffer(const ArgumentId& id) override;
    void GetUnifiedMemoryBufferHandle(const ArgumentId& id, UnifiedBufferMemory& memoryHandle) override;
    bool GetProfiling(const KernelId id) override;

    void SetActiveKernel(const KernelId id);
    void ClearActiveKernel();
    void AddData(const Kernel& kernel, const KernelConfiguration& configuration, const KernelDimensions& dimensions,
        const KernelRunMode mode);
    void ClearData(const KernelId id);
    void ClearComputeEngineData();
    KernelResult GenerateResult(const KernelId id, const Nanoseconds launcherDuration) const;

private:
    ComputeEngine& m_ComputeEngine;
    KernelArgumentManager& m_ArgumentManager;
    std::map<KernelId, std::unique_ptr<ComputeLayerData>> m_Data;
    KernelId m_ActiveKernel;

    ComputeLayerData& GetData();
    const ComputeLayerData& GetData() const;
};

class KernelRunner;

class ValidationData
{
public:
    explicit ValidationData(KernelRunner& kernelRunner, const KernelArgument& argument);

    void SetValidationRange(const size_t range);
    void SetValueComparator(ValueComparator comparator);
    void SetReferenceComputation(ReferenceComputation computation);
    void SetReferenceKernel(const Kernel& kernel, const KernelConfiguration& configuration, const KernelDimensions& dimensions);
    void SetReferenceArgument(const KernelArgument& argument);

    size_t GetValidationRange() const { return m_ValidationRange; }
    ValueComparator GetValueComparator() const { return m_Comparator; }
    bool HasValueComparator() const { return static_cast<bool>(m_Comparator); }
    bool HasReferenceComputation() const { return static_cast<bool>(m_ReferenceComputation); }
    bool HasReferenceKernel() const { return m_ReferenceKernel != nullptr; }
    bool HasReferenceArgument() const { return m_ReferenceArgument != nullptr; }
    KernelId GetReferenceKernelId() const;

    void ComputeReferenceResults();
    void ClearReferenceResults();
    bool HasReferenceResults() const;
    const void* GetReferenceResult() const;

private:
    KernelRunner& m_KernelRunner;
    const KernelArgument& m_Argument;
    size_t m_ValidationRange;
    ValueComparator m_Comparator;
    ReferenceComputation m_ReferenceComputation;
    const Kernel* m_ReferenceKernel;
    KernelConfiguration m_ReferenceConfiguration;
    KernelDimensions m_ReferenceDimensions;
    const KernelArgument* m_ReferenceArgument;
    std::vector<uint8_t> m_ReferenceResult;
    std::vector<uint8_t> m_ReferenceKernelResult;

    void ResetReferenceData();
};

class ResultValidator
{
public:
    explicit ResultValidator(KernelRunner& kernelRunner);

    void SetValidationMethod(const ValidationMethod method, const double toleranceThreshold);
    void SetValidationMode(const ValidationMode mode);
    void InitializeValidationData(const KernelArgument& argument);
    void SetValidationRange(const ArgumentId& id, const size_t range);
    void SetValueComparator(const ArgumentId& id, ValueComparator comparator);
    void SetReferenceComputation(const ArgumentId& id, ReferenceComputation computation);
    void SetReferenceKernel(const ArgumentId& id, const Kernel& kernel, const KernelConfiguration& configuration,
        const KernelDimensions& dimensions);
    void SetReferenceArgument(const ArgumentId& id, const KernelArgument& argument);
    bool HasValidationData(const ArgumentId& id) const;
    void RemoveValidationData(const ArgumentId& id);
    void RemoveDataWithReferenceKernel(const KernelId id);

    void ComputeReferenceResult(const Kernel& kernel, const KernelRunMode runMode);
    void ClearReferenceResult(const Kernel& kernel);
    bool HasReferenceResult(const Kernel& kernel) const;
    bool ValidateArguments(const Kernel& kernel, const KernelRunMode runMode) const;

private:
    KernelRunner& m_KernelRunner;
    std::map<ArgumentId, std::unique_ptr<ValidationData>> m_ValidationData;
    double m_ToleranceThreshold;
    ValidationMethod m_ValidationMethod;
    ValidationMode m_ValidationMode;

    bool IsRunModeValidated(const KernelRunMode mode) const;
    bool ValidateArgument(const KernelArgument& argument) const;
    template <typename T>
    bool ValidateTyped(const KernelArgument& argument, const T* result, const T* reference, const size_t range) const;
};

class KernelRunner
{
public:
    explicit KernelRunner(ComputeEngine& engine, KernelArgumentManager& argumentManager);

    KernelResult RunKernel(const Kernel& kernel, const KernelConfiguration& configuration, const KernelDimensions& dimensions,
        const KernelRunMode mode, const std::vector<BufferOutputDescriptor>& output, const bool manageBuffers = true);
    void SetupBuffers(const Kernel& kernel);
    void CleanupBuffers(const Kernel& kernel);
    void DownloadBuffers(const std::vector<BufferOutputDescriptor>& output);

    void SetReadOnlyArgumentCache(const bool flag);
    void SetProfiling(const bool flag);
    bool IsProfilingActive() const;
    void SetValidationMethod(const ValidationMethod method, const double toleranceThreshold);
    void SetValidationMode(const ValidationMode mode);
    void SetValidationRange(const ArgumentId& id, const size_t range);
    void SetValueComparator(const ArgumentId& id, ValueComparator comparator);
    void SetReferenceComputation(const ArgumentId& id, ReferenceComputation computation);
    void SetReferenceKernel(const ArgumentId& id, const Kernel& kernel, const KernelConfiguration& configuration,
        const KernelDimensions& dimensions);
    void SetReferenceArgument(const ArgumentId& id, const KernelArgument& argument);
    void ClearReferenceResult(const Kernel& kernel);
    void RemoveKernelData(const KernelId id);
    void RemoveValidationData(const ArgumentId& id);

private:
    std::unique_ptr<ComputeLayer> m_ComputeLayer;
    std::unique_ptr<ResultValidator> m_Validator;
    ComputeEngine& m_Engine;
    KernelArgumentManager& m_ArgumentManager;
    bool m_ReadOnlyCacheFlag;

    KernelLauncher GetKernelLauncher(const Kernel& kernel);
    KernelResult RunKernelInternal(const Kernel& kernel, const KernelConfiguration& configuration, const KernelDimensions& dimensions,
        const KernelRunMode mode, KernelLauncher launcher, const std::vector<BufferOutputDescriptor>& output);
    void PrepareValidationData(const ArgumentId& id);
    void ValidateResult(const Kernel& kernel, KernelResult& result, const KernelRunMode mode);
    static ResultStatus GetStatusFromException(const ExceptionReason reason);
};

} // namespace ktt
