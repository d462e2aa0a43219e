/** @file Tuner.h
  * ktt::Tuner, the public facade (Source/Tuner.h:59-1010 of the reference).
  * Method names, arguments and error behaviour follow the reference: every public method
  * catches KttException, logs it and returns an Invalid* value. This build implements the
  * HIP compute API (ComputeApi::HIP); constructing a tuner for any other API fails loudly.
  * Python-script parameters, constraints, modifiers and generators, and XML output are not
  * provided (they require the embedded Python interpreter of the reference).
  */
#pragma once

#include <memory>
#include <ostream>
#include <string>
#include <type_traits>
#include <vector>

#include <ComputeInterface.h>
#include <DimensionVector.h>
#include <Info.h>
#include <KernelConfiguration.h>
#include <KernelResult.h>
#include <KttException.h>
#include <KttTypes.h>
#include <Searcher.h>
#include <StopCondition.h>

namespace ktt
{

class TunerCore;

class Tuner
{
public:
    explicit Tuner(const PlatformIndex platform, const DeviceIndex device, const ComputeApi api);
    explicit Tuner(const PlatformIndex platform, const DeviceIndex device, const ComputeApi api, const uint32_t computeQueueCount);
    explicit Tuner(const ComputeApi api, const ComputeApiInitializer& initializer);
    explicit Tuner(const ComputeApi api, const ComputeApiInitializer& initializer, std::vector<QueueId>& assignedQueueIds);
    ~Tuner();

    // Kernel definitions and kernels
    KernelDefinitionId AddKernelDefinition(const std::string& name, const std::string& source, const DimensionVector& globalSize,
        const DimensionVector& localSize, const std::vector<std::string>& typeNames = {});
    KernelDefinitionId AddKernelDefinitionFromFile(const std::string& name, const std::string& filePath,
        const DimensionVector& globalSize, const DimensionVector& localSize, const std::vector<std::string>& typeNames = {});
    KernelDefinitionId GetKernelDefinitionId(const std::string& name, const std::vector<std::string>& typeNames = {}) const;
    void RemoveKernelDefinition(const KernelDefinitionId id);
    void SetArguments(const KernelDefinitionId id, const std::vector<ArgumentId>& argumentIds);
    KernelId CreateSimpleKernel(const std::string& name, const KernelDefinitionId definitionId);
    KernelId CreateCompositeKernel(const std::string& name, const std::vector<KernelDefinitionId>& definitionIds,
        KernelLauncher launcher = nullptr);
    void RemoveKernel(const KernelId id);
    void SetLauncher(const KernelId id, KernelLauncher launcher);

    // Tuning parameters, modifiers, constraints
    template <typename T>
    void AddParameter(const KernelId id, const std::string& name, const std::vector<T>& values, const std::string& group = "");
    void AddThreadModifier(const KernelId id, const std::vector<KernelDefinitionId>& definitionIds, const ModifierType type,
        const ModifierDimension dimension, const std::vector<std::string>& parameters, ModifierFunction function);
    void AddThreadModifier(const KernelId id, const std::vector<KernelDefinitionId>& definitionIds, const ModifierType type,
        const ModifierDimension dimension, const std::string& parameter, const ModifierAction action);
    void AddConstraint(const KernelId id, const std::vector<std::string>& parameters, ConstraintFunction function);
    void AddGenericConstraint(const KernelId id, const std::vector<std::string>& parameters, GenericConstraintFunction function);
    void SetProfiledDefinitions(const KernelId id, const std::vector<KernelDefinitionId>& definitionIds);

    // Arguments
    template <typename T>
    ArgumentId AddArgumentVector(const std::vector<T>& data, const ArgumentAccessType accessType, const ArgumentId& customId = "");
    template <typename T>
    ArgumentId AddArgumentVector(std::vector<T>& data, const ArgumentAccessType accessType, const ArgumentMemoryLocation memoryLocation,
        const ArgumentManagementType managementType, const bool referenceUserData, const ArgumentId& customId = "");
    template <typename T>
    ArgumentId AddArgumentVector(ComputeBuffer buffer, const size_t bufferSize, const ArgumentAccessType accessType,
        const ArgumentMemoryLocation memoryLocation, const ArgumentId& customId = "");
    ArgumentId AddArgumentVector(ComputeBuffer buffer, const size_t bufferSize, const size_t elementSize,
        const ArgumentAccessType accessType, const ArgumentMemoryLocation memoryLocation, const ArgumentId& customId = "");
    ArgumentId AddArgumentVectorFromFile(const std::string& filePath, const ArgumentDataType dataType, const size_t elementSize,
        const ArgumentAccessType accessType, const ArgumentMemoryLocation memoryLocation = ArgumentMemoryLocation::Device,
        const ArgumentManagementType managementType = ArgumentManagementType::Framework, const ArgumentId& customId = "");
    void SaveArgumentVector(const ArgumentId& id, const std::string& filePath);
    template <typename T>
    ArgumentId AddArgumentScalar(const T& data, const ArgumentId& customId = "");
    ArgumentId AddArgumentScalar(const void* data, const size_t dataSize, const ArgumentId& customId = "");
    template <typename T>
    ArgumentId AddArgumentLocal(const size_t localMemorySize, const ArgumentId& customId = "");
    /** Value copied into the kernel's __device__ or __constant__ variable symbolName when a
     *  configuration's kernel is loaded (Source/Tuner.h:501-510); the variable takes no kernel
     *  parameter slot. On HIP the symbol name is required. */
    template <typename T>
    ArgumentId AddArgumentSymbol(const T& data, const ArgumentId& customId = "", const std::string& symbolName = "");
    void RemoveArgument(const ArgumentId& id);
    void SetReadOnlyArgumentCache(const bool flag);

    // Running and validation
    KernelResult Run(const KernelId id, const KernelConfiguration& configuration, const std::vector<BufferOutputDescriptor>& output);
    KernelResult Run(const KernelId id, const KernelConfiguration& configuration, const KernelDimensions& dimensions,
        const std::vector<BufferOutputDescriptor>& output);
    void SetProfiling(const bool flag);
    bool GetProfiling();
    void SetValidationMethod(const ValidationMethod method, const double toleranceThreshold);
    void SetValidationMode(const ValidationMode mode);
    void SetValidationRange(const ArgumentId& id, const size_t range);
    void SetValueComparator(const ArgumentId& id, ValueComparator comparator);
    void SetReferenceComputation(const ArgumentId& id, ReferenceComputation computation);
    void SetReferenceKernel(const ArgumentId& id, const KernelId referenceId, const KernelConfiguration& configuration,
        const KernelDimensions& dimensions = {});
    void SetReferenceArgument(const ArgumentId& id, const ArgumentId& referenceId);

    // Tuning
    std::vector<KernelResult> Tune(const KernelId id, std::unique_ptr<StopCondition> stopCondition = nullptr);
    std::vector<KernelResult> Tune(const KernelId id, const KernelDimensions& dimensions,
        std::unique_ptr<StopCondition> stopCondition = nullptr);
    KernelResult TuneIteration(const KernelId id, const std::vector<BufferOutputDescriptor>& output,
        const bool recomputeReference = false);
    KernelResult TuneIteration(const KernelId id, const KernelDimensions& dimensions, const std::vector<BufferOutputDescriptor>& output,
        const bool recomputeReference = false);
    std::vector<KernelResult> SimulateTuning(const KernelId id, const std::vector<KernelResult>& results,
        std::unique_ptr<StopCondition> stopCondition = nullptr);
    void SetSearcher(const KernelId id, std::unique_ptr<Searcher> searcher);
    /** Profile-based searcher (Source/Tuner.h:738-748, see ProfileBasedSearcher.h). `modelPath`
     *  names KTT JSON results with counters of the same kernel space, the searcher's knowledge
     *  base (empty: the searcher learns from its own profiled runs); pickled models are not read.
     *  `useBuiltinModule` has no effect (the searcher is native code). */
    void SetProfileBasedSearcher(const KernelId id, const std::string& modelPath, const bool useBuiltinModule = true,
        const uint32_t batchSize = 5, const uint32_t neighborSize = 100, const uint32_t randomSize = 10);
    void InitializeConfigurationData(const KernelId id);
    void ClearConfigurationData(const KernelId id);
    uint64_t GetConfigurationsCount(const KernelId id) const;
    KernelConfiguration GetBestConfiguration(const KernelId id) const;
    KernelConfiguration CreateConfiguration(const KernelId id, const ParameterInput& parameters) const;
    std::string GetKernelSource(const KernelId id, const KernelConfiguration& configuration) const;
    std::string GetKernelDefinitionSource(const KernelDefinitionId id, const KernelConfiguration& configuration) const;

    // Results
    static void SetTimeUnit(const TimeUnit unit);
    void SaveResults(const std::vector<KernelResult>& results, const std::string& filePath, const OutputFormat format,
        const UserData& data = {}) const;
    std::vector<KernelResult> LoadResults(const std::string& filePath, const OutputFormat format) const;
    std::vector<KernelResult> LoadResults(const std::string& filePath, const OutputFormat format, UserData& data) const;

    // Queues and synchronisation
    QueueId AddComputeQueue(ComputeQueue queue);
    void RemoveComputeQueue(const QueueId id);
    void WaitForComputeAction(const ComputeActionId id);
    void WaitForTransferAction(const TransferActionId id);
    void SynchronizeQueue(const QueueId id);
    void SynchronizeQueues();
    void SynchronizeDevice();

    // Engine settings and information
    void SetProfilingCounters(const std::vector<std::string>& counters);
    void SetCompilerOptions(const std::string& options, const bool overrideDefault = false);
    void SetGlobalSizeType(const GlobalSizeType type);
    void SetAutomaticGlobalSizeCorrection(const bool flag);
    void SetKernelCacheCapacity(const uint64_t capacity);
    std::vector<PlatformInfo> GetPlatformInfo() const;
    std::vector<DeviceInfo> GetDeviceInfo(const PlatformIndex platform) const;
    DeviceInfo GetCurrentDeviceInfo() const;

    static void SetLoggingLevel(const LoggingLevel level);
    static LoggingLevel GetLoggingLevel();
    static void SetLoggingTarget(std::ostream& outputTarget);
    static void SetLoggingTarget(const std::string& filePath);

private:
    std::unique_ptr<TunerCore> m_Tuner;

    /** Where a new argument's data comes from: a private copy, the caller's memory, or a device buffer. */
    enum class DataSource
    {
        Copy,
        Reference,
        DeviceBuffer
    };

    ArgumentId RegisterArgument(const size_t elementSize, const ArgumentDataType dataType, const ArgumentMemoryLocation location,
        const ArgumentAccessType access, const ArgumentMemoryType kind, const ArgumentManagementType management, const DataSource source,
        void* data, const size_t bytes, const ArgumentId& name);
    ArgumentId RegisterSymbolArgument(const size_t elementSize, const ArgumentDataType dataType, const void* data, const size_t bytes,
        const ArgumentId& name, const std::string& symbolName);
    void AddParameterValues(const KernelId id, const std::string& name, const std::vector<ParameterValue>& values,
        const std::string& group);

    template <typename T>
    static ArgumentDataType ElementTypeOf();
};

} // namespace ktt

#include <Tuner.inl>
