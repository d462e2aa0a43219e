// ktt::Tuner facade. Every method forwards to TunerCore and, like Source/Tuner.cpp of the
// reference, catches KttException, logs it at error level and returns the Invalid* value
// (or an empty result) instead of throwing. Constructors throw, as in the reference.
#include <Ktt.h>

#include "Common.h"
#include "HipEngine.h"
#include "TunerCore.h"

namespace ktt
{

#define KTT_TRY try {
#define KTT_CATCH_RETURN(value) } catch (const KttException& exception) { TunerCore::Log(LoggingLevel::Error, exception.what()); return value; }
#define KTT_CATCH } catch (const KttException& exception) { TunerCore::Log(LoggingLevel::Error, exception.what()); }

Tuner::Tuner(const PlatformIndex platform, const DeviceIndex device, const ComputeApi api) :
    m_Tuner(std::make_unique<TunerCore>(platform, device, api, 1))
{}

Tuner::Tuner(const PlatformIndex platform, const DeviceIndex device, const ComputeApi api, const uint32_t computeQueueCount) :
    m_Tuner(std::make_unique<TunerCore>(platform, device, api, computeQueueCount))
{}

Tuner::Tuner(const ComputeApi api, const ComputeApiInitializer& initializer)
{
    std::vector<QueueId> ids;
    m_Tuner = std::make_unique<TunerCore>(api, initializer, ids);
}

Tuner::Tuner(const ComputeApi api, const ComputeApiInitializer& initializer, std::vector<QueueId>& assignedQueueIds) :
    m_Tuner(std::make_unique<TunerCore>(api, initializer, assignedQueueIds))
{}

Tuner::~Tuner() = default;

KernelDefinitionId Tuner::AddKernelDefinition(const std::string& name, const std::string& source, const DimensionVector& globalSize,
    const DimensionVector& localSize, const std::vector<std::string>& typeNames)
{
    KTT_TRY
    return m_Tuner->GetKernelManager().AddKernelDefinition(name, source, globalSize, localSize, typeNames);
    KTT_CATCH_RETURN(InvalidKernelDefinitionId)
}

KernelDefinitionId Tuner::AddKernelDefinitionFromFile(const std::string& name, const std::string& filePath,
    const DimensionVector& globalSize, const DimensionVector& localSize, const std::vector<std::string>& typeNames)
{
    KTT_TRY
    return m_Tuner->GetKernelManager().AddKernelDefinitionFromFile(name, filePath, globalSize, localSize, typeNames);
    KTT_CATCH_RETURN(InvalidKernelDefinitionId)
}

KernelDefinitionId Tuner::GetKernelDefinitionId(const std::string& name, const std::vector<std::string>& typeNames) const
{
    KTT_TRY
    return m_Tuner->GetKernelManager().GetDefinitionId(name, typeNames);
    KTT_CATCH_RETURN(InvalidKernelDefinitionId)
}

void Tuner::RemoveKernelDefinition(const KernelDefinitionId id)
{
    KTT_TRY
    m_Tuner->GetKernelManager().RemoveKernelDefinition(id);
    KTT_CATCH
}

void Tuner::SetArguments(const KernelDefinitionId id, const std::vector<ArgumentId>& argumentIds)
{
    KTT_TRY
    m_Tuner->GetKernelManager().SetArguments(id, argumentIds);
    KTT_CATCH
}

KernelId Tuner::CreateSimpleKernel(const std::string& name, const KernelDefinitionId definitionId)
{
    KTT_TRY
    return m_Tuner->GetKernelManager().CreateKernel(name, {definitionId});
    KTT_CATCH_RETURN(InvalidKernelId)
}

KernelId Tuner::CreateCompositeKernel(const std::string& name, const std::vector<KernelDefinitionId>& definitionIds,
    KernelLauncher launcher)
{
    KTT_TRY
    const KernelId id = m_Tuner->GetKernelManager().CreateKernel(name, definitionIds);
    if (launcher)
    {
        m_Tuner->GetKernelManager().SetLauncher(id, launcher);
    }
    return id;
    KTT_CATCH_RETURN(InvalidKernelId)
}

void Tuner::RemoveKernel(const KernelId id)
{
    KTT_TRY
    m_Tuner->RemoveKernel(id);
    KTT_CATCH
}

void Tuner::SetLauncher(const KernelId id, KernelLauncher launcher)
{
    KTT_TRY
    m_Tuner->GetKernelManager().SetLauncher(id, launcher);
    KTT_CATCH
}

void Tuner::AddThreadModifier(const KernelId id, const std::vector<KernelDefinitionId>& definitionIds, const ModifierType type,
    const ModifierDimension dimension, const std::vector<std::string>& parameters, ModifierFunction function)
{
    KTT_TRY
    m_Tuner->GetKernelManager().AddThreadModifier(id, definitionIds, type, dimension, parameters, function);
    KTT_CATCH
}

void Tuner::AddThreadModifier(const KernelId id, const std::vector<KernelDefinitionId>& definitionIds, const ModifierType type,
    const ModifierDimension dimension, const std::string& parameter, const ModifierAction action)
{
    // Actions as in Source/Tuner.cpp:194-233.
    ModifierFunction function;
    switch (action)
    {
    case ModifierAction::Add:
        function = [](const uint64_t size, const std::vector<uint64_t>& values) { return size + values[0]; };
        break;
    case ModifierAction::Subtract:
        function = [](const uint64_t size, const std::vector<uint64_t>& values) { return size - values[0]; };
        break;
    case ModifierAction::Multiply:
        function = [](const uint64_t size, const std::vector<uint64_t>& values) { return size * values[0]; };
        break;
    case ModifierAction::Divide:
        function = [](const uint64_t size, const std::vector<uint64_t>& values) { return size / values[0]; };
        break;
    case ModifierAction::DivideCeil:
        function = [](const uint64_t size, const std::vector<uint64_t>& values) { return (size + values[0] - 1) / values[0]; };
        break;
    }
    AddThreadModifier(id, definitionIds, type, dimension, std::vector<std::string>{parameter}, function);
}

void Tuner::AddConstraint(const KernelId id, const std::vector<std::string>& parameters, ConstraintFunction function)
{
    KTT_TRY
    m_Tuner->GetKernelManager().AddConstraint(id, parameters, function);
    KTT_CATCH
}

void Tuner::AddGenericConstraint(const KernelId id, const std::vector<std::string>& parameters, GenericConstraintFunction function)
{
    KTT_TRY
    m_Tuner->GetKernelManager().AddGenericConstraint(id, parameters, function);
    KTT_CATCH
}

void Tuner::SetProfiledDefinitions(const KernelId id, const std::vector<KernelDefinitionId>& definitionIds)
{
    KTT_TRY
    m_Tuner->GetKernelManager().SetProfiledDefinitions(id, definitionIds);
    KTT_CATCH
}

ArgumentId Tuner::AddArgumentVector(ComputeBuffer buffer, const size_t bufferSize, const size_t elementSize,
    const ArgumentAccessType accessType, const ArgumentMemoryLocation memoryLocation, const ArgumentId& customId)
{
    return AddUserArgument(buffer, elementSize, ArgumentDataType::Custom, memoryLocation, accessType, bufferSize, customId);
}

ArgumentId Tuner::AddArgumentVectorFromFile(const std::string& filePath, const ArgumentDataType dataType, const size_t elementSize,
    const ArgumentAccessType accessType, const ArgumentMemoryLocation memoryLocation, const ArgumentManagementType managementType,
    const ArgumentId& customId)
{
    KTT_TRY
    const std::string data = LoadFileToString(filePath);
    return m_Tuner->GetArgumentManager().AddArgumentWithOwnedData(elementSize, dataType, memoryLocation, accessType,
        ArgumentMemoryType::Vector, managementType, data.data(), data.size(), customId);
    KTT_CATCH_RETURN(InvalidArgumentId)
}

void Tuner::SaveArgumentVector(const ArgumentId& id, const std::string& filePath)
{
    KTT_TRY
    const auto& argument = m_Tuner->GetArgumentManager().GetArgument(id);
    std::ofstream stream(filePath, std::ios::binary);
    if (!stream.is_open())
    {
        throw KttException("Unable to open file: " + filePath);
    }
    stream.write(static_cast<const char*>(argument.GetData()), static_cast<std::streamsize>(argument.GetDataSize()));
    KTT_CATCH
}

ArgumentId Tuner::AddArgumentScalar(const void* data, const size_t dataSize, const ArgumentId& customId)
{
    return AddArgumentWithOwnedData(dataSize, ArgumentDataType::Custom, ArgumentMemoryLocation::Undefined, ArgumentAccessType::ReadOnly,
        ArgumentMemoryType::Scalar, ArgumentManagementType::Framework, data, dataSize, customId);
}

void Tuner::RemoveArgument(const ArgumentId& id)
{
    KTT_TRY
    m_Tuner->RemoveArgument(id);
    KTT_CATCH
}

void Tuner::SetReadOnlyArgumentCache(const bool flag)
{
    KTT_TRY
    m_Tuner->GetExecutor().KeepReadOnlyBuffers(flag);
    KTT_CATCH
}

KernelResult Tuner::Run(const KernelId id, const KernelConfiguration& configuration, const std::vector<BufferOutputDescriptor>& output)
{
    return Run(id, configuration, {}, output);
}

KernelResult Tuner::Run(const KernelId id, const KernelConfiguration& configuration, const KernelDimensions& dimensions,
    const std::vector<BufferOutputDescriptor>& output)
{
    KTT_TRY
    const auto& kernel = m_Tuner->GetKernelManager().GetKernel(id);
    return m_Tuner->GetExecutor().Execute(kernel, configuration, dimensions, KernelRunMode::Running, output);
    KTT_CATCH_RETURN(KernelResult())
}

void Tuner::SetProfiling(const bool flag)
{
    KTT_TRY
    m_Tuner->GetExecutor().SetProfiling(flag);
    KTT_CATCH
}

bool Tuner::GetProfiling()
{
    KTT_TRY
    return m_Tuner->GetExecutor().IsProfiling();
    KTT_CATCH_RETURN(false)
}

void Tuner::SetValidationMethod(const ValidationMethod method, const double toleranceThreshold)
{
    KTT_TRY
    m_Tuner->GetExecutor().Checker().SetMethod(method, toleranceThreshold);
    KTT_CATCH
}

void Tuner::SetValidationMode(const ValidationMode mode)
{
    KTT_TRY
    m_Tuner->GetExecutor().Checker().SetModes(mode);
    KTT_CATCH
}

void Tuner::SetValidationRange(const ArgumentId& id, const size_t range)
{
    KTT_TRY
    m_Tuner->GetExecutor().Checker().SetRange(m_Tuner->GetArgumentManager().GetArgument(id), range);
    KTT_CATCH
}

void Tuner::SetValueComparator(const ArgumentId& id, ValueComparator comparator)
{
    KTT_TRY
    m_Tuner->GetExecutor().Checker().SetComparator(m_Tuner->GetArgumentManager().GetArgument(id), comparator);
    KTT_CATCH
}

void Tuner::SetReferenceComputation(const ArgumentId& id, ReferenceComputation computation)
{
    KTT_TRY
    m_Tuner->GetExecutor().Checker().SetComputation(m_Tuner->GetArgumentManager().GetArgument(id), computation);
    KTT_CATCH
}

void Tuner::SetReferenceKernel(const ArgumentId& id, const KernelId referenceId, const KernelConfiguration& configuration,
    const KernelDimensions& dimensions)
{
    KTT_TRY
    const auto& kernel = m_Tuner->GetKernelManager().GetKernel(referenceId);
    m_Tuner->GetExecutor().Checker().SetKernel(m_Tuner->GetArgumentManager().GetArgument(id), kernel, configuration, dimensions);
    KTT_CATCH
}

void Tuner::SetReferenceArgument(const ArgumentId& id, const ArgumentId& referenceId)
{
    KTT_TRY
    const auto& argument = m_Tuner->GetArgumentManager().GetArgument(referenceId);
    m_Tuner->GetExecutor().Checker().SetArgument(m_Tuner->GetArgumentManager().GetArgument(id), argument);
    KTT_CATCH
}

std::vector<KernelResult> Tuner::Tune(const KernelId id, std::unique_ptr<StopCondition> stopCondition)
{
    return Tune(id, {}, std::move(stopCondition));
}

std::vector<KernelResult> Tuner::Tune(const KernelId id, const KernelDimensions& dimensions, std::unique_ptr<StopCondition> stopCondition)
{
    KTT_TRY
    const auto& kernel = m_Tuner->GetKernelManager().GetKernel(id);
    return m_Tuner->GetTuningLoop().Tune(kernel, dimensions, std::move(stopCondition));
    KTT_CATCH_RETURN(std::vector<KernelResult>{})
}

KernelResult Tuner::TuneIteration(const KernelId id, const std::vector<BufferOutputDescriptor>& output, const bool recomputeReference)
{
    return TuneIteration(id, {}, output, recomputeReference);
}

KernelResult Tuner::TuneIteration(const KernelId id, const KernelDimensions& dimensions, const std::vector<BufferOutputDescriptor>& output,
    const bool recomputeReference)
{
    KTT_TRY
    const auto& kernel = m_Tuner->GetKernelManager().GetKernel(id);
    return m_Tuner->GetTuningLoop().Step(kernel, dimensions, KernelRunMode::OnlineTuning, output, recomputeReference);
    KTT_CATCH_RETURN(KernelResult())
}

std::vector<KernelResult> Tuner::SimulateTuning(const KernelId id, const std::vector<KernelResult>& results,
    std::unique_ptr<StopCondition> stopCondition)
{
    KTT_TRY
    const auto& kernel = m_Tuner->GetKernelManager().GetKernel(id);
    return m_Tuner->GetTuningLoop().Replay(kernel, results, std::move(stopCondition));
    KTT_CATCH_RETURN(std::vector<KernelResult>{})
}

void Tuner::SetSearcher(const KernelId id, std::unique_ptr<Searcher> searcher)
{
    KTT_TRY
    if (!m_Tuner->GetKernelManager().HasKernel(id))
    {
        throw KttException("Attempting to set searcher for kernel with invalid id: " + std::to_string(id));
    }
    if (searcher == nullptr)
    {
        throw KttException("Searcher must be a valid object");
    }
    m_Tuner->GetTuningLoop().SetSearcher(id, std::move(searcher));
    KTT_CATCH
}

void Tuner::InitializeConfigurationData(const KernelId id)
{
    KTT_TRY
    m_Tuner->GetTuningLoop().Prepare(m_Tuner->GetKernelManager().GetKernel(id));
    KTT_CATCH
}

void Tuner::ClearConfigurationData(const KernelId id)
{
    KTT_TRY
    m_Tuner->GetTuningLoop().Reset(id);
    KTT_CATCH
}

uint64_t Tuner::GetConfigurationsCount(const KernelId id) const
{
    KTT_TRY
    return m_Tuner->GetTuningLoop().SpaceSize(id);
    KTT_CATCH_RETURN(0)
}

KernelConfiguration Tuner::GetBestConfiguration(const KernelId id) const
{
    KTT_TRY
    return m_Tuner->GetTuningLoop().Best(id);
    KTT_CATCH_RETURN(KernelConfiguration())
}

KernelConfiguration Tuner::CreateConfiguration(const KernelId id, const ParameterInput& parameters) const
{
    KTT_TRY
    return m_Tuner->GetKernelManager().GetKernel(id).CreateConfiguration(parameters);
    KTT_CATCH_RETURN(KernelConfiguration())
}

std::string Tuner::GetKernelSource(const KernelId id, const KernelConfiguration& configuration) const
{
    KTT_TRY
    return m_Tuner->GetKernelSource(id, configuration);
    KTT_CATCH_RETURN(std::string())
}

std::string Tuner::GetKernelDefinitionSource(const KernelDefinitionId id, const KernelConfiguration& configuration) const
{
    KTT_TRY
    return m_Tuner->GetKernelDefinitionSource(id, configuration);
    KTT_CATCH_RETURN(std::string())
}

void Tuner::SetTimeUnit(const TimeUnit unit)
{
    TimeConfiguration::GetInstance().SetTimeUnit(unit);
}

void Tuner::SaveResults(const std::vector<KernelResult>& results, const std::string& filePath, const OutputFormat format,
    const UserData& data) const
{
    KTT_TRY
    m_Tuner->SaveResults(results, filePath, format, data);
    KTT_CATCH
}

std::vector<KernelResult> Tuner::LoadResults(const std::string& filePath, const OutputFormat format) const
{
    UserData data;
    return LoadResults(filePath, format, data);
}

std::vector<KernelResult> Tuner::LoadResults(const std::string& filePath, const OutputFormat format, UserData& data) const
{
    KTT_TRY
    return m_Tuner->LoadResults(filePath, format, data);
    KTT_CATCH_RETURN(std::vector<KernelResult>{})
}

QueueId Tuner::AddComputeQueue(ComputeQueue queue)
{
    KTT_TRY
    return m_Tuner->GetComputeEngine().AddComputeQueue(queue);
    KTT_CATCH_RETURN(InvalidQueueId)
}

void Tuner::RemoveComputeQueue(const QueueId id)
{
    KTT_TRY
    m_Tuner->GetComputeEngine().RemoveComputeQueue(id);
    KTT_CATCH
}

void Tuner::WaitForComputeAction(const ComputeActionId id)
{
    KTT_TRY
    m_Tuner->GetComputeEngine().WaitForComputeAction(id);
    KTT_CATCH
}

void Tuner::WaitForTransferAction(const TransferActionId id)
{
    KTT_TRY
    m_Tuner->GetComputeEngine().WaitForTransferAction(id);
    KTT_CATCH
}

void Tuner::SynchronizeQueue(const QueueId id)
{
    KTT_TRY
    m_Tuner->GetComputeEngine().SynchronizeQueue(id);
    KTT_CATCH
}

void Tuner::SynchronizeQueues()
{
    KTT_TRY
    m_Tuner->GetComputeEngine().SynchronizeQueues();
    KTT_CATCH
}

void Tuner::SynchronizeDevice()
{
    KTT_TRY
    m_Tuner->GetComputeEngine().SynchronizeDevice();
    KTT_CATCH
}

void Tuner::SetProfilingCounters(const std::vector<std::string>& counters)
{
    KTT_TRY
    m_Tuner->GetComputeEngine().SetProfilingCounters(counters);
    KTT_CATCH
}

void Tuner::SetCompilerOptions(const std::string& options, const bool overrideDefault)
{
    KTT_TRY
    m_Tuner->GetComputeEngine().SetCompilerOptions(options, overrideDefault);
    KTT_CATCH
}

void Tuner::SetGlobalSizeType(const GlobalSizeType type)
{
    KTT_TRY
    m_Tuner->GetComputeEngine().SetGlobalSizeType(type);
    KTT_CATCH
}

void Tuner::SetAutomaticGlobalSizeCorrection(const bool flag)
{
    KTT_TRY
    m_Tuner->GetComputeEngine().SetAutomaticGlobalSizeCorrection(flag);
    KTT_CATCH
}

void Tuner::SetKernelCacheCapacity(const uint64_t capacity)
{
    KTT_TRY
    m_Tuner->GetComputeEngine().SetKernelCacheCapacity(capacity);
    KTT_CATCH
}

std::vector<PlatformInfo> Tuner::GetPlatformInfo() const
{
    KTT_TRY
    return m_Tuner->GetComputeEngine().GetPlatformInfo();
    KTT_CATCH_RETURN(std::vector<PlatformInfo>{})
}

std::vector<DeviceInfo> Tuner::GetDeviceInfo(const PlatformIndex platform) const
{
    KTT_TRY
    return m_Tuner->GetComputeEngine().GetDeviceInfo(platform);
    KTT_CATCH_RETURN(std::vector<DeviceInfo>{})
}

DeviceInfo Tuner::GetCurrentDeviceInfo() const
{
    KTT_TRY
    return m_Tuner->GetComputeEngine().GetCurrentDeviceInfo();
    KTT_CATCH_RETURN(DeviceInfo())
}

void Tuner::SetLoggingLevel(const LoggingLevel level)
{
    Logger::GetLogger().SetLoggingLevel(level);
}

LoggingLevel Tuner::GetLoggingLevel()
{
    return Logger::GetLogger().GetLoggingLevel();
}

void Tuner::SetLoggingTarget(std::ostream& outputTarget)
{
    Logger::GetLogger().SetLoggingTarget(outputTarget);
}

void Tuner::SetLoggingTarget(const std::string& filePath)
{
    Logger::GetLogger().SetLoggingTarget(filePath);
}

ArgumentId Tuner::AddArgumentWithReferencedData(const size_t elementSize, const ArgumentDataType dataType,
    const ArgumentMemoryLocation memoryLocation, const ArgumentAccessType accessType, const ArgumentMemoryType memoryType,
    const ArgumentManagementType managementType, void* data, const size_t dataSize, const ArgumentId& customId)
{
    KTT_TRY
    return m_Tuner->GetArgumentManager().AddArgumentWithReferencedData(elementSize, dataType, memoryLocation, accessType, memoryType,
        managementType, data, dataSize, customId);
    KTT_CATCH_RETURN(InvalidArgumentId)
}

ArgumentId Tuner::AddArgumentWithOwnedData(const size_t elementSize, const ArgumentDataType dataType,
    const ArgumentMemoryLocation memoryLocation, const ArgumentAccessType accessType, const ArgumentMemoryType memoryType,
    const ArgumentManagementType managementType, const void* data, const size_t dataSize, const ArgumentId& customId)
{
    KTT_TRY
    return m_Tuner->GetArgumentManager().AddArgumentWithOwnedData(elementSize, dataType, memoryLocation, accessType, memoryType,
        managementType, data, dataSize, customId);
    KTT_CATCH_RETURN(InvalidArgumentId)
}

ArgumentId Tuner::AddUserArgument(ComputeBuffer buffer, const size_t elementSize, const ArgumentDataType dataType,
    const ArgumentMemoryLocation memoryLocation, const ArgumentAccessType accessType, const size_t dataSize, const ArgumentId& customId)
{
    KTT_TRY
    return m_Tuner->GetArgumentManager().AddUserArgument(buffer, elementSize, dataType, memoryLocation, accessType, dataSize, customId);
    KTT_CATCH_RETURN(InvalidArgumentId)
}

void Tuner::AddParameterInternal(const KernelId id, const std::string& name, const std::vector<ParameterValue>& values,
    const std::string& group)
{
    KTT_TRY
    m_Tuner->GetKernelManager().AddParameter(id, name, values, group);
    KTT_CATCH
}

} // namespace ktt
