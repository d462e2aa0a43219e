// ktt::Tuner. Each public method forwards to TunerCore inside Guarded(): a KttException is
// logged at error level and the method returns its "invalid" value (InvalidKernelId, an
// empty result, ...) instead of throwing, which is the reference's contract for this class
// (Source/Tuner.cpp:41-52). The constructors do throw.
#include <fstream>

#include <Ktt.h>

#include "Common.h"
#include "TunerCore.h"

namespace ktt
{

namespace
{

template <typename Result, typename Body>
Result Guarded(Result fallback, Body body)
{
    try
    {
        return body();
    }
    catch (const KttException& failure)
    {
        Logger::LogError(failure.what());
        return fallback;
    }
}

template <typename Body>
void Guarded(Body body)
{
    try
    {
        body();
    }
    catch (const KttException& failure)
    {
        Logger::LogError(failure.what());
    }
}

ModifierFunction ActionFunction(const ModifierAction action)
{
    switch (action)
    {
    case ModifierAction::Add:
        return [](const uint64_t size, const std::vector<uint64_t>& v) { return size + v[0]; };
    case ModifierAction::Subtract:
        return [](const uint64_t size, const std::vector<uint64_t>& v) { return size - v[0]; };
    case ModifierAction::Multiply:
        return [](const uint64_t size, const std::vector<uint64_t>& v) { return size * v[0]; };
    case ModifierAction::Divide:
        return [](const uint64_t size, const std::vector<uint64_t>& v) { return size / v[0]; };
    case ModifierAction::DivideCeil:
        return [](const uint64_t size, const std::vector<uint64_t>& v) { return (size + v[0] - 1) / v[0]; };
    }
    throw KttException("Unknown thread modifier action");
}

} // namespace

// ---------------------------------------------------------------- construction

Tuner::Tuner(const PlatformIndex platform, const DeviceIndex device, const ComputeApi api) :
    Tuner(platform, device, api, 1)
{}

Tuner::Tuner(const PlatformIndex platform, const DeviceIndex device, const ComputeApi api, const uint32_t computeQueueCount) :
    m_Tuner(std::make_unique<TunerCore>(platform, device, api, computeQueueCount))
{}

Tuner::Tuner(const ComputeApi api, const ComputeApiInitializer& initializer)
{
    std::vector<QueueId> unused;
    m_Tuner = std::make_unique<TunerCore>(api, initializer, unused);
}

Tuner::Tuner(const ComputeApi api, const ComputeApiInitializer& initializer, std::vector<QueueId>& assignedQueueIds) :
    m_Tuner(std::make_unique<TunerCore>(api, initializer, assignedQueueIds))
{}

Tuner::~Tuner() = default;

// ---------------------------------------------------------------- kernels

KernelDefinitionId Tuner::AddKernelDefinition(const std::string& name, const std::string& source, const DimensionVector& globalSize,
    const DimensionVector& localSize, const std::vector<std::string>& typeNames)
{
    return Guarded(InvalidKernelDefinitionId,
        [&]() { return m_Tuner->GetKernels().AddDefinition(name, source, globalSize, localSize, typeNames); });
}

KernelDefinitionId Tuner::AddKernelDefinitionFromFile(const std::string& name, const std::string& filePath,
    const DimensionVector& globalSize, const DimensionVector& localSize, const std::vector<std::string>& typeNames)
{
    return Guarded(InvalidKernelDefinitionId,
        [&]() { return m_Tuner->GetKernels().AddDefinitionFromFile(name, filePath, globalSize, localSize, typeNames); });
}

KernelDefinitionId Tuner::GetKernelDefinitionId(const std::string& name, const std::vector<std::string>& typeNames) const
{
    return Guarded(InvalidKernelDefinitionId, [&]() { return m_Tuner->GetKernels().FindDefinition(name, typeNames); });
}

void Tuner::RemoveKernelDefinition(const KernelDefinitionId id)
{
    Guarded([&]() { m_Tuner->GetKernels().RemoveDefinition(id); });
}

void Tuner::SetArguments(const KernelDefinitionId id, const std::vector<ArgumentId>& argumentIds)
{
    Guarded([&]() { m_Tuner->GetKernels().Bind(id, argumentIds); });
}

KernelId Tuner::CreateSimpleKernel(const std::string& name, const KernelDefinitionId definitionId)
{
    return Guarded(InvalidKernelId, [&]() { return m_Tuner->GetKernels().AddKernel(name, {definitionId}); });
}

KernelId Tuner::CreateCompositeKernel(const std::string& name, const std::vector<KernelDefinitionId>& definitionIds,
    KernelLauncher launcher)
{
    return Guarded(InvalidKernelId, [&]()
    {
        const KernelId id = m_Tuner->GetKernels().AddKernel(name, definitionIds);
        if (launcher)
        {
            m_Tuner->GetKernels().GetKernel(id).SetLauncher(launcher);
        }
        return id;
    });
}

void Tuner::RemoveKernel(const KernelId id)
{
    Guarded([&]() { m_Tuner->RemoveKernel(id); });
}

void Tuner::SetLauncher(const KernelId id, KernelLauncher launcher)
{
    Guarded([&]() { m_Tuner->GetKernels().GetKernel(id).SetLauncher(launcher); });
}

void Tuner::AddParameterValues(const KernelId id, const std::string& name, const std::vector<ParameterValue>& values,
    const std::string& group)
{
    Guarded([&]() { m_Tuner->GetKernels().GetKernel(id).AddParameter(KernelParameter(name, values, group)); });
}

void Tuner::AddThreadModifier(const KernelId id, const std::vector<KernelDefinitionId>& definitionIds, const ModifierType type,
    const ModifierDimension dimension, const std::vector<std::string>& parameters, ModifierFunction function)
{
    Guarded([&]() { m_Tuner->GetKernels().GetKernel(id).AddThreadModifier(type, dimension, parameters, definitionIds, function); });
}

void Tuner::AddThreadModifier(const KernelId id, const std::vector<KernelDefinitionId>& definitionIds, const ModifierType type,
    const ModifierDimension dimension, const std::string& parameter, const ModifierAction action)
{
    Guarded([&]()
    {
        m_Tuner->GetKernels().GetKernel(id).AddThreadModifier(type, dimension, {parameter}, definitionIds, ActionFunction(action));
    });
}

void Tuner::AddConstraint(const KernelId id, const std::vector<std::string>& parameters, ConstraintFunction function)
{
    Guarded([&]() { m_Tuner->GetKernels().GetKernel(id).AddConstraint(KernelConstraint(parameters, function)); });
}

void Tuner::AddGenericConstraint(const KernelId id, const std::vector<std::string>& parameters, GenericConstraintFunction function)
{
    Guarded([&]() { m_Tuner->GetKernels().GetKernel(id).AddConstraint(KernelConstraint(parameters, function)); });
}

void Tuner::SetProfiledDefinitions(const KernelId id, const std::vector<KernelDefinitionId>& definitionIds)
{
    Guarded([&]()
    {
        KernelRegistry& kernels = m_Tuner->GetKernels();
        kernels.GetKernel(id).SetProfiledDefinitions(kernels.Definitions(definitionIds));
    });
}

// ---------------------------------------------------------------- arguments

ArgumentId Tuner::RegisterArgument(const size_t elementSize, const ArgumentDataType dataType, const ArgumentMemoryLocation location,
    const ArgumentAccessType access, const ArgumentMemoryType kind, const ArgumentManagementType management, const DataSource source,
    void* data, const size_t bytes, const ArgumentId& name)
{
    return Guarded(InvalidArgumentId, [&]()
    {
        ArgumentSpec spec;
        spec.elementSize = elementSize;
        spec.dataType = dataType;
        spec.location = location;
        spec.access = access;
        spec.kind = kind;
        spec.management = management;
        const ArgumentSource from = source == DataSource::Reference ? ArgumentSource::Reference
            : source == DataSource::DeviceBuffer ? ArgumentSource::DeviceBuffer : ArgumentSource::Copy;
        return m_Tuner->GetArguments().Add(spec, from, data, bytes, name);
    });
}

ArgumentId Tuner::RegisterSymbolArgument(const size_t elementSize, const ArgumentDataType dataType, const void* data,
    const size_t bytes, const ArgumentId& name, const std::string& symbolName)
{
    return Guarded(InvalidArgumentId, [&]()
    {
        ArgumentSpec spec;
        spec.elementSize = elementSize;
        spec.dataType = dataType;
        spec.location = ArgumentMemoryLocation::Undefined;
        spec.access = ArgumentAccessType::ReadOnly;
        spec.kind = ArgumentMemoryType::Symbol;
        spec.management = ArgumentManagementType::Framework;
        spec.symbolName = symbolName;
        return m_Tuner->GetArguments().Add(spec, ArgumentSource::Copy, const_cast<void*>(data), bytes, name);
    });
}

ArgumentId Tuner::AddArgumentVector(ComputeBuffer buffer, const size_t bufferSize, const size_t elementSize,
    const ArgumentAccessType accessType, const ArgumentMemoryLocation memoryLocation, const ArgumentId& customId)
{
    return RegisterArgument(elementSize, ArgumentDataType::Custom, memoryLocation, accessType, ArgumentMemoryType::Vector,
        ArgumentManagementType::User, DataSource::DeviceBuffer, buffer, bufferSize, customId);
}

ArgumentId Tuner::AddArgumentVectorFromFile(const std::string& filePath, const ArgumentDataType dataType, const size_t elementSize,
    const ArgumentAccessType accessType, const ArgumentMemoryLocation memoryLocation, const ArgumentManagementType managementType,
    const ArgumentId& customId)
{
    return Guarded(InvalidArgumentId, [&]()
    {
        std::string bytes = ReadTextFile(filePath);
        return RegisterArgument(elementSize, dataType, memoryLocation, accessType, ArgumentMemoryType::Vector, managementType,
            DataSource::Copy, bytes.data(), bytes.size(), customId);
    });
}

void Tuner::SaveArgumentVector(const ArgumentId& id, const std::string& filePath)
{
    Guarded([&]()
    {
        const KernelArgument& argument = m_Tuner->GetArguments().Get(id);
        std::ofstream file(filePath, std::ios::binary);
        if (!file)
        {
            throw KttException("Cannot write " + filePath);
        }
        file.write(static_cast<const char*>(argument.GetData()), static_cast<std::streamsize>(argument.GetDataSize()));
    });
}

ArgumentId Tuner::AddArgumentScalar(const void* data, const size_t dataSize, const ArgumentId& customId)
{
    return RegisterArgument(dataSize, ArgumentDataType::Custom, ArgumentMemoryLocation::Undefined, ArgumentAccessType::ReadOnly,
        ArgumentMemoryType::Scalar, ArgumentManagementType::Framework, DataSource::Copy, const_cast<void*>(data), dataSize, customId);
}

void Tuner::RemoveArgument(const ArgumentId& id)
{
    Guarded([&]() { m_Tuner->RemoveArgument(id); });
}

void Tuner::SetReadOnlyArgumentCache(const bool flag)
{
    Guarded([&]() { m_Tuner->GetExecutor().KeepReadOnlyBuffers(flag); });
}

// ---------------------------------------------------------------- running and validation

KernelResult Tuner::Run(const KernelId id, const KernelConfiguration& configuration, const std::vector<BufferOutputDescriptor>& output)
{
    return Run(id, configuration, {}, output);
}

KernelResult Tuner::Run(const KernelId id, const KernelConfiguration& configuration, const KernelDimensions& dimensions,
    const std::vector<BufferOutputDescriptor>& output)
{
    return Guarded(KernelResult(), [&]()
    {
        return m_Tuner->GetExecutor().Execute(m_Tuner->GetKernels().GetKernel(id), configuration, dimensions, KernelRunMode::Running,
            output);
    });
}

void Tuner::SetProfiling(const bool flag)
{
    Guarded([&]() { m_Tuner->GetExecutor().SetProfiling(flag); });
}

bool Tuner::GetProfiling()
{
    return Guarded(false, [&]() { return m_Tuner->GetExecutor().IsProfiling(); });
}

void Tuner::SetValidationMethod(const ValidationMethod method, const double toleranceThreshold)
{
    Guarded([&]() { m_Tuner->GetExecutor().Checker().SetMethod(method, toleranceThreshold); });
}

void Tuner::SetValidationMode(const ValidationMode mode)
{
    Guarded([&]() { m_Tuner->GetExecutor().Checker().SetModes(mode); });
}

void Tuner::SetValidationRange(const ArgumentId& id, const size_t range)
{
    Guarded([&]() { m_Tuner->GetExecutor().Checker().SetRange(m_Tuner->GetArguments().Get(id), range); });
}

void Tuner::SetValueComparator(const ArgumentId& id, ValueComparator comparator)
{
    Guarded([&]() { m_Tuner->GetExecutor().Checker().SetComparator(m_Tuner->GetArguments().Get(id), comparator); });
}

void Tuner::SetReferenceComputation(const ArgumentId& id, ReferenceComputation computation)
{
    Guarded([&]() { m_Tuner->GetExecutor().Checker().SetComputation(m_Tuner->GetArguments().Get(id), computation); });
}

void Tuner::SetReferenceKernel(const ArgumentId& id, const KernelId referenceId, const KernelConfiguration& configuration,
    const KernelDimensions& dimensions)
{
    Guarded([&]()
    {
        m_Tuner->GetExecutor().Checker().SetKernel(m_Tuner->GetArguments().Get(id), m_Tuner->GetKernels().GetKernel(referenceId),
            configuration, dimensions);
    });
}

void Tuner::SetReferenceArgument(const ArgumentId& id, const ArgumentId& referenceId)
{
    Guarded([&]()
    {
        ArgumentRegistry& arguments = m_Tuner->GetArguments();
        m_Tuner->GetExecutor().Checker().SetArgument(arguments.Get(id), arguments.Get(referenceId));
    });
}

// ---------------------------------------------------------------- tuning

std::vector<KernelResult> Tuner::Tune(const KernelId id, std::unique_ptr<StopCondition> stopCondition)
{
    return Tune(id, {}, std::move(stopCondition));
}

std::vector<KernelResult> Tuner::Tune(const KernelId id, const KernelDimensions& dimensions, std::unique_ptr<StopCondition> stopCondition)
{
    return Guarded(std::vector<KernelResult>{}, [&]()
    {
        return m_Tuner->GetTuningLoop().Tune(m_Tuner->GetKernels().GetKernel(id), dimensions, std::move(stopCondition));
    });
}

KernelResult Tuner::TuneIteration(const KernelId id, const std::vector<BufferOutputDescriptor>& output, const bool recomputeReference)
{
    return TuneIteration(id, {}, output, recomputeReference);
}

KernelResult Tuner::TuneIteration(const KernelId id, const KernelDimensions& dimensions, const std::vector<BufferOutputDescriptor>& output,
    const bool recomputeReference)
{
    return Guarded(KernelResult(), [&]()
    {
        return m_Tuner->GetTuningLoop().Step(m_Tuner->GetKernels().GetKernel(id), dimensions, KernelRunMode::OnlineTuning, output,
            recomputeReference);
    });
}

std::vector<KernelResult> Tuner::SimulateTuning(const KernelId id, const std::vector<KernelResult>& results,
    std::unique_ptr<StopCondition> stopCondition)
{
    return Guarded(std::vector<KernelResult>{}, [&]()
    {
        return m_Tuner->GetTuningLoop().Replay(m_Tuner->GetKernels().GetKernel(id), results, std::move(stopCondition));
    });
}

void Tuner::SetProfileBasedSearcher(const KernelId id, const std::string& modelPath, [[maybe_unused]] const bool useBuiltinModule,
    const uint32_t batchSize, const uint32_t neighborSize, const uint32_t randomSize)
{
    Guarded([&]()
    {
        std::vector<KernelResult> knowledge;
        if (!modelPath.empty())
        {
            auto endsWith = [&](const std::string& suffix)
            {
                return modelPath.size() >= suffix.size() && modelPath.compare(modelPath.size() - suffix.size(), suffix.size(), suffix) == 0;
            };
            if (endsWith(".sav"))
            {
                throw KttException("Pickled models are not loaded: give KTT JSON results with profiling counters as " + modelPath);
            }
            const std::string path = endsWith(".json") ? modelPath.substr(0, modelPath.size() - 5) : modelPath;
            UserData ignored;
            knowledge = m_Tuner->LoadResults(path, OutputFormat::JSON, ignored);
        }
        ProfileSearcherOptions options;
        options.batchSize = batchSize;
        options.neighbourCount = neighborSize;
        options.randomCount = randomSize;
        SetSearcher(id, std::make_unique<ProfileBasedSearcher>(*this, options, knowledge));
    });
}

void Tuner::SetSearcher(const KernelId id, std::unique_ptr<Searcher> searcher)
{
    Guarded([&]()
    {
        if (!m_Tuner->GetKernels().HasKernel(id))
        {
            throw KttException("No kernel with id " + std::to_string(id) + " to set a searcher for");
        }
        if (searcher == nullptr)
        {
            throw KttException("The searcher must not be null");
        }
        m_Tuner->GetTuningLoop().SetSearcher(id, std::move(searcher));
    });
}

void Tuner::InitializeConfigurationData(const KernelId id)
{
    Guarded([&]() { m_Tuner->GetTuningLoop().Prepare(m_Tuner->GetKernels().GetKernel(id)); });
}

void Tuner::ClearConfigurationData(const KernelId id)
{
    Guarded([&]() { m_Tuner->GetTuningLoop().Reset(id); });
}

uint64_t Tuner::GetConfigurationsCount(const KernelId id) const
{
    return Guarded(uint64_t{0}, [&]() { return m_Tuner->GetTuningLoop().SpaceSize(id); });
}

KernelConfiguration Tuner::GetBestConfiguration(const KernelId id) const
{
    return Guarded(KernelConfiguration(), [&]() { return m_Tuner->GetTuningLoop().Best(id); });
}

KernelConfiguration Tuner::CreateConfiguration(const KernelId id, const ParameterInput& parameters) const
{
    return Guarded(KernelConfiguration(), [&]() { return m_Tuner->GetKernels().GetKernel(id).CreateConfiguration(parameters); });
}

std::string Tuner::GetKernelSource(const KernelId id, const KernelConfiguration& configuration) const
{
    return Guarded(std::string(), [&]() { return m_Tuner->GetKernelSource(id, configuration); });
}

std::string Tuner::GetKernelDefinitionSource(const KernelDefinitionId id, const KernelConfiguration& configuration) const
{
    return Guarded(std::string(), [&]() { return m_Tuner->GetKernelDefinitionSource(id, configuration); });
}

// ---------------------------------------------------------------- results

void Tuner::SetTimeUnit(const TimeUnit unit)
{
    TimeConfiguration::GetInstance().SetTimeUnit(unit);
}

void Tuner::SaveResults(const std::vector<KernelResult>& results, const std::string& filePath, const OutputFormat format,
    const UserData& data) const
{
    Guarded([&]() { m_Tuner->SaveResults(results, filePath, format, data); });
}

std::vector<KernelResult> Tuner::LoadResults(const std::string& filePath, const OutputFormat format) const
{
    UserData ignored;
    return LoadResults(filePath, format, ignored);
}

std::vector<KernelResult> Tuner::LoadResults(const std::string& filePath, const OutputFormat format, UserData& data) const
{
    return Guarded(std::vector<KernelResult>{}, [&]() { return m_Tuner->LoadResults(filePath, format, data); });
}

// ---------------------------------------------------------------- queues and engine settings

QueueId Tuner::AddComputeQueue(ComputeQueue queue)
{
    return Guarded(InvalidQueueId, [&]() { return m_Tuner->GetComputeEngine().AddComputeQueue(queue); });
}

void Tuner::RemoveComputeQueue(const QueueId id)
{
    Guarded([&]() { m_Tuner->GetComputeEngine().RemoveComputeQueue(id); });
}

void Tuner::WaitForComputeAction(const ComputeActionId id)
{
    Guarded([&]() { m_Tuner->GetComputeEngine().WaitForComputeAction(id); });
}

void Tuner::WaitForTransferAction(const TransferActionId id)
{
    Guarded([&]() { m_Tuner->GetComputeEngine().WaitForTransferAction(id); });
}

void Tuner::SynchronizeQueue(const QueueId id)
{
    Guarded([&]() { m_Tuner->GetComputeEngine().SynchronizeQueue(id); });
}

void Tuner::SynchronizeQueues()
{
    Guarded([&]() { m_Tuner->GetComputeEngine().SynchronizeQueues(); });
}

void Tuner::SynchronizeDevice()
{
    Guarded([&]() { m_Tuner->GetComputeEngine().SynchronizeDevice(); });
}

void Tuner::SetProfilingCounters(const std::vector<std::string>& counters)
{
    Guarded([&]() { m_Tuner->GetComputeEngine().SetProfilingCounters(counters); });
}

void Tuner::SetCompilerOptions(const std::string& options, const bool overrideDefault)
{
    Guarded([&]() { m_Tuner->GetComputeEngine().SetCompilerOptions(options, overrideDefault); });
}

void Tuner::SetGlobalSizeType(const GlobalSizeType type)
{
    Guarded([&]() { m_Tuner->GetComputeEngine().SetGlobalSizeType(type); });
}

void Tuner::SetAutomaticGlobalSizeCorrection(const bool flag)
{
    Guarded([&]() { m_Tuner->GetComputeEngine().SetAutomaticGlobalSizeCorrection(flag); });
}

void Tuner::SetKernelCacheCapacity(const uint64_t capacity)
{
    Guarded([&]() { m_Tuner->GetComputeEngine().SetKernelCacheCapacity(capacity); });
}

std::vector<PlatformInfo> Tuner::GetPlatformInfo() const
{
    return Guarded(std::vector<PlatformInfo>{}, [&]() { return m_Tuner->GetComputeEngine().GetPlatformInfo(); });
}

std::vector<DeviceInfo> Tuner::GetDeviceInfo(const PlatformIndex platform) const
{
    return Guarded(std::vector<DeviceInfo>{}, [&]() { return m_Tuner->GetComputeEngine().GetDeviceInfo(platform); });
}

DeviceInfo Tuner::GetCurrentDeviceInfo() const
{
    return Guarded(DeviceInfo(), [&]() { return m_Tuner->GetComputeEngine().GetCurrentDeviceInfo(); });
}

// ---------------------------------------------------------------- logging

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

} // namespace ktt
