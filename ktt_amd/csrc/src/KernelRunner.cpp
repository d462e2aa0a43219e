#include <cmath>
#include <cstring>
#include <type_traits>

#include <KttException.h>

#include "Common.h"
#include "KernelRunner.h"

namespace ktt
{

// ---------------------------------------------------------------- KernelComputeData

KernelComputeData::KernelComputeData(const Kernel& kernel, const KernelDefinition& definition, const KernelConfiguration& configuration,
    const KernelDimensions& dimensions) :
    m_Name(definition.GetName()),
    m_DefaultSource(definition.GetSource()),
    m_ConfigurationPrefix(configuration.GeneratePrefix()),
    m_TemplatedName(definition.GetTemplatedName()),
    m_SourceDirectory(definition.GetSourceDirectory()),
    m_Configuration(configuration),
    m_Arguments(definition.GetArguments())
{
    const auto id = definition.GetId();
    const auto& pairs = configuration.GetPairs();
    auto iterator = dimensions.find(id);
    if (iterator != dimensions.cend())
    {
        m_GlobalSize = kernel.GetModifiedSize(id, iterator->second.first, ModifierType::Global, pairs);
        m_LocalSize = kernel.GetModifiedSize(id, iterator->second.second, ModifierType::Local, pairs);
    }
    else
    {
        m_GlobalSize = kernel.GetModifiedSize(id, ModifierType::Global, pairs);
        m_LocalSize = kernel.GetModifiedSize(id, ModifierType::Local, pairs);
    }
}

void KernelComputeData::SwapArguments(const ArgumentId& first, const ArgumentId& second)
{
    size_t firstIndex = m_Arguments.size();
    size_t secondIndex = m_Arguments.size();
    for (size_t i = 0; i < m_Arguments.size(); ++i)
    {
        if (m_Arguments[i]->GetId() == first)
        {
            firstIndex = i;
        }
        else if (m_Arguments[i]->GetId() == second)
        {
            secondIndex = i;
        }
    }
    if (firstIndex == m_Arguments.size() || secondIndex == m_Arguments.size())
    {
        throw KttException("One of the arguments with ids " + first + " and " + second + " is not associated with kernel definition "
            + m_Name);
    }
    std::swap(m_Arguments[firstIndex], m_Arguments[secondIndex]);
}

KernelComputeId KernelComputeData::GetUniqueIdentifier() const
{
    return m_Name + m_TemplatedName + m_ConfigurationPrefix;
}

// ---------------------------------------------------------------- ComputeLayerData

ComputeLayerData::ComputeLayerData(const Kernel& kernel, const KernelConfiguration& configuration, const KernelDimensions& dimensions,
    const KernelRunMode mode) :
    m_Kernel(kernel),
    m_Configuration(configuration),
    m_RunMode(mode),
    m_DataOverhead(0)
{
    for (const auto* definition : kernel.GetDefinitions())
    {
        m_ComputeData.emplace(definition->GetId(), KernelComputeData(kernel, *definition, m_Configuration, dimensions));
    }
}

void ComputeLayerData::AddPartialResult(const ComputationResult& result)
{
    m_PartialResults.push_back(result);
}

void ComputeLayerData::IncreaseOverhead(const Nanoseconds overhead)
{
    m_DataOverhead += overhead;
}

void ComputeLayerData::AddArgumentOverride(const ArgumentId& id, const KernelArgument& argument)
{
    auto copy = std::make_unique<KernelArgument>(argument);
    KernelArgument* replacement = copy.get();
    KernelArgument* previous = nullptr;
    auto iterator = m_Overrides.find(id);
    if (iterator != m_Overrides.end())
    {
        previous = iterator->second.get();
    }

    for (auto& [definitionId, data] : m_ComputeData)
    {
        auto arguments = data.GetArguments();
        for (auto*& pointer : arguments)
        {
            if (pointer->GetId() == id)
            {
                pointer = replacement;
            }
        }
        data.SetArguments(arguments);
    }
    (void)previous;
    m_Overrides[id] = std::move(copy);
}

void ComputeLayerData::ChangeArguments(const KernelDefinitionId id, std::vector<KernelArgument*> arguments)
{
    for (auto*& argument : arguments)
    {
        auto iterator = m_Overrides.find(argument->GetId());
        if (iterator != m_Overrides.end())
        {
            argument = iterator->second.get();
        }
    }
    GetComputeData(id).SetArguments(arguments);
}

void ComputeLayerData::SwapArguments(const KernelDefinitionId id, const ArgumentId& first, const ArgumentId& second)
{
    GetComputeData(id).SwapArguments(first, second);
}

const KernelComputeData& ComputeLayerData::GetComputeData(const KernelDefinitionId id) const
{
    auto iterator = m_ComputeData.find(id);
    if (iterator == m_ComputeData.cend())
    {
        throw KttException("Kernel " + m_Kernel.GetName() + " has no data for definition with id " + std::to_string(id));
    }
    return iterator->second;
}

KernelComputeData& ComputeLayerData::GetComputeData(const KernelDefinitionId id)
{
    auto iterator = m_ComputeData.find(id);
    if (iterator == m_ComputeData.end())
    {
        throw KttException("Kernel " + m_Kernel.GetName() + " has no data for definition with id " + std::to_string(id));
    }
    return iterator->second;
}

Nanoseconds ComputeLayerData::CalculateLauncherOverhead() const
{
    Nanoseconds result = m_DataOverhead;
    for (const auto& partial : m_PartialResults)
    {
        result += partial.GetDuration();
        result += partial.GetOverhead();
    }
    return result;
}

KernelResult ComputeLayerData::GenerateResult(const Nanoseconds launcherDuration) const
{
    // ComputeLayerData.cpp:124-143: with a user launcher the time it spends outside kernels and
    // transfers is ExtraDuration (part of TotalDuration); without one it is overhead.
    KernelResult result(m_Kernel.GetName(), m_Configuration, m_PartialResults);
    const Nanoseconds launcherOverhead = CalculateLauncherOverhead();
    const Nanoseconds actualDuration = launcherDuration > launcherOverhead ? launcherDuration - launcherOverhead : 0;

    if (m_Kernel.HasLauncher())
    {
        result.SetExtraDuration(actualDuration);
        result.SetDataMovementOverhead(m_DataOverhead);
    }
    else
    {
        result.SetDataMovementOverhead(actualDuration + m_DataOverhead);
    }
    return result;
}

// ---------------------------------------------------------------- ComputeLayer

ComputeLayer::ComputeLayer(ComputeEngine& engine, KernelArgumentManager& argumentManager) :
    m_ComputeEngine(engine),
    m_ArgumentManager(argumentManager),
    m_ActiveKernel(InvalidKernelId)
{}

void ComputeLayer::RunKernel(const KernelDefinitionId id)
{
    const auto action = RunKernelAsync(id, GetDefaultQueue());
    WaitForComputeAction(action);
}

void ComputeLayer::RunKernel(const KernelDefinitionId id, const DimensionVector& globalSize, const DimensionVector& localSize)
{
    const auto action = RunKernelAsync(id, GetDefaultQueue(), globalSize, localSize);
    WaitForComputeAction(action);
}

ComputeActionId ComputeLayer::RunKernelAsync(const KernelDefinitionId id, const QueueId queue)
{
    const auto& data = GetData().GetComputeData(id);
    return m_ComputeEngine.RunKernelAsync(data, queue);
}

ComputeActionId ComputeLayer::RunKernelAsync(const KernelDefinitionId id, const QueueId queue, const DimensionVector& globalSize,
    const DimensionVector& localSize)
{
    KernelComputeData data = GetData().GetComputeData(id);
    data.SetGlobalSize(globalSize);
    data.SetLocalSize(localSize);
    return m_ComputeEngine.RunKernelAsync(data, queue);
}

void ComputeLayer::WaitForComputeAction(const ComputeActionId id)
{
    const auto result = m_ComputeEngine.WaitForComputeAction(id);
    GetData().AddPartialResult(result);
}

void ComputeLayer::RunKernelWithProfiling(const KernelDefinitionId id)
{
    const auto& data = GetData().GetComputeData(id);
    const auto result = m_ComputeEngine.RunKernelWithProfiling(data, GetDefaultQueue());
    GetData().AddPartialResult(result);
}

void ComputeLayer::RunKernelWithProfiling(const KernelDefinitionId id, const DimensionVector& globalSize,
    const DimensionVector& localSize)
{
    KernelComputeData data = GetData().GetComputeData(id);
    data.SetGlobalSize(globalSize);
    data.SetLocalSize(localSize);
    const auto result = m_ComputeEngine.RunKernelWithProfiling(data, GetDefaultQueue());
    GetData().AddPartialResult(result);
}

uint64_t ComputeLayer::GetRemainingProfilingRuns(const KernelDefinitionId id) const
{
    return m_ComputeEngine.GetRemainingProfilingRuns(GetData().GetComputeData(id).GetUniqueIdentifier());
}

uint64_t ComputeLayer::GetRemainingProfilingRuns() const
{
    uint64_t result = 0;
    for (const auto* definition : GetData().GetKernel().GetProfiledDefinitions())
    {
        result += GetRemainingProfilingRuns(definition->GetId());
    }
    return result;
}

QueueId ComputeLayer::GetDefaultQueue() const
{
    return m_ComputeEngine.GetDefaultQueue();
}

std::vector<QueueId> ComputeLayer::GetAllQueues() const
{
    return m_ComputeEngine.GetAllQueues();
}

void ComputeLayer::SynchronizeQueue(const QueueId queue)
{
    m_ComputeEngine.SynchronizeQueue(queue);
}

void ComputeLayer::SynchronizeQueues()
{
    m_ComputeEngine.SynchronizeQueues();
}

void ComputeLayer::SynchronizeDevice()
{
    m_ComputeEngine.SynchronizeDevice();
}

const DimensionVector& ComputeLayer::GetCurrentGlobalSize(const KernelDefinitionId id) const
{
    return GetData().GetComputeData(id).GetGlobalSize();
}

const DimensionVector& ComputeLayer::GetCurrentLocalSize(const KernelDefinitionId id) const
{
    return GetData().GetComputeData(id).GetLocalSize();
}

const KernelConfiguration& ComputeLayer::GetCurrentConfiguration() const
{
    return GetData().GetConfiguration();
}

KernelRunMode ComputeLayer::GetRunMode() const
{
    return GetData().GetRunMode();
}

void ComputeLayer::ChangeArguments(const KernelDefinitionId id, const std::vector<ArgumentId>& arguments)
{
    GetData().ChangeArguments(id, m_ArgumentManager.GetArguments(arguments));
}

void ComputeLayer::SwapArguments(const KernelDefinitionId id, const ArgumentId& first, const ArgumentId& second)
{
    GetData().SwapArguments(id, first, second);
}

void ComputeLayer::UpdateScalarArgument(const ArgumentId& id, const void* data)
{
    // Per-run override, as in ComputeLayer.cpp:164-176: the registered argument keeps its value.
    const auto& argument = m_ArgumentManager.GetArgument(id);
    if (argument.GetMemoryType() != ArgumentMemoryType::Scalar)
    {
        throw KttException("Argument with id " + id + " is not a scalar argument");
    }
    KernelArgument copy = argument;
    copy.SetOwnedData(data, argument.GetDataSize());
    GetData().AddArgumentOverride(id, copy);
}

void ComputeLayer::UpdateLocalArgument(const ArgumentId& id, const size_t dataSize)
{
    const auto& argument = m_ArgumentManager.GetArgument(id);
    if (argument.GetMemoryType() != ArgumentMemoryType::Local)
    {
        throw KttException("Argument with id " + id + " is not a local memory argument");
    }
    KernelArgument copy = argument;
    copy.SetLocalSize(dataSize);
    GetData().AddArgumentOverride(id, copy);
}

void ComputeLayer::UploadBuffer(const ArgumentId& id)
{
    Timer timer;
    timer.Start();
    const auto action = UploadBufferAsync(id, GetDefaultQueue());
    WaitForTransferAction(action);
    timer.Stop();
    GetData().IncreaseOverhead(timer.GetElapsedTime());
}

TransferActionId ComputeLayer::UploadBufferAsync(const ArgumentId& id, const QueueId queue)
{
    auto& argument = m_ArgumentManager.GetArgument(id);
    return m_ComputeEngine.UploadArgument(argument, queue);
}

void ComputeLayer::DownloadBuffer(const ArgumentId& id, void* destination, const size_t dataSize)
{
    Timer timer;
    timer.Start();
    const auto action = DownloadBufferAsync(id, GetDefaultQueue(), destination, dataSize);
    WaitForTransferAction(action);
    timer.Stop();
    GetData().IncreaseOverhead(timer.GetElapsedTime());
}

TransferActionId ComputeLayer::DownloadBufferAsync(const ArgumentId& id, const QueueId queue, void* destination, const size_t dataSize)
{
    return m_ComputeEngine.DownloadArgument(id, queue, destination, dataSize);
}

void ComputeLayer::UpdateBuffer(const ArgumentId& id, const void* data, const size_t dataSize)
{
    Timer timer;
    timer.Start();
    const auto action = UpdateBufferAsync(id, GetDefaultQueue(), data, dataSize);
    WaitForTransferAction(action);
    timer.Stop();
    GetData().IncreaseOverhead(timer.GetElapsedTime());
}

TransferActionId ComputeLayer::UpdateBufferAsync(const ArgumentId& id, const QueueId queue, const void* data, const size_t dataSize)
{
    return m_ComputeEngine.UpdateArgument(id, queue, data, dataSize);
}

void ComputeLayer::CopyBuffer(const ArgumentId& destination, const ArgumentId& source, const size_t dataSize)
{
    Timer timer;
    timer.Start();
    const auto action = CopyBufferAsync(destination, source, GetDefaultQueue(), dataSize);
    WaitForTransferAction(action);
    timer.Stop();
    GetData().IncreaseOverhead(timer.GetElapsedTime());
}

TransferActionId ComputeLayer::CopyBufferAsync(const ArgumentId& destination, const ArgumentId& source, const QueueId queue,
    const size_t dataSize)
{
    return m_ComputeEngine.CopyArgument(destination, queue, source, dataSize);
}

void ComputeLayer::WaitForTransferAction(const TransferActionId id)
{
    m_ComputeEngine.WaitForTransferAction(id);
}

void ComputeLayer::ResizeBuffer(const ArgumentId& id, const size_t newDataSize, const bool preserveData)
{
    m_ComputeEngine.ResizeArgument(id, newDataSize, preserveData);
}

void ComputeLayer::ClearBuffer(const ArgumentId& id)
{
    m_ComputeEngine.ClearBuffer(id);
}

bool ComputeLayer::HasBuffer(const ArgumentId& id)
{
    return m_ComputeEngine.HasBuffer(id);
}

void ComputeLayer::GetUnifiedMemoryBufferHandle(const ArgumentId& id, UnifiedBufferMemory& memoryHandle)
{
    m_ComputeEngine.GetUnifiedMemoryBufferHandle(id, memoryHandle);
}

bool ComputeLayer::GetProfiling([[maybe_unused]] const KernelId id)
{
    return m_ComputeEngine.IsProfilingActive();
}

void ComputeLayer::SetActiveKernel(const KernelId id)
{
    m_ActiveKernel = id;
}

void ComputeLayer::ClearActiveKernel()
{
    m_ActiveKernel = InvalidKernelId;
}

void ComputeLayer::AddData(const Kernel& kernel, const KernelConfiguration& configuration, const KernelDimensions& dimensions,
    const KernelRunMode mode)
{
    m_Data[kernel.GetId()] = std::make_unique<ComputeLayerData>(kernel, configuration, dimensions, mode);
}

void ComputeLayer::ClearData(const KernelId id)
{
    m_Data.erase(id);
}

void ComputeLayer::ClearComputeEngineData()
{
    m_ComputeEngine.ClearKernelCache();
}

KernelResult ComputeLayer::GenerateResult(const KernelId id, const Nanoseconds launcherDuration) const
{
    auto iterator = m_Data.find(id);
    if (iterator == m_Data.cend())
    {
        throw KttException("No compute layer data for kernel with id " + std::to_string(id));
    }
    return iterator->second->GenerateResult(launcherDuration);
}

ComputeLayerData& ComputeLayer::GetData()
{
    auto iterator = m_Data.find(m_ActiveKernel);
    if (iterator == m_Data.end())
    {
        throw KttException("Compute interface is used outside of an active kernel run");
    }
    return *iterator->second;
}

const ComputeLayerData& ComputeLayer::GetData() const
{
    auto iterator = m_Data.find(m_ActiveKernel);
    if (iterator == m_Data.cend())
    {
        throw KttException("Compute interface is used outside of an active kernel run");
    }
    return *iterator->second;
}

// ---------------------------------------------------------------- ValidationData

ValidationData::ValidationData(KernelRunner& kernelRunner, const KernelArgument& argument) :
    m_KernelRunner(kernelRunner),
    m_Argument(argument),
    m_ValidationRange(static_cast<size_t>(argument.GetNumberOfElements())),
    m_Comparator(nullptr),
    m_ReferenceComputation(nullptr),
    m_ReferenceKernel(nullptr),
    m_ReferenceArgument(nullptr)
{
    if (argument.GetMemoryType() != ArgumentMemoryType::Vector)
    {
        throw KttException("Kernel argument with id " + argument.GetId() + " is not a vector and cannot be validated");
    }
    if (argument.GetAccessType() == ArgumentAccessType::ReadOnly)
    {
        throw KttException("Kernel argument with id " + argument.GetId() + " is read-only and cannot be validated");
    }
}

void ValidationData::SetValidationRange(const size_t range)
{
    m_ValidationRange = range;
}

void ValidationData::SetValueComparator(ValueComparator comparator)
{
    if (!comparator)
    {
        throw KttException("Value comparator must be a valid function");
    }
    m_Comparator = comparator;
}

void ValidationData::SetReferenceComputation(ReferenceComputation computation)
{
    ResetReferenceData();
    m_ReferenceComputation = computation;
}

void ValidationData::SetReferenceKernel(const Kernel& kernel, const KernelConfiguration& configuration, const KernelDimensions& dimensions)
{
    ResetReferenceData();
    m_ReferenceKernel = &kernel;
    m_ReferenceConfiguration = configuration;
    m_ReferenceDimensions = dimensions;
}

void ValidationData::SetReferenceArgument(const KernelArgument& argument)
{
    if (argument.GetMemoryType() != ArgumentMemoryType::Vector)
    {
        throw KttException("Kernel argument with id " + argument.GetId() + " is not a vector and cannot be used as a reference argument");
    }
    if (argument.GetOwnership() == ArgumentOwnership::User)
    {
        throw KttException("Kernel argument with id " + argument.GetId()
            + " has user ownership and cannot be used as a reference argument");
    }
    ResetReferenceData();
    m_ReferenceArgument = &argument;
}

KernelId ValidationData::GetReferenceKernelId() const
{
    return m_ReferenceKernel->GetId();
}

void ValidationData::ComputeReferenceResults()
{
    const Nanoseconds elapsed = RunScopeTimer([this]()
    {
        const size_t referenceSize = m_ValidationRange * m_Argument.GetElementSize();
        if (HasReferenceComputation())
        {
            Logger::LogInfo("Computing reference computation result for argument with id " + m_Argument.GetId());
            m_ReferenceResult.assign(referenceSize, 0);
            m_ReferenceComputation(m_ReferenceResult.data());
        }
        if (HasReferenceKernel())
        {
            Logger::LogInfo("Computing reference kernel result for argument with id " + m_Argument.GetId());
            const bool profiling = m_KernelRunner.IsProfilingActive();
            m_KernelRunner.SetProfiling(false);
            m_ReferenceKernelResult.assign(referenceSize, 0);
            BufferOutputDescriptor descriptor(m_Argument.GetId(), m_ReferenceKernelResult.data(), referenceSize);
            const auto result = m_KernelRunner.RunKernel(*m_ReferenceKernel, m_ReferenceConfiguration, m_ReferenceDimensions,
                KernelRunMode::ResultValidation, {descriptor});
            m_KernelRunner.SetProfiling(profiling);
            if (!result.IsValid())
            {
                m_ReferenceKernelResult.clear();
                throw KttException("Reference kernel " + m_ReferenceKernel->GetName() + " failed with status "
                    + ResultStatusToString(result.GetStatus()));
            }
        }
    });
    const auto& time = TimeConfiguration::GetInstance();
    Logger::LogInfo("Reference result for argument with id " + m_Argument.GetId() + " was computed in "
        + std::to_string(time.ConvertFromNanoseconds(elapsed)) + time.GetUnitTag());
}

void ValidationData::ClearReferenceResults()
{
    m_ReferenceResult.clear();
    m_ReferenceKernelResult.clear();
}

bool ValidationData::HasReferenceResults() const
{
    return (HasReferenceComputation() && !m_ReferenceResult.empty()) || (HasReferenceKernel() && !m_ReferenceKernelResult.empty())
        || HasReferenceArgument();
}

const void* ValidationData::GetReferenceResult() const
{
    if (HasReferenceComputation())
    {
        return m_ReferenceResult.data();
    }
    if (HasReferenceKernel())
    {
        return m_ReferenceKernelResult.data();
    }
    if (HasReferenceArgument())
    {
        return m_ReferenceArgument->GetData();
    }
    throw KttException("Validation data for argument with id " + m_Argument.GetId() + " has no reference");
}

void ValidationData::ResetReferenceData()
{
    ClearReferenceResults();
    m_ReferenceComputation = nullptr;
    m_ReferenceKernel = nullptr;
    m_ReferenceConfiguration = KernelConfiguration();
    m_ReferenceDimensions.clear();
    m_ReferenceArgument = nullptr;
}

// ---------------------------------------------------------------- ResultValidator

ResultValidator::ResultValidator(KernelRunner& kernelRunner) :
    m_KernelRunner(kernelRunner),
    m_ToleranceThreshold(1e-4),
    m_ValidationMethod(ValidationMethod::SideBySideComparison),
    m_ValidationMode(ValidationMode::OfflineTuning | ValidationMode::OnlineTuning)
{}

void ResultValidator::SetValidationMethod(const ValidationMethod method, const double toleranceThreshold)
{
    if (toleranceThreshold < 0.0)
    {
        throw KttException("Tolerance threshold cannot be negative");
    }
    m_ValidationMethod = method;
    m_ToleranceThreshold = toleranceThreshold;
}

void ResultValidator::SetValidationMode(const ValidationMode mode)
{
    m_ValidationMode = mode;
}

void ResultValidator::InitializeValidationData(const KernelArgument& argument)
{
    m_ValidationData[argument.GetId()] = std::make_unique<ValidationData>(m_KernelRunner, argument);
}

void ResultValidator::SetValidationRange(const ArgumentId& id, const size_t range)
{
    m_ValidationData.at(id)->SetValidationRange(range);
}

void ResultValidator::SetValueComparator(const ArgumentId& id, ValueComparator comparator)
{
    m_ValidationData.at(id)->SetValueComparator(comparator);
}

void ResultValidator::SetReferenceComputation(const ArgumentId& id, ReferenceComputation computation)
{
    m_ValidationData.at(id)->SetReferenceComputation(computation);
}

void ResultValidator::SetReferenceKernel(const ArgumentId& id, const Kernel& kernel, const KernelConfiguration& configuration,
    const KernelDimensions& dimensions)
{
    m_ValidationData.at(id)->SetReferenceKernel(kernel, configuration, dimensions);
}

void ResultValidator::SetReferenceArgument(const ArgumentId& id, const KernelArgument& argument)
{
    m_ValidationData.at(id)->SetReferenceArgument(argument);
}

bool ResultValidator::HasValidationData(const ArgumentId& id) const
{
    return ContainsKey(m_ValidationData, id);
}

void ResultValidator::RemoveValidationData(const ArgumentId& id)
{
    m_ValidationData.erase(id);
}

void ResultValidator::RemoveDataWithReferenceKernel(const KernelId id)
{
    for (auto iterator = m_ValidationData.begin(); iterator != m_ValidationData.end();)
    {
        if (iterator->second->HasReferenceKernel() && iterator->second->GetReferenceKernelId() == id)
        {
            iterator = m_ValidationData.erase(iterator);
        }
        else
        {
            ++iterator;
        }
    }
}

void ResultValidator::ComputeReferenceResult(const Kernel& kernel, const KernelRunMode runMode)
{
    if (!IsRunModeValidated(runMode))
    {
        return;
    }
    for (const auto* argument : kernel.GetVectorArguments())
    {
        auto iterator = m_ValidationData.find(argument->GetId());
        if (iterator != m_ValidationData.end() && !iterator->second->HasReferenceResults())
        {
            iterator->second->ComputeReferenceResults();
        }
    }
}

void ResultValidator::ClearReferenceResult(const Kernel& kernel)
{
    for (const auto* argument : kernel.GetVectorArguments())
    {
        auto iterator = m_ValidationData.find(argument->GetId());
        if (iterator != m_ValidationData.end())
        {
            iterator->second->ClearReferenceResults();
        }
    }
}

bool ResultValidator::HasReferenceResult(const Kernel& kernel) const
{
    bool result = true;
    for (const auto* argument : kernel.GetVectorArguments())
    {
        auto iterator = m_ValidationData.find(argument->GetId());
        if (iterator != m_ValidationData.cend())
        {
            result &= iterator->second->HasReferenceResults();
        }
    }
    return result;
}

bool ResultValidator::ValidateArguments(const Kernel& kernel, const KernelRunMode runMode) const
{
    if (!IsRunModeValidated(runMode))
    {
        return true;
    }
    bool result = true;
    for (const auto* argument : kernel.GetVectorArguments())
    {
        if (HasValidationData(argument->GetId()))
        {
            result &= ValidateArgument(*argument);
        }
    }
    return result;
}

bool ResultValidator::IsRunModeValidated(const KernelRunMode mode) const
{
    switch (mode)
    {
    case KernelRunMode::Running:
        return HasFlag(m_ValidationMode, ValidationMode::Running);
    case KernelRunMode::OfflineTuning:
        return HasFlag(m_ValidationMode, ValidationMode::OfflineTuning);
    case KernelRunMode::OnlineTuning:
        return HasFlag(m_ValidationMode, ValidationMode::OnlineTuning);
    default:
        return false;
    }
}

template <typename T>
bool ResultValidator::ValidateTyped(const KernelArgument& argument, const T* result, const T* reference, const size_t range) const
{
    if constexpr (!std::is_floating_point_v<T>)
    {
        for (size_t i = 0; i < range; ++i)
        {
            if (result[i] != reference[i])
            {
                Logger::LogWarning("Results differ for argument with id " + argument.GetId() + " at index " + std::to_string(i)
                    + ", reference value: " + std::to_string(reference[i]) + ", result value: " + std::to_string(result[i]));
                return false;
            }
        }
        return true;
    }
    else
    {
        switch (m_ValidationMethod)
        {
        case ValidationMethod::AbsoluteDifference:
        {
            double difference = 0.0;
            for (size_t i = 0; i < range; ++i)
            {
                difference += std::fabs(result[i] - reference[i]);
            }
            if (difference > m_ToleranceThreshold)
            {
                Logger::LogWarning("Results differ for argument with id " + argument.GetId() + ", absolute difference is "
                    + std::to_string(difference));
                return false;
            }
            return true;
        }
        case ValidationMethod::SideBySideComparison:
            for (size_t i = 0; i < range; ++i)
            {
                const T difference = std::fabs(result[i] - reference[i]);
                if (difference > m_ToleranceThreshold)
                {
                    Logger::LogWarning("Results differ for argument with id " + argument.GetId() + " at index " + std::to_string(i)
                        + ", reference value: " + std::to_string(reference[i]) + ", result value: " + std::to_string(result[i])
                        + ", difference: " + std::to_string(difference));
                    return false;
                }
            }
            return true;
        case ValidationMethod::SideBySideRelativeComparison:
            for (size_t i = 0; i < range; ++i)
            {
                const T difference = std::fabs(result[i] - reference[i]);
                const T relative = difference / reference[i];
                if (difference > 1e-4 && relative > m_ToleranceThreshold)
                {
                    Logger::LogWarning("Results differ for argument with id " + argument.GetId() + " at index " + std::to_string(i)
                        + ", reference value: " + std::to_string(reference[i]) + ", result value: " + std::to_string(result[i])
                        + ", relative difference: " + std::to_string(relative));
                    return false;
                }
            }
            return true;
        }
        return false;
    }
}

bool ResultValidator::ValidateArgument(const KernelArgument& argument) const
{
    const auto& validationData = *m_ValidationData.at(argument.GetId());
    const size_t range = validationData.GetValidationRange();
    const size_t bufferSize = range * argument.GetElementSize();
    std::vector<uint8_t> data(bufferSize);
    BufferOutputDescriptor descriptor(argument.GetId(), data.data(), bufferSize);
    m_KernelRunner.DownloadBuffers({descriptor});
    const void* reference = validationData.GetReferenceResult();

    if (validationData.HasValueComparator())
    {
        const auto comparator = validationData.GetValueComparator();
        const size_t elementSize = argument.GetElementSize();
        for (size_t i = 0; i < bufferSize; i += elementSize)
        {
            if (!comparator(data.data() + i, static_cast<const uint8_t*>(reference) + i))
            {
                Logger::LogWarning("Results differ for argument with id: " + argument.GetId() + " at index "
                    + std::to_string(i / elementSize));
                return false;
            }
        }
        return true;
    }

    switch (argument.GetDataType())
    {
    case ArgumentDataType::Char:
        return ValidateTyped(argument, reinterpret_cast<const int8_t*>(data.data()), static_cast<const int8_t*>(reference), range);
    case ArgumentDataType::UnsignedChar:
        return ValidateTyped(argument, reinterpret_cast<const uint8_t*>(data.data()), static_cast<const uint8_t*>(reference), range);
    case ArgumentDataType::Short:
        return ValidateTyped(argument, reinterpret_cast<const int16_t*>(data.data()), static_cast<const int16_t*>(reference), range);
    case ArgumentDataType::UnsignedShort:
        return ValidateTyped(argument, reinterpret_cast<const uint16_t*>(data.data()), static_cast<const uint16_t*>(reference), range);
    case ArgumentDataType::Int:
        return ValidateTyped(argument, reinterpret_cast<const int32_t*>(data.data()), static_cast<const int32_t*>(reference), range);
    case ArgumentDataType::UnsignedInt:
        return ValidateTyped(argument, reinterpret_cast<const uint32_t*>(data.data()), static_cast<const uint32_t*>(reference), range);
    case ArgumentDataType::Long:
        return ValidateTyped(argument, reinterpret_cast<const int64_t*>(data.data()), static_cast<const int64_t*>(reference), range);
    case ArgumentDataType::UnsignedLong:
        return ValidateTyped(argument, reinterpret_cast<const uint64_t*>(data.data()), static_cast<const uint64_t*>(reference), range);
    case ArgumentDataType::Float:
        return ValidateTyped(argument, reinterpret_cast<const float*>(data.data()), static_cast<const float*>(reference), range);
    case ArgumentDataType::Double:
        return ValidateTyped(argument, reinterpret_cast<const double*>(data.data()), static_cast<const double*>(reference), range);
    default:
        throw KttException("Validation of kernel arguments with custom data type requires usage of value comparator");
    }
}

// ---------------------------------------------------------------- KernelRunner

KernelRunner::KernelRunner(ComputeEngine& engine, KernelArgumentManager& argumentManager) :
    m_ComputeLayer(std::make_unique<ComputeLayer>(engine, argumentManager)),
    m_Validator(std::make_unique<ResultValidator>(*this)),
    m_Engine(engine),
    m_ArgumentManager(argumentManager),
    m_ReadOnlyCacheFlag(true)
{}

KernelResult KernelRunner::RunKernel(const Kernel& kernel, const KernelConfiguration& configuration, const KernelDimensions& dimensions,
    const KernelRunMode mode, const std::vector<BufferOutputDescriptor>& output, const bool manageBuffers)
{
    m_Engine.EnsureThreadContext();

    if (!m_Validator->HasReferenceResult(kernel))
    {
        m_Validator->ComputeReferenceResult(kernel, mode);
    }

    Nanoseconds dataOverhead = 0;
    if (manageBuffers)
    {
        dataOverhead += RunScopeTimer([this, &kernel]() { SetupBuffers(kernel); });
    }

    if (IsProfilingActive())
    {
        Logger::LogInfo("Profiling kernel " + kernel.GetName() + " with configuration: " + configuration.GetString());
    }
    else
    {
        Logger::LogInfo("Running kernel " + kernel.GetName() + " with configuration: " + configuration.GetString());
    }

    auto launcher = GetKernelLauncher(kernel);
    KernelResult result = RunKernelInternal(kernel, configuration, dimensions, mode, launcher, output);
    ValidateResult(kernel, result, mode);

    if (manageBuffers)
    {
        dataOverhead += RunScopeTimer([this, &kernel]() { CleanupBuffers(kernel); });
    }

    result.SetDataMovementOverhead(result.GetDataMovementOverhead() + dataOverhead);
    return result;
}

void KernelRunner::SetupBuffers(const Kernel& kernel)
{
    for (auto* argument : kernel.GetVectorArguments())
    {
        const auto& id = argument->GetId();
        if (argument->GetManagementType() == ArgumentManagementType::User && !argument->HasUserBuffer())
        {
            if (!kernel.HasLauncher())
            {
                Logger::LogWarning("Kernel argument with id " + id + " has buffer managed by user, but its associated kernel "
                    + kernel.GetName() + " does not have a launcher defined by user");
            }
            continue;
        }
        if (argument->HasUserBuffer())
        {
            if (!m_Engine.HasBuffer(id))
            {
                m_Engine.AddCustomBuffer(*argument, argument->GetUserBuffer());
            }
            continue;
        }
        if (m_ReadOnlyCacheFlag && argument->GetAccessType() == ArgumentAccessType::ReadOnly && m_Engine.HasBuffer(id))
        {
            continue;
        }
        const auto action = m_Engine.UploadArgument(*argument, m_ComputeLayer->GetDefaultQueue());
        m_Engine.WaitForTransferAction(action);
    }
}

void KernelRunner::CleanupBuffers(const Kernel& kernel)
{
    for (const auto* argument : kernel.GetVectorArguments())
    {
        if (argument->GetManagementType() == ArgumentManagementType::User || argument->HasUserBuffer())
        {
            continue;
        }
        if (m_ReadOnlyCacheFlag && argument->GetAccessType() == ArgumentAccessType::ReadOnly)
        {
            continue;
        }
        m_Engine.ClearBuffer(argument->GetId());
    }
}

void KernelRunner::DownloadBuffers(const std::vector<BufferOutputDescriptor>& output)
{
    m_Engine.EnsureThreadContext();
    for (const auto& descriptor : output)
    {
        const auto action = m_Engine.DownloadArgument(descriptor.GetArgumentId(), m_Engine.GetDefaultQueue(),
            descriptor.GetOutputDestination(), descriptor.GetOutputSize());
        m_Engine.WaitForTransferAction(action);
    }
}

void KernelRunner::SetReadOnlyArgumentCache(const bool flag)
{
    m_ReadOnlyCacheFlag = flag;
}

void KernelRunner::SetProfiling(const bool flag)
{
    m_Engine.SetProfiling(flag);
}

bool KernelRunner::IsProfilingActive() const
{
    return m_Engine.IsProfilingActive();
}

void KernelRunner::SetValidationMethod(const ValidationMethod method, const double toleranceThreshold)
{
    m_Validator->SetValidationMethod(method, toleranceThreshold);
}

void KernelRunner::SetValidationMode(const ValidationMode mode)
{
    m_Validator->SetValidationMode(mode);
}

void KernelRunner::SetValidationRange(const ArgumentId& id, const size_t range)
{
    PrepareValidationData(id);
    m_Validator->SetValidationRange(id, range);
}

void KernelRunner::SetValueComparator(const ArgumentId& id, ValueComparator comparator)
{
    PrepareValidationData(id);
    m_Validator->SetValueComparator(id, comparator);
}

void KernelRunner::SetReferenceComputation(const ArgumentId& id, ReferenceComputation computation)
{
    PrepareValidationData(id);
    m_Validator->SetReferenceComputation(id, computation);
}

void KernelRunner::SetReferenceKernel(const ArgumentId& id, const Kernel& kernel, const KernelConfiguration& configuration,
    const KernelDimensions& dimensions)
{
    PrepareValidationData(id);
    m_Validator->SetReferenceKernel(id, kernel, configuration, dimensions);
}

void KernelRunner::SetReferenceArgument(const ArgumentId& id, const KernelArgument& argument)
{
    PrepareValidationData(id);
    m_Validator->SetReferenceArgument(id, argument);
}

void KernelRunner::ClearReferenceResult(const Kernel& kernel)
{
    m_Validator->ClearReferenceResult(kernel);
}

void KernelRunner::RemoveKernelData(const KernelId id)
{
    m_ComputeLayer->ClearData(id);
    m_Validator->RemoveDataWithReferenceKernel(id);
}

void KernelRunner::RemoveValidationData(const ArgumentId& id)
{
    m_Validator->RemoveValidationData(id);
}

KernelLauncher KernelRunner::GetKernelLauncher(const Kernel& kernel)
{
    if (kernel.HasLauncher())
    {
        return kernel.GetLauncher();
    }
    if (kernel.IsComposite())
    {
        throw KttException("Composite kernel must have kernel launcher defined by user");
    }
    const auto primaryId = kernel.GetPrimaryDefinition().GetId();
    if (IsProfilingActive())
    {
        return [primaryId](ComputeInterface& interface) { interface.RunKernelWithProfiling(primaryId); };
    }
    return [primaryId](ComputeInterface& interface) { interface.RunKernel(primaryId); };
}

KernelResult KernelRunner::RunKernelInternal(const Kernel& kernel, const KernelConfiguration& configuration,
    const KernelDimensions& dimensions, const KernelRunMode mode, KernelLauncher launcher,
    const std::vector<BufferOutputDescriptor>& output)
{
    m_ComputeLayer->AddData(kernel, configuration, dimensions, mode);
    const KernelId id = kernel.GetId();
    m_ComputeLayer->SetActiveKernel(id);
    KernelResult result(kernel.GetName(), configuration);

    Timer timer;
    timer.Start();
    try
    {
        Timer launcherTimer;
        launcherTimer.Start();
        launcher(*m_ComputeLayer);
        launcherTimer.Stop();
        result = m_ComputeLayer->GenerateResult(id, launcherTimer.GetElapsedTime());
        timer.Stop();
    }
    catch (const KttException& exception)
    {
        Logger::LogWarning(std::string("Kernel run failed with reason: ") + exception.what());
        try
        {
            m_ComputeLayer->SynchronizeDevice();
        }
        catch (const KttException&)
        {
        }
        m_ComputeLayer->ClearComputeEngineData();
        timer.Stop();
        result.SetFailedKernelOverhead(timer.GetElapsedTime());
        result.SetStatus(GetStatusFromException(exception.GetReason()));
    }

    const Nanoseconds dataOverhead = RunScopeTimer([this, &output]() { DownloadBuffers(output); });
    result.SetDataMovementOverhead(result.GetDataMovementOverhead() + dataOverhead);
    m_ComputeLayer->ClearActiveKernel();
    m_ComputeLayer->ClearData(id);
    return result;
}

void KernelRunner::PrepareValidationData(const ArgumentId& id)
{
    if (!m_Validator->HasValidationData(id))
    {
        m_Validator->InitializeValidationData(m_ArgumentManager.GetArgument(id));
    }
}

void KernelRunner::ValidateResult(const Kernel& kernel, KernelResult& result, const KernelRunMode mode)
{
    if (!result.IsValid())
    {
        return;
    }
    bool valid = false;
    const Nanoseconds overhead = RunScopeTimer([this, &kernel, mode, &valid]() { valid = m_Validator->ValidateArguments(kernel, mode); });
    result.SetValidationOverhead(overhead);

    const auto& time = TimeConfiguration::GetInstance();
    const std::string tag = time.GetUnitTag();
    const std::string durations = std::to_string(time.ConvertFromNanoseconds(result.GetTotalDuration())) + tag
        + ", kernel duration was " + std::to_string(time.ConvertFromNanoseconds(result.GetKernelDuration())) + tag;
    if (valid)
    {
        Logger::LogInfo("Kernel run completed successfully in " + durations);
        return;
    }
    Logger::LogWarning("Kernel run completed in " + durations + ", but results differ");
    result.SetStatus(ResultStatus::ValidationFailed);
}

ResultStatus KernelRunner::GetStatusFromException(const ExceptionReason reason)
{
    switch (reason)
    {
    case ExceptionReason::CompilerError:
        return ResultStatus::CompilationFailed;
    case ExceptionReason::DeviceLimitsExceeded:
        return ResultStatus::DeviceLimitsExceeded;
    default:
        return ResultStatus::ComputationFailed;
    }
}

} // namespace ktt
