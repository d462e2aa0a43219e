#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include <KttException.h>

#include "Common.h"
#include "Execution.h"

namespace ktt
{

// ---------------------------------------------------------------- KernelComputeData

KernelComputeData::KernelComputeData(const KernelDefinition& definition, const KernelConfiguration& configuration,
    const DimensionVector& globalSize, const DimensionVector& localSize) :
    m_Name(definition.GetName()),
    m_Templated(definition.GetTemplatedName()),
    m_Source(definition.GetSource()),
    m_Defines(configuration.GeneratePrefix()),
    m_Directory(definition.GetSourceDirectory()),
    m_Configuration(configuration),
    m_Global(globalSize),
    m_Local(localSize),
    m_Arguments(definition.GetArguments())
{}

void KernelComputeData::SwapArguments(const ArgumentId& first, const ArgumentId& second)
{
    const auto byId = [this](const ArgumentId& id)
    {
        return std::find_if(m_Arguments.begin(), m_Arguments.end(), [&id](const KernelArgument* argument) { return argument->GetId() == id; });
    };
    const auto a = byId(first);
    const auto b = byId(second);
    if (a == m_Arguments.end() || b == m_Arguments.end())
    {
        throw KttException("Arguments " + first + " and " + second + " are not both bound to kernel " + m_Name);
    }
    std::iter_swap(a, b);
}

// ---------------------------------------------------------------- RunRecord

RunRecord::RunRecord(const Kernel& kernel, const KernelConfiguration& configuration, const KernelDimensions& dimensions,
    const KernelRunMode mode) :
    m_Kernel(kernel),
    m_Configuration(configuration),
    m_Mode(mode),
    m_TransferTime(0)
{
    const auto& pairs = configuration.GetPairs();
    for (const KernelDefinition* definition : kernel.GetDefinitions())
    {
        const KernelDefinitionId id = definition->GetId();
        // Sizes given for this run replace the definition's defaults before the modifiers apply.
        const auto given = dimensions.find(id);
        const DimensionVector& global = given != dimensions.end() ? given->second.first : definition->GetGlobalSize();
        const DimensionVector& local = given != dimensions.end() ? given->second.second : definition->GetLocalSize();
        m_Launches.emplace_back(id, KernelComputeData(*definition, configuration,
            kernel.GetModifiedSize(id, global, ModifierType::Global, pairs), kernel.GetModifiedSize(id, local, ModifierType::Local, pairs)));
    }
}

KernelComputeData& RunRecord::Launch(const KernelDefinitionId id)
{
    return const_cast<KernelComputeData&>(static_cast<const RunRecord*>(this)->Launch(id));
}

const KernelComputeData& RunRecord::Launch(const KernelDefinitionId id) const
{
    for (const auto& [definition, data] : m_Launches)
    {
        if (definition == id)
        {
            return data;
        }
    }
    throw KttException("Kernel " + m_Kernel.GetName() + " has no definition with id " + std::to_string(id));
}

KernelArgument* RunRecord::FindSubstitute(const ArgumentId& id) const
{
    for (const auto& substitute : m_Substitutes)
    {
        if (substitute->GetId() == id)
        {
            return substitute.get();
        }
    }
    return nullptr;
}

void RunRecord::Substitute(const ArgumentId& id, std::unique_ptr<KernelArgument> replacement)
{
    KernelArgument* fresh = replacement.get();
    m_Substitutes.push_back(std::move(replacement));
    for (auto& [definition, data] : m_Launches)
    {
        std::vector<KernelArgument*> bound = data.GetArguments();
        std::replace_if(bound.begin(), bound.end(), [&id](const KernelArgument* argument) { return argument->GetId() == id; }, fresh);
        data.SetArguments(bound);
    }
}

void RunRecord::Rebind(const KernelDefinitionId id, std::vector<KernelArgument*> arguments)
{
    for (auto& argument : arguments)
    {
        if (KernelArgument* substitute = FindSubstitute(argument->GetId()))
        {
            argument = substitute;
        }
    }
    Launch(id).SetArguments(arguments);
}

KernelResult RunRecord::Finish(const Nanoseconds launcherTime) const
{
    Nanoseconds accounted = m_TransferTime;
    for (const auto& computation : m_Computations)
    {
        accounted += computation.GetDuration() + computation.GetOverhead();
    }
    const Nanoseconds rest = launcherTime > accounted ? launcherTime - accounted : 0;

    KernelResult result(m_Kernel.GetName(), m_Configuration, m_Computations);
    // A user launcher's own host work counts towards the run's duration; the implicit launcher's
    // bookkeeping does not.
    if (m_Kernel.HasLauncher())
    {
        result.SetExtraDuration(rest);
        result.SetDataMovementOverhead(m_TransferTime);
    }
    else
    {
        result.SetDataMovementOverhead(m_TransferTime + rest);
    }
    return result;
}

// ---------------------------------------------------------------- LaunchContext

LaunchContext::LaunchContext(ComputeEngine& engine, ArgumentRegistry& arguments) :
    m_Engine(engine),
    m_Arguments(arguments),
    m_Record(nullptr)
{}

RunRecord& LaunchContext::Record() const
{
    if (m_Record == nullptr)
    {
        throw KttException("The compute interface can only be used inside a kernel launcher");
    }
    return *m_Record;
}

template <typename Transfer>
void LaunchContext::TimedTransfer(Transfer transfer)
{
    RunRecord& record = Record();
    record.AddTransferTime(RunScopeTimer([&]() { m_Engine.WaitForTransferAction(transfer()); }));
}

void LaunchContext::RunKernel(const KernelDefinitionId id)
{
    WaitForComputeAction(RunKernelAsync(id, GetDefaultQueue()));
}

void LaunchContext::RunKernel(const KernelDefinitionId id, const DimensionVector& globalSize, const DimensionVector& localSize)
{
    WaitForComputeAction(RunKernelAsync(id, GetDefaultQueue(), globalSize, localSize));
}

ComputeActionId LaunchContext::RunKernelAsync(const KernelDefinitionId id, const QueueId queue)
{
    return m_Engine.RunKernelAsync(Record().Launch(id), queue);
}

ComputeActionId LaunchContext::RunKernelAsync(const KernelDefinitionId id, const QueueId queue, const DimensionVector& globalSize,
    const DimensionVector& localSize)
{
    KernelComputeData launch = Record().Launch(id);
    launch.SetGlobalSize(globalSize);
    launch.SetLocalSize(localSize);
    return m_Engine.RunKernelAsync(launch, queue);
}

void LaunchContext::WaitForComputeAction(const ComputeActionId id)
{
    Record().AddComputation(m_Engine.WaitForComputeAction(id));
}

void LaunchContext::RunKernelWithProfiling(const KernelDefinitionId id)
{
    Record().AddComputation(m_Engine.RunKernelWithProfiling(Record().Launch(id), GetDefaultQueue()));
}

void LaunchContext::RunKernelWithProfiling(const KernelDefinitionId id, const DimensionVector& globalSize, const DimensionVector& localSize)
{
    KernelComputeData launch = Record().Launch(id);
    launch.SetGlobalSize(globalSize);
    launch.SetLocalSize(localSize);
    Record().AddComputation(m_Engine.RunKernelWithProfiling(launch, GetDefaultQueue()));
}

uint64_t LaunchContext::GetRemainingProfilingRuns(const KernelDefinitionId id) const
{
    return m_Engine.GetRemainingProfilingRuns(Record().Launch(id).GetUniqueIdentifier());
}

uint64_t LaunchContext::GetRemainingProfilingRuns() const
{
    uint64_t runs = 0;
    for (const KernelDefinition* definition : Record().GetKernel().GetProfiledDefinitions())
    {
        runs += GetRemainingProfilingRuns(definition->GetId());
    }
    return runs;
}

QueueId LaunchContext::GetDefaultQueue() const { return m_Engine.GetDefaultQueue(); }
std::vector<QueueId> LaunchContext::GetAllQueues() const { return m_Engine.GetAllQueues(); }
void LaunchContext::SynchronizeQueue(const QueueId queue) { m_Engine.SynchronizeQueue(queue); }
void LaunchContext::SynchronizeQueues() { m_Engine.SynchronizeQueues(); }
void LaunchContext::SynchronizeDevice() { m_Engine.SynchronizeDevice(); }

const DimensionVector& LaunchContext::GetCurrentGlobalSize(const KernelDefinitionId id) const { return Record().Launch(id).GetGlobalSize(); }
const DimensionVector& LaunchContext::GetCurrentLocalSize(const KernelDefinitionId id) const { return Record().Launch(id).GetLocalSize(); }
const KernelConfiguration& LaunchContext::GetCurrentConfiguration() const { return Record().GetConfiguration(); }
KernelRunMode LaunchContext::GetRunMode() const { return Record().GetMode(); }

void LaunchContext::ChangeArguments(const KernelDefinitionId id, const std::vector<ArgumentId>& arguments)
{
    Record().Rebind(id, m_Arguments.Resolve(arguments));
}

void LaunchContext::SwapArguments(const KernelDefinitionId id, const ArgumentId& first, const ArgumentId& second)
{
    Record().Launch(id).SwapArguments(first, second);
}

void LaunchContext::UpdateScalarArgument(const ArgumentId& id, const void* data)
{
    // The new value holds for this run only; the registered argument is left as it is.
    const KernelArgument& original = m_Arguments.Get(id);
    if (original.GetMemoryType() != ArgumentMemoryType::Scalar)
    {
        throw KttException("Argument " + id + " is not a scalar");
    }
    auto value = std::make_unique<KernelArgument>(original);
    value->SetOwnedData(data, original.GetDataSize());
    Record().Substitute(id, std::move(value));
}

void LaunchContext::UpdateLocalArgument(const ArgumentId& id, const size_t dataSize)
{
    const KernelArgument& original = m_Arguments.Get(id);
    if (original.GetMemoryType() != ArgumentMemoryType::Local)
    {
        throw KttException("Argument " + id + " is not a local-memory argument");
    }
    auto value = std::make_unique<KernelArgument>(original);
    value->SetLocalSize(dataSize);
    Record().Substitute(id, std::move(value));
}

void LaunchContext::UploadBuffer(const ArgumentId& id)
{
    TimedTransfer([&]() { return UploadBufferAsync(id, GetDefaultQueue()); });
}

TransferActionId LaunchContext::UploadBufferAsync(const ArgumentId& id, const QueueId queue)
{
    return m_Engine.UploadArgument(m_Arguments.Get(id), queue);
}

void LaunchContext::DownloadBuffer(const ArgumentId& id, void* destination, const size_t dataSize)
{
    TimedTransfer([&]() { return DownloadBufferAsync(id, GetDefaultQueue(), destination, dataSize); });
}

TransferActionId LaunchContext::DownloadBufferAsync(const ArgumentId& id, const QueueId queue, void* destination, const size_t dataSize)
{
    return m_Engine.DownloadArgument(id, queue, destination, dataSize);
}

void LaunchContext::UpdateBuffer(const ArgumentId& id, const void* data, const size_t dataSize)
{
    TimedTransfer([&]() { return UpdateBufferAsync(id, GetDefaultQueue(), data, dataSize); });
}

TransferActionId LaunchContext::UpdateBufferAsync(const ArgumentId& id, const QueueId queue, const void* data, const size_t dataSize)
{
    return m_Engine.UpdateArgument(id, queue, data, dataSize);
}

void LaunchContext::CopyBuffer(const ArgumentId& destination, const ArgumentId& source, const size_t dataSize)
{
    TimedTransfer([&]() { return CopyBufferAsync(destination, source, GetDefaultQueue(), dataSize); });
}

TransferActionId LaunchContext::CopyBufferAsync(const ArgumentId& destination, const ArgumentId& source, const QueueId queue,
    const size_t dataSize)
{
    return m_Engine.CopyArgument(destination, queue, source, dataSize);
}

void LaunchContext::WaitForTransferAction(const TransferActionId id) { m_Engine.WaitForTransferAction(id); }
void LaunchContext::ResizeBuffer(const ArgumentId& id, const size_t newDataSize, const bool preserveData)
{
    m_Engine.ResizeArgument(id, newDataSize, preserveData);
}
void LaunchContext::ClearBuffer(const ArgumentId& id) { m_Engine.ClearBuffer(id); }
bool LaunchContext::HasBuffer(const ArgumentId& id) { return m_Engine.HasBuffer(id); }
void LaunchContext::GetUnifiedMemoryBufferHandle(const ArgumentId& id, UnifiedBufferMemory& memoryHandle)
{
    m_Engine.GetUnifiedMemoryBufferHandle(id, memoryHandle);
}
bool LaunchContext::GetProfiling([[maybe_unused]] const KernelId id) { return m_Engine.IsProfilingActive(); }

// ---------------------------------------------------------------- OutputChecker

namespace
{

std::string Mismatch(const ArgumentId& id, const size_t index, const double actual, const double expected)
{
    return "Output " + id + " differs at element " + std::to_string(index) + ": got " + std::to_string(actual) + ", expected "
        + std::to_string(expected);
}

/** Comparison of `count` elements under the configured method; integers must match exactly. */
template <typename T>
bool CompareTyped(const ArgumentId& id, const ValidationMethod method, const double tolerance, const T* actual, const T* expected,
    const size_t count)
{
    if constexpr (!std::is_floating_point_v<T>)
    {
        const auto* first = std::mismatch(actual, actual + count, expected).first;
        if (first == actual + count)
        {
            return true;
        }
        const size_t index = static_cast<size_t>(first - actual);
        Logger::LogWarning(Mismatch(id, index, static_cast<double>(actual[index]), static_cast<double>(expected[index])));
        return false;
    }
    else
    {
        if (method == ValidationMethod::AbsoluteDifference)
        {
            // The summed absolute difference over the whole range against the tolerance.
            double sum = 0.0;
            for (size_t i = 0; i < count; ++i)
            {
                sum += std::fabs(static_cast<double>(actual[i]) - static_cast<double>(expected[i]));
            }
            if (sum > tolerance)
            {
                Logger::LogWarning("Output " + id + " differs: summed absolute difference " + std::to_string(sum));
                return false;
            }
            return true;
        }
        for (size_t i = 0; i < count; ++i)
        {
            const T difference = std::fabs(actual[i] - expected[i]);
            bool bad = false;
            if (method == ValidationMethod::SideBySideComparison)
            {
                bad = difference > tolerance;
            }
            else
            {
                // Relative comparison, ignoring absolute differences up to 1e-4.
                bad = difference > static_cast<T>(1e-4) && difference / expected[i] > tolerance;
            }
            if (bad)
            {
                Logger::LogWarning(Mismatch(id, i, static_cast<double>(actual[i]), static_cast<double>(expected[i])));
                return false;
            }
        }
        return true;
    }
}

} // namespace

OutputChecker::OutputChecker(Executor& executor) :
    m_Executor(executor),
    m_Method(ValidationMethod::SideBySideComparison),
    m_Tolerance(1e-4),
    m_Modes(ValidationMode::OfflineTuning | ValidationMode::OnlineTuning)
{}

void OutputChecker::SetMethod(const ValidationMethod method, const double tolerance)
{
    if (tolerance < 0.0)
    {
        throw KttException("The validation tolerance must not be negative");
    }
    m_Method = method;
    m_Tolerance = tolerance;
}

namespace
{

/** Forgets how the expected output was obtained; the range and the comparator stay. */
void ClearSource(OutputExpectation& expectation)
{
    expectation.computation = nullptr;
    expectation.referenceKernel = nullptr;
    expectation.referenceConfiguration = KernelConfiguration();
    expectation.referenceDimensions.clear();
    expectation.referenceArgument = nullptr;
    expectation.expected.clear();
    expectation.ready = false;
}

} // namespace

OutputExpectation& OutputChecker::Expect(const KernelArgument& argument)
{
    auto found = m_Expectations.find(argument.GetId());
    if (found != m_Expectations.end())
    {
        return found->second;
    }
    if (argument.GetMemoryType() != ArgumentMemoryType::Vector || argument.GetAccessType() == ArgumentAccessType::ReadOnly)
    {
        throw KttException("Only writable vector arguments can be validated; " + argument.GetId() + " is not one");
    }
    OutputExpectation& expectation = m_Expectations[argument.GetId()];
    expectation.argument = &argument;
    expectation.elements = static_cast<size_t>(argument.GetNumberOfElements());
    return expectation;
}

void OutputChecker::SetRange(const KernelArgument& argument, const size_t elements)
{
    OutputExpectation& expectation = Expect(argument);
    expectation.elements = elements;
    expectation.expected.clear();
    expectation.ready = false;
}

void OutputChecker::SetComparator(const KernelArgument& argument, ValueComparator comparator)
{
    if (!comparator)
    {
        throw KttException("The value comparator must be callable");
    }
    Expect(argument).comparator = comparator;
}

void OutputChecker::SetComputation(const KernelArgument& argument, ReferenceComputation computation)
{
    OutputExpectation& expectation = Expect(argument);
    ClearSource(expectation);
    expectation.computation = computation;
}

void OutputChecker::SetKernel(const KernelArgument& argument, const Kernel& kernel, const KernelConfiguration& configuration,
    const KernelDimensions& dimensions)
{
    OutputExpectation& expectation = Expect(argument);
    ClearSource(expectation);
    expectation.referenceKernel = &kernel;
    expectation.referenceConfiguration = configuration;
    expectation.referenceDimensions = dimensions;
}

void OutputChecker::SetArgument(const KernelArgument& argument, const KernelArgument& reference)
{
    if (reference.GetMemoryType() != ArgumentMemoryType::Vector || reference.GetOwnership() == ArgumentOwnership::User)
    {
        throw KttException("Reference argument " + reference.GetId() + " must be a vector with host data");
    }
    OutputExpectation& expectation = Expect(argument);
    ClearSource(expectation);
    expectation.referenceArgument = &reference;
    expectation.ready = true;
}

void OutputChecker::ForgetReferenceKernel(const KernelId id)
{
    for (auto entry = m_Expectations.begin(); entry != m_Expectations.end();)
    {
        const bool uses = entry->second.referenceKernel != nullptr && entry->second.referenceKernel->GetId() == id;
        entry = uses ? m_Expectations.erase(entry) : std::next(entry);
    }
}

bool OutputChecker::Checked(const KernelRunMode mode) const
{
    switch (mode)
    {
    case KernelRunMode::Running:
        return HasFlag(m_Modes, ValidationMode::Running);
    case KernelRunMode::OfflineTuning:
        return HasFlag(m_Modes, ValidationMode::OfflineTuning);
    case KernelRunMode::OnlineTuning:
        return HasFlag(m_Modes, ValidationMode::OnlineTuning);
    default:
        return false;
    }
}

void OutputChecker::Compute(OutputExpectation& expectation)
{
    const size_t bytes = expectation.elements * expectation.argument->GetElementSize();
    const ArgumentId& id = expectation.argument->GetId();
    const Nanoseconds took = RunScopeTimer([&]()
    {
        expectation.expected.assign(bytes, 0);
        if (expectation.computation)
        {
            expectation.computation(expectation.expected.data());
        }
        else if (expectation.referenceKernel != nullptr)
        {
            // The reference kernel runs unprofiled and is not itself checked.
            const bool profiling = m_Executor.IsProfiling();
            m_Executor.SetProfiling(false);
            const KernelResult run = m_Executor.Execute(*expectation.referenceKernel, expectation.referenceConfiguration,
                expectation.referenceDimensions, KernelRunMode::ResultValidation, {BufferOutputDescriptor(id, expectation.expected.data(), bytes)});
            m_Executor.SetProfiling(profiling);
            if (!run.IsValid())
            {
                expectation.expected.clear();
                throw KttException("Reference kernel " + expectation.referenceKernel->GetName() + " failed: "
                    + ResultStatusToString(run.GetStatus()));
            }
        }
    });
    expectation.ready = true;
    const auto& time = TimeConfiguration::GetInstance();
    Logger::LogInfo("Expected output of " + id + " computed in " + std::to_string(time.ConvertFromNanoseconds(took)) + time.GetUnitTag());
}

void OutputChecker::Prepare(const Kernel& kernel, const KernelRunMode mode)
{
    if (!Checked(mode))
    {
        return;
    }
    for (const KernelArgument* argument : kernel.GetVectorArguments())
    {
        auto found = m_Expectations.find(argument->GetId());
        if (found == m_Expectations.end() || found->second.ready)
        {
            continue;
        }
        if (found->second.computation || found->second.referenceKernel != nullptr)
        {
            Compute(found->second);
        }
    }
}

void OutputChecker::Drop(const Kernel& kernel)
{
    for (const KernelArgument* argument : kernel.GetVectorArguments())
    {
        auto found = m_Expectations.find(argument->GetId());
        if (found != m_Expectations.end() && found->second.referenceArgument == nullptr)
        {
            found->second.expected.clear();
            found->second.ready = false;
        }
    }
}

bool OutputChecker::Compare(const OutputExpectation& expectation, const std::vector<uint8_t>& actual) const
{
    const KernelArgument& argument = *expectation.argument;
    const ArgumentId& id = argument.GetId();
    const void* expected = expectation.referenceArgument != nullptr ? expectation.referenceArgument->GetData()
                                                                     : static_cast<const void*>(expectation.expected.data());
    const size_t count = expectation.elements;
    if (expectation.comparator)
    {
        const size_t size = argument.GetElementSize();
        for (size_t i = 0; i < count; ++i)
        {
            if (!expectation.comparator(actual.data() + i * size, static_cast<const uint8_t*>(expected) + i * size))
            {
                Logger::LogWarning("Output " + id + " differs at element " + std::to_string(i) + " (user comparator)");
                return false;
            }
        }
        return true;
    }
    const auto typed = [&](auto sample)
    {
        using T = decltype(sample);
        return CompareTyped(id, m_Method, m_Tolerance, reinterpret_cast<const T*>(actual.data()), static_cast<const T*>(expected), count);
    };
    switch (argument.GetDataType())
    {
    case ArgumentDataType::Char: return typed(int8_t{});
    case ArgumentDataType::UnsignedChar: return typed(uint8_t{});
    case ArgumentDataType::Short: return typed(int16_t{});
    case ArgumentDataType::UnsignedShort: return typed(uint16_t{});
    case ArgumentDataType::Int: return typed(int32_t{});
    case ArgumentDataType::UnsignedInt: return typed(uint32_t{});
    case ArgumentDataType::Long: return typed(int64_t{});
    case ArgumentDataType::UnsignedLong: return typed(uint64_t{});
    case ArgumentDataType::Float: return typed(float{});
    case ArgumentDataType::Double: return typed(double{});
    default:
        throw KttException("Output " + id + " has a custom element type: validating it needs a value comparator");
    }
}

bool OutputChecker::Check(const Kernel& kernel, const KernelRunMode mode)
{
    if (!Checked(mode))
    {
        return true;
    }
    bool matches = true;
    for (const KernelArgument* argument : kernel.GetVectorArguments())
    {
        const auto found = m_Expectations.find(argument->GetId());
        if (found == m_Expectations.end())
        {
            continue;
        }
        const OutputExpectation& expectation = found->second;
        if (!expectation.ready)
        {
            continue;
        }
        std::vector<uint8_t> actual(expectation.elements * argument->GetElementSize());
        m_Executor.Fetch({BufferOutputDescriptor(argument->GetId(), actual.data(), actual.size())});
        matches = Compare(expectation, actual) && matches;
    }
    return matches;
}

// ---------------------------------------------------------------- Executor

namespace
{

ResultStatus StatusFor(const ExceptionReason reason)
{
    if (reason == ExceptionReason::CompilerError)
    {
        return ResultStatus::CompilationFailed;
    }
    if (reason == ExceptionReason::DeviceLimitsExceeded)
    {
        return ResultStatus::DeviceLimitsExceeded;
    }
    return ResultStatus::ComputationFailed;
}

} // namespace

Executor::Executor(ComputeEngine& engine, ArgumentRegistry& arguments) :
    m_Engine(engine),
    m_Arguments(arguments),
    m_Context(std::make_unique<LaunchContext>(engine, arguments)),
    m_Checker(std::make_unique<OutputChecker>(*this)),
    m_KeepReadOnly(true)
{}

void Executor::StageBuffers(const Kernel& kernel)
{
    for (KernelArgument* argument : kernel.GetVectorArguments())
    {
        const ArgumentId& id = argument->GetId();
        if (argument->HasUserBuffer())
        {
            // Device memory owned by the caller: registered once, never copied.
            if (!m_Engine.HasBuffer(id))
            {
                m_Engine.AddCustomBuffer(*argument, argument->GetUserBuffer());
            }
            continue;
        }
        if (argument->GetManagementType() == ArgumentManagementType::User)
        {
            if (!kernel.HasLauncher())
            {
                Logger::LogWarning("Argument " + id + " is managed by the user but kernel " + kernel.GetName() + " has no launcher");
            }
            continue;
        }
        const bool cached = m_KeepReadOnly && argument->GetAccessType() == ArgumentAccessType::ReadOnly && m_Engine.HasBuffer(id);
        if (!cached)
        {
            m_Engine.WaitForTransferAction(m_Engine.UploadArgument(*argument, m_Engine.GetDefaultQueue()));
        }
    }
}

void Executor::ReleaseBuffers(const Kernel& kernel)
{
    for (const KernelArgument* argument : kernel.GetVectorArguments())
    {
        const bool external = argument->HasUserBuffer() || argument->GetManagementType() == ArgumentManagementType::User;
        const bool kept = m_KeepReadOnly && argument->GetAccessType() == ArgumentAccessType::ReadOnly;
        if (!external && !kept)
        {
            m_Engine.ClearBuffer(argument->GetId());
        }
    }
}

void Executor::Fetch(const std::vector<BufferOutputDescriptor>& output)
{
    m_Engine.EnsureThreadContext();
    for (const auto& descriptor : output)
    {
        m_Engine.WaitForTransferAction(m_Engine.DownloadArgument(descriptor.GetArgumentId(), m_Engine.GetDefaultQueue(),
            descriptor.GetOutputDestination(), descriptor.GetOutputSize()));
    }
}

KernelLauncher Executor::LauncherFor(const Kernel& kernel) const
{
    if (kernel.HasLauncher())
    {
        return kernel.GetLauncher();
    }
    if (kernel.IsComposite())
    {
        throw KttException("Composite kernel " + kernel.GetName() + " needs a launcher");
    }
    const KernelDefinitionId only = kernel.GetPrimaryDefinition().GetId();
    const bool profiled = m_Engine.IsProfilingActive();
    return [only, profiled](ComputeInterface& interface)
    {
        if (profiled)
        {
            interface.RunKernelWithProfiling(only);
        }
        else
        {
            interface.RunKernel(only);
        }
    };
}

KernelResult Executor::Launch(const Kernel& kernel, const KernelConfiguration& configuration, const KernelDimensions& dimensions,
    const KernelRunMode mode, const std::vector<BufferOutputDescriptor>& output)
{
    KernelResult result(kernel.GetName(), configuration);
    const auto started = std::chrono::steady_clock::now();
    try
    {
        RunRecord record(kernel, configuration, dimensions, mode);
        const KernelLauncher launcher = LauncherFor(kernel);
        m_Context->Begin(record);
        const Nanoseconds launcherTime = RunScopeTimer([&]() { launcher(*m_Context); });
        m_Context->End();
        result = record.Finish(launcherTime);
    }
    catch (const KttException& failure)
    {
        m_Context->End();
        Logger::LogWarning(std::string("Kernel run failed: ") + failure.what());
        // Leave the device idle and drop compiled kernels that may be in a bad state.
        try
        {
            m_Engine.SynchronizeDevice();
        }
        catch (const KttException&)
        {
        }
        m_Engine.ClearKernelCache();
        result.SetStatus(StatusFor(failure.GetReason()));
        result.SetFailedKernelOverhead(static_cast<Nanoseconds>(
            std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - started).count()));
    }
    const Nanoseconds download = RunScopeTimer([&]() { Fetch(output); });
    result.SetDataMovementOverhead(result.GetDataMovementOverhead() + download);
    return result;
}

KernelResult Executor::Execute(const Kernel& kernel, const KernelConfiguration& configuration, const KernelDimensions& dimensions,
    const KernelRunMode mode, const std::vector<BufferOutputDescriptor>& output, const bool manageBuffers)
{
    m_Engine.EnsureThreadContext();
    m_Checker->Prepare(kernel, mode);

    Nanoseconds staging = 0;
    if (manageBuffers)
    {
        staging += RunScopeTimer([&]() { StageBuffers(kernel); });
    }
    Logger::LogInfo(std::string(IsProfiling() ? "Profiling " : "Running ") + kernel.GetName() + " with " + configuration.GetString());

    KernelResult result = Launch(kernel, configuration, dimensions, mode, output);

    if (result.IsValid())
    {
        bool matches = true;
        result.SetValidationOverhead(RunScopeTimer([&]() { matches = m_Checker->Check(kernel, mode); }));
        const auto& time = TimeConfiguration::GetInstance();
        const std::string took = std::to_string(time.ConvertFromNanoseconds(result.GetTotalDuration())) + time.GetUnitTag() + " (kernels "
            + std::to_string(time.ConvertFromNanoseconds(result.GetKernelDuration())) + time.GetUnitTag() + ")";
        if (matches)
        {
            Logger::LogInfo("Run finished in " + took);
        }
        else
        {
            Logger::LogWarning("Run finished in " + took + " but its output does not match the reference");
            result.SetStatus(ResultStatus::ValidationFailed);
        }
    }

    if (manageBuffers)
    {
        staging += RunScopeTimer([&]() { ReleaseBuffers(kernel); });
    }
    result.SetDataMovementOverhead(result.GetDataMovementOverhead() + staging);
    return result;
}

} // namespace ktt
