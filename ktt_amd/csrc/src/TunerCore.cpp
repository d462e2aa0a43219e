#include <chrono>
#include <ctime>
#include <fstream>
#include <sstream>

#include <Ktt.h>

#include "Common.h"
#include "HipEngine.h"
#include "TunerCore.h"

namespace ktt
{

static std::unique_ptr<ComputeEngine> CreateEngine(const PlatformIndex platform, const DeviceIndex device, const ComputeApi api,
    const uint32_t queueCount)
{
    switch (api)
    {
    case ComputeApi::HIP:
        return std::make_unique<HipEngine>(platform, device, queueCount);
    default:
        throw KttException("Compute API " + ComputeApiToString(api) + " is not available in this build; use ComputeApi::HIP");
    }
}

TunerCore::TunerCore(const PlatformIndex platform, const DeviceIndex device, const ComputeApi api, const uint32_t queueCount) :
    m_Arguments(std::make_unique<ArgumentRegistry>()),
    m_Kernels(std::make_unique<KernelRegistry>(*m_Arguments))
{
    m_ComputeEngine = CreateEngine(platform, device, api, queueCount);
    CreateDrivers();
}

TunerCore::TunerCore(const ComputeApi api, const ComputeApiInitializer& initializer, std::vector<QueueId>& assignedQueueIds) :
    m_Arguments(std::make_unique<ArgumentRegistry>()),
    m_Kernels(std::make_unique<KernelRegistry>(*m_Arguments))
{
    if (api != ComputeApi::HIP)
    {
        throw KttException("Compute API " + ComputeApiToString(api) + " is not available in this build; use ComputeApi::HIP");
    }
    m_ComputeEngine = std::make_unique<HipEngine>(initializer, assignedQueueIds);
    CreateDrivers();
}

TunerCore::~TunerCore()
{
    // The loop and the executor hold references to the engine.
    m_TuningLoop.reset();
    m_Executor.reset();
    m_ComputeEngine.reset();
}

void TunerCore::CreateDrivers()
{
    m_Executor = std::make_unique<Executor>(*m_ComputeEngine, *m_Arguments);
    m_TuningLoop = std::make_unique<TuningLoop>(*m_Executor);
}

void TunerCore::RemoveKernel(const KernelId id)
{
    m_Kernels->RemoveKernel(id);
    m_TuningLoop->Reset(id, true);
    m_Executor->ForgetKernel(id);
}

void TunerCore::RemoveArgument(const ArgumentId& id)
{
    if (m_Kernels->UsesArgument(id))
    {
        throw KttException("Argument " + id + " is still bound to a kernel definition");
    }
    m_ComputeEngine->ClearBuffer(id);
    m_Executor->Checker().Forget(id);
    m_Arguments->Remove(id);
}

std::string TunerCore::GetKernelSource(const KernelId id, const KernelConfiguration& configuration) const
{
    const Kernel& kernel = m_Kernels->GetKernel(id);
    if (kernel.IsComposite())
    {
        throw KttException("Kernel " + kernel.GetName() + " has several definitions; ask for one definition's source");
    }
    return configuration.GeneratePrefix() + kernel.GetPrimaryDefinition().GetSource();
}

std::string TunerCore::GetKernelDefinitionSource(const KernelDefinitionId id, const KernelConfiguration& configuration) const
{
    return configuration.GeneratePrefix() + m_Kernels->GetDefinition(id).GetSource();
}

JsonValue TunerCore::CreateMetadata() const
{
    JsonValue metadata = JsonValue::MakeObject();
    metadata["ComputeApi"] = ComputeApiToString(m_ComputeEngine->GetComputeApi());
    metadata["Device"] = m_ComputeEngine->GetCurrentDeviceInfo().GetName();
    metadata["GlobalSizeType"] = GlobalSizeTypeToString(m_ComputeEngine->GetGlobalSizeType());
    metadata["KttVersion"] = KTT_VERSION_STRING;
    metadata["Platform"] = m_ComputeEngine->GetCurrentPlatformInfo().GetName();
    metadata["TimeUnit"] = TimeUnitToString(TimeConfiguration::GetInstance().GetTimeUnit());
    const auto now = std::chrono::system_clock::to_time_t(std::chrono::system_clock::now());
    std::tm utc{};
    gmtime_r(&now, &utc);
    char buffer[64];
    std::strftime(buffer, sizeof(buffer), "%Y-%m-%d %H:%M:%S UTC", &utc);
    metadata["Timestamp"] = std::string(buffer);
    return metadata;
}

void TunerCore::SaveResults(const std::vector<KernelResult>& results, const std::string& filePath, const OutputFormat format,
    const UserData& data) const
{
    if (results.empty())
    {
        throw KttException("Unable to save results because input vector is empty");
    }
    if (format != OutputFormat::JSON)
    {
        throw KttException("Only JSON output is available in this build");
    }
    const std::string file = filePath + ".json";
    Logger::LogInfo("Saving kernel results to file: " + file);
    std::ofstream stream(file);
    if (!stream.is_open())
    {
        throw KttException("Unable to open file: " + file);
    }
    stream << SerializeResults(CreateMetadata(), results, data);
}

std::vector<KernelResult> TunerCore::LoadResults(const std::string& filePath, const OutputFormat format, UserData& data) const
{
    if (format != OutputFormat::JSON)
    {
        throw KttException("Only JSON input is available in this build");
    }
    const std::string file = filePath + ".json";
    Logger::LogInfo("Loading kernel results from file: " + file);
    std::ifstream stream(file);
    if (!stream.is_open())
    {
        throw KttException("Unable to open file: " + file);
    }
    std::stringstream buffer;
    buffer << stream.rdbuf();
    JsonValue metadata;
    auto results = DeserializeResults(buffer.str(), data, &metadata);
    if (metadata.Contains("TimeUnit")
        && TimeUnitFromString(metadata.At("TimeUnit").AsString()) != TimeConfiguration::GetInstance().GetTimeUnit())
    {
        Logger::LogWarning("Loaded kernel results use different time unit than tuner");
    }
    return results;
}

void TunerCore::Log(const LoggingLevel level, const std::string& message)
{
    Logger::GetLogger().Log(level, message);
}

// ---------------------------------------------------------------- JSON conversion

static JsonValue DimensionsToJson(const DimensionVector& vector)
{
    JsonValue value = JsonValue::MakeObject();
    value["X"] = static_cast<uint64_t>(vector.GetSizeX());
    value["Y"] = static_cast<uint64_t>(vector.GetSizeY());
    value["Z"] = static_cast<uint64_t>(vector.GetSizeZ());
    return value;
}

static DimensionVector DimensionsFromJson(const JsonValue& value)
{
    return DimensionVector(value.At("X").AsUint(), value.At("Y").AsUint(), value.At("Z").AsUint());
}

static JsonValue PairToJson(const ParameterPair& pair)
{
    JsonValue value = JsonValue::MakeObject();
    value["Name"] = pair.GetName();
    value["ValueType"] = ParameterValueTypeToString(pair.GetValueType());
    switch (pair.GetValueType())
    {
    case ParameterValueType::Int: value["Value"] = std::get<int64_t>(pair.GetValue()); break;
    case ParameterValueType::UnsignedInt: value["Value"] = pair.GetValueUint(); break;
    case ParameterValueType::Double: value["Value"] = std::get<double>(pair.GetValue()); break;
    case ParameterValueType::Bool: value["Value"] = std::get<bool>(pair.GetValue()); break;
    default: value["Value"] = pair.GetValueString(); break;
    }
    return value;
}

static ParameterPair PairFromJson(const JsonValue& value)
{
    const std::string name = value.At("Name").AsString();
    const auto type = ParameterValueTypeFromString(value.At("ValueType").AsString());
    const auto& raw = value.At("Value");
    switch (type)
    {
    case ParameterValueType::Int: return ParameterPair(name, raw.AsInt());
    case ParameterValueType::UnsignedInt: return ParameterPair(name, raw.AsUint());
    case ParameterValueType::Double:
    case ParameterValueType::Float: return ParameterPair(name, raw.AsDouble());
    case ParameterValueType::Bool: return ParameterPair(name, raw.AsBool());
    default: return ParameterPair(name, raw.AsString());
    }
}

static JsonValue CounterToJson(const KernelProfilingCounter& counter)
{
    JsonValue value = JsonValue::MakeObject();
    value["Name"] = counter.GetName();
    value["Type"] = ProfilingCounterTypeToString(counter.GetType());
    switch (counter.GetType())
    {
    case ProfilingCounterType::Int: value["Value"] = counter.GetValueInt(); break;
    case ProfilingCounterType::UnsignedInt:
    case ProfilingCounterType::Throughput:
    case ProfilingCounterType::UtilizationLevel: value["Value"] = counter.GetValueUint(); break;
    default: value["Value"] = counter.GetValueDouble(); break;
    }
    return value;
}

static KernelProfilingCounter CounterFromJson(const JsonValue& value)
{
    const std::string name = value.At("Name").AsString();
    const auto type = ProfilingCounterTypeFromString(value.At("Type").AsString());
    const auto& raw = value.At("Value");
    switch (type)
    {
    case ProfilingCounterType::Int: return KernelProfilingCounter(name, type, raw.AsInt());
    case ProfilingCounterType::UnsignedInt:
    case ProfilingCounterType::Throughput:
    case ProfilingCounterType::UtilizationLevel: return KernelProfilingCounter(name, type, raw.AsUint());
    default: return KernelProfilingCounter(name, type, raw.AsDouble());
    }
}

static JsonValue ComputationToJson(const ComputationResult& result)
{
    const auto& time = TimeConfiguration::GetInstance();
    JsonValue value = JsonValue::MakeObject();
    value["KernelFunction"] = result.GetKernelFunction();
    value["Duration"] = time.ConvertFromNanosecondsDouble(result.GetDuration() == InvalidDuration ? 0 : result.GetDuration());
    value["Overhead"] = time.ConvertFromNanosecondsDouble(result.GetOverhead());
    value["CompilationOverhead"] = time.ConvertFromNanosecondsDouble(result.GetCompilationOverhead());
    value["GlobalSize"] = DimensionsToJson(result.GetGlobalSize());
    value["LocalSize"] = DimensionsToJson(result.GetLocalSize());
    if (result.HasCompilationData())
    {
        const auto& data = result.GetCompilationData();
        JsonValue compilation = JsonValue::MakeObject();
        compilation["MaxWorkGroupSize"] = data.m_MaxWorkGroupSize;
        compilation["LocalMemorySize"] = data.m_LocalMemorySize;
        compilation["PrivateMemorySize"] = data.m_PrivateMemorySize;
        compilation["ConstantMemorySize"] = data.m_ConstantMemorySize;
        compilation["RegistersCount"] = data.m_RegistersCount;
        value["CompilationData"] = compilation;
    }
    if (result.HasProfilingData())
    {
        const auto& data = result.GetProfilingData();
        JsonValue profiling = JsonValue::MakeObject();
        JsonValue counters = JsonValue::MakeArray();
        for (const auto& counter : data.GetCounters())
        {
            counters.PushBack(CounterToJson(counter));
        }
        profiling["Counters"] = counters;
        profiling["RemainingProfilingRuns"] = data.GetRemainingProfilingRuns();
        value["ProfilingData"] = profiling;
    }
    if (result.HasPowerData())
    {
        value["PowerUsage"] = static_cast<uint64_t>(result.GetPowerUsage());
        value["EnergyConsumption"] = result.GetEnergyConsumption();
    }
    return value;
}

static ComputationResult ComputationFromJson(const JsonValue& value)
{
    const auto& time = TimeConfiguration::GetInstance();
    ComputationResult result(value.At("KernelFunction").AsString());
    // Results written by KTT 2.1.0 carry no per-computation compilation overhead.
    const double compilation = value.Contains("CompilationOverhead") ? value.At("CompilationOverhead").AsDouble() : 0.0;
    result.SetDurationData(time.ConvertToNanosecondsDouble(value.At("Duration").AsDouble()),
        time.ConvertToNanosecondsDouble(value.At("Overhead").AsDouble()), time.ConvertToNanosecondsDouble(compilation));
    result.SetSizeData(DimensionsFromJson(value.At("GlobalSize")), DimensionsFromJson(value.At("LocalSize")));
    if (value.Contains("CompilationData"))
    {
        const auto& compilation = value.At("CompilationData");
        auto data = std::make_unique<KernelCompilationData>();
        data->m_MaxWorkGroupSize = compilation.At("MaxWorkGroupSize").AsUint();
        data->m_LocalMemorySize = compilation.At("LocalMemorySize").AsUint();
        data->m_PrivateMemorySize = compilation.At("PrivateMemorySize").AsUint();
        data->m_ConstantMemorySize = compilation.At("ConstantMemorySize").AsUint();
        data->m_RegistersCount = compilation.At("RegistersCount").AsUint();
        result.SetCompilationData(std::move(data));
    }
    if (value.Contains("ProfilingData"))
    {
        const auto& profiling = value.At("ProfilingData");
        const uint64_t remaining = profiling.At("RemainingProfilingRuns").AsUint();
        std::unique_ptr<KernelProfilingData> data;
        if (remaining == 0)
        {
            std::vector<KernelProfilingCounter> counters;
            const auto& list = profiling.At("Counters");
            for (size_t i = 0; i < list.Size(); ++i)
            {
                counters.push_back(CounterFromJson(list[i]));
            }
            data = std::make_unique<KernelProfilingData>(counters);
        }
        else
        {
            data = std::make_unique<KernelProfilingData>(remaining);
        }
        result.SetProfilingData(std::move(data));
    }
    if (value.Contains("PowerUsage"))
    {
        result.SetPowerUsage(static_cast<uint32_t>(value.At("PowerUsage").AsUint()));
    }
    return result;
}

JsonValue ResultToJson(const KernelResult& result)
{
    const auto& time = TimeConfiguration::GetInstance();
    JsonValue value = JsonValue::MakeObject();
    value["KernelName"] = result.GetKernelName();
    value["Status"] = ResultStatusToString(result.GetStatus());
    value["TotalDuration"] = time.ConvertFromNanosecondsDouble(result.GetTotalDuration());
    value["TotalOverhead"] = time.ConvertFromNanosecondsDouble(result.GetTotalOverhead());
    value["ExtraDuration"] = time.ConvertFromNanosecondsDouble(result.GetExtraDuration());
    value["DataMovementOverhead"] = time.ConvertFromNanosecondsDouble(result.GetDataMovementOverhead());
    value["ValidationOverhead"] = time.ConvertFromNanosecondsDouble(result.GetValidationOverhead());
    value["SearcherOverhead"] = time.ConvertFromNanosecondsDouble(result.GetSearcherOverhead());
    value["CompilationOverhead"] = time.ConvertFromNanosecondsDouble(result.GetCompilationOverhead());
    value["ProfilingOverhead"] = time.ConvertFromNanosecondsDouble(result.GetProfilingTotalOverhead());
    JsonValue configuration = JsonValue::MakeArray();
    for (const auto& pair : result.GetConfiguration().GetPairs())
    {
        configuration.PushBack(PairToJson(pair));
    }
    value["Configuration"] = configuration;
    JsonValue computations = JsonValue::MakeArray();
    for (const auto& computation : result.GetResults())
    {
        computations.PushBack(ComputationToJson(computation));
    }
    value["ComputationResults"] = computations;
    return value;
}

KernelResult ResultFromJson(const JsonValue& value)
{
    const auto& time = TimeConfiguration::GetInstance();
    std::vector<ParameterPair> pairs;
    const auto& configuration = value.At("Configuration");
    for (size_t i = 0; i < configuration.Size(); ++i)
    {
        pairs.push_back(PairFromJson(configuration[i]));
    }
    std::vector<ComputationResult> computations;
    const auto& list = value.At("ComputationResults");
    for (size_t i = 0; i < list.Size(); ++i)
    {
        computations.push_back(ComputationFromJson(list[i]));
    }
    KernelResult result(value.At("KernelName").AsString(), KernelConfiguration(pairs), computations);
    result.SetStatus(ResultStatusFromString(value.At("Status").AsString()));
    result.SetExtraDuration(time.ConvertToNanosecondsDouble(value.At("ExtraDuration").AsDouble()));
    result.SetDataMovementOverhead(time.ConvertToNanosecondsDouble(value.At("DataMovementOverhead").AsDouble()));
    result.SetValidationOverhead(time.ConvertToNanosecondsDouble(value.At("ValidationOverhead").AsDouble()));
    result.SetSearcherOverhead(time.ConvertToNanosecondsDouble(value.At("SearcherOverhead").AsDouble()));
    return result;
}

std::string SerializeResults(const JsonValue& metadata, const std::vector<KernelResult>& results, const UserData& data)
{
    JsonValue output = JsonValue::MakeObject();
    output["Metadata"] = metadata;
    JsonValue list = JsonValue::MakeArray();
    for (const auto& result : results)
    {
        list.PushBack(ResultToJson(result));
    }
    output["Results"] = list;
    if (!data.empty())
    {
        JsonValue user = JsonValue::MakeObject();
        for (const auto& [key, entry] : data)
        {
            user[key] = entry;
        }
        output["UserData"] = user;
    }
    return output.Dump(2);
}

std::vector<KernelResult> DeserializeResults(const std::string& text, UserData& data, JsonValue* metadata)
{
    const JsonValue input = JsonValue::Parse(text);
    if (input.Contains("UserData"))
    {
        for (const auto& [key, value] : input.At("UserData").Items())
        {
            data[key] = value.AsString();
        }
    }
    if (metadata != nullptr && input.Contains("Metadata"))
    {
        *metadata = input.At("Metadata");
    }
    std::vector<KernelResult> results;
    const auto& list = input.At("Results");
    for (size_t i = 0; i < list.Size(); ++i)
    {
        results.push_back(ResultFromJson(list[i]));
    }
    return results;
}

} // namespace ktt
