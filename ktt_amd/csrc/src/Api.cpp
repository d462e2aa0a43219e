// Value types of the public API: DimensionVector, ParameterPair, KernelConfiguration,
// profiling data, ComputationResult, KernelResult; searchers; stop conditions; and the
// internal logger / time configuration.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <limits>
#include <sstream>

#include <DimensionVector.h>
#include <KernelConfiguration.h>
#include <KernelResult.h>
#include <KttException.h>
#include <Searcher.h>
#include <StopCondition.h>

#include "Common.h"

namespace ktt
{

// ---------------------------------------------------------------- DimensionVector

DimensionVector::DimensionVector() : DimensionVector(1, 1, 1) {}
DimensionVector::DimensionVector(const size_t sizeX) : DimensionVector(sizeX, 1, 1) {}
DimensionVector::DimensionVector(const size_t sizeX, const size_t sizeY) : DimensionVector(sizeX, sizeY, 1) {}
DimensionVector::DimensionVector(const size_t sizeX, const size_t sizeY, const size_t sizeZ) :
    m_SizeX(sizeX),
    m_SizeY(sizeY),
    m_SizeZ(sizeZ)
{}

DimensionVector::DimensionVector(const std::vector<size_t>& vector) :
    m_SizeX(1),
    m_SizeY(1),
    m_SizeZ(1)
{
    if (vector.size() > 3)
    {
        throw KttException("Dimension vector can have at most three elements");
    }
    if (vector.size() > 0) m_SizeX = vector[0];
    if (vector.size() > 1) m_SizeY = vector[1];
    if (vector.size() > 2) m_SizeZ = vector[2];
}

void DimensionVector::SetSizeX(const size_t sizeX) { m_SizeX = sizeX; }
void DimensionVector::SetSizeY(const size_t sizeY) { m_SizeY = sizeY; }
void DimensionVector::SetSizeZ(const size_t sizeZ) { m_SizeZ = sizeZ; }

void DimensionVector::SetSize(const ModifierDimension dimension, const size_t size)
{
    switch (dimension)
    {
    case ModifierDimension::X: m_SizeX = size; break;
    case ModifierDimension::Y: m_SizeY = size; break;
    case ModifierDimension::Z: m_SizeZ = size; break;
    }
}

void DimensionVector::Multiply(const DimensionVector& factor)
{
    m_SizeX *= factor.m_SizeX;
    m_SizeY *= factor.m_SizeY;
    m_SizeZ *= factor.m_SizeZ;
}

void DimensionVector::Divide(const DimensionVector& divisor)
{
    m_SizeX /= divisor.m_SizeX;
    m_SizeY /= divisor.m_SizeY;
    m_SizeZ /= divisor.m_SizeZ;
}

void DimensionVector::RoundUp(const DimensionVector& multiple)
{
    m_SizeX = ktt::RoundUp(m_SizeX, multiple.m_SizeX);
    m_SizeY = ktt::RoundUp(m_SizeY, multiple.m_SizeY);
    m_SizeZ = ktt::RoundUp(m_SizeZ, multiple.m_SizeZ);
}

void DimensionVector::DivideCeil(const DimensionVector& divisor)
{
    m_SizeX = (m_SizeX + divisor.m_SizeX - 1) / divisor.m_SizeX;
    m_SizeY = (m_SizeY + divisor.m_SizeY - 1) / divisor.m_SizeY;
    m_SizeZ = (m_SizeZ + divisor.m_SizeZ - 1) / divisor.m_SizeZ;
}

void DimensionVector::ModifyByValue(const size_t value, const ModifierAction action, const ModifierDimension dimension)
{
    size_t size = GetSize(dimension);
    switch (action)
    {
    case ModifierAction::Add: size += value; break;
    case ModifierAction::Subtract: size -= value; break;
    case ModifierAction::Multiply: size *= value; break;
    case ModifierAction::Divide: size /= value; break;
    case ModifierAction::DivideCeil: size = (size + value - 1) / value; break;
    }
    SetSize(dimension, size);
}

size_t DimensionVector::GetSizeX() const { return m_SizeX; }
size_t DimensionVector::GetSizeY() const { return m_SizeY; }
size_t DimensionVector::GetSizeZ() const { return m_SizeZ; }

size_t DimensionVector::GetSize(const ModifierDimension dimension) const
{
    switch (dimension)
    {
    case ModifierDimension::X: return m_SizeX;
    case ModifierDimension::Y: return m_SizeY;
    default: return m_SizeZ;
    }
}

size_t DimensionVector::GetTotalSize() const { return m_SizeX * m_SizeY * m_SizeZ; }
std::vector<size_t> DimensionVector::GetVector() const { return {m_SizeX, m_SizeY, m_SizeZ}; }

std::string DimensionVector::GetString() const
{
    return "(" + std::to_string(m_SizeX) + ", " + std::to_string(m_SizeY) + ", " + std::to_string(m_SizeZ) + ")";
}

bool DimensionVector::operator==(const DimensionVector& other) const
{
    return m_SizeX == other.m_SizeX && m_SizeY == other.m_SizeY && m_SizeZ == other.m_SizeZ;
}

bool DimensionVector::operator!=(const DimensionVector& other) const { return !(*this == other); }

// ---------------------------------------------------------------- ParameterPair

ParameterPair::ParameterPair() :
    m_Value(static_cast<uint64_t>(0))
{}

ParameterPair::ParameterPair(const std::string& name, const ParameterValue& value) :
    m_Name(name),
    m_Value(value)
{}

void ParameterPair::SetValue(const ParameterValue& value) { m_Value = value; }
const std::string& ParameterPair::GetName() const { return m_Name; }
std::string ParameterPair::GetString() const { return m_Name + " " + GetValueString(); }

std::string ParameterPair::GetValueString() const
{
    switch (GetValueType())
    {
    case ParameterValueType::Int: return std::to_string(std::get<int64_t>(m_Value));
    case ParameterValueType::UnsignedInt: return std::to_string(std::get<uint64_t>(m_Value));
    case ParameterValueType::Double: return std::to_string(std::get<double>(m_Value));
    case ParameterValueType::Bool: return std::get<bool>(m_Value) ? "true" : "false";
    case ParameterValueType::String: return std::get<std::string>(m_Value);
    default: return "";
    }
}

const ParameterValue& ParameterPair::GetValue() const { return m_Value; }

uint64_t ParameterPair::GetValueUint() const
{
    if (std::holds_alternative<uint64_t>(m_Value))
    {
        return std::get<uint64_t>(m_Value);
    }
    if (std::holds_alternative<int64_t>(m_Value))
    {
        return static_cast<uint64_t>(std::get<int64_t>(m_Value));
    }
    if (std::holds_alternative<bool>(m_Value))
    {
        return std::get<bool>(m_Value) ? 1 : 0;
    }
    throw KttException("Parameter with name " + m_Name + " does not have an integer value");
}

ParameterValueType ParameterPair::GetValueType() const { return GetTypeFromValue(m_Value); }
bool ParameterPair::HasValueDouble() const { return std::holds_alternative<double>(m_Value); }

bool ParameterPair::HasSameValue(const ParameterPair& other) const
{
    if (GetValueType() != other.GetValueType())
    {
        return false;
    }
    switch (GetValueType())
    {
    case ParameterValueType::Int: return std::get<int64_t>(m_Value) == std::get<int64_t>(other.m_Value);
    case ParameterValueType::UnsignedInt: return std::get<uint64_t>(m_Value) == std::get<uint64_t>(other.m_Value);
    case ParameterValueType::Double: return std::fabs(std::get<double>(m_Value) - std::get<double>(other.m_Value)) < 1e-12;
    case ParameterValueType::Bool: return std::get<bool>(m_Value) == std::get<bool>(other.m_Value);
    default: return GetValueString() == other.GetValueString();
    }
}

ParameterValueType ParameterPair::GetTypeFromValue(const ParameterValue& value)
{
    if (std::holds_alternative<int64_t>(value)) return ParameterValueType::Int;
    if (std::holds_alternative<uint64_t>(value)) return ParameterValueType::UnsignedInt;
    if (std::holds_alternative<double>(value)) return ParameterValueType::Double;
    if (std::holds_alternative<bool>(value)) return ParameterValueType::Bool;
    return ParameterValueType::String;
}

bool ParameterPair::operator==(const ParameterPair& other) const { return m_Name == other.m_Name && HasSameValue(other); }
bool ParameterPair::operator!=(const ParameterPair& other) const { return !(*this == other); }

// ---------------------------------------------------------------- KernelConfiguration

KernelConfiguration::KernelConfiguration() {}
KernelConfiguration::KernelConfiguration(const std::vector<ParameterPair>& pairs) : m_Pairs(pairs) {}
const std::vector<ParameterPair>& KernelConfiguration::GetPairs() const { return m_Pairs; }
bool KernelConfiguration::IsValid() const { return !m_Pairs.empty(); }

std::string KernelConfiguration::GeneratePrefix() const
{
    std::string result;
    for (const auto& pair : m_Pairs)
    {
        result += "#define " + pair.GetString() + "\n";
    }
    return result;
}

std::vector<std::string> KernelConfiguration::GenerateDefines() const
{
    std::vector<std::string> result;
    for (const auto& pair : m_Pairs)
    {
        result.push_back("-D" + pair.GetName() + "=" + pair.GetValueString());
    }
    return result;
}

std::string KernelConfiguration::GetString() const
{
    if (m_Pairs.empty())
    {
        return "empty";
    }
    std::string result;
    for (size_t i = 0; i < m_Pairs.size(); ++i)
    {
        result += m_Pairs[i].GetString();
        if (i + 1 != m_Pairs.size())
        {
            result += ", ";
        }
    }
    return result;
}

void KernelConfiguration::Merge(const KernelConfiguration& other)
{
    for (const auto& otherPair : other.GetPairs())
    {
        const bool present = std::any_of(m_Pairs.cbegin(), m_Pairs.cend(),
            [&otherPair](const auto& pair) { return pair.GetName() == otherPair.GetName(); });
        if (!present)
        {
            m_Pairs.push_back(otherPair);
        }
    }
}

std::vector<KernelConfiguration> KernelConfiguration::GenerateNeighbours(const std::string& parameter,
    const std::vector<ParameterPair>& pairs) const
{
    size_t index = m_Pairs.size();
    for (size_t i = 0; i < m_Pairs.size(); ++i)
    {
        if (m_Pairs[i].GetName() == parameter)
        {
            index = i;
            break;
        }
    }
    if (index == m_Pairs.size())
    {
        throw KttException("Invalid parameter on input: " + parameter);
    }
    std::vector<KernelConfiguration> result;
    for (const auto& pair : pairs)
    {
        if (pair.HasSameValue(m_Pairs[index]))
        {
            continue;
        }
        auto newPairs = m_Pairs;
        newPairs[index] = pair;
        result.emplace_back(newPairs);
    }
    return result;
}

bool KernelConfiguration::operator==(const KernelConfiguration& other) const
{
    if (m_Pairs.size() != other.m_Pairs.size())
    {
        return false;
    }
    for (const auto& pair : m_Pairs)
    {
        const auto iterator = std::find_if(other.m_Pairs.cbegin(), other.m_Pairs.cend(),
            [&pair](const auto& otherPair) { return otherPair.GetName() == pair.GetName(); });
        if (iterator == other.m_Pairs.cend() || !pair.HasSameValue(*iterator))
        {
            return false;
        }
    }
    return true;
}

bool KernelConfiguration::operator!=(const KernelConfiguration& other) const { return !(*this == other); }

// ---------------------------------------------------------------- profiling data

KernelProfilingCounter::KernelProfilingCounter() :
    m_Type(ProfilingCounterType::Int), m_ValueInt(0), m_ValueUint(0), m_ValueDouble(0.0)
{}

KernelProfilingCounter::KernelProfilingCounter(const std::string& name, const ProfilingCounterType type, const int64_t value) :
    m_Name(name), m_Type(type), m_ValueInt(value), m_ValueUint(static_cast<uint64_t>(value)), m_ValueDouble(static_cast<double>(value))
{}

KernelProfilingCounter::KernelProfilingCounter(const std::string& name, const ProfilingCounterType type, const uint64_t value) :
    m_Name(name), m_Type(type), m_ValueInt(static_cast<int64_t>(value)), m_ValueUint(value), m_ValueDouble(static_cast<double>(value))
{}

KernelProfilingCounter::KernelProfilingCounter(const std::string& name, const ProfilingCounterType type, const double value) :
    m_Name(name), m_Type(type), m_ValueInt(static_cast<int64_t>(value)), m_ValueUint(static_cast<uint64_t>(value)), m_ValueDouble(value)
{}

const std::string& KernelProfilingCounter::GetName() const { return m_Name; }
ProfilingCounterType KernelProfilingCounter::GetType() const { return m_Type; }
int64_t KernelProfilingCounter::GetValueInt() const { return m_ValueInt; }
uint64_t KernelProfilingCounter::GetValueUint() const { return m_ValueUint; }
double KernelProfilingCounter::GetValueDouble() const { return m_ValueDouble; }

KernelProfilingData::KernelProfilingData() : m_RemainingRuns(0) {}
KernelProfilingData::KernelProfilingData(const uint64_t remainingRuns) : m_RemainingRuns(remainingRuns) {}
KernelProfilingData::KernelProfilingData(const std::vector<KernelProfilingCounter>& counters) : m_Counters(counters), m_RemainingRuns(0) {}

void KernelProfilingData::AddCounter(const KernelProfilingCounter& counter) { m_Counters.push_back(counter); }

void KernelProfilingData::SetCounters(const std::vector<KernelProfilingCounter>& counters)
{
    m_Counters = counters;
    m_RemainingRuns = 0;
}

void KernelProfilingData::DecreaseRemainingProfilingRuns()
{
    if (m_RemainingRuns > 0)
    {
        --m_RemainingRuns;
    }
}

bool KernelProfilingData::IsValid() const { return m_RemainingRuns == 0; }

bool KernelProfilingData::HasCounter(const std::string& name) const
{
    return std::any_of(m_Counters.cbegin(), m_Counters.cend(), [&name](const auto& counter) { return counter.GetName() == name; });
}

const KernelProfilingCounter& KernelProfilingData::GetCounter(const std::string& name) const
{
    for (const auto& counter : m_Counters)
    {
        if (counter.GetName() == name)
        {
            return counter;
        }
    }
    throw KttException("Profiling counter with name " + name + " was not found");
}

const std::vector<KernelProfilingCounter>& KernelProfilingData::GetCounters() const { return m_Counters; }
uint64_t KernelProfilingData::GetRemainingProfilingRuns() const { return m_RemainingRuns; }
bool KernelProfilingData::HasRemainingProfilingRuns() const { return m_RemainingRuns > 0; }

// ---------------------------------------------------------------- ComputationResult

ComputationResult::ComputationResult() : ComputationResult("") {}

ComputationResult::ComputationResult(const std::string& kernelFunction) :
    m_KernelFunction(kernelFunction), m_Duration(InvalidDuration), m_Overhead(0), m_CompilationOverhead(0), m_PowerUsage(0)
{}

ComputationResult::ComputationResult(const ComputationResult& other) :
    m_KernelFunction(other.m_KernelFunction),
    m_GlobalSize(other.m_GlobalSize),
    m_LocalSize(other.m_LocalSize),
    m_Duration(other.m_Duration),
    m_Overhead(other.m_Overhead),
    m_CompilationOverhead(other.m_CompilationOverhead),
    m_PowerUsage(other.m_PowerUsage)
{
    if (other.m_CompilationData)
    {
        m_CompilationData = std::make_unique<KernelCompilationData>(*other.m_CompilationData);
    }
    if (other.m_ProfilingData)
    {
        m_ProfilingData = std::make_unique<KernelProfilingData>(*other.m_ProfilingData);
    }
}

ComputationResult& ComputationResult::operator=(const ComputationResult& other)
{
    if (this != &other)
    {
        m_KernelFunction = other.m_KernelFunction;
        m_GlobalSize = other.m_GlobalSize;
        m_LocalSize = other.m_LocalSize;
        m_Duration = other.m_Duration;
        m_Overhead = other.m_Overhead;
        m_CompilationOverhead = other.m_CompilationOverhead;
        m_PowerUsage = other.m_PowerUsage;
        m_CompilationData = other.m_CompilationData ? std::make_unique<KernelCompilationData>(*other.m_CompilationData) : nullptr;
        m_ProfilingData = other.m_ProfilingData ? std::make_unique<KernelProfilingData>(*other.m_ProfilingData) : nullptr;
    }
    return *this;
}

void ComputationResult::SetDurationData(const Nanoseconds duration, const Nanoseconds overhead, const Nanoseconds compilationOverhead)
{
    m_Duration = duration;
    m_Overhead = overhead;
    m_CompilationOverhead = compilationOverhead;
}

void ComputationResult::SetSizeData(const DimensionVector& globalSize, const DimensionVector& localSize)
{
    m_GlobalSize = globalSize;
    m_LocalSize = localSize;
}

void ComputationResult::SetCompilationData(std::unique_ptr<KernelCompilationData> data) { m_CompilationData = std::move(data); }
void ComputationResult::SetProfilingData(std::unique_ptr<KernelProfilingData> data) { m_ProfilingData = std::move(data); }
void ComputationResult::SetPowerUsage(const uint32_t powerUsage) { m_PowerUsage = powerUsage; }
const std::string& ComputationResult::GetKernelFunction() const { return m_KernelFunction; }
const DimensionVector& ComputationResult::GetGlobalSize() const { return m_GlobalSize; }
const DimensionVector& ComputationResult::GetLocalSize() const { return m_LocalSize; }
Nanoseconds ComputationResult::GetDuration() const { return m_Duration; }
Nanoseconds ComputationResult::GetOverhead() const { return m_Overhead; }
Nanoseconds ComputationResult::GetCompilationOverhead() const { return m_CompilationOverhead; }
bool ComputationResult::HasCompilationData() const { return m_CompilationData != nullptr; }

const KernelCompilationData& ComputationResult::GetCompilationData() const
{
    if (!m_CompilationData)
    {
        throw KttException("Computation result has no compilation data");
    }
    return *m_CompilationData;
}

bool ComputationResult::HasProfilingData() const { return m_ProfilingData != nullptr; }

const KernelProfilingData& ComputationResult::GetProfilingData() const
{
    if (!m_ProfilingData)
    {
        throw KttException("Computation result has no profiling data");
    }
    return *m_ProfilingData;
}

bool ComputationResult::HasRemainingProfilingRuns() const { return m_ProfilingData && m_ProfilingData->HasRemainingProfilingRuns(); }
bool ComputationResult::HasPowerData() const { return m_PowerUsage != 0; }
uint32_t ComputationResult::GetPowerUsage() const { return m_PowerUsage; }

double ComputationResult::GetEnergyConsumption() const
{
    return static_cast<double>(m_PowerUsage) * static_cast<double>(m_Duration) / 1.0e9;
}

// ---------------------------------------------------------------- KernelResult

KernelResult::KernelResult() : KernelResult("", KernelConfiguration()) {}

KernelResult::KernelResult(const std::string& kernelName, const KernelConfiguration& configuration) :
    KernelResult(kernelName, configuration, {})
{
    m_Status = ResultStatus::ComputationFailed;
}

KernelResult::KernelResult(const std::string& kernelName, const KernelConfiguration& configuration,
    const std::vector<ComputationResult>& results) :
    m_Results(results),
    m_Configuration(configuration),
    m_KernelName(kernelName),
    m_ExtraDuration(0),
    m_DataMovementOverhead(0),
    m_ValidationOverhead(0),
    m_SearcherOverhead(0),
    m_FailedKernelOverhead(0),
    m_ProfilingRunsOverhead(0),
    m_ProfilingOverhead(0),
    m_CompilationOverhead(0),
    m_Status(ResultStatus::Ok)
{}

void KernelResult::SetStatus(const ResultStatus status) { m_Status = status; }
void KernelResult::SetExtraDuration(const Nanoseconds duration) { m_ExtraDuration = duration; }
void KernelResult::SetExtraOverhead(const Nanoseconds overhead) { m_DataMovementOverhead = overhead; }
void KernelResult::SetDataMovementOverhead(const Nanoseconds overhead) { m_DataMovementOverhead = overhead; }
void KernelResult::SetValidationOverhead(const Nanoseconds overhead) { m_ValidationOverhead = overhead; }
void KernelResult::SetSearcherOverhead(const Nanoseconds overhead) { m_SearcherOverhead = overhead; }
void KernelResult::SetFailedKernelOverhead(const Nanoseconds overhead) { m_FailedKernelOverhead = overhead; }
void KernelResult::SetProfilingRunsOverhead(const Nanoseconds overhead) { m_ProfilingRunsOverhead = overhead; }
void KernelResult::SetCompilationOverhead(const Nanoseconds overhead) { m_CompilationOverhead = overhead; }
const std::string& KernelResult::GetKernelName() const { return m_KernelName; }
const std::vector<ComputationResult>& KernelResult::GetResults() const { return m_Results; }
const KernelConfiguration& KernelResult::GetConfiguration() const { return m_Configuration; }
ResultStatus KernelResult::GetStatus() const { return m_Status; }

Nanoseconds KernelResult::GetKernelDuration() const
{
    Nanoseconds duration = 0;
    for (const auto& result : m_Results)
    {
        if (result.GetDuration() != InvalidDuration)
        {
            duration += result.GetDuration();
        }
    }
    return duration;
}

Nanoseconds KernelResult::GetKernelOverhead() const
{
    Nanoseconds overhead = 0;
    for (const auto& result : m_Results)
    {
        overhead += result.GetOverhead();
    }
    return overhead;
}

Nanoseconds KernelResult::GetKernelCompilationOverhead() const
{
    Nanoseconds overhead = 0;
    for (const auto& result : m_Results)
    {
        overhead += result.GetCompilationOverhead();
    }
    return overhead + m_CompilationOverhead;
}

Nanoseconds KernelResult::GetExtraDuration() const { return m_ExtraDuration; }
Nanoseconds KernelResult::GetExtraOverhead() const { return m_DataMovementOverhead; }
Nanoseconds KernelResult::GetDataMovementOverhead() const { return m_DataMovementOverhead; }
Nanoseconds KernelResult::GetValidationOverhead() const { return m_ValidationOverhead; }
Nanoseconds KernelResult::GetSearcherOverhead() const { return m_SearcherOverhead; }
Nanoseconds KernelResult::GetFailedKernelOverhead() const { return m_FailedKernelOverhead; }
Nanoseconds KernelResult::GetProfilingRunsOverhead() const { return m_ProfilingRunsOverhead; }
Nanoseconds KernelResult::GetProfilingOverhead() const { return m_ProfilingOverhead; }
Nanoseconds KernelResult::GetProfilingTotalOverhead() const { return m_ProfilingOverhead + m_ProfilingRunsOverhead; }
Nanoseconds KernelResult::GetCompilationOverhead() const { return GetKernelCompilationOverhead(); }

Nanoseconds KernelResult::GetTotalDuration() const
{
    return m_ExtraDuration + GetKernelDuration();
}

Nanoseconds KernelResult::GetTotalOverhead() const
{
    Nanoseconds overhead = m_DataMovementOverhead + m_ValidationOverhead + m_SearcherOverhead + m_FailedKernelOverhead
        + m_ProfilingRunsOverhead + m_CompilationOverhead;
    if (m_ProfilingRunsOverhead == 0)
    {
        overhead += GetKernelOverhead();
    }
    return overhead;
}

bool KernelResult::IsValid() const { return m_Status == ResultStatus::Ok; }

bool KernelResult::HasRemainingProfilingRuns() const
{
    return std::any_of(m_Results.cbegin(), m_Results.cend(), [](const auto& result) { return result.HasRemainingProfilingRuns(); });
}

void KernelResult::FuseProfilingTimes(const KernelResult& previousResult, bool first)
{
    if (!first)
    {
        m_ProfilingRunsOverhead += previousResult.GetKernelDuration() + previousResult.GetKernelOverhead()
            + previousResult.GetProfilingRunsOverhead();
        m_ProfilingOverhead += previousResult.GetDataMovementOverhead() + previousResult.GetExtraDuration()
            + previousResult.GetProfilingOverhead();
    }
    m_CompilationOverhead += previousResult.GetCompilationOverhead();
    m_ExtraDuration += previousResult.GetExtraDuration();
    m_DataMovementOverhead += previousResult.GetDataMovementOverhead();
    m_ValidationOverhead += previousResult.GetValidationOverhead();
    m_SearcherOverhead += previousResult.GetSearcherOverhead();
    m_FailedKernelOverhead += previousResult.GetFailedKernelOverhead();
}

void KernelResult::CopyProfilingTimes(const KernelResult& originalResult)
{
    m_ProfilingRunsOverhead = originalResult.GetProfilingRunsOverhead();
    m_ProfilingOverhead = originalResult.GetProfilingOverhead();
    m_CompilationOverhead = originalResult.GetCompilationOverhead();
    m_ExtraDuration = originalResult.GetExtraDuration();
    m_DataMovementOverhead = originalResult.GetDataMovementOverhead();
    m_ValidationOverhead = originalResult.GetValidationOverhead();
    m_SearcherOverhead = originalResult.GetSearcherOverhead();
    m_FailedKernelOverhead = originalResult.GetFailedKernelOverhead();
}

void KernelResult::TransferPowerData(const KernelResult& previousResult)
{
    const auto& previous = previousResult.GetResults();
    for (size_t i = m_Results.size(); i < previous.size(); ++i)
    {
        m_Results.push_back(previous[i]);
    }
    for (size_t i = 0; i < previous.size(); ++i)
    {
        if (previous[i].HasPowerData())
        {
            m_Results[i].SetPowerUsage(previous[i].GetPowerUsage());
        }
    }
}

// ---------------------------------------------------------------- logger, time, enums

Logger::Logger() :
    m_Level(LoggingLevel::Info),
    m_Stream(&std::clog)
{
    if (const char* level = std::getenv("KTT_LOG_LEVEL"))
    {
        const std::string value = level;
        if (value == "Off") m_Level = LoggingLevel::Off;
        else if (value == "Error") m_Level = LoggingLevel::Error;
        else if (value == "Warning") m_Level = LoggingLevel::Warning;
        else if (value == "Debug") m_Level = LoggingLevel::Debug;
    }
}

Logger& Logger::GetLogger()
{
    static Logger instance;
    return instance;
}

void Logger::SetLoggingLevel(const LoggingLevel level) { m_Level = level; }

void Logger::SetLoggingTarget(std::ostream& target)
{
    std::lock_guard<std::mutex> lock(m_Mutex);
    m_File.reset();
    m_Stream = &target;
}

void Logger::SetLoggingTarget(const std::string& filePath)
{
    std::lock_guard<std::mutex> lock(m_Mutex);
    m_File = std::make_unique<std::ofstream>(filePath, std::ios::app);
    m_Stream = m_File.get();
}

LoggingLevel Logger::GetLoggingLevel() const { return m_Level; }

void Logger::Log(const LoggingLevel level, const std::string& message) const
{
    if (level == LoggingLevel::Off || static_cast<int>(level) > static_cast<int>(m_Level))
    {
        return;
    }
    static const char* tags[] = {"", "[Error] ", "[Warning] ", "[Info] ", "[Debug] "};
    std::lock_guard<std::mutex> lock(m_Mutex);
    *m_Stream << tags[static_cast<int>(level)] << message << std::endl;
}

void Logger::LogError(const std::string& message) { GetLogger().Log(LoggingLevel::Error, message); }
void Logger::LogWarning(const std::string& message) { GetLogger().Log(LoggingLevel::Warning, message); }
void Logger::LogInfo(const std::string& message) { GetLogger().Log(LoggingLevel::Info, message); }
void Logger::LogDebug(const std::string& message) { GetLogger().Log(LoggingLevel::Debug, message); }

TimeConfiguration::TimeConfiguration() : m_TimeUnit(TimeUnit::Milliseconds) {}

TimeConfiguration& TimeConfiguration::GetInstance()
{
    static TimeConfiguration instance;
    return instance;
}

void TimeConfiguration::SetTimeUnit(const TimeUnit unit) { m_TimeUnit = unit; }
TimeUnit TimeConfiguration::GetTimeUnit() const { return m_TimeUnit; }

static double UnitDivisor(const TimeUnit unit)
{
    switch (unit)
    {
    case TimeUnit::Nanoseconds: return 1.0;
    case TimeUnit::Microseconds: return 1.0e3;
    case TimeUnit::Milliseconds: return 1.0e6;
    default: return 1.0e9;
    }
}

uint64_t TimeConfiguration::ConvertFromNanoseconds(const Nanoseconds duration) const
{
    return static_cast<uint64_t>(static_cast<double>(duration) / UnitDivisor(m_TimeUnit));
}

double TimeConfiguration::ConvertFromNanosecondsDouble(const Nanoseconds duration) const
{
    return static_cast<double>(duration) / UnitDivisor(m_TimeUnit);
}

Nanoseconds TimeConfiguration::ConvertToNanosecondsDouble(const double duration) const
{
    return static_cast<Nanoseconds>(std::llround(duration * UnitDivisor(m_TimeUnit)));
}

std::string TimeConfiguration::GetUnitTag() const
{
    switch (m_TimeUnit)
    {
    case TimeUnit::Nanoseconds: return "ns";
    case TimeUnit::Microseconds: return "us";
    case TimeUnit::Milliseconds: return "ms";
    default: return "s";
    }
}

uint64_t DefaultSearcherSeed()
{
    if (const char* seed = std::getenv("KTT_SEED"))
    {
        return std::strtoull(seed, nullptr, 10);
    }
    std::random_device device;
    return (static_cast<uint64_t>(device()) << 32) ^ device();
}

template <typename E>
struct EnumNames
{
    E value;
    const char* name;
};

template <typename E, size_t N>
static std::string EnumToString(const EnumNames<E> (&names)[N], const E value)
{
    for (const auto& entry : names)
    {
        if (entry.value == value)
        {
            return entry.name;
        }
    }
    throw KttException("Unhandled enumeration value");
}

template <typename E, size_t N>
static E EnumFromString(const EnumNames<E> (&names)[N], const std::string& value)
{
    for (const auto& entry : names)
    {
        if (value == entry.name)
        {
            return entry.value;
        }
    }
    throw KttException("Unknown enumeration value: " + value);
}

static const EnumNames<ComputeApi> kComputeApiNames[] = {{ComputeApi::OpenCL, "OpenCL"}, {ComputeApi::CUDA, "CUDA"},
    {ComputeApi::Vulkan, "Vulkan"}, {ComputeApi::HIP, "HIP"}};
static const EnumNames<GlobalSizeType> kGlobalSizeTypeNames[] = {{GlobalSizeType::OpenCL, "OpenCL"}, {GlobalSizeType::CUDA, "CUDA"},
    {GlobalSizeType::Vulkan, "Vulkan"}};
static const EnumNames<TimeUnit> kTimeUnitNames[] = {{TimeUnit::Nanoseconds, "Nanoseconds"}, {TimeUnit::Microseconds, "Microseconds"},
    {TimeUnit::Milliseconds, "Milliseconds"}, {TimeUnit::Seconds, "Seconds"}};
static const EnumNames<ResultStatus> kResultStatusNames[] = {{ResultStatus::Ok, "Ok"}, {ResultStatus::ComputationFailed, "ComputationFailed"},
    {ResultStatus::ValidationFailed, "ValidationFailed"}, {ResultStatus::CompilationFailed, "CompilationFailed"},
    {ResultStatus::DeviceLimitsExceeded, "DeviceLimitsExceeded"}};
static const EnumNames<ParameterValueType> kParameterValueTypeNames[] = {{ParameterValueType::Int, "Int"},
    {ParameterValueType::UnsignedInt, "UnsignedInt"}, {ParameterValueType::Double, "Double"}, {ParameterValueType::Float, "Float"},
    {ParameterValueType::Bool, "Bool"}, {ParameterValueType::String, "String"}};
static const EnumNames<ProfilingCounterType> kProfilingCounterTypeNames[] = {{ProfilingCounterType::Int, "Int"},
    {ProfilingCounterType::UnsignedInt, "UnsignedInt"}, {ProfilingCounterType::Double, "Double"}, {ProfilingCounterType::Percent, "Percent"},
    {ProfilingCounterType::Throughput, "Throughput"}, {ProfilingCounterType::UtilizationLevel, "UtilizationLevel"}};

std::string ComputeApiToString(const ComputeApi api) { return EnumToString(kComputeApiNames, api); }
std::string GlobalSizeTypeToString(const GlobalSizeType type) { return EnumToString(kGlobalSizeTypeNames, type); }
std::string TimeUnitToString(const TimeUnit unit) { return EnumToString(kTimeUnitNames, unit); }
std::string ResultStatusToString(const ResultStatus status) { return EnumToString(kResultStatusNames, status); }
std::string ParameterValueTypeToString(const ParameterValueType type) { return EnumToString(kParameterValueTypeNames, type); }
std::string ProfilingCounterTypeToString(const ProfilingCounterType type) { return EnumToString(kProfilingCounterTypeNames, type); }
ComputeApi ComputeApiFromString(const std::string& value) { return EnumFromString(kComputeApiNames, value); }
GlobalSizeType GlobalSizeTypeFromString(const std::string& value) { return EnumFromString(kGlobalSizeTypeNames, value); }
TimeUnit TimeUnitFromString(const std::string& value) { return EnumFromString(kTimeUnitNames, value); }
ResultStatus ResultStatusFromString(const std::string& value) { return EnumFromString(kResultStatusNames, value); }
ParameterValueType ParameterValueTypeFromString(const std::string& value) { return EnumFromString(kParameterValueTypeNames, value); }
ProfilingCounterType ProfilingCounterTypeFromString(const std::string& value) { return EnumFromString(kProfilingCounterTypeNames, value); }

} // namespace ktt
