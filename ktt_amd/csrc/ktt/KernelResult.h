/** @file KernelResult.h
  * Results of kernel runs: per-launch ComputationResult and per-configuration KernelResult,
  * with the overhead breakdown of Source/Api/Output/KernelResult.cpp:108-211
  * (TotalDuration = ExtraDuration + sum of kernel durations).
  */
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include <DimensionVector.h>
#include <KernelConfiguration.h>
#include <KttTypes.h>

namespace ktt
{

/** Static resource usage of a compiled kernel (HIP: from the code object's metadata). */
struct KernelCompilationData
{
    uint64_t m_MaxWorkGroupSize = 0;
    uint64_t m_LocalMemorySize = 0;
    uint64_t m_PrivateMemorySize = 0;
    uint64_t m_ConstantMemorySize = 0;
    uint64_t m_RegistersCount = 0;
};

class KernelProfilingCounter
{
public:
    KernelProfilingCounter();
    explicit KernelProfilingCounter(const std::string& name, const ProfilingCounterType type, const int64_t value);
    explicit KernelProfilingCounter(const std::string& name, const ProfilingCounterType type, const uint64_t value);
    explicit KernelProfilingCounter(const std::string& name, const ProfilingCounterType type, const double value);

    const std::string& GetName() const;
    ProfilingCounterType GetType() const;
    int64_t GetValueInt() const;
    uint64_t GetValueUint() const;
    double GetValueDouble() const;

private:
    std::string m_Name;
    ProfilingCounterType m_Type;
    int64_t m_ValueInt;
    uint64_t m_ValueUint;
    double m_ValueDouble;
};

class KernelProfilingData
{
public:
    KernelProfilingData();
    explicit KernelProfilingData(const uint64_t remainingRuns);
    explicit KernelProfilingData(const std::vector<KernelProfilingCounter>& counters);

    void AddCounter(const KernelProfilingCounter& counter);
    void SetCounters(const std::vector<KernelProfilingCounter>& counters);
    void DecreaseRemainingProfilingRuns();

    bool IsValid() const;
    bool HasCounter(const std::string& name) const;
    const KernelProfilingCounter& GetCounter(const std::string& name) const;
    const std::vector<KernelProfilingCounter>& GetCounters() const;
    uint64_t GetRemainingProfilingRuns() const;
    bool HasRemainingProfilingRuns() const;

private:
    std::vector<KernelProfilingCounter> m_Counters;
    uint64_t m_RemainingRuns;
};

class ComputationResult
{
public:
    ComputationResult();
    explicit ComputationResult(const std::string& kernelFunction);
    ComputationResult(const ComputationResult& other);
    ComputationResult& operator=(const ComputationResult& other);

    void SetDurationData(const Nanoseconds duration, const Nanoseconds overhead, const Nanoseconds compilationOverhead = 0);
    void SetSizeData(const DimensionVector& globalSize, const DimensionVector& localSize);
    void SetCompilationData(std::unique_ptr<KernelCompilationData> data);
    void SetProfilingData(std::unique_ptr<KernelProfilingData> data);
    void SetPowerUsage(const uint32_t powerUsage);

    const std::string& GetKernelFunction() const;
    const DimensionVector& GetGlobalSize() const;
    const DimensionVector& GetLocalSize() const;
    Nanoseconds GetDuration() const;
    Nanoseconds GetOverhead() const;
    Nanoseconds GetCompilationOverhead() const;
    bool HasCompilationData() const;
    const KernelCompilationData& GetCompilationData() const;
    bool HasProfilingData() const;
    const KernelProfilingData& GetProfilingData() const;
    bool HasRemainingProfilingRuns() const;
    bool HasPowerData() const;
    uint32_t GetPowerUsage() const;
    double GetEnergyConsumption() const;

private:
    std::string m_KernelFunction;
    DimensionVector m_GlobalSize;
    DimensionVector m_LocalSize;
    Nanoseconds m_Duration;
    Nanoseconds m_Overhead;
    Nanoseconds m_CompilationOverhead;
    std::unique_ptr<KernelCompilationData> m_CompilationData;
    std::unique_ptr<KernelProfilingData> m_ProfilingData;
    uint32_t m_PowerUsage;
};

class KernelResult
{
public:
    KernelResult();
    explicit KernelResult(const std::string& kernelName, const KernelConfiguration& configuration);
    explicit KernelResult(const std::string& kernelName, const KernelConfiguration& configuration,
        const std::vector<ComputationResult>& results);

    void SetStatus(const ResultStatus status);
    void SetExtraDuration(const Nanoseconds duration);
    void SetExtraOverhead(const Nanoseconds overhead);
    void SetDataMovementOverhead(const Nanoseconds overhead);
    void SetValidationOverhead(const Nanoseconds overhead);
    void SetSearcherOverhead(const Nanoseconds overhead);
    void SetFailedKernelOverhead(const Nanoseconds overhead);
    void SetProfilingRunsOverhead(const Nanoseconds overhead);
    void SetCompilationOverhead(const Nanoseconds overhead);

    const std::string& GetKernelName() const;
    const std::vector<ComputationResult>& GetResults() const;
    const KernelConfiguration& GetConfiguration() const;
    ResultStatus GetStatus() const;
    Nanoseconds GetKernelDuration() const;
    Nanoseconds GetKernelOverhead() const;
    Nanoseconds GetKernelCompilationOverhead() const;
    Nanoseconds GetExtraDuration() const;
    Nanoseconds GetExtraOverhead() const;
    Nanoseconds GetDataMovementOverhead() const;
    Nanoseconds GetValidationOverhead() const;
    Nanoseconds GetSearcherOverhead() const;
    Nanoseconds GetFailedKernelOverhead() const;
    Nanoseconds GetProfilingRunsOverhead() const;
    Nanoseconds GetProfilingOverhead() const;
    Nanoseconds GetProfilingTotalOverhead() const;
    Nanoseconds GetCompilationOverhead() const;
    Nanoseconds GetTotalDuration() const;
    Nanoseconds GetTotalOverhead() const;
    bool IsValid() const;
    bool HasRemainingProfilingRuns() const;

    void FuseProfilingTimes(const KernelResult& previousResult, bool first);
    void CopyProfilingTimes(const KernelResult& originalResult);
    void TransferPowerData(const KernelResult& previousResult);

private:
    std::vector<ComputationResult> m_Results;
    KernelConfiguration m_Configuration;
    std::string m_KernelName;
    Nanoseconds m_ExtraDuration;
    Nanoseconds m_DataMovementOverhead;
    Nanoseconds m_ValidationOverhead;
    Nanoseconds m_SearcherOverhead;
    Nanoseconds m_FailedKernelOverhead;
    Nanoseconds m_ProfilingRunsOverhead;
    Nanoseconds m_ProfilingOverhead;
    Nanoseconds m_CompilationOverhead;
    ResultStatus m_Status;
};

/** Where to download a buffer argument after a run (Source/Api/Output/BufferOutputDescriptor.h). */
class BufferOutputDescriptor
{
public:
    explicit BufferOutputDescriptor(const ArgumentId& id, void* outputDestination, const size_t outputSize = 0) :
        m_ArgumentId(id),
        m_OutputDestination(outputDestination),
        m_OutputSize(outputSize)
    {}

    const ArgumentId& GetArgumentId() const { return m_ArgumentId; }
    void* GetOutputDestination() const { return m_OutputDestination; }
    size_t GetOutputSize() const { return m_OutputSize; }

private:
    ArgumentId m_ArgumentId;
    void* m_OutputDestination;
    size_t m_OutputSize;
};

} // namespace ktt
