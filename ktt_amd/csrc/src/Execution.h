// Execution of one kernel configuration on the compute engine.
//
//   RunRecord      what one run accumulates: the launch data of every definition (sizes from
//                  the configuration's thread modifiers), per-run argument substitutions made
//                  by the launcher, finished computations and transfer time.
//   LaunchContext  the ComputeInterface handed to launchers; it forwards to the engine and
//                  records into the active RunRecord.
//   OutputChecker  expected outputs per argument (reference computation, reference kernel or
//                  reference argument) and the comparison rules.
//   Executor       buffer setup and cleanup, the launcher call, failure handling, output
//                  download and checking. TunerCore and the tuning loop drive it.
//
// The reference's observable behaviour is kept (Source/KernelRunner/KernelRunner.cpp:23-343,
// ComputeLayerData.cpp:124-156, ResultValidator.inl:53-108): read-only buffers stay on the
// device between runs, a launcher's own host time is ExtraDuration (part of TotalDuration),
// compiler errors and device limits map to their result statuses, and the comparison rules
// and default tolerance (1e-4) are the reference's.
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

class RunRecord
{
public:
    RunRecord(const Kernel& kernel, const KernelConfiguration& configuration, const KernelDimensions& dimensions, const KernelRunMode mode);

    const Kernel& GetKernel() const { return m_Kernel; }
    const KernelConfiguration& GetConfiguration() const { return m_Configuration; }
    KernelRunMode GetMode() const { return m_Mode; }

    KernelComputeData& Launch(const KernelDefinitionId id);
    const KernelComputeData& Launch(const KernelDefinitionId id) const;
    /** Every launch of this run that names the argument uses `replacement` instead. */
    void Substitute(const ArgumentId& id, std::unique_ptr<KernelArgument> replacement);
    void Rebind(const KernelDefinitionId id, std::vector<KernelArgument*> arguments);

    void AddComputation(const ComputationResult& computation) { m_Computations.push_back(computation); }
    void AddTransferTime(const Nanoseconds time) { m_TransferTime += time; }
    /** The run's result, given the wall time spent in the launcher. */
    KernelResult Finish(const Nanoseconds launcherTime) const;

private:
    const Kernel& m_Kernel;
    KernelConfiguration m_Configuration;
    KernelRunMode m_Mode;
    std::vector<std::pair<KernelDefinitionId, KernelComputeData>> m_Launches;
    std::vector<std::unique_ptr<KernelArgument>> m_Substitutes;
    std::vector<ComputationResult> m_Computations;
    Nanoseconds m_TransferTime;

    KernelArgument* FindSubstitute(const ArgumentId& id) const;
};

class LaunchContext : public ComputeInterface
{
public:
    LaunchContext(ComputeEngine& engine, ArgumentRegistry& arguments);

    void Begin(RunRecord& record) { m_Record = &record; }
    void End() { m_Record = nullptr; }

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

private:
    ComputeEngine& m_Engine;
    ArgumentRegistry& m_Arguments;
    RunRecord* m_Record;

    RunRecord& Record() const;
    template <typename Transfer>
    void TimedTransfer(Transfer transfer);
};

class Executor;

/** Expected contents of one output argument and how they are obtained. */
struct OutputExpectation
{
    const KernelArgument* argument = nullptr;
    size_t elements = 0;
    ValueComparator comparator;
    ReferenceComputation computation;
    const Kernel* referenceKernel = nullptr;
    KernelConfiguration referenceConfiguration;
    KernelDimensions referenceDimensions;
    const KernelArgument* referenceArgument = nullptr;
    std::vector<uint8_t> expected;
    bool ready = false;
};

class OutputChecker
{
public:
    explicit OutputChecker(Executor& executor);

    void SetMethod(const ValidationMethod method, const double tolerance);
    void SetModes(const ValidationMode modes) { m_Modes = modes; }
    void SetRange(const KernelArgument& argument, const size_t elements);
    void SetComparator(const KernelArgument& argument, ValueComparator comparator);
    void SetComputation(const KernelArgument& argument, ReferenceComputation computation);
    void SetKernel(const KernelArgument& argument, const Kernel& kernel, const KernelConfiguration& configuration,
        const KernelDimensions& dimensions);
    void SetArgument(const KernelArgument& argument, const KernelArgument& reference);
    void Forget(const ArgumentId& id) { m_Expectations.erase(id); }
    void ForgetReferenceKernel(const KernelId id);

    /** Computes the expected outputs of the kernel's checked arguments that are not known yet. */
    void Prepare(const Kernel& kernel, const KernelRunMode mode);
    void Drop(const Kernel& kernel);
    /** True when every checked output of the kernel matches (or the mode is not checked). */
    bool Check(const Kernel& kernel, const KernelRunMode mode);

private:
    Executor& m_Executor;
    std::map<ArgumentId, OutputExpectation> m_Expectations;
    ValidationMethod m_Method;
    double m_Tolerance;
    ValidationMode m_Modes;

    OutputExpectation& Expect(const KernelArgument& argument);
    bool Checked(const KernelRunMode mode) const;
    void Compute(OutputExpectation& expectation);
    bool Compare(const OutputExpectation& expectation, const std::vector<uint8_t>& actual) const;
};

class Executor
{
public:
    Executor(ComputeEngine& engine, ArgumentRegistry& arguments);

    KernelResult Execute(const Kernel& kernel, const KernelConfiguration& configuration, const KernelDimensions& dimensions,
        const KernelRunMode mode, const std::vector<BufferOutputDescriptor>& output, const bool manageBuffers = true);
    void Fetch(const std::vector<BufferOutputDescriptor>& output);

    void KeepReadOnlyBuffers(const bool flag) { m_KeepReadOnly = flag; }
    void SetProfiling(const bool flag) { m_Engine.SetProfiling(flag); }
    bool IsProfiling() const { return m_Engine.IsProfilingActive(); }
    OutputChecker& Checker() { return *m_Checker; }
    ArgumentRegistry& Arguments() { return m_Arguments; }
    void ForgetKernel(const KernelId id) { m_Checker->ForgetReferenceKernel(id); }

private:
    ComputeEngine& m_Engine;
    ArgumentRegistry& m_Arguments;
    std::unique_ptr<LaunchContext> m_Context;
    std::unique_ptr<OutputChecker> m_Checker;
    bool m_KeepReadOnly;

    void StageBuffers(const Kernel& kernel);
    void ReleaseBuffers(const Kernel& kernel);
    KernelLauncher LauncherFor(const Kernel& kernel) const;
    KernelResult Launch(const Kernel& kernel, const KernelConfiguration& configuration, const KernelDimensions& dimensions,
        const KernelRunMode mode, const std::vector<BufferOutputDescriptor>& output);
};

} // namespace ktt
