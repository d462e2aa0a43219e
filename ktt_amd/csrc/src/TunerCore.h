// Owner of the argument and kernel registries, the compute engine, the executor and the tuning
// loop. The engine factory accepts ComputeApi::HIP only (the reference dispatches OpenCL, CUDA
// and Vulkan in Source/TunerCore.cpp:485-550).
#pragma once

#include <memory>
#include <ostream>
#include <string>
#include <vector>

#include <ComputeInterface.h>
#include <Info.h>
#include <KernelResult.h>
#include <KttTypes.h>
#include <Searcher.h>
#include <StopCondition.h>

#include "ComputeEngine.h"
#include "Json.h"
#include "Kernel.h"
#include "KernelArgument.h"
#include "Execution.h"
#include "Exploration.h"

namespace ktt
{

class TunerCore
{
public:
    explicit TunerCore(const PlatformIndex platform, const DeviceIndex device, const ComputeApi api, const uint32_t queueCount);
    explicit TunerCore(const ComputeApi api, const ComputeApiInitializer& initializer, std::vector<QueueId>& assignedQueueIds);
    ~TunerCore();

    ArgumentRegistry& GetArguments() { return *m_Arguments; }
    KernelRegistry& GetKernels() { return *m_Kernels; }
    const KernelRegistry& GetKernels() const { return *m_Kernels; }
    ComputeEngine& GetComputeEngine() { return *m_ComputeEngine; }
    const ComputeEngine& GetComputeEngine() const { return *m_ComputeEngine; }
    Executor& GetExecutor() { return *m_Executor; }
    TuningLoop& GetTuningLoop() { return *m_TuningLoop; }
    const TuningLoop& GetTuningLoop() const { return *m_TuningLoop; }

    void RemoveKernel(const KernelId id);
    void RemoveArgument(const ArgumentId& id);
    std::string GetKernelSource(const KernelId id, const KernelConfiguration& configuration) const;
    std::string GetKernelDefinitionSource(const KernelDefinitionId id, const KernelConfiguration& configuration) const;

    void SaveResults(const std::vector<KernelResult>& results, const std::string& filePath, const OutputFormat format,
        const UserData& data) const;
    std::vector<KernelResult> LoadResults(const std::string& filePath, const OutputFormat format, UserData& data) const;

    JsonValue CreateMetadata() const;
    static void Log(const LoggingLevel level, const std::string& message);

private:
    std::unique_ptr<ArgumentRegistry> m_Arguments;
    std::unique_ptr<KernelRegistry> m_Kernels;
    std::unique_ptr<ComputeEngine> m_ComputeEngine;
    std::unique_ptr<Executor> m_Executor;
    std::unique_ptr<TuningLoop> m_TuningLoop;

    void CreateDrivers();
};

/** KTT result JSON (Source/Output/JsonConverters.cpp:8-351 schema: Metadata, Results, UserData). */
JsonValue ResultToJson(const KernelResult& result);
KernelResult ResultFromJson(const JsonValue& value);
std::string SerializeResults(const JsonValue& metadata, const std::vector<KernelResult>& results, const UserData& data);
std::vector<KernelResult> DeserializeResults(const std::string& text, UserData& data, JsonValue* metadata = nullptr);

} // namespace ktt
