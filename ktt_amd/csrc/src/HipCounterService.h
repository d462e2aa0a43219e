// Hardware counters for profiled kernel runs on the HIP engine.
//
// The counters come from libktt_hip_counters.so (rocprofiler-sdk's dispatch counting service),
// which is loaded only when counters are enabled: with the environment variable
// KTT_HIP_COUNTERS=1 when this library loads, or with EnableHipCounters() before the process's
// first HIP call. rocprofiler-sdk must be registered before the HIP runtime initialises, so a
// later request cannot take effect; the engine then reports that counters are unavailable.
//
// This is the counterpart of the reference's CUPTI metric interface
// (Source/ComputeEngine/Cuda/Cupti/CuptiMetricInterface.cpp:25-657, CuptiInstance.cpp): a
// counter list is planned into passes that the hardware collects at once, each profiled run of
// a launch collects one pass, and the values of all passes form the run's profiling data.
#pragma once

#include <cstddef>
#include <string>
#include <vector>

#include <HipCounters.h>

namespace ktt
{

class HipCounterService
{
public:
    static HipCounterService& Get();

    /** Loads the counter library and registers it with rocprofiler-sdk. */
    bool Enable(std::string& error);
    /** True once the service collects counters (the HIP runtime initialised after Enable). */
    bool IsRunning() const;
    /** Why counters are unavailable, for error messages. */
    std::string UnavailableReason() const;

    /** Plans `counters` on GPU `device`: the pass of each counter and the number of passes.
     *  Throws KttException for unknown counters and for counters one pass cannot hold. */
    size_t Plan(const int device, const std::vector<std::string>& counters, std::vector<size_t>& passOf);
    /** Collects `pass` of the last plan on the next dispatch with this workgroup size. */
    void Arm(const size_t pass, const unsigned blockX, const unsigned blockY, const unsigned blockZ);
    /** Waits for the armed pass; fills the values of its counters (indexed like the plan). */
    void Collect(std::vector<double>& values, const int timeoutMs);
    void Disarm();
    /** Counters the GPU of the last plan supports. */
    std::vector<std::string> Supported();

    /** Counters collected when profiling is on and the user named none: the inputs of the
     *  profile-based searcher's bottleneck analysis on CDNA GPUs. */
    static const std::vector<std::string>& DefaultCounters();

private:
    HipCounterService() = default;

    void* m_Library = nullptr;
    std::string m_LoadError;
    int (*m_Enable)(char*, size_t) = nullptr;
    int (*m_Running)() = nullptr;
    int (*m_Plan)(int, const char* const*, size_t, size_t*, size_t*, char*, size_t) = nullptr;
    int (*m_Arm)(size_t, unsigned, unsigned, unsigned, char*, size_t) = nullptr;
    int (*m_Collect)(double*, int, char*, size_t) = nullptr;
    void (*m_Disarm)() = nullptr;
    size_t (*m_List)(char*, size_t) = nullptr;
    std::vector<std::string> m_PlannedCounters;
};

} // namespace ktt
