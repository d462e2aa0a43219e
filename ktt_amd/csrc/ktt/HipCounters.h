/** @file HipCounters.h
  * Hardware performance counters for profiled runs on ComputeApi::HIP (addition: the reference
  * collects CUDA counters through CUPTI, Source/ComputeEngine/Cuda/Cupti/CuptiMetricInterface.cpp).
  *
  * Counters are collected with rocprofiler-sdk, which has to be registered before the HIP
  * runtime initialises: call EnableHipCounters() before the first HIP call of the process (before
  * the Tuner is created), or set KTT_HIP_COUNTERS=1 in the environment. Without it, profiled runs
  * report only the launch's static resource usage and Tuner::SetProfilingCounters fails.
  */
#pragma once

#include <string>

namespace ktt
{

/** Enables hardware counters for this process. Returns false, with the reason in `error`, when
 *  the counter library cannot be loaded or registered. */
bool EnableHipCounters(std::string* error = nullptr);

} // namespace ktt
