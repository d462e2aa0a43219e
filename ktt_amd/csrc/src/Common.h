// Internal utilities: logger, timer, time-unit conversion, container helpers
// (Source/Utility/Logger/Logger.cpp:7-109, Timer/Timer.cpp:11-46,
// Output/TimeConfiguration/TimeConfiguration.cpp of the reference).
#pragma once

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <fstream>
#include <functional>
#include <iostream>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <set>
#include <string>
#include <vector>

#include <KttException.h>
#include <KttTypes.h>

namespace ktt
{

class Logger
{
public:
    static Logger& GetLogger();

    void SetLoggingLevel(const LoggingLevel level);
    void SetLoggingTarget(std::ostream& target);
    void SetLoggingTarget(const std::string& filePath);
    LoggingLevel GetLoggingLevel() const;
    void Log(const LoggingLevel level, const std::string& message) const;

    static void LogError(const std::string& message);
    static void LogWarning(const std::string& message);
    static void LogInfo(const std::string& message);
    static void LogDebug(const std::string& message);

private:
    Logger();

    LoggingLevel m_Level;
    std::ostream* m_Stream;
    std::unique_ptr<std::ofstream> m_File;
    mutable std::mutex m_Mutex;
};

class Timer
{
public:
    void Start() { m_Start = std::chrono::steady_clock::now(); }
    void Stop() { m_End = std::chrono::steady_clock::now(); }
    Nanoseconds GetElapsedTime() const
    {
        return static_cast<Nanoseconds>(std::chrono::duration_cast<std::chrono::nanoseconds>(m_End - m_Start).count());
    }

private:
    std::chrono::steady_clock::time_point m_Start;
    std::chrono::steady_clock::time_point m_End;
};

template <typename F>
Nanoseconds RunScopeTimer(F&& function)
{
    Timer timer;
    timer.Start();
    function();
    timer.Stop();
    return timer.GetElapsedTime();
}

class TimeConfiguration
{
public:
    static TimeConfiguration& GetInstance();

    void SetTimeUnit(const TimeUnit unit);
    TimeUnit GetTimeUnit() const;
    uint64_t ConvertFromNanoseconds(const Nanoseconds duration) const;
    double ConvertFromNanosecondsDouble(const Nanoseconds duration) const;
    Nanoseconds ConvertToNanosecondsDouble(const double duration) const;
    std::string GetUnitTag() const;

private:
    TimeConfiguration();
    TimeUnit m_TimeUnit;
};

template <typename Container, typename Key>
bool ContainsKey(const Container& container, const Key& key)
{
    return container.find(key) != container.cend();
}

template <typename Container, typename Element>
bool ContainsElement(const Container& container, const Element& element)
{
    return std::find(container.cbegin(), container.cend(), element) != container.cend();
}

template <typename Container, typename Predicate>
bool ContainsElementIf(const Container& container, Predicate predicate)
{
    return std::find_if(container.cbegin(), container.cend(), predicate) != container.cend();
}

inline size_t RoundUp(const size_t number, const size_t multiple)
{
    if (multiple == 0)
    {
        return number;
    }
    return ((number + multiple - 1) / multiple) * multiple;
}

/** Process-wide seed for the searchers' random sources: KTT_SEED if set, otherwise random_device. */
uint64_t DefaultSearcherSeed();

std::string ComputeApiToString(const ComputeApi api);
std::string GlobalSizeTypeToString(const GlobalSizeType type);
std::string TimeUnitToString(const TimeUnit unit);
std::string ResultStatusToString(const ResultStatus status);
std::string ParameterValueTypeToString(const ParameterValueType type);
std::string ProfilingCounterTypeToString(const ProfilingCounterType type);
ComputeApi ComputeApiFromString(const std::string& value);
GlobalSizeType GlobalSizeTypeFromString(const std::string& value);
TimeUnit TimeUnitFromString(const std::string& value);
ResultStatus ResultStatusFromString(const std::string& value);
ParameterValueType ParameterValueTypeFromString(const std::string& value);
ProfilingCounterType ProfilingCounterTypeFromString(const std::string& value);

} // namespace ktt
