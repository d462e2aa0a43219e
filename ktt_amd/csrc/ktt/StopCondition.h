/** @file StopCondition.h
  * Tuning stop conditions (Source/Api/StopCondition/ of the reference): configuration count,
  * best configuration duration (ms), explored fraction, tuning wall time (s), and any-of union.
  */
#pragma once

#include <chrono>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include <KernelResult.h>

namespace ktt
{

class StopCondition
{
public:
    virtual ~StopCondition() = default;
    virtual bool IsFulfilled() const = 0;
    virtual void Initialize(const uint64_t configurationsCount) = 0;
    virtual void Update(const KernelResult& result) = 0;
    virtual std::string GetStatusString() const = 0;
};

class ConfigurationCount : public StopCondition
{
public:
    explicit ConfigurationCount(const uint64_t count);
    bool IsFulfilled() const override;
    void Initialize(const uint64_t configurationsCount) override;
    void Update(const KernelResult& result) override;
    std::string GetStatusString() const override;

private:
    uint64_t m_Limit;
    uint64_t m_Seen;
};

class ConfigurationDuration : public StopCondition
{
public:
    explicit ConfigurationDuration(const double duration);
    bool IsFulfilled() const override;
    void Initialize(const uint64_t configurationsCount) override;
    void Update(const KernelResult& result) override;
    std::string GetStatusString() const override;

private:
    double m_TargetMs;
    double m_FastestMs;
};

class ConfigurationFraction : public StopCondition
{
public:
    explicit ConfigurationFraction(const double fraction);
    bool IsFulfilled() const override;
    void Initialize(const uint64_t configurationsCount) override;
    void Update(const KernelResult& result) override;
    std::string GetStatusString() const override;

private:
    double m_Fraction;
    uint64_t m_Seen;
    uint64_t m_Space;
};

class TuningDuration : public StopCondition
{
public:
    explicit TuningDuration(const double duration);
    bool IsFulfilled() const override;
    void Initialize(const uint64_t configurationsCount) override;
    void Update(const KernelResult& result) override;
    std::string GetStatusString() const override;

private:
    double m_LimitSeconds;
    double m_ElapsedSeconds;
    std::chrono::steady_clock::time_point m_Begin;
};

class UnionCondition : public StopCondition
{
public:
    explicit UnionCondition(const std::vector<std::shared_ptr<StopCondition>>& conditions);
    bool IsFulfilled() const override;
    void Initialize(const uint64_t configurationsCount) override;
    void Update(const KernelResult& result) override;
    std::string GetStatusString() const override;

private:
    std::vector<std::shared_ptr<StopCondition>> m_Parts;
};

} // namespace ktt
