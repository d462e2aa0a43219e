// Searchers and stop conditions.
//
// Behaviour kept from the reference (Source/Api/Searcher, Source/Api/StopCondition), since
// users choose searchers and stop conditions by it:
//   Deterministic  every configuration in index order, then done.
//   Random         a uniformly drawn unexplored configuration at every step.
//   MCMC           BootIterations random states whose fastest becomes the chain's state;
//                  then proposals among the unexplored configurations that differ from the
//                  chain's state in at most MaximumDifferences parameters; a proposal that is
//                  not slower replaces the state, a slower one with probability
//                  EscapeProbability; with no unexplored neighbour left the chain restarts
//                  from a random unexplored configuration.
//   Stop conditions: a configuration count (capped by the space), the best valid duration
//   reaching a target in milliseconds, an explored fraction, tuning wall time in seconds,
//   and any-of.
#include <algorithm>
#include <limits>
#include <sstream>

#include <KttException.h>
#include <Searcher.h>
#include <StopCondition.h>

#include "Common.h"
#include "Exploration.h"

namespace ktt
{

namespace
{

constexpr double kUnmeasured = std::numeric_limits<double>::infinity();

std::string Format(const double value)
{
    std::ostringstream stream;
    stream << value;
    return stream.str();
}

} // namespace

// ---------------------------------------------------------------- base

Searcher::Searcher() : m_Space(nullptr) {}
void Searcher::OnInitialize() {}
void Searcher::OnReset() {}

KernelConfiguration Searcher::GetConfiguration(const uint64_t index) const { return m_Space->ConfigurationAt(index); }
uint64_t Searcher::GetIndex(const KernelConfiguration& configuration) const { return m_Space->IndexOf(configuration); }
KernelConfiguration Searcher::GetRandomConfiguration() const { return m_Space->DrawUnexplored(); }

std::vector<KernelConfiguration> Searcher::GetNeighbourConfigurations(const KernelConfiguration& configuration,
    const uint64_t maxDifferences, const size_t maxNeighbours) const
{
    return m_Space->Neighbours(configuration, maxDifferences, maxNeighbours);
}

uint64_t Searcher::GetConfigurationsCount() const { return m_Space->Size(); }

uint64_t Searcher::GetUnexploredConfigurationsCount() const
{
    return m_Space->Size() - m_Space->Explored().size();
}

const std::set<uint64_t>& Searcher::GetExploredIndices() const { return m_Space->Explored(); }
bool Searcher::IsInitialized() const { return m_Space != nullptr; }

bool Searcher::GetLaunchSize(const KernelConfiguration& configuration, double& threads) const
{
    return m_Space != nullptr && m_Space->LaunchSize(configuration, threads);
}

void Searcher::Initialize(const Exploration& space)
{
    m_Space = &space;
    OnInitialize();
}

void Searcher::Reset()
{
    OnReset();
    m_Space = nullptr;
}

// ---------------------------------------------------------------- deterministic

DeterministicSearcher::DeterministicSearcher() : Searcher(), m_Position(0) {}

void DeterministicSearcher::OnReset() { m_Position = 0; }

bool DeterministicSearcher::CalculateNextConfiguration([[maybe_unused]] const KernelResult& previousResult)
{
    return ++m_Position < GetConfigurationsCount();
}

KernelConfiguration DeterministicSearcher::GetCurrentConfiguration() const { return GetConfiguration(m_Position); }

// ---------------------------------------------------------------- random

RandomSearcher::RandomSearcher() : Searcher() {}

void RandomSearcher::OnInitialize() { m_Drawn = GetRandomConfiguration(); }

bool RandomSearcher::CalculateNextConfiguration([[maybe_unused]] const KernelResult& previousResult)
{
    if (GetUnexploredConfigurationsCount() == 0)
    {
        return false;
    }
    m_Drawn = GetRandomConfiguration();
    return true;
}

KernelConfiguration RandomSearcher::GetCurrentConfiguration() const { return m_Drawn; }

// ---------------------------------------------------------------- MCMC

McmcSearcher::McmcSearcher(const KernelConfiguration& start) : McmcSearcher(start, DefaultSearcherSeed()) {}

McmcSearcher::McmcSearcher(const KernelConfiguration& start, const uint64_t seed) :
    Searcher(),
    m_StartPoint(start),
    m_Random(seed),
    m_Proposal(0),
    m_Chain(0),
    m_ChainDuration(kUnmeasured),
    m_WarmupLeft(0)
{}

uint64_t McmcSearcher::DrawIndex(const uint64_t bound)
{
    return std::uniform_int_distribution<uint64_t>(0, bound - 1)(m_Random);
}

uint64_t McmcSearcher::DrawUnexplored()
{
    const auto& explored = GetExploredIndices();
    const uint64_t total = GetConfigurationsCount();
    uint64_t index = DrawIndex(total);
    if (explored.count(index) == 0)
    {
        return index;
    }
    // Mostly explored: pick uniformly among what is left.
    std::vector<uint64_t> left;
    left.reserve(total - explored.size());
    for (uint64_t i = 0; i < total; ++i)
    {
        if (explored.count(i) == 0)
        {
            left.push_back(i);
        }
    }
    return left[DrawIndex(left.size())];
}

void McmcSearcher::OnInitialize()
{
    m_ChainDuration = kUnmeasured;
    if (m_StartPoint.IsValid())
    {
        m_Proposal = GetIndex(m_StartPoint);
        m_WarmupLeft = 0;
    }
    else
    {
        m_Proposal = DrawIndex(GetConfigurationsCount());
        m_WarmupLeft = BootIterations;
    }
    m_Chain = m_Proposal;
}

void McmcSearcher::OnReset()
{
    m_Proposal = 0;
    m_Chain = 0;
    m_ChainDuration = kUnmeasured;
    m_WarmupLeft = 0;
}

bool McmcSearcher::CalculateNextConfiguration(const KernelResult& previousResult)
{
    const double duration = previousResult.IsValid() ? static_cast<double>(previousResult.GetTotalDuration()) : kUnmeasured;
    const bool warmingUp = m_WarmupLeft > 0;
    if (warmingUp)
    {
        --m_WarmupLeft;
    }
    // While warming up the chain keeps the fastest state seen; afterwards it also accepts a
    // slower proposal now and then, to leave a local minimum.
    const bool escape = !warmingUp && std::uniform_real_distribution<double>(0.0, 1.0)(m_Random) < EscapeProbability;
    if (duration <= m_ChainDuration || escape || m_ChainDuration == kUnmeasured)
    {
        m_Chain = m_Proposal;
        m_ChainDuration = duration;
    }

    if (GetUnexploredConfigurationsCount() == 0)
    {
        return false;
    }
    if (m_WarmupLeft > 0)
    {
        m_Proposal = DrawUnexplored();
        return true;
    }
    const auto neighbours = GetNeighbourConfigurations(GetConfiguration(m_Chain), MaximumDifferences,
        std::numeric_limits<size_t>::max());
    if (neighbours.empty())
    {
        // Nothing unexplored near the chain: restart it elsewhere.
        m_Proposal = DrawUnexplored();
        m_Chain = m_Proposal;
        m_ChainDuration = kUnmeasured;
        return true;
    }
    m_Proposal = GetIndex(neighbours[DrawIndex(neighbours.size())]);
    return true;
}

KernelConfiguration McmcSearcher::GetCurrentConfiguration() const { return GetConfiguration(m_Proposal); }

// ---------------------------------------------------------------- stop conditions

ConfigurationCount::ConfigurationCount(const uint64_t count) : m_Limit(std::max<uint64_t>(count, 1)), m_Seen(0) {}

bool ConfigurationCount::IsFulfilled() const { return m_Seen >= m_Limit; }

void ConfigurationCount::Initialize(const uint64_t configurationsCount)
{
    m_Seen = 0;
    m_Limit = std::min(m_Limit, configurationsCount);
}

void ConfigurationCount::Update([[maybe_unused]] const KernelResult& result) { ++m_Seen; }

std::string ConfigurationCount::GetStatusString() const
{
    return "Configurations tested: " + std::to_string(m_Seen) + " of " + std::to_string(m_Limit);
}

ConfigurationDuration::ConfigurationDuration(const double duration) : m_TargetMs(std::max(duration, 0.0)), m_FastestMs(kUnmeasured) {}

bool ConfigurationDuration::IsFulfilled() const { return m_FastestMs <= m_TargetMs; }

void ConfigurationDuration::Initialize([[maybe_unused]] const uint64_t configurationsCount) { m_FastestMs = kUnmeasured; }

void ConfigurationDuration::Update(const KernelResult& result)
{
    if (result.IsValid())
    {
        m_FastestMs = std::min(m_FastestMs, static_cast<double>(result.GetTotalDuration()) * 1.0e-6);
    }
}

std::string ConfigurationDuration::GetStatusString() const
{
    return "Fastest configuration so far: " + Format(m_FastestMs) + " ms (stop at " + Format(m_TargetMs) + " ms)";
}

ConfigurationFraction::ConfigurationFraction(const double fraction) : m_Fraction(std::clamp(fraction, 0.0, 1.0)), m_Seen(0), m_Space(0) {}

bool ConfigurationFraction::IsFulfilled() const
{
    return m_Space > 0 && static_cast<double>(m_Seen) >= m_Fraction * static_cast<double>(m_Space);
}

void ConfigurationFraction::Initialize(const uint64_t configurationsCount)
{
    m_Seen = 0;
    m_Space = std::max<uint64_t>(configurationsCount, 1);
}

void ConfigurationFraction::Update([[maybe_unused]] const KernelResult& result) { ++m_Seen; }

std::string ConfigurationFraction::GetStatusString() const
{
    const double explored = m_Space > 0 ? static_cast<double>(m_Seen) / static_cast<double>(m_Space) : 0.0;
    return "Explored fraction of the space: " + Format(explored) + " (stop at " + Format(m_Fraction) + ")";
}

TuningDuration::TuningDuration(const double duration) :
    m_LimitSeconds(std::max(duration, 0.0)), m_ElapsedSeconds(0.0), m_Begin(std::chrono::steady_clock::now())
{}

bool TuningDuration::IsFulfilled() const { return m_ElapsedSeconds > m_LimitSeconds; }

void TuningDuration::Initialize([[maybe_unused]] const uint64_t configurationsCount)
{
    m_Begin = std::chrono::steady_clock::now();
    m_ElapsedSeconds = 0.0;
}

void TuningDuration::Update([[maybe_unused]] const KernelResult& result)
{
    m_ElapsedSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - m_Begin).count();
}

std::string TuningDuration::GetStatusString() const
{
    return "Tuning time: " + Format(m_ElapsedSeconds) + " s (stop after " + Format(m_LimitSeconds) + " s)";
}

UnionCondition::UnionCondition(const std::vector<std::shared_ptr<StopCondition>>& conditions) : m_Parts(conditions) {}

bool UnionCondition::IsFulfilled() const
{
    for (const auto& part : m_Parts)
    {
        if (part->IsFulfilled())
        {
            return true;
        }
    }
    return false;
}

void UnionCondition::Initialize(const uint64_t configurationsCount)
{
    for (const auto& part : m_Parts)
    {
        part->Initialize(configurationsCount);
    }
}

void UnionCondition::Update(const KernelResult& result)
{
    for (const auto& part : m_Parts)
    {
        part->Update(result);
    }
}

std::string UnionCondition::GetStatusString() const
{
    std::string text;
    for (const auto& part : m_Parts)
    {
        text += (text.empty() ? "" : "; ") + part->GetStatusString();
    }
    return text;
}

} // namespace ktt
