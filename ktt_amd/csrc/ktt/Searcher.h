/** @file Searcher.h
  * Configuration-space searchers. The base class exposes the helpers of
  * Source/Api/Searcher/Searcher.h:22-136; Deterministic, Random and MCMC follow
  * DeterministicSearcher.cpp:16-20, RandomSearcher.cpp:10-19 and McmcSearcher.cpp:11-188
  * (MCMC constants: 2 maximum differences, 10 boot iterations, 0.02 escape probability,
  * McmcSearcher.h:56-58). The random sources are seedable here; the reference seeds
  * them from random_device / the clock (RandomIntGenerator.inl:9-11, McmcSearcher.cpp:20),
  * which remains the default.
  */
#pragma once

#include <cstdint>
#include <map>
#include <random>
#include <set>
#include <vector>

#include <KernelConfiguration.h>
#include <KernelResult.h>

namespace ktt
{

class ConfigurationData;

class Searcher
{
public:
    virtual ~Searcher() = default;

    virtual void OnInitialize();
    virtual void OnReset();
    virtual bool CalculateNextConfiguration(const KernelResult& previousResult) = 0;
    virtual KernelConfiguration GetCurrentConfiguration() const = 0;

    Searcher();

    KernelConfiguration GetConfiguration(const uint64_t index) const;
    uint64_t GetIndex(const KernelConfiguration& configuration) const;
    KernelConfiguration GetRandomConfiguration() const;
    std::vector<KernelConfiguration> GetNeighbourConfigurations(const KernelConfiguration& configuration,
        const uint64_t maxDifferences, const size_t maxNeighbours = 3) const;
    uint64_t GetConfigurationsCount() const;
    uint64_t GetUnexploredConfigurationsCount() const;
    const std::set<uint64_t>& GetExploredIndices() const;
    bool IsInitialized() const;

    void Initialize(const ConfigurationData& data);
    void Reset();

private:
    const ConfigurationData* m_Data;
};

class DeterministicSearcher : public Searcher
{
public:
    DeterministicSearcher();

    void OnReset() override;
    bool CalculateNextConfiguration(const KernelResult& previousResult) override;
    KernelConfiguration GetCurrentConfiguration() const override;

private:
    uint64_t m_Index;
};

class RandomSearcher : public Searcher
{
public:
    RandomSearcher();

    void OnInitialize() override;
    bool CalculateNextConfiguration(const KernelResult& previousResult) override;
    KernelConfiguration GetCurrentConfiguration() const override;

private:
    KernelConfiguration m_CurrentConfiguration;
};

class McmcSearcher : public Searcher
{
public:
    McmcSearcher(const KernelConfiguration& start = {});
    /** Seeded variant (addition): identical walk for identical results. */
    McmcSearcher(const KernelConfiguration& start, const uint64_t seed);

    void OnInitialize() override;
    void OnReset() override;
    bool CalculateNextConfiguration(const KernelResult& previousResult) override;
    KernelConfiguration GetCurrentConfiguration() const override;

    static constexpr size_t MaximumDifferences = 2;
    static constexpr size_t BootIterations = 10;
    static constexpr double EscapeProbability = 0.02;

private:
    uint64_t m_Index;
    size_t m_VisitedStatesCount;
    size_t m_OriginState;
    size_t m_CurrentState;
    size_t m_Boot;
    double m_BestTime;
    KernelConfiguration m_Start;
    std::map<size_t, double> m_ExecutionTimes;
    std::mt19937_64 m_Generator;
    std::uniform_int_distribution<size_t> m_IntDistribution;
    std::uniform_real_distribution<double> m_ProbabilityDistribution;
};

} // namespace ktt
