/** @file Searcher.h
  * Configuration-space searchers. The base class offers the helpers a custom searcher uses
  * (same interface as Source/Api/Searcher/Searcher.h:22-136). The three built-in searchers
  * have the reference's behaviour: Deterministic walks the space in index order, Random
  * draws unexplored configurations uniformly, and MCMC is a Metropolis-style walk over
  * configurations that differ in at most 2 parameters, after 10 random warm-up states,
  * escaping to a slower state with probability 0.02 (McmcSearcher.h:56-58). The random
  * sources take an optional seed; without one they are seeded from random_device, as the
  * reference's are (RandomIntGenerator.inl:9-11, McmcSearcher.cpp:20).
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

class Exploration;

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
    /** Threads one launch of the configuration runs in the kernel's first definition (global size
     *  times local size after the thread modifiers, in the tuner's size convention); in a replay of
     *  saved results, the sizes those results recorded. False when it is not known. This helper has
     *  no counterpart in the reference's Searcher. */
    bool GetLaunchSize(const KernelConfiguration& configuration, double& threads) const;

    void Initialize(const Exploration& space);
    void Reset();

private:
    const Exploration* m_Space;
};

class DeterministicSearcher : public Searcher
{
public:
    DeterministicSearcher();

    void OnReset() override;
    bool CalculateNextConfiguration(const KernelResult& previousResult) override;
    KernelConfiguration GetCurrentConfiguration() const override;

private:
    uint64_t m_Position;
};

class RandomSearcher : public Searcher
{
public:
    RandomSearcher();

    void OnInitialize() override;
    bool CalculateNextConfiguration(const KernelResult& previousResult) override;
    KernelConfiguration GetCurrentConfiguration() const override;

private:
    KernelConfiguration m_Drawn;
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
    KernelConfiguration m_StartPoint;
    std::mt19937_64 m_Random;
    uint64_t m_Proposal;     // configuration handed out, waiting for its result
    uint64_t m_Chain;        // state of the chain
    double m_ChainDuration;  // its measured duration in ns (infinite until measured)
    size_t m_WarmupLeft;     // random states still to draw before the walk starts

    uint64_t DrawIndex(const uint64_t bound);
    uint64_t DrawUnexplored();
};

} // namespace ktt
