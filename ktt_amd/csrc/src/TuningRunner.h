// Configuration space and tuning loops.
//
// ConfigurationData holds the valid configurations of a kernel: parameters are grouped by
// their group name (Kernel.cpp:238-256); each group's space is the cartesian product of its
// parameters' values filtered by the group's constraints (a depth-first enumeration that checks
// a constraint as soon as its last parameter is assigned). Groups are explored one after
// another, every configuration being merged with the best one found so far
// (ConfigurationData.cpp:78-105, 215-279). Indices are global over the concatenated groups.
//
// TuningRunner implements offline tuning, single tuning iterations and simulated tuning over
// saved results (TuningRunner.cpp:17-201).
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <random>
#include <set>
#include <unordered_map>
#include <vector>

#include <KernelConfiguration.h>
#include <KernelResult.h>
#include <Searcher.h>
#include <StopCondition.h>

#include "Kernel.h"
#include "KernelRunner.h"

namespace ktt
{

class ConfigurationData
{
public:
    explicit ConfigurationData(Searcher& searcher, const Kernel& kernel);
    ~ConfigurationData();

    bool CalculateNextConfiguration(const KernelResult& previousResult);
    KernelConfiguration GetConfigurationForIndex(const uint64_t index) const;
    uint64_t GetIndexForConfiguration(const KernelConfiguration& configuration) const;
    KernelConfiguration GetRandomConfiguration() const;
    std::vector<KernelConfiguration> GetNeighbourConfigurations(const KernelConfiguration& configuration,
        const uint64_t maxDifferences, const size_t maxNeighbours) const;
    uint64_t GetTotalConfigurationsCount() const;
    uint64_t GetExploredConfigurationsCount() const;
    const std::set<uint64_t>& GetExploredConfigurations() const;
    bool IsProcessed() const;
    KernelConfiguration GetCurrentConfiguration() const;
    KernelConfiguration GetBestConfiguration() const;

private:
    struct Group
    {
        std::vector<size_t> parameters;                 // indices into the kernel's parameter list
        std::vector<std::vector<uint32_t>> configurations;  // value indices per parameter
        std::unordered_map<std::string, uint64_t> lookup;   // encoded value indices -> local index
        uint64_t offset = 0;
    };

    Searcher& m_Searcher;
    const Kernel& m_Kernel;
    std::vector<Group> m_Groups;
    std::set<uint64_t> m_ExploredConfigurations;
    std::pair<KernelConfiguration, Nanoseconds> m_BestConfiguration;
    bool m_SearcherActive;
    mutable std::mt19937_64 m_Generator;

    void InitializeConfigurations();
    void EnumerateGroup(Group& group);
    void UpdateBestConfiguration(const KernelResult& previousResult);
    const Group& GetGroupForIndex(const uint64_t index, uint64_t& localIndex) const;
    KernelConfiguration BuildConfiguration(const Group& group, const std::vector<uint32_t>& values) const;
    bool DecodeConfiguration(const KernelConfiguration& configuration, size_t& groupIndex, std::vector<uint32_t>& values) const;
    static std::string Encode(const std::vector<uint32_t>& values);
};

class ConfigurationManager
{
public:
    void SetSearcher(const KernelId id, std::unique_ptr<Searcher> searcher);
    void InitializeData(const Kernel& kernel);
    void ClearData(const KernelId id, const bool clearSearcher = false);
    bool CalculateNextConfiguration(const KernelId id, const KernelResult& previousResult);
    bool HasData(const KernelId id) const;
    bool IsDataProcessed(const KernelId id) const;
    uint64_t GetTotalConfigurationsCount(const KernelId id) const;
    uint64_t GetExploredConfigurationsCount(const KernelId id) const;
    KernelConfiguration GetCurrentConfiguration(const KernelId id) const;
    KernelConfiguration GetBestConfiguration(const KernelId id) const;

private:
    std::map<KernelId, std::unique_ptr<Searcher>> m_Searchers;
    std::map<KernelId, std::unique_ptr<ConfigurationData>> m_ConfigurationData;
};

class TuningRunner
{
public:
    explicit TuningRunner(KernelRunner& kernelRunner);

    std::vector<KernelResult> Tune(const Kernel& kernel, const KernelDimensions& dimensions, std::unique_ptr<StopCondition> stopCondition);
    KernelResult TuneIteration(const Kernel& kernel, const KernelDimensions& dimensions, const KernelRunMode mode,
        const std::vector<BufferOutputDescriptor>& output, const bool recomputeReference);
    std::vector<KernelResult> SimulateTuning(const Kernel& kernel, const std::vector<KernelResult>& results,
        std::unique_ptr<StopCondition> stopCondition);
    void SetSearcher(const KernelId id, std::unique_ptr<Searcher> searcher);
    void InitializeConfigurationData(const Kernel& kernel);
    void ClearConfigurationData(const KernelId id, const bool clearSearcher = false);
    uint64_t GetConfigurationsCount(const KernelId id) const;
    KernelConfiguration GetBestConfiguration(const KernelId id) const;

private:
    KernelRunner& m_KernelRunner;
    std::unique_ptr<ConfigurationManager> m_ConfigurationManager;

    static const KernelResult& FindMatchingResult(const std::vector<KernelResult>& results, const KernelConfiguration& configuration);
};

} // namespace ktt
