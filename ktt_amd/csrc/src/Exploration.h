// Configuration spaces and the tuning loop.
//
//   Exploration  the valid configurations of one kernel and the progress of a search over
//                them. Parameters are split by group name; each group's configurations are
//                the cartesian product of its parameters' values that satisfies the group's
//                constraints (a constraint is checked as soon as its last parameter has a
//                value). Groups are searched one after another and a configuration of one
//                group is completed with the best values found so far for the others.
//                Indices run over the groups in order. The searcher bound to it proposes the
//                next configuration; the exploration records results and the best one.
//   TuningLoop   offline tuning, single online steps and replays of saved results, with one
//                searcher and one exploration per kernel.
//
// The behaviour users rely on is the reference's (Source/TuningRunner/TuningRunner.cpp:17-201,
// ConfigurationData.cpp:31-305): the best configuration is the valid one with the smallest
// total duration, a step after the search has ended runs the best configuration, and a replay
// answers each proposal with the saved result of the same configuration.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <random>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

#include <KernelConfiguration.h>
#include <KernelResult.h>
#include <Searcher.h>
#include <StopCondition.h>

#include "Execution.h"
#include "Kernel.h"

namespace ktt
{

class Exploration
{
public:
    /** Enumerates the space and binds the searcher; throws if the space is empty. */
    Exploration(const Kernel& kernel, Searcher& searcher);
    /** Same, with the seed of the draws of unexplored configurations given. */
    Exploration(const Kernel& kernel, Searcher& searcher, const uint64_t seed);
    ~Exploration();
    Exploration(const Exploration&) = delete;
    Exploration& operator=(const Exploration&) = delete;

    uint64_t Size() const { return m_Size; }
    KernelConfiguration ConfigurationAt(const uint64_t index) const;
    uint64_t IndexOf(const KernelConfiguration& configuration) const;
    /** A uniformly drawn configuration that has not been explored. */
    KernelConfiguration DrawUnexplored() const;
    /** Unexplored configurations differing from `configuration` in 1..maxDifferences values, closest first. */
    std::vector<KernelConfiguration> Neighbours(const KernelConfiguration& configuration, const uint64_t maxDifferences,
        const size_t maxCount) const;
    const std::set<uint64_t>& Explored() const { return m_Explored; }

    /** The searcher's proposal, or an empty configuration once the search has ended. */
    KernelConfiguration Proposal() const;
    /** Marks the result's configuration explored, updates the best, asks the searcher for the next. */
    void Record(const KernelResult& result);
    bool Finished() const { return m_Explored.size() >= m_Size || !m_SearcherLive; }
    /** The fastest valid configuration so far (the first proposal before any valid result). */
    KernelConfiguration Best() const;
    /** Threads one launch of `configuration` runs in the kernel's first definition: its global
     *  size times its local size after the thread modifiers. After SetLaunchSizes, only the sizes
     *  given there count. False when the size is not known. */
    bool LaunchSize(const KernelConfiguration& configuration, double& threads) const;
    /** Launch sizes recorded by saved results (by configuration index), for replays. */
    void SetLaunchSizes(std::map<uint64_t, double> sizes);

private:
    struct Block
    {
        std::vector<size_t> parameters;              // positions in the kernel's parameter list
        std::vector<std::vector<uint32_t>> members;  // value positions, one row per configuration
        std::unordered_map<std::string, uint64_t> rows;
        uint64_t first = 0;                          // global index of members[0]
    };

    const Kernel& m_Kernel;
    Searcher& m_Searcher;
    std::vector<Block> m_Blocks;
    uint64_t m_Size;
    std::set<uint64_t> m_Explored;
    KernelConfiguration m_Best;
    Nanoseconds m_BestDuration;
    bool m_SearcherLive;
    mutable std::mt19937_64 m_Random;
    std::map<uint64_t, double> m_LaunchSizes;
    bool m_SavedLaunchSizes = false;

    void Enumerate(Block& block) const;
    KernelConfiguration Compose(const Block& block, const std::vector<uint32_t>& row) const;
    bool Locate(const KernelConfiguration& configuration, size_t& block, std::vector<uint32_t>& row) const;
    const Block& BlockOf(const uint64_t index, uint64_t& row) const;
    static std::string Key(const std::vector<uint32_t>& row);
};

/** Answers the exploration's proposals with saved results until it ends or `stop` holds. */
std::vector<KernelResult> ReplayResults(Exploration& exploration, const Kernel& kernel, const std::vector<KernelResult>& saved,
    StopCondition* stop);

class TuningLoop
{
public:
    explicit TuningLoop(Executor& executor);

    std::vector<KernelResult> Tune(const Kernel& kernel, const KernelDimensions& dimensions, std::unique_ptr<StopCondition> stop);
    KernelResult Step(const Kernel& kernel, const KernelDimensions& dimensions, const KernelRunMode mode,
        const std::vector<BufferOutputDescriptor>& output, const bool recomputeReference);
    std::vector<KernelResult> Replay(const Kernel& kernel, const std::vector<KernelResult>& saved, std::unique_ptr<StopCondition> stop);

    void SetSearcher(const KernelId id, std::unique_ptr<Searcher> searcher);
    void Prepare(const Kernel& kernel);
    void Reset(const KernelId id, const bool dropSearcher = false);
    uint64_t SpaceSize(const KernelId id) const;
    KernelConfiguration Best(const KernelId id) const;

private:
    struct KernelState
    {
        std::unique_ptr<Searcher> searcher;
        std::unique_ptr<Exploration> exploration;
    };

    Executor& m_Executor;
    std::map<KernelId, KernelState> m_States;

    Exploration& Ensure(const Kernel& kernel);
    const Exploration* Find(const KernelId id) const;
};

} // namespace ktt
