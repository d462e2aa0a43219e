#include <algorithm>
#include <functional>

#include <KttException.h>

#include "Common.h"
#include "Exploration.h"

namespace ktt
{

// ---------------------------------------------------------------- Exploration

std::string Exploration::Key(const std::vector<uint32_t>& row)
{
    return std::string(reinterpret_cast<const char*>(row.data()), row.size() * sizeof(uint32_t));
}

Exploration::Exploration(const Kernel& kernel, Searcher& searcher) :
    Exploration(kernel, searcher, DefaultSearcherSeed())
{}

Exploration::Exploration(const Kernel& kernel, Searcher& searcher, const uint64_t seed) :
    m_Kernel(kernel),
    m_Searcher(searcher),
    m_Size(0),
    m_BestDuration(InvalidDuration),
    m_SearcherLive(true),
    m_Random(seed)
{
    const auto& parameters = kernel.GetParameters();
    std::map<std::string, std::vector<size_t>> byGroup;
    for (size_t i = 0; i < parameters.size(); ++i)
    {
        byGroup[parameters[i].GetGroup()].push_back(i);
    }

    const Nanoseconds took = RunScopeTimer([&]()
    {
        for (const auto& [group, members] : byGroup)
        {
            Block block;
            block.parameters = members;
            block.first = m_Size;
            Enumerate(block);
            m_Size += block.members.size();
            if (!block.members.empty())
            {
                m_Blocks.push_back(std::move(block));
            }
        }
        if (parameters.empty())
        {
            // A kernel without parameters has exactly one (empty) configuration.
            Block block;
            block.members.emplace_back();
            block.rows.emplace(Key({}), 0);
            m_Blocks.push_back(std::move(block));
            m_Size = 1;
        }
    });
    if (m_Size == 0)
    {
        throw KttException("Kernel " + kernel.GetName() + " has no configuration that satisfies its constraints");
    }
    const auto& time = TimeConfiguration::GetInstance();
    Logger::LogInfo("Kernel " + kernel.GetName() + ": " + std::to_string(m_Size) + " configurations (enumerated in "
        + std::to_string(time.ConvertFromNanoseconds(took)) + time.GetUnitTag() + ")");

    for (const Block& block : m_Blocks)
    {
        m_Best.Merge(Compose(block, block.members.front()));
    }
    m_Searcher.Initialize(*this);
}

Exploration::~Exploration()
{
    m_Searcher.Reset();
}

void Exploration::Enumerate(Block& block) const
{
    const auto& parameters = m_Kernel.GetParameters();
    const size_t width = block.parameters.size();

    // Each constraint is evaluated at the position of its last parameter in this block.
    struct Check
    {
        const KernelConstraint* constraint;
        std::vector<size_t> positions;
    };
    std::vector<std::vector<Check>> checksAt(width);
    for (const KernelConstraint& constraint : m_Kernel.GetConstraints())
    {
        std::vector<size_t> positions;
        for (const std::string& name : constraint.GetParameters())
        {
            const auto found = std::find_if(block.parameters.begin(), block.parameters.end(),
                [&](const size_t index) { return parameters[index].GetName() == name; });
            if (found == block.parameters.end())
            {
                positions.clear();
                break;
            }
            positions.push_back(static_cast<size_t>(found - block.parameters.begin()));
        }
        if (!positions.empty())
        {
            checksAt[*std::max_element(positions.begin(), positions.end())].push_back({&constraint, positions});
        }
    }

    std::vector<uint32_t> row(width, 0);
    std::vector<const ParameterValue*> values;
    const std::function<void(size_t)> assign = [&](const size_t depth)
    {
        if (depth == width)
        {
            block.rows.emplace(Key(row), block.members.size());
            block.members.push_back(row);
            return;
        }
        const KernelParameter& parameter = parameters[block.parameters[depth]];
        for (uint32_t value = 0; value < parameter.GetValuesCount(); ++value)
        {
            row[depth] = value;
            const bool allowed = std::all_of(checksAt[depth].begin(), checksAt[depth].end(), [&](const Check& check)
            {
                values.clear();
                for (const size_t position : check.positions)
                {
                    values.push_back(&parameters[block.parameters[position]].GetValues()[row[position]]);
                }
                return check.constraint->IsFulfilled(values);
            });
            if (allowed)
            {
                assign(depth + 1);
            }
        }
    };
    assign(0);
}

KernelConfiguration Exploration::Compose(const Block& block, const std::vector<uint32_t>& row) const
{
    const auto& parameters = m_Kernel.GetParameters();
    std::vector<ParameterPair> pairs;
    pairs.reserve(row.size());
    for (size_t position = 0; position < row.size(); ++position)
    {
        pairs.push_back(parameters[block.parameters[position]].GeneratePair(row[position]));
    }
    return KernelConfiguration(pairs);
}

bool Exploration::Locate(const KernelConfiguration& configuration, size_t& blockIndex, std::vector<uint32_t>& row) const
{
    const auto& pairs = configuration.GetPairs();
    const auto& parameters = m_Kernel.GetParameters();
    if (pairs.empty())
    {
        blockIndex = 0;
        row.clear();
        return m_Blocks.front().parameters.empty();
    }
    for (size_t b = 0; b < m_Blocks.size(); ++b)
    {
        const Block& block = m_Blocks[b];
        const bool mine = std::any_of(block.parameters.begin(), block.parameters.end(),
            [&](const size_t index) { return parameters[index].GetName() == pairs.front().GetName(); });
        if (!mine)
        {
            continue;
        }
        row.assign(block.parameters.size(), 0);
        for (size_t position = 0; position < block.parameters.size(); ++position)
        {
            const KernelParameter& parameter = parameters[block.parameters[position]];
            const auto pair = std::find_if(pairs.begin(), pairs.end(), [&](const ParameterPair& p) { return p.GetName() == parameter.GetName(); });
            if (pair == pairs.end())
            {
                return false;
            }
            bool matched = false;
            for (uint32_t value = 0; value < parameter.GetValuesCount() && !matched; ++value)
            {
                if (parameter.GeneratePair(value).HasSameValue(*pair))
                {
                    row[position] = value;
                    matched = true;
                }
            }
            if (!matched)
            {
                return false;
            }
        }
        blockIndex = b;
        return true;
    }
    return false;
}

const Exploration::Block& Exploration::BlockOf(const uint64_t index, uint64_t& row) const
{
    for (const Block& block : m_Blocks)
    {
        if (index >= block.first && index < block.first + block.members.size())
        {
            row = index - block.first;
            return block;
        }
    }
    throw KttException("Configuration index " + std::to_string(index) + " is outside the space of " + m_Kernel.GetName());
}

KernelConfiguration Exploration::ConfigurationAt(const uint64_t index) const
{
    uint64_t row = 0;
    const Block& block = BlockOf(index, row);
    KernelConfiguration configuration = Compose(block, block.members[row]);
    configuration.Merge(m_Best);
    return configuration;
}

uint64_t Exploration::IndexOf(const KernelConfiguration& configuration) const
{
    size_t blockIndex = 0;
    std::vector<uint32_t> row;
    if (!Locate(configuration, blockIndex, row))
    {
        throw KttException("Configuration " + configuration.GetString() + " is not in the space of " + m_Kernel.GetName());
    }
    const Block& block = m_Blocks[blockIndex];
    const auto found = block.rows.find(Key(row));
    if (found == block.rows.end())
    {
        throw KttException("Configuration " + configuration.GetString() + " breaks a constraint of " + m_Kernel.GetName());
    }
    return block.first + found->second;
}

KernelConfiguration Exploration::DrawUnexplored() const
{
    if (m_Explored.size() >= m_Size)
    {
        throw KttException("Every configuration of " + m_Kernel.GetName() + " has been explored");
    }
    std::uniform_int_distribution<uint64_t> any(0, m_Size - 1);
    // Rejection sampling while most of the space is left; an explicit list once it is not.
    if (m_Explored.size() * 2 < m_Size)
    {
        for (;;)
        {
            const uint64_t index = any(m_Random);
            if (m_Explored.count(index) == 0)
            {
                return ConfigurationAt(index);
            }
        }
    }
    std::vector<uint64_t> left;
    for (uint64_t index = 0; index < m_Size; ++index)
    {
        if (m_Explored.count(index) == 0)
        {
            left.push_back(index);
        }
    }
    return ConfigurationAt(left[std::uniform_int_distribution<size_t>(0, left.size() - 1)(m_Random)]);
}

std::vector<KernelConfiguration> Exploration::Neighbours(const KernelConfiguration& configuration, const uint64_t maxDifferences,
    const size_t maxCount) const
{
    size_t blockIndex = 0;
    std::vector<uint32_t> origin;
    if (!Locate(configuration, blockIndex, origin))
    {
        throw KttException("Configuration " + configuration.GetString() + " is not in the space of " + m_Kernel.GetName());
    }
    const Block& block = m_Blocks[blockIndex];
    std::vector<std::pair<uint64_t, uint64_t>> near;  // (differences, row)
    for (uint64_t row = 0; row < block.members.size(); ++row)
    {
        if (m_Explored.count(block.first + row) != 0)
        {
            continue;
        }
        const auto& member = block.members[row];
        uint64_t differences = 0;
        for (size_t position = 0; position < member.size(); ++position)
        {
            differences += member[position] == origin[position] ? 0 : 1;
        }
        if (differences >= 1 && differences <= maxDifferences)
        {
            near.emplace_back(differences, row);
        }
    }
    std::sort(near.begin(), near.end());
    if (near.size() > maxCount)
    {
        near.resize(maxCount);
    }
    std::vector<KernelConfiguration> result;
    for (const auto& [differences, row] : near)
    {
        KernelConfiguration neighbour = Compose(block, block.members[row]);
        neighbour.Merge(m_Best);
        result.push_back(std::move(neighbour));
    }
    return result;
}

KernelConfiguration Exploration::Proposal() const
{
    if (!m_Searcher.IsInitialized() || Finished())
    {
        return KernelConfiguration();
    }
    return m_Searcher.GetCurrentConfiguration();
}

void Exploration::Record(const KernelResult& result)
{
    m_Explored.insert(IndexOf(result.GetConfiguration()));
    if (result.IsValid() && result.GetTotalDuration() < m_BestDuration)
    {
        m_Best = result.GetConfiguration();
        m_BestDuration = result.GetTotalDuration();
    }
    if (Finished())
    {
        return;
    }
    try
    {
        m_SearcherLive = m_Searcher.CalculateNextConfiguration(result);
    }
    catch (const std::exception& failure)
    {
        Logger::LogError(std::string("Searcher failed: ") + failure.what());
        m_SearcherLive = false;
    }
    if (m_SearcherLive)
    {
        Logger::LogInfo("Next configuration for " + m_Kernel.GetName() + ": " + m_Searcher.GetCurrentConfiguration().GetString());
    }
    else
    {
        Logger::LogInfo("The searcher of " + m_Kernel.GetName() + " has no further configuration");
    }
}

KernelConfiguration Exploration::Best() const
{
    return m_BestDuration != InvalidDuration ? m_Best : Proposal();
}

bool Exploration::LaunchSize(const KernelConfiguration& configuration, double& threads) const
{
    if (m_SavedLaunchSizes)
    {
        try
        {
            const auto found = m_LaunchSizes.find(IndexOf(configuration));
            if (found == m_LaunchSizes.end())
            {
                return false;
            }
            threads = found->second;
            return true;
        }
        catch (const KttException&)
        {
            return false;
        }
    }
    if (m_Kernel.GetDefinitions().empty())
    {
        return false;
    }
    const KernelDefinition& definition = *m_Kernel.GetDefinitions().front();
    const auto& pairs = configuration.GetPairs();
    const DimensionVector global = m_Kernel.GetModifiedSize(definition.GetId(), definition.GetGlobalSize(), ModifierType::Global, pairs);
    const DimensionVector local = m_Kernel.GetModifiedSize(definition.GetId(), definition.GetLocalSize(), ModifierType::Local, pairs);
    threads = static_cast<double>(global.GetTotalSize()) * static_cast<double>(local.GetTotalSize());
    return threads > 0.0;
}

void Exploration::SetLaunchSizes(std::map<uint64_t, double> sizes)
{
    m_LaunchSizes = std::move(sizes);
    m_SavedLaunchSizes = true;
}

// ---------------------------------------------------------------- replay

std::vector<KernelResult> ReplayResults(Exploration& exploration, const Kernel& kernel, const std::vector<KernelResult>& saved,
    StopCondition* stop)
{
    std::map<uint64_t, const KernelResult*> byIndex;
    std::map<uint64_t, double> launchSizes;
    for (const KernelResult& result : saved)
    {
        try
        {
            const uint64_t index = exploration.IndexOf(result.GetConfiguration());
            byIndex.emplace(index, &result);
            if (!result.GetResults().empty())
            {
                const ComputationResult& first = result.GetResults().front();
                launchSizes.emplace(index, static_cast<double>(first.GetGlobalSize().GetTotalSize())
                    * static_cast<double>(first.GetLocalSize().GetTotalSize()));
            }
        }
        catch (const KttException&)
        {
            // Saved results outside the current space are never proposed.
        }
    }
    // The kernel of a replay need not carry the saved runs' thread modifiers: the launch sizes
    // the searcher sees are those the saved results recorded.
    exploration.SetLaunchSizes(std::move(launchSizes));
    if (stop != nullptr)
    {
        stop->Initialize(exploration.Size());
    }
    std::vector<KernelResult> replayed;
    while (!exploration.Finished() && (stop == nullptr || !stop->IsFulfilled()))
    {
        const KernelConfiguration proposal = exploration.Proposal();
        const auto found = byIndex.find(exploration.IndexOf(proposal));
        KernelResult result = found != byIndex.end() ? *found->second : KernelResult(kernel.GetName(), proposal);
        if (found == byIndex.end())
        {
            Logger::LogWarning("No saved result for " + proposal.GetString());
            result.SetStatus(ResultStatus::ComputationFailed);
        }
        result.SetSearcherOverhead(RunScopeTimer([&]() { exploration.Record(result); }));
        if (stop != nullptr)
        {
            stop->Update(result);
        }
        replayed.push_back(std::move(result));
    }
    return replayed;
}

// ---------------------------------------------------------------- TuningLoop

TuningLoop::TuningLoop(Executor& executor) : m_Executor(executor) {}

Exploration& TuningLoop::Ensure(const Kernel& kernel)
{
    KernelState& state = m_States[kernel.GetId()];
    if (!state.searcher)
    {
        state.searcher = std::make_unique<DeterministicSearcher>();
    }
    if (!state.exploration)
    {
        state.exploration = std::make_unique<Exploration>(kernel, *state.searcher);
    }
    return *state.exploration;
}

const Exploration* TuningLoop::Find(const KernelId id) const
{
    const auto found = m_States.find(id);
    return found != m_States.end() ? found->second.exploration.get() : nullptr;
}

void TuningLoop::SetSearcher(const KernelId id, std::unique_ptr<Searcher> searcher)
{
    KernelState& state = m_States[id];
    state.exploration.reset();
    state.searcher = std::move(searcher);
}

void TuningLoop::Prepare(const Kernel& kernel)
{
    m_States[kernel.GetId()].exploration.reset();
    Ensure(kernel);
}

void TuningLoop::Reset(const KernelId id, const bool dropSearcher)
{
    const auto found = m_States.find(id);
    if (found == m_States.end())
    {
        return;
    }
    found->second.exploration.reset();
    if (dropSearcher)
    {
        m_States.erase(found);
    }
}

uint64_t TuningLoop::SpaceSize(const KernelId id) const
{
    const Exploration* exploration = Find(id);
    return exploration != nullptr ? exploration->Size() : 0;
}

KernelConfiguration TuningLoop::Best(const KernelId id) const
{
    const Exploration* exploration = Find(id);
    if (exploration == nullptr)
    {
        throw KttException("The best configuration is known only after the kernel's space has been prepared or searched");
    }
    return exploration->Best();
}

std::vector<KernelResult> TuningLoop::Tune(const Kernel& kernel, const KernelDimensions& dimensions, std::unique_ptr<StopCondition> stop)
{
    Logger::LogInfo("Tuning " + kernel.GetName());
    Exploration& exploration = Ensure(kernel);
    if (stop != nullptr)
    {
        stop->Initialize(exploration.Size());
    }
    std::vector<KernelResult> results;
    while (!exploration.Finished())
    {
        const KernelConfiguration configuration = exploration.Proposal();
        KernelResult result = m_Executor.Execute(kernel, configuration, dimensions, KernelRunMode::OfflineTuning, {});
        if (result.HasRemainingProfilingRuns())
        {
            // Counter passes: the same launch runs until every pass has been collected, and the
            // last run carries the counters and the overheads of all runs (TuningRunner.cpp:37-54).
            KernelResult passes(kernel.GetName(), configuration);
            passes.FuseProfilingTimes(result, true);
            while (result.IsValid() && result.HasRemainingProfilingRuns())
            {
                result = m_Executor.Execute(kernel, configuration, dimensions, KernelRunMode::OfflineTuning, {});
                passes.FuseProfilingTimes(result, false);
            }
            result.CopyProfilingTimes(passes);
        }
        result.SetSearcherOverhead(RunScopeTimer([&]() { exploration.Record(result); }));
        results.push_back(result);
        if (stop != nullptr)
        {
            stop->Update(result);
            Logger::LogInfo(stop->GetStatusString());
            if (stop->IsFulfilled())
            {
                break;
            }
        }
    }
    Logger::LogInfo("Tuning of " + kernel.GetName() + " tested " + std::to_string(results.size()) + " configurations");
    m_Executor.Checker().Drop(kernel);
    return results;
}

KernelResult TuningLoop::Step(const Kernel& kernel, const KernelDimensions& dimensions, const KernelRunMode mode,
    const std::vector<BufferOutputDescriptor>& output, const bool recomputeReference)
{
    if (recomputeReference)
    {
        m_Executor.Checker().Drop(kernel);
    }
    Exploration& exploration = Ensure(kernel);
    const bool searching = !exploration.Finished();
    const KernelConfiguration configuration = searching ? exploration.Proposal() : exploration.Best();
    if (searching)
    {
        Logger::LogInfo("Online step " + std::to_string(exploration.Explored().size() + 1) + " of " + std::to_string(exploration.Size())
            + " for " + kernel.GetName());
    }
    KernelResult result = m_Executor.Execute(kernel, configuration, dimensions, mode, output);
    // A profiled step with counter passes left keeps its configuration: the caller steps again
    // until the last pass (TuningRunner.cpp:118).
    if (searching && !result.HasRemainingProfilingRuns())
    {
        result.SetSearcherOverhead(RunScopeTimer([&]() { exploration.Record(result); }));
    }
    return result;
}

std::vector<KernelResult> TuningLoop::Replay(const Kernel& kernel, const std::vector<KernelResult>& saved, std::unique_ptr<StopCondition> stop)
{
    Logger::LogInfo("Replaying " + std::to_string(saved.size()) + " saved results of " + kernel.GetName());
    std::vector<KernelResult> replayed = ReplayResults(Ensure(kernel), kernel, saved, stop.get());
    Reset(kernel.GetId());
    return replayed;
}

} // namespace ktt
