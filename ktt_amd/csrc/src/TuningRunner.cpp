#include <algorithm>
#include <limits>

#include <KttException.h>

#include "Common.h"
#include "TuningRunner.h"

namespace ktt
{

// ---------------------------------------------------------------- ConfigurationData

ConfigurationData::ConfigurationData(Searcher& searcher, const Kernel& kernel) :
    m_Searcher(searcher),
    m_Kernel(kernel),
    m_BestConfiguration({KernelConfiguration(), InvalidDuration}),
    m_SearcherActive(false),
    m_Generator(DefaultSearcherSeed())
{
    InitializeConfigurations();
}

ConfigurationData::~ConfigurationData()
{
    m_Searcher.Reset();
}

std::string ConfigurationData::Encode(const std::vector<uint32_t>& values)
{
    return std::string(reinterpret_cast<const char*>(values.data()), values.size() * sizeof(uint32_t));
}

void ConfigurationData::EnumerateGroup(Group& group)
{
    const auto& parameters = m_Kernel.GetParameters();
    const auto& constraints = m_Kernel.GetConstraints();
    const size_t count = group.parameters.size();

    // Constraints to check once parameter position p is assigned (their last parameter is at p).
    std::vector<std::vector<std::pair<const KernelConstraint*, std::vector<size_t>>>> checks(count);
    for (const auto& constraint : constraints)
    {
        std::vector<size_t> positions;
        bool inGroup = true;
        for (const auto& name : constraint.GetParameters())
        {
            size_t position = count;
            for (size_t p = 0; p < count; ++p)
            {
                if (parameters[group.parameters[p]].GetName() == name)
                {
                    position = p;
                    break;
                }
            }
            if (position == count)
            {
                inGroup = false;
                break;
            }
            positions.push_back(position);
        }
        if (!inGroup || positions.empty())
        {
            continue;
        }
        const size_t last = *std::max_element(positions.begin(), positions.end());
        checks[last].emplace_back(&constraint, positions);
    }

    std::vector<uint32_t> current(count, 0);
    std::vector<const ParameterValue*> values;

    // Iterative depth-first enumeration; the first parameter varies slowest.
    size_t depth = 0;
    std::vector<uint32_t> next(count + 1, 0);
    while (true)
    {
        if (depth == count)
        {
            group.lookup.emplace(Encode(current), group.configurations.size());
            group.configurations.push_back(current);
            if (depth == 0)
            {
                break;
            }
            --depth;
            continue;
        }
        const auto& parameter = parameters[group.parameters[depth]];
        if (next[depth] >= parameter.GetValuesCount())
        {
            next[depth] = 0;
            if (depth == 0)
            {
                break;
            }
            --depth;
            continue;
        }
        current[depth] = next[depth]++;
        bool valid = true;
        for (const auto& [constraint, positions] : checks[depth])
        {
            values.clear();
            for (const size_t position : positions)
            {
                values.push_back(&parameters[group.parameters[position]].GetValues()[current[position]]);
            }
            if (!constraint->IsFulfilled(values))
            {
                valid = false;
                break;
            }
        }
        if (valid)
        {
            ++depth;
        }
    }
}

void ConfigurationData::InitializeConfigurations()
{
    const auto& parameters = m_Kernel.GetParameters();
    std::map<std::string, std::vector<size_t>> grouped;
    for (size_t i = 0; i < parameters.size(); ++i)
    {
        grouped[parameters[i].GetGroup()].push_back(i);
    }

    Timer timer;
    timer.Start();
    uint64_t offset = 0;
    for (const auto& [name, indices] : grouped)
    {
        Group group;
        group.parameters = indices;
        group.offset = offset;
        EnumerateGroup(group);
        offset += group.configurations.size();
        if (!group.configurations.empty())
        {
            m_Groups.push_back(std::move(group));
        }
    }
    if (parameters.empty())
    {
        Group group;
        group.configurations.push_back({});
        group.lookup.emplace(Encode({}), 0);
        m_Groups.push_back(std::move(group));
    }
    timer.Stop();

    if (GetTotalConfigurationsCount() == 0)
    {
        throw KttException("No valid configuration exists for kernel " + m_Kernel.GetName());
    }

    const auto& time = TimeConfiguration::GetInstance();
    Logger::LogInfo("Total count of " + std::to_string(GetTotalConfigurationsCount()) + " configurations was generated in "
        + std::to_string(time.ConvertFromNanoseconds(timer.GetElapsedTime())) + time.GetUnitTag());

    KernelConfiguration initialBest;
    for (const auto& group : m_Groups)
    {
        initialBest.Merge(BuildConfiguration(group, group.configurations[0]));
    }
    m_BestConfiguration = {initialBest, InvalidDuration};
    m_SearcherActive = true;
    m_Searcher.Initialize(*this);
}

KernelConfiguration ConfigurationData::BuildConfiguration(const Group& group, const std::vector<uint32_t>& values) const
{
    const auto& parameters = m_Kernel.GetParameters();
    std::vector<ParameterPair> pairs;
    pairs.reserve(group.parameters.size());
    for (size_t p = 0; p < group.parameters.size(); ++p)
    {
        pairs.push_back(parameters[group.parameters[p]].GeneratePair(values[p]));
    }
    return KernelConfiguration(pairs);
}

bool ConfigurationData::DecodeConfiguration(const KernelConfiguration& configuration, size_t& groupIndex,
    std::vector<uint32_t>& values) const
{
    const auto& pairs = configuration.GetPairs();
    const auto& parameters = m_Kernel.GetParameters();
    if (pairs.empty())
    {
        groupIndex = 0;
        values.clear();
        return m_Groups[0].parameters.empty();
    }

    for (size_t g = 0; g < m_Groups.size(); ++g)
    {
        const auto& group = m_Groups[g];
        if (group.parameters.empty())
        {
            continue;
        }
        bool belongs = false;
        for (const size_t index : group.parameters)
        {
            if (parameters[index].GetName() == pairs[0].GetName())
            {
                belongs = true;
                break;
            }
        }
        if (!belongs)
        {
            continue;
        }

        values.assign(group.parameters.size(), 0);
        for (size_t p = 0; p < group.parameters.size(); ++p)
        {
            const auto& parameter = parameters[group.parameters[p]];
            const ParameterPair* match = nullptr;
            for (const auto& pair : pairs)
            {
                if (pair.GetName() == parameter.GetName())
                {
                    match = &pair;
                    break;
                }
            }
            if (match == nullptr)
            {
                return false;
            }
            bool found = false;
            for (size_t v = 0; v < parameter.GetValuesCount(); ++v)
            {
                if (parameter.GeneratePair(v).HasSameValue(*match))
                {
                    values[p] = static_cast<uint32_t>(v);
                    found = true;
                    break;
                }
            }
            if (!found)
            {
                return false;
            }
        }
        groupIndex = g;
        return true;
    }
    return false;
}

bool ConfigurationData::CalculateNextConfiguration(const KernelResult& previousResult)
{
    const uint64_t index = GetIndexForConfiguration(previousResult.GetConfiguration());
    m_ExploredConfigurations.insert(index);

    if (previousResult.IsValid())
    {
        UpdateBestConfiguration(previousResult);
    }
    if (IsProcessed())
    {
        return false;
    }

    try
    {
        m_SearcherActive = m_Searcher.CalculateNextConfiguration(previousResult);
        if (m_SearcherActive)
        {
            const auto current = GetCurrentConfiguration();
            Logger::LogInfo("Searcher selected configuration " + std::to_string(GetIndexForConfiguration(current)) + ": "
                + current.GetString());
        }
    }
    catch (const std::runtime_error& error)
    {
        Logger::LogError(error.what());
        m_SearcherActive = false;
    }
    if (!m_SearcherActive)
    {
        Logger::LogInfo("Searcher failed to calculate next configuration for kernel " + m_Kernel.GetName());
    }
    return m_SearcherActive;
}

const ConfigurationData::Group& ConfigurationData::GetGroupForIndex(const uint64_t index, uint64_t& localIndex) const
{
    if (index >= GetTotalConfigurationsCount())
    {
        throw KttException("Invalid configuration index");
    }
    localIndex = index;
    for (const auto& group : m_Groups)
    {
        if (localIndex < group.configurations.size())
        {
            return group;
        }
        localIndex -= group.configurations.size();
    }
    throw KttException("Invalid configuration index");
}

KernelConfiguration ConfigurationData::GetConfigurationForIndex(const uint64_t index) const
{
    uint64_t localIndex = 0;
    const auto& group = GetGroupForIndex(index, localIndex);
    auto result = BuildConfiguration(group, group.configurations[localIndex]);
    result.Merge(m_BestConfiguration.first);
    return result;
}

uint64_t ConfigurationData::GetIndexForConfiguration(const KernelConfiguration& configuration) const
{
    size_t groupIndex = 0;
    std::vector<uint32_t> values;
    if (!DecodeConfiguration(configuration, groupIndex, values))
    {
        throw KttException("Configuration " + configuration.GetString() + " does not belong to the space of kernel " + m_Kernel.GetName());
    }
    const auto& group = m_Groups[groupIndex];
    auto iterator = group.lookup.find(Encode(values));
    if (iterator == group.lookup.cend())
    {
        throw KttException("Configuration " + configuration.GetString() + " violates the constraints of kernel " + m_Kernel.GetName());
    }
    return group.offset + iterator->second;
}

KernelConfiguration ConfigurationData::GetRandomConfiguration() const
{
    const uint64_t total = GetTotalConfigurationsCount();
    if (m_ExploredConfigurations.size() >= total)
    {
        throw KttException("All configurations were already explored");
    }
    uint64_t index = 0;
    if (m_ExploredConfigurations.size() * 2 < total)
    {
        std::uniform_int_distribution<uint64_t> distribution(0, total - 1);
        do
        {
            index = distribution(m_Generator);
        }
        while (ContainsKey(m_ExploredConfigurations, index));
    }
    else
    {
        std::vector<uint64_t> unexplored;
        for (uint64_t i = 0; i < total; ++i)
        {
            if (!ContainsKey(m_ExploredConfigurations, i))
            {
                unexplored.push_back(i);
            }
        }
        std::uniform_int_distribution<size_t> distribution(0, unexplored.size() - 1);
        index = unexplored[distribution(m_Generator)];
    }
    return GetConfigurationForIndex(index);
}

std::vector<KernelConfiguration> ConfigurationData::GetNeighbourConfigurations(const KernelConfiguration& configuration,
    const uint64_t maxDifferences, const size_t maxNeighbours) const
{
    size_t groupIndex = 0;
    std::vector<uint32_t> values;
    if (!DecodeConfiguration(configuration, groupIndex, values))
    {
        throw KttException("Configuration " + configuration.GetString() + " does not belong to the space of kernel " + m_Kernel.GetName());
    }
    const auto& group = m_Groups[groupIndex];

    // Valid, unexplored configurations that differ in 1..maxDifferences parameters, nearest first.
    std::vector<std::pair<uint64_t, uint64_t>> candidates;
    for (uint64_t local = 0; local < group.configurations.size(); ++local)
    {
        const auto& other = group.configurations[local];
        uint64_t differences = 0;
        for (size_t p = 0; p < other.size() && differences <= maxDifferences; ++p)
        {
            differences += other[p] != values[p] ? 1 : 0;
        }
        if (differences == 0 || differences > maxDifferences)
        {
            continue;
        }
        if (ContainsKey(m_ExploredConfigurations, group.offset + local))
        {
            continue;
        }
        candidates.emplace_back(differences, local);
    }
    std::sort(candidates.begin(), candidates.end());

    std::vector<KernelConfiguration> result;
    for (const auto& [differences, local] : candidates)
    {
        if (result.size() >= maxNeighbours)
        {
            break;
        }
        auto neighbour = BuildConfiguration(group, group.configurations[local]);
        neighbour.Merge(m_BestConfiguration.first);
        result.push_back(neighbour);
    }
    return result;
}

uint64_t ConfigurationData::GetTotalConfigurationsCount() const
{
    uint64_t result = 0;
    for (const auto& group : m_Groups)
    {
        result += group.configurations.size();
    }
    return result;
}

uint64_t ConfigurationData::GetExploredConfigurationsCount() const
{
    return static_cast<uint64_t>(m_ExploredConfigurations.size());
}

const std::set<uint64_t>& ConfigurationData::GetExploredConfigurations() const
{
    return m_ExploredConfigurations;
}

bool ConfigurationData::IsProcessed() const
{
    return GetExploredConfigurationsCount() >= GetTotalConfigurationsCount() || !m_SearcherActive;
}

KernelConfiguration ConfigurationData::GetCurrentConfiguration() const
{
    if (!m_Searcher.IsInitialized() || IsProcessed())
    {
        return KernelConfiguration();
    }
    return m_Searcher.GetCurrentConfiguration();
}

KernelConfiguration ConfigurationData::GetBestConfiguration() const
{
    if (m_BestConfiguration.second != InvalidDuration)
    {
        return m_BestConfiguration.first;
    }
    return GetCurrentConfiguration();
}

void ConfigurationData::UpdateBestConfiguration(const KernelResult& previousResult)
{
    const Nanoseconds duration = previousResult.GetTotalDuration();
    if (duration < m_BestConfiguration.second)
    {
        m_BestConfiguration.first = previousResult.GetConfiguration();
        m_BestConfiguration.second = duration;
    }
}

// ---------------------------------------------------------------- ConfigurationManager

void ConfigurationManager::SetSearcher(const KernelId id, std::unique_ptr<Searcher> searcher)
{
    ClearData(id);
    m_Searchers[id] = std::move(searcher);
}

void ConfigurationManager::InitializeData(const Kernel& kernel)
{
    const auto id = kernel.GetId();
    if (!ContainsKey(m_Searchers, id))
    {
        m_Searchers[id] = std::make_unique<DeterministicSearcher>();
    }
    m_ConfigurationData.erase(id);
    m_ConfigurationData[id] = std::make_unique<ConfigurationData>(*m_Searchers[id], kernel);
}

void ConfigurationManager::ClearData(const KernelId id, const bool clearSearcher)
{
    m_ConfigurationData.erase(id);
    if (clearSearcher)
    {
        m_Searchers.erase(id);
    }
}

bool ConfigurationManager::CalculateNextConfiguration(const KernelId id, const KernelResult& previousResult)
{
    return m_ConfigurationData.at(id)->CalculateNextConfiguration(previousResult);
}

bool ConfigurationManager::HasData(const KernelId id) const
{
    return ContainsKey(m_ConfigurationData, id);
}

bool ConfigurationManager::IsDataProcessed(const KernelId id) const
{
    if (!HasData(id))
    {
        return true;
    }
    return m_ConfigurationData.find(id)->second->IsProcessed();
}

uint64_t ConfigurationManager::GetTotalConfigurationsCount(const KernelId id) const
{
    return HasData(id) ? m_ConfigurationData.find(id)->second->GetTotalConfigurationsCount() : 0;
}

uint64_t ConfigurationManager::GetExploredConfigurationsCount(const KernelId id) const
{
    return HasData(id) ? m_ConfigurationData.find(id)->second->GetExploredConfigurationsCount() : 0;
}

KernelConfiguration ConfigurationManager::GetCurrentConfiguration(const KernelId id) const
{
    return m_ConfigurationData.at(id)->GetCurrentConfiguration();
}

KernelConfiguration ConfigurationManager::GetBestConfiguration(const KernelId id) const
{
    if (!HasData(id))
    {
        throw KttException("The best configuration can only be retrieved for kernels with initialized configuration data");
    }
    return m_ConfigurationData.find(id)->second->GetBestConfiguration();
}

// ---------------------------------------------------------------- TuningRunner

TuningRunner::TuningRunner(KernelRunner& kernelRunner) :
    m_KernelRunner(kernelRunner),
    m_ConfigurationManager(std::make_unique<ConfigurationManager>())
{}

std::vector<KernelResult> TuningRunner::Tune(const Kernel& kernel, const KernelDimensions& dimensions,
    std::unique_ptr<StopCondition> stopCondition)
{
    Logger::LogInfo("Starting offline tuning for kernel " + kernel.GetName());
    const auto id = kernel.GetId();
    if (!m_ConfigurationManager->HasData(id))
    {
        m_ConfigurationManager->InitializeData(kernel);
    }
    if (stopCondition != nullptr)
    {
        stopCondition->Initialize(m_ConfigurationManager->GetTotalConfigurationsCount(id));
    }

    std::vector<KernelResult> results;
    while (!m_ConfigurationManager->IsDataProcessed(id))
    {
        KernelResult result = TuneIteration(kernel, dimensions, KernelRunMode::OfflineTuning, {}, false);
        const Nanoseconds searcherOverhead = RunScopeTimer([this, id, &result]()
        {
            m_ConfigurationManager->CalculateNextConfiguration(id, result);
        });
        result.SetSearcherOverhead(searcherOverhead);
        results.push_back(result);

        if (stopCondition == nullptr)
        {
            continue;
        }
        stopCondition->Update(result);
        Logger::LogInfo(stopCondition->GetStatusString());
        if (stopCondition->IsFulfilled())
        {
            break;
        }
    }

    Logger::LogInfo("Ending offline tuning for kernel " + kernel.GetName() + ", total number of tested configurations is "
        + std::to_string(results.size()));
    m_KernelRunner.ClearReferenceResult(kernel);
    return results;
}

KernelResult TuningRunner::TuneIteration(const Kernel& kernel, const KernelDimensions& dimensions, const KernelRunMode mode,
    const std::vector<BufferOutputDescriptor>& output, const bool recomputeReference)
{
    if (recomputeReference)
    {
        m_KernelRunner.ClearReferenceResult(kernel);
    }
    const auto id = kernel.GetId();
    if (!m_ConfigurationManager->HasData(id))
    {
        m_ConfigurationManager->InitializeData(kernel);
    }

    KernelConfiguration configuration;
    if (m_ConfigurationManager->IsDataProcessed(id))
    {
        configuration = m_ConfigurationManager->GetBestConfiguration(id);
        Logger::LogInfo("Launching the best configuration for kernel " + kernel.GetName());
    }
    else
    {
        configuration = m_ConfigurationManager->GetCurrentConfiguration(id);
        Logger::LogInfo("Launching configuration " + std::to_string(m_ConfigurationManager->GetExploredConfigurationsCount(id) + 1)
            + " / " + std::to_string(m_ConfigurationManager->GetTotalConfigurationsCount(id)) + " for kernel " + kernel.GetName());
    }

    KernelResult result = m_KernelRunner.RunKernel(kernel, configuration, dimensions, mode, output);

    if (mode != KernelRunMode::OfflineTuning && !m_ConfigurationManager->IsDataProcessed(id))
    {
        const Nanoseconds searcherOverhead = RunScopeTimer([this, id, &result]()
        {
            m_ConfigurationManager->CalculateNextConfiguration(id, result);
        });
        result.SetSearcherOverhead(searcherOverhead);
    }
    return result;
}

const KernelResult& TuningRunner::FindMatchingResult(const std::vector<KernelResult>& results, const KernelConfiguration& configuration)
{
    for (const auto& result : results)
    {
        if (result.GetConfiguration() == configuration)
        {
            return result;
        }
    }
    throw KttException("Matching configuration was not found in tuning results");
}

std::vector<KernelResult> TuningRunner::SimulateTuning(const Kernel& kernel, const std::vector<KernelResult>& results,
    std::unique_ptr<StopCondition> stopCondition)
{
    Logger::LogInfo("Starting simulated tuning for kernel " + kernel.GetName());
    const auto id = kernel.GetId();
    if (!m_ConfigurationManager->HasData(id))
    {
        m_ConfigurationManager->InitializeData(kernel);
    }
    const uint64_t configurationCount = m_ConfigurationManager->GetTotalConfigurationsCount(id);
    if (stopCondition != nullptr)
    {
        stopCondition->Initialize(configurationCount);
    }

    std::vector<KernelResult> output;
    uint64_t passedIterations = 0;
    while (!m_ConfigurationManager->IsDataProcessed(id))
    {
        if (stopCondition != nullptr && stopCondition->IsFulfilled())
        {
            break;
        }
        const auto current = m_ConfigurationManager->GetCurrentConfiguration(id);
        KernelResult result;
        try
        {
            result = FindMatchingResult(results, current);
        }
        catch (const KttException& error)
        {
            Logger::LogWarning(std::string("Kernel run failed, reason: ") + error.what());
            result = KernelResult(kernel.GetName(), current);
            result.SetStatus(ResultStatus::ComputationFailed);
        }
        const Nanoseconds searcherOverhead = RunScopeTimer([this, id, &result]()
        {
            m_ConfigurationManager->CalculateNextConfiguration(id, result);
        });
        result.SetSearcherOverhead(searcherOverhead);
        output.push_back(result);
        ++passedIterations;
        if (stopCondition != nullptr)
        {
            stopCondition->Update(result);
            Logger::LogInfo(stopCondition->GetStatusString());
        }
    }

    Logger::LogInfo("Ending simulated tuning for kernel " + kernel.GetName() + ", total number of tested configurations is "
        + std::to_string(passedIterations));
    m_ConfigurationManager->ClearData(id);
    return output;
}

void TuningRunner::SetSearcher(const KernelId id, std::unique_ptr<Searcher> searcher)
{
    m_ConfigurationManager->SetSearcher(id, std::move(searcher));
}

void TuningRunner::InitializeConfigurationData(const Kernel& kernel)
{
    m_ConfigurationManager->InitializeData(kernel);
}

void TuningRunner::ClearConfigurationData(const KernelId id, const bool clearSearcher)
{
    m_ConfigurationManager->ClearData(id, clearSearcher);
}

uint64_t TuningRunner::GetConfigurationsCount(const KernelId id) const
{
    return m_ConfigurationManager->GetTotalConfigurationsCount(id);
}

KernelConfiguration TuningRunner::GetBestConfiguration(const KernelId id) const
{
    return m_ConfigurationManager->GetBestConfiguration(id);
}

} // namespace ktt
