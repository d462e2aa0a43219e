// C entry points for searchers over saved results and for hardware counters
// (include/ktt_hip.h section 6).
#include <algorithm>
#include <cstdio>
#include <fstream>
#include <memory>
#include <set>
#include <sstream>

#include <Ktt.h>

#include "../../../include/ktt_hip.h"
#include "Common.h"
#include "CoulombRuntime.h"
#include "Exploration.h"
#include "Json.h"
#include "Kernel.h"
#include "TunerCore.h"

using namespace ktt;

namespace
{

std::string CanonicalKey(const std::vector<ParameterPair>& pairs)
{
    std::vector<std::string> parts;
    for (const auto& pair : pairs)
    {
        parts.push_back(pair.GetName() + "=" + pair.GetValueString());
    }
    std::sort(parts.begin(), parts.end());
    std::string key;
    for (const auto& part : parts)
    {
        key += part + ";";
    }
    return key;
}

std::vector<KernelResult> LoadSaved(const std::string& path)
{
    std::ifstream stream(path);
    if (!stream)
    {
        throw KttException("Cannot open results file " + path);
    }
    std::stringstream buffer;
    buffer << stream.rdbuf();
    const JsonValue document = JsonValue::Parse(buffer.str());
    if (document.Contains("Metadata") && document.At("Metadata").Contains("TimeUnit"))
    {
        TimeConfiguration::GetInstance().SetTimeUnit(TimeUnitFromString(document.At("Metadata").At("TimeUnit").AsString()));
    }
    UserData data;
    return DeserializeResults(buffer.str(), data);
}

/** A kernel whose space is exactly the configurations of the saved results: every parameter
 *  takes the values the results use, and one constraint admits only saved combinations. */
KernelId AddSavedSpace(KernelRegistry& registry, const std::vector<KernelResult>& saved)
{
    if (saved.empty())
    {
        throw KttException("The results file holds no results");
    }
    const auto& first = saved.front().GetConfiguration().GetPairs();
    std::vector<std::string> names;
    std::vector<std::vector<ParameterValue>> values(first.size());
    std::vector<std::set<std::string>> seen(first.size());
    for (const auto& pair : first)
    {
        names.push_back(pair.GetName());
    }
    auto admitted = std::make_shared<std::set<std::string>>();
    for (const KernelResult& result : saved)
    {
        const auto& pairs = result.GetConfiguration().GetPairs();
        if (pairs.size() != names.size())
        {
            throw KttException("The saved results do not all configure the same parameters");
        }
        for (size_t i = 0; i < names.size(); ++i)
        {
            const auto pair = std::find_if(pairs.begin(), pairs.end(), [&](const ParameterPair& p) { return p.GetName() == names[i]; });
            if (pair == pairs.end())
            {
                throw KttException("A saved result lacks parameter " + names[i]);
            }
            if (seen[i].insert(pair->GetValueString()).second)
            {
                values[i].push_back(pair->GetValue());
            }
        }
        admitted->insert(CanonicalKey(pairs));
    }
    const std::string kernelName = saved.front().GetKernelName().empty() ? "saved" : saved.front().GetKernelName();
    const KernelDefinitionId definition = registry.AddDefinition(kernelName + "_definition", "", DimensionVector(1, 1, 1),
        DimensionVector(1, 1, 1), {});
    const KernelId id = registry.AddKernel(kernelName, {definition});
    Kernel& kernel = registry.GetKernel(id);
    for (size_t i = 0; i < names.size(); ++i)
    {
        kernel.AddParameter(KernelParameter(names[i], values[i], ""));
    }
    if (!names.empty())
    {
        kernel.AddConstraint(KernelConstraint(names, GenericConstraintFunction([names, admitted](const std::vector<const ParameterValue*>& v)
        {
            std::vector<ParameterPair> pairs;
            for (size_t i = 0; i < names.size(); ++i)
            {
                pairs.emplace_back(names[i], *v[i]);
            }
            return admitted->count(CanonicalKey(pairs)) != 0;
        })));
    }
    return id;
}

} // namespace

extern "C" {

int ktt_simulate_search(const char* resultsFile, const char* searcherName, const char* knowledgeFile, unsigned long long seed,
    long long maxSteps, unsigned computeUnits, double* stepMs, long long capacity, long long* steps)
{
    try
    {
        if (resultsFile == nullptr || searcherName == nullptr)
        {
            return SetLastError("a results file and a searcher name are required");
        }
        const std::vector<KernelResult> saved = LoadSaved(resultsFile);
        std::vector<KernelResult> knowledge;
        if (knowledgeFile != nullptr && knowledgeFile[0] != '\0')
        {
            knowledge = LoadSaved(knowledgeFile);
        }
        ArgumentRegistry arguments;
        KernelRegistry registry(arguments);
        const KernelId id = AddSavedSpace(registry, saved);
        const Kernel& kernel = registry.GetKernel(id);

        const uint64_t effectiveSeed = seed != 0 ? seed : DefaultSearcherSeed();
        const std::string name = searcherName;
        std::unique_ptr<Searcher> searcher;
        if (name == "Deterministic")
        {
            searcher = std::make_unique<DeterministicSearcher>();
        }
        else if (name == "Random")
        {
            searcher = std::make_unique<RandomSearcher>();
        }
        else if (name == "MCMC")
        {
            searcher = std::make_unique<McmcSearcher>(KernelConfiguration(), effectiveSeed);
        }
        else if (name == "ProfileBased")
        {
            ProfileSearcherOptions options;
            options.seed = effectiveSeed;
            options.computeUnits = computeUnits != 0 ? computeUnits : 256;
            searcher = std::make_unique<ProfileBasedSearcher>(options, knowledge);
        }
        else
        {
            return SetLastError("unknown searcher " + name + " (Deterministic, Random, MCMC or ProfileBased)");
        }

        // A replay logs every step at Info level; keep errors only.
        Logger& logger = Logger::GetLogger();
        const LoggingLevel level = logger.GetLoggingLevel();
        logger.SetLoggingLevel(LoggingLevel::Error);
        struct RestoreLevel
        {
            Logger& logger;
            LoggingLevel level;
            ~RestoreLevel() { logger.SetLoggingLevel(level); }
        } restore{logger, level};
        Exploration exploration(kernel, *searcher, effectiveSeed ^ 0x9E3779B97F4A7C15ull);
        std::unique_ptr<StopCondition> stop;
        if (maxSteps > 0)
        {
            stop = std::make_unique<ConfigurationCount>(static_cast<uint64_t>(maxSteps));
        }
        const std::vector<KernelResult> replayed = ReplayResults(exploration, kernel, saved, stop.get());
        if (steps != nullptr)
        {
            *steps = static_cast<long long>(replayed.size());
        }
        if (stepMs != nullptr)
        {
            for (size_t i = 0; i < replayed.size() && static_cast<long long>(i) < capacity; ++i)
            {
                stepMs[i] = replayed[i].IsValid() ? static_cast<double>(replayed[i].GetTotalDuration()) * 1.0e-6 : -1.0;
            }
        }
        return 0;
    }
    catch (const std::exception& failure)
    {
        return SetLastError(failure.what());
    }
}

int ktt_hip_enable_counters(void)
{
    std::string error;
    if (!EnableHipCounters(&error))
    {
        return SetLastError(error);
    }
    return 0;
}

} // extern "C"
