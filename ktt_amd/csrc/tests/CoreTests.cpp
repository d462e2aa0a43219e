// Host-side tests of the ktt core (no device needed): configuration prefix, configuration
// spaces of the reference's Coulomb examples, the results loader on the reference's own RTX
// 2080 Ti full-search dump, simulated tuning with the three searchers, stop conditions and the
// JSON round trip.
//
// Usage: ktt_core_tests <path to tests/golden/coulomb3d_2080_space.json>
// Exit status 0 when every check passes; each failure is printed with its line.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include <Ktt.h>

#include "Common.h"
#include "Json.h"
#include "Kernel.h"
#include "KernelArgument.h"
#include "TunerCore.h"
#include "TuningRunner.h"

using namespace ktt;

namespace
{

int g_Failures = 0;
int g_Checks = 0;

#define CHECK(condition)                                                                                     \
    do                                                                                                       \
    {                                                                                                        \
        ++g_Checks;                                                                                          \
        if (!(condition))                                                                                    \
        {                                                                                                    \
            ++g_Failures;                                                                                    \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #condition);               \
        }                                                                                                    \
    } while (false)

std::vector<ParameterValue> Values(const std::vector<uint64_t>& values)
{
    return std::vector<ParameterValue>(values.begin(), values.end());
}

/** The CoulombSum3d space of Examples/CoulombSum3d/CoulombSum3d.cpp:139-209 (default, neither
 *  dense nor wide parameters), for the CUDA or the OpenCL variant. */
KernelId AddReferenceCoulomb3dSpace(KernelManager& manager, const bool openCl)
{
    const KernelDefinitionId definition = manager.AddKernelDefinition(openCl ? "directCoulombSumCl" : "directCoulombSum", "",
        DimensionVector(256, 256, 256), DimensionVector(1, 1, 1), {});
    const KernelId kernel = manager.CreateKernel(openCl ? "CoulombSumCl" : "CoulombSum", {definition});
    manager.AddParameter(kernel, "WORK_GROUP_SIZE_X", Values({16, 32}), "");
    manager.AddParameter(kernel, "WORK_GROUP_SIZE_Y", Values({1, 2, 4, 8}), "");
    manager.AddParameter(kernel, "WORK_GROUP_SIZE_Z", Values({1}), "");
    manager.AddParameter(kernel, "Z_ITERATIONS", Values({1, 2, 4, 8, 16, 32}), "");
    manager.AddParameter(kernel, "INNER_UNROLL_FACTOR", Values({0, 1, 2, 4, 8, 16, 32}), "");
    if (openCl)
    {
        manager.AddParameter(kernel, "USE_CONSTANT_MEMORY", Values({0, 1}), "");
        manager.AddParameter(kernel, "USE_SOA", Values({0, 1}), "");
        manager.AddParameter(kernel, "VECTOR_SIZE", Values({1, 2, 4, 8, 16}), "");
        manager.AddConstraint(kernel, {"USE_SOA", "VECTOR_SIZE"}, [](const std::vector<uint64_t>& v) { return v[0] || v[1] == 1; });
    }
    else
    {
        manager.AddParameter(kernel, "USE_CONSTANT_MEMORY", Values({0}), "");
        manager.AddParameter(kernel, "USE_SOA", Values({0, 1}), "");
        manager.AddParameter(kernel, "VECTOR_SIZE", Values({1}), "");
    }
    manager.AddConstraint(kernel, {"INNER_UNROLL_FACTOR", "Z_ITERATIONS"}, [](const std::vector<uint64_t>& v) { return v[0] < v[1]; });
    manager.AddConstraint(kernel, {"WORK_GROUP_SIZE_X", "WORK_GROUP_SIZE_Y"},
        [](const std::vector<uint64_t>& v) { return v[0] * v[1] >= 64; });
    return kernel;
}

/** Canonical text of a configuration: pairs sorted by name. */
std::string Canonical(const KernelConfiguration& configuration)
{
    std::map<std::string, std::string> sorted;
    for (const auto& pair : configuration.GetPairs())
    {
        sorted[pair.GetName()] = pair.GetValueString();
    }
    std::string text;
    for (const auto& [name, value] : sorted)
    {
        text += name + "=" + value + ";";
    }
    return text;
}

std::vector<KernelResult> LoadFixture(const std::string& path)
{
    std::ifstream stream(path);
    std::stringstream buffer;
    buffer << stream.rdbuf();
    UserData data;
    JsonValue metadata;
    TimeConfiguration::GetInstance().SetTimeUnit(TimeUnit::Microseconds);
    auto results = DeserializeResults(buffer.str(), data, &metadata);
    CHECK(metadata.At("TimeUnit").AsString() == "Microseconds");
    return results;
}

/** Simulated tuning as in TuningRunner::SimulateTuning: every configuration the searcher asks
 *  for is answered with the stored result. Returns the explored configurations in order. */
std::vector<std::string> Simulate(ConfigurationManager& configurations, const Kernel& kernel, const std::vector<KernelResult>& results,
    StopCondition* stop, std::vector<double>* bestSoFar = nullptr)
{
    std::map<std::string, const KernelResult*> byConfiguration;
    for (const auto& result : results)
    {
        byConfiguration[Canonical(result.GetConfiguration())] = &result;
    }
    configurations.InitializeData(kernel);
    if (stop != nullptr)
    {
        stop->Initialize(configurations.GetTotalConfigurationsCount(kernel.GetId()));
    }
    std::vector<std::string> order;
    double best = 1.0e300;
    while (!configurations.IsDataProcessed(kernel.GetId()))
    {
        if (stop != nullptr && stop->IsFulfilled())
        {
            break;
        }
        const std::string key = Canonical(configurations.GetCurrentConfiguration(kernel.GetId()));
        const auto found = byConfiguration.find(key);
        CHECK(found != byConfiguration.end());
        if (found == byConfiguration.end())
        {
            break;
        }
        order.push_back(key);
        best = std::min(best, static_cast<double>(found->second->GetTotalDuration()));
        if (bestSoFar != nullptr)
        {
            bestSoFar->push_back(best);
        }
        configurations.CalculateNextConfiguration(kernel.GetId(), *found->second);
        if (stop != nullptr)
        {
            stop->Update(*found->second);
        }
    }
    return order;
}

void TestPrefix()
{
    // Tests/KernelManagerTests.cpp:64-81: "#define NAME VALUE\n" per pair, in the pairs' order.
    const KernelConfiguration configuration({ParameterPair("param_two", static_cast<uint64_t>(5)),
        ParameterPair("param_one", static_cast<uint64_t>(2))});
    CHECK(configuration.GeneratePrefix() == "#define param_two 5\n#define param_one 2\n");
}

void TestSpaceSizes()
{
    KernelArgumentManager arguments;
    KernelManager manager(arguments);
    const KernelId cuda = AddReferenceCoulomb3dSpace(manager, false);
    const KernelId openCl = AddReferenceCoulomb3dSpace(manager, true);
    ConfigurationManager configurations;
    configurations.SetSearcher(cuda, std::make_unique<DeterministicSearcher>());
    configurations.SetSearcher(openCl, std::make_unique<DeterministicSearcher>());
    configurations.InitializeData(manager.GetKernel(cuda));
    configurations.InitializeData(manager.GetKernel(openCl));
    // 5 work-group shapes x 21 (unroll < z-iterations) pairs x 2 layouts; the OpenCL variant
    // adds constant memory (x2) and 6 (layout, vector) pairs instead of 2.
    CHECK(configurations.GetTotalConfigurationsCount(cuda) == 210);
    CHECK(configurations.GetTotalConfigurationsCount(openCl) == 1260);
}

void TestFixtureSpaceAndBest(const std::vector<KernelResult>& results)
{
    CHECK(results.size() == 210);
    KernelArgumentManager arguments;
    KernelManager manager(arguments);
    const KernelId kernel = AddReferenceCoulomb3dSpace(manager, false);

    // The dump holds exactly the space this tuner generates from the example's definitions.
    std::set<std::string> dumped;
    for (const auto& result : results)
    {
        dumped.insert(Canonical(result.GetConfiguration()));
        // KernelResult.cpp:194-198: total = extra + sum of kernel durations.
        Nanoseconds kernels = 0;
        for (const auto& computation : result.GetResults())
        {
            kernels += computation.GetDuration();
        }
        CHECK(result.GetTotalDuration() == result.GetExtraDuration() + kernels);
        CHECK(result.IsValid());
    }
    CHECK(dumped.size() == 210);

    ConfigurationManager configurations;
    configurations.SetSearcher(kernel, std::make_unique<DeterministicSearcher>());
    const auto order = Simulate(configurations, manager.GetKernel(kernel), results, nullptr);
    CHECK(order.size() == 210);
    CHECK(std::set<std::string>(order.begin(), order.end()) == dumped);
    const KernelConfiguration best = configurations.GetBestConfiguration(kernel);
    CHECK(Canonical(best) == "INNER_UNROLL_FACTOR=0;USE_CONSTANT_MEMORY=0;USE_SOA=0;VECTOR_SIZE=1;WORK_GROUP_SIZE_X=32;"
        "WORK_GROUP_SIZE_Y=2;WORK_GROUP_SIZE_Z=1;Z_ITERATIONS=16;");
    const KernelResult* bestResult = nullptr;
    for (const auto& result : results)
    {
        if (bestResult == nullptr || result.GetTotalDuration() < bestResult->GetTotalDuration())
        {
            bestResult = &result;
        }
    }
    // 769.696 us, the published best (coulomb_2080_full_search_space.json).
    CHECK(bestResult->GetTotalDuration() == 769696);
    CHECK(Canonical(bestResult->GetConfiguration()) == Canonical(best));
}

void TestRandomAndMcmc(const std::vector<KernelResult>& results)
{
    KernelArgumentManager arguments;
    KernelManager manager(arguments);
    const KernelId kernel = AddReferenceCoulomb3dSpace(manager, false);
    const Kernel& definition = manager.GetKernel(kernel);

    // Random: no revisits, full coverage, same best; seeded runs repeat.
    std::vector<std::vector<std::string>> orders;
    for (int run = 0; run < 2; ++run)
    {
        ConfigurationManager configurations;
        configurations.SetSearcher(kernel, std::make_unique<RandomSearcher>());
        orders.push_back(Simulate(configurations, definition, results, nullptr));
        CHECK(orders.back().size() == 210);
        CHECK(std::set<std::string>(orders.back().begin(), orders.back().end()).size() == 210);
        CHECK(Canonical(configurations.GetBestConfiguration(kernel)).find("Z_ITERATIONS=16;") != std::string::npos);
    }
    CHECK(orders[0] == orders[1]);

    // MCMC: no revisits; the best-so-far duration never increases; coverage when unbounded.
    for (const uint64_t seed : {3ull, 11ull})
    {
        ConfigurationManager configurations;
        configurations.SetSearcher(kernel, std::make_unique<McmcSearcher>(KernelConfiguration(), seed));
        std::vector<double> best;
        const auto order = Simulate(configurations, definition, results, nullptr, &best);
        CHECK(order.size() == 210);
        CHECK(std::set<std::string>(order.begin(), order.end()).size() == order.size());
        for (size_t i = 1; i < best.size(); ++i)
        {
            CHECK(best[i] <= best[i - 1]);
        }
        CHECK(best.back() == 769696.0);
    }
    CHECK(McmcSearcher::MaximumDifferences == 2 && McmcSearcher::BootIterations == 10 && McmcSearcher::EscapeProbability == 0.02);
}

void TestStopConditions(const std::vector<KernelResult>& results)
{
    KernelArgumentManager arguments;
    KernelManager manager(arguments);
    const KernelId kernel = AddReferenceCoulomb3dSpace(manager, false);
    {
        ConfigurationManager configurations;
        configurations.SetSearcher(kernel, std::make_unique<DeterministicSearcher>());
        ConfigurationCount count(50);
        CHECK(Simulate(configurations, manager.GetKernel(kernel), results, &count).size() == 50);
    }
    {
        ConfigurationManager configurations;
        configurations.SetSearcher(kernel, std::make_unique<DeterministicSearcher>());
        ConfigurationFraction fraction(0.1);
        CHECK(Simulate(configurations, manager.GetKernel(kernel), results, &fraction).size() == 21);
    }
    {
        // The target is in milliseconds: stops at the first configuration of at most 1 ms.
        ConfigurationManager configurations;
        configurations.SetSearcher(kernel, std::make_unique<DeterministicSearcher>());
        ConfigurationDuration duration(1.0);
        const auto order = Simulate(configurations, manager.GetKernel(kernel), results, &duration);
        CHECK(!order.empty() && order.size() < 210);
        size_t fast = 0;
        for (const auto& result : results)
        {
            fast += result.GetTotalDuration() <= 1000000 ? 1 : 0;
        }
        CHECK(fast > 0);
        ConfigurationManager again;
        again.SetSearcher(kernel, std::make_unique<DeterministicSearcher>());
        ConfigurationDuration never(0.5);
        CHECK(Simulate(again, manager.GetKernel(kernel), results, &never).size() == 210);
    }
}

void TestJsonRoundTrip(const std::vector<KernelResult>& results)
{
    JsonValue metadata = JsonValue::MakeObject();
    metadata["TimeUnit"] = "Microseconds";
    const std::string text = SerializeResults(metadata, results, {{"grid_side", "256"}});
    UserData data;
    const auto again = DeserializeResults(text, data);
    CHECK(data["grid_side"] == "256");
    CHECK(again.size() == results.size());
    for (size_t i = 0; i < again.size() && i < results.size(); ++i)
    {
        CHECK(Canonical(again[i].GetConfiguration()) == Canonical(results[i].GetConfiguration()));
        CHECK(again[i].GetTotalDuration() == results[i].GetTotalDuration());
        CHECK(again[i].GetStatus() == results[i].GetStatus());
        CHECK(again[i].GetResults().size() == results[i].GetResults().size());
    }
}

} // namespace

int main(int argc, char** argv)
{
    if (argc != 2)
    {
        std::fprintf(stderr, "usage: %s <coulomb3d_2080_space.json>\n", argv[0]);
        return 2;
    }
    setenv("KTT_SEED", "7", 1);
    Logger::GetLogger().SetLoggingLevel(LoggingLevel::Off);
    TestPrefix();
    TestSpaceSizes();
    const auto results = LoadFixture(argv[1]);
    TestFixtureSpaceAndBest(results);
    TestRandomAndMcmc(results);
    TestStopConditions(results);
    TestJsonRoundTrip(results);
    std::printf("%d checks, %d failures\n", g_Checks, g_Failures);
    return g_Failures == 0 ? 0 : 1;
}
