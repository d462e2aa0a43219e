// Host-side tests of the ktt core (no device needed): configuration prefix, configuration
// spaces of the reference's Coulomb examples, the results loader on the reference's own RTX
// 2080 Ti full-search dump, simulated tuning with the three searchers, stop conditions, the
// JSON round trip, and the profile-based searcher's bottleneck analysis and candidate scoring.
//
// Usage: ktt_core_tests <path to tests/golden/coulomb3d_2080_space.json>
// Exit status 0 when every check passes; each failure is printed with its line.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include <Ktt.h>

#include "Common.h"
#include "Json.h"
#include "Kernel.h"
#include "KernelArgument.h"
#include "TunerCore.h"
#include "Exploration.h"
#include "CoulombBatchPlan.h"

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
KernelId AddReferenceCoulomb3dSpace(KernelRegistry& registry, const bool openCl)
{
    const KernelDefinitionId definition = registry.AddDefinition(openCl ? "directCoulombSumCl" : "directCoulombSum", "",
        DimensionVector(256, 256, 256), DimensionVector(1, 1, 1), {});
    const KernelId id = registry.AddKernel(openCl ? "CoulombSumCl" : "CoulombSum", {definition});
    Kernel& kernel = registry.GetKernel(id);
    const auto parameter = [&kernel](const std::string& name, const std::vector<uint64_t>& values)
    {
        kernel.AddParameter(KernelParameter(name, Values(values), ""));
    };
    parameter("WORK_GROUP_SIZE_X", {16, 32});
    parameter("WORK_GROUP_SIZE_Y", {1, 2, 4, 8});
    parameter("WORK_GROUP_SIZE_Z", {1});
    parameter("Z_ITERATIONS", {1, 2, 4, 8, 16, 32});
    parameter("INNER_UNROLL_FACTOR", {0, 1, 2, 4, 8, 16, 32});
    if (openCl)
    {
        parameter("USE_CONSTANT_MEMORY", {0, 1});
        parameter("USE_SOA", {0, 1});
        parameter("VECTOR_SIZE", {1, 2, 4, 8, 16});
        kernel.AddConstraint(KernelConstraint({"USE_SOA", "VECTOR_SIZE"},
            ConstraintFunction([](const std::vector<uint64_t>& v) { return v[0] || v[1] == 1; })));
    }
    else
    {
        parameter("USE_CONSTANT_MEMORY", {0});
        parameter("USE_SOA", {0, 1});
        parameter("VECTOR_SIZE", {1});
    }
    kernel.AddConstraint(KernelConstraint({"INNER_UNROLL_FACTOR", "Z_ITERATIONS"},
        ConstraintFunction([](const std::vector<uint64_t>& v) { return v[0] < v[1]; })));
    kernel.AddConstraint(KernelConstraint({"WORK_GROUP_SIZE_X", "WORK_GROUP_SIZE_Y"},
        ConstraintFunction([](const std::vector<uint64_t>& v) { return v[0] * v[1] >= 64; })));
    return id;
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

/** Simulated tuning over saved results (ReplayResults, the engine-free core of
 *  Tuner::SimulateTuning) with a fresh exploration of the kernel's space. Returns the explored
 *  configurations in order; `best` receives the exploration's best configuration. */
std::vector<std::string> Simulate(const Kernel& kernel, Searcher& searcher, const std::vector<KernelResult>& results,
    StopCondition* stop, std::vector<double>* bestSoFar = nullptr, KernelConfiguration* best = nullptr)
{
    Exploration exploration(kernel, searcher);
    const std::vector<KernelResult> replayed = ReplayResults(exploration, kernel, results, stop);
    std::vector<std::string> order;
    double fastest = 1.0e300;
    for (const auto& result : replayed)
    {
        CHECK(result.IsValid());
        order.push_back(Canonical(result.GetConfiguration()));
        fastest = std::min(fastest, static_cast<double>(result.GetTotalDuration()));
        if (bestSoFar != nullptr)
        {
            bestSoFar->push_back(fastest);
        }
    }
    if (best != nullptr)
    {
        *best = exploration.Best();
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
    ArgumentRegistry arguments;
    KernelRegistry manager(arguments);
    const KernelId cuda = AddReferenceCoulomb3dSpace(manager, false);
    const KernelId openCl = AddReferenceCoulomb3dSpace(manager, true);
    DeterministicSearcher first;
    DeterministicSearcher second;
    const Exploration cudaSpace(manager.GetKernel(cuda), first);
    const Exploration openClSpace(manager.GetKernel(openCl), second);
    // 5 work-group shapes x 21 (unroll < z-iterations) pairs x 2 layouts; the OpenCL variant
    // adds constant memory (x2) and 6 (layout, vector) pairs instead of 2.
    CHECK(cudaSpace.Size() == 210);
    CHECK(openClSpace.Size() == 1260);
    // Index and configuration map onto each other.
    for (uint64_t index = 0; index < cudaSpace.Size(); index += 7)
    {
        CHECK(cudaSpace.IndexOf(cudaSpace.ConfigurationAt(index)) == index);
    }
}

/** Launch sizes a searcher sees: the first definition's global x local size after the thread
 *  modifiers; after a replay, the sizes the saved results recorded. */
void TestLaunchSizes(const std::vector<KernelResult>& results)
{
    ArgumentRegistry arguments;
    KernelRegistry manager(arguments);
    Kernel& kernel = manager.GetKernel(AddReferenceCoulomb3dSpace(manager, false));
    const KernelDefinitionId definition = kernel.GetDefinitions().front()->GetId();
    kernel.AddThreadModifier(ModifierType::Local, ModifierDimension::X, {"WORK_GROUP_SIZE_X"}, {definition},
        [](const uint64_t size, const std::vector<uint64_t>& v) { return size * v[0]; });
    kernel.AddThreadModifier(ModifierType::Global, ModifierDimension::Z, {"Z_ITERATIONS"}, {definition},
        [](const uint64_t size, const std::vector<uint64_t>& v) { return size / v[0]; });
    DeterministicSearcher searcher;
    Exploration exploration(kernel, searcher);
    for (uint64_t index = 0; index < exploration.Size(); index += 11)
    {
        const KernelConfiguration configuration = exploration.ConfigurationAt(index);
        const auto& pairs = configuration.GetPairs();
        const double wgx = static_cast<double>(ParameterPair::GetParameterValue<uint64_t>(pairs, "WORK_GROUP_SIZE_X"));
        const double z = static_cast<double>(ParameterPair::GetParameterValue<uint64_t>(pairs, "Z_ITERATIONS"));
        double threads = 0.0;
        CHECK(searcher.GetLaunchSize(configuration, threads));
        CHECK(threads == 256.0 * 256.0 * (256.0 / z) * wgx);
    }
    exploration.SetLaunchSizes({{0, 1234.0}});
    double threads = 0.0;
    CHECK(searcher.GetLaunchSize(exploration.ConfigurationAt(0), threads) && threads == 1234.0);
    CHECK(!searcher.GetLaunchSize(exploration.ConfigurationAt(1), threads));

    // A replay takes the sizes of the saved runs, whatever the kernel's own sizes.
    KernelRegistry replayManager(arguments);
    const Kernel& replayKernel = replayManager.GetKernel(AddReferenceCoulomb3dSpace(replayManager, false));
    DeterministicSearcher replaySearcher;
    Exploration replay(replayKernel, replaySearcher);
    ReplayResults(replay, replayKernel, results, nullptr);
    for (size_t i = 0; i < results.size(); i += 13)
    {
        const ComputationResult& first = results[i].GetResults().front();
        const double recorded = static_cast<double>(first.GetGlobalSize().GetTotalSize()) * static_cast<double>(first.GetLocalSize().GetTotalSize());
        CHECK(replaySearcher.GetLaunchSize(results[i].GetConfiguration(), threads) && threads == recorded);
    }
}

void TestFixtureSpaceAndBest(const std::vector<KernelResult>& results)
{
    CHECK(results.size() == 210);
    ArgumentRegistry arguments;
    KernelRegistry manager(arguments);
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

    DeterministicSearcher searcher;
    KernelConfiguration best;
    const auto order = Simulate(manager.GetKernel(kernel), searcher, results, nullptr, nullptr, &best);
    CHECK(order.size() == 210);
    CHECK(std::set<std::string>(order.begin(), order.end()) == dumped);
    CHECK(Canonical(best) == "INNER_UNROLL_FACTOR=0;USE_CONSTANT_MEMORY=0;USE_SOA=0;VECTOR_SIZE=1;WORK_GROUP_SIZE_X=32;"
        "WORK_GROUP_SIZE_Y=2;WORK_GROUP_SIZE_Z=1;Z_ITERATIONS=16;");
    const KernelResult* fastest = nullptr;
    for (const auto& result : results)
    {
        if (fastest == nullptr || result.GetTotalDuration() < fastest->GetTotalDuration())
        {
            fastest = &result;
        }
    }
    // 769.696 us, the published best (coulomb_2080_full_search_space.json).
    CHECK(fastest->GetTotalDuration() == 769696);
    CHECK(Canonical(fastest->GetConfiguration()) == Canonical(best));
}

void TestRandomAndMcmc(const std::vector<KernelResult>& results)
{
    ArgumentRegistry arguments;
    KernelRegistry manager(arguments);
    const KernelId kernel = AddReferenceCoulomb3dSpace(manager, false);
    const Kernel& definition = manager.GetKernel(kernel);

    // Random: no revisits, full coverage, same best; seeded runs repeat.
    std::vector<std::vector<std::string>> orders;
    for (int run = 0; run < 2; ++run)
    {
        RandomSearcher searcher;
        KernelConfiguration best;
        orders.push_back(Simulate(definition, searcher, results, nullptr, nullptr, &best));
        CHECK(orders.back().size() == 210);
        CHECK(std::set<std::string>(orders.back().begin(), orders.back().end()).size() == 210);
        CHECK(Canonical(best).find("Z_ITERATIONS=16;") != std::string::npos);
    }
    CHECK(orders[0] == orders[1]);

    // MCMC: no revisits; the best-so-far duration never increases; coverage when unbounded;
    // the same seed gives the same walk.
    std::vector<std::vector<std::string>> walks;
    for (const uint64_t seed : {3ull, 11ull, 3ull})
    {
        McmcSearcher searcher(KernelConfiguration(), seed);
        std::vector<double> best;
        const auto order = Simulate(definition, searcher, results, nullptr, &best);
        CHECK(order.size() == 210);
        CHECK(std::set<std::string>(order.begin(), order.end()).size() == order.size());
        for (size_t i = 1; i < best.size(); ++i)
        {
            CHECK(best[i] <= best[i - 1]);
        }
        CHECK(best.back() == 769696.0);
        walks.push_back(order);
    }
    CHECK(walks[0] == walks[2]);
    CHECK(walks[0] != walks[1]);
    CHECK(McmcSearcher::MaximumDifferences == 2 && McmcSearcher::BootIterations == 10 && McmcSearcher::EscapeProbability == 0.02);

    // After the 10 warm-up states the walk moves to neighbours of its state: with a known start
    // and no warm-up, the second proposal differs from the start in at most 2 parameters.
    {
        DeterministicSearcher probe;
        const Exploration space(definition, probe);
        const KernelConfiguration start = space.ConfigurationAt(100);
        McmcSearcher searcher(start, 5);
        Exploration walk(definition, searcher);
        CHECK(Canonical(walk.Proposal()) == Canonical(start));
        KernelResult first(definition.GetName(), start);
        walk.Record(first);
        const KernelConfiguration next = walk.Proposal();
        size_t differences = 0;
        for (const auto& pair : next.GetPairs())
        {
            for (const auto& other : start.GetPairs())
            {
                differences += pair.GetName() == other.GetName() && !pair.HasSameValue(other) ? 1 : 0;
            }
        }
        CHECK(differences >= 1 && differences <= 2);
    }
}

void TestStopConditions(const std::vector<KernelResult>& results)
{
    ArgumentRegistry arguments;
    KernelRegistry manager(arguments);
    const Kernel& kernel = manager.GetKernel(AddReferenceCoulomb3dSpace(manager, false));
    {
        DeterministicSearcher searcher;
        ConfigurationCount count(50);
        CHECK(Simulate(kernel, searcher, results, &count).size() == 50);
    }
    {
        DeterministicSearcher searcher;
        ConfigurationFraction fraction(0.1);
        CHECK(Simulate(kernel, searcher, results, &fraction).size() == 21);
    }
    {
        // The target is in milliseconds: stops at the first configuration of at most 1 ms.
        DeterministicSearcher searcher;
        ConfigurationDuration duration(1.0);
        const auto order = Simulate(kernel, searcher, results, &duration);
        CHECK(!order.empty() && order.size() < 210);
        DeterministicSearcher again;
        ConfigurationDuration never(0.5);
        CHECK(Simulate(kernel, again, results, &never).size() == 210);
    }
    {
        DeterministicSearcher searcher;
        auto count = std::make_shared<ConfigurationCount>(30);
        auto fraction = std::make_shared<ConfigurationFraction>(0.5);
        UnionCondition either({count, fraction});
        CHECK(Simulate(kernel, searcher, results, &either).size() == 30);
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

bool Near(const double a, const double b)
{
    return std::fabs(a - b) <= 1e-12 * std::max(1.0, std::fabs(b));
}

/** Bottleneck analysis, counter changes and candidate scoring of the profile-based searcher on
 *  hand-made counter sets whose results follow by hand. */
void TestProfileSearcherAnalysis()
{
    ProfileSearcherOptions options;
    options.computeUnits = 68;

    // CUDA metrics: 60 % DRAM throughput split 3:1 between reads and writes, L2 40 % split 1:1.
    std::map<std::string, double> cuda{{"dram__throughput.avg.pct_of_peak_sustained_elapsed", 60.0},
        {"dram__sectors_read.sum", 300.0}, {"dram__sectors_write.sum", 100.0},
        {"lts__t_sectors.avg.pct_of_peak_sustained_elapsed", 40.0}, {"lts__t_sectors_op_read.sum", 50.0},
        {"lts__t_sectors_op_write.sum", 50.0}, {"sm__warps_active.avg.pct_of_peak_sustained_active", 25.0},
        {"smsp__cycles_active.avg.pct_of_peak_sustained_elapsed", 90.0}, {"smsp__issue_active.avg.pct_of_peak_sustained_active", 40.0},
        {"smsp__inst_executed.sum", 100.0}, {"smsp__thread_inst_executed_per_inst_executed.ratio", 32.0},
        {"smsp__thread_inst_executed_per_inst_executed.pct", 100.0}, {"smsp__sass_thread_inst_executed_op_fp32_pred_on.sum", 3200.0},
        {"Global size", 68.0 * 64.0 * 5.0 / 2.0}, {"Local size", 256.0}};
    auto bn = ProfileBasedSearcher::AnalyzeBottlenecks(cuda, options);
    CHECK(Near(bn["bnDRAMRead"], 0.45));
    CHECK(Near(bn["bnDRAMWrite"], 0.15));
    CHECK(Near(bn["bnL2Read"], 0.2));
    CHECK(Near(bn["bnL2Write"], 0.2));
    CHECK(Near(bn["bnMPparal"], 0.75));
    CHECK(Near(bn["bnGparal"], 0.1));
    CHECK(Near(bn["bnThreads"], 0.5));
    // fp32 is every executed thread instruction and issue is at 80 % of its dual-issue ceiling.
    CHECK(Near(bn["bnSP"], 0.8));
    CHECK(Near(bn["bnInstIssue"], 0.6 * (1.0 - 0.7) / 0.3));
    auto changes = ProfileBasedSearcher::ComputeChanges(bn, cuda, options);
    CHECK(Near(changes["dram__sectors_read.sum"], -0.45));
    CHECK(Near(changes["smsp__sass_thread_inst_executed_op_fp32_pred_on.sum"], -0.8));
    CHECK(Near(changes["Global size"], 0.5));

    // CDNA counters: 4 GiB of HBM traffic (3:1) in 1 ms is 4.295e12 B/s, 53.7 % of 8 TB/s.
    std::map<std::string, double> cdna{{"FETCH_SIZE", 3.0 * 1024 * 1024}, {"WRITE_SIZE", 1.0 * 1024 * 1024}, {"VALUBusy", 90.0},
        {"F32_INSTS", 600.0}, {"SQ_INSTS_VALU", 1000.0}, {"SQ_INSTS_VALU_TRANS_F32", 600.0}, {"OccupancyPercent", 40.0},
        {"SIMD_UTILIZATION", 0.95}, {"LdsUtil", 30.0}, {"kernel_duration_ns", 1.0e6}, {"Global size", 1.0e6}, {"Local size", 256.0}};
    options.computeUnits = 256;
    bn = ProfileBasedSearcher::AnalyzeBottlenecks(cdna, options);
    const double hbm = 4.0 * 1024 * 1024 * 1024 / 1.0e-3 / 8.0e12;
    CHECK(Near(bn["bnDRAMRead"], 0.75 * hbm));
    CHECK(Near(bn["bnDRAMWrite"], 0.25 * hbm));
    CHECK(Near(bn["bnSP"], 0.6 * 0.9));
    // 600 transcendentals of 1,000 VALU instructions, each holding 3.2 issue slots: 1,920 of 2,320
    // slots, times the 0.9 VALU busy fraction.
    const double transShare = 3.2 * 600.0 / (3.2 * 600.0 + 400.0);
    CHECK(Near(bn["bnSFU"], transShare * 0.9));
    CHECK(Near(bn["bnMPparal"], 0.6));
    CHECK(Near(bn["bnGparal"], 0.05));
    CHECK(Near(bn["bnSMRead"], 0.3));
    // Every instruction is a VALU one: the VALU class (share 1) is above the 0.7 threshold, so the
    // unused 0.1 of its issue counts fully; the busy VALU asks for fewer weighted VALU slots, and the
    // transcendentals (0.828 of the issue time) for fewer transcendentals.
    CHECK(Near(bn["bnVALU"], 0.9));
    CHECK(Near(bn["bnInstIssue"], 0.1 * (1.0 - 0.7) / 0.3));
    changes = ProfileBasedSearcher::ComputeChanges(bn, cdna, options);
    CHECK(Near(changes["SQ_INSTS_VALU_TRANS_F32"], -(transShare * 0.9 - 0.7) / 0.3));
    CHECK(Near(changes["VALU_ISSUE_SLOTS"], -(0.9 - 0.7) / 0.3));
    CHECK(changes.count("F32_INSTS") == 0);  // 0.54 is below the threshold
    CHECK(changes.count("SQ_ACTIVE_INST_VALU2") == 0);  // no dual-issue counters, no pairing class

    // With the dual-issue counters the issue time is measured: 1,000 VALU instructions of which
    // 300 transcendentals took 1,200 quad-cycles. The transcendentals held 600 of them; 700 plain
    // instructions would take 350 if all paired, so 250 quad-cycles went to unpaired issue.
    std::map<std::string, double> measured = cdna;
    measured["SQ_INSTS_VALU_TRANS_F32"] = 300.0;
    measured["VALU_QUAD_CYCLES"] = 1200.0;
    measured["SQ_ACTIVE_INST_VALU2"] = 100.0;
    bn = ProfileBasedSearcher::AnalyzeBottlenecks(measured, options);
    CHECK(Near(bn["bnSFU"], 600.0 / 1200.0 * 0.9));
    CHECK(Near(bn["bnPairing"], 250.0 / 1200.0 * 0.9));
    changes = ProfileBasedSearcher::ComputeChanges(bn, measured, options);
    CHECK(Near(changes["VALU_QUAD_CYCLES"], -(0.9 - 0.7) / 0.3));
    CHECK(changes.count("VALU_ISSUE_SLOTS") == 0);
    CHECK(Near(changes["SQ_ACTIVE_INST_VALU2"], 250.0 / 1200.0 * 0.9));
    CHECK(changes.count("SQ_INSTS_VALU_TRANS_F32") == 0);  // 0.45 is below the threshold

    // Scoring: one counter that should shrink. The candidate that halves it scores +1/3, the
    // one that doubles it -1/3 (below the -0.25 cut-off), one without counters 0.
    std::map<std::string, double> want{{"A", -1.0}};
    std::map<std::string, double> current{{"A", 10.0}};
    std::map<std::string, double> smaller{{"A", 5.0}};
    std::map<std::string, double> larger{{"A", 20.0}};
    std::map<std::string, double> slightlyLarger{{"A", 11.0}};
    auto weights = ProfileBasedSearcher::ScoreCandidates(want, current, {&smaller, &larger, nullptr}, options);
    CHECK(weights.size() == 3 && Near(weights[0], 256.0) && Near(weights[1], 0.0001) && Near(weights[2], 0.0001));
    // Above the cut-off a negative score maps to 1 - score / lowest before the exponent.
    weights = ProfileBasedSearcher::ScoreCandidates(want, current, {&smaller, &slightlyLarger, &larger}, options);
    const double mild = -1.0 / 21.0;
    CHECK(Near(weights[1], std::pow(1.0 - mild / (-1.0 / 3.0), 8.0)));
}

} // namespace

/** The batch work list (CoulombBatchPlan.h) on a 1,024-system batch shaped like the benchmark's
 *  (atom counts log-uniform in [200, 20000], sides 32/48/64/96/128, the built-in bucket shapes):
 *  the item count is the tile count the device list is sized for, every (system, tile) appears
 *  exactly once with in-range origins, the order is costliest first, and a list one item longer
 *  than the capacity is refused. A wrong count here is what once let the persistent kernel read
 *  past its list. */
void TestBatchWorkList()
{
    using namespace ktt::batch;
    const int sides[5] = {32, 48, 64, 96, 128};
    const std::vector<BucketShape> shapes = {{16, 2, 4, 64, 8}, {16, 3, 4, 64, 8}, {32, 2, 8, 64, 8}, {32, 3, 8, 128, 8},
        {64, 2, 8, 128, 8}};
    std::vector<HostSystem> systems;
    uint64_t state = 12345;
    const auto next = [&state]()
    {
        state = state * 6364136223846793005ull + 1442695040888963407ull;
        return static_cast<double>(state >> 11) * (1.0 / 9007199254740992.0);
    };
    int atomOffset = 0;
    long long gridOffset = 0;
    for (int s = 0; s < 1024; ++s)
    {
        HostSystem system{};
        system.bucket = static_cast<int>(next() * 5.0) % 5;
        system.gridSize = sides[system.bucket];
        system.atomCount = static_cast<int>(std::lround(std::exp(std::log(200.0) + next() * (std::log(20000.0) - std::log(200.0)))));
        system.atomOffset = atomOffset;
        system.spacing = 0.5f;
        system.gridOffset = gridOffset;
        atomOffset += system.atomCount;
        gridOffset += static_cast<long long>(system.gridSize) * system.gridSize * system.gridSize;
        systems.push_back(system);
    }
    // A system with no atoms and one whose side is not a multiple of its bucket's span.
    systems.push_back({atomOffset, 0, 32, 0, 0.5f, 0, gridOffset});
    systems.push_back({atomOffset, 7, 40, 1, 0.5f, 0, gridOffset + 32 * 32 * 32});

    const std::vector<int> flat = BuildWorkItems(systems, shapes);
    const size_t count = CountWorkItems(systems, shapes);
    CHECK(flat.size() == 4 * count);
    size_t expected = 0;
    std::set<std::tuple<int, int, int>> seen;
    bool inRange = true;
    bool ordered = true;
    long long previousCost = -1;
    for (size_t k = 0; k < flat.size() / 4; ++k)
    {
        const int s = flat[4 * k];
        const int x = flat[4 * k + 1];
        const int row = flat[4 * k + 2];
        if (s < 0 || s >= static_cast<int>(systems.size()) || flat[4 * k + 3] != 0)
        {
            inRange = false;
            continue;
        }
        const HostSystem& system = systems[static_cast<size_t>(s)];
        const BucketShape& shape = shapes[static_cast<size_t>(system.bucket)];
        const long long rows = static_cast<long long>(system.gridSize) * system.gridSize;
        inRange = inRange && x >= 0 && x < system.gridSize && x % shape.Span() == 0 && row >= 0 && row < rows
            && row % shape.RowsPerGroup() == 0;
        seen.insert({s, x, row});
        const long long width = std::min<long long>(shape.Span(), system.gridSize - x);
        const long long height = std::min<long long>(shape.RowsPerGroup(), rows - row);
        const long long cost = static_cast<long long>(shape.Span()) * shape.RowsPerGroup()
            * ((system.atomCount + shape.tile - 1) / shape.tile) * shape.tile + width * height;
        ordered = ordered && (previousCost < 0 || cost <= previousCost);
        previousCost = cost;
    }
    for (const auto& system : systems)
    {
        const BucketShape& shape = shapes[static_cast<size_t>(system.bucket)];
        const size_t g = static_cast<size_t>(system.gridSize);
        expected += ((g + shape.Span() - 1) / shape.Span()) * ((g * g + shape.RowsPerGroup() - 1) / shape.RowsPerGroup());
    }
    CHECK(inRange);
    CHECK(ordered);
    CHECK(count == expected);
    CHECK(seen.size() == count);  // no (system, tile) twice, so all `expected` tiles appear
    CHECK(count > 100000);
    bool accepted = true;
    try
    {
        CheckItemCapacity(flat, count);
    }
    catch (const KttException&)
    {
        accepted = false;
    }
    CHECK(accepted);
    bool refused = false;
    try
    {
        CheckItemCapacity(flat, count - 1);
    }
    catch (const KttException&)
    {
        refused = true;
    }
    CHECK(refused);
    bool ragged = false;
    try
    {
        CheckItemCapacity(std::vector<int>(flat.begin(), flat.end() - 1), count);
    }
    catch (const KttException&)
    {
        ragged = true;
    }
    CHECK(ragged);
}

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
    TestLaunchSizes(results);
    TestRandomAndMcmc(results);
    TestStopConditions(results);
    TestJsonRoundTrip(results);
    TestProfileSearcherAnalysis();
    TestBatchWorkList();
    std::printf("%d checks, %d failures\n", g_Checks, g_Failures);
    return g_Failures == 0 ? 0 : 1;
}
