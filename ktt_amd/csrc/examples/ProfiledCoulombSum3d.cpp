// CoulombSum3d with hardware counters on the HIP backend: the counterpart of the reference's
// profiling examples (Tutorials/08ProfileBasedSearcher/ProfileSearcherExecutor.py:158-216,
// Examples/CoulombSum3d/CoulombSum3d.cpp:211-236).
//
// Usage: ProfiledCoulombSum3d counters [grid side] [atoms] [counter,counter,...]
//            profiles one configuration, repeating the launch until every counter pass ran,
//            twice, and prints the counters as JSON;
//        ProfiledCoulombSum3d search [max configurations, 0 = all] [grid side] [atoms] [results file]
//            [knowledge base results file]
//            tunes with the profile-based searcher and saves the results (with counters);
//        ProfiledCoulombSum3d profile-all [max configurations, 0 = all] [grid side] [atoms] [results file]
//            runs every configuration in index order with profiling on (every result with counters).
//        A seventh argument "reduced" limits both to the 288 shapes with at most 64 threads in x,
//        tiles of at most 128 atoms, INNER_UNROLL_FACTOR 1, USE_SOA 0 and the interleaved inner loop
//        (ISSUE_PHASES 0); "reduced-phases" takes the same shapes with ISSUE_PHASES 0 and 1 (576).
// Counters are enabled before the tuner initialises the HIP runtime.
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include <Ktt.h>
#include <KttCoulomb.h>
#include <ktt_hip.h>

namespace
{

struct Problem
{
    std::vector<float> atomInfo, atomInfoX, atomInfoY, atomInfoZ, atomInfoW, grid;
};

Problem MakeProblem(const int gridSize, const int atoms, const float spacing)
{
    std::mt19937_64 generator(20240812);
    std::uniform_real_distribution<float> position(0.0f, spacing * static_cast<float>(gridSize));
    std::uniform_real_distribution<float> charge(-1.0f, 1.0f);
    Problem p;
    p.atomInfo.resize(4 * static_cast<size_t>(atoms));
    p.atomInfoX.resize(atoms);
    p.atomInfoY.resize(atoms);
    p.atomInfoZ.resize(atoms);
    p.atomInfoW.resize(atoms);
    for (int i = 0; i < atoms; ++i)
    {
        p.atomInfoX[i] = p.atomInfo[4 * i] = position(generator);
        p.atomInfoY[i] = p.atomInfo[4 * i + 1] = position(generator);
        p.atomInfoZ[i] = p.atomInfo[4 * i + 2] = position(generator);
        p.atomInfoW[i] = p.atomInfo[4 * i + 3] = charge(generator);
    }
    p.grid.assign(static_cast<size_t>(gridSize) * gridSize * gridSize, 0.0f);
    return p;
}

std::vector<std::string> Split(const std::string& text)
{
    std::vector<std::string> parts;
    std::stringstream stream(text);
    std::string part;
    while (std::getline(stream, part, ','))
    {
        if (!part.empty())
        {
            parts.push_back(part);
        }
    }
    return parts;
}

} // namespace

int main(int argc, char** argv)
{
    const std::string mode = argc > 1 ? argv[1] : "counters";
    std::string error;
    if (!ktt::EnableHipCounters(&error))
    {
        std::fprintf(stderr, "hardware counters unavailable: %s\n", error.c_str());
        return 3;
    }
    const bool search = mode == "search" || mode == "profile-all";
    const uint64_t maxConfigurations = search && argc > 2 ? std::stoull(argv[2]) : 30;
    const int base = search ? 3 : 2;
    const int gridSize = argc > base ? std::stoi(argv[base]) : 128;
    const int atoms = argc > base + 1 ? std::stoi(argv[base + 1]) : 1024;
    const float spacing = 0.5f;
    Problem p = MakeProblem(gridSize, atoms, spacing);

    ktt::Tuner tuner(0, 0, ktt::ComputeApi::HIP);
    tuner.SetGlobalSizeType(ktt::GlobalSizeType::CUDA);
    tuner.SetTimeUnit(ktt::TimeUnit::Microseconds);
    // Quiet by default; KTT_LOG_LEVEL=Info prints each configuration as it runs.
    if (std::getenv("KTT_LOG_LEVEL") == nullptr)
    {
        ktt::Tuner::SetLoggingLevel(ktt::LoggingLevel::Warning);
    }
    const std::string directory = ktt_hip_kernel_directory();
    const auto definition = tuner.AddKernelDefinitionFromFile("directCoulombSum", directory + "/CoulombSum3d.hip",
        ktt::DimensionVector(gridSize, gridSize * gridSize, 1), ktt::DimensionVector(1, 1, 1));
    const auto referenceDefinition = tuner.AddKernelDefinitionFromFile("directCoulombSumReference", directory + "/CoulombSum3dReference.hip",
        ktt::DimensionVector((gridSize + 15) / 16, (gridSize + 15) / 16, gridSize), ktt::DimensionVector(16, 16, 1));
    const auto kernel = tuner.CreateSimpleKernel("CoulombSum", definition);
    const auto referenceKernel = tuner.CreateSimpleKernel("CoulombSumReference", referenceDefinition);
    const auto ai = tuner.AddArgumentVector(p.atomInfo, ktt::ArgumentAccessType::ReadOnly);
    const auto ax = tuner.AddArgumentVector(p.atomInfoX, ktt::ArgumentAccessType::ReadOnly);
    const auto ay = tuner.AddArgumentVector(p.atomInfoY, ktt::ArgumentAccessType::ReadOnly);
    const auto az = tuner.AddArgumentVector(p.atomInfoZ, ktt::ArgumentAccessType::ReadOnly);
    const auto aw = tuner.AddArgumentVector(p.atomInfoW, ktt::ArgumentAccessType::ReadOnly);
    const auto n = tuner.AddArgumentScalar(atoms);
    const auto h = tuner.AddArgumentScalar(spacing);
    const auto g = tuner.AddArgumentScalar(gridSize);
    const auto gridArgument = tuner.AddArgumentVector(p.grid, ktt::ArgumentAccessType::ReadWrite);
    ktt::AddCoulomb3dTuningSpace(tuner, kernel, definition);
    tuner.AddParameter(kernel, "USE_SOA", std::vector<uint64_t>{0, 1});
    if (search && argc > 7 && (std::string(argv[7]) == "reduced" || std::string(argv[7]) == "reduced-phases"))
    {
        // The 288 shapes with at most 64 threads in x, tiles of at most 128 atoms, no unroll,
        // float4 atoms and the compiler's interleaved inner loop (of 15,288); with "reduced-phases"
        // each of them also with the phased inner loop.
        const uint64_t phases = std::string(argv[7]) == "reduced-phases" ? 1 : 0;
        tuner.AddConstraint(kernel, {"WORK_GROUP_SIZE_X", "ATOM_TILE", "INNER_UNROLL_FACTOR", "USE_SOA", "ISSUE_PHASES"},
            [phases](const std::vector<uint64_t>& v) { return v[0] <= 64 && v[1] <= 128 && v[2] == 1 && v[3] == 0 && v[4] <= phases; });
    }
    tuner.SetArguments(definition, {ai, ax, ay, az, aw, n, h, g, gridArgument});
    tuner.SetArguments(referenceDefinition, {ai, n, h, g, gridArgument});
    tuner.SetReferenceKernel(gridArgument, referenceKernel, ktt::KernelConfiguration());
    tuner.SetValidationMethod(ktt::ValidationMethod::SideBySideComparison, 0.01);

    if (!search)
    {
        const std::vector<std::string> counters = argc > 4 ? Split(argv[4]) : std::vector<std::string>{};
        if (!counters.empty())
        {
            tuner.SetProfilingCounters(counters);
        }
        tuner.SetProfiling(true);
        const auto configuration = tuner.CreateConfiguration(kernel, {{"WORK_GROUP_SIZE_X", uint64_t(64)}, {"WORK_GROUP_SIZE_Y", uint64_t(4)},
            {"POINTS_X", uint64_t(2)}, {"ROWS_PER_THREAD", uint64_t(16)}, {"ATOM_TILE", uint64_t(64)}, {"INNER_UNROLL_FACTOR", uint64_t(1)},
            {"USE_SOA", uint64_t(0)}});
        std::printf("{\"mode\": \"counters\", \"grid\": %d, \"atoms\": %d, \"collections\": [", gridSize, atoms);
        for (int collection = 0; collection < 2; ++collection)
        {
            std::vector<uint64_t> remaining;
            ktt::KernelResult result;
            do
            {
                result = tuner.Run(kernel, configuration, {});
                if (!result.IsValid() || result.GetResults().empty())
                {
                    std::printf("]}\n");
                    std::fprintf(stderr, "profiled run failed with status %d\n", static_cast<int>(result.GetStatus()));
                    return 1;
                }
                const auto& computation = result.GetResults().front();
                remaining.push_back(computation.HasProfilingData() ? computation.GetProfilingData().GetRemainingProfilingRuns() : 0);
            }
            while (result.HasRemainingProfilingRuns());
            const auto& computation = result.GetResults().front();
            std::printf("%s{\"runs\": %zu, \"remaining\": [", collection == 0 ? "" : ", ", remaining.size());
            for (size_t i = 0; i < remaining.size(); ++i)
            {
                std::printf("%s%llu", i == 0 ? "" : ", ", static_cast<unsigned long long>(remaining[i]));
            }
            const auto global = computation.GetGlobalSize();
            const auto local = computation.GetLocalSize();
            std::printf("], \"groups\": %llu, \"group_threads\": %llu, \"duration_us\": %.3f, \"counters\": {",
                static_cast<unsigned long long>(global.GetTotalSize()), static_cast<unsigned long long>(local.GetTotalSize()),
                static_cast<double>(computation.GetDuration()) * 1.0e-3);
            const auto& list = computation.GetProfilingData().GetCounters();
            for (size_t i = 0; i < list.size(); ++i)
            {
                const auto& counter = list[i];
                const double value = counter.GetType() == ktt::ProfilingCounterType::Double ? counter.GetValueDouble()
                    : static_cast<double>(counter.GetValueUint());
                std::printf("%s\"%s\": %.17g", i == 0 ? "" : ", ", counter.GetName().c_str(), value);
            }
            std::printf("}}");
        }
        std::printf("]}\n");
        return 0;
    }

    const std::string output = argc > 5 ? argv[5] : "ProfiledCoulombSum3dOutput";
    const std::string knowledge = argc > 6 ? argv[6] : "";
    if (mode == "profile-all")
    {
        // Every configuration in index order, each profiled: a knowledge base for the
        // profile-based searcher, or saved results to replay it on.
        tuner.SetSearcher(kernel, std::make_unique<ktt::DeterministicSearcher>());
        tuner.SetProfiling(true);
    }
    else
    {
        tuner.SetProfileBasedSearcher(kernel, knowledge, true, 5, 100, 10);
    }
    std::unique_ptr<ktt::StopCondition> stop;
    if (maxConfigurations > 0)
    {
        stop = std::make_unique<ktt::ConfigurationCount>(maxConfigurations);
    }
    const auto results = tuner.Tune(kernel, std::move(stop));
    tuner.SaveResults(results, output, ktt::OutputFormat::JSON);
    size_t valid = 0;
    size_t profiled = 0;
    double bestUs = 0.0;
    for (const auto& result : results)
    {
        if (!result.IsValid())
        {
            continue;
        }
        ++valid;
        const double us = static_cast<double>(result.GetKernelDuration()) * 1.0e-3;
        bestUs = bestUs == 0.0 || us < bestUs ? us : bestUs;
        if (!result.GetResults().empty() && result.GetResults().front().HasProfilingData()
            && result.GetResults().front().GetProfilingData().GetCounters().size() > 10)
        {
            ++profiled;
        }
    }
    std::printf("{\"mode\": \"%s\", \"tested\": %zu, \"valid\": %zu, \"profiled\": %zu, \"best_us\": %.3f, \"space\": %llu}\n",
        mode.c_str(), results.size(), valid, profiled, bestUs, static_cast<unsigned long long>(tuner.GetConfigurationsCount(kernel)));
    return valid > 0 ? 0 : 1;
}
