// CoulombSum2d tuned by ktt on the HIP backend: the counterpart of
// Examples/CoulombSum2d/CoulombSum2d.cpp for this library's gfx950 slice kernel.
//
// Usage: CoulombSum2d [device] [searcher: deterministic|random|mcmc] [max configurations, 0 = all]
//                     [grid side] [atoms] [results file, without extension]
// Defaults: device 0, the whole space, a 256 x 256 slice at z = 31.25 A, 4000 atoms in a
// 128 A box (seeded), results in CoulombSum2dOutput.json. As in the reference example, the atoms'
// z component holds the squared z distance to the slice (CoulombSum2d.cl:126) and every
// configuration is validated against the reference kernel with SideBySideComparison at 0.01
// (CoulombSum2d.cpp:133).
#include <cstdio>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include <Ktt.h>
#include <KttCoulomb.h>
#include <ktt_hip.h>

namespace
{

std::unique_ptr<ktt::Searcher> MakeSearcher(const std::string& name)
{
    if (name == "random")
    {
        return std::make_unique<ktt::RandomSearcher>();
    }
    if (name == "mcmc")
    {
        return std::make_unique<ktt::McmcSearcher>();
    }
    return std::make_unique<ktt::DeterministicSearcher>();
}

} // namespace

int main(int argc, char** argv)
{
    const ktt::DeviceIndex device = argc > 1 ? static_cast<ktt::DeviceIndex>(std::stoul(argv[1])) : 0;
    const std::string searcher = argc > 2 ? argv[2] : "deterministic";
    const uint64_t maxConfigurations = argc > 3 ? std::stoull(argv[3]) : 0;
    const int gridSize = argc > 4 ? std::stoi(argv[4]) : 256;
    const int atoms = argc > 5 ? std::stoi(argv[5]) : 4000;
    const std::string output = argc > 6 ? argv[6] : "CoulombSum2dOutput";
    const float gridSpacing = 0.5f;
    const float sliceZ = 31.25f;

    std::mt19937_64 generator(20240813);
    std::uniform_real_distribution<float> position(0.0f, gridSpacing * static_cast<float>(gridSize));
    std::uniform_real_distribution<float> charge(-1.0f, 1.0f);
    std::vector<float> atomInfo(4 * static_cast<size_t>(atoms));
    std::vector<float> atomInfoX(atoms), atomInfoY(atoms), atomInfoZ2(atoms), atomInfoW(atoms);
    for (int i = 0; i < atoms; ++i)
    {
        atomInfoX[i] = position(generator);
        atomInfoY[i] = position(generator);
        const float dz = sliceZ - position(generator);
        atomInfoZ2[i] = dz * dz;
        atomInfoW[i] = charge(generator);
        atomInfo[4 * i] = atomInfoX[i];
        atomInfo[4 * i + 1] = atomInfoY[i];
        atomInfo[4 * i + 2] = atomInfoZ2[i];
        atomInfo[4 * i + 3] = atomInfoW[i];
    }
    std::vector<float> energyGrid(static_cast<size_t>(gridSize) * gridSize, 0.0f);

    ktt::Tuner tuner(0, device, ktt::ComputeApi::HIP);
    tuner.SetGlobalSizeType(ktt::GlobalSizeType::CUDA);
    tuner.SetTimeUnit(ktt::TimeUnit::Microseconds);

    const std::string directory = ktt_hip_kernel_directory();
    const ktt::KernelDefinitionId definition = tuner.AddKernelDefinitionFromFile("directCoulombSum", directory + "/CoulombSum2d.hip",
        ktt::DimensionVector(gridSize, gridSize, 1), ktt::DimensionVector(1, 1, 1));
    const ktt::KernelDefinitionId referenceDefinition = tuner.AddKernelDefinitionFromFile("directCoulombSumReference",
        directory + "/CoulombSum2dReference.hip", ktt::DimensionVector(gridSize / 16, gridSize / 16, 1), ktt::DimensionVector(16, 16, 1));
    if (definition == ktt::InvalidKernelDefinitionId || referenceDefinition == ktt::InvalidKernelDefinitionId)
    {
        std::fprintf(stderr, "kernel sources not found in %s\n", directory.c_str());
        return 1;
    }
    const ktt::KernelId kernel = tuner.CreateSimpleKernel("CoulombSum", definition);
    const ktt::KernelId referenceKernel = tuner.CreateSimpleKernel("CoulombSumReference", referenceDefinition);

    const ktt::ArgumentId aiId = tuner.AddArgumentVector(atomInfo, ktt::ArgumentAccessType::ReadOnly);
    const ktt::ArgumentId axId = tuner.AddArgumentVector(atomInfoX, ktt::ArgumentAccessType::ReadOnly);
    const ktt::ArgumentId ayId = tuner.AddArgumentVector(atomInfoY, ktt::ArgumentAccessType::ReadOnly);
    const ktt::ArgumentId azId = tuner.AddArgumentVector(atomInfoZ2, ktt::ArgumentAccessType::ReadOnly);
    const ktt::ArgumentId awId = tuner.AddArgumentVector(atomInfoW, ktt::ArgumentAccessType::ReadOnly);
    const ktt::ArgumentId nId = tuner.AddArgumentScalar(atoms);
    const ktt::ArgumentId hId = tuner.AddArgumentScalar(gridSpacing);
    const ktt::ArgumentId gridId = tuner.AddArgumentVector(energyGrid, ktt::ArgumentAccessType::ReadWrite);

    ktt::AddCoulombSliceTuningSpace(tuner, kernel, definition, static_cast<uint64_t>(gridSize), static_cast<uint64_t>(gridSize));
    tuner.SetArguments(definition, {aiId, axId, ayId, azId, awId, nId, hId, gridId});
    tuner.SetArguments(referenceDefinition, {aiId, nId, hId, gridId});
    tuner.SetReferenceKernel(gridId, referenceKernel, ktt::KernelConfiguration());
    tuner.SetValidationMethod(ktt::ValidationMethod::SideBySideComparison, 0.01);
    tuner.SetSearcher(kernel, MakeSearcher(searcher));

    std::unique_ptr<ktt::StopCondition> stop;
    if (maxConfigurations > 0)
    {
        stop = std::make_unique<ktt::ConfigurationCount>(maxConfigurations);
    }
    const std::vector<ktt::KernelResult> results = tuner.Tune(kernel, std::move(stop));
    tuner.SaveResults(results, output, ktt::OutputFormat::JSON);

    const ktt::KernelResult* best = nullptr;
    size_t valid = 0;
    for (const auto& result : results)
    {
        if (result.IsValid())
        {
            ++valid;
            if (best == nullptr || result.GetTotalDuration() < best->GetTotalDuration())
            {
                best = &result;
            }
        }
    }
    if (best == nullptr)
    {
        std::fprintf(stderr, "no configuration passed validation\n");
        return 1;
    }
    const double seconds = static_cast<double>(best->GetKernelDuration()) * 1.0e-9;
    const double interactions = static_cast<double>(gridSize) * gridSize * atoms;
    std::printf("tested %zu configurations, %zu valid; best %.3f us (%.3e interactions/s): %s\n", results.size(), valid,
        seconds * 1.0e6, interactions / seconds, best->GetConfiguration().GetString().c_str());
    return 0;
}
