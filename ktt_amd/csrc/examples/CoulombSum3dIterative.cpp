// CoulombSum3dIterative tuned by ktt on the HIP backend: the counterpart of
// Examples/CoulombSum3dIterative/CoulombSum3dIterative.cpp for this library's gfx950 kernels.
//
// The launcher keeps the reference's per-slice scheme (CoulombSum3dIterative.cpp:146-184): for
// every z slice the host writes each atom's squared z distance into the atoms' z component,
// uploads the atoms, sets zIndex and launches the slice kernel. The parameter DEVICE_DZ2 adds
// the fused alternative: the atoms keep their real z and the slice kernel forms dz^2 on the
// device, so the host work and the per-slice uploads disappear. Validation is the reference's:
// the 3D reference kernel with SideBySideComparison at 0.01 (CoulombSum3dIterative.cpp:141).
//
// Usage: CoulombSum3dIterative [device] [searcher: deterministic|random|mcmc]
//                              [max configurations, 0 = all] [grid side] [atoms] [results file]
// Defaults: device 0, random search of 24 configurations, 128^3 grid, 4096 atoms (seeded).
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
    const std::string searcher = argc > 2 ? argv[2] : "random";
    const uint64_t maxConfigurations = argc > 3 ? std::stoull(argv[3]) : 24;
    const int gridSize = argc > 4 ? std::stoi(argv[4]) : 128;
    const int atoms = argc > 5 ? std::stoi(argv[5]) : 4096;
    const std::string output = argc > 6 ? argv[6] : "CoulombSum3dIterativeOutput";
    const float gridSpacing = 0.5f;

    std::mt19937_64 generator(20240814);
    std::uniform_real_distribution<float> position(0.0f, gridSpacing * static_cast<float>(gridSize));
    std::uniform_real_distribution<float> charge(-1.0f, 1.0f);
    std::vector<float> atomInfo(4 * static_cast<size_t>(atoms));
    std::vector<float> atomZ(atoms);
    for (int i = 0; i < atoms; ++i)
    {
        atomInfo[4 * i] = position(generator);
        atomInfo[4 * i + 1] = position(generator);
        atomZ[i] = position(generator);
        atomInfo[4 * i + 2] = atomZ[i];
        atomInfo[4 * i + 3] = charge(generator);
    }
    // Per-slice copy whose z component the launcher overwrites with dz^2.
    std::vector<float> sliceAtoms = atomInfo;
    std::vector<float> energyGrid(static_cast<size_t>(gridSize) * gridSize * gridSize, 0.0f);
    int zIndex = 0;

    ktt::Tuner tuner(0, device, ktt::ComputeApi::HIP);
    tuner.SetGlobalSizeType(ktt::GlobalSizeType::CUDA);
    tuner.SetTimeUnit(ktt::TimeUnit::Microseconds);

    const std::string directory = ktt_hip_kernel_directory();
    const std::string file = directory + "/CoulombSum3dIterative.hip";
    const ktt::DimensionVector slice(gridSize, gridSize, 1);
    const ktt::KernelDefinitionId hostDz2 = tuner.AddKernelDefinitionFromFile("directCoulombSum", file, slice, ktt::DimensionVector(1, 1, 1));
    const ktt::KernelDefinitionId deviceDz2 = tuner.AddKernelDefinitionFromFile("directCoulombSumSlice", file, slice,
        ktt::DimensionVector(1, 1, 1));
    const ktt::KernelDefinitionId referenceDefinition = tuner.AddKernelDefinitionFromFile("directCoulombSumReference",
        directory + "/CoulombSum3dIterativeReference.hip", ktt::DimensionVector(gridSize / 16, gridSize / 16, gridSize),
        ktt::DimensionVector(16, 16, 1));
    if (hostDz2 == ktt::InvalidKernelDefinitionId || deviceDz2 == ktt::InvalidKernelDefinitionId
        || referenceDefinition == ktt::InvalidKernelDefinitionId)
    {
        std::fprintf(stderr, "kernel sources not found in %s\n", directory.c_str());
        return 1;
    }
    const ktt::KernelId kernel = tuner.CreateCompositeKernel("CoulombSum", {hostDz2, deviceDz2});
    const ktt::KernelId referenceKernel = tuner.CreateSimpleKernel("CoulombSumReference", referenceDefinition);

    const ktt::ArgumentId atomsId = tuner.AddArgumentVector(atomInfo, ktt::ArgumentAccessType::ReadOnly);
    const ktt::ArgumentId sliceAtomsId = tuner.AddArgumentVector(sliceAtoms, ktt::ArgumentAccessType::ReadOnly);
    // The SoA arrays are not read with USE_SOA = 0; the kernels still take them.
    std::vector<ktt::ArgumentId> unusedIds;
    for (int k = 0; k < 4; ++k)
    {
        unusedIds.push_back(tuner.AddArgumentVector(atomZ, ktt::ArgumentAccessType::ReadOnly));
    }
    const ktt::ArgumentId nId = tuner.AddArgumentScalar(atoms);
    const ktt::ArgumentId hId = tuner.AddArgumentScalar(gridSpacing);
    const ktt::ArgumentId zId = tuner.AddArgumentScalar(zIndex);
    const ktt::ArgumentId gridId = tuner.AddArgumentVector(energyGrid, ktt::ArgumentAccessType::ReadWrite);
    tuner.SetArguments(hostDz2, {sliceAtomsId, unusedIds[0], unusedIds[1], unusedIds[2], unusedIds[3], nId, hId, zId, gridId});
    tuner.SetArguments(deviceDz2, {atomsId, unusedIds[0], unusedIds[1], unusedIds[2], unusedIds[3], nId, hId, zId, gridId});
    tuner.SetArguments(referenceDefinition, {atomsId, nId, hId, gridId});

    ktt::AddCoulombSliceTuningSpace(tuner, kernel, hostDz2, static_cast<uint64_t>(gridSize), static_cast<uint64_t>(gridSize));
    tuner.AddParameter(kernel, "DEVICE_DZ2", std::vector<uint64_t>{0, 1});
    // The fused definition launches with the same shape as the host-dz^2 one.
    tuner.AddThreadModifier(kernel, {deviceDz2}, ktt::ModifierType::Local, ktt::ModifierDimension::X, "WORK_GROUP_SIZE_X",
        ktt::ModifierAction::Multiply);
    tuner.AddThreadModifier(kernel, {deviceDz2}, ktt::ModifierType::Local, ktt::ModifierDimension::Y, "WORK_GROUP_SIZE_Y",
        ktt::ModifierAction::Multiply);
    tuner.AddThreadModifier(kernel, {deviceDz2}, ktt::ModifierType::Global, ktt::ModifierDimension::X,
        std::vector<std::string>{"WORK_GROUP_SIZE_X", "POINTS_X"},
        [](const uint64_t size, const std::vector<uint64_t>& values) { return size / (values[0] * values[1]); });
    tuner.AddThreadModifier(kernel, {deviceDz2}, ktt::ModifierType::Global, ktt::ModifierDimension::Y,
        std::vector<std::string>{"WORK_GROUP_SIZE_Y", "ROWS_PER_THREAD"},
        [](const uint64_t size, const std::vector<uint64_t>& values) { return size / (values[0] * values[1]); });
    tuner.AddThreadModifier(kernel, {deviceDz2}, ktt::ModifierType::Global, ktt::ModifierDimension::Z, "ATOM_BLOCKS",
        ktt::ModifierAction::Multiply);

    tuner.SetLauncher(kernel, [&, hostDz2, deviceDz2, sliceAtomsId, zId](ktt::ComputeInterface& interface)
    {
        const auto& pairs = interface.GetCurrentConfiguration().GetPairs();
        const bool onDevice = ktt::ParameterPair::GetParameterValue<uint64_t>(pairs, "DEVICE_DZ2") == 1;
        const ktt::KernelDefinitionId definition = onDevice ? deviceDz2 : hostDz2;
        const ktt::DimensionVector global = interface.GetCurrentGlobalSize(definition);
        const ktt::DimensionVector local = interface.GetCurrentLocalSize(definition);
        for (int z = 0; z < gridSize; ++z)
        {
            if (!onDevice)
            {
                const float coordinate = gridSpacing * static_cast<float>(z);
                for (int i = 0; i < atoms; ++i)
                {
                    const float dz = coordinate - atomZ[i];
                    sliceAtoms[4 * static_cast<size_t>(i) + 2] = dz * dz;
                }
                interface.UpdateBuffer(sliceAtomsId, sliceAtoms.data());
            }
            interface.UpdateScalarArgument(zId, &z);
            interface.RunKernel(definition, global, local);
        }
    });

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
    const double interactions = static_cast<double>(gridSize) * gridSize * gridSize * atoms;
    const double kernelSeconds = static_cast<double>(best->GetKernelDuration()) * 1.0e-9;
    const double totalSeconds = static_cast<double>(best->GetTotalDuration()) * 1.0e-9;
    std::printf("tested %zu configurations, %zu valid; best total %.3f ms (kernels %.3f ms, %.3e interactions/s): %s\n",
        results.size(), valid, totalSeconds * 1.0e3, kernelSeconds * 1.0e3, interactions / kernelSeconds,
        best->GetConfiguration().GetString().c_str());
    return 0;
}
