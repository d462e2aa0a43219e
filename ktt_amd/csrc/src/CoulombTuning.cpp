// ktt_coulomb3d_tune: the CoulombSum3d example (Examples/CoulombSum3d/CoulombSum3d.cpp:117-237)
// driven through ktt::Tuner on ComputeApi::HIP, with the tuning space of this library's gfx950
// kernel. Everything goes through the public C++ API: definitions from files, arguments,
// parameters, thread modifiers, constraints, reference-kernel validation, searcher, stop
// condition, JSON results and a final Run of the best configuration.
#include <algorithm>
#include <cstdio>
#include <memory>
#include <vector>

#include <Ktt.h>
#include <KttCoulomb.h>

#include "../../../include/ktt_hip.h"
#include "CoulombRuntime.h"

namespace ktt
{

/** Adds the gfx950 CoulombSum3d tuning space to a kernel whose definition has default global
 *  size (G, G*G, 1) and local size (1, 1, 1). */
void AddCoulomb3dTuningSpace(Tuner& tuner, const KernelId kernel, const KernelDefinitionId definition)
{
    tuner.AddParameter(kernel, "WORK_GROUP_SIZE_X", std::vector<uint64_t>{16, 32, 64, 128, 256});
    tuner.AddParameter(kernel, "WORK_GROUP_SIZE_Y", std::vector<uint64_t>{1, 2, 4, 8});
    tuner.AddParameter(kernel, "POINTS_X", std::vector<uint64_t>{1, 2, 4, 8});
    tuner.AddParameter(kernel, "ROWS_PER_THREAD", std::vector<uint64_t>{4, 8, 16});
    tuner.AddParameter(kernel, "ATOM_TILE", std::vector<uint64_t>{32, 64, 128, 256});
    tuner.AddParameter(kernel, "INNER_UNROLL_FACTOR", std::vector<uint64_t>{1, 2, 4, 8, 16});
    tuner.AddParameter(kernel, "ISSUE_PHASES", std::vector<uint64_t>{0, 1, 2});

    tuner.AddThreadModifier(kernel, {definition}, ModifierType::Local, ModifierDimension::X, "WORK_GROUP_SIZE_X",
        ModifierAction::Multiply);
    tuner.AddThreadModifier(kernel, {definition}, ModifierType::Local, ModifierDimension::Y, "WORK_GROUP_SIZE_Y",
        ModifierAction::Multiply);
    tuner.AddThreadModifier(kernel, {definition}, ModifierType::Global, ModifierDimension::X,
        std::vector<std::string>{"WORK_GROUP_SIZE_X", "POINTS_X"},
        [](const uint64_t size, const std::vector<uint64_t>& values) { return (size + values[0] * values[1] - 1) / (values[0] * values[1]); });
    tuner.AddThreadModifier(kernel, {definition}, ModifierType::Global, ModifierDimension::Y,
        std::vector<std::string>{"WORK_GROUP_SIZE_Y", "ROWS_PER_THREAD"},
        [](const uint64_t size, const std::vector<uint64_t>& values) { return (size + values[0] * values[1] - 1) / (values[0] * values[1]); });

    // At least one full wave64 and at most the 1024-thread limit per workgroup.
    tuner.AddConstraint(kernel, {"WORK_GROUP_SIZE_X", "WORK_GROUP_SIZE_Y"},
        [](const std::vector<uint64_t>& v) { return v[0] * v[1] >= 64 && v[0] * v[1] <= 1024; });
    // Rows per workgroup and threads per workgroup divide one another; register block <= 64 points.
    tuner.AddConstraint(kernel, {"WORK_GROUP_SIZE_X", "WORK_GROUP_SIZE_Y", "POINTS_X", "ROWS_PER_THREAD"},
        [](const std::vector<uint64_t>& v)
        {
            const uint64_t threads = v[0] * v[1];
            const uint64_t rows = v[1] * v[3];
            return (threads % rows == 0 || rows % threads == 0) && v[2] * v[3] <= 64;
        });
    // The atom tile is a multiple of the bracket precompute stride and fits 64 KiB of LDS.
    tuner.AddConstraint(kernel, {"WORK_GROUP_SIZE_X", "WORK_GROUP_SIZE_Y", "ROWS_PER_THREAD", "ATOM_TILE"},
        [](const std::vector<uint64_t>& v)
        {
            const uint64_t threads = v[0] * v[1];
            const uint64_t rows = v[1] * v[2];
            const uint64_t stride = threads > rows ? threads / rows : 1;
            return v[3] % stride == 0 && v[3] * (rows + 4) * 4 <= 64 * 1024;
        });
    tuner.AddConstraint(kernel, {"INNER_UNROLL_FACTOR", "ATOM_TILE"}, [](const std::vector<uint64_t>& v) { return v[0] <= v[1]; });
}

void AddCoulombSliceTuningSpace(Tuner& tuner, const KernelId kernel, const KernelDefinitionId definition, const uint64_t gridX,
    const uint64_t gridY)
{
    tuner.AddParameter(kernel, "WORK_GROUP_SIZE_X", std::vector<uint64_t>{16, 32, 64});
    tuner.AddParameter(kernel, "WORK_GROUP_SIZE_Y", std::vector<uint64_t>{1, 2, 4, 8});
    tuner.AddParameter(kernel, "POINTS_X", std::vector<uint64_t>{1, 2, 4});
    tuner.AddParameter(kernel, "ROWS_PER_THREAD", std::vector<uint64_t>{2, 4, 8});
    tuner.AddParameter(kernel, "ATOM_TILE", std::vector<uint64_t>{32, 64, 128});
    tuner.AddParameter(kernel, "INNER_UNROLL_FACTOR", std::vector<uint64_t>{1, 8, 16});
    tuner.AddParameter(kernel, "ISSUE_PHASES", std::vector<uint64_t>{0, 1, 2});
    tuner.AddParameter(kernel, "ATOM_BLOCKS", std::vector<uint64_t>{1, 4, 16, 64, 128, 256});

    tuner.AddThreadModifier(kernel, {definition}, ModifierType::Local, ModifierDimension::X, "WORK_GROUP_SIZE_X",
        ModifierAction::Multiply);
    tuner.AddThreadModifier(kernel, {definition}, ModifierType::Local, ModifierDimension::Y, "WORK_GROUP_SIZE_Y",
        ModifierAction::Multiply);
    tuner.AddThreadModifier(kernel, {definition}, ModifierType::Global, ModifierDimension::X,
        std::vector<std::string>{"WORK_GROUP_SIZE_X", "POINTS_X"},
        [](const uint64_t size, const std::vector<uint64_t>& values) { return size / (values[0] * values[1]); });
    tuner.AddThreadModifier(kernel, {definition}, ModifierType::Global, ModifierDimension::Y,
        std::vector<std::string>{"WORK_GROUP_SIZE_Y", "ROWS_PER_THREAD"},
        [](const uint64_t size, const std::vector<uint64_t>& values) { return size / (values[0] * values[1]); });
    tuner.AddThreadModifier(kernel, {definition}, ModifierType::Global, ModifierDimension::Z, "ATOM_BLOCKS",
        ModifierAction::Multiply);

    tuner.AddConstraint(kernel, {"WORK_GROUP_SIZE_X", "WORK_GROUP_SIZE_Y"},
        [](const std::vector<uint64_t>& v) { return v[0] * v[1] >= 64 && v[0] * v[1] <= 1024; });
    // The slice kernels take the row length and row count from the launch: both must tile exactly.
    tuner.AddConstraint(kernel, {"WORK_GROUP_SIZE_X", "POINTS_X"},
        [gridX](const std::vector<uint64_t>& v) { return gridX % (v[0] * v[1]) == 0; });
    tuner.AddConstraint(kernel, {"WORK_GROUP_SIZE_X", "WORK_GROUP_SIZE_Y", "ROWS_PER_THREAD"},
        [gridY](const std::vector<uint64_t>& v)
        {
            const uint64_t threads = v[0] * v[1];
            const uint64_t rows = v[1] * v[2];
            return gridY % rows == 0 && (threads % rows == 0 || rows % threads == 0);
        });
    tuner.AddConstraint(kernel, {"WORK_GROUP_SIZE_X", "WORK_GROUP_SIZE_Y", "ROWS_PER_THREAD", "ATOM_TILE"},
        [](const std::vector<uint64_t>& v)
        {
            const uint64_t threads = v[0] * v[1];
            const uint64_t rows = v[1] * v[2];
            const uint64_t stride = threads > rows ? threads / rows : 1;
            return v[3] % stride == 0 && v[3] * (rows + 4) * 4 <= 64 * 1024;
        });
}

} // namespace ktt

using namespace ktt;

extern "C" int ktt_coulomb3d_tune(const float* hostAtoms, int atomCount, float spacing, int gridSize, int searcher,
    int maxConfigurations, double tolerance, const char* resultsFile, char* bestConfig, size_t bestConfigLength, double* bestMs,
    int* validCount, int* testedCount, float* energyGrid)
{
    try
    {
        if (hostAtoms == nullptr || atomCount <= 0 || gridSize <= 0 || gridSize > 1024)
        {
            return SetLastError("tuning needs host atoms, a positive atom count and a grid size in [1, 1024]");
        }
        const std::string directory = CoulombRuntime::Get().GetKernelDirectory();
        const size_t gridPoints = static_cast<size_t>(gridSize) * gridSize * gridSize;

        std::vector<float> atomInfo(hostAtoms, hostAtoms + static_cast<size_t>(atomCount) * 4);
        std::vector<float> atomX(atomCount), atomY(atomCount), atomZ(atomCount), atomW(atomCount);
        for (int i = 0; i < atomCount; ++i)
        {
            atomX[i] = atomInfo[4 * i];
            atomY[i] = atomInfo[4 * i + 1];
            atomZ[i] = atomInfo[4 * i + 2];
            atomW[i] = atomInfo[4 * i + 3];
        }
        std::vector<float> grid(gridPoints, 0.0f);

        Tuner tuner(0, 0, ComputeApi::HIP);
        tuner.SetGlobalSizeType(GlobalSizeType::CUDA);
        tuner.SetTimeUnit(TimeUnit::Microseconds);

        const DimensionVector ndRange(static_cast<size_t>(gridSize), static_cast<size_t>(gridSize) * gridSize, 1);
        const DimensionVector referenceNdRange((gridSize + 15) / 16, (gridSize + 15) / 16, gridSize);
        const KernelDefinitionId definition = tuner.AddKernelDefinitionFromFile("directCoulombSum", directory + "/CoulombSum3d.hip",
            ndRange, DimensionVector(1, 1, 1));
        const KernelDefinitionId referenceDefinition = tuner.AddKernelDefinitionFromFile("directCoulombSumReference",
            directory + "/CoulombSum3dReference.hip", referenceNdRange, DimensionVector(16, 16, 1));
        if (definition == InvalidKernelDefinitionId || referenceDefinition == InvalidKernelDefinitionId)
        {
            return SetLastError("unable to load the CoulombSum3d kernel sources from " + directory);
        }
        const KernelId kernel = tuner.CreateSimpleKernel("CoulombSum", definition);
        const KernelId referenceKernel = tuner.CreateSimpleKernel("CoulombSumReference", referenceDefinition);

        const ArgumentId aiId = tuner.AddArgumentVector(atomInfo, ArgumentAccessType::ReadOnly);
        const ArgumentId axId = tuner.AddArgumentVector(atomX, ArgumentAccessType::ReadOnly);
        const ArgumentId ayId = tuner.AddArgumentVector(atomY, ArgumentAccessType::ReadOnly);
        const ArgumentId azId = tuner.AddArgumentVector(atomZ, ArgumentAccessType::ReadOnly);
        const ArgumentId awId = tuner.AddArgumentVector(atomW, ArgumentAccessType::ReadOnly);
        const ArgumentId nId = tuner.AddArgumentScalar(atomCount);
        const ArgumentId hId = tuner.AddArgumentScalar(spacing);
        const ArgumentId gId = tuner.AddArgumentScalar(gridSize);
        const ArgumentId gridId = tuner.AddArgumentVector(grid, ArgumentAccessType::WriteOnly);

        AddCoulomb3dTuningSpace(tuner, kernel, definition);
        tuner.SetArguments(definition, {aiId, axId, ayId, azId, awId, nId, hId, gId, gridId});
        tuner.SetArguments(referenceDefinition, {aiId, nId, hId, gId, gridId});
        tuner.SetReferenceKernel(gridId, referenceKernel, KernelConfiguration());
        tuner.SetValidationMethod(ValidationMethod::SideBySideComparison, tolerance);

        switch (searcher)
        {
        case KTT_SEARCHER_RANDOM:
            tuner.SetSearcher(kernel, std::make_unique<RandomSearcher>());
            break;
        case KTT_SEARCHER_MCMC:
            tuner.SetSearcher(kernel, std::make_unique<McmcSearcher>());
            break;
        default:
            tuner.SetSearcher(kernel, std::make_unique<DeterministicSearcher>());
            break;
        }

        std::unique_ptr<StopCondition> stop;
        if (maxConfigurations > 0)
        {
            stop = std::make_unique<ConfigurationCount>(static_cast<uint64_t>(maxConfigurations));
        }
        const std::vector<KernelResult> results = tuner.Tune(kernel, std::move(stop));
        if (results.empty())
        {
            return SetLastError("tuning produced no results");
        }
        if (resultsFile != nullptr && resultsFile[0] != '\0')
        {
            tuner.SaveResults(results, resultsFile, OutputFormat::JSON);
        }

        const KernelResult* best = nullptr;
        int valid = 0;
        for (const auto& result : results)
        {
            if (!result.IsValid())
            {
                continue;
            }
            ++valid;
            if (best == nullptr || result.GetTotalDuration() < best->GetTotalDuration())
            {
                best = &result;
            }
        }
        if (validCount != nullptr)
        {
            *validCount = valid;
        }
        if (testedCount != nullptr)
        {
            *testedCount = static_cast<int>(results.size());
        }
        if (best == nullptr)
        {
            return SetLastError("no configuration passed validation");
        }
        if (bestConfig != nullptr && bestConfigLength > 0)
        {
            std::string text;
            for (const auto& pair : best->GetConfiguration().GetPairs())
            {
                text += (text.empty() ? "" : ",") + pair.GetName() + "=" + pair.GetValueString();
            }
            std::snprintf(bestConfig, bestConfigLength, "%s", text.c_str());
        }
        if (bestMs != nullptr)
        {
            *bestMs = static_cast<double>(best->GetKernelDuration()) / 1.0e6;
        }
        if (energyGrid != nullptr)
        {
            const auto run = tuner.Run(kernel, best->GetConfiguration(),
                {BufferOutputDescriptor(gridId, energyGrid, gridPoints * sizeof(float))});
            if (!run.IsValid())
            {
                return SetLastError("the best configuration failed when run again");
            }
        }
        return 0;
    }
    catch (const std::exception& exception)
    {
        return SetLastError(exception.what());
    }
}
