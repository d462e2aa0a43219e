// ktt_coulomb3d_dynamic: dynamic (online) tuning of the compensated CoulombSum3d kernel through
// ktt::Tuner::TuneIteration (Source/Tuner.h:662-698, Source/TuningRunner/TuningRunner.cpp:84-128).
// Every step moves each atom by a seeded Gaussian displacement on the device, clears the map and
// recomputes it. While the tuning budget lasts a step runs the searcher's next configuration;
// afterwards every step runs the best configuration found (Tuner::GetBestConfiguration,
// Source/Tuner.h:776-782), which is the production phase whose kernel time is the converged rate.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <memory>
#include <vector>

#include <hip/hip_runtime_api.h>

#include <Ktt.h>
#include <KttCoulomb.h>

#include "../../../include/ktt_hip.h"
#include "CoulombRuntime.h"

namespace ktt
{

namespace
{

constexpr uint64_t kMaxAtomBlocks = 4;
constexpr uint64_t kMinSpanX = 32;
constexpr uint64_t kMinRowsPerGroup = 8;

/** Device buffers of one dynamic run, released on every exit path. */
struct DynamicBuffers
{
    void* atoms = nullptr;
    void* grid = nullptr;
    void* partials = nullptr;
    void* counters = nullptr;
    size_t partialsBytes = 0;
    size_t countersBytes = 0;

    ~DynamicBuffers()
    {
        for (void* pointer : {atoms, grid, partials, counters})
        {
            if (pointer != nullptr)
            {
                static_cast<void>(hipFree(pointer));
            }
        }
    }
};

void Check(const hipError_t status, const char* what)
{
    if (status != hipSuccess)
    {
        throw KttException(std::string(what) + ": " + hipGetErrorString(status));
    }
}

uint64_t CeilDiv(const uint64_t value, const uint64_t divisor)
{
    return (value + divisor - 1) / divisor;
}

} // namespace

/** Tuning space of the dynamic run: the compensated 3D kernel with the atom range split over
 *  ATOM_BLOCKS workgroups per tile (FP64 merge of the block partials). Spans along x that waste
 *  more than an eighth of a row are excluded. The atom loop is unrolled 16 times in every
 *  configuration: at 192^3 x 131,072 the space's best shape runs 81.7 ms unrolled 16 times,
 *  81.9 ms 8 times and 84.4 ms not unrolled, and a fixed unroll keeps the space small enough
 *  for 48 online steps. */
void AddCoulombDynamicTuningSpace(Tuner& tuner, const KernelId kernel, const KernelDefinitionId definition, const uint64_t gridSize)
{
    tuner.AddParameter(kernel, "WORK_GROUP_SIZE_X", std::vector<uint64_t>{32, 64});
    tuner.AddParameter(kernel, "WORK_GROUP_SIZE_Y", std::vector<uint64_t>{2, 4, 8});
    tuner.AddParameter(kernel, "POINTS_X", std::vector<uint64_t>{1, 2, 3, 4, 6});
    tuner.AddParameter(kernel, "ROWS_PER_THREAD", std::vector<uint64_t>{4, 8, 16});
    tuner.AddParameter(kernel, "ATOM_TILE", std::vector<uint64_t>{64, 128});
    tuner.AddParameter(kernel, "INNER_UNROLL_FACTOR", std::vector<uint64_t>{16});
    tuner.AddParameter(kernel, "ISSUE_PHASES", std::vector<uint64_t>{0, 1});
    tuner.AddParameter(kernel, "ATOM_BLOCKS", std::vector<uint64_t>{1, 2, kMaxAtomBlocks});

    tuner.AddThreadModifier(kernel, {definition}, ModifierType::Local, ModifierDimension::X, "WORK_GROUP_SIZE_X",
        ModifierAction::Multiply);
    tuner.AddThreadModifier(kernel, {definition}, ModifierType::Local, ModifierDimension::Y, "WORK_GROUP_SIZE_Y",
        ModifierAction::Multiply);
    tuner.AddThreadModifier(kernel, {definition}, ModifierType::Global, ModifierDimension::X,
        std::vector<std::string>{"WORK_GROUP_SIZE_X", "POINTS_X"},
        [](const uint64_t size, const std::vector<uint64_t>& values) { return CeilDiv(size, values[0] * values[1]); });
    tuner.AddThreadModifier(kernel, {definition}, ModifierType::Global, ModifierDimension::Y,
        std::vector<std::string>{"WORK_GROUP_SIZE_Y", "ROWS_PER_THREAD"},
        [](const uint64_t size, const std::vector<uint64_t>& values) { return CeilDiv(size, values[0] * values[1]); });
    tuner.AddThreadModifier(kernel, {definition}, ModifierType::Global, ModifierDimension::Z, "ATOM_BLOCKS",
        ModifierAction::Multiply);

    tuner.AddConstraint(kernel, {"WORK_GROUP_SIZE_X", "WORK_GROUP_SIZE_Y"},
        [](const std::vector<uint64_t>& v) { return v[0] * v[1] >= 64 && v[0] * v[1] <= 1024; });
    tuner.AddConstraint(kernel, {"WORK_GROUP_SIZE_X", "POINTS_X"},
        [gridSize](const std::vector<uint64_t>& v)
        {
            const uint64_t span = v[0] * v[1];
            return CeilDiv(gridSize, span) * span * 8 <= gridSize * 9;
        });
    // Compensated accumulation keeps three values per point: at most 32 points per thread.
    tuner.AddConstraint(kernel, {"WORK_GROUP_SIZE_X", "WORK_GROUP_SIZE_Y", "POINTS_X", "ROWS_PER_THREAD"},
        [](const std::vector<uint64_t>& v)
        {
            const uint64_t threads = v[0] * v[1];
            const uint64_t rows = v[1] * v[3];
            return (threads % rows == 0 || rows % threads == 0) && v[2] * v[3] <= 32;
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

extern "C" int ktt_coulomb3d_dynamic(const float* hostAtoms, int atomCount, float spacing, int gridSize, int steps, float sigma,
    unsigned long long seed, int searcher, int tuningBudget, const char* resultsFile, double* stepKernelMs, double* stepWallMs,
    int* stepTuned, char* bestConfig, size_t bestConfigLength, int* testedCount, int* spaceSize, float* finalAtoms,
    float* finalGrid)
{
    try
    {
        if (hostAtoms == nullptr || atomCount <= 0 || gridSize <= 0 || gridSize > 1024 || steps <= 0 || !(sigma >= 0.0f))
        {
            return SetLastError("the dynamic run needs host atoms, a positive atom count, a grid size in [1, 1024], "
                "at least one step and a non-negative sigma");
        }
        CoulombRuntime& runtime = CoulombRuntime::Get();
        runtime.EnsureDevice();
        const std::string directory = runtime.GetKernelDirectory();
        const uint64_t side = static_cast<uint64_t>(gridSize);
        const uint64_t gridPoints = side * side * side;

        DynamicBuffers buffers;
        const size_t atomBytes = static_cast<size_t>(atomCount) * 4 * sizeof(float);
        Check(hipMalloc(&buffers.atoms, atomBytes), "hipMalloc(atoms)");
        Check(hipMemcpy(buffers.atoms, hostAtoms, atomBytes, hipMemcpyHostToDevice), "hipMemcpy(atoms)");
        Check(hipMalloc(&buffers.grid, gridPoints * sizeof(float)), "hipMalloc(grid)");
        // Worst case over the space: x padded by at most an eighth, rows by less than one group.
        const uint64_t paddedX = side + side / 8 + 1;
        const uint64_t paddedRows = side * side + 128;
        buffers.partialsBytes = static_cast<size_t>(kMaxAtomBlocks * paddedX * paddedRows * 2 * sizeof(float));
        buffers.countersBytes = static_cast<size_t>(CeilDiv(side, kMinSpanX) * CeilDiv(side * side, kMinRowsPerGroup) * sizeof(unsigned int));
        Check(hipMalloc(&buffers.partials, buffers.partialsBytes), "hipMalloc(partials)");
        Check(hipMalloc(&buffers.counters, buffers.countersBytes), "hipMalloc(counters)");
        Check(hipMemset(buffers.counters, 0, buffers.countersBytes), "hipMemset(counters)");
        Check(hipDeviceSynchronize(), "hipDeviceSynchronize");

        Tuner tuner(0, 0, ComputeApi::HIP);
        tuner.SetGlobalSizeType(GlobalSizeType::CUDA);
        tuner.SetTimeUnit(TimeUnit::Microseconds);
        tuner.SetValidationMode(ValidationMode::None);

        const DimensionVector one(1, 1, 1);
        const KernelDefinitionId jitter = tuner.AddKernelDefinitionFromFile("kttJitterAtoms", directory + "/coulomb_util.hip",
            DimensionVector(CeilDiv(static_cast<uint64_t>(atomCount), 256)), DimensionVector(256));
        const KernelDefinitionId zero = tuner.AddKernelDefinitionFromFile("kttZeroFill4", directory + "/coulomb_util.hip",
            DimensionVector(std::min<uint64_t>(CeilDiv(gridPoints, 256), 8192)), DimensionVector(256));
        const KernelDefinitionId coulomb = tuner.AddKernelDefinitionFromFile("directCoulombSumCompensated",
            directory + "/CoulombSum3d.hip", DimensionVector(side, side * side, 1), one);
        if (jitter == InvalidKernelDefinitionId || zero == InvalidKernelDefinitionId || coulomb == InvalidKernelDefinitionId)
        {
            return SetLastError("unable to load the kernel sources from " + directory);
        }

        int currentStep = 0;
        const ArgumentId atomsId = tuner.AddArgumentVector<float>(buffers.atoms, atomBytes, ArgumentAccessType::ReadWrite,
            ArgumentMemoryLocation::Device);
        // The X, Y, Z, W arrays are read only when USE_SOA is 1 (not in this space); each still needs
        // an argument of its own, since a definition's arguments must be distinct.
        std::vector<ArgumentId> soaIds;
        for (int k = 0; k < 4; ++k)
        {
            soaIds.push_back(tuner.AddArgumentVector<float>(buffers.atoms, atomBytes, ArgumentAccessType::ReadOnly,
                ArgumentMemoryLocation::Device));
        }
        const ArgumentId gridId = tuner.AddArgumentVector<float>(buffers.grid, gridPoints * sizeof(float), ArgumentAccessType::ReadWrite,
            ArgumentMemoryLocation::Device);
        const ArgumentId partialsId = tuner.AddArgumentVector<float>(buffers.partials, buffers.partialsBytes,
            ArgumentAccessType::ReadWrite, ArgumentMemoryLocation::Device);
        const ArgumentId countersId = tuner.AddArgumentVector<unsigned int>(buffers.counters, buffers.countersBytes,
            ArgumentAccessType::ReadWrite, ArgumentMemoryLocation::Device);
        const ArgumentId countId = tuner.AddArgumentScalar(atomCount);
        const ArgumentId spacingId = tuner.AddArgumentScalar(spacing);
        const ArgumentId sideId = tuner.AddArgumentScalar(gridSize);
        const ArgumentId wordsId = tuner.AddArgumentScalar(static_cast<unsigned long long>(gridPoints));
        const ArgumentId seedId = tuner.AddArgumentScalar(seed);
        const ArgumentId stepId = tuner.AddArgumentScalar(currentStep);
        const ArgumentId sigmaId = tuner.AddArgumentScalar(sigma);
        tuner.SetArguments(jitter, {atomsId, countId, seedId, stepId, sigmaId});
        tuner.SetArguments(zero, {gridId, wordsId});
        // splitFrom 0: ATOM_BLOCKS splits every tile, as the tuning space means it.
        const ArgumentId splitFromId = tuner.AddArgumentScalar(0);
        tuner.SetArguments(coulomb, {atomsId, soaIds[0], soaIds[1], soaIds[2], soaIds[3], countId, spacingId, sideId, gridId,
            partialsId, countersId, splitFromId});

        const KernelId kernel = tuner.CreateCompositeKernel("CoulombSum3dDynamic", {jitter, zero, coulomb});
        AddCoulombDynamicTuningSpace(tuner, kernel, coulomb, side);
        const size_t partialsBytes = buffers.partialsBytes;
        const size_t countersBytes = buffers.countersBytes;
        tuner.SetLauncher(kernel, [&currentStep, jitter, zero, coulomb, stepId, side, partialsBytes, countersBytes](
            ComputeInterface& interface)
        {
            // The atoms move on every step, whichever configuration runs.
            interface.UpdateScalarArgument(stepId, &currentStep);
            interface.RunKernel(jitter);
            interface.RunKernel(zero);
            const auto& pairs = interface.GetCurrentConfiguration().GetPairs();
            const uint64_t wgx = ParameterPair::GetParameterValue<uint64_t>(pairs, "WORK_GROUP_SIZE_X");
            const uint64_t wgy = ParameterPair::GetParameterValue<uint64_t>(pairs, "WORK_GROUP_SIZE_Y");
            const uint64_t px = ParameterPair::GetParameterValue<uint64_t>(pairs, "POINTS_X");
            const uint64_t rows = ParameterPair::GetParameterValue<uint64_t>(pairs, "ROWS_PER_THREAD");
            const uint64_t blocks = ParameterPair::GetParameterValue<uint64_t>(pairs, "ATOM_BLOCKS");
            const uint64_t tiles = CeilDiv(side, wgx * px) * CeilDiv(side * side, wgy * rows);
            // The kernel writes blocks * tiles * threads * points float2 partials and one counter per tile.
            if (tiles * sizeof(unsigned int) > countersBytes
                || blocks * tiles * (wgx * wgy) * (px * rows) * 2 * sizeof(float) > partialsBytes)
            {
                throw KttException("the dynamic run's scratch is too small for this configuration");
            }
            interface.RunKernel(coulomb);
        });

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
        tuner.InitializeConfigurationData(kernel);
        const uint64_t space = tuner.GetConfigurationsCount(kernel);
        if (space == 0)
        {
            return SetLastError("no configuration of the dynamic tuning space fits grid size " + std::to_string(gridSize));
        }
        if (spaceSize != nullptr)
        {
            *spaceSize = static_cast<int>(space);
        }
        const uint64_t budget = tuningBudget > 0 ? std::min<uint64_t>(static_cast<uint64_t>(tuningBudget), space) : space;

        std::vector<KernelResult> tuningResults;
        KernelConfiguration best;
        bool production = false;
        for (int step = 0; step < steps; ++step)
        {
            currentStep = step;
            const auto start = std::chrono::steady_clock::now();
            KernelResult result;
            const bool tuning = tuningResults.size() < budget;
            if (tuning)
            {
                result = tuner.TuneIteration(kernel, {});
                tuningResults.push_back(result);
            }
            else
            {
                if (!production)
                {
                    best = tuner.GetBestConfiguration(kernel);
                    production = true;
                }
                result = tuner.Run(kernel, best, {});
                if (!result.IsValid())
                {
                    return SetLastError("the best configuration failed in the production phase at step " + std::to_string(step));
                }
            }
            const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
            double kernelMs = -1.0;
            if (result.IsValid())
            {
                for (const auto& part : result.GetResults())
                {
                    if (part.GetKernelFunction() == "directCoulombSumCompensated")
                    {
                        kernelMs = static_cast<double>(part.GetDuration()) / 1.0e6;
                    }
                }
            }
            if (stepKernelMs != nullptr)
            {
                stepKernelMs[step] = kernelMs;
            }
            if (stepWallMs != nullptr)
            {
                stepWallMs[step] = wall;
            }
            if (stepTuned != nullptr)
            {
                stepTuned[step] = tuning ? 1 : 0;
            }
        }
        if (!production)
        {
            best = tuner.GetBestConfiguration(kernel);
        }

        if (testedCount != nullptr)
        {
            *testedCount = static_cast<int>(tuningResults.size());
        }
        if (resultsFile != nullptr && resultsFile[0] != '\0' && !tuningResults.empty())
        {
            tuner.SaveResults(tuningResults, resultsFile, OutputFormat::JSON,
                {{"workload", "dynamic"}, {"grid_side", std::to_string(gridSize)}, {"atoms", std::to_string(atomCount)},
                 {"steps", std::to_string(steps)}, {"sigma", std::to_string(sigma)}});
        }
        if (bestConfig != nullptr && bestConfigLength > 0)
        {
            std::string text;
            for (const auto& pair : best.GetPairs())
            {
                text += (text.empty() ? "" : ",") + pair.GetName() + "=" + pair.GetValueString();
            }
            std::snprintf(bestConfig, bestConfigLength, "%s", text.c_str());
        }
        Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        if (finalAtoms != nullptr)
        {
            Check(hipMemcpy(finalAtoms, buffers.atoms, atomBytes, hipMemcpyDeviceToHost), "hipMemcpy(final atoms)");
        }
        if (finalGrid != nullptr)
        {
            Check(hipMemcpy(finalGrid, buffers.grid, gridPoints * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy(final grid)");
        }
        return 0;
    }
    catch (const std::exception& exception)
    {
        return SetLastError(exception.what());
    }
}
