// N-body step on gfx950: direct launches of the tuned and reference kernels and a tuning run
// through ktt::Tuner (include/ktt_hip.h section 7), the counterpart of Examples/Nbody/Nbody.cpp.
#include <algorithm>
#include <cstdio>
#include <memory>
#include <vector>

#include <Ktt.h>
#include <KttNbody.h>

#include "../../../include/ktt_hip.h"
#include "CoulombRuntime.h"

namespace ktt
{

void AddNbodyTuningSpace(Tuner& tuner, const KernelId kernel, const KernelDefinitionId definition, const uint64_t bodies)
{
    tuner.AddParameter(kernel, "WORK_GROUP_SIZE_X", std::vector<uint64_t>{64, 128, 256});
    tuner.AddParameter(kernel, "WORK_GROUP_SIZE_Y", std::vector<uint64_t>{1, 2, 4, 8, 16});
    tuner.AddParameter(kernel, "BODIES_PER_THREAD", std::vector<uint64_t>{1, 2, 4, 8});
    tuner.AddParameter(kernel, "ATOM_TILE", std::vector<uint64_t>{64, 128, 256, 512});
    tuner.AddParameter(kernel, "INNER_UNROLL_FACTOR", std::vector<uint64_t>{1, 2, 4, 8, 16});
    tuner.AddParameter(kernel, "USE_SOA", std::vector<uint64_t>{0, 1});
    tuner.AddParameter(kernel, "ISSUE_PHASES", std::vector<uint64_t>{0, 1});

    tuner.AddThreadModifier(kernel, {definition}, ModifierType::Local, ModifierDimension::X, "WORK_GROUP_SIZE_X",
        ModifierAction::Multiply);
    tuner.AddThreadModifier(kernel, {definition}, ModifierType::Local, ModifierDimension::Y, "WORK_GROUP_SIZE_Y",
        ModifierAction::Multiply);
    tuner.AddThreadModifier(kernel, {definition}, ModifierType::Global, ModifierDimension::X,
        std::vector<std::string>{"WORK_GROUP_SIZE_X", "BODIES_PER_THREAD"},
        [](const uint64_t size, const std::vector<uint64_t>& values) { return size / (values[0] * values[1]); });

    tuner.AddConstraint(kernel, {"WORK_GROUP_SIZE_X", "WORK_GROUP_SIZE_Y"},
        [](const std::vector<uint64_t>& v) { return v[0] * v[1] <= 1024; });
    tuner.AddConstraint(kernel, {"WORK_GROUP_SIZE_X", "BODIES_PER_THREAD"},
        [bodies](const std::vector<uint64_t>& v) { return bodies % (v[0] * v[1]) == 0; });
    tuner.AddConstraint(kernel, {"WORK_GROUP_SIZE_X", "WORK_GROUP_SIZE_Y", "BODIES_PER_THREAD", "ATOM_TILE"},
        [bodies](const std::vector<uint64_t>& v)
        {
            // The source tiles and the rows' partial accelerations share one LDS buffer.
            const uint64_t lds = std::max<uint64_t>(v[3] * v[1] * 16, (v[1] - 1) * 3 * v[2] * v[0] * 4);
            return bodies % (v[3] * v[1]) == 0 && lds <= 64 * 1024;
        });
    tuner.AddConstraint(kernel, {"INNER_UNROLL_FACTOR", "ATOM_TILE"}, [](const std::vector<uint64_t>& v) { return v[1] % v[0] == 0; });
}

namespace
{

// The fastest of the shapes measured at 128K bodies on MI355X, each timed after about 50 ms of
// steps at its own shape so that the engine clock has settled: 4.06 ms with issue phases (64 x 8
// threads with 128-body tiles 4.19 ms; the interleaved loop's best shape, 64 x 4 threads with
// 256-body tiles, 5.75 ms). n must be a multiple of 1,024. Sixteen source rows per workgroup put
// two workgroups and eight waves on every SIMD, and the phased loop's plain phases then find a
// partner in 89% of their quad-cycles (84% with eight rows).
CoulombConfig NbodyDefaults()
{
    CoulombConfig defaults;
    defaults.Set("WORK_GROUP_SIZE_X", "64");
    defaults.Set("WORK_GROUP_SIZE_Y", "16");
    defaults.Set("BODIES_PER_THREAD", "4");
    defaults.Set("ATOM_TILE", "64");
    defaults.Set("INNER_UNROLL_FACTOR", "16");
    defaults.Set("USE_SOA", "0");
    defaults.Set("ISSUE_PHASES", "1");
    return defaults;
}

void CheckBodies(const void* bodies, const void* velocity, const void* newBodies, const void* newVelocity, const int n)
{
    if (n <= 0)
    {
        throw KttException("The number of bodies must be positive");
    }
    if (bodies == nullptr || velocity == nullptr || newBodies == nullptr || newVelocity == nullptr)
    {
        throw KttException("Body and velocity buffers must not be null");
    }
}

} // namespace

} // namespace ktt

using namespace ktt;

extern "C" {

int ktt_nbody_step(const float* bodyInfo, const float* velocity, int n, float timeDelta, float damping, float softeningSqr,
    float* newBodyInfo, float* newVelocity, const char* config, void* stream, float* kernelMs)
{
    try
    {
        CheckBodies(bodyInfo, velocity, newBodyInfo, newVelocity, n);
        const CoulombConfig parsed = CoulombConfig::Parse(config, NbodyDefaults());
        if (parsed.Get("USE_SOA", 0) != 0)
        {
            return SetLastError("ktt_nbody_step takes float4 bodies; USE_SOA=1 is available through the tuner");
        }
        const long x = parsed.Get("WORK_GROUP_SIZE_X", 64);
        const long y = parsed.Get("WORK_GROUP_SIZE_Y", 4);
        const long perThread = parsed.Get("BODIES_PER_THREAD", 4);
        const long tile = parsed.Get("ATOM_TILE", 256);
        const long unroll = parsed.Get("INNER_UNROLL_FACTOR", 8);
        if (x <= 0 || y <= 0 || perThread <= 0 || tile <= 0 || unroll <= 0 || x * y > 1024 || x % 64 != 0)
        {
            return SetLastError("invalid N-body configuration " + parsed.ToString());
        }
        if (n % (x * perThread) != 0 || n % (tile * y) != 0 || tile % unroll != 0)
        {
            return SetLastError("the number of bodies (" + std::to_string(n) + ") must be a multiple of WORK_GROUP_SIZE_X * "
                "BODIES_PER_THREAD and of ATOM_TILE * WORK_GROUP_SIZE_Y");
        }
        if (std::max(tile * y * 16, (y - 1) * 3 * perThread * x * 4) > 64 * 1024)
        {
            return SetLastError("the configuration needs more than 64 KiB of LDS");
        }
        auto kernel = CoulombRuntime::Get().GetKernel("Nbody.hip", "nbody_kernel", parsed.Defines({}));
        const unsigned int grid[3] = {static_cast<unsigned int>(n / (x * perThread)), 1, 1};
        const unsigned int block[3] = {static_cast<unsigned int>(x), static_cast<unsigned int>(y), 1};
        const float* unused = nullptr;
        void* arguments[] = {&timeDelta, &bodyInfo, &unused, &unused, &unused, &unused, &newBodyInfo, &velocity, &unused, &unused,
            &unused, &newVelocity, &damping, &softeningSqr, &n};
        CoulombRuntime::Launch(*kernel, grid, block, arguments, static_cast<hipStream_t>(stream), kernelMs);
        return 0;
    }
    catch (const std::exception& failure)
    {
        return SetLastError(failure.what());
    }
}

int ktt_nbody_step_reference(const float* bodyInfo, const float* velocity, int n, float timeDelta, float damping,
    float softeningSqr, float* newBodyInfo, float* newVelocity, void* stream, float* kernelMs)
{
    try
    {
        CheckBodies(bodyInfo, velocity, newBodyInfo, newVelocity, n);
        if (n % 64 != 0)
        {
            return SetLastError("the reference N-body kernel needs a multiple of 64 bodies");
        }
        auto kernel = CoulombRuntime::Get().GetKernel("NbodyReference.hip", "nbody_kernel_reference", {});
        const unsigned int grid[3] = {static_cast<unsigned int>(n / 64), 1, 1};
        const unsigned int block[3] = {64, 1, 1};
        void* arguments[] = {&timeDelta, &bodyInfo, &newBodyInfo, &velocity, &newVelocity, &damping, &softeningSqr};
        CoulombRuntime::Launch(*kernel, grid, block, arguments, static_cast<hipStream_t>(stream), kernelMs);
        return 0;
    }
    catch (const std::exception& failure)
    {
        return SetLastError(failure.what());
    }
}

int ktt_nbody_tune(const float* hostBodyInfo, const float* hostVelocity, int n, float timeDelta, float damping, float softeningSqr,
    int searcher, int maxConfigurations, double tolerance, const char* resultsFile, char* bestConfig, size_t bestConfigLength,
    double* bestMs, int* validCount, int* testedCount)
{
    try
    {
        if (hostBodyInfo == nullptr || hostVelocity == nullptr || n <= 0 || n % 64 != 0)
        {
            return SetLastError("tuning needs host bodies and velocities and a positive multiple of 64 bodies");
        }
        const std::string directory = CoulombRuntime::Get().GetKernelDirectory();
        std::vector<float> bodyInfo(hostBodyInfo, hostBodyInfo + 4 * static_cast<size_t>(n));
        std::vector<float> velocity(hostVelocity, hostVelocity + 4 * static_cast<size_t>(n));
        std::vector<float> posX(n), posY(n), posZ(n), mass(n), velX(n), velY(n), velZ(n);
        for (int i = 0; i < n; ++i)
        {
            posX[i] = bodyInfo[4 * i];
            posY[i] = bodyInfo[4 * i + 1];
            posZ[i] = bodyInfo[4 * i + 2];
            mass[i] = bodyInfo[4 * i + 3];
            velX[i] = velocity[4 * i];
            velY[i] = velocity[4 * i + 1];
            velZ[i] = velocity[4 * i + 2];
        }
        std::vector<float> newBodyInfo(4 * static_cast<size_t>(n), 0.0f);
        std::vector<float> newVelocity(4 * static_cast<size_t>(n), 0.0f);

        Tuner tuner(0, 0, ComputeApi::HIP);
        tuner.SetGlobalSizeType(GlobalSizeType::CUDA);
        tuner.SetTimeUnit(TimeUnit::Microseconds);
        const KernelDefinitionId definition = tuner.AddKernelDefinitionFromFile("nbody_kernel", directory + "/Nbody.hip",
            DimensionVector(static_cast<size_t>(n), 1, 1), DimensionVector(1, 1, 1));
        const KernelDefinitionId referenceDefinition = tuner.AddKernelDefinitionFromFile("nbody_kernel_reference",
            directory + "/NbodyReference.hip", DimensionVector(static_cast<size_t>(n) / 64, 1, 1), DimensionVector(64, 1, 1));
        if (definition == InvalidKernelDefinitionId || referenceDefinition == InvalidKernelDefinitionId)
        {
            return SetLastError("unable to load the N-body kernel sources from " + directory);
        }
        const KernelId kernel = tuner.CreateSimpleKernel("Nbody", definition);
        const KernelId referenceKernel = tuner.CreateSimpleKernel("NbodyReference", referenceDefinition);

        const ArgumentId dtId = tuner.AddArgumentScalar(timeDelta);
        const ArgumentId bodyId = tuner.AddArgumentVector(bodyInfo, ArgumentAccessType::ReadOnly);
        const ArgumentId xId = tuner.AddArgumentVector(posX, ArgumentAccessType::ReadOnly);
        const ArgumentId yId = tuner.AddArgumentVector(posY, ArgumentAccessType::ReadOnly);
        const ArgumentId zId = tuner.AddArgumentVector(posZ, ArgumentAccessType::ReadOnly);
        const ArgumentId massId = tuner.AddArgumentVector(mass, ArgumentAccessType::ReadOnly);
        const ArgumentId newBodyId = tuner.AddArgumentVector(newBodyInfo, ArgumentAccessType::WriteOnly);
        const ArgumentId velocityId = tuner.AddArgumentVector(velocity, ArgumentAccessType::ReadOnly);
        const ArgumentId vxId = tuner.AddArgumentVector(velX, ArgumentAccessType::ReadOnly);
        const ArgumentId vyId = tuner.AddArgumentVector(velY, ArgumentAccessType::ReadOnly);
        const ArgumentId vzId = tuner.AddArgumentVector(velZ, ArgumentAccessType::ReadOnly);
        const ArgumentId newVelocityId = tuner.AddArgumentVector(newVelocity, ArgumentAccessType::WriteOnly);
        const ArgumentId dampingId = tuner.AddArgumentScalar(damping);
        const ArgumentId softeningId = tuner.AddArgumentScalar(softeningSqr);
        const ArgumentId countId = tuner.AddArgumentScalar(n);

        AddNbodyTuningSpace(tuner, kernel, definition, static_cast<uint64_t>(n));
        tuner.SetArguments(definition, {dtId, bodyId, xId, yId, zId, massId, newBodyId, velocityId, vxId, vyId, vzId, newVelocityId,
            dampingId, softeningId, countId});
        tuner.SetArguments(referenceDefinition, {dtId, bodyId, newBodyId, velocityId, newVelocityId, dampingId, softeningId});
        // Nbody.cpp:246-251: both outputs are validated against the reference kernel.
        tuner.SetValidationMethod(ValidationMethod::SideBySideComparison, tolerance);
        tuner.SetReferenceKernel(newVelocityId, referenceKernel, KernelConfiguration());
        tuner.SetReferenceKernel(newBodyId, referenceKernel, KernelConfiguration());
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
        if (resultsFile != nullptr && resultsFile[0] != '\0')
        {
            tuner.SaveResults(results, resultsFile, OutputFormat::JSON);
        }
        const KernelResult* best = nullptr;
        int valid = 0;
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
        if (validCount != nullptr)
        {
            *validCount = valid;
        }
        if (testedCount != nullptr)
        {
            *testedCount = static_cast<int>(results.size());
        }
        if (best != nullptr)
        {
            if (bestMs != nullptr)
            {
                *bestMs = static_cast<double>(best->GetKernelDuration()) * 1.0e-6;
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
        }
        return 0;
    }
    catch (const std::exception& failure)
    {
        return SetLastError(failure.what());
    }
}

} // extern "C"
