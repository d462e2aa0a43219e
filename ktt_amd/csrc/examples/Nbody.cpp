// N-body step tuned by ktt on the HIP backend: the counterpart of
// Examples/Nbody/Nbody.cpp:59-256 for this library's gfx950 kernel.
//
// Usage: Nbody [searcher: deterministic|random|mcmc] [max configurations, 0 = all] [bodies]
//              [results file, without extension]
// Defaults: random search of 40 configurations over 128K bodies (positions, masses and
// velocities uniform in [0, 20), seeded), results in NbodyOutput.json. Both outputs are
// validated against the reference kernel with SideBySideComparison at 0.001, as the reference
// example does.
#include <cstdio>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include <Ktt.h>
#include <KttNbody.h>
#include <ktt_hip.h>

int main(int argc, char** argv)
{
    const std::string searcher = argc > 1 ? argv[1] : "random";
    const uint64_t maxConfigurations = argc > 2 ? std::stoull(argv[2]) : 40;
    const int n = argc > 3 ? std::stoi(argv[3]) : 128 * 1024;
    const std::string output = argc > 4 ? argv[4] : "NbodyOutput";
    const float timeDelta = 0.001f;
    const float damping = 0.5f;
    const float softeningSqr = 0.1f * 0.1f;

    std::mt19937_64 generator(20240812);
    std::uniform_real_distribution<float> distribution(0.0f, 20.0f);
    std::vector<float> bodyInfo(4 * static_cast<size_t>(n)), velocity(4 * static_cast<size_t>(n), 0.0f);
    std::vector<float> posX(n), posY(n), posZ(n), mass(n), velX(n), velY(n), velZ(n);
    for (int i = 0; i < n; ++i)
    {
        posX[i] = bodyInfo[4 * i] = distribution(generator);
        posY[i] = bodyInfo[4 * i + 1] = distribution(generator);
        posZ[i] = bodyInfo[4 * i + 2] = distribution(generator);
        mass[i] = bodyInfo[4 * i + 3] = distribution(generator);
        velX[i] = velocity[4 * i] = distribution(generator);
        velY[i] = velocity[4 * i + 1] = distribution(generator);
        velZ[i] = velocity[4 * i + 2] = distribution(generator);
    }
    std::vector<float> newBodyInfo(4 * static_cast<size_t>(n), 0.0f), newVelocity(4 * static_cast<size_t>(n), 0.0f);

    ktt::Tuner tuner(0, 0, ktt::ComputeApi::HIP);
    tuner.SetGlobalSizeType(ktt::GlobalSizeType::CUDA);
    tuner.SetTimeUnit(ktt::TimeUnit::Microseconds);
    const std::string directory = ktt_hip_kernel_directory();
    const auto definition = tuner.AddKernelDefinitionFromFile("nbody_kernel", directory + "/Nbody.hip",
        ktt::DimensionVector(static_cast<size_t>(n), 1, 1), ktt::DimensionVector(1, 1, 1));
    const auto referenceDefinition = tuner.AddKernelDefinitionFromFile("nbody_kernel_reference", directory + "/NbodyReference.hip",
        ktt::DimensionVector(static_cast<size_t>(n) / 64, 1, 1), ktt::DimensionVector(64, 1, 1));
    const auto kernel = tuner.CreateSimpleKernel("Nbody", definition);
    const auto referenceKernel = tuner.CreateSimpleKernel("NbodyReference", referenceDefinition);

    const auto dt = tuner.AddArgumentScalar(timeDelta);
    const auto bodies = tuner.AddArgumentVector(bodyInfo, ktt::ArgumentAccessType::ReadOnly);
    const auto x = tuner.AddArgumentVector(posX, ktt::ArgumentAccessType::ReadOnly);
    const auto y = tuner.AddArgumentVector(posY, ktt::ArgumentAccessType::ReadOnly);
    const auto z = tuner.AddArgumentVector(posZ, ktt::ArgumentAccessType::ReadOnly);
    const auto m = tuner.AddArgumentVector(mass, ktt::ArgumentAccessType::ReadOnly);
    const auto newBodies = tuner.AddArgumentVector(newBodyInfo, ktt::ArgumentAccessType::WriteOnly);
    const auto v = tuner.AddArgumentVector(velocity, ktt::ArgumentAccessType::ReadOnly);
    const auto vx = tuner.AddArgumentVector(velX, ktt::ArgumentAccessType::ReadOnly);
    const auto vy = tuner.AddArgumentVector(velY, ktt::ArgumentAccessType::ReadOnly);
    const auto vz = tuner.AddArgumentVector(velZ, ktt::ArgumentAccessType::ReadOnly);
    const auto newV = tuner.AddArgumentVector(newVelocity, ktt::ArgumentAccessType::WriteOnly);
    const auto dampingId = tuner.AddArgumentScalar(damping);
    const auto softening = tuner.AddArgumentScalar(softeningSqr);
    const auto count = tuner.AddArgumentScalar(n);

    ktt::AddNbodyTuningSpace(tuner, kernel, definition, static_cast<uint64_t>(n));
    tuner.SetArguments(definition, {dt, bodies, x, y, z, m, newBodies, v, vx, vy, vz, newV, dampingId, softening, count});
    tuner.SetArguments(referenceDefinition, {dt, bodies, newBodies, v, newV, dampingId, softening});
    tuner.SetValidationMethod(ktt::ValidationMethod::SideBySideComparison, 0.001);
    tuner.SetReferenceKernel(newV, referenceKernel, ktt::KernelConfiguration());
    tuner.SetReferenceKernel(newBodies, referenceKernel, ktt::KernelConfiguration());
    if (searcher == "random")
    {
        tuner.SetSearcher(kernel, std::make_unique<ktt::RandomSearcher>());
    }
    else if (searcher == "mcmc")
    {
        tuner.SetSearcher(kernel, std::make_unique<ktt::McmcSearcher>());
    }

    std::unique_ptr<ktt::StopCondition> stop;
    if (maxConfigurations > 0)
    {
        stop = std::make_unique<ktt::ConfigurationCount>(maxConfigurations);
    }
    const auto results = tuner.Tune(kernel, std::move(stop));
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
    const double interactions = static_cast<double>(n) * n;
    std::printf("tested %zu configurations, %zu valid; best %.3f ms (%.3e interactions/s, %.1f TFLOP/s at 20 flop): %s\n",
        results.size(), valid, seconds * 1.0e3, interactions / seconds, interactions * 20.0 / seconds * 1.0e-12,
        best->GetConfiguration().GetString().c_str());
    return 0;
}
