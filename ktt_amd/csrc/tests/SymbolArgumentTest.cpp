// GPU test program of symbol arguments (ktt::Tuner::AddArgumentSymbol, Source/Tuner.h:501-510)
// on the HIP engine: atoms given as one symbol argument fill a __constant__ array of the kernel,
// which then computes the Coulomb potential of a cubic grid. The grid goes to a file that the
// Python test compares with the oracle.
//
// Usage: ktt_symbol_test <kernel source> <atoms file: float32 x, y, z, q records> <grid side>
//                        <spacing> <output file>
// Exit status 0 when the run is valid and the argument checks hold; the checks are printed.
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include <Ktt.h>

namespace
{

constexpr int kSymbolAtoms = 256;  // SYMBOL_ATOMS of the kernel

struct SymbolAtoms
{
    float values[4 * kSymbolAtoms];
};

bool Check(const bool condition, const char* what)
{
    std::printf("%s: %s\n", condition ? "ok" : "FAILED", what);
    return condition;
}

} // namespace

int main(int argc, char** argv)
{
    if (argc != 6)
    {
        std::fprintf(stderr, "usage: %s <kernel> <atoms file> <grid side> <spacing> <output file>\n", argv[0]);
        return 2;
    }
    const std::string kernelFile = argv[1];
    const int gridSize = std::stoi(argv[3]);
    const float spacing = std::stof(argv[4]);

    std::ifstream input(argv[2], std::ios::binary);
    input.seekg(0, std::ios::end);
    const std::streamoff bytes = input.tellg();
    input.seekg(0);
    if (bytes <= 0 || bytes % 16 != 0 || bytes / 16 > kSymbolAtoms || gridSize <= 0 || gridSize % 16 != 0)
    {
        std::fprintf(stderr, "need 1 to %d atoms of 16 bytes and a grid side that is a multiple of 16\n", kSymbolAtoms);
        return 2;
    }
    SymbolAtoms packed{};
    input.read(reinterpret_cast<char*>(packed.values), bytes);
    const int atomCount = static_cast<int>(bytes / 16);

    ktt::Tuner tuner(0, 0, ktt::ComputeApi::HIP);
    tuner.SetGlobalSizeType(ktt::GlobalSizeType::CUDA);
    ktt::Tuner::SetLoggingLevel(ktt::LoggingLevel::Error);
    const ktt::KernelDefinitionId definition = tuner.AddKernelDefinitionFromFile("coulombFromSymbol", kernelFile,
        ktt::DimensionVector(gridSize / 16, gridSize / 16, gridSize), ktt::DimensionVector(16, 16, 1));
    const ktt::KernelId kernel = tuner.CreateSimpleKernel("CoulombFromSymbol", definition);

    bool passed = true;
    // On HIP the variable's name is required, and symbol names are unique
    // (Source/KernelArgument/KernelArgumentManager.cpp:122-133).
    passed &= Check(tuner.AddArgumentSymbol(packed) == ktt::InvalidArgumentId, "a symbol argument without a name is refused");
    const ktt::ArgumentId symbolId = tuner.AddArgumentSymbol(packed, "", "symbolAtoms");
    passed &= Check(symbolId != ktt::InvalidArgumentId, "the atoms are added as symbol symbolAtoms");
    passed &= Check(tuner.AddArgumentSymbol(packed, "", "symbolAtoms") == ktt::InvalidArgumentId, "a second symbolAtoms is refused");

    std::vector<float> grid(static_cast<size_t>(gridSize) * gridSize * gridSize, 0.0f);
    const ktt::ArgumentId countId = tuner.AddArgumentScalar(atomCount);
    const ktt::ArgumentId spacingId = tuner.AddArgumentScalar(spacing);
    const ktt::ArgumentId sideId = tuner.AddArgumentScalar(gridSize);
    const ktt::ArgumentId gridId = tuner.AddArgumentVector(grid, ktt::ArgumentAccessType::ReadWrite);
    // The symbol argument takes no parameter slot: its place in the list does not matter.
    tuner.SetArguments(definition, {countId, symbolId, spacingId, sideId, gridId});

    const ktt::KernelResult result = tuner.Run(kernel, ktt::KernelConfiguration(), {ktt::BufferOutputDescriptor(gridId, grid.data())});
    passed &= Check(result.IsValid(), "the kernel with the symbol argument runs");

    std::ofstream output(argv[5], std::ios::binary);
    output.write(reinterpret_cast<const char*>(grid.data()), static_cast<std::streamsize>(grid.size() * sizeof(float)));
    passed &= Check(static_cast<bool>(output), "the grid is written");
    return passed ? 0 : 1;
}
