#include <hip/hiprtc.h>

#include <dlfcn.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iterator>
#include <sstream>

#include <KttException.h>

#include "Common.h"
#include "HipCounterService.h"
#include "HipEngine.h"
#include "Kernel.h"

namespace ktt
{

void CheckHipError(const hipError_t value, const std::string& call)
{
    if (value != hipSuccess)
    {
        throw KttException("HIP engine encountered error: " + std::string(hipGetErrorString(value)) + " in " + call);
    }
}

static void CheckRtcError(const hiprtcResult value, const std::string& call)
{
    if (value != HIPRTC_SUCCESS)
    {
        throw KttException("hipRTC error: " + std::string(hiprtcGetErrorString(value)) + " in " + call,
            ExceptionReason::CompilerError);
    }
}

static std::vector<std::string> SplitOptions(const std::string& options)
{
    std::vector<std::string> result;
    std::istringstream stream(options);
    std::string token;
    while (stream >> token)
    {
        result.push_back(token);
    }
    return result;
}

// ---------------------------------------------------------------- HipCompiler

std::string HipCompiler::DeviceArchitecture(const int device)
{
    hipDeviceProp_t properties;
    if (hipGetDeviceProperties(&properties, device) != hipSuccess)
    {
        return "gfx950";
    }
    std::string name = properties.gcnArchName;
    const size_t colon = name.find(':');
    if (colon != std::string::npos)
    {
        name = name.substr(0, colon);
    }
    return name.empty() ? "gfx950" : name;
}

std::string HipCompiler::DeviceName(const int device)
{
    hipDeviceProp_t properties;
    if (hipGetDeviceProperties(&properties, device) == hipSuccess && properties.name[0] != '\0')
    {
        return properties.name;
    }
    char name[256] = {};
    if (hipDeviceGetName(name, static_cast<int>(sizeof(name)), device) == hipSuccess && name[0] != '\0')
    {
        return name;
    }
    return "AMD GPU (" + DeviceArchitecture(device) + ")";
}

std::vector<std::string> HipCompiler::DefaultOptions(const std::string& architecture)
{
    // Packed FP32 instructions are disabled for every kernel this library builds.
    return {"--offload-arch=" + architecture, "-O3", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
}

// ---------------------------------------------------------------- code-object cache
//
// With KTT_HIP_CACHE_DIR set, compiled code objects are kept on disk: <key>.co holds the code
// object and <key>.name the kernel's lowered name. The key hashes the source, the kernel name,
// the options, the hipRTC version and a fingerprint (name, size, modification time) of every
// file in the options' -I directories, so editing an included header invalidates the entry.
// Files are written under a temporary name and renamed, so concurrent writers never expose a
// partial entry.

namespace
{

uint64_t Fnv1a(const std::string& text, uint64_t hash)
{
    for (const unsigned char c : text)
    {
        hash ^= c;
        hash *= 1099511628211ull;
    }
    return hash;
}

std::string IncludeFingerprint(const std::vector<std::string>& options)
{
    std::string fingerprint;
    for (const auto& option : options)
    {
        if (option.rfind("-I", 0) != 0 || option.size() <= 2)
        {
            continue;
        }
        std::error_code error;
        std::vector<std::string> entries;
        for (const auto& entry : std::filesystem::directory_iterator(option.substr(2), error))
        {
            if (!entry.is_regular_file(error))
            {
                continue;
            }
            const auto time = std::filesystem::last_write_time(entry.path(), error).time_since_epoch().count();
            entries.push_back(entry.path().filename().string() + ":" + std::to_string(entry.file_size(error)) + ":"
                + std::to_string(static_cast<long long>(time)));
        }
        std::sort(entries.begin(), entries.end());
        for (const auto& entry : entries)
        {
            fingerprint += entry + ";";
        }
    }
    return fingerprint;
}

// The compiler's identity: hipRTC's version and the real path, size and modification time of the
// hipRTC library and of the code-object manager beside it (which holds the compiler proper), so
// that a patch release with a changed compiler and the same version number misses the cache.
std::string CompilerIdentity()
{
    static const std::string identity = []()
    {
        int major = 0;
        int minor = 0;
        hiprtcVersion(&major, &minor);
        std::string text = "hiprtc " + std::to_string(major) + "." + std::to_string(minor);
        Dl_info info;
        if (dladdr(reinterpret_cast<void*>(&hiprtcVersion), &info) == 0 || info.dli_fname == nullptr)
        {
            return text;
        }
        const std::filesystem::path library = info.dli_fname;
        for (const auto& path : {library, library.parent_path() / "libamd_comgr.so"})
        {
            std::error_code error;
            const auto real = std::filesystem::canonical(path, error);
            if (error)
            {
                continue;
            }
            const auto size = std::filesystem::file_size(real, error);
            const auto time = std::filesystem::last_write_time(real, error).time_since_epoch().count();
            text += ";" + real.string() + ":" + std::to_string(size) + ":" + std::to_string(static_cast<long long>(time));
        }
        return text;
    }();
    return identity;
}

std::string CacheKey(const std::string& source, const std::string& kernelName, const std::vector<std::string>& options)
{
    std::string material = source + '\0' + kernelName + '\0';
    for (const auto& option : options)
    {
        material += option + '\0';
    }
    material += CompilerIdentity() + '\0' + IncludeFingerprint(options);
    char key[40];
    std::snprintf(key, sizeof(key), "%016llx%016llx", static_cast<unsigned long long>(Fnv1a(material, 1469598103934665603ull)),
        static_cast<unsigned long long>(Fnv1a(material, 0x9E3779B97F4A7C15ull)));
    return key;
}

bool ReadFile(const std::string& path, std::string& contents)
{
    std::ifstream stream(path, std::ios::binary);
    if (!stream)
    {
        return false;
    }
    contents.assign(std::istreambuf_iterator<char>(stream), std::istreambuf_iterator<char>());
    return true;
}

void WriteFileAtomically(const std::string& path, const char* data, const size_t size)
{
    const std::string temporary = path + ".tmp" + std::to_string(static_cast<long long>(getpid()));
    {
        std::ofstream stream(temporary, std::ios::binary | std::ios::trunc);
        if (!stream)
        {
            return;
        }
        stream.write(data, static_cast<std::streamsize>(size));
        if (!stream)
        {
            std::remove(temporary.c_str());
            return;
        }
    }
    std::rename(temporary.c_str(), path.c_str());
}

} // namespace

std::vector<char> HipCompiler::Compile(const std::string& source, const std::string& programName, const std::string& kernelName,
    const std::vector<std::string>& options, std::string& loweredName, std::string* log)
{
    std::vector<std::string> unused;
    return Compile(source, programName, kernelName, options, {}, loweredName, unused, log);
}

std::vector<char> HipCompiler::Compile(const std::string& source, const std::string& programName, const std::string& kernelName,
    const std::vector<std::string>& options, const std::vector<std::string>& symbols, std::string& loweredName,
    std::vector<std::string>& loweredSymbols, std::string* log)
{
    // The cache's .name file holds the kernel's lowered name, then one line per symbol.
    const char* cacheDirectory = std::getenv("KTT_HIP_CACHE_DIR");
    std::string cachePath;
    if (cacheDirectory != nullptr && cacheDirectory[0] != '\0')
    {
        std::error_code error;
        std::filesystem::create_directories(cacheDirectory, error);
        std::vector<std::string> keyOptions = options;
        for (const auto& symbol : symbols)
        {
            keyOptions.push_back("symbol:" + symbol);
        }
        cachePath = std::string(cacheDirectory) + "/" + CacheKey(source, kernelName, keyOptions);
        std::string code;
        std::string names;
        if (ReadFile(cachePath + ".name", names) && ReadFile(cachePath + ".co", code) && !code.empty() && !names.empty())
        {
            std::vector<std::string> lines;
            std::stringstream stream(names);
            for (std::string line; std::getline(stream, line);)
            {
                lines.push_back(line);
            }
            if (lines.size() == symbols.size() + 1 && !lines[0].empty())
            {
                loweredName = lines[0];
                loweredSymbols.assign(lines.begin() + 1, lines.end());
                if (log != nullptr)
                {
                    log->clear();
                }
                return std::vector<char>(code.begin(), code.end());
            }
        }
    }
    std::vector<char> code = CompileWithRtc(source, programName, kernelName, options, loweredName, log, symbols, &loweredSymbols);
    if (!cachePath.empty())
    {
        std::string names = loweredName;
        for (const auto& symbol : loweredSymbols)
        {
            names += "\n" + symbol;
        }
        WriteFileAtomically(cachePath + ".co", code.data(), code.size());
        WriteFileAtomically(cachePath + ".name", names.data(), names.size());
    }
    return code;
}

std::vector<char> HipCompiler::CompileWithRtc(const std::string& source, const std::string& programName, const std::string& kernelName,
    const std::vector<std::string>& options, std::string& loweredName, std::string* log, const std::vector<std::string>& symbols,
    std::vector<std::string>* loweredSymbols)
{
    hiprtcProgram program;
    CheckRtcError(hiprtcCreateProgram(&program, source.c_str(), programName.c_str(), 0, nullptr, nullptr), "hiprtcCreateProgram");

    struct ProgramGuard
    {
        hiprtcProgram& program;
        ~ProgramGuard() { hiprtcDestroyProgram(&program); }
    } guard{program};

    const bool templated = kernelName.find('<') != std::string::npos;
    if (templated)
    {
        CheckRtcError(hiprtcAddNameExpression(program, kernelName.c_str()), "hiprtcAddNameExpression");
    }
    // Global variables of symbol arguments are looked up by their lowered names after compilation
    // (as Source/ComputeEngine/Cuda/CudaProgram.cpp:39-41 does with NVRTC); "&name" is the name
    // expression of a variable.
    for (const auto& symbol : symbols)
    {
        CheckRtcError(hiprtcAddNameExpression(program, ("&" + symbol).c_str()), "hiprtcAddNameExpression(" + symbol + ")");
    }

    std::vector<const char*> rawOptions;
    for (const auto& option : options)
    {
        rawOptions.push_back(option.c_str());
    }

    const hiprtcResult result = hiprtcCompileProgram(program, static_cast<int>(rawOptions.size()), rawOptions.data());
    size_t logSize = 0;
    hiprtcGetProgramLogSize(program, &logSize);
    std::string programLog;
    if (logSize > 1)
    {
        programLog.resize(logSize);
        hiprtcGetProgramLog(program, programLog.data());
        programLog.resize(std::strlen(programLog.c_str()));
    }
    if (log != nullptr)
    {
        *log = programLog;
    }
    if (result != HIPRTC_SUCCESS)
    {
        throw KttException("hipRTC compilation of " + kernelName + " failed: " + std::string(hiprtcGetErrorString(result)) + "\n"
            + programLog, ExceptionReason::CompilerError);
    }

    if (templated)
    {
        const char* lowered = nullptr;
        CheckRtcError(hiprtcGetLoweredName(program, kernelName.c_str(), &lowered), "hiprtcGetLoweredName");
        loweredName = lowered;
    }
    else
    {
        loweredName = kernelName;
    }
    if (loweredSymbols != nullptr)
    {
        loweredSymbols->clear();
        for (const auto& symbol : symbols)
        {
            const char* lowered = nullptr;
            CheckRtcError(hiprtcGetLoweredName(program, ("&" + symbol).c_str(), &lowered), "hiprtcGetLoweredName(" + symbol + ")");
            loweredSymbols->push_back(lowered);
        }
    }

    size_t codeSize = 0;
    CheckRtcError(hiprtcGetCodeSize(program, &codeSize), "hiprtcGetCodeSize");
    std::vector<char> code(codeSize);
    CheckRtcError(hiprtcGetCode(program, code.data()), "hiprtcGetCode");
    return code;
}

// ---------------------------------------------------------------- HipKernel

HipKernel::HipKernel(const std::string& name, const std::vector<char>& codeObject, const std::string& loweredName,
    const Nanoseconds compilationTime) :
    m_Name(name),
    m_Module(nullptr),
    m_Function(nullptr),
    m_CompilationTime(compilationTime)
{
    CheckHipError(hipModuleLoadData(&m_Module, codeObject.data()), "hipModuleLoadData");
    const hipError_t status = hipModuleGetFunction(&m_Function, m_Module, loweredName.c_str());
    if (status != hipSuccess)
    {
        IgnoreHipError(hipModuleUnload(m_Module));
        throw KttException("Kernel function " + loweredName + " was not found in the compiled program", ExceptionReason::CompilerError);
    }

    int value = 0;
    if (hipFuncGetAttribute(&value, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK, m_Function) == hipSuccess)
    {
        m_CompilationData.m_MaxWorkGroupSize = static_cast<uint64_t>(value);
    }
    if (hipFuncGetAttribute(&value, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, m_Function) == hipSuccess)
    {
        m_CompilationData.m_LocalMemorySize = static_cast<uint64_t>(value);
    }
    if (hipFuncGetAttribute(&value, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, m_Function) == hipSuccess)
    {
        m_CompilationData.m_PrivateMemorySize = static_cast<uint64_t>(value);
    }
    if (hipFuncGetAttribute(&value, HIP_FUNC_ATTRIBUTE_CONST_SIZE_BYTES, m_Function) == hipSuccess)
    {
        m_CompilationData.m_ConstantMemorySize = static_cast<uint64_t>(value);
    }
    if (hipFuncGetAttribute(&value, HIP_FUNC_ATTRIBUTE_NUM_REGS, m_Function) == hipSuccess)
    {
        m_CompilationData.m_RegistersCount = static_cast<uint64_t>(value);
    }
}

HipKernel::~HipKernel()
{
    if (m_Module != nullptr)
    {
        IgnoreHipError(hipModuleUnload(m_Module));
    }
}

void HipKernel::FillSymbol(const std::string& loweredName, const void* data, const size_t bytes) const
{
    hipDeviceptr_t address = nullptr;
    size_t size = 0;
    const hipError_t status = hipModuleGetGlobal(&address, &size, m_Module, loweredName.c_str());
    if (status != hipSuccess || address == nullptr)
    {
        throw KttException("Symbol " + loweredName + " was not found in the module of " + m_Name);
    }
    if (size < bytes)
    {
        throw KttException("Symbol " + loweredName + " holds " + std::to_string(size) + " bytes, its argument " + std::to_string(bytes));
    }
    CheckHipError(hipMemcpyHtoD(address, const_cast<void*>(data), bytes), "hipMemcpyHtoD(symbol " + loweredName + ")");
}

// ---------------------------------------------------------------- HipEngine

HipEngine::HipEngine(const PlatformIndex platformIndex, const DeviceIndex deviceIndex, const uint32_t queueCount) :
    m_DeviceIndex(static_cast<int>(deviceIndex)),
    m_NextQueueId(0),
    m_NextComputeActionId(0),
    m_NextTransferActionId(0),
    m_KernelCacheCapacity(10),
    m_OverrideDefaultOptions(false),
    m_GlobalSizeType(GlobalSizeType::CUDA),
    m_GlobalSizeCorrection(false),
    m_Profiling(false),
    m_ClockWarmupMs(50.0)
{
    if (platformIndex != 0)
    {
        throw KttException("Invalid platform index: " + std::to_string(platformIndex) + ", HIP exposes a single platform");
    }
    if (queueCount == 0)
    {
        throw KttException("Number of compute queues must be greater than zero");
    }
    Initialize();
    for (uint32_t i = 0; i < queueCount; ++i)
    {
        hipStream_t stream;
        CheckHipError(hipStreamCreate(&stream), "hipStreamCreate");
        m_Streams[m_NextQueueId] = stream;
        m_StreamOwned[m_NextQueueId] = true;
        ++m_NextQueueId;
    }
}

HipEngine::HipEngine(const ComputeApiInitializer& initializer, std::vector<QueueId>& assignedQueueIds) :
    m_DeviceIndex(0),
    m_NextQueueId(0),
    m_NextComputeActionId(0),
    m_NextTransferActionId(0),
    m_KernelCacheCapacity(10),
    m_OverrideDefaultOptions(false),
    m_GlobalSizeType(GlobalSizeType::CUDA),
    m_GlobalSizeCorrection(false),
    m_Profiling(false),
    m_ClockWarmupMs(50.0)
{
    if (initializer.GetQueues().empty())
    {
        throw KttException("Compute API initializer must contain at least one queue");
    }
    CheckHipError(hipGetDevice(&m_DeviceIndex), "hipGetDevice");
    Initialize();
    for (auto* queue : initializer.GetQueues())
    {
        m_Streams[m_NextQueueId] = static_cast<hipStream_t>(queue);
        m_StreamOwned[m_NextQueueId] = false;
        assignedQueueIds.push_back(m_NextQueueId);
        ++m_NextQueueId;
    }
}

void HipEngine::Initialize()
{
    int count = 0;
    const hipError_t status = hipGetDeviceCount(&count);
    if (status != hipSuccess || count == 0)
    {
        throw KttException("No HIP device is available (hipGetDeviceCount: " + std::string(hipGetErrorString(status)) + ")");
    }
    if (m_DeviceIndex < 0 || m_DeviceIndex >= count)
    {
        throw KttException("Invalid device index: " + std::to_string(m_DeviceIndex));
    }
    CheckHipError(hipSetDevice(m_DeviceIndex), "hipSetDevice");
    CheckHipError(hipGetDeviceProperties(&m_Properties, m_DeviceIndex), "hipGetDeviceProperties");
    m_Architecture = HipCompiler::DeviceArchitecture(m_DeviceIndex);
    Logger::LogInfo("Initializing HIP engine on device " + HipCompiler::DeviceName(m_DeviceIndex) + " (" + m_Architecture + ")");
    if (const char* warmup = std::getenv("KTT_HIP_CLOCK_WARMUP_MS"))
    {
        char* end = nullptr;
        const double value = std::strtod(warmup, &end);
        if (end != warmup && value >= 0.0)
        {
            m_ClockWarmupMs = value;
        }
    }
}

// The source of the clock warm-up kernel: two FMA chains per thread, no memory traffic (the store
// never happens: the chains stay near 1 and 2).
static const char* const kClockWarmupSource = R"(
extern "C" __global__ void kttClockWarmup(float* out, int rounds)
{
    float a = 1.0f + 1.0e-3f * static_cast<float>(threadIdx.x & 15u);
    float b = 2.0f + 1.0e-3f * static_cast<float>(threadIdx.x & 7u);
    for (int i = 0; i < rounds; ++i)
    {
        a = fmaf(a, 0.9999f, 1.0e-4f);
        b = fmaf(b, 0.9999f, 2.0e-4f);
    }
    if (a + b == 123456.0f)
    {
        out[threadIdx.x] = a;
    }
}
)";

void HipEngine::WarmClock(hipStream_t stream)
{
    if (m_ClockWarmupMs <= 0.0)
    {
        return;
    }
    if (m_WarmupKernel == nullptr)
    {
        std::string lowered;
        Timer timer;
        timer.Start();
        const auto code = HipCompiler::Compile(kClockWarmupSource, "ktt_clock_warmup.hip", "kttClockWarmup",
            HipCompiler::DefaultOptions(m_Architecture), lowered);
        timer.Stop();
        m_WarmupKernel = std::make_shared<HipKernel>("kttClockWarmup", code, lowered, timer.GetElapsedTime());
        CheckHipError(hipMalloc(&m_WarmupOut, 256 * sizeof(float)), "hipMalloc(clock warm-up)");
    }
    int rounds = 4096;
    void* arguments[] = {&m_WarmupOut, &rounds};
    const unsigned int groups = static_cast<unsigned int>(std::max(m_Properties.multiProcessorCount, 1)) * 4u;
    const auto start = std::chrono::steady_clock::now();
    const auto length = std::chrono::duration<double, std::milli>(m_ClockWarmupMs);
    while (std::chrono::steady_clock::now() - start < length)
    {
        CheckHipError(hipModuleLaunchKernel(m_WarmupKernel->GetFunction(), groups, 1, 1, 256, 1, 1, 0, stream, arguments, nullptr),
            "hipModuleLaunchKernel(clock warm-up)");
        CheckHipError(hipStreamSynchronize(stream), "hipStreamSynchronize(clock warm-up)");
    }
}

HipEngine::~HipEngine()
{
    IgnoreHipError(hipSetDevice(m_DeviceIndex));
    for (auto& [id, action] : m_ComputeActions)
    {
        IgnoreHipError(hipEventDestroy(action.start));
        IgnoreHipError(hipEventDestroy(action.end));
    }
    for (auto& [id, action] : m_TransferActions)
    {
        IgnoreHipError(hipEventDestroy(action.start));
        IgnoreHipError(hipEventDestroy(action.end));
    }
    ClearBuffers();
    m_KernelCache.clear();
    m_WarmupKernel.reset();
    if (m_WarmupOut != nullptr)
    {
        IgnoreHipError(hipFree(m_WarmupOut));
        m_WarmupOut = nullptr;
    }
    for (auto& [id, stream] : m_Streams)
    {
        if (m_StreamOwned[id])
        {
            IgnoreHipError(hipStreamDestroy(stream));
        }
    }
}

std::vector<std::string> HipEngine::BuildOptions(const KernelComputeData& data) const
{
    std::vector<std::string> options;
    if (!m_OverrideDefaultOptions)
    {
        options = HipCompiler::DefaultOptions(m_Architecture);
    }
    for (const auto& option : SplitOptions(m_CompilerOptions))
    {
        options.push_back(option);
    }
    if (!data.GetSourceDirectory().empty())
    {
        options.push_back("-I" + data.GetSourceDirectory());
    }
    for (const auto& define : data.GetConfiguration().GenerateDefines())
    {
        options.push_back(define);
    }
    return options;
}

std::shared_ptr<HipKernel> HipEngine::LoadKernel(const KernelComputeData& data, bool& compiled)
{
    const KernelComputeId id = data.GetUniqueIdentifier();
    compiled = false;

    for (auto iterator = m_KernelCache.begin(); iterator != m_KernelCache.end(); ++iterator)
    {
        if (iterator->first == id)
        {
            auto kernel = iterator->second;
            m_KernelCache.splice(m_KernelCache.begin(), m_KernelCache, iterator);
            return kernel;
        }
    }

    // Symbol arguments are copied into their variables once per loaded module, as the reference
    // does when it creates a kernel (Source/ComputeEngine/Cuda/CudaProgram.cpp:67-78).
    std::vector<const KernelArgument*> symbolArguments;
    std::vector<std::string> symbolNames;
    for (const auto* argument : data.GetArguments())
    {
        if (argument->GetMemoryType() == ArgumentMemoryType::Symbol)
        {
            symbolArguments.push_back(argument);
            symbolNames.push_back(argument->GetSymbolName());
        }
    }
    Timer timer;
    timer.Start();
    std::string loweredName;
    std::vector<std::string> loweredSymbols;
    const auto code = HipCompiler::Compile(data.GetDefaultSource(), data.GetName() + ".hip", data.GetTemplatedName(),
        BuildOptions(data), symbolNames, loweredName, loweredSymbols);
    timer.Stop();
    auto kernel = std::make_shared<HipKernel>(data.GetName(), code, loweredName, timer.GetElapsedTime());
    for (size_t i = 0; i < symbolArguments.size(); ++i)
    {
        kernel->FillSymbol(loweredSymbols[i], symbolArguments[i]->GetData(), symbolArguments[i]->GetDataSize());
    }
    compiled = true;

    if (m_KernelCacheCapacity > 0)
    {
        m_KernelCache.emplace_front(id, kernel);
        while (m_KernelCache.size() > m_KernelCacheCapacity)
        {
            m_KernelCache.pop_back();
        }
    }
    return kernel;
}

ComputeActionId HipEngine::RunKernelAsync(const KernelComputeData& data, const QueueId queueId, [[maybe_unused]] const bool powerMeasurementAllowed)
{
    if (!ContainsKey(m_Streams, queueId))
    {
        throw KttException("Invalid stream index: " + std::to_string(queueId));
    }

    const auto& localSize = data.GetLocalSize();
    const uint64_t localTotal = static_cast<uint64_t>(localSize.GetTotalSize());
    if (localTotal > static_cast<uint64_t>(m_Properties.maxThreadsPerBlock) || localTotal == 0)
    {
        throw KttException("Block size of " + std::to_string(localTotal) + " exceeds current device limit",
            ExceptionReason::DeviceLimitsExceeded);
    }

    Timer timer;
    timer.Start();
    bool compiled = false;
    auto kernel = LoadKernel(data, compiled);
    if (compiled)
    {
        WarmClock(m_Streams[queueId]);
    }

    // Scalars by host pointer, vectors by device-pointer slot, local arguments as dynamic LDS.
    std::vector<void*> devicePointers;
    devicePointers.reserve(data.GetArguments().size());
    std::vector<void*> parameters;
    size_t dynamicLds = 0;
    for (auto* argument : data.GetArguments())
    {
        switch (argument->GetMemoryType())
        {
        case ArgumentMemoryType::Scalar:
            parameters.push_back(argument->GetData());
            break;
        case ArgumentMemoryType::Vector:
        {
            void* pointer = nullptr;
            if (argument->HasUserBuffer())
            {
                pointer = argument->GetUserBuffer();
            }
            else
            {
                pointer = GetBuffer(argument->GetId()).devicePointer;
            }
            devicePointers.push_back(pointer);
            break;
        }
        case ArgumentMemoryType::Local:
            dynamicLds += argument->GetDataSize();
            break;
        case ArgumentMemoryType::Symbol:
            break;  // in the module's variable since the kernel was loaded; no parameter slot
        default:
            throw KttException("Unhandled argument memory type");
        }
    }
    // Second pass: parameter slots in argument order (device pointer storage is now stable).
    parameters.clear();
    size_t vectorIndex = 0;
    for (auto* argument : data.GetArguments())
    {
        if (argument->GetMemoryType() == ArgumentMemoryType::Scalar)
        {
            parameters.push_back(argument->GetData());
        }
        else if (argument->GetMemoryType() == ArgumentMemoryType::Vector)
        {
            parameters.push_back(&devicePointers[vectorIndex++]);
        }
    }

    const size_t totalLds = dynamicLds + kernel->GetStaticLdsSize();
    if (totalLds > m_Properties.sharedMemPerBlock)
    {
        throw KttException("Shared memory usage of " + std::to_string(totalLds) + " bytes exceeds current device limit",
            ExceptionReason::DeviceLimitsExceeded);
    }

    DimensionVector grid = data.GetGlobalSize();
    if (m_GlobalSizeType != GlobalSizeType::CUDA)
    {
        if (m_GlobalSizeCorrection)
        {
            grid.RoundUp(localSize);
        }
        if (grid.GetSizeX() % localSize.GetSizeX() != 0 || grid.GetSizeY() % localSize.GetSizeY() != 0
            || grid.GetSizeZ() % localSize.GetSizeZ() != 0)
        {
            throw KttException("Global size " + grid.GetString() + " is not a multiple of local size " + localSize.GetString());
        }
        grid.Divide(localSize);
    }
    if (grid.GetTotalSize() == 0)
    {
        throw KttException("Launch with empty grid " + grid.GetString());
    }

    hipStream_t stream = m_Streams[queueId];
    ComputeAction action;
    action.kernel = kernel;
    action.globalSize = data.GetGlobalSize();
    action.localSize = localSize;
    action.compilationOverhead = compiled ? kernel->GetCompilationTime() : 0;
    CheckHipError(hipEventCreate(&action.start), "hipEventCreate");
    CheckHipError(hipEventCreate(&action.end), "hipEventCreate");
    timer.Stop();

    CheckHipError(hipEventRecord(action.start, stream), "hipEventRecord");
    const hipError_t launch = hipModuleLaunchKernel(kernel->GetFunction(), static_cast<unsigned int>(grid.GetSizeX()),
        static_cast<unsigned int>(grid.GetSizeY()), static_cast<unsigned int>(grid.GetSizeZ()),
        static_cast<unsigned int>(localSize.GetSizeX()), static_cast<unsigned int>(localSize.GetSizeY()),
        static_cast<unsigned int>(localSize.GetSizeZ()), static_cast<unsigned int>(dynamicLds), stream, parameters.data(), nullptr);
    if (launch != hipSuccess)
    {
        IgnoreHipError(hipEventDestroy(action.start));
        IgnoreHipError(hipEventDestroy(action.end));
        CheckHipError(launch, "hipModuleLaunchKernel(" + data.GetName() + ")");
    }
    CheckHipError(hipEventRecord(action.end, stream), "hipEventRecord");

    action.overhead = timer.GetElapsedTime();
    const ComputeActionId id = m_NextComputeActionId++;
    m_ComputeActions[id] = std::move(action);
    return id;
}

ComputationResult HipEngine::WaitForComputeAction(const ComputeActionId id)
{
    auto iterator = m_ComputeActions.find(id);
    if (iterator == m_ComputeActions.end())
    {
        throw KttException("Compute action with id " + std::to_string(id) + " was not found");
    }
    ComputeAction action = iterator->second;
    m_ComputeActions.erase(iterator);

    const hipError_t status = hipEventSynchronize(action.end);
    float milliseconds = 0.0f;
    const hipError_t elapsed = status == hipSuccess ? hipEventElapsedTime(&milliseconds, action.start, action.end) : status;
    IgnoreHipError(hipEventDestroy(action.start));
    IgnoreHipError(hipEventDestroy(action.end));
    CheckHipError(status, "hipEventSynchronize");
    CheckHipError(elapsed, "hipEventElapsedTime");

    ComputationResult result(action.kernel->GetName());
    result.SetDurationData(static_cast<Nanoseconds>(static_cast<double>(milliseconds) * 1.0e6), action.overhead,
        action.compilationOverhead);
    result.SetSizeData(action.globalSize, action.localSize);
    result.SetCompilationData(std::make_unique<KernelCompilationData>(action.kernel->GetCompilationData()));
    return result;
}

void HipEngine::ClearData(const KernelComputeId& id)
{
    m_KernelCache.remove_if([&id](const auto& entry) { return entry.first == id; });
}

void HipEngine::ClearKernelData(const std::string& kernelName)
{
    m_KernelCache.remove_if([&kernelName](const auto& entry) { return entry.second->GetName() == kernelName; });
}

void HipEngine::AddStaticCounters(const KernelComputeData& data, const KernelCompilationData& compilation,
    KernelProfilingData& profiling) const
{
    // Resource usage of the code object and the occupancy it allows (waves per SIMD).
    profiling.AddCounter(KernelProfilingCounter("vgpr_count", ProfilingCounterType::UnsignedInt, compilation.m_RegistersCount));
    profiling.AddCounter(KernelProfilingCounter("lds_bytes", ProfilingCounterType::UnsignedInt, compilation.m_LocalMemorySize));
    profiling.AddCounter(KernelProfilingCounter("scratch_bytes", ProfilingCounterType::UnsignedInt, compilation.m_PrivateMemorySize));
    const uint64_t threads = static_cast<uint64_t>(data.GetLocalSize().GetTotalSize());
    const uint64_t wavesPerGroup = (threads + 63) / 64;
    const uint64_t vgprs = std::max<uint64_t>(compilation.m_RegistersCount, 1);
    const uint64_t wavesByVgpr = std::min<uint64_t>(8, 512 / ((vgprs + 7) / 8 * 8));
    const uint64_t ldsPerGroup = compilation.m_LocalMemorySize;
    const uint64_t groupsByLds = ldsPerGroup == 0 ? 64 : static_cast<uint64_t>(m_Properties.maxSharedMemoryPerMultiProcessor) / ldsPerGroup;
    const uint64_t wavesByLds = groupsByLds * wavesPerGroup / 4;
    const uint64_t wavesPerSimd = std::max<uint64_t>(1, std::min(wavesByVgpr, std::max<uint64_t>(wavesByLds, 1)));
    profiling.AddCounter(KernelProfilingCounter("occupancy_waves_per_simd", ProfilingCounterType::UnsignedInt, wavesPerSimd));
}

ComputationResult HipEngine::RunWithStaticProfile(const KernelComputeData& data, const QueueId queueId)
{
    const ComputeActionId id = RunKernelAsync(data, queueId, false);
    ComputationResult result = WaitForComputeAction(id);
    auto profiling = std::make_unique<KernelProfilingData>();
    profiling->AddCounter(KernelProfilingCounter("kernel_duration_ns", ProfilingCounterType::UnsignedInt,
        static_cast<uint64_t>(result.GetDuration())));
    AddStaticCounters(data, result.GetCompilationData(), *profiling);
    result.SetProfilingData(std::move(profiling));
    return result;
}

void HipEngine::EnsurePlan(const std::vector<std::string>& counters)
{
    if (counters != m_PlannedCounters)
    {
        HipCounterService::Get().Plan(m_DeviceIndex, counters, m_PassOf);
        m_PlannedCounters = counters;
    }
}

ComputationResult HipEngine::RunKernelWithProfiling(const KernelComputeData& data, const QueueId queueId)
{
    HipCounterService& service = HipCounterService::Get();
    if (!service.IsRunning())
    {
        if (!m_ProfilingCounters.empty())
        {
            throw KttException("Profiling counters were requested but " + service.UnavailableReason());
        }
        // Without the counter service a profiled run is a single timed run annotated with the
        // static resource usage of the code object.
        return RunWithStaticProfile(data, queueId);
    }

    const KernelComputeId id = data.GetUniqueIdentifier();
    auto found = m_ProfilingSessions.find(id);
    if (found == m_ProfilingSessions.end())
    {
        ProfilingSession session;
        session.counters = m_ProfilingCounters.empty() ? HipCounterService::DefaultCounters() : m_ProfilingCounters;
        EnsurePlan(session.counters);
        session.passes = 1 + *std::max_element(m_PassOf.begin(), m_PassOf.end());
        session.values.assign(session.counters.size(), 0.0);
        found = m_ProfilingSessions.emplace(id, std::move(session)).first;
    }
    ProfilingSession& session = found->second;

    if (session.runsDone == 0)
    {
        // The timing run: the kernel's duration is taken without counter collection.
        ComputationResult result = WaitForComputeAction(RunKernelAsync(data, queueId, false));
        session.duration = result.GetDuration();
        session.runsDone = 1;
        result.SetProfilingData(std::make_unique<KernelProfilingData>(static_cast<uint64_t>(session.passes)));
        return result;
    }

    const size_t pass = session.runsDone - 1;
    ComputationResult result;
    try
    {
        EnsurePlan(session.counters);
        const auto& local = data.GetLocalSize();
        service.Arm(pass, static_cast<unsigned>(local.GetSizeX()), static_cast<unsigned>(local.GetSizeY()),
            static_cast<unsigned>(local.GetSizeZ()));
        result = WaitForComputeAction(RunKernelAsync(data, queueId, false));
        std::vector<double> values;
        service.Collect(values, 10000);
        for (size_t i = 0; i < session.counters.size(); ++i)
        {
            if (m_PassOf[i] == pass)
            {
                session.values[i] = values[i];
            }
        }
    }
    catch (const KttException&)
    {
        service.Disarm();
        m_ProfilingSessions.erase(found);
        throw;
    }
    ++session.runsDone;

    const uint64_t remaining = static_cast<uint64_t>(session.passes + 1 - session.runsDone);
    if (remaining > 0)
    {
        result.SetProfilingData(std::make_unique<KernelProfilingData>(remaining));
        return result;
    }
    auto profiling = std::make_unique<KernelProfilingData>();
    for (size_t i = 0; i < session.counters.size(); ++i)
    {
        profiling->AddCounter(KernelProfilingCounter(session.counters[i], ProfilingCounterType::Double, session.values[i]));
    }
    profiling->AddCounter(KernelProfilingCounter("kernel_duration_ns", ProfilingCounterType::UnsignedInt,
        static_cast<uint64_t>(session.duration)));
    AddStaticCounters(data, result.GetCompilationData(), *profiling);
    // The run reports the duration of the timing run; counter passes serialise the kernel.
    result.SetDurationData(session.duration, result.GetOverhead(), result.GetCompilationOverhead());
    result.SetProfilingData(std::move(profiling));
    m_ProfilingSessions.erase(found);
    return result;
}

void HipEngine::SetProfilingCounters(const std::vector<std::string>& counters)
{
    HipCounterService& service = HipCounterService::Get();
    if (!counters.empty())
    {
        if (!service.IsRunning())
        {
            throw KttException("Profiling counters cannot be collected: " + service.UnavailableReason());
        }
        EnsurePlan(counters);
    }
    m_ProfilingCounters = counters;
    m_ProfilingSessions.clear();
}

bool HipEngine::IsProfilingSessionActive(const KernelComputeId& id)
{
    return m_ProfilingSessions.count(id) != 0;
}

uint64_t HipEngine::GetRemainingProfilingRuns(const KernelComputeId& id)
{
    const auto found = m_ProfilingSessions.find(id);
    if (found == m_ProfilingSessions.end())
    {
        return 0;
    }
    return static_cast<uint64_t>(found->second.passes + 1 - found->second.runsDone);
}

bool HipEngine::HasAccurateRemainingProfilingRuns() const
{
    return true;
}

bool HipEngine::SupportsMultiInstanceProfiling() const
{
    return true;
}

bool HipEngine::IsProfilingActive() const
{
    return m_Profiling;
}

void HipEngine::SetProfiling(const bool profiling)
{
    m_Profiling = profiling;
}

// ---------------------------------------------------------------- buffers

HipEngine::Buffer& HipEngine::GetBuffer(const ArgumentId& id)
{
    auto iterator = m_Buffers.find(id);
    if (iterator == m_Buffers.end())
    {
        throw KttException("Buffer for argument with id " + id + " was not found");
    }
    return iterator->second;
}

void HipEngine::FreeBuffer(Buffer& buffer)
{
    if (buffer.userOwned)
    {
        return;
    }
    switch (buffer.location)
    {
    case ArgumentMemoryLocation::Host:
    case ArgumentMemoryLocation::HostZeroCopy:
        if (buffer.hostPointer != nullptr)
        {
            IgnoreHipError(hipHostFree(buffer.hostPointer));
        }
        break;
    default:
        if (buffer.devicePointer != nullptr)
        {
            IgnoreHipError(hipFree(buffer.devicePointer));
        }
        break;
    }
    buffer.devicePointer = nullptr;
    buffer.hostPointer = nullptr;
}

TransferActionId HipEngine::RecordTransfer(hipEvent_t start, hipEvent_t end, const Nanoseconds overhead)
{
    const TransferActionId id = m_NextTransferActionId++;
    TransferAction action;
    action.start = start;
    action.end = end;
    action.overhead = overhead;
    m_TransferActions[id] = action;
    return id;
}

TransferActionId HipEngine::UploadArgument(KernelArgument& kernelArgument, const QueueId queueId)
{
    if (kernelArgument.GetMemoryType() != ArgumentMemoryType::Vector)
    {
        throw KttException("Only vector arguments can be uploaded into buffers");
    }
    if (!ContainsKey(m_Streams, queueId))
    {
        throw KttException("Invalid stream index: " + std::to_string(queueId));
    }
    const ArgumentId& id = kernelArgument.GetId();
    if (HasBuffer(id))
    {
        ClearBuffer(id);
    }

    Timer timer;
    timer.Start();
    Buffer buffer;
    buffer.size = kernelArgument.GetDataSize();
    buffer.location = kernelArgument.GetMemoryLocation();

    if (kernelArgument.HasUserBuffer())
    {
        buffer.devicePointer = kernelArgument.GetUserBuffer();
        buffer.userOwned = true;
        m_Buffers[id] = buffer;
        timer.Stop();
        hipEvent_t start, end;
        CheckHipError(hipEventCreate(&start), "hipEventCreate");
        CheckHipError(hipEventCreate(&end), "hipEventCreate");
        CheckHipError(hipEventRecord(start, m_Streams[queueId]), "hipEventRecord");
        CheckHipError(hipEventRecord(end, m_Streams[queueId]), "hipEventRecord");
        return RecordTransfer(start, end, timer.GetElapsedTime());
    }

    const void* source = kernelArgument.GetData();
    switch (buffer.location)
    {
    case ArgumentMemoryLocation::Host:
    case ArgumentMemoryLocation::HostZeroCopy:
        CheckHipError(hipHostMalloc(&buffer.hostPointer, buffer.size, hipHostMallocMapped), "hipHostMalloc");
        CheckHipError(hipHostGetDevicePointer(&buffer.devicePointer, buffer.hostPointer, 0), "hipHostGetDevicePointer");
        break;
    case ArgumentMemoryLocation::Unified:
        CheckHipError(hipMallocManaged(&buffer.devicePointer, buffer.size), "hipMallocManaged");
        break;
    default:
        CheckHipError(hipMalloc(&buffer.devicePointer, buffer.size), "hipMalloc");
        break;
    }
    m_Buffers[id] = buffer;

    hipEvent_t start, end;
    CheckHipError(hipEventCreate(&start), "hipEventCreate");
    CheckHipError(hipEventCreate(&end), "hipEventCreate");
    hipStream_t stream = m_Streams[queueId];
    CheckHipError(hipEventRecord(start, stream), "hipEventRecord");
    if (source != nullptr)
    {
        if (buffer.hostPointer != nullptr)
        {
            std::memcpy(buffer.hostPointer, source, buffer.size);
        }
        else
        {
            CheckHipError(hipMemcpyAsync(buffer.devicePointer, source, buffer.size, hipMemcpyHostToDevice, stream), "hipMemcpyAsync");
        }
    }
    CheckHipError(hipEventRecord(end, stream), "hipEventRecord");
    timer.Stop();
    return RecordTransfer(start, end, timer.GetElapsedTime());
}

TransferActionId HipEngine::UpdateArgument(const ArgumentId& id, const QueueId queueId, const void* data, const size_t dataSize)
{
    auto& buffer = GetBuffer(id);
    const size_t size = dataSize == 0 ? buffer.size : dataSize;
    if (size > buffer.size)
    {
        throw KttException("Size of data to update is larger than the buffer of argument with id " + id);
    }
    Timer timer;
    timer.Start();
    hipStream_t stream = m_Streams.at(queueId);
    hipEvent_t start, end;
    CheckHipError(hipEventCreate(&start), "hipEventCreate");
    CheckHipError(hipEventCreate(&end), "hipEventCreate");
    CheckHipError(hipEventRecord(start, stream), "hipEventRecord");
    if (buffer.hostPointer != nullptr)
    {
        CheckHipError(hipStreamSynchronize(stream), "hipStreamSynchronize");
        std::memcpy(buffer.hostPointer, data, size);
    }
    else
    {
        CheckHipError(hipMemcpyAsync(buffer.devicePointer, data, size, hipMemcpyHostToDevice, stream), "hipMemcpyAsync");
    }
    CheckHipError(hipEventRecord(end, stream), "hipEventRecord");
    timer.Stop();
    return RecordTransfer(start, end, timer.GetElapsedTime());
}

TransferActionId HipEngine::DownloadArgument(const ArgumentId& id, const QueueId queueId, void* destination, const size_t dataSize)
{
    auto& buffer = GetBuffer(id);
    const size_t size = dataSize == 0 ? buffer.size : dataSize;
    if (size > buffer.size)
    {
        throw KttException("Size of data to download is larger than the buffer of argument with id " + id);
    }
    Timer timer;
    timer.Start();
    hipStream_t stream = m_Streams.at(queueId);
    hipEvent_t start, end;
    CheckHipError(hipEventCreate(&start), "hipEventCreate");
    CheckHipError(hipEventCreate(&end), "hipEventCreate");
    CheckHipError(hipEventRecord(start, stream), "hipEventRecord");
    if (buffer.hostPointer != nullptr)
    {
        CheckHipError(hipStreamSynchronize(stream), "hipStreamSynchronize");
        std::memcpy(destination, buffer.hostPointer, size);
    }
    else
    {
        CheckHipError(hipMemcpyAsync(destination, buffer.devicePointer, size, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
    }
    CheckHipError(hipEventRecord(end, stream), "hipEventRecord");
    timer.Stop();
    return RecordTransfer(start, end, timer.GetElapsedTime());
}

TransferActionId HipEngine::CopyArgument(const ArgumentId& destination, const QueueId queueId, const ArgumentId& source,
    const size_t dataSize)
{
    auto& target = GetBuffer(destination);
    auto& origin = GetBuffer(source);
    const size_t size = dataSize == 0 ? origin.size : dataSize;
    if (size > target.size || size > origin.size)
    {
        throw KttException("Size of data to copy exceeds one of the buffers " + source + ", " + destination);
    }
    Timer timer;
    timer.Start();
    hipStream_t stream = m_Streams.at(queueId);
    hipEvent_t start, end;
    CheckHipError(hipEventCreate(&start), "hipEventCreate");
    CheckHipError(hipEventCreate(&end), "hipEventCreate");
    CheckHipError(hipEventRecord(start, stream), "hipEventRecord");
    CheckHipError(hipMemcpyAsync(target.devicePointer, origin.devicePointer, size, hipMemcpyDeviceToDevice, stream), "hipMemcpyAsync");
    CheckHipError(hipEventRecord(end, stream), "hipEventRecord");
    timer.Stop();
    return RecordTransfer(start, end, timer.GetElapsedTime());
}

TransferResult HipEngine::WaitForTransferAction(const TransferActionId id)
{
    auto iterator = m_TransferActions.find(id);
    if (iterator == m_TransferActions.end())
    {
        throw KttException("Transfer action with id " + std::to_string(id) + " was not found");
    }
    TransferAction action = iterator->second;
    m_TransferActions.erase(iterator);
    const hipError_t status = hipEventSynchronize(action.end);
    float milliseconds = 0.0f;
    if (status == hipSuccess)
    {
        IgnoreHipError(hipEventElapsedTime(&milliseconds, action.start, action.end));
    }
    IgnoreHipError(hipEventDestroy(action.start));
    IgnoreHipError(hipEventDestroy(action.end));
    CheckHipError(status, "hipEventSynchronize");
    return TransferResult(static_cast<Nanoseconds>(static_cast<double>(milliseconds) * 1.0e6), action.overhead);
}

void HipEngine::ResizeArgument(const ArgumentId& id, const size_t newSize, const bool preserveData)
{
    auto& buffer = GetBuffer(id);
    if (buffer.userOwned)
    {
        throw KttException("Buffer of argument with id " + id + " is owned by the user and cannot be resized");
    }
    if (buffer.size == newSize)
    {
        return;
    }
    Buffer resized;
    resized.size = newSize;
    resized.location = buffer.location;
    if (buffer.hostPointer != nullptr)
    {
        CheckHipError(hipHostMalloc(&resized.hostPointer, newSize, hipHostMallocMapped), "hipHostMalloc");
        CheckHipError(hipHostGetDevicePointer(&resized.devicePointer, resized.hostPointer, 0), "hipHostGetDevicePointer");
    }
    else if (buffer.location == ArgumentMemoryLocation::Unified)
    {
        CheckHipError(hipMallocManaged(&resized.devicePointer, newSize), "hipMallocManaged");
    }
    else
    {
        CheckHipError(hipMalloc(&resized.devicePointer, newSize), "hipMalloc");
    }
    if (preserveData)
    {
        CheckHipError(hipDeviceSynchronize(), "hipDeviceSynchronize");
        CheckHipError(hipMemcpy(resized.devicePointer, buffer.devicePointer, std::min(newSize, buffer.size), hipMemcpyDefault),
            "hipMemcpy");
    }
    FreeBuffer(buffer);
    buffer = resized;
}

void HipEngine::GetUnifiedMemoryBufferHandle(const ArgumentId& id, UnifiedBufferMemory& handle)
{
    auto& buffer = GetBuffer(id);
    if (buffer.location != ArgumentMemoryLocation::Unified)
    {
        throw KttException("Buffer of argument with id " + id + " is not unified memory");
    }
    handle = buffer.devicePointer;
}

void HipEngine::AddCustomBuffer(KernelArgument& kernelArgument, ComputeBuffer buffer)
{
    Buffer entry;
    entry.devicePointer = buffer;
    entry.size = kernelArgument.GetDataSize();
    entry.location = kernelArgument.GetMemoryLocation();
    entry.userOwned = true;
    m_Buffers[kernelArgument.GetId()] = entry;
}

void HipEngine::ClearBuffer(const ArgumentId& id)
{
    auto iterator = m_Buffers.find(id);
    if (iterator == m_Buffers.end())
    {
        return;
    }
    IgnoreHipError(hipDeviceSynchronize());
    FreeBuffer(iterator->second);
    m_Buffers.erase(iterator);
}

void HipEngine::ClearBuffers()
{
    if (m_Buffers.empty())
    {
        return;
    }
    IgnoreHipError(hipDeviceSynchronize());
    for (auto& [id, buffer] : m_Buffers)
    {
        FreeBuffer(buffer);
    }
    m_Buffers.clear();
}

bool HipEngine::HasBuffer(const ArgumentId& id)
{
    return ContainsKey(m_Buffers, id);
}

void* HipEngine::GetDeviceBuffer(const ArgumentId& id)
{
    return GetBuffer(id).devicePointer;
}

// ---------------------------------------------------------------- queues

QueueId HipEngine::AddComputeQueue(ComputeQueue queue)
{
    for (const auto& [id, stream] : m_Streams)
    {
        if (stream == static_cast<hipStream_t>(queue))
        {
            throw KttException("Compute queue was already added");
        }
    }
    const QueueId id = m_NextQueueId++;
    m_Streams[id] = static_cast<hipStream_t>(queue);
    m_StreamOwned[id] = false;
    return id;
}

void HipEngine::RemoveComputeQueue(const QueueId id)
{
    if (!ContainsKey(m_Streams, id))
    {
        throw KttException("Invalid compute queue id: " + std::to_string(id));
    }
    if (m_StreamOwned[id])
    {
        throw KttException("Only queues added by the user can be removed");
    }
    m_Streams.erase(id);
    m_StreamOwned.erase(id);
}

QueueId HipEngine::GetDefaultQueue() const
{
    return m_Streams.begin()->first;
}

std::vector<QueueId> HipEngine::GetAllQueues() const
{
    std::vector<QueueId> result;
    for (const auto& [id, stream] : m_Streams)
    {
        result.push_back(id);
    }
    return result;
}

hipStream_t HipEngine::GetStream(const QueueId id) const
{
    auto iterator = m_Streams.find(id);
    if (iterator == m_Streams.end())
    {
        throw KttException("Invalid compute queue id: " + std::to_string(id));
    }
    return iterator->second;
}

void HipEngine::SynchronizeQueue(const QueueId queueId)
{
    CheckHipError(hipStreamSynchronize(GetStream(queueId)), "hipStreamSynchronize");
}

void HipEngine::SynchronizeQueues()
{
    for (const auto& [id, stream] : m_Streams)
    {
        CheckHipError(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }
}

void HipEngine::SynchronizeDevice()
{
    CheckHipError(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

// ---------------------------------------------------------------- information

std::vector<PlatformInfo> HipEngine::GetPlatformInfo() const
{
    PlatformInfo info(0, "AMD HIP");
    info.SetVendor("Advanced Micro Devices, Inc.");
    int version = 0;
    IgnoreHipError(hipRuntimeGetVersion(&version));
    info.SetVersion(std::to_string(version));
    return {info};
}

std::vector<DeviceInfo> HipEngine::GetDeviceInfo(const PlatformIndex platformIndex) const
{
    if (platformIndex != 0)
    {
        throw KttException("Invalid platform index: " + std::to_string(platformIndex));
    }
    int count = 0;
    CheckHipError(hipGetDeviceCount(&count), "hipGetDeviceCount");
    std::vector<DeviceInfo> result;
    for (int device = 0; device < count; ++device)
    {
        hipDeviceProp_t properties;
        CheckHipError(hipGetDeviceProperties(&properties, device), "hipGetDeviceProperties");
        DeviceInfo info(static_cast<DeviceIndex>(device), HipCompiler::DeviceName(device));
        info.SetVendor("Advanced Micro Devices, Inc.");
        info.SetArchitecture(HipCompiler::DeviceArchitecture(device));
        info.SetGlobalMemorySize(properties.totalGlobalMem);
        info.SetLocalMemorySize(properties.sharedMemPerBlock);
        info.SetMaxConstantBufferSize(properties.totalConstMem);
        info.SetMaxWorkGroupSize(static_cast<uint64_t>(properties.maxThreadsPerBlock));
        info.SetMaxComputeUnits(static_cast<uint32_t>(properties.multiProcessorCount));
        info.SetClockRateKHz(static_cast<uint32_t>(properties.clockRate));
        info.SetWavefrontSize(static_cast<uint32_t>(properties.warpSize));
        info.SetDeviceType(DeviceType::GPU);
        result.push_back(info);
    }
    return result;
}

PlatformInfo HipEngine::GetCurrentPlatformInfo() const
{
    return GetPlatformInfo()[0];
}

DeviceInfo HipEngine::GetCurrentDeviceInfo() const
{
    return GetDeviceInfo(0).at(static_cast<size_t>(m_DeviceIndex));
}

ComputeApi HipEngine::GetComputeApi() const
{
    return ComputeApi::HIP;
}

GlobalSizeType HipEngine::GetGlobalSizeType() const
{
    return m_GlobalSizeType;
}

void HipEngine::SetCompilerOptions(const std::string& options, const bool overrideDefault)
{
    m_CompilerOptions = options;
    m_OverrideDefaultOptions = overrideDefault;
    ClearKernelCache();
}

void HipEngine::SetGlobalSizeType(const GlobalSizeType type)
{
    m_GlobalSizeType = type;
}

void HipEngine::SetAutomaticGlobalSizeCorrection(const bool flag)
{
    m_GlobalSizeCorrection = flag;
}

void HipEngine::SetKernelCacheCapacity(const uint64_t capacity)
{
    m_KernelCacheCapacity = capacity;
    ClearKernelCache();
}

void HipEngine::ClearKernelCache()
{
    m_KernelCache.clear();
}

void HipEngine::EnsureThreadContext()
{
    CheckHipError(hipSetDevice(m_DeviceIndex), "hipSetDevice");
}

} // namespace ktt
