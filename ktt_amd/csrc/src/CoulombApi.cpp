// C entry points of include/ktt_hip.h, sections 1-3: library/device queries, memory and timing
// plumbing, and the direct Coulomb launches on device buffers.
#include <dlfcn.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <sstream>

#include <KttException.h>

#include "../../../include/ktt_hip.h"
#include "Common.h"
#include "CoulombRuntime.h"
#include "HipEngine.h"
#include "Kernel.h"

namespace ktt
{

static thread_local std::string g_LastError;

int SetLastError(const std::string& message)
{
    g_LastError = message;
    return -1;
}

// ---------------------------------------------------------------- CoulombConfig

bool CoulombConfig::Has(const std::string& name) const
{
    return std::any_of(values.cbegin(), values.cend(), [&name](const auto& entry) { return entry.first == name; });
}

long CoulombConfig::Get(const std::string& name, const long fallback) const
{
    for (const auto& [key, value] : values)
    {
        if (key == name)
        {
            char* end = nullptr;
            const long result = std::strtol(value.c_str(), &end, 10);
            if (end == value.c_str() || *end != '\0')
            {
                throw KttException("Configuration value " + name + "=" + value + " is not an integer");
            }
            return result;
        }
    }
    return fallback;
}

void CoulombConfig::Set(const std::string& name, const std::string& value)
{
    for (auto& [key, entry] : values)
    {
        if (key == name)
        {
            entry = value;
            return;
        }
    }
    values.emplace_back(name, value);
}

std::vector<std::string> CoulombConfig::Defines(const std::vector<std::string>& launchOnly) const
{
    std::vector<std::string> result;
    for (const auto& [key, value] : values)
    {
        if (std::find(launchOnly.cbegin(), launchOnly.cend(), key) == launchOnly.cend())
        {
            result.push_back("-D" + key + "=" + value);
        }
    }
    return result;
}

std::string CoulombConfig::ToString() const
{
    std::string result;
    for (const auto& [key, value] : values)
    {
        if (!result.empty())
        {
            result += ",";
        }
        result += key + "=" + value;
    }
    return result;
}

CoulombConfig CoulombConfig::Parse(const char* text, const CoulombConfig& defaults)
{
    CoulombConfig result = defaults;
    if (text == nullptr)
    {
        return result;
    }
    std::stringstream stream(text);
    std::string item;
    while (std::getline(stream, item, ','))
    {
        item.erase(std::remove_if(item.begin(), item.end(), [](const char c) { return c == ' ' || c == '\t'; }), item.end());
        if (item.empty())
        {
            continue;
        }
        const size_t equals = item.find('=');
        if (equals == std::string::npos || equals == 0 || equals + 1 == item.size())
        {
            throw KttException("Invalid configuration item '" + item + "', expected NAME=VALUE");
        }
        const std::string name = item.substr(0, equals);
        for (const char c : name)
        {
            if (!(std::isalnum(static_cast<unsigned char>(c)) || c == '_'))
            {
                throw KttException("Invalid configuration name '" + name + "'");
            }
        }
        result.Set(name, item.substr(equals + 1));
    }
    return result;
}

// ---------------------------------------------------------------- CoulombRuntime

static std::string LibraryDirectory()
{
    Dl_info info;
    if (dladdr(reinterpret_cast<void*>(&LibraryDirectory), &info) != 0 && info.dli_fname != nullptr)
    {
        std::string path = info.dli_fname;
        const size_t slash = path.find_last_of('/');
        if (slash != std::string::npos)
        {
            return path.substr(0, slash);
        }
    }
    return ".";
}

CoulombRuntime::CoulombRuntime() :
    m_DeviceChecked(false)
{
    if (const char* directory = std::getenv("KTT_HIP_KERNEL_DIR"))
    {
        m_KernelDirectory = directory;
    }
    else
    {
        m_KernelDirectory = LibraryDirectory() + "/kernels";
    }
    m_Architecture = "gfx950";
}

CoulombRuntime& CoulombRuntime::Get()
{
    static CoulombRuntime runtime;
    return runtime;
}

void CoulombRuntime::EnsureDevice()
{
    std::lock_guard<std::mutex> lock(m_Mutex);
    if (m_DeviceChecked)
    {
        return;
    }
    int count = 0;
    const hipError_t status = hipGetDeviceCount(&count);
    if (status != hipSuccess || count == 0)
    {
        throw KttException("No HIP device is available (hipGetDeviceCount: " + std::string(hipGetErrorString(status)) + ")");
    }
    int device = 0;
    CheckHipError(hipGetDevice(&device), "hipGetDevice");
    m_Architecture = HipCompiler::DeviceArchitecture(device);
    m_DeviceChecked = true;
}

std::shared_ptr<HipKernel> CoulombRuntime::GetKernel(const std::string& file, const std::string& kernel,
    const std::vector<std::string>& defines)
{
    EnsureDevice();
    std::string key = file + "|" + kernel;
    for (const auto& define : defines)
    {
        key += "|" + define;
    }
    {
        std::lock_guard<std::mutex> lock(m_Mutex);
        auto iterator = m_Kernels.find(key);
        if (iterator != m_Kernels.end())
        {
            return iterator->second;
        }
    }

    const std::string path = m_KernelDirectory + "/" + file;
    const std::string source = ReadTextFile(path);
    std::vector<std::string> options = HipCompiler::DefaultOptions(m_Architecture);
    options.push_back("-I" + m_KernelDirectory);
    for (const auto& define : defines)
    {
        options.push_back(define);
    }
    Timer timer;
    timer.Start();
    std::string lowered;
    const auto code = HipCompiler::Compile(source, file, kernel, options, lowered);
    timer.Stop();
    auto compiled = std::make_shared<HipKernel>(kernel, code, lowered, timer.GetElapsedTime());

    std::lock_guard<std::mutex> lock(m_Mutex);
    m_Kernels[key] = compiled;
    return compiled;
}

void* CoulombRuntime::GetScratch(const std::string& name, const size_t bytes)
{
    std::lock_guard<std::mutex> lock(m_Mutex);
    auto& entry = m_Scratch[name];
    if (entry.second < bytes)
    {
        if (entry.first != nullptr)
        {
            CheckHipError(hipDeviceSynchronize(), "hipDeviceSynchronize");
            IgnoreHipError(hipFree(entry.first));
            entry = {nullptr, 0};
        }
        CheckHipError(hipMalloc(&entry.first, bytes), "hipMalloc(scratch " + name + ")");
        CheckHipError(hipMemset(entry.first, 0, bytes), "hipMemset(scratch " + name + ")");
        CheckHipError(hipDeviceSynchronize(), "hipDeviceSynchronize");
        entry.second = bytes;
    }
    return entry.first;
}

CoulombShape CoulombRuntime::RowShape(const CoulombConfig& config, const long spanX, const long rows, const long atomBlocks,
    const bool foldRows)
{
    const long wgx = config.Get("WORK_GROUP_SIZE_X", 64);
    const long wgy = config.Get("WORK_GROUP_SIZE_Y", 4);
    const long px = config.Get("POINTS_X", 4);
    const long pr = config.Get("ROWS_PER_THREAD", 16);
    if (wgx <= 0 || wgy <= 0 || px <= 0 || pr <= 0 || atomBlocks <= 0)
    {
        throw KttException("Workgroup shape, points per thread and atom blocks must be positive");
    }
    if (wgx * wgy > 1024)
    {
        throw KttException("Workgroup of " + std::to_string(wgx * wgy) + " threads exceeds the device limit of 1024",
            ExceptionReason::DeviceLimitsExceeded);
    }
    const long threads = wgx * wgy;
    const long rowsPerGroup = wgy * pr;
    if (threads % rowsPerGroup != 0 && rowsPerGroup % threads != 0)
    {
        throw KttException("Rows per workgroup (" + std::to_string(rowsPerGroup) + ") and threads per workgroup ("
            + std::to_string(threads) + ") must divide one another");
    }
    CoulombShape shape;
    const long spanPerGroup = wgx * px;
    const long groupsX = (spanX + spanPerGroup - 1) / spanPerGroup;
    const long groupsY = (rows + rowsPerGroup - 1) / rowsPerGroup;
    if (groupsX > 0x7fffffffL || groupsY > 0x7fffffffL || atomBlocks > 65535L)
    {
        throw KttException("Launch grid exceeds the device limits", ExceptionReason::DeviceLimitsExceeded);
    }
    long fold = 1;
    if (groupsY > 65535L)
    {
        if (!foldRows)
        {
            throw KttException("Launch grid of " + std::to_string(groupsY) + " row groups exceeds 65535; raise ROWS_PER_THREAD "
                "or WORK_GROUP_SIZE_Y", ExceptionReason::DeviceLimitsExceeded);
        }
        fold = (groupsY + 65534L) / 65535L;
        if (groupsX * fold > 0x7fffffffL)
        {
            throw KttException("Launch grid exceeds the device limits", ExceptionReason::DeviceLimitsExceeded);
        }
    }
    shape.grid[0] = static_cast<unsigned int>(groupsX * fold);
    shape.grid[1] = static_cast<unsigned int>((groupsY + fold - 1) / fold);
    shape.grid[2] = static_cast<unsigned int>(atomBlocks);
    shape.groupsX = groupsX;
    shape.groupsY = groupsY;
    shape.block[0] = static_cast<unsigned int>(wgx);
    shape.block[1] = static_cast<unsigned int>(wgy);
    shape.block[2] = 1;
    shape.threads = threads;
    shape.rowsPerGroup = rowsPerGroup;
    shape.pointsPerThread = px * pr;
    return shape;
}

size_t CoulombRuntime::ResidentWorkgroups(const HipKernel& kernel, const int threads)
{
    int device = 0;
    CheckHipError(hipGetDevice(&device), "hipGetDevice");
    int computeUnits = 0;
    CheckHipError(hipDeviceGetAttribute(&computeUnits, hipDeviceAttributeMultiprocessorCount, device), "hipDeviceGetAttribute");
    int perUnit = 0;
    CheckHipError(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&perUnit, kernel.GetFunction(), threads, 0),
        "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor");
    return static_cast<size_t>(std::max(computeUnits, 0)) * static_cast<size_t>(std::max(perUnit, 0));
}

void CoulombRuntime::Launch(const HipKernel& kernel, const unsigned int grid[3], const unsigned int block[3], void** arguments,
    hipStream_t stream, float* kernelMs)
{
    if (grid[0] == 0 || grid[1] == 0 || grid[2] == 0)
    {
        if (kernelMs != nullptr)
        {
            *kernelMs = 0.0f;
        }
        return;
    }
    hipEvent_t start = nullptr;
    hipEvent_t end = nullptr;
    if (kernelMs != nullptr)
    {
        CheckHipError(hipEventCreate(&start), "hipEventCreate");
        CheckHipError(hipEventCreate(&end), "hipEventCreate");
        CheckHipError(hipEventRecord(start, stream), "hipEventRecord");
    }
    const hipError_t status = hipModuleLaunchKernel(kernel.GetFunction(), grid[0], grid[1], grid[2], block[0], block[1], block[2], 0,
        stream, arguments, nullptr);
    if (status != hipSuccess)
    {
        if (start != nullptr)
        {
            IgnoreHipError(hipEventDestroy(start));
            IgnoreHipError(hipEventDestroy(end));
        }
        CheckHipError(status, "hipModuleLaunchKernel(" + kernel.GetName() + ")");
    }
    if (kernelMs != nullptr)
    {
        CheckHipError(hipEventRecord(end, stream), "hipEventRecord");
        const hipError_t sync = hipEventSynchronize(end);
        float milliseconds = 0.0f;
        if (sync == hipSuccess)
        {
            IgnoreHipError(hipEventElapsedTime(&milliseconds, start, end));
        }
        IgnoreHipError(hipEventDestroy(start));
        IgnoreHipError(hipEventDestroy(end));
        CheckHipError(sync, "hipEventSynchronize(" + kernel.GetName() + ")");
        *kernelMs = milliseconds;
    }
}

// Tuned defaults (gfx950, 256^3 grid x 65,536 atoms): both forms keep 2 x 8 points per thread
// and run the inner loop in issue phases.
static CoulombConfig Defaults3d(const int form, const long gridSize)
{
    // The fastest shapes of each form at 256^3 x 65,536 on MI355X (sweeps of workgroup shape,
    // register block, atom tile and unroll). With ISSUE_PHASES=1 (FMA phase and v_rsq_f32 phase
    // per atom at different wave priorities, so that the FMAs of two waves issue in one
    // quad-cycle) both forms run 2 x 8 points per thread: 92.8 ms compensated (128 x 2 threads,
    // 256-atom tiles) and 92.7 ms plain (128 x 4 threads, 128-atom tiles), against 105.2 and
    // 103.7 ms for the best shapes of the interleaved loop. A row length that is not a multiple
    // of the span narrows the workgroup to 64 threads in x, so that fewer lanes idle past the
    // row's end.
    //
    // TAIL_SPLIT (compensated form, one atom block): the tiles of the last, partial round of
    // workgroups split their atoms so that the round fills the GPU (see ktt_coulomb3d).  The
    // plain form has no split: its blocks would add with FP32 atomics, whose order is not fixed.
    CoulombConfig config;
    const long wgx = gridSize % 256 != 0 ? 64 : 128;
    if (form == KTT_COULOMB_COMPENSATED)
    {
        config.values = {{"WORK_GROUP_SIZE_X", std::to_string(wgx)}, {"WORK_GROUP_SIZE_Y", "2"}, {"POINTS_X", "2"},
            {"ROWS_PER_THREAD", "8"}, {"ATOM_TILE", "256"}, {"INNER_UNROLL_FACTOR", "16"}, {"ISSUE_PHASES", "1"},
            {"ATOM_BLOCKS", "1"}, {"TAIL_SPLIT", "1"}};
    }
    else
    {
        config.values = {{"WORK_GROUP_SIZE_X", std::to_string(wgx)}, {"WORK_GROUP_SIZE_Y", std::to_string(512 / wgx)},
            {"POINTS_X", "2"}, {"ROWS_PER_THREAD", "8"}, {"ATOM_TILE", "128"}, {"INNER_UNROLL_FACTOR", "16"},
            {"ISSUE_PHASES", "1"}, {"ATOM_BLOCKS", "1"}};
    }
    return config;
}

static CoulombConfig DefaultsSlice(const int atomCount, const long gridX, const long gridY)
{
    // One slice (256 x 256 points) has too few points to fill 256 CUs with several waves per
    // SIMD, so the atoms are split over up to 256 blocks (fewer when there are fewer atom tiles).
    // With issue phases (ISSUE_PHASES=1) the 256 slice launches of 256^3 x 65,536 run 97.1 ms
    // (79% of the FP32 peak) with 2 x 8 points per thread and 100.1 ms with 1 x 4; the 2 x 8
    // block needs rows that are a multiple of 128 points and a multiple of 32 rows, since the
    // slice kernels take the row length and count from the launch.
    //
    // A slice of a few thousand atoms (256^2 x 4,000) is one round of workgroups whose fixed
    // costs (the tile load at the start, the atomic adds at the end) weigh more: 64-atom tiles
    // over up to 128 blocks ran 32.7 us against 36.0 us for 128-atom tiles over 32 blocks, the
    // fastest 64-thread-wide shape of a sweep of 102 on MI355X. Issue phases take that slice
    // from 34.7 to 33.0 us.
    CoulombConfig config;
    const bool few = atomCount <= 16384;
    const bool wide = !few && gridX % 128 == 0 && gridY % 32 == 0;
    config.values = {{"WORK_GROUP_SIZE_X", "64"}, {"WORK_GROUP_SIZE_Y", "4"}, {"POINTS_X", wide ? "2" : "1"},
        {"ROWS_PER_THREAD", wide ? "8" : "4"}, {"ATOM_TILE", few ? "64" : "128"}, {"INNER_UNROLL_FACTOR", "16"},
        {"ISSUE_PHASES", "1"}, {"ATOM_BLOCKS", few ? "128" : "256"}};
    return config;
}

static const std::vector<std::string> kLaunchOnly = {"ATOM_BLOCKS", "TAIL_SPLIT"};

// Most blocks a tail tile is split over (the tail round then holds at most this many times its tiles).
constexpr long kMaxTailBlocks = 16;

static void CheckAtoms(const float* atoms, const int atomCount)
{
    if (atomCount < 0)
    {
        throw KttException("Number of atoms must not be negative");
    }
    if (atomCount > 0 && atoms == nullptr)
    {
        throw KttException("Atom buffer is null");
    }
}

static void CheckGrid(const float* grid, const long points)
{
    if (points <= 0)
    {
        throw KttException("Grid must have at least one point");
    }
    if (grid == nullptr)
    {
        throw KttException("Energy grid buffer is null");
    }
}

// A default atom tile larger than the atom count only stages zero-charge padding (64 atoms in a
// 256-atom tile do a quarter of the tile's work): halve it while it still holds every atom and
// stays a multiple of the bracket stage's atom stride (threads / rows per workgroup), down to 16.
static void ShrinkTileToAtoms(CoulombConfig& config, const int atomCount)
{
    const long threads = config.Get("WORK_GROUP_SIZE_X", 64) * config.Get("WORK_GROUP_SIZE_Y", 4);
    const long rows = config.Get("WORK_GROUP_SIZE_Y", 4) * config.Get("ROWS_PER_THREAD", 16);
    const long stride = (threads > rows && rows > 0) ? threads / rows : 1;
    long tile = config.Get("ATOM_TILE", 64);
    while (tile % 2 == 0 && tile / 2 >= 16 && tile / 2 >= atomCount && (tile / 2) % stride == 0)
    {
        tile /= 2;
    }
    config.Set("ATOM_TILE", std::to_string(tile));
}

static long EffectiveAtomBlocks(const CoulombConfig& config, const int atomCount)
{
    const long tile = std::max<long>(1, config.Get("ATOM_TILE", 64));
    const long tiles = std::max<long>(1, (atomCount + tile - 1) / tile);
    return std::max<long>(1, std::min(config.Get("ATOM_BLOCKS", 1), tiles));
}

} // namespace ktt

using namespace ktt;

#define KTT_C_TRY try {
#define KTT_C_CATCH } catch (const std::exception& exception) { return SetLastError(exception.what()); } catch (...) { return SetLastError("unknown error"); }

extern "C" {

int ktt_hip_abi_version(void)
{
    return 100;
}

const char* ktt_hip_last_error(void)
{
    return g_LastError.c_str();
}

int ktt_hip_device_count(int* count)
{
    KTT_C_TRY
    if (count == nullptr)
    {
        return SetLastError("count is null");
    }
    *count = 0;
    if (hipGetDeviceCount(count) != hipSuccess)
    {
        *count = 0;
    }
    return 0;
    KTT_C_CATCH
}

int ktt_hip_device_info(char* name, size_t nameLength, char* architecture, size_t architectureLength, int* computeUnits,
    size_t* globalMemory)
{
    KTT_C_TRY
    CoulombRuntime::Get().EnsureDevice();
    int device = 0;
    CheckHipError(hipGetDevice(&device), "hipGetDevice");
    hipDeviceProp_t properties;
    CheckHipError(hipGetDeviceProperties(&properties, device), "hipGetDeviceProperties");
    if (name != nullptr && nameLength > 0)
    {
        std::snprintf(name, nameLength, "%s", HipCompiler::DeviceName(device).c_str());
    }
    if (architecture != nullptr && architectureLength > 0)
    {
        std::snprintf(architecture, architectureLength, "%s", HipCompiler::DeviceArchitecture(device).c_str());
    }
    if (computeUnits != nullptr)
    {
        *computeUnits = properties.multiProcessorCount;
    }
    if (globalMemory != nullptr)
    {
        *globalMemory = properties.totalGlobalMem;
    }
    return 0;
    KTT_C_CATCH
}

const char* ktt_hip_kernel_directory(void)
{
    return CoulombRuntime::Get().GetKernelDirectory().c_str();
}

int ktt_hip_compile_check(const char* sourceFile, const char* kernelName, const char* config, char* log, size_t logLength)
{
    KTT_C_TRY
    if (sourceFile == nullptr || kernelName == nullptr)
    {
        return SetLastError("source file and kernel name are required");
    }
    std::string path = sourceFile;
    if (path.find('/') == std::string::npos)
    {
        path = CoulombRuntime::Get().GetKernelDirectory() + "/" + path;
    }
    const std::string source = ReadTextFile(path);
    const size_t slash = path.find_last_of('/');
    std::vector<std::string> options = HipCompiler::DefaultOptions("gfx950");
    options.push_back("-I" + (slash == std::string::npos ? std::string(".") : path.substr(0, slash)));
    for (const auto& define : CoulombConfig::Parse(config, CoulombConfig()).Defines(kLaunchOnly))
    {
        options.push_back(define);
    }
    std::string lowered;
    std::string compilerLog;
    try
    {
        HipCompiler::Compile(source, path, kernelName, options, lowered, &compilerLog);
    }
    catch (const KttException& exception)
    {
        if (log != nullptr && logLength > 0)
        {
            std::snprintf(log, logLength, "%s", exception.what());
        }
        throw;
    }
    if (log != nullptr && logLength > 0)
    {
        std::snprintf(log, logLength, "%s", compilerLog.c_str());
    }
    return 0;
    KTT_C_CATCH
}

int ktt_hip_malloc(void** pointer, size_t bytes)
{
    KTT_C_TRY
    if (pointer == nullptr)
    {
        return SetLastError("pointer is null");
    }
    CoulombRuntime::Get().EnsureDevice();
    CheckHipError(hipMalloc(pointer, bytes), "hipMalloc");
    return 0;
    KTT_C_CATCH
}

int ktt_hip_free(void* pointer)
{
    KTT_C_TRY
    CheckHipError(hipFree(pointer), "hipFree");
    return 0;
    KTT_C_CATCH
}

int ktt_hip_memcpy_htod(void* destination, const void* source, size_t bytes)
{
    KTT_C_TRY
    CheckHipError(hipMemcpy(destination, source, bytes, hipMemcpyHostToDevice), "hipMemcpy(HtoD)");
    return 0;
    KTT_C_CATCH
}

int ktt_hip_memcpy_dtoh(void* destination, const void* source, size_t bytes)
{
    KTT_C_TRY
    CheckHipError(hipMemcpy(destination, source, bytes, hipMemcpyDeviceToHost), "hipMemcpy(DtoH)");
    return 0;
    KTT_C_CATCH
}

int ktt_hip_zero(void* pointer, size_t bytes, void* stream)
{
    KTT_C_TRY
    if (bytes == 0)
    {
        return 0;
    }
    if (pointer == nullptr)
    {
        return SetLastError("pointer is null");
    }
    if (bytes % 4 != 0 || reinterpret_cast<uintptr_t>(pointer) % 4 != 0)
    {
        return SetLastError("ktt_hip_zero needs a 4-byte aligned pointer and a multiple of 4 bytes");
    }
    auto& runtime = CoulombRuntime::Get();
    const bool wide = bytes % 16 == 0 && reinterpret_cast<uintptr_t>(pointer) % 16 == 0;
    auto kernel = runtime.GetKernel("coulomb_util.hip", wide ? "kttZeroFill16" : "kttZeroFill4", {});
    unsigned long long count = wide ? bytes / 16 : bytes / 4;
    void* arguments[] = {&pointer, &count};
    const unsigned long long groups = std::min<unsigned long long>((count + 255) / 256, 8192);
    const unsigned int grid[3] = {static_cast<unsigned int>(groups), 1, 1};
    const unsigned int block[3] = {256, 1, 1};
    CoulombRuntime::Launch(*kernel, grid, block, arguments, static_cast<hipStream_t>(stream), nullptr);
    return 0;
    KTT_C_CATCH
}

int ktt_hip_copy(void* destination, const void* source, size_t bytes, void* stream)
{
    KTT_C_TRY
    if (bytes == 0)
    {
        return 0;
    }
    if (destination == nullptr || source == nullptr)
    {
        return SetLastError("pointer is null");
    }
    if (bytes % 16 != 0 || reinterpret_cast<uintptr_t>(destination) % 16 != 0 || reinterpret_cast<uintptr_t>(source) % 16 != 0)
    {
        return SetLastError("ktt_hip_copy needs 16-byte aligned pointers and a multiple of 16 bytes");
    }
    auto kernel = CoulombRuntime::Get().GetKernel("coulomb_util.hip", "kttCopy16", {});
    unsigned long long count = bytes / 16;
    void* arguments[] = {&destination, &source, &count};
    const unsigned long long groups = std::min<unsigned long long>((count + 255) / 256, 8192);
    const unsigned int grid[3] = {static_cast<unsigned int>(groups), 1, 1};
    const unsigned int block[3] = {256, 1, 1};
    CoulombRuntime::Launch(*kernel, grid, block, arguments, static_cast<hipStream_t>(stream), nullptr);
    return 0;
    KTT_C_CATCH
}

int ktt_hip_issue_rates(int iterations, double* fmaRate, double* fmaScalarRate, double* rsqRate, double* mixRate)
{
    KTT_C_TRY
    if (iterations <= 0 || iterations > std::numeric_limits<int>::max() / 2)
    {
        return SetLastError("iterations must be in [1, INT_MAX / 2]");
    }
    auto& runtime = CoulombRuntime::Get();
    int device = 0;
    CheckHipError(hipGetDevice(&device), "hipGetDevice");
    int units = 0;
    CheckHipError(hipDeviceGetAttribute(&units, hipDeviceAttributeMultiprocessorCount, device), "hipDeviceGetAttribute");
    // 8 workgroups of 256 threads per CU: 8 waves per SIMD, enough to hide every chain's latency.
    const unsigned int groups = static_cast<unsigned int>(std::max(units, 1)) * 8u;
    const size_t threads = static_cast<size_t>(groups) * 256;
    float* out = static_cast<float*>(runtime.GetScratch("issue_rates.out", threads * sizeof(float)));
    const char* names[4] = {"kttRateFma", "kttRateFmaScalar", "kttRateRsq", "kttRateMix"};
    const double perRound[4] = {8.0, 8.0, 8.0, 24.0};  // VALU instructions per round and thread
    double* rates[4] = {fmaRate, fmaScalarRate, rsqRate, mixRate};
    const unsigned int grid[3] = {groups, 1, 1};
    const unsigned int block[3] = {256, 1, 1};
    std::shared_ptr<HipKernel> kernels[4];
    for (int k = 0; k < 4; ++k)
    {
        kernels[k] = runtime.GetKernel("coulomb_util.hip", names[k], {});
    }
    // The engine clock falls while the GPU idles (the compilations above, the caller's host work)
    // and climbs back over tens of milliseconds of work: 100 ms of the mix probe run first, untimed.
    {
        int rounds = iterations;
        float scale = 0.9999f;
        void* arguments[] = {&out, &rounds, &scale};
        const auto start = std::chrono::steady_clock::now();
        while (std::chrono::steady_clock::now() - start < std::chrono::milliseconds(100))
        {
            CoulombRuntime::Launch(*kernels[3], grid, block, arguments, nullptr, nullptr);
            CheckHipError(hipDeviceSynchronize(), "hipDeviceSynchronize");
        }
    }
    for (int k = 0; k < 4; ++k)
    {
        if (rates[k] == nullptr)
        {
            continue;
        }
        const auto& kernel = kernels[k];
        // Fastest of 3 launches (after one that warms up) at n and at 2n rounds; the difference
        // is n rounds of every thread without the launch's ramp-up and drain.
        float best[2] = {0.0f, 0.0f};
        for (int length = 0; length < 2; ++length)
        {
            int rounds = iterations * (length + 1);
            float scale = 0.9999f;
            void* arguments[] = {&out, &rounds, &scale};
            for (int repeat = 0; repeat < 4; ++repeat)
            {
                float milliseconds = 0.0f;
                CoulombRuntime::Launch(*kernel, grid, block, arguments, nullptr, &milliseconds);
                if (repeat > 0 && (best[length] == 0.0f || milliseconds < best[length]))
                {
                    best[length] = milliseconds;
                }
            }
        }
        const double extra = static_cast<double>(best[1]) - static_cast<double>(best[0]);
        *rates[k] = extra > 0.0 ? static_cast<double>(threads) * iterations * perRound[k] / (extra * 1.0e-3) : 0.0;
    }
    return 0;
    KTT_C_CATCH
}

int ktt_hip_synchronize(void)
{
    KTT_C_TRY
    CheckHipError(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return 0;
    KTT_C_CATCH
}

int ktt_hip_stream_create(void** stream)
{
    KTT_C_TRY
    CoulombRuntime::Get().EnsureDevice();
    CheckHipError(hipStreamCreate(reinterpret_cast<hipStream_t*>(stream)), "hipStreamCreate");
    return 0;
    KTT_C_CATCH
}

int ktt_hip_stream_destroy(void* stream)
{
    KTT_C_TRY
    CheckHipError(hipStreamDestroy(static_cast<hipStream_t>(stream)), "hipStreamDestroy");
    return 0;
    KTT_C_CATCH
}

int ktt_hip_event_create(void** event)
{
    KTT_C_TRY
    CoulombRuntime::Get().EnsureDevice();
    CheckHipError(hipEventCreate(reinterpret_cast<hipEvent_t*>(event)), "hipEventCreate");
    return 0;
    KTT_C_CATCH
}

int ktt_hip_event_destroy(void* event)
{
    KTT_C_TRY
    CheckHipError(hipEventDestroy(static_cast<hipEvent_t>(event)), "hipEventDestroy");
    return 0;
    KTT_C_CATCH
}

int ktt_hip_event_record(void* event, void* stream)
{
    KTT_C_TRY
    CheckHipError(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)), "hipEventRecord");
    return 0;
    KTT_C_CATCH
}

int ktt_hip_event_elapsed_ms(void* start, void* end, float* milliseconds)
{
    KTT_C_TRY
    CheckHipError(hipEventSynchronize(static_cast<hipEvent_t>(end)), "hipEventSynchronize");
    CheckHipError(hipEventElapsedTime(milliseconds, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(end)),
        "hipEventElapsedTime");
    return 0;
    KTT_C_CATCH
}

int ktt_coulomb_launch_shape(int kind, int gridSize, const char* config, unsigned int grid[3], unsigned int block[3])
{
    KTT_C_TRY
    if (gridSize <= 0 || grid == nullptr || block == nullptr)
    {
        return SetLastError("invalid arguments");
    }
    const CoulombConfig parsed = CoulombConfig::Parse(config, kind == 3 ? Defaults3d(KTT_COULOMB_PLAIN, gridSize) : DefaultsSlice(1 << 30, gridSize, gridSize));
    const long rows = kind == 3 ? static_cast<long>(gridSize) * gridSize : gridSize;
    const CoulombShape shape = CoulombRuntime::RowShape(parsed, gridSize, rows, parsed.Get("ATOM_BLOCKS", 1), kind == 3);
    for (int i = 0; i < 3; ++i)
    {
        grid[i] = shape.grid[i];
        block[i] = shape.block[i];
    }
    return 0;
    KTT_C_CATCH
}

} // extern "C"

namespace ktt
{

// Launch plan of one ktt_coulomb3d call: configuration, kernel, workgroup grid (z = atom blocks),
// the first split tile and the scratch the split needs.
struct Coulomb3dPlan
{
    CoulombConfig config;
    CoulombShape shape;
    std::shared_ptr<HipKernel> kernel;
    unsigned int grid[3];
    int splitFrom = 0;
    size_t partialBytes = 0;
    size_t counterCount = 0;
};

static Coulomb3dPlan PlanCoulomb3d(const int atomCount, const int gridSize, const int form, const char* config)
{
    if (gridSize <= 0 || static_cast<long>(gridSize) * gridSize > 0x7fffffffL)
    {
        throw KttException("grid size must be in [1, 46340]");
    }
    if (form != KTT_COULOMB_PLAIN && form != KTT_COULOMB_COMPENSATED)
    {
        throw KttException("unknown form " + std::to_string(form));
    }
    Coulomb3dPlan plan;
    plan.config = CoulombConfig::Parse(config, Defaults3d(form, gridSize));
    if (!CoulombConfig::Parse(config, CoulombConfig()).Has("ATOM_TILE"))
    {
        ShrinkTileToAtoms(plan.config, atomCount);
    }
    const long atomBlocks = EffectiveAtomBlocks(plan.config, atomCount);
    plan.shape = CoulombRuntime::RowShape(plan.config, gridSize, static_cast<long>(gridSize) * gridSize, atomBlocks, true);
    const std::vector<std::string> defines = plan.config.Defines(kLaunchOnly);
    auto& runtime = CoulombRuntime::Get();
    plan.grid[0] = plan.shape.grid[0];
    plan.grid[1] = plan.shape.grid[1];
    plan.grid[2] = plan.shape.grid[2];
    if (form == KTT_COULOMB_PLAIN)
    {
        plan.kernel = runtime.GetKernel("CoulombSum3d.hip", "directCoulombSum", defines);
        return plan;
    }

    plan.kernel = runtime.GetKernel("CoulombSum3d.hip", "directCoulombSumCompensated", defines);
    const size_t tiles = static_cast<size_t>(plan.shape.groupsX) * static_cast<size_t>(plan.shape.groupsY);
    // ATOM_BLOCKS > 1 splits every tile.  Otherwise, with TAIL_SPLIT, only the tiles of the last
    // round of workgroups are split: with S resident workgroups per launch (CUs x workgroups per
    // CU at this kernel's occupancy), tiles % S tiles are left for that round, and splitting each
    // over about S / (tiles % S) blocks fills it.
    long blocks = atomBlocks;
    if (atomBlocks == 1 && plan.config.Get("TAIL_SPLIT", 0) != 0)
    {
        const size_t slots = CoulombRuntime::ResidentWorkgroups(*plan.kernel, plan.shape.threads);
        const size_t tail = slots > 0 ? tiles % slots : 0;
        const long tile = plan.config.Get("ATOM_TILE", 64);
        const long atomTiles = std::max<long>(1, (atomCount + tile - 1) / tile);
        const long fill = tail > 0 ? static_cast<long>(slots / tail) : 1;
        blocks = std::max<long>(1, std::min<long>({fill, atomTiles, kMaxTailBlocks}));
        plan.splitFrom = blocks > 1 ? static_cast<int>(tiles - tail) : static_cast<int>(tiles);
    }
    plan.grid[2] = static_cast<unsigned int>(blocks);
    if (blocks > 1)
    {
        const size_t splitTiles = tiles - static_cast<size_t>(plan.splitFrom);
        plan.partialBytes = static_cast<size_t>(blocks) * splitTiles * static_cast<size_t>(plan.shape.threads)
            * static_cast<size_t>(plan.shape.pointsPerThread) * 8;
        plan.counterCount = splitTiles;  // one arrival counter per split tile
    }
    return plan;
}

} // namespace ktt

extern "C" {

int ktt_coulomb3d(const float* atoms, int atomCount, float spacing, int gridSize, float* energyGrid, int form, const char* config,
    void* stream, float* kernelMs)
{
    KTT_C_TRY
    if (gridSize <= 0 || static_cast<long>(gridSize) * gridSize > 0x7fffffffL)
    {
        return SetLastError("grid size must be in [1, 46340]");
    }
    CheckAtoms(atoms, atomCount);
    CheckGrid(energyGrid, static_cast<long>(gridSize) * gridSize * gridSize);
    if (form != KTT_COULOMB_PLAIN && form != KTT_COULOMB_COMPENSATED)
    {
        return SetLastError("unknown form " + std::to_string(form));
    }
    Coulomb3dPlan plan = PlanCoulomb3d(atomCount, gridSize, form, config);
    const float* unused = nullptr;
    hipStream_t hipStream = static_cast<hipStream_t>(stream);
    if (form == KTT_COULOMB_PLAIN)
    {
        void* arguments[] = {&atoms, &unused, &unused, &unused, &unused, &atomCount, &spacing, &gridSize, &energyGrid};
        CoulombRuntime::Launch(*plan.kernel, plan.grid, plan.shape.block, arguments, hipStream, kernelMs);
        return 0;
    }
    auto& runtime = CoulombRuntime::Get();
    void* partials = plan.partialBytes > 0 ? runtime.GetScratch("coulomb3d.partials", plan.partialBytes) : nullptr;
    void* counters = runtime.GetScratch("coulomb3d.counters", std::max<size_t>(1, plan.counterCount) * sizeof(unsigned int));
    void* arguments[] = {&atoms, &unused, &unused, &unused, &unused, &atomCount, &spacing, &gridSize, &energyGrid, &partials, &counters,
        &plan.splitFrom};
    CoulombRuntime::Launch(*plan.kernel, plan.grid, plan.shape.block, arguments, hipStream, kernelMs);
    return 0;
    KTT_C_CATCH
}

int ktt_coulomb3d_launch_plan(int atomCount, int gridSize, int form, const char* config, unsigned int grid[3], unsigned int block[3],
    int* splitFrom, size_t* scratchBytes)
{
    KTT_C_TRY
    if (atomCount < 0)
    {
        return SetLastError("Number of atoms must not be negative");
    }
    const Coulomb3dPlan plan = PlanCoulomb3d(atomCount, gridSize, form, config);
    for (int i = 0; i < 3; ++i)
    {
        if (grid != nullptr)
        {
            grid[i] = plan.grid[i];
        }
        if (block != nullptr)
        {
            block[i] = plan.shape.block[i];
        }
    }
    if (splitFrom != nullptr)
    {
        *splitFrom = plan.splitFrom;
    }
    if (scratchBytes != nullptr)
    {
        *scratchBytes = plan.partialBytes + plan.counterCount * sizeof(unsigned int);
    }
    return 0;
    KTT_C_CATCH
}

// The no-allocation variant lives in libktt_hip_noalloc.so (built by hipcc, with its scratch in a
// static device array); it is loaded on the first call, so that processes that never use it never
// register its code object.
using NoAllocEntry = int (*)(const float*, int, float, int, float*, void*, char*, size_t);

static NoAllocEntry LoadNoAllocEntry()
{
    static std::mutex mutex;
    static NoAllocEntry entry = nullptr;
    std::lock_guard<std::mutex> lock(mutex);
    if (entry == nullptr)
    {
        const std::string path = LibraryDirectory() + "/libktt_hip_noalloc.so";
        void* library = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
        if (library == nullptr)
        {
            const char* reason = dlerror();
            throw KttException("cannot load " + path + (reason != nullptr ? std::string(": ") + reason : std::string()));
        }
        entry = reinterpret_cast<NoAllocEntry>(dlsym(library, "ktt_noalloc_coulomb3d_compensated"));
        if (entry == nullptr)
        {
            throw KttException(path + " does not export ktt_noalloc_coulomb3d_compensated");
        }
    }
    return entry;
}

int ktt_coulomb3d_compensated_noalloc(const float* atoms, int atomCount, float spacing, int gridSize, float* energyGrid,
    void* stream)
{
    KTT_C_TRY
    const NoAllocEntry entry = LoadNoAllocEntry();
    char message[512] = {};
    if (entry(atoms, atomCount, spacing, gridSize, energyGrid, stream, message, sizeof(message)) != 0)
    {
        return SetLastError(message);
    }
    return 0;
    KTT_C_CATCH
}

int ktt_coulomb3d_reference(const float* atoms, int atomCount, float spacing, int gridSize, float* energyGrid, void* stream,
    float* kernelMs)
{
    KTT_C_TRY
    if (gridSize <= 0 || static_cast<long>(gridSize) * gridSize > 0x7fffffffL)
    {
        return SetLastError("grid size must be in [1, 46340]");
    }
    CheckAtoms(atoms, atomCount);
    CheckGrid(energyGrid, static_cast<long>(gridSize) * gridSize * gridSize);
    auto kernel = CoulombRuntime::Get().GetKernel("CoulombSum3dReference.hip", "directCoulombSumReference", {});
    const unsigned int grid[3] = {static_cast<unsigned int>((gridSize + 15) / 16), static_cast<unsigned int>((gridSize + 15) / 16),
        static_cast<unsigned int>(gridSize)};
    const unsigned int block[3] = {16, 16, 1};
    void* arguments[] = {&atoms, &atomCount, &spacing, &gridSize, &energyGrid};
    CoulombRuntime::Launch(*kernel, grid, block, arguments, static_cast<hipStream_t>(stream), kernelMs);
    return 0;
    KTT_C_CATCH
}

int ktt_coulomb2d(const float* atoms, int atomCount, float spacing, int gridX, int gridY, float* energyGrid, const char* config,
    void* stream, float* kernelMs)
{
    KTT_C_TRY
    CheckAtoms(atoms, atomCount);
    if (gridX <= 0 || gridY <= 0)
    {
        return SetLastError("grid sizes must be positive");
    }
    CheckGrid(energyGrid, static_cast<long>(gridX) * gridY);
    const CoulombConfig parsed = CoulombConfig::Parse(config, DefaultsSlice(atomCount, gridX, gridY));
    const long atomBlocks = EffectiveAtomBlocks(parsed, atomCount);
    const CoulombShape shape = CoulombRuntime::RowShape(parsed, gridX, gridY, atomBlocks);
    // The slice kernel derives the row length and row count from the launch (as the reference
    // 2D kernel does), so the grid must be covered exactly.
    const long spanX = static_cast<long>(shape.block[0]) * parsed.Get("POINTS_X", 4);
    if (gridX % spanX != 0 || gridY % shape.rowsPerGroup != 0)
    {
        return SetLastError("grid " + std::to_string(gridX) + "x" + std::to_string(gridY) + " is not a multiple of the workgroup span "
            + std::to_string(spanX) + "x" + std::to_string(shape.rowsPerGroup));
    }
    auto kernel = CoulombRuntime::Get().GetKernel("CoulombSum2d.hip", "directCoulombSum", parsed.Defines(kLaunchOnly));
    const float* unused = nullptr;
    void* arguments[] = {&atoms, &unused, &unused, &unused, &unused, &atomCount, &spacing, &energyGrid};
    CoulombRuntime::Launch(*kernel, shape.grid, shape.block, arguments, static_cast<hipStream_t>(stream), kernelMs);
    return 0;
    KTT_C_CATCH
}

int ktt_coulomb2d_reference(const float* atoms, int atomCount, float spacing, int gridX, int gridY, float* energyGrid, void* stream,
    float* kernelMs)
{
    KTT_C_TRY
    CheckAtoms(atoms, atomCount);
    if (gridX <= 0 || gridY <= 0 || gridX % 16 != 0 || gridY % 16 != 0)
    {
        return SetLastError("the 2D reference kernel needs grid sizes that are positive multiples of 16");
    }
    CheckGrid(energyGrid, static_cast<long>(gridX) * gridY);
    auto kernel = CoulombRuntime::Get().GetKernel("CoulombSum2dReference.hip", "directCoulombSumReference", {});
    const unsigned int grid[3] = {static_cast<unsigned int>(gridX / 16), static_cast<unsigned int>(gridY / 16), 1};
    const unsigned int block[3] = {16, 16, 1};
    void* arguments[] = {&atoms, &atomCount, &spacing, &energyGrid};
    CoulombRuntime::Launch(*kernel, grid, block, arguments, static_cast<hipStream_t>(stream), kernelMs);
    return 0;
    KTT_C_CATCH
}

int ktt_coulomb3d_iterative_reference(const float* atoms, int atomCount, float spacing, int gridSize, float* energyGrid,
    void* stream, float* kernelMs)
{
    KTT_C_TRY
    CheckAtoms(atoms, atomCount);
    if (gridSize <= 0 || gridSize % 16 != 0 || static_cast<long>(gridSize) * gridSize > 0x7fffffffL)
    {
        return SetLastError("the iterative reference kernel needs a grid size that is a positive multiple of 16 (at most 46336)");
    }
    CheckGrid(energyGrid, static_cast<long>(gridSize) * gridSize * gridSize);
    auto kernel = CoulombRuntime::Get().GetKernel("CoulombSum3dIterativeReference.hip", "directCoulombSumReference", {});
    const unsigned int grid[3] = {static_cast<unsigned int>(gridSize / 16), static_cast<unsigned int>(gridSize / 16),
        static_cast<unsigned int>(gridSize)};
    const unsigned int block[3] = {16, 16, 1};
    void* arguments[] = {&atoms, &atomCount, &spacing, &energyGrid};
    CoulombRuntime::Launch(*kernel, grid, block, arguments, static_cast<hipStream_t>(stream), kernelMs);
    return 0;
    KTT_C_CATCH
}

static void CheckSliceGrid(const CoulombShape& shape, const CoulombConfig& parsed, const int gridSize)
{
    const long spanX = static_cast<long>(shape.block[0]) * parsed.Get("POINTS_X", 4);
    if (gridSize % spanX != 0 || gridSize % shape.rowsPerGroup != 0)
    {
        throw KttException("grid " + std::to_string(gridSize) + " is not a multiple of the workgroup span " + std::to_string(spanX) + "x"
            + std::to_string(shape.rowsPerGroup));
    }
}

int ktt_coulomb3d_iterative(const float* hostAtoms, int atomCount, float spacing, int gridSize, float* energyGrid, const char* config,
    void* stream, float* kernelMs, float* wallMs)
{
    KTT_C_TRY
    CheckAtoms(hostAtoms, atomCount);
    if (gridSize <= 0 || static_cast<long>(gridSize) * gridSize > 0x7fffffffL)
    {
        return SetLastError("grid size must be in [1, 46340]");
    }
    CheckGrid(energyGrid, static_cast<long>(gridSize) * gridSize * gridSize);
    auto& runtime = CoulombRuntime::Get();
    const CoulombConfig parsed = CoulombConfig::Parse(config, DefaultsSlice(atomCount, gridSize, gridSize));
    const long atomBlocks = EffectiveAtomBlocks(parsed, atomCount);
    const CoulombShape shape = CoulombRuntime::RowShape(parsed, gridSize, gridSize, atomBlocks);
    CheckSliceGrid(shape, parsed, gridSize);
    auto kernel = runtime.GetKernel("CoulombSum3dIterative.hip", "directCoulombSum", parsed.Defines(kLaunchOnly));
    hipStream_t hipStream = static_cast<hipStream_t>(stream);

    // Per slice, as the reference launcher: dz^2 of every atom in FP32 on the host
    // (CoulombSum3dIterative.cpp:160-168), an upload of the whole atom array and a launch
    // (:169-179). Two staging and two device buffers let the host prepare slice z + 1 while the
    // GPU runs slice z, and the upload runs on a second stream beside that kernel; per-slice
    // event pairs give the kernel time the reference reports.
    const size_t bytes = static_cast<size_t>(std::max(atomCount, 1)) * 4 * sizeof(float);
    float* deviceBase = static_cast<float*>(runtime.GetScratch("iterative.atoms", 2 * bytes));
    float* deviceAtoms[2] = {deviceBase, deviceBase + bytes / sizeof(float)};
    void* pinned = nullptr;
    CheckHipError(hipHostMalloc(&pinned, 2 * bytes, hipHostMallocDefault), "hipHostMalloc");
    float* staging[2] = {static_cast<float*>(pinned), static_cast<float*>(pinned) + bytes / sizeof(float)};
    for (float* buffer : staging)
    {
        std::memcpy(buffer, hostAtoms, static_cast<size_t>(atomCount) * 4 * sizeof(float));
    }
    hipEvent_t events[6] = {};  // per slot: kernel start, kernel end, upload done
    hipStream_t copyStream = nullptr;
    const auto release = [&]()
    {
        IgnoreHipError(hipStreamSynchronize(hipStream));
        if (copyStream != nullptr)
        {
            IgnoreHipError(hipStreamSynchronize(copyStream));
            IgnoreHipError(hipStreamDestroy(copyStream));
        }
        for (hipEvent_t& event : events)
        {
            if (event != nullptr)
            {
                IgnoreHipError(hipEventDestroy(event));
            }
        }
        IgnoreHipError(hipHostFree(pinned));
    };

    const auto begin = std::chrono::steady_clock::now();
    float kernelTotal = 0.0f;
    try
    {
        for (hipEvent_t& event : events)
        {
            CheckHipError(hipEventCreateWithFlags(&event, hipEventDefault), "hipEventCreateWithFlags");
        }
        CheckHipError(hipStreamCreateWithFlags(&copyStream, hipStreamNonBlocking), "hipStreamCreateWithFlags");
        bool pending[2] = {false, false};
        // Waits for the kernel of slot s (and so its upload) and adds its time.
        const auto retire = [&](const int s)
        {
            if (!pending[s])
            {
                return;
            }
            CheckHipError(hipEventSynchronize(events[3 * s + 1]), "hipEventSynchronize");
            float sliceMs = 0.0f;
            CheckHipError(hipEventElapsedTime(&sliceMs, events[3 * s], events[3 * s + 1]), "hipEventElapsedTime");
            kernelTotal += sliceMs;
            pending[s] = false;
        };
        const float* unused = nullptr;
        for (int z = 0; z < gridSize; ++z)
        {
            const int s = z & 1;
            retire(s);
            const float coordinate = spacing * static_cast<float>(z);
            float* slot = staging[s];
            for (int i = 0; i < atomCount; ++i)
            {
                const float d = coordinate - hostAtoms[4 * i + 2];
                slot[4 * i + 2] = d * d;
            }
            // The kernel of slot s has ended (retire), so its device buffer is free for the upload.
            CheckHipError(hipMemcpyAsync(deviceAtoms[s], slot, static_cast<size_t>(atomCount) * 4 * sizeof(float), hipMemcpyHostToDevice,
                copyStream), "hipMemcpyAsync");
            CheckHipError(hipEventRecord(events[3 * s + 2], copyStream), "hipEventRecord");
            CheckHipError(hipStreamWaitEvent(hipStream, events[3 * s + 2], 0), "hipStreamWaitEvent");
            const float* atomsArgument = deviceAtoms[s];
            int zIndex = z;
            void* arguments[] = {&atomsArgument, &unused, &unused, &unused, &unused, &atomCount, &spacing, &zIndex, &energyGrid};
            CheckHipError(hipEventRecord(events[3 * s], hipStream), "hipEventRecord");
            CoulombRuntime::Launch(*kernel, shape.grid, shape.block, arguments, hipStream, nullptr);
            CheckHipError(hipEventRecord(events[3 * s + 1], hipStream), "hipEventRecord");
            pending[s] = true;
        }
        retire(gridSize & 1);
        retire((gridSize + 1) & 1);
    }
    catch (...)
    {
        release();
        throw;
    }
    const auto finish = std::chrono::steady_clock::now();
    release();
    if (kernelMs != nullptr)
    {
        *kernelMs = kernelTotal;
    }
    if (wallMs != nullptr)
    {
        *wallMs = std::chrono::duration<float, std::milli>(finish - begin).count();
    }
    return 0;
    KTT_C_CATCH
}

int ktt_coulomb3d_slices(const float* atoms, int atomCount, float spacing, int gridSize, float* energyGrid, const char* config,
    void* stream, float* kernelMs)
{
    KTT_C_TRY
    CheckAtoms(atoms, atomCount);
    if (gridSize <= 0 || static_cast<long>(gridSize) * gridSize > 0x7fffffffL)
    {
        return SetLastError("grid size must be in [1, 46340]");
    }
    CheckGrid(energyGrid, static_cast<long>(gridSize) * gridSize * gridSize);
    auto& runtime = CoulombRuntime::Get();
    const CoulombConfig parsed = CoulombConfig::Parse(config, DefaultsSlice(atomCount, gridSize, gridSize));
    const long atomBlocks = EffectiveAtomBlocks(parsed, atomCount);
    const CoulombShape shape = CoulombRuntime::RowShape(parsed, gridSize, gridSize, atomBlocks);
    CheckSliceGrid(shape, parsed, gridSize);
    auto kernel = runtime.GetKernel("CoulombSum3dIterative.hip", "directCoulombSumSlice", parsed.Defines(kLaunchOnly));
    hipStream_t hipStream = static_cast<hipStream_t>(stream);

    hipEvent_t start = nullptr;
    hipEvent_t end = nullptr;
    if (kernelMs != nullptr)
    {
        CheckHipError(hipEventCreate(&start), "hipEventCreate");
        CheckHipError(hipEventCreate(&end), "hipEventCreate");
        CheckHipError(hipEventRecord(start, hipStream), "hipEventRecord");
    }
    const float* unused = nullptr;
    for (int z = 0; z < gridSize; ++z)
    {
        int zIndex = z;
        void* arguments[] = {&atoms, &unused, &unused, &unused, &unused, &atomCount, &spacing, &zIndex, &energyGrid};
        CoulombRuntime::Launch(*kernel, shape.grid, shape.block, arguments, hipStream, nullptr);
    }
    if (kernelMs != nullptr)
    {
        CheckHipError(hipEventRecord(end, hipStream), "hipEventRecord");
        CheckHipError(hipEventSynchronize(end), "hipEventSynchronize");
        IgnoreHipError(hipEventElapsedTime(kernelMs, start, end));
        IgnoreHipError(hipEventDestroy(start));
        IgnoreHipError(hipEventDestroy(end));
    }
    return 0;
    KTT_C_CATCH
}

} // extern "C"
