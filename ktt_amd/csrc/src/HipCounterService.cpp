#include "HipCounterService.h"

#include <cstdlib>
#include <cstring>
#include <dlfcn.h>

#include <KttException.h>

namespace ktt
{

namespace
{

std::string ThisLibraryDirectory()
{
    Dl_info info;
    if (dladdr(reinterpret_cast<void*>(&ThisLibraryDirectory), &info) != 0 && info.dli_fname != nullptr)
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

template <typename Function>
bool Bind(void* library, const char* name, Function& target, std::string& error)
{
    target = reinterpret_cast<Function>(dlsym(library, name));
    if (target == nullptr)
    {
        error = std::string("libktt_hip_counters.so lacks ") + name;
        return false;
    }
    return true;
}

// KTT_HIP_COUNTERS=1 enables counters when the library loads, ahead of any HIP call the
// process makes through it.
__attribute__((constructor)) void EnableFromEnvironment()
{
    const char* flag = std::getenv("KTT_HIP_COUNTERS");
    if (flag != nullptr && std::strcmp(flag, "1") == 0)
    {
        std::string error;
        HipCounterService::Get().Enable(error);
    }
}

} // namespace

HipCounterService& HipCounterService::Get()
{
    static HipCounterService* service = new HipCounterService();  // outlives static destruction
    return *service;
}

bool HipCounterService::Enable(std::string& error)
{
    if (m_Library == nullptr)
    {
        const std::string path = ThisLibraryDirectory() + "/libktt_hip_counters.so";
        void* library = dlopen(path.c_str(), RTLD_NOW | RTLD_GLOBAL);
        if (library == nullptr)
        {
            const char* reason = dlerror();
            m_LoadError = "cannot load " + path + (reason != nullptr ? std::string(": ") + reason : std::string());
            error = m_LoadError;
            return false;
        }
        std::string bindError;
        if (!Bind(library, "ktt_counters_enable", m_Enable, bindError) || !Bind(library, "ktt_counters_running", m_Running, bindError)
            || !Bind(library, "ktt_counters_plan", m_Plan, bindError) || !Bind(library, "ktt_counters_arm", m_Arm, bindError)
            || !Bind(library, "ktt_counters_collect", m_Collect, bindError)
            || !Bind(library, "ktt_counters_disarm", m_Disarm, bindError) || !Bind(library, "ktt_counters_list", m_List, bindError))
        {
            m_LoadError = bindError;
            error = bindError;
            return false;
        }
        m_Library = library;
    }
    char message[1024] = {};
    if (m_Enable(message, sizeof(message)) != 0)
    {
        m_LoadError = message;
        error = message;
        return false;
    }
    return true;
}

bool HipCounterService::IsRunning() const
{
    return m_Library != nullptr && m_Running() != 0;
}

std::string HipCounterService::UnavailableReason() const
{
    if (m_Library == nullptr)
    {
        return m_LoadError.empty() ? std::string("hardware counters are not enabled (set KTT_HIP_COUNTERS=1 or call "
                                                 "ktt::EnableHipCounters() before the first HIP call of the process)")
                                   : m_LoadError;
    }
    return "the counter service is not running: it was enabled after the HIP runtime initialised";
}

size_t HipCounterService::Plan(const int device, const std::vector<std::string>& counters, std::vector<size_t>& passOf)
{
    if (!IsRunning())
    {
        throw KttException(UnavailableReason());
    }
    std::vector<const char*> names;
    for (const auto& counter : counters)
    {
        names.push_back(counter.c_str());
    }
    passOf.assign(counters.size(), 0);
    size_t passes = 0;
    char message[1024] = {};
    if (m_Plan(device, names.data(), names.size(), passOf.data(), &passes, message, sizeof(message)) != 0)
    {
        throw KttException(std::string("Profiling counters: ") + message);
    }
    m_PlannedCounters = counters;
    return passes;
}

void HipCounterService::Arm(const size_t pass, const unsigned blockX, const unsigned blockY, const unsigned blockZ)
{
    char message[1024] = {};
    if (m_Arm(pass, blockX, blockY, blockZ, message, sizeof(message)) != 0)
    {
        throw KttException(std::string("Profiling counters: ") + message);
    }
}

void HipCounterService::Collect(std::vector<double>& values, const int timeoutMs)
{
    values.resize(m_PlannedCounters.size(), 0.0);
    char message[1024] = {};
    if (m_Collect(values.data(), timeoutMs, message, sizeof(message)) != 0)
    {
        throw KttException(std::string("Profiling counters: ") + message);
    }
}

void HipCounterService::Disarm()
{
    if (m_Library != nullptr)
    {
        m_Disarm();
    }
}

std::vector<std::string> HipCounterService::Supported()
{
    std::vector<std::string> result;
    if (m_Library == nullptr)
    {
        return result;
    }
    std::string text(m_List(nullptr, 0) + 1, '\0');
    m_List(text.data(), text.size());
    size_t begin = 0;
    while (begin < text.size() && text[begin] != '\0')
    {
        const size_t end = text.find('\n', begin);
        if (end == std::string::npos)
        {
            break;
        }
        result.push_back(text.substr(begin, end - begin));
        begin = end + 1;
    }
    return result;
}

const std::vector<std::string>& HipCounterService::DefaultCounters()
{
    // Grouped so that greedy planning needs few passes (at most 8 SQ, 4 TCC, 4 TCP and 2 GRBM
    // hardware counters per pass; every utilisation ratio below reads GRBM_GUI_ACTIVE).
    static const std::vector<std::string> counters{
        // memory hierarchy: HBM traffic (KiB), L2 requests from the vector caches, L1 accesses
        "FETCH_SIZE", "TCP_TCC_READ_REQ_sum", "WRITE_SIZE", "TCP_TCC_WRITE_REQ_sum", "TCC_HIT_sum", "TCC_MISS_sum",
        "TCP_TOTAL_CACHE_ACCESSES_sum",
        // utilisation ratios (percent)
        "VALUBusy", "SALUBusy", "OccupancyPercent", "LdsUtil", "LDSBankConflict", "SIMD_UTILIZATION",
        "VALUUtilization",
        // instruction counts (per wave)
        "SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_SMEM", "SQ_INSTS_LDS", "SQ_INSTS_VMEM_RD", "SQ_INSTS_VMEM_WR",
        "SQ_INSTS_BRANCH", "SQ_INSTS_VALU_FMA_F32", "SQ_INSTS_VALU_ADD_F32", "SQ_INSTS_VALU_MUL_F32", "SQ_INSTS_VALU_TRANS_F32",
        "SQ_INSTS_VALU_INT32", "SQ_INSTS_VALU_CVT", "SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64",
        "SQ_INSTS_MFMA",
        // VALU issue (quad-cycles; a transcendental counts two) and the quad-cycles in which two
        // waves' VALU instructions issued together (gfx950 dual issue)
        "SQ_ACTIVE_INST_VALU", "SQ_ACTIVE_INST_VALU2",
    };
    return counters;
}

bool EnableHipCounters(std::string* error)
{
    std::string message;
    const bool enabled = HipCounterService::Get().Enable(message);
    if (!enabled && error != nullptr)
    {
        *error = message;
    }
    return enabled;
}

} // namespace ktt
