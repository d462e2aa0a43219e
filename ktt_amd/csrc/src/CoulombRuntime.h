// Process-wide runtime behind the C entry points: device check, kernel-source directory,
// hipRTC module cache keyed by (file, kernel, defines), reusable device scratch, and the
// configuration parser for "NAME=VALUE,..." strings.
#pragma once

#include <hip/hip_runtime_api.h>

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "HipEngine.h"

namespace ktt
{

struct CoulombConfig
{
    std::vector<std::pair<std::string, std::string>> values;

    bool Has(const std::string& name) const;
    long Get(const std::string& name, const long fallback) const;
    void Set(const std::string& name, const std::string& value);
    /** Defines for the compiler: every value except the launch-only ones listed. */
    std::vector<std::string> Defines(const std::vector<std::string>& launchOnly) const;
    std::string ToString() const;

    static CoulombConfig Parse(const char* text, const CoulombConfig& defaults);
};

/** Launch shape of the row/tile kernels (coulomb_core.h). */
struct CoulombShape
{
    unsigned int grid[3];
    unsigned int block[3];
    long threads;
    long rowsPerGroup;
    long pointsPerThread;
    long groupsX;  // workgroups along x (tiles = groupsX * groupsY)
    long groupsY;  // row groups; with a fold, grid[0] = fold * groupsX and grid[1] = ceil(groupsY / fold)
};

class CoulombRuntime
{
public:
    static CoulombRuntime& Get();

    /** Throws if no HIP device is available. */
    void EnsureDevice();
    const std::string& GetKernelDirectory() const { return m_KernelDirectory; }
    std::shared_ptr<HipKernel> GetKernel(const std::string& file, const std::string& kernel, const std::vector<std::string>& defines);
    /** Device scratch of at least bytes (grown on demand, zero-filled when allocated). */
    void* GetScratch(const std::string& name, const size_t bytes);

    /** Launch of `rows` rows of spanX points. A launch holds at most 65,535 row groups in y; with
     *  foldRows (the whole-volume kernels, kc::volumeGroup) more row groups fold into x, else
     *  they throw DeviceLimitsExceeded. */
    static CoulombShape RowShape(const CoulombConfig& config, const long spanX, const long rows, const long atomBlocks,
        const bool foldRows = false);
    /** Workgroups of `threads` threads that the current device holds at once for this kernel:
     *  compute units x workgroups per compute unit at the kernel's occupancy. */
    static size_t ResidentWorkgroups(const HipKernel& kernel, const int threads);
    static void Launch(const HipKernel& kernel, const unsigned int grid[3], const unsigned int block[3], void** arguments,
        hipStream_t stream, float* kernelMs);

private:
    CoulombRuntime();

    std::mutex m_Mutex;
    std::string m_KernelDirectory;
    std::string m_Architecture;
    bool m_DeviceChecked;
    std::map<std::string, std::shared_ptr<HipKernel>> m_Kernels;
    std::map<std::string, std::pair<void*, size_t>> m_Scratch;
};

/** Records the message as this thread's last error; returns -1. */
int SetLastError(const std::string& message);

} // namespace ktt
