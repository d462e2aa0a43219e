/** @file Info.h
  * Platform and device descriptions (Source/Api/Info/PlatformInfo.h, DeviceInfo.h:27-111).
  * The HIP backend fills them from hipDeviceProp_t; GetName of the platform is "AMD HIP".
  */
#pragma once

#include <cstdint>
#include <string>

#include <KttTypes.h>

namespace ktt
{

class PlatformInfo
{
public:
    PlatformInfo(const PlatformIndex index = 0, const std::string& name = "") :
        m_Index(index),
        m_Name(name)
    {}

    PlatformIndex GetIndex() const { return m_Index; }
    const std::string& GetName() const { return m_Name; }
    const std::string& GetVendor() const { return m_Vendor; }
    const std::string& GetVersion() const { return m_Version; }
    const std::string& GetExtensions() const { return m_Extensions; }
    std::string GetString() const
    {
        return "Platform " + std::to_string(m_Index) + ": " + m_Name + ", vendor: " + m_Vendor + ", version: " + m_Version;
    }

    void SetVendor(const std::string& vendor) { m_Vendor = vendor; }
    void SetVersion(const std::string& version) { m_Version = version; }
    void SetExtensions(const std::string& extensions) { m_Extensions = extensions; }

private:
    PlatformIndex m_Index;
    std::string m_Name;
    std::string m_Vendor;
    std::string m_Version;
    std::string m_Extensions;
};

class DeviceInfo
{
public:
    DeviceInfo(const DeviceIndex index = 0, const std::string& name = "") :
        m_Index(index),
        m_Name(name),
        m_GlobalMemorySize(0),
        m_LocalMemorySize(0),
        m_MaxConstantBufferSize(0),
        m_MaxWorkGroupSize(0),
        m_MaxComputeUnits(0),
        m_DeviceType(DeviceType::GPU),
        m_CudaComputeCapabilityMajor(0),
        m_CudaComputeCapabilityMinor(0),
        m_ClockRateKHz(0),
        m_WavefrontSize(64)
    {}

    DeviceIndex GetIndex() const { return m_Index; }
    const std::string& GetName() const { return m_Name; }
    const std::string& GetVendor() const { return m_Vendor; }
    const std::string& GetExtensions() const { return m_Extensions; }
    uint64_t GetGlobalMemorySize() const { return m_GlobalMemorySize; }
    uint64_t GetLocalMemorySize() const { return m_LocalMemorySize; }
    uint64_t GetMaxConstantBufferSize() const { return m_MaxConstantBufferSize; }
    uint64_t GetMaxWorkGroupSize() const { return m_MaxWorkGroupSize; }
    uint32_t GetMaxComputeUnits() const { return m_MaxComputeUnits; }
    DeviceType GetDeviceType() const { return m_DeviceType; }
    std::string GetDeviceTypeString() const { return m_DeviceType == DeviceType::GPU ? "GPU" : (m_DeviceType == DeviceType::CPU ? "CPU" : "Custom"); }
    /** gfx architecture name, e.g. "gfx950" (HIP-specific addition). */
    const std::string& GetArchitecture() const { return m_Architecture; }
    uint32_t GetClockRateKHz() const { return m_ClockRateKHz; }
    uint32_t GetWavefrontSize() const { return m_WavefrontSize; }
    unsigned int GetCUDAComputeCapabilityMajor() const { return m_CudaComputeCapabilityMajor; }
    unsigned int GetCUDAComputeCapabilityMinor() const { return m_CudaComputeCapabilityMinor; }
    std::string GetString() const
    {
        return "Device " + std::to_string(m_Index) + ": " + m_Name + " (" + m_Architecture + "), compute units: "
            + std::to_string(m_MaxComputeUnits) + ", global memory: " + std::to_string(m_GlobalMemorySize / (1024 * 1024))
            + " MiB, LDS: " + std::to_string(m_LocalMemorySize / 1024) + " KiB";
    }

    void SetVendor(const std::string& vendor) { m_Vendor = vendor; }
    void SetExtensions(const std::string& extensions) { m_Extensions = extensions; }
    void SetGlobalMemorySize(const uint64_t size) { m_GlobalMemorySize = size; }
    void SetLocalMemorySize(const uint64_t size) { m_LocalMemorySize = size; }
    void SetMaxConstantBufferSize(const uint64_t size) { m_MaxConstantBufferSize = size; }
    void SetMaxWorkGroupSize(const uint64_t size) { m_MaxWorkGroupSize = size; }
    void SetMaxComputeUnits(const uint32_t units) { m_MaxComputeUnits = units; }
    void SetDeviceType(const DeviceType type) { m_DeviceType = type; }
    void SetArchitecture(const std::string& architecture) { m_Architecture = architecture; }
    void SetClockRateKHz(const uint32_t clock) { m_ClockRateKHz = clock; }
    void SetWavefrontSize(const uint32_t size) { m_WavefrontSize = size; }
    void SetCUDAComputeCapabilityMajor(const unsigned int major) { m_CudaComputeCapabilityMajor = major; }
    void SetCUDAComputeCapabilityMinor(const unsigned int minor) { m_CudaComputeCapabilityMinor = minor; }

private:
    DeviceIndex m_Index;
    std::string m_Name;
    std::string m_Vendor;
    std::string m_Extensions;
    std::string m_Architecture;
    uint64_t m_GlobalMemorySize;
    uint64_t m_LocalMemorySize;
    uint64_t m_MaxConstantBufferSize;
    uint64_t m_MaxWorkGroupSize;
    uint32_t m_MaxComputeUnits;
    DeviceType m_DeviceType;
    unsigned int m_CudaComputeCapabilityMajor;
    unsigned int m_CudaComputeCapabilityMinor;
    uint32_t m_ClockRateKHz;
    uint32_t m_WavefrontSize;
};

} // namespace ktt
