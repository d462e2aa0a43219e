// Kernel arguments and their registry.
//
// An argument is a scalar, a vector or a local-memory size. Vector data is held in one of
// three ways: a private copy of the caller's data (the default), a reference to memory the
// caller keeps alive, or a device buffer the caller owns (never copied, never freed). The
// registry hands out ids ("0", "1", ... unless the caller names the argument).
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include <KttTypes.h>

namespace ktt
{

class KernelArgument
{
public:
    KernelArgument(const ArgumentId& id, const size_t elementSize, const ArgumentDataType dataType,
        const ArgumentMemoryLocation memoryLocation, const ArgumentAccessType accessType, const ArgumentMemoryType memoryType,
        const ArgumentManagementType managementType);

    /** Copies `dataSize` bytes (may be null only for local-memory arguments). */
    void SetOwnedData(const void* data, const size_t dataSize);
    /** Uses the caller's memory, which must outlive the argument. */
    void SetReferencedData(void* data, const size_t dataSize);
    /** Wraps a device buffer owned by the caller. */
    void SetUserBuffer(ComputeBuffer buffer, const size_t dataSize);
    void SetLocalSize(const size_t dataSize);

    const ArgumentId& GetId() const { return m_Id; }
    size_t GetElementSize() const { return m_ElementSize; }
    ArgumentDataType GetDataType() const { return m_DataType; }
    ArgumentMemoryLocation GetMemoryLocation() const { return m_Location; }
    ArgumentAccessType GetAccessType() const { return m_Access; }
    ArgumentMemoryType GetMemoryType() const { return m_Kind; }
    ArgumentManagementType GetManagementType() const { return m_Management; }
    ArgumentOwnership GetOwnership() const { return m_Ownership; }
    size_t GetDataSize() const { return m_Bytes; }
    uint64_t GetNumberOfElements() const { return m_Bytes / m_ElementSize; }
    bool HasUserBuffer() const { return m_Ownership == ArgumentOwnership::User; }
    ComputeBuffer GetUserBuffer() const { return m_Device; }
    /** Host data (copy or reference); throws for a user device buffer. */
    const void* GetData() const;
    void* GetData();
    /** Name of the kernel's __device__ / __constant__ variable a symbol argument fills. */
    const std::string& GetSymbolName() const { return m_SymbolName; }
    void SetSymbolName(const std::string& name) { m_SymbolName = name; }

private:
    ArgumentId m_Id;
    size_t m_ElementSize;
    ArgumentDataType m_DataType;
    ArgumentMemoryLocation m_Location;
    ArgumentAccessType m_Access;
    ArgumentMemoryType m_Kind;
    ArgumentManagementType m_Management;
    ArgumentOwnership m_Ownership;
    std::vector<uint8_t> m_Copy;
    void* m_Host;
    ComputeBuffer m_Device;
    size_t m_Bytes;
    std::string m_SymbolName;
};

/** The vector arguments among `arguments`, in order. */
std::vector<KernelArgument*> VectorArguments(const std::vector<KernelArgument*>& arguments);

/** Where the data of a new argument comes from. */
enum class ArgumentSource
{
    Copy,
    Reference,
    DeviceBuffer
};

struct ArgumentSpec
{
    size_t elementSize = 0;
    ArgumentDataType dataType = ArgumentDataType::Custom;
    ArgumentMemoryLocation location = ArgumentMemoryLocation::Undefined;
    ArgumentAccessType access = ArgumentAccessType::ReadOnly;
    ArgumentMemoryType kind = ArgumentMemoryType::Vector;
    ArgumentManagementType management = ArgumentManagementType::Framework;
    std::string symbolName;  // symbol arguments only
};

class ArgumentRegistry
{
public:
    ArgumentRegistry();

    /** Registers an argument; `data` is host memory (Copy, Reference) or a device pointer (DeviceBuffer). */
    ArgumentId Add(const ArgumentSpec& spec, const ArgumentSource source, void* data, const size_t bytes, const ArgumentId& name);
    void Remove(const ArgumentId& id);

    KernelArgument& Get(const ArgumentId& id);
    const KernelArgument& Get(const ArgumentId& id) const;
    std::vector<KernelArgument*> Resolve(const std::vector<ArgumentId>& ids);
    bool Contains(const ArgumentId& id) const { return m_Arguments.count(id) != 0; }

private:
    std::map<ArgumentId, std::unique_ptr<KernelArgument>> m_Arguments;
    uint64_t m_Counter;
};

} // namespace ktt
