// Kernel argument model (Source/KernelArgument/KernelArgument.h, KernelArgumentManager.cpp:15-63):
// memory type (scalar, vector, local), location, access, management and ownership. Vector
// arguments copy user data by default; a referenced argument points at user memory; a user
// buffer argument wraps a device pointer owned by the caller.
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
    explicit KernelArgument(const ArgumentId& id, const size_t elementSize, const ArgumentDataType dataType,
        const ArgumentMemoryLocation memoryLocation, const ArgumentAccessType accessType, const ArgumentMemoryType memoryType,
        const ArgumentManagementType managementType);

    void SetOwnedData(const void* data, const size_t dataSize);
    void SetReferencedData(void* data, const size_t dataSize);
    void SetUserBuffer(ComputeBuffer buffer, const size_t dataSize);
    void UpdateScalarData(const void* data);
    void SetLocalSize(const size_t dataSize);

    const ArgumentId& GetId() const { return m_Id; }
    size_t GetElementSize() const { return m_ElementSize; }
    ArgumentDataType GetDataType() const { return m_DataType; }
    ArgumentMemoryLocation GetMemoryLocation() const { return m_MemoryLocation; }
    ArgumentAccessType GetAccessType() const { return m_AccessType; }
    ArgumentMemoryType GetMemoryType() const { return m_MemoryType; }
    ArgumentManagementType GetManagementType() const { return m_ManagementType; }
    ArgumentOwnership GetOwnership() const { return m_Ownership; }
    size_t GetDataSize() const { return m_DataSize; }
    uint64_t GetNumberOfElements() const { return m_ElementSize == 0 ? 0 : m_DataSize / m_ElementSize; }
    bool HasUserBuffer() const { return m_Ownership == ArgumentOwnership::User; }
    ComputeBuffer GetUserBuffer() const { return m_UserBuffer; }
    const void* GetData() const;
    void* GetData();

    template <typename T>
    const T* GetDataWithType() const
    {
        return static_cast<const T*>(GetData());
    }

    static std::vector<KernelArgument*> GetArgumentsWithMemoryType(const std::vector<KernelArgument*>& arguments,
        const ArgumentMemoryType type);

private:
    ArgumentId m_Id;
    size_t m_ElementSize;
    ArgumentDataType m_DataType;
    ArgumentMemoryLocation m_MemoryLocation;
    ArgumentAccessType m_AccessType;
    ArgumentMemoryType m_MemoryType;
    ArgumentManagementType m_ManagementType;
    ArgumentOwnership m_Ownership;
    std::vector<uint8_t> m_OwnedData;
    void* m_ReferencedData;
    ComputeBuffer m_UserBuffer;
    size_t m_DataSize;
};

class KernelArgumentManager
{
public:
    KernelArgumentManager();

    ArgumentId AddArgumentWithReferencedData(const size_t elementSize, const ArgumentDataType dataType,
        const ArgumentMemoryLocation location, const ArgumentAccessType accessType, const ArgumentMemoryType memoryType,
        const ArgumentManagementType managementType, void* data, const size_t dataSize, const ArgumentId& customId);
    ArgumentId AddArgumentWithOwnedData(const size_t elementSize, const ArgumentDataType dataType,
        const ArgumentMemoryLocation location, const ArgumentAccessType accessType, const ArgumentMemoryType memoryType,
        const ArgumentManagementType managementType, const void* data, const size_t dataSize, const ArgumentId& customId);
    ArgumentId AddUserArgument(ComputeBuffer buffer, const size_t elementSize, const ArgumentDataType dataType,
        const ArgumentMemoryLocation location, const ArgumentAccessType accessType, const size_t dataSize,
        const ArgumentId& customId);
    void RemoveArgument(const ArgumentId& id);

    KernelArgument& GetArgument(const ArgumentId& id);
    const KernelArgument& GetArgument(const ArgumentId& id) const;
    std::vector<KernelArgument*> GetArguments(const std::vector<ArgumentId>& ids);
    bool HasArgument(const ArgumentId& id) const;

private:
    std::map<ArgumentId, std::unique_ptr<KernelArgument>> m_Arguments;
    uint64_t m_NextId;

    ArgumentId ResolveId(const ArgumentId& customId);
    KernelArgument& AddArgument(const ArgumentId& id, const size_t elementSize, const ArgumentDataType dataType,
        const ArgumentMemoryLocation location, const ArgumentAccessType accessType, const ArgumentMemoryType memoryType,
        const ArgumentManagementType managementType, const size_t dataSize);
};

} // namespace ktt
