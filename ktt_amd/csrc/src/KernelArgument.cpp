#include <cstring>

#include <KttException.h>

#include "Common.h"
#include "KernelArgument.h"

namespace ktt
{

KernelArgument::KernelArgument(const ArgumentId& id, const size_t elementSize, const ArgumentDataType dataType,
    const ArgumentMemoryLocation memoryLocation, const ArgumentAccessType accessType, const ArgumentMemoryType memoryType,
    const ArgumentManagementType managementType) :
    m_Id(id),
    m_ElementSize(elementSize),
    m_DataType(dataType),
    m_MemoryLocation(memoryLocation),
    m_AccessType(accessType),
    m_MemoryType(memoryType),
    m_ManagementType(managementType),
    m_Ownership(ArgumentOwnership::Copy),
    m_ReferencedData(nullptr),
    m_UserBuffer(nullptr),
    m_DataSize(0)
{
    if (elementSize == 0)
    {
        throw KttException("Kernel argument cannot have element size equal to zero");
    }
}

void KernelArgument::SetOwnedData(const void* data, const size_t dataSize)
{
    // Same preconditions as Source/KernelArgument/KernelArgument.cpp:67-88.
    if (dataSize == 0 && m_MemoryType != ArgumentMemoryType::Local)
    {
        throw KttException("Kernel argument cannot be initialized with number of elements equal to zero");
    }
    if (data == nullptr && m_MemoryType != ArgumentMemoryType::Local)
    {
        throw KttException("Kernel argument cannot be initialized with null data");
    }

    m_Ownership = ArgumentOwnership::Copy;
    m_DataSize = dataSize;
    m_ReferencedData = nullptr;
    m_OwnedData.clear();

    if (data != nullptr)
    {
        m_OwnedData.resize(dataSize);
        std::memcpy(m_OwnedData.data(), data, dataSize);
    }
}

void KernelArgument::SetReferencedData(void* data, const size_t dataSize)
{
    if (dataSize == 0)
    {
        throw KttException("Kernel argument cannot be initialized with number of elements equal to zero");
    }
    if (data == nullptr)
    {
        throw KttException("Kernel argument cannot be initialized with null data");
    }

    m_Ownership = ArgumentOwnership::Reference;
    m_ReferencedData = data;
    m_DataSize = dataSize;
    m_OwnedData.clear();
}

void KernelArgument::SetUserBuffer(ComputeBuffer buffer, const size_t dataSize)
{
    if (dataSize == 0)
    {
        throw KttException("Kernel argument cannot be initialized with number of elements equal to zero");
    }

    m_Ownership = ArgumentOwnership::User;
    m_UserBuffer = buffer;
    m_DataSize = dataSize;
    m_OwnedData.clear();
    m_ReferencedData = nullptr;
}

void KernelArgument::UpdateScalarData(const void* data)
{
    if (m_MemoryType != ArgumentMemoryType::Scalar)
    {
        throw KttException("Kernel argument with id " + m_Id + " is not a scalar");
    }
    std::memcpy(m_OwnedData.data(), data, m_DataSize);
}

void KernelArgument::SetLocalSize(const size_t dataSize)
{
    if (m_MemoryType != ArgumentMemoryType::Local)
    {
        throw KttException("Kernel argument with id " + m_Id + " is not a local memory argument");
    }
    m_DataSize = dataSize;
}

const void* KernelArgument::GetData() const
{
    switch (m_Ownership)
    {
    case ArgumentOwnership::Copy:
        return m_OwnedData.empty() ? nullptr : m_OwnedData.data();
    case ArgumentOwnership::Reference:
        return m_ReferencedData;
    default:
        throw KttException("Host data of kernel argument with id " + m_Id + " is not accessible, it is a user buffer");
    }
}

void* KernelArgument::GetData()
{
    return const_cast<void*>(static_cast<const KernelArgument&>(*this).GetData());
}

std::vector<KernelArgument*> KernelArgument::GetArgumentsWithMemoryType(const std::vector<KernelArgument*>& arguments,
    const ArgumentMemoryType type)
{
    std::vector<KernelArgument*> result;
    for (auto* argument : arguments)
    {
        if (argument->GetMemoryType() == type)
        {
            result.push_back(argument);
        }
    }
    return result;
}

KernelArgumentManager::KernelArgumentManager() :
    m_NextId(0)
{}

ArgumentId KernelArgumentManager::ResolveId(const ArgumentId& customId)
{
    if (!customId.empty())
    {
        if (HasArgument(customId))
        {
            throw KttException("Kernel argument with id " + customId + " already exists");
        }
        return customId;
    }

    ArgumentId id;
    do
    {
        id = std::to_string(m_NextId++);
    }
    while (HasArgument(id));
    return id;
}

KernelArgument& KernelArgumentManager::AddArgument(const ArgumentId& id, const size_t elementSize, const ArgumentDataType dataType,
    const ArgumentMemoryLocation location, const ArgumentAccessType accessType, const ArgumentMemoryType memoryType,
    const ArgumentManagementType managementType, [[maybe_unused]] const size_t dataSize)
{
    if (memoryType == ArgumentMemoryType::Vector && location == ArgumentMemoryLocation::Undefined)
    {
        throw KttException("Vector kernel arguments must have properly defined memory location");
    }
    auto argument = std::make_unique<KernelArgument>(id, elementSize, dataType, location, accessType, memoryType, managementType);
    auto& reference = *argument;
    m_Arguments[id] = std::move(argument);
    return reference;
}

ArgumentId KernelArgumentManager::AddArgumentWithReferencedData(const size_t elementSize, const ArgumentDataType dataType,
    const ArgumentMemoryLocation location, const ArgumentAccessType accessType, const ArgumentMemoryType memoryType,
    const ArgumentManagementType managementType, void* data, const size_t dataSize, const ArgumentId& customId)
{
    const ArgumentId id = ResolveId(customId);
    auto& argument = AddArgument(id, elementSize, dataType, location, accessType, memoryType, managementType, dataSize);
    try
    {
        argument.SetReferencedData(data, dataSize);
    }
    catch (const KttException&)
    {
        m_Arguments.erase(id);
        throw;
    }
    return id;
}

ArgumentId KernelArgumentManager::AddArgumentWithOwnedData(const size_t elementSize, const ArgumentDataType dataType,
    const ArgumentMemoryLocation location, const ArgumentAccessType accessType, const ArgumentMemoryType memoryType,
    const ArgumentManagementType managementType, const void* data, const size_t dataSize, const ArgumentId& customId)
{
    const ArgumentId id = ResolveId(customId);
    auto& argument = AddArgument(id, elementSize, dataType, location, accessType, memoryType, managementType, dataSize);
    try
    {
        argument.SetOwnedData(data, dataSize);
    }
    catch (const KttException&)
    {
        m_Arguments.erase(id);
        throw;
    }
    return id;
}

ArgumentId KernelArgumentManager::AddUserArgument(ComputeBuffer buffer, const size_t elementSize, const ArgumentDataType dataType,
    const ArgumentMemoryLocation location, const ArgumentAccessType accessType, const size_t dataSize, const ArgumentId& customId)
{
    const ArgumentId id = ResolveId(customId);
    auto& argument = AddArgument(id, elementSize, dataType, location, accessType, ArgumentMemoryType::Vector,
        ArgumentManagementType::User, dataSize);
    try
    {
        argument.SetUserBuffer(buffer, dataSize);
    }
    catch (const KttException&)
    {
        m_Arguments.erase(id);
        throw;
    }
    return id;
}

void KernelArgumentManager::RemoveArgument(const ArgumentId& id)
{
    if (m_Arguments.erase(id) == 0)
    {
        throw KttException("Attempting to remove argument with invalid id: " + id);
    }
}

KernelArgument& KernelArgumentManager::GetArgument(const ArgumentId& id)
{
    auto iterator = m_Arguments.find(id);
    if (iterator == m_Arguments.end())
    {
        throw KttException("Attempting to retrieve argument with invalid id: " + id);
    }
    return *iterator->second;
}

const KernelArgument& KernelArgumentManager::GetArgument(const ArgumentId& id) const
{
    auto iterator = m_Arguments.find(id);
    if (iterator == m_Arguments.end())
    {
        throw KttException("Attempting to retrieve argument with invalid id: " + id);
    }
    return *iterator->second;
}

std::vector<KernelArgument*> KernelArgumentManager::GetArguments(const std::vector<ArgumentId>& ids)
{
    std::vector<KernelArgument*> result;
    for (const auto& id : ids)
    {
        result.push_back(&GetArgument(id));
    }
    return result;
}

bool KernelArgumentManager::HasArgument(const ArgumentId& id) const
{
    return ContainsKey(m_Arguments, id);
}

} // namespace ktt
