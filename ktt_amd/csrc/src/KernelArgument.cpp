#include <cstring>

#include <KttException.h>

#include "KernelArgument.h"

namespace ktt
{

KernelArgument::KernelArgument(const ArgumentId& id, const size_t elementSize, const ArgumentDataType dataType,
    const ArgumentMemoryLocation memoryLocation, const ArgumentAccessType accessType, const ArgumentMemoryType memoryType,
    const ArgumentManagementType managementType) :
    m_Id(id),
    m_ElementSize(elementSize),
    m_DataType(dataType),
    m_Location(memoryLocation),
    m_Access(accessType),
    m_Kind(memoryType),
    m_Management(managementType),
    m_Ownership(ArgumentOwnership::Copy),
    m_Host(nullptr),
    m_Device(nullptr),
    m_Bytes(0)
{
    if (elementSize == 0)
    {
        throw KttException("Argument " + id + ": the element size must be positive");
    }
    if (memoryType == ArgumentMemoryType::Vector && memoryLocation == ArgumentMemoryLocation::Undefined)
    {
        throw KttException("Argument " + id + ": a vector needs a memory location");
    }
}

void KernelArgument::SetOwnedData(const void* data, const size_t dataSize)
{
    const bool local = m_Kind == ArgumentMemoryType::Local;
    if (!local && (dataSize == 0 || data == nullptr))
    {
        throw KttException("Argument " + m_Id + ": data must be non-null and non-empty");
    }
    m_Ownership = ArgumentOwnership::Copy;
    m_Host = nullptr;
    m_Device = nullptr;
    m_Bytes = dataSize;
    m_Copy.assign(static_cast<const uint8_t*>(data), data != nullptr ? static_cast<const uint8_t*>(data) + dataSize : nullptr);
}

void KernelArgument::SetReferencedData(void* data, const size_t dataSize)
{
    if (dataSize == 0 || data == nullptr)
    {
        throw KttException("Argument " + m_Id + ": referenced data must be non-null and non-empty");
    }
    m_Ownership = ArgumentOwnership::Reference;
    m_Copy.clear();
    m_Device = nullptr;
    m_Host = data;
    m_Bytes = dataSize;
}

void KernelArgument::SetUserBuffer(ComputeBuffer buffer, const size_t dataSize)
{
    if (dataSize == 0)
    {
        throw KttException("Argument " + m_Id + ": a device buffer must have a positive size");
    }
    m_Ownership = ArgumentOwnership::User;
    m_Copy.clear();
    m_Host = nullptr;
    m_Device = buffer;
    m_Bytes = dataSize;
}

void KernelArgument::SetLocalSize(const size_t dataSize)
{
    if (m_Kind != ArgumentMemoryType::Local)
    {
        throw KttException("Argument " + m_Id + " does not describe local memory");
    }
    m_Bytes = dataSize;
}

const void* KernelArgument::GetData() const
{
    if (m_Ownership == ArgumentOwnership::User)
    {
        throw KttException("Argument " + m_Id + " is a device buffer and has no host data");
    }
    if (m_Ownership == ArgumentOwnership::Reference)
    {
        return m_Host;
    }
    return m_Copy.empty() ? nullptr : m_Copy.data();
}

void* KernelArgument::GetData()
{
    return const_cast<void*>(static_cast<const KernelArgument*>(this)->GetData());
}

std::vector<KernelArgument*> VectorArguments(const std::vector<KernelArgument*>& arguments)
{
    std::vector<KernelArgument*> vectors;
    for (KernelArgument* argument : arguments)
    {
        if (argument->GetMemoryType() == ArgumentMemoryType::Vector)
        {
            vectors.push_back(argument);
        }
    }
    return vectors;
}

ArgumentRegistry::ArgumentRegistry() : m_Counter(0) {}

ArgumentId ArgumentRegistry::Add(const ArgumentSpec& spec, const ArgumentSource source, void* data, const size_t bytes,
    const ArgumentId& name)
{
    ArgumentId id = name;
    if (id.empty())
    {
        do
        {
            id = std::to_string(m_Counter++);
        }
        while (Contains(id));
    }
    else if (Contains(id))
    {
        throw KttException("An argument named " + id + " exists already");
    }
    if (spec.kind == ArgumentMemoryType::Symbol)
    {
        // Source/KernelArgument/KernelArgumentManager.cpp:122-133: symbol names are unique. On HIP the
        // name is how the engine finds the variable in the compiled module, so it is required.
        if (spec.symbolName.empty())
        {
            throw KttException("Symbol argument " + id + " needs the name of its __device__ or __constant__ variable");
        }
        for (const auto& [existing, argument] : m_Arguments)
        {
            if (argument->GetMemoryType() == ArgumentMemoryType::Symbol && argument->GetSymbolName() == spec.symbolName)
            {
                throw KttException("Kernel argument with symbol name " + spec.symbolName + " already exists");
            }
        }
    }

    auto argument = std::make_unique<KernelArgument>(id, spec.elementSize, spec.dataType, spec.location, spec.access, spec.kind,
        spec.management);
    switch (source)
    {
    case ArgumentSource::Copy:
        argument->SetOwnedData(data, bytes);
        break;
    case ArgumentSource::Reference:
        argument->SetReferencedData(data, bytes);
        break;
    case ArgumentSource::DeviceBuffer:
        argument->SetUserBuffer(data, bytes);
        break;
    }
    argument->SetSymbolName(spec.symbolName);
    m_Arguments.emplace(id, std::move(argument));
    return id;
}

void ArgumentRegistry::Remove(const ArgumentId& id)
{
    if (m_Arguments.erase(id) == 0)
    {
        throw KttException("No argument named " + id + " to remove");
    }
}

KernelArgument& ArgumentRegistry::Get(const ArgumentId& id)
{
    return const_cast<KernelArgument&>(static_cast<const ArgumentRegistry*>(this)->Get(id));
}

const KernelArgument& ArgumentRegistry::Get(const ArgumentId& id) const
{
    const auto found = m_Arguments.find(id);
    if (found == m_Arguments.end())
    {
        throw KttException("No argument named " + id);
    }
    return *found->second;
}

std::vector<KernelArgument*> ArgumentRegistry::Resolve(const std::vector<ArgumentId>& ids)
{
    std::vector<KernelArgument*> arguments;
    arguments.reserve(ids.size());
    for (const ArgumentId& id : ids)
    {
        arguments.push_back(&Get(id));
    }
    return arguments;
}

} // namespace ktt
