#pragma once

#include <cstdint>
#include <type_traits>

namespace ktt
{

template <typename T>
void Tuner::AddParameter(const KernelId id, const std::string& name, const std::vector<T>& values, const std::string& group)
{
    std::vector<ParameterValue> parameterValues;

    for (const auto& value : values)
    {
        if constexpr (std::is_same_v<T, bool>)
        {
            parameterValues.push_back(static_cast<bool>(value));
        }
        else if constexpr (std::is_same_v<T, std::string> || std::is_same_v<T, const char*>)
        {
            parameterValues.push_back(std::string(value));
        }
        else if constexpr (std::is_floating_point_v<T>)
        {
            parameterValues.push_back(static_cast<double>(value));
        }
        else if constexpr (std::is_signed_v<T>)
        {
            parameterValues.push_back(static_cast<int64_t>(value));
        }
        else
        {
            parameterValues.push_back(static_cast<uint64_t>(value));
        }
    }

    AddParameterInternal(id, name, parameterValues, group);
}

template <typename T>
ArgumentId Tuner::AddArgumentVector(const std::vector<T>& data, const ArgumentAccessType accessType, const ArgumentId& customId)
{
    const size_t elementSize = sizeof(T);
    const size_t dataSize = data.size() * elementSize;
    return AddArgumentWithOwnedData(elementSize, DeriveArgumentDataType<T>(), ArgumentMemoryLocation::Device, accessType,
        ArgumentMemoryType::Vector, ArgumentManagementType::Framework, data.data(), dataSize, customId);
}

template <typename T>
ArgumentId Tuner::AddArgumentVector(std::vector<T>& data, const ArgumentAccessType accessType, const ArgumentMemoryLocation memoryLocation,
    const ArgumentManagementType managementType, const bool referenceUserData, const ArgumentId& customId)
{
    const size_t elementSize = sizeof(T);
    const size_t dataSize = data.size() * elementSize;

    if (referenceUserData)
    {
        return AddArgumentWithReferencedData(elementSize, DeriveArgumentDataType<T>(), memoryLocation, accessType,
            ArgumentMemoryType::Vector, managementType, data.data(), dataSize, customId);
    }

    return AddArgumentWithOwnedData(elementSize, DeriveArgumentDataType<T>(), memoryLocation, accessType, ArgumentMemoryType::Vector,
        managementType, data.data(), dataSize, customId);
}

template <typename T>
ArgumentId Tuner::AddArgumentVector(ComputeBuffer buffer, const size_t bufferSize, const ArgumentAccessType accessType,
    const ArgumentMemoryLocation memoryLocation, const ArgumentId& customId)
{
    return AddUserArgument(buffer, sizeof(T), DeriveArgumentDataType<T>(), memoryLocation, accessType, bufferSize, customId);
}

template <typename T>
ArgumentId Tuner::AddArgumentScalar(const T& data, const ArgumentId& customId)
{
    return AddArgumentWithOwnedData(sizeof(T), DeriveArgumentDataType<T>(), ArgumentMemoryLocation::Undefined,
        ArgumentAccessType::ReadOnly, ArgumentMemoryType::Scalar, ArgumentManagementType::Framework, &data, sizeof(T), customId);
}

template <typename T>
ArgumentId Tuner::AddArgumentLocal(const size_t localMemorySize, const ArgumentId& customId)
{
    return AddArgumentWithOwnedData(sizeof(T), DeriveArgumentDataType<T>(), ArgumentMemoryLocation::Undefined,
        ArgumentAccessType::ReadWrite, ArgumentMemoryType::Local, ArgumentManagementType::Framework, nullptr, localMemorySize,
        customId);
}

template <typename T>
ArgumentDataType Tuner::DeriveArgumentDataType()
{
    if constexpr (std::is_same_v<T, int8_t> || std::is_same_v<T, char>)
    {
        return ArgumentDataType::Char;
    }
    else if constexpr (std::is_same_v<T, uint8_t>)
    {
        return ArgumentDataType::UnsignedChar;
    }
    else if constexpr (std::is_same_v<T, int16_t>)
    {
        return ArgumentDataType::Short;
    }
    else if constexpr (std::is_same_v<T, uint16_t>)
    {
        return ArgumentDataType::UnsignedShort;
    }
    else if constexpr (std::is_same_v<T, int32_t>)
    {
        return ArgumentDataType::Int;
    }
    else if constexpr (std::is_same_v<T, uint32_t>)
    {
        return ArgumentDataType::UnsignedInt;
    }
    else if constexpr (std::is_same_v<T, int64_t>)
    {
        return ArgumentDataType::Long;
    }
    else if constexpr (std::is_same_v<T, uint64_t>)
    {
        return ArgumentDataType::UnsignedLong;
    }
    else if constexpr (std::is_same_v<T, float>)
    {
        return ArgumentDataType::Float;
    }
    else if constexpr (std::is_same_v<T, double>)
    {
        return ArgumentDataType::Double;
    }
    else
    {
        return ArgumentDataType::Custom;
    }
}

} // namespace ktt
