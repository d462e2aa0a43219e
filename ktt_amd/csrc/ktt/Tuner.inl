#pragma once

#include <cstdint>
#include <type_traits>

namespace ktt
{

template <typename T>
void Tuner::AddParameter(const KernelId id, const std::string& name, const std::vector<T>& values, const std::string& group)
{
    // Values are stored as one of the ParameterValue alternatives: bool, string, double,
    // signed or unsigned 64-bit integer.
    std::vector<ParameterValue> stored;
    stored.reserve(values.size());
    for (const T& value : values)
    {
        if constexpr (std::is_same_v<T, bool>)
        {
            stored.emplace_back(static_cast<bool>(value));
        }
        else if constexpr (std::is_same_v<T, std::string> || std::is_same_v<T, const char*>)
        {
            stored.emplace_back(std::string(value));
        }
        else if constexpr (std::is_floating_point_v<T>)
        {
            stored.emplace_back(static_cast<double>(value));
        }
        else if constexpr (std::is_signed_v<T>)
        {
            stored.emplace_back(static_cast<int64_t>(value));
        }
        else
        {
            stored.emplace_back(static_cast<uint64_t>(value));
        }
    }
    AddParameterValues(id, name, stored, group);
}

template <typename T>
ArgumentId Tuner::AddArgumentVector(const std::vector<T>& data, const ArgumentAccessType accessType, const ArgumentId& customId)
{
    return RegisterArgument(sizeof(T), ElementTypeOf<T>(), ArgumentMemoryLocation::Device, accessType, ArgumentMemoryType::Vector,
        ArgumentManagementType::Framework, DataSource::Copy, const_cast<T*>(data.data()), data.size() * sizeof(T), customId);
}

template <typename T>
ArgumentId Tuner::AddArgumentVector(std::vector<T>& data, const ArgumentAccessType accessType, const ArgumentMemoryLocation memoryLocation,
    const ArgumentManagementType managementType, const bool referenceUserData, const ArgumentId& customId)
{
    return RegisterArgument(sizeof(T), ElementTypeOf<T>(), memoryLocation, accessType, ArgumentMemoryType::Vector, managementType,
        referenceUserData ? DataSource::Reference : DataSource::Copy, data.data(), data.size() * sizeof(T), customId);
}

template <typename T>
ArgumentId Tuner::AddArgumentVector(ComputeBuffer buffer, const size_t bufferSize, const ArgumentAccessType accessType,
    const ArgumentMemoryLocation memoryLocation, const ArgumentId& customId)
{
    return RegisterArgument(sizeof(T), ElementTypeOf<T>(), memoryLocation, accessType, ArgumentMemoryType::Vector,
        ArgumentManagementType::User, DataSource::DeviceBuffer, buffer, bufferSize, customId);
}

template <typename T>
ArgumentId Tuner::AddArgumentScalar(const T& data, const ArgumentId& customId)
{
    return RegisterArgument(sizeof(T), ElementTypeOf<T>(), ArgumentMemoryLocation::Undefined, ArgumentAccessType::ReadOnly,
        ArgumentMemoryType::Scalar, ArgumentManagementType::Framework, DataSource::Copy, const_cast<T*>(&data), sizeof(T), customId);
}

template <typename T>
ArgumentId Tuner::AddArgumentLocal(const size_t localMemorySize, const ArgumentId& customId)
{
    return RegisterArgument(sizeof(T), ElementTypeOf<T>(), ArgumentMemoryLocation::Undefined, ArgumentAccessType::ReadWrite,
        ArgumentMemoryType::Local, ArgumentManagementType::Framework, DataSource::Copy, nullptr, localMemorySize, customId);
}

template <typename T>
ArgumentId Tuner::AddArgumentSymbol(const T& data, const ArgumentId& customId, const std::string& symbolName)
{
    static_assert(std::is_trivially_copyable_v<T>, "Argument data type must be trivially copyable");
    static_assert(!std::is_reference_v<T> && !std::is_pointer_v<T> && !std::is_null_pointer_v<T>,
        "Pointer and reference argument data types are not supported");
    static_assert(!std::is_same_v<std::remove_cv_t<T>, bool>, "Bool argument data type is not supported");
    return RegisterSymbolArgument(sizeof(T), ElementTypeOf<T>(), &data, sizeof(T), customId, symbolName);
}

template <typename T>
ArgumentDataType Tuner::ElementTypeOf()
{
    // Fixed-width types map to their KTT data types; anything else is opaque (Custom).
    using U = std::remove_cv_t<T>;
    if constexpr (std::is_same_v<U, float>) { return ArgumentDataType::Float; }
    else if constexpr (std::is_same_v<U, double>) { return ArgumentDataType::Double; }
    else if constexpr (std::is_same_v<U, char> || std::is_same_v<U, int8_t>) { return ArgumentDataType::Char; }
    else if constexpr (std::is_same_v<U, uint8_t>) { return ArgumentDataType::UnsignedChar; }
    else if constexpr (std::is_same_v<U, int16_t>) { return ArgumentDataType::Short; }
    else if constexpr (std::is_same_v<U, uint16_t>) { return ArgumentDataType::UnsignedShort; }
    else if constexpr (std::is_same_v<U, int32_t>) { return ArgumentDataType::Int; }
    else if constexpr (std::is_same_v<U, uint32_t>) { return ArgumentDataType::UnsignedInt; }
    else if constexpr (std::is_same_v<U, int64_t>) { return ArgumentDataType::Long; }
    else if constexpr (std::is_same_v<U, uint64_t>) { return ArgumentDataType::UnsignedLong; }
    else { return ArgumentDataType::Custom; }
}

} // namespace ktt
