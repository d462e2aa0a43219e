#pragma once

#include <type_traits>

#include <KttException.h>

namespace ktt
{

template <typename T>
T ParameterPair::GetParameterValue(const std::vector<ParameterPair>& pairs, const std::string& name)
{
    for (const auto& pair : pairs)
    {
        if (pair.GetName() != name)
        {
            continue;
        }

        const auto& value = pair.GetValue();

        if constexpr (std::is_same_v<T, std::string>)
        {
            return pair.GetValueString();
        }
        else if constexpr (std::is_same_v<T, bool>)
        {
            if (std::holds_alternative<bool>(value))
            {
                return std::get<bool>(value);
            }
            return pair.GetValueUint() != 0;
        }
        else if constexpr (std::is_floating_point_v<T>)
        {
            if (std::holds_alternative<double>(value))
            {
                return static_cast<T>(std::get<double>(value));
            }
            if (std::holds_alternative<int64_t>(value))
            {
                return static_cast<T>(std::get<int64_t>(value));
            }
            return static_cast<T>(pair.GetValueUint());
        }
        else
        {
            if (std::holds_alternative<int64_t>(value))
            {
                return static_cast<T>(std::get<int64_t>(value));
            }
            if (std::holds_alternative<double>(value))
            {
                return static_cast<T>(std::get<double>(value));
            }
            return static_cast<T>(pair.GetValueUint());
        }
    }

    throw KttException("Parameter with name " + name + " was not found");
}

template <typename T>
std::vector<T> ParameterPair::GetParameterValues(const std::vector<ParameterPair>& pairs, const std::vector<std::string>& names)
{
    std::vector<T> result;
    result.reserve(names.size());

    for (const auto& name : names)
    {
        result.push_back(GetParameterValue<T>(pairs, name));
    }

    return result;
}

} // namespace ktt
