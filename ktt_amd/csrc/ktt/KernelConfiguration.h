/** @file KernelConfiguration.h
  * A tuning configuration: ordered (parameter name, value) pairs.
  * The define prefix is "#define NAME VALUE\n" per pair, in pair order
  * (Source/Api/Configuration/KernelConfiguration.cpp:25-35, checked by
  * Tests/KernelManagerTests.cpp:57-81); value strings follow
  * Source/Api/Configuration/ParameterPair.cpp:29-53.
  */
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include <KttTypes.h>

namespace ktt
{

class ParameterPair
{
public:
    ParameterPair();
    explicit ParameterPair(const std::string& name, const ParameterValue& value);

    void SetValue(const ParameterValue& value);

    const std::string& GetName() const;
    std::string GetString() const;
    std::string GetValueString() const;
    const ParameterValue& GetValue() const;
    uint64_t GetValueUint() const;
    ParameterValueType GetValueType() const;
    bool HasValueDouble() const;
    bool HasSameValue(const ParameterPair& other) const;

    static ParameterValueType GetTypeFromValue(const ParameterValue& value);

    template <typename T>
    static T GetParameterValue(const std::vector<ParameterPair>& pairs, const std::string& name);

    template <typename T>
    static std::vector<T> GetParameterValues(const std::vector<ParameterPair>& pairs, const std::vector<std::string>& names);

    bool operator==(const ParameterPair& other) const;
    bool operator!=(const ParameterPair& other) const;

private:
    std::string m_Name;
    ParameterValue m_Value;
};

class KernelConfiguration
{
public:
    KernelConfiguration();
    explicit KernelConfiguration(const std::vector<ParameterPair>& pairs);

    const std::vector<ParameterPair>& GetPairs() const;
    bool IsValid() const;
    std::string GeneratePrefix() const;
    /** "-DNAME=VALUE" per pair, the form in which the HIP backend hands the configuration to hipRTC. */
    std::vector<std::string> GenerateDefines() const;
    std::string GetString() const;

    void Merge(const KernelConfiguration& other);
    std::vector<KernelConfiguration> GenerateNeighbours(const std::string& parameter, const std::vector<ParameterPair>& pairs) const;

    bool operator==(const KernelConfiguration& other) const;
    bool operator!=(const KernelConfiguration& other) const;

private:
    std::vector<ParameterPair> m_Pairs;
};

} // namespace ktt

#include <ParameterPair.inl>
