// Kernel model: definitions (source + default sizes + arguments), kernels (one or more
// definitions, tuning parameters, constraints, thread modifiers, launcher) and the
// kernel manager (Source/Kernel/*.cpp of the reference). Thread modifiers are applied
// per dimension in the order they were added (Kernel.cpp:266-308, ThreadModifier.cpp:44-59).
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include <DimensionVector.h>
#include <KernelConfiguration.h>
#include <KttTypes.h>

#include "KernelArgument.h"

namespace ktt
{

class KernelParameter
{
public:
    explicit KernelParameter(const std::string& name, const std::vector<ParameterValue>& values, const std::string& group);

    const std::string& GetName() const { return m_Name; }
    const std::string& GetGroup() const { return m_Group; }
    const std::vector<ParameterValue>& GetValues() const { return m_Values; }
    size_t GetValuesCount() const { return m_Values.size(); }
    ParameterValueType GetValueType() const;
    ParameterPair GeneratePair(const size_t valueIndex) const;
    std::vector<ParameterPair> GeneratePairs() const;

private:
    std::string m_Name;
    std::vector<ParameterValue> m_Values;
    std::string m_Group;
};

class KernelConstraint
{
public:
    explicit KernelConstraint(const std::vector<std::string>& parameters, ConstraintFunction function);
    explicit KernelConstraint(const std::vector<std::string>& parameters, GenericConstraintFunction function);

    const std::vector<std::string>& GetParameters() const { return m_Parameters; }
    bool HasAllParameters(const std::vector<std::string>& parameters) const;
    bool IsFulfilled(const std::vector<const ParameterValue*>& values) const;

private:
    std::vector<std::string> m_Parameters;
    ConstraintFunction m_Function;
    GenericConstraintFunction m_GenericFunction;
};

class ThreadModifier
{
public:
    explicit ThreadModifier(const std::vector<std::string>& parameters, const std::vector<KernelDefinitionId>& definitions,
        ModifierFunction function);

    const std::vector<std::string>& GetParameters() const { return m_Parameters; }
    const std::vector<KernelDefinitionId>& GetDefinitions() const { return m_Definitions; }
    uint64_t GetModifiedSize(const KernelDefinitionId id, const uint64_t originalSize, const std::vector<ParameterPair>& pairs) const;

private:
    std::vector<std::string> m_Parameters;
    std::vector<KernelDefinitionId> m_Definitions;
    ModifierFunction m_Function;
};

class KernelDefinition
{
public:
    explicit KernelDefinition(const KernelDefinitionId id, const std::string& name, const std::string& source,
        const DimensionVector& globalSize, const DimensionVector& localSize, const std::vector<std::string>& typeNames,
        const std::string& sourceDirectory);

    void SetArguments(const std::vector<KernelArgument*> arguments);

    KernelDefinitionId GetId() const { return m_Id; }
    const std::string& GetName() const { return m_Name; }
    const std::string& GetSource() const { return m_Source; }
    const std::string& GetTemplatedName() const { return m_TemplatedName; }
    const std::vector<std::string>& GetTypeNames() const { return m_TypeNames; }
    const DimensionVector& GetGlobalSize() const { return m_GlobalSize; }
    const DimensionVector& GetLocalSize() const { return m_LocalSize; }
    const std::vector<KernelArgument*>& GetArguments() const { return m_Arguments; }
    std::vector<KernelArgument*> GetVectorArguments() const;
    bool HasArgument(const ArgumentId& id) const;
    /** Directory of the file the source came from ("" for sources given as strings); added to the include path. */
    const std::string& GetSourceDirectory() const { return m_SourceDirectory; }
    bool IsTemplated() const { return !m_TypeNames.empty(); }

    static std::string CreateTemplatedName(const std::string& name, const std::vector<std::string>& typeNames);

private:
    KernelDefinitionId m_Id;
    std::string m_Name;
    std::string m_Source;
    std::string m_TemplatedName;
    std::vector<std::string> m_TypeNames;
    DimensionVector m_GlobalSize;
    DimensionVector m_LocalSize;
    std::vector<KernelArgument*> m_Arguments;
    std::string m_SourceDirectory;
};

class Kernel
{
public:
    explicit Kernel(const KernelId id, const std::string& name, const std::vector<const KernelDefinition*>& definitions);

    void AddParameter(const KernelParameter& parameter);
    void AddConstraint(const KernelConstraint& constraint);
    void AddThreadModifier(const ModifierType type, const ModifierDimension dimension, const ThreadModifier& modifier);
    void SetLauncher(KernelLauncher launcher);
    void SetProfiledDefinitions(const std::vector<const KernelDefinition*>& definitions);

    KernelId GetId() const { return m_Id; }
    const std::string& GetName() const { return m_Name; }
    const KernelDefinition& GetPrimaryDefinition() const;
    const KernelDefinition& GetDefinition(const KernelDefinitionId id) const;
    const std::vector<const KernelDefinition*>& GetDefinitions() const { return m_Definitions; }
    const std::vector<const KernelDefinition*>& GetProfiledDefinitions() const { return m_ProfiledDefinitions; }
    const std::vector<KernelParameter>& GetParameters() const { return m_Parameters; }
    const KernelParameter& GetParameter(const std::string& name) const;
    const std::vector<KernelConstraint>& GetConstraints() const { return m_Constraints; }
    std::vector<KernelArgument*> GetVectorArguments() const;
    bool HasDefinition(const KernelDefinitionId id) const;
    bool HasParameter(const std::string& name) const;
    bool HasLauncher() const { return static_cast<bool>(m_Launcher); }
    KernelLauncher GetLauncher() const { return m_Launcher; }
    bool IsComposite() const { return m_Definitions.size() > 1; }

    KernelConfiguration CreateConfiguration(const ParameterInput& parameters) const;
    DimensionVector GetModifiedSize(const KernelDefinitionId id, const ModifierType type, const std::vector<ParameterPair>& pairs) const;
    DimensionVector GetModifiedSize(const KernelDefinitionId id, const DimensionVector& originalSize, const ModifierType type,
        const std::vector<ParameterPair>& pairs) const;

private:
    KernelId m_Id;
    std::string m_Name;
    std::vector<const KernelDefinition*> m_Definitions;
    std::vector<const KernelDefinition*> m_ProfiledDefinitions;
    std::vector<KernelParameter> m_Parameters;
    std::vector<KernelConstraint> m_Constraints;
    std::map<ModifierType, std::map<ModifierDimension, std::vector<ThreadModifier>>> m_Modifiers;
    KernelLauncher m_Launcher;
};

class KernelManager
{
public:
    explicit KernelManager(KernelArgumentManager& argumentManager);

    KernelDefinitionId AddKernelDefinition(const std::string& name, const std::string& source, const DimensionVector& globalSize,
        const DimensionVector& localSize, const std::vector<std::string>& typeNames, const std::string& sourceDirectory = "");
    KernelDefinitionId AddKernelDefinitionFromFile(const std::string& name, const std::string& filePath,
        const DimensionVector& globalSize, const DimensionVector& localSize, const std::vector<std::string>& typeNames);
    KernelDefinitionId GetDefinitionId(const std::string& name, const std::vector<std::string>& typeNames) const;
    void RemoveKernelDefinition(const KernelDefinitionId id);
    void SetArguments(const KernelDefinitionId id, const std::vector<ArgumentId>& argumentIds);

    KernelId CreateKernel(const std::string& name, const std::vector<KernelDefinitionId>& definitionIds);
    void RemoveKernel(const KernelId id);
    void SetLauncher(const KernelId id, KernelLauncher launcher);
    void AddParameter(const KernelId id, const std::string& name, const std::vector<ParameterValue>& values, const std::string& group);
    void AddConstraint(const KernelId id, const std::vector<std::string>& parameters, ConstraintFunction function);
    void AddGenericConstraint(const KernelId id, const std::vector<std::string>& parameters, GenericConstraintFunction function);
    void AddThreadModifier(const KernelId id, const std::vector<KernelDefinitionId>& definitionIds, const ModifierType type,
        const ModifierDimension dimension, const std::vector<std::string>& parameters, ModifierFunction function);
    void SetProfiledDefinitions(const KernelId id, const std::vector<KernelDefinitionId>& definitionIds);

    Kernel& GetKernel(const KernelId id);
    const Kernel& GetKernel(const KernelId id) const;
    KernelDefinition& GetDefinition(const KernelDefinitionId id);
    const KernelDefinition& GetDefinition(const KernelDefinitionId id) const;
    bool HasKernel(const KernelId id) const;
    bool HasDefinition(const KernelDefinitionId id) const;
    bool IsArgumentUsed(const ArgumentId& id) const;

private:
    KernelArgumentManager& m_ArgumentManager;
    std::map<KernelDefinitionId, std::unique_ptr<KernelDefinition>> m_Definitions;
    std::map<KernelId, std::unique_ptr<Kernel>> m_Kernels;
    KernelDefinitionId m_NextDefinitionId;
    KernelId m_NextKernelId;
};

std::string LoadFileToString(const std::string& filePath);

} // namespace ktt
