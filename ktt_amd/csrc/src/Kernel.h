// Kernels and their registry.
//
//   KernelDefinition  one kernel function: name (with template arguments), source, default
//                     global and local sizes, bound arguments.
//   Kernel            what is tuned: one or more definitions, tuning parameters (with groups),
//                     constraints, thread modifiers and an optional launcher.
//   KernelRegistry    owns definitions and kernels and hands out their ids.
//
// A thread modifier maps a size along one dimension through a function of parameter values;
// the modifiers of a dimension apply in the order they were added.
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
    KernelParameter(const std::string& name, const std::vector<ParameterValue>& values, const std::string& group);

    const std::string& GetName() const { return m_Name; }
    const std::string& GetGroup() const { return m_Group; }
    const std::vector<ParameterValue>& GetValues() const { return m_Values; }
    size_t GetValuesCount() const { return m_Values.size(); }
    ParameterValueType GetValueType() const { return m_Type; }
    /** The pair (name, values[index]). */
    ParameterPair GeneratePair(const size_t index) const { return ParameterPair(m_Name, m_Values.at(index)); }

private:
    std::string m_Name;
    std::string m_Group;
    std::vector<ParameterValue> m_Values;
    ParameterValueType m_Type;
};

class KernelConstraint
{
public:
    KernelConstraint(const std::vector<std::string>& parameters, ConstraintFunction function);
    KernelConstraint(const std::vector<std::string>& parameters, GenericConstraintFunction function);

    const std::vector<std::string>& GetParameters() const { return m_Parameters; }
    /** Evaluates the constraint on values given in the order of GetParameters(). */
    bool IsFulfilled(const std::vector<const ParameterValue*>& values) const;

private:
    std::vector<std::string> m_Parameters;
    ConstraintFunction m_OnIntegers;
    GenericConstraintFunction m_OnValues;
};

class KernelDefinition
{
public:
    KernelDefinition(const KernelDefinitionId id, const std::string& name, const std::vector<std::string>& typeNames,
        const std::string& source, const std::string& sourceDirectory, const DimensionVector& globalSize, const DimensionVector& localSize);

    /** Binds arguments in kernel-signature order; an argument may appear only once. */
    void SetArguments(const std::vector<KernelArgument*>& arguments);

    KernelDefinitionId GetId() const { return m_Id; }
    const std::string& GetName() const { return m_Name; }
    const std::string& GetTemplatedName() const { return m_Templated; }
    const std::string& GetSource() const { return m_Source; }
    /** Directory of the source file ("" for sources given as text); it is on the include path. */
    const std::string& GetSourceDirectory() const { return m_Directory; }
    const DimensionVector& GetGlobalSize() const { return m_Global; }
    const DimensionVector& GetLocalSize() const { return m_Local; }
    const std::vector<KernelArgument*>& GetArguments() const { return m_Arguments; }
    std::vector<KernelArgument*> GetVectorArguments() const { return VectorArguments(m_Arguments); }
    bool HasArgument(const ArgumentId& id) const;

    /** "name<T1, T2>" for template arguments T1, T2; the plain name without any. */
    static std::string TemplatedName(const std::string& name, const std::vector<std::string>& typeNames);

private:
    KernelDefinitionId m_Id;
    std::string m_Name;
    std::string m_Templated;
    std::string m_Source;
    std::string m_Directory;
    DimensionVector m_Global;
    DimensionVector m_Local;
    std::vector<KernelArgument*> m_Arguments;
};

class Kernel
{
public:
    Kernel(const KernelId id, const std::string& name, const std::vector<const KernelDefinition*>& definitions);

    void AddParameter(const KernelParameter& parameter);
    /** All of a constraint's parameters must exist and share one group. */
    void AddConstraint(const KernelConstraint& constraint);
    void AddThreadModifier(const ModifierType type, const ModifierDimension dimension, const std::vector<std::string>& parameters,
        const std::vector<KernelDefinitionId>& definitions, ModifierFunction function);
    void SetLauncher(KernelLauncher launcher) { m_Launcher = launcher; }
    void SetProfiledDefinitions(const std::vector<const KernelDefinition*>& definitions);

    KernelId GetId() const { return m_Id; }
    const std::string& GetName() const { return m_Name; }
    const KernelDefinition& GetPrimaryDefinition() const { return *m_Definitions.front(); }
    const KernelDefinition& GetDefinition(const KernelDefinitionId id) const;
    const std::vector<const KernelDefinition*>& GetDefinitions() const { return m_Definitions; }
    const std::vector<const KernelDefinition*>& GetProfiledDefinitions() const { return m_Profiled; }
    const std::vector<KernelParameter>& GetParameters() const { return m_Parameters; }
    const KernelParameter& GetParameter(const std::string& name) const;
    const std::vector<KernelConstraint>& GetConstraints() const { return m_Constraints; }
    /** The vector arguments of all definitions, each once. */
    std::vector<KernelArgument*> GetVectorArguments() const;
    bool HasDefinition(const KernelDefinitionId id) const;
    bool HasParameter(const std::string& name) const { return FindParameter(name) != nullptr; }
    bool HasLauncher() const { return static_cast<bool>(m_Launcher); }
    KernelLauncher GetLauncher() const { return m_Launcher; }
    bool IsComposite() const { return m_Definitions.size() > 1; }

    /** A full configuration: the given values, and each other parameter's first value. */
    KernelConfiguration CreateConfiguration(const ParameterInput& parameters) const;
    /** `size` of definition `id` after the modifiers of `type` under the configuration `pairs`. */
    DimensionVector GetModifiedSize(const KernelDefinitionId id, const DimensionVector& size, const ModifierType type,
        const std::vector<ParameterPair>& pairs) const;

private:
    struct Modifier
    {
        ModifierType type;
        ModifierDimension dimension;
        std::vector<std::string> parameters;
        std::vector<KernelDefinitionId> definitions;
        ModifierFunction function;
    };

    KernelId m_Id;
    std::string m_Name;
    std::vector<const KernelDefinition*> m_Definitions;
    std::vector<const KernelDefinition*> m_Profiled;
    std::vector<KernelParameter> m_Parameters;
    std::vector<KernelConstraint> m_Constraints;
    std::vector<Modifier> m_Modifiers;
    KernelLauncher m_Launcher;

    const KernelParameter* FindParameter(const std::string& name) const;
};

class KernelRegistry
{
public:
    explicit KernelRegistry(ArgumentRegistry& arguments);

    KernelDefinitionId AddDefinition(const std::string& name, const std::string& source, const DimensionVector& globalSize,
        const DimensionVector& localSize, const std::vector<std::string>& typeNames, const std::string& sourceDirectory = "");
    /** Reads the source from `filePath`; its directory goes on the include path. */
    KernelDefinitionId AddDefinitionFromFile(const std::string& name, const std::string& filePath, const DimensionVector& globalSize,
        const DimensionVector& localSize, const std::vector<std::string>& typeNames);
    KernelDefinitionId FindDefinition(const std::string& name, const std::vector<std::string>& typeNames) const;
    void RemoveDefinition(const KernelDefinitionId id);
    void Bind(const KernelDefinitionId id, const std::vector<ArgumentId>& arguments);

    KernelId AddKernel(const std::string& name, const std::vector<KernelDefinitionId>& definitions);
    void RemoveKernel(const KernelId id);
    std::vector<const KernelDefinition*> Definitions(const std::vector<KernelDefinitionId>& ids) const;

    Kernel& GetKernel(const KernelId id);
    const Kernel& GetKernel(const KernelId id) const;
    KernelDefinition& GetDefinition(const KernelDefinitionId id);
    const KernelDefinition& GetDefinition(const KernelDefinitionId id) const;
    bool HasKernel(const KernelId id) const { return m_Kernels.count(id) != 0; }
    bool HasDefinition(const KernelDefinitionId id) const { return m_Definitions.count(id) != 0; }
    /** True when a definition still binds the argument. */
    bool UsesArgument(const ArgumentId& id) const;

private:
    ArgumentRegistry& m_Arguments;
    std::map<KernelDefinitionId, std::unique_ptr<KernelDefinition>> m_Definitions;
    std::map<KernelId, std::unique_ptr<Kernel>> m_Kernels;
    KernelDefinitionId m_NextDefinition;
    KernelId m_NextKernel;
};

/** The whole content of a file; throws if it cannot be read. */
std::string ReadTextFile(const std::string& path);

} // namespace ktt
