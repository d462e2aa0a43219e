#include <fstream>
#include <sstream>

#include <KttException.h>

#include "Common.h"
#include "Kernel.h"

namespace ktt
{

std::string LoadFileToString(const std::string& filePath)
{
    std::ifstream file(filePath, std::ios::binary);
    if (!file.is_open())
    {
        throw KttException("Unable to open file: " + filePath);
    }
    std::stringstream stream;
    stream << file.rdbuf();
    return stream.str();
}

// ---------------------------------------------------------------- KernelParameter

KernelParameter::KernelParameter(const std::string& name, const std::vector<ParameterValue>& values, const std::string& group) :
    m_Name(name),
    m_Values(values),
    m_Group(group)
{
    if (values.empty())
    {
        throw KttException("Kernel parameter with name " + name + " must have at least one value");
    }
    const auto type = ParameterPair::GetTypeFromValue(values[0]);
    for (const auto& value : values)
    {
        if (ParameterPair::GetTypeFromValue(value) != type)
        {
            throw KttException("All values of kernel parameter with name " + name + " must have the same type");
        }
    }
}

ParameterValueType KernelParameter::GetValueType() const
{
    return ParameterPair::GetTypeFromValue(m_Values[0]);
}

ParameterPair KernelParameter::GeneratePair(const size_t valueIndex) const
{
    return ParameterPair(m_Name, m_Values.at(valueIndex));
}

std::vector<ParameterPair> KernelParameter::GeneratePairs() const
{
    std::vector<ParameterPair> result;
    for (size_t i = 0; i < m_Values.size(); ++i)
    {
        result.push_back(GeneratePair(i));
    }
    return result;
}

// ---------------------------------------------------------------- KernelConstraint

KernelConstraint::KernelConstraint(const std::vector<std::string>& parameters, ConstraintFunction function) :
    m_Parameters(parameters),
    m_Function(function),
    m_GenericFunction(nullptr)
{
    if (!function)
    {
        throw KttException("Constraint function must be valid");
    }
}

KernelConstraint::KernelConstraint(const std::vector<std::string>& parameters, GenericConstraintFunction function) :
    m_Parameters(parameters),
    m_Function(nullptr),
    m_GenericFunction(function)
{
    if (!function)
    {
        throw KttException("Constraint function must be valid");
    }
}

bool KernelConstraint::HasAllParameters(const std::vector<std::string>& parameters) const
{
    for (const auto& name : m_Parameters)
    {
        if (!ContainsElement(parameters, name))
        {
            return false;
        }
    }
    return true;
}

bool KernelConstraint::IsFulfilled(const std::vector<const ParameterValue*>& values) const
{
    if (m_GenericFunction)
    {
        return m_GenericFunction(values);
    }

    std::vector<uint64_t> integers;
    integers.reserve(values.size());
    for (const auto* value : values)
    {
        if (std::holds_alternative<uint64_t>(*value))
        {
            integers.push_back(std::get<uint64_t>(*value));
        }
        else if (std::holds_alternative<int64_t>(*value))
        {
            integers.push_back(static_cast<uint64_t>(std::get<int64_t>(*value)));
        }
        else if (std::holds_alternative<bool>(*value))
        {
            integers.push_back(std::get<bool>(*value) ? 1 : 0);
        }
        else if (std::holds_alternative<double>(*value))
        {
            integers.push_back(static_cast<uint64_t>(std::get<double>(*value)));
        }
        else
        {
            throw KttException("Constraints over string parameters require a generic constraint");
        }
    }
    return m_Function(integers);
}

// ---------------------------------------------------------------- ThreadModifier

ThreadModifier::ThreadModifier(const std::vector<std::string>& parameters, const std::vector<KernelDefinitionId>& definitions,
    ModifierFunction function) :
    m_Parameters(parameters),
    m_Definitions(definitions),
    m_Function(function)
{
    if (!function)
    {
        throw KttException("Thread modifier function must be valid");
    }
}

uint64_t ThreadModifier::GetModifiedSize(const KernelDefinitionId id, const uint64_t originalSize,
    const std::vector<ParameterPair>& pairs) const
{
    if (!ContainsElement(m_Definitions, id))
    {
        return originalSize;
    }
    const std::vector<uint64_t> values = ParameterPair::GetParameterValues<uint64_t>(pairs, m_Parameters);
    return m_Function(originalSize, values);
}

// ---------------------------------------------------------------- KernelDefinition

KernelDefinition::KernelDefinition(const KernelDefinitionId id, const std::string& name, const std::string& source,
    const DimensionVector& globalSize, const DimensionVector& localSize, const std::vector<std::string>& typeNames,
    const std::string& sourceDirectory) :
    m_Id(id),
    m_Name(name),
    m_Source(source),
    m_TemplatedName(CreateTemplatedName(name, typeNames)),
    m_TypeNames(typeNames),
    m_GlobalSize(globalSize),
    m_LocalSize(localSize),
    m_SourceDirectory(sourceDirectory)
{}

void KernelDefinition::SetArguments(const std::vector<KernelArgument*> arguments)
{
    for (size_t i = 0; i < arguments.size(); ++i)
    {
        for (size_t j = i + 1; j < arguments.size(); ++j)
        {
            if (arguments[i] == arguments[j])
            {
                throw KttException("Kernel arguments for a single kernel definition must be unique");
            }
        }
    }
    m_Arguments = arguments;
}

std::vector<KernelArgument*> KernelDefinition::GetVectorArguments() const
{
    return KernelArgument::GetArgumentsWithMemoryType(m_Arguments, ArgumentMemoryType::Vector);
}

bool KernelDefinition::HasArgument(const ArgumentId& id) const
{
    return ContainsElementIf(m_Arguments, [&id](const auto* argument) { return argument->GetId() == id; });
}

std::string KernelDefinition::CreateTemplatedName(const std::string& name, const std::vector<std::string>& typeNames)
{
    if (typeNames.empty())
    {
        return name;
    }
    std::string result = name + "<";
    for (size_t i = 0; i < typeNames.size(); ++i)
    {
        result += typeNames[i];
        if (i + 1 < typeNames.size())
        {
            result += ", ";
        }
    }
    return result + ">";
}

// ---------------------------------------------------------------- Kernel

Kernel::Kernel(const KernelId id, const std::string& name, const std::vector<const KernelDefinition*>& definitions) :
    m_Id(id),
    m_Name(name),
    m_Definitions(definitions),
    m_ProfiledDefinitions(definitions.empty() ? std::vector<const KernelDefinition*>{} : std::vector<const KernelDefinition*>{definitions[0]})
{
    if (definitions.empty())
    {
        throw KttException("Kernel must have at least one definition");
    }
}

void Kernel::AddParameter(const KernelParameter& parameter)
{
    if (HasParameter(parameter.GetName()))
    {
        throw KttException("Kernel parameter with name " + parameter.GetName() + " already exists");
    }
    m_Parameters.push_back(parameter);
}

void Kernel::AddConstraint(const KernelConstraint& constraint)
{
    std::string group;
    bool first = true;
    for (const auto& name : constraint.GetParameters())
    {
        const auto& parameter = GetParameter(name);
        if (!first && parameter.GetGroup() != group)
        {
            throw KttException("Constraint can only be added between parameters that belong to the same group");
        }
        group = parameter.GetGroup();
        first = false;
    }
    m_Constraints.push_back(constraint);
}

void Kernel::AddThreadModifier(const ModifierType type, const ModifierDimension dimension, const ThreadModifier& modifier)
{
    for (const auto& name : modifier.GetParameters())
    {
        const auto& parameter = GetParameter(name);
        const auto valueType = parameter.GetValueType();
        if (valueType != ParameterValueType::UnsignedInt && valueType != ParameterValueType::Int)
        {
            throw KttException("Parameter with name " + name + " used by a thread modifier must have integer values");
        }
    }
    for (const auto id : modifier.GetDefinitions())
    {
        if (!HasDefinition(id))
        {
            throw KttException("Kernel with id " + std::to_string(m_Id) + " does not contain definition with id "
                + std::to_string(id));
        }
    }
    m_Modifiers[type][dimension].push_back(modifier);
}

void Kernel::SetLauncher(KernelLauncher launcher)
{
    m_Launcher = launcher;
}

void Kernel::SetProfiledDefinitions(const std::vector<const KernelDefinition*>& definitions)
{
    for (const auto* definition : definitions)
    {
        if (!HasDefinition(definition->GetId()))
        {
            throw KttException("Kernel with id " + std::to_string(m_Id) + " does not contain definition with id "
                + std::to_string(definition->GetId()));
        }
    }
    m_ProfiledDefinitions = definitions;
}

const KernelDefinition& Kernel::GetPrimaryDefinition() const
{
    return *m_Definitions[0];
}

const KernelDefinition& Kernel::GetDefinition(const KernelDefinitionId id) const
{
    for (const auto* definition : m_Definitions)
    {
        if (definition->GetId() == id)
        {
            return *definition;
        }
    }
    throw KttException("Kernel with id " + std::to_string(m_Id) + " does not contain definition with id " + std::to_string(id));
}

const KernelParameter& Kernel::GetParameter(const std::string& name) const
{
    for (const auto& parameter : m_Parameters)
    {
        if (parameter.GetName() == name)
        {
            return parameter;
        }
    }
    throw KttException("Kernel parameter with name " + name + " does not exist");
}

std::vector<KernelArgument*> Kernel::GetVectorArguments() const
{
    std::vector<KernelArgument*> result;
    for (const auto* definition : m_Definitions)
    {
        for (auto* argument : definition->GetVectorArguments())
        {
            if (!ContainsElement(result, argument))
            {
                result.push_back(argument);
            }
        }
    }
    return result;
}

bool Kernel::HasDefinition(const KernelDefinitionId id) const
{
    return ContainsElementIf(m_Definitions, [id](const auto* definition) { return definition->GetId() == id; });
}

bool Kernel::HasParameter(const std::string& name) const
{
    return ContainsElementIf(m_Parameters, [&name](const auto& parameter) { return parameter.GetName() == name; });
}

KernelConfiguration Kernel::CreateConfiguration(const ParameterInput& parameters) const
{
    std::vector<ParameterPair> pairs;
    for (const auto& [name, value] : parameters)
    {
        const auto& parameter = GetParameter(name);
        if (ParameterPair::GetTypeFromValue(value) != parameter.GetValueType())
        {
            throw KttException("Value type mismatch for parameter with name " + name);
        }
        pairs.emplace_back(name, value);
    }
    for (const auto& parameter : m_Parameters)
    {
        const bool included = ContainsElementIf(pairs, [&parameter](const auto& pair) { return pair.GetName() == parameter.GetName(); });
        if (!included)
        {
            pairs.push_back(parameter.GeneratePair(0));
        }
    }
    return KernelConfiguration(pairs);
}

DimensionVector Kernel::GetModifiedSize(const KernelDefinitionId id, const ModifierType type, const std::vector<ParameterPair>& pairs) const
{
    const auto& definition = GetDefinition(id);
    const auto& defaultSize = type == ModifierType::Global ? definition.GetGlobalSize() : definition.GetLocalSize();
    return GetModifiedSize(id, defaultSize, type, pairs);
}

DimensionVector Kernel::GetModifiedSize(const KernelDefinitionId id, const DimensionVector& originalSize, const ModifierType type,
    const std::vector<ParameterPair>& pairs) const
{
    auto typeIterator = m_Modifiers.find(type);
    if (typeIterator == m_Modifiers.cend())
    {
        return originalSize;
    }

    DimensionVector result;
    for (int i = 0; i <= static_cast<int>(ModifierDimension::Z); ++i)
    {
        const auto dimension = static_cast<ModifierDimension>(i);
        uint64_t size = originalSize.GetSize(dimension);
        auto dimensionIterator = typeIterator->second.find(dimension);
        if (dimensionIterator != typeIterator->second.cend())
        {
            for (const auto& modifier : dimensionIterator->second)
            {
                size = modifier.GetModifiedSize(id, size, pairs);
            }
        }
        result.SetSize(dimension, static_cast<size_t>(size));
    }
    return result;
}

// ---------------------------------------------------------------- KernelManager

KernelManager::KernelManager(KernelArgumentManager& argumentManager) :
    m_ArgumentManager(argumentManager),
    m_NextDefinitionId(0),
    m_NextKernelId(0)
{}

KernelDefinitionId KernelManager::AddKernelDefinition(const std::string& name, const std::string& source,
    const DimensionVector& globalSize, const DimensionVector& localSize, const std::vector<std::string>& typeNames,
    const std::string& sourceDirectory)
{
    if (GetDefinitionId(name, typeNames) != InvalidKernelDefinitionId)
    {
        throw KttException("Kernel definition with name " + name + " already exists");
    }
    const KernelDefinitionId id = m_NextDefinitionId++;
    m_Definitions[id] = std::make_unique<KernelDefinition>(id, name, source, globalSize, localSize, typeNames, sourceDirectory);
    return id;
}

KernelDefinitionId KernelManager::AddKernelDefinitionFromFile(const std::string& name, const std::string& filePath,
    const DimensionVector& globalSize, const DimensionVector& localSize, const std::vector<std::string>& typeNames)
{
    const std::string source = LoadFileToString(filePath);
    const size_t slash = filePath.find_last_of('/');
    const std::string directory = slash == std::string::npos ? std::string(".") : filePath.substr(0, slash);
    return AddKernelDefinition(name, source, globalSize, localSize, typeNames, directory);
}

KernelDefinitionId KernelManager::GetDefinitionId(const std::string& name, const std::vector<std::string>& typeNames) const
{
    const std::string templatedName = KernelDefinition::CreateTemplatedName(name, typeNames);
    for (const auto& [id, definition] : m_Definitions)
    {
        if (definition->GetTemplatedName() == templatedName)
        {
            return id;
        }
    }
    return InvalidKernelDefinitionId;
}

void KernelManager::RemoveKernelDefinition(const KernelDefinitionId id)
{
    for (const auto& [kernelId, kernel] : m_Kernels)
    {
        if (kernel->HasDefinition(id))
        {
            throw KttException("Kernel definition with id " + std::to_string(id)
                + " cannot be removed because it is still referenced by kernel with id " + std::to_string(kernelId));
        }
    }
    if (m_Definitions.erase(id) == 0)
    {
        throw KttException("Attempting to remove kernel definition with invalid id: " + std::to_string(id));
    }
}

void KernelManager::SetArguments(const KernelDefinitionId id, const std::vector<ArgumentId>& argumentIds)
{
    auto& definition = GetDefinition(id);
    definition.SetArguments(m_ArgumentManager.GetArguments(argumentIds));
}

KernelId KernelManager::CreateKernel(const std::string& name, const std::vector<KernelDefinitionId>& definitionIds)
{
    for (const auto& [id, kernel] : m_Kernels)
    {
        if (kernel->GetName() == name)
        {
            throw KttException("Kernel with name " + name + " already exists");
        }
    }
    std::vector<const KernelDefinition*> definitions;
    for (const auto definitionId : definitionIds)
    {
        definitions.push_back(&GetDefinition(definitionId));
    }
    const KernelId id = m_NextKernelId++;
    m_Kernels[id] = std::make_unique<Kernel>(id, name, definitions);
    return id;
}

void KernelManager::RemoveKernel(const KernelId id)
{
    if (m_Kernels.erase(id) == 0)
    {
        throw KttException("Attempting to remove kernel with invalid id: " + std::to_string(id));
    }
}

void KernelManager::SetLauncher(const KernelId id, KernelLauncher launcher)
{
    GetKernel(id).SetLauncher(launcher);
}

void KernelManager::AddParameter(const KernelId id, const std::string& name, const std::vector<ParameterValue>& values,
    const std::string& group)
{
    GetKernel(id).AddParameter(KernelParameter(name, values, group));
}

void KernelManager::AddConstraint(const KernelId id, const std::vector<std::string>& parameters, ConstraintFunction function)
{
    GetKernel(id).AddConstraint(KernelConstraint(parameters, function));
}

void KernelManager::AddGenericConstraint(const KernelId id, const std::vector<std::string>& parameters,
    GenericConstraintFunction function)
{
    GetKernel(id).AddConstraint(KernelConstraint(parameters, function));
}

void KernelManager::AddThreadModifier(const KernelId id, const std::vector<KernelDefinitionId>& definitionIds,
    const ModifierType type, const ModifierDimension dimension, const std::vector<std::string>& parameters, ModifierFunction function)
{
    GetKernel(id).AddThreadModifier(type, dimension, ThreadModifier(parameters, definitionIds, function));
}

void KernelManager::SetProfiledDefinitions(const KernelId id, const std::vector<KernelDefinitionId>& definitionIds)
{
    std::vector<const KernelDefinition*> definitions;
    for (const auto definitionId : definitionIds)
    {
        definitions.push_back(&GetDefinition(definitionId));
    }
    GetKernel(id).SetProfiledDefinitions(definitions);
}

Kernel& KernelManager::GetKernel(const KernelId id)
{
    auto iterator = m_Kernels.find(id);
    if (iterator == m_Kernels.end())
    {
        throw KttException("Attempting to retrieve kernel with invalid id: " + std::to_string(id));
    }
    return *iterator->second;
}

const Kernel& KernelManager::GetKernel(const KernelId id) const
{
    auto iterator = m_Kernels.find(id);
    if (iterator == m_Kernels.end())
    {
        throw KttException("Attempting to retrieve kernel with invalid id: " + std::to_string(id));
    }
    return *iterator->second;
}

KernelDefinition& KernelManager::GetDefinition(const KernelDefinitionId id)
{
    auto iterator = m_Definitions.find(id);
    if (iterator == m_Definitions.end())
    {
        throw KttException("Attempting to retrieve kernel definition with invalid id: " + std::to_string(id));
    }
    return *iterator->second;
}

const KernelDefinition& KernelManager::GetDefinition(const KernelDefinitionId id) const
{
    auto iterator = m_Definitions.find(id);
    if (iterator == m_Definitions.end())
    {
        throw KttException("Attempting to retrieve kernel definition with invalid id: " + std::to_string(id));
    }
    return *iterator->second;
}

bool KernelManager::HasKernel(const KernelId id) const
{
    return ContainsKey(m_Kernels, id);
}

bool KernelManager::HasDefinition(const KernelDefinitionId id) const
{
    return ContainsKey(m_Definitions, id);
}

bool KernelManager::IsArgumentUsed(const ArgumentId& id) const
{
    for (const auto& [definitionId, definition] : m_Definitions)
    {
        if (definition->HasArgument(id))
        {
            return true;
        }
    }
    return false;
}

} // namespace ktt
