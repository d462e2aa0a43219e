#include <algorithm>
#include <fstream>
#include <iterator>

#include <KttException.h>

#include "Kernel.h"

namespace ktt
{

std::string ReadTextFile(const std::string& path)
{
    std::ifstream file(path, std::ios::binary);
    if (!file)
    {
        throw KttException("Cannot read " + path);
    }
    return std::string(std::istreambuf_iterator<char>(file), std::istreambuf_iterator<char>());
}

// ---------------------------------------------------------------- parameters and constraints

KernelParameter::KernelParameter(const std::string& name, const std::vector<ParameterValue>& values, const std::string& group) :
    m_Name(name),
    m_Group(group),
    m_Values(values),
    m_Type(ParameterValueType::UnsignedInt)
{
    if (values.empty())
    {
        throw KttException("Parameter " + name + " needs at least one value");
    }
    m_Type = ParameterPair::GetTypeFromValue(values.front());
    const bool mixed = std::any_of(values.begin(), values.end(),
        [this](const ParameterValue& value) { return ParameterPair::GetTypeFromValue(value) != m_Type; });
    if (mixed)
    {
        throw KttException("The values of parameter " + name + " must all have one type");
    }
}

KernelConstraint::KernelConstraint(const std::vector<std::string>& parameters, ConstraintFunction function) :
    m_Parameters(parameters),
    m_OnIntegers(function)
{
    if (!function)
    {
        throw KttException("A constraint needs a callable function");
    }
}

KernelConstraint::KernelConstraint(const std::vector<std::string>& parameters, GenericConstraintFunction function) :
    m_Parameters(parameters),
    m_OnValues(function)
{
    if (!function)
    {
        throw KttException("A constraint needs a callable function");
    }
}

bool KernelConstraint::IsFulfilled(const std::vector<const ParameterValue*>& values) const
{
    if (m_OnValues)
    {
        return m_OnValues(values);
    }
    // Integer constraints see every numeric or boolean value as an unsigned integer.
    std::vector<uint64_t> integers;
    integers.reserve(values.size());
    for (const ParameterValue* value : values)
    {
        integers.push_back(std::visit([](const auto& held) -> uint64_t
        {
            using T = std::decay_t<decltype(held)>;
            if constexpr (std::is_same_v<T, std::string>)
            {
                throw KttException("A constraint over string values must be a generic constraint");
            }
            else
            {
                return static_cast<uint64_t>(held);
            }
        }, *value));
    }
    return m_OnIntegers(integers);
}

// ---------------------------------------------------------------- KernelDefinition

KernelDefinition::KernelDefinition(const KernelDefinitionId id, const std::string& name, const std::vector<std::string>& typeNames,
    const std::string& source, const std::string& sourceDirectory, const DimensionVector& globalSize, const DimensionVector& localSize) :
    m_Id(id),
    m_Name(name),
    m_Templated(TemplatedName(name, typeNames)),
    m_Source(source),
    m_Directory(sourceDirectory),
    m_Global(globalSize),
    m_Local(localSize)
{}

void KernelDefinition::SetArguments(const std::vector<KernelArgument*>& arguments)
{
    std::vector<const KernelArgument*> sorted(arguments.begin(), arguments.end());
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
    {
        throw KttException("Definition " + m_Templated + " binds the same argument twice; pass each argument once");
    }
    m_Arguments = arguments;
}

bool KernelDefinition::HasArgument(const ArgumentId& id) const
{
    return std::any_of(m_Arguments.begin(), m_Arguments.end(), [&id](const KernelArgument* argument) { return argument->GetId() == id; });
}

std::string KernelDefinition::TemplatedName(const std::string& name, const std::vector<std::string>& typeNames)
{
    if (typeNames.empty())
    {
        return name;
    }
    std::string text = name + "<";
    for (size_t i = 0; i < typeNames.size(); ++i)
    {
        text += (i == 0 ? "" : ", ") + typeNames[i];
    }
    return text + ">";
}

// ---------------------------------------------------------------- Kernel

Kernel::Kernel(const KernelId id, const std::string& name, const std::vector<const KernelDefinition*>& definitions) :
    m_Id(id),
    m_Name(name),
    m_Definitions(definitions)
{
    if (definitions.empty())
    {
        throw KttException("Kernel " + name + " needs at least one definition");
    }
    // Only the first definition is profiled unless the caller chooses otherwise.
    m_Profiled = {definitions.front()};
}

const KernelParameter* Kernel::FindParameter(const std::string& name) const
{
    const auto found = std::find_if(m_Parameters.begin(), m_Parameters.end(),
        [&name](const KernelParameter& parameter) { return parameter.GetName() == name; });
    return found != m_Parameters.end() ? &*found : nullptr;
}

const KernelParameter& Kernel::GetParameter(const std::string& name) const
{
    const KernelParameter* parameter = FindParameter(name);
    if (parameter == nullptr)
    {
        throw KttException("Kernel " + m_Name + " has no parameter " + name);
    }
    return *parameter;
}

void Kernel::AddParameter(const KernelParameter& parameter)
{
    if (HasParameter(parameter.GetName()))
    {
        throw KttException("Kernel " + m_Name + " already has a parameter " + parameter.GetName());
    }
    m_Parameters.push_back(parameter);
}

void Kernel::AddConstraint(const KernelConstraint& constraint)
{
    const auto& names = constraint.GetParameters();
    if (names.empty())
    {
        throw KttException("A constraint of kernel " + m_Name + " names no parameter");
    }
    const std::string& group = GetParameter(names.front()).GetGroup();
    for (const std::string& name : names)
    {
        if (GetParameter(name).GetGroup() != group)
        {
            throw KttException("The parameters of a constraint must belong to one group (kernel " + m_Name + ")");
        }
    }
    m_Constraints.push_back(constraint);
}

void Kernel::AddThreadModifier(const ModifierType type, const ModifierDimension dimension, const std::vector<std::string>& parameters,
    const std::vector<KernelDefinitionId>& definitions, ModifierFunction function)
{
    if (!function)
    {
        throw KttException("A thread modifier needs a callable function");
    }
    for (const std::string& name : parameters)
    {
        const ParameterValueType valueType = GetParameter(name).GetValueType();
        if (valueType != ParameterValueType::UnsignedInt && valueType != ParameterValueType::Int)
        {
            throw KttException("Thread modifiers take integer parameters only; " + name + " is not one");
        }
    }
    for (const KernelDefinitionId id : definitions)
    {
        if (!HasDefinition(id))
        {
            throw KttException("Kernel " + m_Name + " has no definition with id " + std::to_string(id));
        }
    }
    m_Modifiers.push_back({type, dimension, parameters, definitions, function});
}

void Kernel::SetProfiledDefinitions(const std::vector<const KernelDefinition*>& definitions)
{
    for (const KernelDefinition* definition : definitions)
    {
        if (!HasDefinition(definition->GetId()))
        {
            throw KttException("Kernel " + m_Name + " has no definition " + definition->GetTemplatedName());
        }
    }
    m_Profiled = definitions;
}

bool Kernel::HasDefinition(const KernelDefinitionId id) const
{
    return std::any_of(m_Definitions.begin(), m_Definitions.end(), [id](const KernelDefinition* definition) { return definition->GetId() == id; });
}

const KernelDefinition& Kernel::GetDefinition(const KernelDefinitionId id) const
{
    const auto found = std::find_if(m_Definitions.begin(), m_Definitions.end(),
        [id](const KernelDefinition* definition) { return definition->GetId() == id; });
    if (found == m_Definitions.end())
    {
        throw KttException("Kernel " + m_Name + " has no definition with id " + std::to_string(id));
    }
    return **found;
}

std::vector<KernelArgument*> Kernel::GetVectorArguments() const
{
    std::vector<KernelArgument*> vectors;
    for (const KernelDefinition* definition : m_Definitions)
    {
        for (KernelArgument* argument : definition->GetVectorArguments())
        {
            if (std::find(vectors.begin(), vectors.end(), argument) == vectors.end())
            {
                vectors.push_back(argument);
            }
        }
    }
    return vectors;
}

KernelConfiguration Kernel::CreateConfiguration(const ParameterInput& parameters) const
{
    std::vector<ParameterPair> pairs;
    for (const KernelParameter& parameter : m_Parameters)
    {
        const auto given = std::find_if(parameters.begin(), parameters.end(),
            [&parameter](const auto& entry) { return entry.first == parameter.GetName(); });
        if (given == parameters.end())
        {
            pairs.push_back(parameter.GeneratePair(0));
            continue;
        }
        if (ParameterPair::GetTypeFromValue(given->second) != parameter.GetValueType())
        {
            throw KttException("The value given for " + parameter.GetName() + " has the wrong type");
        }
        pairs.emplace_back(parameter.GetName(), given->second);
    }
    for (const auto& entry : parameters)
    {
        if (!HasParameter(entry.first))
        {
            throw KttException("Kernel " + m_Name + " has no parameter " + entry.first);
        }
    }
    return KernelConfiguration(pairs);
}

DimensionVector Kernel::GetModifiedSize(const KernelDefinitionId id, const DimensionVector& size, const ModifierType type,
    const std::vector<ParameterPair>& pairs) const
{
    DimensionVector modified = size;
    for (const Modifier& modifier : m_Modifiers)
    {
        const bool applies = modifier.type == type
            && std::find(modifier.definitions.begin(), modifier.definitions.end(), id) != modifier.definitions.end();
        if (!applies)
        {
            continue;
        }
        const std::vector<uint64_t> values = ParameterPair::GetParameterValues<uint64_t>(pairs, modifier.parameters);
        const uint64_t before = modified.GetSize(modifier.dimension);
        modified.SetSize(modifier.dimension, static_cast<size_t>(modifier.function(before, values)));
    }
    return modified;
}

// ---------------------------------------------------------------- KernelRegistry

KernelRegistry::KernelRegistry(ArgumentRegistry& arguments) :
    m_Arguments(arguments),
    m_NextDefinition(0),
    m_NextKernel(0)
{}

KernelDefinitionId KernelRegistry::AddDefinition(const std::string& name, const std::string& source, const DimensionVector& globalSize,
    const DimensionVector& localSize, const std::vector<std::string>& typeNames, const std::string& sourceDirectory)
{
    if (FindDefinition(name, typeNames) != InvalidKernelDefinitionId)
    {
        throw KttException("A definition " + KernelDefinition::TemplatedName(name, typeNames) + " exists already");
    }
    const KernelDefinitionId id = m_NextDefinition++;
    m_Definitions.emplace(id, std::make_unique<KernelDefinition>(id, name, typeNames, source, sourceDirectory, globalSize, localSize));
    return id;
}

KernelDefinitionId KernelRegistry::AddDefinitionFromFile(const std::string& name, const std::string& filePath,
    const DimensionVector& globalSize, const DimensionVector& localSize, const std::vector<std::string>& typeNames)
{
    const size_t slash = filePath.rfind('/');
    const std::string directory = slash == std::string::npos ? "." : filePath.substr(0, slash);
    return AddDefinition(name, ReadTextFile(filePath), globalSize, localSize, typeNames, directory);
}

KernelDefinitionId KernelRegistry::FindDefinition(const std::string& name, const std::vector<std::string>& typeNames) const
{
    const std::string wanted = KernelDefinition::TemplatedName(name, typeNames);
    for (const auto& [id, definition] : m_Definitions)
    {
        if (definition->GetTemplatedName() == wanted)
        {
            return id;
        }
    }
    return InvalidKernelDefinitionId;
}

void KernelRegistry::RemoveDefinition(const KernelDefinitionId id)
{
    const bool inUse = std::any_of(m_Kernels.begin(), m_Kernels.end(), [id](const auto& entry) { return entry.second->HasDefinition(id); });
    if (inUse)
    {
        throw KttException("Definition " + std::to_string(id) + " belongs to a kernel; remove the kernel first");
    }
    if (m_Definitions.erase(id) == 0)
    {
        throw KttException("No definition with id " + std::to_string(id));
    }
}

void KernelRegistry::Bind(const KernelDefinitionId id, const std::vector<ArgumentId>& arguments)
{
    GetDefinition(id).SetArguments(m_Arguments.Resolve(arguments));
}

std::vector<const KernelDefinition*> KernelRegistry::Definitions(const std::vector<KernelDefinitionId>& ids) const
{
    std::vector<const KernelDefinition*> definitions;
    for (const KernelDefinitionId id : ids)
    {
        definitions.push_back(&GetDefinition(id));
    }
    return definitions;
}

KernelId KernelRegistry::AddKernel(const std::string& name, const std::vector<KernelDefinitionId>& definitions)
{
    const bool taken = std::any_of(m_Kernels.begin(), m_Kernels.end(), [&name](const auto& entry) { return entry.second->GetName() == name; });
    if (taken)
    {
        throw KttException("A kernel named " + name + " exists already");
    }
    const KernelId id = m_NextKernel++;
    m_Kernels.emplace(id, std::make_unique<Kernel>(id, name, Definitions(definitions)));
    return id;
}

void KernelRegistry::RemoveKernel(const KernelId id)
{
    if (m_Kernels.erase(id) == 0)
    {
        throw KttException("No kernel with id " + std::to_string(id));
    }
}

Kernel& KernelRegistry::GetKernel(const KernelId id)
{
    return const_cast<Kernel&>(static_cast<const KernelRegistry*>(this)->GetKernel(id));
}

const Kernel& KernelRegistry::GetKernel(const KernelId id) const
{
    const auto found = m_Kernels.find(id);
    if (found == m_Kernels.end())
    {
        throw KttException("No kernel with id " + std::to_string(id));
    }
    return *found->second;
}

KernelDefinition& KernelRegistry::GetDefinition(const KernelDefinitionId id)
{
    return const_cast<KernelDefinition&>(static_cast<const KernelRegistry*>(this)->GetDefinition(id));
}

const KernelDefinition& KernelRegistry::GetDefinition(const KernelDefinitionId id) const
{
    const auto found = m_Definitions.find(id);
    if (found == m_Definitions.end())
    {
        throw KttException("No definition with id " + std::to_string(id));
    }
    return *found->second;
}

bool KernelRegistry::UsesArgument(const ArgumentId& id) const
{
    return std::any_of(m_Definitions.begin(), m_Definitions.end(), [&id](const auto& entry) { return entry.second->HasArgument(id); });
}

} // namespace ktt
