#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>

#include <Ktt.h>

#include "Common.h"
#include "Exploration.h"
#include "Json.h"
#include "Kernel.h"
#include "ScriptExpression.h"
#include "TunerCore.h"
#include "TuningScript.h"

namespace ktt
{

namespace
{

std::string DirectoryOf(const std::string& path)
{
    const size_t slash = path.rfind('/');
    return slash == std::string::npos ? "." : path.substr(0, slash);
}

std::string Resolve(const std::string& directory, const std::string& file)
{
    return !file.empty() && file[0] == '/' ? file : directory + "/" + file;
}

std::string Text(const JsonValue& object, const std::string& key, const std::string& fallback)
{
    return object.Contains(key) ? object.At(key).AsString() : fallback;
}

ScriptValue FromParameter(const ParameterValue& value)
{
    return std::visit([](const auto& held) -> ScriptValue
    {
        using T = std::decay_t<decltype(held)>;
        if constexpr (std::is_same_v<T, bool>)
        {
            return ScriptValue::Boolean(held);
        }
        else if constexpr (std::is_same_v<T, double>)
        {
            return ScriptValue::Real(held);
        }
        else if constexpr (std::is_same_v<T, std::string>)
        {
            throw KttException("String parameters cannot appear in script expressions");
        }
        else
        {
            return ScriptValue::Integer(static_cast<int64_t>(held));
        }
    }, value);
}

/** Where the configuration space goes: a Tuner (for running) or a bare kernel (for counting). */
class SpaceSink
{
public:
    virtual ~SpaceSink() = default;
    virtual void Parameter(const std::string& name, const std::vector<ParameterValue>& values) = 0;
    virtual void Constraint(const std::vector<std::string>& names, GenericConstraintFunction function) = 0;
};

class KernelSink : public SpaceSink
{
public:
    explicit KernelSink(Kernel& kernel) : m_Kernel(kernel) {}
    void Parameter(const std::string& name, const std::vector<ParameterValue>& values) override
    {
        m_Kernel.AddParameter(KernelParameter(name, values, ""));
    }
    void Constraint(const std::vector<std::string>& names, GenericConstraintFunction function) override
    {
        m_Kernel.AddConstraint(KernelConstraint(names, function));
    }

private:
    Kernel& m_Kernel;
};

class TunerSink : public SpaceSink
{
public:
    TunerSink(Tuner& tuner, const KernelId kernel) : m_Tuner(tuner), m_Kernel(kernel) {}
    void Parameter(const std::string& name, const std::vector<ParameterValue>& values) override
    {
        // Tuner::AddParameter is typed; every value of a parameter shares one alternative.
        std::visit([&](const auto& first)
        {
            using T = std::decay_t<decltype(first)>;
            std::vector<T> typed;
            for (const ParameterValue& value : values)
            {
                typed.push_back(std::get<T>(value));
            }
            m_Tuner.AddParameter(m_Kernel, name, typed);
        }, values.front());
    }
    void Constraint(const std::vector<std::string>& names, GenericConstraintFunction function) override
    {
        m_Tuner.AddGenericConstraint(m_Kernel, names, function);
    }

private:
    Tuner& m_Tuner;
    KernelId m_Kernel;
};

void AddSpace(const JsonValue& script, SpaceSink& sink)
{
    const JsonValue& space = script.At("ConfigurationSpace");
    const JsonValue& parameters = space.At("TuningParameters");
    for (size_t p = 0; p < parameters.Size(); ++p)
    {
        const JsonValue& entry = parameters[p];
        const std::string name = entry.At("Name").AsString();
        const std::string type = Text(entry, "Type", "uint");
        const ScriptValue list = ScriptExpression(entry.At("Values").AsString()).Evaluate(
            [&name](const std::string& unknown) -> ScriptValue
            {
                throw KttException("The values of " + name + " refer to " + unknown);
            });
        std::vector<ParameterValue> values;
        for (const ScriptValue& item : list.Items())
        {
            if (type == "uint")
            {
                values.emplace_back(static_cast<uint64_t>(item.ToInt()));
            }
            else if (type == "int")
            {
                values.emplace_back(static_cast<int64_t>(item.ToInt()));
            }
            else if (type == "float")
            {
                values.emplace_back(item.ToDouble());
            }
            else if (type == "bool")
            {
                values.emplace_back(item.Truthy());
            }
            else
            {
                throw KttException("Parameter " + name + " has unsupported type " + type);
            }
        }
        if (values.empty())
        {
            throw KttException("Parameter " + name + " has no values");
        }
        sink.Parameter(name, values);
    }
    if (!space.Contains("Conditions"))
    {
        return;
    }
    const JsonValue& conditions = space.At("Conditions");
    for (size_t c = 0; c < conditions.Size(); ++c)
    {
        const JsonValue& entry = conditions[c];
        std::vector<std::string> names;
        for (size_t k = 0; k < entry.At("Parameters").Size(); ++k)
        {
            names.push_back(entry.At("Parameters")[k].AsString());
        }
        auto expression = std::make_shared<ScriptExpression>(entry.At("Expression").AsString());
        sink.Constraint(names, [expression, names](const std::vector<const ParameterValue*>& values)
        {
            return expression->Evaluate([&](const std::string& name) -> ScriptValue
            {
                const auto found = std::find(names.begin(), names.end(), name);
                if (found == names.end())
                {
                    throw KttException("Condition \"" + expression->Text() + "\" uses " + name + ", which is not among its parameters");
                }
                return FromParameter(*values[static_cast<size_t>(found - names.begin())]);
            }).Truthy();
        });
    }
}

/** Element data of a vector argument of element type T. */
template <typename T>
std::vector<T> Fill(const JsonValue& entry, const size_t count, const std::string& directory)
{
    std::vector<T> data(count);
    const std::string fill = Text(entry, "FillType", "Constant");
    if (fill == "Constant")
    {
        const double value = entry.Contains("FillValue") ? entry.At("FillValue").AsDouble() : 0.0;
        std::fill(data.begin(), data.end(), static_cast<T>(value));
    }
    else if (fill == "Random")
    {
        // Uniform in [0, FillValue), seeded by RandomSeed when given.
        const double upper = entry.Contains("FillValue") ? entry.At("FillValue").AsDouble() : 1.0;
        std::default_random_engine engine(entry.Contains("RandomSeed") ? static_cast<unsigned>(entry.At("RandomSeed").AsDouble())
                                                                      : static_cast<unsigned>(DefaultSearcherSeed()));
        std::uniform_real_distribution<float> distribution(0.0f, static_cast<float>(upper));
        for (T& element : data)
        {
            element = static_cast<T>(distribution(engine));
        }
    }
    else if (fill == "Generator")
    {
        const ScriptExpression generator(entry.At("DataSource").AsString());
        for (size_t i = 0; i < count; ++i)
        {
            const ScriptValue value = generator.Evaluate([i](const std::string& name) -> ScriptValue
            {
                if (name != "i")
                {
                    throw KttException("A generator may only use i, not " + name);
                }
                return ScriptValue::Integer(static_cast<int64_t>(i));
            });
            data[i] = static_cast<T>(value.ToDouble());
        }
    }
    else if (fill == "BinaryRaw")
    {
        const std::string bytes = ReadTextFile(Resolve(directory, entry.At("DataSource").AsString()));
        if (bytes.size() != count * sizeof(T))
        {
            throw KttException("File " + entry.At("DataSource").AsString() + " holds " + std::to_string(bytes.size()) + " bytes, expected "
                + std::to_string(count * sizeof(T)));
        }
        std::memcpy(data.data(), bytes.data(), bytes.size());
    }
    else
    {
        throw KttException("Fill type " + fill + " is not supported");
    }
    return data;
}

struct ElementType
{
    ArgumentDataType type;
    size_t size;
};

ElementType ElementOf(const std::string& name, const JsonValue& entry)
{
    if (name == "float") return {ArgumentDataType::Float, 4};
    if (name == "double") return {ArgumentDataType::Double, 8};
    if (name == "int8") return {ArgumentDataType::Char, 1};
    if (name == "uint8") return {ArgumentDataType::UnsignedChar, 1};
    if (name == "int16") return {ArgumentDataType::Short, 2};
    if (name == "uint16") return {ArgumentDataType::UnsignedShort, 2};
    if (name == "int32") return {ArgumentDataType::Int, 4};
    if (name == "uint32") return {ArgumentDataType::UnsignedInt, 4};
    if (name == "int64") return {ArgumentDataType::Long, 8};
    if (name == "uint64") return {ArgumentDataType::UnsignedLong, 8};
    static_cast<void>(entry);
    throw KttException("Argument type " + name + " is not supported");
}

template <typename T>
ArgumentId AddVector(Tuner& tuner, const JsonValue& entry, const size_t count, const std::string& directory,
    const ArgumentAccessType access, const ArgumentId& id)
{
    const std::vector<T> data = Fill<T>(entry, count, directory);
    return tuner.AddArgumentVector(data, access, id);
}

ArgumentId AddTypedVector(Tuner& tuner, const ArgumentDataType type, const JsonValue& entry, const size_t count,
    const std::string& directory, const ArgumentAccessType access, const ArgumentId& id)
{
    switch (type)
    {
    case ArgumentDataType::Float: return AddVector<float>(tuner, entry, count, directory, access, id);
    case ArgumentDataType::Double: return AddVector<double>(tuner, entry, count, directory, access, id);
    case ArgumentDataType::Char: return AddVector<int8_t>(tuner, entry, count, directory, access, id);
    case ArgumentDataType::UnsignedChar: return AddVector<uint8_t>(tuner, entry, count, directory, access, id);
    case ArgumentDataType::Short: return AddVector<int16_t>(tuner, entry, count, directory, access, id);
    case ArgumentDataType::UnsignedShort: return AddVector<uint16_t>(tuner, entry, count, directory, access, id);
    case ArgumentDataType::Int: return AddVector<int32_t>(tuner, entry, count, directory, access, id);
    case ArgumentDataType::UnsignedInt: return AddVector<uint32_t>(tuner, entry, count, directory, access, id);
    case ArgumentDataType::Long: return AddVector<int64_t>(tuner, entry, count, directory, access, id);
    case ArgumentDataType::UnsignedLong: return AddVector<uint64_t>(tuner, entry, count, directory, access, id);
    default:
        throw KttException("Vector argument " + id + " has an unsupported element type");
    }
}

ArgumentAccessType AccessOf(const std::string& name)
{
    if (name == "ReadOnly") return ArgumentAccessType::ReadOnly;
    if (name == "WriteOnly") return ArgumentAccessType::WriteOnly;
    if (name == "ReadWrite") return ArgumentAccessType::ReadWrite;
    throw KttException("Unknown access type " + name);
}

/** Registers one argument (or reference argument) entry with the tuner; returns its id. */
ArgumentId AddScriptArgument(Tuner& tuner, const JsonValue& entry, const std::string& directory, const ArgumentId& id,
    const ElementType* inherit = nullptr, const size_t inheritCount = 0)
{
    const std::string kind = Text(entry, "MemoryType", "Vector");
    const ElementType element = inherit != nullptr ? *inherit : ElementOf(Text(entry, "Type", "float"), entry);
    if (kind == "Scalar")
    {
        const double value = entry.Contains("FillValue") ? entry.At("FillValue").AsDouble() : 0.0;
        switch (element.type)
        {
        case ArgumentDataType::Float: return tuner.AddArgumentScalar(static_cast<float>(value), id);
        case ArgumentDataType::Double: return tuner.AddArgumentScalar(value, id);
        case ArgumentDataType::Int: return tuner.AddArgumentScalar(static_cast<int32_t>(value), id);
        case ArgumentDataType::UnsignedInt: return tuner.AddArgumentScalar(static_cast<uint32_t>(value), id);
        case ArgumentDataType::Long: return tuner.AddArgumentScalar(static_cast<int64_t>(value), id);
        case ArgumentDataType::UnsignedLong: return tuner.AddArgumentScalar(static_cast<uint64_t>(value), id);
        default: throw KttException("Scalar argument " + id + " has an unsupported type");
        }
    }
    const size_t count = inherit != nullptr ? inheritCount : static_cast<size_t>(entry.At("Size").AsUint());
    if (kind == "Local")
    {
        return tuner.AddArgumentLocal<uint8_t>(count * element.size, id);
    }
    if (kind != "Vector")
    {
        throw KttException("Memory type " + kind + " is not supported");
    }
    return AddTypedVector(tuner, element.type, entry, count, directory, AccessOf(Text(entry, "AccessType", "ReadOnly")), id);
}

} // namespace

uint64_t CountScriptConfigurations(const std::string& scriptPath)
{
    const JsonValue script = JsonValue::Parse(ReadTextFile(scriptPath));
    ArgumentRegistry arguments;
    KernelRegistry kernels(arguments);
    const KernelDefinitionId definition = kernels.AddDefinition("script", "", DimensionVector(), DimensionVector(), {});
    Kernel& kernel = kernels.GetKernel(kernels.AddKernel("script", {definition}));
    KernelSink sink(kernel);
    AddSpace(script, sink);
    DeterministicSearcher searcher;
    return Exploration(kernel, searcher).Size();
}

TuningScriptSummary RunTuningScript(const std::string& scriptPath, const TuningScriptOptions& options)
{
    const JsonValue script = JsonValue::Parse(ReadTextFile(scriptPath));
    const std::string directory = DirectoryOf(scriptPath);
    const JsonValue& general = script.At("General");
    const JsonValue& specification = script.At("KernelSpecification");

    const std::string language = specification.At("Language").AsString();
    if (language != "HIP")
    {
        throw KttException("This build runs scripts whose Language is HIP, not " + language);
    }
    PlatformIndex platform = 0;
    DeviceIndex device = 0;
    if (specification.Contains("Device"))
    {
        const JsonValue& chosen = specification.At("Device");
        platform = chosen.Contains("PlatformId") ? static_cast<PlatformIndex>(chosen.At("PlatformId").AsUint()) : 0;
        device = chosen.Contains("DeviceId") ? static_cast<DeviceIndex>(chosen.At("DeviceId").AsUint()) : 0;
    }
    if (general.Contains("LoggingLevel"))
    {
        const std::string level = general.At("LoggingLevel").AsString();
        const LoggingLevel value = level == "Off" ? LoggingLevel::Off : level == "Error" ? LoggingLevel::Error
            : level == "Warning" ? LoggingLevel::Warning : level == "Debug" ? LoggingLevel::Debug : LoggingLevel::Info;
        Tuner::SetLoggingLevel(value);
    }
    if (general.Contains("TimeUnit"))
    {
        Tuner::SetTimeUnit(TimeUnitFromString(general.At("TimeUnit").AsString()));
    }
    if (Text(general, "OutputFormat", "JSON") != "JSON")
    {
        Logger::LogWarning("Only JSON output is written; OutputFormat " + general.At("OutputFormat").AsString() + " is ignored");
    }

    Tuner tuner(platform, device, ComputeApi::HIP);
    tuner.SetGlobalSizeType(GlobalSizeTypeFromString(Text(specification, "GlobalSizeType", "CUDA")));
    if (specification.Contains("CompilerOptions"))
    {
        std::string joined;
        const JsonValue& list = specification.At("CompilerOptions");
        for (size_t k = 0; k < list.Size(); ++k)
        {
            joined += (k == 0 ? "" : " ") + list[k].AsString();
        }
        tuner.SetCompilerOptions(joined);
    }
    if (specification.Contains("Profiling") && specification.At("Profiling").AsBool())
    {
        tuner.SetProfiling(true);
    }
    if (specification.Contains("SharedMemory") && specification.At("SharedMemory").AsUint() != 0)
    {
        throw KttException("Dynamic shared memory (SharedMemory) is not supported; declare LDS in the kernel");
    }

    std::vector<std::string> typeNames;
    if (specification.Contains("KernelTypeNames"))
    {
        for (size_t k = 0; k < specification.At("KernelTypeNames").Size(); ++k)
        {
            typeNames.push_back(specification.At("KernelTypeNames")[k].AsString());
        }
    }
    const std::string kernelName = specification.At("KernelName").AsString();
    const KernelDefinitionId definition = tuner.AddKernelDefinitionFromFile(kernelName,
        Resolve(directory, specification.At("KernelFile").AsString()), DimensionVector(1, 1, 1), DimensionVector(1, 1, 1), typeNames);
    if (definition == InvalidKernelDefinitionId)
    {
        throw KttException("Cannot load kernel " + kernelName + " from " + specification.At("KernelFile").AsString());
    }
    const KernelId kernel = tuner.CreateSimpleKernel(kernelName, definition);

    TunerSink sink(tuner, kernel);
    AddSpace(script, sink);
    std::vector<std::string> parameterNames;
    const JsonValue& parameters = script.At("ConfigurationSpace").At("TuningParameters");
    for (size_t p = 0; p < parameters.Size(); ++p)
    {
        parameterNames.push_back(parameters[p].At("Name").AsString());
    }

    // Launch sizes: each dimension's expression becomes a thread modifier over the parameters it reads.
    for (const auto& [key, type] : {std::make_pair(std::string("GlobalSize"), ModifierType::Global),
             std::make_pair(std::string("LocalSize"), ModifierType::Local)})
    {
        const JsonValue& sizes = specification.At(key);
        for (const auto& [axis, dimension] : {std::make_pair(std::string("X"), ModifierDimension::X),
                 std::make_pair(std::string("Y"), ModifierDimension::Y), std::make_pair(std::string("Z"), ModifierDimension::Z)})
        {
            if (!sizes.Contains(axis))
            {
                continue;
            }
            auto expression = std::make_shared<ScriptExpression>(sizes.At(axis).IsString() ? sizes.At(axis).AsString()
                                                                                           : std::to_string(sizes.At(axis).AsUint()));
            std::vector<std::string> used;
            for (const std::string& name : expression->Names())
            {
                if (std::find(parameterNames.begin(), parameterNames.end(), name) == parameterNames.end())
                {
                    throw KttException(key + "." + axis + " uses " + name + ", which is not a tuning parameter");
                }
                used.push_back(name);
            }
            tuner.AddThreadModifier(kernel, {definition}, type, dimension, used,
                [expression, used](const uint64_t, const std::vector<uint64_t>& values)
                {
                    const int64_t size = expression->Evaluate([&](const std::string& name)
                    {
                        const size_t position = static_cast<size_t>(std::find(used.begin(), used.end(), name) - used.begin());
                        return ScriptValue::Integer(static_cast<int64_t>(values[position]));
                    }).ToInt();
                    if (size <= 0)
                    {
                        throw KttException("Launch size \"" + expression->Text() + "\" is not positive");
                    }
                    return static_cast<uint64_t>(size);
                });
        }
    }

    std::vector<ArgumentId> argumentIds;
    std::map<std::string, std::pair<ElementType, size_t>> shapes;
    const JsonValue& arguments = specification.At("Arguments");
    for (size_t a = 0; a < arguments.Size(); ++a)
    {
        const JsonValue& entry = arguments[a];
        const std::string id = Text(entry, "Name", "argument" + std::to_string(a));
        const ArgumentId added = AddScriptArgument(tuner, entry, directory, id);
        if (added == InvalidArgumentId)
        {
            throw KttException("Cannot add argument " + id);
        }
        argumentIds.push_back(added);
        if (Text(entry, "MemoryType", "Vector") == "Vector")
        {
            shapes[id] = {ElementOf(Text(entry, "Type", "float"), entry), static_cast<size_t>(entry.At("Size").AsUint())};
        }
    }
    tuner.SetArguments(definition, argumentIds);

    if (specification.Contains("ReferenceArguments"))
    {
        const JsonValue& references = specification.At("ReferenceArguments");
        for (size_t r = 0; r < references.Size(); ++r)
        {
            const JsonValue& entry = references[r];
            const std::string target = entry.At("TargetName").AsString();
            const auto shape = shapes.find(target);
            if (shape == shapes.end())
            {
                throw KttException("Reference target " + target + " is not a vector argument");
            }
            const std::string id = Text(entry, "Name", target + "_reference");
            if (AddScriptArgument(tuner, entry, directory, id, &shape->second.first, shape->second.second) == InvalidArgumentId)
            {
                throw KttException("Cannot add reference argument " + id);
            }
            tuner.SetReferenceArgument(target, id);
            const std::string method = Text(entry, "ValidationMethod", "SideBySideComparison");
            const double threshold = entry.Contains("ValidationThreshold") ? entry.At("ValidationThreshold").AsDouble() : 1e-4;
            tuner.SetValidationMethod(method == "AbsoluteDifference" ? ValidationMethod::AbsoluteDifference
                    : method == "SideBySideRelativeComparison" ? ValidationMethod::SideBySideRelativeComparison
                                                               : ValidationMethod::SideBySideComparison,
                threshold);
        }
    }

    const std::string searcher = script.Contains("Search") ? Text(script.At("Search"), "Name", "Deterministic") : "Deterministic";
    if (searcher == "Random")
    {
        tuner.SetSearcher(kernel, std::make_unique<RandomSearcher>());
    }
    else if (searcher == "MCMC")
    {
        tuner.SetSearcher(kernel, std::make_unique<McmcSearcher>());
    }
    else if (searcher == "ProfileBased")
    {
        // Attributes as in TuningLoader/Commands/SearcherCommand.cpp:27-41; modelPath names KTT
        // JSON results with counters (relative to the script).
        std::map<std::string, std::string> attributes;
        const JsonValue& search = script.At("Search");
        if (search.Contains("Attributes"))
        {
            const JsonValue& list = search.At("Attributes");
            for (size_t k = 0; k < list.Size(); ++k)
            {
                attributes[list[k].At("Name").AsString()] = list[k].At("Value").AsString();
            }
        }
        auto number = [&](const char* name, const uint32_t fallback)
        {
            return attributes.count(name) != 0 ? static_cast<uint32_t>(std::stoul(attributes[name])) : fallback;
        };
        const std::string model = attributes.count("modelPath") != 0 && !attributes["modelPath"].empty()
            ? Resolve(directory, attributes["modelPath"]) : std::string();
        tuner.SetProfileBasedSearcher(kernel, model, true, number("batchSize", 5), number("neighborSize", 100), number("randomSize", 10));
    }
    else if (searcher != "Deterministic")
    {
        throw KttException("Searcher " + searcher + " is not available in this build");
    }

    std::vector<std::shared_ptr<StopCondition>> budget;
    if (options.configurationLimit > 0)
    {
        budget.push_back(std::make_shared<ConfigurationCount>(static_cast<uint64_t>(options.configurationLimit)));
    }
    else if (script.Contains("Budget"))
    {
        const JsonValue& list = script.At("Budget");
        for (size_t b = 0; b < list.Size(); ++b)
        {
            const std::string type = list[b].At("Type").AsString();
            const double value = list[b].At("BudgetValue").AsDouble();
            if (type == "ConfigurationCount")
            {
                budget.push_back(std::make_shared<ConfigurationCount>(static_cast<uint64_t>(value)));
            }
            else if (type == "ConfigurationFraction")
            {
                budget.push_back(std::make_shared<ConfigurationFraction>(value));
            }
            else if (type == "TuningDuration")
            {
                budget.push_back(std::make_shared<TuningDuration>(value));
            }
            else
            {
                throw KttException("Budget type " + type + " is not supported");
            }
        }
    }
    std::unique_ptr<StopCondition> stop;
    if (!budget.empty())
    {
        stop = std::make_unique<UnionCondition>(budget);
    }

    std::vector<KernelResult> results;
    if (specification.Contains("SimulationInput"))
    {
        std::string input = Resolve(directory, specification.At("SimulationInput").AsString());
        if (input.size() > 5 && input.compare(input.size() - 5, 5, ".json") == 0)
        {
            input.resize(input.size() - 5);
        }
        results = tuner.SimulateTuning(kernel, tuner.LoadResults(input, OutputFormat::JSON), std::move(stop));
    }
    else
    {
        results = tuner.Tune(kernel, std::move(stop));
    }

    TuningScriptSummary summary;
    summary.spaceSize = tuner.GetConfigurationsCount(kernel);
    summary.tested = results.size();
    const std::string output = !options.outputFile.empty() ? options.outputFile
                                                           : Resolve(directory, Text(general, "OutputFile", "TuningOutput"));
    if (!results.empty())
    {
        tuner.SaveResults(results, output, OutputFormat::JSON, {{"script", scriptPath}});
        summary.outputFile = output + ".json";
    }
    const KernelResult* best = nullptr;
    for (const KernelResult& result : results)
    {
        if (result.IsValid())
        {
            ++summary.valid;
            if (best == nullptr || result.GetTotalDuration() < best->GetTotalDuration())
            {
                best = &result;
            }
        }
    }
    if (best != nullptr)
    {
        summary.best = best->GetConfiguration();
        summary.bestKernelMs = static_cast<double>(best->GetKernelDuration()) / 1.0e6;
    }
    return summary;
}

} // namespace ktt
