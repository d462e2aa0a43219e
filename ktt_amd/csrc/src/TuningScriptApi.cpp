// C entry points for KTT JSON tuning scripts (include/ktt_hip.h section 4).
#include <cstdio>

#include <KttException.h>

#include "../../../include/ktt_hip.h"
#include "CoulombRuntime.h"
#include "ScriptExpression.h"
#include "TuningScript.h"

using namespace ktt;

extern "C" {

int ktt_run_tuning_script(const char* scriptPath, const char* outputFile, long long configurationLimit, char* bestConfig,
    size_t bestConfigLength, long long* spaceSize, long long* tested, long long* valid, double* bestMs)
{
    try
    {
        if (scriptPath == nullptr)
        {
            return SetLastError("no script path given");
        }
        TuningScriptOptions options;
        options.outputFile = outputFile != nullptr ? outputFile : "";
        options.configurationLimit = configurationLimit;
        const TuningScriptSummary summary = RunTuningScript(scriptPath, options);
        if (bestConfig != nullptr && bestConfigLength > 0)
        {
            std::string text;
            for (const auto& pair : summary.best.GetPairs())
            {
                text += (text.empty() ? "" : ",") + pair.GetName() + "=" + pair.GetValueString();
            }
            std::snprintf(bestConfig, bestConfigLength, "%s", text.c_str());
        }
        if (spaceSize != nullptr)
        {
            *spaceSize = static_cast<long long>(summary.spaceSize);
        }
        if (tested != nullptr)
        {
            *tested = static_cast<long long>(summary.tested);
        }
        if (valid != nullptr)
        {
            *valid = static_cast<long long>(summary.valid);
        }
        if (bestMs != nullptr)
        {
            *bestMs = summary.bestKernelMs;
        }
        return summary.tested > 0 ? 0 : SetLastError("the script tested no configuration");
    }
    catch (const std::exception& failure)
    {
        return SetLastError(failure.what());
    }
}

int ktt_tuning_script_space(const char* scriptPath, long long* size)
{
    try
    {
        if (scriptPath == nullptr || size == nullptr)
        {
            return SetLastError("script path and size pointer are required");
        }
        *size = static_cast<long long>(CountScriptConfigurations(scriptPath));
        return 0;
    }
    catch (const std::exception& failure)
    {
        return SetLastError(failure.what());
    }
}

int ktt_script_generate(const char* expression, long long count, double* values)
{
    try
    {
        if (expression == nullptr || count < 0 || (count > 0 && values == nullptr))
        {
            return SetLastError("an expression, a non-negative count and an output array are required");
        }
        const ScriptExpression generator(expression);
        for (long long i = 0; i < count; ++i)
        {
            values[i] = generator.Evaluate([i](const std::string& name) -> ScriptValue
            {
                if (name != "i")
                {
                    throw KttException("A generator may only use i, not " + name);
                }
                return ScriptValue::Integer(i);
            }).ToDouble();
        }
        return 0;
    }
    catch (const std::exception& failure)
    {
        return SetLastError(failure.what());
    }
}

} // extern "C"
