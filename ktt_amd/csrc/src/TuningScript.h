// KTT JSON tuning scripts on the HIP backend (the format of the reference's TuningLoader,
// TuningLoader/TuningSchema.h; example Examples/JSON/CoulombSum3d.json):
//   ConfigurationSpace  TuningParameters {Name, Type, Values}, Conditions {Parameters, Expression}
//   Search              {Name: Deterministic | Random | MCMC}
//   Budget              [{Type: ConfigurationCount | ConfigurationFraction | TuningDuration, BudgetValue}]
//   General             {LoggingLevel, TimeUnit, OutputFile, OutputFormat}
//   KernelSpecification {Language: HIP, KernelName, KernelFile, KernelTypeNames, CompilerOptions,
//                        Profiling, GlobalSizeType, GlobalSize {X,Y,Z}, LocalSize {X,Y,Z},
//                        Arguments [...], ReferenceArguments [...], SimulationInput, Device}
// Expressions (Values, Conditions, sizes, Generator data) are evaluated by ScriptExpression,
// not by an embedded Python interpreter. Relative file names are relative to the script.
#pragma once

#include <cstdint>
#include <string>

#include <KernelConfiguration.h>

namespace ktt
{

struct TuningScriptOptions
{
    std::string outputFile;          // overrides General.OutputFile when not empty
    int64_t configurationLimit = 0;  // > 0: a ConfigurationCount budget replacing the script's
};

struct TuningScriptSummary
{
    std::string outputFile;          // results file written (with .json)
    uint64_t spaceSize = 0;
    uint64_t tested = 0;
    uint64_t valid = 0;
    KernelConfiguration best;
    double bestKernelMs = -1.0;
};

/** Runs the script's tuning (or its simulation when SimulationInput is given). */
TuningScriptSummary RunTuningScript(const std::string& scriptPath, const TuningScriptOptions& options);

/** The size of the script's configuration space; needs no device. */
uint64_t CountScriptConfigurations(const std::string& scriptPath);

} // namespace ktt
