/** @file KttTypes.h
  * Identifiers, callback types and enumerations of the ktt public API.
  * Names, values and meanings follow Source/KttTypes.h:25-160 and the enum headers
  * of the reference (ComputeEngine/ComputeApi.h:12-25, Kernel/Modifier*.h,
  * KernelArgument/Argument*.h, KernelRunner/Validation*.h, Api/Output/ResultStatus.h:12-33);
  * ComputeApi gains HIP.
  */
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <limits>
#include <map>
#include <string>
#include <utility>
#include <variant>
#include <vector>

#include <DimensionVector.h>

namespace ktt
{

class ComputeInterface;

using PlatformIndex = uint32_t;
using DeviceIndex = uint32_t;
using QueueId = uint32_t;
using KernelDefinitionId = uint64_t;
using KernelId = uint64_t;
using ArgumentId = std::string;
using ParameterValue = std::variant<int64_t, uint64_t, double, bool, std::string>;
using KernelDimensions = std::map<KernelDefinitionId, std::pair<DimensionVector, DimensionVector>>;
using ParameterInput = std::vector<std::pair<std::string, ParameterValue>>;
using UserData = std::map<std::string, std::string>;
using Nanoseconds = uint64_t;
using KernelComputeId = std::string;
using ComputeActionId = uint64_t;
using TransferActionId = uint64_t;

using ModifierFunction = std::function<uint64_t(const uint64_t /*defaultSize*/, const std::vector<uint64_t>& /*parameterValues*/)>;
using ConstraintFunction = std::function<bool(const std::vector<uint64_t>& /*parameterValues*/)>;
using GenericConstraintFunction = std::function<bool(const std::vector<const ParameterValue*>& /*parameterValues*/)>;
using KernelLauncher = std::function<void(ComputeInterface& /*interface*/)>;
using ReferenceComputation = std::function<void(void* /*buffer*/)>;
using ValueComparator = std::function<bool(const void* /*result*/, const void* /*reference*/)>;

using UnifiedBufferMemory = void*;
using ComputeContext = void*;
using ComputeQueue = void*;
using ComputeBuffer = void*;

inline const QueueId InvalidQueueId = std::numeric_limits<QueueId>::max();
inline const KernelDefinitionId InvalidKernelDefinitionId = std::numeric_limits<KernelDefinitionId>::max();
inline const KernelId InvalidKernelId = std::numeric_limits<KernelId>::max();
inline const ArgumentId InvalidArgumentId = ArgumentId();
inline const Nanoseconds InvalidDuration = std::numeric_limits<Nanoseconds>::max();

/** Compute API used by a tuner. HIP is the backend this library implements. */
enum class ComputeApi
{
    OpenCL,
    CUDA,
    Vulkan,
    HIP
};

/** How global sizes are interpreted: OpenCL = total threads, CUDA = number of workgroups. */
enum class GlobalSizeType
{
    OpenCL,
    CUDA,
    Vulkan
};

enum class ModifierType
{
    Global,
    Local
};

enum class ModifierDimension
{
    X,
    Y,
    Z
};

enum class ModifierAction
{
    Add,
    Subtract,
    Multiply,
    Divide,
    DivideCeil
};

enum class ArgumentAccessType
{
    Undefined,
    ReadOnly,
    WriteOnly,
    ReadWrite
};

enum class ArgumentDataType
{
    Char,
    UnsignedChar,
    Short,
    UnsignedShort,
    Int,
    UnsignedInt,
    Long,
    UnsignedLong,
    Half,
    Float,
    Double,
    Custom
};

enum class ArgumentMemoryLocation
{
    Undefined,
    Device,
    Host,
    HostZeroCopy,
    Unified
};

enum class ArgumentMemoryType
{
    Scalar,
    Vector,
    Local,
    Symbol
};

enum class ArgumentManagementType
{
    Framework,
    User
};

enum class ArgumentOwnership
{
    Copy,
    Reference,
    User
};

enum class ValidationMethod
{
    AbsoluteDifference,
    SideBySideComparison,
    SideBySideRelativeComparison
};

/** Bit flags selecting which run modes are validated. */
enum class ValidationMode
{
    None = 0,
    Running = (1 << 0),
    OfflineTuning = (1 << 1),
    OnlineTuning = (1 << 2),
    All = Running | OfflineTuning | OnlineTuning
};

inline ValidationMode operator|(const ValidationMode a, const ValidationMode b)
{
    return static_cast<ValidationMode>(static_cast<int>(a) | static_cast<int>(b));
}

inline bool HasFlag(const ValidationMode value, const ValidationMode flag)
{
    return (static_cast<int>(value) & static_cast<int>(flag)) != 0;
}

enum class KernelRunMode
{
    Running,
    OfflineTuning,
    OnlineTuning,
    ResultValidation
};

enum class TimeUnit
{
    Nanoseconds,
    Microseconds,
    Milliseconds,
    Seconds
};

/** Result file formats. XML is recognised but not produced by this library. */
enum class OutputFormat
{
    JSON,
    XML
};

enum class LoggingLevel
{
    Off,
    Error,
    Warning,
    Info,
    Debug
};

enum class ResultStatus
{
    Ok,
    ComputationFailed,
    ValidationFailed,
    CompilationFailed,
    DeviceLimitsExceeded
};

enum class ExceptionReason
{
    General,
    CompilerError,
    DeviceLimitsExceeded
};

enum class ParameterValueType
{
    Int,
    UnsignedInt,
    Double,
    Float,
    Bool,
    String
};

enum class ProfilingCounterType
{
    Int,
    UnsignedInt,
    Double,
    Percent,
    Throughput,
    UtilizationLevel
};

enum class DeviceType
{
    CPU,
    GPU,
    Custom
};

} // namespace ktt
