// Hardware performance counters of AMD GPUs for the HIP compute engine, collected in-process per
// kernel dispatch with rocprofiler-sdk's dispatch counting service. This is the MI355X
// counterpart of the reference's CUPTI metric interface
// (Source/ComputeEngine/Cuda/Cupti/CuptiMetricInterface.cpp:25-657): counters are named by the
// user, split into passes that the hardware can collect at once, collected over repeated runs of
// the same launch and reported as one set.
//
// Built as its own library (libktt_hip_counters.so). The compute library loads it only when
// counters are enabled, because rocprofiler-sdk has to be registered before the HIP runtime
// initialises and then intercepts every dispatch of the process.
//
// Pass planning. A derived counter (an expression over other counters) needs the hardware
// counters its expression names, recursively. Each hardware counter belongs to the block named
// by its prefix (SQ_WAVES -> SQ). A pass holds counters whose hardware counters fit the blocks'
// limits: at most 8 SQ, 4 TCC, 4 TCP, 2 TA, 2 TD and 2 GRBM counters. Counters that need another
// block are rejected, so a plan never asks the hardware for more than it can collect.
#include <algorithm>
#include <atomic>
#include <cctype>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include <rocprofiler-sdk/registration.h>
#include <rocprofiler-sdk/rocprofiler.h>

namespace
{

struct CounterInfo
{
    rocprofiler_counter_id_t id;
    std::string block;       // empty for derived counters
    std::string expression;  // empty for hardware counters
    bool derived = false;
    bool constant = false;
};

struct Pass
{
    std::vector<size_t> members;  // positions in the plan's counter list
    rocprofiler_counter_config_id_t config{};
    std::map<std::string, size_t> usage;  // block -> hardware counters used
};

struct State
{
    std::mutex lock;
    std::condition_variable arrived;

    std::atomic<bool> configured{false};
    std::atomic<bool> running{false};
    rocprofiler_context_id_t context{};

    bool haveAgent = false;
    rocprofiler_agent_id_t agent{};
    std::map<std::string, CounterInfo> counters;  // every counter the agent supports
    std::map<uint64_t, std::string> names;        // counter id -> name

    std::vector<std::string> plan;  // counters of the current plan
    std::vector<Pass> passes;

    bool armed = false;
    rocprofiler_dim3_t armedBlock{};
    size_t armedPass = 0;
    uint64_t expectedDispatch = 0;
    bool collected = false;
    std::map<uint64_t, double> sums;  // counter id -> value summed over instances
};

State& Global()
{
    static State* state = new State();  // never destroyed: the SDK may call back during exit
    return *state;
}

void Fail(char* error, size_t length, const std::string& message)
{
    if (error != nullptr && length > 0)
    {
        std::snprintf(error, length, "%s", message.c_str());
    }
}

std::string StatusText(const rocprofiler_status_t status)
{
    const char* text = rocprofiler_get_status_string(status);
    return text != nullptr ? text : ("status " + std::to_string(static_cast<int>(status)));
}

const std::map<std::string, size_t>& BlockLimits()
{
    static const std::map<std::string, size_t> limits{{"SQ", 8}, {"TCC", 4}, {"TCP", 4}, {"TA", 2}, {"TD", 2}, {"GRBM", 2}};
    return limits;
}

std::string BlockOf(const std::string& hardwareCounter)
{
    const size_t underscore = hardwareCounter.find('_');
    return underscore == std::string::npos ? hardwareCounter : hardwareCounter.substr(0, underscore);
}

// Hardware counters that `name` needs, by name; throws std::string on unknown names and cycles.
void CollectHardware(const State& state, const std::string& name, std::set<std::string>& out, std::set<std::string>& visiting)
{
    const auto found = state.counters.find(name);
    if (found == state.counters.end())
    {
        throw std::string("counter " + name + " is not supported by this GPU");
    }
    const CounterInfo& info = found->second;
    if (!info.derived)
    {
        if (!info.constant)
        {
            out.insert(name);
        }
        return;
    }
    if (!visiting.insert(name).second)
    {
        throw std::string("counter " + name + " has a cyclic definition");
    }
    const std::string& expression = info.expression;
    for (size_t i = 0; i < expression.size();)
    {
        const unsigned char c = static_cast<unsigned char>(expression[i]);
        if (std::isalpha(c) || c == '_')
        {
            size_t end = i + 1;
            while (end < expression.size() && (std::isalnum(static_cast<unsigned char>(expression[end])) || expression[end] == '_'))
            {
                ++end;
            }
            const std::string token = expression.substr(i, end - i);
            if (state.counters.count(token) != 0)
            {
                CollectHardware(state, token, out, visiting);
            }
            i = end;
        }
        else
        {
            ++i;
        }
    }
    visiting.erase(name);
}

// Whether `name` or a counter it is derived from uses accumulate(). Definitions nest a few
// levels at most; the depth bound also ends a cyclic definition.
bool ExpressionAccumulates(const State& state, const std::string& name, const int depth = 0)
{
    const auto found = state.counters.find(name);
    if (depth > 16 || found == state.counters.end() || !found->second.derived)
    {
        return false;
    }
    const std::string& expression = found->second.expression;
    if (expression.find("accumulate(") != std::string::npos)
    {
        return true;
    }
    for (const auto& [other, info] : state.counters)
    {
        (void)info;
        const size_t at = expression.find(other);
        if (at != std::string::npos && other != name)
        {
            const bool startOk = at == 0 || !(std::isalnum(static_cast<unsigned char>(expression[at - 1])) || expression[at - 1] == '_');
            const size_t end = at + other.size();
            const bool endOk = end == expression.size() || !(std::isalnum(static_cast<unsigned char>(expression[end])) || expression[end] == '_');
            if (startOk && endOk && ExpressionAccumulates(state, other, depth + 1))
            {
                return true;
            }
        }
    }
    return false;
}

rocprofiler_status_t ListCounters(rocprofiler_agent_id_t, rocprofiler_counter_id_t* ids, size_t count, void* user)
{
    auto& state = *static_cast<State*>(user);
    for (size_t i = 0; i < count; ++i)
    {
        rocprofiler_counter_info_v0_t info{};
        if (rocprofiler_query_counter_info(ids[i], ROCPROFILER_COUNTER_INFO_VERSION_0, &info) != ROCPROFILER_STATUS_SUCCESS
            || info.name == nullptr)
        {
            continue;
        }
        CounterInfo entry;
        entry.id = ids[i];
        entry.block = info.block != nullptr ? info.block : "";
        entry.expression = info.expression != nullptr ? info.expression : "";
        entry.derived = info.is_derived != 0;
        entry.constant = info.is_constant != 0;
        state.counters[info.name] = entry;
        state.names[ids[i].handle] = info.name;
    }
    return ROCPROFILER_STATUS_SUCCESS;
}

struct AgentQuery
{
    int device;
    bool found = false;
    rocprofiler_agent_id_t agent{};
};

rocprofiler_status_t FindAgent(rocprofiler_agent_version_t, const void** agents, size_t count, void* user)
{
    auto& query = *static_cast<AgentQuery*>(user);
    for (size_t i = 0; i < count; ++i)
    {
        const auto* agent = static_cast<const rocprofiler_agent_v0_t*>(agents[i]);
        if (agent->type == ROCPROFILER_AGENT_TYPE_GPU && agent->logical_node_type_id == query.device)
        {
            query.found = true;
            query.agent = agent->id;
        }
    }
    return ROCPROFILER_STATUS_SUCCESS;
}

void OnDispatch(rocprofiler_dispatch_counting_service_data_t data, rocprofiler_counter_config_id_t* config,
    rocprofiler_user_data_t*, void*)
{
    State& state = Global();
    std::lock_guard<std::mutex> guard(state.lock);
    if (!state.armed || data.dispatch_info.agent_id.handle != state.agent.handle)
    {
        return;
    }
    const auto& block = data.dispatch_info.workgroup_size;
    if (block.x != state.armedBlock.x || block.y != state.armedBlock.y || block.z != state.armedBlock.z)
    {
        return;
    }
    *config = state.passes[state.armedPass].config;
    state.armed = false;
    state.expectedDispatch = data.dispatch_info.dispatch_id;
}

void OnRecord(rocprofiler_dispatch_counting_service_data_t data, rocprofiler_counter_record_t* records, size_t count,
    rocprofiler_user_data_t, void*)
{
    State& state = Global();
    std::lock_guard<std::mutex> guard(state.lock);
    if (state.expectedDispatch == 0 || data.dispatch_info.dispatch_id != state.expectedDispatch)
    {
        return;
    }
    for (size_t i = 0; i < count; ++i)
    {
        rocprofiler_counter_id_t counter{};
        if (rocprofiler_query_record_counter_id(records[i].id, &counter) == ROCPROFILER_STATUS_SUCCESS)
        {
            state.sums[counter.handle] += records[i].counter_value;
        }
    }
    state.collected = true;
    state.expectedDispatch = 0;
    state.arrived.notify_all();
}

int Initialize(rocprofiler_client_finalize_t, void*)
{
    State& state = Global();
    if (rocprofiler_create_context(&state.context) != ROCPROFILER_STATUS_SUCCESS)
    {
        return -1;
    }
    if (rocprofiler_configure_callback_dispatch_counting_service(state.context, OnDispatch, nullptr, OnRecord, nullptr)
        != ROCPROFILER_STATUS_SUCCESS)
    {
        return -1;
    }
    if (rocprofiler_start_context(state.context) != ROCPROFILER_STATUS_SUCCESS)
    {
        return -1;
    }
    state.running = true;
    return 0;
}

void Finalize(void*)
{
    State& state = Global();
    if (state.running.exchange(false))
    {
        rocprofiler_stop_context(state.context);
    }
}

rocprofiler_tool_configure_result_t* Configure(uint32_t, const char*, uint32_t, rocprofiler_client_id_t* client)
{
    client->name = "ktt-hip-counters";
    static rocprofiler_tool_configure_result_t result{sizeof(rocprofiler_tool_configure_result_t), &Initialize, &Finalize, nullptr};
    return &result;
}

} // namespace

extern "C" {

/** Registers the counter service with rocprofiler-sdk. Has effect only before the HIP runtime
 *  initialises in this process. Returns 0 on success. */
int ktt_counters_enable(char* error, size_t length)
{
    State& state = Global();
    if (state.configured)
    {
        return 0;
    }
    const rocprofiler_status_t status = rocprofiler_force_configure(&Configure);
    if (status != ROCPROFILER_STATUS_SUCCESS)
    {
        Fail(error, length, "rocprofiler_force_configure failed: " + StatusText(status));
        return 1;
    }
    state.configured = true;
    return 0;
}

/** 1 while the service's context collects counters, else 0. */
int ktt_counters_running(void)
{
    return Global().running ? 1 : 0;
}

/** Plans the collection of `count` counters on GPU `device` (index among the GPUs the runtime
 *  sees). `pass_of[i]` receives the pass of counter i; `passes` the number of passes. */
int ktt_counters_plan(int device, const char* const* names, size_t count, size_t* pass_of, size_t* passes, char* error,
    size_t length)
{
    State& state = Global();
    std::lock_guard<std::mutex> guard(state.lock);
    if (!state.running)
    {
        Fail(error, length, "the counter service is not running (enable it before the HIP runtime initialises)");
        return 1;
    }
    if (state.armed || state.expectedDispatch != 0)
    {
        Fail(error, length, "a counter collection is in progress");
        return 1;
    }
    if (!state.haveAgent)
    {
        AgentQuery query{device};
        const rocprofiler_status_t status = rocprofiler_query_available_agents(ROCPROFILER_AGENT_INFO_VERSION_0, FindAgent,
            sizeof(rocprofiler_agent_v0_t), &query);
        if (status != ROCPROFILER_STATUS_SUCCESS || !query.found)
        {
            Fail(error, length, "no rocprofiler agent for GPU " + std::to_string(device));
            return 1;
        }
        state.agent = query.agent;
        if (rocprofiler_iterate_agent_supported_counters(state.agent, ListCounters, &state) != ROCPROFILER_STATUS_SUCCESS)
        {
            Fail(error, length, "cannot list the counters of GPU " + std::to_string(device));
            return 1;
        }
        state.haveAgent = true;
    }

    std::vector<Pass> plan;
    std::vector<size_t> assignment(count);
    try
    {
        for (size_t i = 0; i < count; ++i)
        {
            const std::string name = names[i];
            std::set<std::string> hardware;
            std::set<std::string> visiting;
            CollectHardware(state, name, hardware, visiting);
            std::map<std::string, std::set<std::string>> byBlock;
            for (const std::string& counter : hardware)
            {
                const std::string block = BlockOf(counter);
                if (BlockLimits().count(block) == 0)
                {
                    throw std::string("counter " + name + " needs " + counter + " of block " + block
                        + ", which the pass planner does not collect");
                }
                byBlock[block].insert(counter);
            }
            for (const auto& [block, members] : byBlock)
            {
                if (members.size() > BlockLimits().at(block))
                {
                    throw std::string("counter " + name + " needs more " + block + " counters than one pass can collect");
                }
            }
            if (ExpressionAccumulates(state, name))
            {
                // accumulate() reads the SQ accumulator counter beside the named one.
                byBlock["SQ"].insert("SQ_ACCUM_PREV_HIRES");
                if (byBlock["SQ"].size() > BlockLimits().at("SQ"))
                {
                    throw std::string("counter " + name + " needs more SQ counters than one pass can collect");
                }
            }
            // First pass with room for every block this counter uses. Hardware counters that two
            // counters of a pass share are counted for each of them, which never underestimates.
            size_t chosen = plan.size();
            for (size_t p = 0; p < plan.size() && chosen == plan.size(); ++p)
            {
                bool fits = true;
                for (const auto& [block, members] : byBlock)
                {
                    const size_t used = plan[p].usage.count(block) ? plan[p].usage[block] : 0;
                    if (used + members.size() > BlockLimits().at(block))
                    {
                        fits = false;
                    }
                }
                if (fits)
                {
                    chosen = p;
                }
            }
            if (chosen == plan.size())
            {
                plan.emplace_back();
            }
            plan[chosen].members.push_back(i);
            for (const auto& [block, members] : byBlock)
            {
                plan[chosen].usage[block] += members.size();
            }
            assignment[i] = chosen;
        }
    }
    catch (const std::string& message)
    {
        Fail(error, length, message);
        return 1;
    }

    for (Pass& pass : plan)
    {
        std::vector<rocprofiler_counter_id_t> ids;
        for (const size_t member : pass.members)
        {
            ids.push_back(state.counters.at(names[member]).id);
        }
        pass.config = {};
        const rocprofiler_status_t status = rocprofiler_create_counter_config(state.agent, ids.data(), ids.size(), &pass.config);
        if (status != ROCPROFILER_STATUS_SUCCESS)
        {
            Fail(error, length, "rocprofiler_create_counter_config failed: " + StatusText(status));
            return 1;
        }
    }
    state.plan.assign(names, names + count);
    state.passes = std::move(plan);
    for (size_t i = 0; i < count; ++i)
    {
        pass_of[i] = assignment[i];
    }
    *passes = state.passes.size();
    return 0;
}

/** Collects pass `pass` of the current plan on the next dispatch with this workgroup size. */
int ktt_counters_arm(size_t pass, unsigned block_x, unsigned block_y, unsigned block_z, char* error, size_t length)
{
    State& state = Global();
    std::lock_guard<std::mutex> guard(state.lock);
    if (pass >= state.passes.size())
    {
        Fail(error, length, "no pass " + std::to_string(pass) + " in the counter plan");
        return 1;
    }
    state.armed = true;
    state.armedBlock = {block_x, block_y, block_z};
    state.armedPass = pass;
    state.expectedDispatch = 0;
    state.collected = false;
    state.sums.clear();
    return 0;
}

/** Waits for the armed pass's values after its kernel finished. `values[i]` receives counter i
 *  of the plan when it belongs to that pass (summed over its instances); others are left. */
int ktt_counters_collect(double* values, int timeout_ms, char* error, size_t length)
{
    State& state = Global();
    std::unique_lock<std::mutex> guard(state.lock);
    const bool done = state.arrived.wait_for(guard, std::chrono::milliseconds(timeout_ms), [&]() { return state.collected; });
    if (!done)
    {
        const bool dispatched = !state.armed;
        state.armed = false;
        state.expectedDispatch = 0;
        Fail(error, length, dispatched ? "the counter record of the profiled dispatch did not arrive"
                                       : "no dispatch with the armed workgroup size was seen");
        return 1;
    }
    for (const size_t member : state.passes[state.armedPass].members)
    {
        const uint64_t id = state.counters.at(state.plan[member]).id.handle;
        const auto found = state.sums.find(id);
        values[member] = found != state.sums.end() ? found->second : 0.0;
    }
    state.collected = false;
    return 0;
}

/** Cancels an armed pass that will not be launched. */
void ktt_counters_disarm(void)
{
    State& state = Global();
    std::lock_guard<std::mutex> guard(state.lock);
    state.armed = false;
    state.expectedDispatch = 0;
    state.collected = false;
}

/** Writes the names of the counters the GPU of the last plan supports, one per line. Returns
 *  the length needed (without the terminator). */
size_t ktt_counters_list(char* out, size_t length)
{
    State& state = Global();
    std::lock_guard<std::mutex> guard(state.lock);
    std::string text;
    for (const auto& [name, info] : state.counters)
    {
        (void)info;
        text += name + "\n";
    }
    if (out != nullptr && length > 0)
    {
        std::snprintf(out, length, "%s", text.c_str());
    }
    return text.size();
}

} // extern "C"
