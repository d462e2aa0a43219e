// Profile-based searcher: bottleneck analysis of hardware counters, counter changes wanted, and
// candidate scoring (Source/Api/Searcher/ProfileBasedSearcher.h:123-858).
#include <algorithm>
#include <cmath>
#include <set>

#include <KttException.h>
#include <ProfileBasedSearcher.h>
#include <Tuner.h>

#include "Common.h"

namespace ktt
{

namespace
{

using Features = std::map<std::string, double>;

double Get(const Features& features, const std::string& name)
{
    const auto found = features.find(name);
    return found != features.end() ? found->second : 0.0;
}

double Clamp01(const double value)
{
    return std::isfinite(value) ? std::min(1.0, std::max(0.0, value)) : 0.0;
}

double Share(const double part, const double whole)
{
    return whole > 0.0 ? part / whole : 0.0;
}

double CounterValue(const KernelProfilingCounter& counter)
{
    switch (counter.GetType())
    {
    case ProfilingCounterType::Int: return static_cast<double>(counter.GetValueInt());
    case ProfilingCounterType::UnsignedInt:
    case ProfilingCounterType::Throughput:
    case ProfilingCounterType::UtilizationLevel: return static_cast<double>(counter.GetValueUint());
    default: return counter.GetValueDouble();
    }
}

// Counter sets the analysis understands.
const char* const kCudaDram = "dram__throughput.avg.pct_of_peak_sustained_elapsed";
const char* const kCdnaValu = "VALUBusy";
const char* const kCdnaQuadCycles = "VALU_QUAD_CYCLES";

bool IsCuda(const Features& features) { return features.count(kCudaDram) != 0; }
bool IsCdna(const Features& features) { return features.count("F32_INSTS") != 0; }

// The reference's analyzeBottlenecks for compute capability 7.0 and newer
// (ProfileBasedSearcher.h:123-340), with the local-memory overhead taken from the counter's
// value (the reference divides the counter's position in its name list).
Features CudaBottlenecks(const Features& f, const ProfileSearcherOptions& options)
{
    Features bn;
    const double dramUtil = Get(f, kCudaDram) / 10.0;
    const double dramRead = Get(f, "dram__sectors_read.sum");
    const double dramWrite = Get(f, "dram__sectors_write.sum");
    const bool dram = dramRead + dramWrite > 0.0 && dramUtil > 0.0;
    bn["bnDRAMRead"] = dram ? Share(dramRead, dramRead + dramWrite) * (dramUtil / 10.0) : 0.0;
    bn["bnDRAMWrite"] = dram ? Share(dramWrite, dramRead + dramWrite) * (dramUtil / 10.0) : 0.0;

    const double l2Util = Get(f, "lts__t_sectors.avg.pct_of_peak_sustained_elapsed") / 10.0;
    const double l2Read = Get(f, "lts__t_sectors_op_read.sum");
    const double l2Write = Get(f, "lts__t_sectors_op_write.sum");
    const double texUtil = Get(f, "l1tex__t_requests_pipe_lsu_mem_global_op_ld.avg.pct_of_peak_sustained_active") / 10.0;
    bn["bnL2Read"] = Share(l2Read, l2Read + l2Write) * (l2Util / 10.0);
    bn["bnL2Write"] = Share(l2Write, l2Read + l2Write) * (l2Util / 10.0);
    bn["bnTex"] = texUtil / 10.0;

    const double localOverhead = 100.0 * Share(Get(f, "l1tex__t_sectors_pipe_lsu_mem_local_op_st.sum"), l2Write);
    bn["bnLocal"] = (localOverhead / 100.0) * std::max({dramUtil / 10.0, l2Util / 10.0, texUtil / 10.0});

    const double sharedUtil = Get(f, "l1tex__data_pipe_lsu_wavefronts_mem_shared.avg.pct_of_peak_sustained_elapsed") / 10.0;
    const double sharedRead = Get(f, "l1tex__data_pipe_lsu_wavefronts_mem_shared_op_ld.sum");
    const double sharedWrite = Get(f, "l1tex__data_pipe_lsu_wavefronts_mem_shared_op_st.sum");
    const bool shared = sharedRead + sharedWrite > 0.0;
    bn["bnSMRead"] = shared ? Share(sharedRead, sharedRead + sharedWrite) * (sharedUtil / 10.0) : 0.0;
    bn["bnSMWrite"] = shared ? Share(sharedWrite, sharedRead + sharedWrite) * (sharedUtil / 10.0) : 0.0;

    bn["bnMPparal"] = 1.0 - Get(f, "sm__warps_active.avg.pct_of_peak_sustained_active") / 100.0;
    bn["bnGparal"] = (100.0 - Get(f, "smsp__cycles_active.avg.pct_of_peak_sustained_elapsed")) / 100.0;

    const double units = options.computeUnits > 0 ? options.computeUnits : 1.0;
    const double cores = units * options.lanesPerUnit;
    const double groups = Share(Get(f, "Global size"), Get(f, "Local size"));
    bn["bnTailEffect"] = groups > 0.0 ? 1.0 - groups / (std::ceil(groups / units) * units) : 0.0;
    bn["bnThreads"] = std::max(0.0, (cores * 5.0 - Get(f, "Global size")) / (cores * 5.0));

    const double sp = Get(f, "smsp__sass_thread_inst_executed_op_fp32_pred_on.sum");
    const double dp = Get(f, "smsp__sass_thread_inst_executed_op_fp64_pred_on.sum");
    const double integer = Get(f, "smsp__sass_thread_inst_executed_op_integer_pred_on.sum");
    const double misc = Get(f, "smsp__sass_thread_inst_executed_op_misc_pred_on.sum");
    const double ldst = Get(f, "smsp__sass_thread_inst_executed_op_memory_pred_on.sum");
    const double control = Get(f, "smsp__sass_thread_inst_executed_op_control_pred_on.sum");
    const double conversion = Get(f, "smsp__sass_thread_inst_executed_op_conversion_pred_on.sum");
    const double executed = Get(f, "smsp__inst_executed.sum");
    const double sfuUtil = Get(f, "smsp__inst_executed_pipe_xu.avg.pct_of_peak_sustained_active") / 10.0;
    const double texFuUtil = Get(f, "smsp__inst_executed_pipe_tex.avg.pct_of_peak_sustained_active") / 10.0;
    const double slotUtil = Get(f, "smsp__issue_active.avg.pct_of_peak_sustained_active");
    const double efficiencyExec = Get(f, "smsp__thread_inst_executed_per_inst_executed.ratio") * 100.0 / 32.0;
    const double efficiencyPred = Get(f, "smsp__thread_inst_executed_per_inst_executed.pct");

    const double fitted = executed * 32.0 * Share(100.0, efficiencyExec) * Share(100.0, efficiencyPred);
    const double issue = std::min(1.0, slotUtil / 50.0);  // dual issue: one class reaches at most 50 %
    auto approx = [&](const double count) { return Share(count, fitted) * issue; };
    bn["bnSP"] = approx(sp);
    bn["bnDP"] = approx(dp);
    bn["bnSFU"] = sfuUtil / 10.0;
    bn["bnCF"] = approx(control);
    bn["bnLDST"] = approx(ldst);
    bn["bnTexFu"] = texFuUtil / 10.0;
    bn["bnInt"] = approx(integer);
    bn["bnMisc"] = approx(misc);
    bn["bnBconv"] = approx(conversion);

    const double threshold = options.instructionThreshold;
    const double maxClass = std::max({Share(bn["bnSP"], issue), Share(bn["bnDP"], issue), sfuUtil / 10.0, Share(bn["bnCF"], issue),
        Share(bn["bnLDST"], issue), Share(bn["bnInt"], issue), Share(bn["bnMisc"], issue), Share(bn["bnBconv"], issue)});
    const double weight = maxClass > threshold ? (maxClass - threshold) / (1.0 - threshold) : 0.0;
    bn["bnInstIssue"] = (100.0 - slotUtil) / 100.0 * weight;
    return bn;
}

// Issue slots a single-precision transcendental (v_rsq_f32 and its kind) holds, in units of one
// v_fma_f32. On gfx950 a v_rsq_f32 holds a SIMD for two quad-cycles and never pairs, while a
// v_fma_f32 whose operands are all VGPRs takes half a quad-cycle when it pairs with another
// wave's and a whole one when it does not: the weight lies between 2 and 4. The issue-rate
// probes (ktt_hip_issue_rates: 8 independent chains per thread, 8 waves per SIMD, marginal
// rates) give 3.0-3.5 for an FMA-only stream against v_rsq_f32 on MI355X; an issue-bound kernel's
// time is its issue quad-cycles less the paired ones (SQ_ACTIVE_INST_VALU - SQ_ACTIVE_INST_VALU2,
// 99.7% of the Coulomb kernels' busy cycles).
constexpr double kTransIssueSlots = 3.2;

// The same bottlenecks from CDNA counters. Utilisations of HBM, L2 and vector L1 are bytes
// moved over the kernel's duration against the peak rates of the options; LDS, occupancy and
// SIMD activity come from the ratio counters; instruction classes are their share of VALU (or
// of all) instructions times the VALU busy fraction. Transcendentals are their share of the VALU
// issue time, each weighted kTransIssueSlots; with the dual-issue counters (SQ_ACTIVE_INST_VALU,
// SQ_ACTIVE_INST_VALU2) that time is measured, and the quad-cycles lost to instructions that
// issued alone form a class of their own (bnPairing), answered by candidates that pair more.
Features CdnaBottlenecks(const Features& f, const ProfileSearcherOptions& options)
{
    Features bn;
    const double seconds = Get(f, "kernel_duration_ns") * 1.0e-9;
    const double units = options.computeUnits > 0 ? options.computeUnits : 256.0;
    auto utilisation = [&](const double bytes, const double peak) { return seconds > 0.0 && peak > 0.0 ? Clamp01(bytes / seconds / peak) : 0.0; };

    const double hbmRead = Get(f, "FETCH_SIZE") * 1024.0;
    const double hbmWrite = Get(f, "WRITE_SIZE") * 1024.0;
    const double hbm = utilisation(hbmRead + hbmWrite, options.memoryBandwidth);
    bn["bnDRAMRead"] = Share(hbmRead, hbmRead + hbmWrite) * hbm;
    bn["bnDRAMWrite"] = Share(hbmWrite, hbmRead + hbmWrite) * hbm;

    const double l2Read = Get(f, "TCP_TCC_READ_REQ_sum");
    const double l2Write = Get(f, "TCP_TCC_WRITE_REQ_sum");
    const double l2 = utilisation((l2Read + l2Write) * 64.0, options.l2Bandwidth);
    bn["bnL2Read"] = Share(l2Read, l2Read + l2Write) * l2;
    bn["bnL2Write"] = Share(l2Write, l2Read + l2Write) * l2;

    const double l1Peak = options.l1Bandwidth > 0.0 ? options.l1Bandwidth : units * 64.0 * options.clockHz;
    const double l1 = utilisation(Get(f, "TCP_TOTAL_CACHE_ACCESSES_sum") * 64.0, l1Peak);
    bn["bnTex"] = l1;
    bn["bnLocal"] = Get(f, "scratch_bytes") > 0.0 ? std::max({hbm, l2, l1}) : 0.0;

    bn["bnSMRead"] = Clamp01(Get(f, "LdsUtil") / 100.0);
    bn["bnSMWrite"] = 0.0;
    bn["bnMPparal"] = 1.0 - Clamp01(Get(f, "OccupancyPercent") / 100.0);
    bn["bnGparal"] = 1.0 - Clamp01(Get(f, "SIMD_UTILIZATION"));

    const double cores = units * options.lanesPerUnit;
    const double groups = Share(Get(f, "Global size"), Get(f, "Local size"));
    bn["bnTailEffect"] = groups > 0.0 ? 1.0 - groups / (std::ceil(groups / units) * units) : 0.0;
    bn["bnThreads"] = std::max(0.0, (cores * 5.0 - Get(f, "Global size")) / (cores * 5.0));

    const double issue = Clamp01(Get(f, kCdnaValu) / 100.0);
    const double valu = Get(f, "SQ_INSTS_VALU");
    const double memory = Get(f, "VMEM_INSTS") + Get(f, "SQ_INSTS_LDS");
    const double all = valu + Get(f, "SCALAR_INSTS") + memory + Get(f, "SQ_INSTS_BRANCH");
    bn["bnSP"] = Share(Get(f, "F32_INSTS"), valu) * issue;
    bn["bnDP"] = Share(Get(f, "F64_INSTS"), valu) * issue;
    const double trans = std::min(Get(f, "SQ_INSTS_VALU_TRANS_F32"), valu);
    const double quadCycles = Get(f, kCdnaQuadCycles);
    if (quadCycles > 0.0)
    {
        // Measured issue time: a transcendental holds its SIMD for two quad-cycles and never pairs;
        // the other VALU instructions take half a quad-cycle each when every one pairs, so what
        // the kernel took beyond that was lost to instructions that issued alone.
        bn["bnSFU"] = Clamp01(Share(2.0 * trans, quadCycles)) * issue;
        const double paired = 0.5 * (valu - trans) + 2.0 * trans;
        bn["bnPairing"] = Clamp01(Share(quadCycles - paired, quadCycles)) * issue;
    }
    else
    {
        bn["bnSFU"] = Share(kTransIssueSlots * trans, kTransIssueSlots * trans + (valu - trans)) * issue;
        bn["bnPairing"] = 0.0;
    }
    // FP32, transcendental, integer and conversion instructions share one VALU issue port on CDNA:
    // a busy VALU is a bottleneck of their weighted sum, whichever class dominates it.
    bn["bnVALU"] = issue;
    bn["bnCF"] = Share(Get(f, "SQ_INSTS_BRANCH"), all) * issue;
    bn["bnLDST"] = Share(memory, all) * issue;
    bn["bnTexFu"] = 0.0;
    bn["bnInt"] = Share(Get(f, "SQ_INSTS_VALU_INT32"), valu) * issue;
    bn["bnMisc"] = Clamp01(Get(f, "SALUBusy") / 100.0);
    bn["bnBconv"] = Share(Get(f, "SQ_INSTS_VALU_CVT"), valu) * issue;

    const double threshold = options.instructionThreshold;
    // The VALU's share of all instructions counts as a class of its own: its sub-classes share one
    // issue port.
    const double maxClass = std::max({Share(bn["bnSP"], issue), Share(bn["bnDP"], issue), Share(bn["bnSFU"], issue),
        Share(bn["bnCF"], issue), Share(bn["bnLDST"], issue), Share(bn["bnInt"], issue), Share(bn["bnMisc"], issue),
        Share(bn["bnBconv"], issue), Share(valu, all)});
    const double weight = maxClass > threshold ? (maxClass - threshold) / (1.0 - threshold) : 0.0;
    bn["bnInstIssue"] = (1.0 - issue) * weight;
    return bn;
}

// CDNA features that count work per launched thread (instructions, quad-cycles, cache requests).
// A knowledge base profiled at another problem size launched other thread counts; these scale to
// this search's launch by the ratio of the threads. The bytes moved from memory, the write
// requests of the output, the utilisations and the duration do not grow with launched threads
// that fall outside the problem, and stay as profiled.
const char* const kLaunchCounts[] = {"TCP_TCC_READ_REQ_sum", "TCP_TOTAL_CACHE_ACCESSES_sum", "SQ_INSTS_LDS", "SQ_INSTS_VALU_INT32",
    "SQ_INSTS_VALU_CVT", "SQ_INSTS_VALU_TRANS_F32", "SQ_INSTS_BRANCH", "SQ_INSTS_VALU", "F32_INSTS", "F64_INSTS", "VMEM_INSTS",
    "SCALAR_INSTS", "VALU_ISSUE_SLOTS", "SQ_ACTIVE_INST_VALU2", kCdnaQuadCycles};

// CDNA utilisations averaged over the device: compute units a launch leaves without a workgroup
// count as idle, so these scale with the share of the units each launch occupies.
const char* const kDeviceShares[] = {"VALUBusy", "SALUBusy", "OccupancyPercent", "LdsUtil", "SIMD_UTILIZATION"};

Features ScaledToLaunch(const Features& known, const double threads, const double computeUnits)
{
    Features scaled = known;
    const double recorded = Get(known, "Global size");
    if (recorded <= 0.0 || threads <= 0.0 || !IsCdna(known))
    {
        return scaled;
    }
    const double factor = threads / recorded;
    for (const char* name : kLaunchCounts)
    {
        const auto found = scaled.find(name);
        if (found != scaled.end())
        {
            found->second *= factor;
        }
    }
    const double local = Get(known, "Local size");
    if (local > 0.0 && computeUnits > 0.0)
    {
        const double occupied = std::min(1.0, recorded / local / computeUnits);
        const double target = std::min(1.0, threads / local / computeUnits);
        for (const char* name : kDeviceShares)
        {
            const auto found = scaled.find(name);
            if (found != scaled.end())
            {
                found->second *= target / occupied;
            }
        }
    }
    return scaled;
}

double Above(const double value, const double threshold)
{
    return (value - threshold) / (1.0 - threshold);
}

std::string ConfigurationKey(const KernelConfiguration& configuration)
{
    std::vector<std::string> parts;
    for (const auto& pair : configuration.GetPairs())
    {
        parts.push_back(pair.GetName() + "=" + pair.GetValueString());
    }
    std::sort(parts.begin(), parts.end());
    std::string key;
    for (const auto& part : parts)
    {
        key += part + ";";
    }
    return key;
}

size_t Differences(const KernelConfiguration& a, const KernelConfiguration& b)
{
    size_t count = 0;
    for (const auto& pair : a.GetPairs())
    {
        const auto& others = b.GetPairs();
        const auto found = std::find_if(others.begin(), others.end(), [&](const ParameterPair& other) { return other.GetName() == pair.GetName(); });
        count += found == others.end() || !found->HasSameValue(pair) ? 1 : 0;
    }
    return count;
}

} // namespace

// ---------------------------------------------------------------- analysis

std::map<std::string, double> ProfileBasedSearcher::Features(const ComputationResult& computation)
{
    ::ktt::Features features;
    if (!computation.HasProfilingData() || computation.GetProfilingData().HasRemainingProfilingRuns())
    {
        return features;
    }
    ::ktt::Features raw;
    for (const auto& counter : computation.GetProfilingData().GetCounters())
    {
        raw[counter.GetName()] = CounterValue(counter);
    }
    const double local = static_cast<double>(computation.GetLocalSize().GetTotalSize());
    const double global = static_cast<double>(computation.GetGlobalSize().GetTotalSize()) * local;
    if (IsCuda(raw))
    {
        raw["Global size"] = global;
        raw["Local size"] = local;
        return raw;
    }
    if (raw.count(kCdnaValu) == 0 || raw.count("FETCH_SIZE") == 0)
    {
        return features;
    }
    for (const char* name : {"FETCH_SIZE", "WRITE_SIZE", "TCP_TCC_READ_REQ_sum", "TCP_TCC_WRITE_REQ_sum", "TCP_TOTAL_CACHE_ACCESSES_sum",
             "scratch_bytes", "SQ_INSTS_LDS", "SQ_INSTS_VALU_INT32", "SQ_INSTS_VALU_CVT", "SQ_INSTS_VALU_TRANS_F32", "SQ_INSTS_BRANCH",
             "SQ_INSTS_VALU", "VALUBusy", "SALUBusy", "OccupancyPercent", "LdsUtil", "SIMD_UTILIZATION", "kernel_duration_ns"})
    {
        features[name] = Get(raw, name);
    }
    features["F32_INSTS"] = Get(raw, "SQ_INSTS_VALU_FMA_F32") + Get(raw, "SQ_INSTS_VALU_ADD_F32") + Get(raw, "SQ_INSTS_VALU_MUL_F32");
    features["F64_INSTS"] = Get(raw, "SQ_INSTS_VALU_FMA_F64") + Get(raw, "SQ_INSTS_VALU_ADD_F64") + Get(raw, "SQ_INSTS_VALU_MUL_F64");
    features["VMEM_INSTS"] = Get(raw, "SQ_INSTS_VMEM_RD") + Get(raw, "SQ_INSTS_VMEM_WR");
    features["SCALAR_INSTS"] = Get(raw, "SQ_INSTS_SALU") + Get(raw, "SQ_INSTS_SMEM");
    // VALU issue time in FMA slots: every VALU instruction one, a transcendental kTransIssueSlots.
    features["VALU_ISSUE_SLOTS"] = Get(raw, "SQ_INSTS_VALU") + (kTransIssueSlots - 1.0) * Get(raw, "SQ_INSTS_VALU_TRANS_F32");
    // With the dual-issue counters, the VALU issue time is measured instead: the quad-cycles
    // issued less those in which two waves issued together.
    if (raw.count("SQ_ACTIVE_INST_VALU") != 0 && raw.count("SQ_ACTIVE_INST_VALU2") != 0)
    {
        features["SQ_ACTIVE_INST_VALU2"] = Get(raw, "SQ_ACTIVE_INST_VALU2");
        features[kCdnaQuadCycles] = std::max(0.0, Get(raw, "SQ_ACTIVE_INST_VALU") - Get(raw, "SQ_ACTIVE_INST_VALU2"));
    }
    features["Global size"] = global;
    features["Local size"] = local;
    return features;
}

std::map<std::string, double> ProfileBasedSearcher::AnalyzeBottlenecks(const std::map<std::string, double>& features,
    const ProfileSearcherOptions& options)
{
    if (IsCuda(features))
    {
        return CudaBottlenecks(features, options);
    }
    if (IsCdna(features))
    {
        return CdnaBottlenecks(features, options);
    }
    return {};
}

std::map<std::string, double> ProfileBasedSearcher::ComputeChanges(const std::map<std::string, double>& bn,
    const std::map<std::string, double>& features, const ProfileSearcherOptions& options)
{
    ::ktt::Features changes;
    const double r = options.instructionThreshold;
    auto b = [&](const char* name) { return Get(bn, name); };
    if (IsCuda(features))
    {
        // computeChanges for compute capability 7.0 and newer (ProfileBasedSearcher.h:349-420).
        changes["dram__sectors_read.sum"] = -b("bnDRAMRead");
        changes["dram__sectors_write.sum"] = -b("bnDRAMWrite");
        changes["lts__t_sectors_op_read.sum"] = -b("bnL2Read");
        changes["lts__t_sectors_op_write.sum"] = -b("bnL2Write");
        changes["l1tex__t_requests_pipe_lsu_mem_global_op_ld.sum"] = -b("bnTex");
        changes["l1tex__t_sectors_pipe_lsu_mem_local_op_ld.sum"] = -b("bnLocal");
        changes["l1tex__t_sectors_pipe_lsu_mem_local_op_st.sum"] = -b("bnLocal");
        changes["l1tex__data_pipe_lsu_wavefronts_mem_shared_op_ld.sum"] = -b("bnSMRead");
        changes["l1tex__data_pipe_lsu_wavefronts_mem_shared_op_st.sum"] = -b("bnSMWrite");
        if (b("bnSP") > r) changes["smsp__sass_thread_inst_executed_op_fp32_pred_on.sum"] = -b("bnSP");
        if (b("bnDP") > r) changes["smsp__sass_thread_inst_executed_op_fp64_pred_on.sum"] = -b("bnDP");
        if (b("bnCF") > r) changes["smsp__sass_thread_inst_executed_op_control_pred_on.sum"] = -Above(b("bnCF"), r);
        if (b("bnLDST") > r) changes["smsp__sass_thread_inst_executed_op_memory_pred_on.sum"] = -Above(b("bnLDST"), r);
        if (b("bnInt") > r) changes["smsp__sass_thread_inst_executed_op_integer_pred_on.sum"] = -Above(b("bnInt"), r);
        if (b("bnMisc") > r) changes["smsp__sass_thread_inst_executed_op_misc_pred_on.sum"] = -Above(b("bnMisc"), r);
        if (b("bnBconv") > r) changes["smsp__sass_thread_inst_executed_op_conversion_pred_on.sum"] = -Above(b("bnBconv"), r);
        changes["smsp__issue_active.avg.pct_of_peak_sustained_active"] = b("bnInstIssue");
        changes["smsp__cycles_active.avg.pct_of_peak_sustained_elapsed"] = b("bnGparal");
        changes["Global size"] = b("bnThreads");
        return changes;
    }
    if (IsCdna(features))
    {
        changes["FETCH_SIZE"] = -b("bnDRAMRead");
        changes["WRITE_SIZE"] = -b("bnDRAMWrite");
        changes["TCP_TCC_READ_REQ_sum"] = -b("bnL2Read");
        changes["TCP_TCC_WRITE_REQ_sum"] = -b("bnL2Write");
        changes["TCP_TOTAL_CACHE_ACCESSES_sum"] = -b("bnTex");
        changes["scratch_bytes"] = -b("bnLocal");
        changes["SQ_INSTS_LDS"] = -(b("bnSMRead") + b("bnSMWrite"));
        if (b("bnSP") > r) changes["F32_INSTS"] = -b("bnSP");
        if (b("bnDP") > r) changes["F64_INSTS"] = -b("bnDP");
        if (b("bnSFU") > r) changes["SQ_INSTS_VALU_TRANS_F32"] = -Above(b("bnSFU"), r);
        if (b("bnVALU") > r) changes[features.count(kCdnaQuadCycles) != 0 ? kCdnaQuadCycles : "VALU_ISSUE_SLOTS"] = -Above(b("bnVALU"), r);
        // Quad-cycles lost to unpaired instructions count whatever the VALU load: a candidate that
        // pairs more of the same instructions issues them in less time.
        if (b("bnPairing") > 0.0) changes["SQ_ACTIVE_INST_VALU2"] = b("bnPairing");
        if (b("bnCF") > r) changes["SQ_INSTS_BRANCH"] = -Above(b("bnCF"), r);
        if (b("bnLDST") > r) changes["VMEM_INSTS"] = -Above(b("bnLDST"), r);
        if (b("bnInt") > r) changes["SQ_INSTS_VALU_INT32"] = -Above(b("bnInt"), r);
        if (b("bnMisc") > r) changes["SCALAR_INSTS"] = -Above(b("bnMisc"), r);
        if (b("bnBconv") > r) changes["SQ_INSTS_VALU_CVT"] = -Above(b("bnBconv"), r);
        changes[kCdnaValu] = b("bnInstIssue");
        changes["SIMD_UTILIZATION"] = b("bnGparal");
        changes["Global size"] = b("bnThreads");
    }
    return changes;
}

std::vector<double> ProfileBasedSearcher::ScoreCandidates(const std::map<std::string, double>& changes,
    const std::map<std::string, double>& current, const std::vector<const std::map<std::string, double>*>& candidates,
    const ProfileSearcherOptions& options)
{
    // scoreTuningConfigurationsExact (ProfileBasedSearcher.h:433-496): a candidate scores the sum
    // over counters of importance x relative difference to the current configuration; scores
    // are normalised to (0, 1) below zero and (1, 2] above, raised to the exponent, and floored
    // at 0.0001 (also for candidates below the cut-off and those without counters).
    std::vector<double> scores(candidates.size(), 0.0);
    for (size_t i = 0; i < candidates.size(); ++i)
    {
        if (candidates[i] == nullptr)
        {
            continue;
        }
        for (const auto& [name, importance] : changes)
        {
            const auto mine = candidates[i]->find(name);
            const auto theirs = current.find(name);
            if (mine == candidates[i]->end() || theirs == current.end())
            {
                continue;
            }
            const double sum = mine->second + theirs->second;
            if (sum != 0.0)
            {
                scores[i] += importance * (mine->second - theirs->second) / sum;
            }
        }
    }
    if (scores.empty())
    {
        return scores;
    }
    const double lowest = *std::min_element(scores.begin(), scores.end());
    const double highest = *std::max_element(scores.begin(), scores.end());
    for (double& score : scores)
    {
        if (score < options.cutoff)
        {
            score = 0.0;
        }
        else
        {
            if (score < 0.0)
            {
                score = 1.0 - score / lowest;
            }
            else if (score > 0.0)
            {
                score = 1.0 + score / highest;
            }
            score = std::pow(score, options.exponent);
        }
        score = std::max(score, 0.0001);
    }
    return scores;
}

// ---------------------------------------------------------------- searcher

ProfileBasedSearcher::ProfileBasedSearcher(Tuner& tuner, const ProfileSearcherOptions& options, const std::vector<KernelResult>& knowledge) :
    ProfileBasedSearcher(options, knowledge)
{
    m_Tuner = &tuner;
    if (m_Options.computeUnits == 0)
    {
        m_Options.computeUnits = tuner.GetCurrentDeviceInfo().GetMaxComputeUnits();
    }
}

ProfileBasedSearcher::ProfileBasedSearcher(const ProfileSearcherOptions& options, const std::vector<KernelResult>& knowledge) :
    Searcher(),
    m_Tuner(nullptr),
    m_Options(options),
    m_Random(options.seed != 0 ? options.seed : DefaultSearcherSeed()),
    m_HaveBest(false),
    m_BestDuration(0.0),
    m_Profiled(0)
{
    if (m_Options.batchSize == 0)
    {
        throw KttException("The profile-based searcher needs a batch of at least one configuration");
    }
    for (const KernelResult& result : knowledge)
    {
        if (!result.IsValid() || result.GetResults().empty())
        {
            continue;
        }
        auto features = Features(result.GetResults().front());
        if (!features.empty())
        {
            m_Knowledge.emplace(ConfigurationKey(result.GetConfiguration()), std::move(features));
        }
    }
    if (!knowledge.empty() && m_Knowledge.empty())
    {
        throw KttException("The knowledge base of the profile-based searcher holds no results with counters it understands");
    }
}

void ProfileBasedSearcher::OnInitialize()
{
    m_Batch.clear();
    m_History.clear();
    m_HaveBest = false;
    m_BestDuration = 0.0;
    m_Profiled = 0;
    FillRandomBatch(std::min<uint64_t>(m_Options.batchSize, GetConfigurationsCount()));
    m_Current = m_Batch.front();
    m_Batch.erase(m_Batch.begin());
}

void ProfileBasedSearcher::OnReset()
{
    m_Batch.clear();
    m_HaveBest = false;
}

KernelConfiguration ProfileBasedSearcher::GetCurrentConfiguration() const
{
    return m_Current;
}

void ProfileBasedSearcher::SetProfiling(const bool on)
{
    if (m_Tuner != nullptr)
    {
        m_Tuner->SetProfiling(on);
    }
}

void ProfileBasedSearcher::FillRandomBatch(const size_t count)
{
    std::set<uint64_t> chosen;
    for (const auto& configuration : m_Batch)
    {
        chosen.insert(GetIndex(configuration));
    }
    // Batch members are unexplored themselves, so the unexplored count bounds the batch.
    const size_t target = std::max<uint64_t>(m_Batch.size(), std::min<uint64_t>(m_Batch.size() + count, GetUnexploredConfigurationsCount()));
    while (m_Batch.size() < target)
    {
        const KernelConfiguration drawn = GetRandomConfiguration();
        if (chosen.insert(GetIndex(drawn)).second)
        {
            m_Batch.push_back(drawn);
        }
    }
}

const std::map<std::string, double>* ProfileBasedSearcher::KnowledgeOf(const KernelConfiguration& configuration) const
{
    const auto found = m_Knowledge.find(ConfigurationKey(configuration));
    return found != m_Knowledge.end() ? &found->second : nullptr;
}

std::map<std::string, double> ProfileBasedSearcher::Estimate(const KernelConfiguration& configuration) const
{
    // Counters of the nearest profiled configuration(s); nearest ties are averaged over the
    // counters all of them hold.
    size_t nearest = SIZE_MAX;
    std::vector<const ::ktt::Features*> tied;
    for (const auto& [profiled, features] : m_History)
    {
        const size_t distance = Differences(configuration, profiled);
        if (distance < nearest)
        {
            nearest = distance;
            tied.clear();
        }
        if (distance == nearest)
        {
            tied.push_back(&features);
        }
    }
    ::ktt::Features estimate;
    if (tied.empty())
    {
        return estimate;
    }
    for (const auto& [name, value] : *tied.front())
    {
        double sum = 0.0;
        bool everywhere = true;
        for (const auto* features : tied)
        {
            const auto found = features->find(name);
            everywhere = everywhere && found != features->end();
            sum += found != features->end() ? found->second : 0.0;
        }
        if (everywhere)
        {
            estimate[name] = sum / static_cast<double>(tied.size());
        }
        (void)value;
    }
    return estimate;
}

size_t ProfileBasedSearcher::Draw(std::vector<double>& weights)
{
    double total = 0.0;
    for (const double weight : weights)
    {
        total += weight;
    }
    size_t index = 0;
    if (total > 0.0)
    {
        const double target = std::uniform_real_distribution<double>(0.0, total)(m_Random);
        double running = 0.0;
        index = weights.size() - 1;
        for (size_t i = 0; i < weights.size(); ++i)
        {
            running += weights[i];
            if (weights[i] > 0.0 && target < running)
            {
                index = i;
                break;
            }
        }
        while (weights[index] <= 0.0 && index > 0)
        {
            --index;
        }
    }
    else
    {
        index = std::uniform_int_distribution<size_t>(0, weights.size() - 1)(m_Random);
    }
    weights[index] = 0.0;
    return index;
}

void ProfileBasedSearcher::NextBatch(const KernelResult& profiled)
{
    if (profiled.GetResults().size() > 1)
    {
        Logger::LogWarning("The profile-based searcher analyses the counters of the first kernel of a composition only");
    }
    const ::ktt::Features features = profiled.GetResults().empty() ? ::ktt::Features{} : Features(profiled.GetResults().front());

    // Candidates: unexplored neighbours of the profiled configuration, then random ones.
    std::vector<KernelConfiguration> candidates = GetNeighbourConfigurations(m_BestOfBatch, m_Options.neighbourDistance,
        m_Options.neighbourCount);
    std::set<uint64_t> chosen;
    for (const auto& candidate : candidates)
    {
        chosen.insert(GetIndex(candidate));
    }
    const size_t target = std::min<uint64_t>(candidates.size() + m_Options.randomCount, GetUnexploredConfigurationsCount());
    while (candidates.size() < target)
    {
        const KernelConfiguration drawn = GetRandomConfiguration();
        if (chosen.insert(GetIndex(drawn)).second)
        {
            candidates.push_back(drawn);
        }
    }

    std::vector<double> weights(candidates.size(), 1.0);
    if (features.empty())
    {
        Logger::LogWarning("The profiled run holds no counters the profile-based searcher understands; drawing the next batch at random");
    }
    else
    {
        ++m_Profiled;
        m_History.emplace_back(m_BestOfBatch, features);
        const auto bottlenecks = AnalyzeBottlenecks(features, m_Options);
        const auto changes = ComputeChanges(bottlenecks, features, m_Options);
        std::vector<::ktt::Features> estimates;
        std::vector<const ::ktt::Features*> pointers;
        const ::ktt::Features* current = &features;
        if (!m_Knowledge.empty())
        {
            current = KnowledgeOf(m_BestOfBatch);
            for (const auto& candidate : candidates)
            {
                pointers.push_back(KnowledgeOf(candidate));
            }
            // Counts of the knowledge base scale to this search's launches when the launch of the
            // profiled configuration and of every candidate the base holds is known.
            std::vector<double> threads(candidates.size() + 1, 0.0);
            bool known = current != nullptr && GetLaunchSize(m_BestOfBatch, threads.back());
            for (size_t i = 0; known && i < candidates.size(); ++i)
            {
                known = pointers[i] == nullptr || GetLaunchSize(candidates[i], threads[i]);
            }
            if (known)
            {
                estimates.reserve(candidates.size() + 1);
                for (size_t i = 0; i < candidates.size(); ++i)
                {
                    if (pointers[i] != nullptr)
                    {
                        estimates.push_back(ScaledToLaunch(*pointers[i], threads[i], m_Options.computeUnits));
                        pointers[i] = &estimates.back();
                    }
                }
                estimates.push_back(ScaledToLaunch(*current, threads.back(), m_Options.computeUnits));
                current = &estimates.back();
            }
        }
        else
        {
            estimates.reserve(candidates.size());
            for (const auto& candidate : candidates)
            {
                estimates.push_back(Estimate(candidate));
            }
            for (const auto& estimate : estimates)
            {
                pointers.push_back(&estimate);
            }
        }
        // Without the profiled configuration in the knowledge base every candidate weighs the same.
        if (current != nullptr)
        {
            weights = ScoreCandidates(changes, *current, pointers, m_Options);
        }
    }

    m_Batch.clear();
    if (candidates.size() > m_Options.batchSize)
    {
        while (m_Batch.size() < m_Options.batchSize)
        {
            m_Batch.push_back(candidates[Draw(weights)]);
        }
    }
    else
    {
        m_Batch = candidates;
    }
}

bool ProfileBasedSearcher::CalculateNextConfiguration(const KernelResult& previousResult)
{
    // CalculateNextConfiguration of the reference's searcher (ProfileBasedSearcher.h:728-858).
    if (previousResult.IsValid() && (!m_HaveBest || static_cast<double>(previousResult.GetKernelDuration()) < m_BestDuration))
    {
        m_BestDuration = static_cast<double>(previousResult.GetKernelDuration());
        m_BestOfBatch = m_Current;
        m_HaveBest = true;
    }

    if (m_Batch.empty())
    {
        if (!m_HaveBest)
        {
            // Every configuration of the batch failed: a new random batch.
            FillRandomBatch(std::min<uint64_t>(m_Options.batchSize, GetUnexploredConfigurationsCount()));
        }
        else if (m_BestDuration >= 0.0)
        {
            // Run the batch's fastest configuration again, profiled.
            m_Current = m_BestOfBatch;
            m_BestDuration = -1.0;
            SetProfiling(true);
            return true;
        }
        else
        {
            NextBatch(previousResult);
            m_HaveBest = false;
            SetProfiling(false);
        }
    }
    if (m_Batch.empty())
    {
        return false;
    }
    m_Current = m_Batch.front();
    m_Batch.erase(m_Batch.begin());
    return true;
}

} // namespace ktt
