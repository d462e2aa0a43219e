/** @file ProfileBasedSearcher.h
  * Searcher that steers the search by the bottlenecks hardware counters show
  * (Source/Api/Searcher/ProfileBasedSearcher.h:10-897; J. Filipovic et al., "Using hardware
  * performance counters to speed up autotuning convergence on GPUs", JPDC 160, 2021).
  *
  * The search runs in batches. The configurations of a batch run unprofiled; the fastest is run
  * once more with profiling. Its counters are turned into bottleneck scores (memory levels,
  * local memory, occupancy, parallelism, instruction mix, issue), the bottlenecks into the
  * direction and weight in which each counter should change, and every candidate (neighbours
  * of the profiled configuration within 2 changed parameters, plus random ones) is scored by
  * how its counters differ from the profiled configuration's in those directions. The next
  * batch is drawn from the candidates with probability growing with the score.
  *
  * The reference predicts a candidate's counters with a pickled scikit-learn model. Here they
  * come without a trained model, from saved results that hold counters:
  *   * a knowledge base: KTT JSON results of a profiled search of the same kernel space
  *     (another input size or another GPU), used for the configurations it holds, as the
  *     reference's exact-mapping variant does (scoreTuningConfigurationsExact). Its CDNA counts
  *     that grow with launched threads (instructions, quad-cycles, cache requests) are scaled by
  *     the ratio of a configuration's threads in this search's launch (Searcher::GetLaunchSize)
  *     to its threads in the base's, and its device-wide utilisations by the ratio of the compute
  *     units the two launches occupy (when computeUnits is known), so a base profiled on a
  *     smaller problem, whose launches overhang it by shape, carries over;
  *   * otherwise the configurations profiled so far in this search: a candidate takes the
  *     counters of the nearest profiled configuration (fewest differing parameter values).
  *
  * Two counter sets are understood: the CDNA counters that the HIP engine collects by default
  * (FETCH_SIZE, VALUBusy, SQ_INSTS_*, ...) and the CUDA compute capability 7.x+ metrics of
  * the reference's CUPTI interface, so that the reference's saved profiled searches can be
  * replayed with Tuner::SimulateTuning.
  */
#pragma once

#include <cstdint>
#include <map>
#include <random>
#include <string>
#include <vector>

#include <KernelResult.h>
#include <Searcher.h>

namespace ktt
{

class Tuner;

struct ProfileSearcherOptions
{
    size_t batchSize = 5;              ///< configurations run before the fastest is profiled
    size_t neighbourCount = 100;       ///< neighbours of the profiled configuration scored
    size_t randomCount = 10;           ///< random configurations scored beside them
    uint64_t neighbourDistance = 2;    ///< parameters a neighbour may differ in
    double exponent = 8.0;             ///< weight = normalised score ^ exponent
    double instructionThreshold = 0.7; ///< instruction-class utilisation that counts as a bottleneck (0.5: compute bound)
    double cutoff = -0.25;             ///< candidates scoring below are dropped (exact variant)
    uint32_t computeUnits = 0;         ///< CUs / SMs; 0 takes the tuner's device
    uint32_t lanesPerUnit = 64;        ///< lanes per CU / cores per SM
    double memoryBandwidth = 8.0e12;   ///< peak HBM bytes per second (CDNA counters)
    double l2Bandwidth = 1.6e13;       ///< peak L2 bytes per second (CDNA counters)
    double l1Bandwidth = 0.0;          ///< peak vector L1 bytes per second; 0: CUs x 64 B x clock
    double clockHz = 2.4e9;            ///< engine clock used when l1Bandwidth is 0
    uint64_t seed = 0;                 ///< 0: from KTT_SEED or random_device
};

class ProfileBasedSearcher : public Searcher
{
public:
    /** Searcher that switches profiling on `tuner` on for the profiled runs and off after them.
     *  `knowledge` holds saved results with counters (may be empty). */
    ProfileBasedSearcher(Tuner& tuner, const ProfileSearcherOptions& options = {}, const std::vector<KernelResult>& knowledge = {});
    /** Searcher without a tuner: for replays of saved profiled results. */
    explicit ProfileBasedSearcher(const ProfileSearcherOptions& options = {}, const std::vector<KernelResult>& knowledge = {});

    void OnInitialize() override;
    void OnReset() override;
    bool CalculateNextConfiguration(const KernelResult& previousResult) override;
    KernelConfiguration GetCurrentConfiguration() const override;

    /** Number of profiled runs analysed so far. */
    size_t GetProfiledCount() const { return m_Profiled; }

    /** Features of one profiled run: the counters the analysis reads, keyed by name, with the
     *  pseudo-counters "Global size" (threads) and "Local size" (threads per group). Empty when
     *  the run holds no counter set this searcher understands. */
    static std::map<std::string, double> Features(const ComputationResult& computation);
    /** Bottleneck scores in [0, 1] of a feature set (names as in the reference: bnDRAMRead, bnL2Read,
     *  bnTex, bnLocal, bnSMRead, bnMPparal, bnGparal, bnThreads, bnSP, bnInstIssue, ...; on CDNA
     *  also bnVALU and, from the dual-issue counters, bnPairing). */
    static std::map<std::string, double> AnalyzeBottlenecks(const std::map<std::string, double>& features,
        const ProfileSearcherOptions& options);
    /** Signed importance of changing each feature: negative asks for less, positive for more. */
    static std::map<std::string, double> ComputeChanges(const std::map<std::string, double>& bottlenecks,
        const std::map<std::string, double>& features, const ProfileSearcherOptions& options);
    /** Normalised selection weights of candidates (exact variant: candidates without features score 0). */
    static std::vector<double> ScoreCandidates(const std::map<std::string, double>& changes, const std::map<std::string, double>& current,
        const std::vector<const std::map<std::string, double>*>& candidates, const ProfileSearcherOptions& options);

private:
    Tuner* m_Tuner;
    ProfileSearcherOptions m_Options;
    std::map<std::string, std::map<std::string, double>> m_Knowledge;  // by configuration key
    std::vector<std::pair<KernelConfiguration, std::map<std::string, double>>> m_History;
    std::mt19937_64 m_Random;

    std::vector<KernelConfiguration> m_Batch;
    KernelConfiguration m_Current;
    KernelConfiguration m_BestOfBatch;
    bool m_HaveBest;
    double m_BestDuration;   // kernel duration of the batch's best; negative once its profiled run is proposed
    size_t m_Profiled;

    void SetProfiling(const bool on);
    void FillRandomBatch(const size_t count);
    void NextBatch(const KernelResult& profiled);
    const std::map<std::string, double>* KnowledgeOf(const KernelConfiguration& configuration) const;
    std::map<std::string, double> Estimate(const KernelConfiguration& configuration) const;
    size_t Draw(std::vector<double>& weights);
};

} // namespace ktt
