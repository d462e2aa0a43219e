/* ktt_hip.h - C ABI of the MI355X (gfx950) direct Coulomb-summation library.
 *
 * The reference exports C++ classes only (Source/KttPlatform.h:22-24: KTT_API is empty on
 * Linux) and has no C ABI. Two layers are offered here:
 *   1. The ktt C++ API itself (ktt_amd/csrc/ktt/Ktt.h, ktt::Tuner with ComputeApi::HIP), the
 *      drop-in for Source/Tuner.h; the functions in section 4 drive it for callers that can
 *      only bind C (ctypes, cgo, JNI).
 *   2. Direct entry points for the hot path (section 3): one launch of a Coulomb kernel on
 *      device buffers, compiled through hipRTC with the configuration given as defines.
 *
 * Conventions. Every function returns 0 on success and a negative value on failure; the
 * message of the last failure on the calling thread is available from ktt_hip_last_error().
 * Atoms are float4 records {x, y, z, q} (Examples/CoulombSum3d/CoulombSum3d.cpp:95-98).
 * Grids are float32, z-major: index z*G*G + y*G + x (Examples/CoulombSum3d/CoulombSum3d.cu:10-11);
 * grid point (x, y, z) sits at (h*x, h*y, h*z). All kernels ADD into the output grid ("+="),
 * as the reference kernels do (CoulombSum3d.cu:46, CoulombSum2d.cl:945); clear it first for V.
 * A configuration string is a comma-separated list NAME=VALUE of kernel defines, or NULL for
 * the tuned default; unknown names are passed to the compiler unchanged.
 * A stream argument is a hipStream_t (NULL: the null stream). kernel_ms, when not NULL,
 * receives the kernel time measured with HIP events on that stream (the call then waits).
 */
#ifndef KTT_HIP_H
#define KTT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- 1. Library and device ---------------------------------------------------------------- */

/* ABI version of this header (major * 100 + minor). */
int ktt_hip_abi_version(void);

/* Message of the last failure on this thread ("" if none). */
const char* ktt_hip_last_error(void);

/* Number of visible HIP devices (0 without a GPU; not an error). */
int ktt_hip_device_count(int* count);

/* Name, gfx architecture, compute units and global memory of the current device. */
int ktt_hip_device_info(char* name, size_t name_length, char* architecture, size_t architecture_length,
                        int* compute_units, size_t* global_memory);

/* Directory the kernel sources are read from (defaults to the kernels/ directory next to the
 * library; overridden by the environment variable KTT_HIP_KERNEL_DIR). */
const char* ktt_hip_kernel_directory(void);

/* hipRTC compilation of one kernel of one source file for gfx950, without loading it; needs
 * no GPU. log (optional) receives the compiler log. Replaces nothing in the reference: it is
 * the build check of the kernels. */
int ktt_hip_compile_check(const char* source_file, const char* kernel_name, const char* config,
                          char* log, size_t log_length);

/* ---- 2. Device memory and timing plumbing (callers without their own HIP binding) --------- */

int ktt_hip_malloc(void** pointer, size_t bytes);
int ktt_hip_free(void* pointer);
int ktt_hip_memcpy_htod(void* device_destination, const void* host_source, size_t bytes);
int ktt_hip_memcpy_dtoh(void* host_destination, const void* device_source, size_t bytes);
/* Zero-fills bytes of device memory with a kernel (no hipMemset node, safe to capture). */
int ktt_hip_zero(void* device_pointer, size_t bytes, void* stream);
/* Device-to-device copy with a plain 16-byte load/store kernel (16-byte aligned pointers, bytes a
 * multiple of 16); its profiler counters calibrate the HBM traffic figures of the benchmark. */
int ktt_hip_copy(void* device_destination, const void* device_source, size_t bytes, void* stream);
/* Issue rates of the current device's vector ALU, in lane operations per second (instructions
 * x 64 lanes), from probe kernels of 8 independent chains per thread at 8 waves per SIMD:
 * v_fma_f32 with VGPR operands (dual issue across waves), v_fma_f32 with an SGPR operand
 * (single issue), v_rsq_f32 alone, and the Coulomb inner loop's mix of 2 FMA + 1 v_rsq_f32 in
 * each wave. Each rate is marginal: the probes run `iterations` and 2 x `iterations` rounds and
 * the rate is the extra work over the extra time, free of the launch's fixed costs.
 * Replaces nothing in the reference: it calibrates the transcendental weight of the CDNA
 * bottleneck analysis (the CUDA one reads the XU pipe's utilisation from CUPTI,
 * Source/Api/Searcher/ProfileBasedSearcher.h:123-340). Each output optional. */
int ktt_hip_issue_rates(int iterations, double* fma_rate, double* fma_scalar_rate, double* rsq_rate, double* mix_rate);
int ktt_hip_synchronize(void);
int ktt_hip_stream_create(void** stream);
int ktt_hip_stream_destroy(void* stream);
/* Event pair for timing a region of work on one stream. */
int ktt_hip_event_create(void** event);
int ktt_hip_event_destroy(void* event);
int ktt_hip_event_record(void* event, void* stream);
int ktt_hip_event_elapsed_ms(void* start, void* end, float* milliseconds);

/* ---- 3. Coulomb hot path on device buffers ----------------------------------------------- */

/* Forms of the 3D computation. */
#define KTT_COULOMB_PLAIN 0        /* FP32 accumulation, as the reference kernels */
#define KTT_COULOMB_COMPENSATED 1  /* tile-relative displacements, two-sum per atom tile,
                                      FP64 merge of atom-block partials in the same launch */

/* CoulombSum3d, one launch over the whole grid: energy_grid[G^3] += V.
 * Replaces the launch of directCoulombSum, Examples/CoulombSum3d/CoulombSum3d.cu:1-47
 * (driven by Examples/CoulombSum3d/CoulombSum3d.cpp:117-237).
 * atoms: float4[n_atoms] on the device. Config names: WORK_GROUP_SIZE_X, WORK_GROUP_SIZE_Y,
 * POINTS_X, ROWS_PER_THREAD, ATOM_TILE, INNER_UNROLL_FACTOR, ATOM_BLOCKS (atom blocks for every
 * tile), TAIL_SPLIT (compensated form with one atom block: 1, the default, splits the tiles of
 * the last, partial round of workgroups over atom blocks; 0 does not). Grids of any side whose
 * G^2 rows fit an int: when there are more than 65,535 row groups (G above 1,023 with the
 * compensated defaults, above 1,448 with the plain ones) they fold into the launch's x dimension,
 * and the point index is 64-bit (the reference kernel's int index overflows above G = 1,290). */
int ktt_coulomb3d(const float* atoms, int n_atoms, float spacing, int grid_size, float* energy_grid,
                  int form, const char* config, void* stream, float* kernel_ms);

/* The launch ktt_coulomb3d makes for these arguments, without launching: the grid of workgroups
 * (z = atom blocks; with the compensated form's tail split, the blocks the split tiles use), the
 * workgroup shape, the first split tile (tiles are numbered row group * x groups + x group, which
 * is y * grid[0] + x unless the row groups fold into x; equal to the tile
 * count when nothing is split) and the device scratch of the split in bytes. Compiles the kernel
 * (the split depends on its occupancy), so it needs the device. Each output optional. Replaces
 * nothing in the reference: the occupancy-driven split has no counterpart there. */
int ktt_coulomb3d_launch_plan(int n_atoms, int grid_size, int form, const char* config, unsigned int grid[3],
                              unsigned int block[3], int* split_from, size_t* scratch_bytes);

/* ktt_coulomb3d in the compensated form with the default configuration, allocating nothing.
 * No counterpart in the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc). The kernels are
 * compiled ahead of time (libktt_hip_noalloc.so, loaded on the first call) and their scratch
 * (the tail split's FP64-merged partials and arrival counters) is a static device array of
 * 64 MiB + 64 KiB per device; a launch needing more is rejected. The first call on a device
 * sets up (symbol address, occupancy): make it outside graph capture. Calls that overlap on
 * one device race on the shared scratch. grid_size in [1, 46340] (row groups fold as in
 * ktt_coulomb3d); atoms 16-byte aligned. */
int ktt_coulomb3d_compensated_noalloc(const float* atoms, int n_atoms, float spacing, int grid_size, float* energy_grid,
                                      void* stream);

/* The one-point-per-thread validation kernel (Examples/CoulombSum3d/CoulombSum3dReference.cu:1-28). */
int ktt_coulomb3d_reference(const float* atoms, int n_atoms, float spacing, int grid_size, float* energy_grid,
                            void* stream, float* kernel_ms);

/* CoulombSum2d, one slice: energy_grid[grid_y][grid_x] += V, where atoms[i].z holds the
 * squared z distance of atom i to the slice (Examples/CoulombSum2d/CoulombSum2d.cl:35-1026).
 * grid_x must be a multiple of WORK_GROUP_SIZE_X*POINTS_X and grid_y of
 * WORK_GROUP_SIZE_Y*ROWS_PER_THREAD. */
int ktt_coulomb2d(const float* atoms, int n_atoms, float spacing, int grid_x, int grid_y, float* energy_grid,
                  const char* config, void* stream, float* kernel_ms);

/* The 2D validation kernel, "=" store (Examples/CoulombSum2d/CoulombSum2dReference.cl:1-21). */
int ktt_coulomb2d_reference(const float* atoms, int n_atoms, float spacing, int grid_x, int grid_y,
                            float* energy_grid, void* stream, float* kernel_ms);

/* CoulombSum3dIterative with the reference launcher's semantics
 * (Examples/CoulombSum3dIterative/CoulombSum3dIterative.cpp:146-184): for every slice z the
 * host computes dz^2 = (h*z - Z_i)^2 in FP32 for all atoms, uploads the atom records with .z
 * replaced, and launches the slice kernel (CoulombSum3dIterative.cl:35) with zIndex = z.
 * host_atoms: float4[n_atoms] in host memory with the real z. wall_ms receives the host
 * time of the whole loop, kernel_ms the sum of the slice kernels' event times. */
int ktt_coulomb3d_iterative(const float* host_atoms, int n_atoms, float spacing, int grid_size, float* energy_grid,
                            const char* config, void* stream, float* kernel_ms, float* wall_ms);

/* The iterative example's validation kernel, one point per thread over the whole G^3 grid
 * with the atoms' real z, "+=" store
 * (Examples/CoulombSum3dIterative/CoulombSum3dIterativeReference.cl:1-26, launched over
 * (G, G, G) in workgroups of 16 x 16 x 1 as CoulombSum3dIterative.cpp:79-81 does); grid_size
 * a positive multiple of 16. */
int ktt_coulomb3d_iterative_reference(const float* atoms, int n_atoms, float spacing, int grid_size, float* energy_grid,
                                      void* stream, float* kernel_ms);

/* Fused slice loop: the slice kernel forms dz^2 on the device (atoms with their real z on
 * the device), one launch per slice, no host work between launches. */
int ktt_coulomb3d_slices(const float* atoms, int n_atoms, float spacing, int grid_size, float* energy_grid,
                         const char* config, void* stream, float* kernel_ms);

/* Launch geometry a configuration implies for a grid (for callers that size scratch or report
 * work per workgroup): grid of workgroups and workgroup shape. kind: 3 = CoulombSum3d,
 * 2 = CoulombSum2d / slice kernels (rows = grid_size), with the defaults of a slice of more
 * than 16,384 atoms; z is the configuration's ATOM_BLOCKS, before it is capped at the atom
 * tiles of a call. For kind 3, more than 65,535 row groups fold into x: grid[0] is then
 * fold * x groups and grid[1] the row groups / fold, rounded up. */
int ktt_coulomb_launch_shape(int kind, int grid_size, const char* config, unsigned int grid[3], unsigned int block[3]);

/* ---- 5. Batches of independent systems --------------------------------------------------- */

/* A batch of independent systems, each with its own atoms, cubic grid side and spacing. The
 * reference has no batched form (its examples run one system per tuner); this extends the
 * CoulombSum3d kernel's semantics (Examples/CoulombSum3d/CoulombSum3d.cu:1-47) to many systems.
 * host_atoms packs the float4 atoms of system 0, then system 1, ...; atom_counts, grid_sizes
 * and spacings have system_count entries. At most five distinct grid sides per batch; each
 * side is a tile-shape bucket. Config names: BUCKETb_WGX, BUCKETb_PX, BUCKETb_ROWS,
 * BUCKETb_TILE (b = 0..4, buckets in increasing grid side) and GROUPS_PER_CU. The batch owns
 * the device copies of the atoms and one packed output array of all grids. */
int ktt_coulomb_batch_create(const float* host_atoms, const int* atom_counts, const int* grid_sizes, const float* spacings,
                             int system_count, const char* config, void** batch);
int ktt_coulomb_batch_destroy(void* batch);

/* Device pointer and size of the packed grids, each system's first point in it (grid_offsets has
 * system_count entries), the number of work items, the total grid-point x atom interactions and
 * the tile shapes in use as NAME=VALUE,... (each output optional). */
int ktt_coulomb_batch_info(void* batch, float** grids, size_t* grid_points, long long* grid_offsets, int* items,
                           long long* interactions, char* shapes, size_t shapes_length);

/* Clears all grids of the batch (zero-fill kernel). */
int ktt_coulomb_batch_zero(void* batch, void* stream);

/* Adds V of every system to its grid: one persistent launch in which workgroups take
 * (system, tile) items from a device-side queue, compensated FP32. use_graph != 0 replays a HIP
 * graph (queue reset + persistent kernel) captured on the first call. */
int ktt_coulomb_batch_run(void* batch, int use_graph, void* stream, float* kernel_ms);

/* Baseline: one ktt_coulomb3d launch per system (form KTT_COULOMB_PLAIN or _COMPENSATED), all
 * on one stream; kernel_ms is the event time around the whole sequence. */
int ktt_coulomb_batch_run_per_system(void* batch, int form, void* stream, float* kernel_ms);

/* Tunes the tile shape of every bucket with ktt::Tuner on ComputeApi::HIP (a composite kernel
 * and launcher over the bucket's systems; searcher as in section 4), bucket after bucket, and
 * keeps the winners in the batch. results_file (optional) is a path prefix: the KTT JSON
 * results of bucket b go to <prefix>_bucket<b>.json and the winners to <prefix>_winners.json.
 * best_config (optional) receives the winning shapes as a configuration string. */
int ktt_coulomb_batch_tune(void* batch, int searcher, int max_configurations_per_bucket, const char* results_file,
                           char* best_config, size_t best_config_length);

/* ---- 4. ktt::Tuner on ComputeApi::HIP (host buffers, full framework path) ---------------- */

/* Searchers accepted by the tuning entry points. */
#define KTT_SEARCHER_DETERMINISTIC 0
#define KTT_SEARCHER_RANDOM 1
#define KTT_SEARCHER_MCMC 2

/* Tunes the CoulombSum3d kernel with ktt::Tuner(0, 0, ComputeApi::HIP) on host atom data,
 * mirroring Examples/CoulombSum3d/CoulombSum3d.cpp:117-237 with this library's parameter
 * space: every configuration is compiled by hipRTC, timed with HIP events and validated
 * against the reference kernel with SideBySideComparison at the given tolerance.
 * max_configurations <= 0 explores the whole space. results_file (optional, without
 * extension) receives the KTT JSON results; best_config (optional) the best configuration
 * as NAME=VALUE,...; best_ms (optional) its kernel time; valid_count (optional) the number
 * of configurations with status Ok; tested_count (optional) the number tested.
 * energy_grid (optional, host, G^3 floats) receives the output of the best configuration. */
int ktt_coulomb3d_tune(const float* host_atoms, int n_atoms, float spacing, int grid_size, int searcher,
                       int max_configurations, double tolerance, const char* results_file,
                       char* best_config, size_t best_config_length, double* best_ms,
                       int* valid_count, int* tested_count, float* energy_grid);

/* Runs a KTT JSON tuning script whose KernelSpecification.Language is "HIP" (the job of the
 * reference's KttTuningLauncher, TuningLoader/TuningLauncher.cpp, with the schema of
 * TuningLoader/TuningSchema.h). Script expressions are evaluated without Python. Relative
 * file names are relative to the script. output_file (optional, no extension) overrides
 * General.OutputFile; configuration_limit > 0 replaces the script's Budget with a
 * ConfigurationCount. Outputs, each optional: best_config (NAME=VALUE,...), space_size,
 * tested, valid (configurations that ran and passed validation) and best_ms (kernel time of
 * the best configuration). */
int ktt_run_tuning_script(const char* script_path, const char* output_file, long long configuration_limit,
                          char* best_config, size_t best_config_length, long long* space_size, long long* tested,
                          long long* valid, double* best_ms);

/* Number of configurations in a script's ConfigurationSpace (parameters and conditions);
 * needs no device. */
int ktt_tuning_script_space(const char* script_path, long long* size);

/* Evaluates a script generator expression (a Python expression in the element index i, as in
 * an argument's "Generator" DataSource) for i = 0 .. count-1 into values. */
int ktt_script_generate(const char* expression, long long count, double* values);

/* Dynamic tuning run of the compensated CoulombSum3d kernel with ktt::Tuner::TuneIteration
 * (replaces the online-tuning loop a user writes around Source/Tuner.h:662-698, whose
 * iteration is Source/TuningRunner/TuningRunner.cpp:84-128). The atoms are copied to the
 * device once. Step s (0-based) moves every atom by a N(0, sigma^2 I) displacement drawn from
 * (seed, s, atom index) on the device, clears the G^3 map and recomputes it. The first
 * min(tuning_budget, space) steps each run the searcher's next configuration
 * (tuning_budget <= 0: the whole space); every later step runs Tuner::GetBestConfiguration
 * (Source/Tuner.h:776-782). Outputs, each optional: step_kernel_ms[steps] (Coulomb kernel
 * time, -1 for a configuration that failed), step_wall_ms[steps] (host time of the step,
 * compilation included), step_tuned[steps] (1 for a tuning step), best_config
 * (NAME=VALUE,...), tested_count, space_size, final_atoms (4 * n_atoms floats) and
 * final_grid (G^3 floats, the last step's map). results_file (optional, no extension)
 * receives the tuning steps as KTT JSON results. */
int ktt_coulomb3d_dynamic(const float* host_atoms, int n_atoms, float spacing, int grid_size, int steps,
                          float sigma, unsigned long long seed, int searcher, int tuning_budget,
                          const char* results_file, double* step_kernel_ms, double* step_wall_ms,
                          int* step_tuned, char* best_config, size_t best_config_length,
                          int* tested_count, int* space_size, float* final_atoms, float* final_grid);

/* ---- 6. Searchers over saved results; hardware counters ------------------------------------ */

/* Simulated tuning over saved KTT JSON results (Tuner::SimulateTuning, Source/Tuner.h:700-712,
 * Source/TuningRunner/TuningRunner.cpp:131-201), without a device: the configuration space is
 * the set of configurations the file holds and each proposal is answered with its saved
 * result. searcher: "Deterministic", "Random", "MCMC" or "ProfileBased"
 * (Source/Tuner.h:738-748); knowledge_file (optional) holds KTT JSON results with counters,
 * the profile-based searcher's knowledge base; compute_units sizes its analysis (0: 256).
 * seed 0 draws a seed. max_steps <= 0 runs until the searcher ends. step_ms[capacity] receives
 * each step's total duration in milliseconds (-1 for an invalid result); steps the count. */
int ktt_simulate_search(const char* results_file, const char* searcher, const char* knowledge_file,
                        unsigned long long seed, long long max_steps, unsigned compute_units,
                        double* step_ms, long long capacity, long long* steps);

/* Enables hardware counters for profiled runs (ktt::EnableHipCounters; replaces the CUPTI
 * set-up of Source/ComputeEngine/Cuda/Cupti/CuptiInstance.cpp). Must precede the process's
 * first HIP call; KTT_HIP_COUNTERS=1 in the environment does the same when the library loads. */
int ktt_hip_enable_counters(void);

/* ---- 7. N-body step (sibling kernel of the Coulomb sum) ------------------------------------ */

/* One integration step of n bodies on the device (Examples/Nbody/Nbody.cu:271-410, kernel
 * nbody_kernel): a_i = sum_j m_j d_ij / (|d_ij|^2 + softening_sqr)^(3/2),
 * p' = p + dt v + damping dt^2 a, v' = v + dt a. body_info and new_body_info are n float4
 * (x, y, z, mass); velocity and new_velocity n float4 (vx, vy, vz, 0). config as in section 3
 * (WORK_GROUP_SIZE_X/Y, BODIES_PER_THREAD, ATOM_TILE, INNER_UNROLL_FACTOR; NULL: 64, 4, 4, 256, 8); n
 * must be a multiple of WORK_GROUP_SIZE_X * BODIES_PER_THREAD and ATOM_TILE * WORK_GROUP_SIZE_Y
 * (1,024 for the defaults). */
int ktt_nbody_step(const float* body_info, const float* velocity, int n, float time_delta, float damping,
                   float softening_sqr, float* new_body_info, float* new_velocity, const char* config, void* stream,
                   float* kernel_ms);

/* The same step with the reference kernel (Examples/Nbody/NbodyReference.cu:40-75); n a
 * multiple of 64. */
int ktt_nbody_step_reference(const float* body_info, const float* velocity, int n, float time_delta, float damping,
                             float softening_sqr, float* new_body_info, float* new_velocity, void* stream,
                             float* kernel_ms);

/* The N-body example's tuning run (Examples/Nbody/Nbody.cpp:146-256) through ktt::Tuner on
 * host data: the kernel's space, both outputs validated against the reference kernel with
 * SideBySideComparison at tolerance (the example uses 0.001). Searcher and outputs as for
 * ktt_coulomb3d_tune; best_ms is the best configuration's kernel time. */
int ktt_nbody_tune(const float* host_body_info, const float* host_velocity, int n, float time_delta, float damping,
                   float softening_sqr, int searcher, int max_configurations, double tolerance,
                   const char* results_file, char* best_config, size_t best_config_length, double* best_ms,
                   int* valid_count, int* tested_count);

#ifdef __cplusplus
}
#endif

#endif /* KTT_HIP_H */
