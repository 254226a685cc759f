/*
 * rfmf.h -- C ABI of the MI355X-native RFMF annealer (librfmf.so).
 *
 * Plain pointers and sizes only.  Every entry point returns an rfmf_status;
 * rfmf_last_error() gives the calling thread's last detail message.  Device
 * memory is owned by the library: the one-shot entry points allocate and free
 * inside the call, as the reference's annealGpu does; contexts keep graphs
 * and chain state resident in HBM between calls.
 *
 * Reference interfaces are cited as path:line in the reference checkout.
 */
#ifndef RFMF_H
#define RFMF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rfmf_status {
    RFMF_OK = 0,
    RFMF_ERR_INVALID_ARG = 1,   /* sizes, null pointers, t0 <= 0, P out of range */
    RFMF_ERR_UNSUPPORTED = 2,   /* max weighted degree >= 2^62, P > 4096 (group tables beyond LDS) */
    RFMF_ERR_HIP = 3,           /* HIP runtime error (message has the detail) */
    RFMF_ERR_NO_DEVICE = 4,
    RFMF_ERR_OUT_OF_MEMORY = 5,
    RFMF_ERR_BETA_CAP = 6,      /* more than 33,000,000 beta runs */
    RFMF_ERR_STEP_CAP = 7,      /* more than 10^12 sweeps per chain */
    RFMF_ERR_IO = 8,            /* unreadable or malformed instance file */
    RFMF_ERR_NOT_FOUND = 9,     /* unknown graph / request id */
    RFMF_ERR_ADMISSION = 10,    /* request larger than the server's HBM budget */
    RFMF_ERR_STATE = 11         /* call not valid in the object's state */
} rfmf_status;

const char* rfmf_status_string(rfmf_status s);
const char* rfmf_last_error(void);
/* Library build identifier ("rfmf <version> gfx950"). */
const char* rfmf_version(void);

/* ---------------------------------------------------------------------- */
/* Graph views and host helpers                                            */
/* ---------------------------------------------------------------------- */

/* Dense CSR view: rows sorted by neighbour id, each undirected edge stored
 * twice.  Replaces anneallib::AnnealCsr (src/immutable_graph.hpp:154-157)
 * as the kernel sees it through DevCsr (src/anneal_gpu.cuh:23-28). */
typedef struct rfmf_csr {
    const uint32_t* offsets;      /* n + 1 */
    const int32_t* edge_to;       /* m2 */
    const int32_t* edge_weight;   /* m2 */
    int64_t n;
    int64_t m2;
} rfmf_csr;

/* Loader statistics, anneallib::ReadStats (src/read_gset.hpp:11-18). */
typedef struct rfmf_read_stats {
    int64_t num_vertices;
    int64_t num_edges;
    int64_t zero_dropped;
    int32_t float_weights;
    int32_t pad;
    double scale;
    int64_t max_abs_weighted_degree;
} rfmf_read_stats;

/* Owning host CSR (the frozen AnnealCsr built by the load path). */
typedef struct rfmf_graph rfmf_graph;

/* read_instance_from_file + removeDegreeZeroVertices + AnnealCsr
 * (src/read_gset.hpp:32-34, src/host_common.cpp:184-194,
 * src/immutable_graph.hpp:154-157). */
rfmf_status rfmf_graph_read_gset(const char* path, double forced_scale, rfmf_graph** out,
                                 rfmf_read_stats* stats);
/* Graph::addEdge over an edge list in order (first weight wins, zero weights
 * kept as given), then AnnealCsr (src/graph.cpp:11-17). */
rfmf_status rfmf_graph_from_edges(const int64_t* u, const int64_t* v, const int64_t* w,
                                  int64_t m, rfmf_graph** out);
void rfmf_graph_free(rfmf_graph* g);
/* View valid while the graph lives. */
rfmf_status rfmf_graph_view(const rfmf_graph* g, rfmf_csr* view);
/* Dense id -> original id (n entries). */
const int64_t* rfmf_graph_original_ids(const rfmf_graph* g);

/* anneallib::solver::cutValueCsr (src/max_cut.hpp:16). */
int64_t rfmf_cut_value(const rfmf_csr* g, const int8_t* spin);
/* maxWeightedDegree (src/host_common.h:111-112). */
int64_t rfmf_max_weighted_degree(const rfmf_csr* g);
/* annealScoreSafeInt32 (src/host_common.h:106-109): 1 if the packed int32
 * score path is safe, else 0. */
int32_t rfmf_score_safe_int32(int64_t total_steps, double t0, int64_t delta_bound);
/* betaFixedForStep (src/host_common.h:73-76). */
int32_t rfmf_beta_fixed_for_step(uint64_t step, double t0);
/* annealTotalSteps (src/host_common.h:67-68). */
rfmf_status rfmf_total_steps(const uint64_t* iters, int64_t n, int64_t* total);
/* Rng(raw seed) + generateRandomSpins (src/rng.hpp:11-58,
 * src/host_common.cpp:40-45). */
rfmf_status rfmf_initial_spins(uint64_t raw_seed, int64_t n, int8_t* out);

/* Run-length beta tables: run r of a table covers 0-based steps
 * [run_end[r-1], run_end[r]) with Q20 value run_val[r]; chain c uses the table
 * starting at chain_run_off[c] (src/host_common.h:82-94). */
typedef struct rfmf_beta_tables {
    const int64_t* run_end;
    const int32_t* run_val;
    const int64_t* chain_run_off;
    int64_t n_runs;
} rfmf_beta_tables;

/* buildBetaRunTables (src/host_common.h:101-104).  Call with run_end/run_val
 * NULL to get the run count only.  chain_run_off receives `chains` entries. */
rfmf_status rfmf_beta_tables_log(const double* t0s, int64_t chains, int64_t total_steps,
                                 int64_t* run_end, int32_t* run_val, int64_t cap,
                                 int64_t* n_runs, int64_t* chain_run_off, int64_t* n_tables);
/* Exact-integer linear schedule beta(s) = b0 + ((b1-b0)*(s-1))/(total-1), Q20,
 * one table shared by all chains (no reference counterpart; the reference's
 * only schedule is the log one). */
rfmf_status rfmf_beta_tables_linear(int32_t b0, int32_t b1, int64_t total_steps,
                                    int64_t* run_end, int32_t* run_val, int64_t cap,
                                    int64_t* n_runs);

/* ---------------------------------------------------------------------- */
/* One-shot annealing: drop-in for gpu3::annealGpu                         */
/* ---------------------------------------------------------------------- */

/* gpu3::CorrMode (src/anneal_gpu.cuh:42-45).  Both are output-invariant; the
 * HIP kernels use one corr formulation for both. */
typedef enum rfmf_corr_mode { RFMF_CORR_WALK = 0, RFMF_CORR_WARPWALK = 1 } rfmf_corr_mode;

/* gpu3::LaunchConfig (src/anneal_gpu.cuh:52-58). */
typedef struct rfmf_launch_cfg {
    int32_t corr;      /* rfmf_corr_mode */
    int32_t device;    /* HIP device ordinal */
    int32_t s_width;   /* 32 or 64: accepted; the state word width is chosen from B */
    int32_t reserved;
} rfmf_launch_cfg;

/* gpu3::annealGpu (src/anneal_gpu.cuh:76-88): C chains from initial_spins
 * (C x n, +/-1) with one t0 and one solver seed each, phases/iters shared.
 * Outputs end_spins / best_spins (C x n), the kernel-only time, and the
 * incrementally tracked best / final cuts (may be NULL).  n == 0 or
 * n_phases == 0 returns the initial state without device work, as the
 * reference does (src/anneal_kernels.cu:856-868). */
rfmf_status rfmf_anneal(const rfmf_csr* g, int64_t max_weighted_deg, const int8_t* initial_spins,
                        int64_t chains, const double* t0s, const uint64_t* phases,
                        const uint64_t* iters, int64_t n_phases, const uint64_t* solver_seeds,
                        const rfmf_launch_cfg* cfg, int8_t* end_spins, int8_t* best_spins,
                        float* kernel_ms, int64_t* kernel_best_vals, int64_t* kernel_final_vals);

/* rfmf_anneal with caller-built beta tables (any monotone or arbitrary Q20
 * schedule) instead of t0s.  Each table must cover sum(iters) steps. */
rfmf_status rfmf_anneal_tables(const rfmf_csr* g, int64_t max_weighted_deg,
                               const int8_t* initial_spins, int64_t chains,
                               const rfmf_beta_tables* beta, const uint64_t* phases,
                               const uint64_t* iters, int64_t n_phases,
                               const uint64_t* solver_seeds, const rfmf_launch_cfg* cfg,
                               int8_t* end_spins, int8_t* best_spins, float* kernel_ms,
                               int64_t* kernel_best_vals, int64_t* kernel_final_vals);

/* rfmf_anneal_tables without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, checks and
 * results as rfmf_anneal_tables, bit for bit; no hipMalloc, hipFree, pinned
 * or heap allocation on its path.  All device memory comes from one fixed
 * 64 MiB array per device, held until the process ends; a call whose
 * CSR, spins, chain state, tables and staging need more returns
 * RFMF_ERR_UNSUPPORTED (about C * (3 n + n / 8) + 16 m2 bytes for 8-bit
 * state words) before any device work.  Calls on one device are serialized
 * by a per-device lock and share the array: each call overwrites what the
 * previous one left there, so calls that overlap on one device (from
 * threads that bypass the lock, e.g. other libraries writing the same
 * memory) would race.  Runs on the null stream and waits for its work. */
rfmf_status rfmf_anneal_tables_noalloc(const rfmf_csr* g, int64_t max_weighted_deg,
                                       const int8_t* initial_spins, int64_t chains,
                                       const rfmf_beta_tables* beta, const uint64_t* phases,
                                       const uint64_t* iters, int64_t n_phases,
                                       const uint64_t* solver_seeds, const rfmf_launch_cfg* cfg,
                                       int8_t* end_spins, int8_t* best_spins, float* kernel_ms,
                                       int64_t* kernel_best_vals, int64_t* kernel_final_vals);

/* ---------------------------------------------------------------------- */
/* Resident device context: graphs in HBM, multi-graph batches, serving.   */
/* No reference counterpart (the reference allocates per call and runs one */
/* graph per launch); these serve callers that batch many graphs or serve  */
/* a request stream.                                                       */
/* ---------------------------------------------------------------------- */

typedef struct rfmf_ctx rfmf_ctx;

rfmf_status rfmf_ctx_create(int32_t device, rfmf_ctx** out);
void rfmf_ctx_destroy(rfmf_ctx* ctx);
/* Copies the CSR into HBM; returns a graph id valid until released. */
rfmf_status rfmf_ctx_upload_graph(rfmf_ctx* ctx, const rfmf_csr* g, int32_t* graph_id);
rfmf_status rfmf_ctx_release_graph(rfmf_ctx* ctx, int32_t graph_id);
/* Device properties seen by the context. */
rfmf_status rfmf_ctx_device_info(rfmf_ctx* ctx, int32_t* cu_count, int64_t* lds_per_block,
                                 int64_t* hbm_bytes, char* name, int32_t name_cap);

/* Copies `bytes` (rounded down to 16) between two fresh device buffers with a
 * plain streaming kernel `iters` times; returns the average kernel ms.  A
 * known-size kernel for calibrating HBM traffic counters and measuring the
 * achievable copy bandwidth (no reference counterpart). */
rfmf_status rfmf_ctx_copy_probe(rfmf_ctx* ctx, int64_t bytes, int32_t iters, float* avg_ms);

/* A job: `chains` chains of one resident graph with their own schedule. */
typedef struct rfmf_job {
    int32_t graph_id;
    int32_t pad;
    int64_t chains;
    const int8_t* initial_spins;   /* chains x n */
    const uint64_t* solver_seeds;  /* chains */
    rfmf_beta_tables beta;         /* chain_run_off has `chains` entries */
    const uint64_t* phases;        /* P per phase, shared by the job's chains */
    const uint64_t* iters;
    int64_t n_phases;
    int8_t* end_spins;             /* out, chains x n (may be NULL) */
    int8_t* best_spins;            /* out, chains x n (may be NULL) */
    int64_t* best_vals;            /* out, chains (may be NULL) */
    int64_t* final_vals;           /* out, chains (may be NULL) */
} rfmf_job;

/* Launch knobs; every one is output-invariant.  Zero fields mean "auto"
 * (the autotuner's cached choice, else the built-in default). */
typedef struct rfmf_tuning {
    int64_t slice_sweeps;       /* sweeps per launch (0: whole run in one launch) */
    int32_t grid_cap;           /* max resident workgroups (0: occupancy x CUs) */
    int32_t use_hip_graph;      /* 1: replay the slice launches from a captured hipGraph */
    int32_t threads_per_chain;  /* 128/256/512/1024 (2..16 wave64 per chain); 0: tuned or heuristic */
    int32_t autotune;           /* 1: time every thread count on one short slice of the live run
                                   first (exact: geometry never changes a result), cache the
                                   fastest per launch shape, run the rest with it */
    int32_t score_table;        /* 8-bit state words with 24-bit scores: 0 score from a 2048-entry
                                   LDS Gumbel table when that keeps as many chains resident, 1
                                   whenever the table fits in LDS, -1 never; output-invariant */
    int32_t state_in_hbm;       /* 0: a chain keeps its state in LDS for a slice when it fits there
                                   and works on it in place in HBM otherwise; 1: every chain in
                                   place in HBM; output-invariant */
    int32_t rows_in_lds;        /* graphs whose weights are all +1/-1 (n <= 32768): 0 copy the
                                   graph's offsets and 2-byte rows into LDS beside the chain when
                                   that keeps as many chains resident, 1 whenever they fit, -1
                                   never; output-invariant */
} rfmf_tuning;

/* Geometry of the context's last batch launch: threads per chain, grid,
 * state word bits, resident workgroups per CU, the autotuner's trial ms per
 * sweep for 128/256/512/1024 threads (0 where not tried), the score
 * arithmetic (0 int32, 1 int32 with 24-bit multiplies, 2 int64, 3 24-bit with
 * the LDS score table), whether the chain state stayed in HBM (1) or was
 * LDS-resident (0), and whether the graph's rows were read from LDS (1).  Any
 * output pointer may be NULL.  Replaces the reference's hard-coded V100
 * launch constants (src/tuning.hpp:8-18). */
rfmf_status rfmf_ctx_last_launch(rfmf_ctx* ctx, int32_t* threads, int32_t* grid, int32_t* word_bits,
                                 int32_t* occupancy, float* trial_ms_per_sweep, int32_t* score_mode,
                                 int32_t* state_in_hbm, int32_t* rows_in_lds);

/* Runs every job's chains to completion in persistent launches that pack all
 * jobs together; returns the kernel-only time of all launches. */
rfmf_status rfmf_batch_run(rfmf_ctx* ctx, const rfmf_job* jobs, int64_t n_jobs,
                           const rfmf_tuning* tuning, float* kernel_ms);

/* rfmf_batch_run cut into timed windows: window i advances every chain to
 * sweep min(total, step_bounds[i]) (bounds ascending) and window_ms[i]
 * receives its kernel-only time (every launch of the window, autotuner trials
 * included); after the last window the chains run to completion untimed.
 * Used to time a window of a run with the state already resident (e.g. after
 * warmup sweeps).  `tuning` may be NULL (defaults). */
rfmf_status rfmf_batch_run_windows(rfmf_ctx* ctx, const rfmf_job* jobs, int64_t n_jobs,
                                   const int64_t* step_bounds, int64_t n_windows, const rfmf_tuning* tuning,
                                   float* window_ms);

/* Step-wise control of a batch: admit jobs once (chain state resident in
 * HBM), advance every chain by `sweeps` per call (one persistent launch per
 * kernel variant; with tuning->use_hip_graph the launch is captured once and
 * replayed), and write the current state of every chain into the jobs'
 * output buffers at any point (final spins = current spins, best spins and
 * values = best so far).  The job output pointers must stay valid until
 * rfmf_run_destroy. */
typedef struct rfmf_run rfmf_run;
rfmf_status rfmf_run_create(rfmf_ctx* ctx, const rfmf_job* jobs, int64_t n_jobs, rfmf_run** out);
rfmf_status rfmf_run_advance(rfmf_run* run, int64_t sweeps, const rfmf_tuning* tuning, float* kernel_ms,
                             int32_t* finished);
rfmf_status rfmf_run_collect(rfmf_run* run);
void rfmf_run_destroy(rfmf_run* run);

/* HBM bytes a job of this shape keeps resident (CSR + chain state + tables),
 * for state words of up to 32 bits (max weighted degree below 2^30; the
 * server counts the 64-bit words of heavier graphs itself). */
int64_t rfmf_footprint_bytes(int64_t n, int64_t m2, int64_t chains, int64_t beta_runs,
                             int64_t n_phases);

/* ---------------------------------------------------------------------- */
/* Flips per sweep, picked by measurement.  No reference counterpart: the  */
/* reference takes the group-count ladder as an input (main.cu --phases).  */
/* ---------------------------------------------------------------------- */

#define RFMF_MAX_FLIPS_CANDIDATES 16

/* The ladder of a top group count: phases P, P/2, ..., 1 (P a power of two)
 * with `sweeps` split into equal phases (the first sweeps % len phases get one
 * more).  Call with phases/iters NULL for the length only. */
rfmf_status rfmf_flips_ladder(uint64_t top_p, int64_t sweeps, uint64_t* phases, uint64_t* iters,
                              int64_t cap, int64_t* n_phases);

typedef struct rfmf_flips_tune {
    int64_t sweeps;            /* per-chain sweep budget of the production run */
    int64_t pilot_chains;      /* chains per candidate; 0: 256 */
    uint64_t seed_base;        /* pilot chain c has raw seed seed_base + 1 + c */
    int32_t beta0_q20;         /* linear schedule beta0 -> beta1, Q20 */
    int32_t beta1_q20;
    const uint64_t* top_p;     /* candidate top group counts (powers of two) */
    int32_t n_candidates;      /* 1 .. RFMF_MAX_FLIPS_CANDIDATES */
    int32_t reference;         /* index of the candidate whose median sets the quality floor; -1: the best */
    double quality_tol;        /* relative tolerance on the median best cut, e.g. 0.001 */
} rfmf_flips_tune;

typedef struct rfmf_flips_report {
    int32_t chosen;            /* index into top_p */
    int32_t n_candidates;
    double median_best[RFMF_MAX_FLIPS_CANDIDATES];        /* median incremental best cut of the pilots */
    double chain_steps_per_s[RFMF_MAX_FLIPS_CANDIDATES];  /* pilot throughput */
} rfmf_flips_report;

/* Anneals `pilot_chains` chains of every candidate ladder over the full
 * budget (initial spins and solver seeds by the production rules, autotuned
 * geometry), then chooses: among candidates whose median best cut is at least
 * (1 - quality_tol) x the floor median (the `reference` candidate's, or the
 * best candidate's), the one with the highest chain-steps/s. */
rfmf_status rfmf_tune_flips(rfmf_ctx* ctx, int32_t graph_id, const rfmf_flips_tune* spec,
                            rfmf_flips_report* report);

/* ---------------------------------------------------------------------- */
/* Serving: continuous batching of MAX-CUT solve requests.                 */
/* ---------------------------------------------------------------------- */

typedef struct rfmf_server rfmf_server;

typedef struct rfmf_server_stats {
    int64_t submitted;
    int64_t completed;
    int64_t rejected;          /* refused at submit: over budget or invalid (counted in submitted) */
    int64_t active_chains;
    int64_t resident_bytes;
    int64_t slices;
    double chain_steps;        /* sweeps executed, all chains */
    double busy_kernel_ms;     /* kernel-only time of all slices */
} rfmf_server_stats;

/* A server owns a worker thread that keeps two slices of `slice_sweeps`
 * sweeps over every admitted chain queued on the device, admits queued
 * requests in arrival order while their HBM footprint fits
 * `hbm_budget_bytes` (a request that does not fit yet is passed by at most 8
 * later requests that do, then blocks the queue until it fits), and retires
 * finished requests while the next slice runs. */
rfmf_status rfmf_server_create(rfmf_ctx* ctx, int64_t hbm_budget_bytes, int64_t slice_sweeps,
                               rfmf_server** out);
/* Queues a request (graph copied, inputs copied); outputs are written into
 * the job's output pointers when the request completes (they must stay valid
 * until then).  job->graph_id is ignored.  The request is validated and its
 * host state prepared (CSR checks, int32 gate, initial cuts) on the calling
 * thread, so RFMF_ERR_ADMISSION (footprint over budget), RFMF_ERR_UNSUPPORTED
 * and RFMF_ERR_INVALID_ARG come back from this call. */
rfmf_status rfmf_server_submit(rfmf_server* srv, const rfmf_csr* g, const rfmf_job* job,
                               int64_t* request_id);
/* done = 1 when the request's outputs are written.  Times are seconds since
 * the server's creation (submit, admission, completion). */
rfmf_status rfmf_server_poll(rfmf_server* srv, int64_t request_id, int32_t* done,
                             double* t_submit, double* t_admit, double* t_done);
rfmf_status rfmf_server_wait(rfmf_server* srv, int64_t request_id, double timeout_s);
rfmf_status rfmf_server_stats_get(rfmf_server* srv, rfmf_server_stats* out);
/* Drains the queue (finishes every admitted and queued request), stops the
 * worker and frees device state. */
void rfmf_server_destroy(rfmf_server* srv);

#ifdef __cplusplus
}
#endif

#endif /* RFMF_H */
