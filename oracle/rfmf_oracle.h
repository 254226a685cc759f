/*
 * rfmf_oracle.h -- CPU restatement of the RFMF multi-chain rejection-free
 * multiple-flip annealer, written as TEST INFRASTRUCTURE.
 *
 * This library is the parity oracle for the HIP path and the timed CPU
 * baseline.  Only the test suite, the smoke check and the CPU-baseline leg of
 * the benchmark load it; the product library never links it, never calls it
 * and has no CPU fallback through it.
 *
 * Every function restates the reference's algorithm from its source (cited
 * per function in rfmf_oracle.c as path:line relative to the reference
 * checkout).  The annealing engine deliberately uses the SEQUENTIAL
 * group-order formulation (flip winners one group at a time against the live
 * neighbour field), which is independent of the analytic corr[] formulation
 * used by the reference kernel and by the HIP kernels here.
 *
 * Pinning (what checks each part, so a dropped term or swapped index fails):
 *   philox4x32_10      Random123 published known-answer vectors, plus two
 *                      independent implementations (rocRAND, ATen) compiled
 *                      on the host by tests/golden/make_philox_kat.py.
 *   gumbel LUT         closed form round(-ln(-ln((2i+1)/32)) * 2^20).
 *   splitmix/xoshiro   published splitmix64 first output; initial spins
 *                      generated by the reference's own rng.hpp compiled
 *                      host-side (oracle/_ref, fixtures in tests/golden).
 *   beta run tables    per-step brute force, and the reference's own
 *                      buildBetaRunTables compiled host-side.
 *   int32 gate, B      closed form and overflow edge cases; reference build.
 *   CSR / cut          reference AnnealCsr + cutValueCsr compiled host-side;
 *                      brute force; committed partitions re-scored to the
 *                      committed values.
 *   anneal engine      (i) replay of chains from the reference's committed
 *                      80M-sweep benchmark logs, byte-for-byte partitions;
 *                      (ii) incremental cut == recomputed cut;
 *                      (iii) brute-force tiny-graph checks of the group
 *                      argmax and tie-break.
 */
#ifndef RFMF_ORACLE_H
#define RFMF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based Gumbel source ---- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
int32_t or_gumbel_from_bits(uint32_t r11);
int32_t or_gumbel_lut(int i);
uint64_t or_pack_score_id(int32_t score, int32_t id);
/* Gumbel (Q20) handed to vertex v in 1-based sweep `step` of the chain whose
 * solver seed is `solver_seed`. */
int32_t or_vertex_gumbel(uint64_t solver_seed, uint64_t step, uint64_t v);

/* ---- initial spins: splitmix64-seeded xoshiro256++, 1 bit per vertex ---- */
uint64_t or_splitmix64_next(uint64_t* state);
void or_initial_spins(uint64_t raw_seed, int64_t n, int8_t* spin_out);

/* ---- beta schedule ---- */
int32_t or_beta_fixed(uint64_t step, double t0);
/* Run-length table of or_beta_fixed over 1-based steps [1, total].  Writes
 * at most `cap` runs; returns the run count, or -1 if cap is exceeded. */
int64_t or_beta_runs_log(double t0, int64_t total, int64_t* run_end,
                         int32_t* run_val, int64_t cap);
/* Integer linear schedule b0 -> b1 (Q20) over 1-based steps [1, total]:
 * beta(s) = b0 + ((b1 - b0) * (s - 1)) / (total - 1)   (total > 1). */
int32_t or_beta_linear(uint64_t step, int64_t total, int32_t b0, int32_t b1);
int64_t or_beta_runs_linear(int32_t b0, int32_t b1, int64_t total,
                            int64_t* run_end, int32_t* run_val, int64_t cap);

/* ---- gate and bound ---- */
int or_score_safe_int32(int64_t total_steps, double t0, int64_t delta_bound);
int64_t or_max_weighted_degree(const uint32_t* off, const int32_t* to,
                               const int32_t* w, int64_t n);

/* ---- graph ---- */
int64_t or_cut_value(const uint32_t* off, const int32_t* to, const int32_t* w,
                     int64_t n, const int8_t* spin);
/* Build the dense CSR from an undirected edge list exactly as the reference
 * load path does: first weight wins on duplicates, zero weights already
 * dropped by the caller, degree-zero vertices removed, dense ids by ascending
 * original id, rows sorted by dense neighbour id.  Arrays are malloc'd; the
 * caller frees them with or_free.  Returns dense n, or -1 on error. */
int64_t or_build_csr(const int64_t* eu, const int64_t* ev, const int64_t* ew,
                     int64_t m, uint32_t** off_out, int32_t** to_out,
                     int32_t** w_out, int64_t** new_to_old_out);
/* Integer-weight Gset/MQLib text reader (blank and '#' lines skipped, CRLF
 * tolerated, zero weights dropped).  Returns dense n, or -1. */
int64_t or_read_gset(const char* path, uint32_t** off_out, int32_t** to_out,
                     int32_t** w_out, int64_t* header_n, int64_t* header_m);
void or_free(void* p);

/* ---- the annealing engine (one chain, sequential group-order semantics) ---- */
typedef struct {
    int64_t cur;        /* incremental cut after the last sweep */
    int64_t best;       /* best incremental cut */
    uint64_t steps;     /* sweeps executed */
} or_chain_result;

/* Runs one chain from init_spin through every phase with P > 0.  The beta
 * value of 1-based sweep s is the run table entry covering 0-based step s-1.
 * end_spin/best_spin receive n bytes each.  Returns 0 on success. */
int or_anneal_chain(const uint32_t* off, const int32_t* to, const int32_t* w,
                    int64_t n, const int8_t* init_spin, uint64_t solver_seed,
                    const int64_t* run_end, const int32_t* run_val, int64_t n_runs,
                    const uint64_t* phases, const uint64_t* iters, int64_t n_phases,
                    int8_t* end_spin, int8_t* best_spin, or_chain_result* res);

/* C chains over `threads` POSIX threads (threads <= 0: one per chain, capped
 * at 64).  Chain c uses init_spins + c*n, solver_seeds[c] and the run table
 * starting at chain_run_off[c] (tables are concatenated; each table's last
 * run end is >= the executed step count).  Returns the solver-only wall
 * seconds (steady clock around the sweep loops), or a negative value on error. */
double or_anneal_many(const uint32_t* off, const int32_t* to, const int32_t* w,
                      int64_t n, int64_t C, const int8_t* init_spins,
                      const uint64_t* solver_seeds, const int64_t* run_end,
                      const int32_t* run_val, const int64_t* chain_run_off,
                      int64_t n_runs_total,
                      const uint64_t* phases, const uint64_t* iters, int64_t n_phases,
                      int threads, int8_t* end_spins, int8_t* best_spins,
                      int64_t* cur_vals, int64_t* best_vals);

#ifdef __cplusplus
}
#endif

#endif /* RFMF_ORACLE_H */
