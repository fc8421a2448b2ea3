/*
 * sse_oracle.h -- CPU restatement of the StochasticSeriesExpansion.jl SSE
 * Markov chain (one walker, reference order).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it, and only as the checker of the GPU
 * path.  The product (libssegpu.so) never links or calls it.
 *
 * It reads models in the plain-array layout of include/ssegpu.h (a data
 * format, shared so both sides consume identical tables).
 */
#ifndef SSE_ORACLE_H
#define SSE_ORACLE_H

#include <stdint.h>
#include "../include/ssegpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct so_walker so_walker;

void so_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
uint64_t so_philox_word(uint64_t key, uint64_t index);
double so_word_to_float(uint64_t word);
uint64_t so_range(uint64_t key, uint64_t *counter, uint64_t s);
double so_expm1(double x);
double so_tanh(double x);

so_walker *so_walker_create(const sse_model_desc *model, const sse_walker_desc *walker);
void so_walker_destroy(so_walker *w);
void so_init(so_walker *w);
int so_sweep(so_walker *w);
int64_t so_worm_traverse_from(so_walker *w, int64_t p0, int l0, int wf0);
int so_sweep_phases(so_walker *w, int phases);
void so_measure(so_walker *w, double *out);
int so_measure_width(const so_walker *w);
void so_get_scalars(const so_walker *w, sse_walker_scalars *s);
int64_t so_get_ops(const so_walker *w, uint32_t *ops, int64_t cap);
void so_get_spins(const so_walker *w, uint8_t *spins);
int so_set_state(so_walker *w, const sse_walker_scalars *s, const uint32_t *ops,
                 const uint8_t *spins);
void so_get_vertex_list(const so_walker *w, int64_t *links, int64_t *first, int64_t *last);
void so_make_vertex_list(int32_t n_sites, int32_t bond_n_sites, const int32_t *bond_sites,
                         const uint32_t *ops, int64_t L, int64_t *links, int64_t *first,
                         int64_t *last);
/* Run n sweeps (+ measure when thermalized) and return elapsed seconds;
 * accumulates measurements into sums (width doubles) when non-NULL.        */
double so_run_timed(so_walker *w, int64_t n_sweeps, int measure, double *sums);
int64_t so_run_series(so_walker *w, int64_t n_sweeps, double *out);
void so_work_counters(so_walker *w, int64_t *out4);

/* Parallel tempering in T (sse.jl:390-405) and one exchange round. */
double so_log(double x);
double so_pt_log_weight_ratio(const so_walker *w, double new_T);
void so_pt_change_parameter(so_walker *w, double new_T);
double so_pt_get_T(const so_walker *w);
double so_pt_uniform(uint64_t seed, int32_t ladder, int64_t round, int32_t slot);
void so_pt_exchange(so_walker *const *walkers, int32_t *slot_walker, int32_t len, uint64_t seed,
                    int32_t ladder, int64_t round, int64_t *accepted, int64_t *proposed);

#ifdef __cplusplus
}
#endif
#endif
