/*
 * ssegpu.h -- C ABI of the MI355X SSE sweep library (libssegpu.so).
 *
 * The library runs the stochastic series expansion (SSE) sweep of
 * StochasticSeriesExpansion.jl for a batch of independent walkers (Markov
 * chains) on one gfx950 GPU: diagonal update, vertex-list construction and
 * abstract-loop (worm) update, plus the operator-string measurement.
 *
 * Everything crosses the boundary as plain pointers and sizes.  The caller
 * owns every input array only for the duration of a call; the library copies
 * what it keeps.  Errors are int status codes (SSE_OK == 0) plus a
 * thread-local message from sse_last_error().  One batch handle is used by one
 * host thread at a time; each batch owns its own HIP stream.
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * StochasticSeriesExpansion.jl v0.1.2 source tree):
 *   sse_batch_create      <- MC(params)                  src/sse.jl:26-45
 *                            + generate_sse_data / VertexData (the model and
 *                            its tables arrive already flattened, see below)
 *   sse_init              <- Carlo.init!(mc, ctx, params)  src/sse.jl:47-60
 *   sse_sweep             <- Carlo.sweep!(mc, ctx)         src/sse.jl:62-68
 *   sse_measure           <- Carlo.measure!(mc, ctx)       src/sse.jl:70-87,321-376
 *   sse_get_walker        <- Carlo.write_checkpoint        src/sse.jl:89-97
 *   sse_set_walker        <- Carlo.read_checkpoint         src/sse.jl:99-107
 */
#ifndef SSEGPU_H
#define SSEGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SSEGPU_ABI_VERSION 2

/* ---- status codes ------------------------------------------------------ */
enum sse_status {
    SSE_OK = 0,
    SSE_ERR_INVALID = 1,   /* invalid argument or unsupported model shape      */
    SSE_ERR_HIP = 2,       /* HIP runtime error                               */
    SSE_ERR_OOM = 3,       /* device allocation failed                        */
    SSE_ERR_SCATTER = 4,   /* scatter fell through (r >= last cumprob) or hit
                              an invalid transition; reference behaviour there
                              is undefined (src/vertex_data.jl:124)           */
    SSE_ERR_NO_DEVICE = 5, /* no usable gfx950 device                         */
    SSE_ERR_STATE = 6      /* call out of order (e.g. sweep before init)      */
};

/* ---- operator encoding ---------------------------------------------------
 * Device operator code (uint32):
 *   0                                   identity
 *   1 | vcode << 1 | bond << 11         bond operator
 * vcode = diag | vidx << 1 (vidx 0-based < 512), bond 0-based < 2^21.
 * The reference encodes OperCode as UInt64 1 | vcode' << 1 | bond' << 26 with
 * vcode' = diag | vidx' << 1 and 1-based vidx', bond' (src/opercode.jl:18-47);
 * sse_opcode_to_reference / sse_opcode_from_reference map losslessly.      */
#define SSE_OP_VCODE_BITS 10
#define SSE_OP_BOND_SHIFT 11
#define SSE_VCODE_INVALID (-1)

/* ---- VertexData of one bond type (src/vertex_data.jl:13-28) -------------
 * A bond type acts on n_sites sites (VertexData{NSites}, 1 <= NSites <= 4),
 * so its vertices have leg_count = 2 * n_sites legs, 0-based: legs
 * 0..n_sites-1 are the lower legs of bond sites 0..n_sites-1 (state before
 * the operator), n_sites..2*n_sites-1 the upper legs (state after).  Both
 * kernel families take any n_sites.  In wave mode a batch whose models all
 * have two-site bonds runs the two-site kernels (LDS records, the chase
 * table, the work-queue launch); a batch with any other model runs every
 * walker through the generic wave kernels, which have no work-queue launch
 * (SSE_OPT_PERSISTENT and sse_run_budgets_noalloc are refused;
 * sse_run_budgets issues one launch per sweep index).                      */
#define SSE_MAX_BOND_SITES 4
typedef struct sse_vertex_data_desc {
    int32_t n_sites;                 /* NSites, 1..SSE_MAX_BOND_SITES           */
    int32_t dims[SSE_MAX_BOND_SITES];/* local dimension of each bond site      */
    double energy_offset;            /* src/vertex_data.jl:129-135             */
    int32_t n_vertices;
    const double *weights;           /* [n_vertices] |w|                       */
    const int8_t *signs;             /* [n_vertices] sign of w                 */
    const uint8_t *leg_states;       /* [n_vertices][2*n_sites], 1-based       */
    const int32_t *diagonal_vertices;/* [prod(dims)] vcode or -1; index
                                        sum_k (s_k - 1) * prod_{j<k} dims_j,
                                        the first site fastest (join_idx)      */
    int32_t max_worm_count;
    const int32_t *transition_offset;/* [n_vertices][max_worm_count][2*n_sites]:
                                        first exit index or -1 (invalid)       */
    const int32_t *transition_count; /* same shape: number of exits            */
    int32_t n_exits;
    const double *exit_cumprob;      /* [n_exits] cumulative probabilities     */
    const int32_t *exit_target;      /* [n_exits] target vcode                 */
    const uint8_t *exit_leg;         /* [n_exits] 0-based exit leg (< 2*n_sites) */
    const uint8_t *exit_worm;        /* [n_exits] 1-based exit worm function   */
} sse_vertex_data_desc;

/* ---- SSEData + estimator tables (src/sse_data.jl:15-22) ------------------
 * All bonds of a model act on the same number of sites (SSEData{NSites}). */
typedef struct sse_model_desc {
    int32_t n_sites;
    const uint8_t *site_dims;        /* [n_sites]                              */
    int32_t n_bonds;
    int32_t bond_n_sites;            /* NSites of every bond and bond type     */
    const int32_t *bond_type;        /* [n_bonds] 0-based VertexData index     */
    const int32_t *bond_sites;       /* [n_bonds][bond_n_sites] 0-based sites  */
    int32_t n_types;
    const sse_vertex_data_desc *types;
    double energy_offset;            /* sum over bonds of type offsets         */
    double normalization_site_count; /* normalization_site_count(model)        */
    /* MagnetizationEstimator family (magnetization_estimator.jl:96-230):
     * estimator e contributes coef[e][site] * mval[e][site][state-1].        */
    int32_t n_estimators;
    int32_t est_max_dim;
    const double *est_coef;          /* [n_estimators][n_sites]                */
    const double *est_mval;          /* [n_estimators][n_sites][est_max_dim]   */
} sse_model_desc;

/* ---- per-walker (per Carlo task) parameters (src/sse.jl:26-60) ----------- */
typedef struct sse_walker_desc {
    int32_t model;                   /* index into the models array            */
    int32_t diagonal_warmup_sweeps;  /* default 5                              */
    double T;                        /* temperature                            */
    uint64_t seed;                   /* Philox4x32-10 key of this walker       */
    int64_t init_opstring_cutoff;    /* < 0: round(n_sites*T), ties to even    */
    double init_num_worms;           /* default 5                              */
    double num_worms_attenuation_factor; /* default 0.01                       */
    double target_worm_length_fraction;  /* default 2.0                        */
    int64_t thermalization;          /* Carlo thermalization sweeps            */
} sse_walker_desc;

/* ---- walker scalars for checkpoints and validation ----------------------- */
typedef struct sse_walker_scalars {
    int64_t cutoff;                  /* L = length(operators)                  */
    int64_t num_operators;           /* n                                      */
    double avg_worm_length;
    double num_worms;
    uint64_t rng_counter;            /* next 64-bit word of the Philox stream  */
    int64_t sweeps;                  /* sweeps done (Carlo ctx.sweeps)         */
    int64_t last_scatter_calls;      /* W of the last sweep                    */
    double last_total_worm_length;   /* total_worm_length of the last sweep    */
    double last_worm_length_fraction;/* WormLengthFraction if measured, else NaN */
} sse_walker_scalars;

/* Number of doubles sse_measure writes per walker: 5 + 5*n_estimators.
 * Order: Sign, OperatorCount, SignOperatorCount, SignOperatorCount2,
 * SignEnergy, then per estimator SignMag, SignAbsMag, SignMag2, SignMag4,
 * SignMagChi (src/sse.jl:73-82, magnetization_estimator.jl:218-227).      */

typedef struct sse_batch sse_batch;

/* kernel strategies */
enum sse_mode {
    SSE_MODE_WAVE = 0,    /* one wavefront per walker: parallel diagonal update
                             over 64-slot chunks, fused vertex-list linking,
                             LDS-resident worm records (default)            */
    SSE_MODE_LANE = 1     /* one lane per walker, reference order slot by slot
                             (the plain form)                               */
};

/* Create a batch of n_walkers walkers on the current HIP device.  models and
 * walkers are copied.  Replaces MC(params) (src/sse.jl:26-45).            */
int sse_batch_create(const sse_model_desc *models, int32_t n_models,
                     const sse_walker_desc *walkers, int32_t n_walkers,
                     int32_t mode, sse_batch **out);
int sse_batch_destroy(sse_batch *b);

/* Carlo.init! for every walker: random spins (one int draw per site),
 * identity opstring of the initial cutoff, diagonal warm-up sweeps.       */
int sse_init(sse_batch *b);

/* n_sweeps x Carlo.sweep! for every walker (diagonal update, vertex list,
 * worm update incl. the worm-number controller while not thermalized).
 * Carlo's counter rule: a walker is thermalized during a sweep when
 * sweeps_done > thermalization; sweeps_done is incremented after it.     */
int sse_sweep(sse_batch *b, int32_t n_sweeps);

/* Carlo.measure! for every walker.  out: [n_walkers][5+5*n_estimators] of
 * the walker's model (host pointer), or NULL to keep results on device.   */
int sse_measure(sse_batch *b, double *out);
int32_t sse_measure_width(const sse_batch *b, int32_t walker);

/* Fused sweep+measure loop: n_sweeps sweeps, each followed by a measure once
 * the walker is thermalized; the per-walker observables are accumulated on
 * device into sums and sums of squares ([n_walkers][width] each).  The
 * sequence is captured into a HIP graph and replayed.                      */
int sse_run(sse_batch *b, int32_t n_sweeps, int32_t measure);
int sse_read_accumulators(sse_batch *b, double *sums, double *sumsq,
                          int64_t *counts);
/* One persistent, work-queue-driven launch in which walker w advances by its
 * own sweeps[w] sweeps (each followed by the measurement once thermalized
 * when measure != 0).  A walker's chain and results are those of the same
 * number of Carlo.sweep! calls (src/sse.jl:62-87) however the budgets
 * differ, so a driver can give every Carlo task of a batch its own sweep
 * count, or balance uneven walkers by expected cost.  In lane mode the call
 * is one sweep (+ measure) launch per sweep index, in which the walkers
 * whose budget is used up stay idle.  Not with temperature ladders that
 * exchange between sweeps.  No reference counterpart: the reference sweeps
 * one MC per call.                                                          */
int sse_run_budgets(sse_batch *b, const int64_t *sweeps, int32_t measure);
/* sse_run_budgets without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory during the
 * call (serving loops that forbid it, harnesses that do).  The task counter
 * and the budgets live in a static device array of each GPU the library
 * runs on (about 0.75 MiB, held until the process ends), so the batch may
 * hold at most SSE_NOALLOC_MAX_WALKERS walkers (SSE_ERR_INVALID beyond).
 * Tasks run in the batch's schedule order; results are identical to
 * sse_run_budgets.  No events or graphs are made, so sse_call_time is not
 * updated.  Calls on one GPU share its array: calls that overlap on one GPU
 * (two host threads, two batches) race and must be serialized by the
 * caller.                                                                  */
#define SSE_NOALLOC_MAX_WALKERS 65536
int sse_run_budgets_noalloc(sse_batch *b, const int64_t *sweeps, int32_t measure);
int sse_reset_accumulators(sse_batch *b);
/* WormLengthFraction = (sum of worm lengths) / n, which the reference
 * measures inside the worm update of every thermalized sweep with n != 0
 * (src/sse.jl:200-202): per walker ([n_walkers] each) its sum, sum of
 * squares and sample count over the measured sweeps of sse_run since the
 * last sse_reset_accumulators (slots follow the temperatures under
 * parallel tempering, like the other sums).                                */
int sse_read_worm_fraction(sse_batch *b, double *sums, double *sumsq, int64_t *counts);

/* Walker state (checkpoint / validation).  ops: [cutoff] device operator
 * codes or NULL; spins: [n_sites] 1-based states or NULL.                   */
int sse_get_walker(sse_batch *b, int32_t walker, sse_walker_scalars *s,
                   uint32_t *ops, int64_t ops_capacity, uint8_t *spins);
int sse_set_walker(sse_batch *b, int32_t walker, const sse_walker_scalars *s,
                   const uint32_t *ops, const uint8_t *spins);

/* Validation: run only some phases of a sweep.  phases is a bit mask of
 * 1 = diagonal update, 2 = vertex list, 4 = worm update.  Used to compare
 * intermediate results with a CPU restatement.                            */
int sse_sweep_phases(sse_batch *b, int32_t phases);
/* Vertex list of walker w after the last vertex-list phase: links
 * [leg_count*cutoff] linked leg ids (leg_count*p + leg) or -1, first/last
 * [n_sites]; leg_count = 2 * bond_n_sites.                                */
int sse_get_vertex_list(sse_batch *b, int32_t walker, int32_t *links,
                        int64_t links_capacity, int32_t *first, int32_t *last);

/* Per-phase device time of the last sse_sweep/sse_run call in ms
 * (diagonal+vertex list, worm, measure), measured with HIP events on the
 * batch stream.                                                            */
int sse_phase_times(const sse_batch *b, double *ms3);
/* With SSE_OPT_TIMING 2: per kernel of the last sse_run/sse_sweep/
 * sse_run_budgets call the summed launch duration in ms and the launch count,
 * [0] diagonal update +
 * vertex list, [1] worm update, [2] measurement, [3] persistent launch.
 * No reference counterpart (measurement).                                  */
int sse_kernel_times(const sse_batch *b, double *ms4, int64_t *launches4);
/* Device time (ms, HIP events on the batch stream) of the last sse_sweep /
 * sse_run call and its sweep count (0, 0 with timing off).  No reference
 * counterpart: the reference runs one walker per process.                  */
int sse_call_time(const sse_batch *b, double *ms, int64_t *sweeps);
/* Dynamic LDS bytes per walker of the current diagonal-update and worm-kernel
 * launches (wave mode; sized from the batch's largest operator count) and the
 * worm-kernel workgroups a CU can hold at that size (HIP occupancy query).
 * No reference counterpart (tuning introspection).                         */
int sse_lds_bytes(const sse_batch *b, int64_t *k1, int64_t *k2, int32_t *k2_per_cu);
/* The autotuner's choice for sse_run (which 0) or sse_run_budgets (which 1):
 * persistent launch (0/1), stream groups, chase variant, worm-kernel LDS
 * setting (0 = sized from the operator counts) and the measured vertex
 * updates per second; candidates receives the number of forms measured.
 * SSE_ERR_STATE while still measuring.  No reference counterpart (tuning
 * introspection).                                                          */
int sse_tuning(const sse_batch *b, int32_t which, int32_t *persistent, int32_t *groups,
               int32_t *chase, int64_t *k2_lds, double *rate, int32_t *candidates);
/* Sum over walkers of cutoff L, operators n and scatter calls W since the
 * last reset (counters taken from the chains themselves).                  */
int sse_work_counters(sse_batch *b, int64_t *L_sum, int64_t *n_sum,
                      int64_t *W_sum, int64_t *walker_sweeps);

/* ---- serving: slots ------------------------------------------------------
 * Every walker of a batch is a slot.  A slot can be restarted with a new job
 * (walker parameters, model and a budget of sweeps) while the others keep
 * running; sse_run/sse_sweep initialise pending slots, sweep active ones and
 * retire a slot after its last budgeted sweep (+ measure).  The job's results
 * are the slot's accumulators.  These entry points have no single counterpart
 * in the reference: they replace Carlo's per-task scheduling
 * (JobTools/TaskMaker + one MC per task, examples/bani2v2o8.jl:1-42).       */

/* Upload one more model into the batch (n_sites <= the batch's maximum). */
int sse_add_model(sse_batch *b, const sse_model_desc *model, int32_t *index);
/* Restart slot with walker parameters wd for n_sweeps sweeps (-1 = unlimited);
 * the slot's accumulators are cleared.                                      */
int sse_slot_start(sse_batch *b, int32_t slot, const sse_walker_desc *wd, int64_t n_sweeps);
/* Stop a slot (it is neither initialised nor swept until restarted).       */
int sse_slot_stop(sse_batch *b, int32_t slot);
/* Slots that finished their budget since the last poll.                     */
int sse_slot_poll(sse_batch *b, int32_t *done_slots, int32_t capacity, int32_t *n_done);
/* Bytes of the slab pool in use (opstring + records of every walker).       */
int sse_pool_usage(sse_batch *b, int64_t *used_bytes, int64_t *total_bytes);

/* ---- parallel tempering in T ---------------------------------------------
 * The reference's two exchange hooks, which Carlo's parallel-tempering loop
 * calls per walker on the host:                                              */
/* Carlo.parallel_tempering_log_weight_ratio(mc, :T, new_T) (src/sse.jl:390-396):
 * -n log(new_T / T) of the walker's current configuration.                 */
int sse_pt_log_weight_ratio(sse_batch *b, int32_t walker, double new_T, double *ratio);
/* Carlo.parallel_tempering_change_parameter!(mc, :T, new_T) (src/sse.jl:398-405):
 * set the walker's T; its configuration and worm controller stay.            */
int sse_pt_change_parameter(sse_batch *b, int32_t walker, double new_T);
/* The walker's current T (read back for checkpoints and checks).           */
int sse_pt_get_parameter(sse_batch *b, int32_t walker, double *T);
/* The exchange loop itself, on the device (replaces Carlo's host-side loop
 * over those hooks; no reference counterpart in this package).  slot_walkers
 * holds n_ladders ladders of ladder_len walkers each, in temperature-slot
 * order; every walker at most once, one model per ladder, no serving slots.
 * Round r proposes the neighbour pairs (k, k+1) with k % 2 == r % 2 of every
 * ladder and accepts a pair with probability min(1, exp(lwr(a->T_b) +
 * lwr(b->T_a))) (uniforms from a Philox stream keyed by seed); an accepted
 * pair swaps the walkers' temperatures, slots and accumulator rows, so row w
 * of sse_read_accumulators / sse_measure keeps collecting the temperature
 * walker w was created with.  interval > 0 runs one round after every
 * interval-th sweep of sse_sweep/sse_run (counted across calls); 0 leaves
 * rounds to sse_pt_exchange.  n_ladders = 0 removes the ladders.  Counters
 * and the round number restart at 0.                                        */
int sse_pt_set_ladders(sse_batch *b, const int32_t *slot_walkers, int32_t n_ladders,
                       int32_t ladder_len, uint64_t seed, int32_t interval);
/* One exchange round now (on the batch stream).                             */
int sse_pt_exchange(sse_batch *b);
/* Current slot order, accepted and proposed swaps per lower slot (each
 * n_ladders * ladder_len entries; may be NULL) and rounds run so far.       */
int sse_pt_get_ladders(sse_batch *b, int32_t *slot_walkers, int64_t *accepted,
                       int64_t *proposed, int64_t *rounds);

/* ---- options -------------------------------------------------------------
 * SSE_OPT_GRAPH  (default 1): sse_run/sse_sweep replay captured HIP graphs
 *                instead of launching each kernel: up to 32 sweeps (+
 *                measures) on one stream, the persistent launch of a call.  2
 *                also captures the stream groups' fork/join pipeline (10
 *                sweeps per graph, the group streams joined to the capture by
 *                events); replays of that multi-stream graph measured slower
 *                than direct launches (6.3 against 4.1 ms per sweep of 4096
 *                16x16 beta=8 walkers in 8 groups), so it is not the default.
 * SSE_OPT_TIMING (default 1): every sweep/run call is bracketed by HIP events
 *                (sse_call_time); single-stream calls of <= 256 sweeps also
 *                record events around each kernel (sse_phase_times) and
 *                launch eagerly.  2: in addition, sse_run/sse_sweep/
 *                sse_run_budgets record events around every kernel launch on
 *                the stream it runs on (each stream group's diagonal/vertex-
 *                list, worm and measurement kernels, or the persistent
 *                launch), with direct launches instead of graph replays
 *                (sse_kernel_times).
 * SSE_OPT_WORM_RECORDS (default 0): record layout of the worm kernel.  0 picks
 *                the smallest that fits (8-byte LDS records for n <= 4096 and
 *                <= 256 chase codes, 10-byte LDS records for n <= 16384, chase
 *                codes in LDS with links read from global memory, else global
 *                memory); the kernel's LDS per walker is capped at 64 KiB so
 *                large walkers take the later layouts; 1 skips the 8-byte
 *                layout; 2 forces global memory; 3 forces the hybrid.  Results
 *                are identical; this is a tuning/testing knob.
 * SSE_OPT_PERSISTENT (default 0): 1 = sse_run/sse_sweep issue one persistent,
 *                work-queue-driven launch per call (wave mode): each
 *                workgroup takes a walker from an atomic task counter and runs
 *                all of that walker's sweeps of the call (diagonal update,
 *                vertex list, worm update, measurement) before it takes the
 *                next; walkers never wait for each other.  Results are
 *                identical.
 * SSE_OPT_STREAMS (default 0): wave mode splits the walkers into this many
 *                groups, each swept on its own HIP stream (forked from and
 *                joined into the batch stream), so the slow tail of one
 *                group's worm update overlaps the other groups' sweeps.
 *                0 = automatic: one group per 512 walkers, at least 2 and at
 *                most one per hardware queue (GPU_MAX_HW_QUEUES, 4 when
 *                unset; streams beyond the queues share them and serialize,
 *                and smaller groups leave CUs idle).  1 = one stream (graphs and
 *                per-kernel timing apply only then).  Results are identical
 *                for any value.
 * SSE_OPT_K2_LDS (default 0): LDS bytes per walker of the worm kernel; 0 =
 *                sized from the batch's largest operator count (then widened to
 *                the largest size with the same walkers per CU), and with
 *                stream groups from each group's own largest count.  Walkers
 *                that do not fit take a slower record layout; results are
 *                identical.
 * SSE_OPT_CHASE (default 0): worm traversal over the LDS record layouts.
 *                0 = the scalar chase (one record read per step); 1 = record
 *                lookahead: at each step four lanes read the records linked
 *                to the current operator's four legs while the exit is
 *                selected, so the next record is already in registers (more
 *                instructions per step; slower where the step is bound by
 *                instruction issue, as on the batched Heisenberg workload).
 *                Results are identical.
 * SSE_OPT_TUNE (default 1): autotuner.  The first measuring sse_run calls of
 *                a wave-mode batch run their sweeps in turn (segments of four
 *                sweeps, twice round-robin) with the candidate launch forms: stream
 *                groups at 2, 4, 8 (with more hardware queues than 8) and
 *                hardware-queue-count groups; the
 *                persistent work-queue launch with the worm-kernel LDS sized
 *                from the largest walker or capped at 32 KiB; then the
 *                winner with the record-lookahead chase.  Each is scored by
 *                vertex updates (the chains' own counters) per second; the
 *                fastest is kept.  sse_run_budgets tunes the same way over
 *                its first calls (work-queue forms only).  Setting
 *                SSE_OPT_STREAMS, SSE_OPT_PERSISTENT, SSE_OPT_K2_LDS or
 *                SSE_OPT_CHASE turns tuning off; setting SSE_OPT_TUNE restarts
 *                it.  Results are identical for every choice.
 * SSE_OPT_FUSE_MEASURE (default 0): 1 = measuring sweeps of the per-sweep
 *                kernels run the worm update and the measurement in one kernel
 *                (each walker's wave measures right after its own worm update,
 *                inside the launch's worm tail).  Slower on the batched
 *                Heisenberg workload: the measuring wave keeps the worm
 *                kernel's LDS, so fewer walkers chase at a time.  Results are
 *                identical.
 * SSE_OPT_LANE_WIDTH: lane mode (SSE_MODE_LANE) runs this many walkers per
 *                wavefront (1..64; one lane each).  Default: 1 until the batch
 *                has 16 wavefronts per compute unit, then the power of two
 *                that keeps it near that.  Fewer walkers
 *                per wavefront means more wavefronts for the same batch, so
 *                the compute units hide more of each lane's memory latency,
 *                and less divergence between the chains of one wavefront.
 *                Results are identical.
 * SSE_OPT_ORDER (default 1): launch schedule.  1 = walkers in order of their
 *                expected operator count (bonds / T), largest first, so the
 *                longest chains start first and each stream group holds walkers
 *                of similar size (and gets a worm-kernel LDS size to match);
 *                0 = walker index order.  Results are identical.
 * SSE_OPT_GENERIC (wave mode): 1 = run every model, two-site bonds included,
 *                through the generic wave kernels (the ones that bonds of one,
 *                three and four sites always take: chase codes of two bytes
 *                per operator in LDS, no work-queue launch); 0 = the two-site
 *                kernels (an error while a model has bonds of other sizes).
 *                Default: 0 unless such a model is present.  Restarts tuning.
 *                Results are identical.                                     */
enum sse_option {
    SSE_OPT_GRAPH = 1,
    SSE_OPT_TIMING = 2,
    SSE_OPT_WORM_RECORDS = 3,
    SSE_OPT_PERSISTENT = 4,
    SSE_OPT_STREAMS = 5,
    SSE_OPT_K2_LDS = 6,
    SSE_OPT_ORDER = 7,
    SSE_OPT_CHASE = 8,
    SSE_OPT_TUNE = 9,
    SSE_OPT_FUSE_MEASURE = 10,
    SSE_OPT_LANE_WIDTH = 11,
    SSE_OPT_GENERIC = 12
};
int sse_set_option(sse_batch *b, int32_t option, int64_t value);

/* Encoding helpers (src/opercode.jl:18-71). */
uint64_t sse_opcode_to_reference(uint32_t op);
uint32_t sse_opcode_from_reference(uint64_t op);

/* Transition tables of one bond type: the abstract-loop linear programs of
 * construct_transitions (src/vertex_data.jl:213-416, solved there by HiGHS
 * through JuMP at :267-350), here by a dense two-phase simplex on the host;
 * no device is used.  Same step, variable and class orders, normalisation,
 * clamping, pruning (prob > tolerance / #steps) and cumulative sums.
 *   n_sites, dims[n_sites]          bond sites and their local dimensions
 *   leg_states[2*n_sites][n_vertices]  1-based, as VertexData.leg_states
 *   weights[n_vertices]             |matrix element| of each vertex
 * Out: trans_offset / trans_length [n_vertices][max_worm][2*n_sites]
 * (max_worm = max(dims) - 1): 1-based offset into the exit lists (0 = no
 * entry) and exit count - 1; up to exit_capacity exits in cumprobs,
 * targets (vertex codes diag | idx << 1, 1-based idx), out_leg, out_worm.
 * *n_exits receives the exit count (SSE_ERR_INVALID when it exceeds
 * exit_capacity); *flags: bit 0 a probability was clamped by more than
 * tolerance, bit 1 an exit list misses 1 by more than tolerance, bit 2 some
 * LP optimum may not be unique (another solver may return another valid
 * table).  SSE_ERR_INVALID also when an LP fails or its solution leaves
 * [0, weight] by more than lp_tolerance.                                   */
int sse_vertex_transitions(int32_t n_sites, const int32_t *dims, int32_t n_vertices,
                           const uint8_t *leg_states, const double *weights, double tolerance,
                           double lp_tolerance, int32_t *trans_offset, int32_t *trans_length,
                           int64_t exit_capacity, double *cumprobs, int32_t *targets,
                           int32_t *out_leg, int32_t *out_worm, int64_t *n_exits, int32_t *flags);

const char *sse_last_error(void);
int32_t sse_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif /* SSEGPU_H */
