/*
 * es_abi.h — C ABI of the MI355X (gfx950) implementation of the
 * ElasticScattering hot path: impurity binning, the per-particle lifetime
 * trace, the Simpson integration over phi and the reductions to sigma_xx /
 * sigma_xy.
 *
 * It replaces the OpenCL backend behind the reference's operator API,
 * `class Simulation` (src/sim/Simulation.h:56-147), and the grid it consumes
 * (`class Grid`, src/sim/Grid.h:44-95).  Plain C types only; every call
 * returns ES_OK or a negative status (the reference prints and exit(0)s,
 * src/utils/ErrorMacros.h:22-38); es_last_error() gives the message.
 *
 * Threading: one solver per host thread.  A solver owns its device memory and
 * its HIP stream on the current device at creation.  Host arrays are
 * caller-owned and are only read/written during the call.
 */
#ifndef ES_ABI_H
#define ES_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ES_OK 0
#define ES_ERR_INVALID_ARG (-1)    /* bad size, null pointer, unsupported setting */
#define ES_ERR_HIP (-2)            /* a HIP runtime call failed */
#define ES_ERR_NO_GRID (-3)        /* no grid set on the solver */
#define ES_ERR_NO_SAMPLE (-4)      /* es_solver_init_sample not called */
#define ES_ERR_OOM (-5)            /* device allocation failed or over the size cap */
#define ES_ERR_OUT_OF_GRID (-6)    /* a start position lies outside the cell grid (reference reads out of bounds, lifetime.h:88-91) */
#define ES_ERR_NO_DEVICE (-7)      /* no HIP device */

#define ES_PRECISION_FP64 0        /* FP64 throughout; lifetimes bit-identical to the CPU restatement */
#define ES_PRECISION_MIXED 1       /* FP32 trace, the cases FP32 cannot decide traced again in FP64;
                                      start data, Simpson accumulation and sigma sums stay FP64 */

typedef struct es_solver es_solver;

/* GridInformation (src/sim/es/settings.h:26-33), plus the cell count. */
typedef struct es_grid_info {
    double region_size;
    double region_extends;
    double cell_size;
    int64_t indexed_impurity_count;   /* Grid::GetTotalImpurityCount (with duplicates) */
    int64_t unique_impurity_count;    /* Grid::GetUniqueImpurityCount */
    int32_t index_size;               /* cells_per_row^2 */
    int32_t cells_per_row;
} es_grid_info;

/* Arguments of the Grid constructors (src/sim/Grid.cpp:28-59). */
typedef struct es_grid_desc {
    double region_size;               /* positions span [0, region_size] */
    double region_extends;            /* spawn square [-extends, region_size + extends] */
    double impurity_radius;
    double impurity_density;          /* used by es_grid_generate when count <= 0 */
    int32_t target_cell_population;   /* cells_per_row = max(round(sqrt(N/target)), 1) (Grid.cpp:40) */
    int32_t cells_per_row;            /* > 0: explicit cell grid instead of the rule above */
} es_grid_desc;

/* Settings (src/ConfigSettings.h:22-34) as used by Simulation::InitSample. */
typedef struct es_settings {
    double tau;                       /* coherent tau (Simulation.h:118) */
    double alpha;
    int32_t is_clockwise;
    double particle_speed;
    double incoherent_tau;            /* > 0: incoherent tau instead of HBAR/(KB*T) (Simulation.h:90) */
} es_settings;

typedef struct es_sigma {
    double xx;
    double xy;
} es_sigma;

/* ParticleMetrics (src/sim/es/particle_metrics.h:22-44) + Metrics (src/Metrics.h:27-44). */
typedef struct es_metrics {
    int64_t cells_passed;
    int64_t impurities_tested;
    int64_t particles_inside_impurity;
    int64_t particles_escaped;
    int64_t particles_at_bound;
    int64_t max_impurities_tested;
    int64_t max_cells_passed;
    int64_t trace_cap_hits;           /* traces stopped after 4*cpr^2+16 cells; 0 on every finite trace */
    int64_t real_particles;           /* total - inside (SimulationCPU.cpp:41) */
    double avg_particle_lifetime;     /* Simulation.h:93-106 */
    double time_elapsed_lifetimes;    /* seconds, device events around the lifetime stage */
    double time_elapsed_temperatures; /* seconds, the per-T reduction stage */
    /* Work counters of this implementation (no reference counterpart): the
       inputs of the algorithmic FP64 work model used for roofline figures. */
    int64_t traced_particles;         /* particles that start outside every impurity */
    int64_t edge_hits;                /* accepted cell-edge intersections (cell_grid.h:80-89 calls) */
    int64_t cross_calls;              /* GetFirstCrossTime calls (details.h:182) */
    int64_t sigma_evals;              /* GetSigma evaluations with a non-zero truncated lifetime */
    int64_t cross_exact;              /* GetFirstCrossTime calls evaluated exactly; the others were
                                         screened out by a bound that proves they cannot lower the minimum */
    int64_t fp64_recomputed;          /* ES_PRECISION_MIXED: particles traced again in FP64 */
} es_metrics;

/* Simulation(positions_per_row, particles_per_quadrant, GridInformation) (Simulation.h:131-146). */
typedef struct es_solver_desc {
    int32_t positions_per_row;
    int32_t particles_per_quadrant;
    es_grid_info grid_info;           /* region_size, region_extends, cell_size of sample 0's grid;
                                         cell_size 0: taken from the first grid set on the solver */
    int32_t precision;                /* ES_PRECISION_FP64 or ES_PRECISION_MIXED */
    int32_t device;                   /* HIP device ordinal */
} es_solver_desc;

/* Simulation::Simulation — allocate a solver (Simulation.h:131-146). */
int es_solver_create(const es_solver_desc* desc, es_solver** out);

/* ~SimulationCL (SimulationCL.cpp:477-508). */
void es_solver_destroy(es_solver* s);

/* Last error message of `s` (or of the last failed create when s is NULL). */
const char* es_last_error(const es_solver* s);

/*
 * Grid(std::vector<v2> impurities, region_size, region_extends, radius, target)
 * (Grid.cpp:48-59): uploads n centres and bins them on the GPU (Grid.cpp:77-180).
 */
int es_grid_set_impurities(es_solver* s, const double* x, const double* y, int64_t n,
                           const es_grid_desc* desc, es_grid_info* out_info);

/*
 * Grid(seed, region_size, region_extends, density, radius, target)
 * (Grid.cpp:28-46): draws the centres on the GPU with this implementation's
 * generator (elasticscattering_amd/csrc/es_rng.h; the reference RNG is
 * implementation-defined) and bins them.  count <= 0 derives
 * N = ceil(side^2 * density) as Grid.cpp:35-39 does.
 */
int es_grid_generate(es_solver* s, uint64_t seed, int64_t count, const es_grid_desc* desc,
                     es_grid_info* out_info);

/*
 * The field exactly as the reference's Grid hands it to a Simulation backend
 * (Simulation::ComputeSigmas(B, T, const Grid&, ...), Simulation.h:128):
 * Grid::GetImpurities() — cell-ordered (x, y) pairs, a disk listed once per
 * cell it overlaps — and Grid::GetIndex() — cells_per_row^2 inclusive end
 * offsets (Grid.h:74-75, Grid.cpp:92-110); the indexed count is the last end.
 * spawn_region_start, spawn_region_size, impurity_radius and cells_per_row
 * are Grid::GetSettings() (Grid.h:82-91).  Uploaded as given (no binning), so
 * the indexed count, impurities_tested and every metric are those of the
 * reference's grid.  The solver must have been created with the sample-0
 * GridInformation (es_solver_desc.grid_info: region_size, cell_size), which
 * places the start positions (Simulation.h:131-146).  unique_impurity_count of
 * out_info is 0 (a binned field does not carry it).
 */
int es_grid_set_binned(es_solver* s, const double* ordered_xy, const int32_t* imp_index, int32_t cells_per_row,
                       double spawn_region_start, double spawn_region_size, double impurity_radius,
                       es_grid_info* out_info);

/*
 * Grid::GetImpurities / Grid::GetIndex (Grid.h:74-75) of the current grid:
 * cell-ordered AoS (x, y) pairs and inclusive end offsets per cell.
 * Either pointer may be NULL.
 */
int es_grid_read(es_solver* s, double* ordered_xy, int32_t* imp_index);

/* Simulation::InitSample (Simulation.h:111-126). */
int es_solver_init_sample(es_solver* s, const es_settings* settings, int32_t coherent);

/*
 * Simulation::ComputeSigmas (Simulation.h:128; SimulationCPU.cpp:33-73):
 * lifetimes for field B, then sigma for each of the nT temperatures.
 * out: nT entries.  metrics may be NULL.  lifetimes (may be NULL) receives
 * the P = ppr^2 * 4 * ppq raw lifetimes in the layout [j][i][q][p]
 * (Simulation.h:67).
 */
int es_solver_compute_sigmas(es_solver* s, double B, const double* T, int32_t nT,
                             es_sigma* out, es_metrics* metrics, double* lifetimes);

/*
 * The same for nB fields in one launch (the B loop of SimulationRunner.cpp:129-145).
 * out: nB * nT entries, B-major.  metrics: nB entries or NULL.
 */
int es_solver_compute_sigmas_sweep(es_solver* s, const double* B, int32_t nB, const double* T, int32_t nT,
                                   es_sigma* out, es_metrics* metrics);

/*
 * Simulation::ComputeSigmasWithImages (Simulation.h:129; SimulationCPU.cpp:159-171):
 * the sigmas of es_solver_compute_sigmas plus, per temperature, the images of
 * IterationResult — images[t][0][j * ppr + i] the phi-integrated truncated
 * lifetime, [t][1][...] sigma_xx and [t][2][...] sigma_xy per position
 * (IntegratePosition, SimulationCPU.cpp:224-243): nT * 3 * ppr^2 doubles.
 */
int es_solver_compute_sigmas_with_images(es_solver* s, double B, const double* T, int32_t nT, es_sigma* out,
                                         es_metrics* metrics, double* images);

/* ---------------------------------------------------------------- batches
 *
 * Many samples of uneven size (own field, cell grid, positions, fields B and
 * temperatures) traced in one launch of each stage.  No reference counterpart
 * as one call: it is the sample loop of SimulationRunner::RunSimulation
 * (SimulationRunner.cpp:88-98) and the B loop of RunSample (:129-145) run as
 * one piece of work on the GPU.  Each sample entry is one pass (coherent or
 * incoherent) of one field; two passes of one field share it through
 * share_field_of.  Every entry of one batch has the same is_clockwise.
 */
typedef struct es_batch es_batch;

typedef struct es_batch_sample {
    uint64_t seed;                    /* field drawn with es_rng.h (es_grid_generate) */
    int64_t impurity_count;           /* <= 0: ceil(side^2 * density) (Grid.cpp:35-39) */
    es_grid_desc grid;
    double grid_cell_size;            /* GridInformation.cell_size for the position offset; 0: this field's */
    int32_t positions_per_row;
    int32_t particles_per_quadrant;
    es_settings settings;
    int32_t coherent;
    int32_t share_field_of;           /* >= 0: use the field of that earlier entry instead of drawing one */
    const double* fields;             /* n_fields values of B (copied) */
    int32_t n_fields;
    const double* temperatures;       /* 0..4096 temperatures (copied; run 8 at a time) */
    int32_t n_temperatures;
} es_batch_sample;

/* Launch shape of the trace kernel (what es_batch_tune chooses). */
typedef struct es_batch_shape {
    int32_t waves_per_eu;             /* register budget: 2 or 3 waves per SIMD */
    int32_t chunk;                    /* work items a wave takes from the queue at a time */
    int32_t grid_mult;                /* resident blocks per CU times this = grid size */
    int32_t cooperative;              /* crossing solves: 0 per lane, 1 shared by the wave, 2 shared and
                                         screened (disks that cannot lower the minimum are skipped),
                                         -1 (default) 2 or 1 by the fields' impurities per cell */
} es_batch_shape;

int es_batch_create(int32_t device, es_batch** out);
void es_batch_destroy(es_batch* b);
const char* es_batch_last_error(const es_batch* b);
/*
 * Draws and bins the entry's field on the GPU (unless shared) and runs its
 * start-cell test, without waiting: the results of every new field are read
 * back together by the next es_batch_run or es_batch_counts, which also report
 * a start position outside a field's cell grid (ES_ERR_OUT_OF_GRID).
 */
int es_batch_add_sample(es_batch* b, const es_batch_sample* sample, int32_t* index);
/*
 * Removes every entry and keeps the drawn fields: a sample added afterwards
 * with the same seed, impurity count and grid description takes such a field
 * (not yet used by another entry) instead of drawing and binning it again.
 * Fields are freed by es_batch_destroy.
 */
int es_batch_clear(es_batch* b);
/*
 * Removes every entry and forgets what drew the kept fields: later entries
 * draw and bin their fields again, into the device memory the batch holds.
 * (No counterpart in the reference, whose runner draws each sample's field
 * once, SimulationRunner.cpp:88-98; for callers that run many batches of fresh
 * samples without freeing and allocating per batch.)
 */
int es_batch_reset(es_batch* b);
int es_batch_set_shape(es_batch* b, const es_batch_shape* shape);
int es_batch_get_shape(es_batch* b, es_batch_shape* shape);
/* enable != 0: record the stage launches once in a hipGraph and replay it. */
int es_batch_set_capture(es_batch* b, int32_t enable);
/* ES_PRECISION_FP64 (default) or ES_PRECISION_MIXED for every entry. */
int es_batch_set_precision(es_batch* b, int32_t precision);
/*
 * Runs every entry: out holds, entry by entry, n_fields * n_temperatures
 * sigmas (field-major); metrics (or NULL) one es_metrics per (entry, field);
 * seconds (or NULL) the device time of the whole stage sequence.  Entries
 * with more than 8 temperatures take the temperature stage once per 8 of
 * them, on the same traced lifetimes (the reference takes any number of
 * temperatures, SimulationConfiguration.h:186-196).
 */
int es_batch_run(es_batch* b, es_sigma* out, es_metrics* metrics, double* seconds);
/*
 * es_batch_run in two halves: es_batch_launch enqueues the run on the batch's
 * own stream and returns (it waits only for the read-backs of newly added
 * fields); es_batch_collect waits for it and fills the same outputs.  Between
 * them the host may prepare another batch, whose work then queues behind or
 * beside this one on the GPU; this batch refuses other calls until collected.
 * (No counterpart in the reference, whose ComputeSigmas blocks,
 * SimulationCL.cpp:310-341; for callers that pipeline samples.)
 */
int es_batch_launch(es_batch* b);  /* up to 8 temperatures per entry (else ES_ERR_INVALID_ARG) */
int es_batch_collect(es_batch* b, es_sigma* out, es_metrics* metrics, double* seconds);
/*
 * Enqueues the drawing, binning and start-cell tests of the fields added since
 * the last run, and their read-backs, on the batch's stream, and returns
 * without waiting; es_batch_launch (or es_batch_run) then only waits for those
 * read-backs.  Calling it before another batch's trace is enqueued lets this
 * batch's binning run ahead instead of queuing behind that trace.  A no-op when
 * nothing is pending.  (No counterpart in the reference, which bins on the host
 * when a grid is built, Grid.cpp:28-59; for callers that pipeline samples.)
 */
int es_batch_prepare(es_batch* b);
/* Sizes of the batch: entries, (entry, field) pairs, sigma outputs, traced work items. */
int es_batch_counts(es_batch* b, int32_t* n_samples, int32_t* n_pairs, int32_t* n_outputs, int64_t* traced_items);
/*
 * How a run of the batch is launched: launch segments (runs of whole (entry,
 * field) pairs sharing the per-item buffers), kernel launches per run, and
 * graph launches per run (1 with es_batch_set_capture, the kernels then
 * replayed from one hipGraph; else 0).  No reference counterpart (the
 * reference launches and waits per sample, pass and field,
 * SimulationCL.cpp:113-177).
 */
int es_batch_launches(es_batch* b, int32_t* segments, int32_t* kernels_per_run, int32_t* graph_launches_per_run);
/* Times every trace launch shape on this batch and keeps the fastest. */
int es_batch_tune(es_batch* b, es_batch_shape* best, double* best_trace_seconds);

/*
 * Binning without device allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory at the point of the call (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It bins the n impurity centres x, y (device arrays) as
 * es_grid_set_impurities + es_grid_read do (Grid.cpp:48-59, 77-180) into
 * caller-provided device buffers: ordered_xy (capacity 9 n (x, y) pairs: a disk
 * lands in at most 9 cells), imp_index (cells_per_row^2 int32, inclusive ends)
 * and indexed_count (one int64).  Work is enqueued on `stream` (a hipStream_t;
 * NULL for the default stream) and the call returns without waiting;
 * out_info (may be NULL) receives the geometry, with indexed_impurity_count -1
 * (the count is on the device).  Its scratch is a fixed array in the library
 * (about 16 MB per GPU, held until the process ends): at most
 * ES_NOALLOC_MAX_IMPURITIES centres and ES_NOALLOC_MAX_CELLS cells, larger
 * inputs are ES_ERR_INVALID_ARG.  Calls that overlap on one GPU share that
 * scratch and race: serialise them (one stream, or wait between them).
 * The message of a failure is es_last_error(NULL).
 */
#define ES_NOALLOC_MAX_IMPURITIES 131072
#define ES_NOALLOC_MAX_CELLS 65536
int es_grid_bin_noalloc(const double* x, const double* y, int64_t n, const es_grid_desc* desc, double* ordered_xy,
                        int32_t* imp_index, int64_t* indexed_count, es_grid_info* out_info, void* stream);

/*
 * The device's elementary functions (elasticscattering_amd/csrc/es_math.h:
 * correctly rounded sin, cos, asin, exp, the ones the kernels call) applied to
 * n host values on `device`; y receives the results.  No reference
 * counterpart: the reference calls the C runtime's libm (details.h:100,
 * lifetime.h:43, util.h:34-46); this lets a caller compare the device's
 * results with any other implementation.  The message of a failure is
 * es_last_error(NULL).
 */
#define ES_MATH_SIN 0
#define ES_MATH_COS 1
#define ES_MATH_ASIN 2
#define ES_MATH_EXP 3
int es_math_eval(int32_t device, int32_t fn, const double* x, double* y, int64_t n);

/* Build/version string of the library. */
const char* es_version(void);

#ifdef __cplusplus
}
#endif

#endif /* ES_ABI_H */
