// es_kernels.h — launchers of the gfx950 kernels (es_kernels.hip).
//
// The lifetime and temperature stages work on a table of "pairs": one pair is
// one (sample, pass, field).  A single ComputeSigmas call is one sample with
// one pair per field; a batch is many samples of uneven size in one launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace es {

constexpr int kBlock = 256;        // threads per block (4 wave64)
constexpr int kMaxT = 8;           // temperatures per sample
constexpr int kSigmaItemsPerThread = 32;  // temperature stage: work items per thread (block count rule)
// int64 metric partials per block: 0 cells, 1 tested, 2 escaped, 3 at_bound,
// 4 cap hits, 5 max tested (max), 6 max cells (max), 7 traced particles,
// 8 edge hits, 9 cross-time calls, 10 sigma evaluations, 11 exact cross-time
// solves, 12 items traced again in FP64 (mixed precision).
constexpr int kMetricSlots = 13;

// Binned field on the device (cell-ordered (x, y) pairs; reference index layout).
struct GridDev {
    double2* ixy = nullptr;
    int32_t* cell_end = nullptr;   // inclusive end per cell
    int64_t total = 0;             // indexed impurities
    double lo = 0, hi = 0, size = 0, cell_size = 0, radius = 0;
    int cpr = 0;
};

// One sample and pass: its field, positions and particle settings.
struct SampleDev {
    const double2* ixy;
    const int32_t* cell_end;
    double start, size, L, radius;   // ImpuritySettings (settings.h:36-45)
    int cpr;
    const int32_t* free_pos;         // position ids j*ppr + i outside every impurity
    int nfree;
    int ppr, ppq;
    // Phi points per quadrant that are traced: ppq, or 1 when the window has
    // zero width (phi_step 0, the coherent pass at alpha = pi/4): its ppq
    // particles per quadrant are then the same particle, traced once.
    int tpq;
    double dist, offx, offy;         // distance_between_positions, small_offset
    double phi_start, phi_step, alpha, speed;
    double phi_integrand_factor;     // Simulation.h:116-120
    double inv_L, inv_size, inv_speed;  // RN(1 / cell size), RN(1 / size), RN(1 / speed)
    double cell_population;          // indexed impurities per cell (host: picks the trace variant)
    // (sin, cos) of phi = phi_start + q pi/2 + p phi_step for k = q ppq + p,
    // k < 4 ppq (esm_sincos on the host: the device's bits).
    const double2* phi_sc;
    int coherent;
    int nT;
    double tau[kMaxT];
    double max_lt[kMaxT];            // 15 * tau (Simulation.h:91)
    double inv_tau[kMaxT];           // RN(1 / tau), for esm_div_rcp
};

// One (sample, pass, field).
struct PairDev {
    int sample;
    int out;                         // this pair's row in the outputs
    double w;                        // angular speed E*B/M
    double R, inv_R, inv_w;          // orbit radius speed / w, RN(1 / R), RN(1 / w) (set by the host)
    long long item_off;              // first work item (ascending over pairs)
    long long nitems;                // traced work items: nfree * 4 * tpq
    long long nall;                  // lifetimes: nfree * 4 * ppq (set by the host with the offsets)
    int blk_off;                     // first temperature-stage block (ascending)
    int nblk;
};

struct TraceArgs {
    const SampleDev* samples;
    const PairDev* pairs;
    int npairs;
    long long total_items;
    unsigned long long* queue;       // work counter, reset by the launcher
    int chunk;                       // items a wave takes from the queue at a time
    double* lt_item;                 // raw lifetime per item
    int4* m_item;                    // cells, tested, edge hits, cross calls
    int32_t* x_item;                 // exact cross-time solves
    uint8_t* f_item;                 // ES_F_* flags
    double* lifetimes;               // [pair][P] reference layout, or null
    long long P;                     // ppr^2 * 4 * ppq when lifetimes is set
    // Mixed precision: the FP32 trace appends the items it cannot decide to
    // `list` (count in *list_count); the FP64 trace with list_mode = 1 takes
    // its work items from that list.
    int32_t* list;
    int32_t* list_count;
    int list_mode;
};

struct SigmaArgs {
    const SampleDev* samples;
    const PairDev* pairs;
    int npairs;
    int total_blocks;
    int clockwise;
    const double* lt_item;
    const int4* m_item;
    const int32_t* x_item;
    const uint8_t* f_item;
    int with_metrics;
    double* dpart;                   // [block][2*NT+1][2] double-double partials
    int64_t* mpart;                  // [block][kMetricSlots]
    // Cleared by the stage for the next trace (it runs after this run's trace).
    unsigned long long* queue;
    int32_t* list_count;             // or null
};

// Binning (Grid.cpp:77-180) into cell-ordered (x, y) + inclusive ends.
struct BinScratch {
    int32_t* targets;   // 9 per impurity
    uint8_t* ntarg;     // per impurity
    int32_t* counts;    // per cell; zero on entry, left zero
    int32_t* slot;      // 9 per impurity: its entry's rank in each target cell
    int64_t* begin;     // per cell + 1
    int64_t* bsum;      // per scan block
    int32_t* order;     // per indexed impurity
};

// One field to bin: its centres (drawn first from es_rng.h with `key` over
// [gen_lo, gen_hi) when `generate`), geometry, scratch and outputs.
struct BinJob {
    double* x;
    double* y;
    int64_t n;
    int generate;
    uint64_t key;
    double gen_lo, gen_hi;
    double lo, size, cell_size, radius;
    int cpr;
    BinScratch s;
    int32_t* cell_end;
    double2* ixy;
    int64_t* total_out;  // indexed count (device memory)
};

// One start-cell test (lifetime.h:84-99) of a ppr x ppr position grid on a
// binned field, with the ordered list of free positions.
struct PosJob {
    const double2* ixy;
    const int32_t* cell_end;
    double lo, size, radius;
    int cpr;
    int ppr;
    double dist, offx, offy;
    uint8_t* flags;      // per position
    int32_t* bcount;     // per 256 positions: free count, bit 31 = a position outside the grid
    int32_t* err;        // out: 1 when a position lies outside the cell grid
    int32_t* free_pos;
    int32_t* nfree;      // out
};

// Bins every job's field with one launch per stage for all of them; `jobs_dev`
// is the job table in device memory, `jobs` the same table on the host.
hipError_t launch_bin_jobs(const BinJob* jobs_dev, const BinJob* jobs, int njobs, hipStream_t st);
// Start-cell tests of every job (after the binning of their fields).
hipError_t launch_positions_jobs(const PosJob* jobs_dev, const PosJob* jobs, int njobs, hipStream_t st);

// Trace kernel variants: register budget 2 or 3 waves per SIMD; coop = 0 for
// one lane per particle throughout, 1 for wave-cooperative crossing solves, 2
// for wave-cooperative and screened (disks that cannot lower the minimum are
// not solved).  The FP32 trace has coop 0 and 1 (2 counts as 1).
hipError_t trace_blocks_per_cu(int clockwise, int waves_per_eu, int coop, int* blocks);
// Traces; the work queue must be zero (the temperature stage clears it after
// every run).  In list mode (the FP64 recompute of the mixed path, after the
// FP32 trace has used the queue) the queue is reset by a kernel first, so the
// sequence can be captured in a graph.
hipError_t launch_trace(const TraceArgs& a, int clockwise, int waves_per_eu, int coop, int nblocks, hipStream_t st);
// Mixed precision: FP32 trace of every item, then FP64 trace of the flagged ones.
hipError_t trace32_blocks_per_cu(int clockwise, int coop, int* blocks);
hipError_t launch_trace_mixed(const TraceArgs& a, int clockwise, int waves_per_eu, int coop, int nblocks32,
                              int nblocks64, hipStream_t st);
hipError_t launch_sigma(const SigmaArgs& a, int NT, hipStream_t st);

// Images mode (SimulationCPU.cpp:159-171, 224-243): for one pair, per free
// position and temperature, the phi-integrated truncated lifetime, sigma_xx and
// sigma_xy: img[t][0..2][j * ppr + i]; positions inside an impurity stay 0.
hipError_t launch_images(const SampleDev* sample, const PairDev* pair, int clockwise, const double* lt_item,
                         const double* tau, const double* max_lt, int nT, int nfree, int ppr, double* img,
                         hipStream_t st);

// One block per pair: out_sig [out][kMaxT][2], out_lt_sum [out], out_m [out][kMetricSlots].
hipError_t launch_finalize(const SampleDev* samples, const PairDev* pairs, int npairs, int NT, int with_metrics,
                           const double* dpart, const int64_t* mpart, const double* factor /* [out][kMaxT] */,
                           double* out_sig, double* out_lt_sum, int64_t* out_m, hipStream_t st);

}  // namespace es
