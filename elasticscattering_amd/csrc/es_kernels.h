// es_kernels.h — launchers of the gfx950 kernels (es_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace es {

constexpr int kBlock = 256;        // threads per block (4 wave64)
constexpr int kMaxT = 8;           // temperatures per sigma launch
// int64 metric partials per block: 0 cells, 1 tested, 2 escaped, 3 at_bound,
// 4 cap hits, 5 max tested (max), 6 max cells (max), 7 traced particles,
// 8 edge hits, 9 cross-time calls, 10 sigma evaluations, 11 exact cross-time solves.
constexpr int kMetricSlots = 12;

// Binned field on the device (cell-ordered (x, y) pairs; reference index layout).
struct GridDev {
    double2* ixy = nullptr;
    int32_t* cell_end = nullptr;   // inclusive end per cell
    int64_t total = 0;             // indexed impurities
    double lo = 0, hi = 0, size = 0, cell_size = 0, radius = 0;
    int cpr = 0;
};

// Lifetime stage: one work item per (field, free position, phi).
struct TraceArgs {
    const double2* ixy;
    const int32_t* cell_end;
    double g_start, g_size, g_L, g_radius;
    int g_cpr;
    const int32_t* free_pos;       // position ids j*ppr + i outside every impurity
    int nfree;
    int ppr, ppq;
    double dist, offx, offy;
    double phi_start, phi_step, alpha, speed;
    int coherent;
    int clockwise;
    int nB;
    const double* w;               // [nB] angular speed E*B/M
    unsigned long long* queue;     // work counter, 0 at launch
    int chunk;                     // items taken from the queue at a time per wave
    // outputs per item g = b * nitems + k
    double* lt_item;               // raw lifetime
    int4* m_item;                  // cells, tested, edge hits, cross calls
    int32_t* x_item;               // exact cross-time solves
    uint8_t* f_item;               // ES_F_* flags
    double* lifetimes;             // [nB][P] reference layout, or null
};

// Temperature stage per field: truncation, GetSigma, Simpson weights, sums.
struct SigmaArgs {
    const double* lt_item;
    const int4* m_item;
    const int32_t* x_item;
    const uint8_t* f_item;
    const int32_t* free_pos;
    int nfree;
    int ppr, ppq;
    double phi_start, phi_step;
    int clockwise;
    int nB;
    const double* w;
    int nT;
    double tau[kMaxT];
    double max_lt[kMaxT];          // 15 * tau
    int with_metrics;              // reduce the per-item work counters too
    double* dpart;                 // [nB][nblk][2*nT+1][2] double-double partials
    int64_t* mpart;                // [nB][nblk][kMetricSlots]
    int nblk;
};

hipError_t launch_generate(uint64_t key, int64_t n, double lo, double hi, double* x, double* y, hipStream_t st);

// Binning (Grid.cpp:77-180) into cell-ordered (x, y) + inclusive ends.
struct BinScratch {
    int32_t* targets;   // 9 per impurity
    uint8_t* ntarg;     // per impurity
    int32_t* counts;    // per cell
    int32_t* fill;      // per cell
    int64_t* begin;     // per cell + 1
    int64_t* bsum;      // per scan block
    int32_t* order;     // per indexed impurity
};
hipError_t launch_bin_count(const double* x, const double* y, int64_t n, const GridDev& g, BinScratch& s,
                            hipStream_t st);
hipError_t launch_bin_scan(const GridDev& g, BinScratch& s, hipStream_t st);
hipError_t launch_bin_place(const double* x, const double* y, int64_t n, GridDev& g, BinScratch& s,
                            hipStream_t st);

// Start-cell test per position (lifetime.h:84-99) and ordered compaction.
// bcount: one int per 256 positions.
hipError_t launch_positions(const GridDev& g, int ppr, double dist, double offx, double offy, uint8_t* flags,
                            int32_t* bcount, int32_t* err, int32_t* free_pos, int32_t* nfree, hipStream_t st);

hipError_t launch_trace(const TraceArgs& a, int nblocks, hipStream_t st);
// Resident blocks per CU of the trace kernel (for sizing its persistent grid).
hipError_t trace_blocks_per_cu(int clockwise, int* blocks);
hipError_t launch_sigma(const SigmaArgs& a, hipStream_t st);

// Sum the per-block partials: out_sig [nB][nT][2], out_lt_sum [nB], out_m [nB][kMetricSlots].
hipError_t launch_finalize(const double* dpart, const int64_t* mpart, int nB, int nblk, int nT, int with_metrics,
                           const double* factor /* [nB][nT] device */, double* out_sig, double* out_lt_sum,
                           int64_t* out_m, hipStream_t st);

}  // namespace es
