// es_kernels.h — launchers of the gfx950 kernels (es_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace es {

constexpr int kBlock = 256;        // threads per block of the trace kernel (4 wave64)
constexpr int kMaxT = 8;           // temperatures per trace launch
// int64 metric partials per block: 0 cells, 1 tested, 2 escaped, 3 at_bound,
// 4 cap hits, 5 max tested (max), 6 max cells (max), 7 traced particles,
// 8 edge hits, 9 cross-time calls, 10 sigma evaluations.
constexpr int kMetricSlots = 11;

// Binned field on the device (SoA, cell-ordered; reference index layout).
struct GridDev {
    double* ix = nullptr;
    double* iy = nullptr;
    int32_t* cell_end = nullptr;   // inclusive end per cell
    int64_t total = 0;             // indexed impurities
    double lo = 0, hi = 0, size = 0, cell_size = 0, radius = 0;
    int cpr = 0;
};

// Inputs of one trace launch (one sample, one pass, nB fields, nT temperatures).
struct TraceArgs {
    // grid
    const double* ix;
    const double* iy;
    const int32_t* cell_end;
    double g_start, g_size, g_L, g_radius;
    int g_cpr;
    // positions
    const int32_t* free_pos;       // position ids j*ppr + i outside every impurity
    int nfree;
    int ppr, ppq;
    double dist, offx, offy;
    // particle settings
    double phi_start, phi_step, alpha, speed;
    int coherent;
    int clockwise;
    // fields and temperatures
    int nB;
    const double* w;               // [nB] angular speed E*B/M
    int nT;
    double tau[kMaxT];
    double max_lt[kMaxT];          // 15 * tau
    // outputs
    double* lifetimes;             // [nB][P] or null
    double* dpart;                 // [nB][nblk][2*nT+1][2] double-double partials
    int64_t* mpart;                // [nB][nblk][kMetricSlots]
    int nblk;
};

hipError_t launch_generate(uint64_t key, int64_t n, double lo, double hi, double* x, double* y, hipStream_t st);

// Binning (Grid.cpp:77-180) into cell-ordered SoA + inclusive ends.  The
// scratch buffers are sized by the caller: targets 9n ints, ntarg n bytes,
// counts/fill ncell ints, begin ncell+1 ints (int64 for totals), order 9n ints.
struct BinScratch {
    int32_t* targets;
    uint8_t* ntarg;
    int32_t* counts;
    int32_t* fill;
    int64_t* begin;
    int32_t* order;
};
hipError_t launch_bin_count(const double* x, const double* y, int64_t n, const GridDev& g, BinScratch& s,
                            hipStream_t st);
hipError_t launch_bin_scan(const GridDev& g, BinScratch& s, hipStream_t st);
hipError_t launch_bin_place(const double* x, const double* y, int64_t n, GridDev& g, BinScratch& s,
                            hipStream_t st);

// Start-cell test per position (lifetime.h:84-99) and ordered compaction.
hipError_t launch_positions(const GridDev& g, int ppr, double dist, double offx, double offy, uint8_t* flags,
                            int32_t* err, int32_t* free_pos, int32_t* nfree, hipStream_t st);

hipError_t launch_trace(const TraceArgs& a, hipStream_t st);

// Sum the per-block partials: out_sig [nB][nT][2], out_lt_sum [nB], out_m [nB][kMetricSlots].
hipError_t launch_finalize(const double* dpart, const int64_t* mpart, int nB, int nblk, int nT,
                           const double* factor /* [nB][nT] device */, double* out_sig, double* out_lt_sum,
                           int64_t* out_m, hipStream_t st);

}  // namespace es
