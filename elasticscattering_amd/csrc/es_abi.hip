// es_abi.hip — the C ABI (include/es_abi.h).
//
// es_solver: one sample at a time, the operator API of the reference's
//   Simulation (src/sim/Simulation.h:56-147) in the order of
//   Simulation::ComputeSigmas (src/sim/SimulationCPU.cpp:33-73).
// es_batch: many samples of uneven size (own fields, grids, fields B, T) in one
//   launch of each stage, optionally replayed from a captured hipGraph, with a
//   tuner for the trace kernel's launch shape.
// Both own their device memory and one HIP stream, and share the stage runner
// below (binning, start-cell test, trace, temperature stage, final sums).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/es_abi.h"
#include "es_kernels.h"
#include "es_math.h"
#include "es_rng.h"

namespace {

// constants.h:23-38
constexpr double kM0 = 9.10938188e-31;
constexpr double kM = (5.0 * kM0);
constexpr double kE = 1.60217646e-19;
constexpr double kHBAR = 1.05457148e-34;
constexpr double kC1 = 1.15e-9;
constexpr double kKB = 1.380649e-23;
constexpr double kPI = 3.141592653589793238463;
constexpr double kHALF_PI = 1.570796326794896619231;

thread_local std::string g_create_error;

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

template <typename T>
T* P(DevBuf& b)
{
    return reinterpret_cast<T*>(b.p);
}

// Stream, events, allocation accounting and the last error of one owner.
struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    size_t byte_cap = 0;  // 0: unlimited; else refuse allocations beyond this total
    size_t bytes_held = 0;
    int cus = 0;
    int trace_bpc[2][3][4] = {{{0}}};  // resident trace blocks per CU [clockwise][coop][waves_per_eu]
    int trace32_bpc[2][2] = {{0, 0}, {0, 0}};  // resident FP32 trace blocks per CU [clockwise][coop]
    // Work items per launch segment: bounds the per-item buffers (29 B per item, 31 GB at this size) and
    // keeps the mixed path's int32 recompute list in range.
    long long seg_items = 1ll << 30;
    std::vector<hipEvent_t> ev;       // 4 per segment: trace start/end, temperature start/end
    void* hbuf = nullptr;             // pinned host staging of the stage tables and results
    size_t hbuf_cap = 0;
    DevBuf jobs;                      // binning / start-test job tables (device)
    void* hjobs = nullptr;            // their pinned staging
    size_t hjobs_cap = 0;
    hipEvent_t jobs_ev = nullptr;     // recorded after the last job-table upload
    bool jobs_copy_pending = false;
};



int fail(Ctx& c, int code, const std::string& msg)
{
    c.err = msg;
    return code;
}

int hip_fail(Ctx& c, hipError_t e, const char* what)
{
    return fail(c, ES_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Grows the event pool to n events (not while a stream is being captured).
int ctx_events(Ctx& c, size_t n)
{
    while (c.ev.size() < n) {
        hipEvent_t e = nullptr;
        const hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) return fail(c, ES_ERR_HIP, std::string("hipEventCreate: ") + hipGetErrorString(r));
        c.ev.push_back(e);
    }
    return ES_OK;
}

#define ES_HIP(c, call)                                        \
    do {                                                       \
        hipError_t _e = (call);                                \
        if (_e != hipSuccess) return hip_fail(c, _e, #call);   \
    } while (0)

int ensure(Ctx& c, DevBuf& b, size_t bytes, const char* what)
{
    if (bytes == 0) bytes = 8;
    if (b.cap >= bytes) return ES_OK;
    if (c.byte_cap && c.bytes_held - b.cap + bytes > c.byte_cap)
        return fail(c, ES_ERR_OOM, std::string("device size cap exceeded allocating ") + what);
    if (b.p) {
        // Work already enqueued may still use the old buffer.
        if (c.stream) (void)hipStreamSynchronize(c.stream);
        (void)hipFree(b.p);
        c.bytes_held -= b.cap;
        b.p = nullptr;
        b.cap = 0;
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, ES_ERR_OOM, std::string("hipMalloc failed for ") + what + ": " + hipGetErrorString(e));
    }
    b.cap = bytes;
    c.bytes_held += bytes;
    return ES_OK;
}

// Grows the pinned staging buffer (only between synchronised points: copies
// from or into it must not be in flight).
int ensure_host(Ctx& c, size_t bytes)
{
    if (c.hbuf_cap >= bytes) return ES_OK;
    if (c.hbuf) (void)hipHostFree(c.hbuf);
    c.hbuf = nullptr;
    c.hbuf_cap = 0;
    const size_t cap = std::max<size_t>(bytes, 64 * 1024);
    hipError_t e = hipHostMalloc(&c.hbuf, cap, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        c.hbuf = nullptr;
        return fail(c, ES_ERR_OOM, std::string("hipHostMalloc failed for staging: ") + hipGetErrorString(e));
    }
    c.hbuf_cap = cap;
    return ES_OK;
}

void release(Ctx& c, DevBuf& b)
{
    if (b.p) (void)hipFree(b.p);
    c.bytes_held -= std::min(c.bytes_held, b.cap);
    b.p = nullptr;
    b.cap = 0;
}

int ctx_init(Ctx& c, int device, std::string& err)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        err = "no HIP device";
        return ES_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= ndev) {
        err = "device ordinal out of range";
        return ES_ERR_INVALID_ARG;
    }
    c.device = device;
    if (const char* cap = std::getenv("ES_MAX_DEVICE_BYTES")) c.byte_cap = (size_t)std::strtoull(cap, nullptr, 10);
    if (const char* si = std::getenv("ES_SEGMENT_ITEMS")) {
        const long long v = std::strtoll(si, nullptr, 10);
        if (v > 0) c.seg_items = v;
    }
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking);
    for (int i = 0; i < 4 && e == hipSuccess; i++) {
        hipEvent_t ev = nullptr;
        e = hipEventCreate(&ev);
        if (e == hipSuccess) c.ev.push_back(ev);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c.jobs_ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c.cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) {
        err = std::string("HIP setup failed: ") + hipGetErrorString(e);
        return ES_ERR_HIP;
    }
    return ES_OK;
}

void ctx_destroy(Ctx& c)
{
    (void)hipSetDevice(c.device);
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    for (hipEvent_t ev : c.ev)
        if (ev) (void)hipEventDestroy(ev);
    c.ev.clear();
    if (c.hbuf) (void)hipHostFree(c.hbuf);
    c.hbuf = nullptr;
    c.hbuf_cap = 0;
    release(c, c.jobs);
    if (c.hjobs) (void)hipHostFree(c.hjobs);
    c.hjobs = nullptr;
    c.hjobs_cap = 0;
    if (c.jobs_ev) (void)hipEventDestroy(c.jobs_ev);
    c.jobs_ev = nullptr;
    if (c.stream) (void)hipStreamDestroy(c.stream);
}

// Grid.cpp:40,54 with the optional explicit count.
int cells_per_row_rule(int64_t unique, int target, int override_cpr)
{
    if (override_cpr > 0) return override_cpr;
    const int cpr = (int)std::round(std::sqrt(unique / (double)target));
    return std::max(cpr, 1);
}

// ------------------------------------------------------------------ field

// One binned impurity field and its start-cell test (Grid + the position part
// of TraceOrbit, lifetime.h:84-99, hoisted: it depends on neither phi nor B).
struct Field {
    DevBuf in_x, in_y, ixy, cell_end;
    DevBuf b_targets, b_ntarg, b_counts, b_slot, b_begin, b_bsum, b_order;
    DevBuf flags, bcount, free_pos, small;  // small: [0] out-of-grid flag, [1] nfree
    es::GridDev g;
    int64_t unique = 0;
    bool has_grid = false;
    // start test of the current positions
    bool pos_valid = false;
    int pos_ppr = 0;
    double pos_dist = 0, pos_offx = 0, pos_offy = 0;
    int nfree = 0;
    int64_t n_inside = 0;
    // Read-backs in flight, collected by field_sync (pinned copy of `small`).
    int64_t* hsmall = nullptr;
    bool total_pending = false;
    bool pos_pending = false;
    bool copy_queued = false;   // the copy into hsmall is enqueued behind the last launch
    bool counts_clean = false;  // b_counts all zero (left so by every binning's scan)
};

void field_release(Ctx& c, Field& f)
{
    DevBuf* bufs[] = {&f.in_x, &f.in_y, &f.ixy, &f.cell_end, &f.b_targets, &f.b_ntarg, &f.b_counts, &f.b_slot,
                      &f.b_begin, &f.b_bsum, &f.b_order, &f.flags, &f.bcount, &f.free_pos, &f.small};
    for (DevBuf* b : bufs) release(c, *b);
    if (f.hsmall) (void)hipHostFree(f.hsmall);
    f.hsmall = nullptr;
}

int field_host_small(Ctx& c, Field& f)
{
    if (f.hsmall) return ES_OK;
    void* p = nullptr;
    hipError_t e = hipHostMalloc(&p, 64, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, ES_ERR_OOM, std::string("hipHostMalloc failed: ") + hipGetErrorString(e));
    }
    f.hsmall = static_cast<int64_t*>(p);
    return ES_OK;
}

// Device word block `small` (16 B): [0] out-of-grid flag (int32), [1] free
// positions (int32), [2..3] indexed impurity count (int64); f.hsmall is its
// pinned host copy.
int field_small(Ctx& c, Field& f)
{
    int rc;
    if ((rc = ensure(c, f.small, 16, "field read-back words"))) return rc;
    return field_host_small(c, f);
}

// Enqueues the copy of the field's read-backs in flight (the indexed count of
// a binning, the start test of the positions) behind the launches that
// produce them; field_sync_finish takes them once the stream has reached it.
int field_sync_enqueue(Ctx& c, Field& f)
{
    if ((f.total_pending || f.pos_pending) && !f.copy_queued) {
        ES_HIP(c, hipMemcpyAsync(f.hsmall, f.small.p, 16, hipMemcpyDeviceToHost, c.stream));
        f.copy_queued = true;
    }
    return ES_OK;
}

// With pos_errors = false a start position outside the grid only leaves the
// positions invalid (the next field_positions reports it).
int field_sync_finish(Ctx& c, Field& f, bool pos_errors = true)
{
    f.copy_queued = false;
    const int32_t* h = reinterpret_cast<const int32_t*>(f.hsmall);
    if (f.total_pending) {
        f.total_pending = false;
        const int64_t total = f.hsmall[1];
        if (total > INT32_MAX) return fail(c, ES_ERR_INVALID_ARG, "indexed impurity count exceeds int32");
        f.g.total = total;
    }
    if (f.pos_pending) {
        f.pos_pending = false;
        if (h[0] != 0 && !pos_errors) return ES_OK;
        if (h[0] != 0)
            return fail(c, ES_ERR_OUT_OF_GRID,
                        "a start position lies outside the impurity cell grid (region_extends smaller than the "
                        "position offset); the reference would read out of bounds (lifetime.h:88-91)");
        const int64_t n = (int64_t)f.pos_ppr * f.pos_ppr;
        f.nfree = h[1];
        f.n_inside = n - h[1];
        f.pos_valid = true;
    }
    return ES_OK;
}

// The read-backs in flight with one copy and one wait.
int field_sync(Ctx& c, Field& f, bool pos_errors = true)
{
    int rc;
    if ((rc = field_sync_enqueue(c, f))) return rc;
    ES_HIP(c, hipStreamSynchronize(c.stream));
    return field_sync_finish(c, f, pos_errors);
}

// Binnings and start-cell tests waiting to be launched together (one launch
// per stage for all their fields), and the fields whose read-backs they fill.
struct Jobs {
    std::vector<es::BinJob> bin;
    std::vector<es::PosJob> pos;
    std::vector<Field*> fields;
    bool empty() const { return bin.empty() && pos.empty(); }
    void add_field(Field* f)
    {
        if (std::find(fields.begin(), fields.end(), f) == fields.end()) fields.push_back(f);
    }
};

// Uploads the job tables (one copy from pinned staging), launches the
// binning and then the start tests, and enqueues each field's read-back copy
// behind them; field_sync_finish takes the read-backs after the next wait.
int jobs_launch(Ctx& c, Jobs& q)
{
    if (q.empty()) return ES_OK;
    const size_t bb = q.bin.size() * sizeof(es::BinJob), pb = q.pos.size() * sizeof(es::PosJob);
    const size_t off = (bb + 255) & ~(size_t)255, bytes = off + pb;
    int rc;
    if ((rc = ensure(c, c.jobs, bytes, "binning job table"))) return rc;
    // The staging may still feed the previous upload: wait for that copy first.
    if (c.jobs_copy_pending) ES_HIP(c, hipEventSynchronize(c.jobs_ev));
    if (c.hjobs_cap < bytes) {
        if (c.hjobs) (void)hipHostFree(c.hjobs);
        c.hjobs = nullptr;
        c.hjobs_cap = 0;
        const size_t cap = std::max<size_t>(bytes, 16 * 1024);
        hipError_t e = hipHostMalloc(&c.hjobs, cap, hipHostMallocDefault);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            c.hjobs = nullptr;
            return fail(c, ES_ERR_OOM, std::string("hipHostMalloc failed for the job table: ") + hipGetErrorString(e));
        }
        c.hjobs_cap = cap;
    }
    char* h = static_cast<char*>(c.hjobs);
    if (bb) std::memcpy(h, q.bin.data(), bb);
    if (pb) std::memcpy(h + off, q.pos.data(), pb);
    ES_HIP(c, hipMemcpyAsync(c.jobs.p, h, bytes, hipMemcpyHostToDevice, c.stream));
    ES_HIP(c, hipEventRecord(c.jobs_ev, c.stream));
    c.jobs_copy_pending = true;
    const es::BinJob* dbin = P<es::BinJob>(c.jobs);
    const es::PosJob* dpos = reinterpret_cast<const es::PosJob*>(P<char>(c.jobs) + off);
    ES_HIP(c, es::launch_bin_jobs(dbin, q.bin.data(), (int)q.bin.size(), c.stream));
    ES_HIP(c, es::launch_positions_jobs(dpos, q.pos.data(), (int)q.pos.size(), c.stream));
    for (Field* f : q.fields) {
        f->copy_queued = false;
        if ((rc = field_sync_enqueue(c, *f))) return rc;
    }
    q = Jobs();
    return ES_OK;
}

// Bins the n centres in f.in_x / f.in_y (Grid.cpp:48-59, 77-180), drawing them
// first when `generate` (key, range [lo, hi)): appends the job to q.  The
// indexed count is read back by field_sync after jobs_launch.
int field_bin(Ctx& c, Field& f, int64_t n, const es_grid_desc* d, Jobs& q, bool generate = false, uint64_t key = 0,
              double lo = 0.0, double hi = 0.0)
{
    if (d->target_cell_population <= 0 && d->cells_per_row <= 0)
        return fail(c, ES_ERR_INVALID_ARG, "target_cell_population must be > 0");
    if (!(d->impurity_radius >= 0)) return fail(c, ES_ERR_INVALID_ARG, "impurity_radius must be >= 0");
    if (!(d->region_size + 2 * d->region_extends > 0)) return fail(c, ES_ERR_INVALID_ARG, "empty spawn region");
    const int cpr = cells_per_row_rule(n, d->target_cell_population, d->cells_per_row);
    if ((int64_t)cpr * cpr > (int64_t)1 << 28) return fail(c, ES_ERR_INVALID_ARG, "cell grid too large");
    es::GridDev& g = f.g;
    g.cpr = cpr;
    g.lo = -d->region_extends;                       // Grid.cpp:52
    g.hi = d->region_size + d->region_extends;
    g.size = g.hi - g.lo;
    g.cell_size = g.size / (double)cpr;              // Grid.cpp:55, Grid.h:89
    g.radius = d->impurity_radius;
    const int64_t ncell = (int64_t)cpr * cpr;
    // Every impurity lands in at most 9 cells (Grid.cpp:112-180), so the
    // binned arrays are sized for that and the count is read back later.
    const int64_t max_total = 9 * n;
    int rc;
    if ((rc = ensure(c, f.b_targets, 9 * n * sizeof(int32_t), "bin targets"))) return rc;
    if ((rc = ensure(c, f.b_slot, 9 * n * sizeof(int32_t), "bin slots"))) return rc;
    if ((rc = ensure(c, f.b_ntarg, n, "bin target counts"))) return rc;
    const size_t had = f.b_counts.cap;
    if ((rc = ensure(c, f.b_counts, ncell * sizeof(int32_t), "cell counts"))) return rc;
    if (f.b_counts.cap != had) f.counts_clean = false;
    if ((rc = ensure(c, f.b_begin, (ncell + 1) * sizeof(int64_t), "cell begin"))) return rc;
    if ((rc = ensure(c, f.b_bsum, ((ncell + 255) / 256) * sizeof(int64_t), "scan block sums"))) return rc;
    if ((rc = ensure(c, f.cell_end, ncell * sizeof(int32_t), "cell index"))) return rc;
    if ((rc = ensure(c, f.b_order, max_total * sizeof(int32_t), "cell order"))) return rc;
    if ((rc = ensure(c, f.ixy, max_total * sizeof(double2), "impurities"))) return rc;
    if ((rc = field_small(c, f))) return rc;
    // The counts start at zero: the scan of every binning leaves them so.
    if (!f.counts_clean) ES_HIP(c, hipMemsetAsync(f.b_counts.p, 0, f.b_counts.cap, c.stream));
    g.cell_end = P<int32_t>(f.cell_end);
    g.ixy = P<double2>(f.ixy);
    es::BinJob j{};
    j.x = P<double>(f.in_x);
    j.y = P<double>(f.in_y);
    j.n = n;
    j.generate = generate ? 1 : 0;
    j.key = key;
    j.gen_lo = lo;
    j.gen_hi = hi;
    j.lo = g.lo;
    j.size = g.size;
    j.cell_size = g.cell_size;
    j.radius = g.radius;
    j.cpr = cpr;
    j.s = es::BinScratch{P<int32_t>(f.b_targets), P<uint8_t>(f.b_ntarg), P<int32_t>(f.b_counts),
                         P<int32_t>(f.b_slot),    P<int64_t>(f.b_begin),  P<int64_t>(f.b_bsum),
                         P<int32_t>(f.b_order)};
    j.cell_end = g.cell_end;
    j.ixy = g.ixy;
    j.total_out = P<int64_t>(f.small) + 1;
    q.bin.push_back(j);
    q.add_field(&f);
    f.counts_clean = true;
    f.total_pending = true;
    f.copy_queued = false;
    f.unique = n;
    f.has_grid = true;
    f.pos_valid = false;
    f.pos_pending = false;
    return ES_OK;
}

int field_upload(Ctx& c, Field& f, const double* x, const double* y, int64_t n, const es_grid_desc* d, Jobs& q)
{
    if (!d || n < 0 || (n > 0 && (!x || !y))) return fail(c, ES_ERR_INVALID_ARG, "bad impurity arguments");
    if (n > 10000000) return fail(c, ES_ERR_INVALID_ARG, "more than 1e7 impurities (Grid.cpp:36)");
    int rc;
    if ((rc = ensure(c, f.in_x, n * sizeof(double), "impurity input x"))) return rc;
    if ((rc = ensure(c, f.in_y, n * sizeof(double), "impurity input y"))) return rc;
    if (n > 0) {
        ES_HIP(c, hipMemcpyAsync(f.in_x.p, x, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
        ES_HIP(c, hipMemcpyAsync(f.in_y.p, y, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    }
    return field_bin(c, f, n, d, q);
}

int field_generate(Ctx& c, Field& f, uint64_t seed, int64_t count, const es_grid_desc* d, Jobs& q)
{
    if (!d) return fail(c, ES_ERR_INVALID_ARG, "null grid description");
    const double lo = -d->region_extends, hi = d->region_size + d->region_extends;
    int64_t n = count;
    if (n <= 0) {  // Grid.cpp:33-39
        const double area_length = hi - lo;
        const double proposal = area_length * area_length * d->impurity_density;
        if (proposal > 1e7) return fail(c, ES_ERR_INVALID_ARG, "impurity count proposal > 1e7 (Grid.cpp:36)");
        if (!(proposal >= 1)) return fail(c, ES_ERR_INVALID_ARG, "no impurities (Grid.cpp:37)");
        n = std::max<int64_t>(1, (int64_t)std::ceil(proposal));
    }
    if (n > 10000000) return fail(c, ES_ERR_INVALID_ARG, "more than 1e7 impurities (Grid.cpp:36)");
    int rc;
    if ((rc = ensure(c, f.in_x, n * sizeof(double), "impurity input x"))) return rc;
    if ((rc = ensure(c, f.in_y, n * sizeof(double), "impurity input y"))) return rc;
    return field_bin(c, f, n, d, q, true, esr_key(seed), lo, hi);
}

// Appends the start cell + inside test of a ppr x ppr position grid to q
// unless the field's current (or pending) test is of the same positions;
// field_sync collects the result after jobs_launch.
// A field binned by the caller (Grid::GetImpurities / GetIndex / GetSettings,
// Grid.h:74-91): the cell-ordered (x, y) pairs with duplicates and the
// inclusive end offsets are uploaded as they are.
int field_set_binned(Ctx& c, Field& f, const double* ordered_xy, const int32_t* imp_index, int cpr, double start,
                     double size, double radius)
{
    if (!imp_index || cpr <= 0) return fail(c, ES_ERR_INVALID_ARG, "null index or cells_per_row <= 0");
    if ((int64_t)cpr * cpr > (int64_t)1 << 28) return fail(c, ES_ERR_INVALID_ARG, "cell grid too large");
    if (!(size > 0) || !(radius >= 0) || !std::isfinite(start))
        return fail(c, ES_ERR_INVALID_ARG, "spawn region size must be > 0, radius >= 0, start finite");
    const int64_t ncell = (int64_t)cpr * cpr;
    int32_t prev = 0;
    for (int64_t k = 0; k < ncell; k++) {  // inclusive ends (Grid.cpp:92-110): non-decreasing from 0
        if (imp_index[k] < prev)
            return fail(c, ES_ERR_INVALID_ARG, "imp_index must be non-decreasing inclusive end offsets");
        prev = imp_index[k];
    }
    const int64_t total = prev;
    if (total > 0 && !ordered_xy) return fail(c, ES_ERR_INVALID_ARG, "null ordered impurities");
    int rc;
    if ((rc = ensure(c, f.ixy, std::max<int64_t>(total, 1) * sizeof(double2), "impurities"))) return rc;
    if ((rc = ensure(c, f.cell_end, ncell * sizeof(int32_t), "cell index"))) return rc;
    if ((rc = field_small(c, f))) return rc;
    if (total > 0)
        ES_HIP(c, hipMemcpyAsync(f.ixy.p, ordered_xy, total * sizeof(double2), hipMemcpyHostToDevice, c.stream));
    ES_HIP(c, hipMemcpyAsync(f.cell_end.p, imp_index, ncell * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
    es::GridDev& g = f.g;
    g.cpr = cpr;
    g.lo = start;
    g.size = size;
    g.hi = start + size;
    g.cell_size = size / (double)cpr;  // Grid::GetSettings (Grid.h:82-91)
    g.radius = radius;
    g.total = total;
    g.ixy = P<double2>(f.ixy);
    g.cell_end = P<int32_t>(f.cell_end);
    f.unique = 0;  // not known from a binned field
    f.has_grid = true;
    f.total_pending = false;
    f.pos_valid = false;
    f.pos_pending = false;
    f.copy_queued = false;
    return ES_OK;
}

int field_positions_job(Ctx& c, Field& f, int ppr, double dist, double offx, double offy, Jobs& q)
{
    const bool same = f.pos_ppr == ppr && f.pos_dist == dist && f.pos_offx == offx && f.pos_offy == offy;
    if (same && (f.pos_valid || f.pos_pending)) return ES_OK;
    const int64_t n = (int64_t)ppr * ppr;
    int rc;
    if ((rc = ensure(c, f.flags, n, "position flags"))) return rc;
    if ((rc = ensure(c, f.free_pos, n * sizeof(int32_t), "free positions"))) return rc;
    if ((rc = ensure(c, f.bcount, ((n + 255) / 256) * sizeof(int32_t), "position block counts"))) return rc;
    if ((rc = field_small(c, f))) return rc;
    int32_t* small = P<int32_t>(f.small);
    es::PosJob j{};
    j.ixy = f.g.ixy;
    j.cell_end = f.g.cell_end;
    j.lo = f.g.lo;
    j.size = f.g.size;
    j.radius = f.g.radius;
    j.cpr = f.g.cpr;
    j.ppr = ppr;
    j.dist = dist;
    j.offx = offx;
    j.offy = offy;
    j.flags = P<uint8_t>(f.flags);
    j.bcount = P<int32_t>(f.bcount);
    j.err = small;
    j.free_pos = P<int32_t>(f.free_pos);
    j.nfree = small + 1;
    q.pos.push_back(j);
    q.add_field(&f);
    f.pos_valid = false;
    f.pos_pending = true;
    f.copy_queued = false;
    f.pos_ppr = ppr;
    f.pos_dist = dist;
    f.pos_offx = offx;
    f.pos_offy = offy;
    return ES_OK;
}

// Start cell + inside test for every position of a ppr x ppr grid (cached).
int field_positions(Ctx& c, Field& f, int ppr, double dist, double offx, double offy)
{
    if (f.pos_valid && !f.pos_pending && f.pos_ppr == ppr && f.pos_dist == dist && f.pos_offx == offx &&
        f.pos_offy == offy && !f.total_pending)
        return ES_OK;
    Jobs q;
    int rc;
    if ((rc = field_positions_job(c, f, ppr, dist, offx, offy, q))) return rc;
    if ((rc = jobs_launch(c, q))) return rc;
    return field_sync(c, f);
}

// ----------------------------------------------------------------- stages

// Simulation.h:69-76
double sigma_factor(double speed, double w, double tau)
{
    const double kf = kM * speed / kHBAR;
    double outside = (kE * kE * kf * kf) / (2.0 * kPI * kPI * kM * kC1);
    const double wct = w * tau;
    outside *= tau / (1.0 + wct * wct);
    return outside;
}

// Simulation.h:88
double integrand_factor(int ppr, double phi_integrand_factor, double speed, double w, double tau)
{
    const double a = 1.0 / (3.0 * (ppr - 1));
    return a * a * phi_integrand_factor * sigma_factor(speed, w, tau) * 1e-8;
}

// InitSample quantities (Simulation.h:111-126).
struct PassSettings {
    double integrand_angle_area, phi_integrand_factor, phi_start, phi_step;
};

PassSettings pass_settings(int ppq, double alpha, int coherent)
{
    PassSettings p;
    const double base_area = alpha * 2.0;
    p.integrand_angle_area = !coherent ? base_area : (kHALF_PI - base_area);
    p.phi_integrand_factor = p.integrand_angle_area / ((ppq - 1) * 3.0);
    p.phi_start = !coherent ? -alpha : alpha;
    p.phi_step = p.integrand_angle_area / (ppq - 1);
    return p;
}

// GetTau (Simulation.h:90) with the explicit incoherent tau.
int taus_for(Ctx& c, const es_settings& st, int coherent, const double* T, int nT, std::vector<double>& taus)
{
    taus.resize(nT);
    for (int t = 0; t < nT; t++) {
        if (coherent)
            taus[t] = st.tau;
        else if (st.incoherent_tau > 0)
            taus[t] = st.incoherent_tau;
        else {
            if (!(T[t] > 0)) return fail(c, ES_ERR_INVALID_ARG, "temperature must be > 0 (or set incoherent_tau)");
            taus[t] = kHBAR / (kKB * T[t]);
        }
        if (!(taus[t] > 0) || !std::isfinite(taus[t])) return fail(c, ES_ERR_INVALID_ARG, "tau must be finite and > 0");
    }
    return ES_OK;
}

struct Shape {
    int waves_per_eu = 3;
    int chunk = 256;
    int grid_mult = 1;
    int coop = -1;  // FP64 trace: 0 per lane, 1 wave-cooperative, 2 cooperative + screened, -1 by the field
};

// Screened solves pay where many disks cross the orbit in a cell: dense fields.
constexpr double kScreenPopulation = 10.0;
constexpr int kMaxBatchTemperatures = 4096;  // per batch sample; run in chunks of kMaxT


// Device buffers and the optional captured graph of the stage kernels.
struct Stages {
    // tables: sample table | pair table | factors, uploaded in one copy;
    // outs: sigma | lifetime sums | metrics, downloaded in one copy.
    DevBuf tables, outs, dpart, mpart, lt_item, m_item, x_item, f_item, queue, lifetimes;
    DevBuf list, list_count;  // mixed precision: items the FP32 trace hands to the FP64 one
    size_t t_pairs = 0, t_factor = 0;  // byte offsets in tables
    size_t o_ltsum = 0, o_mout = 0;    // byte offsets in outs
    // The temperature stage leaves the work queue and the recompute count at
    // zero for the next run; after a run that did not complete they are cleared.
    bool queue_dirty = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

es::SampleDev* d_samples(Stages& s) { return P<es::SampleDev>(s.tables); }
es::PairDev* d_pairs(Stages& s) { return reinterpret_cast<es::PairDev*>(P<char>(s.tables) + s.t_pairs); }
double* d_factor(Stages& s) { return reinterpret_cast<double*>(P<char>(s.tables) + s.t_factor); }
double* d_sig(Stages& s) { return P<double>(s.outs); }
double* d_ltsum(Stages& s) { return reinterpret_cast<double*>(P<char>(s.outs) + s.o_ltsum); }
int64_t* d_mout(Stages& s) { return reinterpret_cast<int64_t*>(P<char>(s.outs) + s.o_mout); }

int stages_reset_queue(Ctx& c, Stages& s)
{
    if (!s.queue_dirty) return ES_OK;
    if (s.queue.p) ES_HIP(c, hipMemsetAsync(s.queue.p, 0, sizeof(unsigned long long), c.stream));
    if (s.list_count.p) ES_HIP(c, hipMemsetAsync(s.list_count.p, 0, sizeof(int32_t), c.stream));
    s.queue_dirty = false;
    return ES_OK;
}

void stages_release(Ctx& c, Stages& s)
{
    if (s.exec) (void)hipGraphExecDestroy(s.exec);
    if (s.graph) (void)hipGraphDestroy(s.graph);
    s.exec = nullptr;
    s.graph = nullptr;
    DevBuf* bufs[] = {&s.tables,  &s.outs,   &s.dpart,  &s.mpart,
                      &s.lt_item, &s.m_item, &s.x_item, &s.f_item, &s.queue, &s.lifetimes, &s.list,  &s.list_count};
    for (DevBuf* b : bufs) release(c, *b);
}

// A run of whole pairs launched together; item and block offsets of its pairs
// count from the segment's start, so one set of per-item buffers serves all.
struct Segment {
    int p0 = 0, p1 = 0;       // pairs [p0, p1)
    long long items = 0;
    int blocks = 0;           // temperature-stage blocks
};

// Host description of one stage run: samples, pairs (sorted by item and block
// offsets), the integrand factors [out][kMaxT], and the derived sizes.
struct Plan {
    std::vector<es::SampleDev> samples;
    std::vector<es::PairDev> pairs;
    std::vector<double> factor;
    std::vector<Segment> segs;
    int nout = 0;
    int clockwise = 0;
    int precision = ES_PRECISION_FP64;
    int NT = 1;
    long long total_items = 0;   // over all segments
    long long max_items = 0;     // largest segment
    int max_blocks = 0;
    long long P = 0;  // lifetimes per pass, when raw lifetimes are returned
};

// Fills item / block offsets and cuts the pairs into segments of at most
// seg_items work items (a pair larger than that is a segment of its own).
void plan_offsets(Plan& pl, long long seg_items)
{
    long long item = 0;
    int blk = 0;
    int maxT = 1;
    pl.segs.clear();
    pl.total_items = pl.max_items = 0;
    pl.max_blocks = 0;
    Segment sg;
    for (size_t i = 0; i < pl.pairs.size(); i++) {
        es::PairDev& p = pl.pairs[i];
        // Per-orbit quotients the trace would otherwise form per particle: IEEE
        // divisions here give the same bits as the device's.
        p.R = pl.samples[p.sample].speed / p.w;  // lifetime.h:41-44
        p.inv_R = 1.0 / p.R;
        p.inv_w = 1.0 / p.w;
        // About kSigmaItemsPerThread items per thread, so the block's fixed-order
        // double-double reduction is amortised over many items.
        const es::SampleDev& sd = pl.samples[p.sample];
        p.nall = p.nitems / sd.tpq * sd.ppq;
        // temperature-stage units: lifetimes, or (position, quadrant) for a zero-width window
        const bool dup = sd.tpq == 1 && sd.ppq > 1;
        const long long units = dup ? p.nall / sd.ppq : p.nall;
        const long long per_blk = (long long)es::kBlock * (dup ? 1 : es::kSigmaItemsPerThread);
        const int nblk = (int)std::max<long long>(1, std::min<long long>((units + per_blk - 1) / per_blk, 1024));
        if ((int)i > sg.p0 && item + p.nitems > seg_items) {
            sg.p1 = (int)i;
            sg.items = item;
            sg.blocks = blk;
            pl.segs.push_back(sg);
            sg = Segment();
            sg.p0 = (int)i;
            item = 0;
            blk = 0;
        }
        p.item_off = item;
        item += p.nitems;
        p.nblk = nblk;
        p.blk_off = blk;
        blk += nblk;
        pl.total_items += p.nitems;
        maxT = std::max(maxT, pl.samples[p.sample].nT);
    }
    sg.p1 = (int)pl.pairs.size();
    sg.items = item;
    sg.blocks = blk;
    pl.segs.push_back(sg);
    for (const Segment& g : pl.segs) {
        pl.max_items = std::max(pl.max_items, g.items);
        pl.max_blocks = std::max(pl.max_blocks, g.blocks);
    }
    pl.NT = maxT <= 3 ? std::max(maxT, 1) : maxT <= 4 ? 4 : 8;
}

// The mixed path lists the items it traces again as int32 item numbers of
// their segment: a segment (one pair can exceed seg_items) must stay below 2^31.
int plan_check(Ctx& c, const Plan& pl)
{
    if (pl.precision != ES_PRECISION_MIXED) return ES_OK;
    for (const Segment& g : pl.segs)
        if (g.items > (long long)INT32_MAX)
            return fail(c, ES_ERR_INVALID_ARG,
                        "mixed precision: a launch segment holds 2^31 or more work items (one field of this size); "
                        "use ES_PRECISION_FP64");
    return ES_OK;
}

size_t align_up(size_t b) { return (b + 255) & ~(size_t)255; }

bool same_bits(double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; }

int stages_prepare(Ctx& c, Stages& s, const Plan& pl)
{
    const int NQ = 2 * pl.NT + 1;
    const size_t npairs = pl.pairs.size();
    int rc;
    const size_t bs = pl.samples.size() * sizeof(es::SampleDev), bp = npairs * sizeof(es::PairDev);
    const size_t bf = (size_t)pl.nout * es::kMaxT * sizeof(double);
    s.t_pairs = align_up(bs);
    s.t_factor = s.t_pairs + align_up(bp);
    // then the phi tables: (sin, cos) of the 4 ppq start angles, one per
    // distinct (ppq, phi_start, phi_step) of the samples
    std::vector<size_t> t_phi(pl.samples.size());
    std::vector<int> phi_owner(pl.samples.size());  // the first sample with the same start angles
    size_t tables = align_up(s.t_factor + bf);
    for (size_t i = 0; i < pl.samples.size(); i++) {
        const es::SampleDev& sd = pl.samples[i];
        phi_owner[i] = (int)i;
        for (size_t k = 0; k < i; k++) {
            const es::SampleDev& o = pl.samples[k];
            if (phi_owner[k] == (int)k && o.ppq == sd.ppq && same_bits(o.phi_start, sd.phi_start) &&
                same_bits(o.phi_step, sd.phi_step)) {
                phi_owner[i] = (int)k;
                break;
            }
        }
        if (phi_owner[i] != (int)i) {
            t_phi[i] = t_phi[phi_owner[i]];
            continue;
        }
        t_phi[i] = tables;
        tables += align_up((size_t)4 * sd.ppq * sizeof(double2));
    }
    const size_t bsig = (size_t)pl.nout * es::kMaxT * 2 * sizeof(double), blt = (size_t)pl.nout * sizeof(double);
    s.o_ltsum = align_up(bsig);
    s.o_mout = s.o_ltsum + align_up(blt);
    const size_t outs = s.o_mout + (size_t)pl.nout * es::kMetricSlots * sizeof(int64_t);
    if ((rc = ensure(c, s.tables, tables, "sample / pair tables"))) return rc;
    if ((rc = ensure(c, s.outs, outs, "results"))) return rc;
    if ((rc = ensure(c, s.dpart, (size_t)pl.max_blocks * NQ * 2 * sizeof(double), "partials"))) return rc;
    if ((rc = ensure(c, s.mpart, (size_t)pl.max_blocks * es::kMetricSlots * sizeof(int64_t), "metric partials")))
        return rc;
    const size_t items = (size_t)std::max<long long>(pl.max_items, 1);
    if ((rc = ensure(c, s.lt_item, items * sizeof(double), "item lifetimes"))) return rc;
    if ((rc = ensure(c, s.m_item, items * sizeof(int4), "item counters"))) return rc;
    if ((rc = ensure(c, s.x_item, items * sizeof(int32_t), "item exact solves"))) return rc;
    if ((rc = ensure(c, s.f_item, items, "item flags"))) return rc;
    size_t had = s.queue.cap;
    if ((rc = ensure(c, s.queue, sizeof(unsigned long long), "work queue"))) return rc;
    if (s.queue.cap != had) s.queue_dirty = true;
    if (pl.precision == ES_PRECISION_MIXED) {
        if ((rc = ensure(c, s.list, items * sizeof(int32_t), "recompute list"))) return rc;
        had = s.list_count.cap;
        if ((rc = ensure(c, s.list_count, sizeof(int32_t), "recompute count"))) return rc;
        if (s.list_count.cap != had) s.queue_dirty = true;
    }
    if ((rc = ctx_events(c, 4 * pl.segs.size()))) return rc;
    if ((rc = stages_reset_queue(c, s))) return rc;
    // The tables go up from pinned staging in one asynchronous copy; the same
    // staging later takes the results (stages_collect), so it is sized for both.
    if ((rc = ensure_host(c, std::max(tables, outs)))) return rc;
    char* h = static_cast<char*>(c.hbuf);
    std::memcpy(h, pl.samples.data(), bs);
    std::memcpy(h + s.t_pairs, pl.pairs.data(), bp);
    std::memcpy(h + s.t_factor, pl.factor.data(), bf);
    es::SampleDev* hs = reinterpret_cast<es::SampleDev*>(h);
    for (size_t i = 0; i < pl.samples.size(); i++) {
        const es::SampleDev& sd = pl.samples[i];
        hs[i].phi_sc = reinterpret_cast<const double2*>(P<char>(s.tables) + t_phi[i]);
        if (phi_owner[i] != (int)i) continue;
        double2* tab = reinterpret_cast<double2*>(h + t_phi[i]);
        for (int q = 0; q < 4; q++)
            for (int p = 0; p < sd.ppq; p++) {
                const double phi = sd.phi_start + q * kHALF_PI + p * sd.phi_step;  // lifetime.h:36-58
                double sn, cs;
                esm_sincos(phi, &sn, &cs);
                tab[q * sd.ppq + p] = make_double2(sn, cs);
            }
    }
    ES_HIP(c, hipMemcpyAsync(s.tables.p, h, tables, hipMemcpyHostToDevice, c.stream));
    return ES_OK;
}

// The trace variant of a shape for a plan: `coop` as set, or for -1 the
// screened cooperative trace when the work items' fields hold on average at
// least kScreenPopulation indexed impurities per cell, else the cooperative one.
int resolve_coop(const Shape& sh, const Plan& pl)
{
    if (sh.coop >= 0) return sh.coop;
    double items = 0.0, pop = 0.0;
    for (const es::PairDev& p : pl.pairs) {
        items += (double)p.nitems;
        pop += (double)p.nitems * pl.samples[p.sample].cell_population;
    }
    return (items > 0.0 && pop >= kScreenPopulation * items) ? 2 : 1;
}

// Trace grid size of every segment: enough resident blocks to fill the chip
// (times grid_mult), never more blocks than the segment has items for.  With
// mixed precision, nblocks holds the FP32 trace's grid and nblocks64 the FP64
// recompute's (the full resident grid: its item count is known on the device only).
int trace_blocks(Ctx& c, int cw, const Shape& sh, const Plan& pl, std::vector<int>& nblocks,
                 int* nblocks64 = nullptr)
{
    const int w = (sh.waves_per_eu == 2) ? 2 : 3;
    const int co = resolve_coop(sh, pl);
    int& bpc64 = c.trace_bpc[cw][co][w];
    if (bpc64 <= 0) {
        ES_HIP(c, es::trace_blocks_per_cu(cw, w, co, &bpc64));
        if (bpc64 <= 0) bpc64 = 1;
    }
    int bpc = bpc64;
    if (pl.precision == ES_PRECISION_MIXED) {
        int& bpc32 = c.trace32_bpc[cw][co];
        if (bpc32 <= 0) {
            ES_HIP(c, es::trace32_blocks_per_cu(cw, co, &bpc32));
            if (bpc32 <= 0) bpc32 = 1;
        }
        if (nblocks64) *nblocks64 = bpc64 * c.cus;
        bpc = bpc32;
    }
    nblocks.clear();
    for (const Segment& g : pl.segs) {
        const long long want = (g.items + es::kBlock - 1) / es::kBlock;
        nblocks.push_back((int)std::max<long long>(
            1, std::min<long long>(want, (long long)bpc * c.cus * std::max(1, sh.grid_mult))));
    }
    return ES_OK;
}

// Enqueues the stage kernels (no host copies: the sequence can be captured).
// Enqueues the stage kernels of every segment (no host copies or memsets: the
// sequence can be captured in a graph).  Events (4 per segment) only when not capturing.
int stages_enqueue(Ctx& c, Stages& s, const Plan& pl, const Shape& sh, bool trace, int with_metrics,
                   double* lifetimes_dev, const std::vector<int>& nblocks, bool events = true, int nblocks64 = 0)
{
    s.queue_dirty = true;  // until stages_collect has seen the run through
    for (size_t si = 0; si < pl.segs.size(); si++) {
        const Segment& g = pl.segs[si];
        es::PairDev* pairs = d_pairs(s) + g.p0;
        const int npairs = g.p1 - g.p0;
        if (trace) {
            es::TraceArgs a{};
            a.samples = d_samples(s);
            a.pairs = pairs;
            a.npairs = npairs;
            a.total_items = g.items;
            a.queue = P<unsigned long long>(s.queue);
            // A small launch gives every wave work: at most the items per wave of its grid.
            const long long per_wave = (g.items + 4LL * nblocks[si] - 1) / (4LL * nblocks[si]);
            a.chunk = (int)std::max<long long>(16, std::min<long long>(sh.chunk, per_wave));
            a.lt_item = P<double>(s.lt_item);
            a.m_item = P<int4>(s.m_item);
            a.x_item = P<int32_t>(s.x_item);
            a.f_item = P<uint8_t>(s.f_item);
            a.lifetimes = lifetimes_dev;
            a.P = pl.P;
            if (events) ES_HIP(c, hipEventRecord(c.ev[4 * si], c.stream));
            if (pl.precision == ES_PRECISION_MIXED) {
                a.list = P<int32_t>(s.list);
                a.list_count = P<int32_t>(s.list_count);
                ES_HIP(c, es::launch_trace_mixed(a, pl.clockwise, sh.waves_per_eu, resolve_coop(sh, pl), nblocks[si], nblocks64,
                                                    c.stream));
            } else {
                ES_HIP(c, es::launch_trace(a, pl.clockwise, sh.waves_per_eu, resolve_coop(sh, pl), nblocks[si], c.stream));
            }
            if (events) ES_HIP(c, hipEventRecord(c.ev[4 * si + 1], c.stream));
        }
        es::SigmaArgs sa{};
        sa.samples = d_samples(s);
        sa.pairs = pairs;
        sa.npairs = npairs;
        sa.total_blocks = g.blocks;
        sa.clockwise = pl.clockwise;
        sa.lt_item = P<double>(s.lt_item);
        sa.m_item = P<int4>(s.m_item);
        sa.x_item = P<int32_t>(s.x_item);
        sa.f_item = P<uint8_t>(s.f_item);
        sa.with_metrics = with_metrics;
        sa.dpart = P<double>(s.dpart);
        sa.mpart = P<int64_t>(s.mpart);
        sa.queue = P<unsigned long long>(s.queue);
        sa.list_count = (pl.precision == ES_PRECISION_MIXED) ? P<int32_t>(s.list_count) : nullptr;
        if (events) ES_HIP(c, hipEventRecord(c.ev[4 * si + 2], c.stream));
        ES_HIP(c, es::launch_sigma(sa, pl.NT, c.stream));
        ES_HIP(c, es::launch_finalize(sa.samples, sa.pairs, npairs, pl.NT, with_metrics, sa.dpart, sa.mpart,
                                      d_factor(s), d_sig(s), d_ltsum(s), d_mout(s), c.stream));
        if (events) ES_HIP(c, hipEventRecord(c.ev[4 * si + 3], c.stream));
    }
    return ES_OK;
}

struct StageOut {
    std::vector<double> sig;     // [out][kMaxT][2]
    std::vector<double> ltsum;   // [out]
    std::vector<int64_t> m;      // [out][kMetricSlots]
    double t_trace = 0, t_temp = 0;
};

// Enqueues the copy of the results into pinned staging (sized by
// stages_prepare), one copy; stages_collect_finish takes them after a wait.
int stages_collect_enqueue(Ctx& c, Stages& s, const Plan& pl, bool metrics)
{
    const size_t bsig = (size_t)pl.nout * es::kMaxT * 2 * sizeof(double);
    const size_t bm = (size_t)pl.nout * es::kMetricSlots * sizeof(int64_t);
    const size_t bytes = metrics ? s.o_mout + bm : bsig;
    if (c.hbuf_cap < bytes) return fail(c, ES_ERR_INVALID_ARG, "result staging smaller than the plan's results");
    ES_HIP(c, hipMemcpyAsync(c.hbuf, s.outs.p, bytes, hipMemcpyDeviceToHost, c.stream));
    return ES_OK;
}

// After the stream has passed the copy of stages_collect_enqueue: the results
// and the stage times of the events.
int stages_collect_finish(Ctx& c, Stages& s, const Plan& pl, bool trace, bool metrics, StageOut& o, bool events)
{
    o.sig.resize((size_t)pl.nout * es::kMaxT * 2);
    o.ltsum.resize(pl.nout);
    o.m.resize((size_t)pl.nout * es::kMetricSlots);
    const size_t bsig = o.sig.size() * sizeof(double), blt = o.ltsum.size() * sizeof(double);
    const size_t bm = o.m.size() * sizeof(int64_t);
    s.queue_dirty = false;
    const char* h = static_cast<const char*>(c.hbuf);
    std::memcpy(o.sig.data(), h, bsig);
    if (metrics) {
        std::memcpy(o.ltsum.data(), h + s.o_ltsum, blt);
        std::memcpy(o.m.data(), h + s.o_mout, bm);
    }
    float ms = 0;
    o.t_trace = o.t_temp = 0;
    if (!events) return ES_OK;
    for (size_t si = 0; si < pl.segs.size(); si++) {
        if (trace) {
            ES_HIP(c, hipEventElapsedTime(&ms, c.ev[4 * si], c.ev[4 * si + 1]));
            o.t_trace += ms * 1e-3;
        }
        ES_HIP(c, hipEventElapsedTime(&ms, c.ev[4 * si + 2], c.ev[4 * si + 3]));
        o.t_temp += ms * 1e-3;
    }
    return ES_OK;
}

// The results in one copy and one wait.
int stages_collect(Ctx& c, Stages& s, const Plan& pl, bool trace, bool metrics, StageOut& o, bool events = true)
{
    int rc;
    if ((rc = stages_collect_enqueue(c, s, pl, metrics))) return rc;
    ES_HIP(c, hipStreamSynchronize(c.stream));
    return stages_collect_finish(c, s, pl, trace, metrics, o, events);
}

void fill_metrics(es_metrics& o, const int64_t* m, double ltsum, int64_t P, int64_t inside, double t_life,
                  double t_temp)
{
    o.cells_passed = m[0];
    o.impurities_tested = m[1];
    o.particles_escaped = m[2];
    o.particles_at_bound = m[3];
    o.trace_cap_hits = m[4];
    o.max_impurities_tested = m[5];
    o.max_cells_passed = m[6];
    o.particles_inside_impurity = inside;
    o.real_particles = P - inside;
    o.avg_particle_lifetime = P ? ltsum / (double)P : 0.0;
    o.time_elapsed_lifetimes = t_life;
    o.time_elapsed_temperatures = t_temp;
    o.traced_particles = m[7];
    o.edge_hits = m[8];
    o.cross_calls = m[9];
    o.sigma_evals = m[10];
    o.cross_exact = m[11];
    o.fp64_recomputed = m[12];
}

// One sample / pass as a SampleDev (tau and nT are filled by the caller).
es::SampleDev sample_dev(Field& f, int ppr, int ppq, double dist, double offx, double offy, const PassSettings& ps,
                         const es_settings& st, int coherent)
{
    es::SampleDev d{};
    d.ixy = f.g.ixy;
    d.cell_end = f.g.cell_end;
    d.start = f.g.lo;
    d.size = f.g.size;
    d.L = f.g.size / (double)f.g.cpr;  // ImpuritySettings.cell_size (Grid.h:89)
    d.radius = f.g.radius;
    d.cpr = f.g.cpr;
    d.free_pos = P<int32_t>(f.free_pos);
    d.nfree = f.nfree;
    d.ppr = ppr;
    d.ppq = ppq;
    d.dist = dist;
    d.offx = offx;
    d.offy = offy;
    d.phi_start = ps.phi_start;
    d.phi_step = ps.phi_step;
    d.alpha = st.alpha;
    d.speed = st.particle_speed;
    d.coherent = coherent;
    d.phi_integrand_factor = ps.phi_integrand_factor;
    d.tpq = (ps.phi_step == 0.0) ? 1 : ppq;
    d.inv_L = 1.0 / d.L;
    d.inv_size = 1.0 / d.size;
    d.inv_speed = 1.0 / d.speed;
    d.cell_population = (double)f.g.total / ((double)f.g.cpr * (double)f.g.cpr);
    return d;
}

int check_settings(Ctx& c, const es_settings* st)
{
    if (!st) return fail(c, ES_ERR_INVALID_ARG, "null settings");
    if (!(st->particle_speed > 0)) return fail(c, ES_ERR_INVALID_ARG, "particle_speed must be > 0");
    if (!(st->alpha >= 0) || !(st->alpha <= kPI / 4 + 1e-12))
        return fail(c, ES_ERR_INVALID_ARG, "alpha must lie in [0, pi/4]");
    return ES_OK;
}

}  // namespace

// ----------------------------------------------------------------- solver

struct es_solver {
    Ctx c;
    es_solver_desc desc{};
    Field f;
    Stages s;
    DevBuf img, img_aux;  // images mode: [nT][3][ppr^2] and the taus / caps
    bool has_sample = false;
    es_settings st{};
    int coherent = 0;
    PassSettings ps{};
};

struct es_batch {
    Ctx c;
    struct Entry {
        int field;               // index into fields
        int ppr, ppq;
        double dist, offx, offy;
        es_settings st;
        int coherent;
        PassSettings ps;
        std::vector<double> B, T, taus;
    };
    // What drew each field: es_batch_clear keeps the fields, and a sample added
    // later with the same recipe takes one that no current entry uses.
    struct Recipe {
        uint64_t seed;
        int64_t count;
        es_grid_desc grid;
        bool valid;              // false after es_batch_reset: the memory is reused by the next draw
    };
    std::vector<Field*> fields;
    std::vector<Recipe> recipes;     // per field
    Jobs pending;                    // binnings and start tests of added entries, launched together
    std::vector<Entry> entries;
    Stages s;
    Shape shape;
    int precision = ES_PRECISION_FP64;
    bool capture = false;
    bool graph_ready = false;
    std::vector<int> graph_blocks;
    Plan plan;
    bool plan_ready = false;
    // es_batch_launch: the run in flight (its results staged, collected by es_batch_collect)
    bool launched = false;
    hipEvent_t t0 = nullptr, t1 = nullptr;
};

namespace {

// ComputeSigmas for nB fields of the solver's current sample.
int solver_compute(es_solver* sv, const double* B, int nB, const double* T, int nT, es_sigma* out, es_metrics* metrics,
                   double* lifetimes_host)
{
    Ctx& c = sv->c;
    if (!sv->f.has_grid) return fail(c, ES_ERR_NO_GRID, "no grid set");
    if (!sv->has_sample) return fail(c, ES_ERR_NO_SAMPLE, "es_solver_init_sample not called");
    if (!(sv->desc.grid_info.region_size > 0)) return fail(c, ES_ERR_INVALID_ARG, "grid_info.region_size must be > 0");
    if (nB <= 0 || !B) return fail(c, ES_ERR_INVALID_ARG, "need at least one field");
    if (nT < 0 || (nT > 0 && (!T || !out))) return fail(c, ES_ERR_INVALID_ARG, "bad temperature arguments");
    for (int b = 0; b < nB; b++)
        if (!(B[b] > 0) || !std::isfinite(B[b])) return fail(c, ES_ERR_INVALID_ARG, "magnetic field must be > 0");
    std::vector<double> taus;
    int rc;
    if ((rc = taus_for(c, sv->st, sv->coherent, T, nT, taus))) return rc;

    const int ppr = sv->desc.positions_per_row, ppq = sv->desc.particles_per_quadrant, ppp = 4 * ppq;
    const double dist = sv->desc.grid_info.region_size / (double)(ppr - 1);  // Simulation.h:142-143
    const double offx = sv->desc.grid_info.cell_size * 0.01;
    const double offy = sv->desc.grid_info.cell_size * 0.005;
    if ((rc = field_positions(c, sv->f, ppr, dist, offx, offy))) return rc;
    const int64_t P_ = (int64_t)ppr * ppr * ppp;

    Plan pl;
    pl.samples.push_back(sample_dev(sv->f, ppr, ppq, dist, offx, offy, sv->ps, sv->st, sv->coherent));
    pl.clockwise = sv->st.is_clockwise ? 1 : 0;
    pl.precision = sv->desc.precision;
    pl.nout = nB;
    std::vector<double> w(nB);
    for (int b = 0; b < nB; b++) {
        w[b] = kE * B[b] / kM;  // SimulationCPU.cpp:107
        es::PairDev p{};
        p.sample = 0;
        p.out = b;
        p.w = w[b];
        p.nitems = (long long)sv->f.nfree * 4 * pl.samples[0].tpq;
        pl.pairs.push_back(p);
    }
    if (lifetimes_host) {
        pl.P = P_;
        if ((rc = ensure(c, sv->s.lifetimes, (size_t)nB * P_ * sizeof(double), "lifetimes"))) return rc;
        ES_HIP(c, hipMemsetAsync(sv->s.lifetimes.p, 0, (size_t)nB * P_ * sizeof(double), c.stream));
    }

    Shape sh;
    std::vector<int> nblocks;
    int nblocks64 = 0;
    double t_life = 0, t_temp = 0;
    std::vector<double> ltsum;
    std::vector<int64_t> mh;
    const int nchunks = nT == 0 ? 1 : (nT + es::kMaxT - 1) / es::kMaxT;
    for (int ch = 0; ch < nchunks; ch++) {
        const int t0 = ch * es::kMaxT;
        const int nt = nT == 0 ? 0 : std::min(es::kMaxT, nT - t0);
        es::SampleDev& sd = pl.samples[0];
        sd.nT = nt;
        for (int t = 0; t < es::kMaxT; t++) {
            sd.tau[t] = t < nt ? taus[t0 + t] : 1.0;
            sd.max_lt[t] = 15.0 * sd.tau[t];  // Simulation.h:91
            sd.inv_tau[t] = 1.0 / sd.tau[t];
        }
        pl.factor.assign((size_t)nB * es::kMaxT, 0.0);
        for (int b = 0; b < nB; b++)
            for (int t = 0; t < nt; t++)
                pl.factor[(size_t)b * es::kMaxT + t] =
                    integrand_factor(ppr, sv->ps.phi_integrand_factor, sv->st.particle_speed, w[b], taus[t0 + t]);
        plan_offsets(pl, c.seg_items);
        if ((rc = plan_check(c, pl))) return rc;
        if ((rc = stages_prepare(c, sv->s, pl))) return rc;
        // Later temperature chunks reuse the item lifetimes of a one-segment
        // plan; with several segments only the last one's are still resident.
        const bool trace = (ch == 0) || pl.segs.size() > 1;
        const bool first = (ch == 0);
        if (trace && (rc = trace_blocks(c, pl.clockwise, sh, pl, nblocks, &nblocks64))) return rc;
        if ((rc = stages_enqueue(c, sv->s, pl, sh, trace, first ? 1 : 0,
                                 (lifetimes_host && first) ? P<double>(sv->s.lifetimes) : nullptr, nblocks, true,
                                 nblocks64)))
            return rc;
        StageOut so;
        if ((rc = stages_collect(c, sv->s, pl, trace, first, so))) return rc;
        if (first) {
            t_life = so.t_trace;
            ltsum = so.ltsum;
            mh = so.m;
            if (lifetimes_host)
                ES_HIP(c, hipMemcpy(lifetimes_host, sv->s.lifetimes.p, (size_t)nB * P_ * sizeof(double),
                                    hipMemcpyDeviceToHost));
        }
        t_temp += so.t_temp;
        for (int b = 0; b < nB; b++)
            for (int t = 0; t < nt; t++) {
                out[(size_t)b * nT + t0 + t].xx = so.sig[((size_t)b * es::kMaxT + t) * 2];
                out[(size_t)b * nT + t0 + t].xy = so.sig[((size_t)b * es::kMaxT + t) * 2 + 1];
            }
    }
    if (metrics) {
        const int64_t inside = sv->f.n_inside * ppp;
        for (int b = 0; b < nB; b++)
            fill_metrics(metrics[b], &mh[(size_t)b * es::kMetricSlots], ltsum[b], P_, inside, t_life / nB, t_temp / nB);
    }
    return ES_OK;
}

}  // namespace

// ------------------------------------------------------ binning without allocation
//
// es_grid_bin_noalloc bins into caller-provided device buffers; its scratch
// (per-impurity targets and slots, per-cell counts and offsets, the job table)
// is this file-scope array, one copy per GPU for the life of the process.  Host
// code never passes `scratch` itself to a kernel: its address on each device
// comes from hipGetSymbolAddress (g_noalloc_scratch below).
constexpr int64_t kNoAllocMaxImpurities = ES_NOALLOC_MAX_IMPURITIES;
constexpr int64_t kNoAllocMaxCells = ES_NOALLOC_MAX_CELLS;
constexpr size_t kNoAllocAlign = 256;

constexpr size_t noalloc_round(size_t b) { return (b + kNoAllocAlign - 1) & ~(kNoAllocAlign - 1); }

// Layout of the scratch array (byte offsets).
constexpr size_t kNaJob = 0;
constexpr size_t kNaTargets = kNaJob + noalloc_round(sizeof(es::BinJob));
constexpr size_t kNaSlot = kNaTargets + noalloc_round(9 * kNoAllocMaxImpurities * sizeof(int32_t));
constexpr size_t kNaOrder = kNaSlot + noalloc_round(9 * kNoAllocMaxImpurities * sizeof(int32_t));
constexpr size_t kNaNtarg = kNaOrder + noalloc_round(9 * kNoAllocMaxImpurities * sizeof(int32_t));
constexpr size_t kNaCounts = kNaNtarg + noalloc_round(kNoAllocMaxImpurities * sizeof(uint8_t));
constexpr size_t kNaBegin = kNaCounts + noalloc_round(kNoAllocMaxCells * sizeof(int32_t));
constexpr size_t kNaBsum = kNaBegin + noalloc_round((kNoAllocMaxCells + 1) * sizeof(int64_t));
constexpr size_t kNaBytes = kNaBsum + noalloc_round(((kNoAllocMaxCells + 255) / 256) * sizeof(int64_t));

static __device__ __align__(256) unsigned char scratch[kNaBytes];

// The job table is written by a kernel from its by-value argument (no host
// staging, no copy), and the counts are cleared by a kernel (no memset node if
// a caller captures the stream).
__global__ void k_noalloc_put_job(es::BinJob job, es::BinJob* dst) { *dst = job; }

__global__ __launch_bounds__(256) void k_noalloc_clear(int32_t* p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

std::mutex g_noalloc_mu;
void* g_noalloc_scratch[64] = {nullptr};  // device address of `scratch` per device ordinal

// The device address of `scratch` on the current device, looked up on the
// first call there (hipGetSymbolAddress with that device current).
int noalloc_scratch(int device, unsigned char** out, std::string& err)
{
    if (device < 0 || device >= 64) {
        err = "device ordinal out of range";
        return ES_ERR_INVALID_ARG;
    }
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return ES_ERR_HIP;
        }
        g_noalloc_scratch[device] = p;
    }
    *out = static_cast<unsigned char*>(g_noalloc_scratch[device]);
    return ES_OK;
}

// Device elementary functions over an array (es_math_eval).
__global__ __launch_bounds__(256) void k_math_eval(int fn, const double* x, double* y, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = x[i];
    double r, s, c;
    switch (fn) {
    case ES_MATH_SIN: esm_sincos(v, &s, &c); r = s; break;
    case ES_MATH_COS: esm_sincos(v, &s, &c); r = c; break;
    case ES_MATH_ASIN: r = esm_asin(v); break;
    default: r = esm_exp(v); break;
    }
    y[i] = r;
}

extern "C" {

const char* es_version(void) { return "es-mi355x 0.4 (gfx950, FP64 with correctly rounded sin/cos/asin/exp + mixed FP32/FP64, batched, MLP surrogate)"; }

int es_grid_bin_noalloc(const double* x, const double* y, int64_t n, const es_grid_desc* desc, double* ordered_xy,
                        int32_t* imp_index, int64_t* indexed_count, es_grid_info* out_info, void* stream)
{
    std::string& err = g_create_error;
    if (!desc || n < 0 || (n > 0 && (!x || !y)) || !ordered_xy || !imp_index || !indexed_count) {
        err = "es_grid_bin_noalloc: null pointer or negative count";
        return ES_ERR_INVALID_ARG;
    }
    if (n > kNoAllocMaxImpurities) {
        err = "es_grid_bin_noalloc: more than " + std::to_string(kNoAllocMaxImpurities) +
              " impurities (the size of its fixed scratch)";
        return ES_ERR_INVALID_ARG;
    }
    if (desc->target_cell_population <= 0 && desc->cells_per_row <= 0) {
        err = "target_cell_population must be > 0";
        return ES_ERR_INVALID_ARG;
    }
    if (!(desc->impurity_radius >= 0)) {
        err = "impurity_radius must be >= 0";
        return ES_ERR_INVALID_ARG;
    }
    if (!(desc->region_size + 2 * desc->region_extends > 0)) {
        err = "empty spawn region";
        return ES_ERR_INVALID_ARG;
    }
    const int cpr = cells_per_row_rule(n, desc->target_cell_population, desc->cells_per_row);
    if ((int64_t)cpr * cpr > kNoAllocMaxCells) {
        err = "es_grid_bin_noalloc: more than " + std::to_string(kNoAllocMaxCells) +
              " cells (the size of its fixed scratch)";
        return ES_ERR_INVALID_ARG;
    }
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) {
        (void)hipGetLastError();
        err = "no HIP device";
        return ES_ERR_NO_DEVICE;
    }
    unsigned char* base = nullptr;
    int rc = noalloc_scratch(device, &base, err);
    if (rc != ES_OK) return rc;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t ncell = (int64_t)cpr * cpr;
    es::BinJob j{};
    j.x = const_cast<double*>(x);  // read only: generate = 0
    j.y = const_cast<double*>(y);
    j.n = n;
    j.generate = 0;
    j.lo = -desc->region_extends;  // Grid.cpp:52
    j.size = (desc->region_size + desc->region_extends) - j.lo;
    j.cell_size = j.size / (double)cpr;  // Grid.cpp:55
    j.radius = desc->impurity_radius;
    j.cpr = cpr;
    j.s = es::BinScratch{reinterpret_cast<int32_t*>(base + kNaTargets), reinterpret_cast<uint8_t*>(base + kNaNtarg),
                         reinterpret_cast<int32_t*>(base + kNaCounts),  reinterpret_cast<int32_t*>(base + kNaSlot),
                         reinterpret_cast<int64_t*>(base + kNaBegin),   reinterpret_cast<int64_t*>(base + kNaBsum),
                         reinterpret_cast<int32_t*>(base + kNaOrder)};
    j.cell_end = imp_index;
    j.ixy = reinterpret_cast<double2*>(ordered_xy);
    j.total_out = indexed_count;
    es::BinJob* djob = reinterpret_cast<es::BinJob*>(base + kNaJob);
    hipLaunchKernelGGL(k_noalloc_clear, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, st, j.s.counts, ncell);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_noalloc_put_job, dim3(1), dim3(1), 0, st, j, djob);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = es::launch_bin_jobs(djob, &j, 1, st);
    if (e != hipSuccess) {
        err = std::string("es_grid_bin_noalloc: ") + hipGetErrorString(e);
        return ES_ERR_HIP;
    }
    if (out_info) {
        es_grid_info info{};
        info.region_size = desc->region_size;
        info.region_extends = desc->region_extends;
        info.cell_size = j.cell_size;
        info.indexed_impurity_count = -1;  // on the device: *indexed_count once the stream has run
        info.unique_impurity_count = n;
        info.index_size = (int32_t)ncell;
        info.cells_per_row = cpr;
        *out_info = info;
    }
    return ES_OK;
}

int es_math_eval(int32_t device, int32_t fn, const double* x, double* y, int64_t n)
{
    if (fn < ES_MATH_SIN || fn > ES_MATH_EXP || n < 0 || (n > 0 && (!x || !y))) {
        g_create_error = "es_math_eval: bad function id, size or pointer";
        return ES_ERR_INVALID_ARG;
    }
    if (n == 0) return ES_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        g_create_error = "no HIP device";
        return ES_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= ndev || hipSetDevice(device) != hipSuccess) {
        g_create_error = "es_math_eval: bad device";
        return ES_ERR_INVALID_ARG;
    }
    double* d = nullptr;
    const size_t bytes = (size_t)n * sizeof(double);
    if (hipMalloc(&d, 2 * bytes) != hipSuccess) {
        g_create_error = "es_math_eval: device allocation failed";
        return ES_ERR_OOM;
    }
    hipError_t e = hipMemcpy(d, x, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_math_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, fn, d, d + n, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(y, d + n, bytes, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) {
        g_create_error = std::string("es_math_eval: ") + hipGetErrorString(e);
        return ES_ERR_HIP;
    }
    return ES_OK;
}

int es_solver_create(const es_solver_desc* desc, es_solver** out)
{
    if (!desc || !out) {
        g_create_error = "null argument";
        return ES_ERR_INVALID_ARG;
    }
    *out = nullptr;
    if (desc->positions_per_row < 2 || desc->particles_per_quadrant < 2) {
        g_create_error = "positions_per_row and particles_per_quadrant must be >= 2";
        return ES_ERR_INVALID_ARG;
    }
    if ((int64_t)desc->positions_per_row * desc->positions_per_row > ((int64_t)1 << 30) ||
        desc->particles_per_quadrant > (1 << 20)) {
        g_create_error = "positions_per_row / particles_per_quadrant too large";
        return ES_ERR_INVALID_ARG;
    }
    if (desc->grid_info.region_size < 0 || desc->grid_info.cell_size < 0) {
        g_create_error = "grid_info.region_size and cell_size must be >= 0 (0: taken from the first grid)";
        return ES_ERR_INVALID_ARG;
    }
    if (desc->precision != ES_PRECISION_FP64 && desc->precision != ES_PRECISION_MIXED) {
        g_create_error = "unsupported precision mode";
        return ES_ERR_INVALID_ARG;
    }
    es_solver* s = new es_solver();
    s->desc = *desc;
    const int rc = ctx_init(s->c, desc->device, g_create_error);
    if (rc != ES_OK) {
        es_solver_destroy(s);
        return rc;
    }
    *out = s;
    return ES_OK;
}

void es_solver_destroy(es_solver* s)
{
    if (!s) return;
    (void)hipSetDevice(s->c.device);
    if (s->c.stream) (void)hipStreamSynchronize(s->c.stream);
    field_release(s->c, s->f);
    stages_release(s->c, s->s);
    release(s->c, s->img);
    release(s->c, s->img_aux);
    ctx_destroy(s->c);
    delete s;
}

const char* es_last_error(const es_solver* s) { return s ? s->c.err.c_str() : g_create_error.c_str(); }

// After a new field: the solver's GridInformation (taken from sample 0's grid
// when the solver was created without one, SimulationRunner.cpp:75-84), the
// start test of its positions launched behind the binning, and one wait for
// both read-backs.
int solver_field_ready(es_solver* s, const es_grid_desc* desc, es_grid_info* out_info, Jobs& q)
{
    Field& f = s->f;
    es_grid_info info{};
    info.region_size = desc->region_size;
    info.region_extends = desc->region_extends;
    info.cell_size = f.g.cell_size;
    info.unique_impurity_count = f.unique;
    info.index_size = (int32_t)((int64_t)f.g.cpr * f.g.cpr);
    info.cells_per_row = f.g.cpr;
    const bool adopt = !(s->desc.grid_info.cell_size > 0);
    if (adopt) s->desc.grid_info = info;
    int rc;
    const int ppr = s->desc.positions_per_row;
    if (s->desc.grid_info.region_size > 0) {
        const double dist = s->desc.grid_info.region_size / (double)(ppr - 1);  // Simulation.h:142-143
        if ((rc = field_positions_job(s->c, f, ppr, dist, s->desc.grid_info.cell_size * 0.01,
                                      s->desc.grid_info.cell_size * 0.005, q)))
            return rc;
    }
    if ((rc = jobs_launch(s->c, q))) return rc;
    if ((rc = field_sync(s->c, f, false))) return rc;
    info.indexed_impurity_count = f.g.total;
    if (adopt) s->desc.grid_info.indexed_impurity_count = f.g.total;
    if (out_info) *out_info = info;
    return ES_OK;
}

int es_grid_set_impurities(es_solver* s, const double* x, const double* y, int64_t n, const es_grid_desc* desc,
                           es_grid_info* out_info)
{
    if (!s) return ES_ERR_INVALID_ARG;
    ES_HIP(s->c, hipSetDevice(s->c.device));
    Jobs q;
    const int rc = field_upload(s->c, s->f, x, y, n, desc, q);
    if (rc) return rc;
    return solver_field_ready(s, desc, out_info, q);
}

int es_grid_generate(es_solver* s, uint64_t seed, int64_t count, const es_grid_desc* desc, es_grid_info* out_info)
{
    if (!s) return ES_ERR_INVALID_ARG;
    ES_HIP(s->c, hipSetDevice(s->c.device));
    Jobs q;
    const int rc = field_generate(s->c, s->f, seed, count, desc, q);
    if (rc) return rc;
    return solver_field_ready(s, desc, out_info, q);
}

int es_grid_set_binned(es_solver* s, const double* ordered_xy, const int32_t* imp_index, int32_t cells_per_row,
                       double spawn_region_start, double spawn_region_size, double impurity_radius,
                       es_grid_info* out_info)
{
    if (!s) return ES_ERR_INVALID_ARG;
    Ctx& c = s->c;
    if (!(s->desc.grid_info.region_size > 0) || !(s->desc.grid_info.cell_size > 0))
        return fail(c, ES_ERR_INVALID_ARG,
                    "es_grid_set_binned needs the solver created with its GridInformation (region_size, cell_size; "
                    "Simulation.h:131-146): a binned field does not carry the start-position grid");
    ES_HIP(c, hipSetDevice(c.device));
    int rc = field_set_binned(c, s->f, ordered_xy, imp_index, cells_per_row, spawn_region_start, spawn_region_size,
                              impurity_radius);
    if (rc) return rc;
    Jobs q;
    const es_grid_info& gi = s->desc.grid_info;
    const int ppr = s->desc.positions_per_row;
    const double dist = gi.region_size / (double)(ppr - 1);  // Simulation.h:142-143
    if ((rc = field_positions_job(c, s->f, ppr, dist, gi.cell_size * 0.01, gi.cell_size * 0.005, q))) return rc;
    if ((rc = jobs_launch(c, q))) return rc;
    if ((rc = field_sync(c, s->f, false))) return rc;
    if (out_info) {
        es_grid_info info{};
        info.region_size = gi.region_size;
        info.region_extends = gi.region_extends;
        info.cell_size = s->f.g.cell_size;
        info.indexed_impurity_count = s->f.g.total;
        info.unique_impurity_count = 0;
        info.index_size = (int32_t)((int64_t)cells_per_row * cells_per_row);
        info.cells_per_row = cells_per_row;
        *out_info = info;
    }
    return ES_OK;
}

int es_grid_read(es_solver* s, double* ordered_xy, int32_t* imp_index)
{
    if (!s) return ES_ERR_INVALID_ARG;
    Ctx& c = s->c;
    if (!s->f.has_grid) return fail(c, ES_ERR_NO_GRID, "no grid set");
    ES_HIP(c, hipSetDevice(c.device));
    const int64_t total = s->f.g.total, ncell = (int64_t)s->f.g.cpr * s->f.g.cpr;
    if (ordered_xy && total > 0)
        ES_HIP(c, hipMemcpyAsync(ordered_xy, s->f.g.ixy, total * sizeof(double2), hipMemcpyDeviceToHost, c.stream));
    if (imp_index)
        ES_HIP(c, hipMemcpyAsync(imp_index, s->f.g.cell_end, ncell * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
    ES_HIP(c, hipStreamSynchronize(c.stream));
    return ES_OK;
}

int es_solver_init_sample(es_solver* s, const es_settings* settings, int32_t coherent)
{
    if (!s) return ES_ERR_INVALID_ARG;
    const int rc = check_settings(s->c, settings);
    if (rc) return rc;
    s->st = *settings;
    s->coherent = coherent ? 1 : 0;
    s->ps = pass_settings(s->desc.particles_per_quadrant, settings->alpha, s->coherent);  // Simulation.h:111-126
    s->has_sample = true;
    return ES_OK;
}

int es_solver_compute_sigmas(es_solver* s, double B, const double* T, int32_t nT, es_sigma* out, es_metrics* metrics,
                             double* lifetimes)
{
    if (!s) return ES_ERR_INVALID_ARG;
    ES_HIP(s->c, hipSetDevice(s->c.device));
    return solver_compute(s, &B, 1, T, nT, out, metrics, lifetimes);
}

int es_solver_compute_sigmas_sweep(es_solver* s, const double* B, int32_t nB, const double* T, int32_t nT,
                                   es_sigma* out, es_metrics* metrics)
{
    if (!s) return ES_ERR_INVALID_ARG;
    ES_HIP(s->c, hipSetDevice(s->c.device));
    return solver_compute(s, B, nB, T, nT, out, metrics, nullptr);
}

int es_solver_compute_sigmas_with_images(es_solver* s, double B, const double* T, int32_t nT, es_sigma* out,
                                         es_metrics* metrics, double* images)
{
    if (!s) return ES_ERR_INVALID_ARG;
    Ctx& c = s->c;
    ES_HIP(c, hipSetDevice(c.device));
    if (nT <= 0 || !T || !out || !images) return fail(c, ES_ERR_INVALID_ARG, "need temperatures, out and images");
    int rc = solver_compute(s, &B, 1, T, nT, out, metrics, nullptr);
    if (rc) return rc;
    // The stage buffers still hold this field's single-pair plan: its item
    // lifetimes (one segment) and its sample / pair tables.
    std::vector<double> aux;
    if ((rc = taus_for(c, s->st, s->coherent, T, nT, aux))) return rc;
    for (int t = 0; t < nT; t++) aux.push_back(15.0 * aux[t]);  // Simulation.h:91
    const int ppr = s->desc.positions_per_row;
    const size_t plane = (size_t)ppr * ppr, bytes = (size_t)nT * 3 * plane * sizeof(double);
    if ((rc = ensure(c, s->img, bytes, "images"))) return rc;
    if ((rc = ensure(c, s->img_aux, aux.size() * sizeof(double), "image taus"))) return rc;
    ES_HIP(c, hipMemcpyAsync(s->img_aux.p, aux.data(), aux.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
    ES_HIP(c, hipMemsetAsync(s->img.p, 0, bytes, c.stream));
    ES_HIP(c, es::launch_images(d_samples(s->s), d_pairs(s->s),
                                s->st.is_clockwise ? 1 : 0, P<double>(s->s.lt_item), P<double>(s->img_aux),
                                P<double>(s->img_aux) + nT, nT, s->f.nfree, ppr, P<double>(s->img), c.stream));
    ES_HIP(c, hipMemcpyAsync(images, s->img.p, bytes, hipMemcpyDeviceToHost, c.stream));
    ES_HIP(c, hipStreamSynchronize(c.stream));
    return ES_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ batch

namespace {

bool same_grid_desc(const es_grid_desc& a, const es_grid_desc& b)
{
    return same_bits(a.region_size, b.region_size) && same_bits(a.region_extends, b.region_extends) &&
           same_bits(a.impurity_radius, b.impurity_radius) && same_bits(a.impurity_density, b.impurity_density) &&
           a.target_cell_population == b.target_cell_population && a.cells_per_row == b.cells_per_row;
}

// Temperatures run in chunks of kMaxT: chunk ch of every sample (its
// temperatures ch*kMaxT ... ch*kMaxT + kMaxT - 1, none past its last) goes
// into the plan's sample table and integrand factors.  The number of chunks
// is that of the sample with the most temperatures.
int batch_chunks(const es_batch* bt)
{
    size_t nT = 0;
    for (const es_batch::Entry& en : bt->entries) nT = std::max(nT, en.T.size());
    return nT == 0 ? 1 : (int)((nT + es::kMaxT - 1) / es::kMaxT);
}

void batch_set_chunk(es_batch* bt, int ch)
{
    Plan& pl = bt->plan;
    const size_t t0 = (size_t)ch * es::kMaxT;
    for (size_t e = 0; e < bt->entries.size() && e < pl.samples.size(); e++) {
        const es_batch::Entry& en = bt->entries[e];
        es::SampleDev& sd = pl.samples[e];
        sd.nT = en.T.size() > t0 ? (int)std::min<size_t>(es::kMaxT, en.T.size() - t0) : 0;
        for (int t = 0; t < es::kMaxT; t++) {
            sd.tau[t] = t < sd.nT ? en.taus[t0 + t] : 1.0;
            sd.max_lt[t] = 15.0 * sd.tau[t];
            sd.inv_tau[t] = 1.0 / sd.tau[t];
        }
    }
    pl.factor.assign((size_t)pl.nout * es::kMaxT, 0.0);
    for (const es::PairDev& p : pl.pairs) {
        const es_batch::Entry& en = bt->entries[p.sample];
        const int nt = pl.samples[p.sample].nT;
        for (int t = 0; t < nt; t++)
            pl.factor[(size_t)p.out * es::kMaxT + t] =
                integrand_factor(en.ppr, en.ps.phi_integrand_factor, en.st.particle_speed, p.w, en.taus[t0 + t]);
    }
}

int batch_plan(es_batch* bt)
{
    Ctx& c = bt->c;
    Plan& pl = bt->plan;
    pl = Plan();
    if (bt->entries.empty()) return fail(c, ES_ERR_INVALID_ARG, "empty batch");
    int jrc;
    if ((jrc = jobs_launch(c, bt->pending))) return jrc;
    // The read-backs of the entries' fields still in flight: one wait for all.
    std::vector<int> pending;
    for (const es_batch::Entry& en : bt->entries) {
        Field& f = *bt->fields[en.field];
        if (!(f.total_pending || f.pos_pending) || std::find(pending.begin(), pending.end(), en.field) != pending.end())
            continue;
        int rc;
        if ((rc = field_sync_enqueue(c, f))) return rc;
        pending.push_back(en.field);
    }
    if (!pending.empty()) {
        ES_HIP(c, hipStreamSynchronize(c.stream));
        int first_rc = ES_OK;
        for (int fi : pending) {
            const int rc = field_sync_finish(c, *bt->fields[fi]);
            if (rc && !first_rc) first_rc = rc;
        }
        if (first_rc) return first_rc;
    }
    pl.clockwise = bt->entries[0].st.is_clockwise ? 1 : 0;
    pl.precision = bt->precision;
    int out = 0;
    for (size_t e = 0; e < bt->entries.size(); e++) {
        const es_batch::Entry& en = bt->entries[e];
        if ((en.st.is_clockwise ? 1 : 0) != pl.clockwise)
            return fail(c, ES_ERR_INVALID_ARG, "a batch holds one rotation sense (is_clockwise)");
        Field& f = *bt->fields[en.field];
        if (!f.pos_valid)
            return fail(c, ES_ERR_OUT_OF_GRID,
                        "a start position lies outside the impurity cell grid (region_extends smaller than the "
                        "position offset); the reference would read out of bounds (lifetime.h:88-91)");
        es::SampleDev sd = sample_dev(f, en.ppr, en.ppq, en.dist, en.offx, en.offy, en.ps, en.st, en.coherent);
        pl.samples.push_back(sd);
        for (size_t b = 0; b < en.B.size(); b++) {
            es::PairDev p{};
            p.sample = (int)e;
            p.out = out++;
            p.w = kE * en.B[b] / kM;
            p.nitems = (long long)f.nfree * 4 * sd.tpq;
            pl.pairs.push_back(p);
        }
    }
    pl.nout = out;
    batch_set_chunk(bt, 0);
    plan_offsets(pl, c.seg_items);
    const int rc = plan_check(c, pl);
    if (rc) return rc;
    bt->plan_ready = true;
    return ES_OK;
}

void batch_release_graph(es_batch* bt)
{
    if (bt->s.exec) (void)hipGraphExecDestroy(bt->s.exec);
    if (bt->s.graph) (void)hipGraphDestroy(bt->s.graph);
    bt->s.exec = nullptr;
    bt->s.graph = nullptr;
    bt->graph_ready = false;
}

// Enqueues one run with a given shape: the stage kernels (or the captured
// graph of them) between two events, and the copy of the results.  No wait
// except the one of batch_plan for the read-backs of new fields.
int batch_launch(es_batch* bt, const Shape& sh)
{
    Ctx& c = bt->c;
    int rc;
    if (bt->launched) return fail(c, ES_ERR_INVALID_ARG, "a launched run has not been collected");
    if (!bt->plan_ready) {
        batch_release_graph(bt);
        if ((rc = batch_plan(bt))) return rc;
        if ((rc = stages_prepare(c, bt->s, bt->plan))) return rc;
    }
    if ((rc = stages_reset_queue(c, bt->s))) return rc;
    std::vector<int> nblocks;
    int nblocks64 = 0;
    if ((rc = trace_blocks(c, bt->plan.clockwise, sh, bt->plan, nblocks, &nblocks64))) return rc;
    if (!bt->t0) ES_HIP(c, hipEventCreate(&bt->t0));
    if (!bt->t1) ES_HIP(c, hipEventCreate(&bt->t1));
    ES_HIP(c, hipEventRecord(bt->t0, c.stream));
    if (bt->capture) {
        if (bt->graph_ready && bt->graph_blocks != nblocks) batch_release_graph(bt);
        if (!bt->graph_ready) {
            hipError_t e = hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) return hip_fail(c, e, "hipStreamBeginCapture");
            rc = stages_enqueue(c, bt->s, bt->plan, sh, true, 1, nullptr, nblocks, false, nblocks64);
            hipGraph_t g = nullptr;
            e = hipStreamEndCapture(c.stream, &g);
            if (!rc && e != hipSuccess) rc = hip_fail(c, e, "hipStreamEndCapture");
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            bt->s.graph = g;
            e = hipGraphInstantiate(&bt->s.exec, g, nullptr, nullptr, 0);
            if (e != hipSuccess) return hip_fail(c, e, "hipGraphInstantiate");
            bt->graph_ready = true;
            bt->graph_blocks = nblocks;
        }
        ES_HIP(c, hipGraphLaunch(bt->s.exec, c.stream));
    } else if ((rc = stages_enqueue(c, bt->s, bt->plan, sh, true, 1, nullptr, nblocks, true, nblocks64))) {
        return rc;
    }
    ES_HIP(c, hipEventRecord(bt->t1, c.stream));
    if ((rc = stages_collect_enqueue(c, bt->s, bt->plan, true))) return rc;
    bt->launched = true;
    return ES_OK;
}

// Waits for the launched run; fills `so` (sigma, sums, metrics, event times)
// and the device time between the events around the stage sequence.
int batch_collect(es_batch* bt, StageOut& so, double* wall)
{
    Ctx& c = bt->c;
    if (!bt->launched) return fail(c, ES_ERR_INVALID_ARG, "no launched run to collect");
    bt->launched = false;
    ES_HIP(c, hipStreamSynchronize(c.stream));
    int rc;
    if ((rc = stages_collect_finish(c, bt->s, bt->plan, true, true, so, !bt->capture))) return rc;
    float ms = 0;
    ES_HIP(c, hipEventElapsedTime(&ms, bt->t0, bt->t1));
    if (wall) *wall = ms * 1e-3;
    return ES_OK;
}

// One run with a given shape (launch, then collect).
int batch_run_once(es_batch* bt, const Shape& sh, StageOut& so, double* wall)
{
    int rc;
    if ((rc = batch_launch(bt, sh))) return rc;
    return batch_collect(bt, so, wall);
}

// Outputs of a collected run in the es_batch_run layout.
void batch_outputs(es_batch* bt, const StageOut& so, es_sigma* out, es_metrics* metrics)
{
    const Plan& pl = bt->plan;
    size_t o = 0;
    for (const es::PairDev& p : pl.pairs) {
        const es_batch::Entry& en = bt->entries[p.sample];
        const int nT = (int)en.T.size();
        for (int t = 0; t < std::min(nT, es::kMaxT) && out; t++) {  // the first chunk; es_batch_run adds the others
            out[o + t].xx = so.sig[((size_t)p.out * es::kMaxT + t) * 2];
            out[o + t].xy = so.sig[((size_t)p.out * es::kMaxT + t) * 2 + 1];
        }
        o += nT;
        if (metrics) {
            const Field& f = *bt->fields[en.field];
            const int64_t P_ = (int64_t)en.ppr * en.ppr * 4 * en.ppq;
            fill_metrics(metrics[p.out], &so.m[(size_t)p.out * es::kMetricSlots], so.ltsum[p.out], P_,
                         f.n_inside * 4 * en.ppq, so.t_trace / pl.nout, so.t_temp / pl.nout);
        }
    }
}

// Mutating calls are refused while a launched run has not been collected:
// its results wait in the staging that a new plan would overwrite.
int batch_idle(es_batch* bt)
{
    if (bt->launched) return fail(bt->c, ES_ERR_INVALID_ARG, "a launched run has not been collected (es_batch_collect)");
    return ES_OK;
}

}  // namespace

extern "C" {

int es_batch_create(int32_t device, es_batch** out)
{
    if (!out) {
        g_create_error = "null argument";
        return ES_ERR_INVALID_ARG;
    }
    *out = nullptr;
    es_batch* b = new es_batch();
    const int rc = ctx_init(b->c, device, g_create_error);
    if (rc != ES_OK) {
        es_batch_destroy(b);
        return rc;
    }
    *out = b;
    return ES_OK;
}

void es_batch_destroy(es_batch* b)
{
    if (!b) return;
    (void)hipSetDevice(b->c.device);
    if (b->c.stream) (void)hipStreamSynchronize(b->c.stream);
    batch_release_graph(b);
    if (b->t0) (void)hipEventDestroy(b->t0);
    if (b->t1) (void)hipEventDestroy(b->t1);
    for (Field* f : b->fields) {
        field_release(b->c, *f);
        delete f;
    }
    stages_release(b->c, b->s);
    ctx_destroy(b->c);
    delete b;
}

const char* es_batch_last_error(const es_batch* b) { return b ? b->c.err.c_str() : g_create_error.c_str(); }

int es_batch_add_sample(es_batch* bt, const es_batch_sample* smp, int32_t* index)
{
    if (!bt) return ES_ERR_INVALID_ARG;
    if (batch_idle(bt)) return ES_ERR_INVALID_ARG;
    Ctx& c = bt->c;
    ES_HIP(c, hipSetDevice(c.device));
    if (!smp) return fail(c, ES_ERR_INVALID_ARG, "null sample");
    if (smp->positions_per_row < 2 || smp->particles_per_quadrant < 2)
        return fail(c, ES_ERR_INVALID_ARG, "positions_per_row and particles_per_quadrant must be >= 2");
    if (smp->n_fields <= 0 || !smp->fields) return fail(c, ES_ERR_INVALID_ARG, "need at least one field");
    if (smp->n_temperatures < 0 || smp->n_temperatures > kMaxBatchTemperatures ||
        (smp->n_temperatures > 0 && !smp->temperatures))
        return fail(c, ES_ERR_INVALID_ARG, "a batch sample takes 0 to 4096 temperatures");
    for (int b = 0; b < smp->n_fields; b++)
        if (!(smp->fields[b] > 0) || !std::isfinite(smp->fields[b]))
            return fail(c, ES_ERR_INVALID_ARG, "magnetic field must be > 0");
    int rc = check_settings(c, &smp->settings);
    if (rc) return rc;
    if (!(smp->grid.region_size > 0)) return fail(c, ES_ERR_INVALID_ARG, "region_size must be > 0");
    es_batch::Entry en;
    en.ppr = smp->positions_per_row;
    en.ppq = smp->particles_per_quadrant;
    if (smp->share_field_of >= 0) {
        if (smp->share_field_of >= (int)bt->entries.size())
            return fail(c, ES_ERR_INVALID_ARG, "share_field_of names no earlier sample");
        en.field = bt->entries[smp->share_field_of].field;
    } else {
        en.field = -1;
        int spare = -1;  // a forgotten field no entry uses: its memory takes the new draw
        for (size_t i = 0; i < bt->fields.size() && en.field < 0; i++) {
            bool used = false;
            for (const es_batch::Entry& other : bt->entries) used = used || other.field == (int)i;
            if (used) continue;
            const es_batch::Recipe& r = bt->recipes[i];
            if (!r.valid) {
                if (spare < 0) spare = (int)i;
                continue;
            }
            if (r.seed == smp->seed && r.count == smp->impurity_count && same_grid_desc(r.grid, smp->grid))
                en.field = (int)i;
        }
        if (en.field < 0) {
            // Drawn, binned and start-tested without a wait: es_batch_run (or
            // es_batch_counts) collects the read-backs of every new field at once.
            Field* f = spare >= 0 ? bt->fields[spare] : new Field();
            if (spare >= 0) bt->recipes[spare].valid = false;
            rc = field_generate(c, *f, smp->seed, smp->impurity_count, &smp->grid, bt->pending);
            if (rc) {
                if (spare < 0) {
                    field_release(c, *f);
                    delete f;
                }
                return rc;
            }
            if (spare < 0) {
                bt->fields.push_back(f);
                bt->recipes.push_back({smp->seed, smp->impurity_count, smp->grid, true});
                en.field = (int)bt->fields.size() - 1;
            } else {
                bt->recipes[spare] = {smp->seed, smp->impurity_count, smp->grid, true};
                en.field = spare;
            }
        }
    }
    Field& f = *bt->fields[en.field];
    const double cell = smp->grid_cell_size > 0 ? smp->grid_cell_size : f.g.cell_size;
    en.dist = smp->grid.region_size / (double)(en.ppr - 1);  // Simulation.h:142-143
    en.offx = cell * 0.01;
    en.offy = cell * 0.005;
    for (const es_batch::Entry& other : bt->entries)
        if (other.field == en.field && (other.ppr != en.ppr || other.dist != en.dist || other.offx != en.offx))
            return fail(c, ES_ERR_INVALID_ARG, "samples sharing a field must share its positions");
    en.st = smp->settings;
    en.coherent = smp->coherent ? 1 : 0;
    en.ps = pass_settings(en.ppq, en.st.alpha, en.coherent);
    en.B.assign(smp->fields, smp->fields + smp->n_fields);
    en.T.assign(smp->temperatures, smp->temperatures + smp->n_temperatures);
    if ((rc = taus_for(c, en.st, en.coherent, en.T.data(), (int)en.T.size(), en.taus))) return rc;
    if ((rc = field_positions_job(c, f, en.ppr, en.dist, en.offx, en.offy, bt->pending))) return rc;
    bt->entries.push_back(en);
    bt->plan_ready = false;
    if (index) *index = (int32_t)bt->entries.size() - 1;
    return ES_OK;
}

int es_batch_clear(es_batch* bt)
{
    if (!bt) return ES_ERR_INVALID_ARG;
    if (batch_idle(bt)) return ES_ERR_INVALID_ARG;
    ES_HIP(bt->c, hipSetDevice(bt->c.device));
    const int rc = jobs_launch(bt->c, bt->pending);  // kept fields stay consistent
    if (rc) return rc;
    bt->entries.clear();
    bt->plan_ready = false;
    batch_release_graph(bt);
    return ES_OK;
}

int es_batch_reset(es_batch* bt)
{
    if (!bt) return ES_ERR_INVALID_ARG;
    if (batch_idle(bt)) return ES_ERR_INVALID_ARG;
    ES_HIP(bt->c, hipSetDevice(bt->c.device));
    const int rc = jobs_launch(bt->c, bt->pending);
    if (rc) return rc;
    bt->entries.clear();
    bt->plan_ready = false;
    batch_release_graph(bt);
    for (es_batch::Recipe& r : bt->recipes) r.valid = false;
    return ES_OK;
}

int es_batch_set_shape(es_batch* bt, const es_batch_shape* shape)
{
    if (!bt) return ES_ERR_INVALID_ARG;
    if (batch_idle(bt)) return ES_ERR_INVALID_ARG;
    if (!shape || (shape->waves_per_eu != 2 && shape->waves_per_eu != 3) || shape->chunk < 1 || shape->chunk > 4096 ||
        shape->grid_mult < 1 || shape->grid_mult > 8 || shape->cooperative < -1 || shape->cooperative > 2)
        return fail(bt->c, ES_ERR_INVALID_ARG,
                    "shape: waves_per_eu 2 or 3, chunk 1..4096, grid_mult 1..8, cooperative -1..2");
    bt->shape.waves_per_eu = shape->waves_per_eu;
    bt->shape.chunk = shape->chunk;
    bt->shape.grid_mult = shape->grid_mult;
    bt->shape.coop = shape->cooperative;
    batch_release_graph(bt);
    return ES_OK;
}

int es_batch_get_shape(es_batch* bt, es_batch_shape* shape)
{
    if (!bt || !shape) return ES_ERR_INVALID_ARG;
    shape->waves_per_eu = bt->shape.waves_per_eu;
    shape->chunk = bt->shape.chunk;
    shape->grid_mult = bt->shape.grid_mult;
    shape->cooperative = bt->shape.coop;
    return ES_OK;
}

int es_batch_set_capture(es_batch* bt, int32_t enable)
{
    if (!bt) return ES_ERR_INVALID_ARG;
    if (batch_idle(bt)) return ES_ERR_INVALID_ARG;
    bt->capture = enable != 0;
    batch_release_graph(bt);
    return ES_OK;
}

int es_batch_set_precision(es_batch* bt, int32_t precision)
{
    if (!bt) return ES_ERR_INVALID_ARG;
    if (batch_idle(bt)) return ES_ERR_INVALID_ARG;
    if (precision != ES_PRECISION_FP64 && precision != ES_PRECISION_MIXED)
        return fail(bt->c, ES_ERR_INVALID_ARG, "unsupported precision mode");
    if (precision != bt->precision) {
        bt->precision = precision;
        bt->plan_ready = false;
        batch_release_graph(bt);
    }
    return ES_OK;
}

int es_batch_run(es_batch* bt, es_sigma* out, es_metrics* metrics, double* seconds)
{
    if (!bt) return ES_ERR_INVALID_ARG;
    Ctx& c = bt->c;
    ES_HIP(c, hipSetDevice(c.device));
    StageOut so;
    double wall = 0;
    int rc = batch_run_once(bt, bt->shape, so, &wall);
    if (rc) return rc;
    const int nch = batch_chunks(bt);
    if (nch > 1) {
        // Later temperature chunks: the temperature stage again on the item
        // lifetimes of the run (the trace again only where the plan has several
        // segments, whose item buffers are reused), then the plan's first chunk back.
        Plan& pl = bt->plan;
        const bool trace = pl.segs.size() > 1;
        std::vector<int> nblocks;
        int nblocks64 = 0;
        if (trace && (rc = trace_blocks(c, pl.clockwise, bt->shape, pl, nblocks, &nblocks64))) return rc;
        for (int ch = 1; ch < nch && !rc; ch++) {
            batch_set_chunk(bt, ch);
            StageOut sc;
            if (!(rc = stages_prepare(c, bt->s, pl)) &&
                !(rc = stages_enqueue(c, bt->s, pl, bt->shape, trace, 0, nullptr, nblocks, true, nblocks64)) &&
                !(rc = stages_collect(c, bt->s, pl, trace, false, sc))) {
                so.t_temp += sc.t_temp;
                wall += sc.t_trace + sc.t_temp;
                size_t o = 0;
                for (const es::PairDev& p : pl.pairs) {
                    const int nT = (int)bt->entries[p.sample].T.size(), nt = pl.samples[p.sample].nT;
                    for (int t = 0; t < nt && out; t++) {
                        out[o + (size_t)ch * es::kMaxT + t].xx = sc.sig[((size_t)p.out * es::kMaxT + t) * 2];
                        out[o + (size_t)ch * es::kMaxT + t].xy = sc.sig[((size_t)p.out * es::kMaxT + t) * 2 + 1];
                    }
                    o += nT;
                }
            }
        }
        batch_set_chunk(bt, 0);
        const int rc0 = stages_prepare(c, bt->s, pl);
        if (rc) return rc;
        if (rc0) return rc0;
    }
    if (seconds) *seconds = wall;
    batch_outputs(bt, so, out, metrics);
    return ES_OK;
}

int es_batch_prepare(es_batch* bt)
{
    if (!bt) return ES_ERR_INVALID_ARG;
    if (batch_idle(bt)) return ES_ERR_INVALID_ARG;
    ES_HIP(bt->c, hipSetDevice(bt->c.device));
    return jobs_launch(bt->c, bt->pending);
}

int es_batch_launch(es_batch* bt)
{
    if (!bt) return ES_ERR_INVALID_ARG;
    ES_HIP(bt->c, hipSetDevice(bt->c.device));
    if (batch_chunks(bt) > 1)
        return fail(bt->c, ES_ERR_INVALID_ARG,
                    "es_batch_launch / es_batch_collect take up to 8 temperatures per sample; es_batch_run takes any "
                    "number");
    return batch_launch(bt, bt->shape);
}

int es_batch_collect(es_batch* bt, es_sigma* out, es_metrics* metrics, double* seconds)
{
    if (!bt) return ES_ERR_INVALID_ARG;
    ES_HIP(bt->c, hipSetDevice(bt->c.device));
    StageOut so;
    double wall = 0;
    const int rc = batch_collect(bt, so, &wall);
    if (rc) return rc;
    if (seconds) *seconds = wall;
    batch_outputs(bt, so, out, metrics);
    return ES_OK;
}

int es_batch_counts(es_batch* bt, int32_t* n_samples, int32_t* n_pairs, int32_t* n_outputs, int64_t* traced_items)
{
    if (!bt) return ES_ERR_INVALID_ARG;
    if (batch_idle(bt)) return ES_ERR_INVALID_ARG;
    if (!bt->plan_ready) {
        batch_release_graph(bt);
        const int rc = batch_plan(bt);
        if (rc) return rc;
        const int rc2 = stages_prepare(bt->c, bt->s, bt->plan);
        if (rc2) return rc2;
    }
    if (n_samples) *n_samples = (int32_t)bt->entries.size();
    if (n_pairs) *n_pairs = (int32_t)bt->plan.pairs.size();
    int32_t nout = 0;
    for (const es_batch::Entry& en : bt->entries) nout += (int32_t)(en.B.size() * en.T.size());
    if (n_outputs) *n_outputs = nout;
    if (traced_items) *traced_items = bt->plan.total_items;
    return ES_OK;
}

int es_batch_launches(es_batch* bt, int32_t* segments, int32_t* kernels_per_run, int32_t* graph_launches_per_run)
{
    if (!bt) return ES_ERR_INVALID_ARG;
    if (batch_idle(bt)) return ES_ERR_INVALID_ARG;
    if (!bt->plan_ready) {
        batch_release_graph(bt);
        const int rc = batch_plan(bt);
        if (rc) return rc;
        const int rc2 = stages_prepare(bt->c, bt->s, bt->plan);
        if (rc2) return rc2;
    }
    // per segment: the trace (FP32 trace, queue reset and FP64 recompute when
    // mixed), the temperature stage and the finalize stage (stages_enqueue)
    const bool mixed = bt->plan.precision == ES_PRECISION_MIXED;
    int32_t kernels = 0;
    for (const Segment& g : bt->plan.segs) kernels += (g.items > 0 ? (mixed ? 3 : 1) : 0) + 2;
    if (segments) *segments = (int32_t)bt->plan.segs.size();
    if (kernels_per_run) *kernels_per_run = kernels;
    if (graph_launches_per_run) *graph_launches_per_run = bt->capture ? 1 : 0;
    return ES_OK;
}

int es_batch_tune(es_batch* bt, es_batch_shape* best, double* best_seconds)
{
    if (!bt) return ES_ERR_INVALID_ARG;
    if (batch_idle(bt)) return ES_ERR_INVALID_ARG;
    Ctx& c = bt->c;
    ES_HIP(c, hipSetDevice(c.device));
    const bool cap = bt->capture;
    bt->capture = false;  // the candidates are timed on direct launches
    Shape best_sh = bt->shape;
    double best_t = 1e300;
    const int coops[] = {2, 1};
    const int waves[] = {2, 3};
    const int chunks[] = {64, 128, 256, 512};
    const int mults[] = {1, 2};
    StageOut so;
    double t = 0;
    int rc = batch_run_once(bt, bt->shape, so, &t);  // warm-up (plan, allocation, code load)
    for (int co : coops)
    for (int w : waves)
        for (int ch : chunks)
            for (int m : mults) {
                if (rc) break;
                Shape sh;
                sh.coop = co;
                sh.waves_per_eu = w;
                sh.chunk = ch;
                sh.grid_mult = m;
                double tt = 1e300;
                for (int rep = 0; rep < 2 && !rc; rep++) {
                    rc = batch_run_once(bt, sh, so, &t);
                    tt = std::min(tt, so.t_trace);
                }
                if (!rc && tt < best_t) {
                    best_t = tt;
                    best_sh = sh;
                }
            }
    bt->capture = cap;
    if (rc) return rc;
    bt->shape = best_sh;
    batch_release_graph(bt);
    if (best) {
        best->waves_per_eu = best_sh.waves_per_eu;
        best->chunk = best_sh.chunk;
        best->grid_mult = best_sh.grid_mult;
        best->cooperative = best_sh.coop;
    }
    if (best_seconds) *best_seconds = best_t;
    return ES_OK;
}

}  // extern "C"
