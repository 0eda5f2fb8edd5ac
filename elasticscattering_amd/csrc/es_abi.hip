// es_abi.hip — the C ABI (include/es_abi.h): a solver that owns device
// memory and one HIP stream, and drives the kernels of es_kernels.hip in the
// order of the reference's Simulation::ComputeSigmas
// (src/sim/SimulationCPU.cpp:33-73; src/sim/SimulationCL.cpp:72-91).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/es_abi.h"
#include "es_kernels.h"
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

}  // namespace

struct es_solver {
    int device = 0;
    hipStream_t stream = nullptr;
    es_solver_desc desc{};
    std::string err;
    size_t byte_cap = 0;  // 0: unlimited; else refuse allocations beyond this total
    size_t bytes_held = 0;

    // grid
    bool has_grid = false;
    es::GridDev g;
    int64_t unique = 0;
    DevBuf in_x, in_y, ixy, cell_end;
    DevBuf b_targets, b_ntarg, b_counts, b_fill, b_begin, b_bsum, b_order;

    // positions (per grid and solver)
    bool pos_valid = false;
    int nfree = 0;
    int64_t n_inside_positions = 0;
    DevBuf flags, bcount, free_pos, small_i32;  // small_i32: [0] err, [1] nfree

    // sample (InitSample)
    bool has_sample = false;
    es_settings st{};
    int coherent = 0;
    double integrand_angle_area = 0, phi_integrand_factor = 0, phi_start = 0, phi_step = 0;

    // per call
    DevBuf w, factor, dpart, mpart, sig, ltsum, mout, lifetimes;
    DevBuf lt_item, m_item, x_item, f_item, queue;
    int cus = 0;                      // compute units of the device
    int trace_bpc[2] = {0, 0};        // resident trace blocks per CU (counter-clockwise, clockwise)
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
};

namespace {

int fail(es_solver* s, int code, const std::string& msg)
{
    if (s) s->err = msg;
    return code;
}

int hip_fail(es_solver* s, hipError_t e, const char* what)
{
    return fail(s, ES_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define ES_HIP(s, call)                                 \
    do {                                                \
        hipError_t _e = (call);                         \
        if (_e != hipSuccess) return hip_fail(s, _e, #call); \
    } while (0)

int ensure(es_solver* s, DevBuf& b, size_t bytes, const char* what)
{
    if (bytes == 0) bytes = 8;
    if (b.cap >= bytes) return ES_OK;
    if (s->byte_cap && s->bytes_held - b.cap + bytes > s->byte_cap)
        return fail(s, ES_ERR_OOM, std::string("device size cap exceeded allocating ") + what);
    if (b.p) {
        (void)hipFree(b.p);
        s->bytes_held -= b.cap;
        b.p = nullptr;
        b.cap = 0;
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(s, ES_ERR_OOM, std::string("hipMalloc failed for ") + what + ": " + hipGetErrorString(e));
    }
    b.cap = bytes;
    s->bytes_held += bytes;
    return ES_OK;
}

void release(DevBuf& b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

template <typename T>
T* P(DevBuf& b)
{
    return reinterpret_cast<T*>(b.p);
}

// Grid.cpp:40,54 with the optional explicit count.
int cells_per_row_rule(int64_t unique, int target, int override_cpr)
{
    if (override_cpr > 0) return override_cpr;
    const int c = (int)std::round(std::sqrt(unique / (double)target));
    return std::max(c, 1);
}

// Bins the centres in s->in_x / s->in_y (n of them) into the current grid.
int bin_uploaded(es_solver* s, int64_t n, const es_grid_desc* d, es_grid_info* out)
{
    if (d->target_cell_population <= 0 && d->cells_per_row <= 0)
        return fail(s, ES_ERR_INVALID_ARG, "target_cell_population must be > 0");
    if (!(d->impurity_radius >= 0)) return fail(s, ES_ERR_INVALID_ARG, "impurity_radius must be >= 0");
    const int cpr = cells_per_row_rule(n, d->target_cell_population, d->cells_per_row);
    if ((int64_t)cpr * cpr > (int64_t)1 << 28) return fail(s, ES_ERR_INVALID_ARG, "cell grid too large");
    es::GridDev& g = s->g;
    g.cpr = cpr;
    g.lo = -d->region_extends;                       // Grid.cpp:52
    g.hi = d->region_size + d->region_extends;
    g.size = g.hi - g.lo;
    g.cell_size = g.size / (double)cpr;              // Grid.cpp:55, Grid.h:89
    g.radius = d->impurity_radius;
    const int64_t ncell = (int64_t)cpr * cpr;
    int rc;
    if ((rc = ensure(s, s->b_targets, 9 * n * sizeof(int32_t), "bin targets"))) return rc;
    if ((rc = ensure(s, s->b_ntarg, n, "bin target counts"))) return rc;
    if ((rc = ensure(s, s->b_counts, ncell * sizeof(int32_t), "cell counts"))) return rc;
    if ((rc = ensure(s, s->b_fill, ncell * sizeof(int32_t), "cell fill"))) return rc;
    if ((rc = ensure(s, s->b_begin, (ncell + 1) * sizeof(int64_t), "cell begin"))) return rc;
    if ((rc = ensure(s, s->b_bsum, ((ncell + 255) / 256) * sizeof(int64_t), "scan block sums"))) return rc;
    if ((rc = ensure(s, s->cell_end, ncell * sizeof(int32_t), "cell index"))) return rc;
    g.cell_end = P<int32_t>(s->cell_end);
    es::BinScratch bs{P<int32_t>(s->b_targets), P<uint8_t>(s->b_ntarg), P<int32_t>(s->b_counts),
                      P<int32_t>(s->b_fill), P<int64_t>(s->b_begin), P<int64_t>(s->b_bsum), nullptr};
    const double* x = P<double>(s->in_x);
    const double* y = P<double>(s->in_y);
    ES_HIP(s, es::launch_bin_count(x, y, n, g, bs, s->stream));
    ES_HIP(s, es::launch_bin_scan(g, bs, s->stream));
    int64_t total = 0;
    ES_HIP(s, hipMemcpyAsync(&total, P<int64_t>(s->b_begin) + ncell, sizeof(int64_t), hipMemcpyDeviceToHost,
                             s->stream));
    ES_HIP(s, hipStreamSynchronize(s->stream));
    if (total > INT32_MAX) return fail(s, ES_ERR_INVALID_ARG, "indexed impurity count exceeds int32");
    if ((rc = ensure(s, s->b_order, total * sizeof(int32_t), "cell order"))) return rc;
    if ((rc = ensure(s, s->ixy, total * sizeof(double2), "impurities"))) return rc;
    bs.order = P<int32_t>(s->b_order);
    g.ixy = P<double2>(s->ixy);
    g.total = total;
    ES_HIP(s, es::launch_bin_place(x, y, n, g, bs, s->stream));
    ES_HIP(s, hipStreamSynchronize(s->stream));
    s->unique = n;
    s->has_grid = true;
    s->pos_valid = false;
    // GridInformation of sample 0's grid when the solver was created without one
    // (SimulationRunner.cpp:75-84 builds the Simulation from the first grid).
    if (!(s->desc.grid_info.cell_size > 0)) {
        s->desc.grid_info.cell_size = g.cell_size;
        s->desc.grid_info.region_size = d->region_size;
        s->desc.grid_info.region_extends = d->region_extends;
        s->desc.grid_info.indexed_impurity_count = total;
        s->desc.grid_info.unique_impurity_count = n;
        s->desc.grid_info.index_size = (int32_t)ncell;
        s->desc.grid_info.cells_per_row = cpr;
    }
    if (out) {
        out->region_size = d->region_size;
        out->region_extends = d->region_extends;
        out->cell_size = g.cell_size;
        out->indexed_impurity_count = total;
        out->unique_impurity_count = n;
        out->index_size = (int32_t)ncell;
        out->cells_per_row = cpr;
    }
    return ES_OK;
}

// Start-cell test per position, once per grid (lifetime.h:84-99).
int ensure_positions(es_solver* s)
{
    if (s->pos_valid) return ES_OK;
    const int ppr = s->desc.positions_per_row;
    const int64_t n = (int64_t)ppr * ppr;
    int rc;
    if ((rc = ensure(s, s->flags, n, "position flags"))) return rc;
    if ((rc = ensure(s, s->free_pos, n * sizeof(int32_t), "free positions"))) return rc;
    if ((rc = ensure(s, s->bcount, ((n + 255) / 256) * sizeof(int32_t), "position block counts"))) return rc;
    if ((rc = ensure(s, s->small_i32, 2 * sizeof(int32_t), "flags"))) return rc;
    // Simulation.h:142-143
    const double dist = s->desc.grid_info.region_size / (double)(ppr - 1);
    const double offx = s->desc.grid_info.cell_size * 0.01;
    const double offy = s->desc.grid_info.cell_size * 0.005;
    int32_t* small = P<int32_t>(s->small_i32);
    ES_HIP(s, es::launch_positions(s->g, ppr, dist, offx, offy, P<uint8_t>(s->flags), P<int32_t>(s->bcount), small,
                                   P<int32_t>(s->free_pos), small + 1, s->stream));
    int32_t host[2] = {0, 0};
    ES_HIP(s, hipMemcpyAsync(host, small, sizeof(host), hipMemcpyDeviceToHost, s->stream));
    ES_HIP(s, hipStreamSynchronize(s->stream));
    if (host[0] != 0)
        return fail(s, ES_ERR_OUT_OF_GRID,
                    "a start position lies outside the impurity cell grid (region_extends smaller than the "
                    "position offset); the reference would read out of bounds (lifetime.h:88-91)");
    s->nfree = host[1];
    s->n_inside_positions = n - host[1];
    s->pos_valid = true;
    return ES_OK;
}

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
double integrand_factor(const es_solver* s, double w, double tau)
{
    const double a = 1.0 / (3.0 * (s->desc.positions_per_row - 1));
    return a * a * s->phi_integrand_factor * sigma_factor(s->st.particle_speed, w, tau) * 1e-8;
}

int compute(es_solver* s, const double* B, int nB, const double* T, int nT, es_sigma* out, es_metrics* metrics,
            double* lifetimes_host)
{
    if (!s) return ES_ERR_INVALID_ARG;
    if (!s->has_grid) return fail(s, ES_ERR_NO_GRID, "no grid set");
    if (!s->has_sample) return fail(s, ES_ERR_NO_SAMPLE, "es_solver_init_sample not called");
    if (!(s->desc.grid_info.region_size > 0)) return fail(s, ES_ERR_INVALID_ARG, "grid_info.region_size must be > 0");
    if (nB <= 0 || !B) return fail(s, ES_ERR_INVALID_ARG, "need at least one field");
    if (nT < 0 || (nT > 0 && (!T || !out))) return fail(s, ES_ERR_INVALID_ARG, "bad temperature arguments");
    for (int b = 0; b < nB; b++)
        if (!(B[b] > 0) || !std::isfinite(B[b])) return fail(s, ES_ERR_INVALID_ARG, "magnetic field must be > 0");
    // GetTau (Simulation.h:90) for each T.
    std::vector<double> taus(nT);
    for (int t = 0; t < nT; t++) {
        if (s->coherent)
            taus[t] = s->st.tau;
        else if (s->st.incoherent_tau > 0)
            taus[t] = s->st.incoherent_tau;
        else {
            if (!(T[t] > 0)) return fail(s, ES_ERR_INVALID_ARG, "temperature must be > 0 (or set incoherent_tau)");
            taus[t] = kHBAR / (kKB * T[t]);
        }
        if (!(taus[t] > 0) || !std::isfinite(taus[t])) return fail(s, ES_ERR_INVALID_ARG, "tau must be finite and > 0");
    }
    int rc;
    if ((rc = ensure_positions(s))) return rc;

    const int ppr = s->desc.positions_per_row, ppq = s->desc.particles_per_quadrant, ppp = 4 * ppq;
    const int64_t P_ = (int64_t)ppr * ppr * ppp;
    const int64_t items = (int64_t)s->nfree * ppp;
    const int64_t total_items = items * nB;
    const int nblk = (int)std::max<int64_t>(1, std::min<int64_t>((items + es::kBlock - 1) / es::kBlock, 1024));
    const int nTc = std::max(1, std::min(nT, es::kMaxT));
    const int nq = 2 * nTc + 1;
    const int cw = s->st.is_clockwise ? 1 : 0;

    if ((rc = ensure(s, s->w, nB * sizeof(double), "fields"))) return rc;
    if ((rc = ensure(s, s->factor, (size_t)nB * es::kMaxT * sizeof(double), "factors"))) return rc;
    if ((rc = ensure(s, s->dpart, (size_t)nB * nblk * nq * 2 * sizeof(double), "partials"))) return rc;
    if ((rc = ensure(s, s->mpart, (size_t)nB * nblk * es::kMetricSlots * sizeof(int64_t), "metric partials"))) return rc;
    if ((rc = ensure(s, s->sig, (size_t)nB * es::kMaxT * 2 * sizeof(double), "sigma out"))) return rc;
    if ((rc = ensure(s, s->ltsum, nB * sizeof(double), "lifetime sums"))) return rc;
    if ((rc = ensure(s, s->mout, (size_t)nB * es::kMetricSlots * sizeof(int64_t), "metrics out"))) return rc;
    if ((rc = ensure(s, s->lt_item, total_items * sizeof(double), "item lifetimes"))) return rc;
    if ((rc = ensure(s, s->m_item, total_items * sizeof(int4), "item counters"))) return rc;
    if ((rc = ensure(s, s->f_item, total_items, "item flags"))) return rc;
    if ((rc = ensure(s, s->x_item, total_items * sizeof(int32_t), "item exact solves"))) return rc;
    if ((rc = ensure(s, s->queue, sizeof(unsigned long long), "work queue"))) return rc;
    if (lifetimes_host) {
        if ((rc = ensure(s, s->lifetimes, (size_t)nB * P_ * sizeof(double), "lifetimes"))) return rc;
        ES_HIP(s, hipMemsetAsync(s->lifetimes.p, 0, (size_t)nB * P_ * sizeof(double), s->stream));
    }
    if (s->trace_bpc[cw] <= 0) {
        ES_HIP(s, es::trace_blocks_per_cu(cw, &s->trace_bpc[cw]));
        if (s->trace_bpc[cw] <= 0) s->trace_bpc[cw] = 1;
    }

    std::vector<double> w(nB);
    for (int b = 0; b < nB; b++) w[b] = kE * B[b] / kM;  // SimulationCPU.cpp:107
    ES_HIP(s, hipMemcpyAsync(s->w.p, w.data(), nB * sizeof(double), hipMemcpyHostToDevice, s->stream));

    // Lifetime stage: every (field, free position, phi) item once.
    es::TraceArgs a{};
    a.ixy = s->g.ixy;
    a.cell_end = s->g.cell_end;
    a.g_start = s->g.lo;
    a.g_size = s->g.size;
    a.g_L = s->g.size / (double)s->g.cpr;  // ImpuritySettings.cell_size (Grid.h:89)
    a.g_radius = s->g.radius;
    a.g_cpr = s->g.cpr;
    a.free_pos = P<int32_t>(s->free_pos);
    a.nfree = s->nfree;
    a.ppr = ppr;
    a.ppq = ppq;
    a.dist = s->desc.grid_info.region_size / (double)(ppr - 1);
    a.offx = s->desc.grid_info.cell_size * 0.01;
    a.offy = s->desc.grid_info.cell_size * 0.005;
    a.phi_start = s->phi_start;
    a.phi_step = s->phi_step;
    a.alpha = s->st.alpha;
    a.speed = s->st.particle_speed;
    a.coherent = s->coherent;
    a.clockwise = cw;
    a.nB = nB;
    a.w = P<double>(s->w);
    a.queue = P<unsigned long long>(s->queue);
    a.chunk = 128;
    a.lt_item = P<double>(s->lt_item);
    a.m_item = P<int4>(s->m_item);
    a.x_item = P<int32_t>(s->x_item);
    a.f_item = P<uint8_t>(s->f_item);
    a.lifetimes = lifetimes_host ? P<double>(s->lifetimes) : nullptr;
    const int64_t want_blocks = (total_items + es::kBlock - 1) / es::kBlock;
    const int tblocks = (int)std::max<int64_t>(1, std::min<int64_t>(want_blocks, (int64_t)s->trace_bpc[cw] * s->cus));

    ES_HIP(s, hipEventRecord(s->ev[0], s->stream));
    if (total_items > 0) ES_HIP(s, es::launch_trace(a, tblocks, s->stream));
    ES_HIP(s, hipEventRecord(s->ev[1], s->stream));

    // Temperature stage, at most kMaxT temperatures per launch.
    es::SigmaArgs sa{};
    sa.lt_item = a.lt_item;
    sa.m_item = a.m_item;
    sa.x_item = a.x_item;
    sa.f_item = a.f_item;
    sa.free_pos = a.free_pos;
    sa.nfree = s->nfree;
    sa.ppr = ppr;
    sa.ppq = ppq;
    sa.phi_start = s->phi_start;
    sa.phi_step = s->phi_step;
    sa.clockwise = cw;
    sa.nB = nB;
    sa.w = a.w;
    sa.dpart = P<double>(s->dpart);
    sa.mpart = P<int64_t>(s->mpart);
    sa.nblk = nblk;

    std::vector<double> sig((size_t)nB * es::kMaxT * 2);
    std::vector<double> ltsum(nB);
    std::vector<int64_t> mh((size_t)nB * es::kMetricSlots);
    double t_temp = 0;

    const int nchunks = nT == 0 ? 1 : (nT + es::kMaxT - 1) / es::kMaxT;
    for (int c = 0; c < nchunks; c++) {
        const int t0 = c * es::kMaxT;
        const int nt = nT == 0 ? 0 : std::min(es::kMaxT, nT - t0);
        sa.nT = nt;
        sa.with_metrics = (c == 0) ? 1 : 0;
        for (int t = 0; t < es::kMaxT; t++) {
            sa.tau[t] = t < nt ? taus[t0 + t] : 1.0;
            sa.max_lt[t] = 15.0 * sa.tau[t];  // Simulation.h:91
        }
        std::vector<double> fac((size_t)nB * std::max(nt, 1), 0.0);
        for (int b = 0; b < nB; b++)
            for (int t = 0; t < nt; t++) fac[(size_t)b * nt + t] = integrand_factor(s, w[b], taus[t0 + t]);
        ES_HIP(s, hipMemcpyAsync(s->factor.p, fac.data(), fac.size() * sizeof(double), hipMemcpyHostToDevice,
                                 s->stream));
        ES_HIP(s, hipEventRecord(s->ev[2], s->stream));
        ES_HIP(s, es::launch_sigma(sa, s->stream));
        ES_HIP(s, es::launch_finalize(P<double>(s->dpart), P<int64_t>(s->mpart), nB, nblk, nt, sa.with_metrics,
                                      P<double>(s->factor), P<double>(s->sig), P<double>(s->ltsum), P<int64_t>(s->mout),
                                      s->stream));
        ES_HIP(s, hipEventRecord(s->ev[3], s->stream));
        ES_HIP(s, hipMemcpyAsync(sig.data(), s->sig.p, (size_t)nB * std::max(nt, 1) * 2 * sizeof(double),
                                 hipMemcpyDeviceToHost, s->stream));
        if (c == 0) {
            ES_HIP(s, hipMemcpyAsync(ltsum.data(), s->ltsum.p, nB * sizeof(double), hipMemcpyDeviceToHost, s->stream));
            ES_HIP(s, hipMemcpyAsync(mh.data(), s->mout.p, mh.size() * sizeof(int64_t), hipMemcpyDeviceToHost,
                                     s->stream));
            if (lifetimes_host)
                ES_HIP(s, hipMemcpyAsync(lifetimes_host, s->lifetimes.p, (size_t)nB * P_ * sizeof(double),
                                         hipMemcpyDeviceToHost, s->stream));
        }
        ES_HIP(s, hipStreamSynchronize(s->stream));
        float ms2 = 0;
        ES_HIP(s, hipEventElapsedTime(&ms2, s->ev[2], s->ev[3]));
        t_temp += ms2 * 1e-3;
        for (int b = 0; b < nB; b++)
            for (int t = 0; t < nt; t++) {
                out[(size_t)b * nT + t0 + t].xx = sig[((size_t)b * nt + t) * 2];
                out[(size_t)b * nT + t0 + t].xy = sig[((size_t)b * nt + t) * 2 + 1];
            }
    }
    float ms1 = 0;
    ES_HIP(s, hipEventElapsedTime(&ms1, s->ev[0], s->ev[1]));
    const double t_life = ms1 * 1e-3;
    if (metrics) {
        const int64_t inside = s->n_inside_positions * ppp;
        for (int b = 0; b < nB; b++) {
            const int64_t* m = &mh[(size_t)b * es::kMetricSlots];
            es_metrics& o = metrics[b];
            o.cells_passed = m[0];
            o.impurities_tested = m[1];
            o.particles_escaped = m[2];
            o.particles_at_bound = m[3];
            o.trace_cap_hits = m[4];
            o.max_impurities_tested = m[5];
            o.max_cells_passed = m[6];
            o.particles_inside_impurity = inside;
            o.real_particles = P_ - inside;
            o.avg_particle_lifetime = P_ ? ltsum[b] / (double)P_ : 0.0;
            o.time_elapsed_lifetimes = t_life / nB;  // one launch for all fields: its share
            o.time_elapsed_temperatures = t_temp / nB;
            o.traced_particles = m[7];
            o.edge_hits = m[8];
            o.cross_calls = m[9];
            o.sigma_evals = m[10];
            o.cross_exact = m[11];
        }
    }
    return ES_OK;
}

}  // namespace

extern "C" {

const char* es_version(void) { return "es-mi355x 0.1 (gfx950, FP64)"; }

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
    if (desc->precision != ES_PRECISION_FP64) {
        g_create_error = "unsupported precision mode";
        return ES_ERR_INVALID_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        g_create_error = "no HIP device";
        return ES_ERR_NO_DEVICE;
    }
    if (desc->device < 0 || desc->device >= ndev) {
        g_create_error = "device ordinal out of range";
        return ES_ERR_INVALID_ARG;
    }
    es_solver* s = new es_solver();
    s->desc = *desc;
    s->device = desc->device;
    if (const char* cap = std::getenv("ES_MAX_DEVICE_BYTES")) s->byte_cap = (size_t)std::strtoull(cap, nullptr, 10);
    hipError_t e = hipSetDevice(s->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    for (int i = 0; i < 4 && e == hipSuccess; i++) e = hipEventCreate(&s->ev[i]);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, s->device);
    if (e != hipSuccess) {
        g_create_error = std::string("HIP setup failed: ") + hipGetErrorString(e);
        es_solver_destroy(s);
        return ES_ERR_HIP;
    }
    *out = s;
    return ES_OK;
}

void es_solver_destroy(es_solver* s)
{
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    DevBuf* bufs[] = {&s->in_x, &s->in_y, &s->ixy, &s->cell_end, &s->b_targets, &s->b_ntarg, &s->b_counts,
                      &s->b_fill, &s->b_begin, &s->b_bsum, &s->b_order, &s->flags, &s->bcount, &s->free_pos,
                      &s->small_i32, &s->w, &s->factor, &s->dpart, &s->mpart, &s->sig, &s->ltsum, &s->mout,
                      &s->lifetimes, &s->lt_item, &s->m_item, &s->x_item, &s->f_item, &s->queue};
    for (DevBuf* b : bufs) release(*b);
    for (hipEvent_t& ev : s->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

const char* es_last_error(const es_solver* s) { return s ? s->err.c_str() : g_create_error.c_str(); }

int es_grid_set_impurities(es_solver* s, const double* x, const double* y, int64_t n, const es_grid_desc* desc,
                           es_grid_info* out_info)
{
    if (!s) return ES_ERR_INVALID_ARG;
    if (!desc || n < 0 || (n > 0 && (!x || !y))) return fail(s, ES_ERR_INVALID_ARG, "bad impurity arguments");
    if (n > 10000000) return fail(s, ES_ERR_INVALID_ARG, "more than 1e7 impurities (Grid.cpp:36)");
    ES_HIP(s, hipSetDevice(s->device));
    int rc;
    if ((rc = ensure(s, s->in_x, n * sizeof(double), "impurity input x"))) return rc;
    if ((rc = ensure(s, s->in_y, n * sizeof(double), "impurity input y"))) return rc;
    if (n > 0) {
        ES_HIP(s, hipMemcpyAsync(s->in_x.p, x, n * sizeof(double), hipMemcpyHostToDevice, s->stream));
        ES_HIP(s, hipMemcpyAsync(s->in_y.p, y, n * sizeof(double), hipMemcpyHostToDevice, s->stream));
    }
    return bin_uploaded(s, n, desc, out_info);
}

int es_grid_generate(es_solver* s, uint64_t seed, int64_t count, const es_grid_desc* desc, es_grid_info* out_info)
{
    if (!s) return ES_ERR_INVALID_ARG;
    if (!desc) return fail(s, ES_ERR_INVALID_ARG, "null grid description");
    ES_HIP(s, hipSetDevice(s->device));
    const double lo = -desc->region_extends, hi = desc->region_size + desc->region_extends;
    int64_t n = count;
    if (n <= 0) {  // Grid.cpp:33-39
        const double area_length = hi - lo;
        const double proposal = area_length * area_length * desc->impurity_density;
        if (proposal > 1e7) return fail(s, ES_ERR_INVALID_ARG, "impurity count proposal > 1e7 (Grid.cpp:36)");
        if (!(proposal >= 1)) return fail(s, ES_ERR_INVALID_ARG, "no impurities (Grid.cpp:37)");
        n = std::max<int64_t>(1, (int64_t)std::ceil(proposal));
    }
    if (n > 10000000) return fail(s, ES_ERR_INVALID_ARG, "more than 1e7 impurities (Grid.cpp:36)");
    int rc;
    if ((rc = ensure(s, s->in_x, n * sizeof(double), "impurity input x"))) return rc;
    if ((rc = ensure(s, s->in_y, n * sizeof(double), "impurity input y"))) return rc;
    ES_HIP(s, es::launch_generate(esr_key(seed), n, lo, hi, P<double>(s->in_x), P<double>(s->in_y), s->stream));
    return bin_uploaded(s, n, desc, out_info);
}

int es_grid_read(es_solver* s, double* ordered_xy, int32_t* imp_index)
{
    if (!s) return ES_ERR_INVALID_ARG;
    if (!s->has_grid) return fail(s, ES_ERR_NO_GRID, "no grid set");
    ES_HIP(s, hipSetDevice(s->device));
    const int64_t total = s->g.total, ncell = (int64_t)s->g.cpr * s->g.cpr;
    if (ordered_xy && total > 0) {
        ES_HIP(s, hipMemcpyAsync(ordered_xy, s->g.ixy, total * sizeof(double2), hipMemcpyDeviceToHost, s->stream));
        ES_HIP(s, hipStreamSynchronize(s->stream));
    }
    if (imp_index) {
        ES_HIP(s, hipMemcpyAsync(imp_index, s->g.cell_end, ncell * sizeof(int32_t), hipMemcpyDeviceToHost, s->stream));
        ES_HIP(s, hipStreamSynchronize(s->stream));
    }
    return ES_OK;
}

int es_solver_init_sample(es_solver* s, const es_settings* settings, int32_t coherent)
{
    if (!s) return ES_ERR_INVALID_ARG;
    if (!settings) return fail(s, ES_ERR_INVALID_ARG, "null settings");
    if (!(settings->particle_speed > 0)) return fail(s, ES_ERR_INVALID_ARG, "particle_speed must be > 0");
    if (!(settings->alpha >= 0) || !(settings->alpha <= kPI / 4 + 1e-12))
        return fail(s, ES_ERR_INVALID_ARG, "alpha must lie in [0, pi/4]");
    s->st = *settings;
    s->coherent = coherent ? 1 : 0;
    // Simulation::InitSample (Simulation.h:111-126)
    const int ppq = s->desc.particles_per_quadrant;
    const double base_area = settings->alpha * 2.0;
    s->integrand_angle_area = !coherent ? base_area : (kHALF_PI - base_area);
    s->phi_integrand_factor = s->integrand_angle_area / ((ppq - 1) * 3.0);
    s->phi_start = !coherent ? -settings->alpha : settings->alpha;
    s->phi_step = s->integrand_angle_area / (ppq - 1);
    s->has_sample = true;
    return ES_OK;
}

int es_solver_compute_sigmas(es_solver* s, double B, const double* T, int32_t nT, es_sigma* out, es_metrics* metrics,
                             double* lifetimes)
{
    if (!s) return ES_ERR_INVALID_ARG;
    hipError_t e = hipSetDevice(s->device);
    if (e != hipSuccess) return hip_fail(s, e, "hipSetDevice");
    return compute(s, &B, 1, T, nT, out, metrics, lifetimes);
}

int es_solver_compute_sigmas_sweep(es_solver* s, const double* B, int32_t nB, const double* T, int32_t nT,
                                   es_sigma* out, es_metrics* metrics)
{
    if (!s) return ES_ERR_INVALID_ARG;
    hipError_t e = hipSetDevice(s->device);
    if (e != hipSuccess) return hip_fail(s, e, "hipSetDevice");
    return compute(s, B, nB, T, nT, out, metrics, nullptr);
}

}  // extern "C"
