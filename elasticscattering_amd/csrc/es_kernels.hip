// es_kernels.hip — gfx950 kernels of the lifetime / sigma path.
//
//   k_bin_count     per impurity (drawn first from es_rng.h when generating,
//                   the Grid.cpp:61-75 analogue): target cells (Grid.cpp:
//                   112-180), per-cell counts, each entry's slot in its cell
//   k_scan_one /
//   k_scan_block,
//   k_scan_add      exclusive scan -> cell begin offsets + the reference's
//                   inclusive ends (one block up to 32,768 cells)
//   k_bin_fill      scatter impurity ids into their slots (no atomics)
//   k_bin_sort_gather  ascending impurity id inside each cell (the
//                   reference's order) and the cell-ordered (x, y) pairs, one
//                   wave per cell
//   k_positions     start cell + "inside an impurity" test per position
//                   (lifetime.h:84-99), hoisted out of the phi and B loops
//   k_compact       ordered list of free positions (deterministic)
//   k_trace_w*<CW>  CreateParticle + TraceOrbit (lifetime.h:36-148) per work
//                   item; a lane whose trace ends takes the next item of its
//                   wave's queue chunk, so long and short traces share waves;
//                   per-item raw lifetime + work counters
//   k_sigma<NT>     truncation, GetSigma, Simpson weights (SimulationCPU.cpp:
//                   143-222) per item and temperature; double-double block
//                   partials per field, no atomics
//   k_finalize      fixed-order double-double sum of the block partials
//
// All FP64; compiled with -ffp-contract=off.
#include "es_kernels.h"
#include "es_rng.h"
#include "es_trace.h"
#include "es_trace32.h"

namespace es {

namespace {

struct DD {
    double s, c;
};

__device__ __forceinline__ void dd_add_d(DD& a, double x)
{
    const double t = a.s + x;
    if (fabs(a.s) >= fabs(x))
        a.c += (a.s - t) + x;
    else
        a.c += (x - t) + a.s;
    a.s = t;
}

__device__ __forceinline__ DD dd_add(DD a, DD b)
{
    const double s = a.s + b.s;
    const double bb = s - a.s;
    double e = (a.s - (s - bb)) + (b.s - bb);
    e += a.c + b.c;
    const double r = s + e;
    DD out;
    out.s = r;
    out.c = e - (r - s);
    return out;
}

__device__ __forceinline__ DD dd_wave_reduce(DD v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        DD o;
        o.s = __shfl_xor(v.s, off, 64);
        o.c = __shfl_xor(v.c, off, 64);
        v = dd_add(v, o);
    }
    return v;
}

__device__ __forceinline__ int64_t i64_wave_sum(int64_t v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ int64_t i64_wave_max(int64_t v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const int64_t o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

__device__ __forceinline__ double simpson_weight(int i, int dim)  // util.h:24-28
{
    const double main_multiplier = (i % 2 == 0) ? 2.0 : 4.0;
    const bool is_edge = (i == 0) || i == (dim - 1);
    return is_edge ? 1.0 : main_multiplier;
}

// Sum of the dim Simpson weights (an integer, exact in a double).
__device__ __forceinline__ double simpson_weight_sum(int dim)
{
    double s = 0.0;
    for (int i = 0; i < dim; i++) s += simpson_weight(i, dim);
    return s;
}

// a += x * y with the product's rounding error added too (fma): the double-
// double accumulator receives x * y exactly.
__device__ __forceinline__ void dd_add_exact_product(DD& a, double x, double y)
{
    const double hi = x * y;
    const double lo = fma(x, y, -hi);
    dd_add_d(a, hi);
    dd_add_d(a, lo);
}


// ------------------------------------------------------------------- binning

struct BinGeom {
    double lo, size, cell_size, r;
    double inv_size;  // RN(1 / size): esm_div_rcp gives the IEEE quotient's bits
    int cpr;
};

__device__ __forceinline__ void grid_cell(const BinGeom& g, double x, double y, int& cx, int& cy)
{
    cx = (int)(esm_div_rcp(x - g.lo, g.size, g.inv_size) * (double)(g.cpr));  // Grid.cpp:182-188
    cy = (int)(esm_div_rcp(y - g.lo, g.size, g.inv_size) * (double)(g.cpr));
}

__device__ __forceinline__ bool in_grid(int cx, int cy, int cpr)
{
    return (cx >= 0 && cx < cpr) && (cy >= 0 && cy < cpr);
}

// Grid.cpp:112-180 for impurity (x, y): the up to 9 distinct target cells,
// kept in registers as 9 fixed candidates with a bit each in the returned
// mask (bit k: cell[k] is a target).  Candidates 0-4 are the probe cells
// {c, c+(r,0), c-(r,0), c+(0,r), c-(0,r)}: one is a target when it lies in
// the grid and differs from every earlier probe (the reference's std::set;
// an earlier equal probe outside the grid makes this one outside too).
// Candidates 5-8 are the diagonal cells tl, tr, ll, lr of the centre cell,
// taken when that corner lies within r (strictly) and the cell is in the grid.
__device__ __forceinline__ unsigned int bin_candidates(const BinGeom& g, double x, double y, int32_t (&cell)[9])
{
    int px[5], py[5];
    grid_cell(g, x, y, px[0], py[0]);
    grid_cell(g, x + g.r, y, px[1], py[1]);
    grid_cell(g, x - g.r, y, px[2], py[2]);
    grid_cell(g, x, y + g.r, px[3], py[3]);
    grid_cell(g, x, y - g.r, px[4], py[4]);
    unsigned int mask = 0u;
#pragma unroll
    for (int k = 0; k < 5; k++) {
        bool seen = false;
#pragma unroll
        for (int m = 0; m < k; m++) seen = seen || (px[m] == px[k] && py[m] == py[k]);
        mask |= (!seen && in_grid(px[k], py[k], g.cpr)) ? (1u << k) : 0u;
        cell[k] = py[k] * g.cpr + px[k];
    }
    const double llx = g.lo + px[0] * g.cell_size;  // Grid.h:64-67
    const double lly = g.lo + py[0] * g.cell_size;
    const double lrx = llx + g.cell_size, lry = lly + 0.0;
    const double tlx = llx + 0.0, tly = lly + g.cell_size;
    const double trx = llx + g.cell_size, try_ = lly + g.cell_size;
    const double ir2 = g.r * g.r;
    const int cx = px[0], cy = py[0];
    const double qx[4] = {tlx, trx, llx, lrx}, qy[4] = {tly, try_, lly, lry};
    const int ox[4] = {-1, 1, -1, 1}, oy[4] = {1, 1, -1, -1};
#pragma unroll
    for (int d = 0; d < 4; d++) {
        const double dx = qx[d] - x, dy = qy[d] - y;
        const bool hit = (dx * dx + dy * dy < ir2) && in_grid(cx + ox[d], cy + oy[d], g.cpr);
        mask |= hit ? (1u << (5 + d)) : 0u;
        cell[5 + d] = (cy + oy[d]) * g.cpr + (cx + ox[d]);
    }
    return mask;
}

// Every kernel of the binning and of the start-cell test works on a table of
// jobs, one field each (blockIdx.y, or blockIdx.x for the one-block scan): a
// batch of new fields is binned with one launch per stage.

// Targets and per-cell counts of every impurity; with `generate` the centres
// are drawn first (es_rng.h) and stored in x, y.  counts are zero on entry.
// All of an impurity's counting atomics are issued before their results are
// used; the targets are stored compacted in candidate order.
__global__ __launch_bounds__(256) void k_bin_count(const BinJob* jobs)
{
    const BinJob& J = jobs[blockIdx.y];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= J.n) return;
    double px, py;
    if (J.generate) {
        esr_impurity(J.key, (uint64_t)i, J.gen_lo, J.gen_hi, &px, &py);
        J.x[i] = px;
        J.y[i] = py;
    } else {
        px = J.x[i];
        py = J.y[i];
    }
    const BinGeom g{J.lo, J.size, J.cell_size, J.radius, 1.0 / J.size, J.cpr};
    int32_t cell[9];
    const unsigned int mask = bin_candidates(g, px, py, cell);
    // The count's old value is the entry's slot in its cell (any order: the
    // cell is sorted by impurity id afterwards).
    int32_t slot[9];
#pragma unroll
    for (int k = 0; k < 9; k++) slot[k] = ((mask >> k) & 1u) ? atomicAdd(&J.s.counts[cell[k]], 1) : 0;
    J.s.ntarg[i] = (uint8_t)__popc(mask);
    int32_t* const tg = J.s.targets + 9 * i;
    int32_t* const sl = J.s.slot + 9 * i;
    int n = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        if ((mask >> k) & 1u) {
            tg[n] = cell[k];
            sl[n] = slot[k];
            n++;
        }
    }
}

// Block-local exclusive scan of 256 int64 values held one per thread.
__device__ __forceinline__ int64_t block_exclusive_scan(int64_t v, int64_t* sh /* [256] */, int64_t& total)
{
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int64_t add = (tid >= off) ? sh[tid - off] : 0;
        __syncthreads();
        sh[tid] += add;
        __syncthreads();
    }
    total = sh[255];
    const int64_t r = sh[tid] - v;
    __syncthreads();
    return r;
}

// Sum of bsum[0 .. nb) by one block (fixed order per thread, then LDS tree).
__device__ __forceinline__ int64_t block_prefix_of(const int64_t* bsum, int nb, int64_t* sh)
{
    int64_t acc = 0;
    for (int k = threadIdx.x; k < nb; k += 256) acc += bsum[k];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    const int64_t r = sh[0];
    __syncthreads();
    return r;
}

// Two-level scan for one job of a large cell grid (more than kScanOneCells).
__global__ __launch_bounds__(256) void k_scan_block(const BinJob* job)
{
    const BinJob& J = *job;
    const int64_t n = (int64_t)J.cpr * J.cpr;
    __shared__ int64_t sh[256];
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t v = (c < n) ? J.s.counts[c] : 0;
    int64_t total;
    const int64_t ex = block_exclusive_scan(v, sh, total);
    if (c < n) J.s.begin[c] = ex;
    if (threadIdx.x == 0) J.s.bsum[blockIdx.x] = total;
}

// Second level of the two-level scan; also clears the counts for the next
// binning, and writes the total.
__global__ __launch_bounds__(256) void k_scan_add(const BinJob* job)
{
    const BinJob& J = *job;
    const int64_t n = (int64_t)J.cpr * J.cpr;
    __shared__ int64_t sh[256];
    const int64_t off = block_prefix_of(J.s.bsum, blockIdx.x, sh);
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c < n) {
        const int64_t b = J.s.begin[c] + off;
        const int32_t cnt = J.s.counts[c];
        J.s.begin[c] = b;
        J.cell_end[c] = (int32_t)(b + cnt);
        J.s.counts[c] = 0;
        if (c == n - 1) {
            J.s.begin[n] = b + cnt;
            *J.total_out = b + cnt;
        }
    }
}

// The same scan in one block of 1024 threads per job (blockIdx.x) for cell
// grids up to kScanOneCells cells, staged in LDS: coalesced loads of the
// counts, each thread sums a contiguous run, the block scans the run sums, the
// runs are turned into exclusive offsets in place, and coalesced stores write
// them out.  Jobs with larger grids are left to the two-level scan.
constexpr int kScanOneCells = 32768;

__global__ __launch_bounds__(1024) void k_scan_one(const BinJob* jobs)
{
    const BinJob& J = jobs[blockIdx.x];
    const int64_t n = (int64_t)J.cpr * J.cpr;
    if (n > kScanOneCells) return;
    extern __shared__ int32_t sh_cnt[];  // [n]
    __shared__ int64_t sh[1024];
    const int tid = threadIdx.x;
    const int ni = (int)n;
    int32_t* __restrict__ counts = J.s.counts;
    for (int c = tid; c < ni; c += 1024) sh_cnt[c] = counts[c];
    for (int c = tid; c < ni; c += 1024) counts[c] = 0;
    __syncthreads();
    const int per = (ni + 1023) / 1024;
    const int c0 = tid * per < ni ? tid * per : ni, c1 = (c0 + per < ni) ? c0 + per : ni;
    int64_t run = 0;
    for (int c = c0; c < c1; c++) run += sh_cnt[c];
    sh[tid] = run;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int64_t add = (tid >= off) ? sh[tid - off] : 0;
        __syncthreads();
        sh[tid] += add;
        __syncthreads();
    }
    const int64_t total = sh[1023];
    int64_t b = sh[tid] - run;
    for (int c = c0; c < c1; c++) {
        const int32_t cnt = sh_cnt[c];
        sh_cnt[c] = (int32_t)b;  // exclusive offset (the total fits in int32: checked by the host)
        b += cnt;
    }
    __syncthreads();
    for (int c = tid; c < ni; c += 1024) {
        J.s.begin[c] = sh_cnt[c];
        J.cell_end[c] = (c + 1 < ni) ? sh_cnt[c + 1] : (int32_t)total;
    }
    if (tid == 0) {
        J.s.begin[n] = total;
        *J.total_out = total;
    }
}

__global__ __launch_bounds__(256) void k_bin_fill(const BinJob* jobs)
{
    const BinJob& J = jobs[blockIdx.y];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= J.n) return;
    const int k = J.s.ntarg[i];
    for (int m = 0; m < k; m++) J.s.order[J.s.begin[J.s.targets[9 * i + m]] + J.s.slot[9 * i + m]] = (int32_t)i;
}

__device__ void sift_down(int32_t* a, int64_t root, int64_t n)
{
    while (true) {
        int64_t child = 2 * root + 1;
        if (child >= n) return;
        if (child + 1 < n && a[child + 1] > a[child]) child++;
        if (a[root] >= a[child]) return;
        const int32_t t = a[root];
        a[root] = a[child];
        a[child] = t;
        root = child;
    }
}

// Ascending impurity ids inside each cell — the reference pushes impurities
// into a cell in index order (Grid.cpp:87-89,129) — and the cell's (x, y)
// pairs in that order.  One wave per cell; ids in a cell are distinct, so a
// lane's rank is the count of smaller ids (a heap sort by one lane for cells
// of more than 64).
__global__ __launch_bounds__(256) void k_bin_sort_gather(const BinJob* jobs)
{
    const BinJob& J = jobs[blockIdx.y];
    const int64_t ncell = (int64_t)J.cpr * J.cpr;
    const int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (c >= ncell) return;
    const int64_t b = J.s.begin[c];
    int32_t* a = J.s.order + b;
    const int64_t n = J.s.begin[c + 1] - b;
    if (n <= 0) return;
    const double* x = J.x;
    const double* y = J.y;
    if (n <= 64) {
        const int32_t v = (lane < n) ? a[lane] : INT32_MAX;
        int rank = 0;
        for (int k = 0; k < (int)n; k++) {
            const int32_t o = __shfl(v, k, 64);
            rank += (o < v) ? 1 : 0;
        }
        if (lane < n) {
            a[rank] = v;
            J.ixy[b + rank] = make_double2(x[v], y[v]);
        }
        return;
    }
    if (lane == 0) {
        for (int64_t r = n / 2 - 1; r >= 0; r--) sift_down(a, r, n);
        for (int64_t e = n - 1; e > 0; e--) {
            const int32_t t = a[0];
            a[0] = a[e];
            a[e] = t;
            sift_down(a, 0, e);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int64_t k = lane; k < n; k += 64) {
        const int32_t v = a[k];
        J.ixy[b + k] = make_double2(x[v], y[v]);
    }
}

// ----------------------------------------------------------------- positions

constexpr int32_t kPosErrBit = (int32_t)0x80000000;

// Start cell + inside test per position; per block the free count, with
// kPosErrBit set when a position lies outside the cell grid.
__global__ __launch_bounds__(256) void k_positions(const PosJob* jobs)
{
    const PosJob& J = jobs[blockIdx.y];
    __shared__ int sh[4];
    __shared__ int sh_err;
    const int ppr = J.ppr;
    const int64_t n = (int64_t)ppr * ppr;
    if ((int64_t)blockIdx.x * 256 >= n) return;  // whole block past this job's positions
    if (threadIdx.x == 0) sh_err = 0;
    __syncthreads();
    const int64_t pidx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    uint8_t f = 1;
    if (pidx < n) {
        const int i = (int)(pidx % ppr), j = (int)(pidx / ppr);
        const double px = (double)i * J.dist + J.offx;  // SimulationCPU.cpp:114
        const double py = (double)j * J.dist + J.offy;
        const int cpr = J.cpr;
        const int cx = (int)((px - J.lo) / J.size * (double)(cpr));  // cell_grid.h:52-58
        const int cy = (int)((py - J.lo) / J.size * (double)(cpr));
        if (!(cx >= 0 && cx < cpr && cy >= 0 && cy < cpr)) {
            f = 2;
            sh_err = 1;
        } else {
            const int ci = cy * cpr + cx;
            const int s = (ci == 0) ? 0 : J.cell_end[ci - 1];
            const int e = J.cell_end[ci];
            const double r2 = J.radius * J.radius;
            f = 0;
            for (int k = s; k < e; k++) {  // InsideImpurity, details.h:208-213
                const double2 q = J.ixy[k];
                const double dx = px - q.x, dy = py - q.y;
                if (r2 > (dx * dx + dy * dy)) {
                    f = 1;
                    break;
                }
            }
        }
        J.flags[pidx] = f;
    }
    const unsigned long long m = __ballot(pidx < n && f == 0);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) J.bcount[blockIdx.x] = (sh[0] + sh[1] + sh[2] + sh[3]) | (sh_err ? kPosErrBit : 0);
}

// Ordered compaction of the free positions; the job's last block writes the
// free count and the out-of-grid flag.
__global__ __launch_bounds__(256) void k_compact(const PosJob* jobs)
{
    const PosJob& J = jobs[blockIdx.y];
    const int64_t n = (int64_t)J.ppr * J.ppr;
    const int nblocks = (int)((n + 255) / 256);
    if ((int)blockIdx.x >= nblocks) return;
    __shared__ int64_t sh[256];
    __shared__ int sh_e[256];
    __shared__ int wsum[4];
    const bool last = (int)blockIdx.x == nblocks - 1;
    int64_t acc = 0;
    int err = 0;
    const int upto = last ? nblocks : (int)blockIdx.x;  // the last block also ORs its own error bit
    for (int k = threadIdx.x; k < upto; k += 256) {
        const int32_t v = J.bcount[k];
        if (k < (int)blockIdx.x) acc += v & ~kPosErrBit;
        err |= (v & kPosErrBit) ? 1 : 0;
    }
    sh[threadIdx.x] = acc;
    sh_e[threadIdx.x] = err;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            sh[threadIdx.x] += sh[threadIdx.x + off];
            sh_e[threadIdx.x] |= sh_e[threadIdx.x + off];
        }
        __syncthreads();
    }
    const int64_t base = sh[0];
    const int64_t pidx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool is_free = pidx < n && J.flags[pidx] == 0;
    const unsigned long long m = __ballot(is_free);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wsum[wave] = __popcll(m);
    __syncthreads();
    int woff = 0;
    for (int w = 0; w < wave; w++) woff += wsum[w];
    const unsigned long long lower = (lane == 0) ? 0ull : (m & (~0ull >> (64 - lane)));
    if (is_free) J.free_pos[base + woff + __popcll(lower)] = (int32_t)pidx;
    if (last && threadIdx.x == 0) {
        *J.nfree = (int32_t)(base + wsum[0] + wsum[1] + wsum[2] + wsum[3]);
        *J.err = sh_e[0];
    }
}

// --------------------------------------------------------------------- trace

// Pair owning work item g (pairs are sorted by item_off).
__device__ __forceinline__ int pair_of_item(const PairDev* pairs, int npairs, long long g)
{
    int lo = 0, hi = npairs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pairs[mid].item_off <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Pair owning temperature-stage block k (pairs are sorted by blk_off).
__device__ __forceinline__ int pair_of_block(const PairDev* pairs, int npairs, int k)
{
    int lo = 0, hi = npairs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pairs[mid].blk_off <= k) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The pair of a chunk's first item, found once per chunk (wave-uniform), and
// the item range it covers: items of the chunk inside it skip their own search.
struct PairHint {
    int pair = 0;
    long long begin = 0, end = -1;
};

__device__ __forceinline__ int hinted_pair(const TraceArgs& a, const PairHint& h, long long item)
{
    return (item >= h.begin && item < h.end) ? h.pair : pair_of_item(a.pairs, a.npairs, item);
}

// Work item k of a pair: free-position slot, quadrant q and phi point p of
// the traced points (tpq per quadrant), and the first lifetime slot it fills
// in the reference layout [j][i][q][p] (with tpq = 1 it fills all ppq).
struct ItemPos {
    long long slot;
    int q, p;
    int pidx;
    long long out_idx;
};

__device__ __forceinline__ ItemPos item_pos(const TraceArgs& a, const PairDev& pr, const SampleDev& sd, long long item)
{
    ItemPos r;
    const int tpq = sd.tpq, tpp = 4 * sd.tpq;
    const long long k = item - pr.item_off;
    if (pr.nitems <= 0x7fffffffll) {  // a 32-bit division (a 64-bit one is a long software sequence)
        r.slot = (long long)((unsigned int)k / (unsigned int)tpp);
    } else {
        r.slot = k / tpp;
    }
    const int v = (int)(k - r.slot * tpp);
    r.q = (int)((unsigned int)v / (unsigned int)tpq);
    r.p = v - r.q * tpq;
    r.pidx = sd.free_pos[r.slot];
    r.out_idx = a.lifetimes ? (long long)pr.out * a.P + (long long)r.pidx * (4 * sd.ppq) + r.q * sd.ppq + r.p : -1;
    return r;
}

// CreateParticle (lifetime.h:36-58) and the trace start for work item `item`:
// the lane's grid view, orbit, start angle, bound time and output slot.
template <bool CW>
__device__ __forceinline__ void d_setup_item(const TraceArgs& a, const PairHint& hint, long long item, DevGrid& g,
                                             Orb& o, double& phi0, double& bound_time, long long& out_idx, int& cap,
                                             TraceState& st)
{
    const int pi = hinted_pair(a, hint, item);
    const PairDev& pr = a.pairs[pi];
    const SampleDev& sd = a.samples[pr.sample];
    const ItemPos ip = item_pos(a, pr, sd, item);
    const int q = ip.q, p = ip.p, pidx = ip.pidx;
    const int j = (int)((unsigned int)pidx / (unsigned int)sd.ppr), i = pidx - j * sd.ppr;
    out_idx = ip.out_idx;
    g.ixy = sd.ixy;
    g.cell_end = sd.cell_end;
    g.start = sd.start;
    g.size = sd.size;
    g.L = sd.L;
    g.inv_L = sd.inv_L;
    g.radius = sd.radius;
    g.cpr = sd.cpr;
    cap = 4 * sd.cpr * sd.cpr + 16;
    const double px = (double)i * sd.dist + sd.offx;  // SimulationCPU.cpp:114
    const double py = (double)j * sd.dist + sd.offy;
    const double w = pr.w;
    const double phi = sd.phi_start + q * ES_HALF_PI + p * sd.phi_step;
    const double2 sc = sd.phi_sc[q * sd.ppq + p];
    const double sphi = sc.x, cphi = sc.y;
    const double vx = cphi * sd.speed, vy = sphi * sd.speed;
    const double R = pr.R;  // speed / w
    const double sx = esm_div_rcp(vy * R, sd.speed, sd.inv_speed);
    const double sy = esm_div_rcp(-vx * R, sd.speed, sd.inv_speed);
    const double radd = R + g.radius, rmin = R - g.radius;
    o.cx = CW ? (px - sx) : (px + sx);
    o.cy = CW ? (py - sy) : (py + sy);
    o.R = R;
    o.R2 = R * R;
    o.eps_r = 1e-10 * R;
    o.radd2 = radd * radd;
    o.rmin2 = rmin * rmin;
    o.r2 = g.radius * g.radius;
    o.pa = phi;
    o.w = w;
    o.inv_R = pr.inv_R;
    o.inv_w = pr.inv_w;
    o.u0x = px - o.cx;
    o.u0y = py - o.cy;
    o.r = (float)g.radius;
    phi0 = phi;
    bound_time = ES_INF;  // GetBoundTime (details.h:219-228)
    if (!sd.coherent) {
        const double a = phi + sd.alpha;  // smod(a, pi / 2), details.h:79-83
        const double remaining = a - ES_HALF_PI * floor(esm_div_rcp(a, ES_HALF_PI, ES_INV_HALF_PI));
        const double dphi = CW ? remaining : (2.0 * sd.alpha - remaining);
        bound_time = esm_div_rcp(dphi, w, o.inv_w);
    }
    const int c0x = (int)(esm_div_rcp(px - g.start, g.size, sd.inv_size) * (double)(g.cpr));
    const int c0y = (int)(esm_div_rcp(py - g.start, g.size, sd.inv_size) * (double)(g.cpr));
    d_trace_begin(st, px, py, phi, c0x, c0y);
}

// The lifetime in the reference layout: one slot, or every phi point of the
// quadrant when they coincide (tpq = 1).
__device__ __forceinline__ void store_lifetime(const TraceArgs& a, long long item, long long out_idx, double lt)
{
    const PairDev& pr = a.pairs[pair_of_item(a.pairs, a.npairs, item)];
    const SampleDev& sd = a.samples[pr.sample];
    const int n = (sd.tpq == 1) ? sd.ppq : 1;
    for (int d = 0; d < n; d++) a.lifetimes[out_idx + d] = lt;
}

// Writes the per-item results of an ended trace (lifetime.h:142-147).
__device__ __forceinline__ void d_store_item(const TraceArgs& a, long long item, long long out_idx,
                                             double bound_time, const TraceState& st, int flags)
{
    double lt = st.lifetime;
    if (lt > bound_time) {  // lifetime.h:142-146
        lt = bound_time;
        flags |= ES_F_AT_BOUND;
    }
    a.lt_item[item] = lt;
    a.m_item[item] = make_int4(st.cells, st.tested, st.edge_hits, st.cross_calls);
    a.x_item[item] = st.cross_exact;
    a.f_item[item] = (uint8_t)(flags | (a.list_mode ? ES_F_RECOMPUTE : 0));
    if (out_idx >= 0) store_lifetime(a, item, out_idx, lt);
}

// Lane refill: every idle lane takes the next item of the wave's chunk of the
// work queue (a fresh chunk of a.chunk items when the current one is used up;
// one atomic per chunk on the queue counter).  Returns the item to a lane that
// got one, -1 to the others.  Wave-uniform state: q_next, q_end, exhausted.
__device__ __forceinline__ void d_refill(const TraceArgs& a, long long total, bool active, long long& q_next,
                                         long long& q_end, bool& exhausted, long long& got, PairHint& hint)
{
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    got = -1;
    unsigned long long need = __ballot(!active);
    while (need != 0ull && !exhausted) {
        if (q_next >= q_end) {
            unsigned long long base = 0;
            if (lane == 0) base = atomicAdd(a.queue, (unsigned long long)a.chunk);
            base = ((unsigned long long)(unsigned int)__builtin_amdgcn_readfirstlane((int)(base >> 32)) << 32) |
                   (unsigned long long)(unsigned int)__builtin_amdgcn_readfirstlane((int)(unsigned int)base);
            if ((long long)base >= total) {
                exhausted = true;
                break;
            }
            q_next = (long long)base;
            q_end = (q_next + a.chunk < total) ? q_next + a.chunk : total;
            if (!a.list_mode) {
                hint.pair = pair_of_item(a.pairs, a.npairs, q_next);
                hint.begin = a.pairs[hint.pair].item_off;
                hint.end = hint.begin + a.pairs[hint.pair].nitems;
            }
        }
        const long long left = q_end - q_next;
        const int n = __popcll(need);
        const int take = (int)((long long)n < left ? (long long)n : left);
        const bool idle = (need >> lane) & 1ull;
        if (idle) {
            const int rank = __popcll(need & below);
            if (rank < take) got = a.list_mode ? (long long)a.list[q_next + rank] : q_next + rank;
        }
        q_next += take;
        need = __ballot(!active && got < 0);
    }
}

template <bool CW>
__device__ __forceinline__ void trace_body(const TraceArgs& a)
{
    const long long total = a.list_mode ? (long long)*a.list_count : a.total_items;

    long long q_next = 0, q_end = 0;  // this wave's current chunk of the queue (uniform)
    PairHint hint;
    bool exhausted = false;
    bool active = false;
    long long item = 0, out_idx = -1;
    int cap = 0;
    DevGrid g;
    Orb o;
    double phi0 = 0.0, bound_time = 0.0;
    TraceState st;

    while (true) {
        long long got;
        d_refill(a, total, active, q_next, q_end, exhausted, got, hint);
        if (got >= 0) {
            item = got;
            d_setup_item<CW>(a, hint, item, g, o, phi0, bound_time, out_idx, cap, st);
            active = true;
        }
        if (__ballot(active) == 0ull) break;  // queue drained and every lane done

        if (active) {
            int flags = 0;
            if (d_step<CW>(o, phi0, g, cap, st, flags)) {
                d_store_item(a, item, out_idx, bound_time, st, flags);
                active = false;
            }
        }
    }
}

// ------------------------------------------------ wave-cooperative FP64 trace
//
// Same lane refill and per-lane cell walk as trace_body; what changes is who
// solves which crossing disk, and which disks need solving at all.  Per cell
// step of a wave:
//   1. each lane runs CirclesCross (details.h:141-151) on its cell's disks,
//      32 at a time into a bit mask, then takes its crossing disks one per
//      pass: it drops those whose centre lies outside the window of the
//      cell's angle range and gives the others a lower bound on the travel
//      angle of their meeting points (d_range_window, d_disk_screen); it keeps
//      the disk with the lowest bound and queues the others on the wave's
//      stack in LDS (the screen's cost follows the crossing disks, about half
//      of those tested, and the wave's passes the lane with the most of them);
//   2. each lane solves its lowest-bound disk (GetFirstCrossTime,
//      details.h:153-206) and turns the time into a travel-angle threshold: a
//      disk whose bound reaches it cannot meet the orbit before that time
//      (d_time_threshold);
//   3. a pass over the stack drops those disks and compacts the rest;
//   4. the rest are solved 64 at a time, any lane taking any lane's disk, with
//      the owner's orbit and angle range read from LDS and the time folded
//      into the owner's minimum by an LDS atomic min on an order-preserving
//      integer key.
// Without SCREEN, steps 2 and 3 are skipped and every crossing disk is solved
// (sparse fields, where few disks cross in a cell, are faster so).
// A nearly full stack goes through 2-4 at once and the tests resume.  Every
// solve runs the same operations on the same operands as in trace_body, a
// minimum does not depend on the order, and a dropped disk cannot lower it, so
// lifetimes are bit-identical; the work counter of exact solves (cross_exact)
// is what shrinks.

constexpr int kCoopStack = 320;  // FP32 trace: < 64 left over + up to 4 new disks per lane per group
constexpr int kCoopCap = 448;    // FP64 trace: queued disks; solved past kCoopCap - 64 (screened) or - 256.
                                 // CoopWave x 4 waves must stay under 53.8 KB for 3 blocks per CU

struct CoopWave {
    double orb[14][64];              // per lane: cx cy R r2 pa w inv_R inv_w entry(x y) P0x P0y exit(x y); NaN
                                     // entry x: the first cell, NaN exit x: no next cell
    unsigned long long tmin[64];     // per lane: order key of the cell's minimum time
    const double2* ixy[64];          // per lane: its sample's binned field
    float thr[64];                   // per lane: travel-angle threshold of step 2
    float pe[64], pf[64];            // per lane: travel pseudo-angles of the cell's entry and exit (d_first_cross_time_sel)
    unsigned int nexact[64];         // per lane: cooperative exact solves this cell
    int32_t qidx[kCoopCap];          // queued disk: index in its owner's field
    float qlb[kCoopCap];             // queued disk: lower bound of its travel angles
    uint8_t qown[kCoopCap];          // queued disk: owner lane
};

// Unsigned order of the keys = numeric order of the doubles (signs included).
__device__ __forceinline__ unsigned long long d_order_key(double t)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(t);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ double d_from_order_key(unsigned long long k)
{
    return __longlong_as_double((long long)((k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k));
}

// Orders this wave's LDS accesses before and after (lanes of one wave only).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Pseudo-angle in [0, 4) of the direction (x, y), increasing with the angle
// from +x towards +y: 0 at +x, 1 at +y, 2 at -x, 3 at -y.  Its slope in the
// angle lies in [1/2, 1].
__device__ __forceinline__ float pseudo_angle(float x, float y)
{
    const float ax = fabsf(x), ay = fabsf(y);
    const float sum = ax + ay;
    const float r = (sum > 0.0f) ? ay * __builtin_amdgcn_rcpf(sum) : 0.0f;
    // y >= 0: x >= 0 ? r : 2 - r; y < 0: x < 0 ? 2 + r : 4 - r (selects, no branches)
    const bool up = y >= 0.0f;
    const bool plus = up ? (x >= 0.0f) : (x < 0.0f);
    const float base = up ? (plus ? 0.0f : 2.0f) : (plus ? 2.0f : 4.0f);
    return plus ? base + r : base - r;
}

// Screening of one lane's crossing disks in one cell.  Angles are travel
// angles: from the start, in the direction of travel, in [0, 2 pi).  Two facts
// bound where the orbit can meet a disk of radius r centred at P:
//  * the arc from the start S to a point X is no shorter than the chord, and
//    |S - X| >= |S - P| - r for X on the disk, so every meeting point lies at
//    a travel angle >= (|S - P| - r) / R;
//  * a disk whose centre is at distance d from the orbit centre meets the
//    orbit within asin(r / d) <= 1.0473 r / d (r / d <= 1/2) of the centre's
//    direction, and a crossing disk has d > R - r: within dmax = 1.0474 r /
//    (R - r).  When the start lies outside the disk, those points do not wrap
//    round the start, so they lie within dmax of the centre's travel angle
//    th_c; the pseudo-angle p_c <= th_c of the centre (slope <= 1) gives the
//    lower bound p_c - dmax.
// "Near" disks, whose centre is within 1.001 r + 1.001e-6 R of the start, get
// only the first bound.  A disk is dropped when its lower bound reaches the
// travel angle of the minimum time so far plus 1e-5 rad, or when its centre
// lies outside the cell's angle range widened by dmax + 1e-5 rad: the margin
// covers the FP32 bounds (< 1e-6 rad) and the FP64 crossing angles (~3e-8
// rad), so a dropped disk's time, as the reference computes it, could not
// have lowered the minimum or passed AngleInRange.
struct DiskScreen {
    double u0x, u0y;        // start - orbit centre
    float gmin2;            // (1.001 r + 1.001e-6 R)^2
    float r, inv_R, dmax;   // dmax = +inf when r / (R - r) is not in (0, 1/2]
    float lo_pa, hi_pa;     // centre pseudo-angles kept: (lo_pa, hi_pa)
};

// The window of centre pseudo-angles whose crossings can fall in the cell's
// range: its accepted angles lie on the arc travelled from the entry to the
// exit (AngleInRange of entry and exit angles, details.h:119-136; in a first
// cell from a raw start angle phi0 < 0 the accepted set is part of that arc or
// empty).  The travel angles of the two points are taken in FP32 from the
// start (error < 1e-6 rad, inside the window's 1e-5 margin).  No window in a
// cell without a next one (a range end at the raw start angle), when the arc
// wraps or is empty, or when it ends within dmax + 1e-4 of a full turn.
template <bool CW>
__device__ __forceinline__ float d_travel_f32(double u0x, double u0y, double vx, double vy)
{
    const double dot = u0x * vx + u0y * vy;
    const double crs = u0x * vy - u0y * vx;
    float a = atan2f((float)(CW ? -crs : crs), (float)dot);  // (-pi, pi]
    return (a < 0.0f) ? a + 6.2831855f : a;
}

template <bool CW>
__device__ __forceinline__ void d_range_window(const Orb& o, bool first, bool avail, double phi0, double ex, double ey,
                                               double fx, double fy, DiskScreen& ds)
{
    ds.lo_pa = -2.0f;
    ds.hi_pa = 5.0f;
    if (!avail || !(ds.dmax < 1.0f)) return;
    const float tin = first ? 0.0f : d_travel_f32<CW>(o.u0x, o.u0y, ex - o.cx, ey - o.cy);
    const float tout = d_travel_f32<CW>(o.u0x, o.u0y, fx - o.cx, fy - o.cy);
    if (!(tout > tin)) return;
    const float th = tout + ds.dmax + 1e-5f;
    if (!(th < 6.2831855f - 1e-4f)) return;
    float sn, cs;
    sincosf(th, &sn, &cs);
    ds.hi_pa = pseudo_angle(cs, sn) - 1e-6f;
    const float tl = tin - ds.dmax - 1e-5f;
    if (tl > 0.0f) {
        sincosf(tl, &sn, &cs);
        ds.lo_pa = pseudo_angle(cs, sn) + 1e-6f;
    }
}

// A crossing disk (q = orbit centre - disk, the CirclesCross operands): false
// when its centre is outside the range window, else lb = the lower bound on
// its meeting points' travel angles.
template <bool CW>
__device__ __forceinline__ bool d_disk_screen(const DiskScreen& ds, double qx, double qy, float& lb)
{
    const float gx = (float)(ds.u0x + qx), gy = (float)(ds.u0y + qy);  // start - disk
    const float ux = (float)ds.u0x, uy = (float)ds.u0y;
    const float vx = (float)(-qx), vy = (float)(-qy);                   // disk - centre
    const float dot = ux * vx + uy * vy;
    const float crs = ux * vy - uy * vx;
    const float pc = pseudo_angle(dot, CW ? -crs : crs);
    const float g2 = gx * gx + gy * gy;
    const bool near = g2 < ds.gmin2;
    const float ld = (__builtin_amdgcn_sqrtf(g2) - ds.r) * ds.inv_R;  // v_sqrt_f32 (1 ulp; g2 is a normal number)
    lb = near ? ld : fmaxf(ld, pc - ds.dmax);
    return near || (pc > ds.lo_pa && pc < ds.hi_pa);
}

// Travel angle below which a disk's bound must lie for it to be solved, given
// the minimum time t so far.
__device__ __forceinline__ float d_time_threshold(double w, double t)
{
    return (t < ES_INF) ? (float)(w * t) + 1e-5f : 3.0e38f;
}

// The cell's accepted range (AngleInRange of rx = the entry angle, or the start
// angle phi0 in the first cell, and ry = the exit angle, or phi0 without a
// next cell, details.h:119-136) as travel pseudo-angles from the start, for
// d_first_cross_time_sel, from the entry and exit points alone: (e, f), with
// e = 0 in the first cell and f = 4 (a whole turn) without a next cell, or
// (-1, 5) when it accepts every angle (rx = ry = phi0: the first cell has no
// next one).  A raw start angle phi0 < 0 at a range end makes AngleInRange
// accept another set than the arc travelled (m = -phi0): clockwise, the first
// cell takes (0, theta_e) while the exit comes before travel 2 pi - m (the
// bottom of the orbit, velocity angle 0) and nothing after it, a last cell
// (theta_i, 2 pi - m); counter-clockwise, the first cell [m, theta_e) (m: the
// top of the orbit), a last cell (theta_i, 2 pi) past m and nothing before
// it.  NaN (nothing decided) when a point lies beyond the slope bound or
// within the margin of the start, of a whole turn, of the other end or of one
// of those turning points, and in the clockwise last cell entered after
// 2 pi - m (two ranges).
template <bool CW>
__device__ __forceinline__ void cell_range_pseudo(const Orb& o, const Isect& entry, const Isect& exit, bool avail,
                                                  bool first, double phi0, double e_known, double f_known, float& pe,
                                                  float& pf)
{
    if (first && !avail) {
        pe = -1.0f;
        pf = 5.0f;
        return;
    }
    const double lim = kAngleSlopeLimit * o.R, RM = kRangeMargin;
    const double evy = entry.y - o.cy, fvy = exit.y - o.cy;
    // e_known / f_known: the same pseudo-angles when the cell step took them
    // (NaN otherwise)
    const double e = first ? 0.0 : (e_known == e_known) ? e_known : d_travel_pseudo<CW>(o.u0x, o.u0y, entry.x - o.cx, evy);
    const double f = !avail ? 4.0 : (f_known == f_known) ? f_known : d_travel_pseudo<CW>(o.u0x, o.u0y, exit.x - o.cx, fvy);
    const bool e_ok = first || (fabs(evy) < lim && e > RM && e < 4.0 - RM);
    const bool f_ok = !avail || (fabs(fvy) < lim && f > RM && f < 4.0 - RM);
    const bool ok = e_ok && f_ok && f - e > RM;
    pe = __builtin_nanf("");
    pf = (float)f;
    if (phi0 < 0.0 && (first || !avail)) {
        // The turning point's pseudo-angle is measured from the point itself
        // when it lies at least kTurnMin from the start; nearer, the point is
        // within rounding of the start, and only the slope bounds of its
        // travel angle are used (pseudo(2 pi - m) in [4 - m, 4 - m / 2]).
        const double m = -phi0;
        const bool far = m >= kTurnMin;
        if (CW) {
            const double pb = far ? d_travel_pseudo<CW>(o.u0x, o.u0y, 0.0, -o.R) : 4.0;  // travel 2 pi - m
            const double pb_lo = far ? pb : 4.0 - m, pb_hi = far ? pb : 4.0 - 0.5 * m;
            if (first) {
                if (ok && f < pb_lo - RM) pe = 0.0f;
                else if (f_ok && f > pb_hi + RM) pe = pf = 5.0f;
            } else if (far && e_ok && e < pb - RM) {
                pe = (float)e;
                pf = (float)pb;
            }
        } else if (far) {
            const double pt = d_travel_pseudo<CW>(o.u0x, o.u0y, 0.0, o.R);  // travel m
            if (first) {
                if (ok && f > pt + RM) pe = (float)pt;
            } else if (e_ok && e > pt + RM) {
                pe = (float)e;
            } else if (e_ok && e < pt - RM) {
                pe = pf = 5.0f;
            }
        }
    } else if (ok) {
        pe = (float)e;
    }
}

// GetFirstCrossTime of disk gi of owner `own`'s field, folded into the
// owner's minimum as `lifetime = (t < lifetime) ? t : lifetime` would
// (lifetime.h:117-123): NaN and times >= DBL_MAX never lower it, and no solve
// yields -0.
template <bool CW>
__device__ __forceinline__ void d_coop_solve(CoopWave& L, int own, int gi)
{
    Orb ot;
    ot.cx = L.orb[0][own];
    ot.cy = L.orb[1][own];
    ot.R = L.orb[2][own];
    ot.R2 = ot.R * ot.R;  // as d_setup_item forms them
    ot.eps_r = 1e-10 * ot.R;
    ot.r2 = L.orb[3][own];
    ot.pa = L.orb[4][own];
    ot.w = L.orb[5][own];
    ot.inv_R = L.orb[6][own];
    ot.inv_w = L.orb[7][own];
    ot.radd2 = ot.rmin2 = ot.u0x = ot.u0y = 0.0;
    ot.r = 0.0f;
    const double2 p = L.ixy[own][gi];
    const double t = d_first_cross_time_sel<CW>(ot, p.x, p.y, L.orb[8][own], L.orb[9][own], L.orb[12][own],
                                                L.orb[13][own], L.orb[10][own] - ot.cx, L.orb[11][own] - ot.cy,
                                                L.pe[own], L.pf[own]);
    if (t < ES_INF) atomicMin(&L.tmin[own], d_order_key(t));
    atomicAdd(&L.nexact[own], 1u);
}

// The cell's disks through global-memory loads (the binned field is global
// memory; a generic pointer would make the compiler emit flat loads, which
// wait on the LDS counter too).
typedef double DiskXY __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) const DiskXY GlobalDisk;

__device__ __forceinline__ const GlobalDisk* global_disks(const double2* p)
{
    return (const GlobalDisk*)p;
}

__device__ __forceinline__ double2 load_disk(const GlobalDisk* p, int k)
{
    const DiskXY v = p[k];
    return make_double2(v.x, v.y);
}

// Steps 2-4 over the `top` queued disks (see above); wave-uniform.
template <bool CW, bool SCREEN>
__device__ __forceinline__ void coop_flush(CoopWave& L, int lane, unsigned long long below, bool active, int first_i,
                                           float first_lb, int top)
{
    // 2. Each lane solves its lowest-bound disk, unless the minimum so far
    // already rules it out; the threshold that results.  (The orbit is read
    // back from LDS, as for a shared solve.)
    if (SCREEN && active) {
        const double w = L.orb[5][lane];
        float thr = d_time_threshold(w, d_from_order_key(L.tmin[lane]));
        if (first_i >= 0 && !(first_lb >= thr)) {
            d_coop_solve<CW>(L, lane, first_i);
            thr = d_time_threshold(w, d_from_order_key(L.tmin[lane]));
        }
        L.thr[lane] = thr;
    }
    wave_lds_sync();

    // 3. Drop the queued disks whose bound reaches their owner's threshold; compact.
    int nsurv = SCREEN ? 0 : top;  // wave-uniform
    for (int k0 = 0; SCREEN && k0 < top; k0 += 64) {
        const int k = k0 + lane;
        bool keep = false;
        int own = 0, gi = 0;
        if (k < top) {
            own = L.qown[k];
            gi = L.qidx[k];
            keep = !(L.qlb[k] >= L.thr[own]);
        }
        const unsigned long long m = __ballot(keep);
        if (keep) {
            const int pos = nsurv + __popcll(m & below);
            L.qidx[pos] = gi;
            L.qown[pos] = (uint8_t)own;
        }
        nsurv += __popcll(m);
        wave_lds_sync();
    }

    // 4. The rest, 64 at a time.
    for (int k0 = 0; k0 < nsurv; k0 += 64)
        if (k0 + lane < nsurv) d_coop_solve<CW>(L, L.qown[k0 + lane], L.qidx[k0 + lane]);
    wave_lds_sync();
}

template <bool CW, bool SCREEN>
__device__ __forceinline__ void trace_coop_body(const TraceArgs& a)
{
    __shared__ CoopWave sh_coop[kBlock / 64];
    CoopWave& L = sh_coop[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const long long total = a.list_mode ? (long long)*a.list_count : a.total_items;
    const unsigned long long key_inf = d_order_key(ES_INF);
    L.tmin[lane] = key_inf;
    L.nexact[lane] = 0u;

    long long q_next = 0, q_end = 0;
    PairHint hint;
    bool exhausted = false;
    bool active = false;
    long long item = 0, out_idx = -1;
    int cap = 0;
    DevGrid g;
    double bound_time = 0.0;
    TraceState st;

    while (true) {
        long long got;
        d_refill(a, total, active, q_next, q_end, exhausted, got, hint);
        if (got >= 0) {
            item = got;
            // The orbit lives in LDS between cell steps (registers are scarce).
            Orb os;
            double phi0s;
            d_setup_item<CW>(a, hint, item, g, os, phi0s, bound_time, out_idx, cap, st);
            L.orb[0][lane] = os.cx;
            L.orb[1][lane] = os.cy;
            L.orb[2][lane] = os.R;
            L.orb[3][lane] = os.r2;
            L.orb[4][lane] = os.pa;
            L.orb[5][lane] = os.w;
            L.orb[6][lane] = os.inv_R;
            L.orb[7][lane] = os.inv_w;
            L.orb[10][lane] = st.nx.x;  // start point
            L.orb[11][lane] = st.nx.y;
            L.ixy[lane] = g.ixy;
            active = true;
        }
        if (__ballot(active) == 0ull) break;

        // This cell step's view of the orbit, formed as d_setup_item forms it.
        Orb o;
        o.cx = L.orb[0][lane];
        o.cy = L.orb[1][lane];
        o.R = L.orb[2][lane];
        o.r2 = L.orb[3][lane];
        o.pa = L.orb[4][lane];
        o.w = L.orb[5][lane];
        o.inv_R = L.orb[6][lane];
        o.inv_w = L.orb[7][lane];
        o.R2 = o.R * o.R;
        o.eps_r = 1e-10 * o.R;
        {
            const double radd = o.R + g.radius, rmin = o.R - g.radius;
            o.radd2 = radd * radd;
            o.rmin2 = rmin * rmin;
        }
        o.u0x = L.orb[10][lane] - o.cx;
        o.u0y = L.orb[11][lane] - o.cy;
        o.r = (float)g.radius;
        const double phi0 = o.pa;
        DiskScreen ds;
        ds.lo_pa = -2.0f;
        ds.hi_pa = 5.0f;
        if (SCREEN) {
            ds.u0x = o.u0x;
            ds.u0y = o.u0y;
            const float gm = 1.001f * (float)g.radius + 1.001e-6f * (float)o.R;
            ds.gmin2 = gm * gm;
            ds.r = (float)g.radius;
            ds.inv_R = (float)o.inv_R;
            const float q = (float)g.radius / (float)(o.R - g.radius);
            ds.dmax = (q > 0.0f && q <= 0.5f) ? 1.0474f * q : __builtin_inff();
        }

        // This lane's cell: the next cell and angle range (lifetime.h:106-115).
        int s = 0, n = 0;
        bool avail = false;
        if (active) {
            const Isect entry = st.nx;
            const bool first = st.cells == 0;
            double e_known, f_known;
            avail = d_next_cell_lazy<CW>(o, phi0, first, g, entry, st.nx, st.edge_hits, e_known, f_known);
            const double nan = __builtin_nan("");
            L.orb[8][lane] = first ? nan : entry.x;
            L.orb[9][lane] = entry.y;
            L.orb[12][lane] = avail ? st.nx.x : nan;
            L.orb[13][lane] = st.nx.y;
            cell_range_pseudo<CW>(o, entry, st.nx, avail, first, phi0, e_known, f_known, L.pe[lane], L.pf[lane]);
            if (SCREEN) d_range_window<CW>(o, first, avail, phi0, entry.x, entry.y, st.nx.x, st.nx.y, ds);
            const int ci = entry.cy * g.cpr + entry.cx;
            s = (ci > 0) ? g.cell_end[ci - 1] : 0;
            n = g.cell_end[ci] - s;
            st.cells += 1;
            st.tested += n;
        }

        int i = 0;
        if constexpr (SCREEN) {
            // 1a. CirclesCross of the next <= 32 disks of the lane's cell into
            // a bit mask (bit k: disk s + base + k crosses the orbit); 1b. the
            // crossing ones, one per lane per pass, through the screen: disks
            // outside the range window are dropped, the one with the lowest
            // bound stays with its lane and the others are queued with their
            // bounds.  Only 1a runs over every disk; the screen runs over the
            // crossing ones, packed (a lane's next crossing disk, not its next disk).
            const GlobalDisk* gxy = global_disks(g.ixy);
            unsigned int pend = 0u;
            int base = 0;
            bool more = true;
            while (true) {
                int top = 0;  // wave-uniform
                float first_lb = 3.0e38f;
                int first_i = -1;
                while (top <= kCoopCap - 64) {
                    if (__ballot(pend != 0u) == 0ull) {
                        if (__ballot(i < n) == 0ull) {
                            more = false;
                            break;
                        }
                        base = i;
                        constexpr int GA = 4;
                        for (int k0 = 0; k0 < 32 && __ballot(base + k0 < n) != 0ull; k0 += GA) {
                            double2 p[GA];
#pragma unroll
                            for (int u = 0; u < GA; u++)
                                p[u] = (base + k0 + u < n) ? load_disk(gxy, s + base + k0 + u) : make_double2(0.0, 0.0);
#pragma unroll
                            for (int u = 0; u < GA; u++) {
                                const double qx = o.cx - p[u].x, qy = o.cy - p[u].y;
                                const double d2 = qx * qx + qy * qy;
                                const bool h = (base + k0 + u < n) && (d2 < o.radd2) && (d2 > o.rmin2);
                                pend |= h ? (1u << (k0 + u)) : 0u;
                            }
                        }
                        st.cross_calls += __popc(pend);
                        i = base + 32;
                        continue;
                    }
                    bool h = false;
                    int qi = -1;
                    float lb = 0.0f;
                    if (pend) {
                        const int k = __builtin_ctz(pend);
                        pend &= pend - 1u;
                        qi = s + base + k;
                        const double2 p = load_disk(gxy, qi);
                        h = d_disk_screen<CW>(ds, o.cx - p.x, o.cy - p.y, lb);
                    }
                    int qq = qi;
                    float ql = lb;
                    if (h && lb < first_lb) {  // a new lowest bound: the previous one is queued
                        qq = first_i;
                        ql = first_lb;
                        h = first_i >= 0;
                        first_lb = lb;
                        first_i = qi;
                    }
                    const unsigned long long m = __ballot(h);
                    if (h) {
                        const int pos = top + __popcll(m & below);
                        L.qidx[pos] = qq;
                        L.qlb[pos] = ql;
                        L.qown[pos] = (uint8_t)lane;
                    }
                    top += __popcll(m);
                }
                coop_flush<CW, true>(L, lane, below, active, first_i, first_lb, top);
                if (!more) break;
            }
        }
        while (!SCREEN) {
            // 1. CirclesCross, G disks at a time; every crossing disk is queued.
            constexpr int G = 4;
            int top = 0;  // wave-uniform
            bool more;
            while ((more = (__ballot(i < n) != 0ull)) && top <= kCoopCap - 64 * G) {
                double2 p[G];
#pragma unroll
                for (int u = 0; u < G; u++) p[u] = (i + u < n) ? g.ixy[s + i + u] : make_double2(0.0, 0.0);
                int hits = 0;
#pragma unroll
                for (int u = 0; u < G; u++) {
                    const double qx = o.cx - p[u].x, qy = o.cy - p[u].y;
                    const double d2 = qx * qx + qy * qy;
                    const bool h = (i + u < n) && (d2 < o.radd2) && (d2 > o.rmin2);
                    hits |= h ? (1 << u) : 0;
                }
                st.cross_calls += __popc(hits);
#pragma unroll
                for (int u = 0; u < G; u++) {
                    const bool h = (hits >> u) & 1;
                    const unsigned long long m = __ballot(h);
                    if (h) {
                        const int pos = top + __popcll(m & below);
                        L.qidx[pos] = s + i + u;
                        L.qown[pos] = (uint8_t)lane;
                    }
                    top += __popcll(m);
                }
                i += G;
            }
            coop_flush<CW, false>(L, lane, below, active, -1, 0.0f, top);
            if (!more) break;
        }

        if (active) {
            st.lifetime = d_from_order_key(L.tmin[lane]);
            st.cross_exact += (int)L.nexact[lane];
            L.tmin[lane] = key_inf;
            L.nexact[lane] = 0u;
            int flags = -1;
            if (st.lifetime < ES_INF) {
                flags = 0;
            } else if (!avail) {
                flags = ES_F_ESCAPED;
                st.lifetime = 0.0;
            } else if (st.cells >= cap) {
                flags = ES_F_CAP;
                st.lifetime = 0.0;
            }
            if (flags >= 0) {
                d_store_item(a, item, out_idx, bound_time, st, flags);
                active = false;
            }
        }
    }
}

// CreateParticle for the FP32 trace: the same start data (position, phi,
// orbit centre, bound time in FP64) in the start point's frame.
template <bool CW>
__device__ __forceinline__ void t32_setup_item(const TraceArgs& a, const PairHint& hint, long long item, DevGrid& g,
                                               Orb32& o, double& w, double& bound_time, long long& out_idx, int& cap,
                                               Trace32& st)
{
    const int pi = hinted_pair(a, hint, item);
    const PairDev& pr = a.pairs[pi];
    const SampleDev& sd = a.samples[pr.sample];
    const ItemPos ip = item_pos(a, pr, sd, item);
    const int q = ip.q, p = ip.p, pidx = ip.pidx;
    const int j = (int)((unsigned int)pidx / (unsigned int)sd.ppr), i = pidx - j * sd.ppr;
    out_idx = ip.out_idx;
    g.ixy = sd.ixy;
    g.cell_end = sd.cell_end;
    g.start = sd.start;
    g.size = sd.size;
    g.L = sd.L;
    g.inv_L = sd.inv_L;
    g.radius = sd.radius;
    g.cpr = sd.cpr;
    cap = 4 * sd.cpr * sd.cpr + 16;
    const double px = (double)i * sd.dist + sd.offx;  // SimulationCPU.cpp:114
    const double py = (double)j * sd.dist + sd.offy;
    w = pr.w;
    const double phi = sd.phi_start + q * ES_HALF_PI + p * sd.phi_step;  // lifetime.h:36-58
    const double2 sc = sd.phi_sc[q * sd.ppq + p];
    const double sphi = sc.x, cphi = sc.y;
    const double vx = cphi * sd.speed, vy = sphi * sd.speed;
    const double R = pr.R;
    const double sx = esm_div_rcp(vy * R, sd.speed, sd.inv_speed);
    const double sy = esm_div_rcp(-vx * R, sd.speed, sd.inv_speed);
    const double ocx = CW ? (px - sx) : (px + sx);
    const double ocy = CW ? (py - sy) : (py + sy);
    o.sx = px;
    o.sy = py;
    o.inv_r = pr.inv_R;
    o.nxd = (ocx - px) * o.inv_r;
    o.nyd = (ocy - py) * o.inv_r;
    o.nx = (float)o.nxd;
    o.ny = (float)o.nyd;
    o.r2 = g.radius * g.radius;
    // The reference's travel angle (GetCrossAngle, details.h:111-116)
    // runs along the orbit against (cos phi, sin phi), in both senses.
    o.txd = -cphi;
    o.tyd = -sphi;
    o.tx = (float)o.txd;
    o.ty = (float)o.tyd;
    o.rr = (float)(g.radius * o.inv_r);
    o.rr2 = o.rr * o.rr;
    o.ccw = !CW;
    // A negative raw angle changes the reference's ranges (es_trace32.h).
    o.quirk = phi < 0.0;
    const float hm = tanf((float)(-0.5 * phi));
    o.km = t32_key(Root32{hm, false});
    o.km2 = t32_key(Root32{-hm, false});
    bound_time = ES_INF;  // GetBoundTime (details.h:219-228)
    if (!sd.coherent) {
        const double a0 = phi + sd.alpha;
        const double remaining = a0 - ES_HALF_PI * floor(esm_div_rcp(a0, ES_HALF_PI, ES_INV_HALF_PI));
        const double dphi = CW ? remaining : (2.0 * sd.alpha - remaining);
        bound_time = esm_div_rcp(dphi, w, pr.inv_w);
    }
    const int c0x = (int)(esm_div_rcp(px - g.start, g.size, sd.inv_size) * (double)(g.cpr));
    const int c0y = (int)(esm_div_rcp(py - g.start, g.size, sd.inv_size) * (double)(g.cpr));
    t32_begin(st, c0x, c0y);
}

// Lifetime of an ended FP32 trace; items FP32 cannot decide go on the list
// for the FP64 trace.
__device__ __forceinline__ void t32_store_item(const TraceArgs& a, long long item, long long out_idx, double w,
                                               double bound_time, const Trace32& st, int flags, double theta)
{
    double lt = 0.0;
    if (flags == 0) {
        lt = theta / w;
        if (bound_time < ES_INF && fabs(lt - bound_time) <= 1e-5 * bound_time) {
            flags = ES_F_RECOMPUTE;  // FP32 cannot tell which side of the bound
        } else if (lt > bound_time) {  // lifetime.h:142-146
            lt = bound_time;
            flags = ES_F_AT_BOUND;
        }
    }
    if (flags & ES_F_RECOMPUTE) {
        const int idx = atomicAdd(a.list_count, 1);
        a.list[idx] = (int32_t)item;
    }
    a.lt_item[item] = lt;
    a.m_item[item] = make_int4(st.cells, st.tested, st.edge_hits, st.cross_calls);
    a.x_item[item] = 0;
    a.f_item[item] = (uint8_t)flags;
    if (out_idx >= 0) store_lifetime(a, item, out_idx, lt);
}

// The FP32 trace (es_trace32.h) over the same work items, one lane per
// particle: same lane refill, same start data; items it cannot decide are
// appended to a.list for the FP64 trace.
template <bool CW>
__device__ __forceinline__ void trace32_body(const TraceArgs& a)
{
    const long long total = a.total_items;
    long long q_next = 0, q_end = 0;
    PairHint hint;
    bool exhausted = false;
    bool active = false;
    long long item = 0, out_idx = -1;
    int cap = 0;
    DevGrid g;
    Orb32 o;
    double w = 1.0, bound_time = 0.0;
    Trace32 st;

    while (true) {
        long long got;
        d_refill(a, total, active, q_next, q_end, exhausted, got, hint);
        if (got >= 0) {
            item = got;
            t32_setup_item<CW>(a, hint, item, g, o, w, bound_time, out_idx, cap, st);
            active = true;
        }
        if (__ballot(active) == 0ull) break;

        if (active) {
            int flags = 0;
            double theta = 0.0;
            if (t32_step(o, g, cap, st, flags, theta)) {
                t32_store_item(a, item, out_idx, w, bound_time, st, flags, theta);
                active = false;
            }
        }
    }
}

// ------------------------------------------------ wave-cooperative FP32 trace
//
// As trace_coop_body for the FP32 trace: each lane screens its cell's disks
// in FP32 (t32_screen); the disks that pass go on the wave's stack and their
// FP64-formed test (t32_disk) runs 64 at a time for any lane.  Per owner the
// LDS keeps: the smallest (key, disk position * 2 + root) of the roots in
// range — the order the one-lane loop keeps on equal keys — the smallest key
// within error of a range end, an "undecidable" bit and the crossing count.
// The owner then solves its winning disk again for the root itself.  A
// crossing near a range end is flagged when it comes within error of the
// cell's first hit (the one-lane loop compares with the first hit so far, a
// superset of the same cases).

struct Coop32Wave {
    double dctx[8][64];              // per lane: sx sy inv_r nxd nyd txd tyd r2
    float fctx[10][64];              // per lane: glo ghi rr k_in band_in b_hi a1 b1 a2 b2
    const double2* ixy[64];          // per lane: its sample's binned field
    unsigned long long best[64];     // per lane: (key bits << 32) | (position * 2 + root), ~0 when none
    unsigned int atend[64];          // per lane: key bits of the first root near a range end
    unsigned int bad[64];            // per lane: undecidable disk seen
    unsigned int cross[64];          // per lane: crossing disks (CirclesCross true)
    int32_t qidx[kCoopStack];        // queued disk: position in its owner's field
    uint8_t qown[kCoopStack];        // queued disk: owner lane
};

__device__ __forceinline__ Disk32Ctx t32_ctx_lds(const Coop32Wave& L, int own)
{
    Disk32Ctx d;
    d.sx = L.dctx[0][own];
    d.sy = L.dctx[1][own];
    d.inv_r = L.dctx[2][own];
    d.nxd = L.dctx[3][own];
    d.nyd = L.dctx[4][own];
    d.txd = L.dctx[5][own];
    d.tyd = L.dctx[6][own];
    d.r2 = L.dctx[7][own];
    d.glo = L.fctx[0][own];
    d.ghi = L.fctx[1][own];
    d.rr = L.fctx[2][own];
    d.k_in = L.fctx[3][own];
    d.band_in = L.fctx[4][own];
    d.b_hi = L.fctx[5][own];
    d.a1 = L.fctx[6][own];
    d.b1 = L.fctx[7][own];
    d.a2 = L.fctx[8][own];
    d.b2 = L.fctx[9][own];
    return d;
}

__device__ __forceinline__ void t32_coop_solve(Coop32Wave& L, int k)
{
    const int own = L.qown[k];
    const int gi = L.qidx[k];
    const Disk32Ctx d = t32_ctx_lds(L, own);
    const double2 p = L.ixy[own][gi];
    bool cross;
    Root32 rt[2];
    float kk[2];
    bool in[2], at_end[2];
    const bool ok = t32_disk(d, p.x - d.sx, p.y - d.sy, cross, rt, kk, in, at_end);
    if (cross) atomicAdd(&L.cross[own], 1u);
    if (!ok) {
        atomicOr(&L.bad[own], 1u);
    } else if (cross) {
#pragma unroll
        for (int r = 0; r < 2; r++) {
            if (in[r])
                atomicMin(&L.best[own],
                          ((unsigned long long)__float_as_uint(kk[r]) << 32) | (unsigned long long)(2 * gi + r));
            if (at_end[r]) atomicMin(&L.atend[own], __float_as_uint(kk[r]));
        }
    }
}

template <bool CW>
__device__ __forceinline__ void trace32_coop_body(const TraceArgs& a)
{
    __shared__ Coop32Wave sh_coop32[kBlock / 64];
    Coop32Wave& L = sh_coop32[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const long long total = a.total_items;
    L.best[lane] = ~0ull;
    L.atend[lane] = 0xFFFFFFFFu;
    L.bad[lane] = 0u;
    L.cross[lane] = 0u;

    long long q_next = 0, q_end = 0;
    PairHint hint;
    bool exhausted = false;
    bool active = false;
    long long item = 0, out_idx = -1;
    int cap = 0;
    DevGrid g;
    Orb32 o;
    double w = 1.0, bound_time = 0.0;
    Trace32 st;

    while (true) {
        long long got;
        d_refill(a, total, active, q_next, q_end, exhausted, got, hint);
        if (got >= 0) {
            item = got;
            t32_setup_item<CW>(a, hint, item, g, o, w, bound_time, out_idx, cap, st);
            L.ixy[lane] = g.ixy;
            active = true;
        }
        if (__ballot(active) == 0ull) break;

        // This lane's cell exit and ranges; a cell FP32 cannot decide ends the trace.
        Cell32 c;
        int s = 0, n = 0;
        bool go = false;
        if (active) {
            if (t32_exit(o, g, st, c)) {
                const Disk32Ctx d = t32_ctx(o, c);
                L.dctx[0][lane] = d.sx;
                L.dctx[1][lane] = d.sy;
                L.dctx[2][lane] = d.inv_r;
                L.dctx[3][lane] = d.nxd;
                L.dctx[4][lane] = d.nyd;
                L.dctx[5][lane] = d.txd;
                L.dctx[6][lane] = d.tyd;
                L.dctx[7][lane] = d.r2;
                L.fctx[0][lane] = d.glo;
                L.fctx[1][lane] = d.ghi;
                L.fctx[2][lane] = d.rr;
                L.fctx[3][lane] = d.k_in;
                L.fctx[4][lane] = d.band_in;
                L.fctx[5][lane] = d.b_hi;
                L.fctx[6][lane] = d.a1;
                L.fctx[7][lane] = d.b1;
                L.fctx[8][lane] = d.a2;
                L.fctx[9][lane] = d.b2;
                const int ci = st.cy * g.cpr + st.cx;
                s = (ci > 0) ? g.cell_end[ci - 1] : 0;
                n = g.cell_end[ci] - s;
                st.cells += 1;
                st.tested += n;
                go = true;
            } else {
                t32_store_item(a, item, out_idx, w, bound_time, st, ES_F_RECOMPUTE, 0.0);
                active = false;
            }
        }
        const float glo = -2.0f * o.rr + o.rr2, ghi = 2.0f * o.rr + o.rr2;
        wave_lds_sync();

        int top = 0;  // wave-uniform
        for (int i = 0;; i += 4) {
            const bool more = __ballot(i < n) != 0ull;
            if (more) {
                double2 p[4];
#pragma unroll
                for (int u = 0; u < 4; u++) p[u] = (i + u < n) ? g.ixy[s + i + u] : make_double2(0.0, 0.0);
                int hits = 0;
#pragma unroll
                for (int u = 0; u < 4; u++)
                    hits |= ((i + u < n) && t32_screen(o, glo, ghi, p[u].x - o.sx, p[u].y - o.sy)) ? (1 << u) : 0;
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const bool h = (hits >> u) & 1;
                    const unsigned long long m = __ballot(h);
                    if (h) {
                        const int pos = top + __popcll(m & below);
                        L.qidx[pos] = s + i + u;
                        L.qown[pos] = (uint8_t)lane;
                    }
                    top += __popcll(m);
                }
                wave_lds_sync();
            }
            while (top >= 64 || (!more && top > 0)) {
                const int m = (top < 64) ? top : 64;
                if (lane < m) t32_coop_solve(L, top - m + lane);
                top -= m;
                wave_lds_sync();
            }
            if (!more) break;
        }

        if (go) {
            st.cross_calls += (int)L.cross[lane];
            bool bad = L.bad[lane] != 0u;
            const unsigned long long b = L.best[lane];
            const unsigned int ae = L.atend[lane];
            L.best[lane] = ~0ull;
            L.atend[lane] = 0xFFFFFFFFu;
            L.bad[lane] = 0u;
            L.cross[lane] = 0u;
            float best_k = 2.0f;
            Root32 best = st.best;
            if (b != ~0ull) {
                best_k = __uint_as_float((unsigned int)(b >> 32));
                const int gi = (int)((unsigned int)b >> 1);
                const int r = (int)(b & 1ull);
                const Disk32Ctx d = t32_ctx_lds(L, lane);
                const double2 p = g.ixy[gi];
                bool cross;
                Root32 rt[2];
                float kk[2];
                bool in[2], at_end[2];
                (void)t32_disk(d, p.x - d.sx, p.y - d.sy, cross, rt, kk, in, at_end);
                best = r ? rt[1] : rt[0];
            }
            if (ae != 0xFFFFFFFFu) {
                const float ak = __uint_as_float(ae);
                if (ak < best_k + (2e-6f * ak + T32_END)) bad = true;
            }
            int flags = 0;
            double theta = 0.0;
            if (t32_finish(o, cap, c, st, bad, best_k, best, flags, theta)) {
                t32_store_item(a, item, out_idx, w, bound_time, st, flags, theta);
                active = false;
            }
        }
    }
}

// One kernel per rotation sense and register budget (the tuner picks one).
template <bool CW>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_trace_w3(TraceArgs a)
{
    trace_body<CW>(a);
}

template <bool CW>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2))) void k_trace_w2(TraceArgs a)
{
    trace_body<CW>(a);
}

template <bool CW, bool SCREEN>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_trace_c3(TraceArgs a)
{
    trace_coop_body<CW, SCREEN>(a);
}

template <bool CW, bool SCREEN>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2))) void k_trace_c2(TraceArgs a)
{
    trace_coop_body<CW, SCREEN>(a);
}

template <bool CW>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_trace32(TraceArgs a)
{
    trace32_body<CW>(a);
}

template <bool CW>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_trace32_c(TraceArgs a)
{
    trace32_coop_body<CW>(a);
}

// --------------------------------------------------------------------- sigma

struct SigmaMetrics {
    int64_t cells = 0, tested = 0, max_cells = 0, max_tested = 0, edge = 0, cross = 0, sigma = 0;
    int64_t escaped = 0, at_bound = 0, cap = 0, items = 0, exact = 0, recomp = 0;
};

// One unit of the temperature stage (see k_sigma): DUP for a zero-width phi
// window, whose unit is a (position, quadrant) standing for ppq lifetimes.
template <int NT, bool DUP>
__device__ __forceinline__ void sigma_unit(const SigmaArgs& a, const PairDev& pr, const SampleDev& sd, long long u,
                                           double ws, DD* acc, SigmaMetrics& ms)
{
    const int ppq = sd.ppq, ppp = 4 * ppq, nT = sd.nT;
    const int reps = DUP ? ppq : 1;
    long long slot;
    int q, p0;
    if (DUP) {
        slot = u >> 2;
        q = (int)(u & 3);
        p0 = 0;
    } else {
        slot = (pr.nitems <= 0x7fffffffll) ? (long long)((unsigned int)u / (unsigned int)ppp) : u / ppp;
        const int v = (int)(u - slot * ppp);
        q = (int)((unsigned int)v / (unsigned int)ppq);
        p0 = v - q * ppq;
    }
    const long long gi = pr.item_off + u;
    const double lt_raw = a.lt_item[gi];
    const int pidx = sd.free_pos[slot];
    const int j = (int)((unsigned int)pidx / (unsigned int)sd.ppr), i = pidx - j * sd.ppr;
    if (a.with_metrics) {
        const int4 m = a.m_item[gi];
        const int f = a.f_item[gi];
        ms.cells += (int64_t)reps * m.x;
        ms.tested += (int64_t)reps * m.y;
        ms.edge += (int64_t)reps * m.z;
        ms.cross += (int64_t)reps * m.w;
        ms.exact += (int64_t)reps * a.x_item[gi];
        ms.max_cells = m.x > ms.max_cells ? m.x : ms.max_cells;
        ms.max_tested = m.y > ms.max_tested ? m.y : ms.max_tested;
        ms.escaped += (f & ES_F_ESCAPED) ? reps : 0;
        ms.at_bound += (f & ES_F_AT_BOUND) ? reps : 0;
        ms.cap += (f & ES_F_CAP) ? reps : 0;
        ms.recomp += (f & ES_F_RECOMPUTE) ? reps : 0;
        ms.items += reps;
    }
    if (DUP) {
        dd_add_exact_product(acc[2 * NT], lt_raw, (double)reps);
    } else {
        dd_add_d(acc[2 * NT], lt_raw);
    }

    // ApplySigma + IntegrateResult terms (SimulationCPU.cpp:178-222): a zero
    // lifetime contributes exactly +-0 (util.h:34-46), so it is skipped.
    if (lt_raw == 0.0) return;  // every temperature's term is +-0
    const double phi = sd.phi_start + q * ES_HALF_PI + p0 * sd.phi_step;
    const double2 sc = sd.phi_sc[q * ppq + p0];
    const double sphi = sc.x, cphi = sc.y;
    const double wji = simpson_weight(j, sd.ppr) * simpson_weight(i, sd.ppr);
    // sin, cos of the end angle for the untruncated lifetime, shared by every
    // temperature whose cap does not cut it (the same argument, so the same bits).
    // The sigma terms take the device library's sin, cos and exp (errors of about
    // one ulp), not the correctly rounded ones the trace needs: no branch depends
    // on them, each term only feeds a compensated sum, and the correctly rounded
    // functions make this stage half again as slow for a change far below the
    // sums' 1e-12 agreement with the compensated oracle.  (The per-position
    // images keep the correctly rounded functions: they are compared bit for bit.)
    double sa0, ca0;
    sincos(phi + ws * lt_raw, &sa0, &ca0);
#pragma unroll
    for (int t = 0; t < NT; t++) {
        if (t < nT) {
            const double ml = sd.max_lt[t];
            const double lt = (lt_raw < ml) ? lt_raw : ml;  // Simulation.h:78-86
            if (lt != 0.0) {
                const double tau = sd.tau[t];
                const double z = exp(esm_div_rcp(-lt, tau, sd.inv_tau[t]));
                double sa = sa0, ca = ca0;
                if (lt != lt_raw) sincos(phi + ws * lt, &sa, &ca);
                double r = cphi - ca * z;
                r += ws * tau * sa * z;
                r -= ws * tau * sphi;
                const double rc = r * cphi, rs = r * sphi;
                if (DUP) {
                    // sum over the ppq phi points of rc * wji * w_p: every product is
                    // rc times a power of two, so the sum is exactly rc * (wji * W)
                    const double wsum = wji * simpson_weight_sum(ppq);
                    dd_add_exact_product(acc[2 * t], rc, wsum);
                    dd_add_exact_product(acc[2 * t + 1], rs, wsum);
                } else {
                    const double wgt = wji * simpson_weight(p0, ppq);
                    dd_add_d(acc[2 * t], rc * wgt);
                    dd_add_d(acc[2 * t + 1], rs * wgt);
                }
                ms.sigma += reps;
            }
        }
    }
    }

template <int NT>
__global__ __launch_bounds__(kBlock) void k_sigma(SigmaArgs a)
{
    constexpr int NQ = 2 * NT + 1;
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // ready for the next run's trace
        *a.queue = 0ull;
        if (a.list_count) *a.list_count = 0;
    }
    const int pi = pair_of_block(a.pairs, a.npairs, blockIdx.x);
    const PairDev& pr = a.pairs[pi];
    const SampleDev& sd = a.samples[pr.sample];
    const int ppq = sd.ppq;
    const long long nall = pr.nall;
    const int tpq = sd.tpq;
    const double w = pr.w;
    const double ws = a.clockwise ? -w : w;  // signed_angular_speed (SimulationCPU.cpp:108)

    DD acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) acc[q].s = acc[q].c = 0.0;
    SigmaMetrics ms;

    // A unit is one lifetime, or with tpq = 1 (a zero-width phi window) one
    // (position, quadrant) whose ppq lifetimes are the same particle: its
    // exp / sincos terms are formed once and added ppq times with each phi
    // point's Simpson weight (products by 1, 2, 4: the terms of tracing them
    // one by one).  Either way the traced item of unit u is item_off + u.
    const long long stride = (long long)pr.nblk * kBlock;
    const long long u0 = (long long)(blockIdx.x - pr.blk_off) * kBlock + threadIdx.x;
    if (tpq == 1) {
        for (long long u = u0; u < nall / ppq; u += stride)
            sigma_unit<NT, true>(a, pr, sd, u, ws, acc, ms);
    } else {
        for (long long u = u0; u < nall; u += stride)
            sigma_unit<NT, false>(a, pr, sd, u, ws, acc, ms);
    }
    int64_t m_cells = ms.cells, m_tested = ms.tested, m_max_cells = ms.max_cells, m_max_tested = ms.max_tested;
    int64_t m_edge = ms.edge, m_cross = ms.cross, m_sigma = ms.sigma, m_escaped = ms.escaped;
    int64_t m_at_bound = ms.at_bound, m_cap = ms.cap, m_items = ms.items, m_exact = ms.exact, m_recomp = ms.recomp;

    // Block reduction in a fixed order: wave shuffles, then the 4 waves in LDS.
    __shared__ double sh_s[NQ][kBlock / 64];
    __shared__ double sh_c[NQ][kBlock / 64];
    __shared__ int64_t sh_m[kMetricSlots][kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const DD r = dd_wave_reduce(acc[q]);
        if (lane == 0) {
            sh_s[q][wave] = r.s;
            sh_c[q][wave] = r.c;
        }
    }
    int64_t mv[kMetricSlots] = {m_cells, m_tested, m_escaped, m_at_bound, m_cap, m_max_tested,
                                m_max_cells, m_items, m_edge, m_cross, m_sigma, m_exact, m_recomp};
#pragma unroll
    for (int s = 0; s < kMetricSlots; s++) {
        const int64_t r = (s == 5 || s == 6) ? i64_wave_max(mv[s]) : i64_wave_sum(mv[s]);
        if (lane == 0) sh_m[s][wave] = r;
    }
    __syncthreads();
    const long long base = blockIdx.x;
    if (threadIdx.x < NQ) {
        const int q = threadIdx.x;
        DD r;
        r.s = sh_s[q][0];
        r.c = sh_c[q][0];
        for (int wv = 1; wv < kBlock / 64; wv++) {
            DD o;
            o.s = sh_s[q][wv];
            o.c = sh_c[q][wv];
            r = dd_add(r, o);
        }
        a.dpart[(base * NQ + q) * 2] = r.s;
        a.dpart[(base * NQ + q) * 2 + 1] = r.c;
    } else if (threadIdx.x >= 64 && threadIdx.x < 64 + kMetricSlots) {
        const int s = threadIdx.x - 64;
        int64_t r = sh_m[s][0];
        for (int wv = 1; wv < kBlock / 64; wv++) {
            const int64_t o = sh_m[s][wv];
            r = (s == 5 || s == 6) ? (o > r ? o : r) : r + o;
        }
        a.mpart[base * kMetricSlots + s] = r;
    }
}

__global__ void k_reset(unsigned long long* q) { *q = 0ull; }

// ------------------------------------------------------------------ images

// One thread per free position: IntegratePosition (SimulationCPU.cpp:224-243)
// of the truncated lifetimes (GetTruncatedLifetimes, Simulation.h:78-86) and
// of ApplySigma's sigma_xx / sigma_xy terms (SimulationCPU.cpp:178-204), each
// summed serially over q, p in the reference's order and scaled by
// phi_integrand_factor.  A zero lifetime gives GetSigma = +-0 exactly, so it
// is skipped as in k_sigma.
__global__ __launch_bounds__(256) void k_images(const SampleDev* sample, const PairDev* pair, int clockwise,
                                                const double* lt_item, const double* tau, const double* max_lt,
                                                int nT, int nfree, int ppr, double* img)
{
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= nfree) return;
    const SampleDev& sd = *sample;
    const PairDev& pr = *pair;
    const int ppq = sd.ppq;
    const int pidx = sd.free_pos[slot];
    const double ws = clockwise ? -pr.w : pr.w;
    const int tpq = sd.tpq;
    const long long base = pr.item_off + (long long)slot * 4 * tpq;
    const size_t plane = (size_t)ppr * ppr;
    for (int t = 0; t < nT; t++) {
        const double ml = max_lt[t], tt = tau[t];
        double lt_tot = 0.0, xx_tot = 0.0, xy_tot = 0.0;
        for (int q = 0; q < 4; q++) {
            for (int p = 0; p < ppq; p++) {
                const double lt_raw = lt_item[base + q * tpq + ((tpq == 1) ? 0 : p)];
                const double lt = (lt_raw < ml) ? lt_raw : ml;
                const double w = simpson_weight(p, ppq);
                lt_tot += lt * w;
                if (lt != 0.0) {
                    const double phi = sd.phi_start + q * ES_HALF_PI + p * sd.phi_step;
                    const double2 sc = sd.phi_sc[q * ppq + p];
                    const double sphi = sc.x, cphi = sc.y;
                    double sa, ca;
                    const double z = esm_exp(-lt / tt);
                    esm_sincos(phi + ws * lt, &sa, &ca);
                    double r = cphi - ca * z;
                    r += ws * tt * sa * z;
                    r -= ws * tt * sphi;
                    xx_tot += (r * cphi) * w;
                    xy_tot += (r * sphi) * w;
                }
            }
        }
        double* o = img + (size_t)t * 3 * plane;
        o[pidx] = lt_tot * sd.phi_integrand_factor;
        o[plane + pidx] = xx_tot * sd.phi_integrand_factor;
        o[2 * plane + pidx] = xy_tot * sd.phi_integrand_factor;
    }
}


// ------------------------------------------------------------------ finalize

__global__ __launch_bounds__(256) void k_finalize(const SampleDev* samples, const PairDev* pairs, int NT,
                                                  int with_metrics, const double* dpart, const int64_t* mpart,
                                                  const double* factor, double* out_sig, double* out_lt_sum,
                                                  int64_t* out_m)
{
    const PairDev& pr = pairs[blockIdx.x];
    const int nT = samples[pr.sample].nT;
    const int NQ = 2 * NT + 1;
    const long long o = pr.out;
    __shared__ double sh_s[4], sh_c[4];
    __shared__ int64_t sh_m[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int q = 0; q < NQ; q++) {
        const bool used = (q < 2 * nT) || (q == 2 * NT);
        if (!used) continue;  // uniform across the block
        DD acc;
        acc.s = acc.c = 0.0;
        for (int blk = threadIdx.x; blk < pr.nblk; blk += 256) {
            const long long b = (long long)pr.blk_off + blk;
            DD v;
            v.s = dpart[(b * NQ + q) * 2];
            v.c = dpart[(b * NQ + q) * 2 + 1];
            acc = dd_add(acc, v);
        }
        acc = dd_wave_reduce(acc);
        if (lane == 0) {
            sh_s[wave] = acc.s;
            sh_c[wave] = acc.c;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            DD r;
            r.s = sh_s[0];
            r.c = sh_c[0];
            for (int wv = 1; wv < 4; wv++) {
                DD v;
                v.s = sh_s[wv];
                v.c = sh_c[wv];
                r = dd_add(r, v);
            }
            const double total = r.s + r.c;
            if (q < 2 * NT) {
                const int t = q >> 1;
                out_sig[(o * kMaxT + t) * 2 + (q & 1)] = total * factor[o * kMaxT + t];
            } else {
                out_lt_sum[o] = total;
            }
        }
        __syncthreads();
    }
    if (!with_metrics) return;
    for (int s = 0; s < kMetricSlots; s++) {
        const bool is_max = (s == 5 || s == 6);
        int64_t acc = 0;
        for (int blk = threadIdx.x; blk < pr.nblk; blk += 256) {
            const int64_t v = mpart[((long long)pr.blk_off + blk) * kMetricSlots + s];
            acc = is_max ? (v > acc ? v : acc) : acc + v;
        }
        acc = is_max ? i64_wave_max(acc) : i64_wave_sum(acc);
        if (lane == 0) sh_m[wave] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            int64_t r = sh_m[0];
            for (int wv = 1; wv < 4; wv++) r = is_max ? (sh_m[wv] > r ? sh_m[wv] : r) : r + sh_m[wv];
            out_m[o * kMetricSlots + s] = r;
        }
        __syncthreads();
    }
}

inline unsigned int blocks_for(int64_t n, int bs) { return (unsigned int)((n + bs - 1) / bs); }

}  // namespace

hipError_t launch_bin_jobs(const BinJob* jobs_dev, const BinJob* jobs, int njobs, hipStream_t st)
{
    if (njobs <= 0) return hipSuccess;
    int64_t max_n = 0, max_cell = 0, max_small = 0;
    for (int j = 0; j < njobs; j++) {
        const int64_t nc = (int64_t)jobs[j].cpr * jobs[j].cpr;
        max_n = std::max<int64_t>(max_n, jobs[j].n);
        max_cell = std::max<int64_t>(max_cell, nc);
        if (nc <= kScanOneCells) max_small = std::max<int64_t>(max_small, nc);
    }
    hipError_t e;
    if (max_n > 0) {
        hipLaunchKernelGGL(k_bin_count, dim3(blocks_for(max_n, 256), njobs), dim3(256), 0, st, jobs_dev);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (max_small > 0) {
        static bool lds_set = false;  // allow the LDS staging beyond the 64 KB default (once per process)
        if (!lds_set) {
            if ((e = hipFuncSetAttribute((const void*)k_scan_one, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         kScanOneCells * (int)sizeof(int32_t))) != hipSuccess)
                return e;
            lds_set = true;
        }
        hipLaunchKernelGGL(k_scan_one, dim3(njobs), dim3(1024), (size_t)max_small * sizeof(int32_t), st, jobs_dev);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    for (int j = 0; j < njobs; j++) {
        const int64_t nc = (int64_t)jobs[j].cpr * jobs[j].cpr;
        if (nc <= kScanOneCells) continue;
        const unsigned int nb = blocks_for(nc, 256);
        hipLaunchKernelGGL(k_scan_block, dim3(nb), dim3(256), 0, st, jobs_dev + j);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(256), 0, st, jobs_dev + j);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (max_n > 0) {
        hipLaunchKernelGGL(k_bin_fill, dim3(blocks_for(max_n, 256), njobs), dim3(256), 0, st, jobs_dev);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_bin_sort_gather, dim3(blocks_for(max_cell, 4), njobs), dim3(256), 0, st, jobs_dev);
    return hipGetLastError();
}

hipError_t launch_positions_jobs(const PosJob* jobs_dev, const PosJob* jobs, int njobs, hipStream_t st)
{
    if (njobs <= 0) return hipSuccess;
    int64_t max_n = 0;
    for (int j = 0; j < njobs; j++) max_n = std::max<int64_t>(max_n, (int64_t)jobs[j].ppr * jobs[j].ppr);
    const unsigned int nb = blocks_for(max_n, 256);
    hipLaunchKernelGGL(k_positions, dim3(nb, njobs), dim3(256), 0, st, jobs_dev);
    hipError_t e;
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_compact, dim3(nb, njobs), dim3(256), 0, st, jobs_dev);
    return hipGetLastError();
}

// The trace kernel of a launch shape: rotation sense x register budget x
// crossing solves per lane (coop 0), wave-cooperative (1) or wave-cooperative
// and screened (2).
static const void* trace_kernel(int clockwise, int waves_per_eu, int coop)
{
    if (coop == 2) {
        if (waves_per_eu == 2)
            return clockwise ? (const void*)k_trace_c2<true, true> : (const void*)k_trace_c2<false, true>;
        return clockwise ? (const void*)k_trace_c3<true, true> : (const void*)k_trace_c3<false, true>;
    }
    if (coop) {
        if (waves_per_eu == 2)
            return clockwise ? (const void*)k_trace_c2<true, false> : (const void*)k_trace_c2<false, false>;
        return clockwise ? (const void*)k_trace_c3<true, false> : (const void*)k_trace_c3<false, false>;
    }
    if (waves_per_eu == 2) return clockwise ? (const void*)k_trace_w2<true> : (const void*)k_trace_w2<false>;
    return clockwise ? (const void*)k_trace_w3<true> : (const void*)k_trace_w3<false>;
}

hipError_t trace_blocks_per_cu(int clockwise, int waves_per_eu, int coop, int* blocks)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, trace_kernel(clockwise, waves_per_eu, coop), kBlock,
                                                        0);
}

hipError_t launch_trace(const TraceArgs& a, int clockwise, int waves_per_eu, int coop, int nblocks, hipStream_t st)
{
    if (a.npairs <= 0 || nblocks <= 0 || a.chunk <= 0) return hipErrorInvalidValue;
    if (a.list_mode) {
        hipLaunchKernelGGL(k_reset, dim3(1), dim3(1), 0, st, a.queue);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    } else if (a.total_items <= 0) {
        return hipSuccess;
    }
    TraceArgs arg = a;
    void* params[] = {&arg};
    return hipLaunchKernel(trace_kernel(clockwise, waves_per_eu, coop), dim3(nblocks), dim3(kBlock), params, 0, st);
}

hipError_t launch_images(const SampleDev* sample, const PairDev* pair, int clockwise, const double* lt_item,
                         const double* tau, const double* max_lt, int nT, int nfree, int ppr, double* img,
                         hipStream_t st)
{
    if (nfree <= 0 || nT <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_images, dim3(blocks_for(nfree, 256)), dim3(256), 0, st, sample, pair, clockwise, lt_item,
                       tau, max_lt, nT, nfree, ppr, img);
    return hipGetLastError();
}

static const void* trace32_kernel(int clockwise, int coop)
{
    if (coop != 0) return clockwise ? (const void*)k_trace32_c<true> : (const void*)k_trace32_c<false>;
    return clockwise ? (const void*)k_trace32<true> : (const void*)k_trace32<false>;
}

hipError_t trace32_blocks_per_cu(int clockwise, int coop, int* blocks)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, trace32_kernel(clockwise, coop), kBlock, 0);
}

hipError_t launch_trace_mixed(const TraceArgs& a, int clockwise, int waves_per_eu, int coop, int nblocks32,
                              int nblocks64, hipStream_t st)
{
    if (a.npairs <= 0 || nblocks32 <= 0 || nblocks64 <= 0 || a.chunk <= 0 || !a.list || !a.list_count)
        return hipErrorInvalidValue;
    hipError_t e = hipSuccess;
    if (a.total_items <= 0) return e;
    TraceArgs a32 = a;
    a32.list_mode = 0;
    void* params[] = {&a32};
    if ((e = hipLaunchKernel(trace32_kernel(clockwise, coop), dim3(nblocks32), dim3(kBlock), params, 0, st)) !=
        hipSuccess)
        return e;
    // The flagged items, in FP64: the count is read on the device, so no host
    // round trip separates the two traces.
    TraceArgs a64 = a;
    a64.list_mode = 1;
    a64.chunk = 16;
    return launch_trace(a64, clockwise, waves_per_eu, coop, nblocks64, st);
}

template <int NT>
static hipError_t launch_sigma_t(const SigmaArgs& a, hipStream_t st)
{
    hipLaunchKernelGGL((k_sigma<NT>), dim3(a.total_blocks), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_sigma(const SigmaArgs& a, int NT, hipStream_t st)
{
    if (a.npairs <= 0 || a.total_blocks <= 0) return hipErrorInvalidValue;
    switch (NT) {
    case 1: return launch_sigma_t<1>(a, st);
    case 2: return launch_sigma_t<2>(a, st);
    case 3: return launch_sigma_t<3>(a, st);
    case 4: return launch_sigma_t<4>(a, st);
    case 8: return launch_sigma_t<8>(a, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_finalize(const SampleDev* samples, const PairDev* pairs, int npairs, int NT, int with_metrics,
                           const double* dpart, const int64_t* mpart, const double* factor, double* out_sig,
                           double* out_lt_sum, int64_t* out_m, hipStream_t st)
{
    hipLaunchKernelGGL(k_finalize, dim3(npairs), dim3(256), 0, st, samples, pairs, NT, with_metrics, dpart, mpart,
                       factor, out_sig, out_lt_sum, out_m);
    return hipGetLastError();
}

}  // namespace es
