// sps_kernels.hip - hand-written HIP kernels (gfx950) of the Stress-Particle
// SPH hot path: cell binning, the pair-order split of the neighbour search,
// and the per-stage gather kernels of the RK4 step.
//
// Storage order.  Per-particle state lives in one set of structure-of-arrays
// buffers whose index ranges are the reference's (nodes [0,nn), stress
// points [nn,nt), dummies [nt,nt2)) but, inside each range, sorted by
// (cell, original index).  Within one cell the reference visits particles in
// ascending original index (code/2_SPH_main_2018.f90:1307-1313), i.e. all
// nodes, then all stress points, then all dummies of the cell, so every
// 3x3 stencil walk of a receiver is a short series of contiguous ranges.
//
// Summation order.  Every pair loop of the reference scatters into both ends
// while walking its linked list; one receiver therefore sums its pairs in
// list order.  The list order of a pair follows from its discovery index d
// (closed form in grid_find_NEW, 2_...:1309-1401): with M the high-water mark
// before the search and n the pairs found, pairs d >= M come first in
// descending d, then pairs d < M ascending.  For a receiver, ascending d is
// the order of its partners in (cell, kind, original index) -- the forward
// stencil walk -- so a receiver's partners are visited with a reverse walk
// (pairs d >= M) followed by a forward walk (pairs d < M) and added one by
// one in that order.  The search (k_count_lists) does that walk once per
// step and stores each receiver's accepted partners in forward walk order
// with the pair's single-precision w and grad w (PairLists); the stage
// kernels read the lists forwards, and backwards for the pairs d >= M, from
// windows of the partners staged in LDS: one lane per receiver, no atomics,
// results bitwise equal to the reference.
#include <type_traits>

#include "sps_device.h"

namespace sps {

constexpr int BLOCK = 256;

// Bits of the device error flag (read back at the end of sps_step).
constexpr int ERR_WALL_TERMS = 1;   // more wall-force terms on one node than MAX_WALL_TERMS
constexpr int ERR_BUCKET = 2;       // more particles of one kind in one cell than MAX_BUCKET_SORT
constexpr int ERR_NONFINITE = 4;    // a non-finite position, velocity or stress in the domain box
constexpr int ERR_GRID = 8;         // the cell grid exceeded its capacity (sized from the domain box)

// Launch size of a kernel that loops over a device-sized range (cells of
// the current grid): enough blocks to cover the host's upper bound, capped.
constexpr unsigned MAX_STRIDE_BLOCKS = 4096;

__device__ __forceinline__ bool finite2(double2 v) { return isfinite(v.x) && isfinite(v.y); }
__device__ __forceinline__ bool finite4(double4 v) {
    return isfinite(v.x) && isfinite(v.y) && isfinite(v.z) && isfinite(v.w);
}

// Grid of the current search (device-resident).
struct Grid {
    double xmin[2];
    double deltx;
    int ndivx, ndivy, ndivt;
    int overflow;
};

// Pair counts and the list split of the current search.
struct Split {
    unsigned long long n;       // pairs found
    unsigned long long M;       // high-water mark before this search
    int sstar;                  // processing slot holding discovery index M, or -1
    int pstar;                  // processing slot of its first reversed partner
    int first_rev;              // smallest slot discovering a pair d >= M (INT_MAX: none)
    int pad_;
    unsigned long long n_by_type[10];
};

// ------------------------------------------------------------ out of domain
// Check_Out_Domain (2_...:1170-1194), sticky.
__global__ void k_check_ood(const double2 *__restrict__ x, int *__restrict__ ood, int nt2,
                            double xmin0, double xmin1, double xmax0, double xmax1) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt2) return;
    double2 p = x[i];
    double d0 = (p.x - xmin0) * (p.x - xmax0);
    double d1 = (p.y - xmin1) * (p.y - xmax1);
    // d > 0 as the reference tests it; a NaN coordinate is flagged too, so
    // that no cell index is ever derived from it
    if (!(d0 <= 0.0) || !(d1 <= 0.0)) ood[i] = 1;
}

// ------------------------------------------------------------- grid bounds
// Task 1 of grid_find_NEW (2_...:1245-1258): min/max of x over in-domain
// particles and the largest h.  Min/max are exact, so the reduction order
// does not matter.
__global__ void k_bounds_partial(const double2 *__restrict__ x, const double *__restrict__ hsml,
                                 const int *__restrict__ ood, int nt2, double *__restrict__ part) {
    __shared__ double s[5][BLOCK];
    double mn0 = (double)1.e+10f, mn1 = (double)1.e+10f, mx0 = -(double)1.e+10f,
           mx1 = -(double)1.e+10f, hm = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nt2; i += gridDim.x * blockDim.x) {
        if (ood[i] == 1) continue;
        double2 p = x[i];
        mn0 = fmin(mn0, p.x);
        mx0 = fmax(mx0, p.x);
        mn1 = fmin(mn1, p.y);
        mx1 = fmax(mx1, p.y);
        hm = fmax(hm, hsml[i]);
    }
    int t = threadIdx.x;
    s[0][t] = mn0; s[1][t] = mn1; s[2][t] = mx0; s[3][t] = mx1; s[4][t] = hm;
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (t < w) {
            s[0][t] = fmin(s[0][t], s[0][t + w]);
            s[1][t] = fmin(s[1][t], s[1][t + w]);
            s[2][t] = fmax(s[2][t], s[2][t + w]);
            s[3][t] = fmax(s[3][t], s[3][t + w]);
            s[4][t] = fmax(s[4][t], s[4][t + w]);
        }
        __syncthreads();
    }
    if (t == 0)
        for (int k = 0; k < 5; ++k) part[5 * blockIdx.x + k] = s[k][0];
}

// Grid parameters with the reference's double-precision expressions
// (one block: reduce the partials, then thread 0 derives the grid).  skin
// widens the cells beyond 2 max h for pair lists that are kept over several
// steps (0: the reference's grid).  Launch with exactly BLOCK threads.
__global__ void __launch_bounds__(BLOCK) k_bounds_final(const double *__restrict__ part, int nparts,
                                                        Grid *__restrict__ g, int cell_cap, double skin,
                                                        int *__restrict__ err) {
    __shared__ double s[5][BLOCK];
    int t = threadIdx.x;
    double mn0 = (double)1.e+10f, mn1 = (double)1.e+10f, mx0 = -(double)1.e+10f,
           mx1 = -(double)1.e+10f, hm = 0.0;
    for (int b = t; b < nparts; b += blockDim.x) {
        mn0 = fmin(mn0, part[5 * b + 0]);
        mn1 = fmin(mn1, part[5 * b + 1]);
        mx0 = fmax(mx0, part[5 * b + 2]);
        mx1 = fmax(mx1, part[5 * b + 3]);
        hm = fmax(hm, part[5 * b + 4]);
    }
    s[0][t] = mn0; s[1][t] = mn1; s[2][t] = mx0; s[3][t] = mx1; s[4][t] = hm;
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (t < w) {
            s[0][t] = fmin(s[0][t], s[0][t + w]);
            s[1][t] = fmin(s[1][t], s[1][t + w]);
            s[2][t] = fmax(s[2][t], s[2][t + w]);
            s[3][t] = fmax(s[3][t], s[3][t + w]);
            s[4][t] = fmax(s[4][t], s[4][t + w]);
        }
        __syncthreads();
    }
    if (t != 0) return;
    double mn[2] = {s[0][0], s[1][0]}, mx[2] = {s[2][0], s[3][0]};
    double deltx = s[4][0] * 2;
    if (skin > 0) deltx = deltx + skin;
    int nd[2];
    for (int d = 0; d < 2; ++d) {
        double length = mx[d] - mn[d];
        nd[d] = (int)(length / deltx + 1);
        double length_new = nd[d] * deltx;
        g->xmin[d] = mn[d] - (length_new - length) / 2 - (double)0.001f * length;
    }
    g->deltx = deltx;
    g->ndivx = nd[0];
    g->ndivy = nd[1];
    long long ndivt = (long long)nd[0] * nd[1];
    g->overflow = (ndivt > cell_cap || nd[0] < 1 || nd[1] < 1) ? 1 : 0;
    g->ndivt = g->overflow ? 0 : (int)ndivt;
    if (g->overflow) atomicOr(err, ERR_GRID);   // every particle then sits in the sentinel cell
}

// Clear the per-(kind, cell) counters of the current grid, 3 (ndivt + 1) + 1
// entries each (sized on the device: no host round trip).
__global__ void k_zero_cells(int *__restrict__ a, int *__restrict__ b, const Grid *__restrict__ g) {
    const long long n = 3LL * (g->ndivt + 1) + 1;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        a[i] = 0;
        b[i] = 0;
    }
}

// ------------------------------------------------------------- cell keys
// Buckets are kept in strip-major cell order: the grid is cut into vertical
// strips of STRIP cells, cells run row by row inside a strip.  Walking the
// storage (and so the receivers) up a strip keeps the partner data of three
// consecutive rows in one L2.  The reference's own visiting order (row-major
// cells, 2_...:1322-1324) is kept separately for the list order (k_proc_slot).
constexpr int STRIP = 8;

__device__ __forceinline__ int cell_key(int cx, int cy, const Grid &g) {
    int x0 = (cx / STRIP) * STRIP;
    int w = g.ndivx - x0 < STRIP ? g.ndivx - x0 : STRIP;
    return x0 * g.ndivy + cy * w + (cx - x0);
}

__device__ __forceinline__ void cell_xy(int key, const Grid &g, int &cx, int &cy) {
    int x0 = (key / (STRIP * g.ndivy)) * STRIP;
    int w = g.ndivx - x0 < STRIP ? g.ndivx - x0 : STRIP;
    int rem = key - x0 * g.ndivy;
    cy = rem / w;
    cx = x0 + rem % w;
}

// ------------------------------------------------------------- cell binning
// Runs of equal keys among a wave's lanes.  The particles are stored sorted
// by kind and cell (the previous search's order), so consecutive lanes mostly
// share a (kind, cell) bucket; one atomic per run instead of one per lane
// keeps the bucket counters from serialising in L2.  Every lane of the wave
// calls it (lanes without a particle pass a key of their own, -1).
struct WaveRun {
    int head;   // the run's first lane
    int rank;   // this lane's place in its run
    int len;    // lanes in the run (meaningful at the head)
};

__device__ __forceinline__ WaveRun wave_run(int key) {
    const int lane = (int)(threadIdx.x & 63);
    const int prev = __shfl_up(key, 1, 64);
    const unsigned long long heads = __ballot(lane == 0 || prev != key);
    const unsigned long long upto = (2ull << lane) - 1;   // lanes 0..lane (all for lane 63)
    const int head = 63 - __clzll(heads & upto);
    const unsigned long long after = heads & ~upto;
    const int next = after ? __ffsll(after) - 1 : 64;
    return WaveRun{head, lane - head, next - lane};
}

// Task 2 of grid_find_NEW (2_...:1277-1305): cell of every particle and
// per-(kind, cell) counts.  Out-of-domain particles go to the sentinel cell
// ndivt at the end of their kind's range.
__global__ void k_cell_count(const double2 *__restrict__ x, const int *__restrict__ ood,
                             const Grid *__restrict__ g, int *__restrict__ cellkey,
                             int *__restrict__ kcount, int nn, int nt, int nt2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < nt2;
    const int ndivt = g->ndivt;
    int b = -1;
    if (live) {
        int c = ndivt;
        if (ood[i] != 1 && !g->overflow) {
            double2 p = x[i];
            int ix = (int)((p.x - g->xmin[0]) / g->deltx + 1);
            int iy = (int)((p.y - g->xmin[1]) / g->deltx + 1);
            if (ix > g->ndivx) ix = g->ndivx;
            if (iy > g->ndivy) iy = g->ndivy;
            c = cell_key(ix - 1, iy - 1, *g);
        }
        cellkey[i] = c;
        b = kind_of(i, nn, nt) * (ndivt + 1) + c;
    }
    const WaveRun w = wave_run(b);
    if (live && w.rank == 0) atomicAdd(&kcount[b], w.len);
}

// Scatter into per-(kind, cell) buckets (arbitrary order inside a bucket):
// the run's head claims the run's slots, the lanes take them in lane order.
__global__ void k_cell_scatter(const int *__restrict__ cellkey, const int *__restrict__ orig,
                               const int *__restrict__ kstart, int *__restrict__ kfill,
                               const Grid *__restrict__ g, int *__restrict__ tmp_orig,
                               int *__restrict__ tmp_src, int *__restrict__ new_cell, int nn,
                               int nt, int nt2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < nt2;
    const int ndivt = g->ndivt;
    int b = -1, c = 0;
    if (live) {
        c = cellkey[i];
        b = kind_of(i, nn, nt) * (ndivt + 1) + c;
    }
    const WaveRun w = wave_run(b);
    int first = 0;
    if (live && w.rank == 0) first = atomicAdd(&kfill[b], w.len);
    first = __shfl(first, w.head, 64);
    if (!live) return;
    const int slot = kstart[b] + first + w.rank;
    tmp_orig[slot] = orig[i];
    tmp_src[slot] = i;
    new_cell[slot] = c;
}

// Restore ascending original index inside each (kind, cell) bucket.  The
// sentinel buckets of out-of-domain particles are sorted too when small;
// their order never reaches a pair loop.
constexpr int MAX_BUCKET_SORT = 4096;

__global__ void k_cell_sort(const int *__restrict__ kstart, int *__restrict__ tmp_orig,
                            int *__restrict__ tmp_src, const Grid *__restrict__ g,
                            int *__restrict__ err) {
    const int ndivt = g->ndivt;
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < 3 * (ndivt + 1); b += gridDim.x * blockDim.x) {
        const int lo = kstart[b], hi = kstart[b + 1];
        if (hi - lo < 2) continue;
        if (hi - lo > MAX_BUCKET_SORT) {
            if (b % (ndivt + 1) != ndivt) atomicOr(err, ERR_BUCKET);   // an in-domain cell
            continue;
        }
        for (int a = lo + 1; a < hi; ++a) {
            int ko = tmp_orig[a], ks = tmp_src[a];
            int p = a - 1;
            while (p >= lo && tmp_orig[p] > ko) {
                tmp_orig[p + 1] = tmp_orig[p];
                tmp_src[p + 1] = tmp_src[p];
                --p;
            }
            tmp_orig[p + 1] = ko;
            tmp_src[p + 1] = ks;
        }
    }
}

// Particles per cell in the reference's row-major visiting order.
__global__ void k_cell_totals(const int *__restrict__ kstart, const Grid *__restrict__ g,
                              unsigned *__restrict__ tot) {
    const Grid gg = *g;
    const int ndivt = gg.ndivt, L = ndivt + 1;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c <= ndivt; c += gridDim.x * blockDim.x) {
        if (c == ndivt) {
            tot[c] = 0;
            continue;
        }
        const int key = cell_key(c % gg.ndivx, c / gg.ndivx, gg);
        tot[c] = (unsigned)((kstart[key + 1] - kstart[key]) + (kstart[L + key + 1] - kstart[L + key]) +
                            (kstart[2 * L + key + 1] - kstart[2 * L + key]));
    }
}

// Processing slot (the reference's visiting order: row-major cells, inside
// a cell ascending original index, i.e. nodes, stress points, dummies) of
// each in-domain particle.
__global__ void k_proc_slot(const int *__restrict__ cell, const int *__restrict__ kstart,
                            const unsigned *__restrict__ cstart_rm, const Grid *__restrict__ g,
                            int *__restrict__ proc, int nn, int nt, int nt2) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nt2) return;
    int ndivt = g->ndivt;
    int c = cell[s];
    if (c >= ndivt) {
        proc[s] = -1;
        return;
    }
    const int L = ndivt + 1;
    int cx, cy;
    cell_xy(c, *g, cx, cy);
    int st0 = kstart[c], st1 = kstart[L + c], st2 = kstart[2 * L + c];
    int n0 = kstart[c + 1] - st0, n1 = kstart[L + c + 1] - st1;
    int k = kind_of(s, nn, nt);
    int pos = k == KIND_NODE ? s - st0 : (k == KIND_SP ? n0 + s - st1 : n0 + n1 + s - st2);
    proc[s] = (int)cstart_rm[cy * g->ndivx + cx] + pos;
}

// ------------------------------------------------------------- permutation
struct StateBufs {
    double2 *x, *vel, *x00, *displ;
    double4 *stress;
    double *mass, *rho, *hsml, *ivar1, *fdr;
    int *orig, *ood;
    float *wallpos, *horiz;
    int *mat;                  // material of each slot (null: every particle has DevParams')
    const Material *mtab;      // the materials
};

// The friction / cohesion (and, for material_params, every material
// constant) of slot i: its own material's where a table is set, else the
// DevParams values, unchanged.
__device__ __forceinline__ void material_tc(const StateBufs &S, int i, double &tanfi, double &coh) {
    if (S.mat) {
        const Material &m = S.mtab[S.mat[i]];
        tanfi = m.tanfi;
        coh = m.coh;
    }
}
__device__ __forceinline__ DevParams material_params(const DevParams &P, const StateBufs &S, int i) {
    DevParams q = P;
    if (S.mat) {
        const Material &m = S.mtab[S.mat[i]];
        q.tanfi = m.tanfi; q.coh = m.coh; q.young = m.young; q.poiss = m.poiss;
        q.D11 = m.D11; q.D22 = m.D22; q.D12 = m.D12; q.D33 = m.D33; q.D41 = m.D41; q.D42 = m.D42;
    }
    return q;
}

// The state into the new order: slot t takes slot perm[t]'s particle.  The
// state stays in src: a block of BLOCK slots whose particles all stay in
// place (most blocks after the first step) does nothing; the others gather
// into the scratch set dst here and k_permute_commit copies them back once
// every gather is done (a gather reads any slot of src).
__global__ void k_permute(StateBufs dst, StateBufs src, const int *__restrict__ perm, int nn,
                          int nt, int nt2, unsigned char *__restrict__ moved) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool mine = t < nt2 && perm[t] != t;
    const bool any = __syncthreads_or(mine);
    if (threadIdx.x == 0) moved[blockIdx.x] = any ? 1 : 0;
    if (!any || t >= nt2) return;
    int s = perm[t];
    dst.x[t] = src.x[s];
    dst.vel[t] = src.vel[s];
    dst.x00[t] = src.x00[s];
    dst.stress[t] = src.stress[s];
    dst.mass[t] = src.mass[s];
    dst.rho[t] = src.rho[s];
    dst.hsml[t] = src.hsml[s];
    dst.orig[t] = src.orig[s];
    dst.ood[t] = src.ood[s];
    dst.wallpos[t] = src.wallpos[s];
    dst.horiz[t] = src.horiz[s];
    if (t < nn) dst.displ[t] = src.displ[s];
    if (t < nt) {
        dst.ivar1[t] = src.ivar1[s];
        dst.fdr[t] = src.fdr[s];
    }
    if (src.mat) dst.mat[t] = src.mat[s];
}

__global__ void k_permute_commit(StateBufs dst, StateBufs src, const unsigned char *__restrict__ moved, int nn,
                                 int nt, int nt2) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (!moved[blockIdx.x] || t >= nt2) return;
    dst.x[t] = src.x[t];
    dst.vel[t] = src.vel[t];
    dst.x00[t] = src.x00[t];
    dst.stress[t] = src.stress[t];
    dst.mass[t] = src.mass[t];
    dst.rho[t] = src.rho[t];
    dst.hsml[t] = src.hsml[t];
    dst.orig[t] = src.orig[t];
    dst.ood[t] = src.ood[t];
    dst.wallpos[t] = src.wallpos[t];
    dst.horiz[t] = src.horiz[t];
    if (t < nn) dst.displ[t] = src.displ[t];
    if (t < nt) {
        dst.ivar1[t] = src.ivar1[t];
        dst.fdr[t] = src.fdr[t];
    }
    if (src.mat) dst.mat[t] = src.mat[t];
}

// Materials of the slots from the materials of the original particles.
__global__ void k_material_slots(const int *__restrict__ orig, const int *__restrict__ mat_orig,
                                 int *__restrict__ mat, int nt2) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nt2) mat[t] = mat_orig[orig[t]];
}

// --------------------------------------------------------- stencil helpers
// The 3x3 stencil of a cell: rows y0..y1, columns x0..x1 (clamped at the
// grid edges like 2_...:1326-1327), visited cell by cell.  Kept in scalars:
// an indexed array here would be placed in scratch or LDS by the compiler.
struct Stencil {
    int cx, cy, x0, x1, y0, y1;
};

__device__ __forceinline__ Stencil make_stencil(int key, const Grid &g) {
    Stencil S;
    cell_xy(key, g, S.cx, S.cy);
    S.x0 = S.cx > 0 ? S.cx - 1 : 0;
    S.x1 = S.cx + 1 < g.ndivx ? S.cx + 1 : g.ndivx - 1;
    S.y0 = S.cy > 0 ? S.cy - 1 : 0;
    S.y1 = S.cy + 1 < g.ndivy ? S.cy + 1 : g.ndivy - 1;
    return S;
}


// pint_type of a pair from the kinds of its ends (3_...:1598-1608).
__device__ __forceinline__ int pair_type(int ka, int kb) {
    int isumm = (ka == KIND_NODE ? 2 : ka == KIND_SP ? 1 : 25) + (kb == KIND_NODE ? 2 : kb == KIND_SP ? 1 : 25);
    switch (isumm) {
        case 3: return 1;
        case 2: return 2;
        case 4: return 3;
        case 27: return 6;
        case 26: return 9;
        default: return 0;
    }
}

// ------------------------------------------------------------------- count
// For every receiver: pairs it discovers (partners later in processing
// order; the reference counts n_pairs with these), by pint_type.  The node
// and stress-point receivers are counted by the tiled k_count_lists below.
// (Slots [r0, r1), global storage; the dummy receivers.)
__global__ void k_count(const double2 *__restrict__ x, const double *__restrict__ hsml,
                        const int *__restrict__ cell, const int *__restrict__ kstart,
                        const int *__restrict__ proc, const Grid *__restrict__ g,
                        unsigned *__restrict__ cnt, Split *__restrict__ split, int nn, int nt,
                        int r0, int r1) {
    int r = r0 + xcd_block() * blockDim.x + threadIdx.x;
    unsigned fwd = 0;
    unsigned t0 = 0, t1 = 0, t2 = 0, t3 = 0, t6 = 0, t9 = 0;   // pairs by pint_type
    int ndivt = g->ndivt;
    if (r < r1 && cell[r] < ndivt) {
        const int L = ndivt + 1;
        const Grid gg = *g;
        int kr = kind_of(r, nn, nt);
        double2 xr = x[r];
        double hr = hsml[r];
        Stencil S = make_stencil(cell[r], gg);
        const int crm_r = S.cy * gg.ndivx + S.cx;
        for (int y = S.y0; y <= S.y1; ++y) {
            for (int xx = S.x0; xx <= S.x1; ++xx) {
                const int key = cell_key(xx, y, gg), crm = y * gg.ndivx + xx;
                for (int ks = 0; ks < 3; ++ks) {
                    int lo = kstart[ks * L + key], hi = kstart[ks * L + key + 1];
                    for (int s = lo; s < hi; ++s) {
                        if (s == r) continue;
                        double2 xs = x[s];
                        double mh;
                        double rr = pair_distance(xr.x - xs.x, xr.y - xs.y, hr, hsml[s], mh);
                        if (rr < 0.0) continue;
                        bool after = crm > crm_r || (crm == crm_r && (ks > kr || (ks == kr && s > r)));
                        if (after) {
                            fwd++;
                            switch (pair_type(kr, ks)) {
                                case 1: t1++; break;
                                case 2: t2++; break;
                                case 3: t3++; break;
                                case 6: t6++; break;
                                case 9: t9++; break;
                                default: t0++; break;
                            }
                        }
                    }
                }
            }
        }
        cnt[proc[r]] = fwd;
    }
    // per-type histogram: wave reduction then one atomic per wave
    unsigned tv[6] = {t0, t1, t2, t3, t6, t9};
    const int tid[6] = {0, 1, 2, 3, 6, 9};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        unsigned v = tv[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if ((threadIdx.x & 63) == 0 && v) atomicAdd(&split->n_by_type[tid[k]], (unsigned long long)v);
    }
}

// Fresh split record: M is the list high-water mark before this search,
// the largest pair count of the earlier searches since the list started
// (2_...:1290-1294, 1361-1369): 0 on a fresh list, else max(M, n) of the
// previous record.
__global__ void k_split_reset(Split *__restrict__ split, int fresh) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const unsigned long long M = fresh ? 0ull : (split->M > split->n ? split->M : split->n);
    split->n = 0;
    split->M = M;
    split->sstar = -1;
    split->pstar = -1;
    split->first_rev = 0x7fffffff;
    for (int k = 0; k < 10; ++k) split->n_by_type[k] = 0;
}

// Locate the processing slot whose discoveries straddle index M (n is the
// scan total of the per-slot counts, written by the scan).
__global__ void k_split(const unsigned *__restrict__ cnt, const unsigned *__restrict__ off,
                        int nslots, Split *__restrict__ split) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nslots) return;
    unsigned long long M = split->M;
    unsigned long long o = off[p], c = cnt[p];
    if (o < M && M < o + c) split->sstar = p;
    // off + cnt grows with the slot, so the first slot discovering a pair
    // d >= M is the one slot with off <= M < off + cnt
    if (o <= M && M < o + c) split->first_rev = p;
}

// First reversed partner of the split slot: its (M - off)-th discovery in
// processing order (2_...:1322-1383 visit partners in that order).
__global__ void k_split_partner(const double2 *__restrict__ x, const double *__restrict__ hsml,
                                const int *__restrict__ cell, const int *__restrict__ kstart,
                                const int *__restrict__ proc, const unsigned *__restrict__ off,
                                const int *__restrict__ proc_to_slot, const Grid *__restrict__ g,
                                Split *__restrict__ split, int nn, int nt,
                                unsigned long long *__restrict__ pairs_total) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    *pairs_total += split->n;   // pairs found by every search of the handle
    int ps = split->sstar;
    split->pstar = -1;
    if (ps < 0) return;
    unsigned long long kstar = split->M - off[ps];
    int r = proc_to_slot[ps];
    const Grid gg = *g;
    const int L = gg.ndivt + 1;
    int kr = kind_of(r, nn, nt);
    double2 xr = x[r];
    Stencil S = make_stencil(cell[r], gg);
    const int crm_r = S.cy * gg.ndivx + S.cx;
    unsigned long long local = 0;
    for (int y = S.y0; y <= S.y1; ++y)
        for (int xx = S.x0; xx <= S.x1; ++xx) {
            const int key = cell_key(xx, y, gg), crm = y * gg.ndivx + xx;
            if (crm < crm_r) continue;
            for (int ks = 0; ks < 3; ++ks)
                for (int s = kstart[ks * L + key]; s < kstart[ks * L + key + 1]; ++s) {
                    if (crm == crm_r && !(ks > kr || (ks == kr && s > r))) continue;
                    double2 xs = x[s];
                    double mh;
                    if (pair_distance(xr.x - xs.x, xr.y - xs.y, hsml[r], hsml[s], mh) < 0.0) continue;
                    if (local == kstar) {
                        split->pstar = proc[s];
                        return;
                    }
                    local++;
                }
        }
}

__global__ void k_proc_to_slot(const int *__restrict__ proc, int *__restrict__ proc_to_slot, int nt2) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nt2 && proc[s] >= 0) proc_to_slot[proc[s]] = s;
}

// ---------------------------------------------------------- ordered walk
// Which side of the high-water mark M a pair's discovery index falls on:
// the discoverer is the end visited first (smaller processing slot); only
// the split slot straddles M, and its partners from pstar on are reversed.
__device__ __forceinline__ bool pair_reversed(int pr, int ps, const unsigned *__restrict__ cnt,
                                              const unsigned *__restrict__ off,
                                              const Split &sp) {
    int pd = pr < ps ? pr : ps, po = pr < ps ? ps : pr;
    unsigned long long o = off[pd], c = cnt[pd];
    if (o >= sp.M) return true;
    if (o + c <= sp.M) return false;
    return po >= sp.pstar;
}

// Everything a stage kernel needs to re-find the pairs of the step's search.
// A receiver whose processing slot is below first_rev has only pairs d < M
// (every pair's discoverer is at or below the receiver's slot), so it needs
// the forward walk alone and no per-pair classification.
struct SearchView {
    const int *cell, *kstart, *proc;
    const unsigned *cnt, *off;
    const Grid *g;
    const Split *split;
    double pi;
};

struct WalkCtx {
    Grid g;
    Split sp;
    int L;
    bool any_rev, any_fwd, mixed;
};

__device__ __forceinline__ WalkCtx walk_ctx(const SearchView &V) {
    WalkCtx w;
    w.g = *V.g;
    w.sp = *V.split;
    w.L = w.g.ndivt + 1;
    w.any_rev = w.sp.n > w.sp.M;    // some pairs have d >= M (listed first, descending d)
    w.any_fwd = w.sp.M > 0;         // some pairs have d < M (listed after, ascending d)
    w.mixed = w.any_rev && w.any_fwd;
    return w;
}

constexpr int MASK_NODE = 1 << KIND_NODE, MASK_SP = 1 << KIND_SP, MASK_DUMMY = 1 << KIND_DUMMY;

template <int K> struct KindTag {
    static constexpr int value = K;
};

// Where the partners of kind K in cell (gx, gy) are: global storage slots.
struct GlobalCells {
    const int *kstart;
    Grid g;
    int L;
    __device__ __forceinline__ int2 range(int kind, int gx, int gy) const {
        const int key = cell_key(gx, gy, g);
        const int lo = kstart[kind * L + key];
        return make_int2(lo, kstart[kind * L + key + 1] - lo);
    }
};

// Partner fields from the global state (j is a storage slot).
struct GlobalFields {
    StateBufs S;
    __device__ __forceinline__ double2 xk(int, int j) const { return S.x[j]; }
    __device__ __forceinline__ double hk(int, int j) const { return S.hsml[j]; }
    template <int K> __device__ __forceinline__ double2 x(KindTag<K>, int j) const { return S.x[j]; }
    template <int K> __device__ __forceinline__ double h(KindTag<K>, int j) const { return S.hsml[j]; }
    template <int K> __device__ __forceinline__ double m(KindTag<K>, int j) const { return S.mass[j]; }
    template <int K> __device__ __forceinline__ double rho(KindTag<K>, int j) const { return S.rho[j]; }
    template <int K> __device__ __forceinline__ double2 vel(KindTag<K>, int j) const { return S.vel[j]; }
    template <int K> __device__ __forceinline__ double4 st(KindTag<K>, int j) const { return S.stress[j]; }
    template <int K> __device__ __forceinline__ double iv(KindTag<K>, int j) const { return S.ivar1[j]; }
    template <int K> __device__ __forceinline__ double mr(KindTag<K>, int j) const { return S.mass[j] / S.rho[j]; }
    template <int K> __device__ __forceinline__ double3 sr2(KindTag<K>, int j) const {
        const double rs = S.rho[j], rs2 = rs * rs;
        const double4 st = S.stress[j];
        return make_double3(st.x / rs2, st.y / rs2, st.z / rs2);
    }
    template <int K> __device__ __forceinline__ float wall(KindTag<K>, int j) const { return S.wallpos[j]; }
    template <int K> __device__ __forceinline__ float horiz(KindTag<K>, int j) const { return S.horiz[j]; }
};

// --------------------------------------------------------------- LDS tiles
// A tiled stage kernel gives each workgroup TILE consecutive receivers of one
// kind.  In storage they are a run of cells of one strip, rows cy_first ..
// cy_last; the union of their 3x3 stencils is the window: those rows +-1,
// the strip's columns and one halo column on each side.  Per partner kind
// the strip part of the window is ONE contiguous storage run (rows follow
// each other inside a strip) and each halo cell a short run, so staging is
// a coalesced copy.  A table gives every window cell its LDS run.  A tile
// whose receivers span two strips or too many rows, whose window overflows
// the LDS capacity, or that holds a receiver with pairs on both sides of M
// walks the global storage instead; both paths visit the same pairs in the
// same order.
//
// The tile width and the window capacities are build parameters
// (sps_tiles.h, written by tune_tiles.py from timings of the candidates;
// -DSPS_TILE=... etc. override it).  Results do not depend on them: every
// receiver sums the same pairs in the same order whichever path it takes.
#include "sps_tiles.h"
constexpr int TILE = SPS_TILE;
// list words in flight in the walks of the two register-heaviest stage
// kernels (k_sp_update's gradient walk, k_deriv_nodes' walk): one keeps them
// at 124 and 126 registers, 4 waves per SIMD (two: 131 and 134, 3 waves;
// measured 1 % slower per FP64 step on the 1 M slope)
#ifndef SPS_WALK_DEPTH
#define SPS_WALK_DEPTH 1
#endif
constexpr int TROWS = SPS_TROWS;       // most window rows staged
constexpr int TCOLS = STRIP + 2;
constexpr int TCELLS = TROWS * TCOLS;
constexpr int CAP = SPS_CAP;           // staged partners of a main kind per window
constexpr int CAP_D = SPS_CAP_D;       // staged dummies per window
static_assert(TILE % 64 == 0 && TILE >= 128 && TILE <= 1024,
              "a tile is whole waves, at least two (the stage kernels plan two kinds at once), at most 1024 lanes");
constexpr int TWAVES = TILE / 64;      // waves of a tile
static_assert(CAP > 0 && CAP < (1 << 14) && CAP_D > 0 && CAP_D < (1 << 14),
              "list entries carry a 14-bit window index");
static_assert(CAP % 8 == 0 && CAP_D % 8 == 0, "every staged field starts 16-byte aligned");
// waves per SIMD the FP64 stage kernels k_sp_update and k_deriv_nodes are
// compiled for (their registers otherwise allow 3): 0 leaves it to the
// compiler; more waves may cost register spills
#if SPS_STAGE_WAVES > 0
#define STAGE_OCCUPANCY __attribute__((amdgpu_waves_per_eu(SPS_STAGE_WAVES)))
#else
#define STAGE_OCCUPANCY
#endif
static_assert(TROWS >= 3 && 2 * TROWS <= 64,
              "a window holds at least one receiver row and its neighbours; one wave plans its halo rows");

// Dynamic LDS of the tiled stage kernels.  A kernel's layout gives the byte
// offset of every field it stages per partner kind (-1: not staged) and of
// the per-kind window-cell tables; all offsets are compile-time constants,
// so every access is a direct LDS access.
extern __shared__ __align__(16) unsigned char sps_lds[];

template <typename T> __device__ __forceinline__ T *lds_at(int off) {
    return reinterpret_cast<T *>(sps_lds + off);
}

struct LdsLayout {
    int x[3], v[3], h[3], m[3], rho[3], s0[3], s1[3], s2[3], s3[3], iv[3], tab[3];
    int mr[3], p0[3], p1[3], p2[3];   // m / rho and sigma / rho^2 (xx, yy, xy) of a partner
    int wall, horiz, bytes;
};

constexpr LdsLayout layout_none() {
    LdsLayout L{};
    for (int k = 0; k < 3; ++k)
        L.x[k] = L.v[k] = L.h[k] = L.m[k] = L.rho[k] = L.s0[k] = L.s1[k] = L.s2[k] = L.s3[k] = L.iv[k] =
            L.tab[k] = L.mr[k] = L.p0[k] = L.p1[k] = L.p2[k] = -1;
    L.wall = L.horiz = -1;
    L.bytes = 0;
    return L;
}

// count / list build: positions and smoothing lengths of all three kinds and
// the window-cell tables
// The list build: x of every kind, the cell tables, then h of every kind,
// which a launch allocates only when the smoothing lengths differ (with one
// smoothing length the window is 26 KiB at 256 receivers, six workgroups per
// CU instead of four).
constexpr LdsLayout layout_count() {
    LdsLayout L = layout_none();
    int o = 0;
    for (int k = 0; k < 3; ++k) {
        L.x[k] = o;
        o += 16 * (k == KIND_DUMMY ? CAP_D : CAP);
    }
    for (int k = 0; k < 3; ++k) {
        L.tab[k] = o;
        o += 4 * TCELLS;
    }
    o = (o + 15) & ~15;
    for (int k = 0; k < 3; ++k) {
        L.h[k] = o;
        o += 8 * (k == KIND_DUMMY ? CAP_D : CAP);
    }
    L.bytes = o;
    return L;
}

// The stage kernels read w and grad w from the pair lists and only the
// fields they sum from the window.
// node receivers <- stress points: m, rho, stress (4), plastic strain
constexpr LdsLayout layout_spu_nodes() {
    LdsLayout L = layout_none();
    int o = 0;
    L.m[KIND_SP] = o; o += 8 * CAP;
    L.rho[KIND_SP] = o; o += 8 * CAP;
    L.s0[KIND_SP] = o; o += 8 * CAP;
    L.s1[KIND_SP] = o; o += 8 * CAP;
    L.s2[KIND_SP] = o; o += 8 * CAP;
    L.s3[KIND_SP] = o; o += 8 * CAP;
    L.iv[KIND_SP] = o; o += 8 * CAP;
    L.bytes = o;
    return L;
}

// stress-point receivers <- nodes (v, m, rho, m / rho; x only for the CSPM
// matrix) and, for the derivative, dummies (x, m, rho, wall position and
// orientation).  The fields a launch does not need come last, so a launch
// allocates only the bytes before them (lds_sp_update on the host): the
// update alone stops before the dummies, and without CSPM the nodes' x is
// left out.
constexpr LdsLayout layout_sp_update() {
    LdsLayout L = layout_none();
    int o = 0;
    L.v[KIND_NODE] = o; o += 16 * CAP;
    L.m[KIND_NODE] = o; o += 8 * CAP;
    L.rho[KIND_NODE] = o; o += 8 * CAP;
    L.mr[KIND_NODE] = o; o += 8 * CAP;
    L.x[KIND_DUMMY] = o; o += 16 * CAP_D;
    L.m[KIND_DUMMY] = o; o += 8 * CAP_D;
    L.rho[KIND_DUMMY] = o; o += 8 * CAP_D;
    L.wall = o; o += 4 * CAP_D;
    L.horiz = o; o += 4 * CAP_D;
    L.x[KIND_NODE] = o; o += 16 * CAP;
    L.bytes = o;
    return L;
}

// node receivers' stress gradient <- stress points (m, rho, sigma / rho^2 of
// xx yy xy; x only for the CSPM matrix) and dummies (m, rho).  As above, the
// field a launch may leave out comes last (lds_deriv_nodes on the host).
constexpr LdsLayout layout_deriv_nodes() {
    LdsLayout L = layout_none();
    int o = 0;
    L.m[KIND_SP] = o; o += 8 * CAP;
    L.rho[KIND_SP] = o; o += 8 * CAP;
    L.p0[KIND_SP] = o; o += 8 * CAP;
    L.p1[KIND_SP] = o; o += 8 * CAP;
    L.p2[KIND_SP] = o; o += 8 * CAP;
    L.m[KIND_DUMMY] = o; o += 8 * CAP_D;
    L.rho[KIND_DUMMY] = o; o += 8 * CAP_D;
    L.x[KIND_SP] = o; o += 16 * CAP;
    L.bytes = o;
    return L;
}

// node receivers' artificial viscosity <- nodes (x, v, m, rho; h last, only
// when the smoothing lengths differ: lds_av_nodes on the host)
constexpr LdsLayout layout_av_nodes() {
    LdsLayout L = layout_none();
    int o = 0;
    L.x[KIND_NODE] = o; o += 16 * CAP;
    L.v[KIND_NODE] = o; o += 16 * CAP;
    L.m[KIND_NODE] = o; o += 8 * CAP;
    L.rho[KIND_NODE] = o; o += 8 * CAP;
    L.h[KIND_NODE] = o; o += 8 * CAP;
    L.bytes = o;
    return L;
}

constexpr LdsLayout LAY_COUNT = layout_count();
constexpr LdsLayout LAY_SPU_NODES = layout_spu_nodes();
constexpr LdsLayout LAY_SP_UPDATE = layout_sp_update();
constexpr LdsLayout LAY_DERIV_NODES = layout_deriv_nodes();
constexpr LdsLayout LAY_AV_NODES = layout_av_nodes();

__device__ __forceinline__ constexpr int sel3(int kind, int a, int b, int c) {
    return kind == 0 ? a : (kind == 1 ? b : c);
}

// Window cells of the staged tile: per kind, start | count << 16.
template <const LdsLayout &LY> struct LdsCells {
    int gx0, gy0;   // grid cell of window cell (0, 0)
    __device__ __forceinline__ int2 range(int kind, int gx, int gy) const {
        const int e = lds_at<int>(sel3(kind, LY.tab[0], LY.tab[1], LY.tab[2]))[(gy - gy0) * TCOLS + (gx - gx0)];
        return make_int2(e & 0xffff, e >> 16);
    }
};

// Partner fields of the staged window (j is an LDS index of kind K).
template <const LdsLayout &LY> struct LdsFields {
    __device__ __forceinline__ double2 xk(int kind, int j) const {
        return lds_at<double2>(sel3(kind, LY.x[0], LY.x[1], LY.x[2]))[j];
    }
    __device__ __forceinline__ double hk(int kind, int j) const {
        return lds_at<double>(sel3(kind, LY.h[0], LY.h[1], LY.h[2]))[j];
    }
    template <int K> __device__ __forceinline__ double2 x(KindTag<K>, int j) const {
        return lds_at<double2>(LY.x[K])[j];
    }
    template <int K> __device__ __forceinline__ double h(KindTag<K>, int j) const {
        return lds_at<double>(LY.h[K])[j];
    }
    template <int K> __device__ __forceinline__ double m(KindTag<K>, int j) const {
        return lds_at<double>(LY.m[K])[j];
    }
    template <int K> __device__ __forceinline__ double rho(KindTag<K>, int j) const {
        return lds_at<double>(LY.rho[K])[j];
    }
    template <int K> __device__ __forceinline__ double2 vel(KindTag<K>, int j) const {
        return lds_at<double2>(LY.v[K])[j];
    }
    template <int K> __device__ __forceinline__ double4 st(KindTag<K>, int j) const {
        const double w = LY.s3[K] >= 0 ? lds_at<double>(LY.s3[K])[j] : 0.0;
        return make_double4(lds_at<double>(LY.s0[K])[j], lds_at<double>(LY.s1[K])[j],
                            lds_at<double>(LY.s2[K])[j], w);
    }
    template <int K> __device__ __forceinline__ double iv(KindTag<K>, int j) const {
        return lds_at<double>(LY.iv[K])[j];
    }
    template <int K> __device__ __forceinline__ double mr(KindTag<K>, int j) const {
        return lds_at<double>(LY.mr[K])[j];
    }
    template <int K> __device__ __forceinline__ double3 sr2(KindTag<K>, int j) const {
        return make_double3(lds_at<double>(LY.p0[K])[j], lds_at<double>(LY.p1[K])[j], lds_at<double>(LY.p2[K])[j]);
    }
    template <int K> __device__ __forceinline__ float wall(KindTag<K>, int j) const {
        return lds_at<float>(LY.wall)[j];
    }
    template <int K> __device__ __forceinline__ float horiz(KindTag<K>, int j) const {
        return lds_at<float>(LY.horiz)[j];
    }
};

// Pass classification of a lane that may hold pairs on both sides of M.
struct NoClass {
    __device__ __forceinline__ bool operator()(int, bool) const { return true; }
};
struct SplitClass {
    bool on;
    int pr;
    const int *proc;
    const unsigned *cnt, *off;
    Split sp;
    __device__ __forceinline__ bool operator()(int j, bool rev) const {
        return !on || pair_reversed(pr, proc[j], cnt, off, sp) == rev;
    }
};

// One run of a receiver's walk: the partners of kind K in cell (gx, gy),
// slots ascending (forward) or descending (reverse).  For each accepted pair
// -- the acceptance test and dx = x_r - x_s of grid_find_NEW (2_...:1345) --
// f(kind_tag, fields, j, w, dW/dx_r, dW/dy_r) with kernel()'s values rounded
// to single precision as the reference's pair list stores them.
template <int KMASK, int K, bool REV, typename Cells, typename Fields, typename Cls, typename F>
__device__ __forceinline__ void walk_run(const Cells &Cs, const Fields &A, int gx, int gy, double2 xr,
                                         double hr, int kr, int self, double pi, KernelFactor &kf,
                                         const Cls &cls, F &f) {
    if constexpr (((KMASK >> K) & 1) != 0) {
        const int2 rg = Cs.range(K, gx, gy);
        for (int q = 0; q < rg.y; ++q) {
            const int j = REV ? rg.x + rg.y - 1 - q : rg.x + q;
            if (K == kr && j == self) continue;
            const double2 xs = A.xk(K, j);
            const double dx0 = xr.x - xs.x, dx1 = xr.y - xs.y;
            double mh;
            const double rr = pair_distance(dx0, dx1, hr, A.hk(K, j), mh);
            if (rr < 0.0) continue;
            if (!cls(j, REV)) continue;
            float w, gwx, gwy;
            cubic_kernel_f(rr, dx0, dx1, mh, kf.get(mh, pi), w, gwx, gwy);
            f(KindTag<K>{}, A, j, w, gwx, gwy);
        }
    }
}

// Every pair of the receiver in cell (cx, cy) with a partner of a kind in
// KMASK, in the order the reference's linked list holds them: the reverse
// stencil walk (cells descending row-major, kinds and slots descending) for
// pairs with d >= M, then the forward walk (row-major cells, kinds, original
// index) for pairs with d < M.
template <int KMASK, typename Cells, typename Fields, typename Cls, typename F>
__device__ __forceinline__ void walk_stencil(const Cells &Cs, const Fields &A, const Grid &g, int cx, int cy,
                                             double2 xr, double hr, int kr, int self, bool do_rev,
                                             bool do_fwd, double pi, const Cls &cls, F &&f) {
    KernelFactor kf;
    const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx + 1 < g.ndivx ? cx + 1 : g.ndivx - 1;
    const int y0 = cy > 0 ? cy - 1 : 0, y1 = cy + 1 < g.ndivy ? cy + 1 : g.ndivy - 1;
    if (do_rev)
        for (int y = y1; y >= y0; --y)
            for (int xx = x1; xx >= x0; --xx) {
                walk_run<KMASK, 2, true>(Cs, A, xx, y, xr, hr, kr, self, pi, kf, cls, f);
                walk_run<KMASK, 1, true>(Cs, A, xx, y, xr, hr, kr, self, pi, kf, cls, f);
                walk_run<KMASK, 0, true>(Cs, A, xx, y, xr, hr, kr, self, pi, kf, cls, f);
            }
    if (do_fwd)
        for (int y = y0; y <= y1; ++y)
            for (int xx = x0; xx <= x1; ++xx) {
                walk_run<KMASK, 0, false>(Cs, A, xx, y, xr, hr, kr, self, pi, kf, cls, f);
                walk_run<KMASK, 1, false>(Cs, A, xx, y, xr, hr, kr, self, pi, kf, cls, f);
                walk_run<KMASK, 2, false>(Cs, A, xx, y, xr, hr, kr, self, pi, kf, cls, f);
            }
}

// One workgroup's receivers and window.
struct Tile {
    int first, end, last;   // receivers [first, end); [first, last] are in the domain
    int x0, w, wy0, nrows;  // window: strip columns [x0, x0 + w) +- 1, rows [wy0, wy0 + nrows)
    bool fits;              // one strip, at most TROWS rows
};

__device__ __forceinline__ Tile make_tile(int base, int count, int kr, const SearchView &V, const WalkCtx &C) {
    Tile T;
    T.first = base + tile_index() * TILE;
    T.end = min(T.first + TILE, base + count);
    T.last = min(T.end, V.kstart[kr * C.L + C.g.ndivt]) - 1;
    T.fits = false;
    T.x0 = T.w = T.wy0 = T.nrows = 0;
    if (T.last < T.first) return T;
    int cxf, cyf, cxl, cyl;
    cell_xy(V.cell[T.first], C.g, cxf, cyf);
    cell_xy(V.cell[T.last], C.g, cxl, cyl);
    T.x0 = (cxf / STRIP) * STRIP;
    T.w = C.g.ndivx - T.x0 < STRIP ? C.g.ndivx - T.x0 : STRIP;
    T.wy0 = cyf > 0 ? cyf - 1 : 0;
    const int wy1 = cyl + 1 < C.g.ndivy ? cyl + 1 : C.g.ndivy - 1;
    T.nrows = wy1 - T.wy0 + 1;
    T.fits = (cxl / STRIP) * STRIP == T.x0 && T.nrows <= TROWS;
    return T;
}

// Staging plan of one kind's window, in LDS: the strip's main run and the
// halo cells (left column rows, then right column rows).
struct StagePlan {
    int hglo[2 * TROWS], hpos[2 * TROWS];
    int nh, total, mainlo, mainn;
};

// Plan the window of kind K on wave `wave` of the block (the other waves
// return at once): main run, halo cell runs and their LDS positions, and the
// halo entries of the cell table when tab is given.  The caller puts a
// barrier between planning and copying.
template <int K>
__device__ __forceinline__ void plan_kind(const Tile &T, const SearchView &V, const WalkCtx &C, int wave,
                                          int *__restrict__ tab, StagePlan &sc) {
    if ((int)(threadIdx.x >> 6) != wave) return;
    const int t = threadIdx.x & 63;
    const int *ks = V.kstart + K * C.L;
    const int kbase = T.x0 * C.g.ndivy;
    int n = 0, lo = 0;
    const int side = t / T.nrows, ly = t - side * T.nrows;
    if (t < 2 * T.nrows) {
        const int gx = side ? T.x0 + T.w : T.x0 - 1;
        if (gx >= 0 && gx < C.g.ndivx) {
            const int key = cell_key(gx, T.wy0 + ly, C.g);
            lo = ks[key];
            n = ks[key + 1] - lo;
        }
    }
    const int mainlo = ks[kbase + T.wy0 * T.w];
    const int mainn = ks[kbase + (T.wy0 + T.nrows) * T.w] - mainlo;
    int incl = n;
    for (int o = 1; o < 64; o <<= 1) {
        int v = __shfl_up(incl, o, 64);
        if (t >= o) incl += v;
    }
    if (t < 2 * T.nrows) {
        sc.hglo[t] = lo;
        sc.hpos[t] = mainn + incl - n;
        if (tab) tab[ly * TCOLS + (side ? T.w + 1 : 0)] = (mainn + incl - n) | (n << 16);
    }
    if (t == 63) {
        sc.nh = 2 * T.nrows;
        sc.total = mainn + incl;
        sc.mainlo = mainlo;
        sc.mainn = mainn;
    }
}

constexpr int NPER = (CAP + TILE - 1) / TILE;   // window partners per thread, at most

// Copy the planned window of kind K (sc.total <= NPER * TILE) into LDS:
// fetch(slot) -> item, put(lds_index, item), and the strip entries of the
// cell table when tab is given.  Every thread issues all its loads before
// its first store, so the copy costs one memory latency.  The caller ends
// the staging with a barrier.
template <int K, typename Fetch, typename Put>
__device__ __forceinline__ void copy_kind(const Tile &T, const SearchView &V, const WalkCtx &C,
                                          int *__restrict__ tab, const StagePlan &sc, Fetch &&fetch, Put &&put) {
    const int t = threadIdx.x;
    const int *ks = V.kstart + K * C.L;
    const int total = sc.total, mainlo = sc.mainlo, mainn = sc.mainn;
    const bool has_tab = tab != nullptr && t < T.nrows * T.w;
    int tlo = 0, thi = 0, tpos = 0;
    if (has_tab) {
        const int ly = t / T.w, lx = t - ly * T.w;
        const int key = T.x0 * C.g.ndivy + (T.wy0 + ly) * T.w + lx;
        tlo = ks[key];
        thi = ks[key + 1];
        tpos = ly * TCOLS + lx + 1;
    }
    decltype(fetch(0)) it[NPER];
#pragma unroll
    for (int k = 0; k < NPER; ++k) {
        const int i = t + k * TILE;
        if (i < mainn) {
            it[k] = fetch(mainlo + i);
        } else if (i < total) {   // halo: the last cell starting at or before i
            int lo = 0, hi = sc.nh - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (sc.hpos[mid] <= i) lo = mid;
                else hi = mid - 1;
            }
            it[k] = fetch(sc.hglo[lo] + (i - sc.hpos[lo]));
        }
    }
    if (has_tab) tab[tpos] = (tlo - mainlo) | ((thi - tlo) << 16);
#pragma unroll
    for (int k = 0; k < NPER; ++k)
        if (t + k * TILE < total) put(t + k * TILE, it[k]);
}

// ------------------------------------------------------- list reuse check
// Largest squared displacement of the moving particles [0, n) since the lists
// were built, as the bits of a non-negative double (which order as the
// doubles do): one atomic per block.
__global__ void k_max_disp(const double2 *__restrict__ x, const double2 *__restrict__ xb, int n,
                           unsigned long long *__restrict__ out) {
    __shared__ double s[BLOCK];
    double m = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double2 a = x[i], b = xb[i];
        const double d0 = a.x - b.x, d1 = a.y - b.y;
        m = fmax(m, d0 * d0 + d1 * d1);
    }
    const int t = threadIdx.x;
    s[t] = m;
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (t < w) s[t] = fmax(s[t], s[t + w]);
        __syncthreads();
    }
    if (t == 0) atomicMax(out, (unsigned long long)__double_as_longlong(s[0]));
}

// ------------------------------------------------------------ step kernels
struct Work {
    double2 *vx0, *x0, *vel0, *rk_vel, *rhs2, *f_bound, *art_visc;
    double4 *stress0, *rk_stress, *rhs1, *grad_u;
    double *rk_dev, *ddev;
};

// Start of a step and its first RK stage in one pass: vx0, x0 (2_...:114),
// the RK4 prologue (2_...:681-690: stage-0 state, accumulators and the
// velocities and stresses zeroed, dummies included) and the stage-0 state
// with its first adapt_stress2 (2_...:700-701, 715).  The stage-0 update
// v0 + 0 * rhs of the reference, whose right-hand sides are zero then, is
// v0 + 0.0 (a zero of either sign becomes +0, as there); node stresses start
// from the zeroed ones.  The right-hand sides themselves need no clearing:
// the stage tails write them before any later stage reads them.  vx0 and
// vel0 are the same velocities: Work.vx0 may alias vel0.
__global__ void k_stage_first(StateBufs S, Work W, int adapt, double tanfi, double coh, int nn, int nt,
                              int nt2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt2) return;
    if (i >= nt) {
        S.vel[i] = make_double2(0, 0);
        S.stress[i] = make_double4(0, 0, 0, 0);
        return;
    }
    const double2 v = S.vel[i];
    W.vel0[i] = v;
    if (W.vx0 != W.vel0) W.vx0[i] = v;
    W.x0[i] = S.x[i];
    W.rk_dev[i] = 0.0;
    material_tc(S, i, tanfi, coh);
    if (i < nn) {
        W.rk_vel[i] = make_double2(0, 0);
        S.vel[i] = make_double2(v.x + 0.0, v.y + 0.0);
        double4 s = make_double4(0, 0, 0, 0);
        if (adapt) adapt_stress(s, tanfi, coh);
        S.stress[i] = s;
    } else {
        const double4 s0 = S.stress[i];
        W.stress0[i] = s0;
        W.rk_stress[i] = make_double4(0, 0, 0, 0);
        S.vel[i] = make_double2(0, 0);
        double4 s = make_double4(s0.x + 0.0, s0.y + 0.0, s0.z + 0.0, s0.w + 0.0);
        if (adapt) adapt_stress(s, tanfi, coh);
        S.stress[i] = s;
    }
}

// Stage state of stages 2-4 (2_...:700-701) and the first adapt_stress2 of
// the stage.
__global__ void k_stage_begin(StateBufs S, Work W, double c1, int adapt, double tanfi, double coh,
                              int nn, int nt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    material_tc(S, i, tanfi, coh);
    if (i < nn) {
        double2 v0 = W.vel0[i], r2 = W.rhs2[i];
        S.vel[i] = make_double2(v0.x + c1 * r2.x, v0.y + c1 * r2.y);
        if (adapt) {
            double4 s = S.stress[i];
            adapt_stress(s, tanfi, coh);
            S.stress[i] = s;
        }
    } else {
        double4 s0 = W.stress0[i], r1 = W.rhs1[i];
        double4 s = make_double4(s0.x + c1 * r1.x, s0.y + c1 * r1.y, s0.z + c1 * r1.z, s0.w + c1 * r1.w);
        if (adapt) adapt_stress(s, tanfi, coh);
        S.stress[i] = s;
    }
}

// ------------------------------------------------------------ pair lists
// The pairs the stage kernels need, found once per step by k_count_lists:
// for every receiver of a tile whose window fits, its accepted partners in
// forward walk order as 16-bit entries kind << 14 | window index (the window
// laid out exactly as the stage kernels stage it), with w and grad w rounded
// to single precision as the reference's pair list stores them
// (2_...:1376-1378).  The reverse walk order is the forward list read
// backwards.  List A: a node's stress-point and dummy partners, a stress
// point's node and dummy partners; list B: a node's node partners.  Entries
// come four to a word, words of a tile lane-interleaved, so a wave reads its
// lanes' next four entries with one coalesced load per array.
constexpr int LMAX = 48;            // entries per receiver and list; more marks the tile unlisted
constexpr int LGROUPS = LMAX / 4;

struct PairLists {
    unsigned long long *ia, *ib;    // [tile][LGROUPS][TILE]: four entries per word
    float4 *wa, *gxa, *gya;         // list A: w, dW/dx_r, dW/dy_r of those entries
    float4 *gxb, *gyb;              // list B (nodes): dW/dx_r, dW/dy_r
    float4 *wb;                     // list B: w (XSPH only; null otherwise, and then stress
                                    // points have no list B)
    unsigned short *deg_a, *deg_b;  // per slot: nodes and stress points / nodes
    unsigned char *ok;              // per tile (node tiles, then SP tiles): lists usable
    int ntile_n;                    // node tiles
    // each tile's extent and (when it fits) the staging plan of every partner
    // kind, as the search computed them: the stage kernels load them in one
    // round trip instead of deriving them again through dependent loads
    Tile *tiles;                    // per tile
    StagePlan *plans;               // per tile and kind
    // per node slot (the outside approach; null otherwise): 2 * its node
    // partners + 1 if it has a wall-dummy partner, i.e. n_int and bc_int of
    // the reference (2_...:579, 870-871), counted within 2h at the search
    int *nodeinfo;
};

// Wave `wave` of the block copies the search's plan of kind K for tile
// ltile into the shared plan (the other waves return at once); the caller
// puts a barrier after it, as after plan_kind.
template <int K>
__device__ __forceinline__ void load_plan(const PairLists &PL, int ltile, int wave, StagePlan &sc) {
    if ((int)(threadIdx.x >> 6) != wave) return;
    const int *src = reinterpret_cast<const int *>(PL.plans + (size_t)ltile * 3 + K);
    int *dst = reinterpret_cast<int *>(&sc);
    for (int i = (int)(threadIdx.x & 63); i < (int)(sizeof(StagePlan) / sizeof(int)); i += 64) dst[i] = src[i];
}

__device__ __forceinline__ size_t word_at(int tile, int g) {
    return ((size_t)tile * LGROUPS + g) * TILE + threadIdx.x;
}

__device__ __forceinline__ void set4(float4 &v, int k, float a) {
    if (k == 0) v.x = a;
    else if (k == 1) v.y = a;
    else if (k == 2) v.z = a;
    else v.w = a;
}

__device__ __forceinline__ float get4(const float4 &v, int k) {
    return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w));
}

// One lane's list being written: entries gathered four to a word in a
// register, each full word stored.
struct IdxWriter {
    unsigned long long idx = 0;
    int n = 0;
    __device__ __forceinline__ void put(int tile, unsigned long long *ip, unsigned e) {
        const int k = n & 3;
        if (k == 0) idx = 0;
        idx |= (unsigned long long)e << (16 * k);
        if (k == 3 && n < LMAX) ip[word_at(tile, n >> 2)] = idx;
        ++n;
    }
    __device__ __forceinline__ void finish(int tile, unsigned long long *ip) {
        if ((n & 3) != 0 && n < LMAX) ip[word_at(tile, n >> 2)] = idx;
    }
};

// Kernel values of a lane's n list entries (n <= LMAX), computed a word at a
// time with the lanes of a wave in step, so that the stores are whole
// coalesced words: kernel() of the receiver-side dx, as the stage kernels'
// global walk computes it.
template <typename Fields>
__device__ __forceinline__ void fill_kernel_values(const unsigned long long *ip, float4 *wp, float4 *gxp,
                                                   float4 *gyp, int tile, int n, const Fields &A, double2 xr,
                                                   double hr, double h_uniform, double pi, KernelFactor &kf) {
    const int ng = (n + 3) >> 2;
    // the lane's own words, written by its walk: the next one is loaded
    // while the current one's kernel values are computed
    unsigned long long next = ng > 0 ? ip[word_at(tile, 0)] : 0ull;
    for (int g = 0; g < ng; ++g) {
        const size_t o = word_at(tile, g);
        const unsigned long long word = next;
        if (g + 1 < ng) next = ip[word_at(tile, g + 1)];
        float wv[4] = {0.f, 0.f, 0.f, 0.f}, gxv[4] = {0.f, 0.f, 0.f, 0.f}, gyv[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (4 * g + k < n) {
                const int ent = (int)((word >> (16 * k)) & 0xffff);
                const int kind = ent >> 14, j = ent & 0x3fff;
                const double2 xs = A.xk(kind, j);
                const double dx0 = xr.x - xs.x, dx1 = xr.y - xs.y;
                double mh;
                const double rr = pair_distance(dx0, dx1, hr, h_uniform > 0 ? h_uniform : A.hk(kind, j), mh);
                cubic_kernel_f(rr, dx0, dx1, mh, kf.get(mh, pi), wv[k], gxv[k], gyv[k]);
            }
        }
        if (wp) wp[o] = make_float4(wv[0], wv[1], wv[2], wv[3]);
        if (gxp) gxp[o] = make_float4(gxv[0], gxv[1], gxv[2], gxv[3]);
        if (gyp) gyp[o] = make_float4(gyv[0], gyv[1], gyv[2], gyv[3]);
    }
}

// Positions and smoothing lengths from the global state (j is a slot).
struct XHFields {
    const double2 *x;
    const double *h;
    __device__ __forceinline__ double2 xk(int, int j) const { return x[j]; }
    __device__ __forceinline__ double hk(int, int j) const { return h[j]; }
};

__device__ __forceinline__ bool in_list_a(int kr, int ks) {
    return kr == KIND_NODE ? ks != KIND_NODE : ks != KIND_SP;
}

// Counting of k_count for the node (KR = 0) or stress-point (KR = 1)
// receivers [base, base + count), tiled as the stage kernels are, and the
// tile's pair lists; with FILL, also the lists' kernel values (the
// mixed-precision stage kernels compute them as they go).
template <int KR, bool FILL>
__global__ __launch_bounds__(TILE) void k_count_lists(SearchView V, const double2 *__restrict__ x,
                                                      const double *__restrict__ hsml,
                                                      unsigned *__restrict__ cnt, Split *__restrict__ split,
                                                      PairLists PL, double h_uniform, int base, int count,
                                                      double skin) {
    __shared__ StagePlan pl[3];
    constexpr const LdsLayout &LY = LAY_COUNT;
    const WalkCtx C = walk_ctx(V);
    const Tile T = make_tile(base, count, KR, V, C);
    const int tile = tile_index();
    const int ltile = KR == KIND_NODE ? tile : PL.ntile_n + tile;
    if (threadIdx.x == 0) PL.tiles[ltile] = T;
    if (T.first >= T.end) return;
    const int r = T.first + (int)threadIdx.x;
    // list B: a node's node partners; with XSPH also a stress point's stress-point partners
    const bool list_b = KR == KIND_NODE || PL.wb != nullptr;
    int self_base = 0;
    bool lds = T.fits;
    if (lds) {
        // kind k is planned by wave k mod the tile's waves (two waves: the
        // first plans nodes and dummies)
        plan_kind<KIND_NODE>(T, V, C, 0 % TWAVES, lds_at<int>(LY.tab[KIND_NODE]), pl[KIND_NODE]);
        plan_kind<KIND_SP>(T, V, C, 1 % TWAVES, lds_at<int>(LY.tab[KIND_SP]), pl[KIND_SP]);
        plan_kind<KIND_DUMMY>(T, V, C, 2 % TWAVES, lds_at<int>(LY.tab[KIND_DUMMY]), pl[KIND_DUMMY]);
        __syncthreads();
        // keep the plans for the stage kernels: the same wave stores each
        for (int k = (int)(threadIdx.x >> 6); k < 3; k += TWAVES) {
            const int *src = reinterpret_cast<const int *>(&pl[k]);
            int *dst = reinterpret_cast<int *>(PL.plans + (size_t)ltile * 3 + k);
            for (int i = (int)(threadIdx.x & 63); i < (int)(sizeof(StagePlan) / sizeof(int)); i += 64)
                dst[i] = src[i];
        }
        lds = pl[KIND_NODE].total <= CAP && pl[KIND_SP].total <= CAP && pl[KIND_DUMMY].total <= CAP_D;
        self_base = pl[KR].mainlo;
    }
    if (lds) {
        const bool stage_h = !(h_uniform > 0);   // the launch allocates h only then
        auto fetch = [&](int g) { return make_double3(x[g].x, x[g].y, stage_h ? hsml[g] : 0.0); };
        auto put = [&](auto kt) {
            constexpr int K = decltype(kt)::value;
            return [&](int i, const double3 &v) {
                lds_at<double2>(LAY_COUNT.x[K])[i] = make_double2(v.x, v.y);
                if (stage_h) lds_at<double>(LAY_COUNT.h[K])[i] = v.z;
            };
        };
        copy_kind<KIND_NODE>(T, V, C, lds_at<int>(LY.tab[KIND_NODE]), pl[KIND_NODE], fetch, put(KindTag<KIND_NODE>{}));
        copy_kind<KIND_SP>(T, V, C, lds_at<int>(LY.tab[KIND_SP]), pl[KIND_SP], fetch, put(KindTag<KIND_SP>{}));
        copy_kind<KIND_DUMMY>(T, V, C, lds_at<int>(LY.tab[KIND_DUMMY]), pl[KIND_DUMMY], fetch,
                              put(KindTag<KIND_DUMMY>{}));
        __syncthreads();
    }
    unsigned fk[3] = {0, 0, 0};   // pairs this receiver discovers, by partner kind
    unsigned n_node = 0;          // node receivers: node partners (all, not only discovered)
    bool any_dummy = false;       // node receivers: a wall-dummy partner
    IdxWriter la, lb;
    const bool active = r < T.end && V.cell[r] < C.g.ndivt;
    if (active) {
        int cx, cy;
        cell_xy(V.cell[r], C.g, cx, cy);
        const int crm_r = cy * C.g.ndivx + cx;
        const double2 xr = x[r];
        const double hr = hsml[r];
        const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx + 1 < C.g.ndivx ? cx + 1 : C.g.ndivx - 1;
        const int y0 = cy > 0 ? cy - 1 : 0, y1 = cy + 1 < C.g.ndivy ? cy + 1 : C.g.ndivy - 1;
        const int self = lds ? r - self_base : r;
        // h_uniform > 0: every particle has this smoothing length
        const AcceptBounds ub = accept_bounds(h_uniform, h_uniform);
        auto walk = [&](const auto &Cs, const auto &A, bool lists, auto uni) __attribute__((always_inline)) {
            for (int y = y0; y <= y1; ++y)
                for (int xx = x0; xx <= x1; ++xx) {
                    const int crm = y * C.g.ndivx + xx;
                    const bool cell_after = crm > crm_r, same_cell = crm == crm_r;
#pragma unroll
                    for (int ks = 0; ks < 3; ++ks) {
                        const int2 rg = Cs.range(ks, xx, y);
                        // discovered by this receiver: a later cell, or this cell and a later
                        // kind, or this kind and a later slot (slots ascend with the window
                        // index inside a cell and kind)
                        const bool kind_after = cell_after || (same_cell && ks > KR);
                        const bool slot_after = same_cell && ks == KR;
                        // a stress point's stress-point partners enter no list (no XSPH): only
                        // those it discovers count, so the cells before its own, and in its
                        // own cell the slots up to its own, are not tested
                        const bool count_only = KR == KIND_SP && ks == KIND_SP && !list_b;
                        if (count_only && !cell_after && !same_cell) continue;
                        const int jbeg = count_only && same_cell && self + 1 > rg.x ? self + 1 : rg.x;
                        const int jend = rg.x + rg.y;
                        double2 xn = jbeg < jend ? A.xk(ks, jbeg) : make_double2(0.0, 0.0);
                        for (int j = jbeg; j < jend; ++j) {
                            const double2 xs = xn;
                            if (j + 1 < jend) xn = A.xk(ks, j + 1);   // the next position in flight
                            if (ks == KR && j == self) continue;
                            if (!FILL && skin > 0) {
                                // lists kept over several steps: also the partners within the
                                // skin; only the pairs themselves are counted
                                const double dx0 = xr.x - xs.x, dx1 = xr.y - xs.y;
                                const double hs = decltype(uni)::value ? h_uniform : A.hk(ks, j);
                                const double ls = (hr + hs) + skin;
                                if (dx0 * dx0 + dx1 * dx1 >= ls * ls) continue;
                                const bool acc = decltype(uni)::value ? accepted_by(dx0, dx1, ub)
                                                                      : accepted_by(dx0, dx1, accept_bounds(hr, hs));
                                if (acc) fk[ks] += (kind_after || (slot_after && j > self)) ? 1u : 0u;
                                if constexpr (KR == KIND_NODE) {
                                    n_node += (acc && ks == KIND_NODE) ? 1u : 0u;
                                    any_dummy = any_dummy || (acc && ks == KIND_DUMMY);
                                }
                            } else {
                                if constexpr (decltype(uni)::value) {
                                    if (!accepted_by(xr.x - xs.x, xr.y - xs.y, ub)) continue;
                                } else {
                                    if (!accepted_by(xr.x - xs.x, xr.y - xs.y, accept_bounds(hr, A.hk(ks, j))))
                                        continue;
                                }
                                fk[ks] += (kind_after || (slot_after && j > self)) ? 1u : 0u;
                                if constexpr (KR == KIND_NODE) {
                                    n_node += ks == KIND_NODE ? 1u : 0u;
                                    any_dummy = any_dummy || ks == KIND_DUMMY;
                                }
                            }
                            if (!lists) continue;
                            const unsigned e = (unsigned)((ks << 14) | j);
                            if (in_list_a(KR, ks))
                                la.put(ltile, PL.ia, e);
                            else if (ks == KR && list_b)
                                lb.put(ltile, PL.ib, e);
                        }
                    }
                }
        };
        using Uni = std::integral_constant<bool, true>;
        using Gen = std::integral_constant<bool, false>;
        if (lds) {
            if (h_uniform > 0)
                walk(LdsCells<LY>{T.x0 - 1, T.wy0}, LdsFields<LY>{}, true, Uni{});
            else
                walk(LdsCells<LY>{T.x0 - 1, T.wy0}, LdsFields<LY>{}, true, Gen{});
        } else {
            walk(GlobalCells{V.kstart, C.g, C.L}, XHFields{x, hsml}, false, Gen{});
        }
        if (lds) {
            la.finish(ltile, PL.ia);
            if (list_b) lb.finish(ltile, PL.ib);
        }
        cnt[V.proc[r]] = fk[0] + fk[1] + fk[2];
    }
    // per-type histogram (one pint_type per partner kind of a receiver kind):
    // wave reduction, then one atomic per wave
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        unsigned v = fk[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if ((threadIdx.x & 63) == 0 && v) atomicAdd(&split->n_by_type[pair_type(KR, k)], (unsigned long long)v);
    }
    const bool over = la.n > LMAX || lb.n > LMAX;
    const bool any_over = lds ? __syncthreads_or(over) : true;
    if (threadIdx.x == 0) PL.ok[ltile] = !any_over;
    if (FILL && !any_over && active) {
        const double2 xr = x[r];
        const double hr = hsml[r];
        KernelFactor kf;
        fill_kernel_values(PL.ia, PL.wa, PL.gxa, PL.gya, ltile, la.n, LdsFields<LY>{}, xr, hr, h_uniform, V.pi, kf);
        if (list_b)
            fill_kernel_values(PL.ib, PL.wb, KR == KIND_NODE ? PL.gxb : nullptr, KR == KIND_NODE ? PL.gyb : nullptr,
                               ltile, lb.n, LdsFields<LY>{}, xr, hr, h_uniform, V.pi, kf);
    }
    if (r < T.end) {
        PL.deg_a[r] = (unsigned short)(la.n < LMAX ? la.n : LMAX);
        if (list_b) PL.deg_b[r] = (unsigned short)(lb.n < LMAX ? lb.n : LMAX);
        if (KR == KIND_NODE && PL.nodeinfo) PL.nodeinfo[r] = (int)(2u * n_node + (any_dummy ? 1u : 0u));
    }
}

// The ne entries of one list word, forwards or backwards; every word index
// is a compile-time constant.
template <int KMASK, bool NW, bool NG, bool REV, typename Fields, typename F>
__device__ __forceinline__ void walk_word(unsigned long long word, const float4 &cw, const float4 &cx,
                                          const float4 &cy, int ne, const Fields &A, F &f) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int kk = REV ? 3 - q : q;
        if (kk >= ne) continue;
        const int ent = (int)((word >> (16 * kk)) & 0xffff);
        const int kind = ent >> 14, j = ent & 0x3fff;
        if (!((KMASK >> kind) & 1)) continue;
        const float w = NW ? get4(cw, kk) : 0.0f;
        const float gx = NG ? get4(cx, kk) : 0.0f, gy = NG ? get4(cy, kk) : 0.0f;
        if constexpr ((KMASK & MASK_NODE) != 0)
            if (kind == KIND_NODE) {
                f(KindTag<KIND_NODE>{}, A, j, w, gx, gy);
                continue;
            }
        if constexpr ((KMASK & MASK_SP) != 0)
            if (kind == KIND_SP) {
                f(KindTag<KIND_SP>{}, A, j, w, gx, gy);
                continue;
            }
        if constexpr ((KMASK & MASK_DUMMY) != 0) f(KindTag<KIND_DUMMY>{}, A, j, w, gx, gy);
    }
}

// One list word and its kernel values as loaded.
struct ListWord {
    unsigned long long idx;
    float4 w, gx, gy;
};

// A receiver's list walk with DEPTH words in flight (2, or 1 where registers
// are short): start() issues the loads of the first words (callers start
// before the window staging so that these loads overlap it), run() calls
// f(kind_tag, fields, j, w, dW/dx_r, dW/dy_r) over the entries of the kinds
// in KMASK, forwards or backwards (rev is the same for every lane).  NW / NG:
// the kernel reads w / the gradient.
template <int KMASK, bool NW, bool NG, int DEPTH = 2> struct ListWalk {
    const unsigned long long *ip;
    const float4 *wp, *gxp, *gyp;
    int tile, deg, ng;
    bool rev;
    ListWord a, b;

    __device__ __forceinline__ void load(ListWord &d, int g) const {
        const size_t o = word_at(tile, rev ? ng - 1 - g : g);
        d.idx = ip[o];
        if (NW) d.w = wp[o];
        if (NG) {
            d.gx = gxp[o];
            d.gy = gyp[o];
        }
    }
    __device__ __forceinline__ void start(const unsigned long long *ip_, const float4 *wp_, const float4 *gxp_,
                                          const float4 *gyp_, int tile_, int deg_, bool rev_) {
        ip = ip_;
        wp = wp_;
        gxp = gxp_;
        gyp = gyp_;
        tile = tile_;
        deg = deg_;
        rev = rev_;
        ng = (deg + 3) >> 2;
        const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
        a = b = ListWord{0ull, z, z, z};
        if (ng > 0) load(a, 0);
        if (DEPTH > 1 && ng > 1) load(b, 1);
    }
    template <typename Fields, typename F> __device__ __forceinline__ void run(const Fields &A, F &&f) {
        for (int gi = 0; gi < ng; ++gi) {
            const ListWord c = a;
            if constexpr (DEPTH > 1) {
                a = b;
                if (gi + 2 < ng) load(b, gi + 2);
            } else if (gi + 1 < ng) {
                load(a, gi + 1);
            }
            const int g = rev ? ng - 1 - gi : gi;
            const int ne = deg - 4 * g < 4 ? deg - 4 * g : 4;
            if (rev)
                walk_word<KMASK, NW, NG, true>(c.idx, c.w, c.gx, c.gy, ne, A, f);
            else
                walk_word<KMASK, NW, NG, false>(c.idx, c.w, c.gx, c.gy, ne, A, f);
        }
    }
};

// Workgroup-wide tiling decision of a tiled stage kernel: the tile fits the
// window layout and no receiver of it needs per-pair pass classification.
// check is this lane's need for that classification (global path).
__device__ __forceinline__ bool tile_uses_lds(const Tile &T, const SearchView &V, const WalkCtx &C, int r,
                                              bool &check) {
    check = C.mixed && r <= T.last && V.proc[r] >= C.sp.first_rev;
    const bool any = C.mixed ? __syncthreads_or(check) : false;
    return T.fits && !any;
}

// stress_point_update (2_...:403-482) for node receivers: CSPM-normalised
// stress and plastic strain from the stress points, then adapt_stress2.
__global__ __launch_bounds__(TILE) void k_spu_nodes(StateBufs S, SearchView V, PairLists PL, int adapt,
                                                    double tanfi, double coh, int nn) {
    __shared__ StagePlan pl;
    constexpr const LdsLayout &LY = LAY_SPU_NODES;
    const WalkCtx C = walk_ctx(V);
    const int tile = tile_index();
    const Tile T = PL.tiles[tile];
    if (T.first >= T.end) return;
    const int r = T.first + (int)threadIdx.x;
    const bool inside = r < T.end && V.cell[r] < C.g.ndivt;
    ListWalk<MASK_SP, true, false> lw;
    lw.start(PL.ia, PL.wa, PL.gxa, PL.gya, tile, inside ? PL.deg_a[r] : 0, !C.any_fwd && C.any_rev);
    bool check;
    bool lds = tile_uses_lds(T, V, C, r, check) && PL.ok[tile];
    if (lds) {
        load_plan<KIND_SP>(PL, tile, 0, pl);
        __syncthreads();
        lds = pl.total <= CAP;
    }
    if (lds) {
        struct Item {
            double m, rho, iv;
            double4 st;
        };
        copy_kind<KIND_SP>(T, V, C, nullptr, pl, [&](int g) { return Item{S.mass[g], S.rho[g], S.ivar1[g], S.stress[g]}; },
                           [&](int i, const Item &v) {
                               lds_at<double>(LY.m[KIND_SP])[i] = v.m;
                               lds_at<double>(LY.rho[KIND_SP])[i] = v.rho;
                               lds_at<double>(LY.s0[KIND_SP])[i] = v.st.x;
                               lds_at<double>(LY.s1[KIND_SP])[i] = v.st.y;
                               lds_at<double>(LY.s2[KIND_SP])[i] = v.st.z;
                               lds_at<double>(LY.s3[KIND_SP])[i] = v.st.w;
                               lds_at<double>(LY.iv[KIND_SP])[i] = v.iv;
                           });
        __syncthreads();
    }
    if (r >= T.end) return;
    double t0 = 0, t1 = 0, t2 = 0, t3 = 0, tv = 0, norm = 0;
    auto body = [&](auto kt, const auto &A, int j, float wf, float, float) __attribute__((always_inline)) {
        double w = (double)wf;
        double ms = A.m(kt, j), rs = A.rho(kt, j);
        double h1 = (ms / rs) * w;
        double4 st = A.st(kt, j);
        t0 = t0 + st.x * h1;
        t1 = t1 + st.y * h1;
        t2 = t2 + st.z * h1;
        t3 = t3 + st.w * h1;
        tv = tv + A.iv(kt, j) * h1;
        norm = norm + (w * ms) / rs;
    };
    if (inside) {
        if (lds) {
            lw.run(LdsFields<LY>{}, body);
        } else {
            int cx, cy;
            cell_xy(V.cell[r], C.g, cx, cy);
            const SplitClass cls{check, check ? V.proc[r] : 0, V.proc, V.cnt, V.off, C.sp};
            walk_stencil<MASK_SP>(GlobalCells{V.kstart, C.g, C.L}, GlobalFields{S}, C.g, cx, cy, S.x[r],
                                  S.hsml[r], KIND_NODE, r, C.any_fwd ? check : C.any_rev, C.any_fwd, V.pi, cls,
                                  body);
        }
    }
    double4 s = S.stress[r];
    if (norm != 0) {
        s = make_double4(t0 / norm, t1 / norm, t2 / norm, t3 / norm);
        S.ivar1[r] = tv / norm;
    } else {
        S.vel[r] = make_double2(0, 0);
    }
    material_tc(S, r, tanfi, coh);
    if (adapt) adapt_stress(s, tanfi, coh);
    S.stress[r] = s;
}

// The end of get_derivatives for stress point r (CSPM correction, div1,
// grad_u; 2_...:596-644) from its velocity-gradient sums and CSPM moments,
// then plastic_terms, the Jaumann spin, the stress right-hand side and the RK
// accumulation of the stage (2_...:730-777).  st is its updated stress.
__device__ __forceinline__ void sp_stage_tail(const DevParams &P, const StateBufs &S, const Work &W, int r,
                                              const double4 &st, double g00, double g01, double g10, double g11,
                                              double A0, double A1, double A2, double A3, double time_sph,
                                              double f2) {
    if (P.cspm) {
        double A4 = A0 * A3 - A1 * A2;
        if (fabs(A4) < P.cspm_tol) {
            A4 = 1; A0 = 1; A1 = 0; A2 = 0; A3 = 1;
        } else {
            A4 = 1 / A4;
        }
        g00 = A4 * (A0 * g00 + A1 * g01);
        g01 = A4 * (A2 * g00 + A3 * g01);
        g10 = A4 * (A0 * g10 + A1 * g11);
        g11 = A4 * (A2 * g10 + A3 * g11);
    }
    double4 d1;
    d1.x = -(P.D11 * g00 + P.D12 * g11);
    d1.y = -(P.D12 * g00 + P.D22 * g11);
    d1.z = -(P.D33 * g10 + P.D33 * g01);
    d1.w = -(P.D41 * g00 + P.D42 * g11);
    double4 gu = make_double4(g00, g10, g01, g11);
    W.grad_u[r] = gu;
    double4 Gs;
    double der1;
    double fdr = S.fdr[r];
    plastic_terms(P, st, gu, fdr, time_sph, S.ivar1[r], Gs, der1);
    S.fdr[r] = fdr;
    double4 sp = make_double4(0, 0, 0, 0);
    if (P.update_x) {
        double om1 = 0.5 * (g01 - g10);
        double om2 = -(0.5 * (g01 - g10));
        sp.x = 2 * om1 * st.z;
        sp.y = 2 * om2 * st.z;
        sp.z = om2 * st.x + om1 * st.y;
    }
    double4 rh;
    rh.x = -d1.x + sp.x + Gs.x;
    rh.y = -d1.y + sp.y + Gs.y;
    rh.z = -d1.z + sp.z + Gs.z;
    rh.w = -d1.w + sp.w + Gs.w;
    W.rhs1[r] = rh;
    double4 rk = W.rk_stress[r];
    W.rk_stress[r] = make_double4(rk.x + f2 * rh.x, rk.y + f2 * rh.y, rk.z + f2 * rh.z, rk.w + f2 * rh.w);
    W.rk_dev[r] = W.rk_dev[r] + der1 * f2;
}

// Stress-point receivers: stress_point_update (CSPM-normalised velocity from
// the nodes, then adapt_stress2) and, when deriv is set, get_derivatives
// (2_...:487-648) fused with plastic_terms (loop 2_...:730-734), the Jaumann
// spin (2_...:751-757) and the stress right-hand side and RK accumulation
// (2_...:762, 772).  The derivative of a stress point reads its own updated
// velocity and stress and the nodes' (which the update leaves alone), so one
// pass over the node partners after the update gives what the reference's
// get_derivatives sees after its stress_point_update.
// sig_out: where the stress after the return map goes (null: the state's
// stress; the RK stages give a separate buffer, so that the node side of
// stress_point_update can read the stresses before it at the same time).
__global__ __launch_bounds__(TILE) STAGE_OCCUPANCY void k_sp_update(StateBufs S, Work W, SearchView V, PairLists PL,
                                                    DevParams P, double time_sph, double f2, int deriv,
                                                    int adapt, int nn, int nt, double4 *__restrict__ sig_out) {
    __shared__ StagePlan pl[2];
    constexpr const LdsLayout &LY = LAY_SP_UPDATE;
    const WalkCtx C = walk_ctx(V);
    const int tile = tile_index();
    const int ltile = PL.ntile_n + tile;
    const Tile T = PL.tiles[ltile];
    if (T.first >= T.end) return;
    const int r = T.first + (int)threadIdx.x;
    const bool inside = r < T.end && V.cell[r] < C.g.ndivt;
    const int deg = inside ? PL.deg_a[r] : 0;
    const bool rev_l = !C.any_fwd && C.any_rev;
    ListWalk<MASK_NODE, true, false> lw;
    lw.start(PL.ia, PL.wa, PL.gxa, PL.gya, ltile, deg, rev_l);
    bool check;
    bool lds = tile_uses_lds(T, V, C, r, check) && PL.ok[ltile];
    if (lds) {
        load_plan<KIND_NODE>(PL, ltile, 0, pl[0]);
        if (deriv) load_plan<KIND_DUMMY>(PL, ltile, 1, pl[1]);
        __syncthreads();
        lds = pl[0].total <= CAP && (!deriv || pl[1].total <= CAP_D);
    }
    if (lds) {
        struct NItem {
            double2 x, v;
            double m, rho;
        };
        const bool stage_x = deriv && P.cspm;   // lds_sp_update allocates x only then
        copy_kind<KIND_NODE>(T, V, C, nullptr, pl[0],
                             [&](int g) { return NItem{stage_x ? S.x[g] : double2{}, S.vel[g], S.mass[g], S.rho[g]}; },
                             [&](int i, const NItem &n) {
                                 if (stage_x) lds_at<double2>(LY.x[KIND_NODE])[i] = n.x;
                                 lds_at<double2>(LY.v[KIND_NODE])[i] = n.v;
                                 lds_at<double>(LY.m[KIND_NODE])[i] = n.m;
                                 lds_at<double>(LY.rho[KIND_NODE])[i] = n.rho;
                                 lds_at<double>(LY.mr[KIND_NODE])[i] = n.m / n.rho;
                             });
        if (deriv) {
            struct DItem {
                double2 x;
                double m, rho;
                float wall, horiz;
            };
            copy_kind<KIND_DUMMY>(T, V, C, nullptr, pl[1],
                                  [&](int g) { return DItem{S.x[g], S.mass[g], S.rho[g], S.wallpos[g], S.horiz[g]}; },
                                  [&](int i, const DItem &d) {
                                      lds_at<double2>(LY.x[KIND_DUMMY])[i] = d.x;
                                      lds_at<double>(LY.m[KIND_DUMMY])[i] = d.m;
                                      lds_at<double>(LY.rho[KIND_DUMMY])[i] = d.rho;
                                      lds_at<float>(LY.wall)[i] = d.wall;
                                      lds_at<float>(LY.horiz)[i] = d.horiz;
                                  });
        }
        __syncthreads();
    }
    if (r >= T.end) return;
    int cx = 0, cy = 0;
    if (inside && !lds) cell_xy(V.cell[r], C.g, cx, cy);
    const double2 xr = S.x[r];
    const double hr = S.hsml[r];
    const LdsFields<LY> LF{};
    const SplitClass cls{check, check ? V.proc[r] : 0, V.proc, V.cnt, V.off, C.sp};
    const bool rev_g = C.any_fwd ? check : C.any_rev;

    // stress_point_update
    double v0 = 0, v1 = 0, norm = 0;
    auto spu = [&](auto kt, const auto &A, int j, float wf, float, float) __attribute__((always_inline)) {
        double w = (double)wf;
        double ms = A.m(kt, j), rs = A.rho(kt, j);
        double h2 = A.mr(kt, j) * w;
        double2 vs = A.vel(kt, j);
        v0 = v0 + vs.x * h2;
        v1 = v1 + vs.y * h2;
        norm = norm + (w * ms) / rs;
    };
    if (inside) {
        if (lds)
            lw.run(LF, spu);
        else
            walk_stencil<MASK_NODE>(GlobalCells{V.kstart, C.g, C.L}, GlobalFields{S}, C.g, cx, cy, xr, hr,
                                    KIND_SP, r, rev_g, C.any_fwd, V.pi, cls, spu);
    }
    double2 vr = S.vel[r];
    if (norm != 0) {
        vr = make_double2(v0 / norm, v1 / norm);
        S.vel[r] = vr;
    }
    double4 st = S.stress[r];
    const DevParams Pm = material_params(P, S, r);
    if (adapt) adapt_stress(st, Pm.tanfi, Pm.coh);
    if (sig_out)
        sig_out[r] = st;
    else if (adapt)
        S.stress[r] = st;
    if (!deriv) return;

    // get_derivatives
    double g00 = 0, g01 = 0, g10 = 0, g11 = 0;   // grad1_tmp(a,b): d v_a / d x_b
    double A0 = 0, A1 = 0, A2 = 0, A3 = 0;
    auto grad = [&](auto kt, const auto &A, int j, float, float gxf, float gyf) __attribute__((always_inline)) {
        constexpr int K = decltype(kt)::value;
        double ms = A.m(kt, j), rs = A.rho(kt, j);
        if constexpr (K == KIND_NODE) {             // type 1: pair_i = this SP
            double gx = (double)gxf, gy = (double)gyf;
            double h1 = gx * ms / rs;
            double h2 = gy * ms / rs;
            double2 vs = A.vel(kt, j);
            g00 = g00 + (vs.x - vr.x) * h1;
            g01 = g01 + (vs.x - vr.x) * h2;
            g10 = g10 + (vs.y - vr.y) * h1;
            g11 = g11 + (vs.y - vr.y) * h2;
            if (P.cspm) {
                double2 xs = A.x(kt, j);
                A0 = A0 + (xs.x - xr.x) * h1;
                A1 = A1 + (xs.y - xr.y) * h1;
                A2 = A2 + (xs.x - xr.x) * h2;
                A3 = A3 + (xs.y - xr.y) * h2;
            }
        } else {                                    // type 9: pair_i = dummy
            double gx = (double)(-gxf), gy = (double)(-gyf);
            double wall = (double)A.wall(kt, j);
            double2 xs = A.x(kt, j);
            double da, db;
            if (A.horiz(kt, j) == 1.0f) {
                da = fabs(xr.y - wall);
                db = fabs(xs.y - wall);
            } else {
                da = fabs(xr.x - wall);
                db = fabs(xs.x - wall);
            }
            double bt = 1 + (db / da);
            double beta = 1.5 < bt ? 1.5 : bt;
            double dv0 = vr.x * (1 - beta) + beta * 0.0;
            double dv1 = vr.y * (1 - beta) + beta * 0.0;
            double h1 = gx * ms / rs;
            double h2 = gy * ms / rs;
            g00 = g00 + (vr.x - dv0) * h1;
            g01 = g01 + (vr.x - dv0) * h2;
            g10 = g10 + (vr.y - dv1) * h1;
            g11 = g11 + (vr.y - dv1) * h2;
        }
    };
    if (inside) {
        if (lds) {
            ListWalk<MASK_NODE | MASK_DUMMY, false, true, SPS_WALK_DEPTH> lg;
            lg.start(PL.ia, PL.wa, PL.gxa, PL.gya, ltile, deg, rev_l);
            lg.run(LF, grad);
        } else
            walk_stencil<MASK_NODE | MASK_DUMMY>(GlobalCells{V.kstart, C.g, C.L}, GlobalFields{S}, C.g, cx, cy,
                                                 xr, hr, KIND_SP, r, rev_g, C.any_fwd, V.pi, cls, grad);
    }
    sp_stage_tail(Pm, S, W, r, st, g00, g01, g10, g11, A0, A1, A2, A3, time_sph, f2);
}

// boundary_forces for one node: the nonzero terms of its inner-layer dummies
// in ascending dummy index (the reference's j loop over every dummy,
// 2_...:1117-1161), found by repeated minimum selection over the dummies of
// its 3x3 cell stencil; cells are 2*max(h) wide, beyond the 0.75*dx reach of
// a nonzero term.
constexpr int MAX_WALL_TERMS = 16;

__device__ __forceinline__ double2 wall_force(const StateBufs &S, const SearchView &V, const WalkCtx &C,
                                              int cx, int cy, double2 xr, double hr, const DevParams &P,
                                              int *__restrict__ err) {
    double2 fb = make_double2(0.0, 0.0);
    const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx + 1 < C.g.ndivx ? cx + 1 : C.g.ndivx - 1;
    const int y0 = cy > 0 ? cy - 1 : 0, y1 = cy + 1 < C.g.ndivy ? cy + 1 : C.g.ndivy - 1;
    int last = -1;
    for (int it = 0; it <= MAX_WALL_TERMS; ++it) {
        int best = 0x7fffffff;
        float b0 = 0.0f, b1 = 0.0f;
        for (int y = y0; y <= y1; ++y)
            for (int xx = x0; xx <= x1; ++xx) {
                const int key = cell_key(xx, y, C.g);
                const int lo = V.kstart[2 * C.L + key], hi = V.kstart[2 * C.L + key + 1];
                for (int s = lo; s < hi; ++s) {
                    int o = S.orig[s];
                    if (o >= P.nt + P.nd2 || o <= last || o >= best) continue;
                    double2 xs = S.x[s];
                    float t0, t1;
                    wall_force_term(xr.x - xs.x, xr.y - xs.y, hr, S.hsml[s], P.dx, t0, t1);
                    if (t0 != 0.0f || t1 != 0.0f || t0 != t0 || t1 != t1) {
                        best = o;
                        b0 = t0;
                        b1 = t1;
                    }
                }
            }
        if (best == 0x7fffffff) return fb;
        if (it == MAX_WALL_TERMS) {
            atomicOr(err, ERR_WALL_TERMS);
            return fb;
        }
        fb.x = fb.x + (double)b0;
        fb.y = fb.y + (double)b1;
        last = best;
    }
    return fb;
}

// get_derivatives for node receivers, fused with gravity_force
// (3_...:2640-2702), boundary_forces (2_...:1117-1161), the artificial
// viscosity's share (2_...:826-904, summed over node partners, list B, by
// k_av_nodes) and the momentum right-hand side and RK accumulation
// (2_...:763-764, 769).  The stress gradient sums over stress-point and dummy
// partners (list A) in list order, from a window of both kinds in LDS.
// The end of get_derivatives for node r (CSPM correction, div2; 2_...:617-642)
// from its stress-divergence sums and CSPM moments, then gravity_force,
// boundary_forces (when a dummy is among its pairs), the artificial
// viscosity sums ax, ay, the velocity right-hand side and the RK accumulation
// of the stage (2_...:737-776).
__device__ __forceinline__ void node_stage_tail(const StateBufs &S, const Work &W, const SearchView &V,
                                                const WalkCtx &C, const DevParams &P, int r, int cx, int cy,
                                                double2 xr, double2 vr, double hr, double q00, double q01,
                                                double q10, double q11, double q20, double q21, double A0,
                                                double A1, double A2, double A3, int ndum, float ax, float ay,
                                                double gx_grav, double gy_grav, double f2,
                                                int *__restrict__ err) {
    if (P.cspm) {
        double A4 = A0 * A3 - A1 * A2;
        if (fabs(A4) < P.cspm_tol) {
            A4 = 1; A0 = 1; A1 = 0; A2 = 0; A3 = 1;
        } else {
            A4 = 1 / A4;
        }
        q00 = A4 * (A0 * q00 + A1 * q01);
        q01 = A4 * (A2 * q00 + A3 * q01);
        q10 = A4 * (A0 * q10 + A1 * q11);
        q11 = A4 * (A2 * q10 + A3 * q11);
    }
    const double d2x = -(q00 + q21);
    const double d2y = -(q20 + q11);
    // gravity_force: G = factg*cgrav - DampingTG*v
    const double Gx = gx_grav - P.damping * vr.x;
    const double Gy = gy_grav - P.damping * vr.y;
    // boundary_forces
    double2 fb = W.f_bound[r];
    if (P.do_bforce) {
        fb = ndum ? wall_force(S, V, C, cx, cy, xr, hr, P, err) : make_double2(0.0, 0.0);
        W.f_bound[r] = fb;
    }
    double2 avf = make_double2(0.0, 0.0);
    if (P.do_av) {
        avf = make_double2(-(double)ax, -(double)ay);
        W.art_visc[r] = avf;
    }
    double2 rh;
    rh.x = -d2x + Gx + avf.x + fb.x + 0.0;
    rh.y = -d2y + Gy + avf.y + fb.y + 0.0;
    W.rhs2[r] = rh;
    double2 rk = W.rk_vel[r];
    W.rk_vel[r] = make_double2(rk.x + f2 * rh.x, rk.y + f2 * rh.y);
}

// sp_stress: the stress points' stresses after the stage's second return
// map (k_sp_update's sig_out, or the state's stress).
__global__ __launch_bounds__(TILE) STAGE_OCCUPANCY void k_deriv_nodes(StateBufs S, Work W, SearchView V, PairLists PL,
                                                      DevParams P, double gx_grav, double gy_grav, double f2,
                                                      int *__restrict__ err, int nn,
                                                      const double4 *__restrict__ sp_stress,
                                                      const float2 *__restrict__ av_in) {
    __shared__ StagePlan pl[2];
    constexpr const LdsLayout &LY = LAY_DERIV_NODES;
    const WalkCtx C = walk_ctx(V);
    const int tile = tile_index();
    const Tile T = PL.tiles[tile];
    if (T.first >= T.end) return;
    const int r = T.first + (int)threadIdx.x;
    const bool active = r < T.end;
    const int key_r = active ? V.cell[r] : C.g.ndivt;
    const bool inside = key_r < C.g.ndivt;
    const bool rev_l = !C.any_fwd && C.any_rev;
    ListWalk<MASK_SP | MASK_DUMMY, false, true, SPS_WALK_DEPTH> la;
    la.start(PL.ia, PL.wa, PL.gxa, PL.gya, tile, inside ? PL.deg_a[r] : 0, rev_l);
    bool check;
    const bool tiled = tile_uses_lds(T, V, C, r, check) && PL.ok[tile];
    int cx = 0, cy = 0;
    if (inside) cell_xy(key_r, C.g, cx, cy);
    const int ra = active ? r : T.first;   // a valid slot for the idle lanes' loads
    const double2 xr = S.x[ra];
    const double4 sr = S.stress[ra];
    const double rhor = S.rho[ra], hr = S.hsml[ra];
    const double rr2 = rhor * rhor;
    const double or0 = sr.x / rr2, or1 = sr.y / rr2, or2 = sr.z / rr2;   // own sigma / rho^2
    const SplitClass cls{check, check ? V.proc[r] : 0, V.proc, V.cnt, V.off, C.sp};
    const bool rev_g = C.any_fwd ? check : C.any_rev;
    const LdsFields<LY> LF{};

    // phase 1: stress gradient over stress points (type 1) and dummies (type 6)
    bool lds = tiled;
    if (lds) {
        load_plan<KIND_SP>(PL, tile, 0, pl[0]);
        load_plan<KIND_DUMMY>(PL, tile, 1, pl[1]);
        __syncthreads();
        lds = pl[0].total <= CAP && pl[1].total <= CAP_D;
    }
    if (lds) {
        struct SItem {
            double2 x;
            double m, rho;
            double4 st;
        };
        const bool stage_x = P.cspm;   // lds_deriv_nodes allocates x only then
        copy_kind<KIND_SP>(T, V, C, nullptr, pl[0],
                           [&](int g) { return SItem{stage_x ? S.x[g] : double2{}, S.mass[g], S.rho[g], sp_stress[g]}; },
                           [&](int i, const SItem &v) {
                               if (stage_x) lds_at<double2>(LY.x[KIND_SP])[i] = v.x;
                               lds_at<double>(LY.m[KIND_SP])[i] = v.m;
                               lds_at<double>(LY.rho[KIND_SP])[i] = v.rho;
                               const double rs2 = v.rho * v.rho;
                               lds_at<double>(LY.p0[KIND_SP])[i] = v.st.x / rs2;
                               lds_at<double>(LY.p1[KIND_SP])[i] = v.st.y / rs2;
                               lds_at<double>(LY.p2[KIND_SP])[i] = v.st.z / rs2;
                           });
        copy_kind<KIND_DUMMY>(T, V, C, nullptr, pl[1], [&](int g) { return make_double2(S.mass[g], S.rho[g]); },
                              [&](int i, const double2 &v) {
                                  lds_at<double>(LY.m[KIND_DUMMY])[i] = v.x;
                                  lds_at<double>(LY.rho[KIND_DUMMY])[i] = v.y;
                              });
        __syncthreads();
    }
    double q00 = 0, q01 = 0, q10 = 0, q11 = 0, q20 = 0, q21 = 0;        // grad2_tmp(s,b)
    double A0 = 0, A1 = 0, A2 = 0, A3 = 0;
    int ndum = 0;
    auto grad = [&](auto kt, const auto &A, int j, float, float gxf, float gyf) __attribute__((always_inline)) {
        constexpr int K = decltype(kt)::value;
        double ms = A.m(kt, j), rs = A.rho(kt, j);
        double gx = (double)(-gxf), gy = (double)(-gyf);   // pair_i is the partner
        double p0, p1, p2;
        if constexpr (K == KIND_SP) {               // type 1: sigma_s / rho_s^2
            const double3 ps = A.sr2(kt, j);
            p0 = ps.x; p1 = ps.y; p2 = ps.z;
            if (P.cspm) {
                double h1b = -(gx * ms / rs);
                double h2b = -(gy * ms / rs);
                double2 xs = A.x(kt, j);
                A0 = A0 + (xs.x - xr.x) * h1b;
                A1 = A1 + (xs.y - xr.y) * h1b;
                A2 = A2 + (xs.x - xr.x) * h2b;
                A3 = A3 + (xs.y - xr.y) * h2b;
            }
        } else {                                    // type 6: dummy stress := node stress
            const double rs2 = rs * rs;
            p0 = sr.x / rs2; p1 = sr.y / rs2; p2 = sr.z / rs2;
            ndum++;
        }
        double a;
        a = gx * (or0 + p0); q00 = q00 - ms * a;
        a = gy * (or0 + p0); q01 = q01 - ms * a;
        a = gx * (or1 + p1); q10 = q10 - ms * a;
        a = gy * (or1 + p1); q11 = q11 - ms * a;
        a = gx * (or2 + p2); q20 = q20 - ms * a;
        a = gy * (or2 + p2); q21 = q21 - ms * a;
    };
    if (inside) {
        if (lds)
            la.run(LF, grad);
        else
        {
            StateBufs S2 = S;   // stress points' stresses from sp_stress (the walk reads no node's)
            S2.stress = const_cast<double4 *>(sp_stress);
            walk_stencil<MASK_SP | MASK_DUMMY>(GlobalCells{V.kstart, C.g, C.L}, GlobalFields{S2}, C.g, cx, cy,
                                               xr, hr, KIND_NODE, r, rev_g, C.any_fwd, V.pi, cls, grad);
        }
    }

    // the artificial viscosity over node partners (type 3), summed by k_av_nodes
    float ax = 0.0f, ay = 0.0f;
    if (!active) return;
    if (av_in) {
        const float2 a = av_in[r];
        ax = a.x;
        ay = a.y;
    }
    const double2 vr = S.vel[r];
    node_stage_tail(S, W, V, C, P, r, cx, cy, xr, vr, hr, q00, q01, q10, q11, q20, q21, A0, A1, A2, A3, ndum,
                    ax, ay, gx_grav, gy_grav, f2, err);
}

// artificial_viscosity (2_...:826-904) for the node receivers: the list-order
// sums over node partners (list B) in single precision, as the reference
// accumulates them, written to av_out for k_deriv_nodes.  It reads only node
// data that the stress points' update leaves alone, so the RK stages run it
// beside k_sp_update.
__global__ __launch_bounds__(TILE) void k_av_nodes(StateBufs S, SearchView V, PairLists PL, DevParams P,
                                                   float2 *__restrict__ av_out, int nn) {
    __shared__ StagePlan pl[2];
    constexpr const LdsLayout &LY = LAY_AV_NODES;
    const WalkCtx C = walk_ctx(V);
    const int tile = tile_index();
    const Tile T = PL.tiles[tile];
    if (T.first >= T.end) return;
    const int r = T.first + (int)threadIdx.x;
    const bool active = r < T.end;
    const int key_r = active ? V.cell[r] : C.g.ndivt;
    const bool inside = key_r < C.g.ndivt;
    const bool rev_l = !C.any_fwd && C.any_rev;
    bool check;
    const bool tiled = tile_uses_lds(T, V, C, r, check) && PL.ok[tile];
    int cx = 0, cy = 0;
    if (inside) cell_xy(key_r, C.g, cx, cy);
    const int ra = active ? r : T.first;   // a valid slot for the idle lanes' loads
    const double2 xr = S.x[ra], vr = S.vel[ra];
    const double rhor = S.rho[ra], hr = S.hsml[ra];
    const SplitClass cls{check, check ? V.proc[r] : 0, V.proc, V.cnt, V.off, C.sp};
    const bool rev_g = C.any_fwd ? check : C.any_rev;
    const LdsFields<LY> LF{};
    float ax = 0.0f, ay = 0.0f, visc = 0.0f;
    {
        ListWalk<MASK_NODE, false, true> lb;
        lb.start(PL.ib, nullptr, PL.gxb, PL.gyb, tile, inside ? PL.deg_b[r] : 0, rev_l);
        bool lds2 = tiled;
        if (lds2) {
            load_plan<KIND_NODE>(PL, tile, 0, pl[0]);
            __syncthreads();
            lds2 = pl[0].total <= CAP;
        }
        if (lds2) {
            struct NItem {
                double2 x, v;
                double h, m, rho;
            };
            const bool stage_h = !P.h_uni;   // lds_av_nodes allocates h only then
            copy_kind<KIND_NODE>(T, V, C, nullptr, pl[0],
                                 [&](int g) { return NItem{S.x[g], S.vel[g], stage_h ? S.hsml[g] : 0.0, S.mass[g],
                                                           S.rho[g]}; },
                                 [&](int i, const NItem &n) {
                                     lds_at<double2>(LY.x[KIND_NODE])[i] = n.x;
                                     lds_at<double2>(LY.v[KIND_NODE])[i] = n.v;
                                     if (stage_h) lds_at<double>(LY.h[KIND_NODE])[i] = n.h;
                                     lds_at<double>(LY.m[KIND_NODE])[i] = n.m;
                                     lds_at<double>(LY.rho[KIND_NODE])[i] = n.rho;
                                 });
            __syncthreads();
        }
        auto av = [&](auto kt, const auto &A, int j, float, float gxf, float gyf) __attribute__((always_inline)) {
            double2 xs = A.x(kt, j), vs = A.vel(kt, j);
            float xij = (float)(xr.x - xs.x);
            float yij = (float)(xr.y - xs.y);
            // one smoothing length for all: the partner's is the receiver's
            float hh = (float)(0.5 * (hr + (P.h_uni ? hr : A.h(kt, j))));
            float rho2 = (float)(0.5 * (rhor + A.rho(kt, j)));
            float div_u = (float)((double)xij * (vr.x - vs.x));
            div_u = (float)((double)div_u + (double)yij * (vr.y - vs.y));
            float sr_ = sqrtf(xij * xij + yij * yij);
            float theta = (hh * div_u) / (sr_ * sr_ + 0.01f * (hh * hh));
            if (div_u < 0) {
                visc = (float)((-(P.alpha * (double)600.0f * (double)theta) + P.beta * (double)(theta * theta)) /
                               (double)rho2);
            } else if (div_u >= 0) {
                visc = 0;
            }
            double ms = A.m(kt, j);
            ax = (float)((double)ax + (double)(visc * gxf) * ms);
            ay = (float)((double)ay + (double)(visc * gyf) * ms);
        };
        if (inside) {
            if (lds2)
                lb.run(LF, av);
            else
                walk_stencil<MASK_NODE>(GlobalCells{V.kstart, C.g, C.L}, GlobalFields{S}, C.g, cx, cy, xr, hr,
                                        KIND_NODE, r, rev_g, C.any_fwd, V.pi, cls, av);
        }
    }
    if (active) av_out[r] = make_float2(ax, ay);
}

// XSPH_update (2_...:189-239) for the receivers of kind KR: the list-order
// sum of (m_s / rho_s) (v_s - v_r) w over the same-kind partners (type-3 pairs
// for nodes, type 2 for stress points), written to xv; k_position then moves
// the particles.  The sum reads only velocities and masses, which no kernel
// of the step changes after the last stress_point_update.
constexpr LdsLayout layout_xsph() {
    LdsLayout L = layout_none();
    for (int k = 0; k < 2; ++k) {   // one receiver kind per launch: both share the bytes
        L.v[k] = 0;
        L.mr[k] = 16 * CAP;
    }
    L.bytes = 24 * CAP;
    return L;
}
constexpr LdsLayout LAY_XSPH = layout_xsph();

template <int KR>
__global__ __launch_bounds__(TILE) void k_xsph(StateBufs S, SearchView V, PairLists PL,
                                               double2 *__restrict__ xv) {
    __shared__ StagePlan pl;
    constexpr const LdsLayout &LY = LAY_XSPH;
    const WalkCtx C = walk_ctx(V);
    const int tile = tile_index();
    const int ltile = KR == KIND_NODE ? tile : PL.ntile_n + tile;
    const Tile T = PL.tiles[ltile];
    if (T.first >= T.end) return;
    const int r = T.first + (int)threadIdx.x;
    const bool inside = r < T.end && V.cell[r] < C.g.ndivt;
    const bool rev_l = !C.any_fwd && C.any_rev;
    ListWalk<(1 << KR), true, false> lw;
    lw.start(PL.ib, PL.wb, nullptr, nullptr, ltile, inside ? PL.deg_b[r] : 0, rev_l);
    bool check;
    bool lds = tile_uses_lds(T, V, C, r, check) && PL.ok[ltile];
    if (lds) {
        load_plan<KR>(PL, ltile, 0, pl);
        __syncthreads();
        lds = pl.total <= CAP;
    }
    if (lds) {
        struct Item {
            double2 v;
            double m, rho;
        };
        copy_kind<KR>(T, V, C, nullptr, pl, [&](int g) { return Item{S.vel[g], S.mass[g], S.rho[g]}; },
                      [&](int i, const Item &v) {
                          lds_at<double2>(LY.v[KR])[i] = v.v;
                          lds_at<double>(LY.mr[KR])[i] = v.m / v.rho;
                      });
        __syncthreads();
    }
    if (r >= T.end) return;
    const double2 vr = S.vel[r];
    double s0 = 0.0, s1 = 0.0;
    auto body = [&](auto kt, const auto &A, int j, float wf, float, float) __attribute__((always_inline)) {
        const double w = (double)wf;
        const double mr = A.mr(kt, j);
        const double2 v = A.vel(kt, j);
        s0 = s0 + (mr * (v.x - vr.x)) * w;
        s1 = s1 + (mr * (v.y - vr.y)) * w;
    };
    if (inside) {
        if (lds) {
            lw.run(LdsFields<LY>{}, body);
        } else {
            int cx, cy;
            cell_xy(V.cell[r], C.g, cx, cy);
            const SplitClass cls{check, check ? V.proc[r] : 0, V.proc, V.cnt, V.off, C.sp};
            walk_stencil<(1 << KR)>(GlobalCells{V.kstart, C.g, C.L}, GlobalFields{S}, C.g, cx, cy, S.x[r],
                                    S.hsml[r], KR, r, C.any_fwd ? check : C.any_rev, C.any_fwd, V.pi, cls, body);
        }
    }
    xv[r] = make_double2(s0, s1);
}

// The type-6 loop of get_derivatives overwrites each dummy's stress with
// that of its node partners in list order (2_...:580); the last one stays.
__global__ void k_dummy_stress(StateBufs S, SearchView V, int nt, int nt2) {
    int r = nt + blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nt2) return;
    const WalkCtx C = walk_ctx(V);
    const int key_r = V.cell[r];
    if (key_r >= C.g.ndivt) return;
    int cx, cy;
    cell_xy(key_r, C.g, cx, cy);
    const bool check = C.mixed && V.proc[r] >= C.sp.first_rev;
    const SplitClass cls{check, check ? V.proc[r] : 0, V.proc, V.cnt, V.off, C.sp};
    int last = -1;
    walk_stencil<MASK_NODE>(GlobalCells{V.kstart, C.g, C.L}, GlobalFields{S}, C.g, cx, cy, S.x[r], S.hsml[r],
                            KIND_DUMMY, r, C.any_fwd ? check : C.any_rev, C.any_fwd, V.pi, cls,
                            [&](auto, const auto &, int j, float, float, float) { last = j; });
    if (last >= 0) S.stress[r] = S.stress[last];
}

// End of RK4 (2_...:784-799) and update_strain (3_...:1847-1863).
// hook: record the stress points' plastic trial states (SurrogateDev).
__global__ void k_rk_final(StateBufs S, Work W, double dt, int adapt, double tanfi, double coh,
                           int nn, int nt, SurrogateDev G, int hook, int *__restrict__ err) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    double c = dt / 6;
    material_tc(S, i, tanfi, coh);
    const bool watched = S.ood[i] == 0;
    if (i < nn) {
        double2 v0 = W.vel0[i], rk = W.rk_vel[i];
        const double2 v = make_double2(v0.x + c * rk.x, v0.y + c * rk.y);
        S.vel[i] = v;
        double4 s = S.stress[i];
        if (adapt) {
            adapt_stress(s, tanfi, coh);
            S.stress[i] = s;
        }
        if (watched && !(finite4(s) && finite2(v))) atomicOr(err, ERR_NONFINITE);
        W.ddev[i] = W.rk_dev[i] / 6;
    } else {
        double4 s0 = W.stress0[i], rk = W.rk_stress[i];
        double4 s = make_double4(s0.x + c * rk.x, s0.y + c * rk.y, s0.z + c * rk.z, s0.w + c * rk.w);
        if (adapt && hook) {
            const double4 tr = s;
            adapt_stress(s, tanfi, coh);
            if (s.x != tr.x || s.y != tr.y || s.z != tr.z || s.w != tr.w)
                surrogate_record(G, i, s0, tr, s, tanfi, coh);
        } else if (adapt) {
            adapt_stress(s, tanfi, coh);
        }
        S.stress[i] = s;
        if (watched && !finite4(s)) atomicOr(err, ERR_NONFINITE);
        double dd = W.rk_dev[i] / 6;
        W.ddev[i] = dd;
        S.ivar1[i] = S.ivar1[i] + dt * dd;
    }
}

// Leapfrog (no counterpart in the reference): the kick of the node
// velocities by kdt with the accelerations of the step's state (W.rhs2).
__global__ void k_lf_kick(StateBufs S, Work W, double kdt, int nn) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const double2 v0 = W.vel0[i], a = W.rhs2[i];
    S.vel[i] = make_double2(v0.x + kdt * a.x, v0.y + kdt * a.y);
}

// Leapfrog's stress update: the stress points' stress by dt with the rate
// of the kicked velocities (W.rhs1), adapt_stress2 of every particle, the
// plastic strain rate of the step and update_strain (3_...:1847-1863).
__global__ void k_lf_final(StateBufs S, Work W, double dt, int adapt, double tanfi, double coh, int nn, int nt,
                           int *__restrict__ err) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    material_tc(S, i, tanfi, coh);
    const bool watched = S.ood[i] == 0;
    if (i < nn) {
        double4 s = S.stress[i];
        if (adapt) {
            adapt_stress(s, tanfi, coh);
            S.stress[i] = s;
        }
        if (watched && !(finite4(s) && finite2(S.vel[i]))) atomicOr(err, ERR_NONFINITE);
        W.ddev[i] = W.rk_dev[i];
        return;
    }
    const double4 s0 = W.stress0[i], r = W.rhs1[i];
    double4 s = make_double4(s0.x + dt * r.x, s0.y + dt * r.y, s0.z + dt * r.z, s0.w + dt * r.w);
    if (adapt) adapt_stress(s, tanfi, coh);
    S.stress[i] = s;
    if (watched && !finite4(s)) atomicOr(err, ERR_NONFINITE);
    const double dd = W.rk_dev[i];
    W.ddev[i] = dd;
    S.ivar1[i] = S.ivar1[i] + dt * dd;
}

// Position update (2_...:140-147) through the single-precision vel_half, and
// the displacement of the nodes (2_...:171, 179).
// xv (XSPH): the XSPH sums of k_xsph; the positions then follow
// XSPH_update (2_...:234-235), x = x0 + dt (v + eps sum), eps = 0.5.
__global__ void k_position(StateBufs S, Work W, double dt, int update_x, int nn, int nt,
                           const double2 *__restrict__ xv, int *__restrict__ err, int drift) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    double2 v0 = W.vx0[i], v = S.vel[i];
    double2 xn = make_double2(0.0, 0.0);
    if (drift) {   // leapfrog: the drift with the half-step velocities
        if (update_x) {
            const double2 x0 = W.x0[i];
            xn = make_double2(x0.x + dt * v.x, x0.y + dt * v.y);
            S.x[i] = xn;
            if (i < nn) {
                const double2 x00 = S.x00[i];
                S.displ[i] = make_double2(xn.x - x00.x, xn.y - x00.y);
            }
        } else if (i < nn) {
            const double2 d = S.displ[i];
            S.displ[i] = make_double2(d.x + dt * v.x, d.y + dt * v.y);
        }
    } else if (update_x && xv) {
        const double eps = 0.5;
        const double2 x0 = W.x0[i], s = xv[i];
        xn = make_double2(x0.x + dt * (v.x + eps * s.x), x0.y + dt * (v.y + eps * s.y));
        S.x[i] = xn;
        if (i < nn) {
            double2 x00 = S.x00[i];
            S.displ[i] = make_double2(xn.x - x00.x, xn.y - x00.y);
        }
    } else if (update_x) {
        float hx = (float)(0.5 * (v0.x + v.x));
        float hy = (float)(0.5 * (v0.y + v.y));
        double2 x0 = W.x0[i];
        xn = make_double2(x0.x + (double)hx * dt, x0.y + (double)hy * dt);
        S.x[i] = xn;
        if (i < nn) {
            double2 x00 = S.x00[i];
            S.displ[i] = make_double2(xn.x - x00.x, xn.y - x00.y);
        }
    } else if (i < nn) {
        double2 d = S.displ[i];
        S.displ[i] = make_double2(d.x + 0.5 * (v0.x + v.x) * dt, d.y + 0.5 * (v0.y + v.y) * dt);
    }
    // the watchdog: the step's velocities and positions of the particles in
    // the domain box (one that has left it interacts with nothing)
    if (S.ood[i] == 0 && !(finite2(v) && finite2(xn))) atomicOr(err, ERR_NONFINITE);
}

// Prescribed values of one step: BC row b's value and its variable (5 vx,
// 6 vy, 1 sxx, 2 syy, 3 sxy, 0 none).
struct BcValues {
    int nbc;
    double v[SPS_MAX_BCS];
    signed char var[SPS_MAX_BCS];
};

// The step's BC values into device memory, read by k_bcs: a replayed graph
// holds k_bcs with its arguments fixed, the values change with time.
__global__ void k_bc_values(BcValues V, double *__restrict__ vals) {
    const int b = (int)threadIdx.x;
    if (b < V.nbc) vals[b] = V.v[b];
}

// Normal_BCs (3_...:1644-1753) on the node slots: a node with BC_or_not = 1
// takes the values of the BCs listed in bc_info(3:7) (by original index), in
// that order; vals holds the step's values (k_bc_values), V the variables.
// ifsigman: apply_stress_free (3_...:1758-1841) on the nodes with BC_or_not
// 2, which Normal_BCs leaves alone, so one thread does either: the stress
// keeps sigma_tt of the normal frame, except at nodes with bc_int = 1 (a
// node-dummy pair in the list of the last derivative pass): bit 0 of
// bci_slot (the step's list, by slot) or bci_orig (the previous step's, by
// original index); both null without wall dummies, where bc_int is 0.
__global__ void k_bcs(StateBufs S, const int *__restrict__ bc_info, const int *__restrict__ bc_or_not, BcValues V,
                      const double *__restrict__ vals, int ifsigman, const double2 *__restrict__ normal, int nn,
                      const int *__restrict__ bci_slot, const int *__restrict__ bci_orig) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nn) return;
    const int i = S.orig[r];
    const int *info = bc_info + 8 * (size_t)i;
    if (ifsigman && bc_or_not[i] == 2) {
        const int bc_int = bci_slot ? (bci_slot[r] & 1) : (bci_orig ? bci_orig[i] : 0);
        if (bc_int == 1) return;
        const double2 n = normal[i];
        if (__builtin_isnan(n.x) || __builtin_isnan(n.y)) return;
        const double s2 = n.y * n.y, c2 = n.x * n.x, sc = n.y * n.x;
        double4 s = S.stress[r];
        const double stt = s2 * s.x - 2 * sc * s.z + c2 * s.y;
        S.stress[r] = make_double4(s2 * stt, c2 * stt, -sc * stt, c2 * stt);
        return;
    }
    if (bc_or_not[i] != 1 || info[1] == 0) return;
    double2 v = S.vel[r];
    double4 s = S.stress[r];
    bool dv = false, ds = false;
    for (int k = 0; k < 5; ++k) {
        const int b = info[2 + k];
        if (b <= 0 || b > V.nbc) continue;
        const double val = vals[b - 1];
        switch (V.var[b - 1]) {
        case 5: v.x = val; dv = true; break;
        case 6: v.y = val; dv = true; break;
        case 1: s.x = val; ds = true; break;
        case 3: s.z = val; ds = true; break;
        case 2: s.y = val; ds = true; break;
        default: break;
        }
    }
    if (dv) S.vel[r] = v;
    if (ds) S.stress[r] = s;
}

// bc_int of each node as the step's derivative passes leave it (bit 0 of the
// search's per-node count), by original index, for the next step's first BC
// calls, which run before that step's first derivative pass
__global__ void k_bcint_keep(const int *__restrict__ nodeinfo, const int *__restrict__ orig,
                             int *__restrict__ bci_orig, int nn) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nn) bci_orig[orig[r]] = nodeinfo[r] & 1;
}

// ------------------------------------------------------------ free surface
// get_nodes_on_free_surface (3_...:1116-1411) over the step's pair list, at
// the updated positions.  k_dump_pairs (as_slots) lays the list out in the
// reference's traversal order; every particle's pairs of types 1-3 are then
// gathered in that order (count, scan, fill, sort), so the single-precision
// sums of steps 1 and 4 add in the reference's order.  Its locals are
// single precision and normal / subset double, as in the reference; x**0.5
// is a square root there (the compiler lowers it so) and 2**0.5 the single-
// precision sqrt(2).
struct SurfBufs {
    const int *pi, *pj, *pt;          // per pair, in traversal order: slots and type
    const float *gx, *gy;             // stored dW/dx, dW/dy of pair_i
    const unsigned *start, *ent;      // per particle: its entries 2 k + (0: pair_i, 1: pair_j)
    double2 *nrm0;                    // first normal estimate (step 2)
    float2 *tt, *tau;                 // scan-region centre and tangent (step 3)
    unsigned *flag;                   // step 3: a partner in the scan region (f_int = -1)
    int *sub;                         // subset after 3_...:1341-1349
};

__global__ void k_surf_zero(unsigned *__restrict__ a, long long n) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = 0u;
}

__device__ __forceinline__ bool surf_kind(int t) { return t == 1 || t == 2 || t == 3; }

__global__ void k_surf_count(SurfBufs B, unsigned *__restrict__ cnt, long long np) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np || !surf_kind(B.pt[k])) return;
    atomicAdd(cnt + B.pi[k], 1u);
    atomicAdd(cnt + B.pj[k], 1u);
}

__global__ void k_surf_fill(SurfBufs B, unsigned *__restrict__ cur, unsigned *__restrict__ ent, long long np) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np || !surf_kind(B.pt[k])) return;
    const int i = B.pi[k], j = B.pj[k];
    ent[B.start[i] + atomicAdd(cur + i, 1u)] = 2u * (unsigned)k;
    ent[B.start[j] + atomicAdd(cur + j, 1u)] = 2u * (unsigned)k + 1u;
}

// each particle's entries into traversal order (insertion sort: a few
// dozen per particle)
__global__ void k_surf_sort(const unsigned *__restrict__ start, unsigned *__restrict__ ent, int nt) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nt) return;
    const unsigned b = start[r], e = start[r + 1];
    for (unsigned a = b + 1; a < e; ++a) {
        const unsigned v = ent[a];
        unsigned c = a;
        while (c > b && ent[c - 1] > v) {
            ent[c] = ent[c - 1];
            --c;
        }
        ent[c] = v;
    }
}

// m_p / rho_p dW of one entry, as the reference's hf1 / hf2 of that end
__device__ __forceinline__ void surf_hf(const StateBufs &S, const SurfBufs &B, unsigned e, int &p, float &h1,
                                        float &h2) {
    const unsigned k = e >> 1;
    const bool as_j = e & 1u;
    p = as_j ? B.pi[k] : B.pj[k];
    const double m = S.mass[p], rho = S.rho[p];
    const double g1 = (double)B.gx[k], g2 = (double)B.gy[k];
    h1 = as_j ? (float)(-(m * g1 / rho)) : (float)(m * g1 / rho);
    h2 = as_j ? (float)(-(m * g2 / rho)) : (float)(m * g2 / rho);
}

// steps 1-2 (3_...:1160-1245): renormalisation matrix, first normal, and
// the scan frame of step 3
__global__ void k_surf_a(StateBufs S, SurfBufs B, int nt) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nt) return;
    const double2 xr = S.x[r];
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f, f0 = 0.f, f1 = 0.f;
    for (unsigned q = B.start[r]; q < B.start[r + 1]; ++q) {
        int p;
        float h1, h2;
        surf_hf(S, B, B.ent[q], p, h1, h2);
        const double2 xp = S.x[p];
        const double d0 = xp.x - xr.x, d1 = xp.y - xr.y;
        a0 = (float)((double)a0 + d0 * (double)h1);
        f0 = f0 + h1;
        a1 = (float)((double)a1 + d1 * (double)h1);
        a2 = (float)((double)a2 + d0 * (double)h2);
        a3 = (float)((double)a3 + d1 * (double)h2);
        f1 = f1 + h2;
    }
    if (fabsf(a0) <= 1.e-8f) a0 = 0.f;
    if (fabsf(a1) <= 1.e-8f) a1 = 0.f;
    if (fabsf(a2) <= 1.e-8f) a2 = 0.f;
    if (fabsf(a3) <= 1.e-8f) a3 = 0.f;
    const float v1 = -(a0 * f0 + a1 * f1), v2 = -(a2 * f0 + a3 * f1);
    const float v3 = sqrtf(v1 * v1 + v2 * v2);
    const double n0 = (double)(v1 / v3), n1 = (double)(v2 / v3);
    const double hr = S.hsml[r];
    B.nrm0[r] = make_double2(n0, n1);
    B.tt[r] = make_float2((float)(xr.x + hr * n0), (float)(xr.y + hr * n1));
    B.tau[r] = make_float2((float)(-n1), (float)n0);
}

// step 3 (3_...:1250-1278): does partner p lie in particle r's scan region
__device__ __forceinline__ bool surf_hit(const StateBufs &S, const SurfBufs &B, int r, int p) {
    const double2 xr = S.x[r], xp = S.x[p];
    const double d0 = xp.x - xr.x, d1 = xp.y - xr.y;
    const float dist = (float)sqrt(d0 * d0 + d1 * d1);
    const float2 t = B.tt[r], u = B.tau[r];
    const double2 n = B.nrm0[r];
    const float v0 = (float)(xp.x - (double)t.x), v1 = (float)(xp.y - (double)t.y);
    const float vnorm = sqrtf(v0 * v0 + v1 * v1);
    const float prod = (float)(fabs(n.x * (double)v0 + n.y * (double)v1) + (double)fabsf(u.x * v0 + u.y * v1));
    const double hr = S.hsml[r];
    const float limit = (float)((double)1.41421354f * hr);
    return (dist >= limit && (double)vnorm < hr) || (dist < limit && (double)prod < hr);
}

// step 3 per pair: node and stress-point pairs test both ends, dummy pairs
// their non-dummy end (pair_j); the outcome does not depend on the order
__global__ void k_surf_hit(StateBufs S, SurfBufs B, long long np) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    const int t = B.pt[k], i = B.pi[k], j = B.pj[k];
    if (t == 2 || t == 3) {
        if (surf_hit(S, B, i, j)) B.flag[i] = 1u;
        if (surf_hit(S, B, j, i)) B.flag[j] = 1u;
    } else if (t == 6 || t == 9) {
        if (surf_hit(S, B, j, i)) B.flag[j] = 1u;
    }
}

// XSPH_update's bookkeeping (2_...:224-230): a node paired (type 3) with a
// free-surface node becomes 3, so the detection that follows resets even a
// prescribed node to 0.  Concurrent marks cannot change any test's outcome:
// they only turn a value other than 2 into 3.
__global__ void k_surf_mark(StateBufs S, SurfBufs B, int *__restrict__ bc_or_not, long long np) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np || B.pt[k] != 3) return;
    const int oi = S.orig[B.pi[k]], oj = S.orig[B.pj[k]];
    const int bi = bc_or_not[oi], bj = bc_or_not[oj];
    if (bi == 2 && bj != 2) bc_or_not[oj] = 3;
    if (bj == 2 && bi != 2) bc_or_not[oi] = 3;
}

// 3_...:1321-1349: subsets and BC_or_not (2 on the surface, 0 inside, 1 kept)
__global__ void k_surf_sub(StateBufs S, SurfBufs B, int *__restrict__ bc_or_not, int nt) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nt) return;
    const int o = S.orig[r];
    int sub = B.flag[r] ? 2 : 1, bc = bc_or_not[o];
    if (bc != 1) bc = 0;
    if (sub == 2 && bc != 1) bc = 0;
    else if (sub == 1 && bc != 1) bc = 2;
    else if (sub == 1 && bc == 1) sub = 2;
    bc_or_not[o] = bc;
    B.sub[r] = sub;
}

// step 4 (3_...:1353-1405): the surface normal from the edges to surface
// neighbours, oriented against the gradient of f_int; by original index
__global__ void k_surf_b(StateBufs S, SurfBufs B, double2 *__restrict__ normal, int nt) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nt) return;
    const bool on = B.sub[r] == 1;
    const float fr = B.flag[r] ? -1.f : 0.f;
    double n0 = 0.0, n1 = 0.0;
    float nb = 0.f, g0 = 0.f, g1 = 0.f;
    for (unsigned q = B.start[r]; q < B.start[r + 1]; ++q) {
        const unsigned e = B.ent[q];
        int p;
        float h1, h2;
        surf_hf(S, B, e, p, h1, h2);
        if (on && B.sub[p] == 1) {
            const unsigned k = e >> 1;
            const double2 xi = S.x[B.pi[k]], xj = S.x[B.pj[k]];
            const float xv = (float)xj.x - (float)xi.x, yv = (float)xj.y - (float)xi.y;
            n0 = n0 - (double)yv;
            n1 = n1 + (double)xv;
            nb = nb + 1.f;
        }
        const float df = (B.flag[p] ? -1.f : 0.f) - fr;
        g0 = g0 + df * h1;
        g1 = g1 + df * h2;
    }
    double2 out = make_double2(0.0, 0.0);
    if (on) {
        n0 = n0 / (double)nb;
        n1 = n1 / (double)nb;
        const float nv = (float)sqrt(n0 * n0 + n1 * n1);
        n0 = n0 / (double)nv;
        n1 = n1 / (double)nv;
        const float ps = (float)((double)g0 * n0 + (double)g1 * n1);
        out = ps < 0 ? make_double2(-n0, -n1) : make_double2(n0, n1);
    }
    normal[S.orig[r]] = out;
}

// Slot of each original particle: the outside approach writes a node's
// stress points by their original indices.
__global__ void k_inv_slots(const int *__restrict__ orig, int *__restrict__ inv, int nt2) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nt2) inv[orig[t]] = t;
}

struct ShiftParams {
    int npoints, vel_vector, shift_now;
    double r_x, r_y, disp_tol, dx;
};

// A velocity-vector offset component in (0, dx/5) or (-dx/5, 0) pushed out
// by dx/3, in the reference's mixed single / double precision (2_...:279-282).
__device__ __forceinline__ float push_out(float r, double dx) {
    if (r > 0.0f && (double)r < dx / 5) r = (float)((double)r + dx / 3.0);
    if (r < 0.0f && (double)r > -(dx / 5)) r = (float)((double)r - dx / 3.0);
    return r;
}

// shift_stress_points (2_...:244-370), the outside approach after the
// position update; one lane per node slot, writing the node's stress points
// nn + npoints * i + p (i its original index) through inv.  On shift steps
// (shift_now) the points are placed along the node velocity (vel_vector:
// +-dx/2 (cos, sin) in single precision, pushed out of (-dx/5, dx/5), only
// when the node moved more than disp_tol dx since the last shift; x_10 and
// disp_10 are kept by original index) or at the fixed offsets of 2_...:293-326
// (the second of two points takes r_x for its y offset, npoints 4 has no
// pattern there).  Every step, the first three points of a node with a
// wall-dummy partner moving faster than 0.4 (speed in single precision), or
// with fewer than two node partners, go onto the node (2_...:346-366).
__global__ void k_shift_sp(StateBufs S, const int *__restrict__ inv, const int *__restrict__ nodeinfo,
                           double2 *__restrict__ x10, double *__restrict__ disp10, ShiftParams P, int nn) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nn) return;
    const int i = S.orig[r];
    const double2 xn = S.x[r], v = S.vel[r];
    const int np = P.npoints;
    double2 q0 = xn, q1 = xn, q2 = xn;
    bool put = false;
    if (P.shift_now) {
        if (P.vel_vector) {
            const double2 o = x10[i];
            const double d0 = xn.x - o.x, d1 = xn.y - o.y;
            double d10 = sqrt(d0 * d0 + d1 * d1);
            d10 = d10 / P.dx;
            disp10[i] = d10;
            x10[i] = xn;
            const float ct = (float)(v.x / sqrt(v.x * v.x + v.y * v.y));
            const float sn = (float)(v.y / sqrt(v.x * v.x + v.y * v.y));
            const float rx = push_out((float)((P.dx / 2) * (double)ct), P.dx);
            const float ry = push_out((float)((P.dx / 2) * (double)sn), P.dx);
            if (d10 > P.disp_tol) {
                q0 = make_double2(xn.x + (double)rx, xn.y + (double)ry);
                q1 = make_double2(xn.x - (double)rx, xn.y - (double)ry);
                put = true;
            }
        } else if (np == 1) {
            q0 = make_double2(xn.x + P.r_x, xn.y + P.r_y);
            put = true;
        } else if (np == 2) {
            q0 = make_double2(xn.x + P.r_x, xn.y + P.r_y);
            q1 = make_double2(xn.x - P.r_x, xn.y - P.r_x);
            put = true;
        } else if (np == 3) {
            q0 = make_double2(xn.x, xn.y + P.r_y);
            q1 = make_double2(xn.x - P.r_x, xn.y - P.r_y);
            q2 = make_double2(xn.x + P.r_x, xn.y - P.r_y);
            put = true;
        }
    }
    const int info = nodeinfo[r];
    const float abs_vel = (float)sqrt(v.x * v.x + v.y * v.y);
    if (((info & 1) && abs_vel > 0.4f) || (float)(info >> 1) < 2.0f) {
        q0 = q1 = q2 = xn;
        put = true;
    }
    if (!put) return;
    const int base = nn + np * i;
    S.x[inv[base]] = q0;
    if (np > 1) S.x[inv[base + 1]] = q1;
    if (np > 2) S.x[inv[base + 2]] = q2;
}

// --------------------------------------------------------------- pair dump
// Pairs of the last search in discovery order (for sps_get_pairs): every
// processing slot writes its discovered pairs at off[p] + local, with the
// orientation and type of Pint_Update (3_...:1569-1608), 1-based indices.
__global__ void k_dump_pairs(const double2 *__restrict__ x, const double2 *__restrict__ x0,
                             const double *__restrict__ hsml, const int *__restrict__ cell,
                             const int *__restrict__ kstart, const int *__restrict__ orig,
                             const unsigned *__restrict__ off, const int *__restrict__ proc,
                             const Grid *__restrict__ g, double pi, int *__restrict__ pi_out,
                             int *__restrict__ pj_out, int *__restrict__ pt_out,
                             float *__restrict__ w_out, float *__restrict__ gx_out,
                             float *__restrict__ gy_out, int nn, int nt, int nt2, int as_slots,
                             unsigned long long npairs, unsigned long long mprev) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nt2) return;
    int ndivt = g->ndivt;
    int c = cell[r];
    if (c >= ndivt) return;
    const int L = ndivt + 1;
    int kr = kind_of(r, nn, nt);
    double2 xr = r < nt ? x0[r] : x[r];
    const Grid gg = *g;
    Stencil S = make_stencil(c, gg);
    const int crm_r = S.cy * gg.ndivx + S.cx;
    unsigned d = off[proc[r]];
    for (int y = S.y0; y <= S.y1; ++y)
        for (int xx = S.x0; xx <= S.x1; ++xx) {
            const int key = cell_key(xx, y, gg), crm = y * gg.ndivx + xx;
            if (crm < crm_r) continue;
            for (int ks = 0; ks < 3; ++ks)
                for (int s = kstart[ks * L + key]; s < kstart[ks * L + key + 1]; ++s) {
                    if (crm == crm_r && !(ks > kr || (ks == kr && s > r))) continue;
                    double2 xs = s < nt ? x0[s] : x[s];
                    double dx0 = xr.x - xs.x, dx1 = xr.y - xs.y, mh;
                    double rr = pair_distance(dx0, dx1, hsml[r], hsml[s], mh);
                    if (rr < 0.0) continue;
                    float w, gx, gy;
                    cubic_kernel(rr, dx0, dx1, mh, pi, w, gx, gy);
                    int a = as_slots ? r : orig[r] + 1, b = as_slots ? s : orig[s] + 1;
                    // Pint_Update: stress point or dummy first; gradient negated on swap
                    bool swap = (kr == KIND_NODE && ks != KIND_NODE) || (kr == KIND_SP && ks == KIND_DUMMY);
                    if (swap) {
                        int t = a; a = b; b = t;
                        gx = -gx; gy = -gy;
                    }
                    // as_slots: at the pair's place in the traversal order,
                    // the discovery indices [n-1 .. M_prev] then [0 .. M_prev-1]
                    const unsigned long long q =
                        !as_slots || npairs <= mprev ? d : (d >= mprev ? npairs - 1 - d : npairs - mprev + d);
                    pi_out[q] = a;
                    pj_out[q] = b;
                    pt_out[q] = pair_type(kr, ks);
                    if (w_out) w_out[q] = w;
                    gx_out[q] = gx;
                    gy_out[q] = gy;
                    d++;
                }
        }
}

// Un-permute state arrays for sps_get_state.
__global__ void k_gather_out2(const double2 *__restrict__ src, const int *__restrict__ orig,
                              double2 *__restrict__ dst, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[orig[t]] = src[t];
}
__global__ void k_gather_out4(const double4 *__restrict__ src, const int *__restrict__ orig,
                              double4 *__restrict__ dst, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[orig[t]] = src[t];
}
__global__ void k_gather_out1(const double *__restrict__ src, const int *__restrict__ orig,
                              double *__restrict__ dst, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[orig[t]] = src[t];
}
__global__ void k_gather_outi(const int *__restrict__ src, const int *__restrict__ orig,
                              int *__restrict__ dst, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[orig[t]] = src[t];
}
__global__ void k_gather_outf(const float *__restrict__ src, const int *__restrict__ orig,
                              float *__restrict__ dst, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[orig[t]] = src[t];
}
__global__ void k_scatter_in_identity(int *__restrict__ orig, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) orig[t] = t;
}

// ------------------------------------------------------------------- scans
// Exclusive scan of unsigned values, 1024 per block (256 lanes x 4).
__device__ __forceinline__ unsigned wave_incl_scan(unsigned v) {
    int lane = threadIdx.x & 63;
    for (int off = 1; off < 64; off <<= 1) {
        unsigned t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

__device__ __forceinline__ unsigned block_excl_scan(unsigned v, unsigned *total) {
    __shared__ unsigned wsum[BLOCK / 64];
    unsigned incl = wave_incl_scan(v);
    int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    unsigned base = 0, tot = 0;
    for (int k = 0; k < BLOCK / 64; ++k) {
        if (k < wid) base += wsum[k];
        tot += wsum[k];
    }
    __syncthreads();
    *total = tot;
    return base + incl - v;
}

// Element count of a scan: the host's n, or (with g) a * ndivt + b of the
// current grid, at most the host's n (the launch bound).
struct ScanCount {
    const Grid *g;
    int a, b;
    __device__ __forceinline__ long long get(long long n) const {
        if (!g) return n;
        const long long m = (long long)a * g->ndivt + b;
        return m < n ? m : n;
    }
};

__global__ void k_scan_local(const unsigned *__restrict__ in, unsigned *__restrict__ out,
                             unsigned *__restrict__ bsum, long long n_max, ScanCount sc) {
    const long long n = sc.get(n_max);
    long long base = (long long)blockIdx.x * (BLOCK * 4) + (long long)threadIdx.x * 4;
    unsigned v[4], s = 0;
    for (int k = 0; k < 4; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        s += v[k];
    }
    unsigned tot;
    unsigned ex = block_excl_scan(s, &tot);
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) out[base + k] = ex;
        ex += v[k];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ void k_scan_sums(unsigned *__restrict__ bsum, int nb_max, unsigned long long *__restrict__ total,
                            long long n_max, ScanCount sc) {
    // one block walks the block sums of the blocks that hold elements, in
    // chunks of BLOCK
    const long long nbl = (sc.get(n_max) + BLOCK * 4 - 1) / (BLOCK * 4);
    const int nb = nbl < nb_max ? (int)nbl : nb_max;
    __shared__ unsigned carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += BLOCK) {
        int i = base + threadIdx.x;
        unsigned v = i < nb ? bsum[i] : 0u;
        unsigned tot;
        unsigned ex = block_excl_scan(v, &tot);
        if (i < nb) bsum[i] = carry + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

__global__ void k_scan_add(unsigned *__restrict__ out, const unsigned *__restrict__ bsum, long long n_max,
                           ScanCount sc) {
    const long long n = sc.get(n_max);
    long long base = (long long)blockIdx.x * (BLOCK * 4) + (long long)threadIdx.x * 4;
    unsigned add = bsum[blockIdx.x];
    for (int k = 0; k < 4; ++k)
        if (base + k < n) out[base + k] += add;
}

}  // namespace sps
