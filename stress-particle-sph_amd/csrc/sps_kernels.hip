// sps_kernels.hip - hand-written HIP kernels (gfx950) of the Stress-Particle
// SPH hot path: cell binning, ordered neighbour lists, and the per-stage
// gather kernels of the RK4 step.
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
// stencil walk -- so the kernels emit each receiver's partners with a reverse
// walk (pairs d >= M) followed by a forward walk (pairs d < M), and the
// stage kernels add them one by one in that order: one lane per receiver,
// no atomics, results bitwise equal to the reference.
#include "sps_device.h"

namespace sps {

constexpr int BLOCK = 256;
constexpr int SLICE = 64;   // receivers per sliced-ELL slice (one wave)

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
    if (d0 > 0.0 || d1 > 0.0) ood[i] = 1;
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
// (one block: reduce the partials, then thread 0 derives the grid).
// Launch with exactly BLOCK threads.
__global__ void __launch_bounds__(BLOCK) k_bounds_final(const double *__restrict__ part, int nparts,
                                                        Grid *__restrict__ g, int cell_cap) {
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
}

// ------------------------------------------------------------- cell keys
// Buckets are kept in strip-major cell order: the grid is cut into vertical
// strips of STRIP cells, cells run row by row inside a strip.  Walking the
// storage (and so the receivers) up a strip keeps the partner data of three
// consecutive rows in one L2.  The reference's own visiting order (row-major
// cells, 2_...:1322-1324) is kept separately for the list order (k_proc_slot).
constexpr int STRIP = 64;

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
// Task 2 of grid_find_NEW (2_...:1277-1305): cell of every particle and
// per-(kind, cell) counts.  Out-of-domain particles go to the sentinel cell
// ndivt at the end of their kind's range.
__global__ void k_cell_count(const double2 *__restrict__ x, const int *__restrict__ ood,
                             const Grid *__restrict__ g, int *__restrict__ cellkey,
                             int *__restrict__ kcount, int nn, int nt, int nt2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt2) return;
    int ndivt = g->ndivt;
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
    int k = kind_of(i, nn, nt);
    atomicAdd(&kcount[k * (ndivt + 1) + c], 1);
}

// Scatter into per-(kind, cell) buckets (arbitrary order inside a bucket).
__global__ void k_cell_scatter(const int *__restrict__ cellkey, const int *__restrict__ orig,
                               const int *__restrict__ kstart, int *__restrict__ kfill,
                               const Grid *__restrict__ g, int *__restrict__ tmp_orig,
                               int *__restrict__ tmp_src, int *__restrict__ new_cell, int nn,
                               int nt, int nt2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt2) return;
    int ndivt = g->ndivt;
    int c = cellkey[i];
    int b = kind_of(i, nn, nt) * (ndivt + 1) + c;
    int slot = kstart[b] + atomicAdd(&kfill[b], 1);
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
    int ndivt = g->ndivt;
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= 3 * (ndivt + 1)) return;
    int lo = kstart[b], hi = kstart[b + 1];
    if (hi - lo < 2) return;
    if (hi - lo > MAX_BUCKET_SORT) {
        if (b % (ndivt + 1) != ndivt) atomicOr(err, 2);   // an in-domain cell
        return;
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

// Particles per cell in the reference's row-major visiting order.
__global__ void k_cell_totals(const int *__restrict__ kstart, const Grid *__restrict__ g,
                              unsigned *__restrict__ tot) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    int ndivt = g->ndivt;
    if (c > ndivt) return;
    if (c == ndivt) {
        tot[c] = 0;
        return;
    }
    const int L = ndivt + 1;
    int key = cell_key(c % g->ndivx, c / g->ndivx, *g);
    tot[c] = (unsigned)((kstart[key + 1] - kstart[key]) + (kstart[L + key + 1] - kstart[L + key]) +
                        (kstart[2 * L + key + 1] - kstart[2 * L + key]));
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
};

__global__ void k_permute(StateBufs dst, StateBufs src, const int *__restrict__ perm, int nn,
                          int nt, int nt2) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt2) return;
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

// Relevance of a partner kind for the receiver's lists:
//  list A: node <- stress points (type 1) and dummies (type 6);
//          stress point <- nodes (type 1) and dummies (type 9);
//          dummy <- nodes (type 6, for the dummy stress copy of 2_...:580).
//  list B: node <- nodes (type 3, artificial viscosity).
__device__ __forceinline__ bool in_list_a(int kr, int ks) {
    return kr == KIND_NODE ? ks != KIND_NODE : (kr == KIND_SP ? ks != KIND_SP : ks == KIND_NODE);
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
// order; the reference counts n_pairs with these), and its list degrees.
__global__ void k_count(const double2 *__restrict__ x, const double *__restrict__ hsml,
                        const int *__restrict__ cell, const int *__restrict__ kstart,
                        const int *__restrict__ proc, const Grid *__restrict__ g,
                        unsigned *__restrict__ cnt, int *__restrict__ degA, int *__restrict__ degB,
                        Split *__restrict__ split, int nn, int nt, int nt2) {
    int r = xcd_block() * blockDim.x + threadIdx.x;
    unsigned fwd = 0;
    int da = 0, db = 0;
    unsigned t0 = 0, t1 = 0, t2 = 0, t3 = 0, t6 = 0, t9 = 0;   // pairs by pint_type
    int ndivt = g->ndivt;
    if (r < nt2 && cell[r] < ndivt) {
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
                        if (in_list_a(kr, ks)) da++;
                        if (kr == KIND_NODE && ks == KIND_NODE) db++;
                    }
                }
            }
        }
        cnt[proc[r]] = fwd;
    }
    if (r < nt2) {
        degA[r] = da;
        if (r < nn) degB[r] = db;
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

// Widest row of each 64-receiver slice.
__global__ void k_slice_width(const int *__restrict__ deg, int n, unsigned *__restrict__ width) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    int v = r < n ? deg[r] : 0;
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_down(v, off, 64));
    if ((threadIdx.x & 63) == 0 && (r >> 6) < (n + 63) / 64) width[r >> 6] = (unsigned)v * SLICE;
}

// Fresh split record: M is the list high-water mark before this search.
__global__ void k_split_reset(Split *__restrict__ split, unsigned long long M) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    split->n = 0;
    split->M = M;
    split->sstar = -1;
    split->pstar = -1;
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
}

// First reversed partner of the split slot: its (M - off)-th discovery in
// processing order (2_...:1322-1383 visit partners in that order).
__global__ void k_split_partner(const double2 *__restrict__ x, const double *__restrict__ hsml,
                                const int *__restrict__ cell, const int *__restrict__ kstart,
                                const int *__restrict__ proc, const unsigned *__restrict__ off,
                                const int *__restrict__ proc_to_slot, const Grid *__restrict__ g,
                                Split *__restrict__ split, int nn, int nt) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
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

// ------------------------------------------------------------------- build
// Emit each receiver's partners in reference list order into the sliced-ELL
// lists, with w and d W / d x_receiver rounded to single precision.
struct Lists {
    int *a_s;
    float *a_w, *a_gx, *a_gy;
    const unsigned *a_off;
    int *b_s;
    float *b_gx, *b_gy;
    const unsigned *b_off;
};

__device__ __forceinline__ bool pair_reversed(int pr, int ps, const unsigned *__restrict__ cnt,
                                              const unsigned *__restrict__ off,
                                              const Split &sp) {
    // discoverer: the end visited first (smaller processing slot)
    int pd = pr < ps ? pr : ps, po = pr < ps ? ps : pr;
    unsigned long long o = off[pd], c = cnt[pd];
    if (o >= sp.M) return true;
    if (o + c <= sp.M) return false;
    return po >= sp.pstar;   // pd is the split slot
}

__global__ void k_build(const double2 *__restrict__ x, const double *__restrict__ hsml,
                        const int *__restrict__ cell, const int *__restrict__ kstart,
                        const int *__restrict__ proc, const unsigned *__restrict__ cnt,
                        const unsigned *__restrict__ off, const Grid *__restrict__ g,
                        const Split *__restrict__ split, Lists Lz, double pi, int nn, int nt,
                        int nt2) {
    int r = xcd_block() * blockDim.x + threadIdx.x;
    if (r >= nt2) return;
    int ndivt = g->ndivt;
    int c = cell[r];
    if (c >= ndivt) return;
    const int L = ndivt + 1;
    const Split sp = *split;
    const bool any_rev = sp.n > sp.M;             // some pairs have d >= M
    const bool any_fwd = sp.M > 0;                // some pairs have d < M
    const bool mixed = any_rev && any_fwd;
    int kr = kind_of(r, nn, nt);
    int pr = proc[r];
    double2 xr = x[r];
    double hr = hsml[r];
    const Grid gg = *g;
    Stencil S = make_stencil(c, gg);
    unsigned base_a = Lz.a_off[r >> 6] + (r & 63);
    unsigned base_b = r < nn ? Lz.b_off[r >> 6] + (r & 63) : 0;
    int na = 0, nb = 0;
    const int nx = S.x1 - S.x0 + 1, ny = S.y1 - S.y0 + 1;
    for (int pass = 0; pass < 2; ++pass) {
        const bool rev = pass == 0;
        if (rev && !any_rev) continue;
        if (!rev && !any_fwd) continue;
        for (int ci = 0; ci < nx * ny; ++ci) {
            // cells in ascending (forward) or descending (reverse) row-major order
            const int cj = rev ? nx * ny - 1 - ci : ci;
            const int y = S.y0 + cj / nx, xx = S.x0 + cj % nx;
            const int key = cell_key(xx, y, gg);
            for (int ki = 0; ki < 3; ++ki) {
                int ks = rev ? 2 - ki : ki;
                bool la = in_list_a(kr, ks), lb = kr == KIND_NODE && ks == KIND_NODE;
                if (!la && !lb) continue;
                int lo = kstart[ks * L + key], hi = kstart[ks * L + key + 1];
                int cntk = hi - lo;
                for (int q = 0; q < cntk; ++q) {
                    int s = rev ? hi - 1 - q : lo + q;
                    if (s == r) continue;
                    double2 xs = x[s];
                    double dx0 = xr.x - xs.x, dx1 = xr.y - xs.y;
                    double mh;
                    double rr = pair_distance(dx0, dx1, hr, hsml[s], mh);
                    if (rr < 0.0) continue;
                    if (mixed && pair_reversed(pr, proc[s], cnt, off, sp) != rev) continue;
                    float w, gx, gy;
                    cubic_kernel(rr, dx0, dx1, mh, pi, w, gx, gy);
                    if (la) {
                        unsigned e = base_a + (unsigned)na * SLICE;
                        Lz.a_s[e] = s;
                        Lz.a_w[e] = w;
                        Lz.a_gx[e] = gx;
                        Lz.a_gy[e] = gy;
                        na++;
                    } else {
                        unsigned e = base_b + (unsigned)nb * SLICE;
                        Lz.b_s[e] = s;
                        Lz.b_gx[e] = gx;
                        Lz.b_gy[e] = gy;
                        nb++;
                    }
                }
            }
        }
    }
}

// ------------------------------------------------------------ step kernels
struct Work {
    double2 *vx0, *x0, *vel0, *rk_vel, *rhs2, *f_bound, *art_visc;
    double4 *stress0, *rk_stress, *rhs1, *grad_u;
    double *rk_dev, *ddev;
};

// Start of a step: vx0, x0 (2_...:114) and the RK4 prologue (2_...:681-690).
__global__ void k_rk_init(StateBufs S, Work W, int nn, int nt, int nt2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt2) return;
    if (i < nt) {
        W.vx0[i] = S.vel[i];
        W.x0[i] = S.x[i];
        W.stress0[i] = S.stress[i];
        W.vel0[i] = S.vel[i];
        W.rk_stress[i] = make_double4(0, 0, 0, 0);
        W.rhs1[i] = make_double4(0, 0, 0, 0);
        W.rk_dev[i] = 0.0;
    }
    if (i < nn) {
        W.rk_vel[i] = make_double2(0, 0);
        W.rhs2[i] = make_double2(0, 0);
        W.art_visc[i] = make_double2(0, 0);
    }
    S.vel[i] = make_double2(0, 0);
    S.stress[i] = make_double4(0, 0, 0, 0);
}

// Stage state (2_...:700-701) and the first adapt_stress2 of the stage.
__global__ void k_stage_begin(StateBufs S, Work W, double c1, int adapt, double tanfi, double coh,
                              int nn, int nt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
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

// stress_point_update (2_...:403-482) for node receivers: CSPM-normalised
// stress and plastic strain from the stress points, then adapt_stress2.
__global__ void k_spu_nodes(StateBufs S, Work W, const int *__restrict__ degA, Lists Lz, int adapt,
                            double tanfi, double coh, int nn, int nt) {
    int r = xcd_block() * blockDim.x + threadIdx.x;
    if (r >= nn) return;
    unsigned e = Lz.a_off[r >> 6] + (r & 63);
    int deg = degA[r];
    double t0 = 0, t1 = 0, t2 = 0, t3 = 0, tv = 0, norm = 0;
#pragma unroll 4
    for (int k = 0; k < deg; ++k, e += SLICE) {
        int s = Lz.a_s[e];
        if (s >= nt) continue;                      // type 6 entry
        double w = (double)Lz.a_w[e];
        double ms = S.mass[s], rs = S.rho[s];
        double h1 = (ms / rs) * w;
        double4 st = S.stress[s];
        t0 = t0 + st.x * h1;
        t1 = t1 + st.y * h1;
        t2 = t2 + st.z * h1;
        t3 = t3 + st.w * h1;
        tv = tv + S.ivar1[s] * h1;
        norm = norm + (w * ms) / rs;
    }
    double4 s = S.stress[r];
    if (norm != 0) {
        s = make_double4(t0 / norm, t1 / norm, t2 / norm, t3 / norm);
        S.ivar1[r] = tv / norm;
    } else {
        S.vel[r] = make_double2(0, 0);
    }
    if (adapt) adapt_stress(s, tanfi, coh);
    S.stress[r] = s;
}

// stress_point_update for stress-point receivers: CSPM-normalised velocity
// from the nodes, then adapt_stress2 of the stress point.
__global__ void k_spu_sps(StateBufs S, Work W, const int *__restrict__ degA, Lists Lz, int adapt,
                          double tanfi, double coh, int nn, int nt) {
    int r = nn + xcd_block() * blockDim.x + threadIdx.x;
    if (r >= nt) return;
    unsigned e = Lz.a_off[r >> 6] + (r & 63);
    int deg = degA[r];
    double v0 = 0, v1 = 0, norm = 0;
#pragma unroll 4
    for (int k = 0; k < deg; ++k, e += SLICE) {
        int s = Lz.a_s[e];
        if (s >= nn) continue;                      // type 9 entry
        double w = (double)Lz.a_w[e];
        double ms = S.mass[s], rs = S.rho[s];
        double h2 = (ms / rs) * w;
        double2 vs = S.vel[s];
        v0 = v0 + vs.x * h2;
        v1 = v1 + vs.y * h2;
        norm = norm + (w * ms) / rs;
    }
    if (norm != 0) S.vel[r] = make_double2(v0 / norm, v1 / norm);
    if (adapt) {
        double4 s = S.stress[r];
        adapt_stress(s, tanfi, coh);
        S.stress[r] = s;
    }
}

// get_derivatives (2_...:487-648) for stress-point receivers, fused with
// plastic_terms (loop 2_...:730-734), the Jaumann spin (2_...:751-757) and the
// stress right-hand side and RK accumulation (2_...:762, 772).
__global__ void k_deriv_sps(StateBufs S, Work W, const int *__restrict__ degA, Lists Lz,
                            DevParams P, double time_sph, double f2, int nn, int nt) {
    int r = nn + xcd_block() * blockDim.x + threadIdx.x;
    if (r >= nt) return;
    unsigned e = Lz.a_off[r >> 6] + (r & 63);
    int deg = degA[r];
    double g00 = 0, g01 = 0, g10 = 0, g11 = 0;   // grad1_tmp(a,b): d v_a / d x_b
    double A0 = 0, A1 = 0, A2 = 0, A3 = 0;
    double2 xr = S.x[r], vr = S.vel[r];
#pragma unroll 4
    for (int k = 0; k < deg; ++k, e += SLICE) {
        int s = Lz.a_s[e];
        double ms = S.mass[s], rs = S.rho[s];
        if (s < nn) {                               // type 1: pair_i = this SP
            double gx = (double)Lz.a_gx[e], gy = (double)Lz.a_gy[e];
            double h1 = gx * ms / rs;
            double h2 = gy * ms / rs;
            double2 vs = S.vel[s];
            g00 = g00 + (vs.x - vr.x) * h1;
            g01 = g01 + (vs.x - vr.x) * h2;
            g10 = g10 + (vs.y - vr.y) * h1;
            g11 = g11 + (vs.y - vr.y) * h2;
            if (P.cspm) {
                double2 xs = S.x[s];
                A0 = A0 + (xs.x - xr.x) * h1;
                A1 = A1 + (xs.y - xr.y) * h1;
                A2 = A2 + (xs.x - xr.x) * h2;
                A3 = A3 + (xs.y - xr.y) * h2;
            }
        } else {                                    // type 9: pair_i = dummy
            double gx = (double)(-Lz.a_gx[e]), gy = (double)(-Lz.a_gy[e]);
            double wall = (double)S.wallpos[s];
            double2 xs = S.x[s];
            double da, db;
            if (S.horiz[s] == 1.0f) {
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
    }
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
    double4 st = S.stress[r];
    double4 Gs;
    double der1;
    double fdr = S.fdr[r];
    plastic_terms(P, st, gu, fdr, time_sph, Gs, der1);
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

// get_derivatives for node receivers, fused with gravity_force
// (3_...:2640-2702), boundary_forces (2_...:1117-1161), artificial_viscosity
// (2_...:826-904) and the momentum right-hand side and RK accumulation
// (2_...:763-764, 769).
constexpr int MAX_WALL_TERMS = 16;

// boundary_forces for one node: the nonzero terms of its inner-layer dummies
// in ascending dummy index (the reference's j loop), found by repeated
// minimum selection over the node's dummy partners (a handful near walls).
__device__ __forceinline__ double2 wall_force(const StateBufs &S, const Lists &Lz, unsigned e0, int deg,
                                              double2 xr, double hr, const DevParams &P,
                                              int *__restrict__ err) {
    double2 fb = make_double2(0.0, 0.0);
    int last = -1;
    for (int it = 0; it <= MAX_WALL_TERMS; ++it) {
        int best = 0x7fffffff;
        float b0 = 0.0f, b1 = 0.0f;
        unsigned e = e0;
        for (int k = 0; k < deg; ++k, e += SLICE) {
            int s = Lz.a_s[e];
            if (s < P.nt) continue;
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
        if (best == 0x7fffffff) return fb;
        if (it == MAX_WALL_TERMS) {
            atomicOr(err, 1);
            return fb;
        }
        fb.x = fb.x + (double)b0;
        fb.y = fb.y + (double)b1;
        last = best;
    }
    return fb;
}

__global__ void k_deriv_nodes(StateBufs S, Work W, const int *__restrict__ degA,
                              const int *__restrict__ degB, Lists Lz, DevParams P, double gx_grav,
                              double gy_grav, double f2, int *__restrict__ err, int nn, int nt) {
    int r = xcd_block() * blockDim.x + threadIdx.x;
    if (r >= nn) return;
    double2 xr = S.x[r], vr = S.vel[r];
    double4 sr = S.stress[r];
    double rhor = S.rho[r], hr = S.hsml[r];
    double rr2 = rhor * rhor;
    const double or0 = sr.x / rr2, or1 = sr.y / rr2, or2 = sr.z / rr2;   // own sigma / rho^2
    double q00 = 0, q01 = 0, q10 = 0, q11 = 0, q20 = 0, q21 = 0;  // grad2_tmp(s,b)
    double A0 = 0, A1 = 0, A2 = 0, A3 = 0;
    int ndum = 0;
    unsigned e = Lz.a_off[r >> 6] + (r & 63);
    int deg = degA[r];
#pragma unroll 4
    for (int k = 0; k < deg; ++k, e += SLICE) {
        int s = Lz.a_s[e];
        double ms = S.mass[s], rs = S.rho[s];
        double gx = (double)(-Lz.a_gx[e]), gy = (double)(-Lz.a_gy[e]);   // pair_i is the partner
        double rs2 = rs * rs;
        if (s < nt) {                                // type 1
            double4 ss = S.stress[s];
            const double p0 = ss.x / rs2, p1 = ss.y / rs2, p2 = ss.z / rs2;   // sigma_s / rho_s^2
            double a;
            a = gx * (or0 + p0); q00 = q00 - ms * a;
            a = gy * (or0 + p0); q01 = q01 - ms * a;
            a = gx * (or1 + p1); q10 = q10 - ms * a;
            a = gy * (or1 + p1); q11 = q11 - ms * a;
            a = gx * (or2 + p2); q20 = q20 - ms * a;
            a = gy * (or2 + p2); q21 = q21 - ms * a;
            if (P.cspm) {
                double h1b = -(gx * ms / rs);
                double h2b = -(gy * ms / rs);
                double2 xs = S.x[s];
                A0 = A0 + (xs.x - xr.x) * h1b;
                A1 = A1 + (xs.y - xr.y) * h1b;
                A2 = A2 + (xs.x - xr.x) * h2b;
                A3 = A3 + (xs.y - xr.y) * h2b;
            }
        } else {                                     // type 6: dummy stress := node stress
            double d0 = sr.x / rs2, d1 = sr.y / rs2, d2 = sr.z / rs2;
            double a;
            a = gx * (or0 + d0); q00 = q00 - ms * a;
            a = gy * (or0 + d0); q01 = q01 - ms * a;
            a = gx * (or1 + d1); q10 = q10 - ms * a;
            a = gy * (or1 + d1); q11 = q11 - ms * a;
            a = gx * (or2 + d2); q20 = q20 - ms * a;
            a = gy * (or2 + d2); q21 = q21 - ms * a;
            ndum++;
        }
    }
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
    double d2x = -(q00 + q21);
    double d2y = -(q20 + q11);
    // gravity_force: G = factg*cgrav - DampingTG*v
    double Gx = gx_grav - P.damping * vr.x;
    double Gy = gy_grav - P.damping * vr.y;
    // boundary_forces: nonzero terms in ascending dummy index
    double2 fb = W.f_bound[r];
    if (P.do_bforce) {
        fb = ndum ? wall_force(S, Lz, Lz.a_off[r >> 6] + (r & 63), degA[r], xr, hr, P, err)
                  : make_double2(0.0, 0.0);
        W.f_bound[r] = fb;
    }
    // artificial_viscosity over node-node pairs, single-precision sums
    double2 av = make_double2(0.0, 0.0);
    if (P.do_av) {
        float ax = 0.0f, ay = 0.0f;
        float visc = 0.0f;
        const float c = 600;
        unsigned eb = Lz.b_off[r >> 6] + (r & 63);
        int degb = degB[r];
#pragma unroll 4
        for (int k = 0; k < degb; ++k, eb += SLICE) {
            int s = Lz.b_s[eb];
            double2 xs = S.x[s], vs = S.vel[s];
            float xij = (float)(xr.x - xs.x);
            float yij = (float)(xr.y - xs.y);
            float hh = (float)(0.5 * (hr + S.hsml[s]));
            float rho2 = (float)(0.5 * (rhor + S.rho[s]));
            float div_u = (float)((double)xij * (vr.x - vs.x));
            div_u = (float)((double)div_u + (double)yij * (vr.y - vs.y));
            float sr_ = sqrtf(xij * xij + yij * yij);
            float theta = (hh * div_u) / (sr_ * sr_ + 0.01f * (hh * hh));
            if (div_u < 0) {
                visc = (float)((-(P.alpha * (double)c * (double)theta) + P.beta * (double)(theta * theta)) /
                               (double)rho2);
            } else if (div_u >= 0) {
                visc = 0;
            }
            double ms = S.mass[s];
            ax = (float)((double)ax + (double)(visc * Lz.b_gx[eb]) * ms);
            ay = (float)((double)ay + (double)(visc * Lz.b_gy[eb]) * ms);
        }
        av = make_double2(-(double)ax, -(double)ay);
        W.art_visc[r] = av;
    }
    double2 rh;
    rh.x = -d2x + Gx + av.x + fb.x + 0.0;
    rh.y = -d2y + Gy + av.y + fb.y + 0.0;
    W.rhs2[r] = rh;
    double2 rk = W.rk_vel[r];
    W.rk_vel[r] = make_double2(rk.x + f2 * rh.x, rk.y + f2 * rh.y);
}

// The type-6 loop of get_derivatives overwrites each dummy's stress with
// that of its node partners in list order (2_...:580); the last one stays.
__global__ void k_dummy_stress(StateBufs S, const int *__restrict__ degA, Lists Lz, int nt,
                               int nt2) {
    int r = nt + blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nt2) return;
    int deg = degA[r];
    if (deg == 0) return;
    unsigned e = Lz.a_off[r >> 6] + (r & 63) + (unsigned)(deg - 1) * SLICE;
    S.stress[r] = S.stress[Lz.a_s[e]];
}

// End of RK4 (2_...:784-799) and update_strain (3_...:1847-1863).
__global__ void k_rk_final(StateBufs S, Work W, double dt, int adapt, double tanfi, double coh,
                           int nn, int nt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    double c = dt / 6;
    if (i < nn) {
        double2 v0 = W.vel0[i], rk = W.rk_vel[i];
        S.vel[i] = make_double2(v0.x + c * rk.x, v0.y + c * rk.y);
        if (adapt) {
            double4 s = S.stress[i];
            adapt_stress(s, tanfi, coh);
            S.stress[i] = s;
        }
        W.ddev[i] = W.rk_dev[i] / 6;
    } else {
        double4 s0 = W.stress0[i], rk = W.rk_stress[i];
        double4 s = make_double4(s0.x + c * rk.x, s0.y + c * rk.y, s0.z + c * rk.z, s0.w + c * rk.w);
        if (adapt) adapt_stress(s, tanfi, coh);
        S.stress[i] = s;
        double dd = W.rk_dev[i] / 6;
        W.ddev[i] = dd;
        S.ivar1[i] = S.ivar1[i] + dt * dd;
    }
}

// Position update (2_...:140-147) through the single-precision vel_half, and
// the displacement of the nodes (2_...:171, 179).
__global__ void k_position(StateBufs S, Work W, double dt, int update_x, int nn, int nt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    double2 v0 = W.vx0[i], v = S.vel[i];
    if (update_x) {
        float hx = (float)(0.5 * (v0.x + v.x));
        float hy = (float)(0.5 * (v0.y + v.y));
        double2 x0 = W.x0[i];
        double2 xn = make_double2(x0.x + (double)hx * dt, x0.y + (double)hy * dt);
        S.x[i] = xn;
        if (i < nn) {
            double2 x00 = S.x00[i];
            S.displ[i] = make_double2(xn.x - x00.x, xn.y - x00.y);
        }
    } else if (i < nn) {
        double2 d = S.displ[i];
        S.displ[i] = make_double2(d.x + 0.5 * (v0.x + v.x) * dt, d.y + 0.5 * (v0.y + v.y) * dt);
    }
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
                             float *__restrict__ gy_out, int nn, int nt, int nt2) {
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
                    int a = orig[r] + 1, b = orig[s] + 1;
                    // Pint_Update: stress point or dummy first; gradient negated on swap
                    bool swap = (kr == KIND_NODE && ks != KIND_NODE) || (kr == KIND_SP && ks == KIND_DUMMY);
                    if (swap) {
                        int t = a; a = b; b = t;
                        gx = -gx; gy = -gy;
                    }
                    pi_out[d] = a;
                    pj_out[d] = b;
                    pt_out[d] = pair_type(kr, ks);
                    w_out[d] = w;
                    gx_out[d] = gx;
                    gy_out[d] = gy;
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

__global__ void k_scan_local(const unsigned *__restrict__ in, unsigned *__restrict__ out,
                             unsigned *__restrict__ bsum, long long n) {
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

__global__ void k_scan_sums(unsigned *__restrict__ bsum, int nb, unsigned long long *__restrict__ total) {
    // one block walks the block sums in chunks of BLOCK
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

__global__ void k_scan_add(unsigned *__restrict__ out, const unsigned *__restrict__ bsum, long long n) {
    long long base = (long long)blockIdx.x * (BLOCK * 4) + (long long)threadIdx.x * 4;
    unsigned add = bsum[blockIdx.x];
    for (int k = 0; k < 4; ++k)
        if (base + k < n) out[base + k] += add;
}

}  // namespace sps
