/*
 * sps_oracle.c - CPU oracle of the Stress-Particle SPH hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product and the timed CPU
 * baseline.  It restates code/2_SPH_main_2018.f90 and the hot-path routines of
 * code/3_SPH_material_2018.f90 in plain C, one function per reference routine,
 * with the reference's own algorithm: the pair list is built in discovery
 * order, traversed in the order the reference's linked list has, and every
 * pair loop scatters into both endpoints.  Mixed single/double precision
 * follows Fortran's rules (a default-real operand met by a real(DP) one is
 * widened, the operation is done in DP, the result is rounded when it is
 * assigned to a default-real variable).  Build with -ffp-contract=off.
 *
 * Pinning (see tests/test_oracle_golden.py): every routine is checked
 * bit-for-bit against dumps of the reference Fortran compiled from source
 * (oracle/_ref), and against closed forms on small inputs.  Parity unpinned:
 * with ifsigman = 1 and wall dummies, that the first two BC calls of a step
 * read bc_int from the previous step's list (the reference's order; no
 * fixture has a surface node gain or lose a wall-dummy pair between steps).
 *
 * Threads: with OMP_NUM_THREADS above 1 the pair loops are split by the
 * particle that receives each term (each thread walks, in list order, the
 * pairs touching its own particles), so every particle sums the same terms
 * in the same order and the results are bitwise those of one thread
 * (tests/test_oracle_threads.py).  One thread is the reference's schedule,
 * and the one used below 20,000 particles (SPS_ORACLE_MIN_PARALLEL).
 *
 * Two stated departures, neither observable:
 *  - the dead store `adapt_stress = 0` (code/3_...:1961) is not executed;
 *  - boundary_forces (code/2_...:1117-1161) visits, for each node, only the
 *    inner-layer dummies whose double-precision offset can give a nonzero
 *    term, in ascending dummy index; every skipped term is an exact +-0.
 */
#include "sps_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdarg.h>

static _Thread_local char g_err[512];

static int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

const char *sps_last_error(void) { return g_err; }
const char *sps_backend(void) { return "oracle-cpu"; }

struct sps_handle {
    sps_params p;
    int nn, ns, nd, nd2, nt, nt2, niv;
    int have_state;
    int64_t steps;
    /* state (Fortran layouts, 0-based) */
    double *x, *vel, *stress, *mass, *rho, *hsml, *ivars, *fdr, *x00, *displ;
    double *grad_u, *ddev, *f_bound, *art_visc;
    double *x10, *disp10;   /* outside approach, vel_vector (x_10, disp_10) */
    float *n_int;           /* node-node pairs of each node (outside approach) */
    int32_t *bc_int;        /* node has a node-dummy pair (outside approach) */
    /* prescribed BCs (sps_set_bcs) */
    int bcs_set, nbc, ncurves;
    double bc_list[8 * SPS_MAX_BCS];
    int32_t *bc_info, *bc_or_not;   /* bc_or_not: every particle (Get_BCs_on_node) */
    int ifsigman;
    /* free-surface detection (get_nodes_on_free_surface): the unit normals
     * (double, as the reference's global) and the subsets of every particle */
    double *normal;
    int32_t *subset;
    int32_t curve_npts[SPS_MAX_CURVES];
    double curve_t[SPS_MAX_TCURVE * SPS_MAX_CURVES];
    float curve_f[SPS_MAX_TCURVE * SPS_MAX_CURVES];
    int32_t *itype, *ood;
    float *wallpos, *horiz;
    /* start-of-step copies (2_...:114) */
    double *x0, *vx0;
    /* pair list in discovery order + traversal order */
    int64_t pcap, n_pairs, m_pairs, pairs_total;
    int32_t *pi, *pj, *pt;
    float *pw, *pgx, *pgy;
    int64_t *order;
    int ndivx, ndivy;
    /* grid scratch */
    int *which_cell, *cell_list;
    int *cell_info; /* 3 per cell: count, start, fill */
    size_t cell_cap;
    /* near-wall dummy grid for boundary_forces */
    double bw_cell, bw_x0, bw_y0;
    int bw_nx, bw_ny;
    int *bw_start, *bw_list;
    /* RK4 scratch */
    double *vel0, *stress0, *rk_vel, *rk_stress, *rk_dev, *rhs1, *rhs2;
    double *divf1, *divf2, *source_stress, *g_grav, *spin;
    double *grad1, *grad2, *ae, *dummy_vel, *cspm_norm, *vel_tmp, *stress_tmp, *ivar_tmp;
    /* plastic trial-state log (sps_set_plastic_log) */
    int64_t log_cap, log_n;
    double *plog;
    /* a material per particle (sps_set_materials): rows of 10 doubles
     * tan(phi), c, E, nu, D11, D22, D12, D33, D41, D42 */
    int32_t *mat;
    double *mtab;
    float *av_tmp;
    /* the integrator (sps_set_integrator): RK4, or leapfrog with the
     * velocities at half steps (lf_first: the next step starts from a state
     * whose velocities are at a whole step and kicks by dt/2) */
    int integrator, lf_first;
    /* several threads (OMP_NUM_THREADS > 1): each owns a contiguous range of
     * particle indices and walks, in list order, the pairs that touch it
     * (sub[t]), updating only its own particles */
    int nthreads;
    int64_t **sub, *sub_n, *sub_cap;
};

/* ------------------------------------------------------------------ utils */

static void *xcalloc(size_t n, size_t sz)
{
    if (n == 0) n = 1;
    return calloc(n, sz);
}

static int alloc_all(struct sps_handle *h)
{
    int nt2 = h->nt2, nt = h->nt, nn = h->nn;
#define A(ptr, n, T) do { h->ptr = (T *)xcalloc((size_t)(n), sizeof(T)); if (!h->ptr) return -1; } while (0)
    A(x, 2 * nt2, double); A(vel, 2 * nt2, double); A(stress, 4 * nt2, double);
    A(mass, nt2, double); A(rho, nt2, double); A(hsml, nt2, double);
    A(ivars, (size_t)h->niv * nt, double); A(fdr, nt, double); A(x00, 2 * nt2, double);
    A(displ, 2 * nn, double); A(grad_u, 4 * nt, double); A(ddev, nt, double);
    A(f_bound, 2 * nn, double); A(art_visc, 2 * nn, double);
    A(itype, nt2, int32_t); A(ood, nt2, int32_t);
    A(wallpos, nt2, float); A(horiz, nt2, float);
    A(x0, 2 * nt, double); A(vx0, 2 * nt, double);
    A(which_cell, nt2, int); A(cell_list, nt2, int);
    A(vel0, 2 * nt, double); A(stress0, 4 * nt, double);
    A(rk_vel, 2 * nt, double); A(rk_stress, 4 * nt, double); A(rk_dev, nt, double);
    A(rhs1, 4 * nt, double); A(rhs2, 2 * nt, double);
    A(divf1, 4 * nt, double); A(divf2, 2 * nt, double);
    A(source_stress, 4 * nt, double); A(g_grav, 2 * nn, double); A(spin, 4 * nt, double);
    A(grad1, 4 * nt, double); A(grad2, 6 * nt, double); A(ae, 5 * nt, double);
    A(dummy_vel, 2 * nt2, double); A(cspm_norm, nt, double);
    A(vel_tmp, 2 * nt, double); A(stress_tmp, 4 * nt, double); A(ivar_tmp, nt, double);
    A(av_tmp, 2 * nn, float);
    A(x10, 2 * nn, double); A(disp10, nn, double); A(n_int, nn, float); A(bc_int, nn, int32_t);
    A(bc_info, 8 * nn, int32_t); A(bc_or_not, nt, int32_t);
    A(normal, 2 * nt, double); A(subset, nt, int32_t);
#undef A
    return 0;
}

static void free_all(struct sps_handle *h)
{
    void *ptrs[] = {h->x, h->vel, h->stress, h->mass, h->rho, h->hsml, h->ivars, h->fdr,
                    h->x00, h->displ, h->grad_u, h->ddev, h->f_bound, h->art_visc,
                    h->itype, h->ood, h->wallpos, h->horiz, h->x0, h->vx0,
                    h->pi, h->pj, h->pt, h->pw, h->pgx, h->pgy, h->order,
                    h->which_cell, h->cell_list, h->cell_info, h->bw_start, h->bw_list,
                    h->vel0, h->stress0, h->rk_vel, h->rk_stress, h->rk_dev, h->rhs1,
                    h->rhs2, h->divf1, h->divf2, h->source_stress, h->g_grav, h->spin,
                    h->grad1, h->grad2, h->ae, h->dummy_vel, h->cspm_norm, h->vel_tmp,
                    h->stress_tmp, h->ivar_tmp, h->av_tmp, h->x10, h->disp10, h->n_int, h->bc_int,
                    h->bc_info, h->bc_or_not, h->normal, h->subset};
    for (size_t k = 0; k < sizeof ptrs / sizeof ptrs[0]; ++k) free(ptrs[k]);
    if (h->sub)
        for (int t = 0; t < h->nthreads; ++t) free(h->sub[t]);
    free(h->sub);
    free(h->sub_n);
    free(h->sub_cap);
}

/* ------------------------------------------------------- threads */
/* The restatement runs on one thread as the reference does.  With several
 * (OpenMP), every pair loop keeps the reference's order for each particle:
 * thread t owns the particles [lo_t, hi_t) and walks the pairs that touch
 * them in list order, applying only the updates of its own particles, so
 * each particle receives the same terms in the same order as on one thread
 * and the results are the same bit for bit; loops over particles are split
 * between the threads.  The pair lists are the search's, gathered per
 * thread once per step (pair_sublists). */
static inline int own_lo(const struct sps_handle *h, int t) { return (int)((int64_t)h->nt2 * t / h->nthreads); }
static inline int own_hi(const struct sps_handle *h, int t) { return (int)((int64_t)h->nt2 * (t + 1) / h->nthreads); }

static int pair_sublists(struct sps_handle *h)
{
    if (h->nthreads <= 1) return 0;
    if (!h->sub) {
        h->sub = calloc((size_t)h->nthreads, sizeof *h->sub);
        h->sub_n = calloc((size_t)h->nthreads, sizeof *h->sub_n);
        h->sub_cap = calloc((size_t)h->nthreads, sizeof *h->sub_cap);
        if (!h->sub || !h->sub_n || !h->sub_cap) return fail(SPS_ENOMEM, "thread pair lists");
    }
    int bad = 0;
#pragma omp parallel num_threads(h->nthreads) reduction(| : bad)
    {
        const int t = omp_get_thread_num(), lo = own_lo(h, t), hi = own_hi(h, t);
        int64_t n = 0;
        for (int64_t q = 0; q < h->n_pairs; ++q) {
            const int64_t d = h->order[q];
            const int i = h->pi[d], j = h->pj[d];
            n += (i >= lo && i < hi) || (j >= lo && j < hi);
        }
        if (n > h->sub_cap[t]) {
            int64_t *a = realloc(h->sub[t], (size_t)(n ? n : 1) * sizeof *a);
            if (a) {
                h->sub[t] = a;
                h->sub_cap[t] = n;
            } else {
                bad = 1;
            }
        }
        if (n <= h->sub_cap[t]) {
            int64_t k = 0;
            for (int64_t q = 0; q < h->n_pairs; ++q) {
                const int64_t d = h->order[q];
                const int i = h->pi[d], j = h->pj[d];
                if ((i >= lo && i < hi) || (j >= lo && j < hi)) h->sub[t][k++] = d;
            }
            h->sub_n[t] = k;
        }
    }
    return bad ? fail(SPS_ENOMEM, "thread pair lists") : 0;
}

/* Run body(h, d, lo, hi, state) over the pairs in list order: on one thread
 * every pair with the whole index range, on several each thread its own
 * list and range. */
#define PAIR_LOOP(h, body, STATE_T, state_init)                                              \
    do {                                                                                     \
        if ((h)->nthreads <= 1) {                                                            \
            STATE_T st_ = state_init;                                                        \
            for (int64_t q_ = 0; q_ < (h)->n_pairs; ++q_) body((h), (h)->order[q_], 0, (h)->nt2, &st_); \
        } else {                                                                             \
            _Pragma("omp parallel num_threads((h)->nthreads)")                              \
            {                                                                                \
                const int t_ = omp_get_thread_num();                                         \
                const int lo_ = own_lo((h), t_), hi_ = own_hi((h), t_);                      \
                STATE_T st_ = state_init;                                                    \
                for (int64_t k_ = 0; k_ < (h)->sub_n[t_]; ++k_) body((h), (h)->sub[t_][k_], lo_, hi_, &st_); \
            }                                                                                \
        }                                                                                    \
    } while (0)

static inline int owns(int p, int lo, int hi) { return p >= lo && p < hi; }

/* a loop over particles split between the threads (each iteration writes
 * only its own particle's values) */
#define OMP_FOR _Pragma("omp parallel for num_threads(h->nthreads > 0 ? h->nthreads : 1)")

/* ------------------------------------------------------- Check_Out_Domain */
/* code/2_SPH_main_2018.f90:1170-1194 (sticky flag).  A non-finite
 * coordinate also flags the particle (dxx > 0 is false for NaN, and the
 * reference's grid_find then takes int() of NaN, which Fortran leaves
 * undefined): the search never sees it. */
static void check_out_domain(struct sps_handle *h)
{
#pragma omp parallel for num_threads(h->nthreads > 0 ? h->nthreads : 1)
    for (int i = 0; i < h->nt2; ++i)
        for (int d = 0; d < 2; ++d) {
            double dxx = (h->x[2 * i + d] - h->p.xmin_domain[d]) *
                         (h->x[2 * i + d] - h->p.xmax_domain[d]);
            if (!(dxx <= 0.0)) h->ood[i] = 1;
        }
}

/* ----------------------------------------------------------------- kernel */
/* code/2_SPH_main_2018.f90:1440-1492, skf = 1, ndimn = 2.  Computed in DP
 * with the single-precision literals 2./3. and 3./2. of the source. */
static void kernel_cubic(double r, const double dxv[2], double hs, double pi,
                         double *w, double dw[2])
{
    double q = r / hs;
    double factor = 15.0 / (7.0 * pi * hs * hs);
    if (q >= 0.0 && q <= 1.0) {
        *w = factor * ((double)(2.0f / 3.0f) - q * q + (q * (q * q)) / 2.0);
        for (int k = 0; k < 2; ++k)
            dw[k] = factor * (-2.0 + (double)(3.0f / 2.0f) * q) / (hs * hs) * dxv[k];
    } else if (q > 1.0 && q <= 2.0) {
        double t = 2.0 - q;
        *w = factor * 1.0 / 6.0 * (t * (t * t));
        for (int k = 0; k < 2; ++k)
            dw[k] = -(factor * 1.0 / 6.0 * 3.0 * (t * t) / hs * (dxv[k] / r));
    } else {
        *w = 0.0;
        dw[0] = dw[1] = 0.0;
    }
}

/* ---------------------------------------------------------- grid_find_NEW */
/* code/2_SPH_main_2018.f90:1199-1435 */
static int reserve_pairs(struct sps_handle *h, int64_t n)
{
    if (n <= h->pcap) return 0;
    int64_t nc = h->pcap ? h->pcap : 4096;
    while (nc < n) nc *= 2;
    int32_t *a = realloc(h->pi, nc * sizeof *a), *b, *c;
    if (!a) return -1;
    h->pi = a;
    b = realloc(h->pj, nc * sizeof *b); if (!b) return -1; h->pj = b;
    c = realloc(h->pt, nc * sizeof *c); if (!c) return -1; h->pt = c;
    float *f1 = realloc(h->pw, nc * sizeof *f1); if (!f1) return -1; h->pw = f1;
    float *f2 = realloc(h->pgx, nc * sizeof *f2); if (!f2) return -1; h->pgx = f2;
    float *f3 = realloc(h->pgy, nc * sizeof *f3); if (!f3) return -1; h->pgy = f3;
    int64_t *o = realloc(h->order, nc * sizeof *o); if (!o) return -1; h->order = o;
    h->pcap = nc;
    return 0;
}

/* Task 3 of the search (2_...:1309-1401) for the cells of one row idy
 * (1-based): the half-stencil pairs of each particle of each cell of the
 * row, in the reference's order.  Counts them, and with `write` stores them
 * from index `base` on. */
static int64_t row_pairs(struct sps_handle *h, int idy, const int ndiv[2], const int *cnt, const int *start,
                         int write, int64_t base)
{
    int64_t k = base;
    for (int idx = 1; idx <= ndiv[0]; ++idx) {
        int idivt = ndiv[0] * (idy - 1) + idx;
        int npicsi = cnt[idivt - 1];
        for (int ii = 1; ii <= npicsi; ++ii) {
            int it = h->cell_list[start[idivt - 1] + ii - 1];
            int idx0 = idx - 1 > 1 ? idx - 1 : 1;
            int idx1 = idx + 1 < ndiv[0] ? idx + 1 : ndiv[0];
            int idy1 = idy + 1 < ndiv[1] ? idy + 1 : ndiv[1];
            for (int jdy = idy; jdy <= idy1; ++jdy)
                for (int jdx = idx0; jdx <= idx1; ++jdx) {
                    int jdivt = ndiv[0] * (jdy - 1) + jdx;
                    if (jdivt < idivt) continue;
                    int jj = (idivt == jdivt) ? ii + 1 : 1;
                    int npicsj = cnt[jdivt - 1];
                    for (int jc = jj; jc <= npicsj; ++jc) {
                        int jt = h->cell_list[start[jdivt - 1] + jc - 1];
                        double dxv[2];
                        dxv[0] = h->x[2 * it] - h->x[2 * jt];
                        double driac = dxv[0] * dxv[0];
                        dxv[1] = h->x[2 * it + 1] - h->x[2 * jt + 1];
                        driac = driac + dxv[1] * dxv[1];
                        double mhsml = (h->hsml[it] + h->hsml[jt]) / 2.0;
                        double r = sqrt(driac);
                        if (r < 2 * mhsml) {
                            if (write) {
                                double w, dw[2];
                                kernel_cubic(r, dxv, mhsml, h->p.pi, &w, dw);
                                h->pi[k] = it;
                                h->pj[k] = jt;
                                h->pt[k] = 0;               /* 2_...:1374 */
                                h->pw[k] = (float)w;        /* FP32 pair fields, 5_...:19-21 */
                                h->pgx[k] = (float)dw[0];
                                h->pgy[k] = (float)dw[1];
                            }
                            ++k;
                        }
                    }
                }
        }
    }
    return k - base;
}

static int grid_find(struct sps_handle *h, int itimestep)
{
    const int nt2 = h->nt2;
    double xmin[2], xmax[2], deltx[2];
    int ndiv[2];
    /* Task 1 (2_...:1245-1258) */
    for (int d = 0; d < 2; ++d) {
        xmin[d] = (double)1.e+10f;
        xmax[d] = -xmin[d];
        deltx[d] = 0.0;
    }
    for (int d = 0; d < 2; ++d) {
        for (int i = 0; i < nt2; ++i) {
            if (h->ood[i] == 1) continue;
            double v = h->x[2 * i + d];
            if (v < xmin[d]) xmin[d] = v;
            if (v > xmax[d]) xmax[d] = v;
            if (deltx[d] < h->hsml[i]) deltx[d] = h->hsml[i];
        }
        deltx[d] = deltx[d] * 2;
        double length = xmax[d] - xmin[d];
        ndiv[d] = (int)(length / deltx[d] + 1);
        double length_new = ndiv[d] * deltx[d];
        xmin[d] = xmin[d] - (length_new - length) / 2 - (double)0.001f * length;
        xmax[d] = xmax[d] + (length_new - length) / 2 + (double)0.001f * length;
    }
    size_t ndivt = (size_t)ndiv[0] * (size_t)ndiv[1];
    if (ndivt > ((size_t)1 << 31)) return fail(SPS_ECAPACITY, "cell grid too large (%zu cells)", ndivt);
    if (ndivt > h->cell_cap) {
        free(h->cell_info);
        h->cell_info = malloc(3 * ndivt * sizeof(int));
        if (!h->cell_info) return fail(SPS_ENOMEM, "cell grid allocation");
        h->cell_cap = ndivt;
    }
    h->ndivx = ndiv[0];
    h->ndivy = ndiv[1];
    int *cnt = h->cell_info, *start = h->cell_info + ndivt, *fill = h->cell_info + 2 * ndivt;
    memset(cnt, 0, ndivt * sizeof(int));
    /* Task 2 (2_...:1277-1305): counting sort, ascending index inside a cell */
    for (int i = 0; i < nt2; ++i) {
        h->which_cell[i] = 0;
        if (h->ood[i] == 1) continue;
        int idiv[2];
        for (int d = 0; d < 2; ++d) {
            idiv[d] = (int)((h->x[2 * i + d] - xmin[d]) / deltx[d] + 1);
            if (idiv[d] > ndiv[d]) idiv[d] = ndiv[d];
        }
        int c = ndiv[0] * (idiv[1] - 1) + idiv[0]; /* 1-based */
        h->which_cell[i] = c;
        cnt[c - 1] += 1;
    }
    int ipost = 0;
    for (size_t c = 0; c < ndivt; ++c) {
        start[c] = ipost;
        ipost += cnt[c];
        fill[c] = 0;
    }
    for (int i = 0; i < nt2; ++i) {
        if (h->ood[i] == 1) continue;
        int c = h->which_cell[i] - 1;
        h->cell_list[start[c] + fill[c]++] = i;
    }
    /* Task 3 (2_...:1309-1401): half-stencil enumeration, row by row of
     * cells; a row's pairs follow the rows before it, so the rows are counted
     * first and then written from their offsets, in any thread order */
    if (itimestep == 1) h->m_pairs = 0;   /* fresh list, 2_...:1290-1294 */
    const int nth = h->nthreads > 0 ? h->nthreads : 1;
    int64_t *row_off = malloc(((size_t)ndiv[1] + 1) * sizeof *row_off);
    if (!row_off) return fail(SPS_ENOMEM, "pair list allocation");
    row_off[0] = 0;
#pragma omp parallel for num_threads(nth) schedule(dynamic, 4)
    for (int idy = 1; idy <= ndiv[1]; ++idy) row_off[idy] = row_pairs(h, idy, ndiv, cnt, start, 0, 0);
    for (int idy = 1; idy <= ndiv[1]; ++idy) row_off[idy] += row_off[idy - 1];
    if (reserve_pairs(h, row_off[ndiv[1]])) {
        free(row_off);
        return fail(SPS_ENOMEM, "pair list allocation");
    }
#pragma omp parallel for num_threads(nth) schedule(dynamic, 4)
    for (int idy = 1; idy <= ndiv[1]; ++idy) row_pairs(h, idy, ndiv, cnt, start, 1, row_off[idy - 1]);
    h->n_pairs = row_off[ndiv[1]];
    free(row_off);
    /* traversal order of the reused linked list (2_...:1309-1319, 1361-1369,
     * 1380-1383, 1398-1401): with M the high-water mark before this search,
     * the list from `last` is [n, n-1, ..., M+1, 1, 2, ..., min(n, M)]. */
    int64_t n = h->n_pairs, M = h->m_pairs, k = 0;
    if (n > M) {
        for (int64_t d = n - 1; d >= M; --d) h->order[k++] = d;
        for (int64_t d = 0; d < M; ++d) h->order[k++] = d;
        h->m_pairs = n;
    } else {
#pragma omp parallel for num_threads(nth)
        for (int64_t d = 0; d < n; ++d) h->order[d] = d;
    }
    return 0;
}

/* ------------------------------------------------------------ Pint_Update */
/* code/3_SPH_material_2018.f90:1555-1615 */
static void pint_update(struct sps_handle *h)
{
#pragma omp parallel for num_threads(h->nthreads > 0 ? h->nthreads : 1)
    for (int64_t k = 0; k < h->n_pairs; ++k) {
        int i = h->pi[k], j = h->pj[k];
        int ii = abs(h->itype[i]), jj = abs(h->itype[j]);
        int isumm = h->itype[i] + h->itype[j];
        if ((ii == 2 && jj == 1) || (ii == 2 && jj == 25) || (ii == 1 && jj == 25)) {
            h->pi[k] = j;
            h->pj[k] = i;
            h->pgx[k] = -h->pgx[k];
            h->pgy[k] = -h->pgy[k];
        }
        if (isumm == 3) h->pt[k] = 1;
        else if (isumm == 2) h->pt[k] = 2;
        else if (isumm == 4) h->pt[k] = 3;
        else if (isumm == 27) h->pt[k] = 6;
        else if (isumm == 26) h->pt[k] = 9;
    }
}

/* ---------------------------------------------------------- adapt_stress2 */
/* code/3_SPH_material_2018.f90:2066-2140 for one stress, all DP */
static void adapt_one(double *s, double tanfi, double coh)
{
    double alpha2 = tanfi / (sqrt(9 + 12 * (tanfi * tanfi)));
    double kc = (3 * coh) / (sqrt(9 + 12 * (tanfi * tanfi)));
    {
        double smean = (s[0] + s[1] + s[3]) / 3.0;
        double dv[4] = {s[0] - smean, s[1] - smean, s[2], s[3] - smean};
        double varj2 = dv[2] * dv[2] + 0.5 * (dv[0] * dv[0] + dv[1] * dv[1] + dv[3] * dv[3]);
        double yield = -(alpha2 * 3 * smean) + kc;
        if (yield < 0) {
            s[0] = s[0] - smean + kc / (3 * alpha2);
            s[1] = s[1] - smean + kc / (3 * alpha2);
            s[3] = s[3] - smean + kc / (3 * alpha2);
            smean = (s[0] + s[1] + s[3]) / 3.0;
            dv[0] = s[0] - smean;
            dv[1] = s[1] - smean;
            dv[2] = s[2];
            dv[3] = s[3] - smean;
            varj2 = dv[2] * dv[2] + 0.5 * (dv[0] * dv[0] + dv[1] * dv[1] + dv[3] * dv[3]);
            yield = -(alpha2 * 3 * smean) + kc;
        }
        if (yield < sqrt(varj2)) {
            double rn = (-(3 * alpha2 * smean) + kc) / (sqrt(varj2));
            if (sqrt(varj2) <= (double)10e-06f) rn = 0;
            s[0] = rn * dv[0] + smean;
            s[1] = rn * dv[1] + smean;
            s[3] = rn * dv[3] + smean;
            s[2] = rn * dv[2];
        }
    }
}

/* Material constants of particle i: its row of the material table when one
 * is set, else props(1,:) and the elastic matrix of sps_params. */
static double mat_tanfi(const struct sps_handle *h, int i) { return h->mat ? h->mtab[10 * h->mat[i]] : h->p.props[12]; }
static double mat_coh(const struct sps_handle *h, int i) { return h->mat ? h->mtab[10 * h->mat[i] + 1] : h->p.props[13]; }
static double mat_young(const struct sps_handle *h, int i) { return h->mat ? h->mtab[10 * h->mat[i] + 2] : h->p.props[2]; }
static double mat_poiss(const struct sps_handle *h, int i) { return h->mat ? h->mtab[10 * h->mat[i] + 3] : h->p.props[3]; }
/* D11, D22, D12, D33, D41, D42 of particle i */
static void mat_D(const struct sps_handle *h, int i, double D[6])
{
    if (h->mat) {
        memcpy(D, h->mtab + 10 * h->mat[i] + 4, 6 * sizeof(double));
    } else {
        D[0] = h->p.D11; D[1] = h->p.D22; D[2] = h->p.D12; D[3] = h->p.D33; D[4] = h->p.D41; D[5] = h->p.D42;
    }
}

/* all of 1..ntotal */
static void adapt_stress2(struct sps_handle *h)
{
#pragma omp parallel for num_threads(h->nthreads > 0 ? h->nthreads : 1)
    for (int i = 0; i < h->nt; ++i) adapt_one(h->stress + 4 * i, mat_tanfi(h, i), mat_coh(h, i));
}

/* ---------------------------------------------------- stress_point_update */
/* code/2_SPH_main_2018.f90:403-482, SP_SPH branch */
/* one pair of stress_point_update's loop: the stress point's and the
 * node's sums, each where this thread owns the particle */
static inline void spu_pair(struct sps_handle *h, int64_t d, int lo, int hi, int *unused)
{
    (void)unused;
    if (h->pt[d] != 1) return;
    const int j = h->pi[d]; /* stress point */
    const int i = h->pj[d]; /* node */
    const double w = (double)h->pw[d];
    if (owns(j, lo, hi)) {
        const double h2 = (h->mass[i] / h->rho[i]) * w;
        for (int a = 0; a < 2; ++a) h->vel_tmp[2 * j + a] = h->vel_tmp[2 * j + a] + h->vel[2 * i + a] * h2;
        h->cspm_norm[j] = h->cspm_norm[j] + (w * h->mass[i]) / h->rho[i];
    }
    if (owns(i, lo, hi)) {
        const double h1 = (h->mass[j] / h->rho[j]) * w;
        for (int s = 0; s < 4; ++s)
            h->stress_tmp[4 * i + s] = h->stress_tmp[4 * i + s] + h->stress[4 * j + s] * h1;
        h->ivar_tmp[i] = h->ivar_tmp[i] + h->ivars[(size_t)h->niv * j] * h1;
        h->cspm_norm[i] = h->cspm_norm[i] + (w * h->mass[j]) / h->rho[j];
    }
}

static void stress_point_update(struct sps_handle *h)
{
    const int nn = h->nn, nt = h->nt;
    memset(h->stress_tmp, 0, 4 * (size_t)nt * sizeof(double));
    memset(h->vel_tmp, 0, 2 * (size_t)nt * sizeof(double));
    memset(h->ivar_tmp, 0, (size_t)nt * sizeof(double));
    memset(h->cspm_norm, 0, (size_t)nt * sizeof(double));
    PAIR_LOOP(h, spu_pair, int, 0);
#pragma omp parallel for num_threads(h->nthreads > 0 ? h->nthreads : 1)
    for (int i = nn; i < nt; ++i) {
        if (h->cspm_norm[i] == 0) continue;
        h->vel[2 * i] = h->vel_tmp[2 * i] / h->cspm_norm[i];
        h->vel[2 * i + 1] = h->vel_tmp[2 * i + 1] / h->cspm_norm[i];
    }
#pragma omp parallel for num_threads(h->nthreads > 0 ? h->nthreads : 1)
    for (int i = 0; i < nn; ++i) {
        if (h->cspm_norm[i] != 0) {
            for (int s = 0; s < 4; ++s) h->stress[4 * i + s] = h->stress_tmp[4 * i + s] / h->cspm_norm[i];
            h->ivars[(size_t)h->niv * i] = h->ivar_tmp[i] / h->cspm_norm[i];
        } else {
            h->vel[2 * i] = 0;
            h->vel[2 * i + 1] = 0;
        }
    }
}

/* -------------------------------------------------------- get_derivatives */
/* code/2_SPH_main_2018.f90:487-648.  fi1 is vel, fi2 is stress (the actual
 * arguments at 2_...:725), so the type-6 copy into stress(:,dummy) is seen
 * through fi2 as in the reference. */
#define G1(a, b, p) h->grad1[4 * (p) + 2 * (b) + (a)] /* grad1_tmp(a,b,p) */
#define G2(s, b, p) h->grad2[6 * (p) + 3 * (b) + (s)] /* grad2_tmp(s,b,p) */
#define AE(c, p) h->ae[5 * (p) + (c)]
/* one pair of get_derivatives' loop (2_...:507-590), each end's sums where
 * this thread owns it.  The type-6 pair sets the dummy's stress to the
 * node's and then reads it: the node's sum reads the node's own, the same
 * value; the type-9 pair's mirrored dummy velocity is the pair's own. */
static inline void deriv_pair(struct sps_handle *h, int64_t d, int lo, int hi, int *unused)
{
    (void)unused;
    const double *x = h->x, *m = h->mass, *rho = h->rho;
    double *vel = h->vel, *st = h->stress;
    const int j = h->pi[d], i = h->pj[d];
    const double gx = (double)h->pgx[d], gy = (double)h->pgy[d];
    double h1 = gx * m[i] / rho[i], h2 = 0.0;
    const int t = h->pt[d];
    if (t == 1) {
        if (owns(j, lo, hi)) {
            for (int a = 0; a < 2; ++a) {
                G1(a, 0, j) = G1(a, 0, j) + (vel[2 * i + a] - vel[2 * j + a]) * h1;
                h2 = gy * m[i] / rho[i];
                G1(a, 1, j) = G1(a, 1, j) + (vel[2 * i + a] - vel[2 * j + a]) * h2;
            }
        }
        if (owns(i, lo, hi))
            for (int s = 0; s < 3; ++s) {
                double g1 = gx * (st[4 * i + s] / (rho[i] * rho[i]) + st[4 * j + s] / (rho[j] * rho[j]));
                G2(s, 0, i) = G2(s, 0, i) - m[j] * g1;
                double g2 = gy * (st[4 * i + s] / (rho[i] * rho[i]) + st[4 * j + s] / (rho[j] * rho[j]));
                G2(s, 1, i) = G2(s, 1, i) - m[j] * g2;
            }
        if (h->p.cspm) {
            h2 = gy * m[i] / rho[i];
            if (owns(j, lo, hi)) {
                AE(0, j) = AE(0, j) + (x[2 * i] - x[2 * j]) * h1;
                AE(1, j) = AE(1, j) + (x[2 * i + 1] - x[2 * j + 1]) * h1;
                AE(2, j) = AE(2, j) + (x[2 * i] - x[2 * j]) * h2;
                AE(3, j) = AE(3, j) + (x[2 * i + 1] - x[2 * j + 1]) * h2;
            }
            if (owns(i, lo, hi)) {
                double h1b = -(gx * m[j] / rho[j]);
                double h2b = -(gy * m[j] / rho[j]);
                AE(0, i) = AE(0, i) + (x[2 * j] - x[2 * i]) * h1b;
                AE(1, i) = AE(1, i) + (x[2 * j + 1] - x[2 * i + 1]) * h1b;
                AE(2, i) = AE(2, i) + (x[2 * j] - x[2 * i]) * h2b;
                AE(3, i) = AE(3, i) + (x[2 * j + 1] - x[2 * i + 1]) * h2b;
            }
        }
    } else if (t == 9) { /* i: stress point, j: dummy */
        double beta_max = 1.5, vel_wall = 0.0, da, db;
        double wall = (double)h->wallpos[j];
        if (h->horiz[j] == 1.0f) {
            da = fabs(x[2 * i + 1] - wall);
            db = fabs(x[2 * j + 1] - wall);
        } else {
            da = fabs(x[2 * i] - wall);
            db = fabs(x[2 * j] - wall);
        }
        double bt = 1 + (db / da);
        double beta = beta_max < bt ? beta_max : bt; /* min(beta_max, 1+db/da) */
        double dv[2];
        for (int a = 0; a < 2; ++a) dv[a] = vel[2 * i + a] * (1 - beta) + beta * vel_wall;
        if (owns(j, lo, hi))
            for (int a = 0; a < 2; ++a) h->dummy_vel[2 * j + a] = dv[a];
        if (owns(i, lo, hi)) {
            h1 = gx * m[j] / rho[j];
            for (int a = 0; a < 2; ++a) {
                G1(a, 0, i) = G1(a, 0, i) + (vel[2 * i + a] - dv[a]) * h1;
                h2 = gy * m[j] / rho[j];
                G1(a, 1, i) = G1(a, 1, i) + (vel[2 * i + a] - dv[a]) * h2;
            }
        }
    } else if (t == 6) { /* i: node, j: dummy */
        if (owns(j, lo, hi))
            for (int s = 0; s < 4; ++s) st[4 * j + s] = st[4 * i + s];
        if (owns(i, lo, hi))
            for (int s = 0; s < 3; ++s) {
                const double sj = st[4 * i + s];   /* the dummy's stress, just set to the node's */
                double a1 = gx * (st[4 * i + s] / (rho[i] * rho[i]) + sj / (rho[j] * rho[j]));
                G2(s, 0, i) = G2(s, 0, i) - m[j] * a1;
                double a2 = gy * (st[4 * i + s] / (rho[i] * rho[i]) + sj / (rho[j] * rho[j]));
                G2(s, 1, i) = G2(s, 1, i) - m[j] * a2;
            }
    }
}

static void get_derivatives(struct sps_handle *h)
{
    const int nn = h->nn, nt = h->nt;
    if (h->ifsigman == 1 && h->nd > 0) {
        /* bc_int (2_...:506, 579): 1 for a node with a node-dummy pair in the
         * list; apply_stress_free reads it at the BC calls that follow, so the
         * first BC calls of a step see the previous step's list.  The outside
         * approach sets it again from the same list after the position update
         * (shift_stress_points), its only other reader. */
        memset(h->bc_int, 0, (size_t)nn * sizeof(int32_t));
        for (int64_t k = 0; k < h->n_pairs; ++k)
            if (h->pt[k] == 6) h->bc_int[h->pj[k]] = 1; /* pair_j: the node */
    }
    memset(h->grad1, 0, 4 * (size_t)nt * sizeof(double));
    memset(h->grad2, 0, 6 * (size_t)nt * sizeof(double));
    memset(h->divf1, 0, 4 * (size_t)nt * sizeof(double));
    memset(h->divf2, 0, 2 * (size_t)nt * sizeof(double));
    memset(h->ae, 0, 5 * (size_t)nt * sizeof(double));
    memset(h->dummy_vel, 0, 2 * (size_t)h->nt2 * sizeof(double));
    PAIR_LOOP(h, deriv_pair, int, 0);
    if (h->p.cspm) { /* 2_...:594-604 */
        double tol = h->p.cspm_det_tol;
#pragma omp parallel for num_threads(h->nthreads > 0 ? h->nthreads : 1)
        for (int i = 0; i < nt; ++i) {
            AE(4, i) = AE(0, i) * AE(3, i) - AE(1, i) * AE(2, i);
            if (fabs(AE(4, i)) < tol) {
                AE(4, i) = 1; AE(0, i) = 1; AE(1, i) = 0; AE(2, i) = 0; AE(3, i) = 1;
            } else {
                AE(4, i) = 1 / AE(4, i);
            }
        }
        /* 2_...:616-627: the second statement of each pair reads the already
         * corrected first column, as the array assignments in sequence do */
        for (int a = 0; a < 2; ++a) {
            for (int p = nn; p < nt; ++p) G1(a, 0, p) = AE(4, p) * (AE(0, p) * G1(a, 0, p) + AE(1, p) * G1(a, 1, p));
            for (int p = nn; p < nt; ++p) G1(a, 1, p) = AE(4, p) * (AE(2, p) * G1(a, 0, p) + AE(3, p) * G1(a, 1, p));
            for (int p = 0; p < nn; ++p) G2(a, 0, p) = AE(4, p) * (AE(0, p) * G2(a, 0, p) + AE(1, p) * G2(a, 1, p));
            for (int p = 0; p < nn; ++p) G2(a, 1, p) = AE(4, p) * (AE(2, p) * G2(a, 0, p) + AE(3, p) * G2(a, 1, p));
        }
    }
#pragma omp parallel for num_threads(h->nthreads > 0 ? h->nthreads : 1)
    for (int p = nn; p < nt; ++p) { /* 2_...:633-636 */
        double D[6]; /* D11 D22 D12 D33 D41 D42 */
        mat_D(h, p, D);
        h->divf1[4 * p + 0] = -(D[0] * G1(0, 0, p) + D[2] * G1(1, 1, p));
        h->divf1[4 * p + 1] = -(D[2] * G1(0, 0, p) + D[1] * G1(1, 1, p));
        h->divf1[4 * p + 2] = -(D[3] * G1(1, 0, p) + D[3] * G1(0, 1, p));
        h->divf1[4 * p + 3] = -(D[4] * G1(0, 0, p) + D[5] * G1(1, 1, p));
    }
    for (int p = 0; p < nn; ++p) { /* 2_...:641-642 */
        h->divf2[2 * p + 0] = -(G2(0, 0, p) + G2(2, 1, p));
        h->divf2[2 * p + 1] = -(G2(2, 0, p) + G2(1, 1, p));
    }
    memcpy(h->grad_u, h->grad1, 4 * (size_t)nt * sizeof(double)); /* 2_...:644 */
}
#undef G1
#undef G2
#undef AE

/* --------------------------------------------- plastic_terms / drucker_prager */
/* code/3_SPH_material_2018.f90:1867-1933 -> 1937-2062 -> 2557-2635.
 * Returns the source Gs(4) = -G and der_intvars(1). */
/* ------------------------------------------------- Perzyna viscoplasticity */
/* plastic_terms' ncrit <= 4 branch (3_...:1912-1923 of the strain-
 * localisation fork, whose sources the reference builds from): Get_Vivel
 * (2168-2190) -> invar09 (2196-2336): invariants, yield measure and the
 * current yield size fdatm0 + H evpstn limited below by 0.1 fdatm0;
 * yieldf09 (2340-2472): the flow vector a = c1 a1 + c2 a2 + c3 a3;
 * flowvp09 (2476-2525): vivel = gamma <F/fdatm>^delta a (or gamma (exp(delta
 * F/fdatm) - 1) a when n != 1) past the 0.01 threshold.  sqrt(3.0) and the
 * literals 0.017453292, 57.29577951308, 0.1, 0.01 are default reals, as the
 * reference writes them.  Writes vivel (zero below yield). */
static void perzyna_vivel(const struct sps_handle *h, const double s[4], double evpstn, double vivel[4])
{
    const double *pr = h->p.props;
    const int ncrit = (int)pr[1];
    const double root3 = (double)sqrtf(3.0f);
    const double smean = (s[0] + s[1] + s[3]) / 3.0;
    const double dv[4] = {s[0] - smean, s[1] - smean, s[2], s[3] - smean};
    const double varj2 = dv[2] * dv[2] + 0.5 * (dv[0] * dv[0] + dv[1] * dv[1] + dv[3] * dv[3]);
    const double varj3 = dv[3] * (dv[3] * dv[3] - varj2);
    const double steff = sqrt(varj2);
    double sint3;
    if (steff != 0.0) {
        sint3 = -(3.0 * root3 * varj3 / (2.0 * varj2 * steff));
        if (sint3 > 1.0) sint3 = 1.0;
    } else {
        sint3 = 0.0;
    }
    if (sint3 < -1.0) sint3 = -1.0;
    if (sint3 > 1.0) sint3 = 1.0;
    const double theta = asin(sint3) / 3.0;
    const double snphi = sin(pr[8] * (double)0.017453292f);
    double yield = 0.0;
    if (ncrit == 1) yield = 2.0 * cos(theta) * steff;
    else if (ncrit == 2) yield = root3 * steff;
    else if (ncrit == 3) yield = smean * snphi + steff * (cos(theta) - sin(theta) * snphi / root3);
    else if (ncrit == 4) yield = 6.0 * smean * snphi / (root3 * (3.0 - snphi)) + steff;
    const double fdatm0 = pr[6], hards = pr[7];
    double fdatm = fdatm0 + hards * evpstn;
    double fact = fabs(fdatm) / fabs(fdatm0);
    if (fact < (double)0.1f) fact = (double)0.1f;
    fdatm = fdatm0 * fact;
    vivel[0] = vivel[1] = vivel[2] = vivel[3] = 0.0;
    if (!(yield > fdatm)) return;
    /* yieldf09 */
    const double a1[4] = {1.0, 1.0, 0.0, 1.0};
    double a2[4] = {0.0, 0.0, 0.0, 0.0}, a3[4];
    if (steff > 0) {
        for (int i = 0; i < 4; ++i) a2[i] = dv[i] / (2.0 * steff);
        a2[2] = dv[2] / steff;
    }
    a3[0] = dv[1] * dv[3] + varj2 / 3.0;
    a3[1] = dv[0] * dv[3] + varj2 / 3.0;
    a3[2] = -(2.0 * dv[2] * dv[3]);
    a3[3] = dv[0] * dv[1] - dv[2] * dv[2] + varj2 / 3.0;
    const double tanth = tan(theta), tant3 = tan(3.0 * theta), sinth = sin(theta), costh = cos(theta);
    const double cost3 = cos(3.0 * theta);
    double c1 = 0.0, c2 = 0.0, c3 = 0.0;
    if (ncrit == 1) {
        const double abthe = fabs(theta * (double)57.29577951308f);
        if (abthe >= 29.0) {
            c2 = root3;
            c3 = 0.0;
        } else {
            c2 = 2.0 * (costh + sinth * tant3);
            c3 = root3 * sinth / (varj2 * cost3);
        }
    } else if (ncrit == 2) {
        c2 = root3;
    } else if (ncrit == 3) {
        c1 = sin(pr[8] * (double)0.017453292f) / 3.0;
        const double abthe = fabs(theta * (double)57.29577951308f);
        if (abthe >= 29.0) {
            const double plumi = theta > 0.0 ? -1.0 : 1.0;
            c2 = 0.5 * (root3 + plumi * c1 * root3);
        } else {
            c2 = costh * ((1.0 + tanth * tant3) + c1 * (tant3 - tanth) * root3);
            c3 = (root3 * sinth + 3.0 * c1 * costh) / (2.0 * varj2 * cost3);
        }
    } else if (ncrit == 4) {
        c1 = 2.0 * snphi / (root3 * (3.0 - snphi));
        c2 = 1.0;
    }
    double av[4];
    for (int i = 0; i < 4; ++i) av[i] = c1 * a1[i] + c2 * a2[i] + c3 * a3[i];
    /* flowvp09 */
    const double gamma = pr[9], delta = pr[10], nflow = pr[11];
    const double fnorm = (yield - fdatm) / fdatm;
    if (fnorm >= (double)0.01f) {
        const double cmult = nflow != 1 ? gamma * (exp(delta * fnorm) - 1.0) : gamma * pow(fnorm, delta);
        for (int i = 0; i < 4; ++i) vivel[i] = cmult * av[i];
    }
}

static void plastic_terms(struct sps_handle *h, int ie, double time_sph, double Gs[4], double *der1)
{
    const double *props = h->p.props;
    int ntype_eco = (int)props[0];
    int ncrit = (int)props[1];
    double G[4] = {0, 0, 0, 0}, vivel[4] = {0, 0, 0, 0};
    Gs[0] = Gs[1] = Gs[2] = Gs[3] = 0.0;
    if (ntype_eco > 1) {
        if (ncrit == 12) {
            const double *st = h->stress + 4 * ie;
            double stress2[4] = {st[0], st[1], st[2], st[3]};
            double f0 = h->fdr[ie];
            float tanfi = (float)mat_tanfi(h, ie), coh = (float)mat_coh(h, ie);
            float young = (float)mat_young(h, ie), poiss = (float)mat_poiss(h, ie);
            float smean = (float)((stress2[0] + stress2[1] + stress2[3]) / 3.0);
            float dv[4];
            dv[0] = (float)(stress2[0] - (double)smean);
            dv[1] = (float)(stress2[1] - (double)smean);
            dv[2] = (float)stress2[2];
            dv[3] = (float)(stress2[3] - (double)smean);
            float varj2 = dv[2] * dv[2] + 0.5f * (dv[0] * dv[0] + dv[1] * dv[1] + dv[3] * dv[3]);
            float vari1 = 3 * smean;
            const double *gu = h->grad_u + 4 * ie; /* [g11, g21, g12, g22] */
            float eps11 = (float)gu[0];
            float eps12 = (float)(0.5 * (gu[2] + gu[1]));
            float eps22 = (float)gu[3];
            float emean = eps11 + eps22;
            float alpha2 = tanfi / (sqrtf(9 + 12 * (tanfi * tanfi)));
            float kc = (3 * coh) / (sqrtf(9 + 12 * (tanfi * tanfi)));
            float yield = -(alpha2 * vari1) + kc;
            h->fdr[ie] = (double)(sqrtf(varj2) - yield);
            float f1 = (float)h->fdr[ie];
            double df = (double)f1 - f0;
            float G_mod = young / (2.0f * (1.0f + poiss));
            float K_mod = young / (3.0f * (1.0f - 2.0f * poiss));
            float s_eps = dv[0] * eps11 + 2 * dv[2] * eps12 + dv[1] * eps22;
            if (time_sph > 0 && f1 >= 0 && df >= 0 && sqrtf(varj2) >= 10e-06f) {
                float G1v[4];
                for (int i = 0; i < 4; ++i) G1v[i] = (G_mod / sqrtf(varj2)) * dv[i];
                float lambda_1 = 3 * alpha2 * K_mod * emean;
                float lambda_2 = (G_mod / sqrtf(varj2)) * s_eps;
                float lambda_3 = G_mod;
                float G2 = (lambda_1 + lambda_2) / lambda_3;
                for (int i = 0; i < 4; ++i) G[i] = (double)(G1v[i] * G2);
                vivel[0] = (double)((1.0f / 6.0f) * (1.0f / sqrtf(varj2)) * (2 * dv[0] - dv[1] - dv[3]));
                vivel[1] = (double)((1.0f / 6.0f) * (1.0f / sqrtf(varj2)) * (2 * dv[1] - dv[0] - dv[3]));
                vivel[2] = (double)((1.0f / sqrtf(varj2)) * dv[2]);
                vivel[3] = (double)((1.0f / 6.0f) * (1 / sqrtf(varj2)) * (2 * dv[3] - dv[0] - dv[1]));
                for (int i = 0; i < 4; ++i) vivel[i] = vivel[i] * (double)G2;
            }
            for (int i = 0; i < 4; ++i) Gs[i] = -G[i];
        } else if (ncrit >= 1 && ncrit <= 4) {
            /* Gs = -Dmatx vivel, Get_Dmatx (3_...:2530-2561) */
            perzyna_vivel(h, h->stress + 4 * ie, h->ivars[(size_t)h->niv * ie], vivel);
            const double young = props[2], poiss = props[3];
            const double c = young * (1.0 - poiss) / ((1.0 + poiss) * (1.0 - 2.0 * poiss));
            const double c12 = c * poiss / (1.0 - poiss), c33 = (1.0 - 2.0 * poiss) * c / (2.0 * (1.0 - poiss));
            const double D[4][4] = {{c, c12, 0.0, c12}, {c12, c, 0.0, c12}, {0.0, 0.0, c33, 0.0}, {c12, c12, 0.0, c}};
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) Gs[i] = Gs[i] - D[i][j] * vivel[j];
        }
        /* Get_derivative_intvars (3_...:2614-2617) */
        *der1 = sqrt((2.0 * (vivel[0] * vivel[0] + vivel[1] * vivel[1] + vivel[3] * vivel[3]) +
                      vivel[2] * vivel[2]) / 3.0);
    } else {
        *der1 = 0.0;
    }
}

/* ---------------------------------------------------------- gravity_force */
/* code/3_SPH_material_2018.f90:2640-2702: the time factor at step time */
static double gravity_factor(const sps_params *P, double time_sph)
{
    double factg;
    if (P->ic_grav == 0) return 0.0;
    double tt0 = -1000.0, tt1 = 0.0;
    int ip, found = 0;
    for (ip = 0; ip < P->tcurve_npts - 1; ++ip) {
        tt0 = P->tcurve_t[ip];
        tt1 = P->tcurve_t[ip + 1];
        if (time_sph >= tt0 && time_sph <= tt1) { found = 1; break; }
        tt0 = -1000.0;
    }
    if (tt0 >= 0.0 && found) {
        double xi = (time_sph - tt0) / (tt1 - tt0);
        factg = (1.0 - xi) * (double)P->tcurve_f[ip] + xi * (double)P->tcurve_f[ip + 1];
    } else {
        factg = 0.0;
    }
    return factg * P->ft_grav;
}

static void gravity_force(struct sps_handle *h, double factg)
{
    OMP_FOR
    for (int i = 0; i < h->nn; ++i)
        for (int a = 0; a < 2; ++a) {
            double g = h->p.ic_grav ? factg * h->p.cgrav[a] : 0.0;
            h->g_grav[2 * i + a] = g - h->p.damping * h->vel[2 * i + a];
        }
}

/* -------------------------------------------------------- boundary_forces */
/* code/2_SPH_main_2018.f90:1117-1161 (test = 2): every node against every
 * inner-layer dummy, FP32 terms summed into DP.  A term is nonzero only when
 * 0 < r2 < 1.5*d0 in single precision; dummies are binned on a grid whose
 * cell exceeds that reach, so each node gathers its candidates from 3x3
 * cells and adds them in ascending dummy index like the reference j loop. */
static float bf_d0(const struct sps_handle *h) { return (float)(h->p.dx / 2.); }

static int bw_setup(struct sps_handle *h)
{
    free(h->bw_start); free(h->bw_list);
    h->bw_start = NULL; h->bw_list = NULL;
    int nd2 = h->nd2;
    if (nd2 == 0) return 0;
    double reach = 2.0 * (double)(1.5f * bf_d0(h)) + 1e-12;
    double x0 = 1e300, y0 = 1e300, x1 = -1e300, y1 = -1e300;
    for (int k = 0; k < nd2; ++k) {
        int j = h->nt + k;
        double a = h->x[2 * j], b = h->x[2 * j + 1];
        if (a < x0) x0 = a;
        if (a > x1) x1 = a;
        if (b < y0) y0 = b;
        if (b > y1) y1 = b;
    }
    h->bw_cell = reach;
    h->bw_x0 = x0 - reach;
    h->bw_y0 = y0 - reach;
    h->bw_nx = (int)((x1 - h->bw_x0) / reach) + 3;
    h->bw_ny = (int)((y1 - h->bw_y0) / reach) + 3;
    if ((double)h->bw_nx * h->bw_ny > 5e7) return -1;
    size_t nc = (size_t)h->bw_nx * h->bw_ny;
    h->bw_start = calloc(nc + 1, sizeof(int));
    h->bw_list = calloc((size_t)nd2, sizeof(int));
    if (!h->bw_start || !h->bw_list) return -1;
    for (int k = 0; k < nd2; ++k) {
        int j = h->nt + k;
        int cx = (int)((h->x[2 * j] - h->bw_x0) / reach), cy = (int)((h->x[2 * j + 1] - h->bw_y0) / reach);
        h->bw_start[(size_t)cy * h->bw_nx + cx + 1]++;
    }
    for (size_t c = 0; c < nc; ++c) h->bw_start[c + 1] += h->bw_start[c];
    int *fill = calloc(nc, sizeof(int));
    if (!fill) return -1;
    for (int k = 0; k < nd2; ++k) { /* ascending index inside a cell */
        int j = h->nt + k;
        int cx = (int)((h->x[2 * j] - h->bw_x0) / reach), cy = (int)((h->x[2 * j + 1] - h->bw_y0) / reach);
        size_t c = (size_t)cy * h->bw_nx + cx;
        h->bw_list[h->bw_start[c] + fill[c]++] = j;
    }
    free(fill);
    return 0;
}

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

static float bf_term_scale(void)
{
    float c = 20;
    return 0.01f * (c * c);
}

static void boundary_forces(struct sps_handle *h)
{
    memset(h->f_bound, 0, 2 * (size_t)h->nn * sizeof(double));
    if (h->nd2 == 0) return;
    const float d0 = bf_d0(h);
    const float scale = bf_term_scale();
    const double reach = h->bw_cell;
#pragma omp parallel for num_threads(h->nthreads > 0 ? h->nthreads : 1) schedule(dynamic, 256)
    for (int i = 0; i < h->nn; ++i) {
        int cand[4096];
        double xi = h->x[2 * i], yi = h->x[2 * i + 1];
        int cx = (int)floor((xi - h->bw_x0) / reach), cy = (int)floor((yi - h->bw_y0) / reach);
        int nc = 0;
        for (int by = cy - 1; by <= cy + 1; ++by) {
            if (by < 0 || by >= h->bw_ny) continue;
            for (int bx = cx - 1; bx <= cx + 1; ++bx) {
                if (bx < 0 || bx >= h->bw_nx) continue;
                size_t c = (size_t)by * h->bw_nx + bx;
                for (int q = h->bw_start[c]; q < h->bw_start[c + 1] && nc < 4096; ++q) cand[nc++] = h->bw_list[q];
            }
        }
        qsort(cand, (size_t)nc, sizeof(int), cmp_int);
        for (int q = 0; q < nc; ++q) {
            int j = cand[q];
            float r[2];
            r[0] = (float)(h->x[2 * i] - h->x[2 * j]);
            r[1] = (float)(h->x[2 * i + 1] - h->x[2 * j + 1]);
            float r2 = sqrtf(r[0] * r[0] + r[1] * r[1]);
            float f2;
            float hh = (float)(0.5 * (h->hsml[i] + h->hsml[j]));
            float r_bound = r2 / (0.75f * hh);
            if (r2 > 0 && r2 < 1.5f * d0) f2 = 1 - (r2 / (1.5f * d0));
            else f2 = 0.0f;
            float f;
            if (0 < r_bound && r_bound <= 2.0f / 3.0f) f = 2.0f / 3.0f;
            else if (2.0f / 3.0f < r_bound && r_bound <= 1) f = 2 * r_bound - 1.5f * (r_bound * r_bound);
            else if (1 < r_bound && r_bound < 2) f = 0.5f * ((2 - r_bound) * (2 - r_bound));
            else f = 0.0f;
            for (int a = 0; a < 2; ++a)
                h->f_bound[2 * i + a] = h->f_bound[2 * i + a] + (double)(scale * f2 * f * (r[a] / (r2 * r2)));
        }
    }
}

/* --------------------------------------------------- artificial_viscosity */
/* code/2_SPH_main_2018.f90:826-904, FP32 locals and accumulators */
static inline void av_pair(struct sps_handle *h, int64_t d, int lo, int hi, float *visc)
{
    if (h->pt[d] != 3) return;
    const double *x = h->x, *v = h->vel, *m = h->mass;
    const float c = 600;
    const int i = h->pi[d], j = h->pj[d];
    if (!owns(i, lo, hi) && !owns(j, lo, hi)) return;
    float xij = (float)(x[2 * i] - x[2 * j]);
    float yij = (float)(x[2 * i + 1] - x[2 * j + 1]);
    float hh = (float)(0.5 * (h->hsml[i] + h->hsml[j]));
    float rho2 = (float)(0.5 * (h->rho[i] + h->rho[j]));
    float div_u = (float)((double)xij * (v[2 * i] - v[2 * j]));
    div_u = (float)((double)div_u + (double)yij * (v[2 * i + 1] - v[2 * j + 1]));
    float sr = sqrtf(xij * xij + yij * yij);
    float theta = (hh * div_u) / (sr * sr + 0.01f * (hh * hh));
    if (div_u < 0) {
        *visc = (float)((-(h->p.alpha * (double)c * (double)theta) + h->p.beta * (double)(theta * theta)) /
                        (double)rho2);
    } else if (div_u >= 0) {
        *visc = 0;
    }
    float gx = h->pgx[d], gy = h->pgy[d];
    if (owns(i, lo, hi)) {
        h->av_tmp[2 * i] = (float)((double)h->av_tmp[2 * i] + (double)(*visc * gx) * m[j]);
        h->av_tmp[2 * i + 1] = (float)((double)h->av_tmp[2 * i + 1] + (double)(*visc * gy) * m[j]);
    }
    if (owns(j, lo, hi)) {
        h->av_tmp[2 * j] = (float)((double)h->av_tmp[2 * j] - (double)(*visc * gx) * m[i]);
        h->av_tmp[2 * j + 1] = (float)((double)h->av_tmp[2 * j + 1] - (double)(*visc * gy) * m[i]);
    }
}

static void artificial_viscosity(struct sps_handle *h)
{
    const int nn = h->nn;
    memset(h->av_tmp, 0, 2 * (size_t)nn * sizeof(float));
    /* visc carries over from the last pair with a finite divergence where a
     * NaN divergence leaves both branches untaken (per thread) */
    PAIR_LOOP(h, av_pair, float, 0.0f);
    for (int i = 0; i < 2 * nn; ++i) h->art_visc[i] = -(double)h->av_tmp[i];
}

/* -------------------------------------------------------------------- RK4 */
/* code/2_SPH_main_2018.f90:653-802 */
static void normal_bcs(struct sps_handle *h, double time_sph);

static void rk4(struct sps_handle *h, double dt_sph, double time_sph)
{
    const int nn = h->nn, nt = h->nt;
    const double f1rk[4] = {0., 0.5, 0.5, 1.0}, f2rk[4] = {1., 2., 2., 1.0};
    const int ncrit = (int)h->p.props[1];
    memcpy(h->vel0, h->vel, 2 * (size_t)nt * sizeof(double));
    memcpy(h->stress0, h->stress, 4 * (size_t)nt * sizeof(double));
    memset(h->rk_stress, 0, 4 * (size_t)nt * sizeof(double));
    memset(h->rk_vel, 0, 2 * (size_t)nt * sizeof(double));
    memset(h->rk_dev, 0, (size_t)nt * sizeof(double));
    memset(h->rhs1, 0, 4 * (size_t)nt * sizeof(double));
    memset(h->rhs2, 0, 2 * (size_t)nt * sizeof(double));
    memset(h->source_stress, 0, 4 * (size_t)nt * sizeof(double));
    memset(h->spin, 0, 4 * (size_t)nt * sizeof(double));
    memset(h->art_visc, 0, 2 * (size_t)nn * sizeof(double));
    memset(h->vel, 0, 2 * (size_t)h->nt2 * sizeof(double));
    memset(h->stress, 0, 4 * (size_t)h->nt2 * sizeof(double));
    const double factg = gravity_factor(&h->p, time_sph);
    const int do_bforce = h->p.inside_approach && (h->p.bforce_gate ? 1 : h->p.dummy_nodes);
    for (int s = 0; s < 4; ++s) {
        const double c1 = f1rk[s] * (dt_sph);
        OMP_FOR
        for (int i = 0; i < nn; ++i)
            for (int a = 0; a < 2; ++a)
                h->vel[2 * i + a] = h->vel0[2 * i + a] + c1 * h->rhs2[2 * i + a];
        OMP_FOR
        for (int i = nn; i < nt; ++i)
            for (int q = 0; q < 4; ++q)
                h->stress[4 * i + q] = h->stress0[4 * i + q] + c1 * h->rhs1[4 * i + q];
        if (ncrit == 12) adapt_stress2(h);
        normal_bcs(h, time_sph);
        stress_point_update(h);
        if (ncrit == 12) adapt_stress2(h);
        normal_bcs(h, time_sph);
        get_derivatives(h);
        OMP_FOR
        for (int j = nn; j < nt; ++j) {
            double Gs[4], der1;
            plastic_terms(h, j, time_sph, Gs, &der1);
            for (int q = 0; q < 4; ++q) h->source_stress[4 * j + q] = Gs[q];
            h->rk_dev[j] = h->rk_dev[j] + der1 * f2rk[s];
        }
        gravity_force(h, factg);
        if (do_bforce) boundary_forces(h);
        if (h->p.alpha > 0 || h->p.beta > 0) artificial_viscosity(h);
        if (h->p.update_x) { /* 2_...:751-757 with get_spin_rate_tensor 2_...:1021-1034 */
            OMP_FOR
            for (int i = nn; i < nt; ++i) {
                const double *gu = h->grad_u + 4 * i;
                double om1 = 0.5 * (gu[2] - gu[1]);
                double om2 = -(0.5 * (gu[2] - gu[1]));
                const double *st = h->stress + 4 * i;
                h->spin[4 * i + 0] = 2 * om1 * st[2];
                h->spin[4 * i + 1] = 2 * om2 * st[2];
                h->spin[4 * i + 2] = om2 * st[0] + om1 * st[1];
            }
        }
        OMP_FOR
        for (int i = nn; i < nt; ++i)
            for (int q = 0; q < 4; ++q)
                h->rhs1[4 * i + q] = -h->divf1[4 * i + q] + h->spin[4 * i + q] + h->source_stress[4 * i + q];
        OMP_FOR
        for (int i = 0; i < nn; ++i)
            for (int a = 0; a < 2; ++a)
                h->rhs2[2 * i + a] = -h->divf2[2 * i + a] + h->g_grav[2 * i + a] + h->art_visc[2 * i + a] +
                                     h->f_bound[2 * i + a] + 0.0;
        OMP_FOR
        for (int i = 0; i < nn; ++i)
            for (int a = 0; a < 2; ++a)
                h->rk_vel[2 * i + a] = h->rk_vel[2 * i + a] + f2rk[s] * h->rhs2[2 * i + a];
        OMP_FOR
        for (int i = nn; i < nt; ++i)
            for (int q = 0; q < 4; ++q)
                h->rk_stress[4 * i + q] = h->rk_stress[4 * i + q] + f2rk[s] * h->rhs1[4 * i + q];
    }
    OMP_FOR
    for (int i = 0; i < nn; ++i)
        for (int a = 0; a < 2; ++a)
            h->vel[2 * i + a] = h->vel0[2 * i + a] + (dt_sph / 6) * h->rk_vel[2 * i + a];
    OMP_FOR
    for (int i = nn; i < nt; ++i)
        for (int q = 0; q < 4; ++q)
            h->stress[4 * i + q] = h->stress0[4 * i + q] + (dt_sph / 6) * h->rk_stress[4 * i + q];
    if (ncrit == 12 && h->log_cap > 0) {
        /* the plain method's tuples of the plastic trial states (training data) */
        for (int i = nn; i < nt; ++i) {
            double ret[4];
            const double *tr = h->stress + 4 * i;
            memcpy(ret, tr, sizeof ret);
            adapt_one(ret, mat_tanfi(h, i), mat_coh(h, i));   /* a copy: adapt_stress2 below does it */
            if (memcmp(tr, ret, sizeof ret) != 0 && h->log_n < h->log_cap) {
                double *o = h->plog + SPS_PLOG_ROW * h->log_n++;
                memcpy(o, h->stress0 + 4 * i, 4 * sizeof(double));
                memcpy(o + 4, tr, 4 * sizeof(double));
                memcpy(o + 8, ret, sizeof ret);
                o[12] = mat_tanfi(h, i);
                o[13] = mat_coh(h, i);
            }
        }
    }
    if (ncrit == 12) adapt_stress2(h);
    normal_bcs(h, time_sph);
    OMP_FOR
    for (int i = 0; i < nt; ++i) h->ddev[i] = h->rk_dev[i] / 6;
}

/* --------------------------------------------------------------- leapfrog */
/* No counterpart in the reference (RK4 only, 2_...:653-802).  Kick-drift
 * leapfrog with the node velocities at half steps, stresses and positions
 * at whole steps: the node accelerations of the step's state (the RK4
 * stage with c1 = 0, as its first stage computes them) kick the velocities
 * by dt (dt / 2 on the first step from a state whose velocities are at a
 * whole step); the stress points take the new velocities and their stress
 * rate from them updates the stress by dt; the positions drift by dt with
 * the new velocities (time_integration's position update, 2_...:140-147,
 * with v instead of the trapezoid's half sum).  One pair-loop pass per half
 * instead of RK4's four stages. */
static void leapfrog(struct sps_handle *h, double dt_sph, double time_sph)
{
    const int nn = h->nn, nt = h->nt;
    const int ncrit = (int)h->p.props[1];
    const double kdt = h->lf_first ? 0.5 * dt_sph : dt_sph;
    memcpy(h->vel0, h->vel, 2 * (size_t)nt * sizeof(double));
    memcpy(h->stress0, h->stress, 4 * (size_t)nt * sizeof(double));
    memset(h->rk_dev, 0, (size_t)nt * sizeof(double));
    memset(h->source_stress, 0, 4 * (size_t)nt * sizeof(double));
    memset(h->spin, 0, 4 * (size_t)nt * sizeof(double));
    memset(h->art_visc, 0, 2 * (size_t)nn * sizeof(double));
    memset(h->vel, 0, 2 * (size_t)h->nt2 * sizeof(double));
    memset(h->stress, 0, 4 * (size_t)h->nt2 * sizeof(double));
    const double factg = gravity_factor(&h->p, time_sph);
    const int do_bforce = h->p.inside_approach && (h->p.bforce_gate ? 1 : h->p.dummy_nodes);
    /* the state of the step (RK4's first stage) and its node accelerations */
    for (int i = 0; i < nn; ++i)
        for (int a = 0; a < 2; ++a) h->vel[2 * i + a] = h->vel0[2 * i + a] + 0.0;
    for (int i = nn; i < nt; ++i)
        for (int q = 0; q < 4; ++q) h->stress[4 * i + q] = h->stress0[4 * i + q] + 0.0;
    if (ncrit == 12) adapt_stress2(h);
    normal_bcs(h, time_sph);
    stress_point_update(h);
    if (ncrit == 12) adapt_stress2(h);
    normal_bcs(h, time_sph);
    get_derivatives(h);
    gravity_force(h, factg);
    if (do_bforce) boundary_forces(h);
    if (h->p.alpha > 0 || h->p.beta > 0) artificial_viscosity(h);
    for (int i = 0; i < nn; ++i)
        for (int a = 0; a < 2; ++a) {
            h->rhs2[2 * i + a] = -h->divf2[2 * i + a] + h->g_grav[2 * i + a] + h->art_visc[2 * i + a] +
                                 h->f_bound[2 * i + a] + 0.0;
            h->vel[2 * i + a] = h->vel0[2 * i + a] + kdt * h->rhs2[2 * i + a];   /* kick */
        }
    normal_bcs(h, time_sph);
    /* the stress points' velocities and stress rates from the kicked ones */
    stress_point_update(h);
    if (ncrit == 12) adapt_stress2(h);
    normal_bcs(h, time_sph);
    get_derivatives(h);
    for (int j = nn; j < nt; ++j) {
        double Gs[4], der1;
        plastic_terms(h, j, time_sph, Gs, &der1);
        for (int q = 0; q < 4; ++q) h->source_stress[4 * j + q] = Gs[q];
        h->rk_dev[j] = h->rk_dev[j] + der1 * 1.0;
    }
    if (h->p.update_x) {
        for (int i = nn; i < nt; ++i) {
            const double *gu = h->grad_u + 4 * i;
            double om1 = 0.5 * (gu[2] - gu[1]);
            double om2 = -(0.5 * (gu[2] - gu[1]));
            const double *st = h->stress + 4 * i;
            h->spin[4 * i + 0] = 2 * om1 * st[2];
            h->spin[4 * i + 1] = 2 * om2 * st[2];
            h->spin[4 * i + 2] = om2 * st[0] + om1 * st[1];
        }
    }
    for (int i = nn; i < nt; ++i)
        for (int q = 0; q < 4; ++q) {
            h->rhs1[4 * i + q] = -h->divf1[4 * i + q] + h->spin[4 * i + q] + h->source_stress[4 * i + q];
            h->stress[4 * i + q] = h->stress0[4 * i + q] + dt_sph * h->rhs1[4 * i + q];
        }
    if (ncrit == 12) adapt_stress2(h);
    normal_bcs(h, time_sph);
    for (int i = 0; i < nt; ++i) h->ddev[i] = h->rk_dev[i];
    h->lf_first = 0;
}

/* ------------------------------------------------------------ Normal_BCs */
/* 3_SPH_material_2018.f90:1644-1753 (code/ as it compiles): the value of BC
 * row b at time t, a0 + a1 sin(w t - phi) ramped by 1 - exp(-t/Tf), or a1
 * times the piecewise-linear time curve (0 outside it) ... */
static double bc_value(const struct sps_handle *h, int b, double t)
{
    const double *L = h->bc_list + 8 * b;
    const int it_curves = (int)L[2];
    if (it_curves == 0) {
        const double a0 = L[4], a1 = L[3], w = L[5], phi = L[6], tt = L[7];
        const double fact = 1.0 - exp(-t / tt);
        const double argum = w * t - phi;
        return (a0 + a1 * sin(argum)) * fact;
    }
    const int c = it_curves - 1, npts = h->curve_npts[c];
    const double *T = h->curve_t + SPS_MAX_TCURVE * c;
    const float *F = h->curve_f + SPS_MAX_TCURVE * c;
    double tt0 = -1000.0, tt1 = 0.0, v = 0.0;
    int ip;
    for (ip = 0; ip < npts - 1; ++ip) {
        tt0 = T[ip];
        tt1 = T[ip + 1];
        if (t >= tt0 && t <= tt1) break;
        tt0 = -1000.0;
    }
    if (tt0 >= 0.0) {
        const double xi = (t - tt0) / (tt1 - tt0);
        v = (1.0 - xi) * (double)F[ip] + xi * (double)F[ip + 1];
    }
    return v * L[3];
}

/* ... applied to the nodes with BC_or_not = 1, up to five BCs each from
 * bc_info(3:7,:) in order: variable 5/6 set vx/vy, 1/2/3 set sxx/syy/sxy. */
static void normal_bcs(struct sps_handle *h, double time_sph)
{
    if (h->p.no_bcs <= 0) return;
    double val[SPS_MAX_BCS];
    for (int b = 0; b < h->nbc; ++b) val[b] = bc_value(h, b, time_sph);
    OMP_FOR
    for (int i = 0; i < h->nn; ++i) {
        const int32_t *info = h->bc_info + 8 * i;
        if (h->bc_or_not[i] != 1 || info[1] == 0) continue;
        for (int k = 0; k < 5; ++k) {
            const int b = info[2 + k];
            if (b == 0) continue;
            const double v = val[b - 1], var = h->bc_list[8 * (b - 1) + 1];
            if (var == 5) h->vel[2 * i] = v;
            else if (var == 6) h->vel[2 * i + 1] = v;
            else if (var == 1) h->stress[4 * i] = v;
            else if (var == 3) h->stress[4 * i + 2] = v;
            else if (var == 2) h->stress[4 * i + 1] = v;
        }
    }
    /* apply_stress_free (3_...:1758-1841, BCs 3_...:1633-1636): the free-
     * surface nodes without a wall-dummy partner keep only the tangential
     * stress of their normal frame; nodes whose normal is NaN are skipped */
    if (h->ifsigman != 1) return;
    OMP_FOR
    for (int i = 0; i < h->nn; ++i) {
        if (h->bc_or_not[i] != 2 || h->bc_int[i] == 1) continue;
        const double costh = h->normal[2 * i], sinth = h->normal[2 * i + 1];
        if (isnan(costh) || isnan(sinth)) continue;
        double *st = h->stress + 4 * (size_t)i;
        const double s2 = sinth * sinth, c2 = costh * costh, sc = sinth * costh;
        const double sigmatt = s2 * st[0] - 2 * sc * st[2] + c2 * st[1];
        st[0] = s2 * sigmatt;
        st[1] = c2 * sigmatt;
        st[2] = -sc * sigmatt;
        st[3] = c2 * sigmatt;
    }
}

/* ------------------------------------------------ get_nodes_on_free_surface */
/* code/3_SPH_material_2018.f90:1116-1411, called after the position update
 * (2_...:150-154).  Its locals are single precision, normal and subset are
 * the reference's double-precision globals; every mixed expression below is
 * promoted and rounded where the Fortran assignment does it.  x**0.5 is a
 * square root there (the compiler lowers it to sqrtss / sqrtsd) and 2**0.5
 * the single-precision sqrt(2). */
static int fs_hit(const struct sps_handle *h, const float *tt, const float *tau, int r, int p)
{
    /* does partner p fall in the scan region of particle r (3_...:1250-1278) */
    const double *xr = h->x + 2 * (size_t)r, *xp = h->x + 2 * (size_t)p;
    const double d0 = xp[0] - xr[0], d1 = xp[1] - xr[1];
    const float dist = (float)sqrt(d0 * d0 + d1 * d1);
    const float v0 = (float)(xp[0] - (double)tt[2 * r]), v1 = (float)(xp[1] - (double)tt[2 * r + 1]);
    const float vnorm = sqrtf(v0 * v0 + v1 * v1);
    const float prod = (float)(fabs(h->normal[2 * r] * (double)v0 + h->normal[2 * r + 1] * (double)v1) +
                               (double)fabsf(tau[2 * r] * v0 + tau[2 * r + 1] * v1));
    const double hr = h->hsml[r];
    const float limit = (float)((double)1.41421354f * hr);
    if (dist >= limit && (double)vnorm < hr) return 1;   /* set 1 */
    if (dist < limit && (double)prod < hr) return 1;     /* set 2 */
    return 0;
}

static void fs_hf(const struct sps_handle *h, int64_t k, float *h1i, float *h1j, float *h2i, float *h2j)
{
    const int i = h->pi[k], j = h->pj[k];
    *h1i = (float)(h->mass[j] * (double)h->pgx[k] / h->rho[j]);
    *h1j = (float)(-(h->mass[i] * (double)h->pgx[k] / h->rho[i]));
    *h2i = (float)(h->mass[j] * (double)h->pgy[k] / h->rho[j]);
    *h2j = (float)(-(h->mass[i] * (double)h->pgy[k] / h->rho[i]));
}

/* the work arrays of the detection, and its three pair loops (each end's
 * sums where this thread owns it) */
typedef struct FsWork {
    float *A, *ff, *tt, *tau, *fint, *gf, *nb;
} FsWork;

static inline void fs_pair1(struct sps_handle *h, int64_t k, int lo, int hi, FsWork *w)
{
    const int t = h->pt[k];
    if (t != 1 && t != 2 && t != 3) return;
    const int i = h->pi[k], j = h->pj[k];
    const double *xi = h->x + 2 * (size_t)i, *xj = h->x + 2 * (size_t)j;
    float h1i, h1j, h2i, h2j;
    fs_hf(h, k, &h1i, &h1j, &h2i, &h2j);
    if (owns(i, lo, hi)) {
        float *ai = w->A + 4 * (size_t)i;
        ai[0] = (float)((double)ai[0] + (xj[0] - xi[0]) * (double)h1i);
        w->ff[2 * i] = w->ff[2 * i] + h1i;
        ai[1] = (float)((double)ai[1] + (xj[1] - xi[1]) * (double)h1i);
        ai[2] = (float)((double)ai[2] + (xj[0] - xi[0]) * (double)h2i);
        ai[3] = (float)((double)ai[3] + (xj[1] - xi[1]) * (double)h2i);
        w->ff[2 * i + 1] = w->ff[2 * i + 1] + h2i;
    }
    if (owns(j, lo, hi)) {
        float *aj = w->A + 4 * (size_t)j;
        aj[0] = (float)((double)aj[0] + (xi[0] - xj[0]) * (double)h1j);
        w->ff[2 * j] = w->ff[2 * j] + h1j;
        aj[1] = (float)((double)aj[1] + (xi[1] - xj[1]) * (double)h1j);
        aj[2] = (float)((double)aj[2] + (xi[0] - xj[0]) * (double)h2j);
        aj[3] = (float)((double)aj[3] + (xi[1] - xj[1]) * (double)h2j);
        w->ff[2 * j + 1] = w->ff[2 * j + 1] + h2j;
    }
}

static inline void fs_pair3(struct sps_handle *h, int64_t k, int lo, int hi, FsWork *w)
{
    const int t = h->pt[k], i = h->pi[k], j = h->pj[k];
    int32_t *sub = h->subset;
    if (t == 2 || t == 3) {
        if (owns(i, lo, hi) && sub[i] == 0 && fs_hit(h, w->tt, w->tau, i, j)) { sub[i] = 2; w->fint[i] = -1; }
        if (owns(j, lo, hi) && sub[j] == 0 && fs_hit(h, w->tt, w->tau, j, i)) { sub[j] = 2; w->fint[j] = -1; }
    }
    if ((t == 6 || t == 9) && owns(j, lo, hi) && sub[j] == 0 && fs_hit(h, w->tt, w->tau, j, i)) {
        sub[j] = 2;
        w->fint[j] = -1;
    }
}

static inline void fs_pair4(struct sps_handle *h, int64_t k, int lo, int hi, FsWork *w)
{
    const int t = h->pt[k];
    if (t != 1 && t != 2 && t != 3) return;
    const int i = h->pi[k], j = h->pj[k];
    const int32_t *sub = h->subset;
    double *nrm = h->normal;
    if (sub[i] == 1 && sub[j] == 1) {
        const float x1 = (float)h->x[2 * i], y1 = (float)h->x[2 * i + 1];
        const float x2 = (float)h->x[2 * j], y2 = (float)h->x[2 * j + 1];
        const float xv = x2 - x1, yv = y2 - y1;
        if (owns(i, lo, hi)) {
            nrm[2 * i] = nrm[2 * i] - (double)yv;
            nrm[2 * i + 1] = nrm[2 * i + 1] + (double)xv;
            w->nb[i] = w->nb[i] + 1.0f;
        }
        if (owns(j, lo, hi)) {
            nrm[2 * j] = nrm[2 * j] - (double)yv;
            nrm[2 * j + 1] = nrm[2 * j + 1] + (double)xv;
            w->nb[j] = w->nb[j] + 1.0f;
        }
    }
    float h1i, h1j, h2i, h2j;
    fs_hf(h, k, &h1i, &h1j, &h2i, &h2j);
    const float *fint = w->fint;
    if (owns(i, lo, hi)) {
        w->gf[2 * i] = w->gf[2 * i] + (fint[j] - fint[i]) * h1i;
        w->gf[2 * i + 1] = w->gf[2 * i + 1] + (fint[j] - fint[i]) * h2i;
    }
    if (owns(j, lo, hi)) {
        w->gf[2 * j] = w->gf[2 * j] + (fint[i] - fint[j]) * h1j;
        w->gf[2 * j + 1] = w->gf[2 * j + 1] + (fint[i] - fint[j]) * h2j;
    }
}

static int free_surface(struct sps_handle *h)
{
    const int nt = h->nt;
    const double *x = h->x;
    double *nrm = h->normal;
    int32_t *sub = h->subset, *bc = h->bc_or_not;
    float *A = calloc(4 * (size_t)nt, sizeof(float)), *ff = calloc(2 * (size_t)nt, sizeof(float));
    float *tt = calloc(2 * (size_t)nt, sizeof(float)), *tau = calloc(2 * (size_t)nt, sizeof(float));
    float *fint = calloc((size_t)nt, sizeof(float)), *gf = calloc(2 * (size_t)nt, sizeof(float));
    float *nb = calloc((size_t)nt, sizeof(float));
    if (!A || !ff || !tt || !tau || !fint || !gf || !nb) {
        free(A); free(ff); free(tt); free(tau); free(fint); free(gf); free(nb);
        return fail(SPS_ENOMEM, "free-surface work arrays");
    }
    memset(sub, 0, (size_t)nt * sizeof(int32_t));
    /* step 1 (3_...:1160-1225): renormalisation matrix and sum of m/rho grad W
     * over the node-stress point, stress point and node pairs, in list order */
    FsWork fw = {A, ff, tt, tau, fint, gf, nb};
    PAIR_LOOP(h, fs_pair1, FsWork, fw);
    /* step 2 (3_...:1229-1245): first normal estimate; step 3's scan frame */
#pragma omp parallel for num_threads(h->nthreads > 0 ? h->nthreads : 1)
    for (int i = 0; i < nt; ++i) {
        float *a = A + 4 * (size_t)i;
        for (int q = 0; q < 4; ++q)
            if (fabsf(a[q]) <= 1.e-8f) a[q] = 0.0f;
        const float v1 = -(a[0] * ff[2 * i] + a[1] * ff[2 * i + 1]);
        const float v2 = -(a[2] * ff[2 * i] + a[3] * ff[2 * i + 1]);
        const float v3 = sqrtf(v1 * v1 + v2 * v2);
        nrm[2 * i] = (double)(v1 / v3);
        nrm[2 * i + 1] = (double)(v2 / v3);
        tt[2 * i] = (float)(x[2 * i] + h->hsml[i] * nrm[2 * i]);
        tt[2 * i + 1] = (float)(x[2 * i + 1] + h->hsml[i] * nrm[2 * i + 1]);
        tau[2 * i] = (float)(-nrm[2 * i + 1]);
        tau[2 * i + 1] = (float)nrm[2 * i];
    }
    /* step 3 (3_...:1247-1319): a particle with a partner in its scan region
     * is inside (subset 2, f_int -1); the rest are on the surface */
    PAIR_LOOP(h, fs_pair3, FsWork, fw);
    for (int i = 0; i < nt; ++i)
        if (sub[i] == 0) sub[i] = 1;
    /* 3_...:1329-1349: BC_or_not 2 on the surface, 0 inside; prescribed
     * particles (1) keep it and leave the surface set */
    for (int i = 0; i < nt; ++i)
        if (bc[i] != 1) bc[i] = 0;
    for (int i = 0; i < nt; ++i) {
        if (sub[i] == 2 && bc[i] != 1) bc[i] = 0;
        else if (sub[i] == 1 && bc[i] != 1) bc[i] = 2;
        else if (sub[i] == 1 && bc[i] == 1) sub[i] = 2;
    }
    /* step 4 (3_...:1353-1389): the normal from the surface neighbours'
     * edges, oriented against the gradient of f_int */
    memset(nrm, 0, 2 * (size_t)nt * sizeof(double));
    PAIR_LOOP(h, fs_pair4, FsWork, fw);
#pragma omp parallel for num_threads(h->nthreads > 0 ? h->nthreads : 1)
    for (int i = 0; i < nt; ++i) {   /* 3_...:1391-1405 */
        if (sub[i] != 1) continue;
        nrm[2 * i] = nrm[2 * i] / (double)nb[i];
        nrm[2 * i + 1] = nrm[2 * i + 1] / (double)nb[i];
        const float nv = (float)sqrt(nrm[2 * i] * nrm[2 * i] + nrm[2 * i + 1] * nrm[2 * i + 1]);
        nrm[2 * i] = nrm[2 * i] / (double)nv;
        nrm[2 * i + 1] = nrm[2 * i + 1] / (double)nv;
        const float ps = (float)((double)gf[2 * i] * nrm[2 * i] + (double)gf[2 * i + 1] * nrm[2 * i + 1]);
        if (ps < 0) {
            nrm[2 * i] = -nrm[2 * i];
            nrm[2 * i + 1] = -nrm[2 * i + 1];
        }
    }
    free(A); free(ff); free(tt); free(tau); free(fint); free(gf); free(nb);
    return 0;
}

/* ------------------------------------------------------- time_integration */
/* code/2_SPH_main_2018.f90:78-184 for the supported flag set */
/* ----------------------------------------------------------- XSPH_update */
/* code/2_SPH_main_2018.f90:189-239: positions moved with the velocity plus
 * eps = 0.5 times the XSPH average over stress-point pairs (type 2) and node
 * pairs (type 3), each pair added to both ends in list order.  Its n_int and
 * bc_or_not bookkeeping feed only the boundary conditions and the outside
 * approach, neither on the supported path. */
static void xsph_update(struct sps_handle *h, double dt)
{
    const int nt = h->nt;
    double *vs = h->vel_tmp; /* free again after stress_point_update */
    memset(vs, 0, 2 * (size_t)nt * sizeof(double));
    for (int64_t k = 0; k < h->n_pairs; ++k) {
        const int64_t d = h->order[k];
        const int t = h->pt[d];
        if (t != 2 && t != 3) continue;
        const int i = h->pi[d], j = h->pj[d];
        if (t == 3) {   /* 2_...:224-230: nodes next to the free surface */
            if (h->bc_or_not[i] == 2 && h->bc_or_not[j] != 2) h->bc_or_not[j] = 3;
            if (h->bc_or_not[j] == 2 && h->bc_or_not[i] != 2) h->bc_or_not[i] = 3;
        }
        const double w = (double)h->pw[d];
        for (int a = 0; a < 2; ++a) {
            const double vi = h->vel[2 * i + a], vj = h->vel[2 * j + a];
            vs[2 * i + a] = vs[2 * i + a] + ((h->mass[j] / h->rho[j]) * (vj - vi)) * w;
            vs[2 * j + a] = vs[2 * j + a] + ((h->mass[i] / h->rho[i]) * (vi - vj)) * w;
        }
    }
    const double eps = 0.5;
    for (int i = 0; i < nt; ++i)
        for (int a = 0; a < 2; ++a)
            h->x[2 * i + a] = h->x0[2 * i + a] + dt * (h->vel[2 * i + a] + eps * vs[2 * i + a]);
}

/* ---------------------------------------------------- shift_stress_points */
/* code/2_SPH_main_2018.f90:244-370, the outside approach after the position
 * update.  On shift steps the stress points of each node are placed again
 * relative to it: along its velocity (vel_vector: two points at +-dx/2 in
 * the direction of travel, components in (0, dx/5) pushed out by dx/3, and
 * only when the node moved more than disp_tol*dx since the last shift), or
 * at the fixed offsets r_x, r_y of 2_...:293-326 (whose second point of two
 * takes r_x for its y offset).  Every step, the points of a node that has
 * a wall-dummy pair and moves faster than 0.4, or has fewer than two node
 * pairs, are put on the node (the first three points: 2_...:352-366).
 * n_int and bc_int count the current list's type-3 and type-6 pairs, as
 * artificial_viscosity / XSPH_update / isolated_nodes and get_derivatives
 * leave them (2_...:221-222, 392-393, 579, 870-871). */
static void shift_stress_points(struct sps_handle *h, int itimestep)
{
    const int nn = h->nn, np = h->p.npoints;
    const double dx = h->p.dx;
    double *x = h->x;
    if (itimestep % h->p.shift_update == 0) {
        if (h->p.vel_vector) {
            for (int i = 0; i < nn; ++i) {
                const double d0 = x[2 * i] - h->x10[2 * i], d1 = x[2 * i + 1] - h->x10[2 * i + 1];
                h->disp10[i] = sqrt(d0 * d0 + d1 * d1);
                h->disp10[i] = h->disp10[i] / dx;
            }
            memcpy(h->x10, x, 2 * (size_t)nn * sizeof(double));
            for (int i = 0; i < nn; ++i) {
                const double v0 = h->vel[2 * i], v1 = h->vel[2 * i + 1];
                const float ct = (float)(v0 / sqrt(v0 * v0 + v1 * v1));
                const float sn = (float)(v1 / sqrt(v0 * v0 + v1 * v1));
                float r[2] = {(float)((dx / 2) * (double)ct), (float)((dx / 2) * (double)sn)};
                for (int a = 0; a < 2; ++a)
                    if (r[a] > 0.0f && (double)r[a] < dx / 5) r[a] = (float)((double)r[a] + dx / 3.0);
                for (int a = 0; a < 2; ++a)
                    if (r[a] < 0.0f && (double)r[a] > -(dx / 5)) r[a] = (float)((double)r[a] - dx / 3.0);
                if (h->disp10[i] > h->p.disp_tol) {
                    const int k = nn + 2 * i;
                    for (int a = 0; a < 2; ++a) {
                        x[2 * k + a] = x[2 * i + a] + (double)r[a];
                        x[2 * (k + 1) + a] = x[2 * i + a] - (double)r[a];
                    }
                }
            }
        } else {
            const double rx = h->p.r_x, ry = h->p.r_y;
            for (int i = 0; i < nn && np <= 3; ++i) {
                double *s = x + 2 * (size_t)(nn + np * i);
                const double xi = x[2 * i], yi = x[2 * i + 1];
                if (np == 1) {
                    s[0] = xi + rx; s[1] = yi + ry;
                } else if (np == 2) {
                    s[0] = xi + rx; s[1] = yi + ry;
                    s[2] = xi - rx; s[3] = yi - rx;
                } else {
                    s[0] = xi; s[1] = yi + ry;
                    s[2] = xi - rx; s[3] = yi - ry;
                    s[4] = xi + rx; s[5] = yi - ry;
                }
            }
        }
    }
    memset(h->n_int, 0, (size_t)nn * sizeof(float));
    memset(h->bc_int, 0, (size_t)nn * sizeof(int32_t));
    for (int64_t k = 0; k < h->n_pairs; ++k) {
        if (h->pt[k] == 3) {
            h->n_int[h->pi[k]] = h->n_int[h->pi[k]] + 1.0f;
            h->n_int[h->pj[k]] = h->n_int[h->pj[k]] + 1.0f;
        } else if (h->pt[k] == 6) {
            h->bc_int[h->pj[k]] = 1; /* pair_j: the node (pint_update) */
        }
    }
    for (int i = 0; i < nn; ++i) {
        const double v0 = h->vel[2 * i], v1 = h->vel[2 * i + 1];
        const float abs_vel = (float)sqrt(v0 * v0 + v1 * v1);
        const int onto = (h->bc_int[i] == 1 && abs_vel > 0.4f) || h->n_int[i] < 2.0f;
        for (int p = 0; onto && p < np && p < 3; ++p) {
            const int k = nn + np * i + p;
            x[2 * k] = x[2 * i];
            x[2 * k + 1] = x[2 * i + 1];
        }
    }
}

static int time_integration(struct sps_handle *h, int itimestep, double time_sph, double dt)
{
    const int nn = h->nn, nt = h->nt;
    const int ncrit = (int)h->p.props[1];
    /* threads pay off from some tens of thousands of particles on; below
     * (SPS_ORACLE_MIN_PARALLEL particles, default 20000) one thread runs */
    const char *minp = getenv("SPS_ORACLE_MIN_PARALLEL");
    h->nthreads = h->nt2 >= (minp ? atol(minp) : 20000) ? omp_get_max_threads() : 1;
    check_out_domain(h);
    int rc = grid_find(h, itimestep);
    h->pairs_total += h->n_pairs;
    if (rc) return rc;
    pint_update(h);
    if ((rc = pair_sublists(h))) return rc;
    if (!h->p.inside_approach && h->p.sph_shift && itimestep > 1 && (itimestep - 1) % h->p.shift_update == 0) {
        stress_point_update(h); /* 2_...:99-109 */
        if (ncrit == 12) adapt_stress2(h);
    }
    memcpy(h->vx0, h->vel, 2 * (size_t)nt * sizeof(double));
    memcpy(h->x0, h->x, 2 * (size_t)nt * sizeof(double));
    if (h->integrator == SPS_INTEGRATOR_LEAPFROG)
        leapfrog(h, dt, time_sph);
    else
        rk4(h, dt, time_sph);
    for (int i = nn; i < nt; ++i) /* update_strain, 3_...:1847-1863 */
        h->ivars[(size_t)h->niv * i] = h->ivars[(size_t)h->niv * i] + dt * h->ddev[i];
    stress_point_update(h);
    if (ncrit == 12) adapt_stress2(h);
    normal_bcs(h, time_sph);
    if (h->p.update_x && h->p.xsph) {
        xsph_update(h, dt);
        if ((rc = free_surface(h))) return rc;   /* 2_...:150-154, every step with update_x */
        if (!h->p.inside_approach) shift_stress_points(h, itimestep);
        for (int i = 0; i < nn; ++i)
            for (int a = 0; a < 2; ++a) h->displ[2 * i + a] = h->x[2 * i + a] - h->x00[2 * i + a];
    } else if (h->p.update_x) {
        OMP_FOR
        for (int i = 0; i < nt; ++i)
            for (int a = 0; a < 2; ++a) {
                if (h->integrator == SPS_INTEGRATOR_LEAPFROG) {   /* drift with the half-step velocity */
                    h->x[2 * i + a] = h->x0[2 * i + a] + dt * h->vel[2 * i + a];
                    continue;
                }
                float vel_half = (float)(0.5 * (h->vx0[2 * i + a] + h->vel[2 * i + a]));
                h->x[2 * i + a] = h->x0[2 * i + a] + (double)vel_half * dt;
            }
        if ((rc = free_surface(h))) return rc;   /* 2_...:150-154, every step with update_x */
        if (!h->p.inside_approach) shift_stress_points(h, itimestep);
        for (int i = 0; i < nn; ++i)
            for (int a = 0; a < 2; ++a) h->displ[2 * i + a] = h->x[2 * i + a] - h->x00[2 * i + a];
    } else if (h->integrator == SPS_INTEGRATOR_LEAPFROG) {
        for (int i = 0; i < nn; ++i)
            for (int a = 0; a < 2; ++a) h->displ[2 * i + a] = h->displ[2 * i + a] + dt * h->vel[2 * i + a];
    } else {
        for (int i = 0; i < nn; ++i)
            for (int a = 0; a < 2; ++a)
                h->displ[2 * i + a] = h->displ[2 * i + a] + 0.5 * (h->vx0[2 * i + a] + h->vel[2 * i + a]) * dt;
    }
    return 0;
}

int sps_set_integrator(sps_handle *h, int32_t integrator)
{
    if (!h) return fail(SPS_EINVAL, "null argument");
    if (integrator != SPS_INTEGRATOR_RK4 && integrator != SPS_INTEGRATOR_LEAPFROG)
        return fail(SPS_EINVAL, "unknown integrator %d", integrator);
    if (integrator == SPS_INTEGRATOR_LEAPFROG &&
        (!h->p.inside_approach || h->p.xsph || h->p.mode != SPS_MODE_ORDERED || h->log_cap > 0))
        return fail(SPS_EINVAL, "leapfrog: inside approach, no XSPH, SPS_MODE_ORDERED, no plastic log");
    h->integrator = integrator;
    h->lf_first = 1;
    return 0;
}

/* ------------------------------------------------------------------ C-ABI */

static int check_params(const sps_params *p)
{
    if (p->abi_version != SPS_ABI_VERSION) return fail(SPS_EINVAL, "abi_version %d != %d", p->abi_version, SPS_ABI_VERSION);
    if (p->nnode <= 0 || p->nstress <= 0 || p->ndummy < 0 || p->ndummy2 < 0 || p->ndummy2 > p->ndummy)
        return fail(SPS_EINVAL, "bad particle counts");
    if (p->nint_vars < 1) return fail(SPS_EINVAL, "nint_vars must be >= 1");
    if (p->ndimn != 2 || p->nstre != 4) return fail(SPS_EINVAL, "only ndimn=2, nstre=4 (plane strain) are supported");
    if (p->skf != 1) return fail(SPS_EINVAL, "only the cubic spline kernel (skf=1) is supported");
    if (!p->sp_sph) return fail(SPS_EINVAL, "only stress-particle SPH (sp_sph = 1) is supported");
    if (p->inside_approach && p->sph_shift) return fail(SPS_EINVAL, "sph_shift belongs to the outside approach");
    if (!p->inside_approach) {
        if (p->npoints < 1 || p->npoints > 4 || (long long)p->nstress != (long long)p->npoints * p->nnode)
            return fail(SPS_EINVAL, "outside approach: npoints must be 1..4 and nstress = npoints * nnode");
        if (p->vel_vector && p->npoints != 2) return fail(SPS_EINVAL, "vel_vector needs npoints = 2");
        if (p->vel_vector && p->mode != SPS_MODE_ORDERED)
            return fail(SPS_EINVAL, "the velocity-vector shift runs in SPS_MODE_ORDERED: it puts a node's stress "
                                    "points on the side of the sign of each velocity component, which single "
                                    "precision flips for components near zero");
        if (p->shift_update < 1) return fail(SPS_EINVAL, "shift_update must be >= 1");
    }
    if (p->reserved0 != 0) return fail(SPS_EINVAL, "reserved0 must be 0");
    if (p->cont_density || p->art_stress)
        return fail(SPS_EINVAL, "continuity density and artificial stress are not supported");
    if (p->no_bcs < 0 || p->no_bcs > SPS_MAX_BCS) return fail(SPS_EINVAL, "no_bcs must be 0..%d", SPS_MAX_BCS);
    int ntype_eco = (int)p->props[0], ncrit = (int)p->props[1];
    if (ntype_eco > 1 && ncrit == 5) return fail(SPS_EINVAL, "the Modified Cam Clay model (ncrit 5) is not supported");
    if (ntype_eco > 1 && ncrit <= 4 && (ncrit < 1 || !(p->props[6] > (double)0.001f)))
        return fail(SPS_EINVAL, "Perzyna models need ncrit 1..4 and an initial yield size props(7) > 0.001");

    if (p->ic_grav && (p->tcurve_npts < 2 || p->tcurve_npts > SPS_MAX_TCURVE))
        return fail(SPS_EINVAL, "gravity time curve needs 2..%d points", SPS_MAX_TCURVE);
    if (p->mode != SPS_MODE_ORDERED && p->mode != SPS_MODE_FAST) return fail(SPS_EINVAL, "unknown mode");
    return 0;
}

int sps_create(const sps_params *params, sps_handle **out)
{
    if (!params || !out) return fail(SPS_EINVAL, "null argument");
    int rc = check_params(params);
    if (rc) return rc;
    struct sps_handle *h = calloc(1, sizeof *h);
    if (!h) return fail(SPS_ENOMEM, "handle allocation");
    h->p = *params;
    h->nn = params->nnode;
    h->ns = params->nstress;
    h->nd = params->ndummy;
    h->nd2 = params->ndummy2;
    h->nt = h->nn + h->ns;
    h->nt2 = h->nt + h->nd;
    h->niv = params->nint_vars;
    if (alloc_all(h)) {
        free_all(h);
        free(h);
        return fail(SPS_ENOMEM, "state allocation");
    }
    *out = h;
    return 0;
}

int sps_set_state(sps_handle *h, const sps_state *st)
{
    if (!h || !st) return fail(SPS_EINVAL, "null argument");
    if (!st->x || !st->vel || !st->stress || !st->mass || !st->rho || !st->hsml || !st->itype ||
        !st->if_out_domain || !st->internal_vars || !st->f_drucker || !st->x00 || !st->displ)
        return fail(SPS_EINVAL, "sps_set_state: a required array is NULL");
    if (h->nd > 0 && (!st->wall_position || !st->horizontal_or_not))
        return fail(SPS_EINVAL, "sps_set_state: wall_position/horizontal_or_not required with dummies");
    size_t n2 = (size_t)h->nt2, n = (size_t)h->nt;
    memcpy(h->x, st->x, 2 * n2 * sizeof(double));
    memcpy(h->vel, st->vel, 2 * n2 * sizeof(double));
    memcpy(h->stress, st->stress, 4 * n2 * sizeof(double));
    memcpy(h->mass, st->mass, n2 * sizeof(double));
    memcpy(h->rho, st->rho, n2 * sizeof(double));
    memcpy(h->hsml, st->hsml, n2 * sizeof(double));
    memcpy(h->itype, st->itype, n2 * sizeof(int32_t));
    memcpy(h->ood, st->if_out_domain, n2 * sizeof(int32_t));
    memcpy(h->ivars, st->internal_vars, (size_t)h->niv * n * sizeof(double));
    memcpy(h->fdr, st->f_drucker, n * sizeof(double));
    memcpy(h->x00, st->x00, 2 * n2 * sizeof(double));
    memcpy(h->displ, st->displ, 2 * (size_t)h->nn * sizeof(double));
    if (h->nd > 0) {
        memcpy(h->wallpos, st->wall_position, n2 * sizeof(float));
        memcpy(h->horiz, st->horizontal_or_not, n2 * sizeof(float));
    }
    if (st->x_10) memcpy(h->x10, st->x_10, 2 * (size_t)h->nn * sizeof(double));
    else memcpy(h->x10, st->x, 2 * (size_t)h->nn * sizeof(double));
    if (st->disp_10) memcpy(h->disp10, st->disp_10, (size_t)h->nn * sizeof(double));
    else memset(h->disp10, 0, (size_t)h->nn * sizeof(double));
    if (bw_setup(h)) return fail(SPS_ENOMEM, "wall-dummy grid");
    h->have_state = 1;
    h->lf_first = 1;   /* velocities at a whole step */
    return 0;
}

int sps_set_bcs(sps_handle *h, const sps_bcs *b)
{
    if (!h || !b || !b->bc_list || !b->bc_info || !b->bc_or_not) return fail(SPS_EINVAL, "null argument");
    if (b->no_bcs != h->p.no_bcs || b->no_bcs < 1) return fail(SPS_EINVAL, "no_bcs %d, sps_params says %d", b->no_bcs, h->p.no_bcs);
    if ((b->ifsigman != 0 && b->ifsigman != 1) || b->reserved0 != 0) return fail(SPS_EINVAL, "ifsigman must be 0 or 1, reserved0 0");
    if (b->ncurves < 0 || b->ncurves > SPS_MAX_CURVES || (b->ncurves > 0 && (!b->curve_npts || !b->curve_t || !b->curve_f)))
        return fail(SPS_EINVAL, "bad time curves");
    for (int c = 0; c < b->ncurves; ++c)
        if (b->curve_npts[c] < 2 || b->curve_npts[c] > SPS_MAX_TCURVE) return fail(SPS_EINVAL, "curve %d: bad point count", c + 1);
    for (int k = 0; k < b->no_bcs; ++k) {
        const double cv = b->bc_list[8 * k + 2];
        if (!(cv >= 0 && (int)cv <= b->ncurves)) return fail(SPS_EINVAL, "BC %d: time curve %g of %d", k + 1, cv, b->ncurves);
    }
    for (int i = 0; i < h->nn; ++i)
        for (int k = 2; k < 7; ++k)
            if (b->bc_info[8 * i + k] < 0 || b->bc_info[8 * i + k] > b->no_bcs)
                return fail(SPS_EINVAL, "node %d: BC number %d of %d", i + 1, b->bc_info[8 * i + k], b->no_bcs);
    for (int i = 0; i < h->nt; ++i)
        if (b->bc_or_not[i] != 0 && b->bc_or_not[i] != 1)
            return fail(SPS_EINVAL, "bc_or_not[%d] = %d (0 or 1)", i, b->bc_or_not[i]);
    /* everything checked: nothing is changed by a refused call */
    h->nbc = b->no_bcs;
    h->ncurves = b->ncurves;
    memcpy(h->bc_list, b->bc_list, 8 * (size_t)b->no_bcs * sizeof(double));
    memcpy(h->bc_info, b->bc_info, 8 * (size_t)h->nn * sizeof(int32_t));
    memcpy(h->bc_or_not, b->bc_or_not, (size_t)h->nt * sizeof(int32_t));
    memset(h->normal, 0, 2 * (size_t)h->nt * sizeof(double));   /* no surface detected yet */
    h->ifsigman = b->ifsigman;
    for (int c = 0; c < b->ncurves; ++c) {
        h->curve_npts[c] = b->curve_npts[c];
        memcpy(h->curve_t + SPS_MAX_TCURVE * c, b->curve_t + SPS_MAX_TCURVE * c, SPS_MAX_TCURVE * sizeof(double));
        memcpy(h->curve_f + SPS_MAX_TCURVE * c, b->curve_f + SPS_MAX_TCURVE * c, SPS_MAX_TCURVE * sizeof(float));
    }
    h->bcs_set = 1;
    return 0;
}

/* The watchdog (no counterpart in the reference, which runs on with NaNs):
 * whether every particle in the domain box has a finite position, velocity
 * and stress.  Particles that left the box interact with nothing. */
static int state_finite(const struct sps_handle *h)
{
    for (int i = 0; i < h->nt; ++i) {
        if (h->ood[i]) continue;
        for (int a = 0; a < 2; ++a)
            if (!isfinite(h->x[2 * i + a]) || !isfinite(h->vel[2 * i + a])) return 0;
        for (int c = 0; c < 4; ++c)
            if (!isfinite(h->stress[4 * i + c])) return 0;
    }
    return 1;
}

int sps_step(sps_handle *h, int32_t nsteps, double dt, double *time_sph, int32_t *itimestep_sph)
{
    if (!h || !time_sph || !itimestep_sph) return fail(SPS_EINVAL, "null argument");
    if (!h->have_state) return fail(SPS_ESTATE, "sps_step before sps_set_state");
    if (h->p.no_bcs > 0 && !h->bcs_set) return fail(SPS_ESTATE, "no_bcs > 0: sps_set_bcs before sps_step");
    int finite = 1;
    for (int32_t s = 0; s < nsteps; ++s) {
        *itimestep_sph += 1;
        int rc = time_integration(h, *itimestep_sph, *time_sph, dt);
        if (rc) return rc;
        *time_sph = *time_sph + dt;
        h->steps++;
        finite = finite && state_finite(h);
    }
    if (!finite)
        return fail(SPS_ENUMERIC, "non-finite position, velocity or stress of a particle in the domain box "
                    "after step %d (checked at the end of each sps_step call)", *itimestep_sph);
    return 0;
}

int sps_get_state(sps_handle *h, sps_state *st)
{
    if (!h || !st) return fail(SPS_EINVAL, "null argument");
    size_t n2 = (size_t)h->nt2, n = (size_t)h->nt, nn = (size_t)h->nn;
#define OUT(f, src, cnt) do { if (st->f) memcpy(st->f, h->src, (cnt) * sizeof(*st->f)); } while (0)
    OUT(x, x, 2 * n2); OUT(vel, vel, 2 * n2); OUT(stress, stress, 4 * n2);
    OUT(mass, mass, n2); OUT(rho, rho, n2); OUT(hsml, hsml, n2);
    OUT(itype, itype, n2); OUT(if_out_domain, ood, n2);
    OUT(internal_vars, ivars, (size_t)h->niv * n); OUT(f_drucker, fdr, n);
    OUT(x00, x00, 2 * n2); OUT(displ, displ, 2 * nn);
    OUT(wall_position, wallpos, n2); OUT(horizontal_or_not, horiz, n2);
    OUT(grad_u, grad_u, 4 * n); OUT(ddev_strn, ddev, n);
    OUT(f_bound, f_bound, 2 * nn); OUT(art_visc, art_visc, 2 * nn);
    if (!h->p.inside_approach) {
        OUT(x_10, x10, 2 * nn); OUT(disp_10, disp10, nn);
    }
#undef OUT
    return 0;
}

int sps_get_pairs(sps_handle *h, int64_t cap, int32_t *pair_i, int32_t *pair_j, int32_t *pint_type,
                  float *w, float *dwdx, float *dwdy, int64_t *n)
{
    if (!h || !n) return fail(SPS_EINVAL, "null argument");
    *n = h->n_pairs;
    if (cap < h->n_pairs) return fail(SPS_ECAPACITY, "pair buffer holds %lld < %lld", (long long)cap, (long long)h->n_pairs);
    for (int64_t k = 0; k < h->n_pairs; ++k) {
        int64_t d = h->order[k];
        if (pair_i) pair_i[k] = h->pi[d] + 1;
        if (pair_j) pair_j[k] = h->pj[d] + 1;
        if (pint_type) pint_type[k] = h->pt[d];
        if (w) w[k] = h->pw[d];
        if (dwdx) dwdx[k] = h->pgx[d];
        if (dwdy) dwdy[k] = h->pgy[d];
    }
    return 0;
}

int sps_get_stats(sps_handle *h, sps_stats *out)
{
    if (!h || !out) return fail(SPS_EINVAL, "null argument");
    memset(out, 0, sizeof *out);
    out->n_pairs = h->n_pairs;
    out->m_pairs = h->m_pairs;
    for (int64_t k = 0; k < h->n_pairs; ++k)
        if (h->pt[k] >= 0 && h->pt[k] < 10) out->n_by_type[h->pt[k]]++;
    out->ndivx = h->ndivx;
    out->ndivy = h->ndivy;
    out->steps = h->steps;
    out->pairs_total = h->pairs_total;
    return 0;
}

int sps_set_materials(sps_handle *h, int32_t nmat, const double *table, const int32_t *mat)
{
    if (!h || nmat < 1 || !table || !mat) return fail(SPS_EINVAL, "bad argument");
    if ((int)h->p.props[0] > 1 && (int)h->p.props[1] <= 4)
        return fail(SPS_EINVAL, "a material per particle is for the Drucker-Prager return map, not the Perzyna models");
    for (int i = 0; i < h->nt2; ++i)
        if (mat[i] < 0 || mat[i] >= nmat) return fail(SPS_EINVAL, "particle %d: material %d of %d", i + 1, mat[i], nmat);
    free(h->mat);
    free(h->mtab);
    h->mat = (int32_t *)malloc((size_t)h->nt2 * sizeof(int32_t));
    h->mtab = (double *)malloc((size_t)nmat * 10 * sizeof(double));
    if (!h->mat || !h->mtab) return fail(SPS_ENOMEM, "material allocation");
    memcpy(h->mat, mat, (size_t)h->nt2 * sizeof(int32_t));
    memcpy(h->mtab, table, (size_t)nmat * 10 * sizeof(double));
    return 0;
}

/* The oracle launches nothing: graph replay is accepted and changes nothing. */
int sps_set_graph(sps_handle *h, int32_t enable)
{
    (void)enable;
    return h ? 0 : fail(SPS_EINVAL, "null argument");
}

int sps_set_plastic_log(sps_handle *h, int64_t cap)
{
    if (!h || cap < 0 || cap > ((int64_t)1 << 30)) return fail(SPS_EINVAL, "bad argument");
    if (cap > 0 && h->integrator == SPS_INTEGRATOR_LEAPFROG) return fail(SPS_EINVAL, "the log records RK4 steps");
    free(h->plog);
    h->plog = cap ? (double *)malloc((size_t)cap * SPS_PLOG_ROW * sizeof(double)) : NULL;
    if (cap && !h->plog) return fail(SPS_ENOMEM, "log allocation");
    h->log_cap = cap;
    h->log_n = 0;
    return 0;
}

int sps_get_plastic_log(sps_handle *h, int64_t cap, double *out, int64_t *n)
{
    if (!h || !n || cap < 0 || (cap && !out)) return fail(SPS_EINVAL, "bad argument");
    int64_t k = h->log_n < cap ? h->log_n : cap;
    if (k) memcpy(out, h->plog, (size_t)k * SPS_PLOG_ROW * sizeof(double));
    *n = k;
    return 0;
}

int sps_return_map(int64_t n, const double *in, const double *tanfi, const double *coh, double *out)
{
    if (n < 0 || (n && (!in || !out || !tanfi || !coh))) return fail(SPS_EINVAL, "bad argument");
    for (int64_t i = 0; i < n; ++i) {
        memcpy(out + 4 * i, in + 4 * i, 4 * sizeof(double));
        adapt_one(out + 4 * i, tanfi[i], coh[i]);
    }
    return 0;
}

int sps_return_map_noalloc(int64_t n, const double *in, const double *tanfi, const double *coh, double *out)
{
    if (n > SPS_RETURN_MAP_NOALLOC_MAX)
        return fail(SPS_EINVAL, "%lld stresses exceed SPS_RETURN_MAP_NOALLOC_MAX (%d)", (long long)n,
                    SPS_RETURN_MAP_NOALLOC_MAX);
    return sps_return_map(n, in, tanfi, coh, out);
}

/* The learned stress update runs on the GPU only; the oracle has no network
 * (its CPU check is oracle/mlp_oracle.py). */
int64_t sps_mlp_param_count(void) { return 37264; }
int sps_mlp_create(int32_t batch_rows, uint64_t seed, sps_mlp **out)
{
    (void)batch_rows; (void)seed;
    if (out) *out = NULL;
    return fail(SPS_EINVAL, "the oracle has no network");
}
void sps_mlp_destroy(sps_mlp *ml) { (void)ml; }
int sps_mlp_get_params(sps_mlp *ml, float *o) { (void)ml; (void)o; return fail(SPS_EINVAL, "the oracle has no network"); }
int sps_mlp_set_params(sps_mlp *ml, const float *i) { (void)ml; (void)i; return fail(SPS_EINVAL, "the oracle has no network"); }
int sps_mlp_forward(sps_mlp *ml, int64_t n, const float *x, float *y)
{
    (void)ml; (void)n; (void)x; (void)y;
    return fail(SPS_EINVAL, "the oracle has no network");
}
int sps_mlp_gradient(sps_mlp *ml, int64_t n, const float *x, const float *t, float *g, float *l)
{
    (void)ml; (void)n; (void)x; (void)t; (void)g; (void)l;
    return fail(SPS_EINVAL, "the oracle has no network");
}
int sps_mlp_train(sps_mlp *ml, int64_t n, const float *x, const float *t, int32_t e, int32_t b, float lr,
                  uint64_t seed, float *el)
{
    (void)ml; (void)n; (void)x; (void)t; (void)e; (void)b; (void)lr; (void)seed; (void)el;
    return fail(SPS_EINVAL, "the oracle has no network");
}
int sps_set_surrogate(sps_handle *h, sps_mlp *ml, double s, double c, double t)
{
    (void)s; (void)c; (void)t;
    if (!h) return fail(SPS_EINVAL, "null handle");
    return ml ? fail(SPS_EINVAL, "the oracle has no network") : 0;
}
int sps_get_surrogate_stats(sps_handle *h, int64_t *ev, int64_t *kept)
{
    if (!h || !ev || !kept) return fail(SPS_EINVAL, "null argument");
    *ev = *kept = 0;
    return 0;
}

/* The oracle searches every step whatever the skin (it runs every mode as
 * SPS_MODE_ORDERED); it accepts a skin so that the same driver runs on it. */
int sps_set_list_skin(sps_handle *h, double skin)
{
    if (!h) return fail(SPS_EINVAL, "null argument");
    if (!(skin >= 0.0)) return fail(SPS_EINVAL, "a list skin must be >= 0");
    return 0;
}

int sps_get_search_count(sps_handle *h, int64_t *n)
{
    if (!h || !n) return fail(SPS_EINVAL, "null argument");
    *n = h->steps;
    return 0;
}

int sps_set_profiling(sps_handle *h, int32_t enable)
{
    if (!h) return fail(SPS_EINVAL, "null argument");
    return 0;
}

int sps_get_profile(sps_handle *h, int32_t cap, sps_kernel_time *out, int32_t *n, int32_t reset)
{
    if (!h || !n) return fail(SPS_EINVAL, "null argument");
    *n = 0;
    return 0;
}

void sps_destroy(sps_handle *h)
{
    if (!h) return;
    free_all(h);
    free(h->plog);
    free(h->mat);
    free(h->mtab);
    free(h);
}

int sps_get_surface(sps_handle *h, int32_t *bc_or_not, double *normal)
{
    if (!h) return fail(SPS_EINVAL, "null argument");
    if (h->p.no_bcs > 0 && !h->bcs_set) return fail(SPS_ESTATE, "no_bcs > 0: sps_set_bcs before sps_get_surface");
    if (h->p.no_bcs <= 0 && !h->p.update_x)
        return fail(SPS_ESTATE, "the free surface is tracked with BCs or the position update only");
    if (bc_or_not) memcpy(bc_or_not, h->bc_or_not, (size_t)h->nt * sizeof(int32_t));
    if (normal) memcpy(normal, h->normal, 2 * (size_t)h->nt * sizeof(double));
    return 0;
}
