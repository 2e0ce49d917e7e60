// sps_api.hip - C-ABI of the Stress-Particle SPH hot path (include/sps.h) on
// one gfx950 GPU: handle, device buffers and the launch sequence of a step.
//
// One step is the reference's time_integration (code/2_SPH_main_2018.f90
// :78-184) for the supported flag set:
//   Check_Out_Domain, grid_find_NEW + Pint_Update  -> k_check_ood .. k_split_partner
//   RK4 (2_...:653-802)                            -> k_stage_first, 4 x stage, k_rk_final
//   update_strain, stress_point_update, adapt      -> k_rk_final, k_spu_*
//   position update and displacement               -> k_position
// The search bins the particles and fixes the order of the pair list; every
// stage kernel then walks each receiver's cell stencil in that order and
// recomputes w and grad w, so no pair list is stored.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#include <mutex>

#include "../../include/sps.h"
#include "sps_kernels.hip"
#include "sps_mixed.hip"
#include "sps_surrogate.hip"

using namespace sps;

static thread_local char g_err[512];

static int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(SPS_EDEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                \
    } while (0)

static inline unsigned nblk(long long n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }
// grid of the kernels that place their tiles with xcd_block(): a multiple of 8
static inline unsigned nblk8(long long n) {
    unsigned b = nblk(n);
    return (b + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
}
// tiles of TILE receivers, a multiple of the XCD count
static inline unsigned ntile8(long long n) {
    unsigned b = (unsigned)((n + TILE - 1) / TILE);
    return (b + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
}


template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
};

template <typename T>
static int dalloc(DBuf<T> &b, size_t n) {
    if (b.n >= n && b.p) return 0;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.n = 0;
    size_t cnt = n ? n : 1;
    hipError_t e = hipMalloc((void **)&b.p, cnt * sizeof(T));
    if (e != hipSuccess) return fail(SPS_ENOMEM, "hipMalloc of %zu bytes: %s", cnt * sizeof(T), hipGetErrorString(e));
    b.n = cnt;
    return 0;
}

template <typename T>
static void dfree(DBuf<T> &b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.n = 0;
}

struct StateSet {
    DBuf<double2> x, vel, x00, displ;
    DBuf<double4> stress;
    DBuf<double> mass, rho, hsml, ivar1, fdr;
    DBuf<int> orig, ood;
    DBuf<float> wallpos, horiz;
    DBuf<int> mat;                 // per-slot materials, when a table is set
    const Material *mtab = nullptr;

    StateBufs view() {
        StateBufs s;
        s.x = x.p; s.vel = vel.p; s.x00 = x00.p; s.displ = displ.p; s.stress = stress.p;
        s.mass = mass.p; s.rho = rho.p; s.hsml = hsml.p; s.ivar1 = ivar1.p; s.fdr = fdr.p;
        s.orig = orig.p; s.ood = ood.p; s.wallpos = wallpos.p; s.horiz = horiz.p;
        s.mat = mtab ? mat.p : nullptr;
        s.mtab = mtab;
        return s;
    }
    int alloc(int nn, int nt, int nt2) {
        int rc = 0;
        rc |= dalloc(x, nt2); rc |= dalloc(vel, nt2); rc |= dalloc(x00, nt2); rc |= dalloc(displ, nn);
        rc |= dalloc(stress, nt2); rc |= dalloc(mass, nt2); rc |= dalloc(rho, nt2);
        rc |= dalloc(hsml, nt2); rc |= dalloc(ivar1, nt); rc |= dalloc(fdr, nt);
        rc |= dalloc(orig, nt2); rc |= dalloc(ood, nt2); rc |= dalloc(wallpos, nt2);
        rc |= dalloc(horiz, nt2);
        return rc;
    }
    int mat_alloc(int nt2) { return dalloc(mat, nt2); }
    void release() {
        dfree(x); dfree(vel); dfree(x00); dfree(displ); dfree(stress); dfree(mass); dfree(rho);
        dfree(hsml); dfree(ivar1); dfree(fdr); dfree(orig); dfree(ood); dfree(wallpos); dfree(horiz);
        dfree(mat);
    }
};

struct sps_handle {
    sps_params p;
    DevParams dp;
    int nn, ns, nd, nd2, nt, nt2, niv;
    int device;
    hipStream_t stream = nullptr;
    bool have_state = false;
    bool poisoned = false;
    long long steps = 0;
    // host copy of internal_vars rows the hot path never touches
    std::vector<double> ivars_host;
    // state: current order and a scratch set for the permutation
    StateSet S[2];
    int cur = 0;
    // step work arrays
    DBuf<double2> x0, vel0, rk_vel, rhs2, f_bound, art_visc;   // vel0 is also the step's vx0
    DBuf<double4> stress0, rk_stress, rhs1, grad_u;
    // the stress points' stresses after a stage's second return map, which
    // k_sp_update writes here so that k_spu_nodes, reading the first, can
    // run beside it (on stream2; ev_fork / ev_join order the two)
    DBuf<double4> sig2;
    DBuf<float2> av_nodes;   // the nodes' artificial viscosity of a stage (k_av_nodes)
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    DBuf<double> rk_dev, ddev;
    // search
    int cell_cap = 0;
    DBuf<Grid> grid;
    DBuf<Split> split;
    DBuf<double> bparts;
    int nbparts = 0;
    DBuf<int> cellkey, new_cell, cell, kcount, kstart, kfill, tmp_orig, tmp_src, proc, proc_to_slot;
    DBuf<unsigned char> moved;    // per block of slots: the permutation moves a particle of it
    DBuf<unsigned> cnt, off, scan_tmp, ctot, cstart_rm;
    // per-step pair lists of the tiled stage kernels
    DBuf<unsigned long long> list_ia, list_ib;
    DBuf<float4> list_wa, list_gxa, list_gya, list_gxb, list_gyb, list_wb;
    bool xsph = false;            // XSPH position update (list B also for stress points, with w)
    DBuf<double2> xsph_v;
    // the outside approach: each node's stress points follow it
    bool outside = false;
    DBuf<int> inv, nodeinfo;      // slot of each original particle; per node slot n_int, bc_int
    // ifsigman = 1 with wall dummies: apply_stress_free skips nodes with bc_int
    // = 1 (a node-dummy pair), which the first BC calls of a step take from the
    // previous step's list: kept by original index at the end of each step
    bool bcint = false;
    DBuf<int> bcint_prev;
    DBuf<double2> x10;            // x_10 and disp_10 by original node index
    DBuf<double> disp10;
    // prescribed BCs (sps_set_bcs): tables on the host, node rows on the device
    bool bcs_set = false;
    bool surface_pending = false;  // the last step's free-surface detection is still to run
    int integrator = SPS_INTEGRATOR_RK4;
    bool lf_first = true;          // leapfrog: the velocities are at a whole step (kick by dt / 2)
    int nbc = 0, ncurves = 0;
    double bc_list[8 * SPS_MAX_BCS];
    int curve_npts[SPS_MAX_CURVES];
    double curve_t[SPS_MAX_TCURVE * SPS_MAX_CURVES];
    float curve_f[SPS_MAX_TCURVE * SPS_MAX_CURVES];
    DBuf<int> bc_info, bc_or_not;  // bc_or_not: every particle, by original index
    int ifsigman = 0;
    // free-surface detection: normals by original index; the pair list in
    // traversal order and each particle's entries (free_surface)
    DBuf<double2> normal, surf_nrm0;
    DBuf<int> surf_pi, surf_pj, surf_pt, surf_sub;
    DBuf<float> surf_gx, surf_gy;
    DBuf<unsigned> surf_cnt, surf_start, surf_cur, surf_ent, surf_flag;
    DBuf<float2> surf_tt, surf_tau;
    DBuf<double> bc_vals;         // the step's values (k_bc_values)
    BcValues bc_vars{};           // the BC rows' variables (values filled per step)
    DBuf<unsigned short> deg_a, deg_b;
    DBuf<unsigned char> tile_ok;
    DBuf<Tile> tiles;
    DBuf<StagePlan> plans;
    int ntile_n = 0, ntile_s = 0;
    double h_uniform = 0.0;   // the smoothing length of every particle, or 0 when they differ
    bool mixed = false;       // SPS_MODE_FAST: the mixed-precision stage kernels
    // pair lists kept over steps (mixed form): built with this skin, reused
    // while no moving particle has gone skin / 2 from where it was at the build
    double skin = 0.0;
    bool lists_valid = false;
    long long searches = 0;
    DBuf<double2> x_build;
    DBuf<unsigned long long> dmax;
    hipEvent_t ev_disp = nullptr;
    // learned stress update (mixed form) and the plain method's training log
    sps_mlp *mlp = nullptr;
    int surr_cap = 0, log_cap = 0;
    double surr_S = 1.0, surr_Sc = 1.0, surr_tol = 0.0;
    DBuf<unsigned> surr_count;   // [0] plastic trial states this step, [1] tuples logged
    DBuf<int> surr_slot;
    DBuf<double4> surr_sn, surr_str;
    DBuf<double2> surr_mat;
    // materials (an ensemble): table and each original particle's entry
    DBuf<Material> mtab;
    DBuf<int> mat_orig;
    // the mixed form's step as a graph (per state set)
    bool use_graph = false;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    double graph_dt[2] = {0.0, 0.0};
    long long graph_steps = 0;
    DBuf<double> stepv;
    DBuf<float> surr_feat;
    DBuf<unsigned long long> surr_stats;
    DBuf<double> plog;
    DBuf<int> err;
    // host-visible summary of the last search
    unsigned long long n_pairs = 0, m_pairs = 0, m_prev = 0;
    bool search_stale = false;         // the device holds a search the host has not read (sync_search)
    DBuf<unsigned long long> pairs_total;
    unsigned long long pairs_total_host = 0;
    unsigned long long n_by_type[10] = {0};
    int ndivx = 0, ndivy = 0;
    unsigned long long *pinned = nullptr;   // readback staging
    // kernel timing
    bool prof = false;
    struct PendingEv {
        int kid;
        hipEvent_t a, b;
    };
    std::vector<PendingEv> pending;
    std::vector<hipEvent_t> ev_pool;
    double prof_ms[32] = {0};
    long long prof_n[32] = {0};

    PairLists lists() {
        return PairLists{list_ia.p,  list_ib.p,  list_wa.p, list_gxa.p,          list_gya.p,
                         list_gxb.p, list_gyb.p, xsph ? list_wb.p : nullptr, deg_a.p, deg_b.p,
                         tile_ok.p,  ntile_n,    tiles.p,   plans.p, (outside || bcint) ? nodeinfo.p : nullptr};
    }
};

// ------------------------------------------------------------ profiling
enum KernelId {
    K_CHECK_OOD, K_BOUNDS, K_CELL_COUNT, K_CELL_SCATTER, K_CELL_SORT, K_PERMUTE, K_PROC,
    K_COUNT_PAIRS, K_SCAN, K_SPLIT, K_RK_INIT, K_STAGE_BEGIN, K_SPU_NODES,
    K_SP_UPDATE, K_DERIV_NODES, K_DUMMY_STRESS, K_RK_FINAL, K_POSITION, K_SURROGATE, K_XSPH, K_SHIFT, K_BCS, K_SURF,
    K_AV_NODES, K_COUNT_DUMMY, K_NKERNELS
};
static const char *const kKernelNames[K_NKERNELS] = {
    "k_check_ood", "k_bounds", "k_cell_count", "k_cell_scatter", "k_cell_sort", "k_permute",
    "k_proc_slot", "k_count_lists", "k_scan", "k_split", "k_stage_first",
    "k_stage_begin", "k_spu_nodes", "k_sp_update", "k_deriv_nodes",
    "k_dummy_stress", "k_rk_final", "k_position", "k_surrogate", "k_xsph", "k_shift_sp", "k_bcs", "k_surf",
    "k_av_nodes", "k_count"};

static hipEvent_t prof_begin(sps_handle *h) {
    if (!h->prof) return nullptr;
    hipEvent_t e;
    if (!h->ev_pool.empty()) {
        e = h->ev_pool.back();
        h->ev_pool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        return nullptr;
    }
    (void)hipEventRecord(e, h->stream);
    return e;
}

static void prof_end(sps_handle *h, int kid, hipEvent_t e0) {
    if (!h->prof || !e0) return;
    hipEvent_t e1;
    if (!h->ev_pool.empty()) {
        e1 = h->ev_pool.back();
        h->ev_pool.pop_back();
    } else if (hipEventCreate(&e1) != hipSuccess) {
        h->ev_pool.push_back(e0);
        return;
    }
    (void)hipEventRecord(e1, h->stream);
    h->pending.push_back({kid, e0, e1});
}

// after a stream synchronisation: fold the recorded intervals into totals
static void prof_collect(sps_handle *h) {
    for (auto &pe : h->pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, pe.a, pe.b) == hipSuccess) {
            h->prof_ms[pe.kid] += ms;
            h->prof_n[pe.kid] += 1;
        }
        h->ev_pool.push_back(pe.a);
        h->ev_pool.push_back(pe.b);
    }
    h->pending.clear();
}

#define LAUNCH_LDS(KID, GRID, BLK, LDS, KERN, ...)               \
    do {                                                         \
        hipEvent_t e0_ = prof_begin(h);                          \
        KERN<<<(GRID), (BLK), (LDS), h->stream>>>(__VA_ARGS__);  \
        prof_end(h, (KID), e0_);                                 \
    } while (0)
#define LAUNCH(KID, GRID, BLK, KERN, ...) LAUNCH_LDS(KID, GRID, BLK, 0, KERN, __VA_ARGS__)

extern "C" const char *sps_last_error(void) { return g_err; }

static void drop_graphs(sps_handle *h);

#include "sps_surrogate_host.hip"
extern "C" const char *sps_backend(void) { return "hip-gfx950"; }

// Dynamic LDS of the list build: h of the window only when the smoothing
// lengths differ (its layout puts h last).
static size_t lds_count(const sps_handle *h) {
    return h->h_uniform > 0 ? (size_t)LAY_COUNT.h[KIND_NODE] : (size_t)LAY_COUNT.bytes;
}

// ------------------------------------------------------------------ scans
// Exclusive scan of n unsigned values; optional device total.  With a
// device count (sc.g set) n is only the launch bound: the kernels scan the
// first sc.get(n) values, as many as the current grid has.
static int scan_u32(sps_handle *h, const unsigned *in, unsigned *out, long long n,
                    unsigned long long *dtotal, ScanCount sc = ScanCount{nullptr, 0, 0}) {
    long long nb = (n + BLOCK * 4 - 1) / (BLOCK * 4);
    if (nb == 0) nb = 1;
    int rc = dalloc(h->scan_tmp, (size_t)nb);
    if (rc) return rc;
    LAUNCH(K_SCAN, (unsigned)nb, BLOCK, k_scan_local, in, out, h->scan_tmp.p, n, sc);
    LAUNCH(K_SCAN, 1, BLOCK, k_scan_sums, h->scan_tmp.p, (int)nb, dtotal, n, sc);
    LAUNCH(K_SCAN, (unsigned)nb, BLOCK, k_scan_add, out, h->scan_tmp.p, n, sc);
    HIP_TRY(hipGetLastError());
    return 0;
}

// Blocks of a grid-stride kernel over at most n items.
static inline unsigned nstride(long long n) {
    const unsigned b = nblk(n);
    return b < MAX_STRIDE_BLOCKS ? (b ? b : 1u) : MAX_STRIDE_BLOCKS;
}

// --------------------------------------------------------------- params
static int check_params(const sps_params *p) {
    if (p->abi_version != SPS_ABI_VERSION) return fail(SPS_EINVAL, "abi_version %d != %d", p->abi_version, SPS_ABI_VERSION);
    if (p->nnode <= 0 || p->nstress <= 0 || p->ndummy < 0 || p->ndummy2 < 0 || p->ndummy2 > p->ndummy)
        return fail(SPS_EINVAL, "bad particle counts");
    if ((long long)p->nnode + p->nstress + p->ndummy >= (1LL << 30)) return fail(SPS_EINVAL, "too many particles");
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
    if (p->no_bcs > 0 && p->update_x && p->xsph && p->mode == SPS_MODE_FAST)
        return fail(SPS_EINVAL, "prescribed BCs with XSPH and the position update: SPS_MODE_ORDERED only "
                                "(XSPH_update re-marks BC nodes next to the free surface, which the mixed "
                                "form does not detect)");
    int ntype_eco = (int)p->props[0], ncrit = (int)p->props[1];
    if (ntype_eco > 1 && ncrit == 5) return fail(SPS_EINVAL, "the Modified Cam Clay model (ncrit 5) is not supported");
    if (ntype_eco > 1 && ncrit <= 4 && (ncrit < 1 || !(p->props[6] > (double)0.001f)))
        return fail(SPS_EINVAL, "Perzyna models need ncrit 1..4 and an initial yield size props(7) > 0.001");

    if (p->ic_grav && (p->tcurve_npts < 2 || p->tcurve_npts > SPS_MAX_TCURVE))
        return fail(SPS_EINVAL, "gravity time curve needs 2..%d points", SPS_MAX_TCURVE);
    if (p->mode != SPS_MODE_ORDERED && p->mode != SPS_MODE_FAST) return fail(SPS_EINVAL, "unknown mode %d", p->mode);
    return 0;
}

// Normal_BCs' value of BC row b at time t (3_...:1671-1700): a0 + a1 sin(w t
// - phi) ramped by 1 - exp(-t/Tf), or a1 times the piecewise-linear time
// curve (0 outside it).  Computed on the host once per step: the time does
// not change inside a step.
static double bc_value(const sps_handle *h, int b, double t) {
    const double *L = h->bc_list + 8 * b;
    const int it_curves = (int)L[2];
    if (it_curves == 0) {
        const double a0 = L[4], a1 = L[3], w = L[5], phi = L[6], tt = L[7];
        const double fact = 1.0 - std::exp(-t / tt);
        const double argum = w * t - phi;
        return (a0 + a1 * std::sin(argum)) * fact;
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

static BcValues bc_values(const sps_handle *h, double t) {
    BcValues V{};
    V.nbc = h->nbc;
    for (int b = 0; b < h->nbc; ++b) {
        V.v[b] = bc_value(h, b, t);
        const double var = h->bc_list[8 * b + 1];
        V.var[b] = (signed char)((var == 5 || var == 6 || var == 1 || var == 2 || var == 3) ? (int)var : 0);
    }
    return V;
}

// gravity_force time factor at step time (3_...:2655-2680)
static double gravity_factor(const sps_params *P, double time_sph) {
    if (P->ic_grav == 0) return 0.0;
    double tt0 = -1000.0, tt1 = 0.0;
    int ip, found = 0;
    for (ip = 0; ip < P->tcurve_npts - 1; ++ip) {
        tt0 = P->tcurve_t[ip];
        tt1 = P->tcurve_t[ip + 1];
        if (time_sph >= tt0 && time_sph <= tt1) {
            found = 1;
            break;
        }
        tt0 = -1000.0;
    }
    double factg = 0.0;
    if (tt0 >= 0.0 && found) {
        double xi = (time_sph - tt0) / (tt1 - tt0);
        factg = (1.0 - xi) * (double)P->tcurve_f[ip] + xi * (double)P->tcurve_f[ip + 1];
    }
    return factg * P->ft_grav;
}

extern "C" int sps_create(const sps_params *params, sps_handle **out) {
    if (!params || !out) return fail(SPS_EINVAL, "null argument");
    int rc = check_params(params);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(SPS_EDEVICE, "no HIP device visible: the product library needs an MI355X (gfx950)");
    sps_handle *h = new (std::nothrow) sps_handle();
    if (!h) return fail(SPS_ENOMEM, "handle allocation");
    h->p = *params;
    h->nn = params->nnode;
    h->ns = params->nstress;
    h->nd = params->ndummy;
    h->nd2 = params->ndummy2;
    h->nt = h->nn + h->ns;
    h->nt2 = h->nt + h->nd;
    h->niv = params->nint_vars;
    h->mixed = params->mode == SPS_MODE_FAST;
    h->xsph = params->xsph != 0;
    h->outside = params->inside_approach == 0;
    (void)hipGetDevice(&h->device);
    DevParams &d = h->dp;
    d.nn = h->nn; d.ns = h->ns; d.nd = h->nd; d.nd2 = h->nd2; d.nt = h->nt; d.nt2 = h->nt2;
    d.cspm = params->cspm; d.update_x = params->update_x;
    d.ntype_eco = (int)params->props[0]; d.ncrit = (int)params->props[1];
    d.do_bforce = params->inside_approach && (params->bforce_gate ? 1 : params->dummy_nodes);
    d.do_av = (params->alpha > 0 || params->beta > 0);
    d.ic_grav = params->ic_grav;
    d.pi = params->pi; d.dx = params->dx; d.alpha = params->alpha; d.beta = params->beta;
    d.damping = params->damping; d.cspm_tol = params->cspm_det_tol;
    d.D11 = params->D11; d.D22 = params->D22; d.D12 = params->D12; d.D33 = params->D33;
    d.D41 = params->D41; d.D42 = params->D42;
    d.tanfi = params->props[12]; d.coh = params->props[13];
    d.young = params->props[2]; d.poiss = params->props[3];
    d.pz_fdatm0 = params->props[6]; d.pz_hards = params->props[7]; d.pz_gamma = params->props[9];
    d.pz_delta = params->props[10]; d.pz_nflow = params->props[11];
    d.pz_root3 = (double)sqrtf(3.0f);
    d.pz_snphi = std::sin(params->props[8] * (double)0.017453292f);
    for (int a = 0; a < 2; ++a) {
        d.xmin_dom[a] = params->xmin_domain[a];
        d.xmax_dom[a] = params->xmax_domain[a];
        d.cgrav[a] = params->cgrav[a];
    }
    auto bail = [&](int code) {
        sps_destroy(h);
        return code;
    };
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(SPS_EDEVICE, "hipStreamCreate failed"));
    if (hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming) != hipSuccess)
        return bail(fail(SPS_EDEVICE, "hipEventCreate failed"));
    int nt2 = h->nt2, nt = h->nt, nn = h->nn;
    rc = 0;
    rc |= h->S[0].alloc(nn, nt, nt2);
    rc |= h->S[1].alloc(nn, nt, nt2);
    rc |= dalloc(h->x0, nt); rc |= dalloc(h->vel0, nt);
    rc |= dalloc(h->rk_vel, nn); rc |= dalloc(h->rhs2, nn); rc |= dalloc(h->f_bound, nn);
    rc |= dalloc(h->art_visc, nn); rc |= dalloc(h->stress0, nt); rc |= dalloc(h->rk_stress, nt);
    rc |= dalloc(h->rhs1, nt); rc |= dalloc(h->grad_u, nt); rc |= dalloc(h->rk_dev, nt);
    rc |= dalloc(h->sig2, nt); rc |= dalloc(h->av_nodes, nn > 0 ? nn : 1);
    rc |= dalloc(h->ddev, nt);
    rc |= dalloc(h->grid, 1); rc |= dalloc(h->split, 1); rc |= dalloc(h->err, 1); rc |= dalloc(h->pairs_total, 1);
    h->nbparts = std::min(1024, (int)nblk(nt2));
    rc |= dalloc(h->bparts, 5 * (size_t)h->nbparts);
    rc |= dalloc(h->cellkey, nt2); rc |= dalloc(h->new_cell, nt2); rc |= dalloc(h->cell, nt2);
    rc |= dalloc(h->tmp_orig, nt2); rc |= dalloc(h->tmp_src, nt2); rc |= dalloc(h->proc, nt2);
    rc |= dalloc(h->moved, nblk(nt2));
    rc |= dalloc(h->proc_to_slot, nt2); rc |= dalloc(h->cnt, nt2); rc |= dalloc(h->off, nt2);
    h->ntile_n = (int)ntile8(nn);
    h->ntile_s = (int)ntile8(nt - nn);
    {
        const size_t wa = (size_t)(h->ntile_n + h->ntile_s) * LGROUPS * TILE, wb = (size_t)h->ntile_n * LGROUPS * TILE;
        rc |= dalloc(h->list_ia, wa); rc |= dalloc(h->list_wa, wa); rc |= dalloc(h->list_gxa, wa);
        rc |= dalloc(h->list_gya, wa);
        // list B: node tiles; with XSPH stress-point tiles too, and its w
        rc |= dalloc(h->list_ib, h->xsph ? wa : wb); rc |= dalloc(h->list_gxb, wb); rc |= dalloc(h->list_gyb, wb);
        if (h->xsph) rc |= dalloc(h->list_wb, wa) | dalloc(h->xsph_v, (size_t)nt);
    }
    rc |= dalloc(h->deg_a, nt); rc |= dalloc(h->deg_b, h->xsph ? nt : nn);
    rc |= dalloc(h->tile_ok, h->ntile_n + h->ntile_s + 1);
    rc |= dalloc(h->tiles, h->ntile_n + h->ntile_s + 1);
    rc |= dalloc(h->plans, 3 * (size_t)(h->ntile_n + h->ntile_s + 1));
    rc |= dalloc(h->x_build, nt); rc |= dalloc(h->dmax, 1); rc |= dalloc(h->stepv, 4);
    if (params->no_bcs > 0)
        rc |= dalloc(h->bc_info, 8 * (size_t)nn) | dalloc(h->bc_vals, SPS_MAX_BCS);
    // BC_or_not and the surface normals exist with BCs and with the position
    // update, which runs the free-surface detection (3_...:1045-1047, 918)
    if (params->no_bcs > 0 || params->update_x) rc |= dalloc(h->bc_or_not, nt) | dalloc(h->normal, nt);
    if (h->outside) {
        rc |= dalloc(h->inv, nt2); rc |= dalloc(h->nodeinfo, nn);
        rc |= dalloc(h->x10, nn); rc |= dalloc(h->disp10, nn);
    }
    // bc_int for apply_stress_free (sps_set_bcs with ifsigman = 1 and wall
    // dummies): zero before the first derivative pass (cleared below)
    if (params->no_bcs > 0 && nt2 > nt) rc |= dalloc(h->nodeinfo, nn) | dalloc(h->bcint_prev, nn);
    if (rc) return bail(rc);
    if (hipHostMalloc((void **)&h->pinned, 64 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess)
        return bail(fail(SPS_ENOMEM, "pinned staging allocation"));
    if (hipEventCreateWithFlags(&h->ev_disp, hipEventDisableTiming) != hipSuccess)
        return bail(fail(SPS_EDEVICE, "event creation failed"));
    if ((h->bc_or_not.p && (hipMemsetAsync(h->bc_or_not.p, 0, nt * sizeof(int), h->stream) != hipSuccess ||
                            hipMemsetAsync(h->normal.p, 0, nt * sizeof(double2), h->stream) != hipSuccess)) ||
        (h->bcint_prev.p && hipMemsetAsync(h->bcint_prev.p, 0, nn * sizeof(int), h->stream) != hipSuccess) ||
        hipMemsetAsync(h->f_bound.p, 0, nn * sizeof(double2), h->stream) != hipSuccess ||
        hipMemsetAsync(h->art_visc.p, 0, nn * sizeof(double2), h->stream) != hipSuccess ||
        hipMemsetAsync(h->grad_u.p, 0, nt * sizeof(double4), h->stream) != hipSuccess ||
        hipMemsetAsync(h->ddev.p, 0, nt * sizeof(double), h->stream) != hipSuccess ||
        hipMemsetAsync(h->err.p, 0, sizeof(int), h->stream) != hipSuccess ||
        hipMemsetAsync(h->split.p, 0, sizeof(Split), h->stream) != hipSuccess ||
        hipMemsetAsync(h->pairs_total.p, 0, sizeof(unsigned long long), h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        return bail(fail(SPS_EDEVICE, "initial clears failed"));
    *out = h;
    return 0;
}

// Cell capacity: the grid spans the in-domain particles, which lie inside
// the domain box, with cells of twice the largest h.
static int size_cells(sps_handle *h, const double *hsml) {
    double hmin = 1e300;
    for (int i = 0; i < h->nt2; ++i) hmin = std::min(hmin, hsml[i]);
    if (!(hmin > 0)) return fail(SPS_EINVAL, "smoothing lengths must be positive");
    double lx = h->p.xmax_domain[0] - h->p.xmin_domain[0], ly = h->p.xmax_domain[1] - h->p.xmin_domain[1];
    double cx = lx / (2 * hmin) + 3, cy = ly / (2 * hmin) + 3;
    double cells = cx * cy;
    if (!(cells > 0) || cells > (double)(1 << 28))
        return fail(SPS_EINVAL, "domain box too large for the cell grid (%.3g cells)", cells);
    int cap = (int)cells + 16;
    int rc = 0;
    size_t L = 3 * ((size_t)cap + 1) + 1;
    rc |= dalloc(h->kcount, L); rc |= dalloc(h->kstart, L); rc |= dalloc(h->kfill, L);
    rc |= dalloc(h->ctot, (size_t)cap + 2); rc |= dalloc(h->cstart_rm, (size_t)cap + 2);
    // block sums of the largest scan of a step (the bucket counts or the slots)
    rc |= dalloc(h->scan_tmp, std::max(L, (size_t)h->nt2 + 1) / (BLOCK * 4) + 2);
    if (rc) return rc;
    h->cell_cap = cap;
    return 0;
}

extern "C" int sps_set_state(sps_handle *h, const sps_state *st) {
    if (!h || !st) return fail(SPS_EINVAL, "null argument");
    if (!st->x || !st->vel || !st->stress || !st->mass || !st->rho || !st->hsml || !st->itype ||
        !st->if_out_domain || !st->internal_vars || !st->f_drucker || !st->x00 || !st->displ)
        return fail(SPS_EINVAL, "sps_set_state: a required array is NULL");
    if (h->nd > 0 && (!st->wall_position || !st->horizontal_or_not))
        return fail(SPS_EINVAL, "sps_set_state: wall_position/horizontal_or_not required with dummies");
    for (int i = 0; i < h->nt2; ++i) {
        int want = i < h->nn ? 2 : (i < h->nt ? 1 : 25);
        if (st->itype[i] != want)
            return fail(SPS_EINVAL, "itype(%d) = %d; the hot path needs 2 for nodes, 1 for stress points, 25 for dummies",
                        i + 1, st->itype[i]);
    }
    int rc = size_cells(h, st->hsml);
    // smoothing lengths do not change on the path: one value for all lets
    // the search precompute its acceptance bounds
    h->h_uniform = st->hsml[0] > 0 ? st->hsml[0] : 0.0;
    for (int i = 1; i < h->nt2 && h->h_uniform > 0; ++i)
        if (!(st->hsml[i] == h->h_uniform)) h->h_uniform = 0.0;
    h->dp.h_uni = h->h_uniform > 0;
    if (rc) return rc;
    if (h->dp.do_bforce) {
        double hmin = 1e300;
        for (int i = 0; i < h->nt2; ++i) hmin = std::min(hmin, st->hsml[i]);
        if (!(0.75 * h->p.dx < 2 * hmin))
            return fail(SPS_EINVAL, "wall force reach 0.75*dx must lie inside the kernel support");
    }
    const int nt2 = h->nt2, nt = h->nt, nn = h->nn;
    StateSet &S = h->S[h->cur];
    std::vector<double> tmp((size_t)nt2 * 4);
    hipStream_t s = h->stream;
    HIP_TRY(hipMemcpyAsync(S.x.p, st->x, nt2 * sizeof(double2), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(S.vel.p, st->vel, nt2 * sizeof(double2), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(S.stress.p, st->stress, nt2 * sizeof(double4), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(S.mass.p, st->mass, nt2 * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(S.rho.p, st->rho, nt2 * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(S.hsml.p, st->hsml, nt2 * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(S.ood.p, st->if_out_domain, nt2 * sizeof(int), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(S.fdr.p, st->f_drucker, nt * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(S.x00.p, st->x00, nt2 * sizeof(double2), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(S.displ.p, st->displ, nn * sizeof(double2), hipMemcpyHostToDevice, s));
    if (h->outside) {   // x_10 and disp_10, by original node index
        HIP_TRY(hipMemcpyAsync(h->x10.p, st->x_10 ? st->x_10 : st->x, nn * sizeof(double2), hipMemcpyHostToDevice, s));
        if (st->disp_10)
            HIP_TRY(hipMemcpyAsync(h->disp10.p, st->disp_10, nn * sizeof(double), hipMemcpyHostToDevice, s));
        else
            HIP_TRY(hipMemsetAsync(h->disp10.p, 0, nn * sizeof(double), s));
    }
    h->ivars_host.assign(st->internal_vars, st->internal_vars + (size_t)h->niv * nt);
    for (int i = 0; i < nt; ++i) tmp[i] = st->internal_vars[(size_t)h->niv * i];
    HIP_TRY(hipMemcpyAsync(S.ivar1.p, tmp.data(), nt * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));   // tmp is reused below
    std::vector<float> wp(nt2, 0.0f), hz(nt2, 0.0f);
    if (h->nd > 0) {
        std::memcpy(wp.data(), st->wall_position, nt2 * sizeof(float));
        std::memcpy(hz.data(), st->horizontal_or_not, nt2 * sizeof(float));
    }
    HIP_TRY(hipMemcpyAsync(S.wallpos.p, wp.data(), nt2 * sizeof(float), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(S.horiz.p, hz.data(), nt2 * sizeof(float), hipMemcpyHostToDevice, s));
    k_scatter_in_identity<<<nblk(nt2), BLOCK, 0, s>>>(S.orig.p, nt2);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemsetAsync(h->err.p, 0, sizeof(int), s));   // a new state starts with a clear flag
    HIP_TRY(hipStreamSynchronize(s));
    h->have_state = true;
    h->lf_first = true;
    h->surface_pending = false;
    h->lists_valid = false;
    drop_graphs(h);
    if (h->mtab.p) {   // the slots are in original order again: their materials too
        StateSet &C = h->S[h->cur];   // (the other set gets them from k_permute)
        LAUNCH(K_PERMUTE, nblk(h->nt2), BLOCK, k_material_slots, C.orig.p, h->mat_orig.p, C.mat.p, h->nt2);
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    h->poisoned = false;
    return 0;
}

// Neighbour search of one step: binning, processing order, pair counts, the
// list split and the ordered lists.  Every launch is sized from host bounds
// (the cell capacity of the domain box, the particle counts) and the kernels
// read the current grid on the device, so the search queues without a host
// round trip; its counts reach the host when asked (sync_search).
static int search(sps_handle *h, bool fresh_list) {
    const int nn = h->nn, nt = h->nt, nt2 = h->nt2;
    StateSet &A = h->S[h->cur];
    StateSet &B = h->S[1 - h->cur];
    // domain check and grid
    LAUNCH(K_CHECK_OOD, nblk(nt2), BLOCK, k_check_ood, A.x.p, A.ood.p, nt2, h->p.xmin_domain[0], h->p.xmin_domain[1],
                                           h->p.xmax_domain[0], h->p.xmax_domain[1]);
    LAUNCH(K_BOUNDS, h->nbparts, BLOCK, k_bounds_partial, A.x.p, A.hsml.p, A.ood.p, nt2, h->bparts.p);
    LAUNCH(K_BOUNDS, 1, BLOCK, k_bounds_final, h->bparts.p, h->nbparts, h->grid.p, h->cell_cap, h->skin, h->err.p);
    // bucket arrays: 3 (ndivt + 1) + 1 entries of the current grid, at most
    // those of the capacity
    const long long L3max = 3LL * (h->cell_cap + 1);
    LAUNCH(K_CELL_COUNT, nstride(L3max + 1), BLOCK, k_zero_cells, h->kcount.p, h->kfill.p, h->grid.p);
    LAUNCH(K_CELL_COUNT, nblk(nt2), BLOCK, k_cell_count, A.x.p, A.ood.p, h->grid.p, h->cellkey.p, h->kcount.p, nn, nt, nt2);
    int rc = scan_u32(h, (const unsigned *)h->kcount.p, (unsigned *)h->kstart.p, L3max + 1, nullptr,
                      ScanCount{h->grid.p, 3, 4});
    if (rc) return rc;
    LAUNCH(K_CELL_SCATTER, nblk(nt2), BLOCK, k_cell_scatter, h->cellkey.p, A.orig.p, h->kstart.p, h->kfill.p, h->grid.p,
                                              h->tmp_orig.p, h->tmp_src.p, h->cell.p, nn, nt, nt2);
    LAUNCH(K_CELL_SORT, nstride(L3max), BLOCK, k_cell_sort, h->kstart.p, h->tmp_orig.p, h->tmp_src.p, h->grid.p,
           h->err.p);
    // permute the state into (kind, cell, original index) order: the blocks
    // whose particles moved go through the scratch set and back
    LAUNCH(K_PERMUTE, nblk(nt2), BLOCK, k_permute, B.view(), A.view(), h->tmp_src.p, nn, nt, nt2, h->moved.p);
    LAUNCH(K_PERMUTE, nblk(nt2), BLOCK, k_permute_commit, A.view(), B.view(), h->moved.p, nn, nt, nt2);
    HIP_TRY(hipGetLastError());
    StateSet &C = A;
    if (h->outside) LAUNCH(K_PERMUTE, nblk(nt2), BLOCK, k_inv_slots, C.orig.p, h->inv.p, nt2);
    LAUNCH(K_PROC, nstride((long long)h->cell_cap + 1), BLOCK, k_cell_totals, h->kstart.p, h->grid.p, h->ctot.p);
    rc = scan_u32(h, h->ctot.p, h->cstart_rm.p, (long long)h->cell_cap + 1, nullptr, ScanCount{h->grid.p, 1, 1});
    if (rc) return rc;
    LAUNCH(K_PROC, nblk(nt2), BLOCK, k_proc_slot, h->cell.p, h->kstart.p, h->cstart_rm.p, h->grid.p, h->proc.p,
           nn, nt, nt2);
    LAUNCH(K_PROC, nblk(nt2), BLOCK, k_proc_to_slot, h->proc.p, h->proc_to_slot.p, nt2);
    // counts, offsets, split
    LAUNCH(K_SPLIT, nblk(nt2), BLOCK, k_surf_zero, h->cnt.p, (long long)nt2);
    LAUNCH(K_SPLIT, 1, 64, k_split_reset, h->split.p, fresh_list ? 1 : 0);
    {
        const SearchView V{h->cell.p, h->kstart.p, h->proc.p, h->cnt.p, h->off.p, h->grid.p, h->split.p, h->p.pi};
        const PairLists PL = h->lists();
        // the stress-point receivers' lists (their tiles, counts and list
        // words are their own; the type histogram is summed with atomics)
        // are built beside the nodes' on the second stream, unless kernel
        // timing (one stream) is on
        const bool fork = !h->prof && nn > 0 && nt > nn;
        if (fork) {
            HIP_TRY(hipEventRecord(h->ev_fork, h->stream));
            HIP_TRY(hipStreamWaitEvent(h->stream2, h->ev_fork, 0));
            if (!h->mixed)
                k_count_lists<KIND_SP, true><<<ntile8(nt - nn), TILE, lds_count(h), h->stream2>>>(
                    V, C.x.p, C.hsml.p, h->cnt.p, h->split.p, PL, h->h_uniform, nn, nt - nn, h->skin);
            else
                k_count_lists<KIND_SP, false><<<ntile8(nt - nn), TILE, lds_count(h), h->stream2>>>(
                    V, C.x.p, C.hsml.p, h->cnt.p, h->split.p, PL, h->h_uniform, nn, nt - nn, h->skin);
            HIP_TRY(hipEventRecord(h->ev_join, h->stream2));
        }
        // the mixed-precision stage kernels compute the kernel values themselves
        if (nn > 0 && !h->mixed)
            LAUNCH_LDS(K_COUNT_PAIRS, ntile8(nn), TILE, lds_count(h), (k_count_lists<KIND_NODE, true>), V, C.x.p,
                       C.hsml.p, h->cnt.p, h->split.p, PL, h->h_uniform, 0, nn, h->skin);
        if (nn > 0 && h->mixed)
            LAUNCH_LDS(K_COUNT_PAIRS, ntile8(nn), TILE, lds_count(h), (k_count_lists<KIND_NODE, false>), V, C.x.p,
                       C.hsml.p, h->cnt.p, h->split.p, PL, h->h_uniform, 0, nn, h->skin);
        if (!fork && nt > nn && !h->mixed)
            LAUNCH_LDS(K_COUNT_PAIRS, ntile8(nt - nn), TILE, lds_count(h), (k_count_lists<KIND_SP, true>), V, C.x.p,
                       C.hsml.p, h->cnt.p, h->split.p, PL, h->h_uniform, nn, nt - nn, h->skin);
        if (!fork && nt > nn && h->mixed)
            LAUNCH_LDS(K_COUNT_PAIRS, ntile8(nt - nn), TILE, lds_count(h), (k_count_lists<KIND_SP, false>), V, C.x.p,
                       C.hsml.p, h->cnt.p, h->split.p, PL, h->h_uniform, nn, nt - nn, h->skin);
        if (nt2 > nt)
            LAUNCH(K_COUNT_DUMMY, nblk8(nt2 - nt), BLOCK, k_count, C.x.p, C.hsml.p, h->cell.p, h->kstart.p, h->proc.p,
                   h->grid.p, h->cnt.p, h->split.p, nn, nt, nt, nt2);
        if (fork) HIP_TRY(hipStreamWaitEvent(h->stream, h->ev_join, 0));
    }
    HIP_TRY(hipGetLastError());
    // discovery offsets over all slots (slots past the in-domain count hold 0)
    rc = scan_u32(h, h->cnt.p, h->off.p, nt2, &h->split.p->n);
    if (rc) return rc;
    LAUNCH(K_SPLIT, nblk(nt2), BLOCK, k_split, h->cnt.p, h->off.p, nt2, h->split.p);
    LAUNCH(K_SPLIT, 1, 64, k_split_partner, C.x.p, C.hsml.p, h->cell.p, h->kstart.p, h->proc.p, h->off.p,
                                     h->proc_to_slot.p, h->grid.p, h->split.p, nn, nt, h->pairs_total.p);
    HIP_TRY(hipGetLastError());
    h->search_stale = true;
    return 0;
}

// The last search's counts on the host (n, M, the type histogram, the grid):
// one copy and a wait, only when a caller asks for them.
static int sync_search(sps_handle *h) {
    if (!h->search_stale) return 0;
    HIP_TRY(hipMemcpyAsync(h->pinned, h->split.p, sizeof(Split), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipMemcpyAsync(h->pinned + 16, h->grid.p, sizeof(Grid), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipMemcpyAsync(h->pinned + 24, h->pairs_total.p, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    Split sh;
    Grid gh;
    std::memcpy(&sh, h->pinned, sizeof(Split));
    std::memcpy(&gh, h->pinned + 16, sizeof(Grid));
    h->n_pairs = sh.n;
    h->m_prev = sh.M;
    h->m_pairs = std::max(sh.M, sh.n);
    for (int k = 0; k < 10; ++k) h->n_by_type[k] = sh.n_by_type[k];
    h->ndivx = gh.ndivx;
    h->ndivy = gh.ndivy;
    std::memcpy(&h->pairs_total_host, h->pinned + 24, sizeof(unsigned long long));
    h->search_stale = false;
    return 0;
}

// Whether the mixed form's lists of the last search still hold every pair:
// no moving particle has gone skin / 2 since (read from the check that the
// previous step queued).
static int lists_still_valid(sps_handle *h, bool &valid) {
    valid = false;
    if (!h->mixed || h->skin <= 0 || !h->lists_valid) return 0;
    HIP_TRY(hipEventSynchronize(h->ev_disp));
    double d2;
    std::memcpy(&d2, h->pinned + 32, sizeof d2);
    valid = 2.0 * std::sqrt(d2) < h->skin;
    return 0;
}

// The mixed form's step after the search: every launch, no host
// synchronisation and no memset, so that it can be captured in a graph.
// stepv: the scalars of a replayed step (k_set_step), or null.
static int mixed_step_kernels(sps_handle *h, double dt, double time_sph, double gxg, double gyg,
                              const double *stepv, bool pre_shift = false, int shift = -1) {
    const int nn = h->nn, nt = h->nt, nt2 = h->nt2;
    StateBufs SB = h->S[h->cur].view();
    Work W{h->vel0.p, h->x0.p, h->vel0.p, h->rk_vel.p, h->rhs2.p, h->f_bound.p, h->art_visc.p,
           h->stress0.p, h->rk_stress.p, h->rhs1.p, h->grad_u.p, h->rk_dev.p, h->ddev.p};
    const SearchView V{h->cell.p, h->kstart.p, h->proc.p, h->cnt.p, h->off.p, h->grid.p, h->split.p, h->p.pi};
    const PairLists PL = h->lists();
    const int adapt = h->dp.ncrit == 12;
    const double tanfi = h->dp.tanfi, coh = h->dp.coh;
    const double f1rk[4] = {0., 0.5, 0.5, 1.0}, f2rk[4] = {1., 2., 2., 1.0};
    const float hu = (float)h->h_uniform;
    // prescribed BCs where the reference applies them (as in the FP64 step);
    // a captured step reads the values k_bc_values leaves before each replay
    const bool bcs = h->p.no_bcs > 0 && nn > 0;
    if (bcs && !stepv) LAUNCH(K_BCS, 1, 64, k_bc_values, bc_values(h, time_sph), h->bc_vals.p);
    if (pre_shift) {
        // the outside approach's stress_point_update + adapt_stress2 on the
        // step's lists (2_...:99-109)
        if (nn > 0)
            LAUNCH_LDS(K_SPU_NODES, ntile8(nn), TILE, LAY_SPU_NODES_MIX.bytes, k_spu_nodes_mix, SB, V, PL, adapt,
                       tanfi, coh, hu, nn);
        if (nt > nn)
            LAUNCH_LDS(K_SP_UPDATE, ntile8(nt - nn), TILE, LAY_SP_UPDATE_MIX.bytes, k_sp_update_mix, SB, W, V, PL,
                       h->dp, time_sph, 0.0, 0, adapt, hu, nn, nt, stepv, nullptr);
    }
    LAUNCH(K_RK_INIT, nblk(nt2), BLOCK, k_stage_first, SB, W, adapt, tanfi, coh, nn, nt, nt2);
    for (int st = 0; st < 4; ++st) {
        const double c1 = f1rk[st] * (dt);
        if (st > 0) LAUNCH(K_STAGE_BEGIN, nblk(nt), BLOCK, k_stage_begin, SB, W, c1, adapt, tanfi, coh, nn, nt);
        if (bcs) LAUNCH(K_BCS, nblk(nn), BLOCK, k_bcs, SB, h->bc_info.p, h->bc_or_not.p, h->bc_vars, h->bc_vals.p, h->ifsigman, h->normal.p, nn, nullptr, nullptr);
        // the two halves of stress_point_update side by side, as in the FP64
        // step; one stream when kernels are timed or the step is captured
        const bool fork = !h->prof && !stepv && nn > 0 && nt > nn;
        if (fork) {
            HIP_TRY(hipEventRecord(h->ev_fork, h->stream));
            HIP_TRY(hipStreamWaitEvent(h->stream2, h->ev_fork, 0));
            hipStream_t s2 = h->stream2;
            k_spu_nodes_mix<<<ntile8(nn), TILE, LAY_SPU_NODES_MIX.bytes, s2>>>(SB, V, PL, adapt, tanfi, coh, hu, nn);
            if (bcs)
                k_bcs<<<nblk(nn), BLOCK, 0, s2>>>(SB, h->bc_info.p, h->bc_or_not.p, h->bc_vars, h->bc_vals.p,
                                                 h->ifsigman, h->normal.p, nn, nullptr, nullptr);
            // the nodes' artificial viscosity reads only node data, after the
            // node side and the second BC call (k_deriv_nodes_mix takes it)
            if (h->dp.do_av)
                k_av_nodes_mix<<<ntile8(nn), TILE, LAY_DERIV_NODES_MIX.bytes, s2>>>(SB, V, PL, h->dp, hu,
                                                                                     h->av_nodes.p, nn);
            HIP_TRY(hipEventRecord(h->ev_join, s2));
        } else {
            if (nn > 0)
                LAUNCH_LDS(K_SPU_NODES, ntile8(nn), TILE, LAY_SPU_NODES_MIX.bytes, k_spu_nodes_mix, SB, V, PL, adapt,
                           tanfi, coh, hu, nn);
            if (bcs) LAUNCH(K_BCS, nblk(nn), BLOCK, k_bcs, SB, h->bc_info.p, h->bc_or_not.p, h->bc_vars, h->bc_vals.p, h->ifsigman, h->normal.p, nn, nullptr, nullptr);
        }
        if (nt > nn)
            LAUNCH_LDS(K_SP_UPDATE, ntile8(nt - nn), TILE, LAY_SP_UPDATE_MIX.bytes, k_sp_update_mix, SB, W, V, PL,
                       h->dp, time_sph, f2rk[st], 1, adapt, hu, nn, nt, stepv, h->sig2.p);
        if (fork) HIP_TRY(hipStreamWaitEvent(h->stream, h->ev_join, 0));
        if (nn > 0)
            LAUNCH_LDS(K_DERIV_NODES, ntile8(nn), TILE, LAY_DERIV_NODES_MIX.bytes, k_deriv_nodes_mix, SB, W, V, PL,
                       h->dp, gxg, gyg, f2rk[st], h->err.p, hu, nn, stepv, nt > nn ? h->sig2.p : SB.stress,
                       fork && h->dp.do_av ? h->av_nodes.p : nullptr);
        if (st == 3 && nt2 > nt)
            LAUNCH(K_DUMMY_STRESS, nblk(nt2 - nt), BLOCK, k_dummy_stress, SB, V, nt, nt2);
    }
    const bool hook = h->mlp || h->log_cap > 0;
    if (h->mlp) LAUNCH(K_SURROGATE, 1, 64, k_zero_u32, h->surr_count.p, 1);
    LAUNCH(K_RK_FINAL, nblk(nt), BLOCK, k_rk_final, SB, W, dt, adapt, tanfi, coh, nn, nt, surrogate_dev(h),
           hook ? 1 : 0, h->err.p);
    if (h->mlp && adapt) {
        int rc = surrogate_correct(h, SB);
        if (rc) return rc;
    }
    if (bcs) LAUNCH(K_BCS, nblk(nn), BLOCK, k_bcs, SB, h->bc_info.p, h->bc_or_not.p, h->bc_vars, h->bc_vals.p, h->ifsigman, h->normal.p, nn, nullptr, nullptr);
    if (nn > 0)
        LAUNCH_LDS(K_SPU_NODES, ntile8(nn), TILE, LAY_SPU_NODES_MIX.bytes, k_spu_nodes_mix, SB, V, PL, adapt, tanfi,
                   coh, hu, nn);
    if (nt > nn)
        LAUNCH_LDS(K_SP_UPDATE, ntile8(nt - nn), TILE, LAY_SP_UPDATE_MIX.bytes, k_sp_update_mix, SB, W, V, PL, h->dp,
                   time_sph, 0.0, 0, adapt, hu, nn, nt, stepv, nullptr);
    if (bcs) LAUNCH(K_BCS, nblk(nn), BLOCK, k_bcs, SB, h->bc_info.p, h->bc_or_not.p, h->bc_vars, h->bc_vals.p, h->ifsigman, h->normal.p, nn, nullptr, nullptr);
    const bool xsph = h->xsph && h->p.update_x;
    if (xsph && nn > 0)
        LAUNCH_LDS(K_XSPH, ntile8(nn), TILE, LAY_XSPH_MIX.bytes, (k_xsph_mix<KIND_NODE>), SB, V, PL, hu, h->xsph_v.p);
    if (xsph && nt > nn)
        LAUNCH_LDS(K_XSPH, ntile8(nt - nn), TILE, LAY_XSPH_MIX.bytes, (k_xsph_mix<KIND_SP>), SB, V, PL, hu, h->xsph_v.p);
    LAUNCH(K_POSITION, nblk(nt), BLOCK, k_position, SB, W, dt, h->p.update_x, nn, nt,
           xsph ? (const double2 *)h->xsph_v.p : (const double2 *)nullptr, h->err.p, 0);
    if (shift >= 0 && h->p.update_x && nn > 0) {
        const ShiftParams sp{h->p.npoints, h->p.vel_vector, shift, h->p.r_x, h->p.r_y, h->p.disp_tol, h->p.dx};
        LAUNCH(K_SHIFT, nblk(nn), BLOCK, k_shift_sp, SB, h->inv.p, h->nodeinfo.p, h->x10.p, h->disp10.p, sp, nn);
    }
    if (h->skin > 0) {
        LAUNCH(K_POSITION, 1, 64, k_zero_u64, h->dmax.p, 1);
        LAUNCH(K_POSITION, std::min(1024u, nblk(nt)), BLOCK, k_max_disp, SB.x, h->x_build.p, nt, h->dmax.p);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

static void drop_graphs(sps_handle *h) {
    for (int k = 0; k < 2; ++k)
        if (h->gexec[k]) {
            (void)hipGraphExecDestroy(h->gexec[k]);
            h->gexec[k] = nullptr;
        }
}

// The mixed form's step as a replayed graph (one per state set, captured on
// first use and whenever dt changes).
static int mixed_step_graph(sps_handle *h, double dt, double time_sph, double gxg, double gyg) {
    hipStream_t s = h->stream;
    const int c = h->cur;
    if (h->gexec[c] && h->graph_dt[c] != dt) {
        (void)hipGraphExecDestroy(h->gexec[c]);
        h->gexec[c] = nullptr;
    }
    if (!h->gexec[c]) {
        hipGraph_t g = nullptr;
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = mixed_step_kernels(h, dt, 0.0, 0.0, 0.0, h->stepv.p);
        hipError_t e = hipStreamEndCapture(s, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(SPS_EDEVICE, "step capture failed: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&h->gexec[c], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            h->gexec[c] = nullptr;
            return fail(SPS_EDEVICE, "step graph instantiation failed: %s", hipGetErrorString(e));
        }
        h->graph_dt[c] = dt;
    }
    k_set_step<<<1, 64, 0, s>>>(h->stepv.p, time_sph, gxg, gyg);
    if (h->p.no_bcs > 0 && h->nn > 0) k_bc_values<<<1, 64, 0, s>>>(bc_values(h, time_sph), h->bc_vals.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipGraphLaunch(h->gexec[c], s));
    ++h->graph_steps;
    return 0;
}

// get_nodes_on_free_surface (3_...:1116-1411) after the position update:
// the step's pair list laid out in traversal order by storage slot, each
// particle's pairs of types 1-3 gathered in that order, then the four steps
// of the detection; BC_or_not and the normals by original index.
static int free_surface(sps_handle *h, const StateBufs &SB) {
    int rc0 = sync_search(h);   // the pair count sizes the detection's arrays
    if (rc0) return rc0;
    const long long np = (long long)h->n_pairs;
    const int nn = h->nn, nt = h->nt, nt2 = h->nt2;
    int rc = 0;
    const size_t npa = (size_t)std::max(np, 1LL);
    rc |= dalloc(h->surf_pi, npa) | dalloc(h->surf_pj, npa) | dalloc(h->surf_pt, npa);
    rc |= dalloc(h->surf_gx, npa) | dalloc(h->surf_gy, npa) | dalloc(h->surf_ent, 2 * npa);
    rc |= dalloc(h->surf_cnt, (size_t)nt + 1) | dalloc(h->surf_start, (size_t)nt + 1) | dalloc(h->surf_cur, nt);
    rc |= dalloc(h->surf_flag, nt) | dalloc(h->surf_sub, nt) | dalloc(h->surf_nrm0, nt);
    rc |= dalloc(h->surf_tt, nt) | dalloc(h->surf_tau, nt);
    if (rc) return rc;
    if (np >= (1LL << 31)) return fail(SPS_ECAPACITY, "free surface: %lld pairs", np);
    hipStream_t s = h->stream;
    StateSet &S = h->S[h->cur];
    const SurfBufs B{h->surf_pi.p, h->surf_pj.p, h->surf_pt.p, h->surf_gx.p, h->surf_gy.p, h->surf_start.p,
                     h->surf_ent.p, h->surf_nrm0.p, h->surf_tt.p, h->surf_tau.p, h->surf_flag.p, h->surf_sub.p};
    LAUNCH(K_SURF, nblk(nt + 1), BLOCK, k_surf_zero, h->surf_cnt.p, (long long)nt + 1);
    LAUNCH(K_SURF, nblk(nt), BLOCK, k_surf_zero, h->surf_cur.p, (long long)nt);
    LAUNCH(K_SURF, nblk(nt), BLOCK, k_surf_zero, h->surf_flag.p, (long long)nt);
    if (np > 0) {
        LAUNCH(K_SURF, nblk(nt2), BLOCK, k_dump_pairs, S.x.p, h->x0.p, S.hsml.p, h->cell.p, h->kstart.p, S.orig.p,
               h->off.p, h->proc.p, h->grid.p, h->p.pi, h->surf_pi.p, h->surf_pj.p, h->surf_pt.p, (float *)nullptr,
               h->surf_gx.p, h->surf_gy.p, nn, nt, nt2, 1, (unsigned long long)np, h->m_prev);
        LAUNCH(K_SURF, nblk(np), BLOCK, k_surf_count, B, h->surf_cnt.p, np);
    }
    rc = scan_u32(h, h->surf_cnt.p, h->surf_start.p, (long long)nt + 1, nullptr);
    if (rc) return rc;
    if (np > 0) LAUNCH(K_SURF, nblk(np), BLOCK, k_surf_fill, B, h->surf_cur.p, h->surf_ent.p, np);
    LAUNCH(K_SURF, nblk(nt), BLOCK, k_surf_sort, h->surf_start.p, h->surf_ent.p, nt);
    LAUNCH(K_SURF, nblk(nt), BLOCK, k_surf_a, SB, B, nt);
    if (np > 0) LAUNCH(K_SURF, nblk(np), BLOCK, k_surf_hit, SB, B, np);
    if (h->xsph && np > 0)   // XSPH_update's marks (2_...:224-230) before the detection
        LAUNCH(K_SURF, nblk(np), BLOCK, k_surf_mark, SB, B, h->bc_or_not.p, np);
    LAUNCH(K_SURF, nblk(nt), BLOCK, k_surf_sub, SB, B, h->bc_or_not.p, nt);
    LAUNCH(K_SURF, nblk(nt), BLOCK, k_surf_b, SB, B, h->normal.p, nt);
    HIP_TRY(hipGetLastError());
    return 0;
}

// Dynamic LDS of the FP64 stage kernels: their layouts put the fields a
// launch may leave out last (the partners' x without CSPM, the dummies for the
// update alone, h when every particle has the same smoothing length), and
// a launch allocates the bytes before them.
static size_t lds_sp_update(const sps_handle *h, int deriv) {
    if (!deriv) return (size_t)LAY_SP_UPDATE.x[KIND_DUMMY];
    return (size_t)LAY_SP_UPDATE.x[KIND_NODE] + (h->dp.cspm ? 16 * (size_t)CAP : 0);
}
static size_t lds_av_nodes(const sps_handle *h) {
    return (size_t)LAY_AV_NODES.h[KIND_NODE] + (h->dp.h_uni ? 0 : 8 * (size_t)CAP);
}
static size_t lds_deriv_nodes(const sps_handle *h) {
    return (size_t)LAY_DERIV_NODES.x[KIND_SP] + (h->dp.cspm ? 16 * (size_t)CAP : 0);
}

static int one_step(sps_handle *h, int itimestep, double time_sph, double dt) {
    bool reuse;
    int rc = lists_still_valid(h, reuse);
    if (rc) return rc;
    hipStream_t s = h->stream;
    const int nn = h->nn, nt = h->nt, nt2 = h->nt2;
    if (!reuse) {
        rc = search(h, itimestep == 1);
        if (rc) return rc;
        ++h->searches;
        if (h->mixed && h->skin > 0) {
            HIP_TRY(hipMemcpyAsync(h->x_build.p, h->S[h->cur].x.p, (size_t)nt * sizeof(double2),
                                   hipMemcpyDeviceToDevice, s));
            h->lists_valid = true;
        }
    }
    const double factg = gravity_factor(&h->p, time_sph);
    const double gxg = h->p.ic_grav ? factg * h->p.cgrav[0] : 0.0;
    const double gyg = h->p.ic_grav ? factg * h->p.cgrav[1] : 0.0;
    if (h->mixed) {
        // the outside approach's per-step work depends on the step number
        const bool pre = h->outside && h->p.sph_shift && itimestep > 1 && (itimestep - 1) % h->p.shift_update == 0;
        const int shift = h->outside ? (itimestep % h->p.shift_update == 0 ? 1 : 0) : -1;
        // replayed as a graph when enabled, no kernel timing is asked for and
        // the step's launches do not depend on its number
        rc = (h->use_graph && !h->prof && !h->outside)
                 ? mixed_step_graph(h, dt, time_sph, gxg, gyg)
                 : mixed_step_kernels(h, dt, time_sph, gxg, gyg, nullptr, pre, shift);
        if (rc) return rc;
        if (h->skin > 0) {
            HIP_TRY(hipMemcpyAsync(h->pinned + 32, h->dmax.p, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipEventRecord(h->ev_disp, s));
        }
        return 0;
    }
    StateBufs SB = h->S[h->cur].view();
    Work W{h->vel0.p, h->x0.p, h->vel0.p, h->rk_vel.p, h->rhs2.p, h->f_bound.p, h->art_visc.p,
           h->stress0.p, h->rk_stress.p, h->rhs1.p, h->grad_u.p, h->rk_dev.p, h->ddev.p};
    const SearchView V{h->cell.p, h->kstart.p, h->proc.p, h->cnt.p, h->off.p, h->grid.p, h->split.p, h->p.pi};
    const PairLists PL = h->lists();
    const int adapt = h->dp.ncrit == 12;
    const double tanfi = h->dp.tanfi, coh = h->dp.coh;
    const double f1rk[4] = {0., 0.5, 0.5, 1.0}, f2rk[4] = {1., 2., 2., 1.0};
    const bool bcs = h->p.no_bcs > 0 && nn > 0;
    const BcValues &BV = h->bc_vars;
    if (bcs) LAUNCH(K_BCS, 1, 64, k_bc_values, bc_values(h, time_sph), h->bc_vals.p);
    if (h->outside && h->p.sph_shift && itimestep > 1 && (itimestep - 1) % h->p.shift_update == 0) {
        // stress_point_update + adapt_stress2 on the new lists (2_...:99-109)
        if (nn > 0)
            LAUNCH_LDS(K_SPU_NODES, ntile8(nn), TILE, LAY_SPU_NODES.bytes, k_spu_nodes, SB, V, PL, adapt, tanfi, coh, nn);
        if (nt > nn)
            LAUNCH_LDS(K_SP_UPDATE, ntile8(nt - nn), TILE, lds_sp_update(h, 0), k_sp_update, SB, W, V, PL, h->dp,
                       time_sph, 0.0, 0, adapt, nn, nt, nullptr);
    }
    const bool lf = h->integrator == SPS_INTEGRATOR_LEAPFROG;
    // bc_int for apply_stress_free: the step's list after its first derivative
    // pass, the previous step's before it
    const int *bci_cur = h->bcint ? h->nodeinfo.p : nullptr;
    const int *bci_prev = h->bcint ? h->bcint_prev.p : nullptr;
    LAUNCH(K_RK_INIT, nblk(nt2), BLOCK, k_stage_first, SB, W, adapt, tanfi, coh, nn, nt, nt2);
    if (lf) {
        // leapfrog: the node accelerations of the step's state kick the node
        // velocities; the stress points' rates from the kicked velocities
        // update their stresses (sps_set_integrator)
        const double kdt = h->lf_first ? 0.5 * dt : dt;
        if (bcs) LAUNCH(K_BCS, nblk(nn), BLOCK, k_bcs, SB, h->bc_info.p, h->bc_or_not.p, BV, h->bc_vals.p, h->ifsigman, h->normal.p, nn, nullptr, bci_prev);
        if (nn > 0)
            LAUNCH_LDS(K_SPU_NODES, ntile8(nn), TILE, LAY_SPU_NODES.bytes, k_spu_nodes, SB, V, PL, adapt, tanfi, coh, nn);
        if (bcs) LAUNCH(K_BCS, nblk(nn), BLOCK, k_bcs, SB, h->bc_info.p, h->bc_or_not.p, BV, h->bc_vals.p, h->ifsigman, h->normal.p, nn, nullptr, bci_prev);
        if (nt > nn)
            LAUNCH_LDS(K_SP_UPDATE, ntile8(nt - nn), TILE, lds_sp_update(h, 0), k_sp_update, SB, W, V, PL, h->dp, time_sph,
                       0.0, 0, adapt, nn, nt, nullptr);
        const bool av = h->dp.do_av && nn > 0;
        if (av) LAUNCH_LDS(K_AV_NODES, ntile8(nn), TILE, lds_av_nodes(h), k_av_nodes, SB, V, PL, h->dp, h->av_nodes.p, nn);
        if (nn > 0)
            LAUNCH_LDS(K_DERIV_NODES, ntile8(nn), TILE, lds_deriv_nodes(h), k_deriv_nodes, SB, W, V, PL, h->dp,
                       gxg, gyg, 1.0, h->err.p, nn, SB.stress, av ? h->av_nodes.p : nullptr);
        LAUNCH(K_STAGE_BEGIN, nblk(nn), BLOCK, k_lf_kick, SB, W, kdt, nn);
        if (bcs) LAUNCH(K_BCS, nblk(nn), BLOCK, k_bcs, SB, h->bc_info.p, h->bc_or_not.p, BV, h->bc_vals.p, h->ifsigman, h->normal.p, nn, bci_cur, nullptr);
        if (nn > 0)
            LAUNCH_LDS(K_SPU_NODES, ntile8(nn), TILE, LAY_SPU_NODES.bytes, k_spu_nodes, SB, V, PL, adapt, tanfi, coh, nn);
        if (bcs) LAUNCH(K_BCS, nblk(nn), BLOCK, k_bcs, SB, h->bc_info.p, h->bc_or_not.p, BV, h->bc_vals.p, h->ifsigman, h->normal.p, nn, bci_cur, nullptr);
        if (nt > nn)
            LAUNCH_LDS(K_SP_UPDATE, ntile8(nt - nn), TILE, lds_sp_update(h, 1), k_sp_update, SB, W, V, PL, h->dp, time_sph,
                       1.0, 1, adapt, nn, nt, nullptr);
        if (nt2 > nt) LAUNCH(K_DUMMY_STRESS, nblk(nt2 - nt), BLOCK, k_dummy_stress, SB, V, nt, nt2);
        LAUNCH(K_RK_FINAL, nblk(nt), BLOCK, k_lf_final, SB, W, dt, adapt, tanfi, coh, nn, nt, h->err.p);
        h->lf_first = false;
    }
    for (int st = 0; st < (lf ? 0 : 4); ++st) {
        const double c1 = f1rk[st] * (dt);
        if (st > 0) LAUNCH(K_STAGE_BEGIN, nblk(nt), BLOCK, k_stage_begin, SB, W, c1, adapt, tanfi, coh, nn, nt);
        if (bcs) LAUNCH(K_BCS, nblk(nn), BLOCK, k_bcs, SB, h->bc_info.p, h->bc_or_not.p, BV, h->bc_vals.p, h->ifsigman, h->normal.p, nn, st ? bci_cur : nullptr, st ? nullptr : bci_prev);
        // stress_point_update's two halves are independent: the node side
        // reads the stress points' stresses before their second return map
        // (k_sp_update leaves that in sig2) and changes node velocities only
        // where no stress point reads them (nodes without stress-point
        // partners); the second BC call sets the node velocities the first
        // set.  So they run side by side, unless kernel timing (one stream)
        // is on.
        // The artificial viscosity (k_av_nodes) reads only node data, after
        // the node side and the second BC call, and goes with them.
        const bool fork = !h->prof && nn > 0 && nt > nn;
        const bool av = h->dp.do_av && nn > 0;
        if (fork) {
            HIP_TRY(hipEventRecord(h->ev_fork, s));
            HIP_TRY(hipStreamWaitEvent(h->stream2, h->ev_fork, 0));
            hipStream_t s2 = h->stream2;
            k_spu_nodes<<<ntile8(nn), TILE, LAY_SPU_NODES.bytes, s2>>>(SB, V, PL, adapt, tanfi, coh, nn);
            if (bcs)
                k_bcs<<<nblk(nn), BLOCK, 0, s2>>>(SB, h->bc_info.p, h->bc_or_not.p, BV, h->bc_vals.p, h->ifsigman,
                                                 h->normal.p, nn, st ? bci_cur : nullptr,
                                                 st ? nullptr : bci_prev);
            if (av) k_av_nodes<<<ntile8(nn), TILE, lds_av_nodes(h), s2>>>(SB, V, PL, h->dp, h->av_nodes.p, nn);
            HIP_TRY(hipEventRecord(h->ev_join, s2));
        } else {
            if (nn > 0)
                LAUNCH_LDS(K_SPU_NODES, ntile8(nn), TILE, LAY_SPU_NODES.bytes, k_spu_nodes, SB, V, PL, adapt, tanfi, coh,
                           nn);
            if (bcs) LAUNCH(K_BCS, nblk(nn), BLOCK, k_bcs, SB, h->bc_info.p, h->bc_or_not.p, BV, h->bc_vals.p, h->ifsigman, h->normal.p, nn, st ? bci_cur : nullptr, st ? nullptr : bci_prev);
            if (av) LAUNCH_LDS(K_AV_NODES, ntile8(nn), TILE, lds_av_nodes(h), k_av_nodes, SB, V, PL, h->dp, h->av_nodes.p, nn);
        }
        if (nt > nn)
            LAUNCH_LDS(K_SP_UPDATE, ntile8(nt - nn), TILE, lds_sp_update(h, 1), k_sp_update, SB, W, V, PL, h->dp, time_sph,
                       f2rk[st], 1, adapt, nn, nt, h->sig2.p);
        if (fork) HIP_TRY(hipStreamWaitEvent(s, h->ev_join, 0));
        if (nn > 0)
            LAUNCH_LDS(K_DERIV_NODES, ntile8(nn), TILE, lds_deriv_nodes(h), k_deriv_nodes, SB, W, V, PL, h->dp, gxg,
                       gyg, f2rk[st], h->err.p, nn, nt > nn ? h->sig2.p : SB.stress, av ? h->av_nodes.p : nullptr);
        if (st == 3 && nt2 > nt)
            LAUNCH(K_DUMMY_STRESS, nblk(nt2 - nt), BLOCK, k_dummy_stress, SB, V, nt, nt2);
    }
    const bool hook = h->log_cap > 0;
    if (!lf)
        LAUNCH(K_RK_FINAL, nblk(nt), BLOCK, k_rk_final, SB, W, dt, adapt, tanfi, coh, nn, nt, surrogate_dev(h),
               hook ? 1 : 0, h->err.p);
    if (bcs) LAUNCH(K_BCS, nblk(nn), BLOCK, k_bcs, SB, h->bc_info.p, h->bc_or_not.p, BV, h->bc_vals.p, h->ifsigman, h->normal.p, nn, bci_cur, nullptr);
    if (nn > 0)
        LAUNCH_LDS(K_SPU_NODES, ntile8(nn), TILE, LAY_SPU_NODES.bytes, k_spu_nodes, SB, V, PL, adapt, tanfi, coh, nn);
    if (nt > nn)
        LAUNCH_LDS(K_SP_UPDATE, ntile8(nt - nn), TILE, lds_sp_update(h, 0), k_sp_update, SB, W, V, PL, h->dp, time_sph, 0.0,
                   0, adapt, nn, nt, nullptr);
    if (bcs) LAUNCH(K_BCS, nblk(nn), BLOCK, k_bcs, SB, h->bc_info.p, h->bc_or_not.p, BV, h->bc_vals.p, h->ifsigman, h->normal.p, nn, bci_cur, nullptr);
    if (h->bcint && nn > 0)   // the next step's first BC calls see this step's list
        LAUNCH(K_BCS, nblk(nn), BLOCK, k_bcint_keep, h->nodeinfo.p, SB.orig, h->bcint_prev.p, nn);
    const bool xsph = h->xsph && h->p.update_x;
    if (xsph && nn > 0)
        LAUNCH_LDS(K_XSPH, ntile8(nn), TILE, LAY_XSPH.bytes, (k_xsph<KIND_NODE>), SB, V, PL, h->xsph_v.p);
    if (xsph && nt > nn)
        LAUNCH_LDS(K_XSPH, ntile8(nt - nn), TILE, LAY_XSPH.bytes, (k_xsph<KIND_SP>), SB, V, PL, h->xsph_v.p);
    LAUNCH(K_POSITION, nblk(nt), BLOCK, k_position, SB, W, dt, h->p.update_x, nn, nt, xsph ? h->xsph_v.p : nullptr,
           h->err.p, lf ? 1 : 0);
    if (h->p.update_x) {   // 2_...:150-154
        // The detection's results (BC_or_not, the normals) change the next
        // step only through apply_stress_free (ifsigman = 1) and through
        // XSPH_update's marks on prescribed nodes (which it then drops), and
        // the outside approach's shift follows it; otherwise only the output
        // reads them (prescribed nodes keep BC_or_not = 1 through it), and
        // sps_get_surface runs the detection on the step's pairs and
        // positions when asked: its result depends on nothing else then.
        h->surface_pending = false;
        if (h->outside || (h->p.no_bcs > 0 && (h->ifsigman || h->xsph))) {
            rc = free_surface(h, SB);
            if (rc) return rc;
        } else {
            h->surface_pending = true;
        }
    }
    if (h->outside && h->p.update_x && nn > 0) {
        const ShiftParams sp{h->p.npoints, h->p.vel_vector, itimestep % h->p.shift_update == 0 ? 1 : 0,
                             h->p.r_x, h->p.r_y, h->p.disp_tol, h->p.dx};
        LAUNCH(K_SHIFT, nblk(nn), BLOCK, k_shift_sp, SB, h->inv.p, h->nodeinfo.p, h->x10.p, h->disp10.p, sp, nn);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int sps_set_bcs(sps_handle *h, const sps_bcs *b) {
    if (!h || !b || !b->bc_list || !b->bc_info || !b->bc_or_not) return fail(SPS_EINVAL, "null argument");
    if (b->no_bcs != h->p.no_bcs || b->no_bcs < 1)
        return fail(SPS_EINVAL, "no_bcs %d, sps_params says %d", b->no_bcs, h->p.no_bcs);
    if ((b->ifsigman != 0 && b->ifsigman != 1) || b->reserved0 != 0)
        return fail(SPS_EINVAL, "ifsigman must be 0 or 1, reserved0 0");
    if (b->ifsigman == 1 && h->mixed)
        return fail(SPS_EINVAL, "ifsigman = 1: SPS_MODE_ORDERED only (the mixed form has no free-surface detection)");
    for (int i = 0; i < h->nt; ++i)
        if (b->bc_or_not[i] != 0 && b->bc_or_not[i] != 1)
            return fail(SPS_EINVAL, "bc_or_not[%d] = %d (0 or 1)", i, b->bc_or_not[i]);
    if (b->ncurves < 0 || b->ncurves > SPS_MAX_CURVES || (b->ncurves > 0 && (!b->curve_npts || !b->curve_t || !b->curve_f)))
        return fail(SPS_EINVAL, "bad time curves");
    for (int c = 0; c < b->ncurves; ++c)
        if (b->curve_npts[c] < 2 || b->curve_npts[c] > SPS_MAX_TCURVE)
            return fail(SPS_EINVAL, "curve %d: bad point count", c + 1);
    for (int k = 0; k < b->no_bcs; ++k) {
        const double cv = b->bc_list[8 * k + 2];
        if (!(cv >= 0 && (int)cv <= b->ncurves))
            return fail(SPS_EINVAL, "BC %d: time curve %g of %d", k + 1, cv, b->ncurves);
    }
    const int nn = h->nn;
    for (int i = 0; i < nn; ++i)
        for (int k = 2; k < 7; ++k)
            if (b->bc_info[8 * i + k] < 0 || b->bc_info[8 * i + k] > b->no_bcs)
                return fail(SPS_EINVAL, "node %d: BC number %d of %d", i + 1, b->bc_info[8 * i + k], b->no_bcs);
    HIP_TRY(hipSetDevice(h->device));
    h->nbc = b->no_bcs;
    h->ncurves = b->ncurves;
    h->bcint = b->ifsigman == 1 && h->nt2 > h->nt;
    std::memcpy(h->bc_list, b->bc_list, 8 * (size_t)b->no_bcs * sizeof(double));
    for (int c = 0; c < b->ncurves; ++c) {
        h->curve_npts[c] = b->curve_npts[c];
        std::memcpy(h->curve_t + SPS_MAX_TCURVE * c, b->curve_t + SPS_MAX_TCURVE * c, SPS_MAX_TCURVE * sizeof(double));
        std::memcpy(h->curve_f + SPS_MAX_TCURVE * c, b->curve_f + SPS_MAX_TCURVE * c, SPS_MAX_TCURVE * sizeof(float));
    }
    HIP_TRY(hipMemcpyAsync(h->bc_info.p, b->bc_info, 8 * (size_t)nn * sizeof(int), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->bc_or_not.p, b->bc_or_not, (size_t)h->nt * sizeof(int), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemsetAsync(h->normal.p, 0, (size_t)h->nt * sizeof(double2), h->stream));
    h->ifsigman = b->ifsigman;
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->bc_vars = bc_values(h, 0.0);   // the variables (the values are the step's, k_bc_values)
    h->bcs_set = true;
    return 0;
}

extern "C" int sps_step(sps_handle *h, int32_t nsteps, double dt, double *time_sph, int32_t *itimestep_sph) {
    if (!h || !time_sph || !itimestep_sph) return fail(SPS_EINVAL, "null argument");
    if (!h->have_state) return fail(SPS_ESTATE, "sps_step before sps_set_state");
    if (h->p.no_bcs > 0 && !h->bcs_set) return fail(SPS_ESTATE, "no_bcs > 0: sps_set_bcs before sps_step");
    if (h->poisoned) return fail(SPS_ESTATE, "a previous step failed; set the state again");
    HIP_TRY(hipSetDevice(h->device));
    for (int32_t k = 0; k < nsteps; ++k) {
        *itimestep_sph += 1;
        int rc = one_step(h, *itimestep_sph, *time_sph, dt);
        if (rc) {
            h->poisoned = true;
            return rc;
        }
        *time_sph = *time_sph + dt;
        h->steps++;
    }
    HIP_TRY(hipMemcpyAsync(h->pinned, h->err.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    prof_collect(h);
    const int flag = *(int *)h->pinned;
    if (flag & ERR_GRID) {
        h->poisoned = true;
        (void)sync_search(h);
        return fail(SPS_ECAPACITY, "cell grid %d x %d exceeds the capacity %d sized from the domain box",
                    h->ndivx, h->ndivy, h->cell_cap);
    }
    if (flag & ERR_NONFINITE) {
        h->poisoned = true;
        return fail(SPS_ENUMERIC, "non-finite position, velocity or stress of a particle in the domain box "
                    "after step %d (checked at the end of each sps_step call)", *itimestep_sph);
    }
    if (flag) {
        h->poisoned = true;
        return fail(SPS_ECAPACITY, "device capacity flag %d: more than %d wall-force terms on one node (1) "
                    "or more than %d particles of one kind in one cell (2)", flag, MAX_WALL_TERMS,
                    MAX_BUCKET_SORT);
    }
    return 0;
}

template <typename T, typename K>
static int fetch(sps_handle *h, const T *dsrc, const int *orig, int n, T *dst, K kern) {
    DBuf<T> tmp;
    int rc = dalloc(tmp, (size_t)n);
    if (rc) return rc;
    kern<<<nblk(n), BLOCK, 0, h->stream>>>(dsrc, orig, tmp.p, n);
    hipError_t e = hipMemcpyAsync(dst, tmp.p, (size_t)n * sizeof(T), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    dfree(tmp);
    if (e != hipSuccess) return fail(SPS_EDEVICE, "state download: %s", hipGetErrorString(e));
    return 0;
}

extern "C" int sps_get_state(sps_handle *h, sps_state *st) {
    if (!h || !st) return fail(SPS_EINVAL, "null argument");
    if (!h->have_state) return fail(SPS_ESTATE, "sps_get_state before sps_set_state");
    HIP_TRY(hipSetDevice(h->device));
    StateSet &S = h->S[h->cur];
    const int *orig = S.orig.p;
    const int nn = h->nn, nt = h->nt, nt2 = h->nt2;
    int rc = 0;
    if (st->x) rc = rc ? rc : fetch(h, S.x.p, orig, nt2, (double2 *)st->x, k_gather_out2);
    if (st->vel) rc = rc ? rc : fetch(h, S.vel.p, orig, nt2, (double2 *)st->vel, k_gather_out2);
    if (st->stress) rc = rc ? rc : fetch(h, S.stress.p, orig, nt2, (double4 *)st->stress, k_gather_out4);
    if (st->mass) rc = rc ? rc : fetch(h, S.mass.p, orig, nt2, st->mass, k_gather_out1);
    if (st->rho) rc = rc ? rc : fetch(h, S.rho.p, orig, nt2, st->rho, k_gather_out1);
    if (st->hsml) rc = rc ? rc : fetch(h, S.hsml.p, orig, nt2, st->hsml, k_gather_out1);
    if (st->if_out_domain) rc = rc ? rc : fetch(h, S.ood.p, orig, nt2, st->if_out_domain, k_gather_outi);
    if (st->f_drucker) rc = rc ? rc : fetch(h, S.fdr.p, orig, nt, st->f_drucker, k_gather_out1);
    if (st->x00) rc = rc ? rc : fetch(h, S.x00.p, orig, nt2, (double2 *)st->x00, k_gather_out2);
    if (st->displ) rc = rc ? rc : fetch(h, S.displ.p, orig, nn, (double2 *)st->displ, k_gather_out2);
    if (st->grad_u) rc = rc ? rc : fetch(h, h->grad_u.p, orig, nt, (double4 *)st->grad_u, k_gather_out4);
    if (st->ddev_strn) rc = rc ? rc : fetch(h, h->ddev.p, orig, nt, st->ddev_strn, k_gather_out1);
    if (st->f_bound) rc = rc ? rc : fetch(h, h->f_bound.p, orig, nn, (double2 *)st->f_bound, k_gather_out2);
    if (st->art_visc) rc = rc ? rc : fetch(h, h->art_visc.p, orig, nn, (double2 *)st->art_visc, k_gather_out2);
    if (rc) return rc;
    if (h->outside && (st->x_10 || st->disp_10)) {
        if (st->x_10)
            HIP_TRY(hipMemcpyAsync(st->x_10, h->x10.p, nn * sizeof(double2), hipMemcpyDeviceToHost, h->stream));
        if (st->disp_10)
            HIP_TRY(hipMemcpyAsync(st->disp_10, h->disp10.p, nn * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    if (st->internal_vars) {
        std::vector<double> iv(nt);
        rc = fetch(h, S.ivar1.p, orig, nt, iv.data(), k_gather_out1);
        if (rc) return rc;
        std::memcpy(st->internal_vars, h->ivars_host.data(), h->ivars_host.size() * sizeof(double));
        for (int i = 0; i < nt; ++i) st->internal_vars[(size_t)h->niv * i] = iv[i];
    }
    if (st->itype)
        for (int i = 0; i < nt2; ++i) st->itype[i] = i < nn ? 2 : (i < nt ? 1 : 25);
    if (st->wall_position) rc = fetch(h, S.wallpos.p, orig, nt2, st->wall_position, k_gather_outf);
    if (!rc && st->horizontal_or_not) rc = fetch(h, S.horiz.p, orig, nt2, st->horizontal_or_not, k_gather_outf);
    return rc;
}

extern "C" int sps_get_pairs(sps_handle *h, int64_t cap, int32_t *pair_i, int32_t *pair_j, int32_t *pint_type,
                             float *w, float *dwdx, float *dwdy, int64_t *n) {
    if (!h || !n) return fail(SPS_EINVAL, "null argument");
    HIP_TRY(hipSetDevice(h->device));
    int rc0 = sync_search(h);
    if (rc0) return rc0;
    *n = (int64_t)h->n_pairs;
    if (h->steps == 0) {
        *n = 0;
        return 0;
    }
    if (cap < (int64_t)h->n_pairs)
        return fail(SPS_ECAPACITY, "pair buffer holds %lld < %llu", (long long)cap, h->n_pairs);
    HIP_TRY(hipSetDevice(h->device));
    size_t np = h->n_pairs;
    DBuf<int> di, dj, dt;
    DBuf<float> dw, dgx, dgy;
    int rc = 0;
    rc |= dalloc(di, np); rc |= dalloc(dj, np); rc |= dalloc(dt, np);
    rc |= dalloc(dw, np); rc |= dalloc(dgx, np); rc |= dalloc(dgy, np);
    if (rc) return rc;
    StateSet &S = h->S[h->cur];
    k_dump_pairs<<<nblk(h->nt2), BLOCK, 0, h->stream>>>(S.x.p, h->x0.p, S.hsml.p, h->cell.p, h->kstart.p, S.orig.p,
                                                     h->off.p, h->proc.p, h->grid.p, h->p.pi, di.p, dj.p, dt.p,
                                                     dw.p, dgx.p, dgy.p, h->nn, h->nt, h->nt2, 0, 0ull, 0ull);
    std::vector<int> hi(np), hj(np), ht(np);
    std::vector<float> hw(np), hgx(np), hgy(np);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(hi.data(), di.p, np * sizeof(int), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(hj.data(), dj.p, np * sizeof(int), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(ht.data(), dt.p, np * sizeof(int), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(hw.data(), dw.p, np * sizeof(float), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(hgx.data(), dgx.p, np * sizeof(float), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(hgy.data(), dgy.p, np * sizeof(float), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    dfree(di); dfree(dj); dfree(dt); dfree(dw); dfree(dgx); dfree(dgy);
    if (e != hipSuccess) return fail(SPS_EDEVICE, "pair download: %s", hipGetErrorString(e));
    // list order: [n-1 .. M_prev] then [0 .. M_prev-1] (discovery indices)
    const unsigned long long nn_ = h->n_pairs, mp = h->m_prev;
    size_t k = 0;
    auto put = [&](size_t d) {
        if (pair_i) pair_i[k] = hi[d];
        if (pair_j) pair_j[k] = hj[d];
        if (pint_type) pint_type[k] = ht[d];
        if (w) w[k] = hw[d];
        if (dwdx) dwdx[k] = hgx[d];
        if (dwdy) dwdy[k] = hgy[d];
        k++;
    };
    if (nn_ > mp) {
        for (unsigned long long d = nn_; d-- > mp;) put((size_t)d);
        for (unsigned long long d = 0; d < mp; ++d) put((size_t)d);
    } else {
        for (unsigned long long d = 0; d < nn_; ++d) put((size_t)d);
    }
    return 0;
}

extern "C" int sps_get_stats(sps_handle *h, sps_stats *out) {
    if (!h || !out) return fail(SPS_EINVAL, "null argument");
    HIP_TRY(hipSetDevice(h->device));
    int rc0 = sync_search(h);
    if (rc0) return rc0;
    std::memset(out, 0, sizeof *out);
    out->n_pairs = (int64_t)h->n_pairs;
    out->pairs_total = (int64_t)h->pairs_total_host;
    out->m_pairs = (int64_t)h->m_pairs;
    for (int k = 0; k < 10; ++k) out->n_by_type[k] = (int64_t)h->n_by_type[k];
    out->ndivx = h->ndivx;
    out->ndivy = h->ndivy;
    out->steps = h->steps;
    return 0;
}

extern "C" int sps_set_materials(sps_handle *h, int32_t nmat, const double *table, const int32_t *mat) {
    if (!h || nmat < 1 || !table || !mat) return fail(SPS_EINVAL, "bad argument");
    if (!h->have_state) return fail(SPS_ESTATE, "sps_set_materials before sps_set_state");
    if (h->dp.ntype_eco > 1 && h->dp.ncrit <= 4)
        return fail(SPS_EINVAL, "a material per particle is for the Drucker-Prager return map, not the Perzyna models");
    for (int i = 0; i < h->nt2; ++i)
        if (mat[i] < 0 || mat[i] >= nmat) return fail(SPS_EINVAL, "particle %d: material %d of %d", i + 1, mat[i], nmat);
    HIP_TRY(hipSetDevice(h->device));
    std::vector<Material> tab((size_t)nmat);
    for (int k = 0; k < nmat; ++k) {
        const double *t = table + 10 * k;
        tab[k] = Material{t[0], t[1], t[2], t[3], t[4], t[5], t[6], t[7], t[8], t[9]};
    }
    int rc = dalloc(h->mtab, (size_t)nmat) | dalloc(h->mat_orig, (size_t)h->nt2) | h->S[0].mat_alloc(h->nt2) |
             h->S[1].mat_alloc(h->nt2);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(h->mtab.p, tab.data(), (size_t)nmat * sizeof(Material), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h->mat_orig.p, mat, (size_t)h->nt2 * sizeof(int), hipMemcpyHostToDevice));
    // the slots' materials through the current order (the other state set gets
    // them from k_permute, which copies them with the rest of the state)
    h->S[0].mtab = h->S[1].mtab = h->mtab.p;
    drop_graphs(h);
    StateSet &C = h->S[h->cur];
    LAUNCH(K_PERMUTE, nblk(h->nt2), BLOCK, k_material_slots, C.orig.p, h->mat_orig.p, C.mat.p, h->nt2);
    HIP_TRY(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int sps_set_integrator(sps_handle *h, int32_t integrator) {
    if (!h) return fail(SPS_EINVAL, "null argument");
    if (integrator != SPS_INTEGRATOR_RK4 && integrator != SPS_INTEGRATOR_LEAPFROG)
        return fail(SPS_EINVAL, "unknown integrator %d", integrator);
    if (integrator == SPS_INTEGRATOR_LEAPFROG && (h->outside || h->xsph || h->mixed || h->log_cap > 0))
        return fail(SPS_EINVAL, "leapfrog: inside approach, no XSPH, SPS_MODE_ORDERED, no plastic log");
    h->integrator = integrator;
    h->lf_first = true;
    return 0;
}

extern "C" int sps_set_list_skin(sps_handle *h, double skin) {
    if (!h) return fail(SPS_EINVAL, "null argument");
    if (!(skin >= 0.0) || (skin > 0.0 && !h->mixed))
        return fail(SPS_EINVAL, "a list skin (%g) needs SPS_MODE_FAST and must be >= 0", skin);
    h->skin = skin;
    h->lists_valid = false;
    drop_graphs(h);
    return 0;
}

extern "C" int sps_set_graph(sps_handle *h, int32_t enable) {
    if (!h) return fail(SPS_EINVAL, "null argument");
    if (enable && !h->mixed) return fail(SPS_EINVAL, "graph replay of a step needs SPS_MODE_FAST");
    h->use_graph = enable != 0;
    drop_graphs(h);
    return 0;
}

extern "C" int sps_get_search_count(sps_handle *h, int64_t *n) {
    if (!h || !n) return fail(SPS_EINVAL, "null argument");
    *n = h->searches;
    return 0;
}

extern "C" int sps_set_profiling(sps_handle *h, int32_t enable) {
    if (!h) return fail(SPS_EINVAL, "null argument");
    h->prof = enable != 0;
    return 0;
}

extern "C" int sps_get_profile(sps_handle *h, int32_t cap, sps_kernel_time *out, int32_t *n, int32_t reset) {
    if (!h || !n) return fail(SPS_EINVAL, "null argument");
    HIP_TRY(hipStreamSynchronize(h->stream));
    prof_collect(h);
    int k = 0;
    for (int id = 0; id < K_NKERNELS; ++id) {
        if (h->prof_n[id] == 0) continue;
        if (out && k < cap) {
            std::memset(out[k].name, 0, sizeof out[k].name);
            std::strncpy(out[k].name, kKernelNames[id], sizeof out[k].name - 1);
            out[k].launches = h->prof_n[id];
            out[k].total_ms = h->prof_ms[id];
        }
        k++;
    }
    *n = k;
    if (reset)
        for (int id = 0; id < K_NKERNELS; ++id) {
            h->prof_ms[id] = 0.0;
            h->prof_n[id] = 0;
        }
    return 0;
}

extern "C" void sps_destroy(sps_handle *h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    prof_collect(h);
    for (auto e : h->ev_pool) (void)hipEventDestroy(e);
    if (h->ev_disp) (void)hipEventDestroy(h->ev_disp);
    drop_graphs(h);
    dfree(h->stepv);
    h->S[0].release();
    h->S[1].release();
    dfree(h->x0); dfree(h->vel0); dfree(h->rk_vel); dfree(h->rhs2); dfree(h->f_bound);
    dfree(h->art_visc); dfree(h->stress0); dfree(h->rk_stress); dfree(h->rhs1); dfree(h->grad_u);
    dfree(h->rk_dev); dfree(h->ddev); dfree(h->grid); dfree(h->split); dfree(h->bparts);
    dfree(h->cellkey); dfree(h->new_cell); dfree(h->cell); dfree(h->kcount); dfree(h->kstart);
    dfree(h->kfill); dfree(h->tmp_orig); dfree(h->tmp_src); dfree(h->proc); dfree(h->proc_to_slot); dfree(h->moved);
    dfree(h->cnt); dfree(h->off); dfree(h->scan_tmp); dfree(h->ctot); dfree(h->cstart_rm); dfree(h->err);
    dfree(h->list_ia); dfree(h->list_wa); dfree(h->list_gxa); dfree(h->list_gya); dfree(h->list_ib);
    dfree(h->list_gxb); dfree(h->list_gyb); dfree(h->deg_a); dfree(h->deg_b); dfree(h->tile_ok); dfree(h->tiles); dfree(h->plans);
    dfree(h->list_wb); dfree(h->xsph_v);
    dfree(h->x_build); dfree(h->dmax); dfree(h->surr_count); dfree(h->surr_slot); dfree(h->surr_sn);
    dfree(h->surr_str); dfree(h->surr_feat); dfree(h->surr_stats); dfree(h->plog); dfree(h->surr_mat);
    dfree(h->mtab); dfree(h->mat_orig); dfree(h->pairs_total);
    dfree(h->inv); dfree(h->nodeinfo); dfree(h->x10); dfree(h->disp10); dfree(h->bcint_prev);
    dfree(h->bc_info); dfree(h->bc_or_not); dfree(h->bc_vals); dfree(h->normal); dfree(h->surf_nrm0);
    dfree(h->surf_pi); dfree(h->surf_pj); dfree(h->surf_pt); dfree(h->surf_sub); dfree(h->surf_gx);
    dfree(h->surf_gy); dfree(h->surf_cnt); dfree(h->surf_start); dfree(h->surf_cur); dfree(h->surf_ent);
    dfree(h->surf_flag); dfree(h->surf_tt); dfree(h->surf_tau);
    if (h->pinned) (void)hipHostFree(h->pinned);
    if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
    if (h->ev_join) (void)hipEventDestroy(h->ev_join);
    if (h->stream2) (void)hipStreamDestroy(h->stream2);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

extern "C" int sps_get_surface(sps_handle *h, int32_t *bc_or_not, double *normal) {
    if (!h) return fail(SPS_EINVAL, "null argument");
    if (h->mixed) return fail(SPS_EINVAL, "SPS_MODE_FAST does not track the free surface");
    if (h->p.no_bcs > 0 && !h->bcs_set) return fail(SPS_ESTATE, "no_bcs > 0: sps_set_bcs before sps_get_surface");
    if (!h->bc_or_not.p)
        return fail(SPS_ESTATE, "the free surface is tracked with BCs or the position update only");
    HIP_TRY(hipSetDevice(h->device));
    if (h->surface_pending) {   // the detection of the last step, on its pairs and positions
        int rc = free_surface(h, h->S[h->cur].view());
        if (rc) return rc;
        h->surface_pending = false;
    }
    if (bc_or_not)
        HIP_TRY(hipMemcpyAsync(bc_or_not, h->bc_or_not.p, (size_t)h->nt * sizeof(int), hipMemcpyDeviceToHost,
                               h->stream));
    if (normal)
        HIP_TRY(hipMemcpyAsync(normal, h->normal.p, (size_t)h->nt * sizeof(double2), hipMemcpyDeviceToHost,
                               h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return 0;
}
