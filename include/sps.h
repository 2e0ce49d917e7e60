/*
 * sps.h - C-ABI of the Stress-Particle SPH hot path on one MI355X (gfx950).
 *
 * The reference has no operator/FFI interface: its hot path is the
 * argument-less `call time_integration` (code/1_SPH_2018.f90:82, public at
 * code/2_SPH_main_2018.f90:22) working on module globals
 * (code/5_SPH_global_vars_2018.f90:27-182).  This header is the seam that
 * replaces it: the Fortran host marshals the same globals through
 * iso_c_binding (sps_binding.f90) and calls sps_step once per time step.
 *
 * Conventions
 *  - Plain C types only.  real(DP) <-> double, default real <-> float,
 *    integer <-> int32_t (code/7_variable_types.f90:7).
 *  - Arrays use the reference's Fortran column-major layout: x(2,ntotal2) is
 *    [x1,y1,x2,y2,...], stress(4,ntotal2) is [sxx,syy,sxy,szz] per particle,
 *    grad_u(2,2,ntotal) is [g11,g21,g12,g22] per particle with
 *    g_ab = d v_a / d x_b, internal_vars(nint_vars,ntotal).
 *  - Particle numbering follows the reference (code/3_SPH_material_2018.f90
 *    :942-1007): velocity nodes 1..nnode, stress points nnode+1..ntotal,
 *    wall dummies ntotal+1..ntotal2, the first ndummy2 dummies being the
 *    innermost wall layer (code/3_SPH_material_2018.f90:793-812).  Pair
 *    indices returned by sps_get_pairs are 1-based like Fortran.
 *  - Every call is synchronous on return.  One host thread per handle.
 *  - Return value: SPS_OK (0) or a negative SPS_E* code; the message is
 *    available from sps_last_error() (per thread).
 *
 * Two shared libraries implement this header: the product
 * (libsps_hip.so, hand-written HIP kernels) and the CPU oracle used only as
 * the checker in tests (oracle/libsps_oracle.so).
 */
#ifndef SPS_H
#define SPS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPS_ABI_VERSION 4

enum {
    SPS_OK = 0,
    SPS_EINVAL = -1,      /* bad argument or unsupported option combination */
    SPS_ENOMEM = -2,      /* host or device allocation failed */
    SPS_EDEVICE = -3,     /* HIP runtime error */
    SPS_ECAPACITY = -4,   /* a grid or pair capacity was exceeded */
    SPS_ESTATE = -5,      /* call out of order (e.g. step before set_state) */
    SPS_ENUMERIC = -6     /* a particle in the domain box had a non-finite
                             position, velocity or stress at the end of a
                             step (checked once per sps_step call; the HIP
                             library then needs sps_set_state again) */
};

/* Arithmetic of the pair loops.  SPS_MODE_FAST has no counterpart in the
 * reference: single-precision pair arithmetic (kernel values computed in
 * single precision from window-relative positions) with compensated (Kahan)
 * sums, in the pair lists' forward order; positions, velocities, the stress
 * state, the CSPM solve, the Drucker-Prager terms and the return map stay in
 * double precision.  Its results differ from the reference's by rounding
 * (tolerance-checked against the oracle); the oracle runs every mode as
 * SPS_MODE_ORDERED. */
enum {
    SPS_MODE_ORDERED = 0, /* reference list order, bitwise parity target */
    SPS_MODE_FAST = 1     /* mixed precision, as above */
};

#define SPS_MAX_TCURVE 64

/*
 * Problem constants.  Field names follow the reference globals; each comment
 * names where the reference reads or derives the value.
 */
typedef struct sps_params {
    int32_t abi_version;     /* must be SPS_ABI_VERSION */
    int32_t nnode;           /* nnode   (3_...:421) */
    int32_t nstress;         /* nstress (3_...:424-440) */
    int32_t ndummy;          /* ndummy  (3_...:785), 0 without dummy walls */
    int32_t ndummy2;         /* innermost wall layer (3_...:776-782) */
    int32_t nint_vars;       /* leading dim of internal_vars (3_...:275), >= 1 */
    int32_t ndimn;           /* must be 2 */
    int32_t nstre;           /* must be 4 (plane strain) */
    int32_t skf;             /* kernel: must be 1, cubic spline (2_...:1468) */
    /* control logicals (0/1), read at 3_...:286-299, 364-375 */
    int32_t sp_sph;          /* must be 1 */
    int32_t inside_approach; /* 1: inside approach; 0: outside approach, the
                                stress points attached to their node and
                                shifted with it (2_...:156-160, 244-370) */
    int32_t cspm;            /* CSPM gradient correction (2_...:539) */
    int32_t update_x;        /* position update + Jaumann spin (2_...:140, 751) */
    int32_t xsph;            /* XSPH position update (2_...:142-143, 189-239) */
    int32_t cont_density;    /* must be 0 */
    int32_t art_stress;      /* must be 0 */
    int32_t dummy_nodes;     /* wall dummies present (.pts) */
    int32_t sph_shift;       /* SPH_shift (3_...:373), outside approach only:
                                stress_point_update + adapt_stress2 after the
                                search on shift steps (2_...:99-109) */
    int32_t no_bcs;          /* prescribed BCs (3_...:165): 0, or > 0 with
                                sps_set_bcs (not with xsph and update_x) */
    int32_t bforce_gate;     /* 0: wall force when inside_approach .and.
                                dummy_nodes (code/2_...:742); 1: when
                                inside_approach (Bui fork 2_...:742) */
    int32_t mode;            /* SPS_MODE_* */
    int32_t ic_grav;         /* 0/1 (3_...:302-311) */
    int32_t tcurve_npts;     /* points in the gravity time curve, 2..64 */
    double props[20];        /* props(1,1:20) (3_...:116-139) */
    double pi;               /* 4*atan(1.0) in single precision (3_...:144) */
    double dx;               /* lattice spacing (.pts, 3_...:393) */
    double alpha, beta;      /* artificial viscosity (3_...:296) */
    double damping;          /* DampingTG (3_...:293) */
    double cspm_det_tol;     /* 1e-03 (code/2_...:599) or 1e-07 (Bui fork),
                                as the single-precision literal widened */
    double D11, D22, D12, D33, D41, D42;   /* elastic matrix (3_...:719-730) */
    double xmin_domain[2], xmax_domain[2]; /* domain box (3_...:270) */
    double cgrav[2];         /* gravity vector (3_...:308) */
    double ft_grav;          /* gravity factor (3_...:308) */
    double tcurve_t[SPS_MAX_TCURVE];  /* ttcurves(tcurve_grav,:), real(DP) */
    float  tcurve_f[SPS_MAX_TCURVE];  /* ftcurves(tcurve_grav,:), default real */
    /* the outside approach (read at 3_...:370-378, derived at 3_...:609);
       ignored with the inside approach */
    int32_t npoints;         /* stress points per node, 1..4 (2 with vel_vector):
                                stress point nnode + npoints*(i-1) + p, p = 1..npoints,
                                belongs to node i (3_...:611-660) */
    int32_t vel_vector;      /* orient the two points along the node velocity
                                (2_...:261-289) */
    int32_t shift_update;    /* steps between shifts, >= 1 (3_...:373) */
    int32_t reserved0;       /* must be 0 */
    double r_x, r_y;         /* dx*rx_factor, dx*ry_factor (3_...:609) */
    double disp_tol;         /* displacement tolerance in dx (vel_vector) */
} sps_params;

/*
 * Pointers to host arrays (Fortran layout).  For sps_set_state the pointers
 * marked "in" are required; for sps_get_state any NULL pointer is skipped.
 */
typedef struct sps_state {
    double  *x;                 /* (2,ntotal2)  in/out */
    double  *vel;               /* (2,ntotal2)  in/out */
    double  *stress;            /* (4,ntotal2)  in/out */
    double  *mass;              /* (ntotal2)    in/out */
    double  *rho;               /* (ntotal2)    in/out */
    double  *hsml;              /* (ntotal2)    in/out */
    int32_t *itype;             /* (ntotal2)    in/out: 2 node, 1 SP, 25 dummy */
    int32_t *if_out_domain;     /* (ntotal2)    in/out, sticky 0/1 */
    double  *internal_vars;     /* (nint_vars,ntotal) in/out; only row 1 is
                                   touched by the hot path */
    double  *f_drucker;         /* (ntotal)     in/out, DP yield-function memory */
    double  *x00;               /* (2,ntotal2)  in: initial positions (displ) */
    double  *displ;             /* (2,nnode)    in/out */
    float   *wall_position;     /* (ntotal2)    in, dummies only (may be NULL
                                   without dummies) */
    float   *horizontal_or_not; /* (ntotal2)    in, dummies only */
    double  *grad_u;            /* (2,2,ntotal) out: last stage velocity gradient */
    double  *ddev_strn;         /* (ntotal)     out: Ddev_strn of the last step */
    double  *f_bound;           /* (2,nnode)    out: last stage wall force */
    double  *art_visc;          /* (2,nnode)    out: last stage art. viscosity */
    double  *x_10;              /* (2,nnode)    in/out, outside approach with
                                   vel_vector: node positions at the last shift
                                   (NULL at sps_set_state: the nodes' x, as
                                   3_...:1032 sets them) */
    double  *disp_10;           /* (nnode)      in/out, the same: displacement
                                   over the last shift interval in dx (NULL at
                                   sps_set_state: 0).  With the inside approach
                                   both are ignored and left untouched. */
} sps_state;

typedef struct sps_stats {
    int64_t n_pairs;        /* pairs found by the last search (2_...:1354) */
    int64_t m_pairs;        /* list high-water mark after it (2_...:1369) */
    int64_t n_by_type[10];  /* pint_type histogram 0..9 (3_...:1598-1608) */
    int32_t ndivx, ndivy;   /* cell grid of the last search (2_...:1250) */
    int64_t steps;          /* steps taken by this handle */
    int64_t pairs_total;    /* pairs found by all searches of this handle (ABI 4;
                               no counterpart in the reference: for throughput
                               over many steps in one sps_step call) */
} sps_stats;

typedef struct sps_handle sps_handle;

/* Create a handle for one problem.  Replaces the allocation part of
 * Setup_Global_Arrays (3_...:883-1026) for the hot-path state.  The product
 * library selects the current HIP device. */
int sps_create(const sps_params *params, sps_handle **out);

/* Copy the state in (host -> library).  Replaces the implicit sharing of the
 * module globals listed in sps_state.  The pair-list high-water mark is not
 * touched: like the reference (2_...:1290-1294) it restarts at zero on the
 * step whose itimestep_sph is 1. */
int sps_set_state(sps_handle *h, const sps_state *st);

/* Advance nsteps steps of dt.  One step is one `call time_integration`
 * (2_...:78-184) followed by the driver's time update
 * (1_SPH_2018.f90:80-82): itimestep_sph += 1 before, time_sph += dt after. */
int sps_step(sps_handle *h, int32_t nsteps, double dt,
             double *time_sph, int32_t *itimestep_sph);

/* Copy the state out (library -> host). */
int sps_get_state(sps_handle *h, sps_state *st);

/* Pairs of the last search after Pint_Update (3_...:1555-1615), in the order
 * a traversal of the reference's linked list from `last` visits them.
 * Indices are 1-based.  *n receives the pair count; at most cap entries are
 * written (SPS_ECAPACITY if cap < *n).  Any output pointer may be NULL. */
int sps_get_pairs(sps_handle *h, int64_t cap, int32_t *pair_i, int32_t *pair_j,
                  int32_t *pint_type, float *w, float *dwdx, float *dwdy,
                  int64_t *n);

int sps_get_stats(sps_handle *h, sps_stats *out);

/* A material per particle.  No counterpart in the reference, which has one
 * material (props(1,:)); for an ensemble of instances concatenated into one
 * particle set.  table: nmat rows of 10 doubles: tan(phi), c, E, nu, D11, D22,
 * D12, D33, D41, D42 (the elastic matrix of 3_...:719-730); mat: each
 * particle's row (ntotal2, 0-based, in sps_set_state's order).  Call after
 * sps_set_state; it then replaces props(1,3:4), props(1,13:14) and D11..D42
 * of sps_params for the pair loops, the Drucker-Prager terms and the return
 * map, particle by particle. */
int sps_set_materials(sps_handle *h, int32_t nmat, const double *table, const int32_t *mat);
/* (SPS_EINVAL with the Perzyna models: they read props(1,:) only.) */

/* Prescribed velocity / stress values (Normal_BCs, 3_...:1644-1753) at the
 * points where the reference calls BCs: after each RK stage's update and
 * after its stress_point_update (2_...:715-723), after the final RK update
 * (2_...:788) and after the step's stress_point_update (2_...:134).  Replaces
 * the globals bc_list, bc_info, BC_or_not and the time curves
 * (3_...:160-260, 1030-1113).  Required before sps_step when
 * sps_params.no_bcs > 0 (with XSPH and update_x together in
 * SPS_MODE_ORDERED only: XSPH_update marks the nodes next to the free
 * surface 3, 2_...:224-230, and the detection then resets even a prescribed
 * one to 0).  With update_x, every step runs the free-surface detection
 * (get_nodes_on_free_surface, 3_...:1116-1411) after the position update:
 * BC_or_not becomes 2 on the surface and 0 inside (1 stays 1), with the
 * surface normals; ifsigman = 1 then applies apply_stress_free
 * (3_...:1758-1841) after Normal_BCs at every BC call: each surface node's
 * stress keeps only its tangential part, sigma_tt in the normal frame,
 * except at nodes with a wall-dummy pair (bc_int = 1) in the list of the
 * last derivative pass: the previous step's list for the first two BC
 * calls of a step, the step's own for the rest.  The
 * mixed form (SPS_MODE_FAST) does not run the detection and rejects
 * ifsigman = 1; in it BC_or_not stays as set.  The BC time is
 * time_sph: the reference adds ic_time*dt_sph, ic_time being a local it never
 * sets (3_...:1622-1627).  Values are applied to the nodes 1..nnode with
 * BC_or_not = 1, as code/ does (the vertical-slope fork loops over ntotal;
 * its stress points never lie on a BC segment with the inside approach). */
#define SPS_MAX_BCS 32
#define SPS_MAX_CURVES 16
typedef struct sps_bcs {
    int32_t no_bcs;            /* rows of bc_list, 1..SPS_MAX_BCS, = sps_params.no_bcs */
    int32_t ifsigman;          /* 0, or 1: apply_stress_free on the free-surface nodes */
    int32_t ncurves;           /* time curves, 0..SPS_MAX_CURVES */
    int32_t reserved0;         /* must be 0 */
    const double *bc_list;     /* (8,no_bcs), 3_...:176-180: number; variable (1 sxx,
                                  2 syy, 3 sxy, 5 vx, 6 vy; others act on nothing);
                                  curve (0: (a0 + a1 sin(w t - phi)) (1 - exp(-t/Tf)),
                                  k > 0: a1 times time curve k); a1; a0; w; phi
                                  (already times twopi); Tf */
    const int32_t *bc_info;    /* (8,nnode): Get_BCs_on_node's rows of the nodes */
    const int32_t *bc_or_not;  /* (ntotal) Get_BCs_on_node's BC_or_not of every
                                  particle, nodes then stress points: 0 or 1
                                  (anything else is refused); the call also
                                  clears the free-surface normals, and a
                                  refused call changes nothing */
    const int32_t *curve_npts; /* (ncurves) nptstcurves, 2..SPS_MAX_TCURVE */
    const double *curve_t;     /* (SPS_MAX_TCURVE,ncurves) ttcurves */
    const float *curve_f;      /* (SPS_MAX_TCURVE,ncurves) ftcurves */
} sps_bcs;
int sps_set_bcs(sps_handle *h, const sps_bcs *bcs);

/* The free surface as the last step's detection left it (SPS_MODE_ORDERED,
 * after sps_set_bcs): bc_or_not (ntotal) the reference's BC_or_not (0
 * inside, 1 prescribed, 2 on the free surface), normal (2,ntotal) its
 * global normal: the unit outward normals of the surface particles, 0
 * elsewhere, NaN where the reference divides zero by zero.  Either pointer
 * may be null.  Replaces reading BC_or_not and normal after
 * get_nodes_on_free_surface (3_...:1116-1411; the reference's
 * surface_points output, 3_...:2910-2919, lists the nodes with 2). */
int sps_get_surface(sps_handle *h, int32_t *bc_or_not, double *normal);

/* The time integrator.  SPS_INTEGRATOR_RK4 (the default) is the
 * reference's classical RK4 (2_...:653-802).  SPS_INTEGRATOR_LEAPFROG has no
 * counterpart in the reference: kick-drift leapfrog with the node velocities
 * at half steps and the stresses and positions at whole steps -- the node
 * accelerations of the step's state kick the velocities by dt (by dt / 2 on
 * the first step after sps_set_integrator or sps_set_state, whose
 * velocities are at a whole step), the stress points take the kicked
 * velocities and their stress rate (elastic, plastic source, Jaumann spin)
 * updates the stress by dt, and the positions drift by dt with the kicked
 * velocities.  Two pair-loop passes per step instead of RK4's four stages;
 * second order.  SPS_MODE_ORDERED, inside approach, without XSPH or the
 * plastic log; sps_get_state then returns the half-step velocities. */
enum {
    SPS_INTEGRATOR_RK4 = 0,
    SPS_INTEGRATOR_LEAPFROG = 1
};
int sps_set_integrator(sps_handle *h, int32_t integrator);

/* Pair lists kept over steps.  No counterpart in the reference, which
 * searches every step; SPS_MODE_FAST only.  With skin > 0 the search also
 * lists the partners within 2h + skin (cells 2 max h + skin wide) and later
 * steps reuse those lists while no moving particle has moved skin / 2 since
 * the search (pairs beyond 2h contribute zero); the out-of-domain check runs
 * with each search.  0 (the default) searches every step.
 * sps_get_search_count returns the searches made since sps_create. */
int sps_set_list_skin(sps_handle *h, double skin);

/* Replay of the mixed form's step (SPS_MODE_FAST; no counterpart in the
 * reference): with enable != 0, the launches of a step that reuses the pair
 * lists are captured once into a HIP graph (one per state buffer, again when
 * dt changes) and replayed; steps that search, steps while kernel timing is
 * on and every step of the outside approach (whose shifts depend on the step
 * number) launch directly. */
int sps_set_graph(sps_handle *h, int32_t enable);
int sps_get_search_count(sps_handle *h, int64_t *n);

/* Learned Drucker-Prager stress update.  No counterpart in the reference.
 * A multilayer perceptron 10 -> 128 -> 128 -> 128 -> 4 (SiLU; bfloat16
 * weights and activations, single-precision accumulation on the matrix
 * cores) maps features of a plastic trial state to the correction the exact
 * return map (adapt_stress2, 3_...:2066-2140) makes:
 *   x[0:4] = sigma_n / S, x[4:8] = (sigma_trial - sigma_n) / S,
 *   x[8] = c / Sc, x[9] = tan(phi);
 *   y[0:4] = (R(sigma_trial) - sigma_trial) / F(sigma_trial)
 * with F = sqrt(J2) + 3 alpha p - k_c the yield function (positive at every
 * plastic trial state)
 * (stress components xx, yy, xy, zz; S, Sc the stress and cohesion scales).
 * Training (forward, backward, Adam) and inference run on the GPU.
 * sps_mlp_create: Glorot-uniform weights from `seed`; batch_rows sizes the
 *   work buffers (grown as needed).
 * sps_mlp_get_params / set_params: the sps_mlp_param_count() single-precision
 *   parameters, blocks W1 [16][128], b1 [128], W2 [128][128], b2, W3, b3,
 *   W4 [128][16], b4 [16] (inputs past 10 and outputs past 4 are padding).
 * sps_mlp_forward: y (n x 4) for x (n x 10).
 * sps_mlp_gradient: mean squared error over n <= batch_rows rows and its
 *   gradient (all parameters), without an update.
 * sps_mlp_train: `epochs` shuffled passes of Adam (lr; beta 0.9, 0.999) in
 *   batches of `batch` rows (a multiple of 64); the mean loss of each epoch.
 * sps_return_map: the exact return map of n stresses (n x 4) on the GPU, each
 *   with its own tan(phi) and c (n each).
 * sps_set_surrogate (SPS_MODE_FAST; NULL detaches): each step's final stress
 *   update sends the stress points' plastic trial states through the network
 *   and keeps a result where it is consistent: |F| <= tol * (sqrt(J2) +
 *   |3 alpha p| + k_c) of the trial state (F the yield function), its mean
 *   stress that of the trial state (the apex beyond the tension cutoff) and
 *   its deviator parallel to the trial one, within the same tolerance;
 *   elsewhere the exact return map's result stays.  sps_get_surrogate_stats:
 *   states evaluated and kept.
 * sps_set_plastic_log / sps_get_plastic_log: the plain method's tuples
 *   (sigma_n, sigma_trial, R(sigma_trial), tan(phi), c), SPS_PLOG_ROW doubles
 *   each, of every plastic trial state of the final stress updates, up to cap
 *   (training data; the material is the point's own with sps_set_materials). */
#define SPS_PLOG_ROW 14
typedef struct sps_mlp sps_mlp;
int64_t sps_mlp_param_count(void);
int sps_mlp_create(int32_t batch_rows, uint64_t seed, sps_mlp **out);
void sps_mlp_destroy(sps_mlp *ml);
int sps_mlp_get_params(sps_mlp *ml, float *out);
int sps_mlp_set_params(sps_mlp *ml, const float *in);
int sps_mlp_forward(sps_mlp *ml, int64_t n, const float *x, float *y);
int sps_mlp_gradient(sps_mlp *ml, int64_t n, const float *x, const float *t, float *grad, float *loss);
int sps_mlp_train(sps_mlp *ml, int64_t n, const float *x, const float *t, int32_t epochs, int32_t batch,
                  float lr, uint64_t seed, float *epoch_loss);
int sps_return_map(int64_t n, const double *in, const double *tanfi, const double *coh, double *out);
/* sps_return_map without any allocation.  No counterpart in the reference;
 * for callers that cannot allocate device or host memory per call (serving
 * loops and harnesses that forbid hipMalloc).  Same arguments and results
 * as sps_return_map, bit for bit, for n <= SPS_RETURN_MAP_NOALLOC_MAX
 * (larger n: SPS_EINVAL).  The device staging of the inputs and outputs is
 * one fixed array per GPU (80 bytes per stress, 20 MiB), which each GPU
 * holds until the process ends; every call on a GPU shares that GPU's
 * array, so calls that overlap on one GPU race.  Synchronous on return, on
 * the current HIP device. */
#define SPS_RETURN_MAP_NOALLOC_MAX 262144
int sps_return_map_noalloc(int64_t n, const double *in, const double *tanfi, const double *coh, double *out);
int sps_set_surrogate(sps_handle *h, sps_mlp *ml, double stress_scale, double coh_scale, double tol);
int sps_get_surrogate_stats(sps_handle *h, int64_t *evaluated, int64_t *kept);
int sps_set_plastic_log(sps_handle *h, int64_t cap);
int sps_get_plastic_log(sps_handle *h, int64_t cap, double *out, int64_t *n);

/* Kernel timing.  No counterpart in the reference (which has no timers);
 * for performance reports.  While enabled, every kernel launch of sps_step is
 * bracketed by HIP events on the handle's stream and the elapsed times are
 * accumulated per kernel; sps_get_profile returns them (and clears them when
 * reset != 0).  The CPU oracle reports no kernels. */
typedef struct sps_kernel_time {
    char name[32];
    int64_t launches;
    double total_ms;
} sps_kernel_time;

int sps_set_profiling(sps_handle *h, int32_t enable);
int sps_get_profile(sps_handle *h, int32_t cap, sps_kernel_time *out, int32_t *n,
                    int32_t reset);

/* Implementation name, e.g. "hip-gfx950" or "oracle-cpu". */
const char *sps_backend(void);

const char *sps_last_error(void);

void sps_destroy(sps_handle *h);

#ifdef __cplusplus
}
#endif

#endif /* SPS_H */
