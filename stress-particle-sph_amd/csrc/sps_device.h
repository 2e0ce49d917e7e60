// sps_device.h - per-particle and per-pair arithmetic of the reference hot
// path, as device functions for gfx950.
//
// Every function restates one reference routine with its operation order and
// its mixed precision: a single-precision operand met by a double one is
// widened and the operation done in double; a result assigned to a
// single-precision variable is rounded then.  The translation unit is built
// with -ffp-contract=off, so no product is fused into a sum, and with
// correctly rounded single-precision division and square root.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sps {

// Constants of one problem, read by every kernel.
struct DevParams {
    int nn, ns, nd, nd2, nt, nt2;      // counts (reference ranges)
    int cspm, update_x, ncrit, ntype_eco, do_bforce, do_av, ic_grav;
    int h_uni;                         // every particle has the same smoothing length
    double pi, dx, alpha, beta, damping, cspm_tol;
    double D11, D22, D12, D33, D41, D42;
    double tanfi, coh;                 // props(1,13), props(1,14)
    double young, poiss;               // props(1,3), props(1,4)
    double xmin_dom[2], xmax_dom[2];
    double cgrav[2];
    // Perzyna models (ntype_eco > 1, ncrit 1..4): props(1,7:12), sqrt(3.0) in
    // single precision, sin(props(1,9) * 0.017453292) from the host's libm
    double pz_fdatm0, pz_hards, pz_gamma, pz_delta, pz_nflow, pz_root3, pz_snphi;
};

// One material's constants (an ensemble of instances has one per instance;
// the reference has one, props(1,:)): tan(phi), c (props(1,13:14)), E, nu
// (props(1,3:4)) and the elastic matrix (3_...:719-730).
struct Material {
    double tanfi, coh, young, poiss, D11, D22, D12, D33, D41, D42;
};

// Kind of a storage slot: the storage keeps the reference's index ranges
// (nodes, then stress points, then dummies), only reordered inside each.
enum { KIND_NODE = 0, KIND_SP = 1, KIND_DUMMY = 2 };

__device__ __forceinline__ int kind_of(int s, int nn, int nt) {
    return s < nn ? KIND_NODE : (s < nt ? KIND_SP : KIND_DUMMY);
}

// kernel(): cubic spline, 2-D (code/2_SPH_main_2018.f90:1440-1492).
// dxv = x_receiver - x_partner; the gradient is d W / d x_receiver.  The
// double results are rounded to single precision as the pair list stores
// them (code/5_SPH_global_vars_2018.f90:19-21, 2_...:1376-1378).
__device__ __forceinline__ void cubic_kernel(double r, double dx0, double dx1, double hs,
                                             double pi, float &w, float &gx, float &gy) {
    double q = r / hs;
    double factor = 15.0 / (7.0 * pi * hs * hs);
    double wd, g0, g1;
    if (q >= 0.0 && q <= 1.0) {
        wd = factor * ((double)(2.0f / 3.0f) - q * q + (q * (q * q)) / 2.0);
        double c = factor * (-2.0 + (double)(3.0f / 2.0f) * q) / (hs * hs);
        g0 = c * dx0;
        g1 = c * dx1;
    } else if (q > 1.0 && q <= 2.0) {
        double t = 2.0 - q;
        wd = factor * 1.0 / 6.0 * (t * (t * t));
        double c = factor * 1.0 / 6.0 * 3.0 * (t * t) / hs;
        g0 = -(c * (dx0 / r));
        g1 = -(c * (dx1 / r));
    } else {
        wd = 0.0;
        g0 = 0.0;
        g1 = 0.0;
    }
    w = (float)wd;
    gx = (float)g0;
    gy = (float)g1;
}

// The parts of the kernel that depend on h alone, for the last smoothing
// length seen: the factor 15 / (7 pi h^2), h^2, factor * 1 / 6 and
// factor * 1 / 6 * 3 -- the same operations on the same h as cubic_kernel's
// left-to-right expressions, evaluated once per distinct h.
struct KernelFactor {
    double h = -1.0, f = 0.0, h2 = 0.0, f6 = 0.0, f63 = 0.0;
    __device__ __forceinline__ const KernelFactor &get(double hs, double pi) {
        if (!(hs == h)) {
            h = hs;
            f = 15.0 / (7.0 * pi * hs * hs);
            h2 = hs * hs;
            f6 = f * 1.0 / 6.0;
            f63 = f6 * 3.0;
        }
        return *this;
    }
};

// cubic_kernel with the h-only parts supplied (K.get(hs, pi)).  The two
// regions' divisions of the gradient factor (by h^2 inside q <= 1, by h
// beyond) are one division of per-lane operands, so that a wave whose lanes
// fall in both regions divides once; every value is the one the reference's
// branch computes.
__device__ __forceinline__ void cubic_kernel_f(double r, double dx0, double dx1, double hs, const KernelFactor &K,
                                               float &w, float &gx, float &gy) {
    const double q = r / hs;
    const double factor = K.f;
    const bool in1 = q >= 0.0 && q <= 1.0;
    const bool in2 = q > 1.0 && q <= 2.0;
    const double t = 2.0 - q;
    const double num = in1 ? factor * (-2.0 + (double)(3.0f / 2.0f) * q) : K.f63 * (t * t);
    const double c = num / (in1 ? K.h2 : hs);
    double wd = 0.0, g0 = 0.0, g1 = 0.0;
    if (in1) {
        wd = factor * ((double)(2.0f / 3.0f) - q * q + (q * (q * q)) / 2.0);
        g0 = c * dx0;
        g1 = c * dx1;
    } else if (in2) {
        wd = K.f6 * (t * (t * t));
        g0 = -(c * (dx0 / r));
        g1 = -(c * (dx1 / r));
    }
    w = (float)wd;
    gx = (float)g0;
    gy = (float)g1;
}

// The w of cubic_kernel alone, with the same operations.
__device__ __forceinline__ float cubic_w_f(double r, double hs, double factor) {
    double q = r / hs;
    double wd;
    if (q >= 0.0 && q <= 1.0) {
        wd = factor * ((double)(2.0f / 3.0f) - q * q + (q * (q * q)) / 2.0);
    } else if (q > 1.0 && q <= 2.0) {
        double t = 2.0 - q;
        wd = factor * 1.0 / 6.0 * (t * (t * t));
    } else {
        wd = 0.0;
    }
    return (float)wd;
}

// The w of cubic_kernel alone, with the same operations.
__device__ __forceinline__ float cubic_w(double r, double hs, double pi) {
    double q = r / hs;
    double factor = 15.0 / (7.0 * pi * hs * hs);
    double wd;
    if (q >= 0.0 && q <= 1.0) {
        wd = factor * ((double)(2.0f / 3.0f) - q * q + (q * (q * q)) / 2.0);
    } else if (q > 1.0 && q <= 2.0) {
        double t = 2.0 - q;
        wd = factor * 1.0 / 6.0 * (t * (t * t));
    } else {
        wd = 0.0;
    }
    return (float)wd;
}

// The acceptance decision of pair_distance alone, for a pair whose limit
// 2*mhsml and squared-distance bounds are precomputed (per pair, or once
// per receiver when every h is the same: (h + h)/2 is then h exactly).
// Outside a relative band of 1e-12 around the limit the squared distance
// decides it: the computed bounds (2*mhsml)^2 * (1 -+ 1e-12) are within 3
// units in the last place of the exact ones, so driac below the lower one
// gives sqrt(driac) < 2*mhsml by more than the square root's rounding, and
// driac above the upper one the converse.  Inside the band the square root
// is taken as pair_distance takes it.
struct AcceptBounds {
    double lim, up, lo;
};

__device__ __forceinline__ AcceptBounds accept_bounds(double hi, double hj) {
    AcceptBounds b;
    b.lim = 2 * ((hi + hj) / 2.0);
    const double l2 = b.lim * b.lim;
    b.up = l2 * (1.0 + 1e-12);
    b.lo = l2 * (1.0 - 1e-12);
    return b;
}

__device__ __forceinline__ bool accepted_by(double dx0, double dx1, const AcceptBounds &b) {
    double driac = dx0 * dx0;
    driac = driac + dx1 * dx1;
    if (driac > b.up) return false;
    if (driac < b.lo) return true;
    // no instruction: makes the square root depend on this (rare) branch so
    // that it is not evaluated for every candidate
    asm volatile("" : "+v"(driac));
    return sqrt(driac) < b.lim;
}

// Pair acceptance of grid_find_NEW (2_...:1345-1353): r < 2 * (h_i + h_j)/2.
// Returns r (>= 0) when accepted, -1 otherwise.  Symmetric in i, j.
// A candidate whose squared distance exceeds (2*mhsml)^2 by a relative
// 1e-12 is rejected without the square root: then sqrt(driac) > 2*mhsml
// exactly, and its correctly rounded value cannot fall below the
// representable 2*mhsml.
__device__ __forceinline__ double pair_distance(double dx0, double dx1, double hi, double hj,
                                                double &mhsml) {
    double driac = dx0 * dx0;
    driac = driac + dx1 * dx1;
    mhsml = (hi + hj) / 2.0;
    double lim = 2 * mhsml;
    if (driac > (lim * lim) * (1.0 + 1e-12)) return -1.0;
    double r = sqrt(driac);
    return (r < lim) ? r : -1.0;
}

// Workgroups are dealt round-robin to the 8 XCDs.  With a grid that is a
// multiple of 8 this gives XCD k one contiguous eighth of the tiles, so the
// receivers it handles walk up the same cell strips and share partner data
// in that XCD's L2.
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int xcd_block() {
    int per = gridDim.x / NUM_XCD;
    return (blockIdx.x % NUM_XCD) * per + blockIdx.x / NUM_XCD;
}

// Tile of this workgroup in the tiled pair kernels.  The hardware hands
// consecutive workgroups to the XCDs in turn; each XCD gets runs of TILE_RUN
// consecutive tiles, the runs dealt out round-robin, so that neighbouring
// tiles (which share window data) mostly share an L2 while every XCD gets
// tiles from the whole domain (tiles differ in cost by region).  A bijection
// of [0, gridDim.x) when gridDim.x is a multiple of NUM_XCD: the blocks
// beyond the last whole run of each XCD take the remaining tiles in order.
constexpr int TILE_RUN = 16;
__device__ __forceinline__ int tile_index() {
    const int xcd = blockIdx.x % NUM_XCD, k = blockIdx.x / NUM_XCD;
    const int per = gridDim.x / NUM_XCD;
    const int full = per / TILE_RUN * TILE_RUN;
    if (k < full) return ((k / TILE_RUN) * NUM_XCD + xcd) * TILE_RUN + (k % TILE_RUN);
    return full * NUM_XCD + xcd * (per - full) + (k - full);
}

// adapt_stress2 for one particle (code/3_SPH_material_2018.f90:2066-2140).
__device__ __forceinline__ void adapt_stress(double4 &s, double tanfi, double coh) {
    double alpha2 = tanfi / (sqrt(9 + 12 * (tanfi * tanfi)));
    double kc = (3 * coh) / (sqrt(9 + 12 * (tanfi * tanfi)));
    double smean = (s.x + s.y + s.w) / 3.0;
    double d0 = s.x - smean, d1 = s.y - smean, d2 = s.z, d3 = s.w - smean;
    double varj2 = d2 * d2 + 0.5 * (d0 * d0 + d1 * d1 + d3 * d3);
    double yield = -(alpha2 * 3 * smean) + kc;
    if (yield < 0) {
        s.x = s.x - smean + kc / (3 * alpha2);
        s.y = s.y - smean + kc / (3 * alpha2);
        s.w = s.w - smean + kc / (3 * alpha2);
        smean = (s.x + s.y + s.w) / 3.0;
        d0 = s.x - smean;
        d1 = s.y - smean;
        d2 = s.z;
        d3 = s.w - smean;
        varj2 = d2 * d2 + 0.5 * (d0 * d0 + d1 * d1 + d3 * d3);
        yield = -(alpha2 * 3 * smean) + kc;
    }
    if (yield < sqrt(varj2)) {
        double rn = (-(3 * alpha2 * smean) + kc) / (sqrt(varj2));
        if (sqrt(varj2) <= (double)10e-06f) rn = 0;
        s.x = rn * d0 + smean;
        s.y = rn * d1 + smean;
        s.w = rn * d3 + smean;
        s.z = rn * d2;
    }
}

// Perzyna viscoplastic strain rate of one stress point (ncrit 1 Tresca, 2 von
// Mises, 3 Mohr-Coulomb, 4 Drucker-Prager; 3_...:2168-2525 of the strain-
// localisation fork): invar09's invariants and yield measure, the current
// yield size fdatm0 + H evpstn (at least 0.1 fdatm0), yieldf09's flow vector
// and flowvp09's overstress function, with the reference's single-precision
// literals.  Zero below the yield size or under the 0.01 threshold.
// <F>^delta with delta = 1 is F itself (what a correctly rounded pow gives).
__device__ __forceinline__ void perzyna_vivel(const DevParams &P, const double4 &st, double evpstn, double v[4]) {
    const int ncrit = P.ncrit;
    const double root3 = P.pz_root3;
    const double smean = (st.x + st.y + st.w) / 3.0;
    const double dv0 = st.x - smean, dv1 = st.y - smean, dv2 = st.z, dv3 = st.w - smean;
    const double varj2 = dv2 * dv2 + 0.5 * (dv0 * dv0 + dv1 * dv1 + dv3 * dv3);
    const double varj3 = dv3 * (dv3 * dv3 - varj2);
    const double steff = sqrt(varj2);
    const double snphi = P.pz_snphi;
    double theta = 0.0;
    if (ncrit == 1 || ncrit == 3) {   // the other criteria never use the Lode angle
        double sint3;
        if (steff != 0.0) {
            sint3 = -(3.0 * root3 * varj3 / (2.0 * varj2 * steff));
            if (sint3 > 1.0) sint3 = 1.0;
        } else {
            sint3 = 0.0;
        }
        if (sint3 < -1.0) sint3 = -1.0;
        if (sint3 > 1.0) sint3 = 1.0;
        theta = asin(sint3) / 3.0;
    }
    double yield = 0.0;
    if (ncrit == 1) yield = 2.0 * cos(theta) * steff;
    else if (ncrit == 2) yield = root3 * steff;
    else if (ncrit == 3) yield = smean * snphi + steff * (cos(theta) - sin(theta) * snphi / root3);
    else if (ncrit == 4) yield = 6.0 * smean * snphi / (root3 * (3.0 - snphi)) + steff;
    double fdatm = P.pz_fdatm0 + P.pz_hards * evpstn;
    double fact = fabs(fdatm) / fabs(P.pz_fdatm0);
    if (fact < (double)0.1f) fact = (double)0.1f;
    fdatm = P.pz_fdatm0 * fact;
    v[0] = v[1] = v[2] = v[3] = 0.0;
    if (!(yield > fdatm)) return;
    double a2[4] = {0.0, 0.0, 0.0, 0.0};
    if (steff > 0) {
        a2[0] = dv0 / (2.0 * steff);
        a2[1] = dv1 / (2.0 * steff);
        a2[3] = dv3 / (2.0 * steff);
        a2[2] = dv2 / steff;
    }
    const double a3[4] = {dv1 * dv3 + varj2 / 3.0, dv0 * dv3 + varj2 / 3.0, -(2.0 * dv2 * dv3),
                          dv0 * dv1 - dv2 * dv2 + varj2 / 3.0};
    double c1 = 0.0, c2 = 0.0, c3 = 0.0;
    if (ncrit == 1 || ncrit == 3) {
        const double tanth = tan(theta), tant3 = tan(3.0 * theta), sinth = sin(theta), costh = cos(theta);
        const double cost3 = cos(3.0 * theta);
        const double abthe = fabs(theta * (double)57.29577951308f);
        if (ncrit == 1) {
            if (abthe >= 29.0) {
                c2 = root3;
            } else {
                c2 = 2.0 * (costh + sinth * tant3);
                c3 = root3 * sinth / (varj2 * cost3);
            }
        } else {
            c1 = snphi / 3.0;
            if (abthe >= 29.0) {
                const double plumi = theta > 0.0 ? -1.0 : 1.0;
                c2 = 0.5 * (root3 + plumi * c1 * root3);
            } else {
                c2 = costh * ((1.0 + tanth * tant3) + c1 * (tant3 - tanth) * root3);
                c3 = (root3 * sinth + 3.0 * c1 * costh) / (2.0 * varj2 * cost3);
            }
        }
    } else if (ncrit == 2) {
        c2 = root3;
    } else {
        c1 = 2.0 * snphi / (root3 * (3.0 - snphi));
        c2 = 1.0;
    }
    const double a1[4] = {1.0, 1.0, 0.0, 1.0};
    double av[4];
    for (int i = 0; i < 4; ++i) av[i] = c1 * a1[i] + c2 * a2[i] + c3 * a3[i];
    const double fnorm = (yield - fdatm) / fdatm;
    if (fnorm >= (double)0.01f) {
        double cmult;
        if (P.pz_nflow != 1) cmult = P.pz_gamma * (exp(P.pz_delta * fnorm) - 1.0);
        else cmult = P.pz_gamma * (P.pz_delta == 1.0 ? fnorm : pow(fnorm, P.pz_delta));
        for (int i = 0; i < 4; ++i) v[i] = cmult * av[i];
    }
}

// plastic_terms -> drucker_prager -> Get_derivative_intvars for one stress
// point (code/3_SPH_material_2018.f90:1867-1933, 1937-2062, 2557-2635).
// st: stress, gu: grad_u as [g11, g21, g12, g22]; fdr: f_drucker in/out.
// Gs receives the stress source (-G), der1 the deviatoric plastic strain rate.
__device__ __forceinline__ void plastic_terms(const DevParams &P, const double4 &st,
                                              const double4 &gu, double &fdr, double time_sph,
                                              double evpstn, double4 &Gs, double &der1) {
    double G0 = 0.0, G1 = 0.0, G2a = 0.0, G3 = 0.0;
    double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
    Gs = make_double4(0.0, 0.0, 0.0, 0.0);
    if (P.ntype_eco > 1) {
        if (P.ncrit == 12) {
            double f0 = fdr;
            float tanfi = (float)P.tanfi, coh = (float)P.coh;
            float young = (float)P.young, poiss = (float)P.poiss;
            float smean = (float)((st.x + st.y + st.w) / 3.0);
            float dv0 = (float)(st.x - (double)smean);
            float dv1 = (float)(st.y - (double)smean);
            float dv2 = (float)st.z;
            float dv3 = (float)(st.w - (double)smean);
            float varj2 = dv2 * dv2 + 0.5f * (dv0 * dv0 + dv1 * dv1 + dv3 * dv3);
            float vari1 = 3 * smean;
            float eps11 = (float)gu.x;
            float eps12 = (float)(0.5 * (gu.z + gu.y));
            float eps22 = (float)gu.w;
            float emean = eps11 + eps22;
            float alpha2 = tanfi / (sqrtf(9 + 12 * (tanfi * tanfi)));
            float kc = (3 * coh) / (sqrtf(9 + 12 * (tanfi * tanfi)));
            float yield = -(alpha2 * vari1) + kc;
            fdr = (double)(sqrtf(varj2) - yield);
            float f1 = (float)fdr;
            double df = (double)f1 - f0;
            float G_mod = young / (2.0f * (1.0f + poiss));
            float K_mod = young / (3.0f * (1.0f - 2.0f * poiss));
            float s_eps = dv0 * eps11 + 2 * dv2 * eps12 + dv1 * eps22;
            if (time_sph > 0 && f1 >= 0 && df >= 0 && sqrtf(varj2) >= 10e-06f) {
                float sj = sqrtf(varj2);
                float gq = G_mod / sj;
                float lambda_1 = 3 * alpha2 * K_mod * emean;
                float lambda_2 = gq * s_eps;
                float lambda_3 = G_mod;
                float G2 = (lambda_1 + lambda_2) / lambda_3;
                G0 = (double)((gq * dv0) * G2);
                G1 = (double)((gq * dv1) * G2);
                G2a = (double)((gq * dv2) * G2);
                G3 = (double)((gq * dv3) * G2);
                float inv = 1.0f / sj;
                v0 = (double)((1.0f / 6.0f) * inv * (2 * dv0 - dv1 - dv3));
                v1 = (double)((1.0f / 6.0f) * inv * (2 * dv1 - dv0 - dv3));
                v2 = (double)(inv * dv2);
                v3 = (double)((1.0f / 6.0f) * inv * (2 * dv3 - dv0 - dv1));
                v0 = v0 * (double)G2;
                v1 = v1 * (double)G2;
                v2 = v2 * (double)G2;
                v3 = v3 * (double)G2;
            }
            Gs = make_double4(-G0, -G1, -G2a, -G3);
        } else if (P.ncrit >= 1 && P.ncrit <= 4) {
            // Gs = -Dmatx vivel with Get_Dmatx (3_...:2530-2561), row by row
            double vv[4];
            perzyna_vivel(P, st, evpstn, vv);
            const double c = P.young * (1.0 - P.poiss) / ((1.0 + P.poiss) * (1.0 - 2.0 * P.poiss));
            const double c12 = c * P.poiss / (1.0 - P.poiss);
            const double c33 = (1.0 - 2.0 * P.poiss) * c / (2.0 * (1.0 - P.poiss));
            double g[4] = {0.0, 0.0, 0.0, 0.0};
            const double D[4][4] = {{c, c12, 0.0, c12}, {c12, c, 0.0, c12}, {0.0, 0.0, c33, 0.0}, {c12, c12, 0.0, c}};
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) g[i] = g[i] - D[i][j] * vv[j];
            Gs = make_double4(g[0], g[1], g[2], g[3]);
            v0 = vv[0]; v1 = vv[1]; v2 = vv[2]; v3 = vv[3];
        }
        der1 = sqrt((2.0 * (v0 * v0 + v1 * v1 + v3 * v3) + v2 * v2) / 3.0);
    } else {
        der1 = 0.0;
    }
}

// One term of boundary_forces test = 2 (code/2_SPH_main_2018.f90:1117-1161),
// node i against inner-layer dummy j, single precision.  ri = x_i - x_j.
__device__ __forceinline__ void wall_force_term(double ri0, double ri1, double hi, double hj,
                                                double dx, float &t0, float &t1) {
    float r0 = (float)ri0, r1 = (float)ri1;
    float r2 = sqrtf(r0 * r0 + r1 * r1);
    float d0 = (float)(dx / 2.);
    float hh = (float)(0.5 * (hi + hj));
    float r_bound = r2 / (0.75f * hh);
    float f2 = (r2 > 0 && r2 < 1.5f * d0) ? 1 - (r2 / (1.5f * d0)) : 0.0f;
    float f;
    if (0 < r_bound && r_bound <= 2.0f / 3.0f) f = 2.0f / 3.0f;
    else if (2.0f / 3.0f < r_bound && r_bound <= 1) f = 2 * r_bound - 1.5f * (r_bound * r_bound);
    else if (1 < r_bound && r_bound < 2) f = 0.5f * ((2 - r_bound) * (2 - r_bound));
    else f = 0.0f;
    float c = 20;
    float scale = 0.01f * (c * c);
    t0 = scale * f2 * f * (r0 / (r2 * r2));
    t1 = scale * f2 * f * (r1 / (r2 * r2));
}

// ------------------------------------------------ learned stress update hook
// The final stress update of a step records every stress point whose trial
// stress the return map changes (a plastic trial state): for the network
// that replaces the return map (features as the network takes them) and, when
// a log is kept, as a training tuple of the plain method.
constexpr int SURR_NIN = 10;

struct SurrogateDev {
    unsigned *count;         // plastic trial states logged this step
    int cap;
    int *slot;               // their storage slots
    double4 *s_n, *s_tr;     // stress before the step and trial stress
    double2 *mat;            // their tan(phi), c
    float *feat;             // their features (cap x 10)
    float stress_scale, coh_scale;
    unsigned long long *stats;   // [0] evaluated, [1] kept
    // plain-method log (training tuples), when log_cap > 0
    unsigned *log_count;
    int log_cap;
    double *log;             // log_cap x 12: sigma_n, sigma_trial, R(sigma_trial)
};

// Record a stress point's plastic trial state for the network (and the
// plain method's tuple for training).
__device__ __forceinline__ void surrogate_record(const SurrogateDev &G, int i, const double4 &s_n,
                                                 const double4 &s_tr, const double4 &s_ret, double tanfi,
                                                 double coh) {
    if (G.log_cap > 0) {
        const unsigned k = atomicAdd(G.log_count, 1u);
        if ((int)k < G.log_cap) {
            double *o = G.log + (long long)SPS_PLOG_ROW * k;
            o[0] = s_n.x; o[1] = s_n.y; o[2] = s_n.z; o[3] = s_n.w;
            o[4] = s_tr.x; o[5] = s_tr.y; o[6] = s_tr.z; o[7] = s_tr.w;
            o[8] = s_ret.x; o[9] = s_ret.y; o[10] = s_ret.z; o[11] = s_ret.w;
            o[12] = tanfi; o[13] = coh;
        }
    }
    if (G.cap <= 0) return;
    const unsigned k = atomicAdd(G.count, 1u);
    if ((int)k >= G.cap) return;          // beyond the batch: the exact result stays
    G.slot[k] = i;
    G.mat[k] = make_double2(tanfi, coh);
    G.s_n[k] = s_n;
    G.s_tr[k] = s_tr;
    const float is = 1.0f / G.stress_scale;
    float *f = G.feat + (long long)k * SURR_NIN;
    f[0] = (float)s_n.x * is; f[1] = (float)s_n.y * is; f[2] = (float)s_n.z * is; f[3] = (float)s_n.w * is;
    f[4] = (float)(s_tr.x - s_n.x) * is; f[5] = (float)(s_tr.y - s_n.y) * is;
    f[6] = (float)(s_tr.z - s_n.z) * is; f[7] = (float)(s_tr.w - s_n.w) * is;
    f[8] = (float)coh / G.coh_scale;
    f[9] = (float)tanfi;
}


// ------------------------------------------------------- mixed precision
// The mixed-precision form sums pair terms in single precision with
// compensation (Kahan 1965: the rounding error of every addition is carried
// into the next), keeping positions, the stress state and the return map in
// double precision.
struct KSum {
    float s = 0.0f, c = 0.0f;   // sum; minus the low part lost so far
    __device__ __forceinline__ void add(float v) {
        const float y = v - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ double get() const { return (double)s - (double)c; }
};

// 15 / (7 pi) of the 2-D cubic spline's normalisation.
constexpr float CUBIC_2D_NORM = 0.682092613f;

// The cubic spline (the formula of kernel(), 2_...:1468-1492) evaluated in
// single precision from a relative position dx = x_r - x_s and the pair's
// smoothing length, with the hardware reciprocal and square root: w and
// d W / d x_r.  Zero beyond 2h.
__device__ __forceinline__ void cubic_kernel_mix(float dx0, float dx1, float mh, float &w, float &gx,
                                                 float &gy) {
    const float r2 = dx0 * dx0 + dx1 * dx1;
    const float ih = __builtin_amdgcn_rcpf(mh);
    const float r = __builtin_amdgcn_sqrtf(r2);
    const float q = r * ih;
    const float fac = CUBIC_2D_NORM * (ih * ih);
    if (q <= 1.0f) {
        w = fac * ((2.0f / 3.0f) - q * q + 0.5f * (q * (q * q)));
        const float c = fac * (-2.0f + 1.5f * q) * (ih * ih);
        gx = c * dx0;
        gy = c * dx1;
    } else if (q < 2.0f) {
        const float t = 2.0f - q;
        w = fac * (1.0f / 6.0f) * (t * (t * t));
        const float c = -(fac * 0.5f * (t * t) * ih) * __builtin_amdgcn_rcpf(r);
        gx = c * dx0;
        gy = c * dx1;
    } else {
        w = gx = gy = 0.0f;
    }
}

// A partner position relative to the tile origin o: staged windows hold it
// already in single precision; global-state positions are reduced here.
__device__ __forceinline__ float2 rel_pos(float2 x, double2) { return x; }
__device__ __forceinline__ float2 rel_pos(double2 x, double2 o) {
    return make_float2((float)(x.x - o.x), (float)(x.y - o.y));
}

}  // namespace sps
