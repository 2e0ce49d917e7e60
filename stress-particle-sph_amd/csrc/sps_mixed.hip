// sps_mixed.hip - mixed-precision form of the stage kernels.
//
// The same three tiled gathers as the double-precision form (stress-point
// update at the nodes; stress-point update and derivative at the stress
// points; derivative, artificial viscosity and boundary force at the nodes)
// over the same per-step pair lists, with three differences:
//
// * pair arithmetic in single precision: the window is staged in single
//   precision, positions relative to the tile's first receiver (so a 10 m
//   domain keeps ~1e-8 m resolution inside a 0.3 m window);
// * w and grad w computed per visit from the staged positions
//   (cubic_kernel_mix) instead of read from the lists, whose words then hold
//   only partner indices: 2 bytes per visit instead of 14;
// * every pair sum compensated (KSum).
//
// Positions, velocities, the stress state, the CSPM solve, plastic_terms,
// the return map and the RK update stay in double precision (the shared
// tails sp_stage_tail / node_stage_tail).  Pairs are summed in forward list
// order; the double-precision form keeps the reference's linked-list order.
// A tile whose window does not fit walks the global state (exact kernel
// values, forward order), as in the double-precision form.

namespace sps {

// ----------------------------------------------------- single-precision windows
constexpr LdsLayout layout_spu_nodes_mix() {
    LdsLayout L = layout_none();
    int o = 0;
    L.x[KIND_SP] = o; o += 8 * CAP;
    L.h[KIND_SP] = o; o += 4 * CAP;
    L.mr[KIND_SP] = o; o += 4 * CAP;
    L.s0[KIND_SP] = o; o += 4 * CAP;
    L.s1[KIND_SP] = o; o += 4 * CAP;
    L.s2[KIND_SP] = o; o += 4 * CAP;
    L.s3[KIND_SP] = o; o += 4 * CAP;
    L.iv[KIND_SP] = o; o += 4 * CAP;
    L.bytes = o;
    return L;
}

constexpr LdsLayout layout_sp_update_mix() {
    LdsLayout L = layout_none();
    int o = 0;
    L.x[KIND_NODE] = o; o += 8 * CAP;
    L.v[KIND_NODE] = o; o += 8 * CAP;
    L.h[KIND_NODE] = o; o += 4 * CAP;
    L.mr[KIND_NODE] = o; o += 4 * CAP;
    L.x[KIND_DUMMY] = o; o += 8 * CAP_D;
    L.h[KIND_DUMMY] = o; o += 4 * CAP_D;
    L.mr[KIND_DUMMY] = o; o += 4 * CAP_D;
    L.wall = o; o += 4 * CAP_D;    // wall position relative to the origin
    L.horiz = o; o += 4 * CAP_D;
    L.bytes = o;
    return L;
}

constexpr LdsLayout layout_deriv_nodes_mix() {
    LdsLayout L = layout_none();
    int o = 0;
    L.x[KIND_SP] = o; o += 8 * CAP;
    L.h[KIND_SP] = o; o += 4 * CAP;
    L.m[KIND_SP] = o; o += 4 * CAP;
    L.mr[KIND_SP] = o; o += 4 * CAP;
    L.p0[KIND_SP] = o; o += 4 * CAP;
    L.p1[KIND_SP] = o; o += 4 * CAP;
    L.p2[KIND_SP] = o; o += 4 * CAP;
    L.x[KIND_DUMMY] = o; o += 8 * CAP_D;
    L.h[KIND_DUMMY] = o; o += 4 * CAP_D;
    L.m[KIND_DUMMY] = o; o += 4 * CAP_D;
    L.rho[KIND_DUMMY] = o; o += 4 * CAP_D;
    const int phase1 = o;
    o = 0;
    L.x[KIND_NODE] = o; o += 8 * CAP;
    L.v[KIND_NODE] = o; o += 8 * CAP;
    L.h[KIND_NODE] = o; o += 4 * CAP;
    L.m[KIND_NODE] = o; o += 4 * CAP;
    L.rho[KIND_NODE] = o; o += 4 * CAP;
    L.bytes = o > phase1 ? o : phase1;
    return L;
}

constexpr LdsLayout LAY_SPU_NODES_MIX = layout_spu_nodes_mix();
constexpr LdsLayout LAY_SP_UPDATE_MIX = layout_sp_update_mix();
constexpr LdsLayout LAY_DERIV_NODES_MIX = layout_deriv_nodes_mix();

// Partner fields of a single-precision window.
template <const LdsLayout &LY> struct LdsFieldsMix {
    template <int K> __device__ __forceinline__ float2 x(KindTag<K>, int j) const { return lds_at<float2>(LY.x[K])[j]; }
    template <int K> __device__ __forceinline__ float h(KindTag<K>, int j) const { return lds_at<float>(LY.h[K])[j]; }
    template <int K> __device__ __forceinline__ float m(KindTag<K>, int j) const { return lds_at<float>(LY.m[K])[j]; }
    template <int K> __device__ __forceinline__ float rho(KindTag<K>, int j) const { return lds_at<float>(LY.rho[K])[j]; }
    template <int K> __device__ __forceinline__ float mr(KindTag<K>, int j) const { return lds_at<float>(LY.mr[K])[j]; }
    template <int K> __device__ __forceinline__ float2 vel(KindTag<K>, int j) const {
        return lds_at<float2>(LY.v[K])[j];
    }
    template <int K> __device__ __forceinline__ float4 st(KindTag<K>, int j) const {
        return make_float4(lds_at<float>(LY.s0[K])[j], lds_at<float>(LY.s1[K])[j], lds_at<float>(LY.s2[K])[j],
                           lds_at<float>(LY.s3[K])[j]);
    }
    template <int K> __device__ __forceinline__ float iv(KindTag<K>, int j) const { return lds_at<float>(LY.iv[K])[j]; }
    template <int K> __device__ __forceinline__ float3 sr2(KindTag<K>, int j) const {
        return make_float3(lds_at<float>(LY.p0[K])[j], lds_at<float>(LY.p1[K])[j], lds_at<float>(LY.p2[K])[j]);
    }
    template <int K> __device__ __forceinline__ float wrel(KindTag<K>, int j) const { return lds_at<float>(LY.wall)[j]; }
    template <int K> __device__ __forceinline__ float horiz(KindTag<K>, int j) const {
        return lds_at<float>(LY.horiz)[j];
    }
};

// The same fields from the global state, reduced as the windows hold them
// (the global walk of a tile whose window does not fit).  xk / hk serve the
// walk's acceptance test.
struct GlobalFieldsMix {
    StateBufs S;
    double2 o;
    __device__ __forceinline__ double2 xk(int, int j) const { return S.x[j]; }
    __device__ __forceinline__ double hk(int, int j) const { return S.hsml[j]; }
    template <int K> __device__ __forceinline__ float2 x(KindTag<K>, int j) const { return rel_pos(S.x[j], o); }
    template <int K> __device__ __forceinline__ float h(KindTag<K>, int j) const { return (float)S.hsml[j]; }
    template <int K> __device__ __forceinline__ float m(KindTag<K>, int j) const { return (float)S.mass[j]; }
    template <int K> __device__ __forceinline__ float rho(KindTag<K>, int j) const { return (float)S.rho[j]; }
    template <int K> __device__ __forceinline__ float mr(KindTag<K>, int j) const {
        return (float)(S.mass[j] / S.rho[j]);
    }
    template <int K> __device__ __forceinline__ float2 vel(KindTag<K>, int j) const {
        const double2 v = S.vel[j];
        return make_float2((float)v.x, (float)v.y);
    }
    template <int K> __device__ __forceinline__ float4 st(KindTag<K>, int j) const {
        const double4 s = S.stress[j];
        return make_float4((float)s.x, (float)s.y, (float)s.z, (float)s.w);
    }
    template <int K> __device__ __forceinline__ float iv(KindTag<K>, int j) const { return (float)S.ivar1[j]; }
    template <int K> __device__ __forceinline__ float3 sr2(KindTag<K>, int j) const {
        const double rs = S.rho[j], rs2 = rs * rs;
        const double4 s = S.stress[j];
        return make_float3((float)(s.x / rs2), (float)(s.y / rs2), (float)(s.z / rs2));
    }
    template <int K> __device__ __forceinline__ float wrel(KindTag<K>, int j) const {
        return (float)((double)S.wallpos[j] - (S.horiz[j] == 1.0f ? o.y : o.x));
    }
    template <int K> __device__ __forceinline__ float horiz(KindTag<K>, int j) const { return S.horiz[j]; }
};

// A list walk's body with w and grad w computed from the window: the
// receiver at xr (relative), smoothing length hr; hu > 0 when every
// particle's smoothing length is hu.
template <typename F> struct OnTheFly {
    F &f;
    float2 xr;
    float hr, hu;
    template <typename KT, typename A>
    __device__ __forceinline__ void operator()(KT kt, const A &fields, int j, float, float, float) const {
        const float2 xs = fields.x(kt, j);
        const float mh = hu > 0.0f ? hu : 0.5f * (hr + fields.h(kt, j));
        float w, gx, gy;
        cubic_kernel_mix(xr.x - xs.x, xr.y - xs.y, mh, w, gx, gy);
        f(kt, fields, j, w, gx, gy);
    }
};

template <typename F> __device__ __forceinline__ OnTheFly<F> on_the_fly(F &f, float2 xr, float hr, float hu) {
    return OnTheFly<F>{f, xr, hr, hu};
}

// Staging of a position relative to the origin.
__device__ __forceinline__ float2 to_rel(double2 x, double2 o) { return rel_pos(x, o); }

// ------------------------------------------------------------------ kernels
// stress_point_update for node receivers (as k_spu_nodes).
__global__ __launch_bounds__(TILE) void k_spu_nodes_mix(StateBufs S, SearchView V, PairLists PL, int adapt,
                                                        double tanfi, double coh, float hu, int nn) {
    __shared__ StagePlan pl;
    constexpr const LdsLayout &LY = LAY_SPU_NODES_MIX;
    const WalkCtx C = walk_ctx(V);
    const int tile = tile_index();
    const Tile T = PL.tiles[tile];
    if (T.first >= T.end) return;
    const int r = T.first + (int)threadIdx.x;
    const bool inside = r < T.end && V.cell[r] < C.g.ndivt;
    ListWalk<MASK_SP, false, false> lw;
    lw.start(PL.ia, nullptr, nullptr, nullptr, tile, inside ? PL.deg_a[r] : 0, false);
    const double2 o = S.x[T.first];
    bool lds = T.fits && PL.ok[tile];
    if (lds) {
        load_plan<KIND_SP>(PL, tile, 0, pl);
        __syncthreads();
        lds = pl.total <= CAP;
    }
    if (lds) {
        struct Item {
            double2 x;
            double h, m, rho, iv;
            double4 st;
        };
        copy_kind<KIND_SP>(T, V, C, nullptr, pl,
                           [&](int g) { return Item{S.x[g], S.hsml[g], S.mass[g], S.rho[g], S.ivar1[g], S.stress[g]}; },
                           [&](int i, const Item &v) {
                               lds_at<float2>(LY.x[KIND_SP])[i] = to_rel(v.x, o);
                               lds_at<float>(LY.h[KIND_SP])[i] = (float)v.h;
                               lds_at<float>(LY.mr[KIND_SP])[i] = (float)(v.m / v.rho);
                               lds_at<float>(LY.s0[KIND_SP])[i] = (float)v.st.x;
                               lds_at<float>(LY.s1[KIND_SP])[i] = (float)v.st.y;
                               lds_at<float>(LY.s2[KIND_SP])[i] = (float)v.st.z;
                               lds_at<float>(LY.s3[KIND_SP])[i] = (float)v.st.w;
                               lds_at<float>(LY.iv[KIND_SP])[i] = (float)v.iv;
                           });
        __syncthreads();
    }
    if (r >= T.end) return;
    KSum t0, t1, t2, t3, tv, norm;
    auto body = [&](auto kt, const auto &A, int j, float w, float, float) __attribute__((always_inline)) {
        const float h1 = A.mr(kt, j) * w;
        const float4 st = A.st(kt, j);
        t0.add(st.x * h1);
        t1.add(st.y * h1);
        t2.add(st.z * h1);
        t3.add(st.w * h1);
        tv.add(A.iv(kt, j) * h1);
        norm.add(h1);
    };
    if (inside) {
        const double2 xr = S.x[r];
        const double hr = S.hsml[r];
        if (lds) {
            lw.run(LdsFieldsMix<LY>{}, on_the_fly(body, to_rel(xr, o), (float)hr, hu));
        } else {
            int cx, cy;
            cell_xy(V.cell[r], C.g, cx, cy);
            walk_stencil<MASK_SP>(GlobalCells{V.kstart, C.g, C.L}, GlobalFieldsMix{S, o}, C.g, cx, cy, xr, hr,
                                  KIND_NODE, r, false, true, V.pi, NoClass{}, body);
        }
    }
    double4 s = S.stress[r];
    const double nd = norm.get();
    if (nd != 0) {
        s = make_double4(t0.get() / nd, t1.get() / nd, t2.get() / nd, t3.get() / nd);
        S.ivar1[r] = tv.get() / nd;
    } else {
        S.vel[r] = make_double2(0, 0);
    }
    material_tc(S, r, tanfi, coh);
    if (adapt) adapt_stress(s, tanfi, coh);
    S.stress[r] = s;
}

// Stress-point receivers (as k_sp_update).
// stepv (a captured step): time_sph from stepv[0] instead of the argument.
// sig_out: where the stress after the return map goes (null: the state's
// stress), as in k_sp_update.
__global__ __launch_bounds__(TILE) void k_sp_update_mix(StateBufs S, Work W, SearchView V, PairLists PL,
                                                        DevParams P, double time_sph, double f2, int deriv,
                                                        int adapt, float hu, int nn, int nt,
                                                        const double *__restrict__ stepv,
                                                        double4 *__restrict__ sig_out) {
    if (stepv) time_sph = stepv[0];
    __shared__ StagePlan pl[2];
    constexpr const LdsLayout &LY = LAY_SP_UPDATE_MIX;
    const WalkCtx C = walk_ctx(V);
    const int tile = tile_index();
    const int ltile = PL.ntile_n + tile;
    const Tile T = PL.tiles[ltile];
    if (T.first >= T.end) return;
    const int r = T.first + (int)threadIdx.x;
    const bool inside = r < T.end && V.cell[r] < C.g.ndivt;
    const int deg = inside ? PL.deg_a[r] : 0;
    ListWalk<MASK_NODE, false, false> lw;
    lw.start(PL.ia, nullptr, nullptr, nullptr, ltile, deg, false);
    const double2 o = S.x[T.first];
    bool lds = T.fits && PL.ok[ltile];
    if (lds) {
        load_plan<KIND_NODE>(PL, ltile, 0, pl[0]);
        if (deriv) load_plan<KIND_DUMMY>(PL, ltile, 1, pl[1]);
        __syncthreads();
        lds = pl[0].total <= CAP && (!deriv || pl[1].total <= CAP_D);
    }
    if (lds) {
        struct NItem {
            double2 x, v;
            double h, m, rho;
        };
        copy_kind<KIND_NODE>(T, V, C, nullptr, pl[0],
                             [&](int g) { return NItem{S.x[g], S.vel[g], S.hsml[g], S.mass[g], S.rho[g]}; },
                             [&](int i, const NItem &n) {
                                 lds_at<float2>(LY.x[KIND_NODE])[i] = to_rel(n.x, o);
                                 lds_at<float2>(LY.v[KIND_NODE])[i] = make_float2((float)n.v.x, (float)n.v.y);
                                 lds_at<float>(LY.h[KIND_NODE])[i] = (float)n.h;
                                 lds_at<float>(LY.mr[KIND_NODE])[i] = (float)(n.m / n.rho);
                             });
        if (deriv) {
            struct DItem {
                double2 x;
                double h, m, rho;
                float wall, horiz;
            };
            copy_kind<KIND_DUMMY>(
                T, V, C, nullptr, pl[1],
                [&](int g) { return DItem{S.x[g], S.hsml[g], S.mass[g], S.rho[g], S.wallpos[g], S.horiz[g]}; },
                [&](int i, const DItem &d) {
                    lds_at<float2>(LY.x[KIND_DUMMY])[i] = to_rel(d.x, o);
                    lds_at<float>(LY.h[KIND_DUMMY])[i] = (float)d.h;
                    lds_at<float>(LY.mr[KIND_DUMMY])[i] = (float)(d.m / d.rho);
                    lds_at<float>(LY.wall)[i] = (float)((double)d.wall - (d.horiz == 1.0f ? o.y : o.x));
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
    const float2 xrr = to_rel(xr, o);
    const LdsFieldsMix<LY> LF{};
    const GlobalFieldsMix GF{S, o};

    // stress_point_update
    KSum v0, v1, norm;
    auto spu = [&](auto kt, const auto &A, int j, float w, float, float) __attribute__((always_inline)) {
        const float h2 = A.mr(kt, j) * w;
        const float2 vs = A.vel(kt, j);
        v0.add(vs.x * h2);
        v1.add(vs.y * h2);
        norm.add(h2);
    };
    if (inside) {
        if (lds)
            lw.run(LF, on_the_fly(spu, xrr, (float)hr, hu));
        else
            walk_stencil<MASK_NODE>(GlobalCells{V.kstart, C.g, C.L}, GF, C.g, cx, cy, xr, hr, KIND_SP, r, false,
                                    true, V.pi, NoClass{}, spu);
    }
    double2 vr = S.vel[r];
    const double nd = norm.get();
    if (nd != 0) {
        vr = make_double2(v0.get() / nd, v1.get() / nd);
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
    const float vrx = (float)vr.x, vry = (float)vr.y;
    KSum g00, g01, g10, g11, A0, A1, A2, A3;
    auto grad = [&](auto kt, const auto &A, int j, float, float gxf, float gyf) __attribute__((always_inline)) {
        constexpr int K = decltype(kt)::value;
        const float mr = A.mr(kt, j);
        if constexpr (K == KIND_NODE) {             // type 1: pair_i = this SP
            const float h1 = gxf * mr, h2 = gyf * mr;
            const float2 vs = A.vel(kt, j);
            const float d0 = vs.x - vrx, d1 = vs.y - vry;
            g00.add(d0 * h1);
            g01.add(d0 * h2);
            g10.add(d1 * h1);
            g11.add(d1 * h2);
            if (P.cspm) {
                const float2 xs = A.x(kt, j);
                const float e0 = xs.x - xrr.x, e1 = xs.y - xrr.y;
                A0.add(e0 * h1);
                A1.add(e1 * h1);
                A2.add(e0 * h2);
                A3.add(e1 * h2);
            }
        } else {                                    // type 9: pair_i = dummy, mirrored no-slip velocity
            const float gx = -gxf, gy = -gyf;
            const float wall = A.wrel(kt, j);
            const float2 xs = A.x(kt, j);
            float da, db;
            if (A.horiz(kt, j) == 1.0f) {
                da = fabsf(xrr.y - wall);
                db = fabsf(xs.y - wall);
            } else {
                da = fabsf(xrr.x - wall);
                db = fabsf(xs.x - wall);
            }
            const float bt = 1.0f + db / da;
            const float beta = 1.5f < bt ? 1.5f : bt;
            // v_r - v_dummy with v_dummy = (1 - beta) v_r
            const float d0 = vrx * beta, d1 = vry * beta;
            const float h1 = gx * mr, h2 = gy * mr;
            g00.add(d0 * h1);
            g01.add(d0 * h2);
            g10.add(d1 * h1);
            g11.add(d1 * h2);
        }
    };
    if (inside) {
        if (lds) {
            ListWalk<MASK_NODE | MASK_DUMMY, false, false> lg;
            lg.start(PL.ia, nullptr, nullptr, nullptr, ltile, deg, false);
            lg.run(LF, on_the_fly(grad, xrr, (float)hr, hu));
        } else
            walk_stencil<MASK_NODE | MASK_DUMMY>(GlobalCells{V.kstart, C.g, C.L}, GF, C.g, cx, cy, xr, hr, KIND_SP,
                                                 r, false, true, V.pi, NoClass{}, grad);
    }
    sp_stage_tail(Pm, S, W, r, st, g00.get(), g01.get(), g10.get(), g11.get(), A0.get(), A1.get(), A2.get(),
                  A3.get(), time_sph, f2);
}

// artificial_viscosity (2_...:826-904) of node receiver r in the mixed
// form: its node partners (list B) with kernel values computed in single
// precision from window-relative positions, and compensated sums; rounded
// to single precision as node_stage_tail takes them.  Every thread of the
// block calls it (it stages the node window through LDS).
__device__ __forceinline__ float2 av_mix(const StateBufs &S, const SearchView &V, const PairLists &PL,
                                         const DevParams &P, const WalkCtx &C, const Tile &T, int tile, int r,
                                         bool inside, bool tiled, int cx, int cy, double2 xr, double2 vr,
                                         double hr, double rhor, float2 xrr, double2 o, float hu, StagePlan &pl) {
    constexpr const LdsLayout &LY = LAY_DERIV_NODES_MIX;
    const LdsFieldsMix<LY> LF{};
    const GlobalFieldsMix GF{S, o};
    KSum ax, ay;
    if (!P.do_av) return make_float2(0.0f, 0.0f);
    ListWalk<MASK_NODE, false, false> lb;
    lb.start(PL.ib, nullptr, nullptr, nullptr, tile, inside ? PL.deg_b[r] : 0, false);
    bool lds2 = tiled;
    if (lds2) {
        load_plan<KIND_NODE>(PL, tile, 0, pl);
        __syncthreads();   // also: the phase-1 walks are done with the shared bytes
        lds2 = pl.total <= CAP;
    }
    if (lds2) {
        struct NItem {
            double2 x, v;
            double h, m, rho;
        };
        copy_kind<KIND_NODE>(T, V, C, nullptr, pl,
                             [&](int g) { return NItem{S.x[g], S.vel[g], S.hsml[g], S.mass[g], S.rho[g]}; },
                             [&](int i, const NItem &n) {
                                 lds_at<float2>(LY.x[KIND_NODE])[i] = to_rel(n.x, o);
                                 lds_at<float2>(LY.v[KIND_NODE])[i] = make_float2((float)n.v.x, (float)n.v.y);
                                 lds_at<float>(LY.h[KIND_NODE])[i] = (float)n.h;
                                 lds_at<float>(LY.m[KIND_NODE])[i] = (float)n.m;
                                 lds_at<float>(LY.rho[KIND_NODE])[i] = (float)n.rho;
                             });
        __syncthreads();
    }
    const float vrx = (float)vr.x, vry = (float)vr.y, hrf = (float)hr, rhof = (float)rhor;
    const float alpha = (float)P.alpha, beta = (float)P.beta;
    auto av = [&](auto kt, const auto &A, int j, float, float gxf, float gyf) __attribute__((always_inline)) {
        const float2 xs = A.x(kt, j), vs = A.vel(kt, j);
        const float xij = xrr.x - xs.x, yij = xrr.y - xs.y;
        const float hh = 0.5f * (hrf + A.h(kt, j));
        const float rho2 = 0.5f * (rhof + A.rho(kt, j));
        const float div_u = xij * (vrx - vs.x) + yij * (vry - vs.y);
        const float theta = (hh * div_u) / (xij * xij + yij * yij + 0.01f * (hh * hh));
        const float visc = div_u < 0.0f ? (-(alpha * 600.0f * theta) + beta * (theta * theta)) / rho2 : 0.0f;
        const float ms = A.m(kt, j);
        ax.add(visc * gxf * ms);
        ay.add(visc * gyf * ms);
    };
    if (inside) {
        if (lds2)
            lb.run(LF, on_the_fly(av, xrr, hrf, hu));
        else
            walk_stencil<MASK_NODE>(GlobalCells{V.kstart, C.g, C.L}, GF, C.g, cx, cy, xr, hr, KIND_NODE, r,
                                    false, true, V.pi, NoClass{}, av);
    }
    return make_float2((float)ax.get(), (float)ay.get());
}

// Node receivers (as k_deriv_nodes).
// stepv (a captured step): the gravity terms from stepv[1:3].
// sp_stress: the stress points' stresses after the stage's return map
// (k_sp_update_mix's sig_out, or the state's stress).
__global__ __launch_bounds__(TILE) void k_deriv_nodes_mix(StateBufs S, Work W, SearchView V, PairLists PL,
                                                          DevParams P, double gx_grav, double gy_grav, double f2,
                                                          int *__restrict__ err, float hu, int nn,
                                                          const double *__restrict__ stepv,
                                                          const double4 *__restrict__ sp_stress,
                                                          const float2 *__restrict__ av_in) {
    if (stepv) {
        gx_grav = stepv[1];
        gy_grav = stepv[2];
    }
    __shared__ StagePlan pl[2];
    constexpr const LdsLayout &LY = LAY_DERIV_NODES_MIX;
    const WalkCtx C = walk_ctx(V);
    const int tile = tile_index();
    const Tile T = PL.tiles[tile];
    if (T.first >= T.end) return;
    const int r = T.first + (int)threadIdx.x;
    const bool active = r < T.end;
    const int key_r = active ? V.cell[r] : C.g.ndivt;
    const bool inside = key_r < C.g.ndivt;
    ListWalk<MASK_SP | MASK_DUMMY, false, false> la;
    la.start(PL.ia, nullptr, nullptr, nullptr, tile, inside ? PL.deg_a[r] : 0, false);
    const double2 o = S.x[T.first];
    const bool tiled = T.fits && PL.ok[tile];
    int cx = 0, cy = 0;
    if (inside) cell_xy(key_r, C.g, cx, cy);
    const int ra = active ? r : T.first;   // a valid slot for the idle lanes' loads
    const double2 xr = S.x[ra], vr = S.vel[ra];
    const double4 sr = S.stress[ra];
    const double rhor = S.rho[ra], hr = S.hsml[ra];
    const double rr2 = rhor * rhor;
    const float or0 = (float)(sr.x / rr2), or1 = (float)(sr.y / rr2), or2 = (float)(sr.z / rr2);
    const float srx = (float)sr.x, sry = (float)sr.y, srz = (float)sr.z;
    const float2 xrr = to_rel(xr, o);
    const LdsFieldsMix<LY> LF{};
    const GlobalFieldsMix GF{S, o};

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
            double h, m, rho;
            double4 st;
        };
        copy_kind<KIND_SP>(T, V, C, nullptr, pl[0],
                           [&](int g) { return SItem{S.x[g], S.hsml[g], S.mass[g], S.rho[g], sp_stress[g]}; },
                           [&](int i, const SItem &v) {
                               lds_at<float2>(LY.x[KIND_SP])[i] = to_rel(v.x, o);
                               lds_at<float>(LY.h[KIND_SP])[i] = (float)v.h;
                               lds_at<float>(LY.m[KIND_SP])[i] = (float)v.m;
                               lds_at<float>(LY.mr[KIND_SP])[i] = (float)(v.m / v.rho);
                               const double rs2 = v.rho * v.rho;
                               lds_at<float>(LY.p0[KIND_SP])[i] = (float)(v.st.x / rs2);
                               lds_at<float>(LY.p1[KIND_SP])[i] = (float)(v.st.y / rs2);
                               lds_at<float>(LY.p2[KIND_SP])[i] = (float)(v.st.z / rs2);
                           });
        struct DItem {
            double2 x;
            double h, m, rho;
        };
        copy_kind<KIND_DUMMY>(T, V, C, nullptr, pl[1],
                              [&](int g) { return DItem{S.x[g], S.hsml[g], S.mass[g], S.rho[g]}; },
                              [&](int i, const DItem &d) {
                                  lds_at<float2>(LY.x[KIND_DUMMY])[i] = to_rel(d.x, o);
                                  lds_at<float>(LY.h[KIND_DUMMY])[i] = (float)d.h;
                                  lds_at<float>(LY.m[KIND_DUMMY])[i] = (float)d.m;
                                  lds_at<float>(LY.rho[KIND_DUMMY])[i] = (float)d.rho;
                              });
        __syncthreads();
    }
    KSum q00, q01, q10, q11, q20, q21, A0, A1, A2, A3;
    int ndum = 0;
    auto grad = [&](auto kt, const auto &A, int j, float, float gxf, float gyf) __attribute__((always_inline)) {
        constexpr int K = decltype(kt)::value;
        const float ms = A.m(kt, j);
        const float gx = -gxf, gy = -gyf;            // pair_i is the partner
        float p0, p1, p2;
        if constexpr (K == KIND_SP) {               // type 1: sigma_s / rho_s^2
            const float3 ps = A.sr2(kt, j);
            p0 = ps.x; p1 = ps.y; p2 = ps.z;
            if (P.cspm) {
                const float mr = A.mr(kt, j);
                const float h1b = -(gx * mr), h2b = -(gy * mr);
                const float2 xs = A.x(kt, j);
                const float e0 = xs.x - xrr.x, e1 = xs.y - xrr.y;
                A0.add(e0 * h1b);
                A1.add(e1 * h1b);
                A2.add(e0 * h2b);
                A3.add(e1 * h2b);
            }
        } else {                                    // type 6: dummy stress := node stress
            const float rs = A.rho(kt, j);
            const float irs2 = __builtin_amdgcn_rcpf(rs * rs);
            p0 = srx * irs2; p1 = sry * irs2; p2 = srz * irs2;
            ndum++;
        }
        const float s0 = or0 + p0, s1 = or1 + p1, s2 = or2 + p2;
        q00.add(-(ms * (gx * s0)));
        q01.add(-(ms * (gy * s0)));
        q10.add(-(ms * (gx * s1)));
        q11.add(-(ms * (gy * s1)));
        q20.add(-(ms * (gx * s2)));
        q21.add(-(ms * (gy * s2)));
    };
    if (inside) {
        if (lds)
            la.run(LF, on_the_fly(grad, xrr, (float)hr, hu));
        else
        {
            StateBufs S2 = S;   // stress points' stresses from sp_stress (the walk reads no node's)
            S2.stress = const_cast<double4 *>(sp_stress);
            walk_stencil<MASK_SP | MASK_DUMMY>(GlobalCells{V.kstart, C.g, C.L}, GlobalFieldsMix{S2, o}, C.g, cx,
                                               cy, xr, hr, KIND_NODE, r, false, true, V.pi, NoClass{}, grad);
        }
    }

    // phase 2: artificial viscosity over node partners (type 3), here or,
    // beside k_sp_update_mix, by k_av_nodes_mix (av_in)
    float2 avv = make_float2(0.0f, 0.0f);
    if (av_in) {
        if (active) avv = av_in[r];
    } else {
        avv = av_mix(S, V, PL, P, C, T, tile, r, inside, tiled, cx, cy, xr, vr, hr, rhor, xrr, o, hu, pl[0]);
    }
    if (!active) return;
    node_stage_tail(S, W, V, C, P, r, cx, cy, xr, vr, hr, q00.get(), q01.get(), q10.get(), q11.get(), q20.get(),
                    q21.get(), A0.get(), A1.get(), A2.get(), A3.get(), ndum, avv.x, avv.y, gx_grav, gy_grav, f2, err);
}

// artificial_viscosity of the node receivers alone (av_mix), written to
// av_out for k_deriv_nodes_mix: it reads only node data that the stress
// points' update leaves alone, so the RK stages run it beside
// k_sp_update_mix.
__global__ __launch_bounds__(TILE) void k_av_nodes_mix(StateBufs S, SearchView V, PairLists PL, DevParams P,
                                                       float hu, float2 *__restrict__ av_out, int nn) {
    __shared__ StagePlan pl[1];
    const WalkCtx C = walk_ctx(V);
    const int tile = tile_index();
    const Tile T = PL.tiles[tile];
    if (T.first >= T.end) return;
    const int r = T.first + (int)threadIdx.x;
    const bool active = r < T.end;
    const int key_r = active ? V.cell[r] : C.g.ndivt;
    const bool inside = key_r < C.g.ndivt;
    const double2 o = S.x[T.first];
    const bool tiled = T.fits && PL.ok[tile];
    int cx = 0, cy = 0;
    if (inside) cell_xy(key_r, C.g, cx, cy);
    const int ra = active ? r : T.first;   // a valid slot for the idle lanes' loads
    const double2 xr = S.x[ra], vr = S.vel[ra];
    const double rhor = S.rho[ra], hr = S.hsml[ra];
    const float2 a = av_mix(S, V, PL, P, C, T, tile, r, inside, tiled, cx, cy, xr, vr, hr, rhor, to_rel(xr, o), o,
                            hu, pl[0]);
    if (active) av_out[r] = a;
}

// ------------------------------------------------------ captured steps
// The step's changing scalars for a replayed graph: time_sph and the gravity
// terms (a launch outside the graph, before each replay).
// XSPH_update (2_...:189-239) in the mixed form: a receiver's same-kind
// partners (list B) with kernel values computed in single precision from
// window-relative positions, velocity differences in single precision and
// compensated sums.
constexpr LdsLayout layout_xsph_mix() {
    LdsLayout L = layout_none();
    for (int k = 0; k < 2; ++k) {   // one receiver kind per launch: both share the bytes
        L.x[k] = 0;
        L.v[k] = 8 * CAP;
        L.h[k] = 16 * CAP;
        L.mr[k] = 20 * CAP;
    }
    L.bytes = 24 * CAP;
    return L;
}
constexpr LdsLayout LAY_XSPH_MIX = layout_xsph_mix();

template <int KR>
__global__ __launch_bounds__(TILE) void k_xsph_mix(StateBufs S, SearchView V, PairLists PL, float hu,
                                                   double2 *__restrict__ xv) {
    __shared__ StagePlan pl;
    constexpr const LdsLayout &LY = LAY_XSPH_MIX;
    const WalkCtx C = walk_ctx(V);
    const int tile = tile_index();
    const int ltile = KR == KIND_NODE ? tile : PL.ntile_n + tile;
    const Tile T = PL.tiles[ltile];
    if (T.first >= T.end) return;
    const int r = T.first + (int)threadIdx.x;
    const bool inside = r < T.end && V.cell[r] < C.g.ndivt;
    ListWalk<(1 << KR), false, false> lw;
    lw.start(PL.ib, nullptr, nullptr, nullptr, ltile, inside ? PL.deg_b[r] : 0, false);
    const double2 o = S.x[T.first];
    bool lds = T.fits && PL.ok[ltile];
    if (lds) {
        load_plan<KR>(PL, ltile, 0, pl);
        __syncthreads();
        lds = pl.total <= CAP;
    }
    if (lds) {
        struct Item {
            double2 x, v;
            double h, m, rho;
        };
        copy_kind<KR>(T, V, C, nullptr, pl,
                      [&](int g) { return Item{S.x[g], S.vel[g], S.hsml[g], S.mass[g], S.rho[g]}; },
                      [&](int i, const Item &v) {
                          lds_at<float2>(LY.x[KR])[i] = to_rel(v.x, o);
                          lds_at<float2>(LY.v[KR])[i] = make_float2((float)v.v.x, (float)v.v.y);
                          lds_at<float>(LY.h[KR])[i] = (float)v.h;
                          lds_at<float>(LY.mr[KR])[i] = (float)(v.m / v.rho);
                      });
        __syncthreads();
    }
    if (r >= T.end) return;
    const double2 vr = S.vel[r];
    const float vx = (float)vr.x, vy = (float)vr.y;
    KSum s0, s1;
    auto body = [&](auto kt, const auto &A, int j, float w, float, float) __attribute__((always_inline)) {
        const float mr = A.mr(kt, j);
        const float2 v = A.vel(kt, j);
        s0.add((mr * (v.x - vx)) * w);
        s1.add((mr * (v.y - vy)) * w);
    };
    if (inside) {
        const double2 xr = S.x[r];
        const double hr = S.hsml[r];
        if (lds) {
            lw.run(LdsFieldsMix<LY>{}, on_the_fly(body, to_rel(xr, o), (float)hr, hu));
        } else {
            int cx, cy;
            cell_xy(V.cell[r], C.g, cx, cy);
            walk_stencil<(1 << KR)>(GlobalCells{V.kstart, C.g, C.L}, GlobalFieldsMix{S, o}, C.g, cx, cy, xr, hr, KR,
                                    r, false, true, V.pi, NoClass{}, body);
        }
    }
    xv[r] = make_double2(s0.get(), s1.get());
}

__global__ void k_set_step(double *__restrict__ stepv, double time_sph, double gx, double gy) {
    if (threadIdx.x == 0) {
        stepv[0] = time_sph;
        stepv[1] = gx;
        stepv[2] = gy;
    }
}

// Clears inside a captured step are kernels (a memset recorded into a graph
// is avoided).
__global__ void k_zero_u32(unsigned *__restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}
__global__ void k_zero_u64(unsigned long long *__restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

}  // namespace sps
