// sps_surrogate.hip - learned Drucker-Prager stress update (no counterpart in
// the reference).
//
// A multilayer perceptron 10 -> 128 -> 128 -> 128 -> 4 with SiLU activations
// maps (stress before the step, strain increment of the step expressed as the
// elastic stress increment D:de, cohesion, friction) to the correction that
// the exact return map (adapt_stress2, code/3_SPH_material_2018.f90:2066-2140)
// applies to the elastic trial stress.  Weights and activations are bfloat16,
// every product is accumulated in single precision by the matrix cores
// (v_mfma_f32_16x16x16_bf16); master weights, Adam moments and gradients are
// single precision.
//
// Features (all scaled by the stress scale S and the cohesion scale Sc given
// at set-up):
//   x[0:4] = sigma_n / S        x[4:8] = (sigma_trial - sigma_n) / S
//   x[8]   = c / Sc             x[9]   = tan(phi)
// Output y[0:4] = (R(sigma_trial) - sigma_trial) / F(sigma_trial), with F the
// yield function, positive at every plastic trial state: the correction per
// unit overstress, of order one, so that the network's relative error scales
// the correction, not the stress.
//
// In a step (SPS_MODE_FAST with a network attached) the final stress update
// of the stress points (k_rk_final) logs every plastic trial state; the
// network corrects them in one batch and each result is kept only if it is
// consistent, each residual at most the tolerance relative to the trial
// state's size (sqrt(J2) + |3 alpha p| + k_c): the yield function |F|, and
// the flow rule of the return map -- the mean stress unchanged (or at the
// apex beyond the tension cutoff) and the deviator parallel to the trial
// one.  Otherwise the exact return map's result stays.  Elastic trial states
// never reach the network (R is the identity there).
namespace sps {

typedef unsigned short bf16_t;
typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int MLP_NIN = SURR_NIN, MLP_NOUT = 4;
// smallest overstress F / (sqrt(J2) + |3 alpha p| + k_c) a state needs for the
// network to correct it (and for a training tuple)
constexpr double OVERSTRESS_MIN = 1e-9;
constexpr int MLP_IN = 16, MLP_H = 128, MLP_OUT = 16;   // padded to the 16-wide matrix tiles
// parameter blocks, weights as [in][out]
constexpr int PW1 = 0, PB1 = PW1 + MLP_IN * MLP_H, PW2 = PB1 + MLP_H, PB2 = PW2 + MLP_H * MLP_H,
              PW3 = PB2 + MLP_H, PB3 = PW3 + MLP_H * MLP_H, PW4 = PB3 + MLP_H, PB4 = PW4 + MLP_H * MLP_OUT,
              MLP_NPARAM = PB4 + MLP_OUT;

__device__ __forceinline__ bf16_t f2bf(float f) {   // round to nearest even
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t b) { return __uint_as_float((unsigned)b << 16); }
__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// C = A B for bfloat16 A (M x K) and B (K x N) with element strides, single
// precision accumulation.  A workgroup of four waves covers 64 x 64 of C,
// each wave 32 x 32 as 2 x 2 tiles of the 16 x 16 x 16 instruction; blockIdx.z
// takes K in chunks of kchunk (split K, accumulated with atomics).  M, N, K
// and kchunk are multiples of 16.  Lane l of a wave holds A[l % 16][4 (l / 16)
// + q] and B[4 (l / 16) + q][l % 16], q = 0..3, and receives
// C[4 (l / 16) + q][l % 16].
struct Gemm {
    const bf16_t *A;
    long long sa_i, sa_k;
    const bf16_t *B;
    long long sb_k, sb_j;
    int M, N, K, kchunk;
    const int *Mdev;         // when set: M = min(M, *Mdev rounded up to 16)
    const float *bias;       // per column, or null
    float *Cf;               // single-precision result (or split-K accumulator), or null
    long long ldc;
    bf16_t *Cb;              // bfloat16 of act(result), or null
    long long ldcb;
    int act;                 // 1: SiLU on the bfloat16 result
    int accumulate;          // 1: atomicAdd into Cf, no bias / act
};

__global__ __launch_bounds__(256) void k_gemm_bf16(Gemm g) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int M = g.M;
    if (g.Mdev) {
        const int m = (*g.Mdev + 15) / 16 * 16;
        M = m < M ? m : M;
    }
    const int r0 = blockIdx.x * 64 + 32 * (wave >> 1), c0 = blockIdx.y * 64 + 32 * (wave & 1);
    if (r0 >= M || c0 >= g.N) return;     // the whole wave: no barrier in this kernel
    const bool rv1 = r0 + 16 < M, cv1 = c0 + 16 < g.N;
    const int kb = blockIdx.z * g.kchunk;
    const int ke = kb + g.kchunk < g.K ? kb + g.kchunk : g.K;
    const int li = lane & 15, lk = 4 * (lane >> 4);
    v4f acc00 = {0.f, 0.f, 0.f, 0.f}, acc01 = acc00, acc10 = acc00, acc11 = acc00;
    for (int k = kb; k < ke; k += 16) {
        v4s a0, a1, b0, b1;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const long long kk = k + lk + q;
            a0[q] = (short)g.A[(r0 + li) * g.sa_i + kk * g.sa_k];
            a1[q] = rv1 ? (short)g.A[(r0 + 16 + li) * g.sa_i + kk * g.sa_k] : (short)0;
            b0[q] = (short)g.B[kk * g.sb_k + (c0 + li) * g.sb_j];
            b1[q] = cv1 ? (short)g.B[kk * g.sb_k + (c0 + 16 + li) * g.sb_j] : (short)0;
        }
        acc00 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b0, acc00, 0, 0, 0);
        acc01 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b1, acc01, 0, 0, 0);
        acc10 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b0, acc10, 0, 0, 0);
        acc11 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b1, acc11, 0, 0, 0);
    }
    auto put = [&](const v4f &acc, int ti, int tj) {
        if ((ti && !rv1) || (tj && !cv1)) return;
        const int col = c0 + 16 * tj + li;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = r0 + 16 * ti + lk + q;
            if (g.accumulate) {
                atomicAdd(&g.Cf[row * g.ldc + col], acc[q]);
                continue;
            }
            const float v = acc[q] + (g.bias ? g.bias[col] : 0.0f);
            if (g.Cf) g.Cf[row * g.ldc + col] = v;
            if (g.Cb) g.Cb[row * g.ldcb + col] = f2bf(g.act ? silu(v) : v);
        }
    };
    put(acc00, 0, 0);
    put(acc01, 0, 1);
    put(acc10, 1, 0);
    put(acc11, 1, 1);
}

// Rows idx[r] (or r) of the n x 10 single-precision features (and n x 4
// targets) into the padded bfloat16 batch X (rows >= nb: zero).
// With nbdev, nb is at most *nbdev and only rows below nb rounded up to 64
// are written (the rows a launch of k_gemm_bf16 with Mdev reads).
__global__ void k_mlp_batch(const float *__restrict__ feat, const float *__restrict__ targ,
                            const unsigned *__restrict__ idx, int nb, int rows, bf16_t *__restrict__ X,
                            float *__restrict__ T, const unsigned *__restrict__ nbdev) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (nbdev) {
        nb = min(nb, (int)*nbdev);
        rows = min(rows, (nb + 63) / 64 * 64);
    }
    if (r >= rows) return;
    const long long src = r < nb ? (idx ? (long long)idx[r] : (long long)r) : -1;
#pragma unroll
    for (int c = 0; c < MLP_IN; ++c)
        X[(long long)r * MLP_IN + c] = f2bf(src >= 0 && c < MLP_NIN ? feat[src * MLP_NIN + c] : 0.0f);
    if (T)
#pragma unroll
        for (int c = 0; c < MLP_NOUT; ++c) T[(long long)r * MLP_NOUT + c] = src >= 0 ? targ[src * MLP_NOUT + c] : 0.0f;
}

// Squared-error loss of the nb rows and dL/dY as bfloat16 (mean over rows
// and the 4 outputs; padding columns and rows get 0).
__global__ void k_mlp_loss(const float *__restrict__ Y, const float *__restrict__ T, int nb, int rows,
                           bf16_t *__restrict__ dY, float *__restrict__ loss) {
    __shared__ float s[256];
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (r < rows) {
        const float inv = 1.0f / (float)(nb * MLP_NOUT);
#pragma unroll
        for (int c = 0; c < MLP_OUT; ++c) {
            float d = 0.0f;
            if (r < nb && c < MLP_NOUT) {
                d = Y[(long long)r * MLP_OUT + c] - T[(long long)r * MLP_NOUT + c];
                l += d * d * inv;
            }
            dY[(long long)r * MLP_OUT + c] = f2bf(2.0f * d * inv);
        }
    }
    s[threadIdx.x] = l;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0 && loss) atomicAdd(loss, s[0]);
}

// dZ = dH * SiLU'(Z) as bfloat16 (rows x 128).
__global__ void k_mlp_silu_back(const float *__restrict__ dH, const float *__restrict__ Z, long long n,
                                bf16_t *__restrict__ dZ) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dZ[i] = f2bf(dH[i] * silu_grad(Z[i]));
}

// Column sums of a rows x ncol bfloat16 matrix added to out (bias gradients).
__global__ void k_mlp_colsum(const bf16_t *__restrict__ D, int rows, int ncol, float *__restrict__ out) {
    const int c = blockIdx.x;   // one block per column
    float s = 0.0f;
    for (int r = threadIdx.x; r < rows; r += blockDim.x) s += bf2f(D[(long long)r * ncol + c]);
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] += red[0];
}

// Adam (Kingma & Ba 2015) on the single-precision masters, then their
// bfloat16 copies.
__global__ void k_mlp_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                           const float *__restrict__ g, bf16_t *__restrict__ wb, int n, float lr, float b1,
                           float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    wb[i] = f2bf(wi);
}

__global__ void k_mlp_to_bf16(const float *__restrict__ w, bf16_t *__restrict__ wb, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wb[i] = f2bf(w[i]);
}

// ------------------------------------------------------- use in the step
// Drucker-Prager yield function of the reference's return map:
// F = sqrt(J2) + 3 alpha p - k_c (adapt_stress2, 3_...:2082-2134).
__device__ __forceinline__ double dp_yield(const double4 &s, double alpha2, double kc, double &sj2) {
    const double p = (s.x + s.y + s.w) / 3.0;
    const double d0 = s.x - p, d1 = s.y - p, d3 = s.w - p;
    sj2 = sqrt(s.z * s.z + 0.5 * (d0 * d0 + d1 * d1 + d3 * d3));
    return sj2 + 3 * alpha2 * p - kc;
}

// The network's corrections of the recorded states, kept where consistent.
__global__ void k_surrogate_apply(SurrogateDev G, const float *__restrict__ Y, double4 *__restrict__ stress,
                                  double tol) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = min((int)*G.count, G.cap);
    bool kept = false;
    if (k < n) {
        const double sc = G.stress_scale;
        const double4 t = G.s_tr[k];
        const double tanfi = G.mat[k].x, coh = G.mat[k].y;
        const float *y = Y + (long long)k * MLP_OUT;
        const double alpha2 = tanfi / (sqrt(9 + 12 * (tanfi * tanfi)));
        const double kc = (3 * coh) / (sqrt(9 + 12 * (tanfi * tanfi)));
        double sj2_t, sj2;
        const double ft = dp_yield(t, alpha2, kc, sj2_t);   // > 0: the state is plastic
        const double4 s = make_double4(t.x + ft * y[0], t.y + ft * y[1], t.z + ft * y[2], t.w + ft * y[3]);
        const double f = dp_yield(s, alpha2, kc, sj2);
        const double size = sj2_t + fabs(3 * alpha2 * (t.x + t.y + t.w) / 3.0) + kc;
        (void)sc;
        // consistency with the return map's flow rule: the returned deviator
        // is the trial one scaled down, at the trial mean stress, or at the
        // apex kc / (3 alpha) where the trial state is beyond the tension cutoff
        const double pt = (t.x + t.y + t.w) / 3.0, po = (s.x + s.y + s.w) / 3.0;
        const double p_exp = kc - 3 * alpha2 * pt < 0 ? kc / (3 * alpha2) : pt;
        const double a0 = t.x - pt, a1 = t.y - pt, a3 = t.w - pt;
        const double b0 = s.x - po, b1 = s.y - po, b3 = s.w - po;
        const double dot = t.z * s.z + 0.5 * (a0 * b0 + a1 * b1 + a3 * b3);   // the J2 inner product
        const double cross = sj2_t * sj2 - dot;                               // >= 0, 0 when parallel
        // left to the return map: an overstress at rounding level (it moved a few
        // bits), and sqrt(J2) <= 1e-5, where it zeroes the deviator whatever
        // the overstress (3_...:2122)
        kept = ft > OVERSTRESS_MIN * size && sj2_t > (double)10e-06f && fabs(f) <= tol * size &&
               fabs(po - p_exp) <= tol * size && cross <= tol * sj2_t * sj2_t;
        if (kept) stress[G.slot[k]] = s;
    }
    // per-wave counts, one atomic per wave
    const unsigned long long ev = __ballot(k < n), kp = __ballot(kept);
    if ((threadIdx.x & 63) == 0) {
        if (ev) atomicAdd(&G.stats[0], (unsigned long long)__popcll(ev));
        if (kp) atomicAdd(&G.stats[1], (unsigned long long)__popcll(kp));
    }
}

// The exact return map of n stresses, each with its own material (labels of
// training tuples).
__global__ void k_return_map(const double4 *__restrict__ in, const double2 *__restrict__ mat,
                             double4 *__restrict__ out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 s = in[i];
    const double2 m = mat[i];   // tan(phi), c
    adapt_stress(s, m.x, m.y);
    out[i] = s;
}

}  // namespace sps
