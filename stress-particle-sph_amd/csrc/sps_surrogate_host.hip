// sps_surrogate_host.hip - host side of the learned stress update: the
// network object, its training (forward, backward, Adam) and inference
// launch sequences, and its use inside a step.  Part of the sps_api.hip
// translation unit (it uses the handle, the launch macros and fail()).

struct sps_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    DBuf<float> w, m, v, g;            // master weights, Adam moments, gradient (MLP_NPARAM)
    DBuf<bf16_t> wb;                   // bfloat16 weights, same offsets
    int rows = 0;                      // batch rows the buffers hold (multiple of 64)
    DBuf<bf16_t> X, H1, H2, H3, dY, dZ;
    DBuf<float> Z1, Z2, Z3, Y, dH, T, loss;
    DBuf<float> feat, targ;            // a training set on the device
    DBuf<unsigned> perm;
    long long adam_t = 0;
};

static int mlp_reserve(sps_mlp *ml, int rows) {
    rows = (rows + 63) / 64 * 64;
    if (rows <= ml->rows) return 0;
    const size_t r = (size_t)rows;
    int rc = 0;
    rc |= dalloc(ml->X, r * MLP_IN); rc |= dalloc(ml->H1, r * MLP_H); rc |= dalloc(ml->H2, r * MLP_H);
    rc |= dalloc(ml->H3, r * MLP_H); rc |= dalloc(ml->dY, r * MLP_OUT); rc |= dalloc(ml->dZ, r * MLP_H);
    rc |= dalloc(ml->Z1, r * MLP_H); rc |= dalloc(ml->Z2, r * MLP_H); rc |= dalloc(ml->Z3, r * MLP_H);
    rc |= dalloc(ml->Y, r * MLP_OUT); rc |= dalloc(ml->dH, r * MLP_H); rc |= dalloc(ml->T, r * MLP_NOUT);
    rc |= dalloc(ml->loss, 1);
    if (rc) return rc;
    ml->rows = rows;
    return 0;
}

static Gemm gemm_of(const bf16_t *A, long long sa_i, long long sa_k, const bf16_t *B, long long sb_k,
                    long long sb_j, int M, int N, int K) {
    Gemm g{};
    g.A = A; g.sa_i = sa_i; g.sa_k = sa_k;
    g.B = B; g.sb_k = sb_k; g.sb_j = sb_j;
    g.M = M; g.N = N; g.K = K; g.kchunk = K;
    return g;
}

// Launches on `st`; through the handle's timed launches when h is given.
#define MLAUNCH(GRID, BLK, KERN, ...)                                  \
    do {                                                               \
        if (h)                                                         \
            LAUNCH_LDS(K_SURROGATE, GRID, BLK, 0, KERN, __VA_ARGS__);  \
        else                                                           \
            KERN<<<(GRID), (BLK), 0, st>>>(__VA_ARGS__);               \
    } while (0)

static void mlp_gemm(hipStream_t st, sps_handle *h, const Gemm &g, int zsplit) {
    const dim3 grid((unsigned)((g.M + 63) / 64), (unsigned)((g.N + 63) / 64), (unsigned)zsplit);
    MLAUNCH(grid, 256, k_gemm_bf16, g);
}

// Y = network(X) over `rows` rows (or the first *Mdev of them); Z kept for
// a backward pass when keepZ.
static void mlp_forward(sps_mlp *ml, hipStream_t st, sps_handle *h, int rows, const int *Mdev, bool keepZ) {
    Gemm g = gemm_of(ml->X.p, MLP_IN, 1, ml->wb.p + PW1, MLP_H, 1, rows, MLP_H, MLP_IN);
    g.Mdev = Mdev; g.bias = ml->w.p + PB1; g.Cf = keepZ ? ml->Z1.p : nullptr; g.ldc = MLP_H;
    g.Cb = ml->H1.p; g.ldcb = MLP_H; g.act = 1;
    mlp_gemm(st, h, g, 1);
    g = gemm_of(ml->H1.p, MLP_H, 1, ml->wb.p + PW2, MLP_H, 1, rows, MLP_H, MLP_H);
    g.Mdev = Mdev; g.bias = ml->w.p + PB2; g.Cf = keepZ ? ml->Z2.p : nullptr; g.ldc = MLP_H;
    g.Cb = ml->H2.p; g.ldcb = MLP_H; g.act = 1;
    mlp_gemm(st, h, g, 1);
    g = gemm_of(ml->H2.p, MLP_H, 1, ml->wb.p + PW3, MLP_H, 1, rows, MLP_H, MLP_H);
    g.Mdev = Mdev; g.bias = ml->w.p + PB3; g.Cf = keepZ ? ml->Z3.p : nullptr; g.ldc = MLP_H;
    g.Cb = ml->H3.p; g.ldcb = MLP_H; g.act = 1;
    mlp_gemm(st, h, g, 1);
    g = gemm_of(ml->H3.p, MLP_H, 1, ml->wb.p + PW4, MLP_OUT, 1, rows, MLP_OUT, MLP_H);
    g.Mdev = Mdev; g.bias = ml->w.p + PB4; g.Cf = ml->Y.p; g.ldc = MLP_OUT;
    mlp_gemm(st, h, g, 1);
}

// Gradient of the mean squared error of the first nb rows (of `rows`) with
// respect to every parameter, into ml->g (cleared first); the loss is added
// to ml->loss.
static int mlp_backward(sps_mlp *ml, hipStream_t st, int nb, int rows) {
    sps_handle *h = nullptr;
    HIP_TRY(hipMemsetAsync(ml->g.p, 0, MLP_NPARAM * sizeof(float), st));
    MLAUNCH((unsigned)((rows + 255) / 256), 256, k_mlp_loss, ml->Y.p, ml->T.p, nb, rows, ml->dY.p, ml->loss.p);
    const int kchunk = 256, zs = (rows + kchunk - 1) / kchunk;
    struct Layer {
        const bf16_t *Hin;   // layer input (rows x fin)
        int fin, fout, pw, pb;
        const float *Zin;    // pre-activation of the input (null for the first layer)
    };
    const Layer L[4] = {{ml->X.p, MLP_IN, MLP_H, PW1, PB1, nullptr},
                        {ml->H1.p, MLP_H, MLP_H, PW2, PB2, ml->Z1.p},
                        {ml->H2.p, MLP_H, MLP_H, PW3, PB3, ml->Z2.p},
                        {ml->H3.p, MLP_H, MLP_OUT, PW4, PB4, ml->Z3.p}};
    const bf16_t *dOut = ml->dY.p;   // gradient at the layer's output (rows x fout), bfloat16
    for (int l = 3; l >= 0; --l) {
        const Layer &Ly = L[l];
        // dW += Hin^T dOut (split over the rows)
        Gemm g = gemm_of(Ly.Hin, 1, Ly.fin, dOut, Ly.fout, 1, Ly.fin, Ly.fout, rows);
        g.kchunk = kchunk; g.Cf = ml->g.p + Ly.pw; g.ldc = Ly.fout; g.accumulate = 1;
        mlp_gemm(st, h, g, zs);
        MLAUNCH((unsigned)Ly.fout, 256, k_mlp_colsum, dOut, rows, Ly.fout, ml->g.p + Ly.pb);
        if (l == 0) break;
        // dHin = dOut W^T, then through the activation
        g = gemm_of(dOut, Ly.fout, 1, ml->wb.p + Ly.pw, 1, Ly.fout, rows, Ly.fin, Ly.fout);
        g.Cf = ml->dH.p; g.ldc = Ly.fin;
        mlp_gemm(st, h, g, 1);
        const long long n = (long long)rows * Ly.fin;
        MLAUNCH((unsigned)((n + 255) / 256), 256, k_mlp_silu_back, ml->dH.p, Ly.Zin, n, ml->dZ.p);
        dOut = ml->dZ.p;
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

static int mlp_adam(sps_mlp *ml, hipStream_t st, float lr) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    ++ml->adam_t;
    const float c1 = 1.0f - powf(b1, (float)ml->adam_t), c2 = 1.0f - powf(b2, (float)ml->adam_t);
    sps_handle *h = nullptr;
    MLAUNCH((unsigned)((MLP_NPARAM + 255) / 256), 256, k_mlp_adam, ml->w.p, ml->m.p, ml->v.p, ml->g.p, ml->wb.p,
            MLP_NPARAM, lr, b1, b2, eps, c1, c2);
    HIP_TRY(hipGetLastError());
    return 0;
}

static int mlp_upload_params(sps_mlp *ml, const float *w) {
    hipStream_t st = ml->stream;
    sps_handle *h = nullptr;
    HIP_TRY(hipMemcpyAsync(ml->w.p, w, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemsetAsync(ml->m.p, 0, MLP_NPARAM * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(ml->v.p, 0, MLP_NPARAM * sizeof(float), st));
    MLAUNCH((unsigned)((MLP_NPARAM + 255) / 256), 256, k_mlp_to_bf16, ml->w.p, ml->wb.p, MLP_NPARAM);
    HIP_TRY(hipStreamSynchronize(st));
    ml->adam_t = 0;
    return 0;
}

static uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
static double uniform01(uint64_t &s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

extern "C" int64_t sps_mlp_param_count(void) { return MLP_NPARAM; }

extern "C" int sps_mlp_create(int32_t batch_rows, uint64_t seed, sps_mlp **out) {
    if (!out || batch_rows < 1) return fail(SPS_EINVAL, "sps_mlp_create: null output or batch_rows < 1");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(SPS_EDEVICE, "no HIP device visible: the product library needs an MI355X (gfx950)");
    sps_mlp *ml = new (std::nothrow) sps_mlp();
    if (!ml) return fail(SPS_ENOMEM, "network allocation");
    (void)hipGetDevice(&ml->device);
    auto bail = [&](int code) {
        sps_mlp_destroy(ml);
        return code;
    };
    if (hipStreamCreateWithFlags(&ml->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(SPS_EDEVICE, "hipStreamCreate failed"));
    int rc = dalloc(ml->w, MLP_NPARAM) | dalloc(ml->m, MLP_NPARAM) | dalloc(ml->v, MLP_NPARAM) |
             dalloc(ml->g, MLP_NPARAM) | dalloc(ml->wb, MLP_NPARAM);
    if (rc || (rc = mlp_reserve(ml, batch_rows))) return bail(rc);
    // Glorot-uniform weights over the real (unpadded) fan-in / fan-out, zero biases
    std::vector<float> w(MLP_NPARAM, 0.0f);
    struct Blk {
        int off, fin, fout, rin, rout;
    };
    const Blk blocks[4] = {{PW1, MLP_IN, MLP_H, MLP_NIN, MLP_H}, {PW2, MLP_H, MLP_H, MLP_H, MLP_H},
                           {PW3, MLP_H, MLP_H, MLP_H, MLP_H}, {PW4, MLP_H, MLP_OUT, MLP_H, MLP_NOUT}};
    uint64_t s = seed;
    for (const Blk &b : blocks) {
        const double a = std::sqrt(6.0 / (b.rin + b.rout));
        for (int i = 0; i < b.rin; ++i)
            for (int j = 0; j < b.rout; ++j) w[b.off + i * b.fout + j] = (float)((2 * uniform01(s) - 1) * a);
    }
    if ((rc = mlp_upload_params(ml, w.data()))) return bail(rc);
    *out = ml;
    return 0;
}

extern "C" void sps_mlp_destroy(sps_mlp *ml) {
    if (!ml) return;
    (void)hipSetDevice(ml->device);
    if (ml->stream) (void)hipStreamSynchronize(ml->stream);
    dfree(ml->w); dfree(ml->m); dfree(ml->v); dfree(ml->g); dfree(ml->wb);
    dfree(ml->X); dfree(ml->H1); dfree(ml->H2); dfree(ml->H3); dfree(ml->dY); dfree(ml->dZ);
    dfree(ml->Z1); dfree(ml->Z2); dfree(ml->Z3); dfree(ml->Y); dfree(ml->dH); dfree(ml->T); dfree(ml->loss);
    dfree(ml->feat); dfree(ml->targ); dfree(ml->perm);
    if (ml->stream) (void)hipStreamDestroy(ml->stream);
    delete ml;
}

extern "C" int sps_mlp_get_params(sps_mlp *ml, float *out) {
    if (!ml || !out) return fail(SPS_EINVAL, "null argument");
    HIP_TRY(hipSetDevice(ml->device));
    HIP_TRY(hipMemcpyAsync(out, ml->w.p, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, ml->stream));
    HIP_TRY(hipStreamSynchronize(ml->stream));
    return 0;
}

extern "C" int sps_mlp_set_params(sps_mlp *ml, const float *in) {
    if (!ml || !in) return fail(SPS_EINVAL, "null argument");
    HIP_TRY(hipSetDevice(ml->device));
    return mlp_upload_params(ml, in);
}

// Upload n rows of features (n x 10) and, when given, targets (n x 4).
static int mlp_upload_set(sps_mlp *ml, int64_t n, const float *x, const float *t) {
    int rc = dalloc(ml->feat, (size_t)n * MLP_NIN);
    if (!rc && t) rc = dalloc(ml->targ, (size_t)n * MLP_NOUT);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(ml->feat.p, x, (size_t)n * MLP_NIN * sizeof(float), hipMemcpyHostToDevice, ml->stream));
    if (t)
        HIP_TRY(hipMemcpyAsync(ml->targ.p, t, (size_t)n * MLP_NOUT * sizeof(float), hipMemcpyHostToDevice,
                               ml->stream));
    return 0;
}

extern "C" int sps_mlp_forward(sps_mlp *ml, int64_t n, const float *x, float *y) {
    if (!ml || !x || !y || n < 0) return fail(SPS_EINVAL, "bad argument");
    HIP_TRY(hipSetDevice(ml->device));
    hipStream_t st = ml->stream;
    sps_handle *h = nullptr;
    int rc = mlp_upload_set(ml, n, x, nullptr);
    if (rc) return rc;
    std::vector<float> yb;
    for (int64_t off = 0; off < n; off += ml->rows) {
        const int nb = (int)std::min<int64_t>(ml->rows, n - off);
        const int rows = (nb + 63) / 64 * 64;
        MLAUNCH((unsigned)((rows + 255) / 256), 256, k_mlp_batch, ml->feat.p + off * MLP_NIN, nullptr, nullptr, nb,
                rows, ml->X.p, nullptr, nullptr);
        mlp_forward(ml, st, nullptr, rows, nullptr, false);
        yb.resize((size_t)nb * MLP_OUT);
        HIP_TRY(hipMemcpyAsync(yb.data(), ml->Y.p, yb.size() * sizeof(float), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        for (int r = 0; r < nb; ++r)
            for (int c = 0; c < MLP_NOUT; ++c) y[(off + r) * MLP_NOUT + c] = yb[(size_t)r * MLP_OUT + c];
    }
    return 0;
}

extern "C" int sps_mlp_gradient(sps_mlp *ml, int64_t n, const float *x, const float *t, float *grad, float *loss) {
    if (!ml || !x || !t || !grad || !loss || n < 1 || n > ml->rows)
        return fail(SPS_EINVAL, "sps_mlp_gradient: bad argument (n must be 1..batch rows)");
    HIP_TRY(hipSetDevice(ml->device));
    hipStream_t st = ml->stream;
    sps_handle *h = nullptr;
    int rc = mlp_upload_set(ml, n, x, t);
    if (rc) return rc;
    const int rows = (int)((n + 63) / 64 * 64);
    MLAUNCH((unsigned)((rows + 255) / 256), 256, k_mlp_batch, ml->feat.p, ml->targ.p, nullptr, (int)n, rows,
            ml->X.p, ml->T.p, nullptr);
    mlp_forward(ml, st, nullptr, rows, nullptr, true);
    HIP_TRY(hipMemsetAsync(ml->loss.p, 0, sizeof(float), st));
    if ((rc = mlp_backward(ml, st, (int)n, rows))) return rc;
    HIP_TRY(hipMemcpyAsync(grad, ml->g.p, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(loss, ml->loss.p, sizeof(float), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

extern "C" int sps_mlp_train(sps_mlp *ml, int64_t n, const float *x, const float *t, int32_t epochs, int32_t batch,
                             float lr, uint64_t seed, float *epoch_loss) {
    if (!ml || !x || !t || n < 1 || epochs < 1 || batch < 64 || batch % 64 != 0 || !(lr > 0))
        return fail(SPS_EINVAL, "sps_mlp_train: bad argument (batch a positive multiple of 64)");
    HIP_TRY(hipSetDevice(ml->device));
    hipStream_t st = ml->stream;
    sps_handle *h = nullptr;
    int rc = mlp_reserve(ml, batch);
    if (!rc) rc = mlp_upload_set(ml, n, x, t);
    if (!rc) rc = dalloc(ml->perm, (size_t)n);
    if (rc) return rc;
    std::vector<unsigned> perm((size_t)n);
    uint64_t s = seed;
    for (int e = 0; e < epochs; ++e) {
        for (int64_t i = 0; i < n; ++i) perm[i] = (unsigned)i;
        for (int64_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[splitmix64(s) % (uint64_t)(i + 1)]);
        HIP_TRY(hipMemcpyAsync(ml->perm.p, perm.data(), (size_t)n * sizeof(unsigned), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemsetAsync(ml->loss.p, 0, sizeof(float), st));
        int nbatch = 0;
        for (int64_t off = 0; off < n; off += batch, ++nbatch) {
            const int nb = (int)std::min<int64_t>(batch, n - off);
            const int rows = (nb + 63) / 64 * 64;
            MLAUNCH((unsigned)((rows + 255) / 256), 256, k_mlp_batch, ml->feat.p, ml->targ.p, ml->perm.p + off, nb,
                    rows, ml->X.p, ml->T.p, nullptr);
            mlp_forward(ml, st, nullptr, rows, nullptr, true);
            if ((rc = mlp_backward(ml, st, nb, rows))) return rc;
            if ((rc = mlp_adam(ml, st, lr))) return rc;
        }
        float l = 0.0f;
        HIP_TRY(hipMemcpyAsync(&l, ml->loss.p, sizeof(float), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));   // the host vector perm is rewritten next epoch
        if (epoch_loss) epoch_loss[e] = l / (float)nbatch;
    }
    return 0;
}

// The exact return map of n stresses (n x 4), for training labels.
extern "C" int sps_return_map(int64_t n, const double *in, const double *tanfi, const double *coh, double *out) {
    if (n < 0 || (n && (!in || !out || !tanfi || !coh))) return fail(SPS_EINVAL, "bad argument");
    if (n > (1ll << 31) - 1) return fail(SPS_EINVAL, "too many stresses");
    if (n == 0) return 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(SPS_EDEVICE, "no HIP device visible: the product library needs an MI355X (gfx950)");
    DBuf<double4> a, b;
    DBuf<double2> m;
    int rc = dalloc(a, (size_t)n) | dalloc(b, (size_t)n) | dalloc(m, (size_t)n);
    if (!rc) {
        hipStream_t st = nullptr;
        sps_handle *h = nullptr;
        std::vector<double2> mh((size_t)n);
        for (int64_t i = 0; i < n; ++i) mh[i] = make_double2(tanfi[i], coh[i]);
        if (hipMemcpy(a.p, in, (size_t)n * sizeof(double4), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(m.p, mh.data(), (size_t)n * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(SPS_EDEVICE, "upload failed");
        if (!rc) {
            MLAUNCH((unsigned)((n + 255) / 256), 256, k_return_map, a.p, m.p, b.p, (int)n);
            if (hipMemcpy(out, b.p, (size_t)n * sizeof(double4), hipMemcpyDeviceToHost) != hipSuccess)
                rc = fail(SPS_EDEVICE, "download failed");
        }
    }
    dfree(a);
    dfree(b);
    dfree(m);
    return rc;
}

// ------------------------------------------- the allocation-free return map
// sps_return_map_noalloc stages its inputs and outputs in one fixed device
// array per GPU instead of allocating: in (4 n doubles), out (4 n), tan(phi)
// (n), c (n), each region SPS_RETURN_MAP_NOALLOC_MAX rows long.  Host code
// never uses the array's name as an address: the address on each device
// comes from hipGetSymbolAddress, taken once per device with that device
// current and kept in rm_scratch_dev.
constexpr long long RM_ROWS = SPS_RETURN_MAP_NOALLOC_MAX;
static __device__ __align__(256) double scratch[10 * RM_ROWS];

constexpr int RM_MAX_DEVICES = 64;
static std::mutex rm_scratch_mu;
static double *rm_scratch_dev[RM_MAX_DEVICES] = {};

static int rm_scratch(int dev, double **out) {
    if (dev < 0 || dev >= RM_MAX_DEVICES) return fail(SPS_EDEVICE, "device %d beyond %d", dev, RM_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(rm_scratch_mu);
    if (!rm_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SPS_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        rm_scratch_dev[dev] = static_cast<double *>(p);
    }
    *out = rm_scratch_dev[dev];
    return 0;
}

// The return map of row i with its own material read from two arrays.
__global__ void k_return_map_rows(const double4 *__restrict__ in, const double *__restrict__ tanfi,
                                  const double *__restrict__ coh, double4 *__restrict__ out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 s = in[i];
    adapt_stress(s, tanfi[i], coh[i]);
    out[i] = s;
}

extern "C" int sps_return_map_noalloc(int64_t n, const double *in, const double *tanfi, const double *coh,
                                      double *out) {
    if (n < 0 || (n && (!in || !out || !tanfi || !coh))) return fail(SPS_EINVAL, "bad argument");
    if (n > RM_ROWS)
        return fail(SPS_EINVAL, "%lld stresses exceed SPS_RETURN_MAP_NOALLOC_MAX (%lld)", (long long)n, RM_ROWS);
    if (n == 0) return 0;
    int ndev = 0, dev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(SPS_EDEVICE, "no HIP device visible: the product library needs an MI355X (gfx950)");
    HIP_TRY(hipGetDevice(&dev));
    double *base = nullptr;
    const int rc = rm_scratch(dev, &base);
    if (rc) return rc;
    double4 *in_d = reinterpret_cast<double4 *>(base), *out_d = reinterpret_cast<double4 *>(base + 4 * RM_ROWS);
    double *tf_d = base + 8 * RM_ROWS, *c_d = base + 9 * RM_ROWS;
    // every row the kernel reads is written first; it writes every out row
    // the caller gets back, so nothing needs clearing between calls
    HIP_TRY(hipMemcpy(in_d, in, (size_t)n * sizeof(double4), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(tf_d, tanfi, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c_d, coh, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    k_return_map_rows<<<(unsigned)((n + 255) / 256), 256, 0, nullptr>>>(in_d, tf_d, c_d, out_d, (int)n);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(out, out_d, (size_t)n * sizeof(double4), hipMemcpyDeviceToHost));
    return 0;
}

// ------------------------------------------------------- use in the step
static SurrogateDev surrogate_dev(sps_handle *h) {
    SurrogateDev G{};
    G.count = h->surr_count.p;
    G.cap = h->mlp ? h->surr_cap : 0;
    G.slot = h->surr_slot.p;
    G.s_n = h->surr_sn.p;
    G.s_tr = h->surr_str.p;
    G.mat = h->surr_mat.p;
    G.feat = h->surr_feat.p;
    G.stress_scale = (float)h->surr_S;
    G.coh_scale = (float)h->surr_Sc;
    G.stats = h->surr_stats.p;
    G.log_count = h->surr_count.p ? h->surr_count.p + 1 : nullptr;
    G.log_cap = h->log_cap;
    G.log = h->plog.p;
    return G;
}

// After k_rk_final recorded the plastic trial states: the network's
// corrections, kept where consistent.
static int surrogate_correct(sps_handle *h, const StateBufs &SB) {
    sps_mlp *ml = h->mlp;
    hipStream_t st = h->stream;
    const int cap = h->surr_cap, rows = (cap + 63) / 64 * 64;
    const SurrogateDev G = surrogate_dev(h);
    MLAUNCH((unsigned)((rows + 255) / 256), 256, k_mlp_batch, h->surr_feat.p, nullptr, nullptr, cap, rows, ml->X.p,
            nullptr, h->surr_count.p);
    mlp_forward(ml, st, h, rows, (const int *)h->surr_count.p, false);
    MLAUNCH((unsigned)((cap + 255) / 256), 256, k_surrogate_apply, G, ml->Y.p, SB.stress, h->surr_tol);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int sps_set_surrogate(sps_handle *h, sps_mlp *ml, double stress_scale, double coh_scale, double tol) {
    if (!h) return fail(SPS_EINVAL, "null handle");
    if (ml && !h->mixed) return fail(SPS_EINVAL, "the learned stress update needs SPS_MODE_FAST");
    if (ml && !(stress_scale > 0 && coh_scale > 0 && tol >= 0))
        return fail(SPS_EINVAL, "scales must be > 0 and the tolerance >= 0");
    if (ml && ml->device != h->device) return fail(SPS_EINVAL, "network and handle on different devices");
    HIP_TRY(hipSetDevice(h->device));
    h->mlp = ml;
    drop_graphs(h);
    if (!ml) return 0;
    const int cap = h->ns > 0 ? h->ns : 1;
    int rc = mlp_reserve(ml, cap);
    rc |= dalloc(h->surr_count, 2) | dalloc(h->surr_slot, (size_t)cap) | dalloc(h->surr_sn, (size_t)cap) |
          dalloc(h->surr_str, (size_t)cap) | dalloc(h->surr_feat, (size_t)cap * MLP_NIN) | dalloc(h->surr_stats, 2) |
          dalloc(h->surr_mat, (size_t)cap);
    if (rc) return rc;
    h->surr_cap = cap;
    h->surr_S = stress_scale;
    h->surr_Sc = coh_scale;
    h->surr_tol = tol;
    HIP_TRY(hipMemsetAsync(h->surr_count.p, 0, 2 * sizeof(unsigned), h->stream));
    HIP_TRY(hipMemsetAsync(h->surr_stats.p, 0, 2 * sizeof(unsigned long long), h->stream));
    HIP_TRY(hipStreamSynchronize(ml->stream));   // the network's own uploads are done
    return 0;
}

extern "C" int sps_get_surrogate_stats(sps_handle *h, int64_t *evaluated, int64_t *kept) {
    if (!h || !evaluated || !kept) return fail(SPS_EINVAL, "null argument");
    unsigned long long v[2] = {0, 0};
    if (h->surr_stats.p) {
        HIP_TRY(hipMemcpyAsync(v, h->surr_stats.p, sizeof v, hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    *evaluated = (int64_t)v[0];
    *kept = (int64_t)v[1];
    return 0;
}

extern "C" int sps_set_plastic_log(sps_handle *h, int64_t cap) {
    if (!h || cap < 0 || cap > (1ll << 30)) return fail(SPS_EINVAL, "bad argument");
    if (cap > 0 && h->integrator == SPS_INTEGRATOR_LEAPFROG) return fail(SPS_EINVAL, "the log records RK4 steps");
    HIP_TRY(hipSetDevice(h->device));
    int rc = dalloc(h->surr_count, 2) | dalloc(h->plog, (size_t)(cap ? cap : 1) * SPS_PLOG_ROW);
    if (rc) return rc;
    h->log_cap = (int)cap;
    HIP_TRY(hipMemsetAsync(h->surr_count.p + 1, 0, sizeof(unsigned), h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" int sps_get_plastic_log(sps_handle *h, int64_t cap, double *out, int64_t *n) {
    if (!h || !n || cap < 0 || (cap && !out)) return fail(SPS_EINVAL, "bad argument");
    *n = 0;
    if (!h->surr_count.p || h->log_cap == 0) return 0;
    unsigned c = 0;
    HIP_TRY(hipMemcpyAsync(&c, h->surr_count.p + 1, sizeof c, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    const int64_t k = std::min<int64_t>(std::min<int64_t>((int64_t)c, h->log_cap), cap);
    if (k > 0) HIP_TRY(hipMemcpy(out, h->plog.p, (size_t)k * SPS_PLOG_ROW * sizeof(double), hipMemcpyDeviceToHost));
    *n = k;
    return 0;
}
