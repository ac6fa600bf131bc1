// GRMHD cons2prim on gfx950: the plain Noble 2D Newton solve, a learned
// initial guess (MLP 8 -> 64x4 -> 2, SiLU) trained on the GPU with
// hand-written forward / backward / Adam kernels, BF16-MFMA inference with
// FP32 accumulation, FP64 Newton polish and fallback to the plain guess.
//
// The solve is the reference's (Hydro/grmhd.c:813-1029, via phys_grmhd.h):
// same residuals, Jacobian, safeguards, |d eta / eta| < prec test, Nextra
// extra iterations and max_iter.  Only the starting point differs.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "disco_gpu.h"
#include "disco_c4_states.h"
#include "mlp_infer.h"
#include "phys_grmhd.h"

using namespace disco;
using namespace disco::mlp;

namespace {

thread_local std::string c2p_err;
int c2p_fail(const std::string& m) {
    c2p_err = m;
    return -1;
}
#define CHK(x)                                                                              \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) return c2p_fail(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define DEV __device__ static inline __attribute__((always_inline))

struct C2PConst {
    double gamma_law, M, a, prec;
    int max_iter, nextra;
    float fmean[NIN], fstd[NIN];  // feature standardisation
    float tmean[NOUT], tstd[NOUT];
};

DEV float silu(float z) { return z / (1.0f + __expf(-z)); }
DEV float dsilu(float z) {
    float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// Targets: t0 = ln W = -ln(1 - v2)/2, t1 = ln(eta / e).
DEV void c2p_targets(const grmhd::C2PInvariants& v, double v2, double eta, float* t) {
    t[0] = (float)(-0.5 * log1p(-v2));
    t[1] = (float)(log(eta) - log(v.e));
}

// Cold start for states without previous primitives: fluid at rest (v2 = 0)
// and eta from the energy residual at v2 = 0 with psi dropped:
// eta - n (eta - 1) + Q/2 = e.
DEV void cold_guess(double gamma_law, const grmhd::C2PInvariants& v, double* v20, double* eta0) {
    double n = (gamma_law - 1.0) / gamma_law;
    double eta = (v.e - n - 0.5 * v.Q) / (1.0 - n);
    *v20 = 0.0;
    *eta0 = fmax(eta, 1.0 + 1e-10);
}

// ------------------------------------------------------------------ generator
__global__ __launch_bounds__(256) void k_c4_generate(disco_c4 c, long k0, long n, double* cons, double* x) {
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double xs[3], prim[8], u[8];
    c4_sample(&c, (uint64_t)(k0 + k), xs, prim);
    grmhd::Metric g;
    grmhd::kerr_metric(xs, c.M, c.a, &g);
    grmhd::prim2cons(c.gamma_law, g, prim, u, xs, 1.0);
    for (int q = 0; q < 8; ++q) cons[q * n + k] = u[q];
    for (int d = 0; d < 3; ++d) x[d * n + k] = xs[d];
}

// ------------------------------------------------------------------ plain solve
// Per state: metric, invariants, Newton from the cold start (or, when
// guess != nullptr, from the reference's guess formula applied to the given
// previous primitives), recovery.  Optionally emits the training pair.
__global__ __launch_bounds__(256) void k_c2p_plain(C2PConst C, long n, const double* __restrict__ cons,
                                                   const double* __restrict__ x, const double* __restrict__ guess,
                                                   const float* __restrict__ mlp, double* prim, int* iters,
                                                   int* flags, float* feat, float* tgt) {
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double u[8], xs[3], p[8];
    for (int q = 0; q < 8; ++q) u[q] = cons[q * n + k];
    for (int d = 0; d < 3; ++d) xs[d] = x[d * n + k];
    grmhd::Metric g;
    grmhd::kerr_metric(xs, C.M, C.a, &g);
    grmhd::C2PInvariants v;
    grmhd::c2p_invariants(g, u, xs, &v);
    double v20, eta0;
    if (mlp) {
        // learned start: network outputs (standardised ln W, ln eta/e)
        float y0 = mlp[2 * k] * C.tstd[0] + C.tmean[0];
        float y1 = mlp[2 * k + 1] * C.tstd[1] + C.tmean[1];
        double W = exp((double)fmaxf(y0, 0.0f));
        v20 = 1.0 - 1.0 / (W * W);
        eta0 = fmax(v.e * exp((double)y1), 1.0 + 1e-10);
    } else if (guess) {
        double gp[8];
        for (int q = 0; q < 8; ++q) gp[q] = guess[q * n + k];
        grmhd::c2p_guess_from_prim(C.gamma_law, g, gp, xs, &v20, &eta0);
    } else {
        cold_guess(C.gamma_law, v, &v20, &eta0);
    }
    double v2, eta;
    int ok;
    int it = grmhd::c2p_newton(C.gamma_law, v, v20, eta0, C.prec, C.max_iter, C.nextra, &v2, &eta, &ok);
    int flag = ok ? 0 : 2;
    if (mlp && (!ok || !(v2 >= 0.0 && v2 < 1.0) || !(eta >= 1.0))) {
        // fallback: the plain (cold) start
        double c20, ce0;
        cold_guess(C.gamma_law, v, &c20, &ce0);
        it += grmhd::c2p_newton(C.gamma_law, v, c20, ce0, C.prec, C.max_iter, C.nextra, &v2, &eta, &ok);
        flag = ok ? 1 : 3;
    }
    grmhd::c2p_recover(C.gamma_law, 0.0, 0.0, g, v, v2, eta, xs, p);
    for (int q = 0; q < 8; ++q) prim[q * n + k] = p[q];
    if (iters) iters[k] = it;
    if (flags) flags[k] = flag;
    if (feat) {
        float f[NIN], t[NOUT];
        c2p_features(v, f);
        c2p_targets(v, v2, eta, t);
        for (int i = 0; i < NIN; ++i) feat[(long)i * n + k] = f[i];
        for (int i = 0; i < NOUT; ++i) tgt[(long)i * n + k] = ok ? t[i] : NAN;
    }
}

// ------------------------------------------------------------------ training
// One thread per sample, FP32.  z[l][s][j] holds pre-activations of hidden
// layer l (l = 0..3); y[s][o] the outputs; a0 = standardised inputs.
__global__ __launch_bounds__(256) void k_mlp_fwd(const float* __restrict__ W, const float* __restrict__ feat,
                                                 long nfeat, const int* __restrict__ idx, int B, float* z, float* y,
                                                 C2PConst C) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= B) return;
    long k = idx[s];
    float a[NH], zn[NH];
#pragma unroll
    for (int i = 0; i < NIN; ++i) a[i] = (feat[(long)i * nfeat + k] - C.fmean[i]) / C.fstd[i];
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        const int ni = l_in(l);
        const float* Wl = W + w_off(l);
        const float* bl = W + b_off(l);
#pragma unroll
        for (int j = 0; j < NH; ++j) {
            float acc = bl[j];
#pragma unroll
            for (int i = 0; i < ni; ++i) acc += a[i] * Wl[i * NH + j];
            zn[j] = acc;
            z[((long)l * B + s) * NH + j] = acc;
        }
#pragma unroll
        for (int j = 0; j < NH; ++j) a[j] = silu(zn[j]);
    }
#pragma unroll
    for (int o = 0; o < NOUT; ++o) {
        float acc = W[OFF_B5 + o];
#pragma unroll
        for (int i = 0; i < NH; ++i) acc += a[i] * W[OFF_W5 + i * NOUT + o];
        y[(long)s * NOUT + o] = acc;
    }
}

// dL/dy for L = mean over samples and outputs of (y - t_std)^2, then the
// deltas of the hidden layers: d[l][s][j] = dL/dz.
__global__ __launch_bounds__(256) void k_mlp_bwd(const float* __restrict__ W, const float* __restrict__ tgt,
                                                 long nfeat, const int* __restrict__ idx, int B,
                                                 const float* __restrict__ z, const float* __restrict__ y, float* d,
                                                 float* dy, float* loss, C2PConst C) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    float l2 = 0.0f;
    if (s < B) {
        long k = idx[s];
        float g[NH], gn[NH];
        float go[NOUT];
#pragma unroll
        for (int o = 0; o < NOUT; ++o) {
            float t = (tgt[(long)o * nfeat + k] - C.tmean[o]) / C.tstd[o];
            float r = y[(long)s * NOUT + o] - t;
            l2 += r * r;
            go[o] = 2.0f * r / (float)(B * NOUT);
            dy[(long)s * NOUT + o] = go[o];
        }
        // layer 4 (hidden 3): g = W5 go * silu'(z3)
#pragma unroll
        for (int i = 0; i < NH; ++i) {
            float acc = 0.0f;
#pragma unroll
            for (int o = 0; o < NOUT; ++o) acc += W[OFF_W5 + i * NOUT + o] * go[o];
            float zz = z[((long)3 * B + s) * NH + i];
            g[i] = acc * dsilu(zz);
            d[((long)3 * B + s) * NH + i] = g[i];
        }
#pragma unroll
        for (int l = 3; l >= 1; --l) {
            const float* Wl = W + w_off(l);
#pragma unroll
            for (int i = 0; i < NH; ++i) {
                float acc = 0.0f;
#pragma unroll
                for (int j = 0; j < NH; ++j) acc += Wl[i * NH + j] * g[j];
                float zz = z[((long)(l - 1) * B + s) * NH + i];
                gn[i] = acc * dsilu(zz);
                d[((long)(l - 1) * B + s) * NH + i] = gn[i];
            }
#pragma unroll
            for (int i = 0; i < NH; ++i) g[i] = gn[i];
        }
    }
    // block sum of the squared residual (monitoring only)
    for (int o = 32; o > 0; o >>= 1) l2 += __shfl_xor(l2, o);
    if ((threadIdx.x & 63) == 0 && l2 != 0.0f) atomicAdd(loss, l2 / (float)(B * NOUT));
}

// Weight gradients as batch reductions, one thread per parameter:
// dW_l[i][j] = sum_s a_{l-1}[s][i] d_l[s][j], db_l[j] = sum_s d_l[s][j].
__global__ __launch_bounds__(256) void k_mlp_wgrad(const float* __restrict__ feat, long nfeat,
                                                   const int* __restrict__ idx, int B, const float* __restrict__ z,
                                                   const float* __restrict__ d, const float* __restrict__ dy,
                                                   float* grad, C2PConst C) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM) return;
    int l = 4;
    while (l > 0 && p < w_off(l)) --l;
    int ni = l_in(l), no = l_out(l);
    int rel = p - w_off(l);
    bool is_b = rel >= ni * no;
    int i = is_b ? -1 : rel / no;
    int j = is_b ? rel - ni * no : rel % no;
    float acc = 0.0f;
    for (int s = 0; s < B; ++s) {
        float delta = (l == 4) ? dy[(long)s * NOUT + j] : d[((long)l * B + s) * NH + j];
        float a = 1.0f;
        if (!is_b) {
            if (l == 0)
                a = (feat[(long)i * nfeat + idx[s]] - C.fmean[i]) / C.fstd[i];
            else
                a = silu(z[((long)(l - 1) * B + s) * NH + i]);
        }
        acc += a * delta;
    }
    grad[p] = acc;
}

__global__ void k_adam(float* W, const float* grad, float* m, float* v, int t, float lr, float b1, float b2,
                       float eps) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM) return;
    float g = grad[p];
    float mm = b1 * m[p] + (1.0f - b1) * g;
    float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    float mh = mm / (1.0f - powf(b1, (float)t));
    float vh = vv / (1.0f - powf(b2, (float)t));
    W[p] -= lr * mh / (sqrtf(vh) + eps);
}

__global__ void k_zero_f(float* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// Network outputs only (standardised), for tests and training diagnostics.
__global__ __launch_bounds__(256) void k_mlp_guess(C2PConst C, long n, const double* __restrict__ cons,
                                                   const double* __restrict__ x,
                                                   const unsigned short* __restrict__ Wt_g,
                                                   const float* __restrict__ W32, float* mlp_out) {
    __shared__ MlpLds L;
    __shared__ unsigned short sF[4][64 * K0P];
    __shared__ float sY[4][64 * 2];
    mlp_load(L, Wt_g, W32);
    const int wave = threadIdx.x >> 6;
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float f[NIN];
    if (k < n) {
        double u[8], xs[3];
        for (int q = 0; q < 8; ++q) u[q] = cons[q * n + k];
        for (int d = 0; d < 3; ++d) xs[d] = x[d * n + k];
        grmhd::Metric g;
        grmhd::kerr_metric(xs, C.M, C.a, &g);
        grmhd::C2PInvariants v;
        grmhd::c2p_invariants(g, u, xs, &v);
        c2p_features(v, f);
    } else {
        for (int i = 0; i < NIN; ++i) f[i] = C.fmean[i];
    }
    for (int i = 0; i < NIN; ++i) f[i] = (f[i] - C.fmean[i]) / C.fstd[i];
    float y[2];
    mlp_wave(L, sF[wave], sY[wave], f, y);
    if (k < n) {
        mlp_out[2 * k] = y[0];
        mlp_out[2 * k + 1] = y[1];
    }
}

// Learned solve, fused: invariants once (FP64), network on MFMA, FP64 Newton
// polish from the network's (W, eta), the plain cold start as fallback.
__global__ __launch_bounds__(256) void k_c2p_learned(C2PConst C, long n, const double* __restrict__ cons,
                                                     const double* __restrict__ x,
                                                     const unsigned short* __restrict__ Wt_g,
                                                     const float* __restrict__ W32, double* prim, int* iters,
                                                     int* flags) {
    __shared__ MlpLds L;
    __shared__ unsigned short sF[4][64 * K0P];
    __shared__ float sY[4][64 * 2];
    mlp_load(L, Wt_g, W32);
    const int wave = threadIdx.x >> 6;
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = k < n;
    double u[8], xs[3];
    grmhd::Metric g;
    grmhd::C2PInvariants v;
    float f[NIN];
    if (valid) {
        for (int q = 0; q < 8; ++q) u[q] = cons[q * n + k];
        for (int d = 0; d < 3; ++d) xs[d] = x[d * n + k];
        grmhd::kerr_metric(xs, C.M, C.a, &g);
        grmhd::c2p_invariants(g, u, xs, &v);
        c2p_features(v, f);
    } else {
        for (int i = 0; i < NIN; ++i) f[i] = C.fmean[i];
    }
    for (int i = 0; i < NIN; ++i) f[i] = (f[i] - C.fmean[i]) / C.fstd[i];
    float y[2];
    mlp_wave(L, sF[wave], sY[wave], f, y);
    if (!valid) return;
    float y0 = y[0] * C.tstd[0] + C.tmean[0];
    float y1 = y[1] * C.tstd[1] + C.tmean[1];
    double W = exp((double)fmaxf(y0, 0.0f));
    double v20 = 1.0 - 1.0 / (W * W);
    double eta0 = fmax(v.e * exp((double)y1), 1.0 + 1e-10);
    double v2, eta;
    int ok;
    int it = grmhd::c2p_newton(C.gamma_law, v, v20, eta0, C.prec, C.max_iter, C.nextra, &v2, &eta, &ok);
    int flag = ok ? 0 : 2;
    if (!ok || !(v2 >= 0.0 && v2 < 1.0) || !(eta >= 1.0)) {
        double c20, ce0;
        cold_guess(C.gamma_law, v, &c20, &ce0);
        it += grmhd::c2p_newton(C.gamma_law, v, c20, ce0, C.prec, C.max_iter, C.nextra, &v2, &eta, &ok);
        flag = ok ? 1 : 3;
    }
    double p[8];
    grmhd::c2p_recover(C.gamma_law, 0.0, 0.0, g, v, v2, eta, xs, p);
    for (int q = 0; q < 8; ++q) prim[q * n + k] = p[q];
    if (iters) iters[k] = it;
    if (flags) flags[k] = flag;
}

// Error of recovered primitives against the generator's exact state:
// max over rho, P (relative) and the velocity components (on the state's
// velocity scale); the field components (copied from cons) are exact.
__global__ __launch_bounds__(256) void k_c2p_error(disco_c4 c, long k0, long n, const double* __restrict__ prim,
                                                   float* err) {
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double xs[3], p[8];
    c4_sample(&c, (uint64_t)(k0 + k), xs, p);
    double vs = fmax(fmax(fabs(p[URR]), fabs(p[UPP] * xs[0] * xs[0])), fabs(p[UZZ]));
    double e = fabs(prim[RHO * n + k] - p[RHO]) / p[RHO];
    e = fmax(e, fabs(prim[PPP * n + k] - p[PPP]) / p[PPP]);
    e = fmax(e, fabs(prim[URR * n + k] - p[URR]) / vs);
    e = fmax(e, fabs((prim[UPP * n + k] - p[UPP]) * xs[0] * xs[0]) / vs);
    e = fmax(e, fabs(prim[UZZ * n + k] - p[UZZ]) / vs);
    err[k] = isfinite(e) ? (float)e : INFINITY;
}

inline int nb(long n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

// ====================================================================== C ABI
struct disco_c2p {
    disco_c2p_params par;
    C2PConst C;
    long cap = 0, n = 0;
    double *cons = nullptr, *x = nullptr, *prim = nullptr, *guess = nullptr;
    float* mlpg = nullptr;  // per state: network outputs
    int *iters = nullptr, *flags = nullptr;
    float* err = nullptr;
    // MLP
    float *W = nullptr, *gradW = nullptr, *mW = nullptr, *vW = nullptr;
    unsigned short* Wt = nullptr;
    float *feat = nullptr, *tgt = nullptr;
    long nfeat = 0;
    float *z = nullptr, *d = nullptr, *y = nullptr, *dy = nullptr, *loss = nullptr;
    int* idx = nullptr;
    int bmax = 0;
    int adam_t = 0;
    bool trained = false;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr, em = nullptr;
};

static void c2p_const(disco_c2p* h) {
    C2PConst& C = h->C;
    C.gamma_law = h->par.gamma_law;
    C.M = h->par.M;
    C.a = h->par.a;
    C.prec = h->par.prec;
    C.max_iter = h->par.max_iter;
    C.nextra = h->par.nextra;
}

// Transposed, padded BF16 copy of the FP32 master weights for inference.
static int refresh_bf16(disco_c2p* h) {
    std::vector<float> W(NPARAM);
    CHK(hipMemcpy(W.data(), h->W, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    std::vector<unsigned short> T(WT_TOTAL, 0);
    mlp_pack_bf16(W.data(), T.data());
    CHK(hipMemcpy(h->Wt, T.data(), WT_TOTAL * sizeof(unsigned short), hipMemcpyHostToDevice));
    return 0;
}

extern "C" {

const char* disco_c2p_last_error(void) { return c2p_err.c_str(); }

int disco_c2p_create(const disco_c2p_params* p, long capacity, disco_c2p_t** out) {
    if (!p || !out || capacity <= 0) return c2p_fail("disco_c2p_create: bad arguments");
    *out = nullptr;
    disco_c2p* h = new disco_c2p();
    h->par = *p;
    h->cap = capacity;
    c2p_const(h);
    for (int i = 0; i < NIN; ++i) {
        h->C.fmean[i] = 0.f;
        h->C.fstd[i] = 1.f;
    }
    for (int i = 0; i < NOUT; ++i) {
        h->C.tmean[i] = 0.f;
        h->C.tstd[i] = 1.f;
    }
    size_t N = capacity;
    hipError_t e = hipSuccess;
    auto A = [&](void** ptr, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(ptr, bytes);
    };
    A((void**)&h->cons, 8 * N * sizeof(double));
    A((void**)&h->x, 3 * N * sizeof(double));
    A((void**)&h->prim, 8 * N * sizeof(double));
    A((void**)&h->iters, N * sizeof(int));
    A((void**)&h->flags, N * sizeof(int));
    A((void**)&h->err, N * sizeof(float));
    A((void**)&h->mlpg, 2 * N * sizeof(float));
    A((void**)&h->W, NPARAM * sizeof(float));
    A((void**)&h->gradW, NPARAM * sizeof(float));
    A((void**)&h->mW, NPARAM * sizeof(float));
    A((void**)&h->vW, NPARAM * sizeof(float));
    A((void**)&h->Wt, WT_TOTAL * sizeof(unsigned short));
    A((void**)&h->loss, sizeof(float));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&h->e0);
    if (e == hipSuccess) e = hipEventCreate(&h->e1);
    if (e == hipSuccess) e = hipEventCreate(&h->em);
    if (e != hipSuccess) {
        disco_c2p_destroy(h);
        return c2p_fail(std::string("disco_c2p_create: ") + hipGetErrorString(e));
    }
    // Xavier-style initial weights from a fixed seed
    std::vector<float> W(NPARAM, 0.f);
    uint64_t s = 0x1234567ull;
    auto rnd = [&]() {
        s = disco_splitmix64(s);
        return (float)(2.0 * disco_u01(s) - 1.0);
    };
    for (int l = 0; l < 5; ++l) {
        float sc = std::sqrt(6.0f / (l_in(l) + l_out(l)));
        for (int i = 0; i < l_in(l) * l_out(l); ++i) W[w_off(l) + i] = sc * rnd();
    }
    if (hipMemcpy(h->W, W.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(h->mW, 0, NPARAM * sizeof(float)) != hipSuccess ||
        hipMemset(h->vW, 0, NPARAM * sizeof(float)) != hipSuccess || refresh_bf16(h)) {
        disco_c2p_destroy(h);
        return c2p_fail("disco_c2p_create: initialising the MLP failed");
    }
    *out = h;
    return 0;
}

void disco_c2p_destroy(disco_c2p_t* h) {
    if (!h) return;
    void* ptrs[] = {h->cons, h->x, h->prim, h->guess, h->mlpg, h->iters, h->flags, h->err, h->W, h->gradW, h->mW, h->vW,
                    h->Wt,   h->feat, h->tgt, h->z, h->d, h->y, h->dy, h->loss, h->idx};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (h->e0) (void)hipEventDestroy(h->e0);
    if (h->e1) (void)hipEventDestroy(h->e1);
    if (h->em) (void)hipEventDestroy(h->em);
    if (h->st) (void)hipStreamDestroy(h->st);
    delete h;
}

int disco_c2p_generate(disco_c2p_t* h, unsigned long long seed, long k0, long n) {
    if (!h || n <= 0 || n > h->cap) return c2p_fail("disco_c2p_generate: bad arguments");
    disco_c4 c = {seed, h->par.M, h->par.a, h->par.gamma_law, h->par.rmin, h->par.rmax, h->par.theta_cut};
    hipLaunchKernelGGL(k_c4_generate, dim3(nb(n, 256)), dim3(256), 0, h->st, c, k0, n, h->cons, h->x);
    CHK(hipGetLastError());
    CHK(hipStreamSynchronize(h->st));
    h->n = n;
    return 0;
}

int disco_c2p_upload(disco_c2p_t* h, const double* cons, const double* x, long n) {
    if (!h || !cons || !x || n <= 0 || n > h->cap) return c2p_fail("disco_c2p_upload: bad arguments");
    CHK(hipMemcpy(h->cons, cons, 8 * n * sizeof(double), hipMemcpyHostToDevice));
    CHK(hipMemcpy(h->x, x, 3 * n * sizeof(double), hipMemcpyHostToDevice));
    h->n = n;
    return 0;
}

int disco_c2p_solve(disco_c2p_t* h, int mode, int nextra, double* ms) {
    if (!h || h->n <= 0) return c2p_fail("disco_c2p_solve: no states");
    if (mode == DISCO_C2P_LEARNED && !h->trained) return c2p_fail("disco_c2p_solve: the MLP is not trained");
    h->C.nextra = nextra;
    long n = h->n;
    CHK(hipEventRecord(h->e0, h->st));
    if (mode == DISCO_C2P_LEARNED) {
        hipLaunchKernelGGL(k_c2p_learned, dim3(nb(n, 256)), dim3(256), 0, h->st, h->C, n, h->cons, h->x, h->Wt,
                           h->W, h->prim, h->iters, h->flags);
        CHK(hipEventRecord(h->em, h->st));
        // network alone on the same states, timed separately
        hipLaunchKernelGGL(k_mlp_guess, dim3(nb(n, 256)), dim3(256), 0, h->st, h->C, n, h->cons, h->x, h->Wt, h->W,
                           h->mlpg);
    } else {
        hipLaunchKernelGGL(k_c2p_plain, dim3(nb(n, 256)), dim3(256), 0, h->st, h->C, n, h->cons, h->x,
                           (const double*)nullptr, (const float*)nullptr, h->prim, h->iters, h->flags,
                           (float*)nullptr, (float*)nullptr);
        CHK(hipEventRecord(h->em, h->st));
    }
    CHK(hipGetLastError());
    CHK(hipEventRecord(h->e1, h->st));
    CHK(hipEventSynchronize(h->e1));
    float f = 0.f, fg = 0.f;
    CHK(hipEventElapsedTime(&f, h->e0, h->em));
    CHK(hipEventElapsedTime(&fg, h->em, h->e1));
    if (ms) {
        ms[0] = f;
        ms[1] = fg;
    }
    h->C.nextra = h->par.nextra;
    return 0;
}

int disco_c2p_download(disco_c2p_t* h, double* prim, int* iters, int* flags, long n) {
    if (!h || n > h->n) return c2p_fail("disco_c2p_download: bad arguments");
    CHK(hipStreamSynchronize(h->st));
    if (prim)
        for (int q = 0; q < 8; ++q)
            CHK(hipMemcpy(prim + q * n, h->prim + q * h->n, n * sizeof(double), hipMemcpyDeviceToHost));
    if (iters) CHK(hipMemcpy(iters, h->iters, n * sizeof(int), hipMemcpyDeviceToHost));
    if (flags) CHK(hipMemcpy(flags, h->flags, n * sizeof(int), hipMemcpyDeviceToHost));
    return 0;
}

int disco_c2p_errors(disco_c2p_t* h, unsigned long long seed, long k0, float* err) {
    if (!h || h->n <= 0 || !err) return c2p_fail("disco_c2p_errors: bad arguments");
    disco_c4 c = {seed, h->par.M, h->par.a, h->par.gamma_law, h->par.rmin, h->par.rmax, h->par.theta_cut};
    hipLaunchKernelGGL(k_c2p_error, dim3(nb(h->n, 256)), dim3(256), 0, h->st, c, k0, h->n, h->prim, h->err);
    CHK(hipGetLastError());
    CHK(hipMemcpyAsync(err, h->err, h->n * sizeof(float), hipMemcpyDeviceToHost, h->st));
    CHK(hipStreamSynchronize(h->st));
    return 0;
}

}  // extern "C"

// Trains the network on the first ntrain resident states: (features,
// targets) pairs from the plain solve, standardisation, Adam over seeded
// random minibatches.
static int train_resident(disco_c2p* h, unsigned long long seed, long ntrain, int steps, int batch, float lr,
                          float* loss_curve) {
    // 1. training pairs: the plain solve on the resident states
    if (h->nfeat < ntrain) {
        if (h->feat) (void)hipFree(h->feat);
        if (h->tgt) (void)hipFree(h->tgt);
        CHK(hipMalloc((void**)&h->feat, NIN * ntrain * sizeof(float)));
        CHK(hipMalloc((void**)&h->tgt, NOUT * ntrain * sizeof(float)));
    }
    h->nfeat = ntrain;
    hipLaunchKernelGGL(k_c2p_plain, dim3(nb(ntrain, 256)), dim3(256), 0, h->st, h->C, ntrain, h->cons, h->x,
                       (const double*)nullptr, (const float*)nullptr, h->prim, h->iters, h->flags, h->feat, h->tgt);
    CHK(hipGetLastError());
    // 2. standardisation from the converged pairs (host)
    std::vector<float> F(NIN * ntrain), T(NOUT * ntrain);
    CHK(hipMemcpy(F.data(), h->feat, F.size() * sizeof(float), hipMemcpyDeviceToHost));
    CHK(hipMemcpy(T.data(), h->tgt, T.size() * sizeof(float), hipMemcpyDeviceToHost));
    std::vector<int> good;
    good.reserve(ntrain);
    for (long k = 0; k < ntrain; ++k)
        if (std::isfinite(T[k]) && std::isfinite(T[ntrain + k])) good.push_back((int)k);
    if (good.size() < (size_t)batch) return c2p_fail("disco_mlp_train: too few converged training states");
    auto stats = [&](const std::vector<float>& A, int nf, float* mean, float* sd) {
        for (int i = 0; i < nf; ++i) {
            double s1 = 0, s2 = 0;
            for (int k : good) {
                double v = A[(size_t)i * ntrain + k];
                s1 += v;
                s2 += v * v;
            }
            double m = s1 / good.size();
            mean[i] = (float)m;
            sd[i] = (float)std::sqrt(std::fmax(s2 / good.size() - m * m, 1e-12));
        }
    };
    stats(F, NIN, h->C.fmean, h->C.fstd);
    stats(T, NOUT, h->C.tmean, h->C.tstd);
    // 3. scratch for a batch
    if (h->bmax < batch) {
        void* ptrs[] = {h->z, h->d, h->y, h->dy, h->idx};
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
        CHK(hipMalloc((void**)&h->z, 4 * (size_t)batch * NH * sizeof(float)));
        CHK(hipMalloc((void**)&h->d, 4 * (size_t)batch * NH * sizeof(float)));
        CHK(hipMalloc((void**)&h->y, (size_t)batch * NOUT * sizeof(float)));
        CHK(hipMalloc((void**)&h->dy, (size_t)batch * NOUT * sizeof(float)));
        CHK(hipMalloc((void**)&h->idx, (size_t)batch * sizeof(int)));
        h->bmax = batch;
    }
    // 4. Adam over random minibatches (indices drawn on the host, seeded)
    uint64_t rs = seed ^ 0xA5A5A5A5ull;
    std::vector<int> idx(batch);
    for (int step = 0; step < steps; ++step) {
        for (int b = 0; b < batch; ++b) {
            rs = disco_splitmix64(rs);
            idx[b] = good[rs % good.size()];
        }
        CHK(hipMemcpyAsync(h->idx, idx.data(), batch * sizeof(int), hipMemcpyHostToDevice, h->st));
        hipLaunchKernelGGL(k_zero_f, dim3(1), dim3(64), 0, h->st, h->loss, 1);
        hipLaunchKernelGGL(k_mlp_fwd, dim3(nb(batch, 256)), dim3(256), 0, h->st, h->W, h->feat, ntrain, h->idx,
                           batch, h->z, h->y, h->C);
        hipLaunchKernelGGL(k_mlp_bwd, dim3(nb(batch, 256)), dim3(256), 0, h->st, h->W, h->tgt, ntrain, h->idx,
                           batch, h->z, h->y, h->d, h->dy, h->loss, h->C);
        hipLaunchKernelGGL(k_mlp_wgrad, dim3(nb(NPARAM, 256)), dim3(256), 0, h->st, h->feat, ntrain, h->idx,
                           batch, h->z, h->d, h->dy, h->gradW, h->C);
        h->adam_t += 1;
        hipLaunchKernelGGL(k_adam, dim3(nb(NPARAM, 256)), dim3(256), 0, h->st, h->W, h->gradW, h->mW, h->vW,
                           h->adam_t, lr, 0.9f, 0.999f, 1e-8f);
        CHK(hipGetLastError());
        if (loss_curve) CHK(hipMemcpyAsync(loss_curve + step, h->loss, sizeof(float), hipMemcpyDeviceToHost, h->st));
    }
    CHK(hipStreamSynchronize(h->st));
    h->trained = true;
    return refresh_bf16(h);
}

extern "C" {

int disco_mlp_train(disco_c2p_t* h, unsigned long long seed, long ntrain, int steps, int batch, float lr,
                    float* loss_curve) {
    if (!h || ntrain <= 0 || steps <= 0 || batch <= 0 || ntrain > h->cap)
        return c2p_fail("disco_mlp_train: bad arguments");
    if (disco_c2p_generate(h, seed, 0, ntrain)) return -1;
    return train_resident(h, seed, ntrain, steps, batch, lr, loss_curve);
}

int disco_mlp_train_resident(disco_c2p_t* h, unsigned long long seed, int steps, int batch, float lr,
                             float* loss_curve) {
    if (!h || h->n <= 0 || steps <= 0 || batch <= 0) return c2p_fail("disco_mlp_train_resident: bad arguments");
    return train_resident(h, seed, h->n, steps, batch, lr, loss_curve);
}

int disco_mlp_get(disco_c2p_t* h, float* params, float* norm) {
    if (!h) return c2p_fail("disco_mlp_get: null handle");
    if (params) CHK(hipMemcpy(params, h->W, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    if (norm) {
        std::memcpy(norm, h->C.fmean, NIN * sizeof(float));
        std::memcpy(norm + NIN, h->C.fstd, NIN * sizeof(float));
        std::memcpy(norm + 2 * NIN, h->C.tmean, NOUT * sizeof(float));
        std::memcpy(norm + 2 * NIN + NOUT, h->C.tstd, NOUT * sizeof(float));
    }
    return 0;
}

int disco_mlp_set(disco_c2p_t* h, const float* params, const float* norm) {
    if (!h) return c2p_fail("disco_mlp_set: null handle");
    if (params) CHK(hipMemcpy(h->W, params, NPARAM * sizeof(float), hipMemcpyHostToDevice));
    if (norm) {
        std::memcpy(h->C.fmean, norm, NIN * sizeof(float));
        std::memcpy(h->C.fstd, norm + NIN, NIN * sizeof(float));
        std::memcpy(h->C.tmean, norm + 2 * NIN, NOUT * sizeof(float));
        std::memcpy(h->C.tstd, norm + 2 * NIN + NOUT, NOUT * sizeof(float));
    }
    h->trained = true;
    return refresh_bf16(h);
}

int disco_mlp_grad(disco_c2p_t* h, const float* feat, const float* tgt, int n, float* grad, float* loss) {
    if (!h || !feat || !tgt || n <= 0) return c2p_fail("disco_mlp_grad: bad arguments");
    float *df = nullptr, *dt = nullptr, *z = nullptr, *d = nullptr, *y = nullptr, *dy = nullptr, *ls = nullptr,
          *g = nullptr;
    int* id = nullptr;
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    hipError_t e = hipSuccess;
    auto A = [&](void** p, size_t b) {
        if (e == hipSuccess) e = hipMalloc(p, b);
    };
    A((void**)&df, NIN * n * sizeof(float));
    A((void**)&dt, NOUT * n * sizeof(float));
    A((void**)&z, 4 * (size_t)n * NH * sizeof(float));
    A((void**)&d, 4 * (size_t)n * NH * sizeof(float));
    A((void**)&y, (size_t)n * NOUT * sizeof(float));
    A((void**)&dy, (size_t)n * NOUT * sizeof(float));
    A((void**)&ls, sizeof(float));
    A((void**)&g, NPARAM * sizeof(float));
    A((void**)&id, n * sizeof(int));
    int rc = 0;
    if (e == hipSuccess) {
        e = hipMemcpy(df, feat, NIN * n * sizeof(float), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dt, tgt, NOUT * n * sizeof(float), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(id, idx.data(), n * sizeof(int), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_zero_f, dim3(1), dim3(64), 0, h->st, ls, 1);
        hipLaunchKernelGGL(k_mlp_fwd, dim3(nb(n, 256)), dim3(256), 0, h->st, h->W, df, (long)n, id, n, z, y, h->C);
        hipLaunchKernelGGL(k_mlp_bwd, dim3(nb(n, 256)), dim3(256), 0, h->st, h->W, dt, (long)n, id, n, z, y, d, dy,
                           ls, h->C);
        hipLaunchKernelGGL(k_mlp_wgrad, dim3(nb(NPARAM, 256)), dim3(256), 0, h->st, df, (long)n, id, n, z, d, dy, g,
                           h->C);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(h->st);
        if (e == hipSuccess && grad) e = hipMemcpy(grad, g, NPARAM * sizeof(float), hipMemcpyDeviceToHost);
        if (e == hipSuccess && loss) e = hipMemcpy(loss, ls, sizeof(float), hipMemcpyDeviceToHost);
    }
    if (e != hipSuccess) rc = c2p_fail(std::string("disco_mlp_grad: ") + hipGetErrorString(e));
    void* ptrs[] = {df, dt, z, d, y, dy, ls, g, id};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    return rc;
}

int disco_mlp_infer(disco_c2p_t* h, float* out, long n) {
    if (!h || !out || h->n <= 0 || n > h->n) return c2p_fail("disco_mlp_infer: bad arguments");
    hipLaunchKernelGGL(k_mlp_guess, dim3(nb(h->n, 256)), dim3(256), 0, h->st, h->C, h->n, h->cons, h->x, h->Wt, h->W,
                       h->mlpg);
    CHK(hipGetLastError());
    CHK(hipStreamSynchronize(h->st));
    CHK(hipMemcpy(out, h->mlpg, 2 * n * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int disco_c2p_features(disco_c2p_t* h, float* feat, float* tgt, long n) {
    // features / plain-solve targets of the resident states (for tests)
    if (!h || h->n <= 0 || n > h->n) return c2p_fail("disco_c2p_features: bad arguments");
    float *df = nullptr, *dt = nullptr;
    CHK(hipMalloc((void**)&df, NIN * h->n * sizeof(float)));
    CHK(hipMalloc((void**)&dt, NOUT * h->n * sizeof(float)));
    hipLaunchKernelGGL(k_c2p_plain, dim3(nb(h->n, 256)), dim3(256), 0, h->st, h->C, h->n, h->cons, h->x,
                       (const double*)nullptr, (const float*)nullptr, h->prim, h->iters, h->flags, df, dt);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    for (int i = 0; i < NIN && e == hipSuccess; ++i)
        e = hipMemcpy(feat + i * n, df + i * h->n, n * sizeof(float), hipMemcpyDeviceToHost);
    for (int i = 0; i < NOUT && e == hipSuccess && tgt; ++i)
        e = hipMemcpy(tgt + i * n, dt + i * h->n, n * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(df);
    (void)hipFree(dt);
    if (e != hipSuccess) return c2p_fail(std::string("disco_c2p_features: ") + hipGetErrorString(e));
    return 0;
}

}  // extern "C"
