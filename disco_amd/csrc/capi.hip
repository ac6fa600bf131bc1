// C ABI (include/disco_gpu.h): device-resident batch of runs, graph-captured
// RK2 steps, upload/download and timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "checkpoint.h"
#include "disco_gpu.h"
#include "device_state.h"
#include "host_setup.h"
#include "mlp_infer.h"

namespace disco {
void launch_mindt(const StepShape& sh, const DevState& S, hipStream_t st, const LaunchHook* hook);
int launch_nf(const StepShape& sh);
void launch_dt_final(const DevState& S, int check, hipStream_t st, const LaunchHook* hook);
void launch_partners(const DevState& S, hipStream_t st);
void launch_report(const DevState& S, int a0, int nann, double* out, hipStream_t st);
void launch_sink_reset(const DevState& S, int run, hipStream_t st);
void launch_report_sum(const double* part, int Nr, int Nz, double* sums, hipStream_t st);
void launch_diag(const DevState& S, int nann, int nq, hipStream_t st, const LaunchHook* hook);
void launch_substep(const StepShape& sh, const DevState& S, double RK, double dtfac, int first, int want_dt,
                    hipStream_t st, const LaunchHook* hook, const Fork* fk);
}  // namespace disco

using namespace disco;

// last error message: a fixed buffer, so that setting it allocates nothing
static thread_local char g_err[512];

static int fail_c(const char* m) {
    std::snprintf(g_err, sizeof g_err, "%s", m);
    return -1;
}
static int fail(const std::string& m) { return fail_c(m.c_str()); }

// Device scratch of disco_gpu_report_noalloc: the per-annulus report sums of
// one run (slot j + Nr k, DISCO_NREPORT each) and, after them, their total.
// One copy per GPU for the life of the process; every call on that GPU
// shares it.  In host code `report_scratch` names a host-side stand-in, not
// device memory: host code uses only the address hipGetSymbolAddress gives
// for the current device (report_scratch_dev).
constexpr int DISCO_REPORT_MAX_ANN = 65536;
constexpr size_t DISCO_REPORT_SCRATCH = (size_t)DISCO_REPORT_MAX_ANN * DISCO_NREPORT + DISCO_NREPORT;
static __device__ __align__(256) double report_scratch[DISCO_REPORT_SCRATCH];
constexpr int DISCO_MAX_DEVICES = 64;
static double* report_scratch_dev[DISCO_MAX_DEVICES];

// Looks up this device's copy of the scratch once (with the device current);
// called by disco_gpu_create.
static int init_report_scratch() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= DISCO_MAX_DEVICES)
        return fail_c("report scratch: no current device");
    if (report_scratch_dev[dev]) return 0;
    double* p = nullptr;
    hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(report_scratch));
    if (e != hipSuccess || !p) return fail_c("report scratch: hipGetSymbolAddress failed");
    report_scratch_dev[dev] = p;
    return 0;
}

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) return fail(std::string(#x) + ": " + hipGetErrorString(e_));     \
    } while (0)

struct disco_gpu {
    std::vector<HostRun> runs;
    std::vector<RunDev> hruns;
    std::vector<AnnDev> hann;
    std::vector<int> run_cell0, run_ann0;
    std::vector<std::vector<int>> cell_perm;  // per run: device position of each cell (reference order)
    int ktile = 8;                            // z-layers per annulus tile in HBM
    int blk[KC_N] = {256, 256, 256, 256, 256, 256, 256, 256};  // threads per block per kernel class
    int nruns = 0, ncell = 0, nann = 0, nq = 5;
    int nf = 5;  // stored face-increment components per cell (the policies' NF: NUM_Q, or 5 with CT)
    AnnDev* ann = nullptr;
    int* cell_ann = nullptr;
    RunDev* druns = nullptr;
    StepShape shape{};
    double *prim = nullptr, *cons = nullptr, *rkcons = nullptr, *gradp = nullptr, *gradr = nullptr,
           *gradz = nullptr;
    double *piph[2] = {nullptr, nullptr}, *dphi[2] = {nullptr, nullptr}, *wiph = nullptr;
    int *lbo = nullptr, *ubi = nullptr, *lbu = nullptr, *ubd = nullptr, *pdg = nullptr;
    int* rot[2] = {nullptr, nullptr};
    double *fphi = nullptr, *fout = nullptr, *fin = nullptr, *fup = nullptr, *fdn = nullptr;
    // constrained transport (MHD)
    double *Phi = nullptr, *rkPhi = nullptr, *emf_phi = nullptr, *emf_t = nullptr, *esum = nullptr,
           *efin = nullptr;
    int* fflags = nullptr;
    int* err = nullptr;
    unsigned long long* dtbits = nullptr;
    double *dt = nullptr, *planet_phi = nullptr, *t = nullptr;
    int* final_step = nullptr;  // per run: check_dt's final flag
    std::vector<int> hbcjobs;   // zero-gradient ghost-fill jobs (StepShape::bc_off)
    int* bcjobs = nullptr;
    size_t bcjobs_cap = 0;
    double *diag = nullptr, *t_avg = nullptr;
    double* src = nullptr;  // GRMHD source increments
    // accretion sinks (DevState::sinkc ...; null without them)
    double *sinkc = nullptr, *sink_ann = nullptr, *sink_acc = nullptr, *sink_rk = nullptr;
    int parity = 0;
    hipStream_t stream = nullptr;
    Fork fork{nullptr, {nullptr, nullptr, nullptr}};  // second stream of the captured steps
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    // GRMHD in-loop cons2prim: learned start and counters
    unsigned short* mlp_wt = nullptr;
    float* mlp_w = nullptr;
    float mlp_norm[20] = {};
    bool learned = false;
    unsigned long long* c2p_stats = nullptr;
    bool counting = false;
};

// The kernels' view of the state with the mesh of buffer `cur` at the start
// of the substep; a substep that moves the mesh writes it into buffer `nxt`
// (nxt == cur: the mesh stays where it is).
static DevState make_state(disco_gpu* h, int cur, int nxt) {
    DevState S{};
    S.ann = h->ann;
    S.cell_ann = h->cell_ann;
    S.runs = h->druns;
    S.ncell = h->ncell;
    S.nruns = h->nruns;
    S.nann = h->nann;
    S.prim = h->prim;
    S.cons = h->cons;
    S.rkcons = h->rkcons;
    S.gradp = h->gradp;
    S.gradr = h->gradr;
    S.gradz = h->gradz;
    S.piph = h->piph[cur];
    S.dphi = h->dphi[cur];
    S.piph_n = h->piph[nxt];
    S.dphi_n = h->dphi[nxt];
    S.rot = h->rot[cur];
    S.rot_n = h->rot[nxt];
    S.wiph = h->wiph;
    S.bcjobs = h->bcjobs;
    S.lbo = h->lbo;
    S.ubi = h->ubi;
    S.lbu = h->lbu;
    S.ubd = h->ubd;
    S.pdg = h->pdg;
    S.fphi = h->fphi;
    S.fout = h->fout;
    S.fin = h->fin;
    S.fup = h->fup;
    S.fdn = h->fdn;
    S.dtbits = h->dtbits;
    S.dt = h->dt;
    S.planet_phi = h->planet_phi;
    S.t = h->t;
    S.final_step = h->final_step;
    S.src = h->src;
    S.diag = h->diag;
    S.t_avg = h->t_avg;
    S.sinkc = h->sinkc;
    S.sink_ann = h->sink_ann;
    S.sink_acc = h->sink_acc;
    S.sink_rk = h->sink_rk;
    S.Phi = h->Phi;
    S.rkPhi = h->rkPhi;
    S.emf_phi = h->emf_phi;
    S.emf_t = h->emf_t;
    S.esum = h->esum;
    S.efin = h->efin;
    S.fflags = h->fflags;
    S.err = h->err;
    if (h->learned) {
        S.mlp_wt = h->mlp_wt;
        S.mlp_w = h->mlp_w;
        std::memcpy(S.mlp_norm, h->mlp_norm, sizeof(S.mlp_norm));
    }
    S.c2p_stats = h->counting ? h->c2p_stats : nullptr;
    return S;
}

// One step for the current parity (timestep.c:7-54): [dt from the running
// CFL minimum] + the substeps of the run's Timestep, then add_diagnostics.
//   Timestep 0, 2 (and any other value but 1, 3): onestep(0, dt, first) +
//     onestep(1/2, dt/2, last), RK2 (timestep.c:45-48);
//   Timestep 3: onestep(0, dt, first) + onestep(3/4, dt/4) +
//     onestep(1/3, 2dt/3, last), the three-stage SSP Runge-Kutta (:36-43);
//   Timestep 1: onestep(0, dt, first, last) and then RK2 as above: timestep.c
//     tests 3 with an if of its own, so its else-branch runs after the Euler
//     substep (:29-48).  The mesh moves twice and t advances by 2 dt.
// The RK copy of the conserved state and face fluxes (timestep.c:18-24) is
// taken by the step's first substep (first == 1); the later substeps that
// move the mesh pass first == 2.  The mesh moves from buffer `parity` to the
// other one in a moving substep; step_parity_flip says whether a step ends
// with it in the other buffer.
// Invariant between steps: dtbits holds the CFL minimum over the current
// state (the last substep's update computes it; create/upload seed it).
// fk: run the independent branches of each substep on two streams (graph
// capture); null: every kernel in order on the handle's stream.
static void enqueue_step(disco_gpu* h, bool with_dt, const LaunchHook* hook, const Fork* fk = nullptr) {
    const int p = h->parity, q = 1 - p;
    const StepShape& sh = h->shape;
    DevState S0 = make_state(h, p, q);
    if (with_dt) launch_dt_final(S0, 1, h->stream, hook);
    launch_substep(sh, S0, 0.0, 1.0, 1, 0, h->stream, hook, fk);
    if (sh.stepper == 3) {
        DevState S1 = make_state(h, q, q);
        launch_substep(sh, S1, 0.75, 0.25, 0, 0, h->stream, hook, fk);
        launch_substep(sh, S1, 1.0 / 3.0, 2.0 / 3.0, 0, 1, h->stream, hook, fk);
        launch_diag(S1, h->nann, h->nq, h->stream, hook);
    } else if (sh.stepper == 1) {
        DevState S1 = make_state(h, q, p);
        launch_substep(sh, S1, 0.0, 1.0, 2, 0, h->stream, hook, fk);
        DevState S2 = make_state(h, p, p);
        launch_substep(sh, S2, 0.5, 0.5, 0, 1, h->stream, hook, fk);
        launch_diag(S2, h->nann, h->nq, h->stream, hook);
    } else {
        DevState S1 = make_state(h, q, q);
        launch_substep(sh, S1, 0.5, 0.5, 0, 1, h->stream, hook, fk);
        launch_diag(S1, h->nann, h->nq, h->stream, hook);
    }
}

// 1 when a step leaves the mesh in the other buffer (one moving substep), 0
// for Timestep 1 (two)
static int step_parity_flip(const disco_gpu* h) { return h->shape.stepper == 1 ? 0 : 1; }

// Recomputes the overlap-face partners and the CFL minimum over the current
// state (after create/upload: the steps search partners only after the mesh
// moved).
static int seed_dt(disco_gpu* h) {
    DevState S = make_state(h, h->parity, 1 - h->parity);
    launch_partners(S, h->stream);
    launch_mindt(h->shape, S, h->stream, nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

template <typename T>
static int dalloc(T** p, size_t n) {
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) return fail(std::string("hipMalloc: ") + hipGetErrorString(e));
    return 0;
}

static void build_tables(disco_gpu* h) {
    const int learned = h->shape.c2p_learned;
    build_annulus_tables(h->runs, &h->hann, &h->hruns, &h->run_cell0, &h->run_ann0, &h->shape, &h->cell_perm,
                         h->ktile, &h->hbcjobs);
    h->shape.c2p_learned = learned;
    for (int k = 0; k < KC_N; ++k) h->shape.blk[k] = h->blk[k];
    h->ncell = 0;
    for (const HostRun& R : h->runs) h->ncell += R.ncell;
    h->nann = (int)h->hann.size();
}

// Host arrays are in the reference's (jk, i) order, the device holds each
// run's annuli in tile order: copy nf fields of run r through the permutation.
static int put_fields(disco_gpu* h, int r, double* dev, const double* host, int nf) {
    const HostRun& R = h->runs[r];
    const std::vector<int>& perm = h->cell_perm[r];
    size_t c0 = h->run_cell0[r], N = h->ncell, n = R.ncell;
    std::vector<double> tmp(n);
    for (int f = 0; f < nf; ++f) {
        const double* src = host + (size_t)f * n;
        for (size_t c = 0; c < n; ++c) tmp[perm[c]] = src[c];
        HIPCHK(hipMemcpy(dev + f * N + c0, tmp.data(), n * sizeof(double), hipMemcpyHostToDevice));
    }
    return 0;
}

static int get_fields(disco_gpu* h, int r, double* host, const double* dev, int nf) {
    const HostRun& R = h->runs[r];
    const std::vector<int>& perm = h->cell_perm[r];
    size_t c0 = h->run_cell0[r], N = h->ncell, n = R.ncell;
    std::vector<double> tmp(n);
    for (int f = 0; f < nf; ++f) {
        HIPCHK(hipMemcpy(tmp.data(), dev + f * N + c0, n * sizeof(double), hipMemcpyDeviceToHost));
        double* dst = host + (size_t)f * n;
        for (size_t c = 0; c < n; ++c) dst[c] = tmp[perm[c]];
    }
    return 0;
}

// rotation points in device annulus order
static int put_rot(disco_gpu* h, int r, int* dev) {
    const HostRun& R = h->runs[r];
    int a0 = h->run_ann0[r], na = R.Nr * R.Nz;
    std::vector<int> tmp(na);
    for (int a = 0; a < na; ++a) tmp[a] = R.rot[h->hann[a0 + a].j + R.Nr * h->hann[a0 + a].k];
    HIPCHK(hipMemcpy(dev + a0, tmp.data(), na * sizeof(int), hipMemcpyHostToDevice));
    return 0;
}

static int upload_run_state(disco_gpu* h, int r) {
    const HostRun& R = h->runs[r];
    if (put_fields(h, r, h->prim, R.prim.data(), h->nq) || put_fields(h, r, h->cons, R.cons.data(), h->nq) ||
        put_fields(h, r, h->piph[h->parity], R.piph.data(), 1) ||
        put_fields(h, r, h->dphi[h->parity], R.dphi.data(), 1) || put_fields(h, r, h->wiph, R.wiph.data(), 1) ||
        put_rot(h, r, h->rot[h->parity]))
        return -1;
    if (h->Phi && !R.Phi.empty() && put_fields(h, r, h->Phi, R.Phi.data(), 5)) return -1;
    double pp[DISCO_MAX_PLANETS] = {0.0, 0.0};
    for (size_t p = 0; p < R.planets.size() && p < DISCO_MAX_PLANETS; ++p) pp[p] = R.planets[p].phi;
    HIPCHK(hipMemcpy(h->planet_phi + r * DISCO_MAX_PLANETS, pp, sizeof(pp), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->t + r, &R.t, sizeof(double), hipMemcpyHostToDevice));
    const int not_final = 0;
    HIPCHK(hipMemcpy(h->final_step + r, &not_final, sizeof(int), hipMemcpyHostToDevice));
    return 0;
}

static int capture_graphs(disco_gpu* h) {
    for (int p = 0; p < 2; ++p) {
        if (h->graph[p]) {
            (void)hipGraphExecDestroy(h->graph[p]);
            h->graph[p] = nullptr;
        }
    }
    int saved = h->parity;
    for (int p = 0; p < 2; ++p) {
        h->parity = p;
        hipGraph_t g;
        HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        enqueue_step(h, true, nullptr, &h->fork);
        HIPCHK(hipStreamEndCapture(h->stream, &g));
        HIPCHK(hipGraphInstantiate(&h->graph[p], g, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(g));
    }
    h->parity = saved;
    return 0;
}

// ---------------------------------------------------------------- layout
// Tables from build_tables to the device (create and relayout).
static int upload_tables(disco_gpu* h) {
    std::vector<int> cell_ann(h->ncell);
    for (int a = 0; a < h->nann; ++a)
        for (int i = 0; i < h->hann[a].np; ++i) cell_ann[h->hann[a].off + i] = a;
    HIPCHK(hipMemcpy(h->ann, h->hann.data(), h->nann * sizeof(AnnDev), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->cell_ann, cell_ann.data(), h->ncell * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->druns, h->hruns.data(), h->nruns * sizeof(RunDev), hipMemcpyHostToDevice));
    if (!h->hbcjobs.empty()) {
        if (h->bcjobs_cap < h->hbcjobs.size()) {
            if (h->bcjobs) HIPCHK(hipFree(h->bcjobs));
            h->bcjobs = nullptr;
            HIPCHK(hipMalloc((void**)&h->bcjobs, h->hbcjobs.size() * sizeof(int)));
            h->bcjobs_cap = h->hbcjobs.size();
        }
        HIPCHK(hipMemcpy(h->bcjobs, h->hbcjobs.data(), h->hbcjobs.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    return 0;
}

// The array index of each annulus' smallest front (the rotation point).
static void host_rotation(HostRun& R) {
    for (int jk = 0; jk < R.Nr * R.Nz; ++jk) {
        int o = R.off[jk], np = R.Np[jk];
        R.rot[jk] = 0;
        for (int i = 0; i < np; ++i) {
            int im = i == 0 ? np - 1 : i - 1;
            if (R.piph[o + i] < R.piph[o + im]) R.rot[jk] = i;
        }
    }
}

// Every device buffer indexed by cell or by annulus: (pointer, records per
// cell/annulus, bytes per record).  A relayout moves all of them, so no state
// of any kind -- carried, scratch or hint -- depends on the layout.
struct DevBuf {
    void* p;
    int fields;  // field-major planes of N records
    int rec;     // bytes per record
    bool per_annulus;
};

static std::vector<DevBuf> layout_buffers(disco_gpu* h) {
    std::vector<DevBuf> b;
    const int nq = h->nq;
    auto D = [&](double* p, int f) {
        if (p) b.push_back({p, f, 8, false});
    };
    auto I = [&](int* p, int f) {
        if (p) b.push_back({p, f, 4, false});
    };
    D(h->prim, nq), D(h->cons, nq), D(h->rkcons, nq), D(h->gradp, nq), D(h->gradr, nq), D(h->gradz, nq);
    D(h->piph[0], 1), D(h->piph[1], 1), D(h->dphi[0], 1), D(h->dphi[1], 1), D(h->wiph, 1);
    D(h->fphi, h->nf), D(h->fout, h->nf), D(h->fin, h->nf), D(h->fup, h->nf), D(h->fdn, h->nf);
    D(h->Phi, 5), D(h->rkPhi, 5), D(h->emf_phi, 4), D(h->emf_t, 12), D(h->esum, 8), D(h->efin, 4);
    D(h->src, 4);
    I(h->lbo, 1), I(h->ubi, 1), I(h->lbu, 1), I(h->ubd, 1), I(h->pdg, 4), I(h->fflags, 1);
    b.push_back({h->rot[0], 1, 4, true});
    b.push_back({h->rot[1], 1, 4, true});
    b.push_back({h->diag, 1, 8 * DISCO_NDIAG, true});
    return b;
}

// dst[to[c]] = src[c] for n records of `words` words of type W
template <typename W>
__global__ void k_move(W* __restrict__ dst, const W* __restrict__ src, const int* __restrict__ to, int n, int words) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)n * words) return;
    long c = i / words, b = i - c * words;
    dst[(long)to[c] * words + b] = src[i];
}

// Moves every buffer from the current annulus tile to `ktile`, on the device.
static int relayout(disco_gpu* h, int ktile) {
    if (ktile == h->ktile) return 0;
    HIPCHK(hipStreamSynchronize(h->stream));
    const std::vector<std::vector<int>> old_perm = h->cell_perm;
    const std::vector<AnnDev> old_ann = h->hann;
    h->ktile = ktile;
    build_tables(h);
    // old device position -> new device position, for cells and annuli
    std::vector<int> to_c(h->ncell), to_a(h->nann);
    for (int r = 0; r < h->nruns; ++r) {
        const HostRun& R = h->runs[r];
        const int c0 = h->run_cell0[r], a0 = h->run_ann0[r];
        for (int c = 0; c < R.ncell; ++c) to_c[c0 + old_perm[r][c]] = c0 + h->cell_perm[r][c];
        std::vector<int> pos(R.Nr * R.Nz);
        for (int a = 0; a < R.Nr * R.Nz; ++a) pos[h->hann[a0 + a].j + R.Nr * h->hann[a0 + a].k] = a0 + a;
        for (int a = 0; a < R.Nr * R.Nz; ++a) to_a[a0 + a] = pos[old_ann[a0 + a].j + R.Nr * old_ann[a0 + a].k];
    }
    int *dto_c = nullptr, *dto_a = nullptr;
    char* tmp = nullptr;
    size_t tmp_bytes = std::max((size_t)h->ncell * 8, (size_t)h->nann * 8 * DISCO_NDIAG);
    hipError_t e = hipMalloc((void**)&dto_c, h->ncell * sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void**)&dto_a, h->nann * sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void**)&tmp, tmp_bytes);
    if (e == hipSuccess) e = hipMemcpy(dto_c, to_c.data(), h->ncell * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dto_a, to_a.data(), h->nann * sizeof(int), hipMemcpyHostToDevice);
    for (const DevBuf& b : layout_buffers(h)) {
        const int n = b.per_annulus ? h->nann : h->ncell;
        const size_t plane = (size_t)n * b.rec;
        for (int f = 0; f < b.fields && e == hipSuccess; ++f) {
            char* p = (char*)b.p + f * plane;
            const int* to = b.per_annulus ? dto_a : dto_c;
            if (b.rec % 8 == 0) {
                const int w = b.rec / 8;
                hipLaunchKernelGGL(k_move<unsigned long long>, dim3((unsigned)(((long)n * w + 255) / 256)), dim3(256),
                                   0, h->stream, (unsigned long long*)tmp, (const unsigned long long*)p, to, n, w);
            } else {
                const int w = b.rec / 4;
                hipLaunchKernelGGL(k_move<unsigned int>, dim3((unsigned)(((long)n * w + 255) / 256)), dim3(256), 0,
                                   h->stream, (unsigned int*)tmp, (const unsigned int*)p, to, n, w);
            }
            e = hipGetLastError();
            if (e == hipSuccess) e = hipMemcpyAsync(p, tmp, plane, hipMemcpyDeviceToDevice, h->stream);
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(dto_c);
    (void)hipFree(dto_a);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(std::string("relayout: ") + hipGetErrorString(e));
    return upload_tables(h);
}

// Annulus tile (z-layers per tile in HBM) and threads per block of the cell
// kernels.  Neither changes any result: every kernel gathers per cell, the
// CFL reduction is a minimum and the diagnostics sum one annulus per block,
// whatever order the cells are stored in and however they are grouped.
static int set_layout(disco_gpu* h, int ktile, const int* blk) {
    if (relayout(h, ktile)) return -1;
    for (int k = 0; k < KC_N; ++k) h->shape.blk[k] = h->blk[k] = blk[k];
    return capture_graphs(h);
}

// Kernel class of a kernel name (the launch names of hydro_step.hip).
static int kernel_class(const std::string& n) {
    static const std::pair<const char*, int> m[] = {
        {"k_stage_a", KC_STAGE}, {"k_plm_trans", KC_PLM}, {"k_flux", KC_FLUX},     {"k_emf_sum", KC_EMF},
        {"k_emf_face", KC_EMF},  {"k_ct_update", KC_CT},
        {"k_update_a", KC_CT},   {"k_source", KC_CT},   {"k_flip_flag", KC_FLIP}, {"k_flip_chain", KC_FLIP}, {"k_update_b", KC_UPDATE},
        {"k_update_b_learned", KC_UPDATE}, {"k_update", KC_UPDATE}, {"k_zcopy", KC_COPY}, {"k_bc_copy", KC_COPY}};
    for (auto& e : m)
        if (n == e.first) return e.second;
    return -1;
}

// A device copy of everything a step reads or writes.
struct Snapshot {
    std::vector<std::pair<void*, void*>> bufs;  // (live, copy)
    std::vector<size_t> bytes;
    ~Snapshot() {
        for (auto& b : bufs) (void)hipFree(b.second);
    }
};

static int snapshot(disco_gpu* h, Snapshot* S, bool restore) {
    HIPCHK(hipStreamSynchronize(h->stream));
    if (!restore) {
        std::vector<std::pair<void*, size_t>> list;
        for (const DevBuf& b : layout_buffers(h))
            list.push_back({b.p, (size_t)b.fields * (b.per_annulus ? h->nann : h->ncell) * b.rec});
        const size_t n = h->nruns;
        list.push_back({h->dtbits, n * sizeof(unsigned long long)});
        list.push_back({h->dt, n * sizeof(double)});
        list.push_back({h->t, n * sizeof(double)});
        list.push_back({h->final_step, n * sizeof(int)});
        list.push_back({h->t_avg, n * sizeof(double)});
        list.push_back({h->planet_phi, n * DISCO_MAX_PLANETS * sizeof(double)});
        if (h->c2p_stats) list.push_back({h->c2p_stats, 4 * sizeof(unsigned long long)});
        if (h->sink_acc) {
            list.push_back({h->sink_acc, n * DISCO_NSINKQ * DISCO_MAX_PLANETS * sizeof(double)});
            list.push_back({h->sink_rk, n * DISCO_NSINKQ * DISCO_MAX_PLANETS * sizeof(double)});
        }
        for (auto& l : list) {
            void* c = nullptr;
            HIPCHK(hipMalloc(&c, l.second));
            S->bufs.push_back({l.first, c});
            S->bytes.push_back(l.second);
            HIPCHK(hipMemcpy(c, l.first, l.second, hipMemcpyDeviceToDevice));
        }
    } else {
        for (size_t i = 0; i < S->bufs.size(); ++i)
            HIPCHK(hipMemcpy(S->bufs[i].first, S->bufs[i].second, S->bytes[i], hipMemcpyDeviceToDevice));
    }
    return 0;
}

extern "C" {

const char* disco_gpu_last_error(void) { return g_err; }

int disco_parse_par(const char* text, disco_params* out, char* errbuf, int errlen) {
    std::string err;
    int e = read_par_text(text ? text : "", out, &err);
    if (errbuf && errlen > 0) {
        std::snprintf(errbuf, errlen, "%s", err.c_str());
    }
    return e;
}

int disco_read_par(const char* path, disco_params* out, char* errbuf, int errlen) {
    std::ifstream f(path);
    if (!f) {
        if (errbuf && errlen > 0) std::snprintf(errbuf, errlen, "cannot open %s", path);
        return -1;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    return disco_parse_par(ss.str().c_str(), out, errbuf, errlen);
}

void disco_gpu_destroy(disco_gpu_t* h) {
    if (!h) return;
    for (int p = 0; p < 2; ++p)
        if (h->graph[p]) (void)hipGraphExecDestroy(h->graph[p]);
    void* ptrs[] = {h->ann,   h->cell_ann, h->druns, h->prim,   h->cons,    h->rkcons,     h->gradp,
                    h->gradr, h->gradz,    h->piph[0], h->piph[1], h->dphi[0], h->dphi[1], h->wiph,
                    h->lbo,   h->ubi,      h->lbu,   h->ubd,    h->rot[0],  h->rot[1],     h->fphi,
                    h->fout,  h->fin,      h->fup,   h->fdn,    h->dtbits,  h->dt,         h->planet_phi,
                    h->t,     h->Phi,      h->rkPhi, h->emf_phi, h->emf_t,  h->esum,
                    h->efin,  h->fflags,   h->err,   h->pdg,   h->mlp_wt,  h->mlp_w,      h->c2p_stats, h->diag,
                    h->t_avg, h->src,      h->final_step, h->bcjobs, h->sinkc, h->sink_ann, h->sink_acc,
                    h->sink_rk};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    if (h->fork.st2) (void)hipStreamDestroy(h->fork.st2);
    for (hipEvent_t e : h->fork.ev)
        if (e) (void)hipEventDestroy(e);
    delete h;
}

int disco_gpu_create(const disco_params* pars, int nruns, disco_gpu_t** out) {
    if (!out || !pars || nruns <= 0) return fail("disco_gpu_create: bad arguments");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail("no HIP device available");
    disco_gpu* h = new disco_gpu();
    h->nruns = nruns;
    if (const char* kt = std::getenv("DISCO_ANNULUS_TILE")) h->ktile = std::atoi(kt);
    h->runs.resize(nruns);
    for (int r = 0; r < nruns; ++r) {
        std::string err;
        if (setup_run(pars[r], &h->runs[r], &err)) {
            delete h;
            return fail("run " + std::to_string(r) + ": " + err);
        }
        if (h->runs[r].par.hydro != h->runs[0].par.hydro) {
            delete h;
            return fail("all runs of a batch must use the same Hydro module");
        }
        if (h->runs[r].par.num_n != h->runs[0].par.num_n) {
            delete h;
            return fail("all runs of a batch must have the same Num_N");
        }
        if (stepper_of(h->runs[r].par.Timestep) != stepper_of(h->runs[0].par.Timestep)) {
            delete h;
            return fail("all runs of a batch must use the same Timestep scheme");
        }
    }
    h->nq = h->runs[0].nq;
    // constrained-transport buffers: ideal MHD and GRMHD
    const int hyd = h->runs[0].par.hydro;
    const bool grmhd = hyd == DISCO_HYDRO_GRMHD || hyd == DISCO_HYDRO_GRMHD2;
    const bool mhd = hyd == DISCO_HYDRO_MHD || grmhd;
    // with constrained transport the field's face increments are never
    // stored (the cell field is rebuilt from the face fluxes): 5 components
    // and the passive scalars'; otherwise every conserved field's
    h->nf = mhd ? 5 + h->runs[0].par.num_n : h->nq;
    build_tables(h);
    // the kernels that will run store exactly this many components per cell
    if (launch_nf(h->shape) != h->nf) {
        delete h;
        return fail("internal: the face-increment buffers do not match the flux kernel's components");
    }
    size_t N = h->ncell;
    size_t Q = (size_t)h->nq * N;
    int rc = 0;
    rc |= dalloc(&h->ann, h->nann);
    rc |= dalloc(&h->cell_ann, N);
    rc |= dalloc(&h->druns, nruns);
    rc |= dalloc(&h->prim, Q);
    rc |= dalloc(&h->cons, Q);
    rc |= dalloc(&h->rkcons, Q);
    rc |= dalloc(&h->gradp, Q);
    rc |= dalloc(&h->gradr, Q);
    rc |= dalloc(&h->gradz, Q);
    rc |= dalloc(&h->piph[0], N);
    rc |= dalloc(&h->piph[1], N);
    rc |= dalloc(&h->dphi[0], N);
    rc |= dalloc(&h->dphi[1], N);
    rc |= dalloc(&h->wiph, N);
    rc |= dalloc(&h->lbo, N);
    rc |= dalloc(&h->ubi, N);
    rc |= dalloc(&h->lbu, N);
    rc |= dalloc(&h->ubd, N);
    rc |= dalloc(&h->rot[0], h->nann);
    rc |= dalloc(&h->rot[1], h->nann);
    rc |= dalloc(&h->err, 1);
    if (mhd) {
        rc |= dalloc(&h->Phi, 5 * N);
        rc |= dalloc(&h->rkPhi, 5 * N);
        rc |= dalloc(&h->emf_phi, 4 * N);
        rc |= dalloc(&h->emf_t, 12 * N);
        rc |= dalloc(&h->esum, 8 * N);
        rc |= dalloc(&h->efin, 4 * N);
        rc |= dalloc(&h->fflags, N);
        rc |= dalloc(&h->pdg, 4 * N);
    }
    const size_t QF = (size_t)h->nf * N;  // stored face-increment components (NF)
    rc |= dalloc(&h->fphi, QF);
    rc |= dalloc(&h->fout, QF);
    rc |= dalloc(&h->fin, QF);
    rc |= dalloc(&h->fup, QF);
    rc |= dalloc(&h->fdn, QF);
    rc |= dalloc(&h->dtbits, nruns);
    rc |= dalloc(&h->dt, nruns);
    rc |= dalloc(&h->planet_phi, (size_t)nruns * DISCO_MAX_PLANETS);
    rc |= dalloc(&h->t, nruns);
    rc |= dalloc(&h->final_step, nruns);
    rc |= dalloc(&h->diag, (size_t)h->nann * DISCO_NDIAG);
    rc |= dalloc(&h->t_avg, nruns);
    if (grmhd) rc |= dalloc(&h->src, 4 * N);
    constexpr size_t SQ = (size_t)DISCO_NSINKQ * DISCO_MAX_PLANETS;
    if (h->shape.accrete) {
        rc |= dalloc(&h->sinkc, SQ * N);
        rc |= dalloc(&h->sink_ann, SQ * h->nann);
        rc |= dalloc(&h->sink_acc, SQ * nruns);
        rc |= dalloc(&h->sink_rk, SQ * nruns);
    }
    if (rc) {
        std::string m = g_err;
        disco_gpu_destroy(h);
        return fail(m);
    }
    auto cleanup = [&](const std::string& m) {
        disco_gpu_destroy(h);
        return fail(m);
    };
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return cleanup("stream");
    if (hipStreamCreateWithFlags(&h->fork.st2, hipStreamNonBlocking) != hipSuccess) return cleanup("stream");
    for (hipEvent_t& e : h->fork.ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return cleanup("event");
    if (init_report_scratch()) return cleanup(g_err);
    if (upload_tables(h)) return cleanup(g_err);
    // zero the scratch fields once (eager, outside any graph)
    double* zq[] = {h->rkcons, h->gradp, h->gradr, h->gradz};
    for (double* p : zq)
        if (hipMemset(p, 0, Q * sizeof(double)) != hipSuccess) return cleanup("hipMemset");
    double* zf[] = {h->fphi, h->fout, h->fin, h->fup, h->fdn};
    for (double* p : zf)
        if (hipMemset(p, 0, QF * sizeof(double)) != hipSuccess) return cleanup("hipMemset");
    int* zi[] = {h->lbo, h->ubi, h->lbu, h->ubd};
    for (int* p : zi)
        if (hipMemset(p, 0, N * sizeof(int)) != hipSuccess) return cleanup("hipMemset");
    if (hipMemset(h->dt, 0, nruns * sizeof(double)) != hipSuccess) return cleanup("hipMemset");
    if (hipMemset(h->diag, 0, (size_t)h->nann * DISCO_NDIAG * sizeof(double)) != hipSuccess ||
        hipMemset(h->t_avg, 0, nruns * sizeof(double)) != hipSuccess)
        return cleanup("hipMemset");
    if (hipMemset(h->err, 0, sizeof(int)) != hipSuccess) return cleanup("hipMemset");
    if (h->shape.accrete && (hipMemset(h->sink_acc, 0, SQ * nruns * sizeof(double)) != hipSuccess ||
                             hipMemset(h->sink_rk, 0, SQ * nruns * sizeof(double)) != hipSuccess))
        return cleanup("hipMemset");
    if (mhd) {
        struct {
            double* p;
            size_t n;
        } zc[] = {{h->Phi, 5 * N},  {h->rkPhi, 5 * N}, {h->emf_phi, 4 * N}, {h->emf_t, 12 * N},
                  {h->esum, 8 * N}, {h->efin, 4 * N}};
        for (auto& z : zc)
            if (hipMemset(z.p, 0, z.n * sizeof(double)) != hipSuccess) return cleanup("hipMemset");
        if (hipMemset(h->fflags, 0, N * sizeof(int)) != hipSuccess) return cleanup("hipMemset");
        if (hipMemset(h->pdg, 0, 4 * N * sizeof(int)) != hipSuccess) return cleanup("hipMemset");
    }
    for (int r = 0; r < nruns; ++r) {
        if (upload_run_state(h, r)) return cleanup(g_err);
        // the other mesh buffer starts equal (substep 2 reads it)
        const HostRun& R = h->runs[r];
        if (put_fields(h, r, h->piph[1], R.piph.data(), 1) || put_fields(h, r, h->dphi[1], R.dphi.data(), 1) ||
            put_rot(h, r, h->rot[1]))
            return cleanup("upload");
    }
    if (capture_graphs(h)) return cleanup(g_err);
    if (seed_dt(h)) return cleanup(g_err);
    if (hipDeviceSynchronize() != hipSuccess) return cleanup("sync after create");
    *out = h;
    return 0;
}

int disco_setup_run(const disco_params* par, int* info, int* Np, double* r_jph, double* z_kph, double* prim,
                    double* cons, double* piph, double* dphi, double* wiph, double* planet_phi) {
    if (!par || !info) return fail("disco_setup_run: bad arguments");
    HostRun R;
    std::string err;
    if (setup_run(*par, &R, &err)) return fail(err);
    int v[12] = {R.Nr, R.Nz, R.NgRa, R.NgRb, R.NgZa, R.NgZb, R.ncell, (int)R.nreal, R.nq, (int)R.planets.size(), 1,
                 R.ncell};
    std::memcpy(info, v, sizeof v);
    if (Np) std::memcpy(Np, R.Np.data(), R.Np.size() * sizeof(int));
    if (r_jph) std::memcpy(r_jph, R.r_jph.data(), R.r_jph.size() * sizeof(double));
    if (z_kph) std::memcpy(z_kph, R.z_kph.data(), R.z_kph.size() * sizeof(double));
    if (prim) std::memcpy(prim, R.prim.data(), R.prim.size() * sizeof(double));
    if (cons) std::memcpy(cons, R.cons.data(), R.cons.size() * sizeof(double));
    if (piph) std::memcpy(piph, R.piph.data(), R.piph.size() * sizeof(double));
    if (dphi) std::memcpy(dphi, R.dphi.data(), R.dphi.size() * sizeof(double));
    if (wiph) std::memcpy(wiph, R.wiph.data(), R.wiph.size() * sizeof(double));
    if (planet_phi)
        for (size_t p = 0; p < R.planets.size(); ++p) planet_phi[p] = R.planets[p].phi;
    return 0;
}

int disco_gpu_info(disco_gpu_t* h, int run, int* info) {
    if (!h || run < 0 || run >= h->nruns || !info) return fail("disco_gpu_info: bad arguments");
    const HostRun& R = h->runs[run];
    int v[12] = {R.Nr, R.Nz, R.NgRa, R.NgRb, R.NgZa, R.NgZb, R.ncell, (int)R.nreal, R.nq, (int)R.planets.size(),
                 h->nruns, h->ncell};
    std::memcpy(info, v, sizeof v);
    return 0;
}

int disco_gpu_grid(disco_gpu_t* h, int run, int* Np, double* r_jph, double* z_kph) {
    if (!h || run < 0 || run >= h->nruns) return fail("disco_gpu_grid: bad arguments");
    const HostRun& R = h->runs[run];
    if (Np) std::memcpy(Np, R.Np.data(), R.Np.size() * sizeof(int));
    if (r_jph) std::memcpy(r_jph, R.r_jph.data(), R.r_jph.size() * sizeof(double));
    if (z_kph) std::memcpy(z_kph, R.z_kph.data(), R.z_kph.size() * sizeof(double));
    return 0;
}

int disco_gpu_mindt(disco_gpu_t* h, double* dt_out) {
    if (!h) return fail("disco_gpu_mindt: null handle");
    DevState S = make_state(h, h->parity, 1 - h->parity);
    launch_mindt(h->shape, S, h->stream, nullptr);
    launch_dt_final(S, 0, h->stream, nullptr);
    HIPCHK(hipGetLastError());
    if (dt_out) {
        HIPCHK(hipMemcpyAsync(dt_out, h->dt, h->nruns * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int disco_gpu_step(disco_gpu_t* h, int nsteps) {
    if (!h || nsteps < 0) return fail("disco_gpu_step: bad arguments");
    for (int n = 0; n < nsteps; ++n) {
        HIPCHK(hipGraphLaunch(h->graph[h->parity], h->stream));
        h->parity ^= step_parity_flip(h);
    }
    return 0;
}

static int check_err(disco_gpu* h);

// main.c:68-75 over the batch: getmindt + check_dt + timestep until every
// run has taken its step to T_End (check_dt's clip runs on the device, at
// the start of every step).  Steps go out in runs of graph launches sized
// from the slowest run's remaining (T_End - t) / dt, so the host reads the
// final flags a few times per run rather than every step; a run that is
// final holds (dt = 0) while the others finish.  poll_files: check_dt's
// file controls in the working directory before every step: "abort" makes
// that step the last (status bit 2), "latest" is removed and the call
// returns before the step (status bit 4: the caller writes its checkpoint,
// as possiblyOutput does before timestep).  Status bit 1: every run is final.
int disco_gpu_advance(disco_gpu_t* h, int max_steps, int poll_files, int* steps_done, int* status) {
    if (!h || max_steps < 0) return fail("disco_gpu_advance: bad arguments");
    const int n = h->nruns;
    std::vector<double> t(n), dt(n);
    std::vector<int> fin(n);
    int done = 0, st = 0;
    bool stop = false;
    while (!stop) {
        HIPCHK(hipMemcpyAsync(fin.data(), h->final_step, n * sizeof(int), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(t.data(), h->t, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(dt.data(), h->dt, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        if (check_err(h)) return -1;
        // steps every unfinished run can take before its clipped one
        long ahead = -1;
        for (int r = 0; r < n; ++r) {
            if (fin[r]) continue;
            double left = h->runs[r].par.t_max - t[r];
            long k = (dt[r] > 0.0 && left > 0.0) ? (long)std::min(left / dt[r], 1e9) : 0;
            ahead = std::max(ahead, k);
        }
        if (ahead < 0) {
            st |= 1;
            break;
        }
        if (done >= max_steps) break;
        // launch most of them (dt drifts between reads), at least one
        long burst = std::max(1L, std::min(256L, (ahead * 3) / 4));
        burst = std::min(burst, (long)(max_steps - done));
        for (long s = 0; s < burst; ++s) {
            if (poll_files) {
                if (FILE* f = std::fopen("latest", "r")) {
                    std::fclose(f);
                    std::remove("latest");
                    st |= 4;
                    stop = true;
                    break;
                }
                if (FILE* f = std::fopen("abort", "r")) {
                    std::fclose(f);
                    st |= 2;
                    stop = true;
                }
            }
            HIPCHK(hipGraphLaunch(h->graph[h->parity], h->stream));
            h->parity ^= step_parity_flip(h);
            ++done;
            if (stop) break;
        }
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    if (check_err(h)) return -1;
    if (stop && !(st & 1)) {
        // the flags after the last step launched
        HIPCHK(hipMemcpy(fin.data(), h->final_step, n * sizeof(int), hipMemcpyDeviceToHost));
        if (std::all_of(fin.begin(), fin.end(), [](int f) { return f != 0; })) st |= 1;
    }
    if (steps_done) *steps_done = done;
    if (status) *status = st;
    return 0;
}

// The report line's accretion fields of run `run` (report.c:278-303: gravL,
// dM, accL, Ls, kin, therm, linXmom, linYmom, each for every planet) into
// out[8 * npl], then the accumulators restart from zero (report.c:91-106;
// a kernel, not a memset, so that the call may be captured).  No host
// allocation.  who: the entry point for error messages (null: fail_c).
static int sink_line(disco_gpu* h, int run, int npl, double* out, const char* who) {
    for (int p = 0; p < 8 * npl; ++p) out[p] = 0.0;
    if (!h->sink_acc) return 0;
    double acc[DISCO_NSINKQ * DISCO_MAX_PLANETS];
    hipError_t e = hipMemcpyAsync(acc, h->sink_acc + (size_t)run * DISCO_NSINKQ * DISCO_MAX_PLANETS, sizeof acc,
                                  hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) {
        launch_sink_reset(make_state(h, h->parity, 1 - h->parity), run, h->stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        if (who) return fail(std::string(who) + ": " + hipGetErrorString(e));
        return fail_c("disco_gpu_report_noalloc: a kernel or copy failed");
    }
    for (int k = 0; k < DISCO_NSINKQ; ++k)
        for (int p = 0; p < npl; ++p) out[k * npl + p] = acc[p * DISCO_NSINKQ + k];
    return 0;
}

int disco_gpu_report(disco_gpu_t* h, int run, double* out, int nmax, int* nout) {
    if (!h || run < 0 || run >= h->nruns || !out) return fail("disco_gpu_report: bad arguments");
    const HostRun& R = h->runs[run];
    const int npl = (int)R.planets.size();
    const int n = 15 + 8 * npl;
    if (nout) *nout = n;
    if (nmax < n) return fail("disco_gpu_report: out holds " + std::to_string(nmax) + " values, the line has " +
                              std::to_string(n));
    const int na = R.Nr * R.Nz, a0 = h->run_ann0[run];
    std::vector<double> part((size_t)na * DISCO_NREPORT);
    double* dpart = nullptr;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMalloc((void**)&dpart, part.size() * sizeof(double)));
    DevState S = make_state(h, h->parity, 1 - h->parity);
    launch_report(S, a0, na, dpart, h->stream);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(part.data(), dpart, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                                            h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(dpart);
    if (e != hipSuccess) return fail(std::string("disco_gpu_report: ") + hipGetErrorString(e));
    // the annuli in report.c's loop order: j outer, k inner (slot j + Nr k)
    double sum[DISCO_NREPORT] = {};
    for (int j = 0; j < R.Nr; ++j)
        for (int k = 0; k < R.Nz; ++k)
            for (int q = 0; q < DISCO_NREPORT; ++q) sum[q] += part[((size_t)j + (size_t)R.Nr * k) * DISCO_NREPORT + q];
    double t = 0.0;
    HIPCHK(hipMemcpy(&t, h->t + run, sizeof(double), hipMemcpyDeviceToHost));
    int o = 0;
    out[o++] = t;
    out[o++] = npl > 1 ? R.planets[1].M / R.planets[0].M : 0.0;
    // the planets' accretion accumulators (sink.c; zero without sinks), then
    // zeroed as report.c:91-106 does
    if (sink_line(h, run, npl, out + o, "disco_gpu_report")) return -1;
    o += 8 * npl;
    for (int q = 0; q < 3; ++q) out[o++] = sum[q];
    for (int q = 3; q < DISCO_NREPORT; ++q) out[o++] = sum[q];
    return 0;
}

// disco_gpu_report without allocating: the annulus sums go to this GPU's
// report_scratch and are added on the device in the same order as the host
// adds them in disco_gpu_report (so both return the same bits); only the
// handle's stream, kernel launches and copies into the caller's `out`.
int disco_gpu_report_noalloc(disco_gpu_t* h, int run, double* out, int nmax, int* nout) {
    if (!h || run < 0 || run >= h->nruns || !out) return fail_c("disco_gpu_report_noalloc: bad arguments");
    const HostRun& R = h->runs[run];
    const int npl = (int)R.planets.size();
    const int n = 15 + 8 * npl;
    if (nout) *nout = n;
    if (nmax < n) return fail_c("disco_gpu_report_noalloc: out holds fewer values than the report line");
    const int na = R.Nr * R.Nz, a0 = h->run_ann0[run];
    if (na > DISCO_REPORT_MAX_ANN)
        return fail_c("disco_gpu_report_noalloc: the run has more annuli than the scratch holds (65536)");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= DISCO_MAX_DEVICES || !report_scratch_dev[dev])
        return fail_c("disco_gpu_report_noalloc: no scratch on the current device (create a handle on it first)");
    double* part = report_scratch_dev[dev];
    double* sums = part + (size_t)DISCO_REPORT_MAX_ANN * DISCO_NREPORT;
    // every slot of the run's Nr x Nz annuli is written by k_report (zeros
    // for ghost annuli) and every total by k_report_sum: nothing left from an
    // earlier call is read
    DevState S = make_state(h, h->parity, 1 - h->parity);
    launch_report(S, a0, na, part, h->stream);
    launch_report_sum(part, R.Nr, R.Nz, sums, h->stream);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, h->t + run, sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out + 2 + 8 * npl, sums, DISCO_NREPORT * sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail_c("disco_gpu_report_noalloc: a kernel or copy failed");
    out[1] = npl > 1 ? R.planets[1].M / R.planets[0].M : 0.0;
    return sink_line(h, run, npl, out + 2, nullptr);
}

// ---------------------------------------------------------------- checkpoints
static const char* hydro_name(int v) {
    return v == DISCO_HYDRO_MHD       ? "mhd"
           : v == DISCO_HYDRO_GRMHD   ? "grmhd"
           : v == DISCO_HYDRO_GRMHD2  ? "grmhd2"
           : v == DISCO_HYDRO_GREULER ? "greuler"
                                      : "euler";
}
static const char* boundary_name(int v) {
    static const char* n[] = {"fixed", "fixed3d", "zerograd", "outflow3d"};
    return v >= 0 && v < 4 ? n[v] : "?";
}
static const char* planets_name(int v) {
    static const char* n[] = {"none", "sun", "binary"};
    return v >= 0 && v < 3 ? n[v] : "?";
}
static const char* initial_name(int v) {
    static const char* n[] = {"kepler_noise", "iso_binary_disk", "mhd_disk", "gr_torus"};
    return v >= 0 && v < 4 ? n[v] : "?";
}

// Output/h5out.c:366-622 for run `run`, in the container of checkpoint.h:
// Grid/{T, r_jph, z_kph, Index, Np, Id_phi0}, Data/Cells[Ntot][NUM_Q +
// NUM_FACES + 1] (prim, Phi, piph per cell in theCells order), Data/Planets
// [Npl][7], Data/Diagnostics[Nz][Nr][8] (avg_diagnostics, then the
// accumulators are zeroed as output() does), Pars/<in.par key>, Opts/<module
// slot>; and, beyond the reference's layout, Data/Cons[Ntot][NUM_Q] and
// Data/Wiph[Ntot], which let disco_gpu_restart resume the run bit for bit.
int disco_gpu_write_checkpoint(disco_gpu_t* h, int run, const char* path) {
    if (!h || run < 0 || run >= h->nruns || !path) return fail("disco_gpu_write_checkpoint: bad arguments");
    const HostRun& R = h->runs[run];
    const int nq = h->nq, nf = h->Phi ? 5 : 0, nc = R.ncell, na = R.Nr * R.Nz;
    std::vector<double> prim((size_t)nq * nc), cons((size_t)nq * nc), piph(nc), dphi(nc), wiph(nc), Phi((size_t)nf * nc);
    if (disco_gpu_download(h, run, prim.data(), cons.data(), piph.data(), dphi.data(), wiph.data())) return -1;
    if (nf && disco_gpu_download_faces(h, run, Phi.data())) return -1;
    double t = 0.0, pph[DISCO_MAX_PLANETS] = {0.0, 0.0};
    HIPCHK(hipMemcpy(&t, h->t + run, sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(pph, h->planet_phi + run * DISCO_MAX_PLANETS, sizeof pph, hipMemcpyDeviceToHost));
    std::vector<double> Qrz((size_t)na * DISCO_NDIAG);
    double t_avg = 0.0;
    if (disco_gpu_diagnostics(h, run, Qrz.data(), &t_avg, 1)) return -1;
    for (double& v : Qrz) v /= t_avg;  // avg_diagnostics (analysis.c:29-50)

    const int nd = nq + nf + 1;
    std::vector<double> cells((size_t)nc * nd), consT((size_t)nc * nq);
    for (int c = 0; c < nc; ++c) {
        for (int q = 0; q < nq; ++q) cells[(size_t)c * nd + q] = prim[(size_t)q * nc + c];
        for (int f = 0; f < nf; ++f) cells[(size_t)c * nd + nq + f] = Phi[(size_t)f * nc + c];
        cells[(size_t)c * nd + nq + nf] = piph[c];
        for (int q = 0; q < nq; ++q) consT[(size_t)c * nq + q] = cons[(size_t)q * nc + c];
    }
    // Index, Np, Id_phi0 per (k, j): h5out.c:526-547
    std::vector<int32_t> index(na), np(na), id0(na);
    for (int k = 0; k < R.Nz; ++k)
        for (int j = 0; j < R.Nr; ++j) {
            const int jk = j + R.Nr * k, o = R.off[jk];
            index[jk] = o;
            np[jk] = R.Np[jk];
            double phi0 = M_PI;
            int id = 0;
            for (int i = 0; i < R.Np[jk]; ++i) {
                const double phi = piph[o + i] - .5 * dphi[o + i];
                if (std::cos(phi0) < std::cos(phi)) {
                    phi0 = phi;
                    id = o + i;
                }
            }
            id0[jk] = id;
        }
    const int npl = (int)R.planets.size();
    std::vector<double> pl((size_t)npl * 7);
    for (int p = 0; p < npl; ++p) {
        const Planet& P = R.planets[p];
        const double row[7] = {P.M, 0.0, P.omega, P.r, p < DISCO_MAX_PLANETS ? pph[p] : P.phi, P.eps, (double)P.type};
        std::memcpy(&pl[(size_t)p * 7], row, sizeof row);
    }
    using namespace ckpt;
    std::vector<Dataset> d;
    const uint64_t Nr = R.Nr, Nz = R.Nz;
    d.push_back(make_f64("Grid/T", {1}, &t));
    d.push_back(make_f64("Grid/r_jph", {Nr + 1}, R.r_jph.data()));
    d.push_back(make_f64("Grid/z_kph", {Nz + 1}, R.z_kph.data()));
    d.push_back(make_i32("Grid/Index", {Nz, Nr}, index.data()));
    d.push_back(make_i32("Grid/Np", {Nz, Nr}, np.data()));
    d.push_back(make_i32("Grid/Id_phi0", {Nz, Nr}, id0.data()));
    d.push_back(make_f64("Data/Cells", {(uint64_t)nc, (uint64_t)nd}, cells.data()));
    d.push_back(make_f64("Data/Planets", {(uint64_t)npl, 7}, pl.data()));
    d.push_back(make_f64("Data/Diagnostics", {Nz, Nr, DISCO_NDIAG}, Qrz.data()));
    d.push_back(make_f64("Data/Cons", {(uint64_t)nc, (uint64_t)nq}, consT.data()));
    d.push_back(make_f64("Data/Wiph", {(uint64_t)nc}, wiph.data()));
    const disco_params& P = R.par;
    auto pi = [&](const char* k, int v) { d.push_back(make_i32(std::string("Pars/") + k, {1}, &v)); };
    auto pd = [&](const char* k, double v) { d.push_back(make_f64(std::string("Pars/") + k, {1}, &v)); };
    // writePars (h5out.c:164-364), the keys this build reads
    pi("Num_R", P.Num_R); pi("Num_Z", P.Num_Z); pd("aspect", P.aspect); pd("T_Start", P.t_min);
    pd("T_End", P.t_max); pd("R_Min", P.rmin); pd("R_Max", P.rmax); pd("Z_Min", P.zmin); pd("Z_Max", P.zmax);
    pd("Phi_Max", P.phimax); pi("Log_Zoning", P.LogZoning); pi("R_Periodic", P.R_Periodic);
    pi("Z_Periodic", P.Z_Periodic); pi("NoBC_Rmin", P.NoBC_Rmin); pi("NoBC_Rmax", P.NoBC_Rmax);
    pi("NoBC_Zmin", P.NoBC_Zmin); pi("NoBC_Zmax", P.NoBC_Zmax); pd("Log_Radius", P.LogRadius); pd("CFL", P.CFL);
    pd("PLM", P.PLM); pd("Max_DT", P.maxDT); pd("Adiabatic_Index", P.Adiabatic_Index);
    pi("Isothermal", P.isothermal_flag); pi("Cs2_Profile", P.Cs2_Profile); pd("Cs2_Par", P.Cs2_Par);
    pd("Density_Floor", P.Density_Floor); pd("Pressure_Floor", P.Pressure_Floor); pi("Mesh_Motion", P.Mesh_Motion);
    pi("Riemann_Solver", P.Riemann_Solver); pi("Exact_Mesh_Omega", P.Exact_Mesh_Omega);
    pd("Exact_Mesh_Omega_Par", P.Exact_Mesh_Omega_Par); pi("Energy_Omega", P.Energy_Omega);
    pd("Energy_Omega_Par", P.Energy_Omega_Par); pi("RotFrame", P.RotFrame); pd("RotOmega", P.RotOmega);
    pd("RotD", P.RotD); pi("Use_Viscosity", P.visc_flag); pd("Mach_Number", P.Disk_Mach);
    pd("Mass_Ratio", P.Mass_Ratio); pi("Constrained_Transport", P.CT); pi("Metric_Par0", P.metricPar0);
    pd("Metric_Par1", P.metricPar1); pd("Metric_Par2", P.metricPar2); pd("Metric_Par3", P.metricPar3);
    pd("Metric_Par4", P.metricPar4); pi("Init_Par0", P.initPar0); pd("Init_Par1", P.initPar1);
    pd("Init_Par2", P.initPar2); pd("Init_Par3", P.initPar3); pd("Init_Par4", P.initPar4); pd("Init_Par5", P.initPar5);
    pd("Init_Par6", P.initPar6); pd("Init_Par7", P.initPar7); pd("Init_Par8", P.initPar8);
    pi("Noise_Type", P.noiseType); pd("Noise_Abs", P.noiseAbs); pd("Noise_Rel", P.noiseRel);
    pi("Sink_Type", P.sinkType); pd("Sink_Par1", P.sinkPar1); pd("Sink_Par2", P.sinkPar2);
    pd("Sink_Par3", P.sinkPar3); pd("Sink_Par4", P.sinkPar4); pd("Sink_Par5", P.sinkPar5);
    pi("Nozzle_Type", P.nozzleType); pd("Nozzle_Par1", P.nozzlePar1); pd("Nozzle_Par2", P.nozzlePar2);
    pd("Nozzle_Par3", P.nozzlePar3); pd("Nozzle_Par4", P.nozzlePar4); pi("Cool_Type", P.coolType);
    pd("Cool_Par1", P.coolPar1); pd("Cool_Par2", P.coolPar2); pd("Cool_Par3", P.coolPar3); pd("Cool_Par4", P.coolPar4);
    pd("Softening", P.grav_eps);
    // writeOpts (h5out.c:96-162): the module slots
    d.push_back(make_str("Opts/HYDRO", hydro_name(P.hydro)));
    d.push_back(make_str("Opts/BOUNDARY", boundary_name(P.boundary)));
    d.push_back(make_str("Opts/PLANETS", planets_name(P.planets)));
    d.push_back(make_str("Opts/INITIAL", initial_name(P.initial)));
    d.push_back(make_str("Opts/GEOMETRY", "cylindrical"));
    pi("NUM_C", P.num_c);
    d.back().name = "Opts/NUM_C";
    pi("NUM_N", P.num_n);
    d.back().name = "Opts/NUM_N";
    pi("CT_MODE", nf ? 2 : 0);
    d.back().name = "Opts/CT_MODE";
    std::string err;
    if (write_file(path, d, &err)) return fail("disco_gpu_write_checkpoint: " + err);
    return 0;
}

// Restart/h5in.c:101-368 into run `run` of a handle created from the same
// parameters: the grid of the file must be the run's (Nr, Nz, Np, r_jph,
// z_kph); prims, face fluxes and fronts come from Data/Cells, the time from
// Grid/T, the planet phases from Data/Planets.  With Data/Cons and Data/Wiph
// in the file (this library's checkpoints) the run continues bit for bit as
// the run that wrote it; otherwise the conserved state, ghosts and mesh
// speeds are rebuilt as the reference's restart does.  The diagnostics start
// from zero.
int disco_gpu_restart(disco_gpu_t* h, int run, const char* path) {
    if (!h || run < 0 || run >= h->nruns || !path) return fail("disco_gpu_restart: bad arguments");
    using namespace ckpt;
    std::vector<Dataset> d;
    std::string err;
    if (read_file(path, &d, &err)) return fail("disco_gpu_restart: " + err);
    HostRun& R = h->runs[run];
    const int nq = h->nq, nf = h->Phi ? 5 : 0, nc = R.ncell, na = R.Nr * R.Nz;
    const Dataset *T = find(d, "Grid/T"), *Ix = find(d, "Grid/Index"), *Np = find(d, "Grid/Np"),
                  *rj = find(d, "Grid/r_jph"), *zk = find(d, "Grid/z_kph"), *C = find(d, "Data/Cells"),
                  *Pl = find(d, "Data/Planets"), *Cs = find(d, "Data/Cons"), *Wi = find(d, "Data/Wiph");
    if (!T || !Ix || !Np || !rj || !zk || !C) return fail("disco_gpu_restart: missing Grid/ or Data/Cells datasets");
    if (Ix->type != I32 || Np->type != I32 || Ix->dims.size() != 2 || Ix->dims[0] != (uint64_t)R.Nz ||
        Ix->dims[1] != (uint64_t)R.Nr || Np->count() != (size_t)na)
        return fail("disco_gpu_restart: the file's grid is " + std::to_string(Ix->dims.size() == 2 ? Ix->dims[1] : 0) +
                    " x " + std::to_string(Ix->dims.size() == 2 ? Ix->dims[0] : 0) + " annuli, the run's " +
                    std::to_string(R.Nr) + " x " + std::to_string(R.Nz));
    for (int jk = 0; jk < na; ++jk)
        if (Np->i32()[jk] != R.Np[jk]) return fail("disco_gpu_restart: Np differs from the run's at annulus " + std::to_string(jk));
    if (rj->count() != R.r_jph.size() || zk->count() != R.z_kph.size() ||
        std::memcmp(rj->f64(), R.r_jph.data(), rj->data.size()) || std::memcmp(zk->f64(), R.z_kph.data(), zk->data.size()))
        return fail("disco_gpu_restart: r_jph / z_kph differ from the run's");
    const int nd = nq + nf + 1;
    if (C->type != F64 || C->dims.size() != 2 || C->dims[0] != (uint64_t)nc || C->dims[1] != (uint64_t)nd)
        return fail("disco_gpu_restart: Data/Cells is not [" + std::to_string(nc) + "][" + std::to_string(nd) +
                    "] (NUM_Q + NUM_FACES + 1 of this run)");
    const bool have_cons = Cs && Cs->type == F64 && Cs->count() == (size_t)nc * nq;
    const bool have_wiph = Wi && Wi->type == F64 && Wi->count() == (size_t)nc;
    for (int jk = 0; jk < na; ++jk) {
        const int o = R.off[jk], src = Ix->i32()[jk];
        if (src < 0 || src + R.Np[jk] > nc) return fail("disco_gpu_restart: Grid/Index out of range");
        for (int i = 0; i < R.Np[jk]; ++i) {
            const double* rec = C->f64() + (size_t)(src + i) * nd;
            const int c = o + i;
            for (int q = 0; q < nq; ++q) R.prim[(size_t)q * nc + c] = rec[q];
            for (int f = 0; f < nf; ++f) R.Phi[(size_t)f * nc + c] = rec[nq + f];
            R.piph[c] = rec[nq + nf];
            if (have_cons)
                for (int q = 0; q < nq; ++q) R.cons[(size_t)q * nc + c] = Cs->f64()[(size_t)(src + i) * nq + q];
            if (have_wiph) R.wiph[c] = Wi->f64()[src + i];
        }
    }
    host_restart_fill(&R, have_cons, have_wiph);
    R.t = T->f64()[0];
    if (Pl && Pl->dims.size() == 2 && Pl->dims[1] >= 5)
        for (size_t p = 0; p < R.planets.size() && p < Pl->dims[0]; ++p) R.planets[p].phi = Pl->f64()[p * Pl->dims[1] + 4];
    HIPCHK(hipStreamSynchronize(h->stream));
    if (upload_run_state(h, run)) return -1;
    if (disco_gpu_diagnostics(h, run, nullptr, nullptr, 1)) return -1;
    return seed_dt(h);
}

int disco_gpu_final(disco_gpu_t* h, int* final_step) {
    if (!h || !final_step) return fail("disco_gpu_final: bad arguments");
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(final_step, h->final_step, h->nruns * sizeof(int), hipMemcpyDeviceToHost));
    return 0;
}

int disco_gpu_timestep(disco_gpu_t* h, const double* dt) {
    if (!h || !dt) return fail("disco_gpu_timestep: bad arguments");
    HIPCHK(hipMemcpyAsync(h->dt, dt, h->nruns * sizeof(double), hipMemcpyHostToDevice, h->stream));
    enqueue_step(h, false, nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    h->parity ^= step_parity_flip(h);
    return 0;
}

// Device error word: set by kernels that meet a mesh configuration they do
// not replay (constrained transport: adjacent flux flips in one track).
static int check_err(disco_gpu* h) {
    int e = 0;
    HIPCHK(hipMemcpy(&e, h->err, sizeof(int), hipMemcpyDeviceToHost));
    if (e & 1)
        return fail("constrained transport: adjacent flux flips in one track (a front overtook two others "
                    "in one step); this pattern is not replayed");
    if (e) return fail("device error word " + std::to_string(e));
    return 0;
}

int disco_gpu_sync(disco_gpu_t* h) {
    if (!h) return fail("disco_gpu_sync: null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    return check_err(h);
}

int disco_gpu_download(disco_gpu_t* h, int run, double* prim, double* cons, double* piph, double* dphi,
                       double* wiph) {
    if (!h || run < 0 || run >= h->nruns) return fail("disco_gpu_download: bad arguments");
    HIPCHK(hipStreamSynchronize(h->stream));
    if (prim && get_fields(h, run, prim, h->prim, h->nq)) return -1;
    if (cons && get_fields(h, run, cons, h->cons, h->nq)) return -1;
    if (piph && get_fields(h, run, piph, h->piph[h->parity], 1)) return -1;
    if (dphi && get_fields(h, run, dphi, h->dphi[h->parity], 1)) return -1;
    if (wiph && get_fields(h, run, wiph, h->wiph, 1)) return -1;
    return check_err(h);
}

int disco_gpu_download_faces(disco_gpu_t* h, int run, double* Phi) {
    if (!h || run < 0 || run >= h->nruns || !Phi) return fail("disco_gpu_download_faces: bad arguments");
    if (!h->Phi) return fail("disco_gpu_download_faces: the run has no face fluxes (not MHD)");
    HIPCHK(hipStreamSynchronize(h->stream));
    if (get_fields(h, run, Phi, h->Phi, 5)) return -1;
    return check_err(h);
}

int disco_gpu_upload_faces(disco_gpu_t* h, int run, const double* Phi) {
    if (!h || run < 0 || run >= h->nruns || !Phi) return fail("disco_gpu_upload_faces: bad arguments");
    if (!h->Phi) return fail("disco_gpu_upload_faces: the run has no face fluxes (not MHD)");
    HIPCHK(hipStreamSynchronize(h->stream));
    HostRun& R = h->runs[run];
    size_t n = R.ncell;
    std::memcpy(R.Phi.data(), Phi, 5 * n * sizeof(double));
    return put_fields(h, run, h->Phi, R.Phi.data(), 5);
}

int disco_gpu_upload(disco_gpu_t* h, int run, const double* prim, const double* cons, const double* piph,
                     const double* dphi, const double* wiph) {
    if (!h || run < 0 || run >= h->nruns) return fail("disco_gpu_upload: bad arguments");
    HIPCHK(hipStreamSynchronize(h->stream));
    HostRun& R = h->runs[run];
    size_t n = R.ncell;
    if (prim) std::memcpy(R.prim.data(), prim, h->nq * n * sizeof(double));
    if (cons) std::memcpy(R.cons.data(), cons, h->nq * n * sizeof(double));
    if (piph) std::memcpy(R.piph.data(), piph, n * sizeof(double));
    if (dphi) std::memcpy(R.dphi.data(), dphi, n * sizeof(double));
    if (wiph) std::memcpy(R.wiph.data(), wiph, n * sizeof(double));
    // recompute the rotation point of every annulus for the new mesh
    host_rotation(R);
    if (upload_run_state(h, run)) return -1;
    return seed_dt(h);
}

int disco_gpu_time(disco_gpu_t* h, double* t, double* dt, double* planet_phi) {
    if (!h) return fail("disco_gpu_time: null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    if (t) HIPCHK(hipMemcpy(t, h->t, h->nruns * sizeof(double), hipMemcpyDeviceToHost));
    if (dt) HIPCHK(hipMemcpy(dt, h->dt, h->nruns * sizeof(double), hipMemcpyDeviceToHost));
    if (planet_phi)
        HIPCHK(hipMemcpy(planet_phi, h->planet_phi, h->nruns * DISCO_MAX_PLANETS * sizeof(double),
                         hipMemcpyDeviceToHost));
    return 0;
}

struct TimerCtx {
    hipStream_t st;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> ev;
};

static void timer_before(void* ctx, const char* name) {
    TimerCtx* T = (TimerCtx*)ctx;
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    (void)hipEventRecord(a, T->st);
    T->ev.push_back({name, {a, b}});
}

static void timer_after(void* ctx, const char* name) {
    (void)name;
    TimerCtx* T = (TimerCtx*)ctx;
    (void)hipEventRecord(T->ev.back().second.second, T->st);
}

int disco_gpu_kernel_times(disco_gpu_t* h, int nsteps, int maxk, char* names, double* avg_ms, int* count) {
    if (!h || nsteps <= 0) return fail("disco_gpu_kernel_times: bad arguments");
    TimerCtx T;
    T.st = h->stream;
    LaunchHook hook{timer_before, timer_after, &T};
    for (int n = 0; n < nsteps; ++n) {
        enqueue_step(h, true, &hook);
        h->parity ^= step_parity_flip(h);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    std::vector<std::string> order;
    std::map<std::string, std::pair<double, int>> acc;
    for (auto& e : T.ev) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e.second.first, e.second.second);
        if (!acc.count(e.first)) order.push_back(e.first);
        acc[e.first].first += ms;
        acc[e.first].second += 1;
        (void)hipEventDestroy(e.second.first);
        (void)hipEventDestroy(e.second.second);
    }
    int k = 0;
    for (auto& nm : order) {
        if (k >= maxk) break;
        if (names) std::snprintf(names + 32 * k, 32, "%s", nm.c_str());
        if (avg_ms) avg_ms[k] = acc[nm].first / acc[nm].second;
        if (count) count[k] = acc[nm].second;
        ++k;
    }
    return k;
}

int disco_gpu_diagnostics(disco_gpu_t* h, int run, double* Qrz, double* t_avg, int reset) {
    if (!h || run < 0 || run >= h->nruns) return fail("disco_gpu_diagnostics: bad arguments");
    HIPCHK(hipStreamSynchronize(h->stream));
    const HostRun& R = h->runs[run];
    const int a0 = h->run_ann0[run], na = R.Nr * R.Nz;
    if (Qrz) {
        std::vector<double> tmp((size_t)na * DISCO_NDIAG);
        HIPCHK(hipMemcpy(tmp.data(), h->diag + (size_t)a0 * DISCO_NDIAG, tmp.size() * sizeof(double),
                         hipMemcpyDeviceToHost));
        // device annulus order -> the reference's jk = j + Nr*k
        for (int a = 0; a < na; ++a) {
            const AnnDev& A = h->hann[a0 + a];
            std::memcpy(Qrz + (size_t)(A.j + R.Nr * A.k) * DISCO_NDIAG, tmp.data() + (size_t)a * DISCO_NDIAG,
                        DISCO_NDIAG * sizeof(double));
        }
    }
    if (t_avg) HIPCHK(hipMemcpy(t_avg, h->t_avg + run, sizeof(double), hipMemcpyDeviceToHost));
    if (reset) {
        HIPCHK(hipMemset(h->diag + (size_t)a0 * DISCO_NDIAG, 0, (size_t)na * DISCO_NDIAG * sizeof(double)));
        HIPCHK(hipMemset(h->t_avg + run, 0, sizeof(double)));
    }
    return check_err(h);
}

int disco_gpu_autotune(disco_gpu_t* h, int nsteps, int* choice, double* ms_table) {
    if (!h || nsteps <= 0) return fail("disco_gpu_autotune: bad arguments");
    // candidates: annulus tiles (3D runs only); per tile, every kernel class
    // takes the width (64, 128, 256 threads per block) at which it ran fastest
    std::vector<int> kts = {h->ktile};
    if (h->shape.has_z) kts = {1, 2, 4, 8, 16};
    const int bss[3] = {64, 128, 256};
    const int kt0 = h->ktile, parity0 = h->parity;
    int blk0[KC_N];
    std::memcpy(blk0, h->blk, sizeof blk0);
    Snapshot S;  // everything the timing steps change, in the current layout
    if (snapshot(h, &S, false)) return -1;
    double best = 1e300;
    int bk = h->ktile, bblk[KC_N], idx = 0;
    std::memcpy(bblk, h->blk, sizeof bblk);
    const int maxk = 32;
    std::vector<char> names(32 * maxk);
    std::vector<double> avg(maxk);
    std::vector<int> cnt(maxk);
    for (int kt : kts) {
        double cls[3][KC_N] = {};
        double other = 0.0;
        for (int b = 0; b < 3; ++b) {
            int w[KC_N];
            for (int k = 0; k < KC_N; ++k) w[k] = bss[b];
            if (set_layout(h, kt, w)) return -1;
            if (disco_gpu_step(h, 1)) return -1;
            int nk = disco_gpu_kernel_times(h, nsteps, maxk, names.data(), avg.data(), cnt.data());
            if (nk < 0) return -1;
            double total = 0.0;
            for (int k = 0; k < nk; ++k) {
                double per_step = avg[k] * cnt[k] / nsteps;
                total += per_step;
                int c = kernel_class(std::string(names.data() + 32 * k));
                if (c >= 0)
                    cls[b][c] += per_step;
                else if (b == 2)
                    other += per_step;  // per-step kernels of fixed width
            }
            if (ms_table) ms_table[idx] = total;
            ++idx;
        }
        int w[KC_N];
        double total = other;
        for (int c = 0; c < KC_N; ++c) {
            int bi = 2;
            for (int b = 0; b < 3; ++b)
                if (cls[b][c] < cls[bi][c]) bi = b;
            w[c] = bss[bi];
            total += cls[bi][c];
        }
        if (total < best) {
            best = total;
            bk = kt;
            std::memcpy(bblk, w, sizeof w);
        }
    }
    // back to the snapshot's layout and state, then move it to the winner
    if (set_layout(h, kt0, blk0)) return -1;
    h->parity = parity0;
    if (snapshot(h, &S, true) || set_layout(h, bk, bblk)) return -1;
    if (choice) {
        choice[0] = bk;
        for (int k = 0; k < KC_N; ++k) choice[1 + k] = bblk[k];
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    return check_err(h);
}

int disco_gpu_set_c2p_net(disco_gpu_t* h, const float* params, const float* norm) {
    if (!h) return fail("disco_gpu_set_c2p_net: null handle");
    if (h->runs[0].par.hydro != DISCO_HYDRO_GRMHD) return fail("disco_gpu_set_c2p_net: the runs are not GRMHD");
    if (!params) {
        h->learned = false;
        h->shape.c2p_learned = 0;
        return capture_graphs(h);
    }
    if (!norm) return fail("disco_gpu_set_c2p_net: norm is NULL");
    for (int i = 0; i < 2 * mlp::NIN + 2 * mlp::NOUT; ++i)
        if (!std::isfinite(norm[i]) || (((i >= mlp::NIN && i < 2 * mlp::NIN) || i >= 2 * mlp::NIN + mlp::NOUT) &&
                                        !(norm[i] > 0.0f)))
            return fail("disco_gpu_set_c2p_net: bad standardisation constants");
    if (!h->mlp_wt && (dalloc(&h->mlp_wt, mlp::WT_TOTAL) || dalloc(&h->mlp_w, mlp::NPARAM))) return -1;
    std::vector<unsigned short> T(mlp::WT_TOTAL);
    mlp::mlp_pack_bf16(params, T.data());
    HIPCHK(hipMemcpy(h->mlp_wt, T.data(), T.size() * sizeof(unsigned short), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->mlp_w, params, mlp::NPARAM * sizeof(float), hipMemcpyHostToDevice));
    std::memcpy(h->mlp_norm, norm, sizeof(h->mlp_norm));
    h->learned = true;
    h->shape.c2p_learned = 1;
    return capture_graphs(h);
}

int disco_gpu_c2p_count(disco_gpu_t* h, int enable) {
    if (!h) return fail("disco_gpu_c2p_count: null handle");
    if (h->runs[0].par.hydro != DISCO_HYDRO_GRMHD) return fail("disco_gpu_c2p_count: the runs are not GRMHD");
    if (!h->c2p_stats) {
        if (dalloc(&h->c2p_stats, 4)) return -1;
        HIPCHK(hipMemset(h->c2p_stats, 0, 4 * sizeof(unsigned long long)));
    }
    h->counting = enable != 0;
    return capture_graphs(h);
}

int disco_gpu_c2p_stats(disco_gpu_t* h, long long* out, int reset) {
    if (!h || !out) return fail("disco_gpu_c2p_stats: bad arguments");
    unsigned long long v[4] = {0, 0, 0, 0};
    HIPCHK(hipStreamSynchronize(h->stream));
    if (h->c2p_stats) {
        HIPCHK(hipMemcpy(v, h->c2p_stats, sizeof(v), hipMemcpyDeviceToHost));
        if (reset) HIPCHK(hipMemset(h->c2p_stats, 0, sizeof(v)));
    }
    for (int i = 0; i < 4; ++i) out[i] = (long long)v[i];
    return 0;
}

int disco_gpu_time_steps(disco_gpu_t* h, int nsteps, double* ms) {
    if (!h || nsteps < 0 || !ms) return fail("disco_gpu_time_steps: bad arguments");
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, h->stream));
    if (disco_gpu_step(h, nsteps)) return -1;
    HIPCHK(hipEventRecord(b, h->stream));
    HIPCHK(hipEventSynchronize(b));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, a, b));
    *ms = f;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return 0;
}

}  // extern "C"
