// HIP kernels of one RK substep of the moving-mesh finite-volume update
// (gfx950, FP64): one thread per cell, 256-thread blocks, around the per-cell
// bodies of step_cells.h / ct_cells.h.
//
// Newtonian hydro (euler.c) runs four cell-parallel kernels per substep:
//   k_stage_a -> k_plm_trans -> k_flux -> k_update (+ k_zcopy, k_planets)
// Ideal MHD with constrained transport (mhd.c, hlld.c, bfields.c) adds the
// EMF averaging and the Stokes update between the fluxes and the update, and
// splits the update around the flux flips of the moving mesh:
//   k_stage_a -> k_plm_trans -> k_flux -> k_emf_sum -> k_emf_face ->
//   k_ct_update -> k_update_a -> [k_flip_flag -> k_flip_chain]
//   -> k_update_b (+ k_zcopy, k_planets)
// GRMHD runs the same sequence with its geometric sources in a kernel of
// their own (k_source, before k_ct_update); with a trained guess network attached,
// k_update_b_learned replaces k_update_b (network on MFMA, then the Newton
// solve from its guess).
// The CFL reduction of the next step is fused into the last substep's update.
#include <hip/hip_runtime.h>

#include "mlp_infer.h"
#include "step_cells.h"

namespace disco {

#define DEV __device__ static inline __attribute__((always_inline))

// Minimum waves per SIMD requested for the register-heavy kernels (0: the
// compiler's choice).  The FP64 Riemann solvers and slope gathers need
// 130-180 VGPRs, i.e. 2-3 waves per SIMD, which leaves little latency hiding
// for their neighbour gathers; the build may trade registers for occupancy.
#ifndef DISCO_OCC_FLUX
#define DISCO_OCC_FLUX 3
#endif
#ifndef DISCO_OCC_PLM
#define DISCO_OCC_PLM 0
#endif
// The conserved update (k_update_a) runs uncapped: held to 4 waves it
// spilled 100 B per lane to scratch and took 10% longer than at the 3 waves
// its 162 VGPRs allow.  The Stokes update (k_ct_update) is held to 6 waves.
#ifndef DISCO_OCC_CT
#define DISCO_OCC_CT 0
#endif
#ifndef DISCO_OCC_CTU
#define DISCO_OCC_CTU 6
#endif
#define DISCO_OCC(n) __attribute__((amdgpu_waves_per_eu((n) > 0 ? (n) : 1, 8)))

// Per module: the GRMHD leaves (metric, its derivatives, the relativistic
// fluxes) need 240-280 VGPRs; capping them spills hundreds of registers to
// scratch, whose traffic costs more than the occupancy buys.
template <class Ph>
struct Occ {
    static constexpr int FLUX = DISCO_OCC_FLUX, PLM = DISCO_OCC_PLM, CT = DISCO_OCC_CT;
};
template <>
struct Occ<PhysGRMHD> {
    static constexpr int FLUX = 0, PLM = 0, CT = 0;
};
template <>
struct Occ<PhysGRMHD2> {
    static constexpr int FLUX = 0, PLM = 0, CT = 0;
};
template <>
struct Occ<PhysGREuler> {
    static constexpr int FLUX = 0, PLM = 0, CT = 0;
};

// Blocks are dispatched round-robin over the 8 XCDs, each with its own L2.
// Remap so that XCD x works on one contiguous band of cells: partner cells in
// the neighbouring annuli (about Np cells away in r, Nr*Np in z) are then
// mostly in the same L2 as the block that needs them.
DEV int cell_index() {
    const int nb = gridDim.x, b = blockIdx.x;
    const int x = b & 7, i = b >> 3;
    const int per = nb >> 3, rem = nb & 7;
    const int lb = x * per + (x < rem ? x : rem) + i;
    return lb * blockDim.x + threadIdx.x;
}

// ------------------------------------------------------------------ dt
__global__ void k_dt_init(DevState S) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= S.nruns) return;
    const PhysParams& P = S.runs[r].ph;
    double dt = P.max_dt / P.cfl;  // misc.c:53-55
    if (dt <= 0.0) dt = 1.0e100;
    S.dtbits[r] = (unsigned long long)__double_as_longlong(dt);
}

// Min-reduction of a per-cell CFL candidate into the run's running minimum
// (bit pattern of a positive double, so unsigned atomicMin orders it).  All
// threads of the block must call it.  A wavefront whose lanes belong to one
// run reduces with cross-lane shuffles; a block whose waves agree reduces
// through LDS and issues one atomic; mixed blocks fall back to per-wave or
// per-lane atomics (only at run boundaries of a batch).
DEV void dt_reduce(const DevState& S, int run, double dtc) {
    __shared__ double wmin[16];
    __shared__ int wrun[16];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwave = (blockDim.x + 63) >> 6;
    int run0 = __shfl(run, 0);
    bool uniform = __all(run == run0);
    if (uniform) {
        for (int o = 32; o > 0; o >>= 1) dtc = fmin(dtc, __shfl_xor(dtc, o));
        if (lane == 0) {
            wmin[wave] = dtc;
            wrun[wave] = run0;
        }
    } else {
        if (run >= 0 && dtc < 1.0e300)
            atomicMin(&S.dtbits[run], (unsigned long long)__double_as_longlong(dtc));
        if (lane == 0) wrun[wave] = -2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        bool same = true;
        for (int w = 1; w < nwave; ++w) same = same && (wrun[w] == wrun[0]);
        if (same && wrun[0] >= 0) {
            double m = wmin[0];
            for (int w = 1; w < nwave; ++w) m = fmin(m, wmin[w]);
            if (m < 1.0e300) atomicMin(&S.dtbits[wrun[0]], (unsigned long long)__double_as_longlong(m));
        } else {
            for (int w = 0; w < nwave; ++w)
                if (wrun[w] >= 0 && wmin[w] < 1.0e300)
                    atomicMin(&S.dtbits[wrun[w]], (unsigned long long)__double_as_longlong(wmin[w]));
        }
    }
}

template <class Ph>
__global__ __launch_bounds__(256) void k_mindt(DevState S) {
    int c = cell_index();
    int run = -1;
    double dtc = 1.0e300;
    if (c < S.ncell) {
        run = S.ann[S.cell_ann[c]].run;
        dtc = mindt_cell<Ph>(S, c);
    }
    dt_reduce(S, run, dtc);
}

__global__ void k_dt_final(DevState S, int check) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= S.nruns) return;
    double m = __longlong_as_double((long long)S.dtbits[r]);
    double dt = m * S.runs[r].ph.cfl;  // misc.c:79
    if (check) {
        // check_dt (domain.c:283-292): the step that would pass T_End ends
        // on it and is the run's last; a run past it holds (dt = 0) while
        // the other runs of the batch go on.
        if (S.final_step[r]) {
            dt = 0.0;
        } else {
            double t = S.t[r], tmax = S.runs[r].t_fin;
            if (t + dt > tmax) {
                dt = tmax - t;
                S.final_step[r] = 1;
            }
        }
    }
    S.dt[r] = dt;
}

// ------------------------------------------------------------------ stages
template <class Ph>
__global__ __launch_bounds__(256) void k_stage_a(DevState S, int first, int search) {
    int c = cell_index();
    if (c < S.ncell) stage_a_cell<Ph>(S, c, first, search);
}

__global__ __launch_bounds__(256) void k_partners(DevState S) {
    int c = cell_index();
    if (c < S.ncell) partners_cell(S, c);
}

// set_wcell's Mesh_Motion 3 (misc.c:137-150): every cell of an annulus
// moves at the mean of its cells' speeds, summed in array order (one block
// per annulus; the annuli of other runs keep theirs)
__global__ void k_wmean(DevState S) {
    const int a = blockIdx.x;
    const AnnDev& A = S.ann[a];
    if (S.runs[A.run].ph.mesh_motion != 3) return;
    __shared__ double wm;
    if (threadIdx.x == 0) {
        double w = 0.0;
        for (int i = 0; i < A.np; ++i) w += S.wiph[A.off + i];
        wm = w / (double)A.np;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < A.np; i += blockDim.x) S.wiph[A.off + i] = wm;
}

template <class Ph>
__global__ __launch_bounds__(256) DISCO_OCC(Occ<Ph>::PLM) void k_plm_trans(DevState S, int has_z) {
    int c = cell_index();
    if (c < S.ncell) plm_trans_cell<Ph>(S, c, has_z);
}

// VISC: with euler.c's viscous fluxes (an instantiation of its own, launched
// only for viscous runs)
// DC: with HLLC_DAMPCENTER's per-face choice of HLL (launched only when a run
// uses that solver)
template <class Ph, bool VISC, bool DC = false>
__global__ __launch_bounds__(256) DISCO_OCC(Occ<Ph>::FLUX) void k_flux(DevState S, double dtfac) {
    int c = cell_index();
    if (c < S.ncell) flux_cell<Ph, VISC, DC>(S, c, dtfac);
}

// Fused update for hydro: conserved state stays in registers between the
// gather/source half and the cons2prim/boundary half.
template <class Ph, bool SINK>
__global__ __launch_bounds__(256) void k_update(DevState S, double RK, double dtfac, int first, int want_dt) {
    int c = cell_index();
    int run = -1;
    double dtc = 1.0e300;
    if (c < S.ncell) {
        double cons[Ph::NQ];
        run = S.ann[S.cell_ann[c]].run;
        update_a_cell<Ph, SINK>(S, c, RK, dtfac, first, cons);
        dtc = update_b_cell<Ph>(S, c, first, want_dt, cons);
    }
    if (want_dt) dt_reduce(S, run, dtc);
}

// Wavefront sums of the in-loop cons2prim counters into S.c2p_stats (when
// counting is on).  Every lane of the wavefront must call it.
DEV void c2p_count(const DevState& S, int iters, int solves, int fallbacks, int unconverged) {
    if (!S.c2p_stats) return;
    for (int o = 32; o > 0; o >>= 1) {
        iters += __shfl_xor(iters, o);
        solves += __shfl_xor(solves, o);
        fallbacks += __shfl_xor(fallbacks, o);
        unconverged += __shfl_xor(unconverged, o);
    }
    if ((threadIdx.x & 63) == 0 && solves > 0) {
        atomicAdd(&S.c2p_stats[0], (unsigned long long)iters);
        atomicAdd(&S.c2p_stats[1], (unsigned long long)solves);
        atomicAdd(&S.c2p_stats[2], (unsigned long long)fallbacks);
        atomicAdd(&S.c2p_stats[3], (unsigned long long)unconverged);
    }
}

template <class Ph>
__global__ __launch_bounds__(256) void k_update_b(DevState S, int first, int want_dt) {
    constexpr int NQ = Ph::NQ;
    int c = cell_index();
    int run = -1;
    double dtc = 1.0e300;
    int it = -1, ok = 1;
    if (c < S.ncell) {
        const size_t N = S.ncell;
        double cons[NQ];
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) cons[q] = S.cons[q * N + c];
        run = S.ann[S.cell_ann[c]].run;
        dtc = update_b_cell<Ph>(S, c, first, want_dt, cons, &it, &ok);
    }
    if (Ph::C2P_GUESS) c2p_count(S, it > 0 ? it : 0, it >= 0, 0, it >= 0 && !ok);
    if (want_dt) dt_reduce(S, run, dtc);
}

// GRMHD second half with the learned start: B faces to cells and the
// invariants of every real cell, the guess network on MFMA for the 64 cells
// of each wavefront (BF16, FP32 accumulate), then the reference's Newton
// solve (same residuals, tolerance, 10 extra iterations, max_iter) from the
// network's (W, eta).  A start that does not converge, or converges outside
// 0 <= v^2 < 1, eta >= 1, falls back to the plain start (the cell's previous
// prims, grmhd.c:885-893).  Non-real lanes feed the network its mean input.
__global__ __launch_bounds__(256) void k_update_b_learned(DevState S, int first, int want_dt) {
    __shared__ mlp::MlpLds L;
    __shared__ unsigned short sF[4][64 * mlp::K0P];
    __shared__ float sY[4][64 * 2];
    mlp::mlp_load(L, S.mlp_wt, S.mlp_w);
    const int wave = threadIdx.x >> 6;
    const int c = cell_index();
    const bool valid = c < S.ncell;
    const size_t N = S.ncell;
    int run = -1;
    double dtc = 1.0e300;
    bool real = false;
    double cons[8];
    CellGeom G;
    if (valid) {
        _Pragma("unroll") for (int q = 0; q < 8; ++q) cons[q] = S.cons[q * N + c];
        run = S.ann[S.cell_ann[c]].run;
        real = update_b_pre<PhysGRMHD>(S, c, first, cons, &G);
    }
    grmhd::Metric g;
    grmhd::C2PInvariants v;
    float f[mlp::NIN];
    if (real) {
        const PhysParams& P = S.runs[run].ph;
        double cons1[8];
        _Pragma("unroll") for (int q = 0; q < 8; ++q) cons1[q] = cons[q] / G.dV;  // grmhd.c:232-234
        grmhd::kerr_metric(G.x, P.gr_M, P.gr_a, &g);
        grmhd::c2p_invariants(g, cons1, G.x, &v);
        mlp::c2p_features(v, f);
    } else {
        _Pragma("unroll") for (int i = 0; i < mlp::NIN; ++i) f[i] = S.mlp_norm[i];
    }
    _Pragma("unroll") for (int i = 0; i < mlp::NIN; ++i) f[i] = (f[i] - S.mlp_norm[i]) / S.mlp_norm[mlp::NIN + i];
    float y[2];
    mlp::mlp_wave(L, sF[wave], sY[wave], f, y);
    int iters = 0, fell = 0, bad = 0;
    if (real) {
        const PhysParams& P = S.runs[run].ph;
        const float* tn = S.mlp_norm + 2 * mlp::NIN;
        float y0 = y[0] * tn[2] + tn[0];
        float y1 = y[1] * tn[3] + tn[1];
        double W = exp((double)fmaxf(y0, 0.0f));
        double v20 = 1.0 - 1.0 / (W * W);
        double eta0 = fmax(v.e * exp((double)y1), 1.0 + 1e-10);
        double v2, eta;
        int ok;
        iters = grmhd::c2p_newton(P.gamma_law, v, v20, eta0, grmhd::C2P_PREC, grmhd::C2P_MAX_ITER,
                                  grmhd::C2P_NEXTRA, &v2, &eta, &ok);
        if (!ok || !(v2 >= 0.0 && v2 < 1.0) || !(eta >= 1.0)) {
            double prev[8];
            _Pragma("unroll") for (int q = 0; q < 8; ++q) prev[q] = S.prim[q * N + c];
            grmhd::c2p_guess_from_prim(P.gamma_law, g, prev, G.x, &v20, &eta0);
            iters += grmhd::c2p_newton(P.gamma_law, v, v20, eta0, grmhd::C2P_PREC, grmhd::C2P_MAX_ITER,
                                       grmhd::C2P_NEXTRA, &v2, &eta, &ok);
            fell = 1;
        }
        bad = !ok;
        double prim[8];
        grmhd::c2p_recover(P.gamma_law, P.rho_floor, P.pre_floor, g, v, v2, eta, G.x, prim);
        dtc = update_b_post<PhysGRMHD>(S, c, prim, G, want_dt);
    } else if (valid) {
        update_b_ghost<PhysGRMHD>(S, c);
    }
    c2p_count(S, iters, real, fell, bad);
    if (want_dt) dt_reduce(S, run, dtc);
}

template <class Ph>
__global__ __launch_bounds__(256) void k_zcopy(DevState S, int first) {
    int c = cell_index();
    if (c < S.ncell) zcopy_cell<Ph>(S, c, first);
}

// ---- constrained transport
__global__ __launch_bounds__(256) void k_emf_sum(DevState S, int has_z) {
    int c = cell_index();
    if (c < S.ncell) emf_sum_cell(S, c, has_z);
}

__global__ __launch_bounds__(256) void k_emf_face(DevState S, int has_z) {
    int c = cell_index();
    if (c < S.ncell) emf_face_cell(S, c, has_z);
}

template <class Ph>
__global__ __launch_bounds__(256) void k_source(DevState S, double dtfac) {
    int c = cell_index();
    if (c < S.ncell) source_cell<Ph>(S, c, dtfac);
}

// update_B_fluxes + make_edge_adjust (face fluxes), then the first half of
// the cell update (conserved state, sources, mesh move) in a launch of its
// own: fused, the two need 127 VGPRs (4 waves per SIMD) and run 19% slower
// than apart.  The Stokes update is held to 6 waves per SIMD (80 VGPRs; 5%
// faster than its natural 83 at 5 waves).
__global__ __launch_bounds__(256) DISCO_OCC(DISCO_OCC_CTU) void k_ct_update(DevState S, double RK, double dtfac, int first,
                                                               int has_z) {
    int c = cell_index();
    if (c < S.ncell) ct_update_cell(S, c, RK, dtfac, first, has_z);
}

template <class Ph, bool SINK>
__global__ __launch_bounds__(256) DISCO_OCC(Occ<Ph>::CT) void k_update_a(DevState S, double RK, double dtfac,
                                                                        int first) {
    int c = cell_index();
    if (c < S.ncell) {
        double cons[Ph::NQ];
        update_a_cell<Ph, SINK>(S, c, RK, dtfac, first, cons);
    }
}

__global__ __launch_bounds__(256) void k_flip_flag(DevState S, int has_z) {
    int c = cell_index();
    if (c < S.ncell) flip_flag_cell(S, c, has_z);
}

__global__ __launch_bounds__(256) void k_flip_chain(DevState S, int has_z) {
    int c = cell_index();
    if (c < S.ncell) flip_chain_cell(S, c, has_z);
}

// add_diagnostics (analysis.c:52-107), every annulus incl. ghosts: the
// volume-weighted annulus average of get_diagnostics' quantities, times dt,
// added to the annulus' accumulator.  One block per annulus; the block's sum
// runs in a fixed order over its threads (the reference sums in array order).
__global__ __launch_bounds__(256) void k_diag(DevState S) {
    __shared__ double red[DISCO_NDIAG + 1][4];
    const int a = blockIdx.x;
    const AnnDev& A = S.ann[a];
    const RunDev& Rn = S.runs[A.run];
    const double Pmax = Rn.ph.phi_max;
    const size_t N = S.ncell;
    double acc[DISCO_NDIAG + 1];
    for (int q = 0; q <= DISCO_NDIAG; ++q) acc[q] = 0.0;
    for (int i = threadIdx.x; i < A.np; i += blockDim.x) {
        const int c = A.off + i;
        double phip = S.piph_n[c];
        double phim = phip - S.dphi_n[c];
        double xp[3] = {A.rp, phip, A.zp};
        double xm[3] = {A.rm, phim, A.zm};
        double x[3] = {get_centroid(xp[0], xm[0], 1), get_centroid(xp[1], xm[1], 0), get_centroid(xp[2], xm[2], 2)};
        double dV = get_dV(xp, xm, Pmax);
        double prim[5];  // diag_euler reads rho, P, v_r, Omega, v_z only
        for (int q = 0; q < 5; ++q) prim[q] = S.prim[q * N + c];
        double Q[DISCO_NDIAG];
        diag_quantities(S, Rn, A.run, x, prim, Q);
        for (int q = 0; q < DISCO_NDIAG; ++q) acc[q] += Q[q] * dV;
        acc[DISCO_NDIAG] += dV;
    }
    // wavefront sums (fixed butterfly order), then the block's 4 partials
    for (int q = 0; q <= DISCO_NDIAG; ++q)
        for (int o = 32; o > 0; o >>= 1) acc[q] += __shfl_xor(acc[q], o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int q = 0; q <= DISCO_NDIAG; ++q) red[q][wave] = acc[q];
    __syncthreads();
    if (threadIdx.x < DISCO_NDIAG) {
        const int q = threadIdx.x;
        const int nw = (blockDim.x + 63) >> 6;
        double sq = red[q][0], sv = red[DISCO_NDIAG][0];
        for (int w = 1; w < nw; ++w) {
            sq += red[q][w];
            sv += red[DISCO_NDIAG][w];
        }
        S.diag[(size_t)a * DISCO_NDIAG + q] += sq * S.dt[A.run] / sv;
    }
}

// Accretion sinks: the substep's contributions of the cells of each real
// annulus summed per planet quantity (one block per annulus; fixed butterfly
// and wavefront order), zero for the other annuli.
constexpr int DISCO_SINK_NQ = DISCO_NSINKQ * DISCO_MAX_PLANETS;

__global__ __launch_bounds__(256) void k_sink_ann(DevState S) {
    __shared__ double red[DISCO_SINK_NQ][4];
    const int a = blockIdx.x;
    const AnnDev& A = S.ann[a];
    const size_t N = S.ncell;
    double acc[DISCO_SINK_NQ];
    for (int q = 0; q < DISCO_SINK_NQ; ++q) acc[q] = 0.0;
    if (A.flags & ANN_REAL)
        for (int i = threadIdx.x; i < A.np; i += blockDim.x)
            for (int q = 0; q < DISCO_SINK_NQ; ++q) acc[q] += S.sinkc[q * N + A.off + i];
    for (int q = 0; q < DISCO_SINK_NQ; ++q)
        for (int o = 32; o > 0; o >>= 1) acc[q] += __shfl_xor(acc[q], o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int q = 0; q < DISCO_SINK_NQ; ++q) red[q][wave] = acc[q];
    __syncthreads();
    if (threadIdx.x < DISCO_SINK_NQ) {
        const int q = threadIdx.x;
        const int nw = (blockDim.x + 63) >> 6;
        double sq = red[q][0];
        for (int w = 1; w < nw; ++w) sq += red[q][w];
        S.sink_ann[(size_t)a * DISCO_SINK_NQ + q] = sq;
    }
}

// The planets' accumulators of each run (one block per run, one thread per
// quantity): planet_RK_copy at the start of the step (first == 1),
// planet_RK_adjust_aux's average (planet.c:236-270, onestep.c:94), then the
// substep's sum over the run's annuli in table order.
__global__ void k_sink_run(DevState S, double RK, int first) {
    const int r = blockIdx.x, q = threadIdx.x;
    if (r >= S.nruns || q >= DISCO_SINK_NQ) return;
    const RunDev& Rn = S.runs[r];
    double sum = 0.0;
    for (int a = Rn.ann0; a < Rn.ann0 + Rn.nann; ++a) sum += S.sink_ann[(size_t)a * DISCO_SINK_NQ + q];
    double* acc = S.sink_acc + (size_t)r * DISCO_SINK_NQ;
    double* rk = S.sink_rk + (size_t)r * DISCO_SINK_NQ;
    double v = acc[q];
    if (first == 1) rk[q] = v;
    v = (1. - RK) * v + RK * rk[q];
    acc[q] = v + sum;
}

// report.c:91-106: the accumulators (and their RK copies) of run r restart
// from zero after a report
__global__ void k_sink_reset(DevState S, int r) {
    const int q = threadIdx.x;
    if (q < DISCO_SINK_NQ) {
        S.sink_acc[(size_t)r * DISCO_SINK_NQ + q] = 0.0;
        S.sink_rk[(size_t)r * DISCO_SINK_NQ + q] = 0.0;
    }
}

// report.c's gas sums (report_quantities) of the annuli a0, a0+1, ... of one
// run: one block per annulus, each writing its DISCO_NREPORT sums (zero for
// ghost annuli) to the annulus' slot j + Nr k of `out`; the annuli are then
// added in the reference's loop order (host, or k_report_sum).
__global__ __launch_bounds__(256) void k_report(DevState S, int a0, double* out) {
    __shared__ double red[DISCO_NREPORT][4];
    const int a = a0 + blockIdx.x;
    const AnnDev& A = S.ann[a];
    double acc[DISCO_NREPORT];
    for (int q = 0; q < DISCO_NREPORT; ++q) acc[q] = 0.0;
    if (A.flags & ANN_REAL) {
        for (int i = threadIdx.x; i < A.np; i += blockDim.x) {
            double Q[DISCO_NREPORT];
            report_quantities(S, A, A.off + i, Q);
            for (int q = 0; q < DISCO_NREPORT; ++q) acc[q] += Q[q];
        }
    }
    for (int q = 0; q < DISCO_NREPORT; ++q)
        for (int o = 32; o > 0; o >>= 1) acc[q] += __shfl_xor(acc[q], o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int q = 0; q < DISCO_NREPORT; ++q) red[q][wave] = acc[q];
    __syncthreads();
    if (threadIdx.x < DISCO_NREPORT) {
        const int q = threadIdx.x;
        double s = red[q][0];
        for (int w = 1; w < (int)((blockDim.x + 63) >> 6); ++w) s += red[q][w];
        const size_t slot = (size_t)A.j + (size_t)S.runs[A.run].Nr * A.k;
        out[slot * DISCO_NREPORT + q] = s;
    }
}

// The annulus sums of k_report added in report.c's loop order (j outer, k
// inner), one thread per quantity: sums[q] = sum over the Nr x Nz slots.
__global__ void k_report_sum(const double* part, int Nr, int Nz, double* sums) {
    const int q = threadIdx.x;
    if (q >= DISCO_NREPORT) return;
    double s = 0.0;
    for (int j = 0; j < Nr; ++j)
        for (int k = 0; k < Nz; ++k) s += part[((size_t)j + (size_t)Nr * k) * DISCO_NREPORT + q];
    sums[q] = s;
}

// Zero-gradient ghost fill, one block per job of one level (bc_copy_cell).
template <class Ph>
__global__ __launch_bounds__(256) void k_bc_copy(DevState S, int j0) {
    const int* J = S.bcjobs + (size_t)(j0 + blockIdx.x) * DISCO_BC_JOB;
    const AnnDev& G = S.ann[J[0]];
    for (int i = threadIdx.x; i < G.np; i += blockDim.x) bc_copy_cell<Ph>(S, G, G.off + i, J[1], J[2], J[3]);
}

__global__ void k_diag_time(DevState S) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < S.nruns) S.t_avg[r] += S.dt[r];
}

__global__ void k_planets(DevState S) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= S.nruns) return;
    planets_run(S, r);
    // open the next step's CFL reduction (filled by the second substep)
    const RunDev& Rn = S.runs[r];
    double d0 = Rn.ph.max_dt / Rn.ph.cfl;
    if (d0 <= 0.0) d0 = 1.0e100;
    S.dtbits[r] = (unsigned long long)__double_as_longlong(d0);
}

// ------------------------------------------------------------------ launchers
static inline int nblk(int n, int b) { return (n + b - 1) / b; }

#define DISCO_LAUNCH(hook, name, kern, grid, block, st, ...)       \
    do {                                                           \
        if (hook) hook->before(hook->ctx, name);                   \
        hipLaunchKernelGGL(kern, grid, block, 0, st, __VA_ARGS__); \
        if (hook) hook->after(hook->ctx, name);                    \
    } while (0)

// the accretion sinks' per-planet sums of a substep (after the cell update)
static void sink_sums(const DevState& S, double RK, int first, hipStream_t st, const LaunchHook* hook) {
    DISCO_LAUNCH(hook, "k_sink_ann", k_sink_ann, dim3(S.nann), dim3(256), st, S);
    DISCO_LAUNCH(hook, "k_sink_run", k_sink_run, dim3(S.nruns), dim3(64), st, S, RK, first);
}

template <class Ph>
struct Launch {
    // Full CFL reduction over the current state (first step, getmindt calls).
    static void mindt(const DevState& S, hipStream_t st, const LaunchHook* hook) {
        const int B = 256;
        DISCO_LAUNCH(hook, "k_dt_init", k_dt_init, dim3(nblk(S.nruns, 64)), dim3(64), st, S);
        DISCO_LAUNCH(hook, "k_mindt", k_mindt<Ph>, dim3(nblk(S.ncell, B)), dim3(B), st, S);
    }
    // With a Fork (graph capture), the cell update of the conserved state
    // (and GRMHD's geometric sources) runs on a second stream beside the EMF
    // averaging and the Stokes update: after the fluxes the two branches
    // touch disjoint arrays (cons, rkcons and the moved mesh on one; EMF
    // sums, edge EMFs and face fluxes on the other), and the flip and
    // face-to-cell kernels wait for both.  Without a Fork every kernel runs in
    // order on `st`.
    static void substep(const DevState& S, double RK, double dtfac, int first, int want_dt, const StepShape& sh,
                        hipStream_t st, const LaunchHook* hook, const Fork* fk) {
        auto B = [&](int k) { return sh.blk[k] > 0 ? sh.blk[k] : 256; };
        auto G = [&](int k) { return dim3(nblk(S.ncell, B(k))); };
        const bool fork = fk != nullptr && Ph::MHD && sh.ct;
        hipStream_t su = fork ? fk->st2 : st;  // stream of the cell-update branch
        if (fork) {
            (void)hipEventRecord(fk->ev[0], st);
            (void)hipStreamWaitEvent(su, fk->ev[0], 0);
        }
        // GRMHD's sources read only the substep's prims and mesh
        if (Ph::MHD && sh.ct && Ph::SOURCE_PASS)
            DISCO_LAUNCH(hook, "k_source", k_source<Ph>, G(KC_CT), dim3(B(KC_CT)), su, S, dtfac);
        // partners: the step's first substep (first == 1) keeps the last
        // substep's (the mesh has not moved since); every later substep
        // searches (a moving substep before it moved the mesh)
        DISCO_LAUNCH(hook, "k_stage_a", k_stage_a<Ph>, G(KC_STAGE), dim3(B(KC_STAGE)), st, S, first, first != 1);
        if (first && sh.wmean) DISCO_LAUNCH(hook, "k_wmean", k_wmean, dim3(S.nann), dim3(64), st, S);
        DISCO_LAUNCH(hook, "k_plm_trans", k_plm_trans<Ph>, G(KC_PLM), dim3(B(KC_PLM)), st, S, sh.has_z);
        if constexpr (!Ph::MHD && Ph::NEWTONIAN) {  // euler.c's viscous fluxes
            if (sh.visc && sh.dampc)
                DISCO_LAUNCH(hook, "k_flux", (k_flux<Ph, true, true>), G(KC_FLUX), dim3(B(KC_FLUX)), st, S, dtfac);
            else if (sh.visc)
                DISCO_LAUNCH(hook, "k_flux", (k_flux<Ph, true>), G(KC_FLUX), dim3(B(KC_FLUX)), st, S, dtfac);
            else if (sh.dampc)
                DISCO_LAUNCH(hook, "k_flux", (k_flux<Ph, false, true>), G(KC_FLUX), dim3(B(KC_FLUX)), st, S, dtfac);
            else
                DISCO_LAUNCH(hook, "k_flux", (k_flux<Ph, false>), G(KC_FLUX), dim3(B(KC_FLUX)), st, S, dtfac);
        } else {
            if (sh.dampc)
                DISCO_LAUNCH(hook, "k_flux", (k_flux<Ph, false, true>), G(KC_FLUX), dim3(B(KC_FLUX)), st, S, dtfac);
            else
                DISCO_LAUNCH(hook, "k_flux", (k_flux<Ph, false>), G(KC_FLUX), dim3(B(KC_FLUX)), st, S, dtfac);
        }
        if (Ph::MHD && sh.ct) {
            if (fork) {
                (void)hipEventRecord(fk->ev[1], st);
                (void)hipStreamWaitEvent(su, fk->ev[1], 0);
            }
            if (Ph::NEWTONIAN && sh.sink)
                DISCO_LAUNCH(hook, "k_update_a", (k_update_a<Ph, true>), G(KC_CT), dim3(B(KC_CT)), su, S, RK, dtfac,
                             first);
            else
                DISCO_LAUNCH(hook, "k_update_a", (k_update_a<Ph, false>), G(KC_CT), dim3(B(KC_CT)), su, S, RK, dtfac,
                             first);
            if (sh.accrete) sink_sums(S, RK, first, su, hook);
            DISCO_LAUNCH(hook, "k_emf_sum", k_emf_sum, G(KC_EMF), dim3(B(KC_EMF)), st, S, sh.has_z);
            DISCO_LAUNCH(hook, "k_emf_face", k_emf_face, G(KC_EMF), dim3(B(KC_EMF)), st, S, sh.has_z);
            DISCO_LAUNCH(hook, "k_ct_update", k_ct_update, G(KC_CT), dim3(B(KC_CT)), st, S, RK, dtfac, first,
                         sh.has_z);
            if (fork) {
                (void)hipEventRecord(fk->ev[2], su);
                (void)hipStreamWaitEvent(st, fk->ev[2], 0);
            }
            if (first) {
                DISCO_LAUNCH(hook, "k_flip_flag", k_flip_flag, G(KC_FLIP), dim3(B(KC_FLIP)), st, S, sh.has_z);
                DISCO_LAUNCH(hook, "k_flip_chain", k_flip_chain, G(KC_FLIP), dim3(B(KC_FLIP)), st, S, sh.has_z);
            }
            if (Ph::C2P_GUESS && sh.c2p_learned)
                DISCO_LAUNCH(hook, "k_update_b_learned", k_update_b_learned, G(KC_UPDATE), dim3(B(KC_UPDATE)), st, S,
                             first, want_dt);
            else
                DISCO_LAUNCH(hook, "k_update_b", k_update_b<Ph>, G(KC_UPDATE), dim3(B(KC_UPDATE)), st, S, first,
                             want_dt);
        } else {
            if (Ph::NEWTONIAN && sh.sink)
                DISCO_LAUNCH(hook, "k_update", (k_update<Ph, true>), G(KC_UPDATE), dim3(B(KC_UPDATE)), st, S, RK,
                             dtfac, first, want_dt);
            else
                DISCO_LAUNCH(hook, "k_update", (k_update<Ph, false>), G(KC_UPDATE), dim3(B(KC_UPDATE)), st, S, RK,
                             dtfac, first, want_dt);
            if (sh.accrete) sink_sums(S, RK, first, st, hook);
        }
        // boundary_trans (onestep.c:166-173): zero-gradient ghosts level by level, then the periodic z copy
        for (int l = 0; l < sh.bc_nlev; ++l)
            DISCO_LAUNCH(hook, "k_bc_copy", k_bc_copy<Ph>, dim3(sh.bc_off[l + 1] - sh.bc_off[l]), dim3(B(KC_COPY)), st,
                         S, sh.bc_off[l]);
        if (sh.zcopy) DISCO_LAUNCH(hook, "k_zcopy", k_zcopy<Ph>, G(KC_COPY), dim3(B(KC_COPY)), st, S, first);
        if (first) DISCO_LAUNCH(hook, "k_planets", k_planets, dim3(nblk(S.nruns, 64)), dim3(64), st, S);
    }
};

// the step's add_diagnostics (timestep.c:51), after the second substep
void launch_diag(const DevState& S, int nann, int nq, hipStream_t st, const LaunchHook* hook) {
    (void)nq;
    DISCO_LAUNCH(hook, "k_diag", k_diag, dim3(nann), dim3(256), st, S);
    DISCO_LAUNCH(hook, "k_diag_time", k_diag_time, dim3(nblk(S.nruns, 64)), dim3(64), st, S);
}

void launch_sink_reset(const DevState& S, int run, hipStream_t st) {
    hipLaunchKernelGGL(k_sink_reset, dim3(1), dim3(64), 0, st, S, run);
}

void launch_report(const DevState& S, int a0, int nann, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_report, dim3(nann), dim3(256), 0, st, S, a0, out);
}

void launch_report_sum(const double* part, int Nr, int Nz, double* sums, hipStream_t st) {
    hipLaunchKernelGGL(k_report_sum, dim3(1), dim3(64), 0, st, part, Nr, Nz, sums);
}

// partners of every cell on the current mesh (after it is set from the host)
void launch_partners(const DevState& S, hipStream_t st) {
    hipLaunchKernelGGL(k_partners, dim3(nblk(S.ncell, 256)), dim3(256), 0, st, S);
}

// dt = CFL x running minimum (misc.c:79); check: then check_dt's clip
void launch_dt_final(const DevState& S, int check, hipStream_t st, const LaunchHook* hook) {
    DISCO_LAUNCH(hook, "k_dt_final", k_dt_final, dim3(nblk(S.nruns, 64)), dim3(64), st, S, check);
}

// Face-increment components per cell that the flux kernel of this shape
// stores (its policy's NF), for the host's check of the buffers it sized.
int launch_nf(const StepShape& sh) {
    if (sh.hydro == DISCO_HYDRO_MHD) return sh.nn == 1 ? PhysMHD1::NF : PhysMHD::NF;
    if (sh.hydro == DISCO_HYDRO_GRMHD) return PhysGRMHD::NF;
    if (sh.hydro == DISCO_HYDRO_GRMHD2) return PhysGRMHD2::NF;
    if (sh.hydro == DISCO_HYDRO_GREULER) return PhysGREuler::NF;
    return sh.nn == 1 ? PhysEuler1::NF : PhysEuler::NF;
}

void launch_mindt(const StepShape& sh, const DevState& S, hipStream_t st, const LaunchHook* hook) {
    if (sh.hydro == DISCO_HYDRO_MHD && sh.nn == 1)
        Launch<PhysMHD1>::mindt(S, st, hook);
    else if (sh.hydro == DISCO_HYDRO_MHD)
        Launch<PhysMHD>::mindt(S, st, hook);
    else if (sh.hydro == DISCO_HYDRO_GRMHD)
        Launch<PhysGRMHD>::mindt(S, st, hook);
    else if (sh.hydro == DISCO_HYDRO_GRMHD2)
        Launch<PhysGRMHD2>::mindt(S, st, hook);
    else if (sh.hydro == DISCO_HYDRO_GREULER)
        Launch<PhysGREuler>::mindt(S, st, hook);
    else if (sh.nn == 1)
        Launch<PhysEuler1>::mindt(S, st, hook);
    else
        Launch<PhysEuler>::mindt(S, st, hook);
}

void launch_substep(const StepShape& sh, const DevState& S, double RK, double dtfac, int first, int want_dt,
                    hipStream_t st, const LaunchHook* hook, const Fork* fk) {
    if (sh.hydro == DISCO_HYDRO_MHD && sh.nn == 1)
        Launch<PhysMHD1>::substep(S, RK, dtfac, first, want_dt, sh, st, hook, fk);
    else if (sh.hydro == DISCO_HYDRO_MHD)
        Launch<PhysMHD>::substep(S, RK, dtfac, first, want_dt, sh, st, hook, fk);
    else if (sh.hydro == DISCO_HYDRO_GRMHD)
        Launch<PhysGRMHD>::substep(S, RK, dtfac, first, want_dt, sh, st, hook, fk);
    else if (sh.hydro == DISCO_HYDRO_GRMHD2)
        Launch<PhysGRMHD2>::substep(S, RK, dtfac, first, want_dt, sh, st, hook, fk);
    else if (sh.hydro == DISCO_HYDRO_GREULER)
        Launch<PhysGREuler>::substep(S, RK, dtfac, first, want_dt, sh, st, hook, fk);
    else if (sh.nn == 1)
        Launch<PhysEuler1>::substep(S, RK, dtfac, first, want_dt, sh, st, hook, fk);
    else
        Launch<PhysEuler>::substep(S, RK, dtfac, first, want_dt, sh, st, hook, fk);
}

}  // namespace disco
