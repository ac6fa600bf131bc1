// HIP kernels of one RK substep of the moving-mesh finite-volume update for
// the Newtonian hydro module (gfx950, FP64).
//
// Reference sequence (onestep.c:86-187) and where each step lives here:
//   set_wcell, adjust_RK_cons, plm_phi       -> k_stage_a
//   setup_faces + plm_trans (r)              -> k_stage_a (partner search) + k_plm_r
//   phi_flux, trans_flux (r)                 -> k_flux
//   add_source, move_cells, clean_pi,
//   calc_dp, calc_prim, boundary_trans       -> k_update
//   movePlanets + clean_pi (planets)         -> k_planets
//   getmindt                                 -> fused into the last substep's
//                                               k_update (k_mindt at start-up)
//
// Overlap faces (faces_alt.c:36-179).  The reference merges the two cyclic,
// sorted front-edge lists of an annulus pair into exactly Np_a + Np_b faces,
// each face ending at the front edge of one of its two cells (LRtype).  Here
// every cell owns the one face of each neighbouring track that ends at its
// own front edge; its partner is the neighbour-annulus cell that is current
// when the reference's merge consumes this front: the first outer cell whose
// front is >= mine (ties go to the inner list first) or the first inner cell
// whose front is > mine.  Both are binary searches over the neighbour's
// fronts in rotated (ascending) order.  The faces a cell shares but does not
// own are those owned by the neighbour cells whose fronts fall inside its
// span: a cyclic range bounded by its own partner and its predecessor's.
// Each face increment is computed once, by its owner, stored per cell, and
// gathered by both cells: deterministic, no atomics, no face lists in HBM.
#include <hip/hip_runtime.h>

#include "device_state.h"
#include "phys_euler.h"

namespace disco {

#define DEV __device__ static inline __attribute__((always_inline))

// Blocks are dispatched round-robin over the 8 XCDs, each with its own L2.
// Remap so that XCD x works on one contiguous band of cells: partner cells in
// the neighbouring annuli (about Np cells away) are then mostly in the same
// L2 as the block that needs them.
DEV int cell_index() {
    const int nb = gridDim.x, b = blockIdx.x;
    const int x = b & 7, i = b >> 3;
    const int per = nb >> 3, rem = nb & 7;
    const int lb = x * per + (x < rem ? x : rem) + i;
    return lb * blockDim.x + threadIdx.x;
}

// plm.c:6-13
DEV double minmod(double a, double b, double c) {
    double m = a;
    if (a * b < 0.0) m = 0.0;
    if (fabs(b) < fabs(m)) m = b;
    if (b * c < 0.0) m = 0.0;
    if (fabs(c) < fabs(m)) m = c;
    return m;
}

// Count of fronts in annulus B strictly below x (strict = true) or at or
// below x (strict = false), in B's rotated order; returns the array index of
// the cell at that rotated position (mod Np).
DEV int front_search(const double* __restrict__ piph, int off, int n, int rot, double x, bool strict) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        int idx = rot + mid;
        if (idx >= n) idx -= n;
        double f = piph[off + idx];
        bool below = strict ? (f < x) : (f <= x);
        if (below)
            lo = mid + 1;
        else
            hi = mid;
    }
    int idx = rot + lo;
    while (idx >= n) idx -= n;
    return idx;
}

// Geometry of the r-face between L (inner annulus) and R (outer annulus),
// faces_alt.c:10-34 and :146-168, plus the phi offsets of plm.c:86-91.
struct RFace {
    double phic, dA, dpL, dpR;
};

DEV RFace r_face(double piphL, double dphiL, double piphR, double dphiR, int LR, double r_face,
                 double dz, double Pmax) {
    RFace F;
    double xp1 = LR == 0 ? piphL : piphR;
    double dd = piphL - dphiL - piphR + dphiR;
    while (dd > .5 * Pmax) dd -= Pmax;
    while (dd < -.5 * Pmax) dd += Pmax;
    double xm1 = dd < 0. ? piphR - dphiR : piphL - dphiL;
    double dp = get_dp(xp1, xm1, Pmax);
    F.phic = get_dp(xp1, .5 * dp, Pmax);
    F.dA = r_face * dp * dz;
    double phiL = piphL - .5 * dphiL;
    double phiR = piphR - .5 * dphiR;
    F.dpL = get_signed_dp(F.phic, phiL, Pmax);
    F.dpR = get_signed_dp(phiR, F.phic, Pmax);
    return F;
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

// misc.c:39-84 for one real cell: mindt(prim, w = mean of the two phi-face
// mesh speeds, xp, xm)
template <int NQ>
DEV double cell_mindt(const DevState& S, const AnnDev& A, const PhysParams& P, int c, const double* prim) {
    int i = c - A.off;
    int cm = A.off + (i == 0 ? A.np - 1 : i - 1);
    double phip = S.piph[c];
    double phim = phip - S.dphi[c];
    double xp[3] = {A.rp, phip, A.zp};
    double xm[3] = {A.rm, phim, A.zm};
    double w = .5 * (S.wiph[cm] + S.wiph[c]);
    return euler::mindt(P, prim, w, xp, xm);
}

template <int NQ>
__global__ __launch_bounds__(256) void k_mindt(DevState S) {
    int c = cell_index();
    int run = -1;
    double dtc = 1.0e300;
    if (c < S.ncell) {
        const AnnDev& A = S.ann[S.cell_ann[c]];
        run = A.run;
        if (A.flags & ANN_REAL) {
            double prim[NQ];
            _Pragma("unroll") for (int q = 0; q < NQ; ++q) prim[q] = S.prim[(size_t)q * S.ncell + c];
            dtc = cell_mindt<NQ>(S, A, S.runs[run].ph, c, prim);
        }
    }
    dt_reduce(S, run, dtc);
}

__global__ void k_dt_final(DevState S) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= S.nruns) return;
    double m = __longlong_as_double((long long)S.dtbits[r]);
    S.dt[r] = m * S.runs[r].ph.cfl;  // misc.c:79
}

// ------------------------------------------------------------------ stage A
// set_wcell (misc.c:101-168, first substep), adjust_RK_cons (misc.c:226-244,
// with timestep.c:18-24's copy folded into the first substep), plm_phi
// (plm.c:15-51) and the partner search of the overlap faces.
template <int NQ>
__global__ __launch_bounds__(256) void k_stage_a(DevState S, double RK, int first) {
    int c = cell_index();
    if (c >= S.ncell) return;
    const AnnDev& A = S.ann[S.cell_ann[c]];
    const PhysParams& P = S.runs[A.run].ph;
    const size_t N = S.ncell;
    int n = A.np;
    int i = c - A.off;
    int cm = A.off + (i == 0 ? n - 1 : i - 1);
    int cp = A.off + (i + 1 == n ? 0 : i + 1);

    if (first) {
        double w = 0.0;
        if (P.mesh_motion == 1) {
            w = .5 * (S.prim[UPP * N + c] + S.prim[UPP * N + cp]);
        } else if (P.mesh_motion == 4) {
            double phip = S.piph[c];
            double phim = phip - S.dphi[c];
            double x[3] = {A.rc, 0.5 * (phim + phip), A.zc};
            w = mesh_om(P, x);
        }
        S.wiph[c] = w;
    }
    if (A.flags & ANN_REAL) {
        if (first) {
            _Pragma("unroll") for (int q = 0; q < NQ; ++q) S.rkcons[q * N + c] = S.cons[q * N + c];
        } else {
            _Pragma("unroll") for (int q = 0; q < NQ; ++q)
                S.cons[q * N + c] = (1. - RK) * S.cons[q * N + c] + RK * S.rkcons[q * N + c];
        }
    }
    double dpL = S.dphi[cm];
    double dpC = S.dphi[c];
    double dpR = S.dphi[cp];
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
        double pL = S.prim[q * N + cm];
        double pC = S.prim[q * N + c];
        double pR = S.prim[q * N + cp];
        double sL = pC - pL;
        sL /= .5 * (dpC + dpL);
        double sR = pR - pC;
        sR /= .5 * (dpR + dpC);
        double sC = pR - pL;
        sC /= .5 * (dpL + dpR) + dpC;
        S.gradp[q * N + c] = minmod(P.plm * sL, sC, P.plm * sR);
    }
    double front = S.piph[c];
    if (A.outer >= 0) {
        const AnnDev& B = S.ann[A.outer];
        S.lbo[c] = front_search(S.piph, B.off, B.np, S.rot[A.outer], front, true);
    }
    if (A.inner >= 0) {
        const AnnDev& W = S.ann[A.inner];
        S.ubi[c] = front_search(S.piph, W.off, W.np, S.rot[A.inner], front, false);
    }
}

// ------------------------------------------------------------------ plm_trans
// plm.c:56-195 for dim = 1: area-weighted mean of the face slopes, divided by
// the cell's total r-face area, then the per-face limiter (order
// independent: zero on any sign change, else the smallest |PLM*S|).
template <int NQ>
struct SlopeAcc {
    double sum[NQ];
    double smin[NQ];
    unsigned pos, neg, zero;  // bit q set when some face slope of field q is >0, <0, ==0
};

template <int NQ>
DEV void acc_init(SlopeAcc<NQ>& a) {
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
        a.sum[q] = 0.0;
        a.smin[q] = 0.0;
    }
    a.pos = a.neg = a.zero = 0u;
}

template <int NQ>
DEV void acc_add(SlopeAcc<NQ>& a, const double* S, double dA, bool firstface) {
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
        a.sum[q] += S[q] * dA;
        a.pos |= (S[q] > 0.0 ? 1u : 0u) << q;
        a.neg |= (S[q] < 0.0 ? 1u : 0u) << q;
        a.zero |= (S[q] == 0.0 ? 1u : 0u) << q;
        a.smin[q] = (firstface || fabs(S[q]) < fabs(a.smin[q])) ? S[q] : a.smin[q];
    }
}

template <int NQ>
DEV void face_slope(const DevState& S, int cL, int cR, const RFace& F, double dxL, double dxR, double* Sq) {
    const size_t N = S.ncell;
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
        double WL = S.prim[q * N + cL] + F.dpL * S.gradp[q * N + cL];
        double WR = S.prim[q * N + cR] - F.dpR * S.gradp[q * N + cR];
        Sq[q] = (WR - WL) / (dxR + dxL);
    }
}

template <int NQ>
__global__ __launch_bounds__(256) void k_plm_r(DevState S) {
    int c = cell_index();
    if (c >= S.ncell) return;
    const AnnDev& A = S.ann[S.cell_ann[c]];
    const PhysParams& P = S.runs[A.run].ph;
    const size_t N = S.ncell;
    const double Pmax = P.phi_max;
    if (P.plm == 0.0) {
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) S.gradr[q * N + c] = 0.0;
        return;
    }
    int n = A.np;
    int i = c - A.off;
    int cm = A.off + (i == 0 ? n - 1 : i - 1);
    double phip = S.piph[c];
    double dph = S.dphi[c];
    double phim = phip - dph;
    double dz = A.zp - A.zm;

    SlopeAcc<NQ> acc;
    acc_init(acc);
    bool firstface = true;
    double Sq[NQ];
    if (A.outer >= 0) {  // this cell is L of the (this | outer) track
        const AnnDev& B = S.ann[A.outer];
        int nb = B.np;
        int own = S.lbo[c];
        int start = S.lbo[cm];
        int cnt = own - start;
        if (cnt < 0) cnt += nb;
        int cb = B.off + own;
        RFace F = r_face(phip, dph, S.piph[cb], S.dphi[cb], 0, A.rp, dz, Pmax);
        face_slope<NQ>(S, c, cb, F, A.dxL, A.dxR, Sq);
        acc_add(acc, Sq, F.dA, firstface);
        firstface = false;
        for (int t = 0; t < cnt; ++t) {
            int b = start + t;
            if (b >= nb) b -= nb;
            cb = B.off + b;
            F = r_face(phip, dph, S.piph[cb], S.dphi[cb], 1, A.rp, dz, Pmax);
            face_slope<NQ>(S, c, cb, F, A.dxL, A.dxR, Sq);
            acc_add(acc, Sq, F.dA, false);
        }
    }
    if (A.inner >= 0) {  // this cell is R of the (inner | this) track
        const AnnDev& W = S.ann[A.inner];
        int nw = W.np;
        int own = S.ubi[c];
        int start = S.ubi[cm];
        int cnt = own - start;
        if (cnt < 0) cnt += nw;
        int ca = W.off + own;
        RFace F = r_face(S.piph[ca], S.dphi[ca], phip, dph, 1, W.rp, dz, Pmax);
        face_slope<NQ>(S, ca, c, F, W.dxL, W.dxR, Sq);
        acc_add(acc, Sq, F.dA, firstface);
        firstface = false;
        for (int t = 0; t < cnt; ++t) {
            int a = start + t;
            if (a >= nw) a -= nw;
            ca = W.off + a;
            F = r_face(S.piph[ca], S.dphi[ca], phip, dph, 0, W.rp, dz, Pmax);
            face_slope<NQ>(S, ca, c, F, W.dxL, W.dxR, Sq);
            acc_add(acc, Sq, F.dA, false);
        }
    }
    // plm.c:112-143: total r-face area of the cell
    double dpp = get_dp(phip, phim, Pmax);
    double dAp = A.rp * dpp * dz;
    double dAm = A.rm * dpp * dz;
    double dAtot = dAp + dAm;
    if (A.inner < 0) dAtot = dAp;
    if (A.outer < 0) dAtot = dAm;
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
        double g = acc.sum[q] / dAtot;
        if (firstface) {
            g = 0.0;
        } else if (g != 0.0) {
            bool flip = g > 0.0 ? (((acc.neg | acc.zero) >> q) & 1u) : (((acc.pos | acc.zero) >> q) & 1u);
            if (flip) {
                g = 0.0;
            } else {
                double lim = P.plm * acc.smin[q];
                if (fabs(lim) < fabs(g)) g = lim;
            }
        }
        S.gradr[q * N + c] = g;
    }
}

// ------------------------------------------------------------------ fluxes
// riemann_phi (riemann.c:39-112) on the cell's phi face, riemann_trans
// (riemann.c:114-213) on its owned r-faces; solve_riemann's increment
// (F - w U*) dA dt is stored per owner cell.
template <int NQ>
DEV void r_face_flux(const DevState& S, const PhysParams& P, int cL, int cR, const RFace& F, double dxL,
                     double dxR, double r_face, double zc, double dt, double* dU) {
    const size_t N = S.ncell;
    double primL[NQ], primR[NQ];
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
        primL[q] = S.prim[q * N + cL] + S.gradr[q * N + cL] * dxL + S.gradp[q * N + cL] * F.dpL;
        primR[q] = S.prim[q * N + cR] - S.gradr[q * N + cR] * dxR - S.gradp[q * N + cR] * F.dpR;
    }
    double x[3] = {r_face, F.phic, zc};
    const double n[3] = {1.0, 0.0, 0.0};
    euler::riemann<NQ>(P, primL, primR, x, n, 0.0, F.dA * dt, dU);
}

template <int NQ>
__global__ __launch_bounds__(256) void k_flux(DevState S, double dtfac) {
    int c = cell_index();
    if (c >= S.ncell) return;
    const AnnDev& A = S.ann[S.cell_ann[c]];
    if (!(A.flags & (ANN_PHIFLUX | ANN_OUTFLUX | ANN_INFLUX))) return;
    const PhysParams& P = S.runs[A.run].ph;
    const size_t N = S.ncell;
    const double Pmax = P.phi_max;
    const double dt = S.dt[A.run] * dtfac;
    int n = A.np;
    int i = c - A.off;
    double dU[NQ];
    double phip = S.piph[c];
    double dph = S.dphi[c];
    double dz = A.zp - A.zm;

    if (A.flags & ANN_PHIFLUX) {
        int cp = A.off + (i + 1 == n ? 0 : i + 1);
        double primL[NQ], primR[NQ];
        double dphR = S.dphi[cp];
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
            primL[q] = S.prim[q * N + c] + .5 * S.gradp[q * N + c] * dph;
            primR[q] = S.prim[q * N + cp] - .5 * S.gradp[q * N + cp] * dphR;
        }
        double x[3] = {A.rc, phip, A.zc};
        const double nv[3] = {0.0, 1.0, 0.0};
        double dA = (A.rp - A.rm) * dz;  // get_dA(xp, xm, 0)
        double w = A.rc * S.wiph[c];     // get_scale_factor(x, 0) * wiph
        euler::riemann<NQ>(P, primL, primR, x, nv, w, dA * dt, dU);
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) S.fphi[q * N + c] = dU[q];
    }
    if (A.flags & ANN_OUTFLUX) {
        const AnnDev& B = S.ann[A.outer];
        int cb = B.off + S.lbo[c];
        RFace F = r_face(phip, dph, S.piph[cb], S.dphi[cb], 0, A.rp, dz, Pmax);
        r_face_flux<NQ>(S, P, c, cb, F, A.dxL, A.dxR, A.rp, A.zc, dt, dU);
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) S.fout[q * N + c] = dU[q];
    }
    if (A.flags & ANN_INFLUX) {
        const AnnDev& W = S.ann[A.inner];
        int ca = W.off + S.ubi[c];
        RFace F = r_face(S.piph[ca], S.dphi[ca], phip, dph, 1, W.rp, dz, Pmax);
        r_face_flux<NQ>(S, P, ca, c, F, W.dxL, W.dxR, W.rp, W.zc, dt, dU);
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) S.fin[q * N + c] = dU[q];
    }
}

// ------------------------------------------------------------------ update
// Conservative update from the gathered face increments (riemann.c:314-317),
// add_source (misc.c:550-593), move_cells/clean_pi/calc_dp (misc.c:266-298,
// 8-36; first substep only), calc_prim (misc.c:300-335) and the fixed
// boundary (Boundary/fixed.c:11-60, boundary_functions.c:161-266).
template <int NQ>
DEV double update_cell(const DevState& S, int c, double dtfac, int first, int want_dt);

// The second substep also produces the next step's CFL candidate from the
// prims it has just written (misc.c:39-84 runs on exactly this state at the
// start of the next step, with this step's mesh speeds).
template <int NQ>
__global__ __launch_bounds__(256) void k_update(DevState S, double dtfac, int first, int want_dt) {
    int c = cell_index();
    int run = -1;
    double dtc = 1.0e300;
    if (c < S.ncell) {
        run = S.ann[S.cell_ann[c]].run;
        dtc = update_cell<NQ>(S, c, dtfac, first, want_dt);
    }
    if (want_dt) dt_reduce(S, run, dtc);
}

template <int NQ>
DEV double update_cell(const DevState& S, int c, double dtfac, int first, int want_dt) {
    const AnnDev& A = S.ann[S.cell_ann[c]];
    const RunDev& Rn = S.runs[A.run];
    const PhysParams& P = Rn.ph;
    const size_t N = S.ncell;
    const double Pmax = P.phi_max;
    const double dt = S.dt[A.run] * dtfac;
    int n = A.np;
    int i = c - A.off;
    int cm = A.off + (i == 0 ? n - 1 : i - 1);
    double phip = S.piph[c];
    double dph = S.dphi[c];

    double prim[NQ], cons[NQ];
    if (A.flags & ANN_REAL) {
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
            cons[q] = S.cons[q * N + c];
            prim[q] = S.prim[q * N + c];
        }
        // phi faces i-1 (this cell is R) and i (this cell is L)
        if (A.flags & ANN_PHIFLUX) {
            if (i == 0) {
                _Pragma("unroll") for (int q = 0; q < NQ; ++q) cons[q] -= S.fphi[q * N + c];
                _Pragma("unroll") for (int q = 0; q < NQ; ++q) cons[q] += S.fphi[q * N + cm];
            } else {
                _Pragma("unroll") for (int q = 0; q < NQ; ++q) cons[q] += S.fphi[q * N + cm];
                _Pragma("unroll") for (int q = 0; q < NQ; ++q) cons[q] -= S.fphi[q * N + c];
            }
        }
        // (inner | this) track: this cell is R, gains the flux
        if (A.inner >= 0 && (A.flags & ANN_INFLUX)) {
            const AnnDev& W = S.ann[A.inner];
            _Pragma("unroll") for (int q = 0; q < NQ; ++q) cons[q] += S.fin[q * N + c];
            int start = S.ubi[cm];
            int cnt = S.ubi[c] - start;
            if (cnt < 0) cnt += W.np;
            for (int t = 0; t < cnt; ++t) {
                int a = start + t;
                if (a >= W.np) a -= W.np;
                _Pragma("unroll") for (int q = 0; q < NQ; ++q) cons[q] += S.fout[q * N + W.off + a];
            }
        }
        // (this | outer) track: this cell is L, loses the flux
        if (A.outer >= 0 && (A.flags & ANN_OUTFLUX)) {
            const AnnDev& B = S.ann[A.outer];
            _Pragma("unroll") for (int q = 0; q < NQ; ++q) cons[q] -= S.fout[q * N + c];
            int start = S.lbo[cm];
            int cnt = S.lbo[c] - start;
            if (cnt < 0) cnt += B.np;
            for (int t = 0; t < cnt; ++t) {
                int b = start + t;
                if (b >= B.np) b -= B.np;
                _Pragma("unroll") for (int q = 0; q < NQ; ++q) cons[q] -= S.fin[q * N + B.off + b];
            }
        }
        // add_source with the pre-move geometry and the substep's prims
        double xp[3] = {A.rp, phip, A.zp};
        double xm[3] = {A.rm, phip - dph, A.zm};
        double dV = get_dV(xp, xm, Pmax);
        euler::source(P, prim, cons, xp, xm, dV * dt);
        for (int p = 0; p < P.npl; ++p) {
            Planet pl = Rn.planets[p];
            pl.phi = S.planet_phi[A.run * DISCO_MAX_PLANETS + p];
            euler::planet_src(P, pl, prim, cons, xp, xm, dV * dt);
        }
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) S.cons[q * N + c] = cons[q];
    }

    // moving mesh (first substep): every annulus, ghosts included
    if (first) {
        double pn = phip + S.wiph[c] * dt;
        while (pn > Pmax) pn -= Pmax;
        while (pn < 0.0) pn += Pmax;
        double pm = S.piph[cm] + S.wiph[cm] * dt;
        while (pm > Pmax) pm -= Pmax;
        while (pm < 0.0) pm += Pmax;
        phip = pn;
        dph = get_dp(pn, pm, Pmax);
        S.piph_n[c] = pn;
        S.dphi_n[c] = dph;
        if (pn < pm) S.rot[S.cell_ann[c]] = i;
    }

    if (A.flags & ANN_REAL) {
        double phim = phip - dph;
        double xp[3] = {A.rp, phip, A.zp};
        double xm[3] = {A.rm, phim, A.zm};
        double dV = get_dV(xp, xm, Pmax);
        double x[3] = {A.rc, 0.5 * (phim + phip), A.zc};
        euler::cons2prim<NQ>(P, cons, prim, x, dV);
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) S.prim[q * N + c] = prim[q];
        if (want_dt) return cell_mindt<NQ>(S, A, P, c, prim);
    } else if (A.flags & ANN_GHOST_FILL) {
        double x[3];
        x[0] = A.rc_b;
        x[2] = A.zc_b;
        if (Rn.bc == DISCO_BC_FIXED)
            x[1] = phip - .5 * dph;
        else
            x[1] = 0.5 * (phip + (phip - dph));
        double p8[8];
        disco_initial(&Rn.ic, p8, x);
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) S.prim[q * N + c] = p8[q];
    }
    return 1.0e300;
}

// movePlanets (Planets/sun.c: none; seeded binary: phi += omega dt) then the
// planet wrap of clean_pi (misc.c:26-33)
__global__ void k_planets(DevState S) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= S.nruns) return;
    const RunDev& Rn = S.runs[r];
    double Pmax = Rn.ph.phi_max;
    double dt = S.dt[r];
    for (int p = 0; p < Rn.ph.npl; ++p) {
        double phi = S.planet_phi[r * DISCO_MAX_PLANETS + p];
        if (Rn.planets_mod == DISCO_PLANETS_BINARY) phi += Rn.planets[p].omega * dt;
        while (phi > Pmax) phi -= Pmax;
        while (phi < 0.0) phi += Pmax;
        S.planet_phi[r * DISCO_MAX_PLANETS + p] = phi;
    }
    S.t[r] += dt;  // timestep.c:47
    // open the next step's CFL reduction (filled by the second substep)
    double d0 = Rn.ph.max_dt / Rn.ph.cfl;
    if (d0 <= 0.0) d0 = 1.0e100;
    S.dtbits[r] = (unsigned long long)__double_as_longlong(d0);
}

// ------------------------------------------------------------------ launchers
static inline int nblk(int n, int b) { return (n + b - 1) / b; }

#define DISCO_LAUNCH(hook, name, kern, grid, block, st, ...)                       \
    do {                                                                           \
        if (hook) hook->before(hook->ctx, name);                                   \
        hipLaunchKernelGGL(kern, grid, block, 0, st, __VA_ARGS__);                 \
        if (hook) hook->after(hook->ctx, name);                                    \
    } while (0)

// Full CFL reduction over the current state (first step, getmindt calls).
void launch_mindt(const DevState& S, hipStream_t st, const LaunchHook* hook) {
    const int B = 256;
    DISCO_LAUNCH(hook, "k_dt_init", k_dt_init, dim3(nblk(S.nruns, 64)), dim3(64), st, S);
    DISCO_LAUNCH(hook, "k_mindt", k_mindt<5>, dim3(nblk(S.ncell, B)), dim3(B), st, S);
}

// dt = CFL x running minimum (misc.c:79)
void launch_dt_final(const DevState& S, hipStream_t st, const LaunchHook* hook) {
    DISCO_LAUNCH(hook, "k_dt_final", k_dt_final, dim3(nblk(S.nruns, 64)), dim3(64), st, S);
}

void launch_substep(const DevState& S, double RK, double dtfac, int first, int want_dt, hipStream_t st,
                    const LaunchHook* hook) {
    const int B = 256;
    dim3 g(nblk(S.ncell, B));
    DISCO_LAUNCH(hook, "k_stage_a", k_stage_a<5>, g, dim3(B), st, S, RK, first);
    DISCO_LAUNCH(hook, "k_plm_r", k_plm_r<5>, g, dim3(B), st, S);
    DISCO_LAUNCH(hook, "k_flux", k_flux<5>, g, dim3(B), st, S, dtfac);
    DISCO_LAUNCH(hook, "k_update", k_update<5>, g, dim3(B), st, S, dtfac, first, want_dt);
    if (first) DISCO_LAUNCH(hook, "k_planets", k_planets, dim3(nblk(S.nruns, 64)), dim3(64), st, S);
}

}  // namespace disco
