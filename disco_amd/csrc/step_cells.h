// Per-cell bodies of one RK substep of the moving-mesh finite-volume update
// (FP64), for 2D and 3D cylindrical grids.  Every function here works on one
// cell of the flat SoA state (device_state.h) and is __host__ __device__; the
// HIP kernels in hydro_step.hip are one-thread-per-cell wrappers around them.
//
// Reference sequence (onestep.c:86-187) and where each step lives:
//   set_wcell, adjust_RK_cons, plm_phi       -> stage_a_cell
//   setup_faces + plm_trans (r and z)        -> stage_a_cell (partner search) + plm_trans_cell
//   phi_flux, trans_flux (r and z)           -> flux_cell
//   CT: avg_Efields, subtract_advective,
//       update_B_fluxes, make_edge_adjust    -> ct_cells.h
//   add_source, move_cells, clean_pi,
//   calc_dp                                  -> update_a_cell
//   CT: check_flipped, flip_fluxes           -> ct_cells.h
//   B_faces_to_cells, calc_prim,
//   boundary_trans                           -> update_b_cell (+ zcopy_cell: periodic z)
//   getmindt                                 -> cell_mindt (fused into the last substep)
//
// Overlap faces (faces_alt.c:36-179).  The reference merges the two cyclic,
// sorted front-edge lists of an annulus pair into exactly Np_a + Np_b faces,
// each face ending at the front edge of one of its two cells (LRtype).  Here
// every cell owns the one face of each neighbouring track that ends at its
// own front edge; its partner is the neighbour-annulus cell that is current
// when the reference's merge consumes this front: the first upper cell whose
// front is >= mine (ties go to the lower list first) or the first lower cell
// whose front is > mine.  Both are binary searches over the neighbour's
// fronts in rotated (ascending) order.  The faces a cell shares but does not
// own are those owned by the neighbour cells whose fronts fall inside its
// span: a cyclic range bounded by its own partner and its predecessor's.
// Each face increment is computed once, by its owner, stored per cell, and
// gathered by both cells: deterministic, no atomics, no face lists in HBM.
// One code path serves the r-tracks (j | j+1) and the z-tracks (k | k+1),
// selected by the template direction D.
#pragma once

#include "device_state.h"
#include "phys_euler.h"
#include "phys_mhd.h"
#include "phys_greuler.h"
#include "phys_grmhd_fv.h"

namespace disco {

// ------------------------------------------------------------------ physics
// Field of slot s of a cell's stored face increments: the first five
// conserved fields, then (after a gap of the field components that
// constrained transport does not store) the passive scalars.
DHD constexpr int fslot(int s, int gap) { return s < 5 ? s : s + gap; }

// The Hydro module's per-face and per-cell leaves (hydro.h:6-35) as policies.
// NN: passive scalars (NUM_N, paul.h:21 NUM_Q = NUM_C + NUM_N), advected as
// prim[q] * rho at q = NUM_C.. (euler.c:60-62, 97-99, 136-138, 166-168).
template <int NN>
struct PhysEulerT {
    static constexpr int NC = 5;  // NUM_C: index of the first passive scalar
    static constexpr int NQ = NC + NN;
    static constexpr int NF = NQ;  // stored face-increment components
    static constexpr int FGAP = 0;  // slot s of the stored increments is field fslot(s, FGAP)
    static constexpr bool MHD = false;
    static constexpr bool C2P_GUESS = false;  // cons2prim reads the previous prims
    static constexpr bool SOURCE_PASS = false;  // sources are added inline in update_a_cell
    template <bool DC = false>
    DHD FaceEMF riemann(const PhysParams& P, const double* pL, const double* pR, const double* x, const double* n,
                        double w, double dAdt, int dim, double* dU) {
        (void)dim;
        euler::riemann<NQ, DC>(P, pL, pR, x, n, w, dAdt, dU);
        return FaceEMF{0.0, 0.0, 0.0, 0.0};
    }
    DHD void source(const PhysParams& P, const double* prim, double* cons, const double* xp, const double* xm,
                    double dVdt) {
        euler::source(P, prim, cons, xp, xm, dVdt);
    }
    DHD int cons2prim(const PhysParams& P, const double* cons, double* prim, const double* x, double dV,
                      int* okp = nullptr) {
        euler::cons2prim<NQ>(P, cons, prim, x, dV);
        if (okp) *okp = 1;
        return 0;
    }
    static constexpr bool NEWTONIAN = true;  // sink.c's cooling and damping apply
    DHD void prim2cons(const PhysParams& P, const double* prim, double* cons, const double* x, double dV) {
        euler::prim2cons<NQ>(P, prim, cons, x, dV);
    }
    DHD double mindt(const PhysParams& P, const double* prim, double w, const double* xp, const double* xm) {
        return euler::mindt(P, prim, w, xp, xm);
    }
};
using PhysEuler = PhysEulerT<0>;
using PhysEuler1 = PhysEulerT<1>;  // one passive scalar (the reference's templates mostly build NUM_N = 1)

template <int NN>
struct PhysMHDT {
    static constexpr int NC = 8;
    static constexpr int NQ = NC + NN;
    // with constrained transport the cell field is rebuilt from the face
    // fluxes (B_faces_to_cells), so the face increments of cons[BRR..BZZ]
    // (riemann.c:314-317) are never used: only the first five and the
    // passive scalars' are stored (slot s holds field fslot(s, FGAP))
    static constexpr int NF = 5 + NN;
    static constexpr int FGAP = 3;
    static constexpr bool MHD = true;
    static constexpr bool C2P_GUESS = false;
    static constexpr bool SOURCE_PASS = false;
    template <bool DC = false>
    DHD FaceEMF riemann(const PhysParams& P, const double* pL, const double* pR, const double* x, const double* n,
                        double w, double dAdt, int dim, double* dU) {
        return mhd::riemann<DC, NQ>(P, pL, pR, x, n, w, dAdt, dim, dU);
    }
    DHD void source(const PhysParams& P, const double* prim, double* cons, const double* xp, const double* xm,
                    double dVdt) {
        mhd::source(P, prim, cons, xp, xm, dVdt);
    }
    DHD int cons2prim(const PhysParams& P, const double* cons, double* prim, const double* x, double dV,
                      int* okp = nullptr) {
        mhd::cons2prim<NQ>(P, cons, prim, x, dV);
        if (okp) *okp = 1;
        return 0;
    }
    static constexpr bool NEWTONIAN = true;
    DHD void prim2cons(const PhysParams& P, const double* prim, double* cons, const double* x, double dV) {
        mhd::prim2cons<NQ>(P, prim, cons, x, dV);
    }
    DHD double mindt(const PhysParams& P, const double* prim, double w, const double* xp, const double* xm) {
        return mhd::mindt(P, prim, w, xp, xm);
    }
};
using PhysMHD = PhysMHDT<0>;
using PhysMHD1 = PhysMHDT<1>;

// GRMHD (grmhd.c) in the Kerr-Schild metric: the same CT layout as ideal MHD;
// its Noble 2D Newton cons2prim starts from the cell's previous primitives,
// so update_b_cell hands them in through prim[].  L2: grmhd2.c, the same
// module with prim[UPP] = u_phi (grmhd.c: u_phi / R^2) and without the
// source's cell-average corrections.
template <bool L2>
struct PhysGRMHDT {
    static constexpr int NC = 8;
    static constexpr int NQ = 8;
    static constexpr int NF = 5;
    static constexpr int FGAP = 0;
    static constexpr bool MHD = true;
    static constexpr bool C2P_GUESS = true;
    static constexpr bool SOURCE_PASS = true;
    static constexpr bool NEWTONIAN = false;  // sink.c's terms are rejected at set-up
    template <bool DC = false>
    DHD FaceEMF riemann(const PhysParams& P, const double* pL, const double* pR, const double* x, const double* n,
                        double w, double dAdt, int dim, double* dU) {
        return grmhd::riemann<L2, DC>(P, pL, pR, x, n, w, dAdt, dim, dU);
    }
    DHD void source(const PhysParams& P, const double* prim, double* cons, const double* xp, const double* xm,
                    double dVdt) {
        grmhd::source<L2>(P, prim, cons, xp, xm, dVdt);
    }
    DHD int cons2prim(const PhysParams& P, const double* cons, double* prim, const double* x, double dV,
                      int* okp = nullptr) {
        return grmhd::cons2prim<L2>(P, cons, prim, x, dV, okp);
    }
    DHD double mindt(const PhysParams& P, const double* prim, double w, const double* xp, const double* xm) {
        return grmhd::mindt<L2>(P, prim, w, xp, xm);
    }
};
using PhysGRMHD = PhysGRMHDT<false>;
using PhysGRMHD2 = PhysGRMHDT<true>;

// GR hydrodynamics (greuler.c) in the same metric: no field, no constrained
// transport; its cons2prim also starts from the cell's previous primitives.
struct PhysGREuler {
    static constexpr int NC = 5;
    static constexpr int NQ = 5;
    static constexpr int NF = 5;
    static constexpr int FGAP = 0;
    static constexpr bool MHD = false;
    static constexpr bool C2P_GUESS = true;
    static constexpr bool SOURCE_PASS = false;
    static constexpr bool NEWTONIAN = false;  // sink.c's terms are rejected at set-up
    template <bool DC = false>
    DHD FaceEMF riemann(const PhysParams& P, const double* pL, const double* pR, const double* x, const double* n,
                        double w, double dAdt, int dim, double* dU) {
        (void)dim;
        greuler::riemann<DC>(P, pL, pR, x, n, w, dAdt, dU);
        return FaceEMF{0.0, 0.0, 0.0, 0.0};
    }
    DHD void source(const PhysParams& P, const double* prim, double* cons, const double* xp, const double* xm,
                    double dVdt) {
        greuler::source(P, prim, cons, xp, xm, dVdt);
    }
    DHD int cons2prim(const PhysParams& P, const double* cons, double* prim, const double* x, double dV,
                      int* okp = nullptr) {
        return greuler::cons2prim(P, cons, prim, x, dV, okp);
    }
    DHD double mindt(const PhysParams& P, const double* prim, double w, const double* xp, const double* xm) {
        return greuler::mindt(P, prim, w, xp, xm);
    }
};

// ------------------------------------------------------------------ helpers
// plm.c:6-13
DHD double minmod(double a, double b, double c) {
    double m = a;
    if (a * b < 0.0) m = 0.0;
    if (fabs(b) < fabs(m)) m = b;
    if (b * c < 0.0) m = 0.0;
    if (fabs(c) < fabs(m)) m = c;
    return m;
}

// Number of fronts of annulus B (n cells from off, smallest front at `rot`)
// strictly below x (strict = true) or at or below x (strict = false).
DHD int front_count(const double* __restrict__ piph, int off, int n, int rot, double x, bool strict) {
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
    return lo;
}

// The same count, returned as the array index of the cell at that rotated
// position (mod Np): the first cell whose front is not counted.
DHD int front_search(const double* __restrict__ piph, int off, int n, int rot, double x, bool strict) {
    int idx = rot + front_count(piph, off, n, rot, x, strict);
    while (idx >= n) idx -= n;
    return idx;
}

// front_search's answer found by walking from a previous one.  The count of
// fronts below x is monotone in the rotated position, so a walk from any
// start ends at the same boundary (and the same array index) as the binary
// search; between substeps a partner moves by at most a cell or two, so the
// walk from the cell's previous partner takes a few probes instead of
// log2(n) dependent ones.  Walks longer than 8 probes fall back to the
// binary search.
DHD int front_search_near(const double* __restrict__ piph, int off, int n, int rot, double x, bool strict,
                          int hint) {
    if (hint >= 0 && hint < n) {
        int c = hint - rot;
        if (c < 0) c += n;
        for (int k = 0; k < 8; ++k) {
            if (c > 0) {
                int i = rot + c - 1;
                if (i >= n) i -= n;
                double f = piph[off + i];
                if (!(strict ? (f < x) : (f <= x))) {
                    --c;
                    continue;
                }
            }
            if (c < n) {
                int i = rot + c;
                if (i >= n) i -= n;
                double f = piph[off + i];
                if (strict ? (f < x) : (f <= x)) {
                    ++c;
                    continue;
                }
            }
            int idx = rot + c;
            if (idx >= n) idx -= n;
            return idx;
        }
    }
    return front_search(piph, off, n, rot, x, strict);
}

DHD int wrap_idx(int i, int n) {
    if (i >= n) i -= n;
    if (i < 0) i += n;
    return i;
}

// Per-direction accessors: the neighbouring annuli, partner arrays, gradient,
// owned-face increments, flags, centroid distances, and the constrained
// transport slots of direction D.  "hi" is the track (this | outer or up) in
// which this cell is L, "lo" the track (inner or down | this) in which it is R.
template <int D>
struct Dir;
template <>
struct Dir<DIR_R> {
    DHD int hi(const AnnDev& A) { return A.outer; }
    DHD int lo(const AnnDev& A) { return A.inner; }
    DHD int* part_hi(const DevState& S) { return S.lbo; }
    DHD int* part_lo(const DevState& S) { return S.ubi; }
    DHD double* grad(const DevState& S) { return S.gradr; }
    DHD double* f_hi(const DevState& S) { return S.fout; }
    DHD double* f_lo(const DevState& S) { return S.fin; }
    static constexpr int FLAG_HI = ANN_OUTFLUX, FLAG_LO = ANN_INFLUX;
    DHD double dxL(const AnnDev& A) { return A.dxL; }
    DHD double dxR(const AnnDev& A) { return A.dxR; }
    // constrained transport
    static constexpr int PHI_HI = 2, PHI_LO = 1;     // Phi slots of the owned faces
    static constexpr int EMF_HI = 0, EMF_LO = 3;     // emf_t base slots
    static constexpr int SUM_HI = 1, SUM_LO = 0;     // esum slots (E; +4 for B)
    static constexpr int E_HI = 1, E_LO = 0;         // edge slots (E1, E0) of edge_avg1 / efin
    static constexpr int FL_HI = 1, FL_LO = 2;       // fflags bits
    static constexpr double SGN = 1.0;               // update_B_fluxes orientation (bfields.c:321-346)
    DHD double dl(const AnnDev& A) { return A.zp - A.zm; }  // face dl = get_dL(.., 2) (faces_alt.c:26-27)
};
template <>
struct Dir<DIR_Z> {
    DHD int hi(const AnnDev& A) { return A.up; }
    DHD int lo(const AnnDev& A) { return A.down; }
    DHD int* part_hi(const DevState& S) { return S.lbu; }
    DHD int* part_lo(const DevState& S) { return S.ubd; }
    DHD double* grad(const DevState& S) { return S.gradz; }
    DHD double* f_hi(const DevState& S) { return S.fup; }
    DHD double* f_lo(const DevState& S) { return S.fdn; }
    static constexpr int FLAG_HI = ANN_UPFLUX, FLAG_LO = ANN_DNFLUX;
    DHD double dxL(const AnnDev& A) { return A.dzL; }
    DHD double dxR(const AnnDev& A) { return A.dzR; }
    static constexpr int PHI_HI = 4, PHI_LO = 3;
    static constexpr int EMF_HI = 6, EMF_LO = 9;
    static constexpr int SUM_HI = 3, SUM_LO = 2;
    static constexpr int E_HI = 3, E_LO = 2;
    static constexpr int FL_HI = 4, FL_LO = 8;
    static constexpr double SGN = -1.0;                     // bfields.c:352-375
    DHD double dl(const AnnDev& A) { return A.rp - A.rm; }  // get_dL(.., 1)
};

// Geometry of the face between L (annulus AL) and R (the annulus above AL in
// direction D): addFace and the merge step of faces_alt.c:10-34, :146-168,
// plus the phi offsets of plm.c:86-91.  x is the face centroid (face cm).
struct TFace {
    double phic, dA, dpL, dpR, dp;
    double x[3];
};

template <int D>
DHD TFace t_face(double piphL, double dphiL, double piphR, double dphiR, int LR, const AnnDev& AL, double Pmax) {
    TFace F;
    double xp1 = LR == 0 ? piphL : piphR;
    double dd = piphL - dphiL - piphR + dphiR;
    while (dd > .5 * Pmax) dd -= Pmax;
    while (dd < -.5 * Pmax) dd += Pmax;
    double xm1 = dd < 0. ? piphR - dphiR : piphL - dphiL;
    double dp = get_dp(xp1, xm1, Pmax);
    F.dp = dp;
    F.phic = get_dp(xp1, .5 * dp, Pmax);
    if (D == DIR_R) {
        F.dA = AL.rp * dp * (AL.zp - AL.zm);  // get_dA(.., 1) = r dphi dz, r = r_jph[j]
        F.x[0] = AL.rp;
        F.x[2] = AL.zc;
    } else {
        F.dA = .5 * (AL.rp + AL.rm) * (AL.rp - AL.rm) * dp;  // get_dA(.., 2) = r dr dphi
        F.x[0] = AL.rc;
        F.x[2] = AL.zp;
    }
    F.x[1] = F.phic;
    double phiL = piphL - .5 * dphiL;
    double phiR = piphR - .5 * dphiR;
    F.dpL = get_signed_dp(F.phic, phiL, Pmax);
    F.dpR = get_signed_dp(phiR, F.phic, Pmax);
    return F;
}

// ------------------------------------------------------------------ viscosity
// omega.c:132-155 (get_height_om): the frequency sqrt(sum_p f_p / (d_p +
// eps_p / 16)) of the run's point-mass planets at x (r, phi) in the plane
DHD double height_om(const DevState& S, const RunDev& Rn, int run, const double* x) {
    double omtot2 = 0.0;
    const double r = x[0], phi = x[1];
    const double gx = r * cos(phi), gy = r * sin(phi);
    for (int p = 0; p < Rn.ph.npl; ++p) {
        const Planet& pl = Rn.planets[p];
        const double pphi = S.planet_phi[run * DISCO_MAX_PLANETS + p];
        const double px = pl.r * cos(pphi), py = pl.r * sin(pphi);
        const double sr = sqrt((px - gx) * (px - gx) + (py - gy) * (py - gy));
        const double f1 = fgrav_pointmass(pl.M, sr, pl.eps);
        omtot2 += f1 / (sr + 0.0625 * pl.eps);
    }
    return sqrt(omtot2);
}

// euler.c:220-225: the kinematic viscosity nu, or alpha c_s^2 / Omega_height
// with Use_As_Alpha, for the primitives prim at x
DHD double visc_nu(const DevState& S, const RunDev& Rn, int run, const double* prim, const double* x) {
    const PhysParams& P = Rn.ph;
    if (!P.visc_alpha) return P.visc_nu;
    const double c2 = P.gamma_law * prim[PPP] / prim[RHO];
    return P.visc_nu * c2 / height_om(S, Rn, run, x);
}

// ------------------------------------------------------------------ dt
// euler.c:351-373: a viscous run's time step is at most 0.5 dx^2 / nu
template <class Ph>
DHD double with_visc_dt(const DevState& S, const AnnDev& A, double dt, const double* prim, const double* xp,
                        const double* xm) {
    if constexpr (!Ph::MHD) {
        const RunDev& Rn = S.runs[A.run];
        if (Rn.ph.visc) {
            const double x[3] = {get_centroid(xp[0], xm[0], 1), get_centroid(xp[1], xm[1], 0),
                                 get_centroid(xp[2], xm[2], 2)};
            const double dtv = euler::visc_dt(Rn.ph, visc_nu(S, Rn, A.run, prim, x), xp, xm);
            if (dt > dtv) dt = dtv;
        }
    }
    return dt;
}

// misc.c:39-84 for one real cell: mindt(prim, w = mean of the two phi-face
// mesh speeds, xp, xm)
template <class Ph>
DHD double cell_mindt(const DevState& S, const AnnDev& A, const PhysParams& P, int c, const double* prim) {
    int i = c - A.off;
    int cm = A.off + (i == 0 ? A.np - 1 : i - 1);
    double phip = S.piph[c];
    double phim = phip - S.dphi[c];
    double xp[3] = {A.rp, phip, A.zp};
    double xm[3] = {A.rm, phim, A.zm};
    double w = .5 * (S.wiph[cm] + S.wiph[c]);
    return with_visc_dt<Ph>(S, A, Ph::mindt(P, prim, w, xp, xm), prim, xp, xm);
}

// CFL candidate of cell c over the current state (1e300 for ghosts)
template <class Ph>
DHD double mindt_cell(const DevState& S, int c) {
    constexpr int NQ = Ph::NQ;
    const AnnDev& A = S.ann[S.cell_ann[c]];
    if (!(A.flags & ANN_REAL)) return 1.0e300;
    double prim[NQ];
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) prim[q] = S.prim[(size_t)q * S.ncell + c];
    return cell_mindt<Ph>(S, A, S.runs[A.run].ph, c, prim);
}

// ------------------------------------------------------------------ partners
// The partner of cell c in each neighbouring annulus (and, with CT, in the
// diagonal ones): the first cell whose front does not precede c's front, as
// the overlap faces of faces_alt.c:36-179 pair them, walked from the last
// substep's partner.  The mesh only moves in the first substep, so the
// partners found in a step's second substep are those of the next step's
// first substep: the search runs in the second substep, and once after the
// mesh is set (creation, upload).
DHD void partners_cell(const DevState& S, int c) {
    const AnnDev& A = S.ann[S.cell_ann[c]];
    const size_t N = S.ncell;
    double front = S.piph[c];
    if (A.outer >= 0) {
        const AnnDev& B = S.ann[A.outer];
        S.lbo[c] = front_search_near(S.piph, B.off, B.np, S.rot[A.outer], front, true, S.lbo[c]);
    }
    if (A.inner >= 0) {
        const AnnDev& W = S.ann[A.inner];
        S.ubi[c] = front_search_near(S.piph, W.off, W.np, S.rot[A.inner], front, false, S.ubi[c]);
    }
    if (A.up >= 0) {
        const AnnDev& B = S.ann[A.up];
        S.lbu[c] = front_search_near(S.piph, B.off, B.np, S.rot[A.up], front, true, S.lbu[c]);
    }
    if (A.down >= 0) {
        const AnnDev& W = S.ann[A.down];
        S.ubd[c] = front_search_near(S.piph, W.off, W.np, S.rot[A.down], front, false, S.ubd[c]);
    }
    if (S.pdg) {
        // diagonal annuli of the azimuthal-edge merges; a front of an annulus
        // higher in (k, j) order counts after an equal front of this one
        _Pragma("unroll") for (int d = 0; d < 4; ++d) {
            int a = A.diag[d];
            if (a < 0) continue;
            const AnnDev& Q = S.ann[a];
            const size_t k = (size_t)d * N + c;
            S.pdg[k] = front_search_near(S.piph, Q.off, Q.np, S.rot[a], front, d < 2, S.pdg[k]);
        }
    }
}

// ------------------------------------------------------------------ stage A
// set_wcell (misc.c:101-168, first substep), plm_phi (plm.c:15-51) and, when
// `search`, the partner search of the overlap faces.  adjust_RK_cons
// (misc.c:226-244) and timestep.c:18-24's copy run where the state is next
// written: in update_a_cell for the conserved state, in ct_update_cell for
// the face fluxes (nothing reads either in between).
template <class Ph>
DHD void stage_a_cell(const DevState& S, int c, int first, int search) {
    constexpr int NQ = Ph::NQ;
    const AnnDev& A = S.ann[S.cell_ann[c]];
    const PhysParams& P = S.runs[A.run].ph;
    const size_t N = S.ncell;
    int n = A.np;
    int i = c - A.off;
    int cm = A.off + (i == 0 ? n - 1 : i - 1);
    int cp = A.off + (i + 1 == n ? 0 : i + 1);

    if (first) {
        double w = 0.0;
        if (P.mesh_motion == 1 || P.mesh_motion == 2 || P.mesh_motion == 3) {
            // misc.c:117-132 (every non-zero mode; 3 then takes the annulus
            // mean in k_wmean)
            w = .5 * (S.prim[UPP * N + c] + S.prim[UPP * N + cp]);
        } else if (P.mesh_motion == 4) {
            double phip = S.piph[c];
            double phim = phip - S.dphi[c];
            double x[3] = {A.rc, 0.5 * (phim + phip), A.zc};
            w = mesh_om(P, x);
        }
        S.wiph[c] = w;
    }
    double dpL = S.dphi[cm];
    double dpC = S.dphi[c];
    double dpR = S.dphi[cp];
    // the three centre distances are shared by every field: one division
    // each, then the exactly rounded quotients of div_rn
    const double dL = .5 * (dpC + dpL);
    const double dR = .5 * (dpR + dpC);
    const double dC = .5 * (dpL + dpR) + dpC;
    const double yL = 1.0 / dL, yR = 1.0 / dR, yC = 1.0 / dC;
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
        double pL = S.prim[q * N + cm];
        double pC = S.prim[q * N + c];
        double pR = S.prim[q * N + cp];
        double sL = div_rn(pC - pL, dL, yL);
        double sR = div_rn(pR - pC, dR, yR);
        double sC = div_rn(pR - pL, dC, yC);
        st_stream(&S.gradp[q * N + c], minmod(P.plm * sL, sC, P.plm * sR));
    }
    if (search) partners_cell(S, c);
}

// ------------------------------------------------------------------ plm_trans
// plm.c:56-195: area-weighted mean of the face slopes, divided by the cell's
// total face area in the direction, then the per-face limiter (order
// independent: zero on any sign change, else the smallest |PLM*S|).
template <int NQ>
struct SlopeAcc {
    double sum[NQ];
    double smin[NQ];
    unsigned pos, neg, zero;  // bit q set when some face slope of field q is >0, <0, ==0
    bool any;
};

template <int NQ>
DHD void acc_add(SlopeAcc<NQ>& a, const double* S, double dA) {
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
        a.sum[q] += S[q] * dA;
        a.pos |= (S[q] > 0.0 ? 1u : 0u) << q;
        a.neg |= (S[q] < 0.0 ? 1u : 0u) << q;
        a.zero |= (S[q] == 0.0 ? 1u : 0u) << q;
        a.smin[q] = (!a.any || fabs(S[q]) < fabs(a.smin[q])) ? S[q] : a.smin[q];
    }
    a.any = true;
}

template <int NQ>
DHD void face_slope(const DevState& S, int cL, int cR, const TFace& F, double dxL, double dxR, double* Sq) {
    const size_t N = S.ncell;
    const double dx = dxR + dxL;
    const double y = 1.0 / dx;
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
        double WL = S.prim[q * N + cL] + F.dpL * S.gradp[q * N + cL];
        double WR = S.prim[q * N + cR] - F.dpR * S.gradp[q * N + cR];
        Sq[q] = div_rn(WR - WL, dx, y);
    }
}

// Gradient of cell c in direction D (plm_trans with dim = D).
template <int NQ, int D>
DHD void trans_gradient(const DevState& S, const AnnDev& A, const PhysParams& P, int c, int cm) {
    using Dr = Dir<D>;
    const size_t N = S.ncell;
    const double Pmax = P.phi_max;
    double* grad = Dr::grad(S);
    if (P.plm == 0.0) {
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) grad[q * N + c] = 0.0;
        return;
    }
    double phip = S.piph[c];
    double dph = S.dphi[c];
    SlopeAcc<NQ> acc;
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
        acc.sum[q] = 0.0;
        acc.smin[q] = 0.0;
    }
    acc.pos = acc.neg = acc.zero = 0u;
    acc.any = false;
    double Sq[NQ];
    const int ahi = Dr::hi(A), alo = Dr::lo(A);
    if (ahi >= 0) {  // this cell is L of the (this | hi) track
        const AnnDev& B = S.ann[ahi];
        const int* part = Dr::part_hi(S);
        int nb = B.np;
        int own = part[c];
        int start = part[cm];
        int cnt = own - start;
        if (cnt < 0) cnt += nb;
        int cb = B.off + own;
        TFace F = t_face<D>(phip, dph, S.piph[cb], S.dphi[cb], 0, A, Pmax);
        face_slope<NQ>(S, c, cb, F, Dr::dxL(A), Dr::dxR(A), Sq);
        acc_add(acc, Sq, F.dA);
        for (int t = 0; t < cnt; ++t) {
            int b = start + t;
            if (b >= nb) b -= nb;
            cb = B.off + b;
            F = t_face<D>(phip, dph, S.piph[cb], S.dphi[cb], 1, A, Pmax);
            face_slope<NQ>(S, c, cb, F, Dr::dxL(A), Dr::dxR(A), Sq);
            acc_add(acc, Sq, F.dA);
        }
    }
    if (alo >= 0) {  // this cell is R of the (lo | this) track
        const AnnDev& W = S.ann[alo];
        const int* part = Dr::part_lo(S);
        int nw = W.np;
        int own = part[c];
        int start = part[cm];
        int cnt = own - start;
        if (cnt < 0) cnt += nw;
        int ca = W.off + own;
        TFace F = t_face<D>(S.piph[ca], S.dphi[ca], phip, dph, 1, W, Pmax);
        face_slope<NQ>(S, ca, c, F, Dr::dxL(W), Dr::dxR(W), Sq);
        acc_add(acc, Sq, F.dA);
        for (int t = 0; t < cnt; ++t) {
            int a = start + t;
            if (a >= nw) a -= nw;
            ca = W.off + a;
            F = t_face<D>(S.piph[ca], S.dphi[ca], phip, dph, 0, W, Pmax);
            face_slope<NQ>(S, ca, c, F, Dr::dxL(W), Dr::dxR(W), Sq);
            acc_add(acc, Sq, F.dA);
        }
    }
    // plm.c:112-143: the cell's total face area in this direction
    double phim = phip - dph;
    double dpp = get_dp(phip, phim, Pmax);
    double dAp, dAm;
    if (D == DIR_R) {
        double dz = A.zp - A.zm;
        dAp = A.rp * dpp * dz;
        dAm = A.rm * dpp * dz;
    } else {
        dAp = .5 * (A.rp + A.rm) * (A.rp - A.rm) * dpp;
        dAm = dAp;
    }
    double dAtot = dAp + dAm;
    if (alo < 0) dAtot = dAp;
    if (ahi < 0) dAtot = dAm;
    const bool geom = D == DIR_R && (A.flags & ANN_GEOM_R);
    const double yA = 1.0 / dAtot;
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
        double g = div_rn(acc.sum[q], dAtot, yA);
        if (!acc.any) {
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
        if (geom) {
            // plm.c:184-189 + geom_grad (cylindrical.c:141-173) at r = 0: the
            // radial components limited by prim / r, every other field flat
            if (q == URR || (NQ > BZZ && q == BRR)) {
                const double SL = S.prim[q * N + c] / A.rc;
                if (g * SL < 0.0)
                    g = 0.0;
                else if (fabs(P.plm * SL) < fabs(g))
                    g = P.plm * SL;
            } else {
                g = 0.0;
            }
        }
        st_stream(&grad[q * N + c], g);
    }
}

template <class Ph>
DHD void plm_trans_cell(const DevState& S, int c, int has_z) {
    constexpr int NQ = Ph::NQ;
    const AnnDev& A = S.ann[S.cell_ann[c]];
    const PhysParams& P = S.runs[A.run].ph;
    int i = c - A.off;
    int cm = A.off + (i == 0 ? A.np - 1 : i - 1);
    trans_gradient<NQ, DIR_R>(S, A, P, c, cm);
    if (has_z) trans_gradient<NQ, DIR_Z>(S, A, P, c, cm);
}

// ------------------------------------------------------------------ fluxes
// riemann_phi (riemann.c:39-112) on the cell's phi face, riemann_trans
// (riemann.c:114-213) on its owned faces; solve_riemann's increment
// (F - w U*) dA dt is stored per owner cell, and for MHD the face's EMFs
// (flux_to_E) beside it.
template <class Ph, int D, bool VISC = false, bool DC = false>
DHD FaceEMF t_face_flux(const DevState& S, const PhysParams& P, int cL, int cR, const TFace& F, const AnnDev& AL,
                        double dt, double* dU) {
    constexpr int NQ = Ph::NQ;
    using Dr = Dir<D>;
    const size_t N = S.ncell;
    const double* grad = Dr::grad(S);
    const double dxL = Dr::dxL(AL), dxR = Dr::dxR(AL);
    double primL[NQ], primR[NQ];
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
        primL[q] = S.prim[q * N + cL] + grad[q * N + cL] * dxL + S.gradp[q * N + cL] * F.dpL;
        primR[q] = S.prim[q * N + cR] - grad[q * N + cR] * dxR - S.gradp[q * N + cR] * F.dpR;
    }
    if (Ph::MHD) {  // riemann.c:146-152: the normal field is single-valued on the face
        const int BT = D == DIR_R ? BRR : BZZ;
        double Bavg = .5 * (primL[BT] + primR[BT]);
        primL[BT] = Bavg;
        primR[BT] = Bavg;
    }
    const double n[3] = {D == DIR_R ? 1.0 : 0.0, 0.0, D == DIR_Z ? 1.0 : 0.0};
    if constexpr (VISC && !Ph::MHD) {
        // riemann.c:114-140, 297-312: the face's mean state; across the face
        // the difference of the two phi-reconstructed values, along it the
        // mean of the cells' gradients
        const AnnDev& Ann = S.ann[S.cell_ann[cL]];
        const RunDev& Rn = S.runs[Ann.run];
        double pm[NQ], g[3][NQ], vF[NQ];
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
            pm[q] = .5 * (primL[q] + primR[q]);
            const double gf = ((S.prim[q * N + cR] - S.gradp[q * N + cR] * F.dpR) -
                               (S.prim[q * N + cL] + S.gradp[q * N + cL] * F.dpL)) /
                              (dxL + dxR);
            g[0][q] = D == DIR_R ? gf : .5 * (S.gradr[q * N + cL] + S.gradr[q * N + cR]);
            g[1][q] = .5 * (S.gradp[q * N + cL] + S.gradp[q * N + cR]);
            g[2][q] = D == DIR_Z ? gf : .5 * (S.gradz[q * N + cL] + S.gradz[q * N + cR]);
        }
        const double nu = visc_nu(S, Rn, Ann.run, pm, F.x);
        euler::visc_flux<NQ>(P, nu, pm, g[0], g[1], g[2], vF, F.x, n);
        euler::riemann<NQ, DC>(P, primL, primR, F.x, n, 0.0, F.dA * dt, dU, vF);
        return FaceEMF{0.0, 0.0, 0.0, 0.0};
    } else {
        return Ph::template riemann<DC>(P, primL, primR, F.x, n, 0.0, F.dA * dt, D, dU);
    }
}

template <class Ph, int D, bool VISC = false, bool DC = false>
DHD void owned_faces_flux(const DevState& S, const AnnDev& A, const PhysParams& P, int c, double dt) {
    constexpr int NQ = Ph::NQ;
    using Dr = Dir<D>;
    const size_t N = S.ncell;
    const double Pmax = P.phi_max;
    double dU[NQ];
    double phip = S.piph[c];
    double dph = S.dphi[c];
    if (A.flags & Dr::FLAG_HI) {
        const AnnDev& B = S.ann[Dr::hi(A)];
        int cb = B.off + Dr::part_hi(S)[c];
        TFace F = t_face<D>(phip, dph, S.piph[cb], S.dphi[cb], 0, A, Pmax);
        FaceEMF e = t_face_flux<Ph, D, VISC, DC>(S, P, c, cb, F, A, dt, dU);
        double* f = Dr::f_hi(S);
        _Pragma("unroll") for (int q = 0; q < Ph::NF; ++q) st_stream(&f[q * N + c], dU[fslot(q, Ph::FGAP)]);
        if (Ph::MHD && S.emf_t) {
            st_stream(&S.emf_t[(Dr::EMF_HI + 0) * N + c], e.E1);
            st_stream(&S.emf_t[(Dr::EMF_HI + 1) * N + c], e.B1);
            st_stream(&S.emf_t[(Dr::EMF_HI + 2) * N + c], e.E2);
        }
    }
    if (A.flags & Dr::FLAG_LO) {
        const AnnDev& W = S.ann[Dr::lo(A)];
        int ca = W.off + Dr::part_lo(S)[c];
        TFace F = t_face<D>(S.piph[ca], S.dphi[ca], phip, dph, 1, W, Pmax);
        FaceEMF e = t_face_flux<Ph, D, VISC, DC>(S, P, ca, c, F, W, dt, dU);
        double* f = Dr::f_lo(S);
        _Pragma("unroll") for (int q = 0; q < Ph::NF; ++q) st_stream(&f[q * N + c], dU[fslot(q, Ph::FGAP)]);
        if (Ph::MHD && S.emf_t) {
            st_stream(&S.emf_t[(Dr::EMF_LO + 0) * N + c], e.E1);
            st_stream(&S.emf_t[(Dr::EMF_LO + 1) * N + c], e.B1);
            st_stream(&S.emf_t[(Dr::EMF_LO + 2) * N + c], e.E2);
        }
    }
}

template <class Ph, bool VISC = false, bool DC = false>
DHD void flux_cell(const DevState& S, int c, double dtfac) {
    constexpr int NQ = Ph::NQ;
    const AnnDev& A = S.ann[S.cell_ann[c]];
    if (!(A.flags & (ANN_PHIFLUX | ANN_OUTFLUX | ANN_INFLUX | ANN_UPFLUX | ANN_DNFLUX))) return;
    const PhysParams& P = S.runs[A.run].ph;
    const size_t N = S.ncell;
    const double dt = S.dt[A.run] * dtfac;
    int n = A.np;
    int i = c - A.off;

    if (A.flags & ANN_PHIFLUX) {
        int cp = A.off + (i + 1 == n ? 0 : i + 1);
        double dU[NQ];
        double primL[NQ], primR[NQ];
        double dph = S.dphi[c];
        double dphR = S.dphi[cp];
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
            primL[q] = S.prim[q * N + c] + .5 * S.gradp[q * N + c] * dph;
            primR[q] = S.prim[q * N + cp] - .5 * S.gradp[q * N + cp] * dphR;
        }
        if (Ph::MHD) {  // riemann.c:55-59
            double Bp = .5 * (primL[BPP] + primR[BPP]);
            primL[BPP] = Bp;
            primR[BPP] = Bp;
        }
        double x[3] = {A.rc, S.piph[c], A.zc};
        const double nv[3] = {0.0, 1.0, 0.0};
        double dA = (A.rp - A.rm) * (A.zp - A.zm);  // get_dA(xp, xm, 0)
        double w = A.rc * S.wiph[c];                // get_scale_factor(x, 0) * wiph
        FaceEMF e;
        if constexpr (VISC && !Ph::MHD) {
            // riemann.c:39-70, 297-312: across the face the difference of the
            // two cells' prims, along it the mean of their gradients
            double pm[NQ], g[3][NQ], vF[NQ];
            _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
                pm[q] = .5 * (primL[q] + primR[q]);
                g[0][q] = .5 * (S.gradr[q * N + c] + S.gradr[q * N + cp]);
                g[1][q] = (S.prim[q * N + cp] - S.prim[q * N + c]) / (0.5 * (dphR + dph));
                g[2][q] = .5 * (S.gradz[q * N + c] + S.gradz[q * N + cp]);
            }
            const double nu = visc_nu(S, S.runs[A.run], A.run, pm, x);
            euler::visc_flux<NQ>(P, nu, pm, g[0], g[1], g[2], vF, x, nv);
            euler::riemann<NQ, DC>(P, primL, primR, x, nv, w, dA * dt, dU, vF);
            e = FaceEMF{0.0, 0.0, 0.0, 0.0};
        } else {
            e = Ph::template riemann<DC>(P, primL, primR, x, nv, w, dA * dt, 0, dU);
        }
        _Pragma("unroll") for (int q = 0; q < Ph::NF; ++q) st_stream(&S.fphi[q * N + c], dU[fslot(q, Ph::FGAP)]);
        if (Ph::MHD && S.emf_phi) {
            st_stream(&S.emf_phi[0 * N + c], e.E1);  // E_z
            st_stream(&S.emf_phi[1 * N + c], e.B1);  // B_r
            st_stream(&S.emf_phi[2 * N + c], e.E2);  // E_r
            st_stream(&S.emf_phi[3 * N + c], e.B2);  // B_z
        }
    }
    owned_faces_flux<Ph, DIR_R, VISC, DC>(S, A, P, c, dt);
    owned_faces_flux<Ph, DIR_Z, VISC, DC>(S, A, P, c, dt);
}

// ------------------------------------------------------------------ update
// Gather of the increments of every face of cell c in direction D: the
// reference subtracts a face's increment from its L cell and adds it to its
// R cell (riemann.c:314-317).
template <int NQ, int D, int GAP = 0>
DHD void gather_trans(const DevState& S, const AnnDev& A, int c, int cm, double* cons) {
    using Dr = Dir<D>;
    const size_t N = S.ncell;
    const int alo = Dr::lo(A), ahi = Dr::hi(A);
    if (alo >= 0 && (A.flags & Dr::FLAG_LO)) {  // this cell is R: gains
        const AnnDev& W = S.ann[alo];
        const double* fo = Dr::f_lo(S);
        const double* fw = Dr::f_hi(S);
        const int* part = Dr::part_lo(S);
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) cons[fslot(q, GAP)] += fo[q * N + c];
        int start = part[cm];
        int cnt = part[c] - start;
        if (cnt < 0) cnt += W.np;
        for (int t = 0; t < cnt; ++t) {
            int a = start + t;
            if (a >= W.np) a -= W.np;
            _Pragma("unroll") for (int q = 0; q < NQ; ++q) cons[fslot(q, GAP)] += fw[q * N + W.off + a];
        }
    }
    if (ahi >= 0 && (A.flags & Dr::FLAG_HI)) {  // this cell is L: loses
        const AnnDev& B = S.ann[ahi];
        const double* fo = Dr::f_hi(S);
        const double* fb = Dr::f_lo(S);
        const int* part = Dr::part_hi(S);
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) cons[fslot(q, GAP)] -= fo[q * N + c];
        int start = part[cm];
        int cnt = part[c] - start;
        if (cnt < 0) cnt += B.np;
        for (int t = 0; t < cnt; ++t) {
            int b = start + t;
            if (b >= B.np) b -= B.np;
            _Pragma("unroll") for (int q = 0; q < NQ; ++q) cons[fslot(q, GAP)] -= fb[q * N + B.off + b];
        }
    }
}

// omega_tot of sink.c's cooling and damping: the square root of the sum over
// the planets of |force| / distance in the plane (planetaryForce mode 1).
DHD double omega_tot(const DevState& S, const RunDev& Rn, int run, const double* x) {
    const PhysParams& P = Rn.ph;
    const double r = x[0], phi = x[1], z = x[2];
    const double gx = r * cos(phi), gy = r * sin(phi);
    double omtot = 0.0;
    for (int p = 0; p < P.npl; ++p) {
        Planet pl = Rn.planets[p];
        pl.phi = S.planet_phi[run * DISCO_MAX_PLANETS + p];
        const double dx = gx - pl.r * cos(pl.phi), dy = gy - pl.r * sin(pl.phi);
        const double mag = sqrt(dx * dx + dy * dy);
        double fr, fp, fz;
        planetary_force(pl, P.grav2D, r, phi, z, &fr, &fp, &fz, 1);
        omtot += sqrt(fr * fr + fp * fp + fz * fz) / mag;
    }
    return sqrt(omtot);
}

// sink.c:274-322 (cooling): beta cooling of the specific internal energy
// toward get_cs2 / (gamma (gamma - 1)) on the time beta / omega_tot, with
// omega_tot^2 the sum over the planets of |force| / distance in the plane,
// integrated over the substep (expm1); Cool_Type 1 cools only above the
// target, 2 relaxes toward it either way.
DHD void cooling_cell(const DevState& S, const RunDev& Rn, int run, const double* prim, double* cons,
                      const double* x, double dV, double dt) {
    const PhysParams& P = Rn.ph;
    const double press = prim[PPP], rho = prim[RHO];
    const double gm1 = P.gamma_law - 1.0;
    const double enTarget = get_cs2(P, x) / (P.gamma_law * gm1);
    const double enCurrent = press / (rho * gm1);
    if (P.cool_type == 2 || enCurrent > enTarget) {
        const double omtot = omega_tot(S, Rn, run, x);
        const double Tm1 = expm1(-dt * omtot / P.cool_beta);
        cons[TAU] += rho * dV * (enCurrent - enTarget) * Tm1;
    }
}

// sink.c:329-426 (damping): the conserved state relaxes toward the initial
// condition's in the zones within Damp*_Len of the inner and outer radius
// and of the upper and lower boundary (the lower zone measured from Z_Max,
// as sink.c has it), with the ramp of the last enabled zone, over the
// substep.  sink.c also turns each zone's Damp*_Time (type 1: code units,
// type 2: units of 1/omega_tot) into a rate that it never applies: the
// factor is expm1(-dt x ramp) whatever the time and the type, and so it is
// here.
template <class Ph>
DHD void damping_cell(const DevState& S, const RunDev& Rn, int run, const double* prim, double* cons,
                      const double* x, double dV, double dt) {
    constexpr int NQ = Ph::NQ;
    (void)S;
    (void)run;
    const PhysParams& P = Rn.ph;
    double dampFactor = 0.0;
    for (int z = 0; z < 4; ++z) {
        if (P.damp_type[z] <= 0 || (z >= 2 && P.two_d)) continue;
        const double theta = (z == 0 ? x[0] - P.rmin : z == 1 ? P.rmax - x[0] : P.zmax - x[2]) / P.damp_len[z];
        if (theta > 1.0)
            dampFactor = 0.0;
        else if (z == 0 || z == 3)
            dampFactor = 1.0 - pow(1.0 - pow(theta, 2.0), 2.0);
        else
            dampFactor = pow(1.0 - pow(theta, 2.0), 2.0);
    }
    dampFactor = expm1(-dt * dampFactor);
    double p0[NQ > 8 ? NQ : 8], c0[NQ], c1[NQ];
    disco_initial(&Rn.ic, p0, x);
    Ph::prim2cons(P, p0, c0, x, dV);
    Ph::prim2cons(P, prim, c1, x, dV);
    cons[DDD] += (c1[DDD] - c0[DDD]) * dampFactor;
    cons[SRR] += (c1[SRR] - c0[SRR]) * dampFactor;
    cons[LLL] += (c1[LLL] - c0[LLL]) * dampFactor;
    cons[SZZ] += (c1[SZZ] - c0[SZZ]) * dampFactor;
    cons[TAU] += (c1[TAU] - c0[TAU]) * dampFactor;
}

// sink.c:107-139 (nozzle): mass injected at the rate 3 / (4 pi dR^3) per unit
// volume (times the chord 2 sqrt(dR^2 - d^2) in 2D) inside the ball of radius
// dR about (r, phi, z) = (R0, 0, 0), d measured with the midpoint
// coordinates and phi folded into (-pi, pi], moving at azimuthal speed v with
// the thermal energy of sound speed v / mach, its kinetic energy counted
// relative to the frame speed r get_om.  The clip compares the injected mass
// with the cell's mass and then sets the rate to cons[DDD] / dV * dt, as the
// reference writes it.
// Returns false for a cell outside the ball: sink_src returns there before
// its accretion sinks.
template <int NC, int NQ>
DHD bool nozzle_cell(const PhysParams& P, double* cons, const double* xp, const double* xm, double dV, double dt) {
    const double R0 = P.nozzle_par[0], dR = P.nozzle_par[1], v = P.nozzle_par[2], mach = P.nozzle_par[3];
    const double r = 0.5 * (xp[0] + xm[0]);
    double phi = 0.5 * (xp[1] + xm[1]);
    const double z = 0.5 * (xp[2] + xm[2]);
    if (phi > M_PI) phi -= 2 * M_PI;
    if ((r - R0) * (r - R0) + r * r * phi * phi + z * z > dR * dR) return false;
    double rhodot = 1.0 * 3.0 / (4.0 * M_PI * dR * dR * dR);
    if (P.two_d) rhodot *= 2 * sqrt(dR * dR - (r - R0) * (r - R0) - r * r * phi * phi);
    if (rhodot * dV * dt > cons[DDD]) rhodot = cons[DDD] / dV * dt;
    const double x[3] = {r, phi, z};
    const double om = get_om(P, x);
    const double cs2 = v * v / (mach * mach);
    const double eps = cs2 / (P.gamma_law * (P.gamma_law - 1));
    cons[DDD] += rhodot * dV * dt;
    cons[LLL] += rhodot * r * v * dV * dt;
    cons[TAU] += (0.5 * rhodot * (v - r * om) * (v - r * om) + rhodot * eps) * dV * dt;
    if constexpr (NQ > NC) cons[NC] += rhodot * dV * dt;  // sink.c:137-138: the first passive scalar tags it
    return true;
}

// sink.c:140-270 (the accretion sinks) for one real cell: each accreting
// planet removes mass at rate Sink_Par1 x its omega x the kernel of its
// distance (1: (1 - (d/eps)^M)^N inside eps, 2: exp(-d^P / (eps^P f)), 3: 1
// inside Sink_Par3), with the momentum of the gas velocity relative to the
// planet's (Sink_Par2 = delta of it removed in the planet's direction, the
// rest in the tangential one), its kinetic and thermal energy.  The planet's
// accumulators gain the removed quantities and the torque accounting
// -(rho - rho0) r_p f_phi dV dt; this cell's contributions go to
// S.sinkc[(p * DISCO_NSINKQ + k) * ncell + c] (zero for planets that do not
// accrete, and for every planet when `active` is false: the cell lies outside
// the nozzle's ball).  The planets' radial speed is zero for the built
// planet modules.
DHD void sink_cell(const DevState& S, const RunDev& Rn, int run, int c, bool active, const double* prim, double* cons,
                   const double* xp, const double* xm, double dV, double dt) {
    const PhysParams& P = Rn.ph;
    const size_t N = S.ncell;
    const double x[3] = {get_centroid(xp[0], xm[0], 1), get_centroid(xp[1], xm[1], 0), get_centroid(xp[2], xm[2], 2)};
    const double r = x[0], phi = x[1];
    const double rho = prim[RHO], vr = prim[URR], vp = prim[UPP] * r, vz = prim[UZZ];
    const double cosg = cos(phi), sing = sin(phi);
    const double gx = r * cosg, gy = r * sing;
    for (int p = 0; p < DISCO_MAX_PLANETS; ++p) {
        double a[DISCO_NSINKQ] = {};
        if (active && p < P.sink_n) {
            Planet pl = Rn.planets[p];
            pl.phi = S.planet_phi[run * DISCO_MAX_PLANETS + p];
            const double cosp = cos(pl.phi), sinp = sin(pl.phi);
            const double px = pl.r * cosp, py = pl.r * sinp;
            const double dx = gx - px, dy = gy - py;
            double mag = dx * dx + dy * dy;
            mag = sqrt(mag);
            double arg = 0.0;
            if (P.sink_type == 3) {
                if (mag <= P.sink_par[2]) arg = 1.0;
            } else if (P.sink_type == 2) {
                const double eps = pow(P.sink_par[2], P.sink_par[3]);
                double epsfactor = P.sink_par[4];
                if (epsfactor <= 0.0) epsfactor = 1.0;
                const double magPow = pow(mag, P.sink_par[3]);
                arg = exp(-magPow / (eps * epsfactor));
            } else if (P.sink_type == 1) {
                const double eps = P.sink_par[2];
                arg = 1.0 - pow((mag / eps), P.sink_par[3]);
                arg = pow(arg, P.sink_par[4]);
                if (mag >= eps) arg = 0.0;
            }
            const double rate = P.sink_par[0] * pl.omega;
            const double surfdiff = rho * rate * arg;
            const double delta = fmax(0.0, P.sink_par[1]);  // sink.c:206-207 keeps only the lower clip
            const double vp_p = pl.r * pl.omega, vp_r = 0.0;
            const double vxp = vp_r * cosp - vp_p * sinp;
            const double vyp = vp_r * sinp + vp_p * cosp;
            double vxg = vr * cosg - vp * sing;
            double vyg = vr * sing + vp * cosg;
            const double vxr = vxg - vxp, vyr = vyg - vyp;
            const double cphi = dx / mag, sphi = dy / mag;
            const double acc_factor = dV * dt * surfdiff;
            const double vpr = cphi * vyr - sphi * vxr;
            a[3] = (1.0 - delta) * mag * vpr * acc_factor;  // Ls
            const double vxn = (cphi * cphi + (1.0 - delta) * sphi * sphi) * vxr + delta * sphi * cphi * vyr;
            const double vyn = delta * cphi * sphi * vxr + (sphi * sphi + (1.0 - delta) * cphi * cphi) * vyr;
            vxg = vxn + vxp;
            vyg = vyn + vyp;
            double vg_r = vxg * cosg + vyg * sing;
            double vg_p = -vxg * sing + vyg * cosg;
            a[2] = vg_p * r * acc_factor;  // accL
            cons[DDD] -= acc_factor;
            if (delta == 0.0) {
                vg_r = vr;
                vg_p = vp;
            }
            cons[SRR] -= vg_r * acc_factor;
            cons[LLL] -= r * vg_p * acc_factor;
            cons[SZZ] -= vz * acc_factor;
            const double v2 = vg_p * vg_p + vg_r * vg_r + vz * vz;
            cons[TAU] -= acc_factor * (0.5 * v2 + prim[PPP] / (rho * (P.gamma_law - 1.0)));
            a[1] = acc_factor;                                           // dM
            a[4] = 0.5 * v2 * acc_factor;                                // kin
            a[5] = prim[PPP] * acc_factor / (rho * (P.gamma_law - 1.0));  // therm
            a[6] = acc_factor * vxg;                                     // linXmom
            a[7] = acc_factor * vyg;                                     // linYmom
            double fr, fp, fz;
            planetary_force(pl, P.grav2D, r, phi, 0.0, &fr, &fp, &fz, 1);
            a[0] = -((rho - P.sink_rho0) * pl.r * fp * dV * dt);  // gravL
        }
        _Pragma("unroll") for (int k = 0; k < DISCO_NSINKQ; ++k) S.sinkc[(p * DISCO_NSINKQ + k) * N + c] = a[k];
    }
}

// First half of the update: the substep's start state (timestep.c:18-24's
// copy in the step's first substep (first == 1), adjust_RK_cons's average in
// the others,
// misc.c:226-244), the conservative update from the gathered face increments
// (riemann.c:314-317), add_source (misc.c:550-593) and, in the first substep,
// move_cells/clean_pi/calc_dp (misc.c:266-298, 8-36).  Returns the updated
// conserved state of a real cell in cons[] (and stores it).  SINK compiles in
// sink.c's cooling, damping and nozzle and rotframe.c's omega_src (a separate
// instantiation, so that the update without them keeps its registers).
template <class Ph, bool SINK = false>
DHD void update_a_cell(const DevState& S, int c, double RK, double dtfac, int first, double* cons) {
    constexpr int NQ = Ph::NQ;
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

    if (A.flags & ANN_REAL) {
        double prim[NQ];
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
            const double u = S.cons[q * N + c];
            if (first == 1) {
                st_stream(&S.rkcons[q * N + c], u);
                cons[q] = u;
            } else {
                cons[q] = (1. - RK) * u + RK * S.rkcons[q * N + c];
            }
            prim[q] = S.prim[q * N + c];
        }
        // phi faces i-1 (this cell is R) and i (this cell is L)
        constexpr int NF = Ph::NF;
        if (A.flags & ANN_PHIFLUX) {
            if (i == 0) {
                _Pragma("unroll") for (int q = 0; q < NF; ++q) cons[fslot(q, Ph::FGAP)] -= S.fphi[q * N + c];
                _Pragma("unroll") for (int q = 0; q < NF; ++q) cons[fslot(q, Ph::FGAP)] += S.fphi[q * N + cm];
            } else {
                _Pragma("unroll") for (int q = 0; q < NF; ++q) cons[fslot(q, Ph::FGAP)] += S.fphi[q * N + cm];
                _Pragma("unroll") for (int q = 0; q < NF; ++q) cons[fslot(q, Ph::FGAP)] -= S.fphi[q * N + c];
            }
        }
        gather_trans<NF, DIR_R, Ph::FGAP>(S, A, c, cm, cons);
        gather_trans<NF, DIR_Z, Ph::FGAP>(S, A, c, cm, cons);
        // add_source with the pre-move geometry and the substep's prims
        double xp[3] = {A.rp, phip, A.zp};
        double xm[3] = {A.rm, phip - dph, A.zm};
        double dV = get_dV(xp, xm, Pmax);
        if constexpr (Ph::NEWTONIAN && SINK) {  // misc.c:579-580
            // get_centroid_arr (geometry_functions.c:5-10) with the host's
            // r and z centroids: the initial condition damping relaxes to
            // hashes these bits for its noise
            const double xc[3] = {A.rc, get_centroid(xp[1], xm[1], 0), A.zc};
            if (P.cool_type) cooling_cell(S, Rn, A.run, prim, cons, xc, dV, dt);
            if (P.damp_type[0] | P.damp_type[1] | P.damp_type[2] | P.damp_type[3])
                damping_cell<Ph>(S, Rn, A.run, prim, cons, xc, dV, dt);
        }
        if constexpr (Ph::SOURCE_PASS) {
            // computed by source_cell (same operands, added to zero first)
            cons[SRR] += S.src[0 * N + c];
            cons[LLL] += S.src[1 * N + c];
            cons[SZZ] += S.src[2 * N + c];
            cons[TAU] += S.src[3 * N + c];
        } else {
            Ph::source(P, prim, cons, xp, xm, dV * dt);
        }
        for (int p = 0; p < P.npl; ++p) {
            Planet pl = Rn.planets[p];
            pl.phi = S.planet_phi[A.run * DISCO_MAX_PLANETS + p];
            euler::planet_src(P, pl, prim, cons, xp, xm, dV * dt);  // planet.c:138-158 (module independent)
        }
        if constexpr (Ph::NEWTONIAN && SINK && !Ph::MHD) {
            if (P.visc) {  // misc.c:585-587
                const double xc[3] = {get_centroid(xp[0], xm[0], 1), get_centroid(xp[1], xm[1], 0),
                                      get_centroid(xp[2], xm[2], 2)};
                double g[3][NQ];
                _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
                    g[0][q] = S.gradr[q * N + c];
                    g[1][q] = S.gradp[q * N + c];
                    g[2][q] = S.gradz[q * N + c];
                }
                const double nu = visc_nu(S, Rn, A.run, prim, xc);
                euler::visc_source(P, nu, prim, g[0], g[1], g[2], cons, xc, dV * dt);
            }
        }
        if constexpr (Ph::NEWTONIAN && SINK) {
            omega_src(P, prim, cons, xp, xm, dV * dt);  // misc.c:588
            // misc.c:589 (sink_src): the nozzle, then the accretion sinks
            bool in_ball = true;
            if (P.nozzle_type) in_ball = nozzle_cell<Ph::NC, NQ>(P, cons, xp, xm, dV, dt);
            if (S.sinkc) sink_cell(S, Rn, A.run, c, P.sink_type != 0 && in_ball, prim, cons, xp, xm, dV, dt);
        }
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) st_stream(&S.cons[q * N + c], cons[q]);
    }

    // moving mesh (first substep): every annulus, ghosts included
    if (first) {
        double pn = add_mul_rn(phip, S.wiph[c], dt);  // misc.c:275
        while (pn > Pmax) pn -= Pmax;
        while (pn < 0.0) pn += Pmax;
        double pm = add_mul_rn(S.piph[cm], S.wiph[cm], dt);
        while (pm > Pmax) pm -= Pmax;
        while (pm < 0.0) pm += Pmax;
        S.piph_n[c] = pn;
        S.dphi_n[c] = get_dp(pn, pm, Pmax);
        if (pn < pm && !(A.flags & ANN_ZCOPY)) S.rot_n[S.cell_ann[c]] = i;
    }
}

// GRMHD: add_source's geometric terms of a real cell (grmhd.c:339-454) with
// the operands update_a_cell would use (prims of the substep, the mesh
// before the move), as increments of S_r, L, S_z, tau.  A kernel of its own
// because the metric derivatives need ~190 registers on top of the update's
// gathers.
template <class Ph>
DHD void source_cell(const DevState& S, int c, double dtfac) {
    constexpr int NQ = Ph::NQ;
    const AnnDev& A = S.ann[S.cell_ann[c]];
    if (!(A.flags & ANN_REAL)) return;
    const PhysParams& P = S.runs[A.run].ph;
    const size_t N = S.ncell;
    const double dt = S.dt[A.run] * dtfac;
    double prim[NQ], src[NQ];
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
        prim[q] = S.prim[q * N + c];
        src[q] = 0.0;
    }
    double phip = S.piph[c];
    double xp[3] = {A.rp, phip, A.zp};
    double xm[3] = {A.rm, phip - S.dphi[c], A.zm};
    double dV = get_dV(xp, xm, P.phi_max);
    Ph::source(P, prim, src, xp, xm, dV * dt);
    S.src[0 * N + c] = src[SRR];
    S.src[1 * N + c] = src[LLL];
    S.src[2 * N + c] = src[SZZ];
    S.src[3 * N + c] = src[TAU];
}

template <class Ph>
DHD void b_faces_to_cell(const DevState& S, const AnnDev& A, int c, int first, double* cons);

// Geometry of a real cell on the moved mesh, as calc_prim builds it
// (misc.c:314-331).
struct CellGeom {
    double xp[3], xm[3], x[3], dV;
};

// Second half of the update, split so that a caller can run the guess
// network between its halves: update_b_pre does (MHD with CT)
// B_faces_to_cells (bfields.c:88-309) and the cell geometry; the caller
// solves cons2prim (calc_prim, misc.c:300-335); update_b_post stores the
// prims and, in the second substep, returns the next step's CFL candidate
// from them (misc.c:39-84 runs on exactly this state at the start of the next
// step, with this step's mesh speeds); update_b_ghost re-initialises fixed
// boundary cells (Boundary/fixed.c:11-60, boundary_functions.c:161-266).
// update_b_pre returns false for cells that are not real.
template <class Ph>
DHD bool update_b_pre(const DevState& S, int c, int first, double* cons, CellGeom* G) {
    const AnnDev& A = S.ann[S.cell_ann[c]];
    if (!(A.flags & ANN_REAL)) return false;
    const double Pmax = S.runs[A.run].ph.phi_max;
    const double phip = S.piph_n[c];
    const double phim = phip - S.dphi_n[c];
    if (Ph::MHD && S.Phi) b_faces_to_cell<Ph>(S, A, c, first, cons);
    G->xp[0] = A.rp;
    G->xp[1] = phip;
    G->xp[2] = A.zp;
    G->xm[0] = A.rm;
    G->xm[1] = phim;
    G->xm[2] = A.zm;
    G->dV = get_dV(G->xp, G->xm, Pmax);
    G->x[0] = A.rc;
    G->x[1] = 0.5 * (phim + phip);
    G->x[2] = A.zc;
    return true;
}

template <class Ph>
DHD double update_b_post(const DevState& S, int c, const double* prim, const CellGeom& G, int want_dt) {
    const AnnDev& A = S.ann[S.cell_ann[c]];
    const size_t N = S.ncell;
    _Pragma("unroll") for (int q = 0; q < Ph::NQ; ++q) st_stream(&S.prim[q * N + c], prim[q]);
    if (!want_dt) return 1.0e300;
    // the CFL step of the next step reads this mesh (piph_n is the current
    // mesh after the first substep)
    int i = c - A.off;
    int cm = A.off + (i == 0 ? A.np - 1 : i - 1);
    double w = .5 * (S.wiph[cm] + S.wiph[c]);
    return with_visc_dt<Ph>(S, A, Ph::mindt(S.runs[A.run].ph, prim, w, G.xp, G.xm), prim, G.xp, G.xm);
}

template <class Ph>
DHD void update_b_ghost(const DevState& S, int c) {
    const AnnDev& A = S.ann[S.cell_ann[c]];
    if (!(A.flags & ANN_GHOST_FILL)) return;
    const RunDev& Rn = S.runs[A.run];
    const size_t N = S.ncell;
    const double phip = S.piph_n[c];
    const double dph = S.dphi_n[c];
    double x[3];
    x[0] = A.rc_b;
    x[2] = A.zc_b;
    if (Rn.bc == DISCO_BC_FIXED)
        x[1] = phip - .5 * dph;
    else
        x[1] = 0.5 * (phip + (phip - dph));
    double p8[Ph::NQ > 8 ? Ph::NQ : 8];
    disco_initial(&Rn.ic, p8, x);
    if (Rn.ph.rot_frame) p8[UPP] -= Rn.ph.omega0;  // subtract_omega (fixed.c:37,54; boundary_functions.c:30)
    _Pragma("unroll") for (int q = 0; q < Ph::NQ; ++q) st_stream(&S.prim[q * N + c], p8[q]);
}

// set_cells_copy_distant (boundary_functions.c:90-159) for cell c of a ghost
// annulus from annulus `src` (the layer inside it), on the mesh after the
// substep's move: the overlap-weighted mean of src's prims, the overlaps
// summed in phi order from the src cell where the reference's loop starts
// for this cell (array index 0: the one holding its back edge; later cells:
// the first whose front is not behind the front of the cell before), then
// divided by the cell's dphi; the outflow diode sets the velocity field dq
// to zero where its sign is dsgn (boundary_functions.c:269-438).  Products
// are rounded before the adds, as the reference's host build does.
template <class Ph>
DHD void bc_copy_cell(const DevState& S, const AnnDev& G, int c, int src, int dq, int dsgn) {
    constexpr int NQ = Ph::NQ;
    const AnnDev& B = S.ann[src];
    const size_t N = S.ncell;
    const double Pmax = S.runs[B.run].ph.phi_max;
    const double phip = S.piph_n[c];
    const double phim = phip - S.dphi_n[c];
    const int n1 = B.np, rot = S.rot_n[src];
    const bool head = c == G.off;
    double key = head ? phim : S.piph_n[c - 1];
    while (key < 0.0) key += Pmax;
    while (key >= Pmax) key -= Pmax;
    int lo = 0, hi = n1;  // first rank (from the rotation point) whose front is > key (head) or >= key
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        int i = rot + mid;
        if (i >= n1) i -= n1;
        const double f = S.piph_n[B.off + i];
        if (head ? f > key : f >= key)
            hi = mid;
        else
            lo = mid + 1;
    }
    int i1 = rot + (lo == n1 ? 0 : lo);
    if (i1 >= n1) i1 -= n1;
    double sum[NQ];
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) sum[q] = 0.0;
    double phip1;
    do {
        phip1 = S.piph_n[B.off + i1];
        double phim1 = phip1 - S.dphi_n[B.off + i1];
        while (phip1 < phim) {
            phip1 += Pmax;
            phim1 += Pmax;
        }
        while (phim1 > phip) {
            phip1 -= Pmax;
            phim1 -= Pmax;
        }
        const double phi1 = phip < phip1 ? phip : phip1;
        const double phi2 = phim > phim1 ? phim : phim1;
        const double dphi = phi1 - phi2;
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) sum[q] = add_mul_rn(sum[q], S.prim[q * N + B.off + i1], dphi);
        i1 = i1 < n1 - 1 ? i1 + 1 : 0;
    } while (phip1 < phip);
    const double dph = S.dphi_n[c];
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
        double v = sum[q] / dph;
        if (q == dq && (dsgn > 0 ? v > 0 : v < 0)) v = 0.0;
        st_stream(&S.prim[q * N + c], v);
    }
}

// The whole second half with the module's own cons2prim.  cons[] holds the
// real cell's conserved state from update_a_cell; *c2p_it and *c2p_ok (if
// given) receive the Newton iteration count of a real cell's cons2prim and
// whether it converged.
template <class Ph>
DHD double update_b_cell(const DevState& S, int c, int first, int want_dt, double* cons, int* c2p_it = nullptr,
                         int* c2p_ok = nullptr) {
    constexpr int NQ = Ph::NQ;
    CellGeom G;
    if (!update_b_pre<Ph>(S, c, first, cons, &G)) {
        update_b_ghost<Ph>(S, c);
        return 1.0e300;
    }
    const PhysParams& P = S.runs[S.ann[S.cell_ann[c]].run].ph;
    const size_t N = S.ncell;
    double prim[NQ];
    if (Ph::C2P_GUESS) {  // calc_prim hands cons2prim the cell's current prims (misc.c:300-335)
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) prim[q] = S.prim[q * N + c];
    }
    int ok = 1;
    int it = Ph::cons2prim(P, cons, prim, G.x, G.dV, &ok);
    if (c2p_it) *c2p_it = it;
    if (c2p_ok) *c2p_ok = ok;
    return update_b_post<Ph>(S, c, prim, G, want_dt);
}

// Periodic z: what a one-rank MPI run's self-wraparound exchange does
// (exchange.c:173-279 with mpisetup.c:17,38; copy_lite_to_cell
// exchange.c:58-70).  Every ghost z-annulus becomes a copy of the real
// annulus one period away -- primitives, mesh speeds, face fluxes and, after
// the mesh moved, the mesh -- and takes over that annulus' rotation point.
// Runs after the update, so r-ghost columns copy the freshly re-initialised
// source ghosts exactly as boundary_trans(1) before exchangeData(1) does
// (onestep.c:166-173).
template <class Ph>
DHD void zcopy_cell(const DevState& S, int c, int first) {
    constexpr int NQ = Ph::NQ;
    const int a = S.cell_ann[c];
    const AnnDev& A = S.ann[a];
    if (!(A.flags & ANN_ZCOPY)) return;
    const AnnDev& Sr = S.ann[A.zsrc];
    const size_t N = S.ncell;
    int i = c - A.off;
    int s = Sr.off + i;
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) S.prim[q * N + c] = S.prim[q * N + s];
    S.wiph[c] = S.wiph[s];
    if (Ph::MHD && S.Phi) {
        _Pragma("unroll") for (int f = 0; f < 5; ++f) {
            S.Phi[f * N + c] = S.Phi[f * N + s];
            S.rkPhi[f * N + c] = S.rkPhi[f * N + s];
        }
    }
    if (first) {
        S.piph_n[c] = S.piph_n[s];
        S.dphi_n[c] = S.dphi_n[s];
        int sm = Sr.off + (i == 0 ? Sr.np - 1 : i - 1);
        if (S.piph_n[s] < S.piph_n[sm]) S.rot_n[a] = i;
    }
}

// movePlanets (Planets/sun.c: none; seeded binary: phi += omega dt) then the
// planet wrap of clean_pi (misc.c:26-33), t += dt (timestep.c:47), and the
// opening of the next step's CFL reduction.
DHD void planets_run(const DevState& S, int r) {
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
    S.t[r] += dt;
}

// report.c:117-207 for one cell of a real annulus (mesh, prims and planet
// phases of the current state): Q[0] = rho dV; with two planets or more,
// Q[1] = -(rho - 1) r_1 f_phi,1 dV and Q[2] = -(rho - 1) r_0 f_phi,0 dV (the
// forces of planets 1 and 0 at (r, phi, 0) projected on the planet's
// direction, planetaryForce mode 1), and Q[3 + 2n], Q[4 + 2n] the same two
// where the cell lies farther than (n + 1) x Sink_Par3 from planet 1.
DHD void report_quantities(const DevState& S, const AnnDev& A, int c, double* Q) {
    const RunDev& Rn = S.runs[A.run];
    const PhysParams& P = Rn.ph;
    const size_t N = S.ncell;
    for (int q = 0; q < DISCO_NREPORT; ++q) Q[q] = 0.0;
    const double phip = S.piph[c];
    const double dph = S.dphi[c];
    const double phi = phip - .5 * dph;
    const double rho = S.prim[RHO * N + c];
    double xp[3] = {A.rp, phip, A.zp};
    double xm[3] = {A.rm, phip - dph, A.zm};
    const double dV = get_dV(xp, xm, P.phi_max);
    const double r = get_centroid(xp[0], xm[0], 1);
    Q[0] = rho * dV;
    if (P.npl > 1) {
        Planet p0 = Rn.planets[0], p1 = Rn.planets[1];
        p0.phi = S.planet_phi[A.run * DISCO_MAX_PLANETS + 0];
        p1.phi = S.planet_phi[A.run * DISCO_MAX_PLANETS + 1];
        double fr, fz, fp, fp2;
        planetary_force(p0, P.grav2D, r, phi, 0.0, &fr, &fp2, &fz, 1);
        planetary_force(p1, P.grav2D, r, phi, 0.0, &fr, &fp, &fz, 1);
        const double rS = p0.r, rp = p1.r, pp = p1.phi;
        double dx = r * cos(phi) - rp * cos(pp);
        double dy = r * sin(phi) - rp * sin(pp);
        double script_r = sqrt(dx * dx + dy * dy);
        double t1v = (rho - 1.0) * rp * fp * dV;
        double t2v = (rho - 1.0) * rS * fp2 * dV;
        Q[1] = -t1v;
        Q[2] = -t2v;
        for (int n = 0; n < 5; ++n) {
            double r_cut = (double)(n + 1.) * P.report_soft;
            if (script_r > r_cut) {
                Q[3 + 2 * n] = -t1v;
                Q[4 + 2 * n] = -t2v;
            }
        }
    }
}

// get_diagnostics of Diagnostics/diag_euler.c:17-44 at the cell's centroid:
// rho, 2 pi r rho v_r, 2 pi r rho v_z, the torque density of planet 1 (when
// there are two), the advective angular-momentum flux, Omega, the angular
// momentum per unit r and P.
DHD void diag_quantities(const DevState& S, const RunDev& Rn, int run, const double* x, const double* prim,
                         double* Q) {
    double r = x[0];
    double rho = prim[RHO];
    double vr = prim[URR];
    double omega = prim[UPP];
    double vz = prim[UZZ];
    double Pp = prim[PPP];
    Q[0] = rho;
    Q[1] = 2. * M_PI * r * rho * vr;
    Q[2] = 2. * M_PI * r * rho * vz;
    double Fr, Fp = 0.0, Fz;
    if (Rn.ph.npl > 1) {
        Planet pl = Rn.planets[1];
        pl.phi = S.planet_phi[run * DISCO_MAX_PLANETS + 1];
        planetary_force(pl, Rn.ph.grav2D, r, x[1], x[2], &Fr, &Fp, &Fz);
    }
    Q[3] = 2. * M_PI * r * rho * (r * Fp);
    Q[4] = 2. * M_PI * r * rho * (r * r * omega * vr);
    Q[5] = omega;
    Q[6] = 2 * M_PI * r * rho * r * r * omega;
    Q[7] = Pp;
}

}  // namespace disco

#include "ct_cells.h"
