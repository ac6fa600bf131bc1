// Constrained transport on the sheared annulus mesh (bfields.c), per cell.
//
// The reference keeps one magnetic flux per face (Phi[5] per cell: its phi
// face and the r/z faces its front edge ends, paul.h:196-216) and updates it
// by Stokes' theorem from edge EMFs.  It walks explicit face lists in merge
// order; here every stage is a gather owned by one cell:
//
//   emf_sum_cell    riemann_trans's EMF accumulation (riemann.c:172-212):
//                   sums of .5*E*frac over the faces of each track a cell touches
//   emf_face_cell   avg_Efields: phi-neighbour averages (bfields.c:400-444,
//                   evaluated where needed), per-face average across the
//                   sheared neighbour (bfields.c:446-538) and
//                   subtract_advective_B_fluxes (bfields.c:646-690)
//   ct_update_cell  update_B_fluxes (bfields.c:313-382) and make_edge_adjust
//                   (bfields.c:888-1127): each owned face sums the EMFs of its
//                   edges; the azimuthal edges come from a 4-way merge of the
//                   fronts of the four annuli around the edge, restricted to
//                   the face's own span
//   flip_flag_cell  check_flipped (bfields.c:693-736)
//   flip_chain_cell flip_fluxes (bfields.c:738-814): order dependent, so each
//                   chain of interacting flips is replayed by one thread in the
//                   reference's loop order
//   b_faces_to_cell B_faces_to_cells (bfields.c:88-309)
//
// Faces are named by (owner cell, side): side 0 = the owner is L of the track
// (this | outer or up) and its front ends the face (LRtype 0); side 1 = the
// owner is R of the track (inner or down | this) (LRtype 1).
#pragma once

namespace disco {

struct FaceId {
    int cell;
    int side;  // 0: owner is L (hi track of the owner); 1: owner is R (lo track)
};

DHD bool same_face(FaceId a, FaceId b) { return a.cell == b.cell && a.side == b.side; }

// Sets bits of the device error word (atomic on the device).
DHD void err_or(int* p, int v) {
#if defined(__HIP_DEVICE_COMPILE__)
    atomicOr(p, v);
#else
    *p |= v;
#endif
}

DHD int cell_pred(const AnnDev& A, int c) {
    int i = c - A.off;
    return A.off + (i == 0 ? A.np - 1 : i - 1);
}
DHD int cell_succ(const AnnDev& A, int c) {
    int i = c - A.off;
    return A.off + (i + 1 == A.np ? 0 : i + 1);
}

// Neighbouring faces in the track's merge order (faces_alt.c:135-173, ties:
// the lower annulus' front first), from the partner indices of stage A.
template <int D>
DHD FaceId face_prev(const DevState& S, FaceId F) {
    using Dr = Dir<D>;
    const AnnDev& A = S.ann[S.cell_ann[F.cell]];
    int Fm = cell_pred(A, F.cell);
    if (F.side == 0) {
        const AnnDev& B = S.ann[Dr::hi(A)];
        const int* part = Dr::part_hi(S);
        int cnt = wrap_idx(part[F.cell] - part[Fm], B.np);
        if (cnt == 0) return FaceId{Fm, 0};
        return FaceId{B.off + wrap_idx(part[F.cell] - 1, B.np), 1};
    }
    const AnnDev& W = S.ann[Dr::lo(A)];
    const int* part = Dr::part_lo(S);
    int cnt = wrap_idx(part[F.cell] - part[Fm], W.np);
    if (cnt == 0) return FaceId{Fm, 1};
    return FaceId{W.off + wrap_idx(part[F.cell] - 1, W.np), 0};
}

template <int D>
DHD FaceId face_next(const DevState& S, FaceId F) {
    using Dr = Dir<D>;
    const AnnDev& A = S.ann[S.cell_ann[F.cell]];
    int Fp = cell_succ(A, F.cell);
    if (F.side == 0) {
        const AnnDev& B = S.ann[Dr::hi(A)];
        const int* part = Dr::part_hi(S);
        int cnt = wrap_idx(part[Fp] - part[F.cell], B.np);
        if (cnt == 0) return FaceId{Fp, 0};
        return FaceId{B.off + part[F.cell], 1};
    }
    const AnnDev& W = S.ann[Dr::lo(A)];
    const int* part = Dr::part_lo(S);
    int cnt = wrap_idx(part[Fp] - part[F.cell], W.np);
    if (cnt == 0) return FaceId{Fp, 1};
    return FaceId{W.off + part[F.cell], 0};
}

// The face's other cell, as the owner of its own face in the same track
// (what get_phi_pointer returns after the LRtype toggle, bfields.c:738-758).
template <int D>
DHD FaceId face_other(const DevState& S, FaceId F) {
    using Dr = Dir<D>;
    const AnnDev& A = S.ann[S.cell_ann[F.cell]];
    if (F.side == 0) {
        const AnnDev& B = S.ann[Dr::hi(A)];
        return FaceId{B.off + Dr::part_hi(S)[F.cell], 1};
    }
    const AnnDev& W = S.ann[Dr::lo(A)];
    return FaceId{W.off + Dr::part_lo(S)[F.cell], 0};
}

// First face of the track in the reference's face list: the merge starts at
// the first cell in quadrant I of each annulus (faces_alt.c:57-79), which is
// the annulus' smallest front (its rotation point).
template <int D>
DHD FaceId face_zero(const DevState& S, FaceId F) {
    using Dr = Dir<D>;
    const AnnDev& A = S.ann[S.cell_ann[F.cell]];
    int la, ua;
    if (F.side == 0) {
        la = S.cell_ann[F.cell];
        ua = Dr::hi(A);
    } else {
        la = Dr::lo(A);
        ua = S.cell_ann[F.cell];
    }
    const AnnDev& L = S.ann[la];
    const AnnDev& U = S.ann[ua];
    int cl = L.off + S.rot[la];
    int cu = U.off + S.rot[ua];
    if (S.piph[cl] <= S.piph[cu]) return FaceId{cl, 0};
    return FaceId{cu, 1};
}

template <int D>
DHD size_t face_slot(const DevState& S, FaceId F) {
    using Dr = Dir<D>;
    return (size_t)(F.side == 0 ? Dr::PHI_HI : Dr::PHI_LO) * S.ncell + F.cell;
}

template <int D>
DHD bool face_flag(const DevState& S, FaceId F) {
    using Dr = Dir<D>;
    return (S.fflags[F.cell] & (F.side == 0 ? Dr::FL_HI : Dr::FL_LO)) != 0;
}

// ------------------------------------------------------------------ EMF sums
// riemann_trans (riemann.c:172-212): every face of a flux-carrying track adds
// .5*E*frac (frac = face dphi / cell dphi) to the edge slots of its two cells.
template <int D>
DHD void emf_sum_dir(const DevState& S, const AnnDev& A, const PhysParams& P, int c, int cm) {
    using Dr = Dir<D>;
    const size_t N = S.ncell;
    const double Pmax = P.phi_max;
    const double phip = S.piph[c], dph = S.dphi[c];
    double sE = 0.0, sB = 0.0;
    if (Dr::lo(A) >= 0 && (A.flags & Dr::FLAG_LO)) {  // this cell is R
        const AnnDev& W = S.ann[Dr::lo(A)];
        const int* part = Dr::part_lo(S);
        int ca = W.off + part[c];
        TFace F = t_face<D>(S.piph[ca], S.dphi[ca], phip, dph, 1, W, Pmax);
        double fr = F.dp / dph;
        sE += .5 * S.emf_t[(Dr::EMF_LO + 0) * N + c] * fr;
        sB += .5 * S.emf_t[(Dr::EMF_LO + 1) * N + c] * fr;
        int start = part[cm];
        int cnt = wrap_idx(part[c] - start, W.np);
        for (int t = 0; t < cnt; ++t) {
            ca = W.off + wrap_idx(start + t, W.np);
            F = t_face<D>(S.piph[ca], S.dphi[ca], phip, dph, 0, W, Pmax);
            fr = F.dp / dph;
            sE += .5 * S.emf_t[(Dr::EMF_HI + 0) * N + ca] * fr;
            sB += .5 * S.emf_t[(Dr::EMF_HI + 1) * N + ca] * fr;
        }
    }
    st_stream(&S.esum[Dr::SUM_LO * N + c], sE);
    st_stream(&S.esum[(Dr::SUM_LO + 4) * N + c], sB);
    sE = 0.0;
    sB = 0.0;
    if (Dr::hi(A) >= 0 && (A.flags & Dr::FLAG_HI)) {  // this cell is L
        const AnnDev& B = S.ann[Dr::hi(A)];
        const int* part = Dr::part_hi(S);
        int cb = B.off + part[c];
        TFace F = t_face<D>(phip, dph, S.piph[cb], S.dphi[cb], 0, A, Pmax);
        double fr = F.dp / dph;
        sE += .5 * S.emf_t[(Dr::EMF_HI + 0) * N + c] * fr;
        sB += .5 * S.emf_t[(Dr::EMF_HI + 1) * N + c] * fr;
        int start = part[cm];
        int cnt = wrap_idx(part[c] - start, B.np);
        for (int t = 0; t < cnt; ++t) {
            cb = B.off + wrap_idx(start + t, B.np);
            F = t_face<D>(phip, dph, S.piph[cb], S.dphi[cb], 1, A, Pmax);
            fr = F.dp / dph;
            sE += .5 * S.emf_t[(Dr::EMF_LO + 0) * N + cb] * fr;
            sB += .5 * S.emf_t[(Dr::EMF_LO + 1) * N + cb] * fr;
        }
    }
    st_stream(&S.esum[Dr::SUM_HI * N + c], sE);
    st_stream(&S.esum[(Dr::SUM_HI + 4) * N + c], sB);
}

DHD void emf_sum_cell(const DevState& S, int c, int has_z) {
    const AnnDev& A = S.ann[S.cell_ann[c]];
    const PhysParams& P = S.runs[A.run].ph;
    int cm = cell_pred(A, c);
    emf_sum_dir<DIR_R>(S, A, P, c, cm);
    if (has_z) emf_sum_dir<DIR_Z>(S, A, P, c, cm);
}

// avg_Efields, first loop (bfields.c:400-444): edge slot k at the front of
// cell c (whose phi-successor is cp) is the mean of the cell's and its
// successor's slots, each holding .5 x the cell's phi-face EMF plus the track
// sums.  Slots: E0, E1, E4, E5 (phi-face E_z, E_z, E_r, E_r), then B0, B1, B4,
// B5 (phi-face B_r, B_r, B_z, B_z).  Evaluated where the face stage needs it
// instead of being stored.
DHD double edge_avg1(const DevState& S, int k, int c, int cp) {
    const size_t N = S.ncell;
    const int src = k < 4 ? (k < 2 ? 0 : 2) : (k < 6 ? 1 : 3);
    const double X = S.emf_phi[src * N + c];
    const double* s = S.esum;
    return .5 * ((.5 * X + s[k * N + c]) + (.5 * X + s[k * N + cp]));
}

// avg_Efields, face loops (bfields.c:446-538): the edge at the owner's front
// averages the owner's slot with the other cell's two edge values
// interpolated to that front; then the mesh-advection part of the EMF is
// removed (bfields.c:646-690).
template <int D>
DHD void emf_face_dir(const DevState& S, const AnnDev& A, int c, double* Ehi, double* Bhi, double* Elo,
                      double* Blo) {
    using Dr = Dir<D>;
    const double Pmax = S.runs[A.run].ph.phi_max;
    const int cp = cell_succ(A, c);
    *Ehi = edge_avg1(S, Dr::E_HI, c, cp);
    *Bhi = edge_avg1(S, Dr::E_HI + 4, c, cp);
    *Elo = edge_avg1(S, Dr::E_LO, c, cp);
    *Blo = edge_avg1(S, Dr::E_LO + 4, c, cp);
    if (Dr::hi(A) >= 0) {  // owned face, LRtype 0: other cell c2 = R
        const AnnDev& B = S.ann[Dr::hi(A)];
        int c2 = B.off + Dr::part_hi(S)[c];
        int c2m = cell_pred(B, c2);
        int c2p = cell_succ(B, c2);
        double dp1 = get_dp(S.piph[c2], S.piph[c], Pmax);
        double dp2 = S.dphi[c2] - dp1;
        double Eavg = (dp2 * edge_avg1(S, Dr::E_LO, c2, c2p) + dp1 * edge_avg1(S, Dr::E_LO, c2m, c2)) / (dp1 + dp2);
        double Bavg =
            (dp2 * edge_avg1(S, Dr::E_LO + 4, c2, c2p) + dp1 * edge_avg1(S, Dr::E_LO + 4, c2m, c2)) / (dp1 + dp2);
        *Ehi = .5 * (*Ehi + Eavg);
        *Bhi = .5 * (*Bhi + Bavg);
    }
    if (Dr::lo(A) >= 0) {  // owned face, LRtype 1: other cell c2 = L
        const AnnDev& W = S.ann[Dr::lo(A)];
        int c2 = W.off + Dr::part_lo(S)[c];
        int c2m = cell_pred(W, c2);
        int c2p = cell_succ(W, c2);
        double dp1 = get_dp(S.piph[c2], S.piph[c], Pmax);
        double dp2 = S.dphi[c2] - dp1;
        double Eavg = (dp2 * edge_avg1(S, Dr::E_HI, c2, c2p) + dp1 * edge_avg1(S, Dr::E_HI, c2m, c2)) / (dp1 + dp2);
        double Bavg =
            (dp2 * edge_avg1(S, Dr::E_HI + 4, c2, c2p) + dp1 * edge_avg1(S, Dr::E_HI + 4, c2m, c2)) / (dp1 + dp2);
        *Elo = .5 * (*Elo + Eavg);
        *Blo = .5 * (*Blo + Bavg);
    }
}

DHD void emf_face_cell(const DevState& S, int c, int has_z) {
    const AnnDev& A = S.ann[S.cell_ann[c]];
    const size_t N = S.ncell;
    const double w = S.wiph[c];
    double E1, B1, E0, B0;
    emf_face_dir<DIR_R>(S, A, c, &E1, &B1, &E0, &B0);
    E0 -= A.rm * w * B0;  // hL = r_jph[j-1]
    E1 -= A.rp * w * B1;  // hR = r_jph[j]
    st_stream(&S.efin[0 * N + c], E0);
    st_stream(&S.efin[1 * N + c], E1);
    if (has_z) {
        double E5, B5, E4, B4;
        emf_face_dir<DIR_Z>(S, A, c, &E5, &B5, &E4, &B4);
        E4 += A.rc * w * B4;  // hD = hU = r centroid
        E5 += A.rc * w * B5;
        st_stream(&S.efin[2 * N + c], E4);
        st_stream(&S.efin[3 * N + c], E5);
    }
}

// ------------------------------------------------------------------ edges
// One azimuthal edge (r_jph[j], z_kph[k]) and its four annuli by rank:
// 0 = (j,k), 1 = (j+1,k), 2 = (j,k+1), 3 = (j+1,k+1) (bfields.c:922-935).
struct Corner {
    int ann[4];
    double r;  // r_jph[j]
};

// EMF along the part of the edge that ends at the front of the cell of rank
// `w4` (make_edge_adjust's segment, bfields.c:946-1108), given the four
// current cells and their fronts ph and backs bk; returns E * dl * dt.
DHD double edge_segment(const DevState& S, const int* cur, const double* ph, const double* bk, int w4, double r,
                        double dt, double Pmax) {
    const size_t N = S.ncell;
    // get_which4 mode 0 toggles (bfields.c:834-885)
    int LR_D = get_signed_dp(ph[0], ph[1], Pmax) > 0.0;
    int LR_U = get_signed_dp(ph[2], ph[3], Pmax) > 0.0;
    int LR_alt = (w4 == 0 || w4 == 1) ? LR_U : LR_D;
    int UD_alt = (w4 == 0 || w4 == 2) ? (get_signed_dp(ph[1], ph[3], Pmax) > 0.0)
                                      : (get_signed_dp(ph[0], ph[2], Pmax) > 0.0);
    // owned-face E_phi: r-out [1], r-in [0], z-up [3], z-down [2]
    const double* Ep = S.emf_t;
#define EO(k) Ep[2 * N + cur[k]]
#define EI(k) Ep[5 * N + cur[k]]
#define EU(k) Ep[8 * N + cur[k]]
#define ED(k) Ep[11 * N + cur[k]]
    double E;
    if (w4 == 0) {
        E = .25 * (EU(0) + EO(0));
        E += .25 * (LR_alt == 0 ? EO(2) : EI(3));
        E += .25 * (UD_alt == 0 ? EU(1) : ED(3));
    } else if (w4 == 1) {
        E = .25 * (EU(1) + EI(1));
        E += .25 * (LR_alt == 0 ? EO(2) : EI(3));
        E += .25 * (UD_alt == 0 ? EU(0) : ED(2));
    } else if (w4 == 2) {
        E = .25 * (ED(2) + EO(2));
        E += .25 * (LR_alt == 0 ? EO(0) : EI(1));
        E += .25 * (UD_alt == 0 ? EU(1) : ED(3));
    } else {
        E = .25 * (ED(3) + EI(3));
        E += .25 * (LR_alt == 0 ? EO(0) : EI(1));
        E += .25 * (UD_alt == 0 ? EU(0) : ED(2));
    }
#undef EO
#undef EI
#undef EU
#undef ED
    // get_which4 mode 1 on the back edges: the latest back is the segment start
    double b1 = get_signed_dp(bk[0], bk[1], Pmax) < 0.0 ? bk[1] : bk[0];
    double b2 = get_signed_dp(bk[2], bk[3], Pmax) < 0.0 ? bk[3] : bk[2];
    double xm1 = get_signed_dp(b1, b2, Pmax) < 0.0 ? b2 : b1;
    double dl = r * get_dp(ph[w4], xm1, Pmax);
    return E * dl * dt;
}

// Partner of cell x in the annulus at offset (dj, dk) from x's own: the
// first cell of that annulus whose front does not precede x's front in the
// (k, j) order of the merges (stage A's binary searches).
DHD int nbr_part(const DevState& S, int x, int dj, int dk) {
    if (dk == 0) return dj > 0 ? S.lbo[x] : S.ubi[x];
    if (dj == 0) return dk > 0 ? S.lbu[x] : S.ubd[x];
    int d = (dk > 0 ? 0 : 2) + (dj > 0 ? 0 : 1);
    return S.pdg[(size_t)d * S.ncell + x];
}

// Sum of E dl dt over the segments of corner K inside face F, whose two
// cells are pc0 (rank p0) and pc1 (rank p1), which ends at the front of its
// owner X (rank orank) and starts at the front of cell P (rank prank).  The
// other two annuli (ranks q0 < q1) contribute the fronts that split the
// face; their positions come from X's and P's partners in those annuli.
DHD double edge_face_sum(const DevState& S, const Corner& K, int p0, int p1, int pc0, int pc1, int X, int orank,
                         int Pc, int prank, double dt, double Pmax) {
    int q0, q1;
    if (p0 == 0 && p1 == 1) { q0 = 2; q1 = 3; }
    else if (p0 == 2 && p1 == 3) { q0 = 0; q1 = 1; }
    else if (p0 == 0 && p1 == 2) { q0 = 1; q1 = 3; }
    else { q0 = 0; q1 = 2; }
    const int aq0 = K.ann[q0], aq1 = K.ann[q1];
    const int n0 = S.ann[aq0].np, n1 = S.ann[aq1].np;
    const int r0 = S.rot[aq0], r1 = S.rot[aq1];
    // rank r sits at (dj, dk) = (r & 1, r >> 1) in the corner
    int a0 = wrap_idx(nbr_part(S, Pc, (q0 & 1) - (prank & 1), (q0 >> 1) - (prank >> 1)) - r0, n0);
    int b0 = wrap_idx(nbr_part(S, X, (q0 & 1) - (orank & 1), (q0 >> 1) - (orank >> 1)) - r0, n0);
    int a1 = wrap_idx(nbr_part(S, Pc, (q1 & 1) - (prank & 1), (q1 >> 1) - (prank >> 1)) - r1, n1);
    int b1 = wrap_idx(nbr_part(S, X, (q1 & 1) - (orank & 1), (q1 >> 1) - (orank >> 1)) - r1, n1);
    int m0 = wrap_idx(b0 - a0, n0);
    int m1 = wrap_idx(b1 - a1, n1);
    // the face's two cells stay; the other two annuli's cells advance
    const int o0 = S.ann[aq0].off, o1 = S.ann[aq1].off;
    int cur[4];
    double ph[4], bk[4];
    cur[p0] = pc0;
    cur[p1] = pc1;
    ph[p0] = S.piph[pc0];
    bk[p0] = ph[p0] - S.dphi[pc0];
    ph[p1] = S.piph[pc1];
    bk[p1] = ph[p1] - S.dphi[pc1];
    auto at = [&](int q, int o, int n, int rot, int pos) {  // corner_cell, loads kept
        int i = rot + pos;
        if (i >= n) i -= n;
        if (i >= n) i -= n;
        const int cc = o + i;
        cur[q] = cc;
        ph[q] = S.piph[cc];
        bk[q] = ph[q] - S.dphi[cc];
    };
    double sum = 0.0;
    int i0 = 0, i1 = 0;
    at(q0, o0, n0, r0, a0);
    at(q1, o1, n1, r1, a1);
    while (i0 < m0 || i1 < m1) {
        bool take0;
        if (i0 >= m0)
            take0 = false;
        else if (i1 >= m1)
            take0 = true;
        else
            take0 = get_signed_dp(ph[q0], ph[q1], Pmax) <= 0.0;  // tie: lower rank q0 first
        sum += edge_segment(S, cur, ph, bk, take0 ? q0 : q1, K.r, dt, Pmax);
        if (take0)
            at(q0, o0, n0, r0, a0 + (++i0));
        else
            at(q1, o1, n1, r1, a1 + (++i1));
    }
    sum += edge_segment(S, cur, ph, bk, orank, K.r, dt, Pmax);
    return sum;
}

// ------------------------------------------------------------------ Stokes
// update_B_fluxes (bfields.c:313-382): the edge at each face's end changes
// the owned face, the owner's phi face and the next face of the track; as a
// gather, an owned face gains its predecessor's end EMF and loses its own.
// make_edge_adjust (3D) adds the azimuthal edges on the face's two other
// sides: D +, U -, L -, R + (add_E_phi, bfields.c:384-389).
template <int D>
DHD double ct_face_update(const DevState& S, const AnnDev& A, int c, int side, double dt, int has_z,
                          double* dPhi0) {
    using Dr = Dir<D>;
    const size_t N = S.ncell;
    const double Pmax = S.runs[A.run].ph.phi_max;
    const double dl = Dr::dl(A);
    FaceId F{c, side};
    FaceId Pv = face_prev<D>(S, F);
    double Eown = S.efin[(side == 0 ? Dr::E_HI : Dr::E_LO) * N + c];
    double Eprev = S.efin[(Pv.side == 0 ? Dr::E_HI : Dr::E_LO) * N + Pv.cell];
    double dOwn = Eown * dl * dt;
    double dPrev = Eprev * dl * dt;
    // r tracks: own -= E, next += E, phi face (LR0 +, LR1 -); z tracks opposite
    double delta = Dr::SGN * (dPrev - dOwn);
    *dPhi0 += (side == 0 ? Dr::SGN : -Dr::SGN) * dOwn;
    if (!has_z) return delta;

    // azimuthal edges of this face (make_edge_adjust)
    int other = side == 0 ? S.ann[Dr::hi(A)].off + Dr::part_hi(S)[c] : S.ann[Dr::lo(A)].off + Dr::part_lo(S)[c];
    int la = side == 0 ? S.cell_ann[c] : Dr::lo(A);    // lower annulus of the track
    int ua = side == 0 ? Dr::hi(A) : S.cell_ann[c];    // upper annulus
    int lcell = side == 0 ? c : other;
    int ucell = side == 0 ? other : c;
    const AnnDev& L = S.ann[la];
    const AnnDev& U = S.ann[ua];
    // ranks of the owner / predecessor within the track pair (0 = lower annulus)
    int o_up = side;                     // owner in the upper annulus of the track?
    int p_up = Pv.side;
    double sum = 0.0;
    if (D == DIR_R) {
        // r face between L=(j,k) and U=(j+1,k): top edge corner (j,k) [D],
        // bottom edge corner (j,k-1) [U]
        if (L.up >= 0 && U.up >= 0) {
            Corner K{{la, ua, L.up, U.up}, L.rp};
            sum += edge_face_sum(S, K, 0, 1, lcell, ucell, c, o_up, Pv.cell, p_up, dt, Pmax);
        }
        if (L.down >= 0 && U.down >= 0) {
            Corner K{{L.down, U.down, la, ua}, L.rp};
            sum -= edge_face_sum(S, K, 2, 3, lcell, ucell, c, 2 + o_up, Pv.cell, 2 + p_up, dt, Pmax);
        }
    } else {
        // z face between L=(j,k) and U=(j,k+1): edge at r_jph[j] is corner
        // (j,k) [L], edge at r_jph[j-1] is corner (j-1,k) [R]
        if (L.outer >= 0 && U.outer >= 0) {
            Corner K{{la, L.outer, ua, U.outer}, L.rp};
            sum -= edge_face_sum(S, K, 0, 2, lcell, ucell, c, 2 * o_up, Pv.cell, 2 * p_up, dt, Pmax);
        }
        if (L.inner >= 0 && U.inner >= 0) {
            Corner K{{L.inner, la, U.inner, ua}, L.rm};
            sum += edge_face_sum(S, K, 1, 3, lcell, ucell, c, 1 + 2 * o_up, Pv.cell, 1 + 2 * p_up, dt, Pmax);
        }
    }
    return delta + sum;
}

DHD void ct_update_cell(const DevState& S, int c, double RK, double dtfac, int first, int has_z) {
    const AnnDev& A = S.ann[S.cell_ann[c]];
    const size_t N = S.ncell;
    const double dt = S.dt[A.run] * dtfac;
    // the face fluxes at the substep's start, every cell: timestep.c:22's copy
    // in the step's first substep (first == 1), adjust_RK_cons's average
    // (misc.c:239-241) in the others; no kernel reads them between the stage
    // and here
    _Pragma("unroll") for (int f = 0; f < 5; ++f) {
        const double p = S.Phi[f * N + c];
        if (first == 1)
            st_stream(&S.rkPhi[f * N + c], p);
        else
            S.Phi[f * N + c] = (1. - RK) * p + RK * S.rkPhi[f * N + c];
    }
    double dPhi0 = 0.0;
    if (A.outer >= 0) S.Phi[2 * N + c] += ct_face_update<DIR_R>(S, A, c, 0, dt, has_z, &dPhi0);
    if (A.inner >= 0) S.Phi[1 * N + c] += ct_face_update<DIR_R>(S, A, c, 1, dt, has_z, &dPhi0);
    if (has_z) {
        if (A.up >= 0) S.Phi[4 * N + c] += ct_face_update<DIR_Z>(S, A, c, 0, dt, has_z, &dPhi0);
        if (A.down >= 0) S.Phi[3 * N + c] += ct_face_update<DIR_Z>(S, A, c, 1, dt, has_z, &dPhi0);
    }
    S.Phi[0 * N + c] += dPhi0;
}

// ------------------------------------------------------------------ flips
// check_flipped (bfields.c:693-736): after the mesh moved, a face whose
// owner's front now lies behind the previous face's owner's front is flagged.
// Reads the partners of the mesh before the move and both meshes.
template <int D>
DHD int flip_flag_dir(const DevState& S, const AnnDev& A, int c, const PhysParams& P) {
    using Dr = Dir<D>;
    int f = 0;
    for (int side = 0; side < 2; ++side) {
        int nb = side == 0 ? Dr::hi(A) : Dr::lo(A);
        if (nb < 0) continue;
        FaceId Pv = face_prev<D>(S, FaceId{c, side});
        double dp = get_signed_dp(S.piph_n[c], S.piph_n[Pv.cell], P.phi_max);
        if (dp < 0.) f |= side == 0 ? Dr::FL_HI : Dr::FL_LO;
    }
    return f;
}

DHD void flip_flag_cell(const DevState& S, int c, int has_z) {
    const AnnDev& A = S.ann[S.cell_ann[c]];
    const PhysParams& P = S.runs[A.run].ph;
    int f = flip_flag_dir<DIR_R>(S, A, c, P);
    if (has_z) f |= flip_flag_dir<DIR_Z>(S, A, c, P);
    S.fflags[c] = f;
}

// One flip at face F (bfields.c:781-808).  tm, tf, tp: whether the faces
// before F, F itself and the face after F have had their LRtype toggled by
// the flips replayed so far; a toggled face's flux is read and written
// through the slot of its other cell (get_phi_pointer after the toggle,
// bfields.c:738-758).  The loads and stores keep the reference's order, so
// slots that coincide through toggles behave as its pointers do.
template <int D>
DHD void flip_one(const DevState& S, FaceId F, bool tm, bool tf, bool tp) {
    FaceId Fm = face_prev<D>(S, F);
    FaceId Fp = face_next<D>(S, F);
    if (tm) Fm = face_other<D>(S, Fm);
    if (tp) Fp = face_other<D>(S, Fp);
    if (tf) F = face_other<D>(S, F);
    size_t sP = face_slot<D>(S, F), sPm = face_slot<D>(S, Fm), sPp = face_slot<D>(S, Fp);
    double* Ph = S.Phi;
    double* Rk = S.rkPhi;
    double Phi = Ph[sP];
    double RK_Phi = Rk[sP];
    Ph[sP] = Ph[sPm] + Phi;
    Rk[sP] = Rk[sPm] + RK_Phi;
    Ph[sPm] = -Phi;
    Rk[sPm] = -RK_Phi;
    Ph[sPp] = Ph[sPp] + Phi;
    Rk[sPp] = Rk[sPp] + RK_Phi;
}

// Replays the flips of one chain part in list order, from its first flagged
// face `first` to its last one, and returns the number of flips.  A flip
// toggles its face and the face before it, so when flip n runs, the faces
// toggled so far around it are: the face before n (if it was flagged, and n
// is not the list's first face), and -- only at the end of the list, after
// the start of the list was replayed -- the first two faces of the list.
template <int D>
DHD int flip_replay(const DevState& S, FaceId first, FaceId z0, int maxlen) {
    FaceId M = first;
    int len = 0;
    for (;;) {
        const FaceId m0 = face_prev<D>(S, M);
        const FaceId p1 = face_next<D>(S, M);
        const FaceId p2 = face_next<D>(S, p1);
        const bool f_z0 = face_flag<D>(S, z0);
        const bool done1 = same_face(p1, z0) && f_z0;  // the face after M: replayed already
        const bool done2 = (same_face(p2, z0) && f_z0) || (same_face(p1, z0) && face_flag<D>(S, face_next<D>(S, z0)));
        const bool tm = !same_face(M, z0) && face_flag<D>(S, m0);
        flip_one<D>(S, M, tm, done1, done1 != done2);
        if (++len > maxlen) {
            err_or(S.err, 1);
            return len;
        }
        // the next flagged face within two, not past the end of the list
        if (same_face(p1, z0)) break;
        if (face_flag<D>(S, p1)) {
            M = p1;
            continue;
        }
        if (same_face(p2, z0) || !face_flag<D>(S, p2)) break;
        M = p2;
    }
    return len;
}

// flip_fluxes (bfields.c:760-814).  A flip reads and writes the fluxes of the
// faces before and after it and toggles the LRtype of itself and the face
// before it, so flips one or two faces apart interact and must run in the
// reference's loop order (face list from its first face, the quadrant-I
// start of faces_alt.c:57-79, i.e. the lower annulus' smallest front); flips
// at least three faces apart are independent.  Each chain of interacting
// flips is replayed by one thread: the thread of its first flagged face
// counting from the start of the list.  A chain that runs over the end of
// the list into its start belongs to the chain at the start: that thread
// replays its own part, then the part at the end of the list (which sees the
// start's toggles).  A chain that closes on itself around the whole track is
// replayed from its first face to the end of the list.  This handles any
// pattern of flags, adjacent ones included (per-cell mesh motion).
template <int D>
DHD void flip_chain_face(const DevState& S, FaceId F) {
    const AnnDev& A = S.ann[S.cell_ann[F.cell]];
    const AnnDev& B = S.ann[F.side == 0 ? Dir<D>::hi(A) : Dir<D>::lo(A)];
    const int maxlen = A.np + B.np + 2;  // faces in the track
    const FaceId z0 = face_zero<D>(S, F);
    const FaceId b1 = face_prev<D>(S, F);
    const FaceId b2 = face_prev<D>(S, b1);
    // not the first flagged face of its chain from the start of the list
    if (!same_face(F, z0) && face_flag<D>(S, b1)) return;
    if (!same_face(F, z0) && !same_face(b1, z0) && face_flag<D>(S, b2)) return;
    // the chain's last face before the end of the list, and whether the chain
    // continues over the end into the start
    FaceId M = F;
    int len = 0;
    for (;;) {
        const FaceId m1 = face_next<D>(S, M), m2 = face_next<D>(S, m1);
        if (same_face(m1, z0)) break;
        if (face_flag<D>(S, m1)) {
            M = m1;
        } else if (!same_face(m2, z0) && face_flag<D>(S, m2)) {
            M = m2;
        } else {
            break;
        }
        if (++len > maxlen) {
            err_or(S.err, 1);
            return;
        }
    }
    const FaceId m1 = face_next<D>(S, M), m2 = face_next<D>(S, m1);
    FaceId c;  // the first flagged face of the list's start that M's chain reaches
    bool over_end = false;
    if (same_face(m1, z0)) {
        if (face_flag<D>(S, z0)) {
            c = z0;
            over_end = true;
        } else if (face_flag<D>(S, m2)) {
            c = m2;
            over_end = true;
        }
    } else if (same_face(m2, z0) && face_flag<D>(S, z0)) {
        c = z0;
        over_end = true;
    }
    if (over_end && !same_face(c, F)) return;  // the end part of a chain that starts at the list's start
    flip_replay<D>(S, F, z0, maxlen);
    if (over_end) return;  // closed around the track: replayed to the end of the list
    // the end of the list holds the rest of this chain when its last faces
    // reach F (which is then the list's first or second face)
    const FaceId e1 = face_prev<D>(S, z0), e2 = face_prev<D>(S, e1);
    const bool f_first = same_face(F, z0), f_second = same_face(b1, z0);
    FaceId E;
    if (face_flag<D>(S, e1) && (f_first || f_second))
        E = e1;
    else if (face_flag<D>(S, e2) && f_first)
        E = e2;
    else
        return;
    // back to the first face of that end part
    len = 0;
    for (;;) {
        const FaceId q1 = face_prev<D>(S, E), q2 = face_prev<D>(S, q1);
        if (face_flag<D>(S, q1))
            E = q1;
        else if (face_flag<D>(S, q2))
            E = q2;
        else
            break;
        if (++len > maxlen) {
            err_or(S.err, 1);
            return;
        }
    }
    flip_replay<D>(S, E, z0, maxlen);
}

template <int D>
DHD void flip_chain_dir(const DevState& S, const AnnDev& A, int c) {
    using Dr = Dir<D>;
    int f = S.fflags[c];
    if (Dr::hi(A) >= 0 && (f & Dr::FL_HI)) flip_chain_face<D>(S, FaceId{c, 0});
    if (Dr::lo(A) >= 0 && (f & Dr::FL_LO)) flip_chain_face<D>(S, FaceId{c, 1});
}

DHD void flip_chain_cell(const DevState& S, int c, int has_z) {
    const AnnDev& A = S.ann[S.cell_ann[c]];
    if (!(S.fflags[c])) return;
    flip_chain_dir<DIR_R>(S, A, c);
    if (has_z) flip_chain_dir<DIR_Z>(S, A, c);
}

// ------------------------------------------------------------------ B_faces_to_cells
// bfields.c:88-309 (type 1): cell field from the face fluxes, using the face
// list of the mesh before the move (with the flips' LRtype toggles) and the
// cell volume after it.
template <int D>
DHD bool face_toggled(const DevState& S, FaceId G, int first) {
    if (!first) return false;
    return face_flag<D>(S, G) != face_flag<D>(S, face_next<D>(S, G));
}

template <int D>
DHD void b_sum_dir(const DevState& S, const AnnDev& A, int c, int cm, int first, double Pmax, double* sPhi,
                   double* sdA) {
    using Dr = Dir<D>;
    const size_t N = S.ncell;
    const double phip = S.piph[c], dph = S.dphi[c];
    double P = 0.0, dA = 0.0;
    if (Dr::lo(A) >= 0) {  // this cell is R of the (lo | this) track
        const AnnDev& W = S.ann[Dr::lo(A)];
        const int* part = Dr::part_lo(S);
        int ca = W.off + part[c];
        TFace F = t_face<D>(S.piph[ca], S.dphi[ca], phip, dph, 1, W, Pmax);
        dA += F.dA;
        FaceId G{c, 1};
        P += face_toggled<D>(S, G, first) ? S.Phi[Dr::PHI_HI * N + ca] : S.Phi[Dr::PHI_LO * N + c];
        int start = part[cm];
        int cnt = wrap_idx(part[c] - start, W.np);
        for (int t = 0; t < cnt; ++t) {
            ca = W.off + wrap_idx(start + t, W.np);
            F = t_face<D>(S.piph[ca], S.dphi[ca], phip, dph, 0, W, Pmax);
            dA += F.dA;
            FaceId H{ca, 0};
            P += face_toggled<D>(S, H, first) ? S.Phi[Dr::PHI_LO * N + c] : S.Phi[Dr::PHI_HI * N + ca];
        }
    }
    if (Dr::hi(A) >= 0) {  // this cell is L of the (this | hi) track
        const AnnDev& B = S.ann[Dr::hi(A)];
        const int* part = Dr::part_hi(S);
        int cb = B.off + part[c];
        TFace F = t_face<D>(phip, dph, S.piph[cb], S.dphi[cb], 0, A, Pmax);
        dA += F.dA;
        FaceId G{c, 0};
        P += face_toggled<D>(S, G, first) ? S.Phi[Dr::PHI_LO * N + cb] : S.Phi[Dr::PHI_HI * N + c];
        int start = part[cm];
        int cnt = wrap_idx(part[c] - start, B.np);
        for (int t = 0; t < cnt; ++t) {
            cb = B.off + wrap_idx(start + t, B.np);
            F = t_face<D>(phip, dph, S.piph[cb], S.dphi[cb], 1, A, Pmax);
            dA += F.dA;
            FaceId H{cb, 1};
            P += face_toggled<D>(S, H, first) ? S.Phi[Dr::PHI_HI * N + c] : S.Phi[Dr::PHI_LO * N + cb];
        }
    }
    *sPhi = P;
    *sdA = dA;
}

template <class Ph>
DHD void b_faces_to_cell(const DevState& S, const AnnDev& A, int c, int first, double* cons) {
    const size_t N = S.ncell;
    const double Pmax = S.runs[A.run].ph.phi_max;
    int cm = cell_pred(A, c);
    double sPr, sAr;
    b_sum_dir<DIR_R>(S, A, c, cm, first, Pmax, &sPr, &sAr);
    double xp[3] = {A.rp, S.piph_n[c], A.zp};
    double xm[3] = {A.rm, S.piph_n[cm], A.zm};
    double dV = get_dV(xp, xm, Pmax);
    double rc = A.rc;
    double rfac = 1.0 / rc;  // bfield_scale_factor(r, 0)
    double hr = 1.0, hp = rc, hz = 1.0;
    double dA0 = (A.rp - A.rm) * (A.zp - A.zm);
    cons[BRR] = sPr * (dV * rfac / (sAr * hr));
    cons[BPP] = .5 * (S.Phi[0 * N + c] + S.Phi[0 * N + cm]) * dV / (dA0 * hp);
    if (A.up >= 0 || A.down >= 0) {
        double sPz, sAz;
        b_sum_dir<DIR_Z>(S, A, c, cm, first, Pmax, &sPz, &sAz);
        double zfac = 1.0;
        cons[BZZ] = sPz * (dV * zfac / (sAz * hz));
    } else {
        cons[BZZ] = 0.0;
    }
    S.cons[BRR * N + c] = cons[BRR];
    S.cons[BPP * N + c] = cons[BPP];
    S.cons[BZZ * N + c] = cons[BZZ];
}

}  // namespace disco
