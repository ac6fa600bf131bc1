// Test probe (built by tests/test_flip_cpu.py with the host compiler): the
// library's flux-flip replay (ct_cells.h flip_chain_cell, the code the GPU
// kernel k_flip_chain runs per cell) against a literal sequential
// restatement of the reference's flip_fluxes loop (bfields.c:760-814) on the
// same tracks, for random flag patterns -- adjacent flips, flips two apart,
// chains over the end of a face list and flags on every face included.
//
// usage: flip_probe PARFILE SEED PROB
//   builds the run's mesh (host set-up), its overlap partners and face lists,
//   flags each owned r and z face with probability PROB, random face fluxes,
//   then replays the flips (a) cell by cell in reverse cell order and (b)
//   track by track in list order with explicit LRtype toggles; prints
//   "flags N mismatches M err E".
#include <cstdio>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "host_setup.h"
#include "step_cells.h"

using namespace disco;

template <int D>
static size_t slot_t(const DevState& S, FaceId f, bool tog) {
    return face_slot<D>(S, tog ? face_other<D>(S, f) : f);
}

// bfields.c:760-814 over one track, faces listed from the list's first face
template <int D>
static void reference_loop(const DevState& S, FaceId any, int nfaces) {
    const FaceId z0 = face_zero<D>(S, any);
    std::vector<FaceId> list;
    FaceId f = z0;
    for (int n = 0; n < nfaces; ++n) {
        list.push_back(f);
        f = face_next<D>(S, f);
    }
    if (!same_face(f, z0)) {
        std::printf("face list does not close\n");
        std::exit(3);
    }
    std::vector<char> tog(nfaces, 0);
    for (int n = 0; n < nfaces; ++n) {
        if (!face_flag<D>(S, list[n])) continue;
        const int np = (n + 1) % nfaces, nm = (n + nfaces - 1) % nfaces;
        const size_t sP = slot_t<D>(S, list[n], tog[n]), sPm = slot_t<D>(S, list[nm], tog[nm]),
                     sPp = slot_t<D>(S, list[np], tog[np]);
        double Phi = S.Phi[sP], RK_Phi = S.rkPhi[sP];
        S.Phi[sP] = S.Phi[sPm] + Phi;
        S.rkPhi[sP] = S.rkPhi[sPm] + RK_Phi;
        S.Phi[sPm] = -Phi;
        S.rkPhi[sPm] = -RK_Phi;
        S.Phi[sPp] = S.Phi[sPp] + Phi;
        S.rkPhi[sPp] = S.rkPhi[sPp] + RK_Phi;
        tog[n] ^= 1;
#ifndef PLANT_NO_PREV_TOGGLE
        tog[nm] ^= 1;
#endif
    }
}

int main(int argc, char** argv) {
    if (argc < 4) return 2;
    std::ifstream pf(argv[1]);
    std::stringstream ss;
    ss << pf.rdbuf();
    disco_params par;
    std::string err;
    if (read_par_text(ss.str(), &par, &err)) return 1;
    std::vector<HostRun> runs(1);
    if (setup_run(par, &runs[0], &err)) {
        std::fprintf(stderr, "%s\n", err.c_str());
        return 1;
    }
    std::vector<AnnDev> hann;
    std::vector<RunDev> hruns;
    std::vector<int> cell0, ann0;
    std::vector<std::vector<int>> perm;
    StepShape sh{};
    build_annulus_tables(runs, &hann, &hruns, &cell0, &ann0, &sh, &perm, 1);
    const HostRun& R = runs[0];
    const int N = R.ncell, na = (int)hann.size();
    std::vector<int> cell_ann(N), rot(na);
    std::vector<double> piph(N), dphi(N);
    for (int a = 0; a < na; ++a) {
        for (int i = 0; i < hann[a].np; ++i) cell_ann[hann[a].off + i] = a;
        rot[a] = R.rot[hann[a].j + R.Nr * hann[a].k];
    }
    for (int c = 0; c < N; ++c) {
        piph[perm[0][c]] = R.piph[c];
        dphi[perm[0][c]] = R.dphi[c];
    }
    std::vector<int> lbo(N, -1), ubi(N, -1), lbu(N, -1), ubd(N, -1), fflags(N, 0);
    std::vector<double> Phi(5 * (size_t)N), rkPhi(5 * (size_t)N);
    int errw = 0;
    DevState S{};
    S.ann = hann.data();
    S.cell_ann = cell_ann.data();
    S.runs = hruns.data();
    S.ncell = N;
    S.nruns = 1;
    S.piph = piph.data();
    S.dphi = dphi.data();
    S.lbo = lbo.data();
    S.ubi = ubi.data();
    S.lbu = lbu.data();
    S.ubd = ubd.data();
    S.rot = rot.data();
    S.Phi = Phi.data();
    S.rkPhi = rkPhi.data();
    S.fflags = fflags.data();
    S.err = &errw;
    for (int c = 0; c < N; ++c) partners_cell(S, c);
    std::mt19937_64 rng(std::stoull(argv[2]));
    std::uniform_real_distribution<double> U(-1.0, 1.0), U01(0.0, 1.0);
    const double prob = std::stod(argv[3]);
    int nflags = 0;
    for (int c = 0; c < N; ++c) {
        const AnnDev& A = hann[cell_ann[c]];
        int f = 0;
        if (A.outer >= 0 && U01(rng) < prob) f |= 1;
        if (A.inner >= 0 && U01(rng) < prob) f |= 2;
        if (A.up >= 0 && U01(rng) < prob) f |= 4;
        if (A.down >= 0 && U01(rng) < prob) f |= 8;
        fflags[c] = f;
        nflags += __builtin_popcount(f);
    }
    for (size_t k = 0; k < Phi.size(); ++k) {
        Phi[k] = U(rng);
        rkPhi[k] = U(rng);
    }
    const std::vector<double> Phi0 = Phi, rk0 = rkPhi;
    // (a) the library: every cell, in reverse order (chains are independent)
    for (int c = N - 1; c >= 0; --c) flip_chain_cell(S, c, 1);
    const std::vector<double> Pa = Phi, Ra = rkPhi;
    // (b) the reference's loop, track by track
    Phi = Phi0;
    rkPhi = rk0;
    for (int a = 0; a < na; ++a) {
        const AnnDev& A = hann[a];
        if (A.outer >= 0) reference_loop<DIR_R>(S, FaceId{A.off, 0}, A.np + hann[A.outer].np);
        if (A.up >= 0) reference_loop<DIR_Z>(S, FaceId{A.off, 0}, A.np + hann[A.up].np);
    }
    int mism = 0;
    for (size_t k = 0; k < Phi.size(); ++k)
        if (Phi[k] != Pa[k] || rkPhi[k] != Ra[k]) ++mism;
    std::printf("flags %d mismatches %d err %d\n", nflags, mism, errw);
    return 0;
}
