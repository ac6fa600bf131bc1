// Test probe (built by tests/test_boundary_cpu.py with the host compiler):
// the per-cell zero-gradient ghost fill the GPU kernels compile
// (step_cells.h:bc_copy_cell, one thread per ghost cell, its start found by
// a search) against the host's copy_distant (host_setup.cpp), which runs
// set_cells_copy_distant's sequential loop in the reference's order and is
// pinned bit for bit to the reference's own step-0 ghost states.  Meshes:
// random fronts of two annuli of random sizes, identical fronts (two
// z-neighbours of one column, which rotate together), and identical fronts
// moved by a few ulps.  Prints one line per case: "case n n1 mismatches".
//
// usage: bc_probe SEED NCASES
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "host_setup.h"
#include "step_cells.h"

using namespace disco;

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    std::mt19937_64 rng(std::strtoull(argv[1], nullptr, 10));
    const int ncases = std::atoi(argv[2]);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const double Pmax = 2.0 * M_PI;
    int bad_total = 0;
    for (int cs = 0; cs < ncases; ++cs) {
        const int kind = cs % 3;  // 0 random, 1 identical fronts, 2 identical fronts moved by ulps
        const int n1 = 4 + (int)(U(rng) * 60);
        const int n = kind ? n1 : 4 + (int)(U(rng) * 60);
        // fronts of an annulus: cumulative random widths from a random phase, wrapped, in array order
        auto ring = [&](int m, std::vector<double>& piph, std::vector<double>& dphi) {
            std::vector<double> w(m);
            double s = 0.0;
            for (int i = 0; i < m; ++i) s += (w[i] = 0.5 + U(rng));
            double p = U(rng) * Pmax;
            piph.resize(m);
            dphi.resize(m);
            for (int i = 0; i < m; ++i) {
                p += w[i] * Pmax / s;
                while (p > Pmax) p -= Pmax;
                piph[i] = p;
            }
            for (int i = 0; i < m; ++i) dphi[i] = get_dp(piph[i], piph[i == 0 ? m - 1 : i - 1], Pmax);
        };
        HostRun R;
        R.nq = 5;
        R.ph.phi_max = Pmax;
        R.Np = {n, n1};
        R.off = {0, n, n + n1};
        R.ncell = n + n1;
        std::vector<double> pg, dg, ps, ds;
        ring(n1, ps, ds);
        if (kind == 0) {
            ring(n, pg, dg);
        } else {
            pg = ps;
            if (kind == 2)
                for (int i = 0; i < n; ++i) {
                    double f = pg[i];
                    int k = (int)(U(rng) * 5) - 2;
                    for (; k > 0; --k) f = std::nextafter(f, 10.0);
                    for (; k < 0; ++k) f = std::nextafter(f, -1.0);
                    pg[i] = f;
                }
            dg.resize(n);
            for (int i = 0; i < n; ++i) dg[i] = get_dp(pg[i], pg[i == 0 ? n - 1 : i - 1], Pmax);
        }
        R.piph = pg;
        R.piph.insert(R.piph.end(), ps.begin(), ps.end());
        R.dphi = dg;
        R.dphi.insert(R.dphi.end(), ds.begin(), ds.end());
        R.prim.assign((size_t)R.nq * R.ncell, 0.0);
        for (int q = 0; q < R.nq; ++q)
            for (int c = 0; c < R.ncell; ++c) R.prim[(size_t)q * R.ncell + c] = U(rng) * 2.0 - 0.5;
        // rotation points (array index of the smallest front)
        std::vector<int> rot(2, 0);
        for (int a = 0; a < 2; ++a)
            for (int i = 0; i < R.Np[a]; ++i) {
                int im = i == 0 ? R.Np[a] - 1 : i - 1;
                if (R.piph[R.off[a] + i] < R.piph[R.off[a] + im]) rot[a] = i;
            }
        std::vector<double> prim_dev = R.prim;  // the per-cell fill writes here
        copy_distant(R, 0, 1);                  // the reference's loop order
        AnnDev ann[2] = {};
        for (int a = 0; a < 2; ++a) {
            ann[a].np = R.Np[a];
            ann[a].off = R.off[a];
            ann[a].run = 0;
        }
        RunDev run{};
        run.ph = R.ph;
        DevState S{};
        S.ann = ann;
        S.runs = &run;
        S.ncell = R.ncell;
        S.nruns = 1;
        S.prim = prim_dev.data();
        S.piph_n = R.piph.data();
        S.dphi_n = R.dphi.data();
        S.rot_n = rot.data();
        for (int i = 0; i < n; ++i) bc_copy_cell<PhysEuler>(S, ann[0], i, 1, -1, 0);
        int bad = 0;
        for (int q = 0; q < R.nq; ++q)
            for (int i = 0; i < n; ++i)
                if (prim_dev[(size_t)q * R.ncell + i] != R.prim[(size_t)q * R.ncell + i]) ++bad;
        std::printf("%d %d %d %d\n", kind, n, n1, bad);
        bad_total += bad;
    }
    return bad_total ? 1 : 0;
}
