// Test probe (built by tests/test_grmhd_spin_cpu.py with the host compiler):
// the library's GRMHD leaves -- the same phys_grmhd.h / phys_grmhd_fv.h the
// GPU kernels compile -- at any spin, on records read from stdin, one record
// per line:
//   M a gamma riemann  prim1[8] prim2[8]  R phi z  dir
// and per record one output line of 62 numbers:
//   alpha, beta[3], gamma_ij[9], gamma^ij[9]                  (22)
//   prim2cons(prim1, dV = 1)                                  (8)
//   flux(prim1) in direction dir (0 r, 1 phi, 2 z)            (8)
//   vel(prim1, prim2) in direction dir: Sl, Sr, Ss            (3)
//   riemann(prim1, prim2, w = 0, dAdt = 1): dU[8], E1, B1, E2, B2   (12)
//   source(prim1) over the cell x +- (0.05, 0.02, 0.03), dVdt = 0.7 (8)
//   mindt(prim1, wc = 0.1) over the same cell                 (1)
// With riemann >= 20 the record is grmhd2.c's (prim[UPP] = u_phi; riemann -
// 20 the solver), in the same 62 columns.
// With riemann >= 10 the record is greuler.c's (GR hydrodynamics; riemann
// - 10 the solver): prim[0..4] with prim[UPP] = u_phi, and the same 62
// columns with zeros where greuler has no value (cons[5..7], F[5..7],
// dU[5..7], the EMFs); the source's columns hold cons2prim(prim2cons(prim1))
// instead (the round trip, started from prim2).
#include <cstdio>
#include <iostream>
#include <sstream>
#include <string>

#include "step_cells.h"

using namespace disco;

int main() {
    std::string line;
    while (std::getline(std::cin, line)) {
        std::istringstream is(line);
        double M, a, gam_law;
        int rs;
        double p1[8], p2[8], x[3];
        int dir;
        if (!(is >> M >> a >> gam_law >> rs)) continue;
        for (double& v : p1) is >> v;
        for (double& v : p2) is >> v;
        is >> x[0] >> x[1] >> x[2] >> dir;
        PhysParams P{};
        P.gamma_law = gam_law;
        P.gr_M = M;
        P.gr_a = a;
        P.riemann = rs % 10;
        const bool l2 = rs >= 20;
        P.phi_max = 2.0 * M_PI;
        grmhd::Metric g;
        grmhd::kerr_metric(x, M, a, &g);
        if (rs >= 10 && !l2) {
            std::printf("%.17g", g.al);
            for (double v : g.be) std::printf(" %.17g", v);
            for (double v : g.gam) std::printf(" %.17g", v);
            for (double v : g.igam) std::printf(" %.17g", v);
            double cons[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            greuler::prim2cons(gam_law, g, p1, cons, 1.0);
            for (double v : cons) std::printf(" %.17g", v);
            const double n[3] = {dir == 0 ? 1.0 : 0.0, dir == 1 ? 1.0 : 0.0, dir == 2 ? 1.0 : 0.0};
            double F[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            greuler::flux(gam_law, g, p1, F, x, n);
            for (double v : F) std::printf(" %.17g", v);
            double Sl, Sr, Ss;
            greuler::vel(gam_law, g, p1, p2, &Sl, &Sr, &Ss, n, x);
            std::printf(" %.17g %.17g %.17g", Sl, Sr, Ss);
            double dU[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            greuler::riemann(P, p1, p2, x, n, 0.0, 1.0, dU);
            for (double v : dU) std::printf(" %.17g", v);
            std::printf(" 0 0 0 0");
            double back[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int q = 0; q < 5; ++q) back[q] = p2[q];
            greuler::cons2prim(P, cons, back, x, 1.0);
            for (double v : back) std::printf(" %.17g", v);
            std::printf(" 0\n");
            continue;
        }
        std::printf("%.17g", g.al);
        for (double v : g.be) std::printf(" %.17g", v);
        for (double v : g.gam) std::printf(" %.17g", v);
        for (double v : g.igam) std::printf(" %.17g", v);
        double cons[8];
        if (l2)
            grmhd::prim2cons<true>(gam_law, g, p1, cons, x, 1.0);
        else
            grmhd::prim2cons(gam_law, g, p1, cons, x, 1.0);
        for (double v : cons) std::printf(" %.17g", v);
        const double n[3] = {dir == 0 ? 1.0 : 0.0, dir == 1 ? 1.0 : 0.0, dir == 2 ? 1.0 : 0.0};
        double F[8];
        if (l2)
            grmhd::flux<true>(gam_law, g, p1, F, x, n);
        else
            grmhd::flux(gam_law, g, p1, F, x, n);
        for (double v : F) std::printf(" %.17g", v);
        double Sl, Sr, Ss;
        if (l2)
            grmhd::vel<true>(gam_law, g, p1, p2, &Sl, &Sr, &Ss, n, x);
        else
            grmhd::vel(gam_law, g, p1, p2, &Sl, &Sr, &Ss, n, x);
        std::printf(" %.17g %.17g %.17g", Sl, Sr, Ss);
        double dU[8];
        const int dim = dir == 0 ? 1 : (dir == 1 ? 0 : 2);  // the reference's dim: 0 phi, 1 r, 2 z
        FaceEMF e = l2 ? PhysGRMHD2::riemann(P, p1, p2, x, n, 0.0, 1.0, dim, dU)
                       : PhysGRMHD::riemann(P, p1, p2, x, n, 0.0, 1.0, dim, dU);
        for (double v : dU) std::printf(" %.17g", v);
        std::printf(" %.17g %.17g %.17g %.17g", e.E1, e.B1, e.E2, e.B2);
        double xp[3] = {x[0] + 0.05, x[1] + 0.02, x[2] + 0.03};
        double xm[3] = {x[0] - 0.05, x[1] - 0.02, x[2] - 0.03};
        double src[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (l2)
            PhysGRMHD2::source(P, p1, src, xp, xm, 0.7);
        else
            PhysGRMHD::source(P, p1, src, xp, xm, 0.7);
        for (double v : src) std::printf(" %.17g", v);
        std::printf(" %.17g\n", l2 ? PhysGRMHD2::mindt(P, p1, 0.1, xp, xm) : PhysGRMHD::mindt(P, p1, 0.1, xp, xm));
    }
    return 0;
}
