// Test probe (built by tests/test_grmhd_metric_cpu.py with the host
// compiler): evaluates the GRMHD module's metric and metric derivatives
// (disco_amd/csrc/phys_grmhd_fv.h, the same source the GPU kernels compile)
// at points read from stdin.
//
// stdin:  lines "M a R z mode"; mode 0 = metric_der (closed form at a = 0,
//         dual numbers otherwise), mode 1 = dual numbers at any a,
//         mode 2 = the metric itself
// stdout: mode 0/1: dg_r[16] dg_z[16] da[4] db[12]; mode 2: al be[3] gam[9]
#include <cstdio>

#include "phys_grmhd_fv.h"

using namespace disco::grmhd;

int main() {
    double M, a, R, z;
    int mode;
    while (std::scanf("%lf %lf %lf %lf %d", &M, &a, &R, &z, &mode) == 5) {
        double x[3] = {R, 0.0, z};
        if (mode == 2) {
            Metric g;
            kerr_metric(x, M, a, &g);
            std::printf("%.17g", g.al);
            for (int i = 0; i < 3; ++i) std::printf(" %.17g", g.be[i]);
            for (int i = 0; i < 9; ++i) std::printf(" %.17g", g.gam[i]);
            std::printf("\n");
            continue;
        }
        MetricDer D;
        if (mode == 0) {
            metric_der(x, M, a, &D);
        } else {
            for (int k = 0; k < 16; ++k) D.dg_r[k] = D.dg_z[k] = 0.0;
            for (int k = 0; k < 4; ++k) D.da[k] = 0.0;
            for (int k = 0; k < 12; ++k) D.db[k] = 0.0;
            for (int dir = 0; dir < 2; ++dir) {
                Dual Rd{R, dir == 0 ? 1.0 : 0.0};
                Dual Zd{z, dir == 1 ? 1.0 : 0.0};
                Dual al, be[3], g4[16];
                kerr_parts_dual(Rd, Zd, M, a, &al, be, g4);
                int mu = dir == 0 ? 1 : 3;
                double* dg = dir == 0 ? D.dg_r : D.dg_z;
                for (int k = 0; k < 16; ++k) dg[k] = g4[k].d;
                D.da[mu] = al.d;
                for (int j = 0; j < 3; ++j) D.db[3 * mu + j] = be[j].d;
            }
        }
        for (int k = 0; k < 16; ++k) std::printf("%.17g ", D.dg_r[k]);
        for (int k = 0; k < 16; ++k) std::printf("%.17g ", D.dg_z[k]);
        for (int k = 0; k < 4; ++k) std::printf("%.17g ", D.da[k]);
        for (int k = 0; k < 12; ++k) std::printf("%.17g%c", D.db[k], k == 11 ? '\n' : ' ');
    }
    return 0;
}
