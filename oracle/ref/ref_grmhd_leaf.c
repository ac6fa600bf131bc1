/* Leaf-function driver for the reference's GRMHD module (test infrastructure).
 *
 * Links the reference's own Hydro/grmhd.c, Hydro/Metric/schw_ks_rpz.c and
 * Hydro/Frame/normal.c and evaluates, for each input record, the metric
 * (metric_lapse/shift/gam/igam/jacobian, schw_ks_rpz.c:20-80), prim2cons
 * (grmhd.c:65-155, dV = 1) and cons2prim (grmhd.c:228-245 -> noble2d,
 * :813-1029) started from a given guess primitive state.
 *
 * usage: disco_ref_grmhd_leaf GAMMA M IN.bin OUT.bin
 *   IN:  int32 n, then n records of x[3], prim[8], guess[8] (doubles)
 *   OUT: n records of cons[8], prim_out[8], metric[23] (doubles)
 */
#include "paul.h"

void setHydroParams(struct domain*);
void setMetricParams(struct domain*);
void prim2cons(const double*, double*, const double*, double);
void cons2prim(const double*, double*, const double*, double);
double metric_lapse(const double*);
void metric_shift(const double*, double*);
void metric_gam(const double*, double*);
void metric_igam(const double*, double*);
double metric_jacobian(const double*);

int main(int argc, char** argv) {
    if (argc < 5) return 2;
    struct domain d;
    memset(&d, 0, sizeof d);
    d.theParList.Adiabatic_Index = atof(argv[1]);
    d.theParList.metricPar2 = atof(argv[2]);
    setHydroParams(&d);
    setMetricParams(&d);
    FILE* fi = fopen(argv[3], "rb");
    FILE* fo = fopen(argv[4], "wb");
    if (!fi || !fo) return 1;
    int n, k;
    if (fread(&n, sizeof n, 1, fi) != 1) return 1;
    for (k = 0; k < n; ++k) {
        double in[19], cons[8], p[8], m[23];
        if (fread(in, sizeof in, 1, fi) != 1) return 1;
        double* x = in;
        prim2cons(in + 3, cons, x, 1.0);
        memcpy(p, in + 11, sizeof p);
        cons2prim(cons, p, x, 1.0);
        m[0] = metric_lapse(x);
        metric_shift(x, m + 1);
        metric_gam(x, m + 4);
        metric_igam(x, m + 13);
        m[22] = metric_jacobian(x);
        fwrite(cons, sizeof cons, 1, fo);
        fwrite(p, sizeof p, 1, fo);
        fwrite(m, sizeof m, 1, fo);
    }
    fclose(fi);
    fclose(fo);
    return 0;
}
