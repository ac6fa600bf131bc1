/* Leaf-function driver for the reference build (test infrastructure).
 *
 * Runs the reference's own set-up (read_par_file, mpiSetup, setupGrid,
 * setupDomain: every module's set*Params from ./in.par) and then evaluates,
 * for each input record, the Hydro module's leaves and the Riemann solver
 * exactly as the reference calls them:
 *   solve_riemann (riemann.c:216-320) from zero conserved increments, so
 *     consR = (F - w U*) dA dt and the flux_to_E values come back;
 *   prim2cons (dV = 1.37) and cons2prim of the result back (the GRMHD solve
 *     starts from primL x 1.001);
 *   mindt(primL, w, xp, xm) and source(primL, cons = 0, xp, xm, dVdt = 0.7)
 *     with xp/xm = x +- (0.05, 0.02, 0.03).
 *
 * usage: disco_ref_leaf_<variant> IN.bin OUT.bin
 *   IN:  int32 n, then n records of primL[NUM_Q], primR[NUM_Q], x[3], n[3],
 *        w, dAdt, dim (doubles)
 *   OUT: n records of dU[NUM_Q], E1, B1, E2, B2, cons[NUM_Q], prim[NUM_Q],
 *        dt, src[NUM_Q] (doubles)
 */
#include "paul.h"

int read_par_file(struct domain*);
int mpiSetup(struct domain*, int, char*[]);
void setupGrid(struct domain*);
void setupDomain(struct domain*);
void solve_riemann(const double*, const double*, double*, double*, const double*, const double*, const double*,
                   const double*, const double*, const double*, const double*, const double*, double, double, int,
                   double*, double*, double*, double*);
void prim2cons(const double*, double*, const double*, double);
void cons2prim(const double*, double*, const double*, double);
double mindt(const double*, double, const double*, const double*);
void source(const double*, double*, const double*, const double*, double);

int main(int argc, char* argv[]) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s IN.bin OUT.bin\n", argv[0]);
        return 2;
    }
    struct domain d;
    memset(&d, 0, sizeof d);
    if (read_par_file(&d)) return 1;
    if (mpiSetup(&d, 1, argv) == 1) return 0;
    setupGrid(&d);
    setupDomain(&d);
    FILE* fi = fopen(argv[1], "rb");
    FILE* fo = fopen(argv[2], "wb");
    if (!fi || !fo) return 1;
    int n, k, q;
    if (fread(&n, sizeof n, 1, fi) != 1) return 1;
    for (k = 0; k < n; ++k) {
        double in[2 * NUM_Q + 9];
        if (fread(in, sizeof in, 1, fi) != 1) return 1;
        const double* pL = in;
        const double* pR = in + NUM_Q;
        const double* x = in + 2 * NUM_Q;
        const double* nn = x + 3;
        double w = nn[3], dAdt = nn[4];
        int dim = (int)nn[5];
        double zero[NUM_Q], cL[NUM_Q], cR[NUM_Q];
        for (q = 0; q < NUM_Q; ++q) zero[q] = cL[q] = cR[q] = 0.0;
        double E[4] = {0.0, 0.0, 0.0, 0.0};
        solve_riemann(pL, pR, cL, cR, zero, zero, zero, zero, zero, zero, x, nn, w, dAdt, dim, E, E + 1, E + 2,
                      E + 3);
        double cons[NUM_Q], prim[NUM_Q];
        prim2cons(pL, cons, x, 1.37);
        for (q = 0; q < NUM_Q; ++q) prim[q] = pL[q] * 1.001;
        cons2prim(cons, prim, x, 1.37);
        double xp[3] = {x[0] + 0.05, x[1] + 0.02, x[2] + 0.03};
        double xm[3] = {x[0] - 0.05, x[1] - 0.02, x[2] - 0.03};
        double dt = mindt(pL, w, xp, xm);
        double src[NUM_Q];
        for (q = 0; q < NUM_Q; ++q) src[q] = 0.0;
        source(pL, src, xp, xm, 0.7);
        fwrite(cR, sizeof cR, 1, fo);
        fwrite(E, sizeof E, 1, fo);
        fwrite(cons, sizeof cons, 1, fo);
        fwrite(prim, sizeof prim, 1, fo);
        fwrite(&dt, sizeof dt, 1, fo);
        fwrite(src, sizeof src, 1, fo);
    }
    fclose(fi);
    fclose(fo);
    return 0;
}
