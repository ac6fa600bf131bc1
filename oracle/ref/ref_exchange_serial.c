/* Serial stand-in for the reference's exchangeData (exchange.c:173-279),
 * linked into the reference harness build in place of exchange.c.
 *
 * Test infrastructure only.  With USE_MPI=0 the reference's exchangeData is
 * empty, so ghost layers of a periodic direction are never filled.  This file
 * does what a one-rank MPI run does through its self-wraparound Cartesian
 * communicator (mpisetup.c:17,38): for a periodic direction, ghost layer
 * n < Ng receives real layer NN-2Ng+n and ghost layer NN-Ng+n receives real
 * layer Ng+n, for every annulus of the other direction.  The copied fields are
 * exactly the cell_lite set (exchange.c:7-16): prim, cons, RKcons, piph, dphi,
 * wiph, Phi, RK_Phi.  Annuli keep their own Np, which is equal along z and
 * therefore along the copy.  Non-periodic directions are left alone, as in
 * the reference's serial build.
 */
#include <string.h>

#include "paul.h"

static void ref_copy_cell(const struct cell* s, struct cell* d) {
    memcpy(d->prim, s->prim, sizeof(d->prim));
    memcpy(d->cons, s->cons, sizeof(d->cons));
    memcpy(d->RKcons, s->RKcons, sizeof(d->RKcons));
    memcpy(d->Phi, s->Phi, sizeof(d->Phi));
    memcpy(d->RK_Phi, s->RK_Phi, sizeof(d->RK_Phi));
    d->piph = s->piph;
    d->dphi = s->dphi;
    d->wiph = s->wiph;
}

static void ref_copy_annulus(struct domain* dom, int jk_src, int jk_dst) {
    int n = dom->Np[jk_src] < dom->Np[jk_dst] ? dom->Np[jk_src] : dom->Np[jk_dst];
    int i;
    for (i = 0; i < n; ++i) ref_copy_cell(&dom->theCells[jk_src][i], &dom->theCells[jk_dst][i]);
}

void exchangeData(struct domain* theDomain, int dim) {
    int Nr = theDomain->Nr, Nz = theDomain->Nz, Ng = theDomain->Ng;
    int j, k, n;
    if (dim == 0) {
        if (!theDomain->theParList.R_Periodic || Nr < 2 * Ng) return;
        for (k = 0; k < Nz; ++k)
            for (n = 0; n < Ng; ++n) {
                ref_copy_annulus(theDomain, (Nr - 2 * Ng + n) + Nr * k, n + Nr * k);
                ref_copy_annulus(theDomain, (Ng + n) + Nr * k, (Nr - Ng + n) + Nr * k);
            }
    } else {
        if (!theDomain->theParList.Z_Periodic || Nz < 2 * Ng) return;
        for (j = 0; j < Nr; ++j)
            for (n = 0; n < Ng; ++n) {
                ref_copy_annulus(theDomain, j + Nr * (Nz - 2 * Ng + n), j + Nr * n);
                ref_copy_annulus(theDomain, j + Nr * (Ng + n), j + Nr * (Nz - Ng + n));
            }
    }
}
