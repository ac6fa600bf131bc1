/* Link stub for the reference's MHD build (test infrastructure only).
 *
 * At this snapshot Hydro/mhd.c defines no visc_source, although misc.c:586
 * calls it under `if(visc_flag)`; without a definition the mhd build does not
 * link.  The benchmark configurations run with Use_Viscosity 0, so the call
 * is never reached and this empty body changes no result.  Aborts if it is
 * ever called, so a configuration that needs viscosity cannot pass silently.
 */
#include <stdio.h>
#include <stdlib.h>

void visc_source(const double* prim, const double* gradr, const double* gradp, const double* gradz,
                 double* cons, const double* xp, const double* xm, double dVdt) {
    (void)prim; (void)gradr; (void)gradp; (void)gradz; (void)cons; (void)xp; (void)xm; (void)dVdt;
    fprintf(stderr, "visc_source: viscosity is not part of the MHD reference build\n");
    abort();
}
