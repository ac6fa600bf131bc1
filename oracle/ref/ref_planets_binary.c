/* Planets-slot plugin for the reference build (test infrastructure).
 *
 * Equal-mass-style binary on a fixed circular orbit, modelled on the
 * reference's Planets/cmA.c:8-86 (analytic motion, phi += omega dt in
 * movePlanets), except that the orbital phase is seeded: planet 0 sits at
 * phi0 + pi and planet 1 at phi0, with phi0 = Init_Par6 (radians).
 * Separation a = 1, mass ratio q = Mass_Ratio, softening = Softening.
 */
#include "paul.h"

static double q_planet = 1.0;
static double eps = 0.0;
static double phase0 = 0.0;

void setPlanetParams(struct domain* theDomain) {
    theDomain->Npl = 2;
    q_planet = theDomain->theParList.Mass_Ratio;
    eps = theDomain->theParList.grav_eps;
    phase0 = theDomain->theParList.initPar6;
}

int planet_motion_analytic(void) { return 1; }

static void zero_planet(struct planet* p) { memset(p, 0, sizeof(*p)); }

void initializePlanets(struct planet* thePlanets) {
    double a = 1.0;
    double mu = q_planet / (1. + q_planet);
    double om = pow(a, -1.5);
    zero_planet(thePlanets + 0);
    zero_planet(thePlanets + 1);
    thePlanets[0].M = 1. - mu;
    thePlanets[0].omega = om;
    thePlanets[0].r = a * mu;
    thePlanets[0].phi = phase0 + M_PI;
    thePlanets[0].eps = eps;
    thePlanets[0].type = PLPOINTMASS;
    thePlanets[1].M = mu;
    thePlanets[1].omega = om;
    thePlanets[1].r = a * (1. - mu);
    thePlanets[1].phi = phase0;
    thePlanets[1].eps = eps;
    thePlanets[1].type = PLPOINTMASS;
}

void movePlanets(struct planet* thePlanets, double t, double dt) {
    (void)t;
    thePlanets[0].phi += thePlanets[0].omega * dt;
    thePlanets[1].phi += thePlanets[1].omega * dt;
}

void forcePlanets(struct planet* thePlanets, double dt) {
    (void)thePlanets;
    (void)dt;
}
