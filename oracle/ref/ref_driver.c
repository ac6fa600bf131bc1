/* Step-count driver for the reference CPU build (test infrastructure).
 *
 * Replaces the reference's main.c:27-89, whose loop only stops at T_End or on
 * an `abort` file.  It runs the same set-up sequence (read_par_file,
 * mpiSetup, setupGrid, setupDomain, setupCells, then the ghost fill of
 * main.c:52-61), then calls the reference's own getmindt() + timestep() a
 * fixed number of times, and writes full-precision binary dumps of every
 * annulus (ghosts included) so the build's fixtures can be taken at exact
 * step counts.  Version 3 dumps end with the time-weighted diagnostics the
 * reference's timestep() accumulates (analysis.c:52-107): t_avg and
 * Qrz[Nr*Nz*num_tools].
 *
 * usage: disco_ref NSTEPS OUTPREFIX [DUMPSTEP ...]
 *   reads ./in.par; writes OUTPREFIX_<step>.bin at step 0, at every DUMPSTEP
 *   and at NSTEPS; prints one line "ref steps=N cells=C seconds=S" with the
 *   wall time of the step loop alone; after the loop the reference's own
 *   report() (report.c) appends its line to ./report.dat.
 *   NSTEPS = 0: the reference's own loop condition instead (main.c:68-75
 *   without output): getmindt, check_dt (domain.c:283-310, the T_End clip),
 *   timestep until check_dt marks the final step; the last dump is
 *   OUTPREFIX_<steps taken>.bin.
 *   With the environment variable DISCO_REF_TOTALS set, every step (and
 *   step 0) also prints "tot STEP DT Q0 .. Q(NUM_Q-1)": the conserved totals
 *   over the real cells (jk order, then array order) with 17 digits.
 *   With DISCO_REF_SINKS set, every step also prints, per planet p,
 *   "sink STEP p gravL dM accL Ls kin therm linXmom linYmom": the accretion
 *   sinks' accumulators (sink.c:143-270) as report.c would print them.
 */
#include "paul.h"
#include <stdint.h>
#include <time.h>

int read_par_file(struct domain*);
int mpiSetup(struct domain*, int, char*[]);
void setupGrid(struct domain*);
void setupDomain(struct domain*);
void setupCells(struct domain*);
void boundary_trans(struct domain*, int);
void exchangeData(struct domain*, int);
double getmindt(struct domain*);
void check_dt(struct domain*, double*);
void timestep(struct domain*, double);
void freeDomain(struct domain*);
void report(struct domain*);

static void dump(struct domain* d, const char* prefix, int step, double dt_last) {
    char name[1024];
    snprintf(name, sizeof name, "%s_%d.bin", prefix, step);
    FILE* f = fopen(name, "wb");
    if (!f) { perror(name); exit(1); }
    int32_t hdr[12] = {0x4F435344, 3, d->Nr, d->Nz, NUM_Q, NUM_FACES, d->Npl,
                       d->NgRa, d->NgRb, d->NgZa, d->NgZb, step};
    fwrite(hdr, sizeof hdr, 1, f);
    double tt[2] = {d->t, dt_last};
    fwrite(tt, sizeof tt, 1, f);
    fwrite(d->r_jph - 1, sizeof(double), d->Nr + 1, f);
    fwrite(d->z_kph - 1, sizeof(double), d->Nz + 1, f);
    int jk, i;
    for (jk = 0; jk < d->Nr * d->Nz; ++jk) {
        int32_t np = d->Np[jk];
        fwrite(&np, sizeof np, 1, f);
    }
    for (jk = 0; jk < d->Nr * d->Nz; ++jk) {
        for (i = 0; i < d->Np[jk]; ++i) {
            struct cell* c = &d->theCells[jk][i];
            double rec[4];
            rec[0] = c->piph;
            rec[1] = c->dphi;
            rec[2] = c->wiph;
            rec[3] = (double)c->real;
            fwrite(rec, sizeof rec, 1, f);
            fwrite(c->prim, sizeof(double), NUM_Q, f);
            fwrite(c->cons, sizeof(double), NUM_Q, f);
#if NUM_FACES > 0
            fwrite(c->Phi, sizeof(double), NUM_FACES, f);
#endif
        }
    }
    int p;
    for (p = 0; p < d->Npl; ++p) {
        struct planet* pl = d->thePlanets + p;
        double rec[5] = {pl->r, pl->phi, pl->M, pl->omega, pl->eps};
        fwrite(rec, sizeof rec, 1, f);
    }
    int32_t nt = d->num_tools;
    fwrite(&nt, sizeof nt, 1, f);
    fwrite(&d->theTools.t_avg, sizeof(double), 1, f);
    fwrite(d->theTools.Qrz, sizeof(double), (size_t)d->Nr * d->Nz * nt, f);
    fclose(f);
}

static void print_totals(struct domain* d, int step, double dt) {
    double tot[NUM_Q] = {0};
    int j, k, i, q;
    for (k = d->NgZa; k < d->Nz - d->NgZb; ++k)
        for (j = d->NgRa; j < d->Nr - d->NgRb; ++j) {
            int jk = j + d->Nr * k;
            for (i = 0; i < d->Np[jk]; ++i)
                for (q = 0; q < NUM_Q; ++q) tot[q] += d->theCells[jk][i].cons[q];
        }
    printf("tot %d %.17e", step, dt);
    for (q = 0; q < NUM_Q; ++q) printf(" %.17e", tot[q]);
    printf("\n");
    fflush(stdout);
}

static void print_sinks(struct domain* d, int step) {
    int p;
    for (p = 0; p < d->Npl; ++p) {
        const struct planet* pl = d->thePlanets + p;
        printf("sink %d %d %.17e %.17e %.17e %.17e %.17e %.17e %.17e %.17e\n", step, p, pl->gravL, pl->dM, pl->accL,
               pl->Ls, pl->kin, pl->therm, pl->linXmom, pl->linYmom);
    }
    fflush(stdout);
}

static double now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char* argv[]) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s NSTEPS OUTPREFIX [DUMPSTEP ...]\n", argv[0]);
        return 2;
    }
    int nsteps = atoi(argv[1]);
    const char* prefix = argv[2];
    int ndump = argc - 3;

    struct domain theDomain = {0};
    if (read_par_file(&theDomain)) return 1;
    if (mpiSetup(&theDomain, 1, argv) == 1) return 0;
    setupGrid(&theDomain);
    setupDomain(&theDomain);
    setupCells(&theDomain);
    if (theDomain.Nr > 1) {
        exchangeData(&theDomain, 0);
        if (!theDomain.theParList.R_Periodic) boundary_trans(&theDomain, 1);
    }
    if (theDomain.Nz > 1) {
        exchangeData(&theDomain, 1);
        if (!theDomain.theParList.Z_Periodic) boundary_trans(&theDomain, 2);
    }
    if (strcmp(prefix, "-") != 0) dump(&theDomain, prefix, 0, 0.0);
    const int want_totals = getenv("DISCO_REF_TOTALS") != NULL;
    const int want_sinks = getenv("DISCO_REF_SINKS") != NULL;
    if (want_totals) print_totals(&theDomain, 0, 0.0);

    long ncells = 0;
    int j, k;
    for (k = theDomain.NgZa; k < theDomain.Nz - theDomain.NgZb; ++k)
        for (j = theDomain.NgRa; j < theDomain.Nr - theDomain.NgRb; ++j)
            ncells += theDomain.Np[j + theDomain.Nr * k];

    double t0 = now();
    double seconds = 0.0;
    double dt = 0.0;
    int n;
    const int to_end = (nsteps == 0);
    for (n = 1; to_end ? !theDomain.final_step : n <= nsteps; ++n) {
        dt = getmindt(&theDomain);
        if (to_end) check_dt(&theDomain, &dt);
        timestep(&theDomain, dt);
        if (want_totals) {
            seconds += now() - t0;
            print_totals(&theDomain, n, dt);
            t0 = now();
        }
        if (want_sinks) print_sinks(&theDomain, n);
        int want = to_end ? theDomain.final_step : (n == nsteps);
        int a;
        for (a = 0; a < ndump; ++a)
            if (atoi(argv[3 + a]) == n) want = 1;
        if (want && strcmp(prefix, "-") != 0) {
            seconds += now() - t0;
            dump(&theDomain, prefix, n, dt);
            t0 = now();
        }
    }
    seconds += now() - t0;
    if (to_end) nsteps = n - 1;
    printf("ref steps=%d cells=%ld seconds=%.6f\n", nsteps, ncells, seconds);
    report(&theDomain);
    freeDomain(&theDomain);
    return 0;
}
