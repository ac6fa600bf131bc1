/* The reference's main loop (main.c:27-89) as a C host of the gfx950
 * library, through the C ABI of include/disco_gpu.h only.
 *
 *   disco_main [PARFILE] [STEPS_PER_CHECKPOINT]
 *
 * read_par_file -> set-up (disco_gpu_create) -> restart from input.ckp when
 * the parameter file says "Restart 1" (domain.c:141-142) -> the loop of
 * getmindt, check_dt and timestep until T_End (disco_gpu_advance, which also
 * polls the "abort" and "latest" files like check_dt) -> a checkpoint every
 * STEPS_PER_CHECKPOINT steps (checkpoint_NNNN.ckp), on "latest"
 * (checkpoint_latest.ckp) and at the end (checkpoint_final.ckp), and the
 * report line (report.c) to stdout after each checkpoint.
 * Errors print the library's message and exit(1), as the reference does.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "disco_gpu.h"

static void die(const char* what) {
    printf("%s: %s\n", what, disco_gpu_last_error());
    exit(1);
}

/* readpar.c:6-42 for one integer key, enough for "Restart" */
static int par_int(const char* path, const char* key, int dflt) {
    FILE* f = fopen(path, "r");
    char line[512], name[256];
    int v = dflt;
    if (!f) return dflt;
    while (fgets(line, sizeof line, f)) {
        int x;
        if (sscanf(line, "%255s %d", name, &x) == 2 && strcmp(name, key) == 0) {
            v = x;
            break;
        }
    }
    fclose(f);
    return v;
}

static void checkpoint(disco_gpu_t* h, const char* name) {
    double line[64];
    int n = 0, i;
    if (disco_gpu_write_checkpoint(h, 0, name)) die("checkpoint");
    if (disco_gpu_report(h, 0, line, 64, &n)) die("report");
    printf("%s:", name);
    for (i = 0; i < n; ++i) printf(" %.10le", line[i]);
    printf("\n");
    fflush(stdout);
}

int main(int argc, char** argv) {
    const char* par = argc > 1 ? argv[1] : "in.par";
    const int every = argc > 2 ? atoi(argv[2]) : 100;
    disco_params p;
    char err[512];
    disco_gpu_t* h = NULL;
    int steps = 0, total = 0, status = 0, nck = 0;
    if (disco_read_par(par, &p, err, sizeof err)) {
        printf("%s\n", err);
        return 1;
    }
    if (disco_gpu_create(&p, 1, &h)) die("set-up");
    if (par_int(par, "Restart", 0) && disco_gpu_restart(h, 0, "input.ckp")) die("restart");
    while (!(status & 1)) {
        if (disco_gpu_advance(h, every > 0 ? every : 100, 1, &steps, &status)) die("advance");
        total += steps;
        if (status & 4) {
            checkpoint(h, "checkpoint_latest.ckp");
        } else if (!(status & 1)) {
            char name[64];
            snprintf(name, sizeof name, "checkpoint_%04d.ckp", nck++);
            checkpoint(h, name);
        }
        if (status & 2) break; /* abort */
    }
    checkpoint(h, "checkpoint_final.ckp");
    printf("steps %d\n", total);
    disco_gpu_destroy(h);
    return 0;
}
