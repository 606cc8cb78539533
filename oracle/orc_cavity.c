/*
 * ORACLE — TEST INFRASTRUCTURE ONLY.
 * Command-line driver of the C restatement, the counterpart of
 * DrivenCavity.cpp:21-43 on one partition:
 *   orc_cavity <root> <name> <steps> [gamma] [pcg_maxit] [pcg_tol] [--full] [--dump DIR]
 * Prints one line per step in the reference's format (GIP_DNS.h:403-404);
 * --full prints 17 significant digits and the PCG iteration count.
 * --dump writes u, v, w, p (all cells) as .vec files into DIR.
 */
#define _POSIX_C_SOURCE 199309L
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "gip_oracle.h"

static void dump(const char *dir, const char *nm, const double *v, int n) {
    char path[4096];
    snprintf(path, sizeof path, "%s/%s.vec", dir, nm);
    FILE *fp = fopen(path, "wb");
    if (!fp) { fprintf(stderr, "cannot write %s\n", path); exit(2); }
    fwrite(&n, sizeof(int), 1, fp);
    fwrite(v, sizeof(double), (size_t)n, fp);
    fclose(fp);
}

int main(int argc, char **argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s <root> <name> <steps> [gamma] [pcg_maxit] [pcg_tol] [--full] [--dump DIR]\n", argv[0]);
        return 1;
    }
    const char *root = argv[1], *name = argv[2], *dumpdir = NULL;
    int steps = atoi(argv[3]), full = 0, maxit = 100, pos = 0;
    double gamma = 1e-3, tol = 0.0;
    for (int a = 4; a < argc; ++a) {
        if (!strcmp(argv[a], "--full")) full = 1;
        else if (!strcmp(argv[a], "--dump") && a + 1 < argc) dumpdir = argv[++a];
        else if (pos == 0) { gamma = atof(argv[a]); pos++; }
        else if (pos == 1) { maxit = atoi(argv[a]); pos++; }
        else if (pos == 2) { tol = atof(argv[a]); pos++; }
    }
    orc_dns *s;
    if (orc_dns_open(root, name, &s)) {
        fprintf(stderr, "cannot open case %s under %s\n", name, root);
        return 2;
    }
    orc_dns_params(s, gamma, 1.0, maxit, tol, 1);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int k = 0; k < steps; ++k) {
        double t, dt, mr;
        int it;
        int st = orc_dns_step(s, &t, &dt, &mr, &it);
        if (full)
            printf("time %.17e [dt::%.17e] massr %.17e ite %d pcg %d status %d\n", t, dt, mr, k + 1, it, st);
        else
            printf("time %e [dt::%e] massr %e ite %d\n", t, dt, mr, k + 1);
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    fprintf(stderr, "integrate_ms %.3f\n", (t1.tv_sec - t0.tv_sec) * 1e3 + (t1.tv_nsec - t0.tv_nsec) * 1e-6);
    if (dumpdir) {
        dump(dumpdir, "u", s->u, s->all);
        dump(dumpdir, "v", s->v, s->all);
        dump(dumpdir, "w", s->w, s->all);
        dump(dumpdir, "p", s->p, s->all);
    }
    orc_dns_close(s);
    return 0;
}
