// Command-line front end of the operator generator (include/gip_cavity.h).
//   gip_cavity_tool gen   <root> <name> <nx> [ny nz] [--beta B] [--gamma G] [--lid U] [--nopin]
//   gip_cavity_tool merge <src_root> <name> <nparts> <dst_root> <dst_name>
// Both write a one-partition case in the reference's InputFiles layout.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "gip_cavity.h"

static int usage() {
    std::fprintf(stderr,
                 "usage: gip_cavity_tool gen <root> <name> <nx> [ny nz] [--beta B] [--gamma G] [--lid U] [--nopin]\n"
                 "       gip_cavity_tool merge <src_root> <name> <nparts> <dst_root> <dst_name>\n");
    return 1;
}

int main(int argc, char **argv) {
    if (argc < 2) return usage();
    gip_cavity *c = nullptr;
    int rc;
    if (!std::strcmp(argv[1], "gen")) {
        if (argc < 5) return usage();
        gip_cavity_desc d{};
        d.nx = d.ny = d.nz = std::atoi(argv[4]);
        d.gamma = 1e-2;
        d.lid_u = 1.0;
        d.pin = 1;
        d.ops = GIP_OP_ALL;
        int a = 5;
        if (argc >= 7 && argv[5][0] != '-') {
            d.ny = std::atoi(argv[5]);
            d.nz = std::atoi(argv[6]);
            a = 7;
        }
        for (; a < argc; ++a) {
            if (!std::strcmp(argv[a], "--beta") && a + 1 < argc) d.beta = std::atof(argv[++a]);
            else if (!std::strcmp(argv[a], "--gamma") && a + 1 < argc) d.gamma = std::atof(argv[++a]);
            else if (!std::strcmp(argv[a], "--lid") && a + 1 < argc) d.lid_u = std::atof(argv[++a]);
            else if (!std::strcmp(argv[a], "--nopin")) d.pin = 0;
            else return usage();
        }
        rc = gip_cavity_generate(&d, &c);
        if (!rc) rc = gip_cavity_write(c, argv[2], argv[3]);
    } else if (!std::strcmp(argv[1], "merge")) {
        if (argc != 7) return usage();
        rc = gip_cavity_merge(argv[2], argv[3], std::atoi(argv[4]), &c);
        if (!rc) rc = gip_cavity_write(c, argv[5], argv[6]);
    } else {
        return usage();
    }
    if (rc) {
        std::fprintf(stderr, "gip_cavity_tool: %s\n", gip_cavity_last_error());
        gip_cavity_free(c);
        return 2;
    }
    std::printf("cells %d ghosts %d faces %d\n", c->n_inner, c->n_ghost, c->n_faces);
    gip_cavity_free(c);
    return 0;
}
