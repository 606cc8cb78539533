#include "GIP_Topo.h"

#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <stdexcept>

void GIP_Topo::postConstruct(std::string filename) {
    FILE *fp = std::fopen(filename.c_str(), "r");
    if (!fp) throw std::runtime_error("GIP_Topo: cannot open topology " + filename);
    int s[9];
    bool ok = true;
    for (int i = 0; i < 9 && ok; ++i) ok = std::fscanf(fp, " %d", &s[i]) == 1;
    auto rd = [&](std::vector<int> &v, int n) {
        if (!ok || n < 0) {
            ok = false;
            return;
        }
        v.assign((size_t)n, 0);
        for (int i = 0; i < n && ok; ++i) ok = std::fscanf(fp, " %d", &v[(size_t)i]) == 1;
    };
    if (ok) {
        rd(send_indx, s[0]);
        rd(send_row, s[2]);
        rd(recv_indx, s[1]);
        rd(recv_row, s[2]);
    }
    std::fclose(fp);
    if (!ok) throw std::runtime_error("GIP_Topo: malformed topology " + filename);
    ownedsize = s[3];
    allsize = s[4];
    interiors = s[5];
    interfaces = s[6];
    boundaries = s[7];
    halos = s[8];
    innersize = interiors + interfaces;
    // the lists must stay inside the ranges they name (host-side checks
    // before any kernel uses them)
    const int np = getPartitions();
    bool sane = np >= 1 && (int)recv_row.size() == np + 1 && send_row[0] == 0 && recv_row[0] == 0 &&
                send_row[(size_t)np] == s[0] && recv_row[(size_t)np] == s[1];
    for (int q = 0; q < np && sane; ++q)
        sane = send_row[(size_t)q] <= send_row[(size_t)q + 1] && recv_row[(size_t)q] <= recv_row[(size_t)q + 1];
    for (int i : send_indx) sane = sane && i >= 0 && i < innersize;
    for (int i : recv_indx) sane = sane && i >= ownedsize && i < allsize;
    if (!sane) throw std::runtime_error("GIP_Topo: inconsistent halo lists in " + filename);
}

void GIP_Topo::postConstructSizes(int inner, int ghosts) {
    send_indx.clear();
    recv_indx.clear();
    send_row.assign(2, 0);
    recv_row.assign(2, 0);
    interiors = inner;
    interfaces = 0;
    boundaries = ghosts;
    halos = 0;
    innersize = inner;
    ownedsize = allsize = inner + ghosts;
}

int GIP_Topo::rangeSize(RUN_DOMAIN d) {
    switch (d) {
        case _ALL_: return allsize;
        case _OWNED_: return ownedsize;
        case _INNER_: return innersize;
        default: return 0;
    }
}

void GIP_Topo::printTopoInfo() {
    std::cout << " Topo Info  " << std::endl
              << " inners     :" << innersize << std::endl
              << " interiors  :" << interiors << std::endl
              << " interfaces :" << interfaces << std::endl
              << " boundaries :" << boundaries << std::endl
              << " halos      :" << halos << std::endl
              << " Ownedsize  :" << ownedsize << std::endl
              << " Allsize    :" << allsize << std::endl;
}
