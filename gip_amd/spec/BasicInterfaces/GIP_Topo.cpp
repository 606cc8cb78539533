#include "GIP_Topo.h"

#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <stdexcept>

void GIP_Topo::postConstruct(std::string filename) {
    FILE *fp = std::fopen(filename.c_str(), "r");
    if (!fp) throw std::runtime_error("GIP_Topo: cannot open topology " + filename);
    int s[9];
    for (int i = 0; i < 9; ++i)
        if (std::fscanf(fp, " %d", &s[i]) != 1) {
            std::fclose(fp);
            throw std::runtime_error("GIP_Topo: malformed topology " + filename);
        }
    std::fclose(fp);
    ownedsize = s[3];
    allsize = s[4];
    interiors = s[5];
    interfaces = s[6];
    boundaries = s[7];
    halos = s[8];
    innersize = interiors + interfaces;
    if (s[0] != 0 || s[1] != 0 || allsize != ownedsize)
        throw std::runtime_error("GIP_Topo: " + filename +
                                 " describes a partition with a halo; this backend runs one partition per device");
}

void GIP_Topo::postConstructSizes(int inner, int ghosts) {
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
