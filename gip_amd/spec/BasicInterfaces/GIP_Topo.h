// Partition layout of one rank (interface of GIP_Topo.h:19-92 in the
// reference): index ranges ALL/OWNED/INNER/... and the halo update.
//
//   [0, interiors)            interior cells
//   [interiors, inner)        interface cells (read by neighbours)
//   [inner, owned)            boundary ghosts
//   [owned, all)              halo (copies of neighbours' interface cells)
//
// This backend runs one partition on one device: the halo is empty and
// update() has nothing to exchange.  A topology with a halo is rejected at
// postConstruct.
#ifndef _gip_topo_
#define _gip_topo_

#include <string>

enum RUN_DOMAIN {
    _ALL_,
    _OWNED_,
    _HALO_,
    _INNER_,
    _INTER_,
    _IFACE_,
    _BOUND_
};

class GIP_Topo {
   public:
    GIP_Topo() {}
    ~GIP_Topo() {}
    // Reads the ASCII .topo file (format of GIP_Topo.cpp:12-49).
    void postConstruct(std::string name);
    // Sets the sizes of a one-partition layout directly (in-memory cases).
    void postConstructSizes(int inner, int ghosts);
    void update(double *) {}

    int getAllSize() { return allsize; }
    int getOwnedSize() { return ownedsize; }
    int getInnerSize() { return innersize; }
    int getInterSize() { return interiors; }
    int getIfaceSize() { return interfaces; }
    int getHaloSize() { return halos; }
    int getBoundarySize() { return boundaries; }
    int getSendSize() { return 0; }
    int getRecvSize() { return 0; }
    // number of elements a RUN_DOMAIN covers from index 0 (0 for the others,
    // as the reference's switch statements leave them untouched)
    int rangeSize(RUN_DOMAIN d);
    void printTopoInfo();

   protected:
    int allsize = 0, ownedsize = 0, innersize = 0;
    int interiors = 0, interfaces = 0, boundaries = 0, halos = 0;
};

#endif
