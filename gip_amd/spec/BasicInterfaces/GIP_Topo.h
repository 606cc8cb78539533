// Partition layout of one rank (interface of GIP_Topo.h:19-92 in the
// reference): index ranges ALL/OWNED/INNER/... and the halo lists.
//
//   [0, interiors)            interior cells
//   [interiors, inner)        interface cells (read by neighbours)
//   [inner, owned)            boundary ghosts
//   [owned, all)              halo (copies of neighbours' interface cells)
//
// The halo lists are those of GIP_Topo.cpp:3-57: send_indx grouped by
// destination rank through send_row, recv_indx (halo slots) grouped by
// source rank through recv_row; block q of one rank's send list and block r
// of rank q's receive list name the same cells in the same order.  The
// exchange itself is the specification's (device pack/unpack kernels and
// the partitions' rendezvous, GIP_PartitionHIP.h), not this class's.
#ifndef _gip_topo_
#define _gip_topo_

#include <string>
#include <vector>

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

    int getAllSize() { return allsize; }
    int getOwnedSize() { return ownedsize; }
    int getInnerSize() { return innersize; }
    int getInterSize() { return interiors; }
    int getIfaceSize() { return interfaces; }
    int getHaloSize() { return halos; }
    int getBoundarySize() { return boundaries; }
    // the halo lists, under the reference's accessor names (GIP_Topo.h:78-83)
    int getSendSize() { return (int)send_indx.size(); }
    int getRecvSize() { return (int)recv_indx.size(); }
    int *getSendIndx() { return send_indx.data(); }
    int *getSendRow() { return send_row.data(); }
    int *getRecvIndx() { return recv_indx.data(); }
    int *getRecvRow() { return recv_row.data(); }
    // ranks the lists are grouped by (cols_size - 1 in the file)
    int getPartitions() { return send_row.empty() ? 1 : (int)send_row.size() - 1; }
    // number of elements a RUN_DOMAIN covers from index 0 (0 for the others,
    // as the reference's switch statements leave them untouched)
    int rangeSize(RUN_DOMAIN d);
    void printTopoInfo();

   protected:
    int allsize = 0, ownedsize = 0, innersize = 0;
    int interiors = 0, interfaces = 0, boundaries = 0, halos = 0;
    std::vector<int> send_indx, send_row, recv_indx, recv_row;
};

#endif
