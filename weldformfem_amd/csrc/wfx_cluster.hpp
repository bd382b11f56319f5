// Element clusters of the fused element pass (host plan, no HIP types).
//
// The elements are stored in cluster order: each connected component of the
// mesh is laid out brick by brick (bricks of about `cs` elements, longest
// along x, in Morton order) by its element centroids and cut into runs of at
// most `cs` elements; within a run the elements keep their ascending
// original order.  One work-group of the fused element pass
// handles one cluster and adds the forces of its element-nodes per cluster
// node in LDS, in that order, so each (cluster, node) pair, a "slot", gets
// one partial sum; the node pass adds a node's slots in ascending cluster
// order.  Clusters never span two components, so a specimen packed with
// others gets the same clusters, slots and sums as on its own.
#pragma once

#include <cstdint>
#include <vector>

namespace wfx {

struct ClusterPlan {
  std::vector<int> s2o;        // stored element -> original element
  std::vector<int> o2s;        // original element -> stored element
  std::vector<int> eoff;       // [C+1] first stored element of each cluster
  std::vector<int> sbase;      // [C+1] first slot of each cluster
  std::vector<uint16_t> loff;  // per cluster its M+1 entry offsets (relative), cluster c at sbase[c] + c
  std::vector<uint16_t> ent;   // per element-node entry t*8 + a (t: element within the cluster),
                               // grouped by slot; cluster c's entries at eoff[c] * n
  std::vector<int> nsl_off;    // [N+1] node -> its slots, ascending cluster
  std::vector<int> nsl;        // [S]
  std::vector<int> spos;       // [S] slot -> its position k in nsl (nsl[spos[s]] = s)
  std::vector<int> snode;      // [S] slot -> its node (ascending node within a cluster)
  std::vector<uint16_t> lslot; // [E][n] stored element's nodes as slot indices within its cluster
  int max_slots = 0;           // largest slot count of a cluster (<= slot_cap(n))
  int clusters() const { return (int)eoff.size() - 1; }
  long slots() const { return sbase.empty() ? 0 : sbase.back(); }
};

// Most slots a cluster may have: the element pass stages each slot's nodal
// data in the LDS its element-node forces use later, which holds this many
// slots for every precision (clusters with more are split).
constexpr int slot_cap(int n) { return n == 8 ? 512 : 256; }

// Plan for elnod [E][n] (original order, nodes in [0, N)) with node
// coordinates x [N][3]; cs <= 256 elements per cluster.  Returns 0.
int plan_clusters(int N, int E, int n, const int *elnod, const double *x, int cs, ClusterPlan &p);

}  // namespace wfx
