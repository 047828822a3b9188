// Host-side cluster tree and block tree of the H-matrix. These stay on the
// CPU: they are O(N log N) integer/geometry work that decides which blocks the
// GPU compresses and multiplies.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace htg {

// One node of the PCA cluster tree. Nodes are stored in a flat vector; the sons
// of a node are nodes[first_son .. first_son + nb_sons).
struct ClusterNode {
    int offset = 0, size = 0, depth = 0, rank = -1;
    int nb_sons = 0, first_son = -1;
    double ctr[3] = {0, 0, 0};
    double rad    = 0;
    bool leaf() const { return nb_sons == 0; }
};

struct ClusterTree {
    std::vector<ClusterNode> nodes; // nodes[0] is the root
    std::vector<int> perm;          // cluster position -> original point index
    int minclustersize = 10;
};

// PCA bisection with geometric splitting for one process: the root has a single
// son covering every point (reference: include/htool/clustering/pca.hpp:15-128,
// splitting.hpp:63-120, cluster.hpp:57-92, misc/evp.hpp:43-133).
void build_cluster_tree(int n, const double *x, int minclustersize, ClusterTree &tree);

// Rjasanow-Steinbach admissibility (include/htool/blocks/admissibility_conditions.hpp:16-19).
bool admissible(const ClusterNode &t, const ClusterNode &s, double eta);

struct BlockTask {
    int t, s;        // cluster node indices in the target / source trees
    bool admissible; // admissibility of (t, s)
};

// Block tree traversal producing the sorted list of tasks of the (single)
// process (include/htool/blocks/blocks.hpp:37-139, 237-278). uplo 'L' or 'U'
// selects the symmetric traversal (blocks.hpp:140-223: a block is a task or a
// merged dense block only when its clusters coincide or their index ranges do
// not overlap, and both sides split at once) and keeps the tasks of one
// triangle (s.offset <= t.offset for 'L', >= for 'U'); any other value the
// ordinary traversal.
// mintargetdepth / minsourcedepth: a block is a task or a merged dense block
// only when its clusters are at least that deep (blocks.hpp:51, 68, 154, 169).
std::vector<BlockTask> build_block_tasks(const ClusterTree &T, const ClusterTree &S, double eta,
                                         std::size_t maxblocksize, char uplo = 'N', int mintargetdepth = 0,
                                         int minsourcedepth = 0);

// Both clusters are the same, or their index ranges are disjoint: the blocks
// a symmetric block tree may keep (blocks.hpp:152, 170).
bool sym_compatible(const ClusterNode &t, const ClusterNode &s);

} // namespace htg
