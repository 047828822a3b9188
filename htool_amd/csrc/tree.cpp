// Host cluster tree and block tree. The arithmetic of the centre, radius,
// covariance, principal direction and split test is kept in the same order as
// the reference so that the same points give the same permutation bit for bit
// (reference: include/htool/clustering/pca.hpp:30-104, misc/evp.hpp:43-133,
// clustering/splitting.hpp:63-120).
#include "tree.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <utility>

namespace htg {
namespace {

inline double dot3(const double *a, const double *b) {
    double s = 0; // summed from zero in x, y, z order
    s = s + a[0] * b[0];
    s = s + a[1] * b[1];
    s = s + a[2] * b[2];
    return s;
}

// Direction of the largest eigenvalue of a symmetric 3x3 matrix by the
// trigonometric closed form (Smith, "Eigenvalues of a symmetric 3x3 matrix",
// CACM 4(4), 1961), followed by a column of (C - l2 I)(C - l3 I).
// c is column-major.
void principal_direction(const double c[9], double dir[3]) {
    auto C = [c](int i, int j) { return c[i + 3 * j]; };
    dir[0] = dir[1] = dir[2] = 0.0;
    const double off = std::pow(C(0, 1), 2) + std::pow(C(0, 2), 2) + std::pow(C(1, 2), 2);
    if (off < 1e-16) { // diagonal: pick the axis of the largest entry, first axis on ties
        double e0 = C(0, 0);
        int axis  = 0;
        if (e0 < C(1, 1)) {
            e0   = C(1, 1);
            axis = 1;
        }
        if (e0 < C(2, 2))
            axis = 2;
        dir[axis] = 1.0;
        return;
    }
    const double q  = (C(0, 0) + C(1, 1) + C(2, 2)) / 3.;
    const double p2 = std::pow(C(0, 0) - q, 2) + std::pow(C(1, 1) - q, 2) + std::pow(C(2, 2) - q, 2) + 2. * off;
    const double p  = std::sqrt(p2 / 6.);
    const double ip = 1. / p;
    double b[9];
    for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++)
            b[i + 3 * j] = (C(i, j) - (i == j ? 1.0 : 0.0) * q) * ip;
    auto B            = [&b](int i, int j) { return b[i + 3 * j]; };
    const double detb = B(0, 0) * (B(1, 1) * B(2, 2) - B(1, 2) * B(2, 1)) - B(0, 1) * (B(1, 0) * B(2, 2) - B(1, 2) * B(2, 0)) + B(0, 2) * (B(1, 0) * B(2, 1) - B(1, 1) * B(2, 0));
    const double half = detb / 2.;
    const double pi   = 3.14159265358979323846;
    const double phi  = half <= -1 ? pi / 3. : (half >= 1 ? 0.0 : std::acos(half) / 3.);
    const double l1   = q + 2. * p * std::cos(phi);
    double l3         = q + 2. * p * std::cos(phi + (2. * pi / 3.));
    double l2         = 3. * q - l1 - l3;
    if (std::abs(l2) < 1.e-12)
        l2 = 0.;
    if (std::abs(l3) < 1.e-12)
        l3 = 0.;
    if (std::abs(l1) < 1.e-16)
        return;
    double m1[9], m2[9];
    for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++) {
            m1[i + 3 * j] = C(i, j) - (i == j ? 1.0 : 0.0) * l2;
            m2[i + 3 * j] = C(i, j) - (i == j ? 1.0 : 0.0) * l3;
        }
    for (int col = 0; col < 3; col++) {
        double v[3];
        for (int i = 0; i < 3; i++) {
            double s = 0;
            for (int k = 0; k < 3; k++)
                s += m1[i + 3 * k] * m2[k + 3 * col];
            v[i] = s;
        }
        const double nrm = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (nrm >= 1.e-15) {
            dir[0] = v[0] / nrm;
            dir[1] = v[1] / nrm;
            dir[2] = v[2] / nrm;
            return;
        }
    }
    dir[0] = 1.0;
    dir[1] = dir[2] = 0.0;
}

} // namespace

void build_cluster_tree(int n, const double *x, int mcs, ClusterTree &tree) {
    if (n <= 0)
        throw std::invalid_argument("cluster tree needs at least one point");
    if (mcs <= 0)
        throw std::invalid_argument("minclustersize must be positive");
    tree.minclustersize = mcs;
    tree.nodes.assign(1, ClusterNode());
    tree.nodes[0].size = n;
    tree.perm.resize(n);
    std::iota(tree.perm.begin(), tree.perm.end(), 0);

    // Each pending node owns perm[offset, offset+size); a split reorders that
    // range stably into its sons, and is undone if a son is below minclustersize.
    std::vector<int> pending = {0};
    std::vector<int> scratch;
    std::vector<char> side;
    while (!pending.empty()) {
        const int ci = pending.back();
        pending.pop_back();
        ClusterNode node = tree.nodes[ci];
        int *num         = tree.perm.data() + node.offset;
        const int np     = node.size;
        const int nsons  = node.depth == 0 ? 1 : 2;

        double G = 0;
        for (int j = 0; j < np; j++)
            G += 1.0;
        double xc[3] = {0, 0, 0};
        for (int j = 0; j < np; j++)
            for (int d = 0; d < 3; d++)
                xc[d] += x[3 * num[j] + d];
        const double ig = 1. / G;
        for (int d = 0; d < 3; d++)
            xc[d] = xc[d] * ig;
        double cov[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        double rad    = 0;
        for (int j = 0; j < np; j++) {
            double u[3];
            for (int d = 0; d < 3; d++)
                u[d] = x[3 * num[j] + d] - xc[d];
            rad = std::max(rad, std::sqrt(std::abs(dot3(u, u))));
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++)
                    cov[a + 3 * b] += u[a] * u[b];
        }
        std::copy(xc, xc + 3, tree.nodes[ci].ctr);
        tree.nodes[ci].rad = rad;

        if (nsons == 1) { // the one-process level: a single son with every point, same order
            ClusterNode son;
            son.offset = node.offset;
            son.size   = np;
            son.depth  = node.depth + 1;
            son.rank   = 0;
            tree.nodes[ci].nb_sons   = 1;
            tree.nodes[ci].first_son = (int)tree.nodes.size();
            tree.nodes.push_back(son);
            pending.push_back(tree.nodes[ci].first_son);
            continue;
        }

        double dir[3];
        principal_direction(cov, dir);
        side.resize(np);
        int n0 = 0;
        for (int j = 0; j < np; j++) {
            double dx[3];
            for (int d = 0; d < 3; d++)
                dx[d] = x[3 * num[j] + d] - xc[d];
            side[j] = dot3(dir, dx) > 0 ? 0 : 1;
            n0 += side[j] == 0;
        }
        const int n1 = np - n0;
        if (n0 < mcs || n1 < mcs)
            continue; // leaf: keeps its points in the order they arrived
        scratch.resize(np);
        int a = 0, b = n0;
        for (int j = 0; j < np; j++)
            scratch[side[j] == 0 ? a++ : b++] = num[j];
        std::copy(scratch.begin(), scratch.begin() + np, num);
        tree.nodes[ci].nb_sons   = 2;
        tree.nodes[ci].first_son = (int)tree.nodes.size();
        for (int p = 0; p < 2; p++) {
            ClusterNode son;
            son.offset = node.offset + (p == 0 ? 0 : n0);
            son.size   = p == 0 ? n0 : n1;
            son.depth  = node.depth + 1;
            son.rank   = node.rank;
            tree.nodes.push_back(son);
        }
        pending.push_back(tree.nodes[ci].first_son);
        pending.push_back(tree.nodes[ci].first_son + 1);
    }
}

bool admissible(const ClusterNode &t, const ClusterNode &s, double eta) {
    double d[3] = {t.ctr[0] - s.ctr[0], t.ctr[1] - s.ctr[1], t.ctr[2] - s.ctr[2]};
    const double dist = std::sqrt(std::abs(dot3(d, d)));
    return 2 * std::min(t.rad, s.rad) < eta * std::max(dist - t.rad - s.rad, 0.);
}

namespace {

struct TaskBuilder {
    const ClusterTree &T, &S;
    double eta;
    std::size_t maxblocksize;
    std::vector<BlockTask> tasks;
    int mintd = 0, minsd = 0;
    bool deep(const ClusterNode &a, const ClusterNode &b) const { return a.depth >= mintd && b.depth >= minsd; }

    // Returns true when (t, s) was not turned into tasks (it may be merged into
    // a larger dense block by its parent).
    bool visit(int t, int s) {
        const ClusterNode &a = T.nodes[t], &b = S.nodes[s];
        const bool adm       = admissible(a, b, eta);
        if (adm && a.rank >= 0 && deep(a, b)) {
            tasks.push_back({t, s, true});
            return false;
        }
        bool split_t;
        if (b.leaf()) {
            if (a.leaf())
                return true;
            split_t = true;
        } else {
            split_t = a.leaf() ? false : (a.size > b.size);
        }
        const ClusterNode &owner = split_t ? a : b;
        bool all_not_pushed      = true;
        bool np[2]               = {false, false};
        for (int p = 0; p < owner.nb_sons; p++) {
            const int son = owner.first_son + p;
            np[p]         = split_t ? visit(son, s) : visit(t, son);
            all_not_pushed = all_not_pushed && np[p];
        }
        const std::size_t bsize = std::size_t(a.size) * std::size_t(b.size);
        if (bsize <= maxblocksize && all_not_pushed && a.rank >= 0 && deep(a, b))
            return true;
        for (int p = 0; p < owner.nb_sons; p++) {
            if (!np[p])
                continue;
            const int son = owner.first_son + p;
            const int ts = split_t ? son : t, ss = split_t ? s : son;
            tasks.push_back({ts, ss, admissible(T.nodes[ts], S.nodes[ss], eta)});
        }
        return false;
    }
};

// The symmetric traversal (blocks.hpp:140-223): tasks and merges only on
// coinciding or disjoint clusters; when neither side is a leaf every pair of
// sons is visited (son index p + l * nt for target son p, source son l).
struct SymTaskBuilder {
    const ClusterTree &T, &S;
    double eta;
    std::size_t maxblocksize;
    std::vector<BlockTask> tasks;
    int mintd = 0, minsd = 0;
    bool deep(const ClusterNode &a, const ClusterNode &b) const { return a.depth >= mintd && b.depth >= minsd; }

    bool visit(int t, int s) {
        const ClusterNode &a = T.nodes[t], &b = S.nodes[s];
        const bool ok        = sym_compatible(a, b);
        if (admissible(a, b, eta) && a.rank >= 0 && deep(a, b) && ok) {
            tasks.push_back({t, s, true});
            return false;
        }
        if (a.leaf() && b.leaf())
            return true;
        std::vector<std::pair<int, int>> kids;
        if (b.leaf()) {
            for (int p = 0; p < a.nb_sons; p++)
                kids.push_back({a.first_son + p, s});
        } else if (a.leaf()) {
            for (int p = 0; p < b.nb_sons; p++)
                kids.push_back({t, b.first_son + p});
        } else {
            for (int l = 0; l < b.nb_sons; l++)
                for (int p = 0; p < a.nb_sons; p++)
                    kids.push_back({a.first_son + p, b.first_son + l});
        }
        std::vector<char> np(kids.size());
        bool all_not_pushed = true;
        for (size_t k = 0; k < kids.size(); k++) {
            np[k]          = visit(kids[k].first, kids[k].second);
            all_not_pushed = all_not_pushed && np[k];
        }
        const std::size_t bsize = std::size_t(a.size) * std::size_t(b.size);
        if (bsize <= maxblocksize && all_not_pushed && a.rank >= 0 && deep(a, b) && ok)
            return true;
        for (size_t k = 0; k < kids.size(); k++)
            if (np[k])
                tasks.push_back({kids[k].first, kids[k].second,
                                 admissible(T.nodes[kids[k].first], S.nodes[kids[k].second], eta)});
        return false;
    }
};

} // namespace

bool sym_compatible(const ClusterNode &t, const ClusterNode &s) {
    if (t.offset == s.offset)
        return t.size == s.size;
    return t.offset < s.offset ? s.offset - t.offset >= t.size : t.offset - s.offset >= s.size;
}

std::vector<BlockTask> build_block_tasks(const ClusterTree &T, const ClusterTree &S, double eta,
                                         std::size_t maxblocksize, char uplo, int mintargetdepth, int minsourcedepth) {
    if (maxblocksize == 0)
        throw std::invalid_argument("maxblocksize cannot be zero");
    std::vector<BlockTask> all;
    if (uplo == 'L' || uplo == 'U') {
        SymTaskBuilder tb{T, S, eta, maxblocksize, {}, mintargetdepth, minsourcedepth};
        if (tb.visit(0, 0))
            tb.tasks.push_back({0, 0, admissible(T.nodes[0], S.nodes[0], eta)});
        all.swap(tb.tasks);
    } else {
        TaskBuilder tb{T, S, eta, maxblocksize, {}, mintargetdepth, minsourcedepth};
        if (tb.visit(0, 0))
            tb.tasks.push_back({0, 0, admissible(T.nodes[0], S.nodes[0], eta)});
        all.swap(tb.tasks);
    }
    // Only blocks whose target lies in this process's depth-1 cluster (rank 0)
    // are local (blocks.hpp:258-270); symmetric storage keeps one triangle.
    std::vector<BlockTask> local;
    for (const BlockTask &bt : all) {
        if (T.nodes[bt.t].rank != 0)
            continue;
        const int to = T.nodes[bt.t].offset, so = S.nodes[bt.s].offset;
        if ((uplo == 'L' && so > to) || (uplo == 'U' && so < to))
            continue;
        local.push_back(bt);
    }
    std::sort(local.begin(), local.end(), [&](const BlockTask &u, const BlockTask &v) {
        const int tu = T.nodes[u.t].offset, tv = T.nodes[v.t].offset;
        return tu == tv ? S.nodes[u.s].offset < S.nodes[v.s].offset : tu < tv;
    });
    return local;
}

} // namespace htg
