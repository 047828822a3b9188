// H-matrix assembly on the GPU: block tasks from the host trees, dense leaves
// left for generation in the product layout, low-rank leaves compressed by the
// batched ACA into a compressed store, optionally recompressed (FP64 QR + SVD)
// and split into precision classes.
#pragma once

#include "common.hpp"
#include "kernels.hpp"
#include "tree.hpp"

namespace htg {

struct PointsOnDevice {
    DevArray<double> x, y, z;
    DevArray<int> perm;
    // Cluster-ordered copy of interleaved points: entry i is point perm[i].
    void upload(const double *pts, const std::vector<int> &perm_h);
    DevPoints view() const { return DevPoints{x.p, y.p, z.p, perm.p}; }
};

struct Assembly {
    int nr = 0, nc = 0;
    std::vector<int> perm_t, perm_s;
    std::vector<Leaf> leaves; // in no particular order
    int false_positive = 0;
    DevArray<double> store;   // U and V of the low-rank leaves (Leaf::src_u, src_v; complex: 2 doubles per entry)
    KernelDesc kd{};
    bool same = true;
    bool cplx = false; // complex kernel family: the store holds complex<double> entries
    bool delay_dense = false; // dense leaves not generated (filled by the caller later)
    bool host_gen = false;    // entries come from a host generator (no device kernel family)
    PointsOnDevice pT, pS;
    // ACA runs of admissible blocks (first attempts): features, terms needed,
    // first workspace terms; retries and workspace totals.
    std::vector<float> aca_feat;
    std::vector<int> aca_terms, aca_cap;
    int aca_retry = 0, aca_wide = 0, aca_int_wrap = 0;
    int64_t aca_ws_entries = 0, aca_need_entries = 0;
    double aca_seconds = 0;
    DevPoints dT() const { return pT.view(); }
    DevPoints dS() const { return same ? pT.view() : pS.view(); }
};

KernelDesc kernel_desc(const htg_generator &g, bool same);

// Kept triangle of a symmetric or Hermitian build ('L' or 'U'), 'N' for a general one.
inline char sym_uplo(const htg_params &p) {
    const int s = p.symmetry == 0 ? 'N' : p.symmetry;
    return s == 'N' ? 'N' : (char)p.uplo;
}
// Storage kind: 'N', 'S' or 'H' (the block tree and assembly of 'H' are those of 'S').
inline char sym_kind(const htg_params &p) {
    const int s = p.symmetry == 0 ? 'N' : p.symmetry;
    return s == 'H' ? 'H' : (s == 'S' ? 'S' : 'N');
}

// HMatrix::build (reference: include/htool/types/hmatrix.hpp:285-341, 422-607).
void assemble(const htg_params &prm, const htg_generator &gen, int nt, const double *xt, int ns, const double *xs,
              Assembly &out);
// The same with either a device kernel family (gen) or a host generator
// (hgen), and optionally the caller's cluster trees (NULL: built from the points).
void assemble_ex(const htg_params &prm, const htg_generator *gen, const htg_host_generator *hgen, bool same, int nt,
                 const double *xt, int ns, const double *xs, const ClusterTree *tree_t, const ClusterTree *tree_s,
                 Assembly &out);

// Recompresses the low-rank leaves in the store (FP64) and sets each leaf's
// per-term scales; then assigns precision classes (Leaf::k_hi).
// Without recompression (tol 0) each term is scaled to unit factors instead.
void recompress_and_split(const htg_params &prm, std::vector<Leaf> &leaves, DevArray<double> &store, bool cplx);

} // namespace htg
