// Host side of libhtool_amd: product layout, storage modes and the C ABI
// declared in include/htool_amd.h. Assembly lives in assemble.cpp.
#include "assemble.hpp"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

namespace htg {
namespace {


// Default work decomposition (htg_layout): rows per row block, target entries
// per V chunk and per phase-1 wavefront, most chunks per wavefront, stacked V
// rows per chunk (one matrix-core tile).
constexpr htg_layout DEFAULT_LAYOUT = {128, 8192, 2048, 4, 16};
constexpr int CLASS_ALIGN[2] = {4, 8}; // entries per aligned group: class 1 may hold 16-bit values
// Row strides and block bases in entries: every stored row (a row of a row
// block, a V row) starts on a 128-byte line at every storage width of its
// class (32-bit class 0: 32 entries; 16-bit class 1: 64), so no line holds the
// end of one row and the start of the next: a row read in several passes
// fetches each line once.
constexpr int ROW_ALIGN[2] = {32, 64};
constexpr int CAP_MAX        = 512;    // must match the ACA kernel's CAPMAX

inline uint64_t align_to(uint64_t v, int a) { return (v + a - 1) / a * a; }
inline int round_up(int v, int a) { return (v + a - 1) / a * a; }
// V row of a low-rank leaf in class c: global columns [oj & ~(a-1), round_a(oj + n))
// (v_extent entries, read by the products), stored with stride v_stride.
inline int v_first(const Leaf &L, int c) { return L.oj & ~(CLASS_ALIGN[c] - 1); }
inline int v_extent(const Leaf &L, int c) { return round_up(L.oj + L.n, CLASS_ALIGN[c]) - v_first(L, c); }
inline int v_stride(const Leaf &L, int c) { return round_up(v_extent(L, c), ROW_ALIGN[c]); }
inline int terms(const Leaf &L, int c) { return L.k <= 0 ? 0 : (c == 0 ? L.k_hi : L.k - L.k_hi); }

// Entries per multi-RHS row-range work item (V phase): the column range of a
// range of rows is cut so that items hold about this many. Each cut costs a
// partial sum of rows x RHS FP64 entries written and read back (for 64
// complex RHS, 1 KB per row against 4 bytes per FP16 entry), so items are
// large; the lists still hold thousands of items.
constexpr int RR_WORK = 131072;

// V chunks grouped per wavefront, and the first chunk of every row group cut
// into several chunks (its partial sums are added by a second pass).
struct VTileList {
    std::vector<VTile> tiles, multi;
    int n0 = 0; // row-range lists: tiles [0, n0) are class 0 (XCD-ordered), the rest class 1
    std::vector<int> vg_off;
    DevArray<VTile> d_tiles, d_multi;
    DevArray<int> d_vg_off;
};

// Per precision class: the row blocks' column lists and blocks, the V arena and its chunks.
struct ClassLayout {
    std::vector<int> rb_col0, colidx;
    std::vector<uint64_t> rb_aoff;
    uint64_t arena_size = 0, varena_size = 0;
    int nvtiles = 0;
    std::vector<uint64_t> leaf_v_off; // per leaf
    DevArray<int> d_rb_col0, d_colidx;
    DevArray<uint64_t> d_rb_aoff;
    // storage variants of the arenas (FP64 master copy first)
    DevArray<double> a64, v64;
    DevArray<float> a32, v32;
    DevArray<__half> a16, v16;
    DevArray<__hip_bfloat16> ab16, vb16;
    bool empty() const { return colidx.empty() && nvtiles == 0; }
};

} // namespace
} // namespace htg

using namespace htg;

// One assembled instance inside an H-matrix: its leaves are [leaf_begin,
// leaf_end) of the (sorted) leaf list, its rows and columns start at `shift`.
// A transposed part (the H-matrix of A^T) shares the assembly of A: its
// leaves have rows and columns exchanged, U and V exchanged in the store, and
// dense leaves come from the generator with the point sets exchanged.
struct Part {
    std::shared_ptr<htg::Assembly> A;
    size_t leaf_begin = 0, leaf_end = 0;
    int shift       = 0;
    bool transposed = false;
};

struct htg_hmatrix {
    int nr = 0, nc = 0, tpad = 0;
    htg_layout lay = htg::DEFAULT_LAYOUT;
    std::vector<Part> parts; // assemblies kept for re-layout (empty for leaves from outside)
    std::vector<int> perm_t, perm_s;
    std::vector<Leaf> leaves; // sorted by (oi, oj)
    int false_positive = 0;
    double build_seconds = 0;
    // ACA runs of the build (first attempts): features, terms needed, first workspace terms
    std::vector<float> aca_feat;
    std::vector<int> aca_terms, aca_cap;
    int aca_retry = 0, aca_wide = 0, aca_int_wrap = 0;
    int64_t aca_ws_entries = 0, aca_need_entries = 0;
    double aca_seconds = 0;
    char symmetry = 'N', uplo = 'N'; // symmetric storage keeps one triangle; `leaves` adds mirrored copies
    std::vector<double> user_dense;  // dense leaves supplied by the caller (set_dense_blocks), get_leaves order
    std::vector<std::array<int, 4>> user_keys; // (off_i, off_j, m, n) of each block of user_dense, in its order
    // mvprod_subrhs: the H-matrix of the leaves whose columns meet a window,
    // per (offset, size), made on first use
    std::map<std::pair<int, int>, std::unique_ptr<htg_hmatrix>> subrhs_views;
    bool scaled = false; // some low-rank terms carry scales (recompressed)

    std::vector<int> rb_row0, rb_order;
    // single-RHS row phase: every row block whole, in XCD order (padding
    // entries 0xffffff)
    std::vector<int> rb_parts;
    // multi-RHS row-range row phase: every row block cut into parts of at most
    // RR_ITEM_ROWS rows (entry = rb | p << 24), in XCD order by stripes of
    // consecutive row blocks (padding entries 0xffffff)
    std::vector<int> rb_items;
    // per leaf, per covered row block: (row block, class-0 tile offset, class-1 tile offset)
    std::vector<std::vector<std::tuple<int, uint64_t, uint64_t>>> leaf_tiles;
    std::vector<std::vector<int>> panels; // low-rank leaves per source cluster
    std::vector<int> t_first[2];          // per class, per leaf: first t entry (-1 if none)
    std::vector<double> tscale; // per z entry (x part unused)
    int t_size = 0, part_size = 0;
    ClassLayout cls[2];
    // V chunks of both classes, grouped per wavefront, in three cuts of the
    // same V arena: vt[0] cuts every panel into row groups of vrows rows and
    // wide groups into column chunks of about vchunk_el entries (single right-
    // hand side in FP64); vt[1] keeps every panel no wider than vchunk_el whole
    // and shrinks its row groups instead (single right-hand side in FP32
    // arithmetic); vt[2] is vt[0] with chunks of twice vchunk_el entries
    // (several right-hand sides: longer matrix-core passes per chunk). vt[3]
    // and vt[4] are the 32-RHS row ranges (FP32 arithmetic, FP64): at most
    // RR_ITEM_ROWS rows of a panel's class over a column range of about
    // RR_WORK entries, one workgroup each, in XCD order (vg_off unused).
    VTileList vt[5];

    DevArray<int> d_perm_t, d_perm_s, d_rb_row0, d_rb_order, d_rb_parts, d_rb_items;
    DevArray<double> d_tscale;
    DevArray<double> d_z, d_part; // z: cluster-ordered x (tpad entries) then t, per RHS
    DevArray<float> d_xmax;       // 32-RHS FP32 passes: largest |x| per 32-entry block of z
    DevArray<unsigned> d_tmax;    // 32-RHS FP32 passes: largest class-1 |t| (FP32 bits)
    DevArray<double> d_hx, d_hy;  // staging for host-pointer products
    DevArray<int> d_iperm_s;      // inverse column permutation (32-RHS gather)
    DevArray<float> d_xstage;     // 32-RHS FP32 z, RHS-interleaved: x (gather staging) then t, (tpad + t) x 32
    int storage = HTG_STORAGE_FP64;
    bool cplx   = false; // complex<double> entries, x and y
    // infos["nb_mat_vec_prod"], infos["total_time_mat_vec_prod"] (hmatrix.hpp:1180-1181)
    int64_t n_products     = 0;
    double product_seconds = 0;
    // Layout generation: changes whenever device buffers a captured product
    // reads are replaced (layout rebuilds, storage changes); a graph keeps the
    // value it was captured at, and the token is set to ~0 when the H-matrix
    // is destroyed.
    std::shared_ptr<std::atomic<uint64_t>> generation = std::make_shared<std::atomic<uint64_t>>(0);
    ~htg_hmatrix() { generation->store(~uint64_t(0)); }
    int esz() const { return cplx ? 2 : 1; }
};

namespace htg {
namespace {

// z entry of term j of low-rank leaf l.
int t_entry(const htg_hmatrix &h, int l, int j) {
    const int kh = h.leaves[l].k_hi;
    return j < kh ? h.tpad + h.t_first[0][l] + j : h.tpad + h.t_first[1][l] + (j - kh);
}

// A work list in XCD order. Workgroup b runs on XCD b % 8, each XCD with its
// own L2; a group is work that reads the same vector entries (the parts of a
// row block, neighbouring row blocks, the row ranges of a V panel), so all of a
// group goes to one XCD, its entries adjacent so that they run at about the
// same time. Groups, heaviest first, go to the least loaded XCD; XCD j's list
// takes positions 8 i + j, shorter lists padded with `pad` (an entry the
// kernels skip).
template <typename E>
std::vector<E> xcd_order(const std::vector<std::vector<E>> &groups, const std::vector<int64_t> &weight, const E &pad) {
    constexpr int NX = 8;
    std::vector<int> order(groups.size());
    for (size_t g = 0; g < groups.size(); g++)
        order[g] = (int)g;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return weight[a] > weight[b]; });
    std::vector<std::vector<E>> lst(NX);
    std::vector<int64_t> load(NX, 0);
    for (int g : order) {
        const int j = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        lst[j].insert(lst[j].end(), groups[g].begin(), groups[g].end());
        load[j] += weight[g];
    }
    size_t len = 0;
    for (const auto &l : lst)
        len = std::max(len, l.size());
    std::vector<E> out(len * NX, pad);
    for (int j = 0; j < NX; j++)
        for (size_t i = 0; i < lst[j].size(); i++)
            out[i * NX + j] = lst[j][i];
    return out;
}

// Product layout.
//  * Row blocks: the segments between consecutive leaf row boundaries (every
//    leaf covers a segment entirely or not at all), split to <= rb_max rows.
//  * Per precision class, each row block owns one contiguous row-major R x C
//    block (row stride C rounded up to the class alignment, zero padded): the
//    class-0 block holds the rows of the dense leaves (n columns each) and of
//    the low-rank leaves' leading k_hi U columns; the class-1 block holds the
//    remaining U columns. colidx gives, per column, the z entry it multiplies.
//  * Each low-rank leaf stores its V rows (class 0: the first k_hi, class 1 the
//    rest) row-major and aligned to the global column index, stacked with the
//    other leaves of its source cluster (a panel); panel rows are cut into
//    groups of at most vrows rows and about vchunk_el entries (columns are
//    cut only in panels wider than vchunk_el), grouped into wavefront work
//    lists (htg_layout).
void build_layout(htg_hmatrix &h) {
    std::vector<int> bnd = {0, h.nr};
    for (const Leaf &L : h.leaves) {
        bnd.push_back(L.oi);
        bnd.push_back(L.oi + L.m);
    }
    std::sort(bnd.begin(), bnd.end());
    bnd.erase(std::unique(bnd.begin(), bnd.end()), bnd.end());
    h.rb_row0.clear();
    for (size_t b = 0; b + 1 < bnd.size(); b++) {
        const int a = bnd[b], len = bnd[b + 1] - a;
        const int parts = (len + h.lay.rb_max - 1) / h.lay.rb_max;
        for (int p = 0; p < parts; p++)
            h.rb_row0.push_back(a + (int)((long long)len * p / parts));
    }
    h.rb_row0.push_back(h.nr);
    const int nrb = (int)h.rb_row0.size() - 1;
    if (nrb > 0xffffff)
        throw Error(HTG_ERR_UNSUPPORTED, "more than 2^24 row blocks");
    h.tpad        = round_up(h.nc, 8);

    // Low-rank leaves sharing a source cluster (same column range) form one
    // panel per class: their V rows are stacked with a common stride, so a
    // phase-1 chunk covers rows of several leaves and reads its x slice once.
    std::map<std::pair<int, int>, int> panel_of;
    h.panels.clear();
    for (size_t l = 0; l < h.leaves.size(); l++) {
        const Leaf &L = h.leaves[l];
        if (L.k <= 0)
            continue;
        auto it = panel_of.emplace(std::make_pair(L.oj, L.n), (int)panel_of.size()).first;
        if (it->second == (int)h.panels.size())
            h.panels.emplace_back();
        h.panels[it->second].push_back((int)l);
    }
    // panels in column order
    std::vector<int> porder;
    for (const auto &kv : panel_of)
        porder.push_back(kv.second);
    {
        std::vector<std::vector<int>> sorted;
        for (int p : porder)
            sorted.push_back(std::move(h.panels[p]));
        h.panels.swap(sorted);
    }

    // t entries (one per low-rank term): class by class, panel by panel, leaf
    // by leaf, so the class-c terms of a panel are contiguous. Scales per entry.
    h.t_size = 0;
    h.scaled = false;
    for (int c = 0; c < 2; c++) {
        h.t_first[c].assign(h.leaves.size(), -1);
        for (const auto &P : h.panels)
            for (int l : P) {
                const int kc = terms(h.leaves[l], c);
                if (kc > 0) {
                    h.t_first[c][l] = h.t_size;
                    h.t_size += kc;
                }
            }
    }
    for (const Leaf &L : h.leaves)
        h.scaled = h.scaled || !L.sig.empty();
    h.tscale.clear();
    if (h.scaled) {
        h.tscale.assign(h.tpad + h.t_size, 1.0);
        for (size_t l = 0; l < h.leaves.size(); l++)
            for (int j = 0; j < (int)h.leaves[l].sig.size(); j++)
                h.tscale[t_entry(h, (int)l, j)] = h.leaves[l].sig[j];
    }

    // Leaves of each row block, in leaf order.
    std::vector<int> cnt(nrb + 1, 0);
    std::vector<std::pair<int, int>> inc(h.leaves.size());
    for (size_t l = 0; l < h.leaves.size(); l++) {
        const Leaf &L = h.leaves[l];
        const int r0 = (int)(std::upper_bound(h.rb_row0.begin(), h.rb_row0.end(), L.oi) - h.rb_row0.begin()) - 1;
        int r1       = r0;
        while (r1 < nrb && h.rb_row0[r1] < L.oi + L.m)
            r1++;
        inc[l] = {r0, r1};
        if (L.k != 0)
            for (int r = r0; r < r1; r++)
                cnt[r + 1]++;
    }
    for (int r = 0; r < nrb; r++)
        cnt[r + 1] += cnt[r];
    std::vector<int> rb_leaves(cnt[nrb]);
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (size_t l = 0; l < h.leaves.size(); l++)
        if (h.leaves[l].k != 0)
            for (int r = inc[l].first; r < inc[l].second; r++)
                rb_leaves[fill[r]++] = (int)l;

    h.leaf_tiles.assign(h.leaves.size(), {});
    for (int c = 0; c < 2; c++) {
        ClassLayout &C = h.cls[c];
        C.rb_col0.assign(nrb + 1, 0);
        C.rb_aoff.assign(nrb, 0);
        C.colidx.clear();
        C.arena_size = 0;
    }
    for (int r = 0; r < nrb; r++) {
        const int R = h.rb_row0[r + 1] - h.rb_row0[r];
        int col[2]  = {0, 0};
        for (int c = 0; c < 2; c++)
            h.cls[c].rb_aoff[r] = h.cls[c].arena_size;
        for (int q = cnt[r]; q < cnt[r + 1]; q++) {
            const int l   = rb_leaves[q];
            const Leaf &L = h.leaves[l];
            const uint64_t off0 = h.cls[0].rb_aoff[r] + (uint64_t)col[0];
            const uint64_t off1 = h.cls[1].rb_aoff[r] + (uint64_t)col[1];
            h.leaf_tiles[l].emplace_back(r, off0, off1);
            if (L.k < 0) {
                for (int j = 0; j < L.n; j++)
                    h.cls[0].colidx.push_back(L.oj + j);
                col[0] += L.n;
            } else {
                for (int j = 0; j < L.k_hi; j++)
                    h.cls[0].colidx.push_back(t_entry(h, l, j));
                for (int j = L.k_hi; j < L.k; j++)
                    h.cls[1].colidx.push_back(t_entry(h, l, j));
                col[0] += L.k_hi;
                col[1] += L.k - L.k_hi;
            }
        }
        for (int c = 0; c < 2; c++) {
            ClassLayout &C   = h.cls[c];
            C.rb_col0[r + 1] = C.rb_col0[r] + col[c];
            C.arena_size     = align_to(C.arena_size + (uint64_t)R * round_up(col[c], ROW_ALIGN[c]), ROW_ALIGN[c]);
        }
    }
    // Heaviest row blocks are dispatched first.
    h.rb_order.resize(nrb);
    for (int r = 0; r < nrb; r++)
        h.rb_order[r] = r;
    auto work = [&](int r) {
        int64_t cols = 0;
        for (int c = 0; c < 2; c++)
            cols += h.cls[c].rb_col0[r + 1] - h.cls[c].rb_col0[r];
        return (int64_t)(h.rb_row0[r + 1] - h.rb_row0[r]) * cols;
    };
    std::stable_sort(h.rb_order.begin(), h.rb_order.end(), [&](int a, int b) { return work(a) > work(b); });
    // Multi-RHS row phase: parts of RR_ITEM_ROWS rows, in stripes of
    // consecutive row blocks (rows under the same leaves read the same t
    // entries, neighbours share near-field columns) of about 1/128 of the work,
    // each stripe on one XCD.
    {
        constexpr int NSTRIPE = 128;
        int64_t total = 0;
        for (int r = 0; r < nrb; r++)
            total += work(r);
        const int64_t cap = std::max<int64_t>(1, total / NSTRIPE);
        std::vector<std::vector<int>> stripes;
        std::vector<int64_t> sw;
        for (int r = 0; r < nrb; r++) {
            if (stripes.empty() || sw.back() >= cap) {
                stripes.emplace_back();
                sw.push_back(0);
            }
            for (int p = 0; p * RR_ITEM_ROWS < h.rb_row0[r + 1] - h.rb_row0[r]; p++)
                stripes.back().push_back(r | p << 24);
            sw.back() += work(r);
        }
        h.rb_items = xcd_order(stripes, sw, 0xffffff);
    }
    // Single right-hand side: whole row blocks, each on one XCD (its column
    // indices and vector entries fetched into one L2), heaviest first per XCD.
    // Measured at N = 1e5 against parts of <= 32 rows (the previous form): mixed
    // row phase 75.8 -> 69.7 us, FP64 214 -> 213 us, FP32 unchanged.
    {
        std::vector<std::vector<int>> blocks(nrb);
        std::vector<int64_t> bw(nrb);
        for (int r = 0; r < nrb; r++) {
            blocks[r] = {r};
            bw[r]     = work(r);
        }
        h.rb_parts = xcd_order(blocks, bw, 0xffffff);
    }

    // V arena: per class, panel by panel, the stacked V rows of the panel's
    // leaves with a common stride.
    for (int c = 0; c < 2; c++) {
        ClassLayout &C = h.cls[c];
        C.varena_size  = 0;
        C.leaf_v_off.assign(h.leaves.size(), 0);
        for (const auto &P : h.panels) {
            int K = 0, ld = 0;
            for (int l : P) {
                const int kc = terms(h.leaves[l], c);
                if (kc > 0) {
                    if (K == 0)
                        ld = v_stride(h.leaves[l], c);
                    C.leaf_v_off[l] = C.varena_size + (uint64_t)K * ld;
                    K += kc;
                }
            }
            if (K > 0)
                C.varena_size = align_to(C.varena_size + (uint64_t)K * ld, ROW_ALIGN[c]);
        }
    }
    // V chunks, class 0 then class 1, grouped per wavefront, in the three
    // cuts. A panel's stacked rows are cut into groups of at most vrows rows;
    // in cuts 0 and 2 (and for a panel wider than vchunk_el in cut 1) each
    // group's columns are cut into chunks of about vchunk_el entries (twice
    // that in cut 2), in cut 1 a narrower panel's groups keep all columns and
    // take at most vchunk_el entries.
    h.part_size = 0;
    for (int cut = 0; cut < 3; cut++) {
        VTileList &T = h.vt[cut];
        T.tiles.clear();
        T.multi.clear();
        T.vg_off.assign(1, 0);
        int part_size    = 0;
        int64_t group_el = 0;
        for (int c = 0; c < 2; c++) {
            ClassLayout &C = h.cls[c];
            if (cut == 0)
                C.nvtiles = 0;
            for (const auto &P : h.panels) {
                int K = 0, first = -1;
                for (int l : P)
                    if (terms(h.leaves[l], c) > 0) {
                        if (first < 0)
                            first = l;
                        K += terms(h.leaves[l], c);
                    }
                if (K == 0)
                    continue;
                const Leaf &L0       = h.leaves[first];
                const int ld         = v_stride(L0, c);
                const int ext        = v_extent(L0, c);
                const uint64_t base  = C.leaf_v_off[first];
                const int tfirst     = t_entry(h, first, c == 0 ? 0 : L0.k_hi);
                const bool whole     = cut == 1 && ext <= h.lay.vchunk_el;
                const int chunk_el   = cut == 2 ? 2 * h.lay.vchunk_el : h.lay.vchunk_el;
                const int rows_g     = whole ? std::max(1, std::min(h.lay.vrows, h.lay.vchunk_el / ext)) : h.lay.vrows;
                for (int g0 = 0; g0 < K; g0 += rows_g) {
                    const int kr     = std::min(rows_g, K - g0);
                    const int cpc    = whole ? ext : round_up(std::max(64, chunk_el / kr), ROW_ALIGN[c]);
                    const int nparts = (ext + cpc - 1) / cpc;
                    const int pbase  = part_size;
                    if (nparts > 1)
                        part_size += nparts * kr;
                    for (int q = 0; q < nparts; q++) {
                        VTile v{};
                        const int j0 = q * cpc;
                        v.v_off      = base + (uint64_t)g0 * ld + (uint64_t)j0;
                        v.ld         = ld;
                        v.k          = kr;
                        v.ncols      = std::min(cpc, ext - j0);
                        v.x_off      = v_first(L0, c) + j0;
                        v.t_index    = tfirst + g0;
                        v.part_off   = pbase + q * kr;
                        v.part_base  = pbase;
                        v.nparts     = nparts;
                        v.cls        = c;
                        if (nparts > 1 && q == 0)
                            T.multi.push_back(v);
                        const int64_t el   = (int64_t)v.k * v.ncols;
                        const int in_group = (int)T.tiles.size() - T.vg_off.back();
                        if (in_group > 0 && (group_el + el > h.lay.vgroup_el || in_group >= h.lay.vgroup_tiles)) {
                            T.vg_off.push_back((int)T.tiles.size());
                            group_el = 0;
                        }
                        T.tiles.push_back(v);
                        if (cut == 0)
                            C.nvtiles++;
                        group_el += el;
                    }
                }
            }
        }
        if ((int)T.tiles.size() > T.vg_off.back())
            T.vg_off.push_back((int)T.tiles.size());
        {
            // wavefront groups in decreasing work (16-bit rows count half):
            // the long groups start first and the short ones fill the end
            const int ng = (int)T.vg_off.size() - 1;
            std::vector<int64_t> work(ng, 0);
            for (int g = 0; g < ng; g++)
                for (int t = T.vg_off[g]; t < T.vg_off[g + 1]; t++)
                    work[g] += (int64_t)T.tiles[t].k * T.tiles[t].ncols * (T.tiles[t].cls == 0 ? 2 : 1);
            std::vector<int> ord(ng);
            for (int g = 0; g < ng; g++)
                ord[g] = g;
            std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return work[a] > work[b]; });
            std::vector<VTile> nt;
            std::vector<int> no(1, 0);
            for (int g : ord) {
                for (int t = T.vg_off[g]; t < T.vg_off[g + 1]; t++)
                    nt.push_back(T.tiles[t]);
                no.push_back((int)nt.size());
            }
            T.tiles.swap(nt);
            T.vg_off.swap(no);
        }
        h.part_size = std::max(h.part_size, part_size); // the cuts never run at once
    }
    // Multi-RHS row ranges: rows [r0, r0 + RR_ITEM_ROWS) of a panel's class,
    // their columns cut into ranges of whole staged chunks holding about RR_WORK
    // entries each (fewer, longer ranges for fewer rows), in XCD order by panel.
    for (int cut = 3; cut < 5; cut++) {
        VTileList &T = h.vt[cut];
        T.tiles.clear();
        T.multi.clear();
        T.vg_off.assign(1, 0);
        const int cw  = cut == 3 ? RR_ITEM_CHUNK32 : RR_ITEM_CHUNK64;
        int part_size = 0;
        std::vector<std::vector<VTile>> groups[2]; // per class, per panel
        groups[0].resize(h.panels.size());
        groups[1].resize(h.panels.size());
        for (int c = 0; c < 2; c++) {
            const ClassLayout &C = h.cls[c];
            for (size_t pi = 0; pi < h.panels.size(); pi++) {
                const auto &P = h.panels[pi];
                int K = 0, first = -1;
                for (int l : P)
                    if (terms(h.leaves[l], c) > 0) {
                        if (first < 0)
                            first = l;
                        K += terms(h.leaves[l], c);
                    }
                if (K == 0)
                    continue;
                const Leaf &L0   = h.leaves[first];
                const int ext    = v_extent(L0, c);
                const int ld     = v_stride(L0, c);
                const int tfirst = t_entry(h, first, c == 0 ? 0 : L0.k_hi);
                for (int r0 = 0; r0 < K; r0 += RR_ITEM_ROWS) {
                    const int rows = std::min(RR_ITEM_ROWS, K - r0);
                    int width      = std::max(cw, round_up((RR_WORK + rows - 1) / rows, cw));
                    const int np0  = (ext + width - 1) / width;
                    width          = round_up((ext + np0 - 1) / np0, cw);
                    const int np   = (ext + width - 1) / width;
                    const int pbase = part_size;
                    if (np > 1)
                        part_size += np * rows;
                    for (int q = 0; q < np; q++) {
                        VTile v{};
                        v.v_off     = C.leaf_v_off[first] + (uint64_t)r0 * ld + (uint64_t)q * width;
                        v.ld        = ld;
                        v.k         = rows;
                        v.ncols     = std::min(width, ext - q * width);
                        v.x_off     = v_first(L0, c) + q * width;
                        v.t_index   = tfirst + r0;
                        v.part_off  = pbase + q * rows;
                        v.part_base = pbase;
                        v.nparts    = np;
                        v.cls       = c;
                        if (np > 1 && q == 0)
                            T.multi.push_back(v);
                        groups[c][pi].push_back(v);
                    }
                }
            }
        }
        // per class (class 0 first; their lists can run as separate launches), a
        // panel's items on one XCD, those reading the same x columns adjacent
        T.tiles.clear();
        for (int c = 0; c < 2; c++) {
            std::vector<int64_t> gw(groups[c].size(), 0);
            for (size_t pi = 0; pi < groups[c].size(); pi++) {
                std::stable_sort(groups[c][pi].begin(), groups[c][pi].end(),
                                 [](const VTile &a, const VTile &b) { return a.x_off < b.x_off; });
                for (const VTile &v : groups[c][pi])
                    gw[pi] += (int64_t)v.k * v.ncols;
            }
            const std::vector<VTile> l = xcd_order(groups[c], gw, VTile{}); // padding: k = 0
            T.tiles.insert(T.tiles.end(), l.begin(), l.end());
            if (c == 0)
                T.n0 = (int)T.tiles.size();
        }
        h.part_size = std::max(h.part_size, part_size);
    }
}

void upload_layout(htg_hmatrix &h) {
    h.d_perm_t.upload(h.perm_t);
    h.d_perm_s.upload(h.perm_s);
    h.d_rb_row0.upload(h.rb_row0);
    h.d_rb_order.upload(h.rb_order);
    h.d_rb_parts.upload(h.rb_parts);
    h.d_rb_items.upload(h.rb_items);
    for (ClassLayout &C : h.cls) {
        C.d_rb_col0.upload(C.rb_col0);
        C.d_rb_aoff.upload(C.rb_aoff);
        C.d_colidx.upload(C.colidx);
    }
    for (VTileList &T : h.vt) {
        T.d_tiles.upload(T.tiles);
        T.d_vg_off.upload(T.vg_off);
        T.d_multi.upload(T.multi);
    }
    if (h.scaled)
        h.d_tscale.upload(h.tscale);
    // z = [x (nc entries, zero padded to tpad) | t], per RHS; zeroed once so the
    // padding entries are always finite zeros.
    h.d_z.alloc((size_t)(h.tpad + h.t_size) * 64);
    h.d_z.zero();
    h.d_part.alloc((size_t)std::max(h.part_size, 1) * 64 * h.esz()); // complex FP32 passes: 64 (re, im) per row
    h.d_xmax.alloc((size_t)h.tpad / 32 + 1);
    // multi-RHS gathers: the inverse column permutation and an nc x 32 FP32
    // staging buffer, allocated here so that products allocate nothing
    std::vector<int> ip(h.nc);
    for (int j = 0; j < h.nc; j++)
        ip[h.perm_s[j]] = j;
    h.d_iperm_s.upload(ip);
    h.d_xstage.alloc((size_t)(h.tpad + h.t_size) * 32 * (h.cplx ? 4 : 1)); // complex: 64 (re, im) pairs
    h.d_xstage.zero(); // padding entries [nc, tpad) stay zero
    h.d_tmax.alloc(TMAX_SLOTS);
    h.generation->fetch_add(1); // graphs captured before read freed buffers
}

// FP64 master arenas of both classes, zero filled (row padding stays zero).
void alloc_master(htg_hmatrix &h) {
    for (ClassLayout &C : h.cls) {
        C.a64.alloc(C.arena_size * h.esz());
        C.v64.alloc(C.varena_size * h.esz());
        C.a64.zero();
        C.v64.zero();
    }
}

int row_block_rows(const htg_hmatrix &h, int rb) { return h.rb_row0[rb + 1] - h.rb_row0[rb]; }
int row_stride(const htg_hmatrix &h, int c, int rb) {
    return round_up(h.cls[c].rb_col0[rb + 1] - h.cls[c].rb_col0[rb], ROW_ALIGN[c]);
}

// Right-hand sides per pass: at most 32, the width of the matrix-core passes
// (FP32 and FP64 arithmetic).
int pick_mu_chunk(int remaining, int cap) {
    int c = cap;
    while (c > remaining)
        c >>= 1;
    return c;
}

template <typename T>
const T *arena_of(const ClassLayout &C, bool v);
template <>
const double *arena_of<double>(const ClassLayout &C, bool v) { return v ? C.v64.p : C.a64.p; }
template <>
const float *arena_of<float>(const ClassLayout &C, bool v) { return v ? C.v32.p : C.a32.p; }
template <>
const __half *arena_of<__half>(const ClassLayout &C, bool v) { return v ? C.v16.p : C.a16.p; }
template <>
const __hip_bfloat16 *arena_of<__hip_bfloat16>(const ClassLayout &C, bool v) { return v ? C.vb16.p : C.ab16.p; }

template <typename TS0, typename TS1, typename TC>
void product(htg_hmatrix &h, const double *d_x, double *d_y, int mu, hipStream_t st, hipEvent_t *ev) {
    TC *z = reinterpret_cast<TC *>(h.d_z.p);
    const double *tscale = h.scaled ? h.d_tscale.p : nullptr;
    const bool has1      = !h.cls[1].empty();
    const RowClass rc0{h.cls[0].d_rb_col0.p, h.cls[0].d_rb_aoff.p, h.cls[0].d_colidx.p, CLASS_ALIGN[0], ROW_ALIGN[0]};
    const RowClass rc1{h.cls[1].d_rb_col0.p, h.cls[1].d_rb_aoff.p, h.cls[1].d_colidx.p, CLASS_ALIGN[1], ROW_ALIGN[1]};
    int done = 0;
    while (done < mu) {
        const int MU = pick_mu_chunk(mu - done, 32);
        if (ev)
            HTG_HIP(hipEventRecord(ev[0], st));
        if (std::is_same<TC, float>::value && MU == 32) {
            launch_gather_x32(d_x, h.d_iperm_s.p, h.nc, h.tpad, done, h.d_xstage.p, reinterpret_cast<float *>(z),
                              h.d_xmax.p, h.d_tmax.p, st);
        } else if (std::is_same<TC, double>::value && MU == 32) {
            launch_gather_x64(d_x, h.d_iperm_s.p, h.nc, h.tpad, done, reinterpret_cast<double *>(z), st);
        } else {
            launch_gather_x<TC>(d_x, h.d_perm_s.p, h.nc, h.tpad, done, MU, z, st);
        }
        if (ev)
            HTG_HIP(hipEventRecord(ev[1], st));
        const ClassLayout &C0 = h.cls[0], &C1 = h.cls[1];
        if (MU == 32) {
            const VTileList &S = h.vt[std::is_same<TC, float>::value ? 3 : 4];
            launch_vx_rr<TS0, TS1, TC>(S.d_tiles.p, (int)S.tiles.size(), S.d_multi.p, (int)S.multi.size(),
                                       arena_of<TS0>(C0, true), has1 ? arena_of<TS1>(C1, true) : nullptr, z,
                                       h.d_xstage.p, h.d_xmax.p, h.d_tmax.p, h.d_part.p, tscale, st);
        } else {
            const VTileList &T = h.vt[MU > 1 ? 2 : std::is_same<TC, float>::value ? 1 : 0];
            launch_vx<TS0, TS1, TC>(T.d_vg_off.p, (int)T.vg_off.size() - 1, T.d_tiles.p, T.d_multi.p,
                                    (int)T.multi.size(), arena_of<TS0>(C0, true),
                                    has1 ? arena_of<TS1>(C1, true) : nullptr, z, h.d_part.p, tscale, MU, st);
        }
        if (ev)
            HTG_HIP(hipEventRecord(ev[2], st));
        if (MU == 32) {
            launch_rows_rr<TS0, TS1, TC>(h.d_rb_items.p, (int)h.rb_items.size(), h.d_rb_row0.p, rc0,
                                         arena_of<TS0>(C0, false), rc1, has1 ? arena_of<TS1>(C1, false) : nullptr, z,
                                         h.d_xstage.p, h.d_tmax.p, h.d_perm_t.p, h.nr, done, d_y, st);
        } else {
            // single RHS: the XCD-ordered row-block list
            const bool parts = MU == 1;
            launch_rows<TS0, TS1, TC>(parts ? h.d_rb_parts.p : h.d_rb_order.p, h.d_rb_row0.p,
                                      parts ? (int)h.rb_parts.size() : (int)h.rb_row0.size() - 1, rc0,
                                      arena_of<TS0>(C0, false), rc1, has1 ? arena_of<TS1>(C1, false) : nullptr, z,
                                      h.d_perm_t.p, h.nr, done, MU, d_y, st);
        }
        if (ev)
            HTG_HIP(hipEventRecord(ev[3], st));
        done += MU;
    }
    HTG_HIP(hipGetLastError());
}

// Complex product: the same layout with (re, im) pairs; at most 32 RHS per pass
// (FP64 arithmetic) or 16 (FP32 arithmetic).
template <typename TS0, typename TS1, typename R>
void product_c(htg_hmatrix &h, const double *d_x, double *d_y, int mu, hipStream_t st, hipEvent_t *ev) {
    R *z                 = reinterpret_cast<R *>(h.d_z.p);
    const double *tscale = h.scaled ? h.d_tscale.p : nullptr;
    const bool has1      = !h.cls[1].empty();
    const RowClass rc0{h.cls[0].d_rb_col0.p, h.cls[0].d_rb_aoff.p, h.cls[0].d_colidx.p, CLASS_ALIGN[0], ROW_ALIGN[0]};
    const RowClass rc1{h.cls[1].d_rb_col0.p, h.cls[1].d_rb_aoff.p, h.cls[1].d_colidx.p, CLASS_ALIGN[1], ROW_ALIGN[1]};
    int done = 0;
    while (done < mu) {
        // complex RHS per pass: 32 (FP64) or 64 (FP32 arithmetic) on the row-range
        // kernels, fewer for the remainder
        const int MU = pick_mu_chunk(mu - done, std::is_same<R, float>::value ? 64 : 32);
        if (ev)
            HTG_HIP(hipEventRecord(ev[0], st));
        if constexpr (std::is_same<R, float>::value) {
            if (MU >= 32) {
                const ClassLayout &C0 = h.cls[0], &C1 = h.cls[1];
                const VTileList &S = h.vt[4];
                launch_rrc32<TS0, TS1>(MU, d_x, h.d_iperm_s.p, h.nc, h.tpad, done, h.d_xstage.p, h.d_xmax.p, h.d_tmax.p,
                                       S.d_tiles.p, (int)S.tiles.size(), S.n0, S.d_multi.p, (int)S.multi.size(),
                                       h.d_rb_items.p, (int)h.rb_items.size(), h.d_rb_row0.p, rc0,
                                       arena_of<TS0>(C0, false), rc1, has1 ? arena_of<TS1>(C1, false) : nullptr,
                                       arena_of<TS0>(C0, true), has1 ? arena_of<TS1>(C1, true) : nullptr, h.d_part.p,
                                       tscale, h.d_perm_t.p, h.nr, d_y, st, ev ? ev[1] : nullptr,
                                       ev ? ev[2] : nullptr);
                if (ev)
                    HTG_HIP(hipEventRecord(ev[3], st));
                done += MU;
                continue;
            }
        }
        if (std::is_same<R, double>::value && (MU == 16 || MU == 32)) {
            launch_gather_xc64(d_x, h.d_iperm_s.p, h.nc, h.tpad, done, MU, reinterpret_cast<double *>(z), st);
        } else if (std::is_same<R, float>::value && MU == 16) {
            launch_gather_xc32(d_x, h.d_iperm_s.p, h.nc, h.tpad, done, h.d_xstage.p, reinterpret_cast<float *>(z), st);
        } else {
            launch_gather_xc<R>(d_x, h.d_perm_s.p, h.nc, h.tpad, done, MU, z, st);
        }
        if (ev)
            HTG_HIP(hipEventRecord(ev[1], st));
        const ClassLayout &C0 = h.cls[0], &C1 = h.cls[1];
        if (std::is_same<R, double>::value && MU == 32) {
            const VTileList &S = h.vt[4];
            launch_rrc64(S.d_tiles.p, (int)S.tiles.size(), S.d_multi.p, (int)S.multi.size(), h.d_rb_items.p,
                         (int)h.rb_items.size(), h.d_rb_row0.p, rc0, arena_of<double>(C0, false), rc1,
                         has1 ? arena_of<double>(C1, false) : nullptr, arena_of<double>(C0, true),
                         has1 ? arena_of<double>(C1, true) : nullptr, reinterpret_cast<double *>(z), h.d_part.p,
                         tscale, h.d_perm_t.p, h.nr, done, d_y, st, ev ? ev[2] : nullptr);
            if (ev)
                HTG_HIP(hipEventRecord(ev[3], st));
            done += MU;
            continue;
        }
        // longer chunks measured faster for the FP32-arithmetic passes only
        const VTileList &T = h.vt[MU > 1 && std::is_same<R, float>::value ? 2 : 0];
        launch_vx_c<TS0, TS1, R>(T.d_vg_off.p, (int)T.vg_off.size() - 1, T.d_tiles.p, T.d_multi.p,
                                 (int)T.multi.size(), arena_of<TS0>(C0, true),
                                 has1 ? arena_of<TS1>(C1, true) : nullptr, z, h.d_part.p, tscale, MU, st);
        if (ev)
            HTG_HIP(hipEventRecord(ev[2], st));
        launch_rows_c<TS0, TS1, R>(h.d_rb_order.p, h.d_rb_row0.p, (int)h.rb_row0.size() - 1, rc0,
                                   arena_of<TS0>(C0, false), rc1, has1 ? arena_of<TS1>(C1, false) : nullptr, z,
                                   h.d_perm_t.p, h.nr, done, MU, d_y, st);
        if (ev)
            HTG_HIP(hipEventRecord(ev[3], st));
        done += MU;
    }
    HTG_HIP(hipGetLastError());
}

void run_product(htg_hmatrix &h, const double *d_x, double *d_y, int mu, hipStream_t st, hipEvent_t *ev = nullptr) {
    if (mu <= 0)
        throw Error(HTG_ERR_INVALID, "mu must be positive");
    if (h.cplx) {
        switch (h.storage) {
        case HTG_STORAGE_FP32:
            product_c<float, float, float>(h, d_x, d_y, mu, st, ev);
            break;
        case HTG_STORAGE_FP32_FP16:
            product_c<float, __half, float>(h, d_x, d_y, mu, st, ev);
            break;
        case HTG_STORAGE_FP32_BF16:
            product_c<float, __hip_bfloat16, float>(h, d_x, d_y, mu, st, ev);
            break;
        default:
            product_c<double, double, double>(h, d_x, d_y, mu, st, ev);
        }
        return;
    }
    switch (h.storage) {
    case HTG_STORAGE_FP32:
        product<float, float, float>(h, d_x, d_y, mu, st, ev);
        break;
    case HTG_STORAGE_FP32_FP16:
        product<float, __half, float>(h, d_x, d_y, mu, st, ev);
        break;
    case HTG_STORAGE_FP32_BF16:
        product<float, __hip_bfloat16, float>(h, d_x, d_y, mu, st, ev);
        break;
    default:
        product<double, double, double>(h, d_x, d_y, mu, st, ev);
    }
}

void require_device() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        throw Error(HTG_ERR_HIP, "no HIP device available: libhtool_amd runs on the GPU only");
}

void sort_leaves(std::vector<Leaf> &v) {
    std::stable_sort(v.begin(), v.end(), [](const Leaf &a, const Leaf &b) { return a.oi == b.oi ? a.oj < b.oj : a.oi < b.oi; });
}

} // namespace
} // namespace htg

namespace htg {
namespace {

// Dense tiles from the generator and low-rank tiles from the compressed store,
// into the FP64 master arenas.
void fill_from_assembly(htg_hmatrix &h, const Part &part) {
    const Assembly &A = *part.A;
    // the generator always sees (target point, source point) of the assembly:
    // tiles of a transposed part evaluate G(x_t[col], x_s[row]) (FillTile::tr),
    // which needs no symmetry of the kernel function
    const DevPoints rowsP = A.dT(), colsP = A.dS();
    std::vector<FillTile> fills;
    std::vector<LrPack> packs[2];
    std::vector<UTileDst> udst[2];
    for (size_t l = part.leaf_begin; l < part.leaf_end; l++) {
        const Leaf &L = h.leaves[l];
        if (L.k < 0) {
            // the generator sees the instance's own point numbering
            for (const auto &rt : h.leaf_tiles[l]) {
                const int rb = std::get<0>(rt);
                fills.push_back(FillTile{std::get<1>(rt), h.rb_row0[rb] - part.shift, row_block_rows(h, rb),
                                         L.oj - part.shift, L.n, row_stride(h, 0, rb), part.transposed ? 1 : 0});
            }
            continue;
        }
        for (int c = 0; c < 2; c++) {
            const int kc = terms(L, c);
            if (kc <= 0)
                continue;
            const int first = c == 0 ? 0 : L.k_hi;
            LrPack pk{};
            pk.src_u       = L.src_u + (uint64_t)L.m * first;
            pk.src_v       = L.src_v + (uint64_t)L.n * first;
            pk.dst_v       = h.cls[c].leaf_v_off[l] + (uint64_t)(L.oj - v_first(L, c));
            pk.M           = L.m;
            pk.N           = L.n;
            pk.k           = kc;
            pk.src_ldv     = L.n;
            pk.dst_ldv     = v_stride(L, c);
            pk.piece_first = (int)udst[c].size();
            pk.npieces     = (int)h.leaf_tiles[l].size();
            pk.conj        = L.mirror && h.symmetry == 'H' && h.cplx; // K^H of the kept leaf K
            for (const auto &rt : h.leaf_tiles[l]) {
                const int rb = std::get<0>(rt);
                udst[c].push_back(UTileDst{c == 0 ? std::get<1>(rt) : std::get<2>(rt), h.rb_row0[rb] - L.oi,
                                           row_block_rows(h, rb), row_stride(h, c, rb), 1});
            }
            packs[c].push_back(pk);
        }
    }
    DevArray<FillTile> d_fills;
    d_fills.upload(fills);
    if (A.delay_dense) // left zero (alloc_master) until the caller supplies the blocks
        fills.clear();
    if (fills.empty()) {
    } else if (h.cplx)
        launch_dense_fill_c(A.kd, rowsP, colsP, d_fills.p, (int)fills.size(),
                            reinterpret_cast<cdbl *>(h.cls[0].a64.p), nullptr);
    else
        launch_dense_fill(A.kd, rowsP, colsP, d_fills.p, (int)fills.size(), h.cls[0].a64.p, nullptr);
    HTG_HIP(hipGetLastError());
    for (int c = 0; c < 2; c++) {
        if (packs[c].empty())
            continue;
        DevArray<LrPack> d_packs;
        DevArray<UTileDst> d_udst;
        d_packs.upload(packs[c]);
        d_udst.upload(udst[c]);
        if (h.cplx)
            launch_lr_pack_c(d_packs.p, (int)packs[c].size(), d_udst.p, reinterpret_cast<const cdbl *>(A.store.p),
                             reinterpret_cast<cdbl *>(h.cls[c].a64.p), reinterpret_cast<cdbl *>(h.cls[c].v64.p),
                             nullptr);
        else
            launch_lr_pack(d_packs.p, (int)packs[c].size(), d_udst.p, A.store.p, h.cls[c].a64.p, h.cls[c].v64.p,
                       nullptr);
        HTG_HIP(hipGetLastError());
        HTG_HIP(hipDeviceSynchronize());
    }
    HTG_HIP(hipDeviceSynchronize());
}

template <typename TS>
void ensure_copy(DevArray<TS> &dst, const DevArray<double> &src, uint64_t n) {
    if (dst.p)
        return;
    dst.alloc(n);
    launch_convert<TS>(src.p, dst.p, n, nullptr);
    HTG_HIP(hipGetLastError());
}

// Converts an existing storage copy again from the FP64 master, in place.
template <typename TS>
void reconvert(DevArray<TS> &dst, const DevArray<double> &src) {
    if (!dst.p)
        return;
    launch_convert<TS>(src.p, dst.p, std::min<uint64_t>(dst.n, src.n), nullptr);
    HTG_HIP(hipGetLastError());
}

// Adds the storage variant's copies of both classes (converted from FP64).
void apply_storage(htg_hmatrix &h, int storage) {
    for (int c = 0; c < 2; c++) {
        ClassLayout &C = h.cls[c];
        const bool low = c == 1 && (storage == HTG_STORAGE_FP32_FP16 || storage == HTG_STORAGE_FP32_BF16);
        if (storage == HTG_STORAGE_FP64)
            continue;
        if (storage != HTG_STORAGE_FP32 && storage != HTG_STORAGE_FP32_FP16 && storage != HTG_STORAGE_FP32_BF16)
            throw Error(HTG_ERR_UNSUPPORTED, "unknown storage mode " + std::to_string(storage));
        if (!low) {
            ensure_copy(C.a32, C.a64, C.a64.n);
            ensure_copy(C.v32, C.v64, C.v64.n);
        } else if (storage == HTG_STORAGE_FP32_FP16) {
            ensure_copy(C.a16, C.a64, C.a64.n);
            ensure_copy(C.v16, C.v64, C.v64.n);
        } else {
            ensure_copy(C.ab16, C.a64, C.a64.n);
            ensure_copy(C.vb16, C.v64, C.v64.n);
        }
    }
    HTG_HIP(hipDeviceSynchronize());
    if (h.storage != storage)
        h.generation->fetch_add(1);
    h.storage = storage;
}

// (off_i, off_j, m, n) of the kept dense leaves in get_leaves order: the list
// the caller's blocks of set_dense_blocks follow.
std::vector<std::array<int, 4>> kept_dense_keys(const htg_hmatrix &h) {
    std::vector<std::array<int, 4>> k;
    for (const Leaf &L : h.leaves)
        if (!L.mirror && L.k < 0)
            k.push_back({L.oi, L.oj, L.m, L.n});
    return k;
}

// Writes the caller's dense leaves (h.user_dense) into the FP64 master arena:
// each kept dense leaf into its row-block tiles, and its transpose into the
// tiles of its mirrored copy (symmetric storage).
void write_user_dense(htg_hmatrix &h) {
    if (h.user_dense.empty())
        return;
    const int e = h.esz();
    ClassLayout &C = h.cls[0];
    std::vector<double> arena(C.a64.n);
    HTG_HIP(hipMemcpy(arena.data(), C.a64.p, arena.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::map<std::pair<int, int>, size_t> where; // (off_i, off_j) of a supplied dense block -> data offset
    {
        size_t pos = 0;
        for (const auto &key : h.user_keys) {
            where[{key[0], key[1]}] = pos;
            pos += (size_t)key[2] * key[3] * e;
        }
        if (pos != h.user_dense.size())
            throw Error(HTG_ERR_INVALID, "supplied dense blocks do not match their list");
    }
    for (size_t l = 0; l < h.leaves.size(); l++) {
        const Leaf &L = h.leaves[l];
        if (L.k >= 0)
            continue;
        const auto it = where.find(L.mirror ? std::make_pair(L.oj, L.oi) : std::make_pair(L.oi, L.oj));
        if (it == where.end())
            throw Error(HTG_ERR_INVALID, "dense leaf without supplied data");
        const double *d = h.user_dense.data() + it->second;
        for (const auto &rt : h.leaf_tiles[l]) {
            const int rb = std::get<0>(rt), R = row_block_rows(h, rb), i0 = h.rb_row0[rb] - L.oi;
            const size_t ld = row_stride(h, 0, rb);
            double *dst     = arena.data() + e * std::get<1>(rt);
            for (int i = 0; i < R; i++)
                for (int j = 0; j < L.n; j++) {
                    // kept leaf: column-major m x n; mirror: the kept n x m leaf's (j, i)
                    const size_t s = L.mirror ? (size_t)j + (size_t)L.n * (i0 + i) : (size_t)(i0 + i) + (size_t)L.m * j;
                    for (int u = 0; u < e; u++) // Hermitian storage: the mirror is the conjugate transpose
                        dst[e * (i * ld + j) + u] = (u == 1 && L.mirror && h.symmetry == 'H') ? -d[e * s + u] : d[e * s + u];
                }
        }
    }
    HTG_HIP(hipMemcpy(C.a64.p, arena.data(), arena.size() * sizeof(double), hipMemcpyHostToDevice));
}

// Default work decomposition for freshly assembled leaves. With a 16-bit class
// every leaf's terms are split between two V arenas (the leading ones and the
// trailing ones), so each class's panels stack fewer rows per leaf: row groups
// of 32 rows bring its chunks back to about the single-class size (measured
// at N = 1e5: 2.4 us faster single-vector products, 32-RHS unchanged; with one
// class 16 rows stay faster).
void pick_default_layout(htg_hmatrix &h) {
    h.lay = DEFAULT_LAYOUT;
    for (const Leaf &L : h.leaves)
        if (L.k > 0 && L.k_hi < L.k) {
            h.lay.vrows = 32;
            break;
        }
}

// Lays out the product for h.lay and fills it from the kept assemblies.
void relayout(htg_hmatrix &h) {
    if (h.parts.empty())
        throw Error(HTG_ERR_UNSUPPORTED,
                    "a layout change needs an H-matrix assembled by htg_hmatrix_build or htg_hmatrix_build_batch");
    for (ClassLayout &C : h.cls) {
        C.a64.reset(), C.v64.reset(), C.a32.reset(), C.v32.reset();
        C.a16.reset(), C.v16.reset(), C.ab16.reset(), C.vb16.reset();
    }
    build_layout(h);
    upload_layout(h);
    alloc_master(h);
    for (const Part &part : h.parts)
        fill_from_assembly(h, part);
    write_user_dense(h);
    apply_storage(h, h.storage);
}

void check_layout(const htg_layout &L) {
    if (L.rb_max < 16 || L.rb_max > 128 || L.vchunk_el < 256 || L.vchunk_el > (1 << 20) || L.vgroup_el < 64 ||
        L.vgroup_el > (1 << 20) || L.vgroup_tiles < 1 || L.vgroup_tiles > 1024 || L.vrows < 1 || L.vrows > 512)
        throw Error(HTG_ERR_INVALID, "layout out of range (rb_max 16..128, vchunk_el 256..2^20, vgroup_el "
                                     "64..2^20, vgroup_tiles 1..1024, vrows 1..512)");
}

// Leaves of part p: the contiguous range of the sorted list whose rows lie in
// [shift, shift + n).
void locate_part(htg_hmatrix &h, Part &p, int n) {
    auto lo = std::lower_bound(h.leaves.begin(), h.leaves.end(), p.shift,
                               [](const Leaf &L, int v) { return L.oi < v; });
    auto hi = std::lower_bound(h.leaves.begin(), h.leaves.end(), p.shift + n,
                               [](const Leaf &L, int v) { return L.oi < v; });
    p.leaf_begin = (size_t)(lo - h.leaves.begin());
    p.leaf_end   = (size_t)(hi - h.leaves.begin());
}

// Symmetric storage: the product also applies the transpose of every kept
// off-diagonal leaf (hmatrix.hpp:876-904). Each gets a mirrored copy that
// shares its data: dense entries are regenerated from the symmetric kernel,
// low-rank copies exchange U and V (V, k x n row-major, is the n x k
// column-major U of the transpose, and back).
void add_mirrors(std::vector<Leaf> &leaves, int first) {
    const size_t n = leaves.size();
    for (size_t l = (size_t)first; l < n; l++) {
        if (leaves[l].oi == leaves[l].oj)
            continue;
        Leaf M = leaves[l];
        std::swap(M.oi, M.oj);
        std::swap(M.m, M.n);
        std::swap(M.src_u, M.src_v);
        M.mirror = true;
        leaves.push_back(std::move(M));
    }
}

// The H-matrix of the leaves (kept and mirrored) of h whose columns meet
// [offset, offset + size], sharing h's assemblies and supplied dense blocks:
// what HMatrix::mvprod_subrhs applies (hmatrix.hpp:1457-1554).
htg_hmatrix &subrhs_view(htg_hmatrix &h, int offset, int size) {
    auto it = h.subrhs_views.find({offset, size});
    if (it != h.subrhs_views.end())
        return *it->second;
    if (h.parts.empty())
        throw Error(HTG_ERR_UNSUPPORTED, "mvprod_subrhs needs an H-matrix assembled by the library");
    std::unique_ptr<htg_hmatrix> v(new htg_hmatrix);
    v->nr = h.nr, v->nc = h.nc, v->cplx = h.cplx, v->lay = h.lay;
    v->perm_t = h.perm_t, v->perm_s = h.perm_s;
    v->symmetry = h.symmetry, v->uplo = h.uplo;
    v->user_dense = h.user_dense;
    v->user_keys  = h.user_keys;
    for (const Leaf &L : h.leaves)
        if (L.oj <= offset + size && offset <= L.oj + L.n)
            v->leaves.push_back(L);
    for (const Part &p : h.parts) {
        Part q = p;
        locate_part(*v, q, p.transposed ? p.A->nc : p.A->nr);
        v->parts.push_back(std::move(q));
    }
    build_layout(*v);
    upload_layout(*v);
    alloc_master(*v);
    for (const Part &part : v->parts)
        fill_from_assembly(*v, part);
    write_user_dense(*v);
    htg_hmatrix &ref = *v;
    h.subrhs_views[{offset, size}] = std::move(v);
    return ref;
}

void take_aca_stats(htg_hmatrix &h, Assembly &A) {
    h.false_positive += A.false_positive;
    h.aca_feat.insert(h.aca_feat.end(), A.aca_feat.begin(), A.aca_feat.end());
    h.aca_terms.insert(h.aca_terms.end(), A.aca_terms.begin(), A.aca_terms.end());
    h.aca_cap.insert(h.aca_cap.end(), A.aca_cap.begin(), A.aca_cap.end());
    h.aca_retry += A.aca_retry;
    h.aca_wide += A.aca_wide;
    h.aca_int_wrap += A.aca_int_wrap;
    h.aca_ws_entries += A.aca_ws_entries;
    h.aca_need_entries += A.aca_need_entries;
    h.aca_seconds += A.aca_seconds;
    std::vector<float>().swap(A.aca_feat);
    std::vector<int>().swap(A.aca_terms);
    std::vector<int>().swap(A.aca_cap);
}

} // namespace
} // namespace htg

namespace htg {
namespace {
// Hermitian storage applies the conjugate transpose of each kept block: the
// kernel family must be Hermitian (every real family is symmetric, so 'H' is
// 'S' there, as the reference's real HMatrix with 'H' is).
void check_hermitian(char symmetry, const htg_generator &g) {
    if (symmetry == 'H' && !kernel_is_hermitian(g.kernel))
        throw Error(HTG_ERR_UNSUPPORTED, "Hermitian storage needs a Hermitian kernel family (HTG_KERNEL_HERMITIAN_TEST "
                                         "or a real family)");
}

void validate_params(const htg_params &prm) {
    // symmetry and storage triangle as HMatrix's constructor checks them (hmatrix.hpp:116-121)
    const int sy = prm.symmetry == 0 ? 'N' : prm.symmetry, up = prm.uplo == 0 ? 'N' : prm.uplo;
    if (!((sy == 'N' || sy == 'S' || sy == 'H') && (up == 'N' || up == 'L' || up == 'U') &&
          ((sy == 'N' && up == 'N') || (sy != 'N' && up != 'N'))))
        throw Error(HTG_ERR_INVALID, "[Htool error] Invalid arguments to create HMatrix");
    if (prm.minclustersize <= 0)
        throw Error(HTG_ERR_INVALID, "[Htool error] MinClusterSize parameter cannot be zero");
    if (prm.maxblocksize < 1)
        throw Error(HTG_ERR_INVALID, "[Htool error] MaxBlockSize parameter cannot be zero");
    if (prm.compressor != HTG_PARTIAL_ACA && prm.compressor != HTG_SYM_PARTIAL_ACA)
        throw Error(HTG_ERR_UNSUPPORTED, "unknown compressor");
    if (prm.recompress_tol < 0 || prm.lowprec_tol < 0)
        throw Error(HTG_ERR_INVALID, "[Htool error] negative tolerance");
}
} // namespace
} // namespace htg

struct htg_graph {
    hipGraph_t graph    = nullptr;
    hipGraphExec_t exec = nullptr;
    std::shared_ptr<std::atomic<uint64_t>> generation; // the H-matrix's token
    uint64_t captured_at = 0;
    ~htg_graph() {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (graph)
            (void)hipGraphDestroy(graph);
    }
};

namespace htg {
namespace {

// Arguments of htg_aca_batch / htg_aca_batch_noalloc (host arrays; the
// meaning of each is in include/htool_amd.h).
struct AcaBatchArgs {
    const htg_generator *gen;
    int compressor;
    double epsilon;
    int reqrank, nt;
    const double *xt;
    int ns;
    const double *xs;
    int nblocks;
    const int *M, *N, *rows, *row_off, *cols, *col_off;
    const double *ctr_t, *ctr_s;
    const int *off_t, *off_s;
    int cap;
    int *ranks;
    double *U;
    const int64_t *u_off;
    double *V;
    const int64_t *v_off;
};

// Host side of one batched compression: each block's points gathered into
// contiguous stretches (the kernel's cluster order), the block records, and
// where each device array sits in one byte buffer (256-byte aligned parts).
struct AcaBatchPlan {
    std::vector<int> rp, cp;
    std::vector<AcaBlock> blocks;
    std::vector<uint64_t> vis_off;
    uint64_t ws = 0, vis = 0;
    size_t off_pt[4] = {}, off_ps[4] = {}, off_blocks = 0, off_visoff = 0, off_ranks = 0, off_ws = 0, off_vis = 0;
    size_t bytes = 0;
};

AcaBatchPlan plan_aca_batch(const AcaBatchArgs &a) {
    if (!a.gen || !a.xt || a.nblocks < 0 || !a.M || !a.N || !a.rows || !a.row_off || !a.cols || !a.col_off ||
        !a.ctr_t || !a.ctr_s || !a.off_t || !a.off_s || !a.ranks || !a.U || !a.u_off || !a.V || !a.v_off || a.cap <= 0)
        throw Error(HTG_ERR_INVALID, "[Htool error] Invalid arguments to htg_aca_batch");
    if (a.cap > CAP_MAX)
        throw Error(HTG_ERR_INVALID, "cap above " + std::to_string(CAP_MAX));
    if (kernel_is_complex(a.gen->kernel))
        throw Error(HTG_ERR_UNSUPPORTED, "htg_aca_batch compresses real kernel families only");
    const bool same = a.gen->same_points || a.xs == nullptr;
    AcaBatchPlan P;
    P.blocks.resize(a.nblocks);
    P.vis_off.resize(a.nblocks);
    for (int b = 0; b < a.nblocks; b++) {
        if (a.M[b] <= 0 || a.N[b] <= 0)
            throw Error(HTG_ERR_INVALID, "empty block");
        AcaBlock &B = P.blocks[b];
        B.off_t     = (int)P.rp.size();
        B.M         = a.M[b];
        B.off_s     = (int)P.cp.size();
        B.N         = a.N[b];
        for (int i = 0; i < a.M[b]; i++) {
            const int r = a.rows[a.row_off[b] + i];
            if (r < 0 || r >= a.nt)
                throw Error(HTG_ERR_INVALID, "row index out of range");
            P.rp.push_back(r);
        }
        for (int j = 0; j < a.N[b]; j++) {
            const int c = a.cols[a.col_off[b] + j];
            if (c < 0 || c >= (same ? a.nt : a.ns))
                throw Error(HTG_ERR_INVALID, "column index out of range");
            P.cp.push_back(c);
        }
        B.swap = (a.compressor == HTG_SYM_PARTIAL_ACA && a.off_t[b] < a.off_s[b]) ? 1 : 0;
        B.cap  = a.cap;
        B.mn   = aca_size_bound(B.M, B.N, false);
        const int n1 = B.swap ? B.N : B.M;
        B.ws1        = P.ws;
        B.ws2        = P.ws + (uint64_t)n1 * a.cap;
        P.ws += (uint64_t)(B.M + B.N) * a.cap;
        const double *c = B.swap ? a.ctr_s + 3 * b : a.ctr_t + 3 * b;
        B.ctr[0] = c[0], B.ctr[1] = c[1], B.ctr[2] = c[2];
        P.vis_off[b] = P.vis;
        P.vis += (uint64_t)(B.M + B.N);
    }
    size_t o = 0;
    auto take = [&](size_t n) {
        const size_t at = o;
        o               = (o + std::max<size_t>(n, 1) + 255) / 256 * 256;
        return at;
    };
    for (int d = 0; d < 3; d++)
        P.off_pt[d] = take(P.rp.size() * sizeof(double));
    P.off_pt[3] = take(P.rp.size() * sizeof(int));
    for (int d = 0; d < 3; d++)
        P.off_ps[d] = take(P.cp.size() * sizeof(double));
    P.off_ps[3]  = take(P.cp.size() * sizeof(int));
    P.off_blocks = take(P.blocks.size() * sizeof(AcaBlock));
    P.off_visoff = take(P.vis_off.size() * sizeof(uint64_t));
    P.off_ranks  = take((size_t)a.nblocks * sizeof(int));
    P.off_ws     = take(P.ws * sizeof(double));
    P.off_vis    = take(P.vis);
    P.bytes      = o;
    return P;
}

// Uploads the plan into `dev` (P.bytes of device memory), runs the batched ACA
// and copies ranks and factors out. Synchronous.
void run_aca_batch(const AcaBatchArgs &a, const AcaBatchPlan &P, unsigned char *dev) {
    require_device();
    const bool same    = a.gen->same_points || a.xs == nullptr;
    const double *xsrc = same ? a.xt : a.xs;
    auto put_points = [&](const double *pts, const std::vector<int> &idx, const size_t *off) {
        std::vector<double> c(idx.size());
        for (int d = 0; d < 3; d++) {
            for (size_t i = 0; i < idx.size(); i++)
                c[i] = pts[3 * (size_t)idx[i] + d];
            if (!c.empty())
                HTG_HIP(hipMemcpy(dev + off[d], c.data(), c.size() * sizeof(double), hipMemcpyHostToDevice));
        }
        if (!idx.empty())
            HTG_HIP(hipMemcpy(dev + off[3], idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice));
        return DevPoints{reinterpret_cast<const double *>(dev + off[0]), reinterpret_cast<const double *>(dev + off[1]),
                         reinterpret_cast<const double *>(dev + off[2]), reinterpret_cast<const int *>(dev + off[3])};
    };
    const DevPoints pT = put_points(a.xt, P.rp, P.off_pt), pS = put_points(xsrc, P.cp, P.off_ps);
    if (a.nblocks > 0) {
        HTG_HIP(hipMemcpy(dev + P.off_blocks, P.blocks.data(), P.blocks.size() * sizeof(AcaBlock),
                          hipMemcpyHostToDevice));
        HTG_HIP(hipMemcpy(dev + P.off_visoff, P.vis_off.data(), P.vis_off.size() * sizeof(uint64_t),
                          hipMemcpyHostToDevice));
    }
    const KernelDesc kd = kernel_desc(*a.gen, same);
    double *ws          = reinterpret_cast<double *>(dev + P.off_ws);
    int *d_ranks        = reinterpret_cast<int *>(dev + P.off_ranks);
    // the kernel clears each block's visited flags itself; the workspace is
    // written before it is read
    launch_aca(kd, pT, pS, reinterpret_cast<const AcaBlock *>(dev + P.off_blocks), a.nblocks, a.epsilon, a.reqrank,
               ws, dev + P.off_vis, reinterpret_cast<const uint64_t *>(dev + P.off_visoff), d_ranks, nullptr);
    HTG_HIP(hipGetLastError());
    if (a.nblocks == 0)
        return;
    HTG_HIP(hipMemcpy(a.ranks, d_ranks, a.nblocks * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<double> hws(P.ws);
    HTG_HIP(hipMemcpy(hws.data(), ws, P.ws * sizeof(double), hipMemcpyDeviceToHost));
    for (int b = 0; b < a.nblocks; b++) {
        const AcaBlock &B = P.blocks[b];
        if (a.ranks[b] == -3)
            throw Error(HTG_ERR_RANK_CAP, "block " + std::to_string(b) + " needs more than cap terms");
        const int k = a.ranks[b];
        if (k <= 0)
            continue;
        const double *su = hws.data() + (B.swap ? B.ws2 : B.ws1);
        const double *sv = hws.data() + (B.swap ? B.ws1 : B.ws2);
        std::copy(su, su + (size_t)B.M * k, a.U + a.u_off[b]);
        double *vo = a.V + a.v_off[b];
        for (int r = 0; r < k; r++)
            for (int j = 0; j < B.N; j++)
                vo[r + (size_t)k * j] = sv[(size_t)r * B.N + j];
    }
}

// Device scratch of htg_aca_batch_noalloc: one array per GPU the library runs
// on, held until the process ends. Host code never uses the array's name as an
// address: noalloc_scratch() returns this device's address, looked up once per
// device with hipGetSymbolAddress while that device is current.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex &noalloc_lock() {
    static std::mutex m;
    return m;
}

unsigned char *noalloc_scratch() {
    static std::mutex lock;
    static unsigned char *per_device[64] = {};
    int dev = 0;
    HTG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64)
        throw Error(HTG_ERR_UNSUPPORTED, "more than 64 devices");
    std::lock_guard<std::mutex> hold(lock);
    if (!per_device[dev]) {
        void *ptr = nullptr;
        HTG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr)
            throw Error(HTG_ERR_HIP, "no device address for the scratch array");
        per_device[dev] = static_cast<unsigned char *>(ptr);
    }
    return per_device[dev];
}

} // namespace
} // namespace htg

namespace htg {
namespace {

// A caller's cluster tree (htg_cluster_tree) as the library's, checked: a
// permutation of [0, n), the root over every point, sons consecutive and
// tiling their parent's range, the rank-0 clusters of depth 1 covering all.
ClusterTree tree_from(const htg_cluster_tree &c, int n, const char *which) {
    const std::string w = which;
    if (c.n != n || !c.perm || c.nnodes < 1 || !c.nodes_int || !c.nodes_dbl)
        throw Error(HTG_ERR_INVALID, "[Htool error] " + w + " cluster tree does not match the points");
    ClusterTree T;
    T.perm.assign(c.perm, c.perm + n);
    std::vector<char> seen(n, 0);
    for (int p : T.perm) {
        if (p < 0 || p >= n || seen[p])
            throw Error(HTG_ERR_INVALID, "[Htool error] " + w + " cluster permutation is not a permutation");
        seen[p] = 1;
    }
    T.nodes.resize(c.nnodes);
    for (int i = 0; i < c.nnodes; i++) {
        const int *v = c.nodes_int + 6 * i;
        ClusterNode &d = T.nodes[i];
        d.offset = v[0], d.size = v[1], d.depth = v[2], d.rank = v[3], d.nb_sons = v[4], d.first_son = v[5];
        for (int k = 0; k < 3; k++)
            d.ctr[k] = c.nodes_dbl[4 * i + k];
        d.rad = c.nodes_dbl[4 * i + 3];
    }
    auto bad = [&](const std::string &m) { throw Error(HTG_ERR_INVALID, "[Htool error] " + w + " cluster tree: " + m); };
    if (T.nodes[0].offset != 0 || T.nodes[0].size != n)
        bad("the root does not cover every point");
    long long rank0 = 0;
    for (int i = 0; i < c.nnodes; i++) {
        const ClusterNode &d = T.nodes[i];
        if (d.size <= 0 || d.offset < 0 || d.offset + d.size > n || d.nb_sons < 0)
            bad("node " + std::to_string(i) + " out of range");
        if (d.depth == 1 && d.rank == 0)
            rank0 += d.size;
        if (d.nb_sons == 0)
            continue;
        if (d.first_son <= i || d.first_son + d.nb_sons > c.nnodes)
            bad("node " + std::to_string(i) + " has sons outside the node list");
        int o = d.offset;
        for (int p = 0; p < d.nb_sons; p++) {
            const ClusterNode &sn = T.nodes[d.first_son + p];
            if (sn.offset != o || sn.depth != d.depth + 1)
                bad("the sons of node " + std::to_string(i) + " do not tile it");
            o += sn.size;
        }
        if (o != d.offset + d.size)
            bad("the sons of node " + std::to_string(i) + " do not tile it");
    }
    if (rank0 != n)
        bad("its rank-0 clusters do not cover every point (a tree of one process is needed)");
    return T;
}

// The caller's dense leaves from the host generator, in get_leaves order.
void host_dense_blocks(htg_hmatrix &h, const htg_host_generator &g) {
    const int e = h.esz();
    h.user_dense.clear();
    for (const Leaf &L : h.leaves) {
        if (L.mirror || L.k >= 0)
            continue;
        const size_t at = h.user_dense.size();
        h.user_dense.resize(at + (size_t)L.m * L.n * e);
        g.copy_submatrix(L.m, L.n, h.perm_t.data() + L.oi, h.perm_s.data() + L.oj, h.user_dense.data() + at, g.ctx);
    }
    h.user_keys = kept_dense_keys(h);
}

// Calls f(i, j, v) with v the e doubles of entry (i, j) (cluster ordering) of
// every leaf, kept and mirrored, whose rows and columns meet `want(L)`, from
// the FP64 master arenas; low-rank entries are formed from U, the scales and V.
template <typename Want, typename F>
void for_leaf_entries(const htg_hmatrix &h, Want want, F f) {
    const int e = h.esz();
    std::vector<double> arena[2], varena[2];
    for (int c = 0; c < 2; c++) {
        arena[c].resize(h.cls[c].a64.n);
        varena[c].resize(h.cls[c].v64.n);
        HTG_HIP(hipMemcpy(arena[c].data(), h.cls[c].a64.p, arena[c].size() * sizeof(double), hipMemcpyDeviceToHost));
        HTG_HIP(hipMemcpy(varena[c].data(), h.cls[c].v64.p, varena[c].size() * sizeof(double),
                          hipMemcpyDeviceToHost));
    }
    double v[2];
    for (size_t l = 0; l < h.leaves.size(); l++) {
        const Leaf &L = h.leaves[l];
        if (L.k == 0 || !want(L))
            continue;
        for (const auto &rt : h.leaf_tiles[l]) {
            const int rb = std::get<0>(rt), R = row_block_rows(h, rb), i0 = h.rb_row0[rb] - L.oi;
            for (int i = 0; i < R; i++) {
                const int gi = L.oi + i0 + i;
                for (int j = 0; j < L.n; j++) {
                    if (!want.entry(gi, L.oj + j))
                        continue;
                    if (L.k < 0) {
                        const double *src = arena[0].data() + e * (std::get<1>(rt) + (size_t)i * row_stride(h, 0, rb) + j);
                        for (int u = 0; u < e; u++)
                            v[u] = src[u];
                    } else {
                        v[0] = v[1] = 0;
                        for (int r = 0; r < L.k; r++) {
                            const int c = r < L.k_hi ? 0 : 1, rc = c == 0 ? r : r - L.k_hi;
                            const double *us = arena[c].data() + e * ((c == 0 ? std::get<1>(rt) : std::get<2>(rt)) +
                                                                      (size_t)i * row_stride(h, c, rb) + rc);
                            const double sc  = L.sig.empty() ? 1.0 : L.sig[r];
                            const double *vs = varena[c].data() + e * (h.cls[c].leaf_v_off[l] + (L.oj - v_first(L, c)) +
                                                                       (size_t)rc * v_stride(L, c) + j);
                            if (e == 1) {
                                v[0] += us[0] * sc * vs[0];
                            } else {
                                const double ar = us[0] * sc, ai = us[1] * sc;
                                v[0] += ar * vs[0] - ai * vs[1];
                                v[1] += ar * vs[1] + ai * vs[0];
                            }
                        }
                    }
                    f(gi, L.oj + j, v);
                }
            }
        }
    }
}

struct AllEntries {
    bool operator()(const Leaf &) const { return true; }
    bool entry(int, int) const { return true; }
};
struct DiagonalEntries {
    bool operator()(const Leaf &L) const { return L.oi < L.oj + L.n && L.oj < L.oi + L.m; }
    bool entry(int i, int j) const { return i == j; }
};

// HMatrix::build(A, xt, xs) for htg_hmatrix_build and htg_hmatrix_build_ex.
void build_one(const htg_params &prm, const htg_generator *gen, const htg_host_generator *hgen, int nt,
               const double *xt, int ns, const double *xs, const htg_cluster_tree *tree_t,
               const htg_cluster_tree *tree_s, htg_hmatrix **out) {
    validate_params(prm);
    if (gen)
        check_hermitian(sym_kind(prm), *gen);
    const bool same = (gen && gen->same_points) || xs == nullptr;
    if (sym_uplo(prm) != 'N' && !same)
        throw Error(HTG_ERR_INVALID, "[Htool error] symmetric storage needs a single point set");
    if (same && ns != nt && ns > 0)
        throw Error(HTG_ERR_INVALID, "[Htool error] Invalid size in arguments for building HMatrix");
    if (!same && ns <= 0)
        throw Error(HTG_ERR_INVALID, "[Htool error] Invalid size in arguments for building HMatrix");
    if (prm.no_permutation && !tree_t)
        throw Error(HTG_ERR_INVALID, "[Htool error] no_permutation needs the caller's cluster trees (their "
                                     "ordering is the one of the points and vectors)");
    ClusterTree Tt, Ts;
    if (tree_t)
        Tt = tree_from(*tree_t, nt, "target");
    if (tree_s && !same)
        Ts = tree_from(*tree_s, ns, "source");
    if (prm.no_permutation && !same && !tree_s)
        throw Error(HTG_ERR_INVALID, "[Htool error] no_permutation needs the source cluster tree too");
    require_device();
    const auto t0 = std::chrono::steady_clock::now();
    std::unique_ptr<htg_hmatrix> h(new htg_hmatrix);
    Part part;
    part.A.reset(new Assembly);
    Assembly &A = *part.A;
    assemble_ex(prm, gen, hgen, same, nt, xt, ns, xs, tree_t ? &Tt : nullptr, tree_s && !same ? &Ts : nullptr, A);
    h->nr       = A.nr;
    h->nc       = A.nc;
    h->cplx     = A.cplx;
    h->perm_t   = A.perm_t;
    h->perm_s   = A.perm_s;
    h->leaves   = A.leaves;
    h->uplo     = sym_uplo(prm);
    h->symmetry = sym_kind(prm);
    if (h->uplo != 'N')
        add_mirrors(h->leaves, 0);
    take_aca_stats(*h, A);
    sort_leaves(h->leaves);
    part.leaf_begin = 0;
    part.leaf_end   = h->leaves.size();
    h->parts.push_back(std::move(part));
    pick_default_layout(*h);
    build_layout(*h);
    upload_layout(*h);
    alloc_master(*h);
    fill_from_assembly(*h, h->parts[0]);
    if (hgen && !prm.delay_dense) {
        host_dense_blocks(*h, *hgen);
        write_user_dense(*h);
    }
    h->build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    *out             = h.release();
}

} // namespace
} // namespace htg

extern "C" {

const char *htg_last_error(void) { return htg::last_error().c_str(); }

const char *htg_status_string(int status) {
    switch (status) {
    case HTG_OK: return "ok";
    case HTG_ERR_INVALID: return "invalid argument";
    case HTG_ERR_HIP: return "HIP error";
    case HTG_ERR_OOM: return "out of memory";
    case HTG_ERR_UNSUPPORTED: return "unsupported";
    case HTG_ERR_RANK_CAP: return "ACA rank cap exceeded";
    default: return "unknown status";
    }
}

int htg_hmatrix_build(const htg_params *prm, const htg_generator *gen, int nt, const double *xt, int ns,
                      const double *xs, htg_hmatrix **out) {
    return guard([&] {
        if (!prm || !gen || !xt || !out || nt <= 0)
            throw Error(HTG_ERR_INVALID, "[Htool error] Invalid arguments to build HMatrix");
        build_one(*prm, gen, nullptr, nt, xt, ns, xs, nullptr, nullptr, out);
    });
}

int htg_hmatrix_build_ex(const htg_params *prm, const htg_generator *gen, const htg_host_generator *host_gen,
                         int nt, const double *xt, int ns, const double *xs, const htg_cluster_tree *tree_t,
                         const htg_cluster_tree *tree_s, htg_hmatrix **out) {
    return guard([&] {
        if (!prm || (!gen == !host_gen) || !xt || !out || nt <= 0)
            throw Error(HTG_ERR_INVALID, "[Htool error] Invalid arguments to build HMatrix (one generator, points)");
        if (host_gen && !host_gen->copy_submatrix)
            throw Error(HTG_ERR_INVALID, "[Htool error] host generator without copy_submatrix");
        build_one(*prm, gen, host_gen, nt, xt, ns, xs, tree_t, tree_s, out);
    });
}

int htg_hmatrix_from_leaves(int nr, int nc, const int *perm_t, const int *perm_s, int nleaves, const int *leaves5,
                            const double *data, htg_hmatrix **out) {
    return guard([&] {
        if (nr <= 0 || nc <= 0 || !perm_t || !perm_s || nleaves < 0 || (nleaves && (!leaves5 || !data)) || !out)
            throw Error(HTG_ERR_INVALID, "[Htool error] Invalid arguments to htg_hmatrix_from_leaves");
        for (int l = 0; l < nleaves; l++) {
            const int *L = leaves5 + 5 * l;
            if (L[1] <= 0 || L[3] <= 0 || L[0] < 0 || L[2] < 0 || L[0] + L[1] > nr || L[2] + L[3] > nc || L[4] < -1)
                throw Error(HTG_ERR_INVALID, "[Htool error] leaf " + std::to_string(l) + " out of range");
        }
        require_device();
        std::unique_ptr<htg_hmatrix> h(new htg_hmatrix);
        h->nr = nr;
        h->nc = nc;
        h->perm_t.assign(perm_t, perm_t + nr);
        h->perm_s.assign(perm_s, perm_s + nc);
        std::vector<Leaf> leaves(nleaves);
        const double *p = data;
        for (int l = 0; l < nleaves; l++) {
            const int *L = leaves5 + 5 * l;
            leaves[l]    = Leaf{L[0], L[1], L[2], L[3], L[4]};
            leaves[l].k_hi  = L[4] > 0 ? L[4] : 0;
            leaves[l].src_u = (uint64_t)(p - data); // offset of this leaf's data
            p += L[4] < 0 ? (size_t)L[1] * L[3] : (size_t)L[4] * (L[1] + L[3]);
        }
        sort_leaves(leaves);
        h->leaves = leaves;
        build_layout(*h);
        ClassLayout &C = h->cls[0];
        std::vector<double> arena(C.arena_size, 0.0), varena(C.varena_size, 0.0);
        for (size_t l = 0; l < h->leaves.size(); l++) {
            const Leaf &L  = h->leaves[l];
            const double *d = data + L.src_u;
            if (L.k == 0)
                continue;
            for (const auto &rt : h->leaf_tiles[l]) {
                const int rb      = std::get<0>(rt);
                const int R       = row_block_rows(*h, rb);
                const int i0      = h->rb_row0[rb] - L.oi;
                const size_t ld   = row_stride(*h, 0, rb);
                double *dst       = arena.data() + std::get<1>(rt);
                const int nc_tile = L.k < 0 ? L.n : L.k; // dense: A columns; low rank: U columns
                for (int j = 0; j < nc_tile; j++)
                    for (int i = 0; i < R; i++)
                        dst[i * ld + j] = d[(size_t)(i0 + i) + (size_t)L.m * j];
            }
            if (L.k > 0) { // V: column-major k x n in, row-major aligned rows on the device
                const double *V = d + (size_t)L.m * L.k;
                double *dst     = varena.data() + C.leaf_v_off[l] + (L.oj - v_first(L, 0));
                const size_t np = v_stride(L, 0);
                for (int j = 0; j < L.n; j++)
                    for (int r = 0; r < L.k; r++)
                        dst[r * np + j] = V[r + (size_t)L.k * j];
            }
        }
        for (Leaf &L : h->leaves)
            L.src_u = 0;
        upload_layout(*h);
        C.a64.upload(arena);
        C.v64.upload(varena);
        h->cls[1].a64.alloc(1);
        h->cls[1].v64.alloc(1);
        HTG_HIP(hipDeviceSynchronize());
        *out = h.release();
    });
}

int htg_hmatrix_destroy(htg_hmatrix *h) {
    return guard([&] { delete h; });
}

int htg_hmatrix_set_storage(htg_hmatrix *h, int storage) {
    return guard([&] {
        if (!h)
            throw Error(HTG_ERR_INVALID, "null handle");
        apply_storage(*h, storage);
    });
}

int htg_hmatrix_build_batch(const htg_params *prm, int count, const htg_generator *gens, const int *n,
                            const double *const *pts, htg_hmatrix **out) {
    return guard([&] {
        if (!prm || count <= 0 || !gens || !n || !pts || !out)
            throw Error(HTG_ERR_INVALID, "[Htool error] Invalid arguments for building a batch of HMatrix");
        long long total = 0;
        for (int i = 0; i < count; i++) {
            if (n[i] <= 0 || !pts[i])
                throw Error(HTG_ERR_INVALID, "[Htool error] instance " + std::to_string(i) + " has no points");
            total += n[i];
        }
        if (total > (1LL << 30))
            throw Error(HTG_ERR_INVALID, "batch too large");
        validate_params(*prm);
        for (int i = 0; i < count; i++)
            check_hermitian(sym_kind(*prm), gens[i]);
        require_device();
        const auto t0 = std::chrono::steady_clock::now();
        std::unique_ptr<htg_hmatrix> h(new htg_hmatrix);
        int shift = 0;
        for (int i = 0; i < count; i++) {
            Part part;
            part.A.reset(new Assembly);
            htg_generator g = gens[i];
            g.same_points   = 1;
            assemble(*prm, g, n[i], pts[i], n[i], nullptr, *part.A);
            Assembly &A = *part.A;
            if (i == 0)
                h->cplx = A.cplx;
            else if (A.cplx != h->cplx)
                throw Error(HTG_ERR_INVALID, "a batch mixes real and complex kernel families");
            for (int v : A.perm_t)
                h->perm_t.push_back(v + shift);
            for (int v : A.perm_s)
                h->perm_s.push_back(v + shift);
            const size_t first = h->leaves.size();
            for (Leaf L : A.leaves) {
                L.oi += shift;
                L.oj += shift;
                h->leaves.push_back(L);
            }
            if (sym_uplo(*prm) != 'N')
                add_mirrors(h->leaves, (int)first);
            std::vector<Leaf>().swap(A.leaves);
            take_aca_stats(*h, A);
            part.shift = shift;
            h->parts.push_back(std::move(part));
            shift += n[i];
        }
        h->nr = h->nc = shift;
        h->uplo     = sym_uplo(*prm);
        h->symmetry = sym_kind(*prm);
        sort_leaves(h->leaves);
        for (int i = 0; i < count; i++)
            locate_part(*h, h->parts[i], n[i]);
        pick_default_layout(*h);
        build_layout(*h);
        upload_layout(*h);
        alloc_master(*h);
        for (const Part &part : h->parts)
            fill_from_assembly(*h, part);
        h->build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        *out             = h.release();
    });
}

int htg_hmatrix_transpose(const htg_hmatrix *h, htg_hmatrix **out) {
    return guard([&] {
        if (!h || !out)
            throw Error(HTG_ERR_INVALID, "null argument");
        if (h->parts.empty())
            throw Error(HTG_ERR_UNSUPPORTED, "the transpose needs an H-matrix assembled by htg_hmatrix_build or "
                                             "htg_hmatrix_build_batch");
        std::unique_ptr<htg_hmatrix> t(new htg_hmatrix);
        t->nr             = h->nc;
        t->nc             = h->nr;
        t->cplx           = h->cplx;
        t->lay            = h->lay;
        t->perm_t         = h->perm_s;
        t->perm_s         = h->perm_t;
        t->false_positive = h->false_positive;
        t->aca_int_wrap   = h->aca_int_wrap;
        t->symmetry       = h->symmetry;
        t->uplo           = h->uplo == 'L' ? 'U' : (h->uplo == 'U' ? 'L' : 'N');
        t->leaves.reserve(h->leaves.size());
        for (Leaf L : h->leaves) {
            std::swap(L.oi, L.oj);
            std::swap(L.m, L.n);
            std::swap(L.src_u, L.src_v); // V (k x n row-major) is U^T's column-major block, and back
            t->leaves.push_back(std::move(L));
        }
        sort_leaves(t->leaves);
        for (const Part &p : h->parts) {
            Part q;
            q.A          = p.A;
            q.shift      = p.shift;
            q.transposed = !p.transposed;
            locate_part(*t, q, q.transposed ? p.A->nc : p.A->nr);
            t->parts.push_back(std::move(q));
        }
        // Dense blocks the caller supplied to h (set_dense_blocks) become t's
        // supplied blocks: each kept dense leaf of t is the transpose of a kept
        // dense leaf of h (no conjugation), in t's get_leaves order.
        if (!h->user_dense.empty()) {
            const int e = h->esz();
            std::map<std::pair<int, int>, size_t> where; // (off_i, off_j) of h's kept dense leaf -> data offset
            size_t pos = 0;
            for (const Leaf &L : h->leaves) {
                if (L.mirror || L.k >= 0)
                    continue;
                where[{L.oi, L.oj}] = pos;
                pos += (size_t)L.m * L.n * e;
            }
            t->user_dense.reserve(pos);
            for (const Leaf &L : t->leaves) {
                if (L.mirror || L.k >= 0)
                    continue;
                const auto it = where.find({L.oj, L.oi});
                if (it == where.end())
                    throw Error(HTG_ERR_INVALID, "transposed dense leaf without supplied data");
                const double *d = h->user_dense.data() + it->second; // h's n x m leaf, column-major
                for (int j = 0; j < L.n; j++)
                    for (int i = 0; i < L.m; i++)
                        for (int u = 0; u < e; u++)
                            t->user_dense.push_back(d[e * ((size_t)j + (size_t)L.n * i) + u]);
            }
            t->user_keys = kept_dense_keys(*t);
        }
        build_layout(*t);
        upload_layout(*t);
        alloc_master(*t);
        for (const Part &part : t->parts)
            fill_from_assembly(*t, part);
        write_user_dense(*t);
        *out = t.release();
    });
}

int htg_hmatrix_get_layout(const htg_hmatrix *h, htg_layout *layout) {
    return guard([&] {
        if (!h || !layout)
            throw Error(HTG_ERR_INVALID, "null argument");
        *layout = h->lay;
    });
}

int htg_hmatrix_set_layout(htg_hmatrix *h, const htg_layout *layout) {
    return guard([&] {
        if (!h || !layout)
            throw Error(HTG_ERR_INVALID, "null argument");
        check_layout(*layout);
        const htg_layout old = h->lay;
        h->lay               = *layout;
        try {
            relayout(*h);
        } catch (...) {
            h->lay = old;
            throw;
        }
    });
}

int htg_hmatrix_autotune(htg_hmatrix *h, int mu, int reps, htg_layout *best, double *best_ms, int *ncandidates) {
    return guard([&] {
        if (!h || mu <= 0 || reps <= 0)
            throw Error(HTG_ERR_INVALID, "invalid autotune arguments");
        if (h->parts.empty())
            throw Error(HTG_ERR_UNSUPPORTED, "autotuning needs an H-matrix assembled by htg_hmatrix_build or "
                                             "htg_hmatrix_build_batch");
        DevArray<double> dx, dy;
        dx.alloc((size_t)h->nc * mu * h->esz());
        dy.alloc((size_t)h->nr * mu * h->esz());
        dx.zero();
        hipEvent_t e0, e1;
        HTG_HIP(hipEventCreate(&e0));
        HTG_HIP(hipEventCreate(&e1));
        struct EvGuard {
            hipEvent_t a, b;
            ~EvGuard() {
                (void)hipEventDestroy(a);
                (void)hipEventDestroy(b);
            }
        } eg{e0, e1};
        const htg_layout start = h->lay;
        htg_layout top         = start;
        double top_ms          = 1e300;
        int tried              = 0;
        // the current layout, then rows per row block x (V chunk, wavefront
        // share) x stacked V rows
        std::vector<htg_layout> cand{start};
        for (int rb : {64, 128})
            for (auto cg : {std::pair<int, int>{8192, 2048}, {8192, 4096}, {16384, 4096}, {16384, 16384}})
                for (int vrows : {8, 16, 32})
                    cand.push_back(htg_layout{rb, cg.first, cg.second, 4, vrows});
        auto time_layout = [&](const htg_layout &L, int n) {
            h->lay = L;
            relayout(*h);
            run_product(*h, dx.p, dy.p, mu, nullptr); // warm-up
            HTG_HIP(hipEventRecord(e0, nullptr));
            for (int r = 0; r < n; r++)
                run_product(*h, dx.p, dy.p, mu, nullptr);
            HTG_HIP(hipEventRecord(e1, nullptr));
            HTG_HIP(hipEventSynchronize(e1));
            float ms = 0;
            HTG_HIP(hipEventElapsedTime(&ms, e0, e1));
            return (double)ms / n;
        };
        std::vector<std::pair<double, int>> first;
        for (size_t c = 0; c < cand.size(); c++) {
            first.emplace_back(time_layout(cand[c], reps), (int)c);
            tried++;
        }
        // the fastest of one pass over many candidates is biased low (the
        // least favourable noise wins): the three fastest are timed again,
        // with twice the repetitions, and the second timing decides and is
        // the one reported
        std::stable_sort(first.begin(), first.end());
        for (size_t i = 0; i < first.size() && i < 3; i++) {
            const double ms = time_layout(cand[first[i].second], 2 * reps);
            if (ms < top_ms) {
                top_ms = ms;
                top    = cand[first[i].second];
            }
        }
        h->lay = top;
        relayout(*h);
        if (best)
            *best = top;
        if (best_ms)
            *best_ms = top_ms;
        if (ncandidates)
            *ncandidates = tried;
    });
}

int htg_graph_capture(htg_hmatrix *h, const double *d_x, double *d_y, int mu, void *stream, htg_graph **out) {
    return guard([&] {
        if (!h || !d_x || !d_y || mu <= 0 || !stream || !out)
            throw Error(HTG_ERR_INVALID, "invalid graph capture arguments (a stream is required)");
        hipStream_t st = static_cast<hipStream_t>(stream);
        // one eager product first: every lazily allocated buffer exists before capture
        run_product(*h, d_x, d_y, mu, st);
        HTG_HIP(hipStreamSynchronize(st));
        std::unique_ptr<htg_graph> g(new htg_graph);
        HTG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            run_product(*h, d_x, d_y, mu, st);
        } catch (...) {
            hipGraph_t dummy = nullptr;
            (void)hipStreamEndCapture(st, &dummy);
            if (dummy)
                (void)hipGraphDestroy(dummy);
            throw;
        }
        HTG_HIP(hipStreamEndCapture(st, &g->graph));
        HTG_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
        g->generation  = h->generation;
        g->captured_at = h->generation->load();
        *out           = g.release();
    });
}

int htg_graph_launch(htg_graph *g, void *stream) {
    return guard([&] {
        if (!g || !g->exec)
            throw Error(HTG_ERR_INVALID, "null graph");
        if (!g->generation || g->generation->load() != g->captured_at)
            throw Error(HTG_ERR_INVALID, "stale graph: the H-matrix was destroyed, or its layout, storage or "
                                         "buffers changed after the capture (capture again)");
        HTG_HIP(hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)));
    });
}

int htg_graph_destroy(htg_graph *g) {
    return guard([&] { delete g; });
}

int htg_hmatrix_get_info(const htg_hmatrix *h, htg_info *info) {
    return guard([&] {
        if (!h || !info)
            throw Error(HTG_ERR_INVALID, "null argument");
        htg_info r{};
        r.nr = h->nr, r.nc = h->nc;
        r.n_false_positive = h->false_positive;
        long long ksum = 0;
        bool first     = true;
        for (const Leaf &L : h->leaves) {
            if (L.mirror) {
                r.mirrored_entries += L.k < 0 ? (int64_t)L.m * L.n : (int64_t)L.k * (L.m + L.n);
                if (L.k > 0)
                    r.mirrored_lowprec_entries += (int64_t)(L.k - L.k_hi) * (L.m + L.n);
                continue;
            }
            if (L.k < 0) {
                r.n_dense++;
                r.dense_entries += (int64_t)L.m * L.n;
            } else {
                r.n_lowrank++;
                r.lowrank_entries += (int64_t)L.k * (L.m + L.n);
                r.lowprec_entries += (int64_t)(L.k - L.k_hi) * (L.m + L.n);
                r.lowrank_u_entries += (int64_t)L.k * L.m;
                r.lowprec_u_entries += (int64_t)(L.k - L.k_hi) * L.m;
                ksum += L.k;
                r.rank_min = first ? L.k : std::min(r.rank_min, L.k);
                r.rank_max = first ? L.k : std::max(r.rank_max, L.k);
                first      = false;
            }
        }
        r.rank_mean         = r.n_lowrank ? double(ksum) / r.n_lowrank : 0.0;
        const double stored = double(r.dense_entries) + double(r.lowrank_entries);
        r.compression_ratio = stored > 0 ? double(h->nr) * double(h->nc) / stored : 0.0;
        r.n_row_blocks      = (int)h->rb_row0.size() - 1;
        r.n_columns         = (int)(h->cls[0].colidx.size() + h->cls[1].colidx.size());
        r.n_vtiles          = (int)h->vt[0].tiles.size();
        r.build_seconds     = h->build_seconds;
        r.n_aca             = (int)h->aca_terms.size();
        r.n_aca_retry       = h->aca_retry;
        r.aca_ws_entries    = h->aca_ws_entries;
        r.aca_need_entries  = h->aca_need_entries;
        r.aca_seconds       = h->aca_seconds;
        r.is_complex        = h->cplx ? 1 : 0;
        r.n_aca_wide        = h->aca_wide;
        r.symmetry          = h->symmetry;
        r.uplo              = h->uplo;
        r.n_aca_int_wrap    = h->aca_int_wrap;
        r.nb_mat_vec_prod   = h->n_products;
        r.total_time_mat_vec_prod = h->product_seconds;
        *info               = r;
    });
}

int htg_hmatrix_aca_records(const htg_hmatrix *h, int *n, float *feat, int *terms, int *cap_first) {
    return guard([&] {
        if (!h || !n)
            throw Error(HTG_ERR_INVALID, "null argument");
        const int cnt = (int)h->aca_terms.size();
        if (feat) {
            if (*n < cnt)
                throw Error(HTG_ERR_INVALID, "record buffers hold fewer than the build's ACA runs");
            std::memcpy(feat, h->aca_feat.data(), h->aca_feat.size() * sizeof(float));
            if (terms)
                std::memcpy(terms, h->aca_terms.data(), (size_t)cnt * sizeof(int));
            if (cap_first)
                std::memcpy(cap_first, h->aca_cap.data(), (size_t)cnt * sizeof(int));
        }
        *n = cnt;
    });
}

int htg_hmatrix_nleaves(const htg_hmatrix *h, int *nleaves, int64_t *data_size) {
    return guard([&] {
        if (!h)
            throw Error(HTG_ERR_INVALID, "null handle");
        int64_t s = 0;
        int cnt   = 0;
        for (const Leaf &L : h->leaves) {
            if (L.mirror)
                continue;
            s += L.k < 0 ? (int64_t)L.m * L.n : (int64_t)L.k * (L.m + L.n);
            cnt++;
        }
        if (nleaves)
            *nleaves = cnt;
        if (data_size)
            *data_size = s;
    });
}

int htg_hmatrix_get_leaves(const htg_hmatrix *h, int *leaves5, double *data) {
    return guard([&] {
        if (!h || !leaves5)
            throw Error(HTG_ERR_INVALID, "null argument");
        size_t o = 0;
        for (const Leaf &L : h->leaves) {
            if (L.mirror)
                continue;
            int v[5] = {L.oi, L.m, L.oj, L.n, L.k};
            std::memcpy(leaves5 + 5 * o++, v, sizeof(v));
        }
        if (!data)
            return;
        std::vector<double> arena[2], varena[2];
        for (int c = 0; c < 2; c++) {
            arena[c].resize(h->cls[c].a64.n);
            varena[c].resize(h->cls[c].v64.n);
            HTG_HIP(hipMemcpy(arena[c].data(), h->cls[c].a64.p, arena[c].size() * sizeof(double),
                              hipMemcpyDeviceToHost));
            HTG_HIP(hipMemcpy(varena[c].data(), h->cls[c].v64.p, varena[c].size() * sizeof(double),
                              hipMemcpyDeviceToHost));
        }
        const int e = h->esz(); // doubles per entry
        double *out = data;
        auto put    = [&](double *dst, const double *src, double sc) {
            for (int u = 0; u < e; u++)
                dst[u] = src[u] * sc;
        };
        for (size_t l = 0; l < h->leaves.size(); l++) {
            const Leaf &L = h->leaves[l];
            if (L.k == 0 || L.mirror)
                continue;
            const int ncol = L.k < 0 ? L.n : L.k;
            // U (or the dense block): column j lives in class 0 (j < k_hi, or dense) or class 1.
            for (const auto &rt : h->leaf_tiles[l]) {
                const int rb = std::get<0>(rt);
                const int R  = row_block_rows(*h, rb);
                const int i0 = h->rb_row0[rb] - L.oi;
                for (int j = 0; j < ncol; j++) {
                    const int c       = (L.k < 0 || j < L.k_hi) ? 0 : 1;
                    const int jc      = c == 0 ? j : j - L.k_hi;
                    const size_t ld   = row_stride(*h, c, rb);
                    const double *src = arena[c].data() + e * (c == 0 ? std::get<1>(rt) : std::get<2>(rt));
                    const double sc   = (L.k > 0 && !L.sig.empty()) ? L.sig[j] : 1.0;
                    for (int i = 0; i < R; i++)
                        put(out + e * ((size_t)(i0 + i) + (size_t)L.m * j), src + e * (i * ld + jc), sc);
                }
            }
            out += e * (size_t)L.m * ncol;
            if (L.k > 0) {
                for (int r = 0; r < L.k; r++) {
                    const int c       = r < L.k_hi ? 0 : 1;
                    const int rc      = c == 0 ? r : r - L.k_hi;
                    const size_t np   = v_stride(L, c);
                    const double *src = varena[c].data() +
                                        e * (h->cls[c].leaf_v_off[l] + (L.oj - v_first(L, c)) + (size_t)rc * np);
                    for (int j = 0; j < L.n; j++)
                        put(out + e * (r + (size_t)L.k * j), src + e * j, 1.0);
                }
                out += e * (size_t)L.k * L.n;
            }
        }
    });
}

int htg_hmatrix_set_dense_blocks(htg_hmatrix *h, const double *data) {
    return guard([&] {
        if (!h || !data)
            throw Error(HTG_ERR_INVALID, "null argument");
        size_t n = 0;
        for (const Leaf &L : h->leaves)
            if (!L.mirror && L.k < 0)
                n += (size_t)L.m * L.n * h->esz();
        h->user_dense.assign(data, data + n);
        h->user_keys = kept_dense_keys(*h);
        h->subrhs_views.clear();
        write_user_dense(*h);
        // storage copies are converted again from the new FP64 entries, in
        // place: captured graphs keep valid pointers and see the new blocks
        for (ClassLayout &C : h->cls) {
            reconvert(C.a32, C.a64);
            reconvert(C.a16, C.a64);
            reconvert(C.ab16, C.a64);
        }
        HTG_HIP(hipDeviceSynchronize());
    });
}

int htg_cluster_build(int n, const double *x, int minclustersize, int *perm, int *nnodes, int *nodes_int,
                      double *nodes_dbl) {
    return guard([&] {
        if (n <= 0 || !x || !nnodes)
            throw Error(HTG_ERR_INVALID, "[Htool error] Invalid arguments to build a cluster tree");
        if (minclustersize <= 0)
            throw Error(HTG_ERR_INVALID, "[Htool error] MinClusterSize parameter cannot be zero");
        ClusterTree T;
        build_cluster_tree(n, x, minclustersize, T);
        if (perm)
            std::copy(T.perm.begin(), T.perm.end(), perm);
        if (nodes_int) {
            if (*nnodes < (int)T.nodes.size())
                throw Error(HTG_ERR_INVALID, "node arrays too small");
            for (size_t i = 0; i < T.nodes.size(); i++) {
                const ClusterNode &c = T.nodes[i];
                const int v[6]       = {c.offset, c.size, c.depth, c.rank, c.nb_sons, c.first_son};
                std::copy(v, v + 6, nodes_int + 6 * i);
                if (nodes_dbl) {
                    const double d[4] = {c.ctr[0], c.ctr[1], c.ctr[2], c.rad};
                    std::copy(d, d + 4, nodes_dbl + 4 * i);
                }
            }
        }
        *nnodes = (int)T.nodes.size();
    });
}

int htg_hmatrix_get_perms(const htg_hmatrix *h, int *perm_t, int *perm_s) {
    return guard([&] {
        if (!h)
            throw Error(HTG_ERR_INVALID, "null handle");
        if (perm_t)
            std::copy(h->perm_t.begin(), h->perm_t.end(), perm_t);
        if (perm_s)
            std::copy(h->perm_s.begin(), h->perm_s.end(), perm_s);
    });
}

int htg_hmatrix_mvprod_device(htg_hmatrix *h, const double *d_x, double *d_y, int mu, void *stream) {
    return guard([&] {
        if (!h || !d_x || !d_y)
            throw Error(HTG_ERR_INVALID, "null argument");
        run_product(*h, d_x, d_y, mu, static_cast<hipStream_t>(stream));
        h->n_products++;
    });
}

int htg_hmatrix_mvprod(htg_hmatrix *h, const double *x, double *y, int mu) {
    return guard([&] {
        if (!h || !x || !y || mu <= 0)
            throw Error(HTG_ERR_INVALID, "null argument");
        const auto t0 = std::chrono::steady_clock::now();
        const size_t nx = (size_t)h->nc * mu * h->esz(), ny = (size_t)h->nr * mu * h->esz();
        if (h->d_hx.n < nx)
            h->d_hx.alloc(nx);
        if (h->d_hy.n < ny)
            h->d_hy.alloc(ny);
        HTG_HIP(hipMemcpy(h->d_hx.p, x, nx * sizeof(double), hipMemcpyHostToDevice));
        run_product(*h, h->d_hx.p, h->d_hy.p, mu, nullptr);
        HTG_HIP(hipMemcpy(y, h->d_hy.p, ny * sizeof(double), hipMemcpyDeviceToHost));
        h->n_products++;
        h->product_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    });
}

int htg_hmatrix_mymvprod_local_to_local_device(htg_hmatrix *h, const double *d_in, double *d_out, int mu,
                                               double *d_work, void *stream) {
    return guard([&] {
        if (!h || !d_in || !d_out || mu <= 0)
            throw Error(HTG_ERR_INVALID, "null argument");
        hipStream_t st  = static_cast<hipStream_t>(stream);
        const int e     = h->esz();
        const size_t nx = (size_t)h->nc * mu * e, ny = (size_t)h->nr * mu * e;
        DevArray<double> own; // the reference allocates `work` when the caller passes none (hmatrix.hpp:981-985)
        if (!d_work) {
            own.alloc(nx + ny);
            d_work = own.p;
        }
        launch_to_original(d_in, h->d_perm_s.p, h->nc, mu, e, d_work, st);
        run_product(*h, d_work, d_work + nx, mu, st);
        launch_to_cluster(d_work + nx, h->d_perm_t.p, h->nr, mu, e, d_out, st);
        HTG_HIP(hipGetLastError());
        if (own.p)
            HTG_HIP(hipStreamSynchronize(st)); // before the per-call workspace is freed
        h->n_products++;
    });
}

int htg_hmatrix_mymvprod_local_to_local(htg_hmatrix *h, const double *in, double *out, int mu) {
    return guard([&] {
        if (!h || !in || !out || mu <= 0)
            throw Error(HTG_ERR_INVALID, "null argument");
        const auto t0   = std::chrono::steady_clock::now();
        const int e     = h->esz();
        const size_t nx = (size_t)h->nc * mu * e, ny = (size_t)h->nr * mu * e;
        if (h->d_hx.n < 2 * nx)
            h->d_hx.alloc(2 * nx);
        if (h->d_hy.n < 2 * ny)
            h->d_hy.alloc(2 * ny);
        HTG_HIP(hipMemcpy(h->d_hx.p + nx, in, nx * sizeof(double), hipMemcpyHostToDevice));
        launch_to_original(h->d_hx.p + nx, h->d_perm_s.p, h->nc, mu, e, h->d_hx.p, nullptr);
        run_product(*h, h->d_hx.p, h->d_hy.p, mu, nullptr);
        launch_to_cluster(h->d_hy.p, h->d_perm_t.p, h->nr, mu, e, h->d_hy.p + ny, nullptr);
        HTG_HIP(hipGetLastError());
        HTG_HIP(hipMemcpy(out, h->d_hy.p + ny, ny * sizeof(double), hipMemcpyDeviceToHost));
        h->n_products++;
        h->product_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    });
}

int htg_hmatrix_mvprod_subrhs(htg_hmatrix *h, const double *in, double *out, int mu, int offset, int size,
                              int margin) {
    return guard([&] {
        if (!h || !in || !out || mu <= 0 || size < 0 || margin < 0 || offset < 0 || offset > h->nc)
            throw Error(HTG_ERR_INVALID, "invalid mvprod_subrhs arguments");
        // in holds the cluster-ordered entries [offset - margin, offset + size + margin)
        // (interleaved); every leaf whose columns meet [offset, offset + size] reads its
        // columns from there (hmatrix.hpp:1457-1554), the others are not applied: the
        // product of the H-matrix of those leaves (made once per window) with in
        // placed in a vector of zeros.
        htg_hmatrix &v = subrhs_view(*h, offset, size);
        apply_storage(v, h->storage);
        const int e = h->esz();
        std::vector<double> x((size_t)h->nc * mu * e, 0.0);
        const int lo = std::max(0, offset - margin), hi = std::min(h->nc, offset + size + margin);
        for (int j = lo; j < hi; j++)
            std::copy(in + (size_t)(j - offset + margin) * mu * e, in + (size_t)(j - offset + margin + 1) * mu * e,
                      x.data() + (size_t)j * mu * e);
        const int st = htg_hmatrix_mymvprod_local_to_local(&v, x.data(), out, mu);
        if (st != HTG_OK)
            throw Error(st, last_error());
        h->n_products++;
        h->product_seconds += v.product_seconds;
        v.product_seconds = 0;
    });
}

int htg_hmatrix_get_dense(const htg_hmatrix *h, double *out, int original_order) {
    return guard([&] {
        if (!h || !out)
            throw Error(HTG_ERR_INVALID, "null argument");
        const int e = h->esz();
        std::fill(out, out + (size_t)h->nr * h->nc * e, 0.0);
        // get_local_dense (cluster ordering) or copy_local_dense_perm (original ordering)
        for_leaf_entries(*h, AllEntries{}, [&](int i, int j, const double *v) {
            const size_t at = original_order ? (size_t)h->perm_t[i] + (size_t)h->nr * h->perm_s[j]
                                             : (size_t)i + (size_t)h->nr * j;
            for (int u = 0; u < e; u++)
                out[e * at + u] = v[u];
        });
    });
}

int htg_hmatrix_get_diagonal(const htg_hmatrix *h, double *diag, int original_order) {
    return guard([&] {
        if (!h || !diag)
            throw Error(HTG_ERR_INVALID, "null argument");
        if (h->nr != h->nc)
            throw Error(HTG_ERR_INVALID, "[Htool error] the diagonal needs a square matrix");
        const int e = h->esz();
        std::fill(diag, diag + (size_t)h->nr * e, 0.0);
        for_leaf_entries(*h, DiagonalEntries{}, [&](int i, int, const double *v) {
            for (int u = 0; u < e; u++)
                diag[e * (size_t)(original_order ? h->perm_t[i] : i) + u] = v[u];
        });
    });
}

int htg_hmatrix_profile(htg_hmatrix *h, const double *d_x, double *d_y, int mu, int reps, void *stream, double *ms) {
    return guard([&] {
        if (!h || !d_x || !d_y || !ms || reps <= 0 || mu <= 0 || mu > 64)
            throw Error(HTG_ERR_INVALID, "profile needs 1 <= mu <= 64 and reps > 0");
        hipStream_t st = static_cast<hipStream_t>(stream);
        hipEvent_t ev[4];
        for (auto &e : ev)
            HTG_HIP(hipEventCreate(&e));
        double acc[4] = {0, 0, 0, 0};
        for (int r = 0; r < reps; r++) {
            run_product(*h, d_x, d_y, mu, st, ev);
            HTG_HIP(hipEventSynchronize(ev[3]));
            float a = 0, b = 0, c = 0, d = 0;
            HTG_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
            HTG_HIP(hipEventElapsedTime(&b, ev[1], ev[2]));
            HTG_HIP(hipEventElapsedTime(&c, ev[2], ev[3]));
            HTG_HIP(hipEventElapsedTime(&d, ev[0], ev[3]));
            acc[0] += a, acc[1] += b, acc[2] += c, acc[3] += d;
        }
        for (auto &e : ev)
            (void)hipEventDestroy(e);
        for (int i = 0; i < 4; i++)
            ms[i] = acc[i] / reps;
    });
}

int htg_aca_batch(const htg_generator *gen, int compressor, double epsilon, int reqrank, int nt, const double *xt,
                  int ns, const double *xs, int nblocks, const int *M, const int *N, const int *rows,
                  const int *row_off, const int *cols, const int *col_off, const double *ctr_t, const double *ctr_s,
                  const int *off_t, const int *off_s, int cap, int *ranks, double *U, const int64_t *u_off, double *V,
                  const int64_t *v_off) {
    return guard([&] {
        AcaBatchArgs a{gen,    compressor, epsilon, reqrank, nt,    xt,    ns,    xs,    nblocks, M,     N,
                       rows,   row_off,    cols,    col_off, ctr_t, ctr_s, off_t, off_s, cap,     ranks, U,
                       u_off,  V,          v_off};
        AcaBatchPlan plan = plan_aca_batch(a);
        // device buffers of this call, allocated here and freed on return
        DevArray<unsigned char> buf;
        buf.alloc(plan.bytes);
        run_aca_batch(a, plan, buf.p);
    });
}

int htg_aca_batch_noalloc(const htg_generator *gen, int compressor, double epsilon, int reqrank, int nt,
                          const double *xt, int ns, const double *xs, int nblocks, const int *M, const int *N,
                          const int *rows, const int *row_off, const int *cols, const int *col_off,
                          const double *ctr_t, const double *ctr_s, const int *off_t, const int *off_s, int cap,
                          int *ranks, double *U, const int64_t *u_off, double *V, const int64_t *v_off) {
    return guard([&] {
        AcaBatchArgs a{gen,    compressor, epsilon, reqrank, nt,    xt,    ns,    xs,    nblocks, M,     N,
                       rows,   row_off,    cols,    col_off, ctr_t, ctr_s, off_t, off_s, cap,     ranks, U,
                       u_off,  V,          v_off};
        AcaBatchPlan plan = plan_aca_batch(a);
        require_device();
        if (plan.bytes > NOALLOC_SCRATCH_BYTES)
            throw Error(HTG_ERR_INVALID, "htg_aca_batch_noalloc: the call needs " + std::to_string(plan.bytes) +
                                             " bytes of device scratch, more than its " +
                                             std::to_string(NOALLOC_SCRATCH_BYTES) +
                                             " (fewer or smaller blocks, a smaller cap, or htg_aca_batch)");
        // one call at a time per device: the calls share the device's array
        unsigned char *dev = noalloc_scratch();
        std::lock_guard<std::mutex> hold(noalloc_lock());
        run_aca_batch(a, plan, dev);
    });
}

int htg_generator_block(const htg_generator *gen, int nt, const double *xt, int ns, const double *xs, int M, int N,
                        const int *rows, const int *cols, double *out) {
    return guard([&] {
        if (!gen || !xt || nt <= 0 || M < 0 || N < 0 || ((M > 0 && N > 0) && (!rows || !cols || !out)))
            throw Error(HTG_ERR_INVALID, "[Htool error] Invalid arguments to copy_submatrix");
        if (M == 0 || N == 0)
            return;
        const bool same = gen->same_points || xs == nullptr;
        const double *xsrc = same ? xt : xs;
        const int nsrc     = same ? nt : ns;
        std::vector<int> r(rows, rows + M), c(cols, cols + N);
        for (int v : r)
            if (v < 0 || v >= nt)
                throw Error(HTG_ERR_INVALID, "row index out of range");
        for (int v : c)
            if (v < 0 || v >= nsrc)
                throw Error(HTG_ERR_INVALID, "column index out of range");
        require_device();
        const KernelDesc kd = kernel_desc(*gen, same);
        const bool cplx     = kernel_is_complex(kd.kind);
        const int e         = cplx ? 2 : 1;
        PointsOnDevice pT, pS;
        pT.upload(xt, r); // point i of the tile is xt[rows[i]], its perm entry rows[i] (the diagonal rule)
        pS.upload(xsrc, c);
        std::vector<FillTile> tiles;
        for (int i0 = 0; i0 < M; i0 += 64)
            tiles.push_back(FillTile{(uint64_t)i0 * N, i0, std::min(64, M - i0), 0, N, N, 0});
        DevArray<FillTile> d_tiles;
        DevArray<double> d_out;
        d_tiles.upload(tiles);
        d_out.alloc((size_t)M * N * e);
        if (cplx)
            launch_dense_fill_c(kd, pT.view(), pS.view(), d_tiles.p, (int)tiles.size(), reinterpret_cast<cdbl *>(d_out.p),
                                nullptr);
        else
            launch_dense_fill(kd, pT.view(), pS.view(), d_tiles.p, (int)tiles.size(), d_out.p, nullptr);
        HTG_HIP(hipGetLastError());
        std::vector<double> h((size_t)M * N * e);
        HTG_HIP(hipMemcpy(h.data(), d_out.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int i = 0; i < M; i++) // row-major tile -> column-major ptr[i + M j]
            for (int j = 0; j < N; j++)
                for (int u = 0; u < e; u++)
                    out[e * ((size_t)i + (size_t)M * j) + u] = h[e * ((size_t)i * N + j) + u];
    });
}

} // extern "C"
