// Host side of libhtool_amd: H-matrix assembly driver, product layout and the
// C ABI declared in include/htool_amd.h.
//
// Assembly follows HMatrix::build / ComputeBlocks / ComputeAdmissibleBlock
// (reference: include/htool/types/hmatrix.hpp:285-341, 422-607): the block
// tasks come from the host block tree; dense tasks are generated on the GPU;
// admissible tasks are compressed on the GPU in batches ("waves"), and a block
// whose ACA reports rank -1 is split into its sons exactly as the reference's
// recursion does, the sons going into the next wave. The merge of all-dense
// sons into one dense block is decided afterwards, bottom up, with the
// reference's rule (hmatrix.hpp:536-546).
#include "../../include/htool_amd.h"
#include "kernels.hpp"
#include "tree.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace htg {
namespace {

thread_local std::string g_last_error;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HTG_HIP(expr)                                                                                 \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            throw Error(e_ == hipErrorOutOfMemory ? HTG_ERR_OOM : HTG_ERR_HIP,                        \
                        std::string("HIP error in ") + #expr + ": " + hipGetErrorString(e_));         \
    } while (0)

template <typename T>
struct DevArray {
    T *p     = nullptr;
    size_t n = 0;
    DevArray() = default;
    DevArray(const DevArray &) = delete;
    DevArray &operator=(const DevArray &) = delete;
    ~DevArray() { reset(); }
    void reset() {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        reset();
        if (count == 0)
            count = 1;
        HTG_HIP(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    void upload(const T *src, size_t count, hipStream_t st = nullptr) {
        alloc(count);
        if (count)
            HTG_HIP(hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice));
        (void)st;
    }
    void upload(const std::vector<T> &v, hipStream_t st = nullptr) { upload(v.data(), v.size(), st); }
};

constexpr int RB_MAX     = 128;   // rows per row block
constexpr int ALIGN      = 32;    // tile alignment in elements
constexpr int VCHUNK_EL  = 8192;  // target entries per V column chunk
constexpr int VGROUP_EL  = 8192;  // target entries per phase-1 wavefront
constexpr int VGROUP_TILES = 32;  // most V chunks per phase-1 wavefront
constexpr int CAP_FIRST  = 64;    // ACA terms reserved per block in the first attempt
constexpr int CAP_MAX    = 512;   // must match the ACA kernel's CAPMAX
constexpr size_t WS_MAX  = size_t(1) << 30; // doubles of ACA workspace per sub-batch (8 GiB)

inline uint64_t align_up(uint64_t v) { return (v + ALIGN - 1) / ALIGN * ALIGN; }
inline int pad4(int v) { return (v + 3) & ~3; }

struct Leaf {
    int oi, m, oj, n, k; // k = -1: dense
    uint64_t src_u = 0, src_v = 0; // build path: offsets of U and V in the compressed store
};

// Stored row length of a low-rank leaf's V: the global columns
// [oj & ~3, round4(oj + n)), so rows and the matching x entries are 16-byte aligned.
inline int v_ld(const Leaf &L) { return pad4(L.oj + L.n) - (L.oj & ~3); }

} // namespace
} // namespace htg

using namespace htg;

struct htg_hmatrix {
    int nr = 0, nc = 0;
    std::vector<int> perm_t, perm_s;
    std::vector<Leaf> leaves; // sorted by (oi, oj)
    int false_positive = 0;
    double build_seconds = 0;

    // product layout (see build_layout)
    std::vector<int> rb_row0, rb_col0, rb_order, colidx;
    std::vector<uint64_t> rb_aoff;
    std::vector<int> rb_ld; // row stride of each row block's block (columns rounded up to 4)
    std::vector<int> vg_off, multi_first; // multi_first: first chunk of each leaf with several chunks
    std::vector<VTile> vtiles;
    std::vector<std::vector<std::pair<int, uint64_t>>> leaf_tiles; // per leaf: (row block, tile offset)
    std::vector<uint64_t> leaf_v_off;
    uint64_t arena_size = 0, varena_size = 0;
    int t_size = 0, part_size = 0, n_lr = 0;

    DevArray<int> d_perm_t, d_perm_s, d_rb_row0, d_rb_col0, d_rb_order, d_colidx, d_vg_off, d_multi_first;
    DevArray<uint64_t> d_rb_aoff;
    DevArray<VTile> d_vtiles;
    DevArray<double> d_arena64, d_varena64;
    DevArray<float> d_arena32, d_varena32;
    DevArray<double> d_z, d_part; // z: cluster-ordered x then t (FP64 or FP32 view)
    DevArray<double> d_hx, d_hy; // staging for host-pointer products
    int storage = HTG_STORAGE_FP64;
};

namespace htg {
namespace {

// Product layout.
//  * Row blocks: the segments between consecutive leaf row boundaries (every
//    leaf covers a segment entirely or not at all), split to <= RB_MAX rows.
//  * Each row block owns one contiguous R x C column-major block in the tile
//    arena: the rows of its dense leaves (n columns each) and of its low-rank
//    leaves' U (k columns each), in leaf order. colidx gives, per column, the
//    z entry it multiplies: an x entry (dense) or a t entry (low rank).
//  * Each low-rank leaf stores V column-major (k x n) in the V arena; its
//    columns are cut into chunks of about VCHUNK_EL entries, and chunks are
//    grouped into wavefront work lists of about VGROUP_EL entries.
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
        const int parts = (len + RB_MAX - 1) / RB_MAX;
        for (int p = 0; p < parts; p++)
            h.rb_row0.push_back(a + (int)((long long)len * p / parts));
    }
    h.rb_row0.push_back(h.nr);
    const int nrb = (int)h.rb_row0.size() - 1;

    // t offsets (low-rank terms) and V arena offsets.
    h.t_size      = 0;
    h.varena_size = 0;
    std::vector<int> t_off(h.leaves.size(), -1);
    h.leaf_v_off.assign(h.leaves.size(), 0);
    for (size_t l = 0; l < h.leaves.size(); l++) {
        const Leaf &L = h.leaves[l];
        if (L.k > 0) {
            t_off[l] = h.t_size;
            h.t_size += L.k;
            h.leaf_v_off[l] = h.varena_size;
            h.varena_size   = align_up(h.varena_size + (uint64_t)L.k * v_ld(L));
        }
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
    h.rb_col0.assign(nrb + 1, 0);
    h.rb_aoff.assign(nrb, 0);
    h.rb_ld.assign(nrb, 0);
    h.colidx.clear();
    h.arena_size = 0;
    for (int r = 0; r < nrb; r++) {
        const int R   = h.rb_row0[r + 1] - h.rb_row0[r];
        h.rb_aoff[r]  = h.arena_size;
        int col       = 0;
        for (int q = cnt[r]; q < cnt[r + 1]; q++) {
            const int l   = rb_leaves[q];
            const Leaf &L = h.leaves[l];
            h.leaf_tiles[l].emplace_back(r, h.rb_aoff[r] + (uint64_t)col);
            if (L.k < 0) {
                for (int j = 0; j < L.n; j++)
                    h.colidx.push_back(L.oj + j);
                col += L.n;
            } else {
                for (int j = 0; j < L.k; j++)
                    h.colidx.push_back(pad4(h.nc) + t_off[l] + j);
                col += L.k;
            }
        }
        h.rb_col0[r + 1] = h.rb_col0[r] + col;
        h.rb_ld[r]       = pad4(col);
        h.arena_size     = align_up(h.arena_size + (uint64_t)R * h.rb_ld[r]);
    }
    // Heaviest row blocks are dispatched first.
    h.rb_order.resize(nrb);
    for (int r = 0; r < nrb; r++)
        h.rb_order[r] = r;
    std::stable_sort(h.rb_order.begin(), h.rb_order.end(), [&](int a, int b) {
        const int64_t wa = (int64_t)(h.rb_row0[a + 1] - h.rb_row0[a]) * (h.rb_col0[a + 1] - h.rb_col0[a]);
        const int64_t wb = (int64_t)(h.rb_row0[b + 1] - h.rb_row0[b]) * (h.rb_col0[b + 1] - h.rb_col0[b]);
        return wa > wb;
    });

    // V column chunks, grouped per workgroup.
    h.vtiles.clear();
    h.multi_first.clear();
    h.vg_off.assign(1, 0);
    h.part_size = 0;
    int lr_id = 0;
    int64_t group_el = 0;
    for (size_t l = 0; l < h.leaves.size(); l++) {
        const Leaf &L = h.leaves[l];
        if (L.k <= 0)
            continue;
        const int ld     = v_ld(L);
        const int cpc    = pad4(std::max(64, VCHUNK_EL / L.k));
        const int nparts = (ld + cpc - 1) / cpc;
        const int pbase  = h.part_size;
        if (nparts > 1)
            h.part_size += nparts * L.k;
        for (int c = 0; c < nparts; c++) {
            VTile v{};
            const int j0 = c * cpc;
            v.v_off     = h.leaf_v_off[l] + (uint64_t)j0;
            v.ld        = ld;
            v.k         = L.k;
            v.ncols     = std::min(cpc, ld - j0);
            v.x_off     = (L.oj & ~3) + j0;
            v.t_index   = pad4(h.nc) + t_off[l];
            v.part_off  = pbase + c * L.k;
            v.part_base = pbase;
            v.nparts    = nparts;
            if (nparts > 1 && c == 0)
                h.multi_first.push_back((int)h.vtiles.size());
            const int64_t el = (int64_t)v.k * v.ncols;
            const int in_group = (int)h.vtiles.size() - h.vg_off.back();
            if (in_group > 0 && (group_el + el > VGROUP_EL || in_group >= VGROUP_TILES)) {
                h.vg_off.push_back((int)h.vtiles.size());
                group_el = 0;
            }
            h.vtiles.push_back(v);
            group_el += el;
        }
        lr_id++;
    }
    if ((int)h.vtiles.size() > h.vg_off.back())
        h.vg_off.push_back((int)h.vtiles.size());
    h.n_lr = lr_id;
}

void upload_layout(htg_hmatrix &h) {
    h.d_perm_t.upload(h.perm_t);
    h.d_perm_s.upload(h.perm_s);
    h.d_rb_row0.upload(h.rb_row0);
    h.d_rb_col0.upload(h.rb_col0);
    h.d_rb_order.upload(h.rb_order);
    h.d_rb_aoff.upload(h.rb_aoff);
    h.d_colidx.upload(h.colidx);
    h.d_vg_off.upload(h.vg_off);
    h.d_vtiles.upload(h.vtiles);
    h.d_multi_first.upload(h.multi_first);
    // z = [x (nc entries, zero padded to a multiple of 4) | t], per RHS; zeroed
    // once so the padding entries are always finite zeros.
    h.d_z.alloc((size_t)(pad4(h.nc) + h.t_size) * 64);
    HTG_HIP(hipMemset(h.d_z.p, 0, h.d_z.n * sizeof(double)));
    h.d_part.alloc((size_t)std::max(h.part_size, 1) * 64);
}

struct PointsOnDevice {
    DevArray<double> x, y, z;
    DevArray<int> perm;
    void upload(const double *pts, const std::vector<int> &perm_h) {
        const size_t n = perm_h.size();
        std::vector<double> a(n), b(n), c(n);
        for (size_t i = 0; i < n; i++) {
            a[i] = pts[3 * (size_t)perm_h[i] + 0];
            b[i] = pts[3 * (size_t)perm_h[i] + 1];
            c[i] = pts[3 * (size_t)perm_h[i] + 2];
        }
        x.upload(a);
        y.upload(b);
        z.upload(c);
        perm.upload(perm_h);
    }
    DevPoints view() const { return DevPoints{x.p, y.p, z.p, perm.p}; }
};

KernelDesc kernel_desc(const htg_generator &g, bool same) {
    if (g.kernel < 0 || g.kernel > 3)
        throw Error(HTG_ERR_UNSUPPORTED, "unknown kernel family " + std::to_string(g.kernel));
    return KernelDesc{g.kernel, same ? 1 : 0, g.delta, g.diag};
}

// Nodes of the false-positive recursion (hmatrix.hpp:503-607).
struct FpNode {
    int t, s;
    int state; // 0 pending ACA, 1 low rank, 2 split into kids, 3 pair of leaf clusters
    int lr = -1;
    int kid0 = -1, nkids = 0;
};

struct Assembler {
    const ClusterTree &T, &S;
    const htg_params &prm;
    const KernelDesc kd;
    DevPoints dT, dS;
    std::vector<FpNode> nodes;
    std::vector<int> pending;
    std::vector<Leaf> out;
    int false_positive = 0;
    // compressed store: U (M x k col-major) then V (k x N row-major) per low-rank leaf
    std::vector<DevArray<double> *> store_chunks;
    DevArray<double> store;
    uint64_t store_used = 0;

    void expand(int id) {
        const ClusterNode &a = T.nodes[nodes[id].t], &b = S.nodes[nodes[id].s];
        if (a.leaf() && b.leaf()) {
            nodes[id].state = 3;
            return;
        }
        const bool split_t       = b.leaf() ? true : (a.leaf() ? false : a.size > b.size);
        const ClusterNode &owner = split_t ? a : b;
        const int kid0           = (int)nodes.size();
        for (int p = 0; p < owner.nb_sons; p++) {
            FpNode c{};
            c.t = split_t ? owner.first_son + p : nodes[id].t;
            c.s = split_t ? nodes[id].s : owner.first_son + p;
            nodes.push_back(c);
        }
        nodes[id].state = 2;
        nodes[id].kid0  = kid0;
        nodes[id].nkids = owner.nb_sons;
        for (int p = 0; p < owner.nb_sons; p++) {
            const int c = kid0 + p;
            if (admissible(T.nodes[nodes[c].t], S.nodes[nodes[c].s], prm.eta)) {
                nodes[c].state = 0;
                pending.push_back(c);
            } else {
                expand(c);
            }
        }
    }

    void dense(int t, int s) {
        const ClusterNode &a = T.nodes[t], &b = S.nodes[s];
        out.push_back(Leaf{a.offset, a.size, b.offset, b.size, -1});
    }

    // true = not pushed (may merge into its parent's dense block)
    bool resolve(int id) {
        const FpNode nd = nodes[id];
        if (nd.state == 1)
            return false;
        if (nd.state == 3)
            return true;
        bool all = true;
        std::vector<char> np(nd.nkids);
        for (int p = 0; p < nd.nkids; p++) {
            np[p] = resolve(nd.kid0 + p);
            all   = all && np[p];
        }
        const ClusterNode &a = T.nodes[nd.t], &b = S.nodes[nd.s];
        const double bsize   = double(a.size) * double(b.size);
        if (bsize <= prm.maxblocksize && all)
            return true;
        for (int p = 0; p < nd.nkids; p++)
            if (np[p])
                dense(nodes[nd.kid0 + p].t, nodes[nd.kid0 + p].s);
        return false;
    }

    // Runs ACA on every pending node; returns the low-rank leaves found.
    void run_wave(std::vector<int> wave, std::vector<Leaf> &lr_out, std::vector<int> &lr_node) {
        std::vector<int> caps(wave.size());
        for (size_t w = 0; w < wave.size(); w++) {
            const ClusterNode &a = T.nodes[nodes[wave[w]].t], &b = S.nodes[nodes[wave[w]].s];
            const long long M = a.size, N = b.size;
            long long kmax    = (M * N) / (M + N);
            if (prm.reqrank > 0)
                kmax = std::min<long long>(kmax, std::min<long long>(prm.reqrank, std::min(M, N)));
            caps[w] = (int)std::max<long long>(1, std::min<long long>(kmax, CAP_FIRST));
        }
        std::vector<int> todo(wave.size());
        for (size_t w = 0; w < wave.size(); w++)
            todo[w] = (int)w;
        while (!todo.empty()) {
            // sub-batch bounded by workspace size
            size_t b0 = 0;
            std::vector<int> retry;
            while (b0 < todo.size()) {
                std::vector<AcaBlock> blocks;
                std::vector<uint64_t> vis_off;
                uint64_t ws_used = 0, vis_used = 0;
                size_t b1        = b0;
                while (b1 < todo.size()) {
                    const int w          = todo[b1];
                    const FpNode &nd     = nodes[wave[w]];
                    const ClusterNode &a = T.nodes[nd.t], &b = S.nodes[nd.s];
                    const uint64_t need  = (uint64_t)(a.size + b.size) * caps[w];
                    if (!blocks.empty() && ws_used + need > WS_MAX)
                        break;
                    AcaBlock B{};
                    B.off_t = a.offset;
                    B.M     = a.size;
                    B.off_s = b.offset;
                    B.N     = b.size;
                    B.swap  = (prm.compressor == HTG_SYM_PARTIAL_ACA && a.offset < b.offset) ? 1 : 0;
                    B.cap   = caps[w];
                    const int n1 = B.swap ? B.N : B.M;
                    B.ws1 = ws_used;
                    B.ws2 = ws_used + (uint64_t)n1 * B.cap;
                    ws_used += need;
                    const double *c = B.swap ? b.ctr : a.ctr;
                    B.ctr[0] = c[0], B.ctr[1] = c[1], B.ctr[2] = c[2];
                    vis_off.push_back(vis_used);
                    vis_used += (uint64_t)(a.size + b.size);
                    blocks.push_back(B);
                    b1++;
                }
                DevArray<AcaBlock> d_blocks;
                DevArray<uint64_t> d_vis_off;
                DevArray<double> d_ws;
                DevArray<unsigned char> d_vis;
                DevArray<int> d_ranks;
                d_blocks.upload(blocks);
                d_vis_off.upload(vis_off);
                d_ws.alloc(ws_used);
                d_vis.alloc(vis_used);
                d_ranks.alloc(blocks.size());
                launch_aca(kd, dT, dS, d_blocks.p, (int)blocks.size(), prm.epsilon, prm.reqrank, d_ws.p, d_vis.p,
                           d_vis_off.p, d_ranks.p, nullptr);
                HTG_HIP(hipGetLastError());
                std::vector<int> ranks(blocks.size());
                HTG_HIP(hipMemcpy(ranks.data(), d_ranks.p, ranks.size() * sizeof(int), hipMemcpyDeviceToHost));

                // Move the accepted factors into the compressed store.
                std::vector<LrPack> packs;
                std::vector<UTileDst> udst;
                uint64_t add = 0;
                for (size_t i = 0; i < blocks.size(); i++)
                    if (ranks[i] > 0)
                        add += (uint64_t)ranks[i] * (blocks[i].M + blocks[i].N);
                grow_store(store_used + add);
                for (size_t i = 0; i < blocks.size(); i++) {
                    const int w      = todo[b0 + i];
                    const AcaBlock &B = blocks[i];
                    const int k      = ranks[i];
                    if (k == -3) {
                        if (caps[w] >= CAP_MAX)
                            throw Error(HTG_ERR_RANK_CAP, "ACA needed more than " + std::to_string(CAP_MAX) +
                                                              " terms in a " + std::to_string(B.M) + "x" +
                                                              std::to_string(B.N) + " block");
                        const long long kmax = ((long long)B.M * B.N) / (B.M + B.N);
                        caps[w]              = (int)std::min<long long>(kmax, CAP_MAX);
                        retry.push_back(w);
                        continue;
                    }
                    if (k == 0) { // fixed rank 0: a zero block with no factors (lrmat.hpp:37-44)
                        lr_out.push_back(Leaf{B.off_t, B.M, B.off_s, B.N, 0});
                        lr_node.push_back(wave[w]);
                    } else if (k > 0) {
                        LrPack p{};
                        p.src_u = B.swap ? B.ws2 : B.ws1;
                        p.src_v = B.swap ? B.ws1 : B.ws2;
                        p.M = B.M, p.N = B.N, p.k = k;
                        p.src_ldv = B.N;
                        p.dst_ldv = B.N;
                        Leaf L{B.off_t, B.M, B.off_s, B.N, k};
                        L.src_u = store_used;
                        L.src_v = store_used + (uint64_t)B.M * k;
                        p.dst_v = L.src_v;
                        p.piece_first = (int)udst.size();
                        p.npieces     = 1;
                        udst.push_back(UTileDst{L.src_u, 0, B.M, 1, B.M});
                        store_used += (uint64_t)k * (B.M + B.N);
                        packs.push_back(p);
                        lr_out.push_back(L);
                        lr_node.push_back(wave[w]);
                    } else {
                        lr_node.push_back(-1 - wave[w]); // false positive marker
                        lr_out.push_back(Leaf{B.off_t, B.M, B.off_s, B.N, k});
                    }
                }
                if (!packs.empty()) {
                    DevArray<LrPack> d_packs;
                    DevArray<UTileDst> d_udst;
                    d_packs.upload(packs);
                    d_udst.upload(udst);
                    launch_lr_pack(d_packs.p, (int)packs.size(), d_udst.p, d_ws.p, store.p, store.p, nullptr);
                    HTG_HIP(hipGetLastError());
                    HTG_HIP(hipDeviceSynchronize());
                }
                b0 = b1;
            }
            todo.swap(retry);
        }
    }

    void grow_store(uint64_t need) {
        if (need <= store.n)
            return;
        uint64_t cap = std::max<uint64_t>(need, store.n + store.n / 2 + (1u << 20));
        double *np   = nullptr;
        HTG_HIP(hipMalloc(&np, cap * sizeof(double)));
        if (store.p && store_used)
            HTG_HIP(hipMemcpy(np, store.p, store_used * sizeof(double), hipMemcpyDeviceToDevice));
        store.reset();
        store.p = np;
        store.n = cap;
    }
};

void finish_build(htg_hmatrix &h, const std::vector<Leaf> &leaves) {
    h.leaves = leaves;
    std::sort(h.leaves.begin(), h.leaves.end(),
              [](const Leaf &a, const Leaf &b) { return a.oi == b.oi ? a.oj < b.oj : a.oi < b.oi; });
    build_layout(h);
    upload_layout(h);
}

int pick_mu_chunk(int remaining) {
    int c = 64;
    while (c > remaining)
        c >>= 1;
    return c;
}

template <typename TS, typename TC>
void product(htg_hmatrix &h, const TS *arena, const TS *varena, const double *d_x, double *d_y, int mu,
             hipStream_t st, hipEvent_t *ev) {
    TC *z = reinterpret_cast<TC *>(h.d_z.p);
    int done = 0;
    while (done < mu) {
        const int MU = pick_mu_chunk(mu - done);
        if (ev)
            HTG_HIP(hipEventRecord(ev[0], st));
        launch_gather_x<TC>(d_x, h.d_perm_s.p, h.nc, done, MU, z, st);
        if (ev)
            HTG_HIP(hipEventRecord(ev[1], st));
        launch_vx<TS, TC>(h.d_vg_off.p, (int)h.vg_off.size() - 1, h.d_vtiles.p, h.d_multi_first.p,
                          (int)h.multi_first.size(), varena, z, h.d_part.p, MU, st);
        if (ev)
            HTG_HIP(hipEventRecord(ev[2], st));
        launch_rows<TS, TC>(h.d_rb_order.p, h.d_rb_row0.p, h.d_rb_col0.p, h.d_rb_aoff.p, (int)h.rb_row0.size() - 1,
                            h.d_colidx.p, arena, z, h.d_perm_t.p, h.nr, done, MU, d_y, st);
        if (ev)
            HTG_HIP(hipEventRecord(ev[3], st));
        done += MU;
    }
    HTG_HIP(hipGetLastError());
}

void run_product(htg_hmatrix &h, const double *d_x, double *d_y, int mu, hipStream_t st, hipEvent_t *ev = nullptr) {
    if (mu <= 0)
        throw Error(HTG_ERR_INVALID, "mu must be positive");
    if (h.storage == HTG_STORAGE_FP32)
        product<float, float>(h, h.d_arena32.p, h.d_varena32.p, d_x, d_y, mu, st, ev);
    else
        product<double, double>(h, h.d_arena64.p, h.d_varena64.p, d_x, d_y, mu, st, ev);
}

template <typename F>
int guard(F &&f) {
    try {
        f();
        return HTG_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return HTG_ERR_OOM;
    } catch (const std::invalid_argument &e) {
        g_last_error = std::string("[Htool error] ") + e.what();
        return HTG_ERR_INVALID;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return HTG_ERR_HIP;
    }
}

void require_device() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        throw Error(HTG_ERR_HIP, "no HIP device available: libhtool_amd runs on the GPU only");
}

} // namespace
} // namespace htg

extern "C" {

const char *htg_last_error(void) { return htg::g_last_error.c_str(); }

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
        if (prm->minclustersize <= 0)
            throw Error(HTG_ERR_INVALID, "[Htool error] MinClusterSize parameter cannot be zero");
        if (prm->maxblocksize < 1)
            throw Error(HTG_ERR_INVALID, "[Htool error] MaxBlockSize parameter cannot be zero");
        if (prm->compressor != HTG_PARTIAL_ACA && prm->compressor != HTG_SYM_PARTIAL_ACA)
            throw Error(HTG_ERR_UNSUPPORTED, "unknown compressor");
        const bool same = gen->same_points || xs == nullptr;
        if (same && ns != nt && ns > 0)
            throw Error(HTG_ERR_INVALID, "[Htool error] Invalid size in arguments for building HMatrix");
        require_device();
        auto t0 = std::chrono::steady_clock::now();
        std::unique_ptr<htg_hmatrix> h(new htg_hmatrix);
        ClusterTree T, Sown;
        build_cluster_tree(nt, xt, prm->minclustersize, T);
        if (!same)
            build_cluster_tree(ns, xs, prm->minclustersize, Sown);
        const ClusterTree &S = same ? T : Sown;
        h->nr     = nt;
        h->nc     = same ? nt : ns;
        h->perm_t = T.perm;
        h->perm_s = S.perm;

        PointsOnDevice pT, pS;
        pT.upload(xt, T.perm);
        if (!same)
            pS.upload(xs, S.perm);
        Assembler as{T, S, *prm, kernel_desc(*gen, same), pT.view(), same ? pT.view() : pS.view(), {}, {}, {}, 0,
                     {}, {}, 0};

        const std::vector<BlockTask> tasks = build_block_tasks(T, S, prm->eta, (size_t)prm->maxblocksize);
        std::vector<int> roots(tasks.size(), -1);
        for (size_t p = 0; p < tasks.size(); p++) {
            if (!tasks[p].admissible) {
                as.dense(tasks[p].t, tasks[p].s);
            } else {
                FpNode nd{};
                nd.t     = tasks[p].t;
                nd.s     = tasks[p].s;
                nd.state = 0;
                roots[p] = (int)as.nodes.size();
                as.nodes.push_back(nd);
                as.pending.push_back(roots[p]);
            }
        }
        std::vector<Leaf> lr_leaves;
        while (!as.pending.empty()) {
            std::vector<int> wave;
            wave.swap(as.pending);
            std::vector<Leaf> got;
            std::vector<int> got_node;
            as.run_wave(wave, got, got_node);
            for (size_t i = 0; i < got.size(); i++) {
                if (got_node[i] >= 0) {
                    as.nodes[got_node[i]].state = 1;
                    lr_leaves.push_back(got[i]);
                } else {
                    as.false_positive++;
                    as.expand(-1 - got_node[i]);
                }
            }
        }
        for (size_t p = 0; p < tasks.size(); p++)
            if (roots[p] >= 0 && as.resolve(roots[p]))
                as.dense(tasks[p].t, tasks[p].s);

        std::vector<Leaf> all = as.out;
        all.insert(all.end(), lr_leaves.begin(), lr_leaves.end());
        h->false_positive = as.false_positive;
        finish_build(*h, all);

        // Dense tiles straight from the generator; low-rank tiles from the store.
        h->d_arena64.alloc(h->arena_size);
        h->d_varena64.alloc(h->varena_size);
        HTG_HIP(hipMemset(h->d_arena64.p, 0, h->d_arena64.n * sizeof(double))); // row padding stays zero
        HTG_HIP(hipMemset(h->d_varena64.p, 0, h->d_varena64.n * sizeof(double)));
        std::vector<FillTile> fills;
        std::vector<LrPack> packs;
        std::vector<UTileDst> udst;
        for (size_t l = 0; l < h->leaves.size(); l++) {
            const Leaf &L = h->leaves[l];
            if (L.k < 0) {
                for (const auto &rt : h->leaf_tiles[l]) {
                    const int rb = rt.first;
                    const int R  = h->rb_row0[rb + 1] - h->rb_row0[rb];
                    fills.push_back(FillTile{rt.second, h->rb_row0[rb], R, L.oj, L.n, h->rb_ld[rb]});
                }
            } else if (L.k > 0) {
                LrPack pk{};
                pk.src_u = L.src_u;
                pk.src_v = L.src_v;
                pk.dst_v = h->leaf_v_off[l] + (L.oj & 3);
                pk.M = L.m, pk.N = L.n, pk.k = L.k;
                pk.src_ldv = L.n;
                pk.dst_ldv = v_ld(L);
                pk.piece_first = (int)udst.size();
                pk.npieces     = (int)h->leaf_tiles[l].size();
                for (const auto &rt : h->leaf_tiles[l]) {
                    const int rb = rt.first;
                    const int R  = h->rb_row0[rb + 1] - h->rb_row0[rb];
                    udst.push_back(UTileDst{rt.second, h->rb_row0[rb] - L.oi, R, h->rb_ld[rb], 1});
                }
                packs.push_back(pk);
            }
        }
        DevArray<FillTile> d_fills;
        DevArray<LrPack> d_packs;
        DevArray<UTileDst> d_udst;
        d_fills.upload(fills);
        d_packs.upload(packs);
        d_udst.upload(udst);
        launch_dense_fill(as.kd, as.dT, as.dS, d_fills.p, (int)fills.size(), h->d_arena64.p, nullptr);
        HTG_HIP(hipGetLastError());
        launch_lr_pack(d_packs.p, (int)packs.size(), d_udst.p, as.store.p, h->d_arena64.p, h->d_varena64.p, nullptr);
        HTG_HIP(hipGetLastError());
        HTG_HIP(hipDeviceSynchronize());
        h->build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        *out             = h.release();
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
        std::vector<const double *> src(nleaves);
        const double *p = data;
        for (int l = 0; l < nleaves; l++) {
            const int *L = leaves5 + 5 * l;
            if (L[1] <= 0 || L[3] <= 0 || L[0] < 0 || L[2] < 0 || L[0] + L[1] > nr || L[2] + L[3] > nc || L[4] < -1)
                throw Error(HTG_ERR_INVALID, "[Htool error] leaf " + std::to_string(l) + " out of range");
            leaves[l] = Leaf{L[0], L[1], L[2], L[3], L[4]};
            src[l]    = p;
            p += L[4] < 0 ? (size_t)L[1] * L[3] : (size_t)L[4] * (L[1] + L[3]);
        }
        // keep the caller's data reachable through the sort
        std::vector<int> order(nleaves);
        for (int l = 0; l < nleaves; l++)
            order[l] = l;
        std::sort(order.begin(), order.end(), [&](int a, int b) {
            return leaves[a].oi == leaves[b].oi ? leaves[a].oj < leaves[b].oj : leaves[a].oi < leaves[b].oi;
        });
        std::vector<Leaf> sorted(nleaves);
        std::vector<const double *> ssrc(nleaves);
        for (int l = 0; l < nleaves; l++) {
            sorted[l] = leaves[order[l]];
            ssrc[l]   = src[order[l]];
        }
        h->leaves = sorted;
        build_layout(*h);
        std::vector<double> arena(h->arena_size, 0.0), varena(h->varena_size, 0.0);
        for (int l = 0; l < nleaves; l++) {
            const Leaf &L  = h->leaves[l];
            const double *d = ssrc[l];
            for (const auto &rt : h->leaf_tiles[l]) {
                const int rb = rt.first;
                const int R  = h->rb_row0[rb + 1] - h->rb_row0[rb];
                const int i0 = h->rb_row0[rb] - L.oi;
                double *dst  = arena.data() + rt.second;
                const size_t ld = h->rb_ld[rb];
                const int nc_tile = L.k < 0 ? L.n : L.k; // dense: A columns; low rank: U columns
                for (int j = 0; j < nc_tile; j++)
                    for (int i = 0; i < R; i++)
                        dst[i * ld + j] = d[(size_t)(i0 + i) + (size_t)L.m * j];
            }
            if (L.k > 0) { // V: column-major k x n in, row-major (padded rows) on the device
                const double *V = d + (size_t)L.m * L.k;
                double *dst     = varena.data() + h->leaf_v_off[l] + (L.oj & 3);
                const size_t np = v_ld(L);
                for (int j = 0; j < L.n; j++)
                    for (int r = 0; r < L.k; r++)
                        dst[r * np + j] = V[r + (size_t)L.k * j];
            }
        }
        upload_layout(*h);
        h->d_arena64.upload(arena);
        h->d_varena64.upload(varena);
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
        if (storage == HTG_STORAGE_FP32) {
            if (!h->d_arena32.p) {
                h->d_arena32.alloc(h->arena_size);
                h->d_varena32.alloc(h->varena_size);
                launch_convert_f32(h->d_arena64.p, h->d_arena32.p, h->arena_size, nullptr);
                launch_convert_f32(h->d_varena64.p, h->d_varena32.p, h->varena_size, nullptr);
                HTG_HIP(hipGetLastError());
                HTG_HIP(hipDeviceSynchronize());
            }
        } else if (storage != HTG_STORAGE_FP64) {
            throw Error(HTG_ERR_UNSUPPORTED, "unknown storage mode " + std::to_string(storage));
        }
        h->storage = storage;
    });
}

int htg_hmatrix_get_info(const htg_hmatrix *h, htg_info *info) {
    return guard([&] {
        if (!h || !info)
            throw Error(HTG_ERR_INVALID, "null argument");
        htg_info r{};
        r.nr = h->nr, r.nc = h->nc;
        r.n_false_positive = h->false_positive;
        r.rank_min         = 0;
        long long ksum = 0;
        bool first     = true;
        for (const Leaf &L : h->leaves) {
            if (L.k < 0) {
                r.n_dense++;
                r.dense_entries += (int64_t)L.m * L.n;
            } else {
                r.n_lowrank++;
                r.lowrank_entries += (int64_t)L.k * (L.m + L.n);
                ksum += L.k;
                r.rank_min = first ? L.k : std::min(r.rank_min, L.k);
                r.rank_max = first ? L.k : std::max(r.rank_max, L.k);
                first      = false;
            }
        }
        r.rank_mean          = r.n_lowrank ? double(ksum) / r.n_lowrank : 0.0;
        const double stored  = double(r.dense_entries) + double(r.lowrank_entries);
        r.compression_ratio  = stored > 0 ? double(h->nr) * double(h->nc) / stored : 0.0;
        r.n_row_blocks       = (int)h->rb_row0.size() - 1;
        r.n_columns           = (int)h->colidx.size();
        r.n_vtiles           = (int)h->vtiles.size();
        r.build_seconds      = h->build_seconds;
        *info                = r;
    });
}

int htg_hmatrix_nleaves(const htg_hmatrix *h, int *nleaves, int64_t *data_size) {
    return guard([&] {
        if (!h)
            throw Error(HTG_ERR_INVALID, "null handle");
        int64_t s = 0;
        for (const Leaf &L : h->leaves)
            s += L.k < 0 ? (int64_t)L.m * L.n : (int64_t)L.k * (L.m + L.n);
        if (nleaves)
            *nleaves = (int)h->leaves.size();
        if (data_size)
            *data_size = s;
    });
}

int htg_hmatrix_get_leaves(const htg_hmatrix *h, int *leaves5, double *data) {
    return guard([&] {
        if (!h || !leaves5)
            throw Error(HTG_ERR_INVALID, "null argument");
        for (size_t l = 0; l < h->leaves.size(); l++) {
            const Leaf &L = h->leaves[l];
            int v[5]      = {L.oi, L.m, L.oj, L.n, L.k};
            std::memcpy(leaves5 + 5 * l, v, sizeof(v));
        }
        if (!data)
            return;
        std::vector<double> arena(h->arena_size), varena(h->varena_size);
        HTG_HIP(hipMemcpy(arena.data(), h->d_arena64.p, arena.size() * sizeof(double), hipMemcpyDeviceToHost));
        HTG_HIP(hipMemcpy(varena.data(), h->d_varena64.p, varena.size() * sizeof(double), hipMemcpyDeviceToHost));
        double *out = data;
        for (size_t l = 0; l < h->leaves.size(); l++) {
            const Leaf &L = h->leaves[l];
            const int ncol = L.k < 0 ? L.n : L.k;
            if (L.k == 0)
                continue;
            for (const auto &rt : h->leaf_tiles[l]) {
                const int rb = rt.first;
                const int R  = h->rb_row0[rb + 1] - h->rb_row0[rb];
                const int i0 = h->rb_row0[rb] - L.oi;
                const double *src = arena.data() + rt.second;
                const size_t ld   = h->rb_ld[rb];
                for (int j = 0; j < ncol; j++)
                    for (int i = 0; i < R; i++)
                        out[(size_t)(i0 + i) + (size_t)L.m * j] = src[i * ld + j];
            }
            out += (size_t)L.m * ncol;
            if (L.k > 0) {
                const double *src = varena.data() + h->leaf_v_off[l] + (L.oj & 3);
                const size_t np   = v_ld(L);
                for (int j = 0; j < L.n; j++)
                    for (int r = 0; r < L.k; r++)
                        out[r + (size_t)L.k * j] = src[r * np + j];
                out += (size_t)L.k * L.n;
            }
        }
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
    });
}

int htg_hmatrix_mvprod(htg_hmatrix *h, const double *x, double *y, int mu) {
    return guard([&] {
        if (!h || !x || !y || mu <= 0)
            throw Error(HTG_ERR_INVALID, "null argument");
        const size_t nx = (size_t)h->nc * mu, ny = (size_t)h->nr * mu;
        if (h->d_hx.n < nx)
            h->d_hx.alloc(nx);
        if (h->d_hy.n < ny)
            h->d_hy.alloc(ny);
        HTG_HIP(hipMemcpy(h->d_hx.p, x, nx * sizeof(double), hipMemcpyHostToDevice));
        run_product(*h, h->d_hx.p, h->d_hy.p, mu, nullptr);
        HTG_HIP(hipMemcpy(y, h->d_hy.p, ny * sizeof(double), hipMemcpyDeviceToHost));
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
        if (!gen || !xt || nblocks < 0 || !M || !N || !rows || !row_off || !cols || !col_off || !ctr_t || !ctr_s ||
            !off_t || !off_s || !ranks || !U || !u_off || !V || !v_off || cap <= 0)
            throw Error(HTG_ERR_INVALID, "[Htool error] Invalid arguments to htg_aca_batch");
        if (cap > CAP_MAX)
            throw Error(HTG_ERR_INVALID, "cap above " + std::to_string(CAP_MAX));
        require_device();
        const bool same = gen->same_points || xs == nullptr;
        const double *xsrc = same ? xt : xs;
        // Gather each block's points into contiguous stretches (the kernel's cluster order).
        std::vector<int> rp, cp;
        std::vector<AcaBlock> blocks(nblocks);
        std::vector<uint64_t> vis_off(nblocks);
        uint64_t ws = 0, vis = 0;
        for (int b = 0; b < nblocks; b++) {
            if (M[b] <= 0 || N[b] <= 0)
                throw Error(HTG_ERR_INVALID, "empty block");
            AcaBlock &B = blocks[b];
            B.off_t     = (int)rp.size();
            B.M         = M[b];
            B.off_s     = (int)cp.size();
            B.N         = N[b];
            for (int i = 0; i < M[b]; i++) {
                const int r = rows[row_off[b] + i];
                if (r < 0 || r >= nt)
                    throw Error(HTG_ERR_INVALID, "row index out of range");
                rp.push_back(r);
            }
            for (int j = 0; j < N[b]; j++) {
                const int c = cols[col_off[b] + j];
                if (c < 0 || c >= (same ? nt : ns))
                    throw Error(HTG_ERR_INVALID, "column index out of range");
                cp.push_back(c);
            }
            B.swap = (compressor == HTG_SYM_PARTIAL_ACA && off_t[b] < off_s[b]) ? 1 : 0;
            B.cap  = cap;
            const int n1 = B.swap ? B.N : B.M;
            B.ws1        = ws;
            B.ws2        = ws + (uint64_t)n1 * cap;
            ws += (uint64_t)(B.M + B.N) * cap;
            const double *c = B.swap ? ctr_s + 3 * b : ctr_t + 3 * b;
            B.ctr[0] = c[0], B.ctr[1] = c[1], B.ctr[2] = c[2];
            vis_off[b] = vis;
            vis += (uint64_t)(B.M + B.N);
        }
        PointsOnDevice pT, pS;
        pT.upload(xt, rp);
        pS.upload(xsrc, cp);
        const KernelDesc kd = kernel_desc(*gen, same);
        DevArray<AcaBlock> d_blocks;
        DevArray<uint64_t> d_vis_off;
        DevArray<double> d_ws;
        DevArray<unsigned char> d_vis;
        DevArray<int> d_ranks;
        d_blocks.upload(blocks);
        d_vis_off.upload(vis_off);
        d_ws.alloc(ws);
        d_vis.alloc(vis);
        d_ranks.alloc(nblocks);
        launch_aca(kd, pT.view(), pS.view(), d_blocks.p, nblocks, epsilon, reqrank, d_ws.p, d_vis.p, d_vis_off.p,
                   d_ranks.p, nullptr);
        HTG_HIP(hipGetLastError());
        HTG_HIP(hipMemcpy(ranks, d_ranks.p, nblocks * sizeof(int), hipMemcpyDeviceToHost));
        std::vector<double> hws(ws);
        HTG_HIP(hipMemcpy(hws.data(), d_ws.p, ws * sizeof(double), hipMemcpyDeviceToHost));
        for (int b = 0; b < nblocks; b++) {
            const AcaBlock &B = blocks[b];
            if (ranks[b] == -3)
                throw Error(HTG_ERR_RANK_CAP, "block " + std::to_string(b) + " needs more than cap terms");
            const int k = ranks[b];
            if (k <= 0)
                continue;
            const double *su = hws.data() + (B.swap ? B.ws2 : B.ws1);
            const double *sv = hws.data() + (B.swap ? B.ws1 : B.ws2);
            std::copy(su, su + (size_t)B.M * k, U + u_off[b]);
            double *vo = V + v_off[b];
            for (int r = 0; r < k; r++)
                for (int j = 0; j < B.N; j++)
                    vo[r + (size_t)k * j] = sv[(size_t)r * B.N + j];
        }
    });
}

} // extern "C"
