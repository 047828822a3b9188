// H-matrix assembly on the GPU (see assemble.hpp).
//
// Follows HMatrix::build / ComputeBlocks / ComputeAdmissibleBlock (reference:
// include/htool/types/hmatrix.hpp:285-341, 422-607): the block tasks come from
// the host block tree; admissible tasks are compressed on the GPU in batches
// ("waves"), and a block whose ACA reports rank -1 is split into its sons
// exactly as the reference's recursion does, the sons going into the next wave.
// The merge of all-dense sons into one dense block is decided afterwards,
// bottom up, with the reference's rule (hmatrix.hpp:536-546).
#include "assemble.hpp"
#include "predictor.hpp"

#include <chrono>

#include <algorithm>
#include <cmath>
#include <mutex>

namespace htg {

// ACA workspace and visited flags, one grow-only pair per device kept for the
// life of the process and lent to one build at a time: allocating and freeing
// the multi-GiB workspace build after build made the ACA of later builds up to
// 20x slower (N = 5e5: 0.15 s, then 1.4-3.0 s from the fourth build on).
struct AcaScratch {
    std::mutex lock;
    DevArray<double> ws;
    DevArray<unsigned char> vis;
};
AcaScratch &aca_scratch() {
    // never destroyed: freeing device memory from a static destructor can run
    // after the HIP runtime has shut down; the process exit releases it
    static AcaScratch *per_device = new AcaScratch[64];
    int dev = 0;
    HTG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64)
        throw Error(HTG_ERR_UNSUPPORTED, "more than 64 devices");
    return per_device[dev];
}

constexpr int CAP_FIRST = 64;                       // ACA terms reserved per block in the first attempt
constexpr int CAP_MAX   = 512;                      // must match the ACA kernel's CAPMAX
constexpr size_t WS_MAX = size_t(1) << 30;          // entries of ACA workspace per sub-batch (8 GiB of
                                                    // real entries, 16 GiB of complex ones)
constexpr int ACAW_MIN_DEFAULT = 16384;             // M + N from which a block runs the multi-workgroup ACA
constexpr int ACAW_MIN_DEFAULT_C = 4096;            // the same for complex entries (costlier per entry: the
                                                    // one-workgroup kernel's largest blocks set its length)

void PointsOnDevice::upload(const double *pts, const std::vector<int> &perm_h) {
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

KernelDesc kernel_desc(const htg_generator &g, bool same) {
    if (g.kernel < 0 || g.kernel > HTG_KERNEL_HERMITIAN_TEST)
        throw Error(HTG_ERR_UNSUPPORTED, "unknown kernel family " + std::to_string(g.kernel));
    return KernelDesc{g.kernel, same ? 1 : 0, g.delta, g.diag, g.wavenumber};
}

namespace {

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
    DevArray<double> store;
    uint64_t store_used = 0;     // entries (complex entries for the complex families)
    Assembly *rec = nullptr; // ACA records and counters
    int esz       = 1;       // doubles per entry
    // ACA workspace and visited flags of a sub-batch: the device's grow-only
    // pair, held for this build (see AcaScratch)
    AcaScratch &scratch = aca_scratch();
    std::lock_guard<std::mutex> scratch_hold{scratch.lock};
    DevArray<double> &d_ws        = scratch.ws;
    DevArray<unsigned char> &d_vis = scratch.vis;
    bool sym = false; // symmetric storage (htg_params.symmetry 'S')
    // A user generator on the host (htg_hmatrix_build_ex): every ACA runs the
    // multi-workgroup kernels with the generator's rows and columns fetched
    // through its copy_submatrix between the two halves of each cross term;
    // prow / pcol give the original index of each cluster position.
    const htg_host_generator *hg = nullptr;
    const int *prow = nullptr, *pcol = nullptr;

    // Sons of a false positive (hmatrix.hpp:543-606): the larger side splits;
    // in a symmetric build both sides split into every pair of sons
    // (ComputeAdmissibleBlocksSym, hmatrix.hpp:610-690).
    void expand(int id) {
        const int t = nodes[id].t, s = nodes[id].s;
        const ClusterNode &a = T.nodes[t], &b = S.nodes[s];
        if (a.leaf() && b.leaf()) {
            nodes[id].state = 3;
            return;
        }
        std::vector<std::pair<int, int>> kids;
        const bool split_t = b.leaf() ? true : (a.leaf() ? false : a.size > b.size);
        if (sym && !a.leaf() && !b.leaf()) {
            for (int l = 0; l < b.nb_sons; l++)
                for (int p = 0; p < a.nb_sons; p++)
                    kids.push_back({a.first_son + p, b.first_son + l});
        } else if (split_t) {
            for (int p = 0; p < a.nb_sons; p++)
                kids.push_back({a.first_son + p, s});
        } else {
            for (int p = 0; p < b.nb_sons; p++)
                kids.push_back({t, b.first_son + p});
        }
        const int kid0 = (int)nodes.size();
        for (const auto &k : kids) {
            FpNode c{};
            c.t = k.first;
            c.s = k.second;
            nodes.push_back(c);
        }
        nodes[id].state = 2;
        nodes[id].kid0  = kid0;
        nodes[id].nkids = (int)kids.size();
        for (int p = 0; p < (int)kids.size(); p++) {
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
        const size_t nw = wave.size();
        std::vector<int> caps(nw), kmaxs(nw);
        std::vector<float> feat(nw * 8);
        for (size_t w = 0; w < nw; w++) {
            const ClusterNode &a = T.nodes[nodes[wave[w]].t], &b = S.nodes[nodes[wave[w]].s];
            const long long M = a.size, N = b.size;
            long long kmax    = (M * N) / (M + N);
            if (prm.reqrank > 0)
                kmax = std::min<long long>(kmax, std::min<long long>(prm.reqrank, std::min(M, N)));
            kmaxs[w] = (int)std::max<long long>(1, std::min<long long>(kmax, CAP_MAX));
            caps[w]  = (int)std::max<long long>(1, std::min<long long>(kmax, CAP_FIRST));
            block_features(a, b, prm.eta, prm.epsilon, &feat[w * 8]);
        }
        std::vector<int> todo(nw);
        for (size_t w = 0; w < nw; w++)
            todo[w] = (int)w;
        if (prm.predictor && nw > 0) {
            // presize each block's workspace from the predicted terms, heaviest blocks first
            std::vector<float> t(nw);
            predictor_predict(*prm.predictor, feat.data(), (int)nw, t.data());
            const double f = 1.0 + std::max(prm.predictor_margin, 0.0);
            for (size_t w = 0; w < nw; w++) {
                const double c = std::ceil(f * t[w]) + 2;
                caps[w]        = (int)std::max(1.0, std::min<double>(c, kmaxs[w]));
            }
            auto work = [&](int w) {
                const FpNode &nd = nodes[wave[w]];
                return (int64_t)(T.nodes[nd.t].size + S.nodes[nd.s].size) * caps[w];
            };
            std::stable_sort(todo.begin(), todo.end(), [&](int x, int y) { return work(x) > work(y); });
        }
        std::vector<int> first_cap(caps);
        for (size_t w = 0; w < nw; w++) {
            const FpNode &nd = nodes[wave[w]];
            rec->aca_ws_entries += (int64_t)(T.nodes[nd.t].size + S.nodes[nd.s].size) * caps[w];
        }
        const size_t rec0 = rec->aca_terms.size();
        rec->aca_feat.insert(rec->aca_feat.end(), feat.begin(), feat.end());
        rec->aca_terms.resize(rec0 + nw, 0);
        rec->aca_cap.insert(rec->aca_cap.end(), first_cap.begin(), first_cap.end());
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
                    B.mn    = aca_size_bound(B.M, B.N, prm.ref_int_overflow != 0);
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
                d_ws.reserve(ws_used * esz); // kept across sub-batches and waves
                d_vis.reserve(vis_used);
                const std::vector<int> ranks = run_aca(blocks, vis_off);

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
                        if (caps[w] >= kmaxs[w])
                            throw Error(HTG_ERR_RANK_CAP, "ACA needed more than " + std::to_string(CAP_MAX) +
                                                              " terms in a " + std::to_string(B.M) + "x" +
                                                              std::to_string(B.N) + " block");
                        caps[w] = kmaxs[w];
                        rec->aca_retry++;
                        retry.push_back(w);
                        continue;
                    }
                    const int need               = aca_terms_needed(B.M, B.N, k);
                    rec->aca_terms[rec0 + w]     = need;
                    if (prm.ref_int_overflow ? (k == -1 && B.mn < (int64_t)B.M * B.N && B.M + B.N <= (int64_t)B.M * B.N)
                                             : aca_int32_wrap_differs(B.M, B.N, k))
                        rec->aca_int_wrap++;
                    rec->aca_need_entries += (int64_t)(B.M + B.N) * need;
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
                    if (esz == 2)
                        launch_lr_pack_c(d_packs.p, (int)packs.size(), d_udst.p, reinterpret_cast<cdbl *>(d_ws.p),
                                         reinterpret_cast<cdbl *>(store.p), reinterpret_cast<cdbl *>(store.p),
                                         nullptr);
                    else
                        launch_lr_pack(d_packs.p, (int)packs.size(), d_udst.p, d_ws.p, store.p, store.p, nullptr);
                    HTG_HIP(hipGetLastError());
                    HTG_HIP(hipDeviceSynchronize());
                }
                b0 = b1;
            }
            todo.swap(retry);
        }
    }

    // ACA of one sub-batch (workspace and visited flags in d_ws, d_vis): blocks
    // with M + N >= aca_wide_min on the multi-workgroup path, the others one
    // workgroup each. Ranks in block order.
    std::vector<int> run_aca(const std::vector<AcaBlock> &blocks, const std::vector<uint64_t> &vis_off) {
        const int wmin = hg ? 1 : prm.aca_wide_min != 0 ? prm.aca_wide_min : esz == 2 ? ACAW_MIN_DEFAULT_C
                                                                                           : ACAW_MIN_DEFAULT;
        std::vector<int> narrow, wide;
        for (size_t i = 0; i < blocks.size(); i++)
            ((wmin > 0 && blocks[i].M + blocks[i].N >= wmin) ? wide : narrow).push_back((int)i);
        std::vector<int> ranks(blocks.size());
        if (!narrow.empty()) {
            std::vector<AcaBlock> nb;
            std::vector<uint64_t> nv;
            for (int i : narrow) {
                nb.push_back(blocks[i]);
                nv.push_back(vis_off[i]);
            }
            DevArray<AcaBlock> d_b;
            DevArray<uint64_t> d_v;
            DevArray<int> d_r;
            d_b.upload(nb);
            d_v.upload(nv);
            d_r.alloc(nb.size());
            if (esz == 2)
                launch_aca_c(kd, dT, dS, d_b.p, (int)nb.size(), prm.epsilon, prm.reqrank,
                             reinterpret_cast<cdbl *>(d_ws.p), d_vis.p, d_v.p, d_r.p, nullptr);
            else
                launch_aca(kd, dT, dS, d_b.p, (int)nb.size(), prm.epsilon, prm.reqrank, d_ws.p, d_vis.p, d_v.p,
                           d_r.p, nullptr);
            HTG_HIP(hipGetLastError());
            std::vector<int> r(nb.size());
            HTG_HIP(hipMemcpy(r.data(), d_r.p, r.size() * sizeof(int), hipMemcpyDeviceToHost));
            for (size_t k = 0; k < narrow.size(); k++)
                ranks[narrow[k]] = r[k];
        }
        if (wide.empty())
            return ranks;
        rec->aca_wide += (int)wide.size();
        std::vector<AcaBlock> wb;
        std::vector<uint64_t> wv;
        std::vector<AcaWide> st;
        std::vector<AcawTask> t1, t2;
        for (size_t k = 0; k < wide.size(); k++) {
            const AcaBlock &B = blocks[wide[k]];
            const int n1 = B.swap ? B.N : B.M, n2 = B.swap ? B.M : B.N;
            AcaWide w{};
            w.nch1  = (n1 + ACAW_CHUNK - 1) / ACAW_CHUNK;
            w.nch2  = (n2 + ACAW_CHUNK - 1) / ACAW_CHUNK;
            w.task1 = (int)t1.size();
            w.task2 = (int)t2.size();
            for (int c = 0; c < w.nch1; c++)
                t1.push_back(AcawTask{(int)k, c});
            for (int c = 0; c < w.nch2; c++)
                t2.push_back(AcawTask{(int)k, c});
            wb.push_back(B);
            wv.push_back(vis_off[wide[k]]);
            st.push_back(w);
        }
        for (AcaWide &w : st)
            w.task1 += (int)t2.size(); // partial results: side-2 chunks, then side-1 chunks
        const size_t nslots = t1.size() + t2.size();
        DevArray<AcaBlock> d_b;
        DevArray<uint64_t> d_v;
        DevArray<AcaWide> d_st;
        DevArray<AcawTask> d_t1, d_t2;
        DevArray<double> d_pv, d_dd, d_xs;
        DevArray<int> d_pi;
        d_b.upload(wb);
        d_v.upload(wv);
        d_st.upload(st);
        d_t1.upload(t1);
        d_t2.upload(t2);
        d_pv.alloc(nslots);
        d_pi.alloc(nslots);
        d_dd.alloc(nslots);
        d_xs.alloc(nslots * CAP_MAX * esz);
        const int nw = (int)wide.size();
        launch_acaw_init(kd, dT, dS, d_b.p, d_st.p, nw, prm.reqrank, prm.epsilon, d_vis.p, d_v.p, nullptr);
        HTG_HIP(hipGetLastError());
        if (hg)
            host_terms(wb, st, d_b, d_st, d_v, d_t1, t1.size(), d_t2, t2.size(), d_pv, d_pi, d_dd, d_xs);
        // one cross term per iteration; every block stops by its cap (<= CAP_MAX terms)
        for (int it = 0; it <= CAP_MAX && !hg; it++) {
            launch_acaw_term(kd, dT, dS, d_b.p, d_st.p, nw, d_t1.p, (int)t1.size(), d_t2.p, (int)t2.size(),
                             prm.reqrank, prm.epsilon, d_ws.p, esz == 2, d_vis.p, d_v.p, d_pv.p, d_pi.p, d_dd.p,
                             d_xs.p, nullptr);
            HTG_HIP(hipGetLastError());
            if (it % 8 == 7 || it == CAP_MAX) {
                HTG_HIP(hipMemcpy(st.data(), d_st.p, st.size() * sizeof(AcaWide), hipMemcpyDeviceToHost));
                if (std::none_of(st.begin(), st.end(), [](const AcaWide &w) { return w.active != 0; }))
                    break;
            }
        }
        HTG_HIP(hipMemcpy(st.data(), d_st.p, st.size() * sizeof(AcaWide), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < wide.size(); k++) {
            if (st[k].active)
                throw Error(HTG_ERR_HIP, "multi-workgroup ACA did not finish");
            ranks[wide[k]] = st[k].rank;
        }
        return ranks;
    }

    // Cross terms of the multi-workgroup ACA with the host generator: before
    // each half the block states come back, the generator fills the vector
    // the half needs for every active block (side 2 at the side-1 pivot, then
    // side 1 at the side-2 pivot; with `swap` the sides are columns and rows),
    // and the values go to the device.
    void host_terms(const std::vector<AcaBlock> &wb, std::vector<AcaWide> &st, DevArray<AcaBlock> &d_b,
                    DevArray<AcaWide> &d_st, DevArray<uint64_t> &d_v, DevArray<AcawTask> &d_t1, size_t n1t,
                    DevArray<AcawTask> &d_t2, size_t n2t, DevArray<double> &d_pv, DevArray<int> &d_pi,
                    DevArray<double> &d_dd, DevArray<double> &d_xs) {
        const int nw = (int)wb.size();
        std::vector<uint64_t> gv_off(nw);
        uint64_t tot = 0;
        for (int k = 0; k < nw; k++) {
            gv_off[k] = tot;
            tot += (uint64_t)std::max(wb[k].M, wb[k].N);
        }
        std::vector<double> h_gv(tot * esz);
        DevArray<double> d_gv;
        DevArray<uint64_t> d_go;
        d_gv.alloc(tot * esz);
        d_go.upload(gv_off);
        for (int it = 0; it <= CAP_MAX; it++) {
            for (int half = 0; half < 2; half++) {
                HTG_HIP(hipMemcpy(st.data(), d_st.p, st.size() * sizeof(AcaWide), hipMemcpyDeviceToHost));
                bool any = false;
                for (int k = 0; k < nw; k++) {
                    if (!st[k].active)
                        continue;
                    any = true;
                    const AcaBlock &B = wb[k];
                    const int *rows = prow + B.off_t, *cols = pcol + B.off_s;
                    double *dst     = h_gv.data() + esz * gv_off[k];
                    // half 0: side-2 vector at I1; half 1: side-1 vector at I2
                    const int piv      = half == 0 ? st[k].I1 : st[k].I2;
                    const bool row_vec = (half == 0) != (B.swap != 0); // a row of the block: 1 x N
                    if (row_vec)
                        hg->copy_submatrix(1, B.N, rows + piv, cols, dst, hg->ctx);
                    else
                        hg->copy_submatrix(B.M, 1, rows, cols + piv, dst, hg->ctx);
                }
                if (!any)
                    return;
                HTG_HIP(hipMemcpy(d_gv.p, h_gv.data(), h_gv.size() * sizeof(double), hipMemcpyHostToDevice));
                launch_acaw_half(half == 0, dT, dS, d_b.p, d_st.p, nw, half == 0 ? d_t2.p : d_t1.p,
                                 (int)(half == 0 ? n2t : n1t), half == 0 ? 0 : (int)n2t, prm.reqrank, prm.epsilon,
                                 d_ws.p, esz == 2, d_vis.p, d_v.p, d_pv.p, d_pi.p, d_dd.p, d_xs.p, d_gv.p, d_go.p,
                                 nullptr);
                HTG_HIP(hipGetLastError());
            }
        }
    }

    void grow_store(uint64_t need) { // need: entries
        if (need * esz <= store.n)
            return;
        // doubling: few reallocations (each frees a large buffer, which later
        // builds in the same process pay for)
        uint64_t cap = std::max<uint64_t>(need * esz, std::max<uint64_t>(2 * store.n, uint64_t(1) << 20));
        double *np   = nullptr;
        HTG_HIP(hipMalloc(&np, cap * sizeof(double)));
        if (store.p && store_used)
            HTG_HIP(hipMemcpy(np, store.p, store_used * esz * sizeof(double), hipMemcpyDeviceToDevice));
        store.reset();
        store.p = np;
        store.n = cap;
    }
};

} // namespace

void assemble(const htg_params &prm, const htg_generator &gen, int nt, const double *xt, int ns, const double *xs,
              Assembly &res) {
    assemble_ex(prm, &gen, nullptr, gen.same_points || xs == nullptr, nt, xt, ns, xs, nullptr, nullptr, res);
}

void assemble_ex(const htg_params &prm, const htg_generator *gen, const htg_host_generator *hgen, bool same, int nt,
                 const double *xt, int ns, const double *xs, const ClusterTree *tree_t, const ClusterTree *tree_s,
                 Assembly &res) {
    same = same || xs == nullptr;
    ClusterTree T, Sown;
    if (tree_t)
        T = *tree_t;
    else
        build_cluster_tree(nt, xt, prm.minclustersize, T);
    if (!same) {
        if (tree_s)
            Sown = *tree_s;
        else
            build_cluster_tree(ns, xs, prm.minclustersize, Sown);
    }
    const ClusterTree &S = same ? T : Sown;
    res.nr     = nt;
    res.nc     = same ? nt : ns;
    res.perm_t = T.perm;
    res.perm_s = S.perm;
    if (prm.no_permutation) {
        // set_use_permutation(false) (hmatrix.hpp:218, 306-311): matrix indices,
        // points and product vectors are in the cluster ordering already
        for (int i = 0; i < res.nr; i++)
            res.perm_t[i] = i;
        for (int j = 0; j < res.nc; j++)
            res.perm_s[j] = j;
    }
    res.same = same;
    if (hgen) {
        res.kd      = KernelDesc{-1, same ? 1 : 0, 0.0, 0.0, 0.0}; // no device family: entries from the host
        res.cplx    = hgen->is_complex != 0;
        res.host_gen = true;
    } else {
        res.kd   = kernel_desc(*gen, same);
        res.cplx = kernel_is_complex(res.kd.kind);
    }
    res.pT.upload(xt, res.perm_t);
    if (!same)
        res.pS.upload(xs, res.perm_s);

    // dense leaves of a host generator are filled after the build (from the caller's generator)
    res.delay_dense = prm.delay_dense != 0 || hgen != nullptr;
    Assembler as{T, S, prm, res.kd, res.dT(), res.dS(), {}, {}, {}, 0, {}, 0, &res, res.cplx ? 2 : 1};
    const char uplo = sym_uplo(prm);
    as.sym          = uplo != 'N';
    as.hg           = hgen;
    as.prow         = res.perm_t.data();
    as.pcol         = res.perm_s.data();
    const std::vector<BlockTask> tasks = build_block_tasks(T, S, prm.eta, (size_t)prm.maxblocksize, uplo,
                                                           prm.mintargetdepth, prm.minsourcedepth);
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
    const auto aca_t0 = std::chrono::steady_clock::now();
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
    res.aca_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - aca_t0).count();
    for (size_t p = 0; p < tasks.size(); p++)
        if (roots[p] >= 0 && as.resolve(roots[p]))
            as.dense(tasks[p].t, tasks[p].s);

    for (Leaf &L : lr_leaves)
        L.k_hi = L.k;
    if (prm.recompress_tol > 0 || prm.lowprec_all || prm.lowprec_tol > 0)
        recompress_and_split(prm, lr_leaves, as.store, res.cplx);
    res.leaves = as.out;
    res.leaves.insert(res.leaves.end(), lr_leaves.begin(), lr_leaves.end());
    res.false_positive = as.false_positive;
    res.store.take(as.store);
}

void recompress_and_split(const htg_params &prm, std::vector<Leaf> &leaves, DevArray<double> &store, bool cplx) {
    if (prm.recompress_tol <= 0) {
        // No recompression: each ACA term gets unit factors and a scale, so a
        // 16-bit class holds every term near unit magnitude.
        std::vector<ScaleLeaf> sl;
        std::vector<int> which;
        int off = 0;
        for (size_t l = 0; l < leaves.size(); l++) {
            const Leaf &L = leaves[l];
            if (L.k > 0) {
                sl.push_back(ScaleLeaf{L.src_u, L.src_v, L.m, L.n, L.k, off});
                which.push_back((int)l);
                off += L.k;
            }
        }
        if (!sl.empty()) {
            DevArray<ScaleLeaf> d_sl;
            DevArray<double> d_sig;
            d_sl.upload(sl);
            d_sig.alloc(off);
            if (cplx)
                launch_term_scale_c(d_sl.p, (int)sl.size(), reinterpret_cast<cdbl *>(store.p), d_sig.p, nullptr);
            else
                launch_term_scale(d_sl.p, (int)sl.size(), store.p, d_sig.p, nullptr);
            HTG_HIP(hipGetLastError());
            std::vector<double> sig(off);
            HTG_HIP(hipMemcpy(sig.data(), d_sig.p, sig.size() * sizeof(double), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < sl.size(); i++) {
                Leaf &L = leaves[which[i]];
                L.sig.assign(sig.begin() + sl[i].sig_off, sig.begin() + sl[i].sig_off + L.k);
            }
        }
    } else {
        std::vector<RcLeaf> rc;
        std::vector<int> which;
        for (size_t l = 0; l < leaves.size(); l++) {
            const Leaf &L = leaves[l];
            if (L.k >= 2 && L.k <= (cplx ? RC_KMAX_C : RC_KMAX)) {
                rc.push_back(RcLeaf{L.src_u, L.src_v, L.m, L.n, L.k, 0});
                which.push_back((int)l);
            }
        }
        if (!rc.empty()) {
            DevArray<RcLeaf> d_rc;
            DevArray<int> d_rank, d_ok;
            DevArray<double> d_sig;
            d_rc.upload(rc);
            d_rank.alloc(rc.size());
            d_ok.alloc(rc.size());
            d_sig.alloc(rc.size() * RC_KMAX);
            if (cplx)
                launch_recompress_c(d_rc.p, (int)rc.size(), reinterpret_cast<cdbl *>(store.p), prm.recompress_tol,
                                    d_rank.p, d_ok.p, d_sig.p, nullptr);
            else
                launch_recompress(d_rc.p, (int)rc.size(), store.p, prm.recompress_tol, d_rank.p, d_ok.p, d_sig.p,
                                  nullptr);
            HTG_HIP(hipGetLastError());
            std::vector<int> rank(rc.size()), ok(rc.size());
            std::vector<double> sig(rc.size() * RC_KMAX);
            HTG_HIP(hipMemcpy(rank.data(), d_rank.p, rank.size() * sizeof(int), hipMemcpyDeviceToHost));
            HTG_HIP(hipMemcpy(ok.data(), d_ok.p, ok.size() * sizeof(int), hipMemcpyDeviceToHost));
            HTG_HIP(hipMemcpy(sig.data(), d_sig.p, sig.size() * sizeof(double), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < rc.size(); i++) {
                if (!ok[i])
                    continue;
                Leaf &L = leaves[which[i]];
                // V now holds r rows of N entries at src_v with row stride N: unchanged layout.
                L.k = rank[i];
                L.sig.assign(sig.begin() + i * RC_KMAX, sig.begin() + i * RC_KMAX + rank[i]);
            }
        }
    }
    // Precision classes. The trailing terms (after recompression, the
    // smallest singular values) go to class 1 (16-bit storage) while the error
    // they add, 2 * 2^-11 * sigma_j summed, stays within lowprec_tol * ||sigma||_2.
    const double u16 = std::ldexp(1.0, -11);
    for (Leaf &L : leaves) {
        if (L.k <= 0)
            continue;
        if (prm.lowprec_all) {
            L.k_hi = 0;
            continue;
        }
        L.k_hi = L.k;
        if (prm.lowprec_tol <= 0 || L.sig.empty())
            continue;
        double tot = 0;
        for (double s : L.sig)
            tot += s * s;
        const double budget = prm.lowprec_tol * std::sqrt(tot);
        double acc          = 0;
        int hi              = L.k;
        while (hi > 0 && acc + 2 * u16 * L.sig[hi - 1] <= budget) {
            acc += 2 * u16 * L.sig[hi - 1];
            hi--;
        }
        L.k_hi = hi;
    }
}

} // namespace htg

extern "C" int htg_release_workspace(void) {
    return htg::guard([&] {
        htg::AcaScratch &s = htg::aca_scratch();
        std::lock_guard<std::mutex> hold(s.lock); // a build on this device holds it until it ends
        s.ws.reset();
        s.vis.reset();
    });
}
