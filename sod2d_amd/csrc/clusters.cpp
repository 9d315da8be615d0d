// clusters.cpp -- element clusters for the brick-local assembly of the fused
// p=3 stage (host, set-up).
//
// Clusters are grown greedily from the lowest-numbered free element: the next
// member is the free element that shares the most nodes with the cluster
// (ties: the one whose nodes the cluster's elements share most often, then
// the lowest id), as long as the cluster has at most kClusterG elements and
// kClusterLMax distinct nodes.  On a box mesh numbered along i, then j, then
// k this yields 2x2x2 bricks.  A cluster's elements are evaluated in
// ascending id, so an interior node sums its element contributions in the
// reference's serial element order.
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace sodhip {

void free_clusters(MeshCache *m)
{
    if (!m || !m->clusters) return;
    ClusterTables *c = m->clusters.get();
    (void)hipFree(c->cel);
    (void)hipFree(c->cloc);
    (void)hipFree(c->cnb);
    (void)hipFree(c->cnode);
    (void)hipFree(c->ckind);
    (void)hipFree(c->snode);
    (void)hipFree(c->srow);
    (void)hipFree(c->ssrc);
    m->clusters.reset();
}

namespace {

template <typename V> int upload(sodhip_ctx ctx, int **dst, const std::vector<V> &v)
{
    const size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(int);
    if (hipMalloc(dst, bytes) != hipSuccess) {
        *dst = nullptr;
        return set_error(SODHIP_ERR_ALLOC, "cluster table allocation (%zu bytes) failed", bytes);
    }
    if (!v.empty())
        SODHIP_TRY_HIP(hipMemcpyAsync(*dst, v.data(), v.size() * sizeof(int),
                                      hipMemcpyHostToDevice, ctx->stream));
    return SODHIP_OK;
}

} // namespace

// The cluster tables on the host.  cT: element-major 0-based connectivity;
// row/src: the node-ordered incidence (src = e*64 + lexicographic position,
// ascending (e, a) per node); wflag: working flags; lw: the working list
// (1-based) in the caller's order.
struct HostClusters {
    int ncl = 0;
    std::vector<int> cel, cloc, cnb, cnode, ckind, snode, srow, ssrc;
};

static int build_clusters_host(int nelem, int npoin, const std::vector<int> &cT,
                               const std::vector<int> &row, const std::vector<int> &src,
                               const std::vector<unsigned char> &wflag, const std::vector<int> &lw,
                               HostClusters &H)
{
    constexpr int NN = 64, G = kClusterG, LMAX = kClusterLMax;
    const int ne = nelem, np = npoin;
    // lexicographic position of each element-local node
    std::vector<int> lexpos(NN);
    {
        std::vector<int> gI(NN), gJ(NN), gK(NN);
        if (sodhip_basis_tables(3, nullptr, nullptr, gI.data(), gJ.data(), gK.data(), nullptr,
                                nullptr, nullptr, nullptr) != SODHIP_OK)
            return SODHIP_ERR_ARG;
        for (int a = 0; a < NN; ++a) lexpos[a] = (gI[a] - 1) + 4 * ((gJ[a] - 1) + 4 * (gK[a] - 1));
    }
    auto elem_of_slot = [&](int j) { return src[j] / NN; };
    // the accumulation gives each node of an element its own lane: an element
    // that lists a node twice (a periodic mesh one element wide) has no
    // cluster tables, and its stage runs the node-ordered gather
    {
        std::vector<int> mark(np, -1);
        for (int e = 0; e < ne; ++e)
            for (int a = 0; a < NN; ++a) {
                const int n = cT[(size_t)e * NN + a];
                if (mark[n] == e)
                    return set_error(SODHIP_ERR_UNSUPPORTED,
                                     "element %d lists node %d twice: no element clusters", e + 1,
                                     n + 1);
                mark[n] = e;
            }
    }

    // ---- greedy growth
    std::vector<char> taken(ne, 0);
    std::vector<int> in_node(np, -1); // cluster id holding the node in its node set
    std::vector<int> score(ne, 0), mult(ne, 0), seen(ne, -1);
    std::vector<int> &cel = H.cel;
    cel.clear();
    cel.reserve((size_t)ne + G);
    std::vector<int> members, cand, nodes;
    int ncl = 0;
    for (int seed = 0; seed < ne; ++seed) {
        if (taken[seed]) continue;
        const int cid = ncl++;
        members.clear();
        cand.clear();
        nodes.clear();
        auto add = [&](int e) {
            taken[e] = 1;
            members.push_back(e);
            for (int a = 0; a < NN; ++a) {
                const int n = cT[(size_t)e * NN + a];
                const bool fresh = in_node[n] != cid;
                if (fresh) {
                    in_node[n] = cid;
                    nodes.push_back(n);
                }
                for (int j = row[n]; j < row[n + 1]; ++j) {
                    const int f = elem_of_slot(j);
                    if (taken[f]) continue;
                    if (seen[f] != cid) {
                        seen[f] = cid;
                        score[f] = 0;
                        mult[f] = 0;
                        cand.push_back(f);
                    }
                    if (fresh) ++score[f];
                    ++mult[f];
                }
            }
        };
        add(seed);
        while ((int)members.size() < G) {
            int best = -1;
            for (int f : cand) {
                if (taken[f]) continue;
                // every element has NN distinct nodes: NN - score of them are new
                if ((int)nodes.size() + NN - score[f] > LMAX) continue;
                if (best < 0 || score[f] > score[best] ||
                    (score[f] == score[best] &&
                     (mult[f] > mult[best] || (mult[f] == mult[best] && f < best))))
                    best = f;
            }
            if (best < 0) break;
            add(best);
        }
        std::sort(members.begin(), members.end());
        for (int i = 0; i < G; ++i) cel.push_back(i < (int)members.size() ? members[i] : -1);
    }
    H.ncl = ncl;

    // ---- local nodes, kinds and element-node local indices
    std::vector<int> &cnb = H.cnb, &cnode = H.cnode, &ckind = H.ckind, &cloc = H.cloc;
    cnb.assign((size_t)ncl + 1, 0);
    cloc.assign((size_t)ncl * G * NN, 0);
    cnode.clear();
    ckind.clear();
    std::vector<int> cnt(np, 0), loc(np, -1);
    cnode.reserve((size_t)ncl * 343);
    for (int c = 0; c < ncl; ++c) {
        const size_t b = cnode.size();
        for (int i = 0; i < G; ++i) {
            const int e = cel[(size_t)c * G + i];
            if (e < 0) continue;
            for (int a = 0; a < NN; ++a) {
                const int n = cT[(size_t)e * NN + a];
                if (cnt[n]++ == 0) cnode.push_back(n);
            }
        }
        std::sort(cnode.begin() + b, cnode.end());
        for (size_t l = b; l < cnode.size(); ++l) {
            const int n = cnode[l];
            loc[n] = (int)(l - b);
            const bool working = wflag[n] != 0;
            const bool interior = cnt[n] == row[n + 1] - row[n];
            ckind.push_back(!working ? 2 : interior ? 0 : 1);
        }
        for (int i = 0; i < G; ++i) {
            const int e = cel[(size_t)c * G + i];
            if (e < 0) continue;
            for (int a = 0; a < NN; ++a)
                cloc[((size_t)c * G + i) * NN + lexpos[a]] = loc[cT[(size_t)e * NN + a]];
        }
        for (size_t l = b; l < cnode.size(); ++l) cnt[cnode[l]] = 0;
        cnb[c + 1] = (int)cnode.size();
    }
    const size_t nloc = cnode.size();
    SODHIP_REQUIRE(nloc < ((size_t)1 << 31), "too many cluster-local nodes");

    // ---- surface gather over the working nodes not finished in a cluster,
    // in the caller's working-list order; a node's records in cluster order
    std::vector<int> interior_node(np, 0), nrec(np, 0);
    for (size_t l = 0; l < nloc; ++l) {
        if (ckind[l] == 0) interior_node[cnode[l]] = 1;
        if (ckind[l] == 1) ++nrec[cnode[l]];
    }
    std::vector<int> &snode = H.snode, &srow = H.srow, &ssrc = H.ssrc;
    snode.clear();
    srow.assign(1, 0);
    std::vector<int> pos(np, -1);
    for (size_t i = 0; i < lw.size(); ++i) {
        const int n = lw[i] - 1;
        if (interior_node[n]) continue;
        pos[n] = (int)snode.size();
        snode.push_back(n + 1);
        srow.push_back(srow.back() + nrec[n]);
    }
    ssrc.assign(srow.back(), 0);
    std::vector<int> fill(srow.begin(), srow.end() - 1);
    for (int c = 0; c < ncl; ++c)
        for (int l = cnb[c]; l < cnb[c + 1]; ++l)
            if (ckind[l] == 1) ssrc[fill[pos[cnode[l]]]++] = l;
    return SODHIP_OK;
}

int get_clusters(sodhip_ctx ctx, MeshCache *m, const ClusterTables **out)
{
    if (m->clusters) {
        *out = m->clusters.get();
        return SODHIP_OK;
    }
    SODHIP_REQUIRE(m->porder == 3, "element clusters are built for porder 3 only");
    SODHIP_REQUIRE(m->wrow && (int)m->h_wflag.size() == m->npoin, "working tables missing");
    std::vector<int> lw(m->npoin_w);
    SODHIP_TRY_HIP(hipMemcpyAsync(lw.data(), m->lpw_key, lw.size() * sizeof(int),
                                  hipMemcpyDeviceToHost, ctx->stream));
    SODHIP_TRY_HIP(hipStreamSynchronize(ctx->stream));
    HostClusters H;
    int rc = build_clusters_host(m->nelem, m->npoin, m->h_cT, m->h_row, m->h_src, m->h_wflag, lw, H);
    if (rc) return rc;
    auto t = std::make_unique<ClusterTables>();
    t->ncl = H.ncl;
    t->nloc = (int)H.cnode.size();
    t->ns = (int)H.snode.size();
    if ((rc = upload(ctx, &t->cel, H.cel)) || (rc = upload(ctx, &t->cloc, H.cloc)) ||
        (rc = upload(ctx, &t->cnb, H.cnb)) || (rc = upload(ctx, &t->cnode, H.cnode)) ||
        (rc = upload(ctx, &t->ckind, H.ckind)) || (rc = upload(ctx, &t->snode, H.snode)) ||
        (rc = upload(ctx, &t->srow, H.srow)) || (rc = upload(ctx, &t->ssrc, H.ssrc))) {
        (void)hipStreamSynchronize(ctx->stream);
        m->clusters = std::move(t);
        free_clusters(m);
        return rc;
    }
    SODHIP_TRY_HIP(hipStreamSynchronize(ctx->stream));
    m->clusters = std::move(t);
    *out = m->clusters.get();
    return SODHIP_OK;
}

} // namespace sodhip

using namespace sodhip;

extern "C" int sodhip_element_clusters(int porder, int nelem, int npoin, const int *connec,
                                       int npoin_w, const int *lpoin_w, int *sizes, int *cel,
                                       int *cloc, int *cnb, int *cnode, int *ckind, int *snode,
                                       int *srow, int *ssrc)
{
    SODHIP_REQUIRE(porder == 3, "element clusters are built for porder 3 only");
    SODHIP_REQUIRE(nelem > 0 && npoin > 0 && npoin_w > 0 && npoin_w <= npoin && connec &&
                       lpoin_w && sizes,
                   "invalid arguments");
    constexpr int NN = 64;
    std::vector<int> gI(NN), gJ(NN), gK(NN), perm(NN);
    if (sodhip_basis_tables(3, nullptr, nullptr, gI.data(), gJ.data(), gK.data(), nullptr, nullptr,
                            nullptr, nullptr) != SODHIP_OK)
        return SODHIP_ERR_ARG;
    for (int a = 0; a < NN; ++a) perm[a] = (gI[a] - 1) + 4 * ((gJ[a] - 1) + 4 * (gK[a] - 1));
    const size_t en = (size_t)nelem * NN;
    std::vector<int> cT(en), row((size_t)npoin + 1, 0), src(en);
    for (long e = 0; e < nelem; ++e)
        for (int a = 0; a < NN; ++a) {
            const int v = connec[(size_t)a * nelem + e];
            SODHIP_REQUIRE(v >= 1 && v <= npoin, "connec entry %d outside 1..npoin", v);
            cT[(size_t)e * NN + a] = v - 1;
            row[(size_t)v]++;
        }
    for (long n = 0; n < npoin; ++n) row[n + 1] += row[n];
    std::vector<int> fill(row.begin(), row.end() - 1);
    for (long e = 0; e < nelem; ++e)
        for (int a = 0; a < NN; ++a) src[fill[cT[(size_t)e * NN + a]]++] = (int)(e * NN + perm[a]);
    std::vector<unsigned char> wflag(npoin, 0);
    std::vector<int> lw(lpoin_w, lpoin_w + npoin_w);
    for (int i = 0; i < npoin_w; ++i) {
        SODHIP_REQUIRE(lw[i] >= 1 && lw[i] <= npoin, "lpoin_w entry %d outside 1..npoin", lw[i]);
        wflag[lw[i] - 1] = 1;
    }
    HostClusters H;
    int rc = build_clusters_host(nelem, npoin, cT, row, src, wflag, lw, H);
    if (rc) return rc;
    sizes[0] = H.ncl;
    sizes[1] = (int)H.cnode.size();
    sizes[2] = (int)H.snode.size();
    sizes[3] = (int)H.ssrc.size();
    auto put = [](int *dst, const std::vector<int> &v) {
        if (dst) std::copy(v.begin(), v.end(), dst);
    };
    put(cel, H.cel);
    put(cloc, H.cloc);
    put(cnb, H.cnb);
    put(cnode, H.cnode);
    put(ckind, H.ckind);
    put(snode, H.snode);
    put(srow, H.srow);
    put(ssrc, H.ssrc);
    return SODHIP_OK;
}
