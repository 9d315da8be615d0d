// runtime.hip -- device context, OpenACC-style present table, cached
// connectivity derivatives (element-major connectivity and node-ordered
// incidence), launch timing and error reporting.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace sodhip {

static thread_local char g_err[1024] = "";

int set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int ensure_buf(DevBuf &b, size_t bytes)
{
    if (b.bytes >= bytes && b.ptr) return SODHIP_OK;
    if (b.pinned)
        return set_error(SODHIP_ERR_ARG,
                         "a captured graph holds this scratch buffer (%zu bytes); %zu bytes "
                         "requested: release the graph or reserve more before capturing",
                         b.bytes, bytes);
    if (b.ptr) {
        SODHIP_TRY_HIP(hipFree(b.ptr));
        b.ptr = nullptr;
        b.bytes = 0;
    }
    size_t want = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&b.ptr, want);
    if (e != hipSuccess) {
        b.ptr = nullptr;
        return set_error(SODHIP_ERR_ALLOC, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    b.bytes = want;
    return SODHIP_OK;
}

int diag_grid(sodhip_ctx ctx, int nelem, int blocks_per_cu)
{
    const int g = std::min((nelem + 7) / 8 * 8, ctx->num_cu * blocks_per_cu);
    return std::max(8, g / 8 * 8);
}

int elem_grid(sodhip_ctx ctx, int nelem, int blocks_per_cu)
{
    // a multiple of 8 so that every XCD receives the same number of blocks
    long g = (long)ctx->num_cu * blocks_per_cu;
    long need = ((long)nelem + 7) / 8 * 8;
    g = std::min(g, need);
    g = std::max<long>(8, (g + 7) / 8 * 8);
    return (int)g;
}

static void free_mesh(MeshCache *m)
{
    if (!m) return;
    free_clusters(m);
    (void)hipFree(m->connecT);
    (void)hipFree(m->row);
    (void)hipFree(m->src);
    (void)hipFree(m->perm);
    (void)hipFree(m->wrow);
    (void)hipFree(m->wsrc);
    (void)hipFree(m->wflag);
}

int get_mesh(sodhip_ctx ctx, int porder, int nelem, int npoin, const int *connec,
             MeshCache **out)
{
    for (auto &m : ctx->meshes)
        if (m->connec_key == connec && m->porder == porder && m->nelem == nelem &&
            m->npoin == npoin) {
            *out = m.get();
            return SODHIP_OK;
        }
    const int nn = nnode_of(porder);
    const size_t en = (size_t)nelem * nn;
    SODHIP_REQUIRE(en < (size_t)1 << 31, "nelem*nnode exceeds 2^31");
    std::vector<int> h(en);
    SODHIP_TRY_HIP(hipMemcpyAsync(h.data(), connec, en * sizeof(int), hipMemcpyDeviceToHost,
                                  ctx->stream));
    SODHIP_TRY_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<int> cT(en), row((size_t)npoin + 1, 0), src(en), perm(nn);
    {
        // element-local scratch position: lexicographic in the (i, j, k) line
        // indices of the standard tables of this order
        std::vector<int> gI(nn), gJ(nn), gK(nn);
        if (sodhip_basis_tables(porder, nullptr, nullptr, gI.data(), gJ.data(), gK.data(),
                                nullptr, nullptr, nullptr, nullptr) != SODHIP_OK)
            return SODHIP_ERR_ARG;
        const int n1 = porder + 1;
        for (int a = 0; a < nn; ++a) perm[a] = (gI[a] - 1) + n1 * ((gJ[a] - 1) + n1 * (gK[a] - 1));
    }
    for (size_t i = 0; i < en; ++i) {
        int v = h[i];
        if (v < 1 || v > npoin)
            return set_error(SODHIP_ERR_ARG, "connec entry %d outside 1..npoin=%d", v, npoin);
    }
    for (long e = 0; e < nelem; ++e)
        for (int a = 0; a < nn; ++a) {
            int n = h[(size_t)a * nelem + e] - 1;
            cT[(size_t)e * nn + a] = n;
            row[(size_t)n + 1]++;
        }
    for (long n = 0; n < npoin; ++n) row[n + 1] += row[n];
    std::vector<int> fill(row.begin(), row.end() - 1);
    for (long e = 0; e < nelem; ++e) // ascending (e,a): the reference's serial assembly order
        for (int a = 0; a < nn; ++a) {
            int n = cT[(size_t)e * nn + a];
            src[fill[n]++] = (int)((size_t)e * nn + perm[a]);
        }
    auto m = std::make_unique<MeshCache>();
    m->connec_key = connec;
    m->porder = porder;
    m->nelem = nelem;
    m->npoin = npoin;
    hipError_t e1 = hipMalloc(&m->connecT, en * sizeof(int));
    hipError_t e2 = hipMalloc(&m->row, ((size_t)npoin + 1) * sizeof(int));
    hipError_t e3 = hipMalloc(&m->src, en * sizeof(int));
    hipError_t e4 = hipMalloc(&m->perm, nn * sizeof(int));
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess) {
        free_mesh(m.get());
        return set_error(SODHIP_ERR_ALLOC, "mesh cache allocation failed");
    }
    SODHIP_TRY_HIP(hipMemcpyAsync(m->connecT, cT.data(), en * sizeof(int), hipMemcpyHostToDevice,
                                  ctx->stream));
    SODHIP_TRY_HIP(hipMemcpyAsync(m->row, row.data(), row.size() * sizeof(int),
                                  hipMemcpyHostToDevice, ctx->stream));
    SODHIP_TRY_HIP(hipMemcpyAsync(m->src, src.data(), en * sizeof(int), hipMemcpyHostToDevice,
                                  ctx->stream));
    SODHIP_TRY_HIP(hipMemcpyAsync(m->perm, perm.data(), nn * sizeof(int), hipMemcpyHostToDevice,
                                  ctx->stream));
    SODHIP_TRY_HIP(hipStreamSynchronize(ctx->stream));
    m->h_row = std::move(row);
    m->h_src = std::move(src);
    m->h_cT = std::move(cT);
    *out = m.get();
    ctx->meshes.push_back(std::move(m));
    return SODHIP_OK;
}

int find_mesh(sodhip_ctx ctx, int porder, int nelem, int npoin, const int *connec, int npoin_w,
              const int *lpoin_w, const MeshCache **out)
{
    for (auto &m : ctx->meshes)
        if (m->connec_key == connec && m->porder == porder && m->nelem == nelem &&
            m->npoin == npoin) {
            if (m->lpw_key != lpoin_w || m->npoin_w != npoin_w || !m->wrow)
                return set_error(SODHIP_ERR_ARG,
                                 "no working-node tables for this lpoin_w: an allocating call "
                                 "(sodhip_rk_stage_f64) must build them first");
            *out = m.get();
            return SODHIP_OK;
        }
    return set_error(SODHIP_ERR_ARG, "no mesh tables for this connec: an allocating call "
                                     "(sodhip_rk_stage_f64) must build them first");
}

int get_working(sodhip_ctx ctx, MeshCache *m, int npoin_w, const int *lpoin_w)
{
    if (m->lpw_key == lpoin_w && m->npoin_w == npoin_w && m->wrow) return SODHIP_OK;
    std::vector<int> lw(npoin_w);
    SODHIP_TRY_HIP(hipMemcpyAsync(lw.data(), lpoin_w, (size_t)npoin_w * sizeof(int),
                                  hipMemcpyDeviceToHost, ctx->stream));
    SODHIP_TRY_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<int> wrow((size_t)npoin_w + 1, 0);
    for (int i = 0; i < npoin_w; ++i) {
        const int n = lw[i] - 1;
        if (n < 0 || n >= m->npoin)
            return set_error(SODHIP_ERR_ARG, "lpoin_w entry %d outside 1..npoin=%d", lw[i], m->npoin);
        wrow[(size_t)i + 1] = wrow[i] + (m->h_row[(size_t)n + 1] - m->h_row[n]);
    }
    std::vector<unsigned char> wflag((size_t)m->npoin, 0);
    for (int i = 0; i < npoin_w; ++i) wflag[(size_t)lw[i] - 1] = 1;
    std::vector<int> wsrc(wrow[npoin_w]);
    for (int i = 0; i < npoin_w; ++i) {
        const int n = lw[i] - 1;
        std::copy(m->h_src.begin() + m->h_row[n], m->h_src.begin() + m->h_row[(size_t)n + 1],
                  wsrc.begin() + wrow[i]);
    }
    free_clusters(m);
    (void)hipFree(m->wrow);
    (void)hipFree(m->wsrc);
    (void)hipFree(m->wflag);
    m->wrow = m->wsrc = nullptr;
    m->wflag = nullptr;
    m->lpw_key = nullptr;
    if (hipMalloc(&m->wrow, wrow.size() * sizeof(int)) != hipSuccess ||
        hipMalloc(&m->wsrc, std::max<size_t>(1, wsrc.size()) * sizeof(int)) != hipSuccess ||
        hipMalloc(&m->wflag, wflag.size()) != hipSuccess)
        return set_error(SODHIP_ERR_ALLOC, "working incidence allocation failed");
    SODHIP_TRY_HIP(hipMemcpyAsync(m->wflag, wflag.data(), wflag.size(), hipMemcpyHostToDevice,
                                  ctx->stream));
    SODHIP_TRY_HIP(hipMemcpyAsync(m->wrow, wrow.data(), wrow.size() * sizeof(int),
                                  hipMemcpyHostToDevice, ctx->stream));
    if (!wsrc.empty())
        SODHIP_TRY_HIP(hipMemcpyAsync(m->wsrc, wsrc.data(), wsrc.size() * sizeof(int),
                                      hipMemcpyHostToDevice, ctx->stream));
    SODHIP_TRY_HIP(hipStreamSynchronize(ctx->stream));
    m->lpw_key = lpoin_w;
    m->npoin_w = npoin_w;
    m->h_wflag = std::move(wflag);
    return SODHIP_OK;
}

int prof_begin(sodhip_ctx ctx, int kind)
{
    if (!ctx->prof) return -1;
    hipEvent_t ev[2];
    for (int i = 0; i < 2; ++i) {
        if (!ctx->event_pool.empty()) {
            ev[i] = ctx->event_pool.back();
            ctx->event_pool.pop_back();
        } else if (hipEventCreate(&ev[i]) != hipSuccess) {
            return -1;
        }
    }
    (void)hipEventRecord(ev[0], ctx->stream);
    ctx->pending.push_back({kind, ev[0], ev[1]});
    return (int)ctx->pending.size() - 1;
}

void prof_end(sodhip_ctx ctx, int token)
{
    if (token < 0) return;
    (void)hipEventRecord(ctx->pending[token].b, ctx->stream);
}

static void prof_drain(sodhip_ctx ctx)
{
    if (ctx->pending.empty()) return;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &p : ctx->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            ctx->slots[p.kind].total_ms += ms;
            ctx->slots[p.kind].count += 1;
        }
        ctx->event_pool.push_back(p.a);
        ctx->event_pool.push_back(p.b);
    }
    ctx->pending.clear();
}

} // namespace sodhip

using namespace sodhip;

extern "C" int sodhip_version(void) { return 10000; }

extern "C" int sodhip_last_error(char *buf, int len)
{
    if (buf && len > 0) {
        strncpy(buf, g_err, (size_t)len - 1);
        buf[len - 1] = '\0';
    }
    return g_err[0] ? 1 : 0;
}

extern "C" int sodhip_create(int device, sodhip_ctx *out)
{
    SODHIP_REQUIRE(out, "null context pointer");
    int n = 0;
    SODHIP_TRY_HIP(hipGetDeviceCount(&n));
    SODHIP_REQUIRE(device >= 0 && device < n, "device %d not present (%d devices)", device, n);
    SODHIP_TRY_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    SODHIP_TRY_HIP(hipGetDeviceProperties(&prop, device));
    int rc = noalloc_init(device);
    if (rc) return rc;
    auto *c = new sodhip_ctx_s();
    c->device = device;
    c->stream = nullptr;
    c->num_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    if (const char *v = getenv("SODHIP_STAGE_CLUSTERS")) c->cluster_stage = atoi(v) != 0;
    *out = c;
    return SODHIP_OK;
}

extern "C" int sodhip_destroy(sodhip_ctx ctx)
{
    if (!ctx) return SODHIP_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &m : ctx->meshes) free_mesh(m.get());
    for (auto &kv : ctx->present) (void)hipFree(kv.second.dev);
    (void)hipFree(ctx->scratch.ptr);
    (void)hipFree(ctx->scratch2.ptr);
    (void)hipFree(ctx->diag.ptr);
    if (ctx->hostbuf.ptr) (void)hipHostFree(ctx->hostbuf.ptr);
    for (auto &p : ctx->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    delete ctx;
    return SODHIP_OK;
}

// Drops every cached table derived from device memory in [lo, hi): meshes keyed
// on a connectivity there, and working-node incidences keyed on a list there.
static int forget_range(sodhip_ctx ctx, uintptr_t lo, uintptr_t hi)
{
    bool synced = false;
    auto sync = [&]() -> int {
        if (!synced) SODHIP_TRY_HIP(hipStreamSynchronize(ctx->stream));
        synced = true;
        return SODHIP_OK;
    };
    int rc;
    for (auto it = ctx->meshes.begin(); it != ctx->meshes.end();) {
        MeshCache *m = it->get();
        const uintptr_t k = (uintptr_t)m->connec_key;
        if (k >= lo && k < hi) {
            if ((rc = sync())) return rc;
            free_mesh(m);
            it = ctx->meshes.erase(it);
            continue;
        }
        const uintptr_t w = (uintptr_t)m->lpw_key;
        if (m->wrow && w >= lo && w < hi) {
            if ((rc = sync())) return rc;
            free_clusters(m);
            (void)hipFree(m->wrow);
            (void)hipFree(m->wsrc);
            (void)hipFree(m->wflag);
            m->wrow = m->wsrc = nullptr;
            m->wflag = nullptr;
            m->lpw_key = nullptr;
            m->npoin_w = 0;
        }
        ++it;
    }
    return SODHIP_OK;
}

extern "C" int sodhip_forget_mesh(sodhip_ctx ctx, const void *dev)
{
    SODHIP_REQUIRE(ctx, "null context");
    if (!dev) return SODHIP_OK;
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    return forget_range(ctx, (uintptr_t)dev, (uintptr_t)dev + 1);
}

extern "C" int sodhip_reserve_scratch(sodhip_ctx ctx, size_t scratch_bytes, size_t scratch2_bytes)
{
    SODHIP_REQUIRE(ctx, "null context");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    int rc = scratch_bytes ? ensure_buf(ctx->scratch, scratch_bytes) : SODHIP_OK;
    if (!rc && scratch2_bytes) rc = ensure_buf(ctx->scratch2, scratch2_bytes);
    return rc;
}

extern "C" int sodhip_pin_scratch(sodhip_ctx ctx, int pin)
{
    SODHIP_REQUIRE(ctx, "null context");
    ctx->scratch.pinned = pin != 0;
    ctx->scratch2.pinned = pin != 0;
    return SODHIP_OK;
}

extern "C" int sodhip_set_stage_config(sodhip_ctx ctx, int elem_blocks_per_cu, int geo_rows16,
                                       int gather_threads)
{
    SODHIP_REQUIRE(ctx, "null context");
    SODHIP_REQUIRE(elem_blocks_per_cu >= 1 && elem_blocks_per_cu <= 64,
                   "elem_blocks_per_cu %d outside 1..64", elem_blocks_per_cu);
    SODHIP_REQUIRE(geo_rows16 == 0 || geo_rows16 == 1, "geo_rows16 must be 0 or 1");
    SODHIP_REQUIRE(gather_threads == 64 || gather_threads == 128 || gather_threads == 256,
                   "gather_threads %d not 64, 128 or 256", gather_threads);
    ctx->stage_bpc = elem_blocks_per_cu;
    ctx->stage_geo16 = geo_rows16;
    ctx->gather_threads = gather_threads;
    return SODHIP_OK;
}

extern "C" int sodhip_get_stage_config(sodhip_ctx ctx, int *elem_blocks_per_cu, int *geo_rows16,
                                       int *gather_threads)
{
    SODHIP_REQUIRE(ctx, "null context");
    if (elem_blocks_per_cu) *elem_blocks_per_cu = ctx->stage_bpc;
    if (geo_rows16) *geo_rows16 = ctx->stage_geo16;
    if (gather_threads) *gather_threads = ctx->gather_threads;
    return SODHIP_OK;
}

extern "C" int sodhip_set_stream(sodhip_ctx ctx, void *stream)
{
    SODHIP_REQUIRE(ctx, "null context");
    ctx->stream = (hipStream_t)stream;
    return SODHIP_OK;
}

extern "C" int sodhip_synchronize(sodhip_ctx ctx)
{
    SODHIP_REQUIRE(ctx, "null context");
    SODHIP_TRY_HIP(hipStreamSynchronize(ctx->stream));
    return SODHIP_OK;
}

// ---------------------------------------------------------------- present table
static PresentEntry *find_present(sodhip_ctx ctx, const void *host, size_t *offset)
{
    uintptr_t h = (uintptr_t)host;
    auto it = ctx->present.upper_bound(h);
    if (it == ctx->present.begin()) return nullptr;
    --it;
    if (h >= it->first && h < it->first + it->second.bytes) {
        if (offset) *offset = h - it->first;
        return &it->second;
    }
    return nullptr;
}

extern "C" int sodhip_enter_data(sodhip_ctx ctx, const void *host, size_t bytes, int copyin)
{
    SODHIP_REQUIRE(ctx && host && bytes > 0, "invalid enter_data arguments");
    if (find_present(ctx, host, nullptr))
        return set_error(SODHIP_ERR_ARG, "host range at %p is already present", host);
    PresentEntry e;
    e.bytes = bytes;
    hipError_t rc = hipMalloc(&e.dev, bytes);
    if (rc != hipSuccess) return set_error(SODHIP_ERR_ALLOC, "hipMalloc(%zu) failed", bytes);
    if (copyin) {
        SODHIP_TRY_HIP(hipMemcpyAsync(e.dev, host, bytes, hipMemcpyHostToDevice, ctx->stream));
        SODHIP_TRY_HIP(hipStreamSynchronize(ctx->stream));
    }
    ctx->present[(uintptr_t)host] = e;
    return SODHIP_OK;
}

extern "C" int sodhip_update_device(sodhip_ctx ctx, const void *host, size_t bytes)
{
    size_t off = 0;
    PresentEntry *e = ctx ? find_present(ctx, host, &off) : nullptr;
    if (!e) return set_error(SODHIP_ERR_NOT_PRESENT, "host address %p is not present", host);
    SODHIP_REQUIRE(off + bytes <= e->bytes, "update_device beyond the present range");
    SODHIP_TRY_HIP(hipMemcpyAsync((char *)e->dev + off, host, bytes, hipMemcpyHostToDevice,
                                  ctx->stream));
    SODHIP_TRY_HIP(hipStreamSynchronize(ctx->stream));
    return SODHIP_OK;
}

extern "C" int sodhip_update_host(sodhip_ctx ctx, void *host, size_t bytes)
{
    size_t off = 0;
    PresentEntry *e = ctx ? find_present(ctx, host, &off) : nullptr;
    if (!e) return set_error(SODHIP_ERR_NOT_PRESENT, "host address %p is not present", host);
    SODHIP_REQUIRE(off + bytes <= e->bytes, "update_host beyond the present range");
    SODHIP_TRY_HIP(hipMemcpyAsync(host, (char *)e->dev + off, bytes, hipMemcpyDeviceToHost,
                                  ctx->stream));
    SODHIP_TRY_HIP(hipStreamSynchronize(ctx->stream));
    return SODHIP_OK;
}

extern "C" int sodhip_exit_data(sodhip_ctx ctx, const void *host, int copyout)
{
    SODHIP_REQUIRE(ctx, "null context");
    auto it = ctx->present.find((uintptr_t)host);
    if (it == ctx->present.end())
        return set_error(SODHIP_ERR_NOT_PRESENT, "host address %p is not the start of a present range",
                         host);
    if (copyout) {
        SODHIP_TRY_HIP(hipMemcpyAsync((void *)host, it->second.dev, it->second.bytes,
                                      hipMemcpyDeviceToHost, ctx->stream));
    }
    SODHIP_TRY_HIP(hipStreamSynchronize(ctx->stream));
    {
        const uintptr_t lo = (uintptr_t)it->second.dev;
        int rc = forget_range(ctx, lo, lo + it->second.bytes);
        if (rc) return rc;
    }
    SODHIP_TRY_HIP(hipFree(it->second.dev));
    ctx->present.erase(it);
    return SODHIP_OK;
}

extern "C" int sodhip_copy_present(sodhip_ctx ctx, const void *src_host, void *dst_host,
                                   size_t bytes)
{
    size_t so = 0, d_off = 0;
    PresentEntry *s = ctx ? find_present(ctx, src_host, &so) : nullptr;
    PresentEntry *d = ctx ? find_present(ctx, dst_host, &d_off) : nullptr;
    if (!s || !d)
        return set_error(SODHIP_ERR_NOT_PRESENT, "host address %p is not present",
                         s ? dst_host : src_host);
    SODHIP_REQUIRE(so + bytes <= s->bytes && d_off + bytes <= d->bytes,
                   "copy_present beyond a present range");
    SODHIP_TRY_HIP(hipSetDevice(ctx->device));
    SODHIP_TRY_HIP(hipMemcpyAsync((char *)d->dev + d_off, (const char *)s->dev + so, bytes,
                                  hipMemcpyDeviceToDevice, ctx->stream));
    return SODHIP_OK;
}

extern "C" void *sodhip_deviceptr(sodhip_ctx ctx, const void *host)
{
    if (!ctx || !host) return nullptr;
    size_t off = 0;
    PresentEntry *e = find_present(ctx, host, &off);
    if (!e) {
        set_error(SODHIP_ERR_NOT_PRESENT, "host address %p is not present", host);
        return nullptr;
    }
    return (char *)e->dev + off;
}

// ---------------------------------------------------------------- timing
extern "C" int sodhip_prof_enable(sodhip_ctx ctx, int enable)
{
    SODHIP_REQUIRE(ctx, "null context");
    ctx->prof = enable != 0;
    return SODHIP_OK;
}

extern "C" int sodhip_prof_get(sodhip_ctx ctx, int kind, double *total_ms, long long *count)
{
    SODHIP_REQUIRE(ctx && kind >= 0 && kind < kProfKinds, "invalid prof query");
    prof_drain(ctx);
    if (total_ms) *total_ms = ctx->slots[kind].total_ms;
    if (count) *count = ctx->slots[kind].count;
    return SODHIP_OK;
}

extern "C" int sodhip_prof_reset(sodhip_ctx ctx)
{
    SODHIP_REQUIRE(ctx, "null context");
    prof_drain(ctx);
    for (auto &s : ctx->slots) s = ProfSlot();
    return SODHIP_OK;
}
