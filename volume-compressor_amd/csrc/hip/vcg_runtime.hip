// vcg_runtime.hip — device selection, per-thread streams and scratch, device
// memory helpers and the device-resident archive.
#include <hip/hip_runtime.h>
#include <mutex>
#include <vector>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "vcg_dev.hpp"
#include "vcg_xform.hpp"

vcg_dtab *g_dtab = nullptr;
static std::once_flag g_init_once;
static int g_init_rc = -1;

void vcg_fail(const char *what, hipError_t e, const char *file, int line) {
    char msg[512];
    snprintf(msg, sizeof msg, "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
    vc__panic(msg);
    __builtin_unreachable();
}

extern "C" int vcg_copy(void *dst, const void *src, size_t bytes, int kind);

extern "C" int vcg_init(const vcg_tables *t) {
    std::call_once(g_init_once, [t] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) { g_init_rc = -1; return; }
        VCG_CHECK(hipSetDevice(0));
        // the transform kernels stage one atom as a 32x32x33 i32 tile (132 KiB + a
        // few shared words); gfx950 provides 160 KiB of LDS per workgroup
        int lds_block = 0, lds_optin = 0, lds_cu = 0;
        VCG_CHECK(hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, 0));
        (void)hipDeviceGetAttribute(&lds_optin, hipDeviceAttributeSharedMemPerBlockOptin, 0);
        (void)hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, 0);
        (void)hipGetLastError();
        const int need = 32 * 32 * 33 * 4 + 64;
        if (getenv("VC_VERBOSE"))
            fprintf(stderr, "vcg: LDS per block %d, opt-in %d, per CU %d, need %d\n", lds_block, lds_optin, lds_cu, need);
        if ((lds_block > lds_optin ? lds_block : lds_optin) < need) {
            char msg[160];
            snprintf(msg, sizeof msg, "device LDS per workgroup %d/%d bytes < %d required by the transform kernels",
                     lds_block, lds_optin, need);
            vc__panic(msg);
        }
        // the forward and inverse transforms regroup their even-row sums by
        // these exact symmetries of the basis (they hold for the Q14 table by
        // construction)
        for (int k = 0; k < 32; ++k)
            for (int n = 0; n < 16; ++n) {
                const int sgn = (k & 1) ? -1 : 1;
                bool ok = t->cmat[k][31 - n] == sgn * t->cmat[k][n];
                if ((k & 1) == 0 && n < 8) ok = ok && t->cmat[k][15 - n] == ((k & 2) ? -1 : 1) * t->cmat[k][n];
                if ((k & 3) == 0 && n < 4) ok = ok && t->cmat[k][7 - n] == ((k & 4) ? -1 : 1) * t->cmat[k][n];
                if ((k & 7) == 0 && n < 2) ok = ok && t->cmat[k][3 - n] == ((k & 8) ? -1 : 1) * t->cmat[k][n];
                if (!ok) vc__panic("DCT basis lacks the symmetries the inverse transform relies on");
            }
        // the table image is 240 KB: built in static storage, not on the
        // calling thread's stack (this runs once)
        static vcg_dtab h;
        memcpy(h.cmat, t->cmat, sizeof h.cmat);
        memcpy(h.gstep, t->gstep, sizeof h.gstep);
        memcpy(h.band_start, t->band_start, sizeof h.band_start);
        memcpy(h.scan, t->scan, sizeof h.scan);
        for (int q = 0; q < VCG_A3; ++q)
            h.fscan[q] = (uint8_t)((h.scan[q] >> 10) + ((h.scan[q] >> 5) & 31) + (h.scan[q] & 31));
        vcg_fwd1_tables(h.cmat, h.fa, h.rowsum);
        vcg_im_tables(h.cmat, h.ib);
        for (int q = 0; q < VCG_A3; ++q) h.imtab[q] = vcg_im_tab(h.scan[q]);
        // the allocation-free decode's static scratch: its device address
        if (vcg_noalloc_setup() != 0) vc__panic("hipGetSymbolAddress failed for the decoder's static scratch");
        VCG_CHECK(hipMalloc(&g_dtab, sizeof(vcg_dtab)));
        // through this thread's stream: the library never touches the null
        // stream, which would take one of the device's few hardware queues
        vcg_copy(g_dtab, &h, sizeof h, 1);
        g_init_rc = 0;
    });
    return g_init_rc;
}

// ---- per-thread contexts: created on first use, recycled when the thread exits
static std::mutex g_pool_mu;
static std::vector<vcg_ctx *> g_pool;

struct ctx_holder {
    vcg_ctx *c = nullptr;
    ~ctx_holder() {
        if (c) { std::lock_guard<std::mutex> g(g_pool_mu); g_pool.push_back(c); }
    }
};
static thread_local ctx_holder tl_ctx;

__global__ void k_reset_word(uint32_t *w);

vcg_ctx *vcg_get_ctx() {
    if (tl_ctx.c) return tl_ctx.c;
    vcg_ctx *r = nullptr;
    {
        std::lock_guard<std::mutex> g(g_pool_mu);
        if (!g_pool.empty()) { r = g_pool.back(); g_pool.pop_back(); }
    }
    if (r) {
        // a recycled context starts clean: its last owner's work has finished,
        // its status word no longer holds that owner's corrupt atoms (list
        // launches report per item and never collect the word), and the
        // owner's stage events (which it may have destroyed) are dropped
        for (int k = 0; k < VCG_SLOTS; ++k) VCG_CHECK(hipStreamSynchronize(r->sE[k]));
        if (r->sC) VCG_CHECK(hipStreamSynchronize(r->sC));
        for (int k = 0; k < 3; ++k) r->ev[k] = nullptr;
        hipLaunchKernelGGL(k_reset_word, dim3(1), dim3(1), 0, r->s, r->status_d);
        VCG_CHECK(hipGetLastError());
        VCG_CHECK(hipStreamSynchronize(r->s));
        tl_ctx.c = r;
        return r;
    }
    vcg_ctx *c = new vcg_ctx();
    VCG_CHECK(hipSetDevice(0));
    VCG_CHECK(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    for (int k = 0; k < VCG_SLOTS; ++k) VCG_CHECK(hipStreamCreateWithFlags(&c->sE[k], hipStreamNonBlocking));
    VCG_CHECK(hipEventCreateWithFlags(&c->evF, hipEventDisableTiming));
    for (int k = 0; k < VCG_SLOTS; ++k) VCG_CHECK(hipEventCreateWithFlags(&c->evI[k], hipEventDisableTiming));
    VCG_CHECK(hipMalloc(&c->status_d, 4));
    VCG_CHECK(hipHostMalloc((void **)&c->status_h, 4, hipHostMallocDefault));
    VCG_CHECK(hipMemsetAsync(c->status_d, 0, 4, c->s));
    VCG_CHECK(hipStreamSynchronize(c->s));
    tl_ctx.c = c;
    return c;
}

// Grow a device buffer to hold `need` elements (contents are not preserved).
void vcg_grow(void **p, size_t *cap, size_t need, size_t elem) {
    if (need <= *cap) return;
    vcg_ctx *c = vcg_get_ctx();
    if (*p) {
        VCG_CHECK(hipStreamSynchronize(c->s));
        VCG_CHECK(hipFree(*p));
        *p = nullptr;
    }
    size_t n = need + need / 4;
    if (hipMalloc(p, n * elem) != hipSuccess) {
        (void)hipGetLastError();
        *cap = 0; *p = nullptr;
        vc__panic("out of device memory (scratch)");
    }
    *cap = n;
}

extern "C" void *vcg_stream(void) { return (void *)vcg_get_ctx()->s; }

extern "C" void *vcg_alloc(size_t bytes) {
    void *p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) { (void)hipGetLastError(); return nullptr; }
    return p;
}
extern "C" void vcg_free(void *p) { if (p) VCG_CHECK(hipFree(p)); }

extern "C" int vcg_copy(void *dst, const void *src, size_t bytes, int kind) {
    vcg_ctx *c = vcg_get_ctx();
    hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    VCG_CHECK(hipMemcpyAsync(dst, src, bytes, k, c->s));
    VCG_CHECK(hipStreamSynchronize(c->s));
    return 0;
}

static hipStream_t slot_stream(vcg_ctx *c, int slot) {
    return slot >= 0 && slot < VCG_SLOTS ? c->sE[slot] : c->s;
}

extern "C" int vcg_copy_async(void *dst, const void *src, size_t bytes, int kind, int slot) {
    vcg_ctx *c = vcg_get_ctx();
    hipMemcpyKind k = kind == 0 ? hipMemcpyDefault : kind == 1 ? hipMemcpyHostToDevice
                    : kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    VCG_CHECK(hipMemcpyAsync(dst, src, bytes, k, slot_stream(c, slot)));
    return 0;
}

extern "C" int vcg_ptr_is_device(const void *p) {
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess) { (void)hipGetLastError(); return 0; }
    return at.type == hipMemoryTypeDevice ? 1 : 0;
}

extern "C" int vcg_mem_info(size_t *free_bytes, size_t *total_bytes) {
    vcg_get_ctx();                      // the library's device is current
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) { (void)hipGetLastError(); return -1; }
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return 0;
}

extern "C" void *vcg_event_create(void) {
    hipEvent_t e = nullptr;
    VCG_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return (void *)e;
}
extern "C" void vcg_event_destroy(void *e) { if (e) VCG_CHECK(hipEventDestroy((hipEvent_t)e)); }
extern "C" void vcg_event_record(void *e, int slot) {
    vcg_ctx *c = vcg_get_ctx();
    VCG_CHECK(hipEventRecord((hipEvent_t)e, slot_stream(c, slot)));
}
extern "C" void vcg_event_sync(void *e) { VCG_CHECK(hipEventSynchronize((hipEvent_t)e)); }
extern "C" int vcg_event_done(void *e) {
    const hipError_t r = hipEventQuery((hipEvent_t)e);
    if (r == hipSuccess) return 1;
    if (r == hipErrorNotReady) { (void)hipGetLastError(); return 0; }
    vcg_fail("hipEventQuery", r, __FILE__, __LINE__);
    return 0;
}
extern "C" void *vcg_host_alloc(size_t bytes) {
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) { (void)hipGetLastError(); return nullptr; }
    return p;
}
extern "C" void vcg_host_free(void *p) { if (p) VCG_CHECK(hipHostFree(p)); }

extern "C" void vcg_set_stage_events(void *ev[3]) {
    vcg_ctx *c = vcg_get_ctx();
    for (int i = 0; i < 3; ++i) c->ev[i] = ev ? ev[i] : nullptr;
}

__global__ void k_reset_word(uint32_t *w) { *w = 0; }

// plain 16-byte-per-lane copy: the reference point for HBM byte counters
__global__ __launch_bounds__(256) void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

extern "C" int vcg_copy_kernel(void *dst, const void *src, size_t bytes) {
    if (bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16) return -1;
    vcg_ctx *c = vcg_get_ctx();
    const size_t n = bytes / 16;
    size_t blocks = (n + 255) / 256;
    if (blocks > 256 * 64) blocks = 256 * 64;
    if (!blocks) return 0;
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, c->s, (const uint4 *)src, (uint4 *)dst, n);
    VCG_CHECK(hipGetLastError());
    VCG_CHECK(hipStreamSynchronize(c->s));
    return 0;
}

extern "C" int vcg_sync(void) {
    vcg_ctx *c = vcg_get_ctx();
    VCG_CHECK(hipMemcpyAsync(c->status_h, c->status_d, 4, hipMemcpyDeviceToHost, c->s));
    hipLaunchKernelGGL(k_reset_word, dim3(1), dim3(1), 0, c->s, c->status_d);
    VCG_CHECK(hipGetLastError());
    VCG_CHECK(hipStreamSynchronize(c->s));
    return (*c->status_h & 1u) ? 1 : 0;
}

// ---- device-resident archive
struct vcg_arc {
    uint8_t  *buf = nullptr;   // archive bytes in HBM
    vcg_geom *geom = nullptr;  // geometry + dequant tables in HBM
    vcg_geom  host;            // host copy (launch bounds)
};

extern "C" vcg_arc *vcg_arc_create(const uint8_t *buf, const vcg_geom *g) {
    vcg_ctx *c = vcg_get_ctx();
    vcg_arc *a = new vcg_arc();
    a->host = *g;
    // padded: the entropy kernel reads 16-byte look-ahead blocks up to 15 bytes past a payload
    if (hipMalloc(&a->buf, g->len + 64) != hipSuccess ||
        hipMalloc(&a->geom, sizeof(vcg_geom)) != hipSuccess) {
        (void)hipGetLastError();
        if (a->buf) (void)hipFree(a->buf);
        delete a;
        return nullptr;
    }
    VCG_CHECK(hipMemsetAsync(a->buf + g->len, 0, 64, c->s));
    VCG_CHECK(hipMemcpyAsync(a->buf, buf, g->len, hipMemcpyHostToDevice, c->s));
    VCG_CHECK(hipMemcpyAsync(a->geom, g, sizeof(vcg_geom), hipMemcpyHostToDevice, c->s));
    VCG_CHECK(hipStreamSynchronize(c->s));
    return a;
}

// Region-blob arena of a streaming archive: g->len bytes of which the first
// head_len come from head (header and directory); blobs arrive by vcg_arc_put.
extern "C" vcg_arc *vcg_arc_create_arena(const vcg_geom *g, const uint8_t *head, uint64_t head_len) {
    if (head_len > g->len) return nullptr;
    vcg_ctx *c = vcg_get_ctx();
    vcg_arc *a = new vcg_arc();
    a->host = *g;
    if (hipMalloc(&a->buf, g->len + 64) != hipSuccess ||
        hipMalloc(&a->geom, sizeof(vcg_geom)) != hipSuccess) {
        (void)hipGetLastError();
        if (a->buf) (void)hipFree(a->buf);
        delete a;
        return nullptr;
    }
    VCG_CHECK(hipMemsetAsync(a->buf + g->len, 0, 64, c->s));
    VCG_CHECK(hipMemcpyAsync(a->buf, head, head_len, hipMemcpyHostToDevice, c->s));
    VCG_CHECK(hipMemcpyAsync(a->geom, g, sizeof(vcg_geom), hipMemcpyHostToDevice, c->s));
    VCG_CHECK(hipStreamSynchronize(c->s));
    return a;
}

extern "C" int vcg_arc_put(vcg_arc *a, uint64_t off, const void *src, uint64_t n) {
    if (off > a->host.len || n > a->host.len - off) return -1;
    vcg_ctx *c = vcg_get_ctx();
    VCG_CHECK(hipMemcpyAsync(a->buf + off, src, n, hipMemcpyHostToDevice, c->s));
    VCG_CHECK(hipStreamSynchronize(c->s));
    return 0;
}

// A blob read from file offset file_off now sits at arena offset at: point its
// present slots at the arena.  A slot whose payload the reader would reject
// (outside the file, or longer than the streaming bound src/vc/vc.c:1276-1277)
// gets the invalid offset, which the decoder reports as a corrupt atom; one
// whose payload lies in the file but outside this blob sets *irr.
__global__ __launch_bounds__(1024) void k_blob_rebase(uint8_t *__restrict__ blob, uint64_t at, uint64_t file_off,
                                                      uint32_t blob_len, uint64_t file_len, uint32_t *irr) {
    const uint32_t s = blockIdx.x * 1024u + threadIdx.x;
    uint4 *p = (uint4 *)(blob + (uint64_t)s * 16u);
    uint4 v = *p;
    if ((v.w & 0xFFu) != 1u) return;
    const uint64_t o = (uint64_t)v.x | (uint64_t)v.y << 32;
    const uint32_t l = v.z;
    if (o == ~0ull || l < 1u) return;
    uint64_t n;
    if (o > file_len || (uint64_t)l > file_len - o || l > VCG_RC_CAP) {
        n = ~0ull;
    } else if (o < file_off + 32768ull * 16u || o + l > file_off + blob_len) {
        atomicOr(irr, 1u);
        return;
    } else {
        n = o - file_off + at;
    }
    v.x = (uint32_t)n; v.y = (uint32_t)(n >> 32);
    *p = v;
}

extern "C" int vcg_arc_rebase(vcg_arc *a, uint64_t at, uint64_t file_off, uint32_t blob_len, uint64_t file_len) {
    if ((at & 15u) || blob_len < 32768u * 16u || at > a->host.len || blob_len > a->host.len - at) return -1;
    vcg_ctx *c = vcg_get_ctx();
    uint32_t *dirr = nullptr, irr = 0;
    if (hipMalloc(&dirr, 4) != hipSuccess) { (void)hipGetLastError(); return -1; }
    VCG_CHECK(hipMemsetAsync(dirr, 0, 4, c->s));
    hipLaunchKernelGGL(k_blob_rebase, dim3(32), dim3(1024), 0, c->s, a->buf + at, at, file_off, blob_len, file_len, dirr);
    VCG_CHECK(hipGetLastError());
    VCG_CHECK(hipMemcpyAsync(&irr, dirr, 4, hipMemcpyDeviceToHost, c->s));
    VCG_CHECK(hipStreamSynchronize(c->s));
    VCG_CHECK(hipFree(dirr));
    return irr ? 1 : 0;
}

extern "C" void vcg_device_sync(void) { VCG_CHECK(hipDeviceSynchronize()); }

extern "C" void vcg_arc_destroy(vcg_arc *a) {
    if (!a) return;
    VCG_CHECK(hipDeviceSynchronize());
    VCG_CHECK(hipFree(a->buf));
    VCG_CHECK(hipFree(a->geom));
    delete a;
}

// accessors for the other translation units
const uint8_t *vcg_arc_buf(const vcg_arc *a) { return a->buf; }
const vcg_geom *vcg_arc_geom_dev(const vcg_arc *a) { return a->geom; }
const vcg_geom *vcg_arc_geom_host(const vcg_arc *a) { return &a->host; }
