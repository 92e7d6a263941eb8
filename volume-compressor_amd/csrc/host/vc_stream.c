/* vc_stream.c — streaming v2 archives (vc_open_streaming) decoded from HBM.
 *
 * A v2 region is one contiguous blob [512 KiB slot array][payloads] that the
 * directory locates (src/vc/vc.c:1451-1471, vc_region_blob_range), so a
 * reader can fetch it with one range read.  Here every decode call first makes
 * the regions it touches resident in a device arena [header][directory]
 * [blobs ...]: each missing blob is read by one call of the caller's read
 * function into pinned memory, copied to the arena and rebased there (its
 * present slots then address the arena), and its directory entry in the arena
 * switched from ABSENT to DATA.  The decode kernels then run on the arena
 * exactly as on an in-memory archive.  When the blobs a call needs do not fit
 * beside the resident ones, the arena is emptied first (after the device has
 * finished every decode in flight).
 *
 * A call falls back to the host-resolved path (per-atom reads, as before)
 * when a region it needs cannot be served from one blob: a blob outside the
 * file, a read that fails, a payload that lies in the file outside its
 * region's blob, or blobs that do not fit the arena at all.  Results are the
 * same either way.
 */
#include <stdlib.h>
#include <string.h>

#include "vc_host.h"

#define ARENA_DEFAULT (8ull << 30)

#define SDEV_EVENTS 64

typedef struct vc_sdev {
    pthread_mutex_t mu;            /* held from a call's residency check through its launches */
    vcg_arc *arc;
    uint64_t cap, head, cursor;    /* arena bytes, header + directory bytes, next free byte */
    uint64_t nreg;
    uint8_t *dir;                  /* the file's directory */
    uint8_t *dir0;                 /* the arena's directory when empty (data regions ABSENT) */
    uint64_t *res;                 /* arena offset of each resident blob, 0 = not resident */
    uint8_t *irr;                  /* region served by the host path only */
    uint32_t *mark; uint32_t gen;  /* regions touched by the current call */
    uint64_t *need; uint64_t nneed;
    uint8_t *stage; uint64_t stage_cap;   /* pinned */
    uint64_t reads, evictions;     /* statistics */
    /* one event per recent call that launched on the arena, recorded on the
     * calling thread's stream after its launches: an eviction waits for these
     * (not for the whole device) before the arena is overwritten */
    void *ev[SDEV_EVENTS];
    uint32_t ev_next;
} vc_sdev;

static uint64_t align16(uint64_t x) { return (x + 15u) & ~15ull; }

void vc__stream_free(vc_archive *a) {
    vc_sdev *s = (vc_sdev *)a->sdev;
    if (!s) return;
    for (int k = 0; k < SDEV_EVENTS; ++k) if (s->ev[k]) vcg_event_destroy(s->ev[k]);
    if (s->arc) vcg_arc_destroy(s->arc);
    if (s->stage) vcg_host_free(s->stage);
    free(s->dir); free(s->dir0); free(s->res); free(s->irr); free(s->mark); free(s->need);
    pthread_mutex_destroy(&s->mu);
    free(s);
    a->sdev = NULL;
}

/* the arena of a streaming v2 archive, created on first use; NULL when the
 * archive is not one or the arena cannot be set up */
static vc_sdev *sdev_get(vc_archive *a) {
    if (a->buf || a->version != 2 || !a->read) return NULL;
    pthread_mutex_lock(&a->mu);
    vc_sdev *s = (vc_sdev *)a->sdev;
    if (s || a->sdev_failed) { pthread_mutex_unlock(&a->mu); return s; }
    vc__require_gpu();
    const uint64_t nreg = a->region_prefix[a->nlod];
    const uint64_t head = align16(FILE_HDR + nreg * 16);
    /* default: the whole archive fits (blobs are 16-byte aligned here) */
    const uint64_t whole = a->len + 16 * nreg + 64;
    uint64_t cap = a->stream_cache ? a->stream_cache : (whole < ARENA_DEFAULT ? whole : ARENA_DEFAULT);
    if (cap < head + L2_BLOCK) cap = head + L2_BLOCK;
    s = (vc_sdev *)calloc(1, sizeof *s);
    uint8_t *hb = (uint8_t *)calloc(1, head);
    if (!s || !hb) goto fail;
    s->nreg = nreg; s->head = head; s->cap = cap; s->cursor = head;
    s->dir = (uint8_t *)malloc(nreg * 16 + 1); s->dir0 = (uint8_t *)malloc(nreg * 16 + 1);
    s->res = (uint64_t *)calloc(nreg + 1, 8); s->irr = (uint8_t *)calloc(nreg + 1, 1);
    s->mark = (uint32_t *)calloc(nreg + 1, 4); s->need = (uint64_t *)malloc((nreg + 1) * 8);
    if (!s->dir || !s->dir0 || !s->res || !s->irr || !s->mark || !s->need) goto fail;
    /* header and directory: two reads */
    if (vc__rd(a, 0, FILE_HDR, hb) != 0) goto fail;
    if (nreg && (nreg * 16 > 0xFFFFFFFFull || vc__rd(a, a->dir_off, (uint32_t)(nreg * 16), s->dir) != 0)) goto fail;
    memcpy(s->dir0, s->dir, nreg * 16);
    for (uint64_t r = 0; r < nreg; ++r) {
        uint8_t *e = s->dir0 + 16 * r;
        const uint32_t f = rd32(e + 12);
        if (f != 0 && f != 1) memset(e, 0, 16);                  /* data: ABSENT until fetched */
    }
    memcpy(hb + FILE_HDR, s->dir0, nreg * 16);
    vcg_geom *g = (vcg_geom *)calloc(1, sizeof *g);
    if (!g) goto fail;
    g->len = cap; g->version = 2; g->nlod = a->nlod; g->dir_off = FILE_HDR;
    for (int l = 0; l <= VC_NLOD; ++l) g->region_prefix[l] = a->region_prefix[l];
    for (int l = 0; l < a->nlod; ++l) {
        vc_dims d = vc__lod_dims(a->dims0, l);
        g->lnx[l] = d.nx; g->lny[l] = d.ny; g->lnz[l] = d.nz;
        g->acx[l] = (d.nx + VC_A - 1) / VC_A; g->acy[l] = (d.ny + VC_A - 1) / VC_A; g->acz[l] = (d.nz + VC_A - 1) / VC_A;
        vc__region_grid(a->dims0, l, &g->nrz[l], &g->nry[l], &g->nrx[l]);
        vc__dq(a->base_q[l], g->dq_step[l], g->dq_bias[l]);
    }
    s->arc = vcg_arc_create_arena(g, hb, head);
    free(g);
    if (!s->arc) goto fail;
    free(hb);
    pthread_mutex_init(&s->mu, NULL);
    a->sdev = s;
    pthread_mutex_unlock(&a->mu);
    return s;
fail:
    free(hb);
    if (s) {
        free(s->dir); free(s->dir0); free(s->res); free(s->irr); free(s->mark); free(s->need);
        free(s);
    }
    a->sdev_failed = 1;
    pthread_mutex_unlock(&a->mu);
    return NULL;
}

static void touch(vc_sdev *s, const vc_archive *a, int lod, uint32_t rz, uint32_t ry, uint32_t rx) {
    uint32_t nrz, nry, nrx;
    vc__region_grid(a->dims0, lod, &nrz, &nry, &nrx);
    if (rz >= nrz || ry >= nry || rx >= nrx) return;
    const uint64_t ri = a->region_prefix[lod] + ((uint64_t)rz * nry + ry) * nrx + rx;
    if (ri >= s->nreg || s->mark[ri] == s->gen) return;
    s->mark[ri] = s->gen;
    s->need[s->nneed++] = ri;
}

/* every launch that read the arena has finished (called with s->mu held) */
static void arena_idle(vc_sdev *s) {
    for (int k = 0; k < SDEV_EVENTS; ++k) if (s->ev[k]) vcg_event_sync(s->ev[k]);
}

static void evict(vc_sdev *s) {
    arena_idle(s);                              /* no decode in flight still reads the arena */
    if (vcg_arc_put(s->arc, FILE_HDR, s->dir0, s->nreg * 16) != 0) vc__panic("stream arena reset");
    memset(s->res, 0, s->nreg * 8);
    s->cursor = s->head;
    s->evictions++;
}

/* make the touched regions resident; 0 = the arena serves the call, 1 = use
 * the host path.  Called with s->mu held. */
static int make_resident(vc_archive *a, vc_sdev *s) {
    uint64_t want = 0;
    for (uint64_t k = 0; k < s->nneed; ++k) {
        const uint64_t ri = s->need[k];
        if (s->irr[ri]) return 1;
        const uint8_t *e = s->dir + 16 * ri;
        const uint32_t f = rd32(e + 12);
        if (f == 0 || f == 1) continue;         /* ABSENT / ZERO: the arena directory says so */
        const uint64_t off = rd64(e), len = rd32(e + 8);
        if (len < L2_BLOCK || off > a->len || len > a->len - off) { s->irr[ri] = 1; return 1; }
        want += align16(len);
    }
    if (want > s->cap - s->head) return 1;
    uint64_t missing = 0;
    for (uint64_t k = 0; k < s->nneed; ++k) {
        const uint64_t ri = s->need[k];
        const uint32_t f = rd32(s->dir + 16 * ri + 12);
        if (f != 0 && f != 1 && !s->res[ri]) missing += align16(rd32(s->dir + 16 * ri + 8));
    }
    if (missing > s->cap - s->cursor) evict(s);
    for (uint64_t k = 0; k < s->nneed; ++k) {
        const uint64_t ri = s->need[k];
        const uint8_t *e = s->dir + 16 * ri;
        const uint32_t f = rd32(e + 12);
        if (f == 0 || f == 1 || s->res[ri]) continue;
        const uint64_t off = rd64(e);
        const uint32_t len = rd32(e + 8);
        if (s->stage_cap < len) {
            if (s->stage) vcg_host_free(s->stage);
            s->stage = (uint8_t *)vcg_host_alloc(len);
            s->stage_cap = s->stage ? len : 0;
            if (!s->stage) return 1;
        }
        s->reads++;
        if (a->read(a->ud, off, len, s->stage) != VC_OK) return 1;
        const uint64_t at = s->cursor;
        if (vcg_arc_put(s->arc, at, s->stage, len) != 0) return 1;
        const int rb = vcg_arc_rebase(s->arc, at, off, len, a->len);
        if (rb != 0) { if (rb > 0) s->irr[ri] = 1; return 1; }
        uint8_t de[16];
        wr64(de, at); wr32(de + 8, len); wr32(de + 12, 2u);
        if (vcg_arc_put(s->arc, FILE_HDR + 16 * ri, de, 16) != 0) return 1;
        s->res[ri] = at;
        s->cursor = align16(at + len);
    }
    return 0;
}

/* Atoms (lod, ax, ay, az quads) or a box of one LOD: on success the arena is
 * returned with s->mu held (vc__stream_end releases it after the launches). */
vcg_arc *vc__stream_begin(vc_archive *a, int lod, const vc_box *box, size_t n, const int32_t *q) {
    vc_sdev *s = sdev_get(a);
    if (!s) return NULL;
    pthread_mutex_lock(&s->mu);
    if (++s->gen == 0) { memset(s->mark, 0, s->nreg * 4); s->gen = 1; }
    s->nneed = 0;
    if (box) {
        for (uint32_t rz = box->z0 / (VC_A * R_ATOMS); rz <= (box->z1 - 1) / (VC_A * R_ATOMS); ++rz)
            for (uint32_t ry = box->y0 / (VC_A * R_ATOMS); ry <= (box->y1 - 1) / (VC_A * R_ATOMS); ++ry)
                for (uint32_t rx = box->x0 / (VC_A * R_ATOMS); rx <= (box->x1 - 1) / (VC_A * R_ATOMS); ++rx)
                    touch(s, a, lod, rz, ry, rx);
    } else {
        for (size_t i = 0; i < n; ++i)
            touch(s, a, q[4 * i], (uint32_t)q[4 * i + 3] / R_ATOMS, (uint32_t)q[4 * i + 2] / R_ATOMS,
                  (uint32_t)q[4 * i + 1] / R_ATOMS);
    }
    if (make_resident(a, s) != 0) { pthread_mutex_unlock(&s->mu); return NULL; }
    return s->arc;
}

void vc__stream_end(vc_archive *a) {
    vc_sdev *s = (vc_sdev *)a->sdev;
    if (!s) return;
    /* the slot's previous record (64 calls ago) must be done before the
     * event is reused, or an eviction could miss that call's launches */
    void **e = &s->ev[s->ev_next];
    s->ev_next = (s->ev_next + 1) % SDEV_EVENTS;
    if (!*e) *e = vcg_event_create();
    else vcg_event_sync(*e);
    vcg_event_record(*e, -1);
    pthread_mutex_unlock(&s->mu);
}

vc_status vcx_stream_cache(vc_archive *a, size_t bytes) {
    if (!a || a->buf) return VC_ERR_RANGE;
    pthread_mutex_lock(&a->mu);
    vc_sdev *s = (vc_sdev *)a->sdev;
    pthread_mutex_unlock(&a->mu);
    if (s) {
        pthread_mutex_lock(&s->mu);
        arena_idle(s);
        pthread_mutex_unlock(&s->mu);
        vc__stream_free(a);
    }
    pthread_mutex_lock(&a->mu);
    a->stream_cache = bytes;
    a->sdev_failed = 0;
    pthread_mutex_unlock(&a->mu);
    return VC_OK;
}

vc_status vcx_stream_stats(const vc_archive *a, uint64_t *blob_reads, uint64_t *evictions) {
    if (!a) return VC_ERR_RANGE;
    const vc_sdev *s = (const vc_sdev *)a->sdev;
    if (blob_reads) *blob_reads = s ? s->reads : 0;
    if (evictions) *evictions = s ? s->evictions : 0;
    return VC_OK;
}
