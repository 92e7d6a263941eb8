/* vc_serve.c — device-resident request loop with continuous batching.
 *
 * Requests (a list of atoms, or a voxel box of one LOD) are queued by any
 * thread and served by one worker thread that owns a GPU stream.  The worker
 * keeps up to SRV_SLOTS batches in flight: whenever a slot is free it admits
 * the queued requests, atoms and boxes alike, that fit the slot's share of
 * the device-memory budget (in arrival order) as one batched decode, and
 * launches it without waiting; consecutive batches overlap on the device.  A
 * box request becomes the list of its atoms, each clipped into the box raster
 * by the transform kernel (a box of n atoms needs at most n atoms of output
 * space).  Finished batches are retired by polling their events, and each
 * request is reported with its status and completion time.  Results are
 * copied into the caller's buffer (pinned memory from vcx_host_alloc keeps
 * the copies asynchronous).
 *
 * Batches are LOD-aware.  Coarser LODs are quantized finer, so their atoms
 * carry several times the bins of an LOD0 atom, and a batch lasts as long as
 * its slowest atom.  Requests that touch LOD0 only (class 0) and the others
 * (class 1) therefore go to separate batches: a free slot takes the class of
 * the oldest queued request and admits the queued requests of that class in
 * arrival order, and class 1 holds at most SRV_COARSE_SLOTS slots, so LOD0
 * requests always find one.
 *
 * Decoding a request in a batch gives the same bytes as decoding it alone:
 * every atom is decoded independently (the batch only shares launches).
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "vc_host.h"

#define SRV_SLOTS 3
/* most slots class 1 may hold: one, so LOD0 requests (the common case,
 * with batches several times shorter) keep two */
#ifndef SRV_COARSE_SLOTS
#define SRV_COARSE_SLOTS 1
#endif
#define SRV_MAX_BATCH 32768u

typedef struct srv_req {
    int64_t id;
    int kind;                   /* 0: atoms, 1: region */
    int cls;                    /* 0: LOD0 only, 1: touches a coarser LOD */
    size_t n;                   /* atoms decoded */
    int32_t *quads;             /* kind 0: owned copy, 4 per atom */
    int lod;
    vc_box box;
    uint8_t *out;
    size_t out_bytes;
    size_t first;               /* index of its first atom in the batch */
    struct srv_req *next;
} srv_req;

typedef struct {
    int busy;
    int cls;                    /* class of the admitted requests */
    void *ev;
    srv_req *reqs;              /* admitted requests, in batch order */
    size_t n;                   /* atoms in the batch */
    int32_t *q_pin, *st_pin;    /* pinned staging: quads, then box placements */
    int32_t *q_dev, *st_dev;
    uint8_t *out_dev;
} srv_slot;

struct vcx_server {
    vc_archive *a;
    vcg_arc *dev;
    size_t cap_atoms;           /* per slot */
    pthread_t th;
    pthread_mutex_t mu;
    pthread_cond_t cv_work, cv_done;
    int stop;
    int ready;                  /* worker: 1 scratch reserved, -1 failed */
    srv_req *qhead, *qtail;
    int64_t next_id;
    size_t pending, inflight;
    /* completions: ring of (id, status, t_done_ns) */
    int64_t *d_id; int32_t *d_st; uint64_t *d_t;
    size_t d_cap, d_head, d_len;
    srv_slot slot[SRV_SLOTS];
};

uint64_t vcx_now_ns(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (uint64_t)t.tv_sec * 1000000000ull + (uint64_t)t.tv_nsec;
}

static void push_done(vcx_server *s, int64_t id, int32_t st, uint64_t t) {
    if (s->d_len == s->d_cap) {
        size_t nc = s->d_cap ? 2 * s->d_cap : 1024;
        int64_t *ni = (int64_t *)malloc(nc * 8); int32_t *ns = (int32_t *)malloc(nc * 4);
        uint64_t *nt = (uint64_t *)malloc(nc * 8);
        if (!ni || !ns || !nt) vc__panic("out of memory (server completions)");
        for (size_t k = 0; k < s->d_len; ++k) {
            size_t j = (s->d_head + k) % (s->d_cap ? s->d_cap : 1);
            ni[k] = s->d_id[j]; ns[k] = s->d_st[j]; nt[k] = s->d_t[j];
        }
        free(s->d_id); free(s->d_st); free(s->d_t);
        s->d_id = ni; s->d_st = ns; s->d_t = nt; s->d_cap = nc; s->d_head = 0;
    }
    size_t j = (s->d_head + s->d_len) % s->d_cap;
    s->d_id[j] = id; s->d_st[j] = st; s->d_t[j] = t;
    s->d_len++;
}

/* worker: enqueue one admitted batch (no lock held) */
static void launch(vcx_server *s, srv_slot *b) {
    vcg_place *pl = (vcg_place *)(b->q_pin + 4 * b->n);
    size_t np = 0;
    for (srv_req *r = b->reqs; r; r = r->next) {
        int32_t *q = b->q_pin + 4 * r->first;
        if (r->kind == 0) { memcpy(q, r->quads, r->n * 16); continue; }
        /* a box: its atoms, clipped into the raster at the request's output space */
        const vc_box bx = r->box;
        pl[np] = (vcg_place){bx.x0, bx.y0, bx.z0, bx.x1, bx.y1, bx.z1, r->first * (uint64_t)VC_ATOM3};
        const int32_t tag = r->lod | (int32_t)(np + 1) << 8;
        for (uint32_t az = bx.z0 / VC_A; az <= (bx.z1 - 1) / VC_A; ++az)
            for (uint32_t ay = bx.y0 / VC_A; ay <= (bx.y1 - 1) / VC_A; ++ay)
                for (uint32_t ax = bx.x0 / VC_A; ax <= (bx.x1 - 1) / VC_A; ++ax) {
                    q[0] = tag; q[1] = (int32_t)ax; q[2] = (int32_t)ay; q[3] = (int32_t)az;
                    q += 4;
                }
        np++;
    }
    /* the batch, its result copies and its event all go on the stream of
     * this slot's scratch: batches in other slots neither wait for it nor
     * delay it, and the slot's next batch follows its copies */
    const int k = (int)(b - s->slot);
    if (vcg_list_launch(s->dev, b->n, b->q_pin, b->q_dev, b->out_dev, b->st_dev, np, k) != 0)
        vc__panic("server batch launch rejected");
    for (srv_req *r = b->reqs; r; r = r->next)
        vcg_copy_async(r->out, b->out_dev + r->first * (size_t)VC_ATOM3, r->out_bytes, 2, k);
    vcg_copy_async(b->st_pin, b->st_dev, b->n * 4, 2, k);
    vcg_event_record(b->ev, k);
}

static void retire(vcx_server *s, srv_slot *b, uint64_t now) {
    srv_req *r = b->reqs;
    while (r) {
        int32_t st = 0;
        for (size_t k = 0; k < r->n; ++k) st |= b->st_pin[r->first + k];
        push_done(s, r->id, st ? VC_ERR_FORMAT : VC_OK, now);
        srv_req *nx = r->next;
        free(r->quads);
        free(r);
        r = nx;
    }
    s->inflight -= 1;
    b->reqs = NULL;
    b->n = 0;
    b->busy = 0;
}

static void *worker(void *arg) {
    vcx_server *s = (vcx_server *)arg;
    /* the decoder scratch of a full batch in every slot, reserved on this
     * thread (it owns the decode context) before the first request: no batch
     * then waits on an allocation and the device synchronisation it takes */
    const int ok = vcg_reserve_slots(s->cap_atoms) == 0;
    pthread_mutex_lock(&s->mu);
    s->ready = ok ? 1 : -1;
    pthread_cond_broadcast(&s->cv_done);
    if (!ok) { pthread_mutex_unlock(&s->mu); return NULL; }
    for (;;) {
        /* retire finished batches */
        int retired = 0;
        for (int k = 0; k < SRV_SLOTS; ++k) {
            srv_slot *b = &s->slot[k];
            if (b->busy && vcg_event_done(b->ev)) { retire(s, b, vcx_now_ns()); retired = 1; }
        }
        if (retired) pthread_cond_broadcast(&s->cv_done);
        /* admit queued requests into a free slot: the class of the oldest
         * queued request, unless class 1 already holds SRV_COARSE_SLOTS slots */
        srv_slot *fb = NULL;
        int busy1 = 0;
        for (int k = 0; k < SRV_SLOTS; ++k) {
            if (!s->slot[k].busy) { if (!fb) fb = &s->slot[k]; }
            else busy1 += s->slot[k].cls;
        }
        int cls = -1;
        if (fb && s->qhead) {
            cls = s->qhead->cls;
            if (cls == 1 && busy1 >= SRV_COARSE_SLOTS) {
                cls = -1;
                for (srv_req *r = s->qhead; r; r = r->next) if (r->cls == 0) { cls = 0; break; }
            }
        }
        if (cls >= 0) {
            srv_req *tail = NULL, *prev = NULL, *r = s->qhead;
            fb->n = 0;
            while (r) {
                srv_req *nx = r->next;
                if (r->cls != cls) { prev = r; r = nx; continue; }
                if (fb->n + r->n > s->cap_atoms) break;
                /* unlink r (queue order of the rest is kept) */
                if (prev) prev->next = nx; else s->qhead = nx;
                if (s->qtail == r) s->qtail = prev;
                r->next = NULL;
                r->first = fb->n;
                fb->n += r->n;
                if (tail) tail->next = r; else fb->reqs = r;
                tail = r;
                r = nx;
            }
            fb->cls = cls;
            fb->busy = 1;
            s->inflight += 1;
            pthread_mutex_unlock(&s->mu);
            launch(s, fb);
            pthread_mutex_lock(&s->mu);
            continue;
        }
        if (s->stop && !s->qhead && s->inflight == 0) break;
        /* queued requests that no free slot may take wait for a retirement */
        if (s->inflight == 0) {
            pthread_cond_wait(&s->cv_work, &s->mu);
        } else {
            /* batches in flight: poll their events at a 20 us cadence */
            pthread_mutex_unlock(&s->mu);
            struct timespec ts = {0, 20000};
            nanosleep(&ts, NULL);
            pthread_mutex_lock(&s->mu);
        }
    }
    pthread_mutex_unlock(&s->mu);
    return NULL;
}

vcx_server *vcx_server_create(vc_archive *a, size_t device_bytes) {
    if (!a || !a->buf) return NULL;             /* needs the device-resident archive */
    vc__require_gpu();
    vcg_arc *d = vc__arc_dev(a);
    if (!d) return NULL;
    size_t per = device_bytes / SRV_SLOTS / vcx_server_atom_bytes();
    if (per > SRV_MAX_BATCH) per = SRV_MAX_BATCH;
    if (per < 1) return NULL;
    vcx_server *s = (vcx_server *)calloc(1, sizeof *s);
    if (!s) return NULL;
    s->a = a; s->dev = d; s->cap_atoms = per;
    pthread_mutex_init(&s->mu, NULL);
    pthread_cond_init(&s->cv_work, NULL);
    pthread_cond_init(&s->cv_done, NULL);
    for (int k = 0; k < SRV_SLOTS; ++k) {
        srv_slot *b = &s->slot[k];
        b->ev = vcg_event_create();
        /* quads plus at most one box placement per atom */
        b->q_pin = (int32_t *)vcg_host_alloc(per * (16 + sizeof(vcg_place)));
        b->st_pin = (int32_t *)vcg_host_alloc(per * 4);
        b->q_dev = (int32_t *)vcg_alloc(per * (16 + sizeof(vcg_place)));
        b->st_dev = (int32_t *)vcg_alloc(per * 4);
        b->out_dev = (uint8_t *)vcg_alloc(per * (size_t)VC_ATOM3);
        if (!b->q_pin || !b->st_pin || !b->q_dev || !b->st_dev || !b->out_dev) {
            s->stop = 1;
            vcx_server_destroy(s);
            return NULL;
        }
    }
    if (pthread_create(&s->th, NULL, worker, s) != 0) { vcx_server_destroy(s); return NULL; }
    pthread_mutex_lock(&s->mu);
    while (!s->ready) pthread_cond_wait(&s->cv_done, &s->mu);
    const int ready = s->ready;
    pthread_mutex_unlock(&s->mu);
    if (ready < 0) { vcx_server_destroy(s); return NULL; }
    return s;
}

void vcx_server_destroy(vcx_server *s) {
    if (!s) return;
    pthread_mutex_lock(&s->mu);
    int running = s->th != 0;
    s->stop = 1;
    pthread_cond_broadcast(&s->cv_work);
    pthread_mutex_unlock(&s->mu);
    if (running) pthread_join(s->th, NULL);
    for (srv_req *r = s->qhead; r;) { srv_req *nx = r->next; free(r->quads); free(r); r = nx; }
    for (int k = 0; k < SRV_SLOTS; ++k) {
        srv_slot *b = &s->slot[k];
        vcg_event_destroy(b->ev);
        vcg_host_free(b->q_pin); vcg_host_free(b->st_pin);
        vcg_free(b->q_dev); vcg_free(b->st_dev); vcg_free(b->out_dev);
    }
    free(s->d_id); free(s->d_st); free(s->d_t);
    pthread_cond_destroy(&s->cv_done);
    pthread_cond_destroy(&s->cv_work);
    pthread_mutex_destroy(&s->mu);
    free(s);
}

static int64_t enqueue(vcx_server *s, srv_req *r) {
    pthread_mutex_lock(&s->mu);
    if (s->stop) { pthread_mutex_unlock(&s->mu); free(r->quads); free(r); return -1; }
    r->id = s->next_id++;
    r->next = NULL;
    if (s->qtail) s->qtail->next = r; else s->qhead = r;
    s->qtail = r;
    int64_t id = r->id;
    pthread_cond_signal(&s->cv_work);
    pthread_mutex_unlock(&s->mu);
    return id;
}

int64_t vcx_server_submit_atoms(vcx_server *s, size_t n, const int32_t *quads, uint8_t *out) {
    if (!s || !n || !quads || !out || n > s->cap_atoms) return -1;
    for (size_t i = 0; i < n; ++i)
        if (!vc__atom_ok(s->a, quads[4 * i], quads[4 * i + 1], quads[4 * i + 2], quads[4 * i + 3])) return -1;
    srv_req *r = (srv_req *)calloc(1, sizeof *r);
    if (!r) return -1;
    r->quads = (int32_t *)malloc(n * 16);
    if (!r->quads) { free(r); return -1; }
    memcpy(r->quads, quads, n * 16);
    r->kind = 0; r->n = n; r->out = out; r->out_bytes = n * (size_t)VC_ATOM3;
    for (size_t i = 0; i < n; ++i) if (quads[4 * i] != 0) r->cls = 1;
    return enqueue(s, r);
}

int64_t vcx_server_submit_region(vcx_server *s, int lod, vc_box box, uint8_t *out) {
    if (!s || !out || lod < 0 || lod >= s->a->nlod) return -1;
    if (vc__region_check(s->a, lod, box) != VC_OK) return -1;
    const size_t nax = (box.x1 - 1) / VC_A - box.x0 / VC_A + 1, nay = (box.y1 - 1) / VC_A - box.y0 / VC_A + 1,
                 naz = (box.z1 - 1) / VC_A - box.z0 / VC_A + 1;
    const size_t bytes = (size_t)(box.x1 - box.x0) * (box.y1 - box.y0) * (box.z1 - box.z0);
    if (nax * nay * naz > s->cap_atoms || bytes > s->cap_atoms * (size_t)VC_ATOM3) return -1;
    srv_req *r = (srv_req *)calloc(1, sizeof *r);
    if (!r) return -1;
    r->kind = 1; r->n = nax * nay * naz; r->lod = lod; r->box = box; r->out = out; r->out_bytes = bytes;
    r->cls = lod != 0;
    return enqueue(s, r);
}

size_t vcx_server_poll(vcx_server *s, int64_t *ids, int32_t *status, uint64_t *t_done_ns, size_t max, int wait_ms) {
    if (!s || !max) return 0;
    pthread_mutex_lock(&s->mu);
    if (s->d_len == 0 && wait_ms > 0) {
        struct timespec ts;
        clock_gettime(CLOCK_REALTIME, &ts);
        ts.tv_sec += wait_ms / 1000;
        ts.tv_nsec += (long)(wait_ms % 1000) * 1000000L;
        if (ts.tv_nsec >= 1000000000L) { ts.tv_sec++; ts.tv_nsec -= 1000000000L; }
        pthread_cond_timedwait(&s->cv_done, &s->mu, &ts);
    }
    size_t k = 0;
    while (k < max && s->d_len) {
        size_t j = s->d_head;
        if (ids) ids[k] = s->d_id[j];
        if (status) status[k] = s->d_st[j];
        if (t_done_ns) t_done_ns[k] = s->d_t[j];
        s->d_head = (s->d_head + 1) % s->d_cap;
        s->d_len--;
        k++;
    }
    pthread_mutex_unlock(&s->mu);
    return k;
}

size_t vcx_server_capacity(const vcx_server *s) { return s ? s->cap_atoms : 0; }

/* output atom, quad, box placement, status word and decoder scratch */
size_t vcx_server_atom_bytes(void) {
    return (size_t)VC_ATOM3 + 16 + sizeof(vcg_place) + 4 + vcg_slot_scratch_bytes();
}

void *vcx_host_alloc(size_t bytes) { vc__require_gpu(); return vcg_host_alloc(bytes); }
void vcx_host_free(void *p) { vcg_host_free(p); }
