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
 * copied into the caller's buffer (host or device memory; pinned memory from
 * vcx_host_alloc keeps host copies asynchronous).  A request larger than one
 * slot is served as parts over consecutive batches and reported once.
 *
 * A batch runs as two persistent kernels (entropy waves claiming 64 atoms at
 * a time, transform workgroups one atom at a time, the batch size read from
 * device memory), captured once per slot, LOD class and size bucket as a
 * hipGraph.  A batch replays its graph while SRV_REPLAY_BACKLOG or more
 * requests are queued (replays of concurrent slots run one after another, so
 * the saturated loop has no launch gaps); otherwise it launches the same
 * kernels directly, so a short batch overlaps the batches in flight.  Launch
 * shapes come from a table that vcx_server_autotune fills by timing
 * candidates on the archive's own atoms.
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
#include <stdio.h>
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
/* queued requests from which batches replay their captured graphs */
#ifndef SRV_REPLAY_BACKLOG
#define SRV_REPLAY_BACKLOG 8
#endif
/* launch-shape table: one captured persistent launch per LOD class and batch
 * size bucket (a batch takes the smallest bucket holding it) */
#define SRV_NB 4
static const size_t srv_bucket[SRV_NB] = {64, 512, 4096, 32768};

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
    int out_dev;                /* out is device memory */
    size_t first;               /* index of its first atom in the batch */
    uint64_t t_sub, t_adm;      /* submitted, admitted into a batch (vcx_now_ns) */
    struct srv_parent *parent;  /* a part of a request split over several batches, or NULL */
    struct srv_req *next;
} srv_req;

/* a request larger than one batch slot, served as parts in consecutive
 * batches and reported once, when its last part completes */
typedef struct srv_parent {
    int64_t id;
    size_t left;                /* parts not yet retired */
    int32_t st;
    uint64_t t_adm;             /* admission of its first part */
} srv_parent;

typedef struct {
    int busy;
    int cls;                    /* class of the admitted requests */
    void *ev;
    srv_req *reqs;              /* admitted requests, in batch order */
    size_t n;                   /* atoms in the batch */
    uint8_t *q_pin, *q_dev;     /* [16-byte header][cap quads][cap placements], pinned / device */
    int32_t *st_pin, *st_dev;
    uint8_t *out_dev;
    void *scratch;              /* decoder scratch of this slot's batches */
    vcg_qgraph *g[2][SRV_NB];   /* captured launches, built on first use */
} srv_slot;

struct vcx_server {
    vc_archive *a;
    vcg_arc *dev;
    size_t cap_atoms;           /* per slot */
    pthread_t th;
    pthread_mutex_t mu;
    pthread_cond_t cv_work, cv_done;
    int stop;
    int ready;                  /* worker: 1 once its launches are captured */
    srv_req *qhead, *qtail;
    int64_t next_id;
    size_t pending, inflight;
    /* completions, in retirement order */
    vcx_served *done;
    size_t d_cap, d_head, d_len;
    srv_slot slot[SRV_SLOTS];
    /* launch shapes per (class, bucket): the built-in default until
     * vcx_server_autotune measures better ones */
    vcg_qcfg cfg[2][SRV_NB];
    float ms_default[2][SRV_NB], ms_tuned[2][SRV_NB];
    int tune_req, tune_done, tuned;
    int rebuild;                /* shapes changed: recapture before the next batch */
};

uint64_t vcx_now_ns(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (uint64_t)t.tv_sec * 1000000000ull + (uint64_t)t.tv_nsec;
}

static void push_done(vcx_server *s, const vcx_served *d) {
    if (s->d_len == s->d_cap) {
        size_t nc = s->d_cap ? 2 * s->d_cap : 1024;
        vcx_served *nd = (vcx_served *)malloc(nc * sizeof *nd);
        if (!nd) vc__panic("out of memory (server completions)");
        for (size_t k = 0; k < s->d_len; ++k) nd[k] = s->done[(s->d_head + k) % s->d_cap];
        free(s->done);
        s->done = nd; s->d_cap = nc; s->d_head = 0;
    }
    s->done[(s->d_head + s->d_len) % s->d_cap] = *d;
    s->d_len++;
}

static int bucket_of(size_t n) {
    int k = 0;
    while (k + 1 < SRV_NB && n > srv_bucket[k]) ++k;
    return k;
}

/* the built-in launch shape for batches of up to S atoms: the one-shot
 * kernels' geometry (four-wave entropy workgroups covering every atom at
 * once, one transform workgroup per atom) */
static vcg_qcfg default_cfg(size_t S) {
    vcg_qcfg c;
    c.ent_waves = 4;
    c.ent_groups = (uint32_t)((S + 255) / 256);
    c.idct_groups = (uint32_t)S;
    return c;
}

static size_t slot_qbytes(size_t cap) { return 16 + cap * 16 + cap * sizeof(vcg_place); }

static vcg_qgraph *slot_graph(vcx_server *s, srv_slot *b, int cls, int bk) {
    if (!b->g[cls][bk]) {
        const int k = (int)(b - s->slot);
        b->g[cls][bk] = vcg_qgraph_create(s->dev, k, &s->cfg[cls][bk], b->q_dev, s->cap_atoms, b->out_dev, b->st_dev,
                                          b->scratch);
        if (!b->g[cls][bk]) vc__panic("server: capturing a batch launch failed");
    }
    return b->g[cls][bk];
}

/* capture every reachable (slot, class, bucket) launch up front */
static void prebuild(vcx_server *s) {
    for (int k = 0; k < SRV_SLOTS; ++k)
        for (int c = 0; c < 2; ++c)
            for (int bk = 0; bk < SRV_NB; ++bk)
                if (bk == 0 || srv_bucket[bk - 1] < s->cap_atoms) (void)slot_graph(s, &s->slot[k], c, bk);
}

static void drop_graphs(vcx_server *s) {
    for (int k = 0; k < SRV_SLOTS; ++k)
        for (int c = 0; c < 2; ++c)
            for (int bk = 0; bk < SRV_NB; ++bk) {
                vcg_qgraph_destroy(s->slot[k].g[c][bk]);
                s->slot[k].g[c][bk] = NULL;
            }
}

/* worker: enqueue one admitted batch (no lock held).  replay: from its
 * captured graph, else the same kernels launched directly */
static void launch(vcx_server *s, srv_slot *b, int replay) {
    const size_t cap = s->cap_atoms;
    vcg_place *pl = (vcg_place *)(b->q_pin + 16 + cap * 16);
    uint32_t hdr[4] = {(uint32_t)b->n, 0, 0, 0};
    memcpy(b->q_pin, hdr, 16);
    size_t np = 0;
    for (srv_req *r = b->reqs; r; r = r->next) {
        int32_t *q = (int32_t *)(b->q_pin + 16) + 4 * r->first;
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
    if (vcg_qgraph_launch(slot_graph(s, b, b->cls, bucket_of(b->n)), b->q_pin, b->n, np, !replay) != 0)
        vc__panic("server batch launch rejected");
    /* to host memory (pinned or pageable) or device memory (no PCIe
     * transfer), with the copy kind of each destination */
    for (srv_req *r = b->reqs; r; r = r->next)
        vcg_copy_async(r->out, b->out_dev + r->first * (size_t)VC_ATOM3, r->out_bytes, r->out_dev ? 3 : 2, k);
    vcg_copy_async(b->st_pin, b->st_dev, b->n * 4, 2, k);
    vcg_event_record(b->ev, k);
}

/* Launch-shape candidates for batches of up to S atoms: entropy workgroups
 * of 1, 2 or 4 waves (64 to 256 atoms per workgroup in flight), enough of
 * them for every atom at once or for half (waves then take two rounds of 64),
 * and 256, 512 or 1024 transform workgroups (one, two or four per CU
 * claiming atoms in turn) or one per atom.  The default is among them. */
static int tune_candidates(size_t S, vcg_qcfg *out) {
    static const int ws[3] = {1, 2, 4};
    int n = 0;
    for (int a = 0; a < 3; ++a)
        for (int half = 0; half < 2; ++half) {
            uint32_t ge = (uint32_t)((S + 64u * ws[a] - 1) / (64u * ws[a]));
            if (half) { if (ge < 2) continue; ge = (ge + 1) / 2; }
            const uint32_t gts[4] = {256, 512, 1024, (uint32_t)S};
            for (int t = 0; t < 4; ++t) {
                if (gts[t] > S || (t < 3 && gts[t] == S)) continue;
                out[n].ent_waves = ws[a]; out[n].ent_groups = ge; out[n].idct_groups = gts[t];
                n++;
            }
        }
    return n;
}

/* S random atoms of LOD class cls (0: LOD0; 1: LODs 1..nlod-1, uniform)
 * into slot b's pinned batch buffer */
static void tune_sample(const vcx_server *s, srv_slot *b, int cls, size_t S, unsigned *seed) {
    uint32_t hdr[4] = {(uint32_t)S, 0, 0, 0};
    memcpy(b->q_pin, hdr, 16);
    int32_t *q = (int32_t *)(b->q_pin + 16);
    for (size_t i = 0; i < S; ++i) {
        const int lod = cls == 0 ? 0 : 1 + (int)(rand_r(seed) % (unsigned)(s->a->nlod - 1));
        const vc_dims d = vc__lod_dims(s->a->dims0, lod);
        q[4 * i] = lod;
        q[4 * i + 1] = (int32_t)((unsigned)rand_r(seed) % ((d.nx + VC_A - 1) / VC_A));
        q[4 * i + 2] = (int32_t)((unsigned)rand_r(seed) % ((d.ny + VC_A - 1) / VC_A));
        q[4 * i + 3] = (int32_t)((unsigned)rand_r(seed) % ((d.nz + VC_A - 1) / VC_A));
    }
}

/* worker, no batch in flight: time every candidate shape per (class,
 * bucket) on sample batches of the archive's own atoms (the fastest of three
 * replays each) and keep the fastest; the captured launches are rebuilt on
 * next use */
static void autotune(vcx_server *s) {
    drop_graphs(s);
    srv_slot *b = &s->slot[0];
    unsigned seed = 20261204u;
    vcg_qcfg cand[24];
    for (int cls = 0; cls < 2; ++cls)
        for (int bk = 0; bk < SRV_NB; ++bk) {
            const size_t S = srv_bucket[bk] < s->cap_atoms ? srv_bucket[bk] : s->cap_atoms;
            if ((cls == 1 && s->a->nlod < 2) || (bk > 0 && srv_bucket[bk - 1] >= s->cap_atoms)) {
                s->ms_default[cls][bk] = s->ms_tuned[cls][bk] = 0.0f;     /* class or bucket never used */
                continue;
            }
            tune_sample(s, b, cls, S, &seed);
            const vcg_qcfg def = default_cfg(S);
            const int nc = tune_candidates(S, cand);
            float best = 1e30f;
            for (int c = 0; c < nc; ++c) {
                vcg_qgraph *g = vcg_qgraph_create(s->dev, 0, &cand[c], b->q_dev, s->cap_atoms, b->out_dev, b->st_dev,
                                                  b->scratch);
                if (!g) continue;
                const float ms = vcg_qgraph_time(g, b->q_pin, S, 0, 3);
                vcg_qgraph_destroy(g);
                if (ms < 0) continue;
                if (cand[c].ent_waves == def.ent_waves && cand[c].ent_groups == def.ent_groups &&
                    cand[c].idct_groups == def.idct_groups)
                    s->ms_default[cls][bk] = ms;
                if (ms < best) { best = ms; s->cfg[cls][bk] = cand[c]; }
            }
            s->ms_tuned[cls][bk] = best;
        }
    s->tuned = 1;
}

static void retire(vcx_server *s, srv_slot *b, uint64_t now) {
    srv_req *r = b->reqs;
    while (r) {
        int32_t st = 0;
        for (size_t k = 0; k < r->n; ++k) st |= b->st_pin[r->first + k];
        vcx_served d;
        d.id = r->id; d.cls = b->cls; d.t_submit_ns = r->t_sub; d.t_admit_ns = r->t_adm; d.t_done_ns = now;
        d.batch_atoms = b->n;
        if (r->parent) {
            srv_parent *p = r->parent;
            p->st |= st;
            if (--p->left == 0) {
                d.status = p->st ? VC_ERR_FORMAT : VC_OK;
                d.t_admit_ns = p->t_adm;
                push_done(s, &d);
                free(p);
            }
        } else {
            d.status = st ? VC_ERR_FORMAT : VC_OK;
            push_done(s, &d);
        }
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
    /* every slot's launches captured before the first request (the slots'
     * scratch was allocated with the server), on this thread's streams */
    prebuild(s);
    pthread_mutex_lock(&s->mu);
    s->ready = 1;
    pthread_cond_broadcast(&s->cv_done);
    for (;;) {
        if (s->rebuild && s->inflight == 0) {
            s->rebuild = 0;
            pthread_mutex_unlock(&s->mu);
            drop_graphs(s);
            prebuild(s);
            pthread_mutex_lock(&s->mu);
            continue;
        }
        /* a tuning request runs once no batch is in flight */
        if (s->tune_req && s->inflight == 0) {
            pthread_mutex_unlock(&s->mu);
            autotune(s);
            prebuild(s);
            pthread_mutex_lock(&s->mu);
            s->tune_req = 0;
            s->tune_done = 1;
            pthread_cond_broadcast(&s->cv_done);
            continue;
        }
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
            const uint64_t t_adm = vcx_now_ns();
            fb->n = 0;
            while (r) {
                srv_req *nx = r->next;
                if (r->cls != cls) { prev = r; r = nx; continue; }
                if (fb->n + r->n > s->cap_atoms) break;
                r->t_adm = t_adm;
                if (r->parent && !r->parent->t_adm) r->parent->t_adm = t_adm;
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
            /* Saturated (at least SRV_REPLAY_BACKLOG requests still queued
             * after this admission), the batch replays its captured graph:
             * graph replays of concurrent slots run one after another, which
             * keeps a saturated device from splitting its SIMDs over every
             * batch in flight.  Otherwise the same kernels go directly, so a
             * short LOD0 batch overlaps a long coarse one. */
            size_t backlog = 0;
            for (srv_req *q = s->qhead; q && backlog < SRV_REPLAY_BACKLOG; q = q->next) backlog++;
            const int replay = backlog >= SRV_REPLAY_BACKLOG;
            pthread_mutex_unlock(&s->mu);
            launch(s, fb, replay);
            pthread_mutex_lock(&s->mu);
            continue;
        }
        if (s->stop && !s->qhead && s->inflight == 0 && !s->tune_req) break;
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
    drop_graphs(s);
    return NULL;
}

vcx_server *vcx_server_create(vc_archive *a, size_t device_bytes) {
    if (!a || !a->buf) return NULL;             /* needs the device-resident archive */
    vc__require_gpu();
    vcg_arc *d = vc__arc_dev(a);
    if (!d) return NULL;
    if (device_bytes == 0) {
        /* admission by free device memory: half of what is free now, at
         * most what full batches in every slot take */
        size_t fr = 0;
        if (vcg_mem_info(&fr, NULL) != 0) return NULL;
        device_bytes = fr / 2;
        const size_t most = (size_t)SRV_SLOTS * SRV_MAX_BATCH * vcx_server_atom_bytes();
        if (device_bytes > most) device_bytes = most;
    }
    size_t per = device_bytes / SRV_SLOTS / vcx_server_atom_bytes();
    if (per > SRV_MAX_BATCH) per = SRV_MAX_BATCH;
    if (per < 1) return NULL;
    vcx_server *s = (vcx_server *)calloc(1, sizeof *s);
    if (!s) return NULL;
    s->a = a; s->dev = d; s->cap_atoms = per;
    for (int c = 0; c < 2; ++c)
        for (int bk = 0; bk < SRV_NB; ++bk)
            s->cfg[c][bk] = default_cfg(srv_bucket[bk] < per ? srv_bucket[bk] : per);
    pthread_mutex_init(&s->mu, NULL);
    pthread_cond_init(&s->cv_work, NULL);
    pthread_cond_init(&s->cv_done, NULL);
    for (int k = 0; k < SRV_SLOTS; ++k) {
        srv_slot *b = &s->slot[k];
        b->ev = vcg_event_create();
        /* quads plus at most one box placement per atom */
        b->q_pin = (uint8_t *)vcg_host_alloc(slot_qbytes(per));
        b->st_pin = (int32_t *)vcg_host_alloc(per * 4);
        b->q_dev = (uint8_t *)vcg_alloc(slot_qbytes(per));
        b->st_dev = (int32_t *)vcg_alloc(per * 4);
        b->out_dev = (uint8_t *)vcg_alloc(per * (size_t)VC_ATOM3);
        b->scratch = vcg_alloc(per * vcg_slot_scratch_bytes());
        if (!b->q_pin || !b->st_pin || !b->q_dev || !b->st_dev || !b->out_dev || !b->scratch) {
            s->stop = 1;
            vcx_server_destroy(s);
            return NULL;
        }
    }
    if (pthread_create(&s->th, NULL, worker, s) != 0) { vcx_server_destroy(s); return NULL; }
    pthread_mutex_lock(&s->mu);
    while (!s->ready) pthread_cond_wait(&s->cv_done, &s->mu);
    pthread_mutex_unlock(&s->mu);
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
    for (srv_req *r = s->qhead; r;) {
        srv_req *nx = r->next;
        if (r->parent && --r->parent->left == 0) free(r->parent);
        free(r->quads);
        free(r);
        r = nx;
    }
    for (int k = 0; k < SRV_SLOTS; ++k) {
        srv_slot *b = &s->slot[k];
        vcg_event_destroy(b->ev);
        vcg_host_free(b->q_pin); vcg_host_free(b->st_pin);
        vcg_free(b->q_dev); vcg_free(b->st_dev); vcg_free(b->out_dev); vcg_free(b->scratch);
    }
    free(s->done);
    pthread_cond_destroy(&s->cv_done);
    pthread_cond_destroy(&s->cv_work);
    pthread_mutex_destroy(&s->mu);
    free(s);
}

/* queue the parts of one request (one part: the request itself) in order,
 * under one id */
static int64_t enqueue(vcx_server *s, srv_req **parts, size_t np, srv_parent *par) {
    pthread_mutex_lock(&s->mu);
    if (s->stop) {
        pthread_mutex_unlock(&s->mu);
        for (size_t k = 0; k < np; ++k) { free(parts[k]->quads); free(parts[k]); }
        free(par);
        return -1;
    }
    const int64_t id = s->next_id++;
    const uint64_t t_sub = vcx_now_ns();
    if (par) { par->id = id; par->left = np; }
    for (size_t k = 0; k < np; ++k) {
        srv_req *r = parts[k];
        r->id = id;
        r->t_sub = t_sub;
        r->parent = par;
        r->next = NULL;
        if (s->qtail) s->qtail->next = r; else s->qhead = r;
        s->qtail = r;
    }
    pthread_cond_signal(&s->cv_work);
    pthread_mutex_unlock(&s->mu);
    return id;
}

static void free_parts(srv_req **parts, size_t np) {
    for (size_t k = 0; k < np; ++k) if (parts[k]) { free(parts[k]->quads); free(parts[k]); }
    free(parts);
}

/* A list longer than one batch slot goes as parts of at most a slot's
 * capacity, each writing its atoms at their place in out. */
int64_t vcx_server_submit_atoms(vcx_server *s, size_t n, const int32_t *quads, uint8_t *out) {
    if (!s || !n || !quads || !out) return -1;
    for (size_t i = 0; i < n; ++i)
        if (!vc__atom_ok(s->a, quads[4 * i], quads[4 * i + 1], quads[4 * i + 2], quads[4 * i + 3])) return -1;
    const size_t cap = s->cap_atoms, np = (n + cap - 1) / cap;
    const int dev = vcg_ptr_is_device(out);
    srv_req **parts = (srv_req **)calloc(np, sizeof *parts);
    srv_parent *par = np > 1 ? (srv_parent *)calloc(1, sizeof *par) : NULL;
    if (!parts || (np > 1 && !par)) { free(parts); free(par); return -1; }
    for (size_t k = 0; k < np; ++k) {
        const size_t i0 = k * cap, m = n - i0 < cap ? n - i0 : cap;
        srv_req *r = (srv_req *)calloc(1, sizeof *r);
        if (!r || !(r->quads = (int32_t *)malloc(m * 16))) { free(r); free_parts(parts, np); free(par); return -1; }
        parts[k] = r;
        memcpy(r->quads, quads + 4 * i0, m * 16);
        r->kind = 0; r->n = m; r->out = out + i0 * (size_t)VC_ATOM3; r->out_bytes = m * (size_t)VC_ATOM3;
        r->out_dev = dev;
        for (size_t i = 0; i < m; ++i) if (r->quads[4 * i] != 0) r->cls = 1;
    }
    const int64_t id = enqueue(s, parts, np, par);
    free(parts);
    return id;
}

/* A box of more atoms than one batch slot goes as z-slabs of whole atom
 * layers (the first and last clipped to the box): each slab's raster is the
 * contiguous piece of the box raster at its first z.  Rejected only when one
 * atom layer of the box exceeds a slot. */
int64_t vcx_server_submit_region(vcx_server *s, int lod, vc_box box, uint8_t *out) {
    if (!s || !out || lod < 0 || lod >= s->a->nlod) return -1;
    if (vc__region_check(s->a, lod, box) != VC_OK) return -1;
    const size_t nax = (box.x1 - 1) / VC_A - box.x0 / VC_A + 1, nay = (box.y1 - 1) / VC_A - box.y0 / VC_A + 1;
    const uint32_t az0 = box.z0 / VC_A, az1 = (box.z1 - 1) / VC_A + 1;
    const size_t layer = nax * nay, cap = s->cap_atoms;
    if (layer > cap) return -1;
    const uint32_t per = (uint32_t)(cap / layer);                 /* atom layers per part */
    const size_t np = (az1 - az0 + per - 1) / per;
    const uint64_t plane = (uint64_t)(box.x1 - box.x0) * (box.y1 - box.y0);
    const int dev = vcg_ptr_is_device(out);
    srv_req **parts = (srv_req **)calloc(np, sizeof *parts);
    srv_parent *par = np > 1 ? (srv_parent *)calloc(1, sizeof *par) : NULL;
    if (!parts || (np > 1 && !par)) { free(parts); free(par); return -1; }
    for (size_t k = 0; k < np; ++k) {
        const uint32_t la = az0 + (uint32_t)k * per, lb = la + per < az1 ? la + per : az1;
        vc_box b = box;
        b.z0 = la * VC_A > box.z0 ? la * VC_A : box.z0;
        b.z1 = lb * VC_A < box.z1 ? lb * VC_A : box.z1;
        srv_req *r = (srv_req *)calloc(1, sizeof *r);
        if (!r) { free_parts(parts, np); free(par); return -1; }
        parts[k] = r;
        r->kind = 1; r->n = layer * (lb - la); r->lod = lod; r->box = b;
        r->out = out + (uint64_t)(b.z0 - box.z0) * plane; r->out_bytes = (uint64_t)(b.z1 - b.z0) * plane;
        r->out_dev = dev;
        r->cls = lod != 0;
    }
    const int64_t id = enqueue(s, parts, np, par);
    free(parts);
    return id;
}

/* up to max completions in retirement order into the caller's arrays,
 * waiting up to wait_ms for the first one */
static size_t take_done(vcx_server *s, size_t max, int wait_ms, int64_t *ids, int32_t *status, uint64_t *t_done_ns,
                        vcx_served *rec) {
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
        const vcx_served *d = &s->done[s->d_head];
        if (ids) ids[k] = d->id;
        if (status) status[k] = d->status;
        if (t_done_ns) t_done_ns[k] = d->t_done_ns;
        if (rec) rec[k] = *d;
        s->d_head = (s->d_head + 1) % s->d_cap;
        s->d_len--;
        k++;
    }
    pthread_mutex_unlock(&s->mu);
    return k;
}

size_t vcx_server_poll(vcx_server *s, int64_t *ids, int32_t *status, uint64_t *t_done_ns, size_t max, int wait_ms) {
    return take_done(s, max, wait_ms, ids, status, t_done_ns, NULL);
}

size_t vcx_server_poll_served(vcx_server *s, vcx_served *out, size_t max, int wait_ms) {
    return take_done(s, out ? max : 0, wait_ms, NULL, NULL, NULL, out);
}

size_t vcx_server_capacity(const vcx_server *s) { return s ? s->cap_atoms : 0; }

vc_status vcx_server_autotune(vcx_server *s) {
    if (!s) return VC_ERR_RANGE;
    pthread_mutex_lock(&s->mu);
    if (s->stop) { pthread_mutex_unlock(&s->mu); return VC_ERR_RANGE; }
    s->tune_done = 0;
    s->tune_req = 1;
    pthread_cond_signal(&s->cv_work);
    while (!s->tune_done) pthread_cond_wait(&s->cv_done, &s->mu);
    pthread_mutex_unlock(&s->mu);
    return VC_OK;
}

vc_status vcx_test_server_shape(vcx_server *s, int ent_waves, uint32_t ent_groups, uint32_t idct_groups) {
    if (!s || (ent_waves != 1 && ent_waves != 2 && ent_waves != 4) || !ent_groups || !idct_groups) return VC_ERR_RANGE;
    pthread_mutex_lock(&s->mu);
    while (s->inflight || s->qhead) {           /* drain: the slots' graphs are rebuilt */
        pthread_mutex_unlock(&s->mu);
        struct timespec ts = {0, 200000};
        nanosleep(&ts, NULL);
        pthread_mutex_lock(&s->mu);
    }
    for (int c = 0; c < 2; ++c)
        for (int bk = 0; bk < SRV_NB; ++bk) {
            s->cfg[c][bk].ent_waves = ent_waves;
            s->cfg[c][bk].ent_groups = ent_groups;
            s->cfg[c][bk].idct_groups = idct_groups;
        }
    s->rebuild = 1;
    pthread_mutex_unlock(&s->mu);
    return VC_OK;
}

size_t vcx_server_tuning(vcx_server *s, char *buf, size_t cap) {
    if (!s) return 0;
    char tmp[4096];
    size_t n = 0;
    pthread_mutex_lock(&s->mu);
    n += (size_t)snprintf(tmp + n, sizeof tmp - n, "# tuned %d; class bucket ent_waves ent_groups idct_groups default_ms tuned_ms\n",
                          s->tuned);
    for (int c = 0; c < 2; ++c)
        for (int bk = 0; bk < SRV_NB; ++bk)
            n += (size_t)snprintf(tmp + n, sizeof tmp - n, "%d %zu %d %u %u %.4f %.4f\n", c, srv_bucket[bk],
                                  s->cfg[c][bk].ent_waves, s->cfg[c][bk].ent_groups, s->cfg[c][bk].idct_groups,
                                  (double)s->ms_default[c][bk], (double)s->ms_tuned[c][bk]);
    pthread_mutex_unlock(&s->mu);
    if (buf && cap) {
        const size_t m = n < cap - 1 ? n : cap - 1;
        memcpy(buf, tmp, m);
        buf[m] = 0;
    }
    return n;
}

/* output atom, quad, box placement, status word and decoder scratch */
size_t vcx_server_atom_bytes(void) {
    return (size_t)VC_ATOM3 + 16 + sizeof(vcg_place) + 4 + vcg_slot_scratch_bytes();
}

void *vcx_host_alloc(size_t bytes) { vc__require_gpu(); return vcg_host_alloc(bytes); }
void vcx_host_free(void *p) { vcg_host_free(p); }
