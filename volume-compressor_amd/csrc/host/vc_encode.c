/* vc_encode.c — whole-volume encode.
 *
 * vc_encode is sketched but absent in the reference (README.md:51-52,
 * plan.txt:206-207); here it is defined by composing the reference's own
 * primitives:
 *   pyramid   : LOD l+1 = integer rounded mean of the in-bounds voxels of each
 *               2x2x2 cell of LOD l (vc_test_downsample2x, src/vc/vc.c:1562-1577),
 *               LOD count by vc_create's rule (src/vc/vc.c:1001-1006);
 *   base_q    : rate control below, then q[l] = q0*0.6^l by iterated float
 *               multiply (tools/vc_export_stream.c:704-716), each clamped to
 *               [0.05, 4096] (src/vc/vc.c:1072-1073);
 *   atoms     : every atom of every LOD, zero padded past the edge
 *               (tests/test_realdata.c:108-113), encode_atom (src/vc/vc.c:595-622);
 *   container : v2, laid out exactly as tools/vc_repack.c:96-270 lays out the
 *               v1 archive of a single-threaded raster append of those atoms.
 * The codec work (pyramid, transforms, quantization, range coding, image
 * assembly) runs on the GPU; rate control and layout are integer/float host
 * arithmetic evaluated in a fixed order, so archives are deterministic.
 */
#include <math.h>
#include <stdlib.h>

#include "vc_host.h"

/* ---- rate control ----------------------------------------------------------
 * Calibration sample: LOD0 atoms on a lattice of at most 8 indices per axis
 * (every index when the axis has <= 8 atoms, else floor((2j+1)n/16)).
 * ratio(q) = in-bounds voxels of the sample's non-zero atoms / their payload
 * bytes (present-atom ratio, docs/SPEC.md:181-187).  meets(q): voxels >=
 * target*bytes in double.  Geometric bisection with sqrtf over [0.05, 4096],
 * 16 steps; the result is the upper bracket (finest q meeting the target).
 * A sample whose encode overflows the coder bound does not meet the target. */
static int lattice(uint32_t n, uint32_t *idx) {
    if (n <= 8) { for (uint32_t i = 0; i < n; ++i) idx[i] = i; return (int)n; }
    for (uint32_t j = 0; j < 8; ++j) idx[j] = (uint32_t)(((uint64_t)(2 * j + 1) * n) / 16);
    return 8;
}

/* The search below is exactly the sequential one (finest q, coarsest q, then
 * 16 bisection steps mid = sqrtf(lo*hi); meets -> hi = mid, else lo = mid),
 * evaluated `depth` steps per launch: each round encodes the sample at every
 * midpoint of the next `depth` levels of the bisection tree (plus both bounds
 * in the first round) and then follows the path the sequential search takes,
 * so every depth gives the same q.  A candidate's encode stops once its byte
 * total can no longer meet the target (the largest meeting total is fixed by
 * the voxel count), which decides it the same way.  The rounds of many
 * volumes share launches. */
#define RC_MAX_DEPTH 4
#define RC_NODES ((1 << RC_MAX_DEPTH) - 1)
#define RC_STEPS 16

typedef struct {
    vcg_enc *e;
    vc_dims d;
    float target;
    uint32_t *list, *vox;
    uint8_t *zero;
    int ns;
    uint64_t v, budget;
    float lo, hi, q0;
    int first, steps, done;
    int nq, base, depth;                  /* the current round's candidates */
    float qs[RC_NODES + 2], mid[RC_NODES + 1];
} rc_vol;

/* the calibration lattice and each sample atom's in-bounds voxel count */
static void rc_lattice(rc_vol *r) {
    uint32_t ix[8], iy[8], iz[8];
    const vc_dims d = r->d;
    int nx = lattice((d.nx + 31) / 32, ix), ny = lattice((d.ny + 31) / 32, iy), nz = lattice((d.nz + 31) / 32, iz);
    r->ns = nx * ny * nz;
    r->list = (uint32_t *)malloc((size_t)r->ns * 12);
    r->vox = (uint32_t *)malloc((size_t)r->ns * 4);
    r->zero = (uint8_t *)malloc((size_t)r->ns);
    if (!r->list || !r->vox || !r->zero) vc__panic("out of memory");
    int k = 0;
    for (int a = 0; a < nz; ++a) for (int b = 0; b < ny; ++b) for (int c = 0; c < nx; ++c, ++k) {
        r->list[3 * k] = iz[a]; r->list[3 * k + 1] = iy[b]; r->list[3 * k + 2] = ix[c];
        uint32_t ex = d.nx - ix[c] * 32, ey = d.ny - iy[b] * 32, ez = d.nz - iz[a] * 32;
        r->vox[k] = (ex < 32 ? ex : 32) * (ey < 32 ? ey : 32) * (ez < 32 ? ez : 32);
    }
}

/* after the zero masks: the voxel count, the byte budget, the search bounds */
static void rc_start(rc_vol *r) {
    r->v = 0;
    for (int i = 0; i < r->ns; ++i) if (!r->zero[i]) r->v += r->vox[i];
    r->lo = 0.05f; r->hi = 4096.0f;
    r->first = 1; r->steps = 0; r->done = 0;
    if (r->v == 0) { r->q0 = 1.0f; r->done = 1; return; }   /* all-zero sample: nothing to calibrate */
    /* meets(q): no overflow and voxels >= target * bytes (double); budget =
     * the largest byte total that meets */
    const double v = (double)r->v, t = (double)r->target;
    uint64_t budget = (uint64_t)(v / t);
    while (v >= t * (double)(budget + 1)) budget++;
    while (budget > 0 && !(v >= t * (double)budget)) budget--;
    r->budget = budget;
    r->q0 = r->hi;
}

/* level-order midpoints of the next r->depth levels from (lo, hi) */
static void rc_propose(rc_vol *r, int depth) {
    if (depth > RC_STEPS - r->steps) depth = RC_STEPS - r->steps;
    const int nodes = (1 << depth) - 1;
    float L[RC_NODES + 1], H[RC_NODES + 1];
    L[1] = r->lo; H[1] = r->hi;
    for (int t = 1; t <= nodes; ++t) {
        r->mid[t] = sqrtf(L[t] * H[t]);
        if (2 * t + 1 <= nodes) {
            L[2 * t] = L[t]; H[2 * t] = r->mid[t];          /* meets: hi = mid */
            L[2 * t + 1] = r->mid[t]; H[2 * t + 1] = H[t];  /* else:  lo = mid */
        }
    }
    r->depth = depth;
    r->nq = 0;
    if (r->first) { r->qs[r->nq++] = r->lo; r->qs[r->nq++] = r->hi; }
    r->base = r->nq;
    for (int t = 1; t <= nodes; ++t) r->qs[r->nq++] = r->mid[t];
}

static void rc_accept(rc_vol *r, const uint64_t *bytes, const int *flags) {
    int meets[RC_NODES + 2];
    for (int j = 0; j < r->nq; ++j)
        meets[j] = flags[j] == 0 && (double)r->v >= (double)r->target * (double)bytes[j];
    if (r->first) {
        r->first = 0;
        if (meets[0]) { r->q0 = r->lo; r->done = 1; return; }
        if (!meets[1]) { r->q0 = r->hi; r->done = 1; return; }
    }
    int t = 1;
    for (int l = 0; l < r->depth; ++l, ++r->steps) {
        if (meets[r->base + t - 1]) { r->hi = r->mid[t]; t = 2 * t; }
        else { r->lo = r->mid[t]; t = 2 * t + 1; }
    }
    if (r->steps >= RC_STEPS) { r->q0 = r->hi; r->done = 1; }
}

/* base_q of nv volumes (r[v].e, r[v].d, r[v].target set); depth = bisection
 * levels per launch (1..RC_MAX_DEPTH) */
static void rate_control(int nv, rc_vol *r, int depth) {
    vcg_enc **es = (vcg_enc **)malloc((size_t)nv * sizeof *es);
    const uint32_t **lists = (const uint32_t **)malloc((size_t)nv * sizeof *lists);
    uint8_t **zeros = (uint8_t **)malloc((size_t)nv * sizeof *zeros);
    int *ns = (int *)malloc((size_t)nv * sizeof *ns);
    const size_t cap = (size_t)nv * (RC_NODES + 2);
    vcg_enc **ge = (vcg_enc **)malloc(cap * sizeof *ge);
    float *gq = (float *)malloc(cap * sizeof *gq);
    uint64_t *gb = (uint64_t *)malloc(cap * 8), *bytes = (uint64_t *)malloc(cap * 8);
    int *flags = (int *)malloc(cap * sizeof *flags);
    if (!es || !lists || !zeros || !ns || !ge || !gq || !gb || !bytes || !flags) vc__panic("out of memory");
    for (int v = 0; v < nv; ++v) {
        rc_lattice(&r[v]);
        es[v] = r[v].e; lists[v] = r[v].list; zeros[v] = r[v].zero; ns[v] = r[v].ns;
    }
    vcg_enc_samples(nv, es, lists, ns, zeros);
    for (int v = 0; v < nv; ++v) rc_start(&r[v]);
    for (;;) {
        int ng = 0;
        for (int v = 0; v < nv; ++v) {
            if (r[v].done) continue;
            rc_propose(&r[v], depth);
            for (int j = 0; j < r[v].nq; ++j, ++ng) { ge[ng] = r[v].e; gq[ng] = r[v].qs[j]; gb[ng] = r[v].budget; }
        }
        if (!ng) break;
        vcg_enc_candidates(ng, ge, gq, gb, bytes, flags);
        for (int v = 0, g = 0; v < nv; ++v) {
            if (r[v].done) continue;
            rc_accept(&r[v], bytes + g, flags + g);
            g += r[v].nq;
        }
    }
    for (int v = 0; v < nv; ++v) { free(r[v].list); free(r[v].vox); free(r[v].zero); }
    free(es); free(lists); free(zeros); free(ns); free(ge); free(gq); free(gb); free(bytes); free(flags);
}

static void lod_schedule(float q0, int nlod, float qs[VC_NLOD]) {
    float q = q0;
    for (int l = 0; l < VC_NLOD; ++l) { qs[l] = l < nlod ? vc__clamp_q(q) : 0.0f; q *= 0.6f; }
}

/* ---- layout + assembly ---- */
/* An archive laid out on the host, for vcg_enc_assemble_many to fill: its
 * header and directory, each atom's payload and slot offsets, its length and
 * its host buffer. */
typedef struct {
    uint8_t *head, *img;
    size_t head_len;
    uint64_t *pay_off, *slot_off;
    uint64_t len;
} asm_plan;

/* pre: a buffer of pre_len bytes allocated for the archive ahead of time
 * (NULL if none); layout owns it: the plan's buffer, or freed.  On VC_OK the
 * caller assembles the plan and then releases it with plan_done. */
static vc_status layout(vc_dims d0, int nlod, float target, const float qs[VC_NLOD],
                        const uint32_t *lens, uint64_t total, uint8_t *pre, size_t pre_len, asm_plan *P) {
    uint64_t prefix[VC_NLOD + 1], apre[VC_NLOD + 1], nreg = 0, na = 0;
    uint32_t acx[VC_NLOD], acy[VC_NLOD], acz[VC_NLOD], nrx[VC_NLOD], nry[VC_NLOD], nrz[VC_NLOD];
    for (int l = 0; l < nlod; ++l) {
        vc_dims d = vc__lod_dims(d0, l);
        acx[l] = (d.nx + 31) / 32; acy[l] = (d.ny + 31) / 32; acz[l] = (d.nz + 31) / 32;
        vc__region_grid(d0, l, &nrz[l], &nry[l], &nrx[l]);
        prefix[l] = nreg; nreg += (uint64_t)nrz[l] * nry[l] * nrx[l];
        apre[l] = na; na += (uint64_t)acx[l] * acy[l] * acz[l];
    }
    prefix[nlod] = nreg; apre[nlod] = na;
    uint8_t *cls = (uint8_t *)calloc(nreg, 1);          /* 0 absent, 1 zero, 2 data */
    uint64_t *blen = (uint64_t *)calloc(nreg, 8), *boff = (uint64_t *)calloc(nreg, 8);
    uint64_t *pay_off = (uint64_t *)calloc(total, 8), *slot_off = (uint64_t *)calloc(total, 8);
    if (!cls || !blen || !boff || !pay_off || !slot_off) vc__panic("out of memory");
    for (int l = 0; l < nlod; ++l)
        for (uint32_t az = 0; az < acz[l]; ++az) for (uint32_t ay = 0; ay < acy[l]; ++ay) for (uint32_t ax = 0; ax < acx[l]; ++ax) {
            uint64_t ri = prefix[l] + ((uint64_t)(az / 32) * nry[l] + ay / 32) * nrx[l] + ax / 32;
            uint32_t len = lens[apre[l] + ((uint64_t)az * acy[l] + ay) * acx[l] + ax];
            if (len) { cls[ri] = 2; blen[ri] += len; }
            else if (cls[ri] < 1) cls[ri] = 1;
        }
    /* tools/vc_repack.c:213-247: directory after the header, blobs from the next
     * 4 KiB boundary, coarse LOD first, regions in raster order within a LOD */
    const uint64_t dir_off = FILE_HDR;
    uint64_t cursor = (dir_off + nreg * 16 + 4095) & ~4095ull;
    const uint64_t head_len = cursor;
    int too_big = 0;
    for (int l = nlod - 1; l >= 0; --l)
        for (uint64_t ri = prefix[l]; ri < prefix[l + 1]; ++ri) {
            if (cls[ri] != 2) continue;
            blen[ri] += L2_BLOCK;
            if (blen[ri] > 0xFFFFFFFFull) too_big = 1;      /* dir blob_len is u32 */
            boff[ri] = cursor; cursor += blen[ri];
        }
    vc_status rc = VC_OK;
    memset(P, 0, sizeof *P);
    if (too_big) { free(pre); rc = VC_ERR_RANGE; goto done; }
    /* per-atom slot and payload offsets: payloads follow the slot array in slot order */
    for (int l = 0; l < nlod; ++l)
        for (uint32_t rz = 0; rz < nrz[l]; ++rz) for (uint32_t ry = 0; ry < nry[l]; ++ry) for (uint32_t rx = 0; rx < nrx[l]; ++rx) {
            uint64_t ri = prefix[l] + ((uint64_t)rz * nry[l] + ry) * nrx[l] + rx;
            if (cls[ri] != 2) continue;
            uint64_t pc = boff[ri] + L2_BLOCK;
            for (uint32_t dz = 0; dz < 32; ++dz) { uint32_t az = rz * 32 + dz; if (az >= acz[l]) break;
            for (uint32_t dy = 0; dy < 32; ++dy) { uint32_t ay = ry * 32 + dy; if (ay >= acy[l]) break;
            for (uint32_t dx = 0; dx < 32; ++dx) { uint32_t ax = rx * 32 + dx; if (ax >= acx[l]) break;
                uint64_t i = apre[l] + ((uint64_t)az * acy[l] + ay) * acx[l] + ax;
                slot_off[i] = boff[ri] + (((uint64_t)dz * 32 + dy) * 32 + dx) * ATOM_SLOT;
                if (lens[i]) { pay_off[i] = pc; pc += lens[i]; }
            } } }
        }
    {
        uint8_t *head = (uint8_t *)calloc(head_len, 1);
        if (!head) vc__panic("out of memory");
        wr32(head + FH_MAGIC, VC_MAGIC); wr32(head + FH_VER, 2); wr32(head + FH_ATOM, VC_A);
        wr32(head + FH_RGN, R_ATOMS); wr32(head + FH_NLOD, (uint32_t)nlod);
        wr32(head + FH_NX, d0.nx); wr32(head + FH_NY, d0.ny); wr32(head + FH_NZ, d0.nz);
        wrf(head + FH_TRATIO, target);
        wr64(head + FH2_DIROFF, dir_off); wr64(head + FH2_DIRCNT, nreg); wr64(head + FH2_TOTLEN, cursor);
        for (int l = 0; l < VC_NLOD; ++l) wrf(head + FH_BASEQ + 4 * l, qs[l]);
        for (uint64_t ri = 0; ri < nreg; ++ri) {
            uint8_t *en = head + dir_off + ri * 16;
            if (cls[ri] == 2) { wr64(en, boff[ri]); wr32(en + 8, (uint32_t)blen[ri]); wr32(en + 12, 2); }
            else if (cls[ri] == 1) wr32(en + 12, 1);
        }
        uint8_t *img = pre && pre_len >= cursor ? pre : NULL;
        if (!img) { free(pre); pre = NULL; img = (uint8_t *)malloc(cursor); }
        if (!img) { free(head); rc = VC_ERR_OOM; goto done; }
        P->head = head; P->head_len = head_len; P->img = img; P->len = cursor;
        P->pay_off = pay_off; P->slot_off = slot_off;
        pay_off = slot_off = NULL;                     /* the plan's now */
    }
done:
    free(cls); free(blen); free(boff); free(pay_off); free(slot_off);
    return rc;
}

/* after the assembly: the archive (trimmed to its length) to *out */
static void plan_done(asm_plan *P, size_t pre_len, uint8_t **out, size_t *out_len) {
    uint8_t *img = P->img;
    if (pre_len > P->len) {                             /* give back the estimate's excess */
        uint8_t *t = (uint8_t *)realloc(img, P->len);
        if (t) img = t;
    }
    *out = img; *out_len = (size_t)P->len;
    free(P->head); free(P->pay_off); free(P->slot_off);
    memset(P, 0, sizeof *P);
}

/* Bisection levels per launch for a group: the most whose rounds (2^depth - 1
 * midpoints per volume) fit one launch of vcg_rc_rows() sample rows; the
 * first round, which adds both bounds, may take two. */
static int rc_depth(int nv, const vc_dims *dims) {
    uint64_t ns = 0;
    uint32_t ix[8];
    for (int v = 0; v < nv; ++v)
        ns += (uint64_t)lattice((dims[v].nx + 31) / 32, ix) * lattice((dims[v].ny + 31) / 32, ix) *
              lattice((dims[v].nz + 31) / 32, ix);
    int d = RC_MAX_DEPTH;
    while (d > 1 && ns * (uint64_t)((1 << d) - 1) > vcg_rc_rows()) --d;
    return d;
}

/* An archive size that the encode is expected to stay under: header and
 * directory, every region's slot array, and twice the LOD0 payload the
 * target ratio allows (all LODs together came to 1.3-1.8 times that on
 * synthetic volumes of 40-512 voxels per axis).  0 when larger than
 * PREFAULT_MAX: such an archive is not allocated ahead. */
#define PREFAULT_MAX (4ull << 30)
static uint64_t archive_estimate(vc_dims d0, int nlod, float target) {
    uint64_t nreg = 0;
    for (int l = 0; l < nlod; ++l) {
        uint32_t rz, ry, rx;
        vc__region_grid(d0, l, &rz, &ry, &rx);
        nreg += (uint64_t)rz * ry * rx;
    }
    const double pay = 2.0 * (double)d0.nx * d0.ny * d0.nz / (double)target;
    const uint64_t head = (FILE_HDR + nreg * 16 + 4095) & ~4095ull;
    if (pay > (double)PREFAULT_MAX) return 0;
    const uint64_t est = head + nreg * L2_BLOCK + (uint64_t)pay + (64u << 10);
    return est > PREFAULT_MAX ? 0 : est;
}

static uint64_t atom_total(vc_dims dims, int nlod) {
    uint64_t total = 0;
    for (int l = 0; l < nlod; ++l) {
        vc_dims dl = vc__lod_dims(dims, l);
        total += (uint64_t)((dl.nx + 31) / 32) * ((dl.ny + 31) / 32) * ((dl.nz + 31) / 32);
    }
    return total;
}

/* Encode nv volumes together: pyramids, rate control rounds shared across
 * volumes, one set of encode launches over every atom, then each archive.
 * qs_in (one volume only): explicit base_q per LOD instead of rate control. */
static void encode_group(int nv, const uint8_t *const *vols, int on_dev, const vc_dims *dims, const float *targets,
                         const float *qs_in, uint8_t **outs, size_t *out_lens, vc_status *st, float *q0_only) {
    rc_vol *r = (rc_vol *)calloc((size_t)nv, sizeof *r);
    vcg_enc **es = (vcg_enc **)malloc((size_t)nv * sizeof *es);
    float (*qs)[VC_NLOD] = (float (*)[VC_NLOD])malloc((size_t)nv * sizeof *qs);
    uint32_t **lens = (uint32_t **)calloc((size_t)nv, sizeof *lens);
    int *nlod = (int *)malloc((size_t)nv * sizeof *nlod), *ovf = (int *)malloc((size_t)nv * sizeof *ovf);
    uint64_t *total = (uint64_t *)malloc((size_t)nv * 8);
    uint8_t **pre = (uint8_t **)calloc((size_t)nv, sizeof *pre);
    size_t *pre_len = (size_t *)calloc((size_t)nv, sizeof *pre_len);
    if (!r || !es || !qs || !lens || !nlod || !ovf || !total || !pre || !pre_len) vc__panic("out of memory");
    uint32_t (*d3)[3] = (uint32_t (*)[3])malloc((size_t)nv * sizeof *d3);
    if (!d3) vc__panic("out of memory");
    for (int v = 0; v < nv; ++v) {
        nlod[v] = vc__nlod(dims[v]);
        d3[v][0] = dims[v].nx; d3[v][1] = dims[v].ny; d3[v][2] = dims[v].nz;
    }
    vcg_enc_begin_many(nv, vols, on_dev, (const uint32_t (*)[3])d3, nlod, es);
    free(d3);
    for (int v = 0; v < nv; ++v) { r[v].e = es[v]; r[v].d = dims[v]; r[v].target = targets[v]; }
    if (!q0_only) {
        /* the archives' host memory, allocated at its estimate and faulted in
         * by host threads while the device runs rate control and the encode
         * (fresh pages otherwise dominate the copy-out); at most 32 GiB */
        uint64_t sum = 0;
        for (int v = 0; v < nv; ++v) {
            const uint64_t est = archive_estimate(dims[v], nlod[v], targets[v]);
            if (!est || sum + est > (32ull << 30)) continue;
            pre[v] = (uint8_t *)malloc((size_t)est);
            if (pre[v]) { pre_len[v] = (size_t)est; sum += est; }
        }
        vcg_host_prefault_begin(pre, pre_len, nv);
    }
    if (qs_in) {
        for (int l = 0; l < VC_NLOD; ++l) qs[0][l] = l < nlod[0] ? vc__clamp_q(qs_in[l]) : 0.0f;
    } else {
        rate_control(nv, r, rc_depth(nv, dims));
        for (int v = 0; v < nv; ++v) lod_schedule(r[v].q0, nlod[v], qs[v]);
    }
    if (q0_only) {
        *q0_only = r[0].q0;
    } else {
        for (int v = 0; v < nv; ++v) {
            total[v] = atom_total(dims[v], nlod[v]);
            lens[v] = (uint32_t *)malloc(total[v] * 4);
            if (!lens[v]) vc__panic("out of memory");
        }
        vcg_enc_all(nv, es, (const float (*)[VC_NLOD])qs, lens, ovf);
        vcg_host_prefault_wait();
        /* every archive laid out, then all assembled in one device pass */
        asm_plan *P = (asm_plan *)calloc((size_t)nv, sizeof *P);
        vcg_asm_job *J = (vcg_asm_job *)calloc((size_t)nv, sizeof *J);
        if (!P || !J) vc__panic("out of memory");
        int nj = 0;
        for (int v = 0; v < nv; ++v) {
            outs[v] = NULL; out_lens[v] = 0;
            if (ovf[v]) { free(pre[v]); st[v] = VC_ERR_RANGE; continue; }   /* a payload exceeded the coder bound */
            st[v] = layout(dims[v], nlod[v], targets[v], qs[v], lens[v], total[v], pre[v], pre_len[v], &P[v]);
            if (st[v] != VC_OK) continue;
            J[nj++] = (vcg_asm_job){es[v], P[v].head, P[v].head_len, P[v].pay_off, P[v].slot_off, lens[v],
                                    total[v], P[v].img, P[v].len};
        }
        if (nj) vcg_enc_assemble_many(nj, J);
        for (int v = 0; v < nv; ++v)
            if (st[v] == VC_OK) plan_done(&P[v], P[v].img == pre[v] ? pre_len[v] : 0, &outs[v], &out_lens[v]);
        free(P); free(J);
    }
    for (int v = 0; v < nv; ++v) { vcg_enc_end(es[v]); free(lens[v]); }
    free(r); free(es); free(qs); free(lens); free(nlod); free(ovf); free(total); free(pre); free(pre_len);
}

static vc_status encode_impl(const uint8_t *vol, int on_dev, vc_dims dims, float target, const float *qs_in,
                             uint8_t **out, size_t *out_len, float *q0_only) {
    if (!vol || (!q0_only && (!out || !out_len))) return VC_ERR_RANGE;
    if (!vc__dims_valid(dims) || !(target >= 1.0f)) return VC_ERR_RANGE;
    vc__require_gpu();
    uint8_t *o = NULL; size_t ol = 0;
    vc_status st = VC_OK;
    encode_group(1, &vol, on_dev, &dims, &target, qs_in, &o, &ol, &st, q0_only);
    if (!q0_only) { *out = o; *out_len = ol; }
    return st;
}

void vcx_test_rc_rows(uint32_t rows) { vcg_set_rc_rows(rows); }

/* device bytes a volume holds while its group encodes: the pyramid (at most
 * 8/7 of LOD0) and the sample coefficients */
static uint64_t group_bytes(vc_dims d) {
    uint32_t ix[8];
    uint64_t s = (uint64_t)lattice((d.nx + 31) / 32, ix) * lattice((d.ny + 31) / 32, ix) *
                 lattice((d.nz + 31) / 32, ix);
    return (uint64_t)d.nx * d.ny * d.nz / 7 * 8 + s * (32768u * 4u + 16u);
}
#define VC_GROUP_BYTES (24ull << 30)

vc_status vcx_encode_batch(size_t n, const uint8_t *const *vols, int vols_on_device, const vc_dims *dims,
                           const float *target_ratios, uint8_t **outs, size_t *out_lens, vc_status *statuses) {
    if (n == 0) return VC_OK;
    if (!vols || !dims || !target_ratios || !outs || !out_lens || !statuses) return VC_ERR_RANGE;
    vc__require_gpu();
    vc_status first = VC_OK;
    size_t *idx = (size_t *)malloc(n * sizeof *idx);
    const uint8_t **gv = (const uint8_t **)malloc(n * sizeof *gv);
    vc_dims *gd = (vc_dims *)malloc(n * sizeof *gd);
    float *gt = (float *)malloc(n * sizeof *gt);
    uint8_t **go = (uint8_t **)malloc(n * sizeof *go);
    size_t *gl = (size_t *)malloc(n * sizeof *gl);
    vc_status *gs = (vc_status *)malloc(n * sizeof *gs);
    if (!idx || !gv || !gd || !gt || !go || !gl || !gs) vc__panic("out of memory");
    size_t i = 0;
    while (i < n) {
        /* the next group: valid volumes in order while their device bytes fit */
        int ng = 0;
        uint64_t bytes = 0;
        for (; i < n; ++i) {
            outs[i] = NULL; out_lens[i] = 0;
            if (!vols[i] || !vc__dims_valid(dims[i]) || !(target_ratios[i] >= 1.0f)) {
                statuses[i] = VC_ERR_RANGE;
                continue;
            }
            const uint64_t b = group_bytes(dims[i]);
            if (ng && bytes + b > VC_GROUP_BYTES) break;
            bytes += b;
            idx[ng] = i; gv[ng] = vols[i]; gd[ng] = dims[i]; gt[ng] = target_ratios[i];
            ++ng;
        }
        if (!ng) break;
        vcg_enc_arena_open();
        encode_group(ng, gv, vols_on_device, gd, gt, NULL, go, gl, gs, NULL);
        vcg_enc_arena_close();
        for (int k = 0; k < ng; ++k) { outs[idx[k]] = go[k]; out_lens[idx[k]] = gl[k]; statuses[idx[k]] = gs[k]; }
    }
    for (size_t k = 0; k < n; ++k) if (statuses[k] != VC_OK && first == VC_OK) first = statuses[k];
    free(idx); free(gv); free(gd); free(gt); free(go); free(gl); free(gs);
    return first;
}

vc_status vc_encode(const uint8_t *vol, vc_dims dims, float target_ratio, uint8_t **out_archive, size_t *out_len) {
    return encode_impl(vol, 0, dims, target_ratio, NULL, out_archive, out_len, NULL);
}
vc_status vcx_encode_q(const uint8_t *vol, vc_dims dims, float target_ratio, const float base_q[VC_NLOD],
                       uint8_t **out_archive, size_t *out_len) {
    if (!base_q) return VC_ERR_RANGE;
    return encode_impl(vol, 0, dims, target_ratio, base_q, out_archive, out_len, NULL);
}
vc_status vcx_encode_device(const void *dev_vol, vc_dims dims, float target_ratio, uint8_t **out_archive,
                            size_t *out_len) {
    return encode_impl((const uint8_t *)dev_vol, 1, dims, target_ratio, NULL, out_archive, out_len, NULL);
}
float vcx_rate_control(const uint8_t *vol, vc_dims dims, float target_ratio) {
    float q0 = -1.0f;
    if (encode_impl(vol, 0, dims, target_ratio, NULL, NULL, NULL, &q0) != VC_OK) return -1.0f;
    return q0;
}
