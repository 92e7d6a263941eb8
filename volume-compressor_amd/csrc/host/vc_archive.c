/* vc_archive.c — reader: open/validate (host), index queries (host), and the
 * decode entry points, which run on the GPU from a device copy of the archive. */
#include <stdlib.h>

#include "vc_host.h"

/* ---- open (src/vc/vc.c:1285-1338) ---- */
static int parse_header(vc_archive *a, const uint8_t *h, uint64_t len) {
    if (rd32(h + FH_MAGIC) != VC_MAGIC) return -1;
    if (rd32(h + FH_ATOM) != VC_A) return -1;
    if (rd32(h + FH_RGN) != R_ATOMS) return -1;
    uint32_t ver = rd32(h + FH_VER), nlod = rd32(h + FH_NLOD);
    if (nlod == 0 || nlod > VC_NLOD) return -1;
    a->len = len; a->nlod = (int)nlod;
    a->dims0.nx = rd32(h + FH_NX); a->dims0.ny = rd32(h + FH_NY); a->dims0.nz = rd32(h + FH_NZ);
    for (int l = 0; l < VC_NLOD; ++l) a->base_q[l] = rdf(h + FH_BASEQ + 4 * l);
    if (ver == 2) {
        a->version = 2;
        a->dir_off = rd64(h + FH2_DIROFF);
        uint64_t cnt = rd64(h + FH2_DIRCNT), acc = 0;
        if (a->dir_off > len || cnt > (len - a->dir_off) / 16) return -1;
        for (int l = 0; l < a->nlod; ++l) {
            a->region_prefix[l] = acc;
            uint32_t rz, ry, rx; vc__region_grid(a->dims0, l, &rz, &ry, &rx);
            acc += (uint64_t)rz * ry * rx;
        }
        a->region_prefix[a->nlod] = acc;
        return acc == cnt ? 0 : -1;          /* the directory covers every region */
    }
    if (ver != 1 || rd32(h + FH_LAYOUT) != 1) return -1;
    a->version = 1;
    a->l1_off = rd64(h + FH_L1OFF);
    a->l1_cap = rd64(h + FH_L1CAP);
    if (a->l1_cap == 0 || (a->l1_cap & (a->l1_cap - 1))) return -1;
    if (a->l1_off > len || a->l1_cap > (len - a->l1_off) / 16) return -1;
    return 0;
}

static vc_archive *new_archive(void) {
    vc_archive *a = (vc_archive *)calloc(1, sizeof *a);
    if (!a) vc__panic("out of memory");
    pthread_mutex_init(&a->mu, NULL);
    return a;
}

vc_archive *vc_open(const uint8_t *archive, size_t len) {
    if (!archive || len < FILE_HDR) return NULL;
    vc_archive *a = new_archive();
    a->buf = archive;
    if (parse_header(a, archive, len) != 0) { pthread_mutex_destroy(&a->mu); free(a); return NULL; }
    return a;
}

vc_archive *vc_open_streaming(vc_read_fn read, void *ud, uint64_t total_len) {
    if (!read || total_len < FILE_HDR) return NULL;
    uint8_t hdr[FILE_HDR];
    if (read(ud, 0, FILE_HDR, hdr) != VC_OK) return NULL;
    vc_archive *a = new_archive();
    a->read = read; a->ud = ud;
    if (parse_header(a, hdr, total_len) != 0) { pthread_mutex_destroy(&a->mu); free(a); return NULL; }
    return a;
}

void vc_close(vc_archive *a) {
    if (!a) return;
    if (a->dev) vcg_arc_destroy(a->dev);
    vc__stream_free(a);
    pthread_mutex_destroy(&a->mu);
    free(a);
}

vc_status vc_lod_dims(const vc_archive *a, int lod, vc_dims *out) {
    if (!a || lod < 0 || lod >= a->nlod || !out) return VC_ERR_RANGE;
    *out = vc__lod_dims(a->dims0, lod);
    return VC_OK;
}

int vc_archive_version(const vc_archive *a) { return a ? a->version : 0; }

/* ---- host-side index (src/vc/vc.c:1264-1280, 1349-1402) ---- */
int vc__rd(const vc_archive *a, uint64_t off, uint32_t len, uint8_t *dst) {
    if (a->read) return a->read(a->ud, off, len, dst) == VC_OK ? 0 : -1;
    if (off > a->len || len > a->len - off) return -1;
    memcpy(dst, a->buf + off, len);
    return 0;
}

static uint8_t resolve_slot(const vc_archive *a, uint64_t so, uint64_t *off, uint32_t *len) {
    uint8_t s[ATOM_SLOT];
    if (vc__rd(a, so, ATOM_SLOT, s) != 0) return AF_ABSENT;
    uint8_t f = s[12];
    if (f == AF_PRESENT) { *off = rd64(s); *len = rd32(s + 8); }
    return f;
}

int vc__region_block(const vc_archive *a, int lod, uint32_t rz, uint32_t ry, uint32_t rx, uint64_t *blk_out) {
    uint64_t blk;
    *blk_out = 0;
    if (a->version == 2) {
        uint32_t nrz, nry, nrx; vc__region_grid(a->dims0, lod, &nrz, &nry, &nrx);
        if (rz >= nrz || ry >= nry || rx >= nrx) return AF_ABSENT;
        uint64_t ri = a->region_prefix[lod] + ((uint64_t)rz * nry + ry) * nrx + rx;
        if (ri >= a->region_prefix[a->nlod]) return AF_ABSENT;
        uint8_t e[16];
        if (vc__rd(a, a->dir_off + ri * 16, 16, e) != 0) return AF_ABSENT;
        uint32_t f = rd32(e + 12);
        if (f == 0) return AF_ABSENT;
        if (f == 1) return AF_ZERO;
        blk = rd64(e);
    } else {
        uint64_t key = vc__region_key(lod, rz, ry, rx);
        uint64_t mask = a->l1_cap - 1, i = vc__mix64(key) & mask;
        blk = 0;
        for (uint64_t probe = 0; probe < a->l1_cap; ++probe) {
            uint8_t e[16];
            if (vc__rd(a, a->l1_off + i * 16, 16, e) != 0) return AF_ABSENT;
            uint64_t br = rd64(e + 8);
            if (br == 0) return AF_ABSENT;
            if (rd64(e) == key) { blk = br; break; }
            i = (i + 1) & mask;
        }
        if (blk == 0) return AF_ABSENT;
        if (blk == ZERO_REGION) return AF_ZERO;
    }
    if (blk > a->len || a->len - blk < L2_BLOCK) return AF_ABSENT;     /* corrupt guard */
    *blk_out = blk;
    return AF_PRESENT;
}

uint8_t vc__resolve(const vc_archive *a, int lod, uint32_t az, uint32_t ay, uint32_t ax, uint64_t *off,
                    uint32_t *len) {
    uint64_t blk;
    const int kind = vc__region_block(a, lod, az / R_ATOMS, ay / R_ATOMS, ax / R_ATOMS, &blk);
    if (kind != AF_PRESENT) return (uint8_t)kind;
    uint64_t si = ((uint64_t)(az & (R_ATOMS - 1)) * R_ATOMS + (ay & (R_ATOMS - 1))) * R_ATOMS + (ax & (R_ATOMS - 1));
    return resolve_slot(a, blk + si * ATOM_SLOT, off, len);
}

vc_cover vc_atom_coverage(const vc_archive *a, int lod, uint32_t az, uint32_t ay, uint32_t ax) {
    if (!a || lod < 0 || lod >= a->nlod) return VC_ABSENT;
    uint64_t off; uint32_t len;
    uint8_t f = vc__resolve(a, lod, az, ay, ax, &off, &len);
    return f == AF_PRESENT ? VC_PRESENT : f == AF_ZERO ? VC_KNOWN_ZERO : VC_ABSENT;
}

vc_cover vc_atom_payload_range(const vc_archive *a, int lod, uint32_t az, uint32_t ay, uint32_t ax,
                               uint64_t *off, uint32_t *len) {
    if (off) *off = 0;
    if (len) *len = 0;
    if (!a || lod < 0 || lod >= a->nlod) return VC_ABSENT;
    uint64_t o = 0; uint32_t l = 0;
    uint8_t f = vc__resolve(a, lod, az, ay, ax, &o, &l);
    if (f == AF_PRESENT) { if (off) *off = o; if (len) *len = l; return VC_PRESENT; }
    return f == AF_ZERO ? VC_KNOWN_ZERO : VC_ABSENT;
}

vc_cover vc_region_blob_range(const vc_archive *a, int lod, uint32_t rz, uint32_t ry, uint32_t rx,
                              uint64_t *off, uint64_t *len) {
    if (off) *off = 0;
    if (len) *len = 0;
    if (!a || a->version != 2 || lod < 0 || lod >= a->nlod) return VC_ABSENT;
    uint32_t nrz, nry, nrx; vc__region_grid(a->dims0, lod, &nrz, &nry, &nrx);
    if (rz >= nrz || ry >= nry || rx >= nrx) return VC_ABSENT;
    uint64_t ri = a->region_prefix[lod] + ((uint64_t)rz * nry + ry) * nrx + rx;
    if (ri >= a->region_prefix[a->nlod]) return VC_ABSENT;
    uint8_t e[16];
    if (vc__rd(a, a->dir_off + ri * 16, 16, e) != 0) return VC_ABSENT;
    uint32_t f = rd32(e + 12);
    if (f == 1) return VC_KNOWN_ZERO;
    if (f != 2) return VC_ABSENT;
    if (off) *off = rd64(e);
    if (len) *len = rd32(e + 8);
    return VC_PRESENT;
}

/* ---- device copy ---- */
vcg_arc *vc__arc_dev(vc_archive *a) {
    pthread_mutex_lock(&a->mu);
    if (!a->dev) {
        vc__require_gpu();
        vcg_geom *g = (vcg_geom *)calloc(1, sizeof *g);
        g->len = a->len; g->version = a->version; g->nlod = a->nlod;
        g->l1_off = a->l1_off; g->l1_cap = a->l1_cap; g->dir_off = a->dir_off;
        for (int l = 0; l <= VC_NLOD; ++l) g->region_prefix[l] = a->region_prefix[l];
        for (int l = 0; l < a->nlod; ++l) {
            vc_dims d = vc__lod_dims(a->dims0, l);
            g->lnx[l] = d.nx; g->lny[l] = d.ny; g->lnz[l] = d.nz;
            g->acx[l] = (d.nx + VC_A - 1) / VC_A; g->acy[l] = (d.ny + VC_A - 1) / VC_A; g->acz[l] = (d.nz + VC_A - 1) / VC_A;
            vc__region_grid(a->dims0, l, &g->nrz[l], &g->nry[l], &g->nrx[l]);
            vc__dq(a->base_q[l], g->dq_step[l], g->dq_bias[l]);
        }
        a->dev = vcg_arc_create(a->buf, g);
        free(g);
    }
    vcg_arc *d = a->dev;
    pthread_mutex_unlock(&a->mu);
    return d;
}

/* ---- host-resolved path: gather payload bytes, decode on the GPU ---- */
vc_status vc__decode_atoms_host(vc_archive *a, size_t n, const int32_t *q, uint8_t *out, vc_status *per) {
    vc__require_gpu();
    uint64_t *offs = (uint64_t *)calloc(n ? n : 1, 8);
    uint32_t *lens = (uint32_t *)calloc(n ? n : 1, 4);
    int32_t *st = (int32_t *)calloc(n ? n : 1, 4);
    size_t cap = 1 << 20, used = 0;
    uint8_t *pool = (uint8_t *)malloc(cap);
    int *pre = (int *)calloc(n ? n : 1, sizeof(int));        /* pre-decode failure */
    int *lod_of = (int *)calloc(n ? n : 1, sizeof(int));
    if (!offs || !lens || !st || !pool || !pre || !lod_of) vc__panic("out of memory");
    for (size_t i = 0; i < n; ++i) {
        int lod = q[4 * i];
        lod_of[i] = lod;
        uint64_t off = 0; uint32_t len = 0;
        uint8_t f = vc__resolve(a, lod, (uint32_t)q[4 * i + 3], (uint32_t)q[4 * i + 2], (uint32_t)q[4 * i + 1], &off, &len);
        if (f != AF_PRESENT) continue;                                  /* zeros, OK */
        if (off == SLOT_ABSENT || len < 1 || off > a->len || len > a->len - off) { pre[i] = 1; continue; }
        if (a->read && len > VCG_RC_CAP) { pre[i] = 1; continue; }   /* streaming buffer bound, vc.c:1276-1277 */
        if (used + len > cap) {
            while (used + len > cap) cap *= 2;
            pool = (uint8_t *)realloc(pool, cap);
            if (!pool) vc__panic("out of memory");
        }
        if (vc__rd(a, off, len, pool + used) != 0) { pre[i] = 1; continue; }
        offs[i] = used; lens[i] = len; used += len;
    }
    vc_status rc = VC_OK;
    /* decode LOD by LOD (one dequant table per launch) */
    for (int lod = 0; lod < VC_NLOD; ++lod) {
        size_t m = 0;
        for (size_t i = 0; i < n; ++i) if (lod_of[i] == lod) m++;
        if (!m) continue;
        uint64_t *o2 = (uint64_t *)malloc(m * 8); uint32_t *l2 = (uint32_t *)malloc(m * 4);
        size_t *ix = (size_t *)malloc(m * sizeof(size_t));
        uint8_t *dec = (uint8_t *)malloc(m * (size_t)VC_A3);
        int32_t *s2 = (int32_t *)calloc(m, 4);
        if (!o2 || !l2 || !ix || !dec || !s2) vc__panic("out of memory");
        size_t k = 0;
        for (size_t i = 0; i < n; ++i) if (lod_of[i] == lod) { ix[k] = i; o2[k] = offs[i]; l2[k] = pre[i] ? 0 : lens[i]; k++; }
        int32_t step[VCG_NFREQ], bias[VCG_NFREQ];
        vc__dq(a->base_q[lod], step, bias);
        vcg_decode_payloads(m, pool, o2, l2, step, bias, dec, s2);
        for (k = 0; k < m; ++k) {
            size_t i = ix[k];
            vc_status s = pre[i] || s2[k] ? VC_ERR_FORMAT : VC_OK;
            if (s != VC_OK) memset(out + i * (size_t)VC_A3, 0, VC_A3);
            else memcpy(out + i * (size_t)VC_A3, dec + k * (size_t)VC_A3, VC_A3);
            if (per) per[i] = s;
            if (s != VC_OK && rc == VC_OK) rc = s;
        }
        free(o2); free(l2); free(ix); free(dec); free(s2);
    }
    free(offs); free(lens); free(st); free(pool); free(pre); free(lod_of);
    return rc;
}

/* ---- decode entry points ---- */
static int atom_in_range(const vc_archive *a, int lod, int ax, int ay, int az) {
    vc_dims d = vc__lod_dims(a->dims0, lod);
    uint32_t acx = (d.nx + VC_A - 1) / VC_A, acy = (d.ny + VC_A - 1) / VC_A, acz = (d.nz + VC_A - 1) / VC_A;
    return (uint32_t)ax < acx && (uint32_t)ay < acy && (uint32_t)az < acz;
}

vc_status vcx_decode_atoms(vc_archive *a, size_t n, const int32_t *q, uint8_t *out, vc_status *per) {
    if (!a || (!out && n) || (!q && n)) return VC_ERR_RANGE;
    for (size_t i = 0; i < n; ++i)
        if (q[4 * i] < 0 || q[4 * i] >= a->nlod || !atom_in_range(a, q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]))
            return VC_ERR_RANGE;
    if (n == 0) return VC_OK;
    vcg_arc *d;
    if (!a->buf) {
        d = vc__stream_begin(a, 0, NULL, n, q);
        if (!d) return vc__decode_atoms_host(a, n, q, out, per);
    } else {
        d = vc__arc_dev(a);
        if (!d) return VC_ERR_OOM;
    }
    int32_t *st = (int32_t *)malloc(n * 4);
    if (!st) { if (!a->buf) vc__stream_end(a); return VC_ERR_OOM; }
    int rc = vcg_decode_list(d, n, q, out, st);
    if (!a->buf) vc__stream_end(a);
    for (size_t i = 0; i < n; ++i) if (per) per[i] = st[i] ? VC_ERR_FORMAT : VC_OK;
    free(st);
    return rc ? VC_ERR_FORMAT : VC_OK;
}

_Static_assert(VCX_NOALLOC_MAX_ATOMS == VCG_NOALLOC_MAX, "no-alloc batch limit");

/* No counterpart in the reference: vcx_decode_atoms for callers that cannot
 * allocate (device-resident quads and output, static decoder scratch). */
vc_status vcx_decode_atoms_noalloc(vc_archive *a, size_t n, const int32_t *quads_dev, uint8_t *out_dev,
                                   int32_t *status_dev, void *stream) {
    if (!a || !a->buf) return VC_ERR_RANGE;
    if (n == 0) return VC_OK;
    if (!quads_dev || !out_dev || n > VCX_NOALLOC_MAX_ATOMS) return VC_ERR_RANGE;
    pthread_mutex_lock(&a->mu);
    vcg_arc *d = a->dev;
    pthread_mutex_unlock(&a->mu);
    if (!d) return VC_ERR_RANGE;               /* creating the device copy would allocate */
    return vcg_list_launch_noalloc(d, n, quads_dev, out_dev, status_dev, stream) == 0 ? VC_OK : VC_ERR_RANGE;
}

/* src/vc/vc.c:1404-1421 */
vc_status vc_decode_atom(vc_archive *a, int lod, int ax, int ay, int az, uint8_t out[VC_ATOM3]) {
    if (!a || lod < 0 || lod >= a->nlod || !out) return VC_ERR_RANGE;
    if (!atom_in_range(a, lod, ax, ay, az)) return VC_ERR_RANGE;
    int32_t q[4] = {lod, ax, ay, az};
    return vcx_decode_atoms(a, 1, q, out, NULL);
}

static vc_status region_check(const vc_archive *a, int lod, vc_box b) {
    vc_dims d = vc__lod_dims(a->dims0, lod);
    if (b.x1 > d.nx || b.y1 > d.ny || b.z1 > d.nz) return VC_ERR_RANGE;
    if (b.x0 >= b.x1 || b.y0 >= b.y1 || b.z0 >= b.z1) return VC_ERR_RANGE;
    return VC_OK;
}
vc_status vc__region_check(const vc_archive *a, int lod, vc_box b) { return region_check(a, lod, b); }
int vc__atom_ok(const vc_archive *a, int lod, int ax, int ay, int az) {
    return lod >= 0 && lod < a->nlod && atom_in_range(a, lod, ax, ay, az);
}

/* Same atom set and clipped copy as src/vc/vc.c:1535-1555.  When an atom's
 * slot is invalid the call still writes every other atom and reads that one
 * as zeros, returning VC_ERR_FORMAT (the reference returns at the first such
 * atom, leaving the rest of `out` unwritten). */
static vc_status region_host(vc_archive *a, int lod, vc_box b, uint8_t *out) {
    uint32_t ax0 = b.x0 / VC_A, ax1 = (b.x1 - 1) / VC_A, ay0 = b.y0 / VC_A, ay1 = (b.y1 - 1) / VC_A;
    uint32_t az0 = b.z0 / VC_A, az1 = (b.z1 - 1) / VC_A;
    size_t n = (size_t)(ax1 - ax0 + 1) * (ay1 - ay0 + 1) * (az1 - az0 + 1);
    int32_t *q = (int32_t *)malloc(n * 16);
    uint8_t *atoms = (uint8_t *)malloc(n * (size_t)VC_A3);
    if (!q || !atoms) vc__panic("out of memory");
    size_t k = 0;
    for (uint32_t az = az0; az <= az1; ++az) for (uint32_t ay = ay0; ay <= ay1; ++ay) for (uint32_t ax = ax0; ax <= ax1; ++ax) {
        q[4 * k] = lod; q[4 * k + 1] = (int)ax; q[4 * k + 2] = (int)ay; q[4 * k + 3] = (int)az; k++;
    }
    vc_status rc = vc__decode_atoms_host(a, n, q, atoms, NULL);
    uint64_t ox = b.x1 - b.x0, oy = b.y1 - b.y0;
    k = 0;
    for (uint32_t az = az0; az <= az1; ++az) for (uint32_t ay = ay0; ay <= ay1; ++ay) for (uint32_t ax = ax0; ax <= ax1; ++ax, ++k) {
        const uint8_t *at = atoms + k * (size_t)VC_A3;
        for (uint32_t z = 0; z < VC_A; ++z) { uint32_t vz = az * VC_A + z; if (vz < b.z0 || vz >= b.z1) continue;
        for (uint32_t y = 0; y < VC_A; ++y) { uint32_t vy = ay * VC_A + y; if (vy < b.y0 || vy >= b.y1) continue;
            uint32_t xs = ax * VC_A < b.x0 ? b.x0 - ax * VC_A : 0;
            uint32_t xe = ax * VC_A + VC_A > b.x1 ? b.x1 - ax * VC_A : VC_A;
            memcpy(out + ((uint64_t)(vz - b.z0) * oy + (vy - b.y0)) * ox + (ax * VC_A + xs - b.x0),
                   at + (z * VC_A + y) * VC_A + xs, xe - xs);
        } }
    }
    free(q); free(atoms);
    return rc;
}

vc_status vc_decode_region(vc_archive *a, int lod, vc_box box, uint8_t *out) {
    if (!a || lod < 0 || lod >= a->nlod || !out) return VC_ERR_RANGE;
    vc_status s = region_check(a, lod, box);
    if (s != VC_OK) return s;
    vcg_box b = {box.x0, box.y0, box.z0, box.x1, box.y1, box.z1};
    if (!a->buf) {
        vcg_arc *d = vc__stream_begin(a, lod, &box, 0, NULL);
        if (!d) return region_host(a, lod, box, out);
        const int rc = vcg_decode_box(d, lod, &b, out, 1, 1);
        vc__stream_end(a);
        return rc ? VC_ERR_FORMAT : VC_OK;
    }
    vcg_arc *d = vc__arc_dev(a);
    if (!d) return VC_ERR_OOM;
    return vcg_decode_box(d, lod, &b, out, 1, 1) ? VC_ERR_FORMAT : VC_OK;
}

vc_status vc_decode_lod(vc_archive *a, int lod, uint8_t *out_vol, vc_dims *out_dims) {
    if (!a || lod < 0 || lod >= a->nlod || !out_vol) return VC_ERR_RANGE;
    vc_dims d = vc__lod_dims(a->dims0, lod);
    if (out_dims) *out_dims = d;
    vc_box b = {0, 0, 0, d.nx, d.ny, d.nz};
    return vc_decode_region(a, lod, b, out_vol);
}

vc_status vcx_decode_region_device(vc_archive *a, int lod, vc_box box, void *dev_out, int sync) {
    if (!a || lod < 0 || lod >= a->nlod || !dev_out) return VC_ERR_RANGE;
    vc_status s = region_check(a, lod, box);
    if (s != VC_OK) return s;
    vcg_box b = {box.x0, box.y0, box.z0, box.x1, box.y1, box.z1};
    if (!a->buf) {
        vcg_arc *d = vc__stream_begin(a, lod, &box, 0, NULL);
        if (d) {
            const int rc = vcg_decode_box(d, lod, &b, (uint8_t *)dev_out, 0, sync);
            vc__stream_end(a);
            return rc ? VC_ERR_FORMAT : VC_OK;
        }
        /* host-resolved decode, then one copy into the device buffer */
        const size_t nb = (size_t)(box.x1 - box.x0) * (box.y1 - box.y0) * (box.z1 - box.z0);
        uint8_t *tmp = (uint8_t *)malloc(nb);
        if (!tmp) return VC_ERR_OOM;
        vc_status st = region_host(a, lod, box, tmp);
        vcg_copy(dev_out, tmp, nb, 1);
        free(tmp);
        return st;
    }
    vcg_arc *d = vc__arc_dev(a);
    if (!d) return VC_ERR_OOM;
    return vcg_decode_box(d, lod, &b, (uint8_t *)dev_out, 0, sync) ? VC_ERR_FORMAT : VC_OK;
}

vc_status vcx_decode_lod_device(vc_archive *a, int lod, void *dev_out, int sync) {
    if (!a || lod < 0 || lod >= a->nlod || !dev_out) return VC_ERR_RANGE;
    vc_dims d = vc__lod_dims(a->dims0, lod);
    vc_box b = {0, 0, 0, d.nx, d.ny, d.nz};
    return vcx_decode_region_device(a, lod, b, dev_out, sync);
}
