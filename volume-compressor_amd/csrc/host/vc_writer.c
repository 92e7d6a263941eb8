/* vc_writer.c — the appendable v1 archive writer (src/vc/vc.c:767-1245).
 *
 * File layout and index semantics are the reference's: 128-byte header,
 * preallocated open-addressed L1 region table (capacity = smallest power of 2
 * >= 1024 and >= 2 x regions), 512 KiB L2 slot blocks and payloads appended at
 * the cursor, cursor persisted and the file trimmed on close.  Appends from
 * one thread produce byte-identical files to the reference's writer.
 *
 * Differences by design: the reference publishes slots lock-free (CAS +
 * release stores); here every operation holds one mutex and atom encoding
 * runs on the GPU (vc_append_box encodes all of its atoms in one batch).
 */
#define _GNU_SOURCE
#include <fcntl.h>
#include <stdlib.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "vc_host.h"

#define RESERVE (1ull << 40)   /* virtual reservation; the file grows under it */
#define GROW    (64ull << 20)

struct vc_writer {
    int fd;
    uint8_t *map;
    uint64_t map_len, file_len, cursor;
    int fixed;
    vc_dims dims0;
    int nlod;
    float base_q[VC_NLOD];
    pthread_mutex_t mu;
};

static int ensure(vc_writer *w, uint64_t need) {
    if (need <= w->file_len) return 0;
    uint64_t nl = w->file_len;
    while (nl < need) nl += GROW;
    if (w->fixed) {
        if (nl > w->map_len) nl = w->map_len;
        if (need > w->map_len) return -1;
        if (ftruncate(w->fd, (off_t)nl) != 0) return -1;
    } else {
        if (ftruncate(w->fd, (off_t)nl) != 0) return -1;
        void *m = mremap(w->map, w->map_len, nl, MREMAP_MAYMOVE);
        if (m == MAP_FAILED) return -1;
        w->map = (uint8_t *)m; w->map_len = nl;
    }
    w->file_len = nl;
    return 0;
}
static uint64_t alloc_bytes(vc_writer *w, uint64_t n) {
    uint64_t off = w->cursor;
    if (ensure(w, off + n) != 0) return 0;
    w->cursor += n;
    return off;
}

vc_writer *vc_create(const char *path, vc_dims d0, float target_ratio) {
    if (!path || !vc__dims_valid(d0) || !(target_ratio >= 1.0f)) return NULL;
    int fd = open(path, O_RDWR | O_CREAT | O_TRUNC, 0644);
    if (fd < 0) return NULL;
    int nlod = vc__nlod(d0);
    uint64_t regions = 0;
    for (int l = 0; l < nlod; ++l) {
        uint32_t rz, ry, rx; vc__region_grid(d0, l, &rz, &ry, &rx);
        regions += (uint64_t)rz * ry * rx;
    }
    uint64_t cap = 1024;                                  /* src/vc/vc.c:1020-1021 */
    while (cap < regions * 2) cap <<= 1;
    uint64_t init_len = FILE_HDR + cap * 16;
    if (ftruncate(fd, (off_t)init_len) != 0) { close(fd); return NULL; }
    vc_writer *w = (vc_writer *)calloc(1, sizeof *w);
    if (!w) vc__panic("out of memory");
    w->fd = fd;
    void *m = mmap(NULL, RESERVE, PROT_READ | PROT_WRITE, MAP_SHARED | MAP_NORESERVE, fd, 0);
    if (m != MAP_FAILED) { w->fixed = 1; w->map_len = RESERVE; }
    else {
        m = mmap(NULL, init_len, PROT_READ | PROT_WRITE, MAP_SHARED, fd, 0);
        if (m == MAP_FAILED) { close(fd); free(w); return NULL; }
        w->map_len = init_len;
    }
    w->map = (uint8_t *)m;
    w->file_len = init_len;
    memset(w->map, 0, init_len);
    wr32(w->map + FH_MAGIC, VC_MAGIC); wr32(w->map + FH_VER, 1); wr32(w->map + FH_ATOM, VC_A);
    wr32(w->map + FH_LAYOUT, 1); wr32(w->map + FH_RGN, R_ATOMS); wr32(w->map + FH_NLOD, (uint32_t)nlod);
    wr32(w->map + FH_NX, d0.nx); wr32(w->map + FH_NY, d0.ny); wr32(w->map + FH_NZ, d0.nz);
    wrf(w->map + FH_TRATIO, target_ratio);
    wr64(w->map + FH_L1CAP, cap); wr64(w->map + FH_L1CNT, 0); wr64(w->map + FH_L1OFF, FILE_HDR);
    wr64(w->map + FH_CURSOR, init_len);
    w->cursor = init_len;
    w->dims0 = d0; w->nlod = nlod;
    pthread_mutex_init(&w->mu, NULL);
    return w;
}

vc_status vc_set_base_q(vc_writer *w, int lod, float q) {
    if (!w || lod < 0 || lod >= w->nlod) return VC_ERR_RANGE;
    q = vc__clamp_q(q);
    pthread_mutex_lock(&w->mu);
    wrf(w->map + FH_BASEQ + 4 * lod, q);
    w->base_q[lod] = q;
    pthread_mutex_unlock(&w->mu);
    return VC_OK;
}

/* find (create) a region's L2 block; 0 on failure / absent, 1 = ZERO_REGION */
static uint64_t region_block(vc_writer *w, uint64_t key, int create) {
    uint64_t cap = rd64(w->map + FH_L1CAP), off = rd64(w->map + FH_L1OFF), mask = cap - 1;
    uint64_t i = vc__mix64(key) & mask;
    for (uint64_t probe = 0; probe < cap; ++probe, i = (i + 1) & mask) {
        uint8_t *e = w->map + off + i * 16;
        uint64_t br = rd64(e + 8);
        if (br != 0) { if (rd64(e) == key) return br; continue; }
        if (!create) return 0;
        uint64_t nb = alloc_bytes(w, L2_BLOCK);
        if (!nb) return 0;
        e = w->map + off + i * 16;
        memset(w->map + nb, 0, L2_BLOCK);
        wr64(e, key); wr64(e + 8, nb);
        wr64(w->map + FH_L1CNT, rd64(w->map + FH_L1CNT) + 1);
        return nb;
    }
    return 0;
}

static uint64_t slot_index(uint32_t az, uint32_t ay, uint32_t ax) {
    return ((uint64_t)(az & 31) * 32 + (ay & 31)) * 32 + (ax & 31);
}

/* store one atom whose payload (len > 0) or zero-ness (len == 0) is known;
 * caller holds the mutex.  src/vc/vc.c:1082-1159 */
static vc_status store_atom(vc_writer *w, int lod, uint32_t az, uint32_t ay, uint32_t ax,
                            const uint8_t *pay, uint32_t len) {
    uint64_t key = vc__region_key(lod, az / R_ATOMS, ay / R_ATOMS, ax / R_ATOMS);
    uint64_t blk = region_block(w, key, 1);
    if (blk == ZERO_REGION) {          /* promote a known-zero region to a real block */
        uint64_t cap = rd64(w->map + FH_L1CAP), off = rd64(w->map + FH_L1OFF), mask = cap - 1;
        uint64_t i = vc__mix64(key) & mask;
        while (rd64(w->map + off + i * 16) != key || rd64(w->map + off + i * 16 + 8) == 0) i = (i + 1) & mask;
        uint64_t nb = alloc_bytes(w, L2_BLOCK);
        if (!nb) return VC_ERR_OOM;
        memset(w->map + nb, 0, L2_BLOCK);
        wr64(w->map + off + i * 16 + 8, nb);
        blk = nb;
    }
    if (blk == 0) return VC_ERR_OOM;
    uint64_t so = blk + slot_index(az, ay, ax) * ATOM_SLOT;
    if (len == 0) {
        uint8_t *s = w->map + so;
        wr64(s, SLOT_ABSENT); wr32(s + 8, 0); s[13] = 0; s[12] = AF_ZERO;
        return VC_OK;
    }
    uint64_t po = alloc_bytes(w, len);
    if (!po) return VC_ERR_OOM;
    memcpy(w->map + po, pay, len);
    uint8_t *s = w->map + so;
    wr32(s + 8, len); s[12] = AF_PRESENT; s[13] = pay[0];
    wr64(s, po);
    return VC_OK;
}

static int all_zero(const uint8_t *v) {
    for (uint32_t i = 0; i < VC_A3; ++i) if (v[i]) return 0;
    return 1;
}

/* encode n atoms (all non-zero) at q on the GPU; payload i at pay + i*VCG_PAY_CAP */
static vc_status encode_batch(size_t n, const uint8_t *vox, float q, uint8_t *pay, uint32_t *lens) {
    vc__require_gpu();
    if (vcg_encode_atoms(n, vox, q, pay, VCG_PAY_CAP, lens)) vc__panic("range encoder overflow");
    return VC_OK;
}

vc_status vc_append_atom(vc_writer *w, int lod, uint32_t az, uint32_t ay, uint32_t ax, const uint8_t vox[VC_ATOM3]) {
    if (!w || lod < 0 || lod >= w->nlod || !vox) return VC_ERR_RANGE;
    vc_dims d = vc__lod_dims(w->dims0, lod);
    if (ax >= (d.nx + 31) / 32 || ay >= (d.ny + 31) / 32 || az >= (d.nz + 31) / 32) return VC_ERR_RANGE;
    uint8_t *pay = NULL; uint32_t len = 0;
    if (!all_zero(vox)) {
        pthread_mutex_lock(&w->mu);
        float q = w->base_q[lod];
        pthread_mutex_unlock(&w->mu);
        if (q <= 0.0f) return VC_ERR_RANGE;
        pay = (uint8_t *)malloc(VCG_PAY_CAP);
        if (!pay) vc__panic("out of memory");
        encode_batch(1, vox, q, pay, &len);
    }
    pthread_mutex_lock(&w->mu);
    vc_status s = store_atom(w, lod, az, ay, ax, pay, len);
    pthread_mutex_unlock(&w->mu);
    free(pay);
    return s;
}

vc_status vc_append_box(vc_writer *w, int lod, vc_box b, const uint8_t *voxels) {
    if (!w || lod < 0 || lod >= w->nlod || !voxels) return VC_ERR_RANGE;       /* src/vc/vc.c:1170-1189 */
    if (b.x0 % 32 || b.y0 % 32 || b.z0 % 32) return VC_ERR_RANGE;
    if ((b.x1 - b.x0) % 32 || (b.y1 - b.y0) % 32 || (b.z1 - b.z0) % 32) return VC_ERR_RANGE;
    if (b.x1 <= b.x0 || b.y1 <= b.y0 || b.z1 <= b.z0) return VC_ERR_RANGE;
    uint32_t bx = b.x1 - b.x0, by = b.y1 - b.y0;
    size_t n = (size_t)(bx / 32) * (by / 32) * ((b.z1 - b.z0) / 32);
    uint8_t *atoms = (uint8_t *)malloc(n * (size_t)VC_A3);
    uint8_t *zero = (uint8_t *)calloc(n, 1);
    if (!atoms || !zero) vc__panic("out of memory");
    size_t k = 0;
    for (uint32_t vz = b.z0; vz < b.z1; vz += 32)
    for (uint32_t vy = b.y0; vy < b.y1; vy += 32)
    for (uint32_t vx = b.x0; vx < b.x1; vx += 32, ++k) {
        uint8_t *a = atoms + k * (size_t)VC_A3;
        for (uint32_t z = 0; z < 32; ++z) for (uint32_t y = 0; y < 32; ++y)
            memcpy(a + (z * 32 + y) * 32, voxels + ((uint64_t)(vz - b.z0 + z) * by + (vy - b.y0 + y)) * bx + (vx - b.x0), 32);
        zero[k] = (uint8_t)all_zero(a);
    }
    pthread_mutex_lock(&w->mu);
    float q = w->base_q[lod];
    pthread_mutex_unlock(&w->mu);
    /* encode every non-zero atom in one GPU batch (only if q is set; without q
     * the zero atoms before the first non-zero one are still stored) */
    size_t nnz = 0;
    for (k = 0; k < n; ++k) nnz += !zero[k];
    uint8_t *pay = NULL; uint32_t *lens = NULL; size_t *slot = NULL;
    if (nnz && q > 0.0f) {
        uint8_t *src = (uint8_t *)malloc(nnz * (size_t)VC_A3);
        pay = (uint8_t *)malloc(nnz * (size_t)VCG_PAY_CAP);
        lens = (uint32_t *)calloc(nnz, 4);
        slot = (size_t *)calloc(n, sizeof(size_t));
        if (!src || !pay || !lens || !slot) vc__panic("out of memory");
        size_t j = 0;
        for (k = 0; k < n; ++k) if (!zero[k]) { memcpy(src + j * (size_t)VC_A3, atoms + k * (size_t)VC_A3, VC_A3); slot[k] = j++; }
        encode_batch(nnz, src, q, pay, lens);
        free(src);
    }
    vc_status s = VC_OK;
    k = 0;
    pthread_mutex_lock(&w->mu);
    for (uint32_t vz = b.z0; vz < b.z1 && s == VC_OK; vz += 32)
    for (uint32_t vy = b.y0; vy < b.y1 && s == VC_OK; vy += 32)
    for (uint32_t vx = b.x0; vx < b.x1; vx += 32, ++k) {
        if (!zero[k] && !(q > 0.0f)) { s = VC_ERR_RANGE; break; }
        s = zero[k] ? store_atom(w, lod, vz / 32, vy / 32, vx / 32, NULL, 0)
                    : store_atom(w, lod, vz / 32, vy / 32, vx / 32, pay + slot[k] * (size_t)VCG_PAY_CAP, lens[slot[k]]);
        if (s != VC_OK) break;
    }
    pthread_mutex_unlock(&w->mu);
    free(atoms); free(zero); free(pay); free(lens); free(slot);
    return s;
}

vc_status vc_mark_zero_atom(vc_writer *w, int lod, uint32_t az, uint32_t ay, uint32_t ax) {
    if (!w || lod < 0 || lod >= w->nlod) return VC_ERR_RANGE;      /* src/vc/vc.c:1191-1205 */
    pthread_mutex_lock(&w->mu);
    uint64_t blk = region_block(w, vc__region_key(lod, az / R_ATOMS, ay / R_ATOMS, ax / R_ATOMS), 1);
    vc_status s = VC_OK;
    if (blk == 0) s = VC_ERR_OOM;
    else if (blk != ZERO_REGION) {
        uint8_t *sl = w->map + blk + slot_index(az, ay, ax) * ATOM_SLOT;
        wr64(sl, SLOT_ABSENT); wr32(sl + 8, 0); sl[13] = 0; sl[12] = AF_ZERO;
    }
    pthread_mutex_unlock(&w->mu);
    return s;
}

vc_status vc_mark_zero_region(vc_writer *w, int lod, uint32_t rz, uint32_t ry, uint32_t rx) {
    if (!w || lod < 0 || lod >= w->nlod) return VC_ERR_RANGE;      /* src/vc/vc.c:1207-1229 */
    pthread_mutex_lock(&w->mu);
    uint64_t key = vc__region_key(lod, rz, ry, rx);
    uint64_t cap = rd64(w->map + FH_L1CAP), off = rd64(w->map + FH_L1OFF), mask = cap - 1;
    uint64_t i = vc__mix64(key) & mask;
    for (uint64_t probe = 0; probe < cap; ++probe, i = (i + 1) & mask) {
        uint8_t *e = w->map + off + i * 16;
        if (rd64(e + 8) != 0) { if (rd64(e) == key) break; continue; }
        wr64(e, key); wr64(e + 8, ZERO_REGION);
        wr64(w->map + FH_L1CNT, rd64(w->map + FH_L1CNT) + 1);
        break;
    }
    pthread_mutex_unlock(&w->mu);
    return VC_OK;
}

/* reader view over the live file (src/vc/vc.c:1475-1482); caller holds the mutex */
static void view(vc_writer *w, vc_archive *a) {
    memset(a, 0, sizeof *a);
    a->buf = w->map; a->len = w->cursor; a->dims0 = w->dims0; a->nlod = w->nlod; a->version = 1;
    a->l1_off = rd64(w->map + FH_L1OFF); a->l1_cap = rd64(w->map + FH_L1CAP);
    for (int l = 0; l < VC_NLOD; ++l) a->base_q[l] = w->base_q[l];
}

vc_cover vc_writer_coverage(vc_writer *w, int lod, uint32_t az, uint32_t ay, uint32_t ax) {
    if (!w || lod < 0 || lod >= w->nlod) return VC_ABSENT;
    pthread_mutex_lock(&w->mu);
    vc_archive a; view(w, &a);
    vc_cover c = vc_atom_coverage(&a, lod, az, ay, ax);
    pthread_mutex_unlock(&w->mu);
    return c;
}

vc_status vc_writer_decode_atom(vc_writer *w, int lod, uint32_t az, uint32_t ay, uint32_t ax, uint8_t out[VC_ATOM3]) {
    if (!w || lod < 0 || lod >= w->nlod || !out) return VC_ERR_RANGE;
    vc_dims d = vc__lod_dims(w->dims0, lod);
    if (ax >= (d.nx + 31) / 32 || ay >= (d.ny + 31) / 32 || az >= (d.nz + 31) / 32) return VC_ERR_RANGE;
    pthread_mutex_lock(&w->mu);
    vc_archive a; view(w, &a);
    int32_t q[4] = {lod, (int32_t)ax, (int32_t)ay, (int32_t)az};
    vc_status s = vc__decode_atoms_host(&a, 1, q, out, NULL);
    pthread_mutex_unlock(&w->mu);
    return s;
}

vc_status vc_writer_decode_2x2x2(vc_writer *w, int lod, uint32_t az0, uint32_t ay0, uint32_t ax0,
                                 uint8_t out8[8][VC_ATOM3], int *any) {
    if (!w || lod < 0 || lod >= w->nlod || !out8) return VC_ERR_RANGE;   /* src/vc/vc.c:1511-1533 */
    vc_dims d = vc__lod_dims(w->dims0, lod);
    uint32_t acx = (d.nx + 31) / 32, acy = (d.ny + 31) / 32, acz = (d.nz + 31) / 32;
    int32_t q[32]; int idx[8]; int n = 0, a_any = 0;
    pthread_mutex_lock(&w->mu);
    vc_archive a; view(w, &a);
    for (int c = 0; c < 8; ++c) {
        uint32_t az = az0 + (c >> 2), ay = ay0 + ((c >> 1) & 1), ax = ax0 + (c & 1);
        memset(out8[c], 0, VC_A3);
        if (az >= acz || ay >= acy || ax >= acx) continue;
        uint64_t off = 0; uint32_t len = 0;
        if (vc__resolve(&a, lod, az, ay, ax, &off, &len) != AF_PRESENT) continue;
        if (off == SLOT_ABSENT || len < 1 || off > a.len || len > a.len - off) continue;
        q[4 * n] = lod; q[4 * n + 1] = (int32_t)ax; q[4 * n + 2] = (int32_t)ay; q[4 * n + 3] = (int32_t)az;
        idx[n++] = c;
        a_any = 1;
    }
    if (n) {
        uint8_t *tmp = (uint8_t *)malloc((size_t)n * VC_A3);
        if (!tmp) vc__panic("out of memory");
        vc__decode_atoms_host(&a, (size_t)n, q, tmp, NULL);
        for (int k = 0; k < n; ++k) memcpy(out8[idx[k]], tmp + (size_t)k * VC_A3, VC_A3);
        free(tmp);
    }
    pthread_mutex_unlock(&w->mu);
    if (any) *any = a_any;
    return VC_OK;
}

void vc_writer_close(vc_writer *w) {
    if (!w) return;
    pthread_mutex_lock(&w->mu);
    wr64(w->map + FH_CURSOR, w->cursor);
    msync(w->map, w->file_len, MS_SYNC);
    munmap(w->map, w->map_len);
    if (w->cursor && w->cursor < w->file_len) { if (ftruncate(w->fd, (off_t)w->cursor) != 0) { /* keep */ } }
    close(w->fd);
    pthread_mutex_unlock(&w->mu);
    pthread_mutex_destroy(&w->mu);
    free(w);
}

/* ---- internal: the exporter cascade (vc_export.c) ---- */
/* store one atom whose payload was encoded elsewhere (len > 0) */
vc_status vc__writer_store(vc_writer *w, int lod, uint32_t az, uint32_t ay, uint32_t ax, const uint8_t *pay,
                           uint32_t len) {
    if (!w || lod < 0 || lod >= w->nlod || !pay || len == 0) return VC_ERR_RANGE;
    pthread_mutex_lock(&w->mu);
    vc_status s = store_atom(w, lod, az, ay, ax, pay, len);
    pthread_mutex_unlock(&w->mu);
    return s;
}

/* decode n atoms (quads) from the live file, as vc_writer_decode_atom does
 * one by one: absent and zero atoms read as zeros */
vc_status vc__writer_decode_atoms(vc_writer *w, size_t n, const int32_t *quads, uint8_t *out) {
    if (!w) return VC_ERR_RANGE;
    pthread_mutex_lock(&w->mu);
    vc_archive a; view(w, &a);
    vc_status s = vc__decode_atoms_host(&a, n, quads, out, NULL);
    pthread_mutex_unlock(&w->mu);
    return s;
}

int vc__writer_nlod(const vc_writer *w) { return w ? w->nlod : 0; }
float vc__writer_q(vc_writer *w, int lod) {
    pthread_mutex_lock(&w->mu);
    const float q = w->base_q[lod];
    pthread_mutex_unlock(&w->mu);
    return q;
}
