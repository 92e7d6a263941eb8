/* vc_repack.c — rewrite an opened archive (v1 or v2) into the canonical v2
 * container, the bytes tools/vc_repack.c writes for the same input.
 *
 * Regions are taken coarse-LOD-first (tools/vc_repack.c:226-232).  The host
 * resolves each region's slot array (the v1 L1 probe or the v2 directory) and
 * the device sizes every slot array (k_repack_size).  A region with a present
 * atom becomes a data blob [512 KiB slot array][payloads in slot order], one
 * with only zero atoms (or a v1 zero region) a ZERO entry, and the rest ABSENT
 * (tools/vc_repack.c:96-121).  A serial prefix sum in the same order places
 * the blobs after the 4 KiB-aligned directory (tools/vc_repack.c:244-256), and
 * the device writes the slot arrays and copies the payloads verbatim
 * (tools/vc_repack.c:131-162).  Payload ranges are checked against the
 * archive (the tool copies them unchecked): one outside it is VC_ERR_FORMAT.
 */
#include <stdlib.h>
#include <string.h>

#include "vc_host.h"

#define V2_DIR_ENTRY 16u
#define V2_RGN_ABSENT 0u
#define V2_RGN_ZERO 1u
#define V2_RGN_DATA 2u

typedef struct {
    uint64_t blobs_off, total_len, njobs;
    uint8_t *head;                       /* header, directory, zero pad: blobs_off bytes */
    size_t ndata;
    uint64_t *blk, *boff;                /* per data region, coarse-first order */
    uint64_t *cdst, *cjob;               /* per plane of a data region: first payload, first job */
} rp_plan;

#define PLANES 32u                       /* 1024-slot planes per slot array */

static void plan_free(rp_plan *p) {
    free(p->head); free(p->blk); free(p->boff); free(p->cdst); free(p->cjob);
    memset(p, 0, sizeof *p);
}

static vc_status plan_build(vc_archive *a, vcg_arc *d, rp_plan *p) {
    memset(p, 0, sizeof *p);
    const int nlod = a->nlod;
    uint64_t prefix[VC_NLOD + 1], nreg = 0;          /* directory index base per LOD (v1 has none) */
    for (int l = 0; l < nlod; ++l) {
        uint32_t nrz, nry, nrx;
        vc__region_grid(a->dims0, l, &nrz, &nry, &nrx);
        prefix[l] = nreg;
        nreg += (uint64_t)nrz * nry * nrx;
    }
    prefix[nlod] = nreg;
    const uint64_t dir_off = FILE_HDR;
    p->blobs_off = (dir_off + nreg * V2_DIR_ENTRY + 4095u) & ~4095ull;
    /* the coarse-first region order, and each region's slot array */
    uint64_t *order = (uint64_t *)malloc((nreg ? nreg : 1) * 8);
    uint8_t *kind = (uint8_t *)malloc(nreg ? nreg : 1);
    uint64_t *rblk = (uint64_t *)malloc((nreg ? nreg : 1) * 8);
    uint64_t *bl = (uint64_t *)malloc((nreg ? nreg : 1) * 8);      /* slot arrays to size */
    uint64_t *sums = (uint64_t *)malloc((nreg ? nreg : 1) * PLANES * 24);
    p->head = (uint8_t *)calloc(1, p->blobs_off);
    p->blk = (uint64_t *)malloc((nreg ? nreg : 1) * 8);
    p->boff = (uint64_t *)malloc((nreg ? nreg : 1) * 8);
    p->cdst = (uint64_t *)malloc((nreg ? nreg : 1) * PLANES * 8);
    p->cjob = (uint64_t *)malloc((nreg ? nreg : 1) * PLANES * 8);
    vc_status st = VC_OK;
    if (!order || !kind || !rblk || !bl || !sums || !p->head || !p->blk || !p->boff || !p->cdst || !p->cjob) {
        st = VC_ERR_OOM;
        goto done;
    }
    uint64_t ni = 0, nb = 0;
    for (int l = nlod - 1; l >= 0; --l) {
        uint32_t nrz, nry, nrx;
        vc__region_grid(a->dims0, l, &nrz, &nry, &nrx);
        for (uint32_t rz = 0; rz < nrz; ++rz)
            for (uint32_t ry = 0; ry < nry; ++ry)
                for (uint32_t rx = 0; rx < nrx; ++rx) {
                    const uint64_t ri = prefix[l] + ((uint64_t)rz * nry + ry) * nrx + rx;
                    order[ni++] = ri;
                    kind[ri] = (uint8_t)vc__region_block(a, l, rz, ry, rx, &rblk[ri]);
                    if (kind[ri] == AF_PRESENT) bl[nb++] = rblk[ri];
                }
    }
    const int rc = vcg_repack_size(d, nb, bl, sums);
    if (rc != 0) { st = rc < 0 ? VC_ERR_OOM : VC_ERR_FORMAT; goto done; }
    /* classify + place, coarse-first */
    uint8_t *dir = p->head + dir_off;
    uint64_t cursor = p->blobs_off, jcur = 0, b = 0;
    for (uint64_t i = 0; i < ni; ++i) {
        const uint64_t ri = order[i];
        uint8_t *e = dir + ri * V2_DIR_ENTRY;
        uint32_t flags = kind[ri] == AF_ZERO ? V2_RGN_ZERO : V2_RGN_ABSENT;
        if (kind[ri] == AF_PRESENT) {
            const uint64_t *sp = sums + 3 * PLANES * b;
            uint64_t bytes = 0, np = 0, nz = 0;
            for (uint32_t z = 0; z < PLANES; ++z) { bytes += sp[3 * z]; np += sp[3 * z + 1]; nz += sp[3 * z + 2]; }
            b++;
            if (np) {
                const uint64_t blen = L2_BLOCK + bytes;
                /* the directory keeps a blob's length in 32 bits (the tool
                 * truncates it silently; vc_encode rejects the same case) */
                if (blen > 0xFFFFFFFFull) { st = VC_ERR_RANGE; goto done; }
                wr64(e, cursor);
                wr32(e + 8, (uint32_t)blen);
                flags = V2_RGN_DATA;
                /* the planes' payloads follow one another in slot order */
                uint64_t pc = cursor + L2_BLOCK, jc = jcur;
                for (uint32_t z = 0; z < PLANES; ++z) {
                    p->cdst[PLANES * p->ndata + z] = pc; p->cjob[PLANES * p->ndata + z] = jc;
                    pc += sp[3 * z]; jc += sp[3 * z + 1];
                }
                p->blk[p->ndata] = rblk[ri]; p->boff[p->ndata] = cursor;
                p->ndata++;
                cursor += blen;
                jcur += np;
            } else if (nz) {
                flags = V2_RGN_ZERO;
            }
        }
        wr32(e + 12, flags);
    }
    p->total_len = cursor;
    p->njobs = jcur;
    /* header (tools/vc_repack.c:272-281): dims, LOD count, target ratio and
     * base_q carried over; every other field zero */
    uint8_t *h = p->head;
    wr32(h + FH_MAGIC, VC_MAGIC); wr32(h + FH_VER, 2u);
    wr32(h + FH_ATOM, VC_A); wr32(h + FH_RGN, R_ATOMS); wr32(h + FH_NLOD, (uint32_t)nlod);
    memcpy(h + FH_NX, a->buf + FH_NX, 12);
    memcpy(h + FH_TRATIO, a->buf + FH_TRATIO, 4);
    memcpy(h + FH_BASEQ, a->buf + FH_BASEQ, VC_NLOD * 4);
    wr64(h + FH2_DIROFF, dir_off); wr64(h + FH2_DIRCNT, nreg); wr64(h + FH2_TOTLEN, p->total_len);
done:
    free(order); free(kind); free(rblk); free(bl); free(sums);
    if (st != VC_OK) plan_free(p);
    return st;
}

static vc_status plan_write(vc_archive *a, vcg_arc *d, const rp_plan *p, uint8_t *dev_out) {
    if (vcg_copy(dev_out, p->head, p->blobs_off, 1) != 0) return VC_ERR_IO;
    if (vcg_repack_write(d, p->ndata, p->blk, p->boff, p->cdst, p->cjob, p->njobs, dev_out) != 0) return VC_ERR_OOM;
    (void)a;
    return VC_OK;
}

static vc_status repack_begin(vc_archive *a, vcg_arc **d, rp_plan *p) {
    if (!a || !a->buf) return VC_ERR_RANGE;          /* an archive opened from memory */
    vc__require_gpu();
    *d = vc__arc_dev(a);
    if (!*d) return VC_ERR_OOM;
    return plan_build(a, *d, p);
}

vc_status vcx_repack_device(vc_archive *a, void *dev_out, size_t cap, size_t *out_len) {
    if (out_len) *out_len = 0;
    vcg_arc *d;
    rp_plan p;
    vc_status st = repack_begin(a, &d, &p);
    if (st != VC_OK) return st;
    if (out_len) *out_len = (size_t)p.total_len;
    if (!dev_out || cap < p.total_len) { plan_free(&p); return VC_ERR_RANGE; }
    st = plan_write(a, d, &p, (uint8_t *)dev_out);
    plan_free(&p);
    return st;
}

vc_status vcx_repack(vc_archive *a, uint8_t **out, size_t *out_len) {
    if (!out || !out_len) return VC_ERR_RANGE;
    *out = NULL; *out_len = 0;
    vcg_arc *d;
    rp_plan p;
    vc_status st = repack_begin(a, &d, &p);
    if (st != VC_OK) return st;
    uint8_t *dev = (uint8_t *)vcg_alloc(p.total_len);
    uint8_t *host = (uint8_t *)malloc(p.total_len);
    if (!dev || !host) { st = VC_ERR_OOM; goto done; }
    st = plan_write(a, d, &p, dev);
    if (st == VC_OK && vcg_copy(host, dev, p.total_len, 2) != 0) st = VC_ERR_IO;
    if (st == VC_OK) { *out = host; *out_len = (size_t)p.total_len; host = NULL; }
done:
    vcg_free(dev);
    free(host);
    plan_free(&p);
    return st;
}
