/* vc_export.c — the exporter's pyramid cascade for a local volume, with the
 * pyramid, the transforms and the range coding on the GPU.
 *
 * Restates tools/vc_export_stream.c for a source volume held in memory whose
 * every source chunk is present (no S3, no zarr parsing, no occupancy map):
 *   - the archive's dims are the source dims rounded up to whole index
 *     regions (1024 voxels; zero padding), nlod by vc_create's rule
 *     (:690-701);
 *   - base_q[l] = q0 * falloff^l by iterated float multiply, raised to 0.05
 *     before each use (:711-719);
 *   - prepass: every region whose LOD0 extent starts past the source is
 *     marked zero (:778-790, region_definite_zero :338-352);
 *   - units (xy tile of SB voxels, z band of BAND = 32 << fine voxels,
 *     fine = min(nlod-1, 2)) in unit order (bz fastest, then x tiles, then y
 *     tiles, :122-124): a band whose footprint starts past the source's chunk
 *     grid stamps all of its LOD0..fine atoms zero, level by level (:505-522);
 *     any other band runs the plane cascade (push_plane / flush_slab,
 *     :395-446), which appends each z-slab of atoms of a level, y-major, as the
 *     slab fills: all-zero atoms marked zero, the rest encoded (:397-416);
 *   - the coarse tail: each level above fine from the archive's own decoded
 *     parents (2x2x2 atoms of the level below, :603-626), downscaled and
 *     appended in raster order.
 * That is the order the tool's single-threaded pipeline produces, so the file
 * equals the reference writer's driven in that order.  The within-cell
 * downscale (tools/downscale.h) of the whole padded level equals the tool's
 * per-plane-pair cascade because every band and tile is a whole number of
 * cells.  Device side: the padded LOD0, LOD1..fine (vcg_ds2x), the atom
 * encode of every level (k_fdct with the pass-1 MFMA, k_rc_encode) and the
 * tail's decode and downscale; host side: the order, and the writes into the
 * v1 file.
 */
#include <math.h>
#include <stdlib.h>

#include "vc_host.h"

#define REGV 1024u              /* region edge in voxels (32 atoms) */

/* the (level, z-slab) flush order of one full band (push_plane's counters
 * over BAND LOD0 planes); returns the number of flushes */
typedef struct { int lod; uint32_t zslab; } flush_ev;
static void cascade_push(int L, int fine, int *slabfill, int *pairfill, uint32_t *zslab, flush_ev *ev, int *n) {
    if (++slabfill[L] == (int)VC_A) { ev[(*n)++] = (flush_ev){L, zslab[L]++}; slabfill[L] = 0; }
    if (L + 1 <= fine && ++pairfill[L] == 2) {
        cascade_push(L + 1, fine, slabfill, pairfill, zslab, ev, n);
        pairfill[L] = 0;
    }
}
static int band_flushes(int fine, uint32_t band, flush_ev *ev) {
    int slabfill[VC_NLOD] = {0}, pairfill[VC_NLOD] = {0}, n = 0;
    uint32_t zslab[VC_NLOD] = {0};
    for (uint32_t p = 0; p < band; ++p) cascade_push(0, fine, slabfill, pairfill, zslab, ev, &n);
    return n;
}

/* encode every atom of the nl device levels at qs; lens[i] (0 = all-zero
 * atom) and the packed payloads (atom i at off[i]) in host memory */
static vc_status encode_levels(const uint8_t *const *lods, const uint32_t (*dims)[3], int nl, const float *qs,
                               uint32_t **lens_out, uint64_t **off_out, uint8_t **pay_out, uint64_t *total_out) {
    vcg_enc *e = vcg_enc_begin_levels(lods, dims, nl);
    if (!e) return VC_ERR_RANGE;
    uint64_t total = 0;
    for (int l = 0; l < nl; ++l)
        total += (uint64_t)((dims[l][0] + 31) / 32) * ((dims[l][1] + 31) / 32) * ((dims[l][2] + 31) / 32);
    uint32_t *lens = (uint32_t *)calloc(total, 4);
    uint64_t *off = (uint64_t *)calloc(total, 8), *slot = (uint64_t *)calloc(total, 8);
    if (!lens || !off || !slot) vc__panic("out of memory");
    float q8[VCG_NLOD] = {0};
    for (int l = 0; l < nl; ++l) q8[l] = qs[l];
    int ovf = 0;
    vcg_enc_all(1, &e, (const float (*)[VCG_NLOD])q8, &lens, &ovf);
    if (ovf) { vcg_enc_end(e); free(lens); free(off); free(slot); return VC_ERR_RANGE; }
    uint64_t acc = 0;
    for (uint64_t i = 0; i < total; ++i) { off[i] = acc; acc += lens[i]; }
    uint8_t *pay = (uint8_t *)malloc(acc ? acc : 1);
    if (!pay) vc__panic("out of memory");
    uint8_t none = 0;
    if (acc) vcg_enc_assemble(e, &none, 0, off, slot, lens, total, pay, acc);
    vcg_enc_end(e);
    free(slot);
    *lens_out = lens; *off_out = off; *pay_out = pay; *total_out = total;
    return VC_OK;
}

static vc_status put_atom(vc_writer *w, int lod, uint32_t az, uint32_t ay, uint32_t ax, const uint8_t *pay,
                          uint32_t len) {
    return len ? vc__writer_store(w, lod, az, ay, ax, pay, len) : vc_mark_zero_atom(w, lod, az, ay, ax);
}

vc_status vcx_export_volume(const uint8_t *vol, vc_dims src, const uint32_t chunk[3],
                            const char *path, float q0, float q_falloff, int method, float alpha, uint32_t tile) {
    if (!vol || !path || !vc__dims_valid(src) || (method != VCX_DS_BOX && method != VCX_DS_CBOX) || isnan(alpha) ||
        !(q0 > 0.0f) || !(q_falloff > 0.0f))
        return VC_ERR_RANGE;
    const uint32_t cx = chunk ? chunk[0] : 128, cy = chunk ? chunk[1] : 128, cz = chunk ? chunk[2] : 128;
    if (!cx || !cy || !cz) return VC_ERR_RANGE;
    vc__require_gpu();
    const vc_dims pd = {(src.nx + REGV - 1) / REGV * REGV, (src.ny + REGV - 1) / REGV * REGV,
                        (src.nz + REGV - 1) / REGV * REGV};
    if (!vc__dims_valid(pd)) return VC_ERR_RANGE;
    vc_writer *w = vc_create(path, pd, 1.0f);
    if (!w) return VC_ERR_IO;
    const int nlod = vc__writer_nlod(w);
    float qs[VC_NLOD] = {0};
    {
        float q = q0;
        for (int l = 0; l < nlod; ++l) {
            if (q < 0.05f) q = 0.05f;
            vc_set_base_q(w, l, q);
            qs[l] = vc__writer_q(w, l);
            q *= q_falloff;
        }
    }
    /* prepass: regions whose LOD0 extent starts past the source */
    for (int L = 0; L < nlod; ++L) {
        uint32_t nrz, nry, nrx;
        vc__region_grid(pd, L, &nrz, &nry, &nrx);
        for (uint32_t rz = 0; rz < nrz; ++rz)
            for (uint32_t ry = 0; ry < nry; ++ry)
                for (uint32_t rx = 0; rx < nrx; ++rx)
                    if (((uint64_t)rx * REGV << L) >= src.nx || ((uint64_t)ry * REGV << L) >= src.ny ||
                        ((uint64_t)rz * REGV << L) >= src.nz)
                        vc_mark_zero_region(w, L, rz, ry, rx);
    }
    const int fine = nlod - 1 < 2 ? nlod - 1 : 2;
    const uint32_t BAND = VC_A << fine, SBmax = VC_A << (nlod - 1);
    uint32_t SB = tile ? tile : 1024;
    if (SB > SBmax) SB = SBmax;
    SB = (SB + BAND - 1) / BAND * BAND;

    /* LOD0 .. fine on the device: the padded volume and its within-cell downscales */
    uint8_t *lv[VC_NLOD] = {0};
    uint32_t ld[VC_NLOD][3];
    vc_status st = VC_OK;
    for (int l = 0; l <= fine; ++l) {
        const vc_dims d = vc__lod_dims(pd, l);
        ld[l][0] = d.nx; ld[l][1] = d.ny; ld[l][2] = d.nz;
        lv[l] = (uint8_t *)vcg_alloc((size_t)d.nx * d.ny * d.nz);
        if (!lv[l]) { st = VC_ERR_OOM; goto done; }
    }
    {
        const uint32_t sd[3] = {src.nx, src.ny, src.nz};   /* host or device memory, by pointer */
        if (vcg_pad_volume(vol, sd, lv[0], ld[0]) != 0) { st = VC_ERR_RANGE; goto done; }
    }
    for (int l = 1; l <= fine; ++l)
        if (vcg_ds2x(lv[l - 1], ld[l - 1], method, alpha, lv[l], 1) != 0) { st = VC_ERR_RANGE; goto done; }
    {
        uint32_t *lens; uint64_t *off, total; uint8_t *pay;
        st = encode_levels((const uint8_t *const *)lv, (const uint32_t (*)[3])ld, fine + 1, qs, &lens, &off, &pay, &total);
        if (st != VC_OK) goto done;
        uint64_t prefix[VC_NLOD + 1] = {0};
        for (int l = 0; l <= fine; ++l)
            prefix[l + 1] = prefix[l] + (uint64_t)((ld[l][0] + 31) / 32) * ((ld[l][1] + 31) / 32) * ((ld[l][2] + 31) / 32);
        flush_ev ev[64];
        const int nev = band_flushes(fine, BAND, ev);
        const uint32_t ntx = (pd.nx + SB - 1) / SB, nty = (pd.ny + SB - 1) / SB, nbz = (pd.nz + BAND - 1) / BAND;
        const uint32_t ncx = (src.nx + cx - 1) / cx, ncy = (src.ny + cy - 1) / cy, ncz = (src.nz + cz - 1) / cz;
        const uint64_t nbands = (uint64_t)ntx * nty * nbz;
        for (uint64_t ti = 0; ti < nbands && st == VC_OK; ++ti) {
            const uint32_t bz = (uint32_t)(ti % nbz), txy = (uint32_t)(ti / nbz), ty = txy / ntx, tx = txy % ntx;
            const uint32_t vx0 = tx * SB, vy0 = ty * SB, vz0 = bz * BAND;
            const uint32_t vx1 = vx0 + SB < pd.nx ? vx0 + SB : pd.nx, vy1 = vy0 + SB < pd.ny ? vy0 + SB : pd.ny;
            const uint32_t vz1 = vz0 + BAND < pd.nz ? vz0 + BAND : pd.nz;
            const int empty = vx0 / cx >= ncx || vy0 / cy >= ncy || vz0 / cz >= ncz;
            if (empty) {                             /* stamp every owned atom zero */
                for (int L = 0; L <= fine && st == VC_OK; ++L) {
                    const uint32_t az0 = (vz0 >> L) / VC_A, ay0 = (vy0 >> L) / VC_A, ax0 = (vx0 >> L) / VC_A;
                    const uint32_t nz = ((vz1 - vz0) >> L) / VC_A, nyc = ((vy1 - vy0) >> L) / VC_A,
                                   nxc = ((vx1 - vx0) >> L) / VC_A;
                    for (uint32_t az = 0; az < nz; ++az)
                        for (uint32_t ay = 0; ay < nyc; ++ay)
                            for (uint32_t ax = 0; ax < nxc && st == VC_OK; ++ax)
                                st = vc_mark_zero_atom(w, L, az0 + az, ay0 + ay, ax0 + ax);
                }
                continue;
            }
            for (int k = 0; k < nev && st == VC_OK; ++k) {
                const int L = ev[k].lod;
                const uint32_t acx = ((vx1 - vx0) >> L) / VC_A, acy = ((vy1 - vy0) >> L) / VC_A;
                const uint32_t ax0 = (vx0 >> L) / VC_A, ay0 = (vy0 >> L) / VC_A, gaz = (vz0 >> L) / VC_A + ev[k].zslab;
                const uint32_t gcx = (ld[L][0] + 31) / 32, gcy = (ld[L][1] + 31) / 32;
                for (uint32_t ay = 0; ay < acy && st == VC_OK; ++ay)
                    for (uint32_t ax = 0; ax < acx && st == VC_OK; ++ax) {
                        const uint64_t gi = prefix[L] + ((uint64_t)gaz * gcy + ay0 + ay) * gcx + ax0 + ax;
                        st = put_atom(w, L, gaz, ay0 + ay, ax0 + ax, pay + off[gi], lens[gi]);
                    }
            }
        }
        free(lens); free(off); free(pay);
    }
    for (int l = 0; l <= fine; ++l) { vcg_free(lv[l]); lv[l] = NULL; }

    /* coarse tail: level L from the archive's decoded level L-1 */
    for (int L = fine + 1; L < nlod && st == VC_OK; ++L) {
        const vc_dims dp = vc__lod_dims(pd, L - 1), dl = vc__lod_dims(pd, L);
        const uint32_t pcx = (dp.nx + 31) / 32, pcy = (dp.ny + 31) / 32, pcz = (dp.nz + 31) / 32;
        const uint32_t dcx = (dl.nx + 31) / 32, dcy = (dl.ny + 31) / 32, dcz = (dl.nz + 31) / 32;
        const size_t np = (size_t)pcx * pcy * pcz;
        int32_t *q = (int32_t *)malloc(np * 16);
        uint8_t *par = (uint8_t *)malloc(np * (size_t)VC_A3);
        const uint32_t vx = 64 * dcx, vy = 64 * dcy, vz = 64 * dcz;       /* whole 2x2x2 parent blocks */
        uint8_t *v = (uint8_t *)calloc((size_t)vx * vy * vz, 1);
        if (!q || !par || !v) vc__panic("out of memory");
        size_t k = 0;
        for (uint32_t az = 0; az < pcz; ++az)
            for (uint32_t ay = 0; ay < pcy; ++ay)
                for (uint32_t ax = 0; ax < pcx; ++ax, ++k) {
                    q[4 * k] = L - 1; q[4 * k + 1] = (int32_t)ax; q[4 * k + 2] = (int32_t)ay; q[4 * k + 3] = (int32_t)az;
                }
        st = vc__writer_decode_atoms(w, np, q, par);
        if (st == VC_ERR_FORMAT) st = VC_OK;          /* a corrupt parent reads as zeros, as in the tool */
        k = 0;
        for (uint32_t az = 0; az < pcz; ++az)
            for (uint32_t ay = 0; ay < pcy; ++ay)
                for (uint32_t ax = 0; ax < pcx; ++ax, ++k)
                    for (uint32_t z = 0; z < 32; ++z)
                        for (uint32_t y = 0; y < 32; ++y)
                            memcpy(v + ((size_t)(az * 32 + z) * vy + ay * 32 + y) * vx + ax * 32,
                                   par + k * (size_t)VC_A3 + (z * 32 + y) * 32, 32);
        free(q); free(par);
        uint8_t *dv = (uint8_t *)vcg_alloc((size_t)vx * vy * vz), *du = (uint8_t *)vcg_alloc((size_t)vx * vy * vz / 8);
        if (!dv || !du) { vcg_free(dv); vcg_free(du); free(v); st = VC_ERR_OOM; break; }
        vcg_copy(dv, v, (size_t)vx * vy * vz, 1);
        free(v);
        const uint32_t dd[3] = {vx, vy, vz}, ud[1][3] = {{vx / 2, vy / 2, vz / 2}};
        vcg_ds2x(dv, dd, method, alpha, du, 1);
        uint32_t *lens; uint64_t *off, total; uint8_t *pay;
        const uint8_t *lvl[1] = {du};
        if (st == VC_OK) st = encode_levels(lvl, ud, 1, &qs[L], &lens, &off, &pay, &total);
        vcg_free(dv); vcg_free(du);
        if (st != VC_OK) break;
        k = 0;
        for (uint32_t az = 0; az < dcz && st == VC_OK; ++az)
            for (uint32_t ay = 0; ay < dcy && st == VC_OK; ++ay)
                for (uint32_t ax = 0; ax < dcx && st == VC_OK; ++ax, ++k)
                    st = put_atom(w, L, az, ay, ax, pay + off[k], lens[k]);
        free(lens); free(off); free(pay);
    }
done:
    for (int l = 0; l < VC_NLOD; ++l) vcg_free(lv[l]);
    vc_writer_close(w);
    return st;
}
