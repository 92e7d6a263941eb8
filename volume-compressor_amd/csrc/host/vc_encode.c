/* vc_encode.c — whole-volume encode.
 *
 * vc_encode is sketched but absent in the reference (README.md:51-52,
 * plan.txt:206-207); this build defines it by composing the reference's own
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
 * evaluated RC_DEPTH steps per launch: each round encodes the sample at every
 * midpoint of the next RC_DEPTH levels of the bisection tree (plus both bounds
 * in the first round) and then follows the path the sequential search takes.
 * A candidate's encode stops once its byte total can no longer meet the
 * target (the largest meeting total is fixed by the voxel count), which
 * decides it the same way. */
#define RC_DEPTH 4
#define RC_NODES ((1 << RC_DEPTH) - 1)

static float rate_control(vcg_enc *e, vc_dims d, float target) {
    uint32_t ix[8], iy[8], iz[8];
    int nx = lattice((d.nx + 31) / 32, ix), ny = lattice((d.ny + 31) / 32, iy), nz = lattice((d.nz + 31) / 32, iz);
    int ns = nx * ny * nz;
    uint32_t *list = (uint32_t *)malloc((size_t)ns * 12);
    uint32_t *vox = (uint32_t *)malloc((size_t)ns * 4);
    uint8_t *zero = (uint8_t *)malloc((size_t)ns);
    if (!list || !vox || !zero) vc__panic("out of memory");
    int k = 0;
    for (int a = 0; a < nz; ++a) for (int b = 0; b < ny; ++b) for (int c = 0; c < nx; ++c, ++k) {
        list[3 * k] = iz[a]; list[3 * k + 1] = iy[b]; list[3 * k + 2] = ix[c];
        uint32_t ex = d.nx - ix[c] * 32, ey = d.ny - iy[b] * 32, ez = d.nz - iz[a] * 32;
        vox[k] = (ex < 32 ? ex : 32) * (ey < 32 ? ey : 32) * (ez < 32 ? ez : 32);
    }
    vcg_enc_sample_zero(e, list, ns, zero);
    uint64_t v = 0;
    for (int i = 0; i < ns; ++i) if (!zero[i]) v += vox[i];
    float q0;
    if (v == 0) {
        q0 = 1.0f;                        /* all-zero sample: nothing to calibrate */
    } else {
        /* meets(q): no overflow and voxels >= target * bytes (double); budget =
         * the largest byte total that meets */
        uint64_t budget = (uint64_t)((double)v / (double)target);
        while ((double)v >= (double)target * (double)(budget + 1)) budget++;
        while (budget > 0 && !((double)v >= (double)target * (double)budget)) budget--;
        float lo = 0.05f, hi = 4096.0f;
        float qs[RC_NODES + 2], mid[RC_NODES + 1], L[RC_NODES + 1], H[RC_NODES + 1];
        uint64_t bytes[RC_NODES + 2];
        int flags[RC_NODES + 2], meets[RC_NODES + 2];
        int first = 1, steps = 0, done = 0;
        q0 = hi;
        while (!done) {
            /* level-order midpoints of the next RC_DEPTH levels from (lo, hi) */
            L[1] = lo; H[1] = hi;
            for (int t = 1; t <= RC_NODES; ++t) {
                mid[t] = sqrtf(L[t] * H[t]);
                if (2 * t + 1 <= RC_NODES) {
                    L[2 * t] = L[t]; H[2 * t] = mid[t];          /* meets: hi = mid */
                    L[2 * t + 1] = mid[t]; H[2 * t + 1] = H[t];  /* else:  lo = mid */
                }
            }
            int nq = 0;
            if (first) { qs[nq++] = lo; qs[nq++] = hi; }
            const int base = nq;
            for (int t = 1; t <= RC_NODES; ++t) qs[nq++] = mid[t];
            vcg_enc_sample_multi(e, list, ns, qs, nq, budget, bytes, flags);
            for (int j = 0; j < nq; ++j)
                meets[j] = flags[j] == 0 && (double)v >= (double)target * (double)bytes[j];
            if (first) {
                first = 0;
                if (meets[0]) { q0 = lo; break; }
                if (!meets[1]) { q0 = hi; break; }
            }
            int t = 1;
            for (int l = 0; l < RC_DEPTH && steps < 16; ++l, ++steps) {
                if (meets[base + t - 1]) { hi = mid[t]; t = 2 * t; }
                else { lo = mid[t]; t = 2 * t + 1; }
            }
            if (steps >= 16) { q0 = hi; done = 1; }
        }
    }
    free(list); free(vox); free(zero);
    return q0;
}

static void lod_schedule(float q0, int nlod, float qs[VC_NLOD]) {
    float q = q0;
    for (int l = 0; l < VC_NLOD; ++l) { qs[l] = l < nlod ? vc__clamp_q(q) : 0.0f; q *= 0.6f; }
}

/* ---- layout + assembly ---- */
static vc_status assemble(vcg_enc *e, vc_dims d0, int nlod, float target, const float qs[VC_NLOD],
                          const uint32_t *lens, uint64_t total, uint8_t **out, size_t *out_len) {
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
    if (too_big) { rc = VC_ERR_RANGE; goto done; }
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
        uint8_t *img = (uint8_t *)malloc(cursor);
        if (!img) { free(head); rc = VC_ERR_OOM; goto done; }
        vcg_enc_assemble(e, head, head_len, pay_off, slot_off, lens, total, img, cursor);
        free(head);
        *out = img; *out_len = (size_t)cursor;
    }
done:
    free(cls); free(blen); free(boff); free(pay_off); free(slot_off);
    return rc;
}

static vc_status encode_impl(const uint8_t *vol, int on_dev, vc_dims dims, float target, const float *qs_in,
                             uint8_t **out, size_t *out_len, float *q0_only) {
    if (!vol || (!q0_only && (!out || !out_len))) return VC_ERR_RANGE;
    if (!vc__dims_valid(dims) || !(target >= 1.0f)) return VC_ERR_RANGE;
    vc__require_gpu();
    int nlod = vc__nlod(dims);
    uint32_t d[3] = {dims.nx, dims.ny, dims.nz};
    vcg_enc *e = vcg_enc_begin(vol, on_dev, d, nlod);
    float qs[VC_NLOD];
    if (qs_in) {
        for (int l = 0; l < VC_NLOD; ++l) qs[l] = l < nlod ? vc__clamp_q(qs_in[l]) : 0.0f;
    } else {
        float q0 = rate_control(e, dims, target);
        if (q0_only) { *q0_only = q0; vcg_enc_end(e); return VC_OK; }
        lod_schedule(q0, nlod, qs);
    }
    uint64_t total = 0;
    for (int l = 0; l < nlod; ++l) {
        vc_dims dl = vc__lod_dims(dims, l);
        total += (uint64_t)((dl.nx + 31) / 32) * ((dl.ny + 31) / 32) * ((dl.nz + 31) / 32);
    }
    uint32_t *lens = (uint32_t *)malloc(total * 4);
    if (!lens) { vcg_enc_end(e); return VC_ERR_OOM; }
    vc_status rc;
    if (vcg_enc_all(e, qs, lens, total)) rc = VC_ERR_RANGE;    /* a payload exceeded the coder bound */
    else rc = assemble(e, dims, nlod, target, qs, lens, total, out, out_len);
    free(lens);
    vcg_enc_end(e);
    return rc;
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
