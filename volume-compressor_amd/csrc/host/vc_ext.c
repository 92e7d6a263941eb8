/* vc_ext.c — vc_gpu.h extensions, the synthetic generator's host half and the
 * reference's test hooks (vc_test_downsample2x, vc_test_atom_rt). */
#include <stdlib.h>

#include "vc_host.h"
#include "vc_testhooks.h"

vc_status vcx_device_init(void) { return vcg_init(vc__tables()) == 0 ? VC_OK : VC_ERR_IO; }
void *vcx_stream(void) { vc__require_gpu(); return vcg_stream(); }
void *vcx_dev_alloc(size_t bytes) { vc__require_gpu(); return vcg_alloc(bytes); }
void vcx_dev_free(void *p) { if (p) vcg_free(p); }
vc_status vcx_h2d(void *dst, const void *src, size_t n) { vc__require_gpu(); vcg_copy(dst, src, n, 1); return VC_OK; }
vc_status vcx_d2h(void *dst, const void *src, size_t n) { vc__require_gpu(); vcg_copy(dst, src, n, 2); return VC_OK; }
vc_status vcx_copy_kernel(void *dst, const void *src, size_t n) {
    vc__require_gpu();
    return vcg_copy_kernel(dst, src, n) ? VC_ERR_RANGE : VC_OK;
}
vc_status vcx_sync(void) { vc__require_gpu(); return vcg_sync() ? VC_ERR_FORMAT : VC_OK; }
void vcx_set_stage_events(void *events[3]) { vc__require_gpu(); vcg_set_stage_events(events); }

vc_status vcx_encode_atoms(size_t n, const uint8_t *vox, float base_q, uint8_t *out, size_t cap, uint32_t *lens) {
    if ((n && (!vox || !out || !lens)) || cap < 1) return VC_ERR_RANGE;
    vc__require_gpu();
    return vcg_encode_atoms(n, vox, base_q, out, cap, lens) ? VC_ERR_RANGE : VC_OK;
}

vc_status vcx_decode_payloads(size_t n, const uint8_t *pay, const uint64_t *offs, const uint32_t *lens,
                              float base_q, uint8_t *out, vc_status *per_atom) {
    if (n && (!pay || !offs || !lens || !out)) return VC_ERR_RANGE;
    vc__require_gpu();
    int32_t step[VCG_NFREQ], bias[VCG_NFREQ];
    vc__dq(base_q, step, bias);
    int32_t *st = (int32_t *)calloc(n ? n : 1, 4);
    if (!st) return VC_ERR_OOM;
    int rc = vcg_decode_payloads(n, pay, offs, lens, step, bias, out, st);
    if (per_atom) for (size_t i = 0; i < n; ++i) per_atom[i] = st[i] ? VC_ERR_FORMAT : VC_OK;
    free(st);
    return rc ? VC_ERR_FORMAT : VC_OK;
}

/* ---- synthetic volume: sheet parameters drawn on the host ----
 * splitmix64 stream seeded with seed ^ 0x5EEDC0DE5EEDC0DE; uniform integers
 * u(n) = (next() >> 11) % n.  Per sheet: dominant axis (0 z, 1 y, 2 x), centre
 * (Q8), two tilt slopes (Q16, +-0.25), two warp amplitudes (Q8, up to M/16 and
 * M/32 voxels), two wavelengths (1 + M/4 + u(M) voxels, as 16-bit turns per
 * voxel), two phases, half thickness (Q8, thickness 3..6 voxels) and
 * intensity 110..190, where M is the largest dimension. */
static uint64_t sm_next(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static uint32_t sm_u(uint64_t *s, uint32_t n) { return n ? (uint32_t)((sm_next(s) >> 11) % n) : 0; }

vc_status vcx_synth_device(uint64_t seed, vc_dims dims, int nsheets, void *dev_out) {
    if (!dev_out || !vc__dims_valid(dims) || nsheets < 0) return VC_ERR_RANGE;
    vc__require_gpu();
    vcg_sheet *sh = (vcg_sheet *)calloc(nsheets > 0 ? nsheets : 1, sizeof *sh);
    if (!sh) return VC_ERR_OOM;
    uint64_t st = seed ^ 0x5EEDC0DE5EEDC0DEull;
    uint32_t M = dims.nx > dims.ny ? dims.nx : dims.ny;
    if (dims.nz > M) M = dims.nz;
    uint32_t dim[3] = {dims.nz, dims.ny, dims.nx};
    for (int s = 0; s < nsheets; ++s) {
        vcg_sheet *p = &sh[s];
        p->axis = (int32_t)sm_u(&st, 3);
        p->c0 = (int32_t)sm_u(&st, dim[p->axis] * 256u);
        p->tu = (int32_t)sm_u(&st, 32769) - 16384;
        p->tv = (int32_t)sm_u(&st, 32769) - 16384;
        p->a1 = (int32_t)sm_u(&st, M * 256u / 16u + 1);
        p->a2 = (int32_t)sm_u(&st, M * 256u / 32u + 1);
        p->w1 = (int32_t)(65536u / (1u + M / 4u + sm_u(&st, M)));
        p->w2 = (int32_t)(65536u / (1u + M / 4u + sm_u(&st, M)));
        p->p1 = (int32_t)sm_u(&st, 65536);
        p->p2 = (int32_t)sm_u(&st, 65536);
        p->half = (int32_t)((768u + sm_u(&st, 769)) / 2u);
        p->inten = 110 + (int32_t)sm_u(&st, 81);
    }
    uint32_t d[3] = {dims.nx, dims.ny, dims.nz};
    int rc = vcg_synth(seed, d, sh, nsheets, dev_out);
    free(sh);
    return rc ? VC_ERR_OOM : VC_OK;
}

/* ---- the exporter's within-cell downscale (tools/downscale.h:32-93) ---- */
vc_status vcx_downscale2x_device(const void *dev_in, vc_dims in, int method, float alpha, void *dev_out, int sync) {
    if (!dev_in || !dev_out || !in.nx || !in.ny || !in.nz) return VC_ERR_RANGE;
    if ((method != VCX_DS_BOX && method != VCX_DS_CBOX) || !(alpha == alpha)) return VC_ERR_RANGE;
    vc__require_gpu();
    const uint32_t d[3] = {in.nx, in.ny, in.nz};
    vcg_ds2x(dev_in, d, method, alpha, dev_out, sync);
    return VC_OK;
}

void vcx_test_wide_index(int on) { vcg_test_wide_index(on); }

/* ---- test hooks (src/vc/vc.c:1557-1591), run on the GPU ---- */
uint8_t *vc_test_downsample2x(const uint8_t *vol, vc_dims in, vc_dims *out) {
    vc__require_gpu();
    vc_dims o = {(in.nx + 1) / 2, (in.ny + 1) / 2, (in.nz + 1) / 2};
    if (!o.nx) o.nx = 1;
    if (!o.ny) o.ny = 1;
    if (!o.nz) o.nz = 1;
    size_t ni = (size_t)in.nx * in.ny * in.nz, no = (size_t)o.nx * o.ny * o.nz;
    uint8_t *h = (uint8_t *)malloc(no ? no : 1);
    if (!h) vc__panic("out of memory");
    if (ni) {
        void *di = vcg_alloc(ni), *dout = vcg_alloc(no);
        if (!di || !dout) vc__panic("out of device memory");
        vcg_copy(di, vol, ni, 1);
        uint32_t d[3] = {in.nx, in.ny, in.nz};
        vcg_downsample(di, d, dout);
        vcg_copy(h, dout, no, 2);
        vcg_free(di); vcg_free(dout);
    } else memset(h, 0, no);
    if (out) *out = o;
    return h;
}

uint32_t vc_test_atom_rt(const uint8_t vox[VC_ATOM3], float q, uint8_t *recon) {
    vc__require_gpu();
    uint8_t *pay = (uint8_t *)malloc(VCG_PAY_CAP);
    if (!pay) vc__panic("out of memory");
    uint32_t len = 0;
    if (vcg_encode_atoms(1, vox, q, pay, VCG_PAY_CAP, &len)) vc__panic("range encoder overflow");
    if (recon) {
        if (len == 0) memset(recon, 0, VC_A3);
        else {
            int32_t step[VCG_NFREQ], bias[VCG_NFREQ], st = 0;
            uint64_t off = 0;
            vc__dq(q, step, bias);
            vcg_decode_payloads(1, pay, &off, &len, step, bias, recon, &st);
        }
    }
    free(pay);
    return len;
}
