/* vc_tables.c — codec tables (host libm), geometry helpers, panic plumbing. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "vc_host.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ---- fail-fast panic (src/vc/vc.c:26-34) ---- */
static void (*g_panic)(const char *) = NULL;
void vc_set_panic_hook(void (*hook)(const char *msg)) { g_panic = hook; }
void vc__panic(const char *msg) {
    if (g_panic) g_panic(msg);
    fprintf(stderr, "vc: panic: %s\n", msg);
    abort();
}

/* ---- tables ---- */
static vcg_tables g_tab;
static pthread_once_t g_tab_once = PTHREAD_ONCE_INIT;

static void build(void) {
    /* Q14 orthonormal DCT-II basis, src/vc/vc.c:77-88 */
    for (int k = 0; k < 32; ++k) {
        double ck = k == 0 ? sqrt(1.0 / 32.0) : sqrt(2.0 / 32.0);
        for (int n = 0; n < 32; ++n)
            g_tab.cmat[k][n] = (int32_t)llround(ck * cos(M_PI * (2.0 * n + 1.0) * k / 64.0) * 16384.0);
    }
    /* HF weight per L1 frequency, src/vc/vc.c:273-279 */
    for (int f = 0; f < VCG_NFREQ; ++f) g_tab.gstep[f] = powf(1.0f + (float)f, 0.65f);
    /* scan: ascending L1 frequency, ties by raster index (src/vc/vc.c:258-287),
     * built by bucketing on frequency */
    uint32_t cnt[VCG_NFREQ + 1] = {0};
    for (uint32_t i = 0; i < VC_A3; ++i) cnt[(i >> 10) + ((i >> 5) & 31) + (i & 31) + 1]++;
    for (int f = 0; f < VCG_NFREQ; ++f) cnt[f + 1] += cnt[f];
    for (uint32_t i = 0; i < VC_A3; ++i) g_tab.scan[cnt[(i >> 10) + ((i >> 5) & 31) + (i & 31)]++] = (uint16_t)i;
    /* band(freq) = min(7, freq*8/96) (src/vc/vc.c:462-468); freq is nondecreasing
     * along the scan, so each band is one contiguous scan range */
    int b = 0;
    g_tab.band_start[0] = 0;
    for (uint32_t p = 0; p < VC_A3; ++p) {
        uint32_t i = g_tab.scan[p];
        int f = (int)((i >> 10) + ((i >> 5) & 31) + (i & 31));
        int bb = f * 8 / 96; if (bb > 7) bb = 7;
        while (b < bb) g_tab.band_start[++b] = p;
    }
    while (b < 8) g_tab.band_start[++b] = VC_A3;
}
const vcg_tables *vc__tables(void) { pthread_once(&g_tab_once, build); return &g_tab; }

/* Q8 dequant tables per frequency for one base_q (src/vc/vc.c:318-326) */
void vc__dq(float base_q, int32_t step[VCG_NFREQ], int32_t bias[VCG_NFREQ]) {
    const vcg_tables *t = vc__tables();
    for (int f = 0; f < VCG_NFREQ; ++f) {
        float st = base_q * t->gstep[f];
        step[f] = (int32_t)lrintf(st * 256.0f);
        bias[f] = (int32_t)lrintf((0.40f - 1.0f + 0.80f) * st * 256.0f);
    }
}

void vc__require_gpu(void) {
    if (vcg_init(vc__tables()) != 0) vc__panic("no usable HIP device (gfx950) for the codec kernels");
}

float vc__clamp_q(float q) {                 /* src/vc/vc.c:1072-1073 */
    if (q < 0.05f) q = 0.05f;
    if (q > 4096.0f) q = 4096.0f;
    return q;
}

/* ---- geometry (src/vc/vc.c:721-765) ---- */
vc_dims vc__lod_dims(vc_dims d, int lod) {
    for (int i = 0; i < lod; ++i) { d.nx = (d.nx + 1) / 2; d.ny = (d.ny + 1) / 2; d.nz = (d.nz + 1) / 2; }
    return d;
}
void vc__region_grid(vc_dims d0, int lod, uint32_t *nrz, uint32_t *nry, uint32_t *nrx) {
    vc_dims d = vc__lod_dims(d0, lod);
    *nrx = ((d.nx + VC_A - 1) / VC_A + R_ATOMS - 1) / R_ATOMS;
    *nry = ((d.ny + VC_A - 1) / VC_A + R_ATOMS - 1) / R_ATOMS;
    *nrz = ((d.nz + VC_A - 1) / VC_A + R_ATOMS - 1) / R_ATOMS;
}
int vc__nlod(vc_dims d0) {
    int nlod = 1;
    for (int l = 1; l < VC_NLOD; ++l) {
        vc_dims d = vc__lod_dims(d0, l);
        nlod = l + 1;
        if (d.nx <= 1 && d.ny <= 1 && d.nz <= 1) break;
    }
    return nlod;
}
int vc__dims_valid(vc_dims d) {
    return d.nx && d.ny && d.nz && d.nx <= VC_MAX_DIM && d.ny <= VC_MAX_DIM && d.nz <= VC_MAX_DIM;
}
uint64_t vc__mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
uint64_t vc__region_key(int lod, uint32_t rz, uint32_t ry, uint32_t rx) {
    return ((uint64_t)(lod & 7) << 45) | ((uint64_t)(rz & 0x7FFF) << 30) | ((uint64_t)(ry & 0x7FFF) << 15) |
           (uint64_t)(rx & 0x7FFF);
}
uint32_t vc_region_atoms(void) { return R_ATOMS; }
