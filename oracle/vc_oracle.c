/* vc_oracle.c — CPU restatement of the volume-compressor atom pipeline.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  See vc_oracle.h.  Built with
 * -fwrapv (the reference's i32 MAC/dequant arithmetic wraps in practice; the
 * oracle makes that explicit) and strict IEEE float (-ffp-contract=off, no
 * fast-math): the quantizer's float expression is evaluated exactly in source
 * order, which is the order the HIP kernel pins.
 *
 * Citations are to the reference tree (SuperOptimizer/volume-compressor).
 */
#include "vc_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <time.h>

typedef uint8_t u8;  typedef uint16_t u16; typedef uint32_t u32; typedef uint64_t u64;
typedef int16_t i16; typedef int32_t i32;  typedef int64_t i64;

#define A   VCO_A
#define A3  VCO_A3

/* ===================================================================== */
/* Tables                                                                */
/* ===================================================================== */
static i32   T_C[32][32];         /* Q14 DCT-II basis, vc.c:75-88 */
static float T_GSTEP[VCO_NFREQ];  /* (1+freq)^0.65f, vc.c:273-279 (depends only on freq) */
static u16   T_SCAN[A3];          /* scan pos -> raster idx, vc.c:258-287 */
static u8    T_BAND[A3];          /* band of scan pos, vc.c:462-468 */
static u8    T_FREQ[A3];          /* L1 freq of raster idx */
static pthread_once_t g_once = PTHREAD_ONCE_INIT;

static void build_tables(void) {
    /* CMATA[k][n] = llround(c_k * cos(pi*(2n+1)*k/64) * 2^14), vc.c:79-86 */
    for (int k = 0; k < A; ++k) {
        double ck = (k == 0) ? sqrt(1.0 / 32.0) : sqrt(2.0 / 32.0);
        for (int n = 0; n < A; ++n)
            T_C[k][n] = (i32)llround(ck * cos(M_PI * (2.0 * n + 1.0) * k / 64.0) * 16384.0);
    }
    for (int f = 0; f < VCO_NFREQ; ++f) T_GSTEP[f] = powf(1.0f + (float)f, 0.65f);
    for (int i = 0; i < A3; ++i) T_FREQ[i] = (u8)((i >> 10) + ((i >> 5) & 31) + (i & 31));
    /* scan = stable order by (freq, raster idx): identical to the qsort with
     * scan_cmp at vc.c:258-264 (a total order), done here as a counting sort. */
    u32 start[VCO_NFREQ + 1] = {0};
    for (int i = 0; i < A3; ++i) start[T_FREQ[i] + 1]++;
    for (int f = 0; f < VCO_NFREQ; ++f) start[f + 1] += start[f];
    for (int i = 0; i < A3; ++i) T_SCAN[start[T_FREQ[i]]++] = (u16)i;
    for (int p = 0; p < A3; ++p) {
        int b = (int)T_FREQ[T_SCAN[p]] * 8 / 96;   /* vc.c:465-466 */
        T_BAND[p] = (u8)(b > 7 ? 7 : b);
    }
}
static void tables(void) { pthread_once(&g_once, build_tables); }

void vco_tables(i32 cmat[32][32], float gstep[VCO_NFREQ], u16 scan[A3], u8 band[A3]) {
    tables();
    if (cmat) memcpy(cmat, T_C, sizeof T_C);
    if (gstep) memcpy(gstep, T_GSTEP, sizeof T_GSTEP);
    if (scan) memcpy(scan, T_SCAN, sizeof T_SCAN);
    if (band) memcpy(band, T_BAND, sizeof T_BAND);
}

/* Q8 dequant tables keyed by freq, vc.c:318-326 */
void vco_dq_tables(float base_q, i32 dq_step[VCO_NFREQ], i32 dq_bias[VCO_NFREQ]) {
    tables();
    for (int f = 0; f < VCO_NFREQ; ++f) {
        float step = base_q * T_GSTEP[f];
        dq_step[f] = (i32)lrintf(step * 256.0f);
        dq_bias[f] = (i32)lrintf((0.40f - 1.0f + 0.80f) * step * 256.0f);
    }
}

/* ===================================================================== */
/* Integer DCT-32                                                        */
/* ===================================================================== */
/* One forward line: even/odd partial butterfly, vc.c:148-169 (gcc form). */
static void fwd_line(i32 *v, int stride) {
    i32 s[16], d[16], o[32];
    for (int n = 0; n < 16; ++n) {
        s[n] = v[n * stride] + v[(31 - n) * stride];
        d[n] = v[n * stride] - v[(31 - n) * stride];
    }
    for (int k = 0; k < 32; ++k) {
        const i32 *src = (k & 1) ? d : s;
        i32 acc = 0;
        for (int n = 0; n < 16; ++n) acc += T_C[k][n] * src[n];
        o[k] = (acc + 8192) >> 14;
    }
    for (int k = 0; k < 32; ++k) v[k * stride] = o[k];
}
/* One inverse line, vc.c:178-184. */
static void inv_line(i32 *v, int stride) {
    i32 o[32];
    for (int n = 0; n < 16; ++n) {
        i32 evn = 0, odd = 0;
        for (int k = 0; k < 32; k += 2) evn += T_C[k][n] * v[k * stride];
        for (int k = 1; k < 32; k += 2) odd += T_C[k][n] * v[k * stride];
        o[n] = (evn + odd + 8192) >> 14;
        o[31 - n] = (evn - odd + 8192) >> 14;
    }
    for (int n = 0; n < 32; ++n) v[n * stride] = o[n];
}
/* The reference transforms the contiguous axis and rotates (z,y,x)->(x,z,y)
 * between passes (vc.c:208-218): passes run along x, then y, then z, and after
 * the third rotate the result is raster (kz,ky,kx).  Here the three passes are
 * applied in place along each axis of the (z,y,x) raster instead. */
static void dct3(i32 *blk, void (*line)(i32 *, int)) {
    for (int z = 0; z < A; ++z) for (int y = 0; y < A; ++y) line(blk + (z * A + y) * A, 1);
    for (int z = 0; z < A; ++z) for (int x = 0; x < A; ++x) line(blk + z * A * A + x, A);
    for (int y = 0; y < A; ++y) for (int x = 0; x < A; ++x) line(blk + y * A + x, A * A);
}

void vco_fdct(const u8 vox[A3], i32 dc, i32 coef[A3]) {
    tables();
    for (int i = 0; i < A3; ++i) coef[i] = (i32)vox[i] - dc;   /* vc.c:214 */
    dct3(coef, fwd_line);
}

void vco_idct(const i32 coef[A3], i32 dc, u8 vox[A3]) {
    tables();
    i32 *b = (i32 *)malloc(A3 * sizeof(i32));
    memcpy(b, coef, A3 * sizeof(i32));
    dct3(b, inv_line);
    for (int i = 0; i < A3; ++i) {             /* vc.c:228 */
        i32 v = b[i] + dc;
        vox[i] = (u8)(v < 0 ? 0 : v > 255 ? 255 : v);
    }
    free(b);
}

/* ===================================================================== */
/* Quantization                                                          */
/* ===================================================================== */
/* Dead-zone quantizer, strict IEEE in source order, vc.c:293-306. */
static i32 quant1(i32 c, float base_q, float g) {
    float step = base_q * g;
    float dz = 0.80f * step;
    float a = (float)(c < 0 ? -c : c);
    i32 level = 0;
    if (a >= dz) level = (i32)((a - dz) / step + 1.0f);
    if (level > 32767) level = 32767;
    return c < 0 ? -level : level;
}
void vco_quantize(const i32 coef[A3], float base_q, i16 q[A3]) {
    tables();
    for (int i = 0; i < A3; ++i) q[i] = (i16)quant1(coef[i], base_q, T_GSTEP[T_FREQ[i]]);
}
void vco_quant_levels(float base_q, int freq, const i32 *coef, size_t n, i32 *out) {
    tables();
    for (size_t i = 0; i < n; ++i) out[i] = quant1(coef[i], base_q, T_GSTEP[freq]);
}

void vco_dequantize(const i16 q[A3], float base_q, i32 coef[A3]) {
    i32 st[VCO_NFREQ], bi[VCO_NFREQ];
    vco_dq_tables(base_q, st, bi);
    for (int i = 0; i < A3; ++i) {             /* vc.c:330-337 */
        i32 lv = q[i];
        i32 mag = lv < 0 ? -lv : lv;
        i32 r = (mag * st[T_FREQ[i]] + bi[T_FREQ[i]] + 128) >> 8;
        coef[i] = lv == 0 ? 0 : (lv < 0 ? -r : r);
    }
}

/* ===================================================================== */
/* Adaptive binary range coder (LZMA-style carry), vc.c:344-441          */
/* ===================================================================== */
typedef struct { u8 *buf; size_t cap, len; u64 low; u32 range; u8 cache; u64 csize; int ovf; } renc;
typedef struct { const u8 *buf; size_t len, pos; u32 code, range; } rdec;

static void re_put(renc *e, u8 b) {
    if (e->len >= e->cap) { e->ovf = 1; return; }
    e->buf[e->len++] = b;
}
static void re_shift_low(renc *e) {            /* vc.c:369-380 */
    if ((u32)(e->low >> 32) != 0 || e->low < 0xFF000000ull) {
        u8 carry = (u8)(e->low >> 32);
        do { re_put(e, (u8)(e->cache + carry)); e->cache = 0xFF; } while (--e->csize);
        e->cache = (u8)(e->low >> 24);
    }
    e->csize++;
    e->low = (e->low << 8) & 0xFFFFFFFFull;
}
static void re_norm(renc *e) { while (e->range < (1u << 24)) { re_shift_low(e); e->range <<= 8; } }
static void re_bit(renc *e, u16 *p, int bit) { /* vc.c:381-393 */
    u32 r0 = (u32)(((u64)e->range * *p) >> 12);
    if (!bit) { e->range = r0; *p = (u16)(*p + ((4096 - *p) >> 5)); }
    else { e->low += r0; e->range -= r0; *p = (u16)(*p - (*p >> 5)); }
    re_norm(e);
}
static void re_bypass(renc *e, int bit) {     /* vc.c:394-398 */
    e->range >>= 1;
    if (bit) e->low += e->range;
    re_norm(e);
}

static u8 rd_byte(rdec *d) { return d->pos < d->len ? d->buf[d->pos++] : 0; } /* vc.c:407,425 */
static void rd_norm(rdec *d) {
    while (d->range < (1u << 24)) { d->code = (d->code << 8) | rd_byte(d); d->range <<= 8; }
}
static int rd_bit(rdec *d, u16 *p) {           /* vc.c:414-430 */
    u32 r0 = (u32)(((u64)d->range * *p) >> 12);
    int bit;
    if (d->code < r0) { d->range = r0; bit = 0; *p = (u16)(*p + ((4096 - *p) >> 5)); }
    else { d->code -= r0; d->range -= r0; bit = 1; *p = (u16)(*p - (*p >> 5)); }
    rd_norm(d);
    return bit;
}
static int rd_bypass(rdec *d) {                /* vc.c:431-441 */
    d->range >>= 1;
    int bit = d->code >= d->range;
    if (bit) d->code -= d->range;
    rd_norm(d);
    return bit;
}

/* Coefficient context coder: 8 sig bands + 12 magnitude contexts, reset per
 * atom, EOB as 16 bypass bits, magnitude unary over mag[0..11] then EG0 bypass,
 * sign bypass.  vc.c:448-560. */
long vco_entropy_encode(const i16 q[A3], u8 *out, size_t cap) {
    tables();
    renc e = { out, cap, 0, 0, 0xFFFFFFFFu, 0, 1, 0 };
    u16 sig[8], mag[12];
    for (int i = 0; i < 8; ++i) sig[i] = 2048;
    for (int i = 0; i < 12; ++i) mag[i] = 2048;
    u32 eob = 0;
    for (u32 p = A3; p-- > 0;) if (q[T_SCAN[p]] != 0) { eob = p + 1; break; }
    for (int i = 15; i >= 0; --i) re_bypass(&e, (eob >> i) & 1);        /* vc.c:513-519 */
    for (u32 p = 0; p < eob; ++p) {
        i16 v = q[T_SCAN[p]];
        u16 *sp = &sig[T_BAND[p]];
        if (v == 0) { re_bit(&e, sp, 0); continue; }
        re_bit(&e, sp, 1);
        u32 m = (u32)(v < 0 ? -v : v);
        /* magnitude, vc.c:471-490 */
        u32 r = m - 1, k = 0;
        int done = 0;
        while (k < 11 && r > 0) {
            re_bit(&e, &mag[k], 1); r -= 1; k++;
            if (r == 0) { re_bit(&e, &mag[k], 0); done = 1; break; }
        }
        if (!done) {
            if (r == 0) re_bit(&e, &mag[k], 0);
            else {
                re_bit(&e, &mag[k], 1);
                u32 nb = 0, t = r + 1;
                while (t > 1) { t >>= 1; nb++; }
                for (u32 i = 0; i < nb; ++i) re_bypass(&e, 1);
                re_bypass(&e, 0);
                for (int i = (int)nb - 1; i >= 0; --i) re_bypass(&e, ((r + 1) >> i) & 1);
            }
        }
        re_bypass(&e, v < 0);
    }
    for (int i = 0; i < 5; ++i) re_shift_low(&e);                      /* vc.c:399-401 */
    return e.ovf ? -1 : (long)e.len;
}

int vco_entropy_decode(const u8 *rc, size_t len, i16 q[A3], u32 *eob_out) {
    tables();
    memset(q, 0, A3 * sizeof(i16));
    rdec d = { rc, len, 0, 0, 0xFFFFFFFFu };
    for (int i = 0; i < 5; ++i) d.code = (d.code << 8) | rd_byte(&d);  /* vc.c:403-409 */
    u32 eob = 0;
    for (int i = 0; i < 16; ++i) eob = (eob << 1) | (u32)rd_bypass(&d);
    if (eob > A3) eob = A3;                                             /* vc.c:549 */
    if (eob_out) *eob_out = eob;
    u16 sig[8], mag[12];
    for (int i = 0; i < 8; ++i) sig[i] = 2048;
    for (int i = 0; i < 12; ++i) mag[i] = 2048;
    for (u32 p = 0; p < eob; ++p) {
        if (!rd_bit(&d, &sig[T_BAND[p]])) continue;
        /* magnitude, vc.c:491-506 */
        u32 v = 0, k = 0, m;
        while (k < 11 && rd_bit(&d, &mag[k])) { v++; k++; }
        if (k < 11) m = v + 1;
        else if (!rd_bit(&d, &mag[11])) m = v + 1;
        else {
            u32 nb = 0;
            while (rd_bypass(&d)) {
                if (++nb > VCO_EG_PREFIX_CAP) { memset(q, 0, A3 * sizeof(i16)); return VCO_ERR_FORMAT; }
            }
            u32 x = 1;
            for (u32 i = 0; i < nb; ++i) x = (x << 1) | (u32)rd_bypass(&d);
            m = v + (x - 1) + 1;
        }
        int neg = rd_bypass(&d);
        q[T_SCAN[p]] = (i16)(neg ? -(i32)m : (i32)m);
    }
    return VCO_OK;
}

/* ===================================================================== */
/* Atom payload                                                          */
/* ===================================================================== */
long vco_encode_atom(const u8 vox[A3], float base_q, u8 *payload, size_t cap, int *is_zero) {
    u32 sum = 0;                                          /* vc.c:599-604 */
    for (int i = 0; i < A3; ++i) sum += vox[i];
    if (sum == 0) { *is_zero = 1; return 0; }
    *is_zero = 0;
    i32 dc = (i32)((sum + A3 / 2) / A3);
    i32 *coef = (i32 *)malloc(A3 * sizeof(i32));
    i16 *q = (i16 *)malloc(A3 * sizeof(i16));
    vco_fdct(vox, dc, coef);
    vco_quantize(coef, base_q, q);
    long n = -1;
    if (cap >= 1) {
        payload[0] = (u8)dc;
        size_t rc_cap = cap - 1 < (size_t)VCO_RC_CAP ? cap - 1 : (size_t)VCO_RC_CAP;
        long r = vco_entropy_encode(q, payload + 1, rc_cap);
        n = r < 0 ? -1 : r + 1;
    }
    free(coef); free(q);
    return n;
}

int vco_decode_payload(const u8 *pay, u32 len, float base_q, u8 vox[A3]) {
    i32 dc = pay[0];                                      /* vc.c:626 */
    i16 *q = (i16 *)malloc(A3 * sizeof(i16));
    i32 *coef = (i32 *)malloc(A3 * sizeof(i32));
    int st = vco_entropy_decode(pay + 1, len - 1, q, NULL);
    if (st == VCO_OK) {
        vco_dequantize(q, base_q, coef);
        vco_idct(coef, dc, vox);
    } else memset(vox, 0, A3);
    free(q); free(coef);
    return st;
}

/* ===================================================================== */
/* Pyramid + geometry                                                    */
/* ===================================================================== */
void vco_lod_dims(u32 nx, u32 ny, u32 nz, int lod, u32 *ox, u32 *oy, u32 *oz) {
    for (int i = 0; i < lod; ++i) { nx = (nx + 1) / 2; ny = (ny + 1) / 2; nz = (nz + 1) / 2; }
    *ox = nx; *oy = ny; *oz = nz;
}
int vco_nlod(u32 nx, u32 ny, u32 nz) {                    /* vc.c:1001-1006 */
    int nlod = 1;
    for (int l = 1; l < VCO_NLOD; ++l) {
        u32 x, y, z; vco_lod_dims(nx, ny, nz, l, &x, &y, &z);
        nlod = l + 1;
        if (x <= 1 && y <= 1 && z <= 1) break;
    }
    return nlod;
}
/* integer rounded mean over the in-bounds voxels of each 2x2x2 cell, vc.c:1562-1577 */
void vco_downsample2x(const u8 *in, u32 nx, u32 ny, u32 nz, u8 *out) {
    u32 ox = (nx + 1) / 2, oy = (ny + 1) / 2, oz = (nz + 1) / 2;
    /* output planes are independent */
    #pragma omp parallel for schedule(static)
    for (u32 z = 0; z < oz; ++z) for (u32 y = 0; y < oy; ++y) for (u32 x = 0; x < ox; ++x) {
        u32 acc = 0, n = 0;
        for (u32 dz = 0; dz < 2; ++dz) { u32 zz = 2 * z + dz; if (zz >= nz) continue;
        for (u32 dy = 0; dy < 2; ++dy) { u32 yy = 2 * y + dy; if (yy >= ny) continue;
        for (u32 dx = 0; dx < 2; ++dx) { u32 xx = 2 * x + dx; if (xx >= nx) continue;
            acc += in[((size_t)zz * ny + yy) * nx + xx]; n++; } } }
        out[((size_t)z * oy + y) * ox + x] = (u8)((acc + n / 2) / (n ? n : 1));
    }
}
/* within-cell 2x downscale of the exporter, tools/downscale.h:32-93: each
 * output voxel from its own 2x2x2 cell, the odd boundary edge-replicated;
 * method 0 box (float mean, +0.5, truncate), 1 cbox (the mean pushed by alpha
 * toward the first voxel of largest deviation); an all-zero cell is 0.
 * Strict float in source order (no contraction). */
static u8 ds_u8(float v) { return (u8)(v < 0 ? 0 : v > 255 ? 255 : (v + 0.5f)); }
void vco_ds2x(const u8 *in, u32 nx, u32 ny, u32 nz, int method, float alpha, u8 *out) {
    u32 ox = (nx + 1) / 2, oy = (ny + 1) / 2, oz = (nz + 1) / 2;
    if (ox < 1) ox = 1;
    if (oy < 1) oy = 1;
    if (oz < 1) oz = 1;
    /* output planes are independent */
    #pragma omp parallel for schedule(static)
    for (u32 z = 0; z < oz; ++z) for (u32 y = 0; y < oy; ++y) for (u32 x = 0; x < ox; ++x) {
        u8 c[8]; int n = 0; unsigned acc = 0;
        for (u32 dz = 0; dz < 2; ++dz) { u32 zz = 2 * z + dz; if (zz >= nz) zz = nz - 1;
        for (u32 dy = 0; dy < 2; ++dy) { u32 yy = 2 * y + dy; if (yy >= ny) yy = ny - 1;
        for (u32 dx = 0; dx < 2; ++dx) { u32 xx = 2 * x + dx; if (xx >= nx) xx = nx - 1;
            c[n] = in[((size_t)zz * ny + yy) * nx + xx]; acc |= c[n]; n++; } } }
        u8 r = 0;
        if (acc) {
            const float mean = (float)(c[0] + c[1] + c[2] + c[3] + c[4] + c[5] + c[6] + c[7]) * 0.125f;
            if (method == 0) {
                r = ds_u8(mean);
            } else {
                float best = mean, bestdev = -1.0f;
                for (int i = 0; i < 8; ++i) {
                    const float d = (float)c[i] - mean, ad = d < 0 ? -d : d;
                    if (ad > bestdev) { bestdev = ad; best = c[i]; }
                }
                r = ds_u8(mean + alpha * (best - mean));
            }
        }
        out[((size_t)z * oy + y) * ox + x] = r;
    }
}

/* atom gather with zero padding outside dims, tests/test_realdata.c:108-113 */
void vco_gather_atom(const u8 *vol, u32 nx, u32 ny, u32 nz, u32 az, u32 ay, u32 ax, u8 atom[A3]) {
    for (u32 z = 0; z < A; ++z) for (u32 y = 0; y < A; ++y) for (u32 x = 0; x < A; ++x) {
        u32 vz = az * A + z, vy = ay * A + y, vx = ax * A + x;
        atom[(z * A + y) * A + x] = (vz < nz && vy < ny && vx < nx) ? vol[((size_t)vz * ny + vy) * nx + vx] : 0;
    }
}

/* ===================================================================== */
/* Container (v2 layout, tools/vc_repack.c)                              */
/* ===================================================================== */
#define R_ATOMS   32u
#define R3        (R_ATOMS * R_ATOMS * R_ATOMS)
#define SLOT      16u
#define L2_BYTES  ((u64)R3 * SLOT)
#define HDR       128u
#define MAGIC     0x00314356u

static void w32(u8 *p, u32 v) { memcpy(p, &v, 4); }
static void w64(u8 *p, u64 v) { memcpy(p, &v, 8); }
static void wf(u8 *p, float v) { memcpy(p, &v, 4); }
static u32 r32(const u8 *p) { u32 v; memcpy(&v, p, 4); return v; }
static u64 r64(const u8 *p) { u64 v; memcpy(&v, p, 8); return v; }
static float rf(const u8 *p) { float v; memcpy(&v, p, 4); return v; }

static void region_grid(u32 nx, u32 ny, u32 nz, int lod, u32 *rz, u32 *ry, u32 *rx) {
    u32 x, y, z; vco_lod_dims(nx, ny, nz, lod, &x, &y, &z);   /* vc.c:743-748 */
    *rx = ((x + A - 1) / A + R_ATOMS - 1) / R_ATOMS;
    *ry = ((y + A - 1) / A + R_ATOMS - 1) / R_ATOMS;
    *rz = ((z + A - 1) / A + R_ATOMS - 1) / R_ATOMS;
}
static float clamp_q(float q) { if (q < 0.05f) q = 0.05f; if (q > 4096.0f) q = 4096.0f; return q; }

typedef struct { u32 len; u32 off; u8 zero; } enc_atom;  /* off into payload pool */

/* Encode every atom of every LOD (payload pool in LOD, raster order). */
static int encode_all(u8 **lods, const u32 (*ld)[3], int nlod, const float *qs,
                      enc_atom **atoms_out, u8 **pool_out, u64 *counts /* per lod */) {
    u64 total = 0;
    for (int l = 0; l < nlod; ++l) {
        u64 c = (u64)((ld[l][0] + 31) / 32) * ((ld[l][1] + 31) / 32) * ((ld[l][2] + 31) / 32);
        counts[l] = c; total += c;
    }
    enc_atom *at = (enc_atom *)calloc(total, sizeof(enc_atom));
    u8 **pay = (u8 **)calloc(total, sizeof(u8 *));
    int fail = 0;
    u64 base = 0;
    for (int l = 0; l < nlod; ++l) {
        u32 acx = (ld[l][0] + 31) / 32, acy = (ld[l][1] + 31) / 32, acz = (ld[l][2] + 31) / 32;
        #pragma omp parallel for schedule(dynamic)
        for (i64 i = 0; i < (i64)counts[l]; ++i) {
            u32 ax = (u32)(i % acx), ay = (u32)((i / acx) % acy), az = (u32)(i / ((u64)acx * acy));
            (void)acz;
            u8 atom[A3]; u8 *buf = (u8 *)malloc(VCO_RC_CAP + 1);
            vco_gather_atom(lods[l], ld[l][0], ld[l][1], ld[l][2], az, ay, ax, atom);
            int z = 0;
            long n = vco_encode_atom(atom, qs[l], buf, VCO_RC_CAP + 1, &z);
            if (n < 0) { fail = 1; free(buf); continue; }
            at[base + i].zero = (u8)z; at[base + i].len = (u32)n;
            pay[base + i] = buf;
        }
        base += counts[l];
    }
    u64 pool_len = 0;
    for (u64 i = 0; i < total; ++i) pool_len += at[i].len;
    u8 *pool = (u8 *)malloc(pool_len ? pool_len : 1);
    u64 cur = 0;
    for (u64 i = 0; i < total; ++i) {
        at[i].off = (u32)cur;
        if (at[i].len) memcpy(pool + cur, pay[i], at[i].len);
        cur += at[i].len; free(pay[i]);
    }
    free(pay);
    if (fail || pool_len > 0xFFFFFFFFull) { free(at); free(pool); return VCO_ERR_RANGE; }
    *atoms_out = at; *pool_out = pool;
    return VCO_OK;
}

int vco_encode_q(const u8 *vol, u32 nx, u32 ny, u32 nz, float target_ratio,
                 const float base_q_in[VCO_NLOD], u8 **out, size_t *out_len) {
    tables();
    if (!vol || !out || !out_len) return VCO_ERR_RANGE;
    if (!nx || !ny || !nz || nx > (1u << 20) || ny > (1u << 20) || nz > (1u << 20) || !(target_ratio >= 1.0f))
        return VCO_ERR_RANGE;
    int nlod = vco_nlod(nx, ny, nz);
    float qs[VCO_NLOD] = {0};
    for (int l = 0; l < nlod; ++l) qs[l] = clamp_q(base_q_in[l]);
    /* pyramid */
    u8 *lods[VCO_NLOD]; u32 ld[VCO_NLOD][3];
    lods[0] = (u8 *)vol; ld[0][0] = nx; ld[0][1] = ny; ld[0][2] = nz;
    for (int l = 1; l < nlod; ++l) {
        vco_lod_dims(nx, ny, nz, l, &ld[l][0], &ld[l][1], &ld[l][2]);
        lods[l] = (u8 *)malloc((size_t)ld[l][0] * ld[l][1] * ld[l][2]);
        vco_downsample2x(lods[l - 1], ld[l - 1][0], ld[l - 1][1], ld[l - 1][2], lods[l]);
    }
    enc_atom *at = NULL; u8 *pool = NULL; u64 counts[VCO_NLOD] = {0};
    int st = encode_all(lods, (const u32 (*)[3])ld, nlod, qs, &at, &pool, counts);
    for (int l = 1; l < nlod; ++l) free(lods[l]);
    if (st != VCO_OK) return st;

    /* region table, LOD-major raster (vc.c:752-761) */
    u64 prefix[VCO_NLOD + 1], nreg = 0, abase[VCO_NLOD], ab = 0;
    for (int l = 0; l < nlod; ++l) {
        u32 rz, ry, rx; region_grid(nx, ny, nz, l, &rz, &ry, &rx);
        prefix[l] = nreg; nreg += (u64)rz * ry * rx; abase[l] = ab; ab += counts[l];
    }
    prefix[nlod] = nreg;
    u64 *blen = (u64 *)calloc(nreg, sizeof(u64));
    u8 *cls = (u8 *)calloc(nreg, 1);         /* 0 absent, 1 zero, 2 data */
    u64 *boff = (u64 *)calloc(nreg, sizeof(u64));
    for (int l = 0; l < nlod; ++l) {
        u32 acx = (ld[l][0] + 31) / 32, acy = (ld[l][1] + 31) / 32, acz = (ld[l][2] + 31) / 32;
        u32 rz, ry, rx; region_grid(nx, ny, nz, l, &rz, &ry, &rx);
        for (u32 az = 0; az < acz; ++az) for (u32 ay = 0; ay < acy; ++ay) for (u32 ax = 0; ax < acx; ++ax) {
            u64 ri = prefix[l] + ((u64)(az / 32) * ry + ay / 32) * rx + ax / 32;
            const enc_atom *e = &at[abase[l] + ((u64)az * acy + ay) * acx + ax];
            if (e->zero) { if (cls[ri] < 1) cls[ri] = 1; }
            else { cls[ri] = 2; blen[ri] += e->len; }
        }
    }
    u64 dir_off = HDR;
    u64 cursor = (dir_off + nreg * 16 + 4095) & ~4095ull;         /* vc_repack.c:213-214 */
    /* coarse-LOD-first blob order (vc_repack.c:216-226, 237-247) */
    int bad = 0;
    for (int l = nlod - 1; l >= 0; --l) {
        for (u64 ri = prefix[l]; ri < prefix[l + 1]; ++ri) {
            if (cls[ri] != 2) continue;
            blen[ri] += L2_BYTES;
            if (blen[ri] > 0xFFFFFFFFull) bad = 1;
            boff[ri] = cursor; cursor += blen[ri];
        }
    }
    if (bad) { free(at); free(pool); free(blen); free(cls); free(boff); return VCO_ERR_RANGE; }
    u64 total_len = cursor;
    u8 *o = (u8 *)calloc(total_len, 1);
    if (!o) { free(at); free(pool); free(blen); free(cls); free(boff); return VCO_ERR_OOM; }
    /* header, vc_repack.c:256-265 */
    w32(o + 0, MAGIC); w32(o + 4, 2); w32(o + 8, A); w32(o + 16, R_ATOMS); w32(o + 20, (u32)nlod);
    w32(o + 24, nx); w32(o + 28, ny); w32(o + 32, nz); wf(o + 36, target_ratio);
    w64(o + 40, dir_off); w64(o + 48, nreg); w64(o + 56, total_len);
    for (int l = 0; l < VCO_NLOD; ++l) wf(o + 72 + 4 * l, l < nlod ? qs[l] : 0.0f);
    for (u64 ri = 0; ri < nreg; ++ri) {
        u8 *e = o + dir_off + ri * 16;
        if (cls[ri] == 2) { w64(e, boff[ri]); w32(e + 8, (u32)blen[ri]); w32(e + 12, 2); }
        else if (cls[ri] == 1) w32(e + 12, 1);
    }
    /* blobs: slot array then payloads in slot order (vc_repack.c:136-162) */
    for (int l = 0; l < nlod; ++l) {
        u32 acx = (ld[l][0] + 31) / 32, acy = (ld[l][1] + 31) / 32, acz = (ld[l][2] + 31) / 32;
        u32 rz, ry, rx; region_grid(nx, ny, nz, l, &rz, &ry, &rx);
        for (u32 qz = 0; qz < rz; ++qz) for (u32 qy = 0; qy < ry; ++qy) for (u32 qx = 0; qx < rx; ++qx) {
            u64 ri = prefix[l] + ((u64)qz * ry + qy) * rx + qx;
            if (cls[ri] != 2) continue;
            u8 *slots = o + boff[ri];
            u64 pc = boff[ri] + L2_BYTES;
            for (u32 dz = 0; dz < R_ATOMS; ++dz) for (u32 dy = 0; dy < R_ATOMS; ++dy) for (u32 dx = 0; dx < R_ATOMS; ++dx) {
                u32 az = qz * 32 + dz, ay = qy * 32 + dy, ax = qx * 32 + dx;
                if (az >= acz || ay >= acy || ax >= acx) continue;      /* ABSENT slot */
                const enc_atom *e = &at[abase[l] + ((u64)az * acy + ay) * acx + ax];
                u8 *s = slots + ((u64)(dz * 32 + dy) * 32 + dx) * SLOT;
                if (e->zero) { s[12] = 2; continue; }
                w64(s, pc); w32(s + 8, e->len); s[12] = 1; s[13] = pool[e->off];
                memcpy(o + pc, pool + e->off, e->len);
                pc += e->len;
            }
        }
    }
    free(at); free(pool); free(blen); free(cls); free(boff);
    *out = o; *out_len = (size_t)total_len;
    return VCO_OK;
}

/* Rate control (build-defined; the reference has none, docs/SPEC.md:172-197,
 * docs/DO_NOT_IMPLEMENT.md:40-45 prescribe real measure-bisection on a small
 * calibration sample).  Sample = LOD0 atoms on a lattice of at most 8 indices
 * per axis (all indices when the axis has <= 8 atoms; else floor((2j+1)n/16)).
 * Ratio over the sample = in-bounds voxels of non-zero atoms / payload bytes.
 * meets(q) := voxels >= target * bytes (double).  Geometric bisection with
 * sqrtf (correctly rounded) over [0.05, 4096], 16 steps, returns the upper
 * bracket (the finest q found that meets the target). */
static int lattice(u32 n, u32 *idx) {
    if (n <= 8) { for (u32 i = 0; i < n; ++i) idx[i] = i; return (int)n; }
    for (u32 j = 0; j < 8; ++j) idx[j] = (u32)(((u64)(2 * j + 1) * n) / 16);
    return 8;
}
typedef struct { i32 *coef; u32 vox; } sample_atom;

static int sample_meets(const sample_atom *s, int ns, float q, float target) {
    u64 bytes = 0, vox = 0;
    int ovf = 0;
    #pragma omp parallel for reduction(+:bytes, vox) reduction(|:ovf) schedule(dynamic)
    for (int i = 0; i < ns; ++i) {
        i16 *lv = (i16 *)malloc(A3 * sizeof(i16));
        u8 *buf = (u8 *)malloc(VCO_RC_CAP);
        vco_quantize(s[i].coef, q, lv);
        long n = vco_entropy_encode(lv, buf, VCO_RC_CAP);
        if (n < 0) ovf = 1; else { bytes += (u64)n + 1; vox += s[i].vox; }
        free(lv); free(buf);
    }
    if (ovf) return 0;
    return (double)vox >= (double)target * (double)bytes;
}

float vco_rate_control(const u8 *vol, u32 nx, u32 ny, u32 nz, float target) {
    tables();
    u32 ix[8], iy[8], iz[8];
    u32 acx = (nx + 31) / 32, acy = (ny + 31) / 32, acz = (nz + 31) / 32;
    int nx_ = lattice(acx, ix), ny_ = lattice(acy, iy), nz_ = lattice(acz, iz);
    sample_atom s[512]; int ns = 0;
    for (int a = 0; a < nz_; ++a) for (int b = 0; b < ny_; ++b) for (int c = 0; c < nx_; ++c) {
        u8 atom[A3];
        vco_gather_atom(vol, nx, ny, nz, iz[a], iy[b], ix[c], atom);
        u32 sum = 0; for (int i = 0; i < A3; ++i) sum += atom[i];
        if (!sum) continue;
        i32 dc = (i32)((sum + A3 / 2) / A3);
        s[ns].coef = (i32 *)malloc(A3 * sizeof(i32));
        vco_fdct(atom, dc, s[ns].coef);
        u32 ex = nx - ix[c] * 32, ey = ny - iy[b] * 32, ez = nz - iz[a] * 32;
        s[ns].vox = (ex < 32 ? ex : 32) * (ey < 32 ? ey : 32) * (ez < 32 ? ez : 32);
        ns++;
    }
    float q0;
    if (ns == 0) q0 = 1.0f;
    else {
        float lo = 0.05f, hi = 4096.0f;
        if (sample_meets(s, ns, lo, target)) q0 = lo;
        else if (!sample_meets(s, ns, hi, target)) q0 = hi;
        else {
            for (int it = 0; it < 16; ++it) {
                float mid = sqrtf(lo * hi);
                if (sample_meets(s, ns, mid, target)) hi = mid; else lo = mid;
            }
            q0 = hi;
        }
    }
    for (int i = 0; i < ns; ++i) free(s[i].coef);
    return q0;
}

/* q[l] = q0 * 0.6^l by iterated float multiply (tools/vc_export_stream.c:704-716,
 * falloff 0.6 as tests/test_realdata.c:71), clamped like vc.c:1072-1073. */
void vco_lod_schedule(float q0, int nlod, float qs[VCO_NLOD]) {
    float q = q0;
    for (int l = 0; l < VCO_NLOD; ++l) {
        qs[l] = l < nlod ? clamp_q(q) : 0.0f;
        q *= 0.6f;
    }
}

int vco_encode(const u8 *vol, u32 nx, u32 ny, u32 nz, float target_ratio, u8 **out, size_t *out_len) {
    if (!vol || !out || !out_len) return VCO_ERR_RANGE;
    if (!nx || !ny || !nz || nx > (1u << 20) || ny > (1u << 20) || nz > (1u << 20) || !(target_ratio >= 1.0f))
        return VCO_ERR_RANGE;
    float q0 = vco_rate_control(vol, nx, ny, nz, target_ratio);
    float qs[VCO_NLOD];
    vco_lod_schedule(q0, vco_nlod(nx, ny, nz), qs);
    return vco_encode_q(vol, nx, ny, nz, target_ratio, qs, out, out_len);
}

/* ===================================================================== */
/* Reader, vc.c:1250-1421                                                */
/* ===================================================================== */
struct vco_archive {
    const u8 *buf; u64 len; int version, nlod; u32 nx, ny, nz;
    u64 l1_off, l1_cap, dir_off, prefix[VCO_NLOD + 1];
    float base_q[VCO_NLOD];
};

static u64 mix64(u64 x) {                                 /* vc.c:721-726 */
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

vco_archive *vco_open(const u8 *buf, size_t len) {      /* vc.c:1285-1323 */
    tables();
    if (!buf || len < HDR) return NULL;
    if (r32(buf) != MAGIC || r32(buf + 8) != A || r32(buf + 16) != R_ATOMS) return NULL;
    u32 ver = r32(buf + 4), nlod = r32(buf + 20);
    if (nlod == 0 || nlod > VCO_NLOD) return NULL;
    vco_archive *a = (vco_archive *)calloc(1, sizeof *a);
    a->buf = buf; a->len = len; a->nlod = (int)nlod;
    a->nx = r32(buf + 24); a->ny = r32(buf + 28); a->nz = r32(buf + 32);
    for (int l = 0; l < VCO_NLOD; ++l) a->base_q[l] = rf(buf + 72 + 4 * l);
    if (ver == 2) {
        a->version = 2; a->dir_off = r64(buf + 40);
        u64 cnt = r64(buf + 48), acc = 0;
        if (a->dir_off > len || cnt > (len - a->dir_off) / 16) { free(a); return NULL; }
        for (int l = 0; l < a->nlod; ++l) {
            a->prefix[l] = acc; u32 rz, ry, rx; region_grid(a->nx, a->ny, a->nz, l, &rz, &ry, &rx);
            acc += (u64)rz * ry * rx;
        }
        a->prefix[a->nlod] = acc;
        if (acc != cnt) { free(a); return NULL; }
        return a;
    }
    if (ver != 1 || r32(buf + 12) != 1) { free(a); return NULL; }
    a->version = 1; a->l1_off = r64(buf + 56); a->l1_cap = r64(buf + 40);
    if (a->l1_cap == 0 || (a->l1_cap & (a->l1_cap - 1))) { free(a); return NULL; }
    if (a->l1_off > len || a->l1_cap > (len - a->l1_off) / 16) { free(a); return NULL; }
    return a;
}
void vco_close(vco_archive *a) { free(a); }
int vco_nlod_of(const vco_archive *a) { return a->nlod; }

/* flags: 0 absent, 1 present, 2 zero (slot flag values, vc.c:573-575) */
static int resolve(const vco_archive *a, int lod, u32 az, u32 ay, u32 ax, u64 *off, u32 *len) {
    u64 blk;
    if (a->version == 2) {                                /* vc.c:1376-1391 */
        u32 rz, ry, rx; region_grid(a->nx, a->ny, a->nz, lod, &rz, &ry, &rx);
        if (az / 32 >= rz || ay / 32 >= ry || ax / 32 >= rx) return 0;
        u64 ri = a->prefix[lod] + ((u64)(az / 32) * ry + ay / 32) * rx + ax / 32;
        if (ri >= a->prefix[a->nlod]) return 0;
        const u8 *e = a->buf + a->dir_off + ri * 16;
        u32 f = r32(e + 12);
        if (f == 0) return 0;
        if (f == 1) return 2;
        blk = r64(e);
    } else {                                              /* vc.c:1349-1360, 1396-1399 */
        u64 key = ((u64)(lod & 7) << 45) | ((u64)((az / 32) & 0x7FFF) << 30)
                | ((u64)((ay / 32) & 0x7FFF) << 15) | (u64)((ax / 32) & 0x7FFF);
        u64 mask = a->l1_cap - 1, i = mix64(key) & mask;
        for (;;) {
            const u8 *e = a->buf + a->l1_off + i * 16;
            u64 br = r64(e + 8);
            if (br == 0) return 0;
            if (r64(e) == key) { blk = br; break; }
            i = (i + 1) & mask;
        }
        if (blk == 1) return 2;
    }
    if (blk + L2_BYTES > a->len) return 0;                /* corrupt guard */
    const u8 *s = a->buf + blk + ((u64)((az & 31) * 32 + (ay & 31)) * 32 + (ax & 31)) * SLOT;
    u8 f = s[12];
    if (f == 1) { *off = r64(s); *len = r32(s + 8); }     /* vc.c:1365-1373 */
    return f;
}

int vco_coverage(const vco_archive *a, int lod, u32 az, u32 ay, u32 ax, u64 *off, u32 *len) {
    u64 o = 0; u32 l = 0;
    if (off) *off = 0;
    if (len) *len = 0;
    if (!a || lod < 0 || lod >= a->nlod) return 0;
    int f = resolve(a, lod, az, ay, ax, &o, &l);
    if (f == 1) { if (off) *off = o; if (len) *len = l; return 2; }
    return f == 2 ? 1 : 0;
}

/* Container rewrite into the canonical v2 layout, serially
 * (tools/vc_repack.c:184-291): regions coarse-LOD-first; pass 1 classifies
 * each region from its atoms' coverage and sizes its blob (:96-121); a
 * prefix sum places the blobs after the 4 KiB-aligned directory (:244-256);
 * pass 2 writes each data region's slot array, payloads in slot order with
 * dc = first payload byte (:131-162); header fields carried over (:272-281).
 * A present payload range outside the archive, which the tool copies
 * unchecked, is VCO_ERR_FORMAT here. */
int vco_repack(const u8 *buf, size_t len, u8 **out, size_t *out_len) {
    if (!out || !out_len) return VCO_ERR_RANGE;
    *out = NULL; *out_len = 0;
    vco_archive *a = vco_open(buf, len);
    if (!a) return VCO_ERR_FORMAT;
    u64 prefix[VCO_NLOD + 1], nreg = 0;                            /* directory index base per LOD */
    for (int l = 0; l < a->nlod; ++l) {
        u32 nrz, nry, nrx; region_grid(a->nx, a->ny, a->nz, l, &nrz, &nry, &nrx);
        prefix[l] = nreg;
        nreg += (u64)nrz * nry * nrx;
    }
    prefix[a->nlod] = nreg;
    u64 *blen = (u64 *)calloc(nreg ? nreg : 1, 8), *boff = (u64 *)calloc(nreg ? nreg : 1, 8);
    u32 *rflag = (u32 *)calloc(nreg ? nreg : 1, 4);
    u64 *order = (u64 *)calloc(nreg ? nreg : 1, 8);
    int *olod = (int *)calloc(nreg ? nreg : 1, sizeof(int));
    u32 *orz = (u32 *)calloc(nreg ? nreg : 1, 4), *ory = (u32 *)calloc(nreg ? nreg : 1, 4),
        *orx = (u32 *)calloc(nreg ? nreg : 1, 4);
    int st = VCO_OK;
    u64 ni = 0;
    for (int l = a->nlod - 1; l >= 0; --l) {
        u32 nrz, nry, nrx; region_grid(a->nx, a->ny, a->nz, l, &nrz, &nry, &nrx);
        for (u32 rz = 0; rz < nrz; ++rz) for (u32 ry = 0; ry < nry; ++ry) for (u32 rx = 0; rx < nrx; ++rx) {
            order[ni] = prefix[l] + ((u64)rz * nry + ry) * nrx + rx;
            olod[ni] = l; orz[ni] = rz; ory[ni] = ry; orx[ni] = rx;
            ni++;
        }
    }
    for (u64 i = 0; i < ni && st == VCO_OK; ++i) {                 /* pass 1 */
        u64 pay = 0; int anyp = 0, anyz = 0;
        for (u32 dz = 0; dz < 32; ++dz) for (u32 dy = 0; dy < 32; ++dy) for (u32 dx = 0; dx < 32; ++dx) {
            u64 o; u32 n;
            int c = vco_coverage(a, olod[i], orz[i] * 32 + dz, ory[i] * 32 + dy, orx[i] * 32 + dx, &o, &n);
            if (c == 2) {
                if (o >= len || n > len - o) st = VCO_ERR_FORMAT;
                anyp = 1; pay += n;
            } else if (c == 1) {
                anyz = 1;
            }
        }
        rflag[order[i]] = anyp ? 2u : anyz ? 1u : 0u;
        if (anyp) blen[order[i]] = L2_BYTES + pay;
    }
    const u64 dir_off = HDR, blobs = (dir_off + nreg * 16 + 4095) & ~4095ull;
    u64 cur = blobs;
    for (u64 i = 0; i < ni; ++i)
        if (rflag[order[i]] == 2) { boff[order[i]] = cur; cur += blen[order[i]]; }
    u8 *o8 = st == VCO_OK ? (u8 *)calloc(cur, 1) : NULL;
    if (st == VCO_OK && !o8) st = VCO_ERR_OOM;
    for (u64 i = 0; i < ni && st == VCO_OK; ++i) {                 /* pass 2 */
        const u64 ri = order[i];
        u8 *e = o8 + dir_off + ri * 16;
        w32(e + 12, rflag[ri]);
        if (rflag[ri] != 2) continue;
        w64(e, boff[ri]); w32(e + 8, (u32)blen[ri]);
        u64 pc = boff[ri] + L2_BYTES;
        for (u32 dz = 0; dz < 32; ++dz) for (u32 dy = 0; dy < 32; ++dy) for (u32 dx = 0; dx < 32; ++dx) {
            u64 so; u32 n;
            int c = vco_coverage(a, olod[i], orz[i] * 32 + dz, ory[i] * 32 + dy, orx[i] * 32 + dx, &so, &n);
            u8 *slot = o8 + boff[ri] + ((u64)(dz * 32 + dy) * 32 + dx) * SLOT;
            if (c == 2) {
                w64(slot, pc); w32(slot + 8, n); slot[12] = 1; slot[13] = buf[so];
                memcpy(o8 + pc, buf + so, n);
                pc += n;
            } else if (c == 1) {
                slot[12] = 2;
            }
        }
    }
    if (st == VCO_OK) {
        w32(o8, MAGIC); w32(o8 + 4, 2); w32(o8 + 8, A); w32(o8 + 16, R_ATOMS); w32(o8 + 20, (u32)a->nlod);
        memcpy(o8 + 24, buf + 24, 12);                 /* nx, ny, nz */
        memcpy(o8 + 36, buf + 36, 4);                  /* target ratio */
        memcpy(o8 + 72, buf + 72, 4 * VCO_NLOD);       /* base_q */
        w64(o8 + 40, dir_off); w64(o8 + 48, nreg); w64(o8 + 56, cur);
        *out = o8; *out_len = (size_t)cur;
    } else {
        free(o8);
    }
    free(blen); free(boff); free(rflag); free(order); free(olod); free(orz); free(ory); free(orx);
    vco_close(a);
    return st;
}

int vco_decode_atom(vco_archive *a, int lod, int ax, int ay, int az, u8 out[A3]) {  /* vc.c:1404-1421 */
    if (!a || lod < 0 || lod >= a->nlod || !out) return VCO_ERR_RANGE;
    u32 x, y, z; vco_lod_dims(a->nx, a->ny, a->nz, lod, &x, &y, &z);
    if ((u32)ax >= (x + 31) / 32 || (u32)ay >= (y + 31) / 32 || (u32)az >= (z + 31) / 32) return VCO_ERR_RANGE;
    u64 off = 0; u32 len = 0;
    int f = resolve(a, lod, (u32)az, (u32)ay, (u32)ax, &off, &len);
    if (f == 1) {
        if (off == ~0ull || len < 1 || off + len > a->len) { memset(out, 0, A3); return VCO_ERR_FORMAT; }
        return vco_decode_payload(a->buf + off, len, a->base_q[lod], out);
    }
    memset(out, 0, A3);
    return VCO_OK;
}

/* Region decode.  Same atom set and clipped copy as vc.c:1535-1555; on a
 * failing atom this keeps going (failing atoms read as zeros) and returns the
 * first error, where the reference stops at the first error. */
int vco_decode_region(vco_archive *a, int lod, const u32 b[6], u8 *out) {
    if (!a || lod < 0 || lod >= a->nlod || !out) return VCO_ERR_RANGE;
    u32 nx, ny, nz; vco_lod_dims(a->nx, a->ny, a->nz, lod, &nx, &ny, &nz);
    if (b[3] > nx || b[4] > ny || b[5] > nz) return VCO_ERR_RANGE;
    if (b[0] >= b[3] || b[1] >= b[4] || b[2] >= b[5]) return VCO_ERR_RANGE;
    u32 ox = b[3] - b[0], oy = b[4] - b[1];
    const u32 ax0 = b[0] / 32, ay0 = b[1] / 32, az0 = b[2] / 32;
    const u32 nax = (b[3] - 1) / 32 - ax0 + 1, nay = (b[4] - 1) / 32 - ay0 + 1, naz = (b[5] - 1) / 32 - az0 + 1;
    const i64 natoms = (i64)nax * nay * naz;
    int st = VCO_OK;
    /* atoms are independent and write disjoint parts of out: decoded in
     * parallel; every failing atom is VCO_ERR_FORMAT, so the first error is
     * any error */
    #pragma omp parallel reduction(max:st)
    {
        u8 *atom = (u8 *)malloc(A3);
        #pragma omp for schedule(dynamic, 4)
        for (i64 k = 0; k < natoms; ++k) {
            const u32 ax = ax0 + (u32)(k % nax), ay = ay0 + (u32)((k / nax) % nay), az = az0 + (u32)(k / ((i64)nax * nay));
            int s = vco_decode_atom(a, lod, (int)ax, (int)ay, (int)az, atom);
            if (s > st) st = s;
            for (u32 z = 0; z < 32; ++z) { u32 vz = az * 32 + z; if (vz < b[2] || vz >= b[5]) continue;
            for (u32 y = 0; y < 32; ++y) { u32 vy = ay * 32 + y; if (vy < b[1] || vy >= b[4]) continue;
            for (u32 x = 0; x < 32; ++x) { u32 vx = ax * 32 + x; if (vx < b[0] || vx >= b[3]) continue;
                out[((size_t)(vz - b[2]) * oy + (vy - b[1])) * ox + (vx - b[0])] = atom[(z * 32 + y) * 32 + x];
            } } }
        }
        free(atom);
    }
    return st;
}

int vco_decode_lod(vco_archive *a, int lod, u8 *out, u32 dims[3]) {
    if (!a || lod < 0 || lod >= a->nlod || !out) return VCO_ERR_RANGE;
    u32 b[6] = {0, 0, 0, 0, 0, 0};
    vco_lod_dims(a->nx, a->ny, a->nz, lod, &b[3], &b[4], &b[5]);
    if (dims) { dims[0] = b[3]; dims[1] = b[4]; dims[2] = b[5]; }
    return vco_decode_region(a, lod, b, out);
}

double vco_decode_atoms_omp(vco_archive *a, int lod, const u32 *zyx, size_t n, int nthreads, u8 *out) {
    struct timespec t0, t1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    clock_gettime(CLOCK_MONOTONIC, &t0);
    #pragma omp parallel
    {
        u8 *tmp = (u8 *)malloc(A3);
        #pragma omp for schedule(dynamic, 4)
        for (i64 i = 0; i < (i64)n; ++i)
            vco_decode_atom(a, lod, (int)zyx[3 * i + 2], (int)zyx[3 * i + 1], (int)zyx[3 * i],
                            out ? out + (size_t)i * A3 : tmp);
        free(tmp);
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* ===================================================================== */
/* Synthetic scroll-like volume (build-defined, integer-exact so the HIP    */
/* generator reproduces it byte for byte).                                  */
/* ===================================================================== */
static u64 sm64_next(u64 *st) {
    u64 z = (*st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static u32 sm_u(u64 *st, u32 n) { return n ? (u32)((sm64_next(st) >> 11) % n) : 0; }
/* Q14 sine of a 16-bit angle (full turn = 65536), Bhaskara I's rational form. */
static i32 isin16(u32 th) {
    i64 h = th & 32767, y = h * (32768 - h);
    i64 s = (16 * y * 16384) / (5ll * (1ll << 30) - 4 * y);
    return (th & 32768) ? -(i32)s : (i32)s;
}
/* sum of four 16-bit uniforms from one 64-bit hash, centered at 0 */
static i64 irwin4(u64 r) {
    return (i64)((r & 0xFFFF) + ((r >> 16) & 0xFFFF) + ((r >> 32) & 0xFFFF) + (r >> 48)) - 131070;
}
static u64 vhash(u64 seed, u32 stream, u64 idx) {
    u64 s = seed ^ ((u64)stream << 56) ^ (idx * 0xD1B54A32D192ED03ull);
    return sm64_next(&s);
}
typedef struct { i32 axis, c0, tu, tv, a1, a2, w1, w2, p1, p2, half, inten; } sheet_t;

static void make_sheets(u64 seed, u32 nx, u32 ny, u32 nz, int ns, sheet_t *sh) {
    u64 st = seed ^ 0x5EEDC0DE5EEDC0DEull;
    u32 M = nx > ny ? nx : ny; if (nz > M) M = nz;
    u32 dim[3] = {nz, ny, nx};
    for (int s = 0; s < ns; ++s) {
        sheet_t *p = &sh[s];
        p->axis = (i32)sm_u(&st, 3);
        p->c0 = (i32)sm_u(&st, dim[p->axis] * 256u);
        p->tu = (i32)sm_u(&st, 32769) - 16384;
        p->tv = (i32)sm_u(&st, 32769) - 16384;
        p->a1 = (i32)sm_u(&st, M * 256u / 16u + 1);
        p->a2 = (i32)sm_u(&st, M * 256u / 32u + 1);
        p->w1 = (i32)(65536u / (1u + M / 4u + sm_u(&st, M)));
        p->w2 = (i32)(65536u / (1u + M / 4u + sm_u(&st, M)));
        p->p1 = (i32)sm_u(&st, 65536);
        p->p2 = (i32)sm_u(&st, 65536);
        p->half = (i32)((768u + sm_u(&st, 769)) / 2u);
        p->inten = 110 + (i32)sm_u(&st, 81);
    }
}

void vco_synth(u64 seed, u32 nx, u32 ny, u32 nz, int ns, u8 *out) {
    size_t N = (size_t)nx * ny * nz;
    sheet_t *sh = (sheet_t *)malloc(sizeof(sheet_t) * (ns > 0 ? ns : 1));
    make_sheets(seed, nx, ny, nz, ns, sh);
    u16 *a = (u16 *)malloc(N * 2), *b = (u16 *)malloc(N * 2);
    /* base: background N(12,6) (Irwin-Hall of 4 uniforms) max sheet intensity */
    #pragma omp parallel for schedule(static)
    for (i64 z = 0; z < (i64)nz; ++z) for (u32 y = 0; y < ny; ++y) for (u32 x = 0; x < nx; ++x) {
        u64 idx = ((u64)z * ny + y) * nx + x;
        i64 bg = 12 * 256 + ((irwin4(vhash(seed, 0, idx)) * 174353943ll + (1ll << 31)) >> 32);
        if (bg < 0) bg = 0;
        if (bg > 255 * 256) bg = 255 * 256;
        i32 v = (i32)bg;
        u32 c[3] = {(u32)z, y, x}, dm[3] = {nz, ny, nx};
        for (int s = 0; s < ns; ++s) {
            const sheet_t *p = &sh[s];
            int ua = p->axis == 0 ? 1 : 0, va = p->axis == 2 ? 1 : 2;
            i64 u = c[ua], w = c[va];
            i64 surf = (i64)p->c0 + (((i64)p->tu * (u * 256 - (i64)dm[ua] * 128)) >> 16)
                     + (((i64)p->tv * (w * 256 - (i64)dm[va] * 128)) >> 16)
                     + (((i64)p->a1 * isin16((u32)(p->p1 + u * p->w1) & 0xFFFF)) >> 14)
                     + (((i64)p->a2 * isin16((u32)(p->p2 + w * p->w2) & 0xFFFF)) >> 14);
            i64 d = (i64)c[p->axis] * 256 + 128 - surf;
            if (d < 0) d = -d;
            if (d <= p->half && p->inten * 256 > v) v = p->inten * 256;
        }
        a[idx] = (u16)v;
    }
    /* separable 7-tap Gaussian sigma=1 (Q12 taps), edge-replicate, x then y then z */
    static const u32 W[7] = {18, 221, 991, 1636, 991, 221, 18};
    u32 dmx[3] = {nx, ny, nz};
    for (int ax = 0; ax < 3; ++ax) {
        size_t stride = ax == 0 ? 1 : ax == 1 ? nx : (size_t)nx * ny;
        i64 n = dmx[ax];
        #pragma omp parallel for schedule(static)
        for (i64 i = 0; i < (i64)N; ++i) {
            i64 pos = (i64)((i / (i64)stride) % n);
            u32 acc = 2048;
            for (int k = 0; k < 7; ++k) {
                i64 q = pos + k - 3; if (q < 0) q = 0; if (q >= n) q = n - 1;
                acc += W[k] * a[i + (q - pos) * (i64)stride];
            }
            b[i] = (u16)(acc >> 12);
        }
        u16 *t = a; a = b; b = t;
    }
    /* additive N(0,8) noise, round, clip */
    #pragma omp parallel for schedule(static)
    for (i64 i = 0; i < (i64)N; ++i) {
        i64 nq = (irwin4(vhash(seed, 1, (u64)i)) * 232471924ll + (1ll << 31)) >> 32;
        i64 t = (i64)a[i] + nq + 128;
        out[i] = (u8)(t < 0 ? 0 : (t >> 8) > 255 ? 255 : (t >> 8));
    }
    free(a); free(b); free(sh);
}
