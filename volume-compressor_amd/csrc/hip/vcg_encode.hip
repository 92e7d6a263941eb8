// vcg_encode.hip — batched atom encode on gfx950.
//
//   k_downsample : LOD l -> l+1, integer rounded mean of the in-bounds voxels
//                  of each 2x2x2 cell (src/vc/vc.c:1562-1577).
//   k_fdct       : one 256-thread workgroup per atom.  Gathers the 32^3 atom
//                  (zero padded), sums it (zero atom -> no payload, dc =
//                  (sum+16384)/32768, src/vc/vc.c:599-604), runs the three
//                  forward partial-butterfly passes along x, y, z in LDS
//                  (src/vc/vc.c:142-171, 212-218) and either stores raster
//                  coefficients (rate-control sample) or quantizes them
//                  (src/vc/vc.c:293-306, strict IEEE in source order) and
//                  stores the levels in scan order plus the EOB.
//   k_quant      : re-quantize stored coefficients at a new base_q.
//   k_rc_encode  : one lane per atom.  Adaptive binary range coder with LZMA
//                  carry and the coefficient context coder
//                  (src/vc/vc.c:360-401, 471-544), payload = [dc][bytes].
//   k_compact / k_assemble : pack payloads, then write the archive image.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>
#include <vector>
#include "vcg_dev.hpp"
#include "vcg_xform.hpp"

#define FDCT_THREADS 256
#define ENC_BATCH 32768u

struct enc_src {
    const uint8_t *lod[VCG_NLOD];
    uint32_t nx[VCG_NLOD], ny[VCG_NLOD], nz[VCG_NLOD];
    uint32_t acx[VCG_NLOD], acy[VCG_NLOD];
    uint64_t prefix[VCG_NLOD + 1];    // atom index base per LOD
    float q[VCG_NLOD];
    int nlod;
    const uint32_t *list;             // optional explicit LOD0 atoms (z,y,x)
    uint64_t base;                    // first global atom of this launch
    uint32_t n;
};

struct enc_item {
    uint32_t eob;
    uint32_t dc;
    uint32_t zero;
    uint32_t lod;
};

__device__ __forceinline__ void enc_atom_of(const enc_src &s, uint64_t gi, int &lod, uint32_t &ax,
                                            uint32_t &ay, uint32_t &az) {
    if (s.list) {
        lod = 0;
        az = s.list[3 * gi]; ay = s.list[3 * gi + 1]; ax = s.list[3 * gi + 2];
        return;
    }
    int l = 0;
    while (l + 1 < s.nlod && gi >= s.prefix[l + 1]) ++l;
    uint64_t r = gi - s.prefix[l];
    lod = l;
    ax = (uint32_t)(r % s.acx[l]);
    ay = (uint32_t)((r / s.acx[l]) % s.acy[l]);
    az = (uint32_t)(r / ((uint64_t)s.acx[l] * s.acy[l]));
}

// ---------------------------------------------------------------------------
__global__ void k_downsample(const uint8_t *__restrict__ in, uint32_t nx, uint32_t ny, uint32_t nz,
                             uint8_t *__restrict__ out) {
    const uint32_t ox = (nx + 1) / 2, oy = (ny + 1) / 2, oz = (nz + 1) / 2;
    const uint64_t total = (uint64_t)ox * oy * oz;
    for (uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; o < total;
         o += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t x = (uint32_t)(o % ox), y = (uint32_t)((o / ox) % oy), z = (uint32_t)(o / ((uint64_t)ox * oy));
        uint32_t acc = 0, n = 0;
        for (uint32_t dz = 0; dz < 2; ++dz) {
            uint32_t zz = 2 * z + dz; if (zz >= nz) continue;
            for (uint32_t dy = 0; dy < 2; ++dy) {
                uint32_t yy = 2 * y + dy; if (yy >= ny) continue;
                const uint8_t *row = in + ((uint64_t)zz * ny + yy) * nx;
                for (uint32_t dx = 0; dx < 2; ++dx) {
                    uint32_t xx = 2 * x + dx; if (xx >= nx) continue;
                    acc += row[xx]; n++;
                }
            }
        }
        out[o] = (uint8_t)((acc + n / 2) / n);
    }
}

enum { FDCT_STORE_COEF = 0, FDCT_QUANT = 1 };

__global__ __launch_bounds__(FDCT_THREADS) void k_fdct(enc_src s, int mode, const vcg_dtab *__restrict__ T,
                                                       const vcg_fwd_tab *__restrict__ F,
                                                       int32_t *__restrict__ coef_out,
                                                       int16_t *__restrict__ lv_out,
                                                       enc_item *__restrict__ items) {
    extern __shared__ int buf[];
    __shared__ uint32_t ssum;
    __shared__ int seob;
    const int tid = threadIdx.x;
    const uint32_t i = blockIdx.x;
    const uint64_t gi = s.base + i;
    int lod; uint32_t ax, ay, az;
    enc_atom_of(s, gi, lod, ax, ay, az);
    if (tid == 0) { ssum = 0; seob = 0; }
    __syncthreads();
    const uint32_t part = vcg_gather(buf, s.lod[lod], s.nx[lod], s.ny[lod], s.nz[lod], ax, ay, az, tid, FDCT_THREADS);
    if (part) atomicAdd(&ssum, part);
    __syncthreads();
    const uint32_t sum = ssum;
    if (sum == 0) {                          // all-zero atom: KNOWN_ZERO, no payload
        if (tid == 0) { enc_item it = {0, 0, 1, (uint32_t)lod}; items[i] = it; }
        return;
    }
    const int dc = (int)((sum + 16384u) / 32768u);
    vcg_level_shift(buf, dc, tid, FDCT_THREADS);
    __syncthreads();
    vcg_fwd_pass_x(buf, F, tid, FDCT_THREADS);
    __syncthreads();
    vcg_fwd_pass_y(buf, F, tid, FDCT_THREADS);
    __syncthreads();
    vcg_fwd_pass_z(buf, F, tid, FDCT_THREADS);
    __syncthreads();
    if (mode == FDCT_STORE_COEF) {
        int32_t *co = coef_out + (size_t)i * VCG_A3;
        for (int k = tid; k < 32768; k += FDCT_THREADS) co[k] = buf[vcg_tidx(k >> 10, (k >> 5) & 31, k & 31)];
        if (tid == 0) { enc_item it = {0, (uint32_t)dc, 0, (uint32_t)lod}; items[i] = it; }
        return;
    }
    const int last = vcg_quant_scan(buf, T->scan, T->gstep, s.q[lod], lv_out + (size_t)i * VCG_A3, tid, FDCT_THREADS);
    if (last) atomicMax(&seob, last);
    __syncthreads();
    if (tid == 0) { enc_item it = {(uint32_t)seob, (uint32_t)dc, 0, (uint32_t)lod}; items[i] = it; }
}

__global__ __launch_bounds__(FDCT_THREADS) void k_quant(const int32_t *__restrict__ coef, float q,
                                                        const vcg_dtab *__restrict__ T,
                                                        int16_t *__restrict__ lv_out,
                                                        enc_item *__restrict__ items) {
    __shared__ int seob;
    const int tid = threadIdx.x;
    const uint32_t i = blockIdx.x;
    if (items[i].zero) return;
    if (tid == 0) seob = 0;
    __syncthreads();
    const int32_t *co = coef + (size_t)i * VCG_A3;
    int16_t *lo = lv_out + (size_t)i * VCG_A3;
    int last = 0;
    for (int p = tid; p < 32768; p += FDCT_THREADS) {
        uint32_t idx = T->scan[p];
        int l = vcg_quant1(co[idx], q, T->gstep[vcg_freq(idx)]);
        lo[p] = (int16_t)l;
        if (l) last = p + 1;
    }
    if (last) atomicMax(&seob, last);
    __syncthreads();
    if (tid == 0) items[i].eob = (uint32_t)seob;
}

// quantize stored sample coefficients at nq base_q values at once:
// block (atom i, candidate c) -> lv_out[(c*n + i) * 32768 ...], eobs[c*n + i]
__global__ __launch_bounds__(FDCT_THREADS) void k_quant_multi(const int32_t *__restrict__ coef,
                                                              const float *__restrict__ qs,
                                                              const vcg_dtab *__restrict__ T,
                                                              int16_t *__restrict__ lv_out,
                                                              const enc_item *__restrict__ items,
                                                              uint32_t *__restrict__ eobs, uint32_t n) {
    __shared__ int seob;
    const int tid = threadIdx.x;
    const uint32_t i = blockIdx.x, c = blockIdx.y;
    const size_t g = (size_t)c * n + i;
    if (items[i].zero) { if (tid == 0) eobs[g] = 0; return; }
    if (tid == 0) seob = 0;
    __syncthreads();
    const float q = qs[c];
    const int32_t *co = coef + (size_t)i * VCG_A3;
    int16_t *lo = lv_out + g * VCG_A3;
    int last = 0;
    for (int p = tid; p < 32768; p += FDCT_THREADS) {
        const uint32_t idx = T->scan[p];
        const int l = vcg_quant1(co[idx], q, T->gstep[vcg_freq(idx)]);
        lo[p] = (int16_t)l;
        if (l) last = p + 1;
    }
    if (last) atomicMax(&seob, last);
    __syncthreads();
    if (tid == 0) eobs[g] = (uint32_t)seob;
}

// byte counts of the sample at nq base_q values (lane per (candidate, atom)),
// payload bytes not stored.  A candidate whose running total passes `budget`
// is abandoned (its lanes report 0xFFFFFFFE); 0xFFFFFFFF = coder bound exceeded.
__global__ __launch_bounds__(64) void k_rc_count(uint32_t n, uint32_t nq, const int16_t *__restrict__ lv,
                                                 const enc_item *__restrict__ items,
                                                 const uint32_t *__restrict__ eobs,
                                                 const vcg_dtab *__restrict__ T, uint32_t budget,
                                                 uint32_t *__restrict__ kb, uint32_t *__restrict__ lens) {
    const uint32_t g = blockIdx.x * 64u + threadIdx.x;
    if (g >= n * nq) return;
    const uint32_t c = g / n, i = g % n;
    if (items[i].zero) { lens[g] = 0; return; }
    int ovf = 0, stopped = 0;
    const uint32_t len = vcg_rc_encode_atom(lv + (size_t)g * VCG_A3, eobs[g], T->band_start, nullptr, VCG_RC_CAP,
                                            &ovf, kb + c, budget, &stopped);
    lens[g] = ovf ? 0xFFFFFFFFu : stopped ? 0xFFFFFFFEu : len + 1;
}

// ---------------------------------------------------------------------------
// one lane per atom: payload = [dc][range-coder bytes]
__global__ __launch_bounds__(64) void k_rc_encode(uint32_t n, const int16_t *__restrict__ lv,
                                                  const enc_item *__restrict__ items,
                                                  const vcg_dtab *__restrict__ T,
                                                  uint8_t *__restrict__ pay, uint32_t pay_cap,
                                                  uint32_t *__restrict__ lens) {
    const uint32_t i = blockIdx.x * 64u + threadIdx.x;
    if (i >= n) return;
    const enc_item it = items[i];
    if (it.zero) { lens[i] = 0; return; }
    uint8_t *out = pay + (size_t)i * pay_cap;
    out[0] = (uint8_t)it.dc;
    const uint32_t cap = pay_cap - 1 < VCG_RC_CAP ? pay_cap - 1 : VCG_RC_CAP;
    int ovf = 0;
    uint32_t len = vcg_rc_encode_atom(lv + (size_t)i * VCG_A3, it.eob, T->band_start, out + 1, cap, &ovf);
    lens[i] = ovf ? 0xFFFFFFFFu : len + 1;
}

// copy payload i (len at lens[i]) from batch scratch into the pool at pool_off[i]
__global__ void k_compact(const uint8_t *__restrict__ pay, uint32_t pay_cap, const uint32_t *__restrict__ lens,
                          const uint64_t *__restrict__ pool_off, uint8_t *__restrict__ pool) {
    const uint32_t i = blockIdx.x;
    const uint32_t len = lens[i];
    if (len == 0 || len == 0xFFFFFFFFu) return;
    const uint8_t *src = pay + (size_t)i * pay_cap;
    uint8_t *dst = pool + pool_off[i];
    for (uint32_t b = threadIdx.x; b < len; b += blockDim.x) dst[b] = src[b];
}

// slot + payload for atom i of the archive image (layout computed on the host)
__global__ void k_assemble(const uint8_t *__restrict__ pool, const uint64_t *__restrict__ pool_off,
                           const uint32_t *__restrict__ lens, const uint64_t *__restrict__ pay_off,
                           const uint64_t *__restrict__ slot_off, uint8_t *__restrict__ img) {
    const uint64_t i = blockIdx.x;
    const uint32_t len = lens[i];
    const uint64_t so = slot_off[i];
    if (threadIdx.x == 0 && so) {
        uint8_t s[16];
        for (int b = 0; b < 16; ++b) s[b] = 0;
        if (len) {
            uint64_t po = pay_off[i];
            for (int b = 0; b < 8; ++b) s[b] = (uint8_t)(po >> (8 * b));
            for (int b = 0; b < 4; ++b) s[8 + b] = (uint8_t)(len >> (8 * b));
            s[12] = 1; s[13] = pool[pool_off[i]];
        } else {
            s[12] = 2;
        }
        for (int b = 0; b < 16; ++b) img[so + b] = s[b];
    }
    if (len) {
        const uint8_t *src = pool + pool_off[i];
        uint8_t *dst = img + pay_off[i];
        for (uint32_t b = threadIdx.x; b < len; b += blockDim.x) dst[b] = src[b];
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static vcg_fwd_tab *g_ftab = nullptr;
static const size_t LDS_BYTES = (size_t)VCG_TILE_INTS * 4;

static void enc_setup() {
    static bool done = false;
    if (done) return;
    vcg_dtab h;
    vcg_copy(&h, g_dtab, sizeof h, 2);
    vcg_fwd_tab f;
    for (int n = 0; n < 16; ++n)
        for (int j = 0; j < 16; ++j) { f.e[n][j] = h.cmat[2 * j][n]; f.o[n][j] = h.cmat[2 * j + 1][n]; }
    VCG_CHECK(hipMalloc(&g_ftab, sizeof f));
    vcg_copy(g_ftab, &f, sizeof f, 1);
    VCG_CHECK(hipFuncSetAttribute((const void *)k_fdct, hipFuncAttributeMaxDynamicSharedMemorySize, (int)LDS_BYTES));
    done = true;
}

struct vcg_enc {
    enc_src src;
    uint8_t *lod_buf[VCG_NLOD] = {nullptr};
    bool own0 = false;
    // rate-control sample
    uint32_t *sample_list = nullptr;
    int32_t *sample_coef = nullptr;
    enc_item *sample_items = nullptr;
    int sample_n = 0;
    // payload pool
    uint8_t *pool = nullptr; uint64_t pool_cap = 0, pool_len = 0;
    std::vector<uint64_t> pool_off;
    uint64_t *pool_off_d = nullptr;
};

static void *dmalloc(size_t n) {
    void *p = nullptr;
    if (hipMalloc(&p, n ? n : 1) != hipSuccess) { (void)hipGetLastError(); vc__panic("out of device memory (encode)"); }
    return p;
}

extern "C" int vcg_downsample(const void *in, const uint32_t d[3], void *out) {
    vcg_ctx *c = vcg_get_ctx();
    uint64_t total = (uint64_t)((d[0] + 1) / 2) * ((d[1] + 1) / 2) * ((d[2] + 1) / 2);
    uint64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_downsample, dim3((uint32_t)blocks), dim3(256), 0, c->s, (const uint8_t *)in, d[0], d[1],
                       d[2], (uint8_t *)out);
    VCG_CHECK(hipGetLastError());
    return 0;
}

extern "C" vcg_enc *vcg_enc_begin(const uint8_t *vol, int on_dev, const uint32_t dims[3], int nlod) {
    vcg_ctx *c = vcg_get_ctx();
    enc_setup();
    vcg_enc *e = new vcg_enc();
    memset(&e->src, 0, sizeof e->src);
    uint32_t x = dims[0], y = dims[1], z = dims[2];
    uint64_t acc = 0;
    for (int l = 0; l < nlod; ++l) {
        e->src.nx[l] = x; e->src.ny[l] = y; e->src.nz[l] = z;
        e->src.acx[l] = (x + 31) / 32; e->src.acy[l] = (y + 31) / 32;
        e->src.prefix[l] = acc;
        acc += (uint64_t)((x + 31) / 32) * ((y + 31) / 32) * ((z + 31) / 32);
        x = (x + 1) / 2; y = (y + 1) / 2; z = (z + 1) / 2;
    }
    e->src.prefix[nlod] = acc;
    e->src.nlod = nlod;
    size_t n0 = (size_t)dims[0] * dims[1] * dims[2];
    if (on_dev) {
        e->src.lod[0] = vol;
    } else {
        e->lod_buf[0] = (uint8_t *)dmalloc(n0);
        VCG_CHECK(hipMemcpyAsync(e->lod_buf[0], vol, n0, hipMemcpyHostToDevice, c->s));
        e->src.lod[0] = e->lod_buf[0];
        e->own0 = true;
    }
    for (int l = 1; l < nlod; ++l) {
        size_t nl = (size_t)e->src.nx[l] * e->src.ny[l] * e->src.nz[l];
        e->lod_buf[l] = (uint8_t *)dmalloc(nl);
        uint32_t d[3] = {e->src.nx[l - 1], e->src.ny[l - 1], e->src.nz[l - 1]};
        vcg_downsample(e->src.lod[l - 1], d, e->lod_buf[l]);
        e->src.lod[l] = e->lod_buf[l];
    }
    return e;
}

// scratch shared by the encode launches of this thread
struct enc_scratch {
    int16_t *lv = nullptr; size_t lv_cap = 0;
    enc_item *items = nullptr; size_t items_cap = 0;
    uint8_t *pay = nullptr; size_t pay_cap = 0;
    uint32_t *lens = nullptr; size_t lens_cap = 0;
    uint32_t *aux = nullptr; size_t aux_cap = 0;      // rate-control candidates: eobs, counters, q values
};
static thread_local enc_scratch tl_es;

// clear n words with a kernel (no memset node in code that may be captured)
__global__ void k_reset_words(uint32_t *w, uint32_t n) {
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) w[k] = 0;
}

static void enc_run(vcg_ctx *c, enc_src s, uint32_t n, const int32_t *coef_in, float q_coef,
                    enc_item *items, uint32_t pay_cap) {
    enc_scratch &es = tl_es;
    vcg_grow((void **)&es.lv, &es.lv_cap, n, (size_t)VCG_A3 * 2);
    vcg_grow((void **)&es.pay, &es.pay_cap, (size_t)n * pay_cap, 1);
    vcg_grow((void **)&es.lens, &es.lens_cap, n, 4);
    if (coef_in) {
        hipLaunchKernelGGL(k_quant, dim3(n), dim3(FDCT_THREADS), 0, c->s, coef_in, q_coef, g_dtab, es.lv, items);
    } else {
        s.n = n;
        hipLaunchKernelGGL(k_fdct, dim3(n), dim3(FDCT_THREADS), LDS_BYTES, c->s, s, (int)FDCT_QUANT, g_dtab,
                           g_ftab, (int32_t *)nullptr, es.lv, items);
    }
    VCG_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_rc_encode, dim3((n + 63) / 64), dim3(64), 0, c->s, n, es.lv, items, g_dtab, es.pay,
                       pay_cap, es.lens);
    VCG_CHECK(hipGetLastError());
}

static void sample_prepare(vcg_enc *e, vcg_ctx *c, const uint32_t *atoms, int n) {
    if (!e->sample_coef) {
        e->sample_n = n;
        e->sample_list = (uint32_t *)dmalloc((size_t)n * 12);
        e->sample_coef = (int32_t *)dmalloc((size_t)n * VCG_A3 * 4);
        e->sample_items = (enc_item *)dmalloc((size_t)n * sizeof(enc_item));
        VCG_CHECK(hipMemcpyAsync(e->sample_list, atoms, (size_t)n * 12, hipMemcpyHostToDevice, c->s));
        enc_src s = e->src;
        s.list = e->sample_list; s.base = 0; s.n = (uint32_t)n;
        hipLaunchKernelGGL(k_fdct, dim3(n), dim3(FDCT_THREADS), LDS_BYTES, c->s, s, (int)FDCT_STORE_COEF, g_dtab,
                           g_ftab, e->sample_coef, (int16_t *)nullptr, e->sample_items);
        VCG_CHECK(hipGetLastError());
    }
}

extern "C" int vcg_enc_sample(vcg_enc *e, const uint32_t *atoms, int n, float q, uint64_t *bytes,
                              uint8_t *zero_mask) {
    vcg_ctx *c = vcg_get_ctx();
    *bytes = 0;
    if (n <= 0) return 0;
    sample_prepare(e, c, atoms, n);
    enc_run(c, e->src, (uint32_t)n, e->sample_coef, q, e->sample_items, VCG_PAY_CAP);
    std::vector<uint32_t> lens(n);
    std::vector<enc_item> its(n);
    VCG_CHECK(hipMemcpyAsync(lens.data(), tl_es.lens, (size_t)n * 4, hipMemcpyDeviceToHost, c->s));
    VCG_CHECK(hipMemcpyAsync(its.data(), e->sample_items, (size_t)n * sizeof(enc_item), hipMemcpyDeviceToHost, c->s));
    VCG_CHECK(hipStreamSynchronize(c->s));
    int ovf = 0;
    uint64_t b = 0;
    for (int i = 0; i < n; ++i) {
        if (zero_mask) zero_mask[i] = (uint8_t)its[i].zero;
        if (lens[i] == 0xFFFFFFFFu) ovf = 1; else b += lens[i];
    }
    *bytes = b;
    return ovf;
}

extern "C" int vcg_enc_sample_zero(vcg_enc *e, const uint32_t *atoms, int n, uint8_t *zero_mask) {
    vcg_ctx *c = vcg_get_ctx();
    if (n <= 0) return 0;
    sample_prepare(e, c, atoms, n);
    std::vector<enc_item> its(n);
    VCG_CHECK(hipMemcpyAsync(its.data(), e->sample_items, (size_t)n * sizeof(enc_item), hipMemcpyDeviceToHost, c->s));
    VCG_CHECK(hipStreamSynchronize(c->s));
    for (int i = 0; i < n; ++i) zero_mask[i] = (uint8_t)its[i].zero;
    return 0;
}

// Sample byte totals at nq base_q values in one launch pair.  bytes[k]: the
// total (dc bytes included); flags[k]: 1 = coder bound exceeded, 2 = total
// passed `budget` (the encode was abandoned; bytes[k] is then a lower bound).
extern "C" int vcg_enc_sample_multi(vcg_enc *e, const uint32_t *atoms, int n, const float *qs, int nq,
                                    uint64_t budget, uint64_t *bytes, int *flags) {
    vcg_ctx *c = vcg_get_ctx();
    for (int k = 0; k < nq; ++k) { bytes[k] = 0; flags[k] = 0; }
    if (n <= 0 || nq <= 0) return 0;
    sample_prepare(e, c, atoms, n);
    enc_scratch &es = tl_es;
    const size_t m = (size_t)n * nq;
    vcg_grow((void **)&es.lv, &es.lv_cap, m, (size_t)VCG_A3 * 2);
    vcg_grow((void **)&es.lens, &es.lens_cap, m, 4);
    vcg_grow((void **)&es.aux, &es.aux_cap, m + 2 * (size_t)nq, 4);
    uint32_t *eobs = es.aux, *kb = es.aux + m;
    float *dq = (float *)(es.aux + m + nq);
    VCG_CHECK(hipMemcpyAsync(dq, qs, (size_t)nq * 4, hipMemcpyHostToDevice, c->s));
    hipLaunchKernelGGL(k_reset_words, dim3(1), dim3(64), 0, c->s, kb, (uint32_t)nq);
    hipLaunchKernelGGL(k_quant_multi, dim3(n, nq), dim3(FDCT_THREADS), 0, c->s, e->sample_coef, (const float *)dq,
                       g_dtab, es.lv, e->sample_items, eobs, (uint32_t)n);
    VCG_CHECK(hipGetLastError());
    // a candidate's counted total stays below 2^32 (at most 512 atoms x the
    // coder bound), so a budget of 2^32-1 or more never stops an encode
    const uint32_t b32 = budget >= 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)budget;
    hipLaunchKernelGGL(k_rc_count, dim3((unsigned)((m + 63) / 64)), dim3(64), 0, c->s, (uint32_t)n, (uint32_t)nq,
                       es.lv, e->sample_items, eobs, g_dtab, b32, kb, es.lens);
    VCG_CHECK(hipGetLastError());
    std::vector<uint32_t> lens(m);
    VCG_CHECK(hipMemcpyAsync(lens.data(), es.lens, m * 4, hipMemcpyDeviceToHost, c->s));
    VCG_CHECK(hipStreamSynchronize(c->s));
    for (int k = 0; k < nq; ++k)
        for (int i = 0; i < n; ++i) {
            const uint32_t L = lens[(size_t)k * n + i];
            if (L == 0xFFFFFFFFu) flags[k] |= 1;
            else if (L == 0xFFFFFFFEu) flags[k] |= 2;
            else bytes[k] += L;
        }
    return 0;
}

extern "C" int vcg_enc_all(vcg_enc *e, const float qs[VCG_NLOD], uint32_t *lens, uint64_t total) {
    vcg_ctx *c = vcg_get_ctx();
    enc_scratch &es = tl_es;
    enc_src s = e->src;
    for (int l = 0; l < VCG_NLOD; ++l) s.q[l] = qs[l];
    s.list = nullptr;
    e->pool_off.assign(total, 0);
    e->pool_len = 0;
    uint32_t maxb = total < ENC_BATCH ? (uint32_t)total : ENC_BATCH;
    vcg_grow((void **)&es.items, &es.items_cap, maxb, sizeof(enc_item));
    uint64_t *poff_d = nullptr; size_t poff_cap = 0;
    int ovf = 0;
    for (uint64_t b = 0; b < total; b += maxb) {
        uint32_t n = (uint32_t)(total - b < maxb ? total - b : maxb);
        s.base = b;
        enc_run(c, s, n, nullptr, 0.f, es.items, VCG_PAY_CAP);
        VCG_CHECK(hipMemcpyAsync(lens + b, es.lens, (size_t)n * 4, hipMemcpyDeviceToHost, c->s));
        VCG_CHECK(hipStreamSynchronize(c->s));
        uint64_t add = 0;
        for (uint32_t k = 0; k < n; ++k) {
            if (lens[b + k] == 0xFFFFFFFFu) { ovf = 1; lens[b + k] = 0; continue; }
            e->pool_off[b + k] = e->pool_len + add;
            add += lens[b + k];
        }
        if (e->pool_len + add > e->pool_cap) {
            uint64_t nc = (e->pool_len + add) + (e->pool_len + add) / 2 + 4096;
            uint8_t *np = (uint8_t *)dmalloc(nc);
            if (e->pool_len) VCG_CHECK(hipMemcpyAsync(np, e->pool, e->pool_len, hipMemcpyDeviceToDevice, c->s));
            VCG_CHECK(hipStreamSynchronize(c->s));
            if (e->pool) VCG_CHECK(hipFree(e->pool));
            e->pool = np; e->pool_cap = nc;
        }
        vcg_grow((void **)&poff_d, &poff_cap, n, 8);
        VCG_CHECK(hipMemcpyAsync(poff_d, e->pool_off.data() + b, (size_t)n * 8, hipMemcpyHostToDevice, c->s));
        hipLaunchKernelGGL(k_compact, dim3(n), dim3(256), 0, c->s, es.pay, (uint32_t)VCG_PAY_CAP, es.lens, poff_d,
                           e->pool);
        VCG_CHECK(hipGetLastError());
        VCG_CHECK(hipStreamSynchronize(c->s));
        e->pool_len += add;
    }
    if (poff_d) VCG_CHECK(hipFree(poff_d));
    return ovf;
}

extern "C" int vcg_enc_assemble(vcg_enc *e, const uint8_t *head, size_t head_len, const uint64_t *pay_off,
                                const uint64_t *slot_off, const uint32_t *lens, uint64_t total,
                                uint8_t *out, uint64_t total_len) {
    vcg_ctx *c = vcg_get_ctx();
    uint8_t *img = (uint8_t *)dmalloc(total_len);
    uint64_t *d_pool_off = (uint64_t *)dmalloc(total * 8);
    uint64_t *d_pay_off = (uint64_t *)dmalloc(total * 8);
    uint64_t *d_slot_off = (uint64_t *)dmalloc(total * 8);
    uint32_t *d_lens = (uint32_t *)dmalloc(total * 4);
    VCG_CHECK(hipMemsetAsync(img, 0, total_len, c->s));
    VCG_CHECK(hipMemcpyAsync(img, head, head_len, hipMemcpyHostToDevice, c->s));
    VCG_CHECK(hipMemcpyAsync(d_pool_off, e->pool_off.data(), total * 8, hipMemcpyHostToDevice, c->s));
    VCG_CHECK(hipMemcpyAsync(d_pay_off, pay_off, total * 8, hipMemcpyHostToDevice, c->s));
    VCG_CHECK(hipMemcpyAsync(d_slot_off, slot_off, total * 8, hipMemcpyHostToDevice, c->s));
    VCG_CHECK(hipMemcpyAsync(d_lens, lens, total * 4, hipMemcpyHostToDevice, c->s));
    for (uint64_t b = 0; b < total; b += 65536) {
        uint64_t n = total - b < 65536 ? total - b : 65536;
        hipLaunchKernelGGL(k_assemble, dim3((uint32_t)n), dim3(256), 0, c->s, e->pool, d_pool_off + b, d_lens + b,
                           d_pay_off + b, d_slot_off + b, img);
        VCG_CHECK(hipGetLastError());
    }
    VCG_CHECK(hipMemcpyAsync(out, img, total_len, hipMemcpyDeviceToHost, c->s));
    VCG_CHECK(hipStreamSynchronize(c->s));
    VCG_CHECK(hipFree(img)); VCG_CHECK(hipFree(d_pool_off)); VCG_CHECK(hipFree(d_pay_off));
    VCG_CHECK(hipFree(d_slot_off)); VCG_CHECK(hipFree(d_lens));
    return 0;
}

extern "C" void vcg_enc_end(vcg_enc *e) {
    if (!e) return;
    vcg_ctx *c = vcg_get_ctx();
    VCG_CHECK(hipStreamSynchronize(c->s));
    for (int l = 0; l < VCG_NLOD; ++l) if (e->lod_buf[l]) VCG_CHECK(hipFree(e->lod_buf[l]));
    if (e->sample_list) VCG_CHECK(hipFree(e->sample_list));
    if (e->sample_coef) VCG_CHECK(hipFree(e->sample_coef));
    if (e->sample_items) VCG_CHECK(hipFree(e->sample_items));
    if (e->pool) VCG_CHECK(hipFree(e->pool));
    delete e;
}

extern "C" int vcg_encode_atoms(size_t n, const uint8_t *vox, float q, uint8_t *out, size_t cap, uint32_t *lens) {
    vcg_ctx *c = vcg_get_ctx();
    enc_setup();
    if (n == 0) return 0;
    // treat the atoms as one LOD0 volume of n atoms along x: 32n x 32 x 32 is not the
    // same memory order, so gather explicitly through a (32, 32, 32n) z-stacked volume
    uint8_t *d = (uint8_t *)dmalloc(n * (size_t)VCG_A3);
    VCG_CHECK(hipMemcpyAsync(d, vox, n * (size_t)VCG_A3, hipMemcpyHostToDevice, c->s));
    enc_src s;
    memset(&s, 0, sizeof s);
    s.lod[0] = d; s.nx[0] = 32; s.ny[0] = 32; s.nz[0] = (uint32_t)(32 * n);
    s.acx[0] = 1; s.acy[0] = 1; s.prefix[0] = 0; s.prefix[1] = n; s.nlod = 1; s.q[0] = q;
    enc_scratch &es = tl_es;
    int rc = 0;
    std::vector<uint32_t> hl;
    for (size_t b = 0; b < n; b += ENC_BATCH) {
        uint32_t cnt = (uint32_t)(n - b < ENC_BATCH ? n - b : ENC_BATCH);
        vcg_grow((void **)&es.items, &es.items_cap, cnt, sizeof(enc_item));
        s.base = b;
        enc_run(c, s, cnt, nullptr, 0.f, es.items, VCG_PAY_CAP);
        hl.resize(cnt);
        VCG_CHECK(hipMemcpyAsync(hl.data(), es.lens, (size_t)cnt * 4, hipMemcpyDeviceToHost, c->s));
        VCG_CHECK(hipStreamSynchronize(c->s));
        for (uint32_t k = 0; k < cnt; ++k) {
            uint32_t L = hl[k];
            if (L == 0xFFFFFFFFu || L > cap) { rc = 1; lens[b + k] = 0; continue; }
            lens[b + k] = L;
            if (L) VCG_CHECK(hipMemcpyAsync(out + (b + k) * cap, es.pay + (size_t)k * VCG_PAY_CAP, L,
                                            hipMemcpyDeviceToHost, c->s));
        }
        VCG_CHECK(hipStreamSynchronize(c->s));
    }
    VCG_CHECK(hipFree(d));
    return rc;
}
