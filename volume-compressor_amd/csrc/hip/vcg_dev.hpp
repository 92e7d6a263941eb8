// vcg_dev.hpp — device-side definitions shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../vcg.h"

#define VCG_CHECK(x)                                                          \
    do {                                                                      \
        hipError_t e_ = (x);                                                  \
        if (e_ != hipSuccess) vcg_fail(#x, e_, __FILE__, __LINE__);           \
    } while (0)

[[noreturn]] void vcg_fail(const char *what, hipError_t e, const char *file, int line);

// Codec tables in device memory (copied from vcg_tables at init).
struct vcg_dtab {
    int32_t  cmat[32][32];
    float    gstep[VCG_NFREQ];
    uint32_t band_start[9];
    uint16_t scan[VCG_A3];
    uint8_t  fscan[VCG_A3];   // frequency cz+cy+cx of scan position p (non-decreasing in p)
    int8_t   fa[3][64][16];   // forward transform on MFMA: per-lane basis planes Ch, Cl, 4 Cl (vcg_fwd1_tables)
    int32_t  rowsum[32];      // basis row sums
    int8_t   ib[3][64][16];   // inverse on MFMA: per-lane basis planes Ch, 4 Cl, Cl (vcg_im_tables)
    uint32_t imtab[VCG_A3];   // scan position p: i16-tile byte offset | frequency << 16 (vcg_im_tab)
};
extern vcg_dtab *g_dtab;  // device pointer

// Per-item result of the entropy stage.
struct vcg_item {
    uint32_t cnt;     // coefficient entries written
    uint32_t status;  // 0: all-zero output, 1: coefficients, 2: corrupt (zeros + FORMAT), 3: out of range (zeros)
    uint32_t dc;      // payload dc byte
    uint16_t lod;     // dequant table
    int16_t fmax;     // highest coded frequency cz+cy+cx (-1: none)
};

// Coefficient entry: scan position << 17 | (2m - 11 + sign) mod 2^17 for a
// level of magnitude m (vcg_ent_make / vcg_ent_level).  Entries of one item
// are written in increasing scan position.
#define VCG_LIST_CAP 32768u
// per-atom list stride in entries: padded past a power of two so the lists of
// the 64 lanes of a wave (written concurrently) do not share cache sets
#define VCG_LIST_STRIDE (VCG_LIST_CAP + 64u)

// Per-thread execution context (one HIP stream, grow-only scratch).
// Per host thread.  s is the thread's public stream (vcx_stream): calls are
// ordered on it.  A decode batch runs its entropy and transform launches on the
// stream of one of VCG_SLOTS scratch slots (sE[slot]); batches use the slots
// in turn, so the entropy passes of consecutive batches or calls run side by
// side (one pass keeps only half the SIMDs busy) and overlap the transforms.
// A transform also waits for earlier work on s (it writes the caller's
// output), and the output is joined back into s by an event.  With the public
// stream that makes four streams: the device's default number of hardware
// queues.
#ifndef VCG_SLOTS
#define VCG_SLOTS 3
#endif
struct vcg_ctx {
    hipStream_t s = nullptr, sE[VCG_SLOTS] = {};
    hipEvent_t evF = nullptr, evI[VCG_SLOTS] = {};
    int slot = 0, last_slot = 0;
    uint8_t  *stage = nullptr;  size_t stage_cap = 0;
    int32_t  *quads = nullptr;  size_t quads_cap = 0;
    int32_t  *stat = nullptr;   size_t stat_cap = 0;
    uint32_t *status_d = nullptr;
    uint32_t *status_h = nullptr;  // pinned
    void     *ev[3] = {nullptr, nullptr, nullptr};
    // encode: the coarse atoms' stream, its start (descriptor upload) and end events
    hipStream_t sC = nullptr;
    hipEvent_t evCup = nullptr, evCdone = nullptr;
};
vcg_ctx *vcg_get_ctx();
void vcg_grow(void **p, size_t *cap, size_t need, size_t elem);

__device__ __forceinline__ uint32_t ld_u32(const uint8_t *p) {
    uint32_t v; __builtin_memcpy(&v, p, 4); return v;
}
__device__ __forceinline__ uint64_t ld_u64(const uint8_t *p) {
    uint64_t v; __builtin_memcpy(&v, p, 8); return v;
}

// splitmix64 finalizer (src/vc/vc.c:721-726), used by the v1 region hash probe.
__device__ __forceinline__ uint64_t vcg_mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Resolve an atom to its payload.  Returns 0: decode to zeros (ABSENT or
// KNOWN_ZERO), 1: present with a valid [off, off+len), 2: present but the
// slot is invalid (zeros + VC_ERR_FORMAT).  src/vc/vc.c:1349-1402, 1412-1413.
__device__ inline int vcg_resolve(const uint8_t *arc, const vcg_geom *g, int lod,
                                  uint32_t az, uint32_t ay, uint32_t ax,
                                  uint64_t *off, uint32_t *len) {
    uint64_t blk;
    if (g->version == 2) {
        uint32_t rz = az >> 5, ry = ay >> 5, rx = ax >> 5;
        if (rz >= g->nrz[lod] || ry >= g->nry[lod] || rx >= g->nrx[lod]) return 0;
        uint64_t ri = g->region_prefix[lod] + ((uint64_t)rz * g->nry[lod] + ry) * g->nrx[lod] + rx;
        if (ri >= g->region_prefix[g->nlod]) return 0;
        const uint8_t *e = arc + g->dir_off + ri * 16;
        uint32_t f = ld_u32(e + 12);
        if (f == 0u || f == 1u) return 0;            // ABSENT / ZERO region; any other flag reads as data
        blk = ld_u64(e);
    } else {
        uint64_t key = ((uint64_t)(lod & 7) << 45) | ((uint64_t)((az >> 5) & 0x7FFF) << 30) |
                       ((uint64_t)((ay >> 5) & 0x7FFF) << 15) | (uint64_t)((ax >> 5) & 0x7FFF);
        uint64_t mask = g->l1_cap - 1, i = vcg_mix64(key) & mask;
        blk = 0;
        for (uint64_t probe = 0; probe < g->l1_cap; ++probe) {   // bounded: a full table has no empty bucket
            const uint8_t *e = arc + g->l1_off + i * 16;
            uint64_t br = ld_u64(e + 8);
            if (br == 0) return 0;
            if (ld_u64(e) == key) { blk = br; break; }
            i = (i + 1) & mask;
        }
        if (blk == 0 || blk == 1) return 0;          // not found / ZERO_REGION
    }
    if (blk > g->len || g->len - blk < 32768ull * 16) return 0;   // corrupt guard (vc.c:1389,1400)
    const uint8_t *s = arc + blk + ((uint64_t)(((az & 31) * 32 + (ay & 31)) * 32 + (ax & 31))) * 16;
    if (s[12] != 1) return 0;                        // AF_ABSENT / AF_ZERO
    uint64_t o = ld_u64(s);
    uint32_t l = ld_u32(s + 8);
    if (o == ~0ull || l < 1 || o > g->len || (uint64_t)l > g->len - o) return 2;
    *off = o; *len = l;
    return 1;
}

__device__ __forceinline__ int wave_max_i32(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}
