// vcg_codec.hpp — per-atom entropy coder and quantizer shared by the HIP
// kernels and by the CPU build of the same code (tests/kernel_cpu), so the
// exact device logic can be exercised against the oracle on the host.
//
//   vcg_rc_decode_atom : adaptive binary range decoder + coefficient context
//                        decoder (src/vc/vc.c:403-441, 491-560)
//   vcg_rc_encode_atom : range encoder with LZMA carry + context coder
//                        (src/vc/vc.c:360-401, 471-544)
//   vcg_quant1         : dead-zone quantizer (src/vc/vc.c:293-306)
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define VCG_HD __host__ __device__ __forceinline__
#define VCG_HDM __host__ __device__ __forceinline__
#else
#define VCG_HD static inline
#define VCG_HDM inline
#endif

#define VCG_EG_PREFIX_CAP 64u

// ---------------------------------------------------------------------------
// decoder.  One renormalisation step per bin is exact: after a context bin
// range >= 2^24*31/4096 and after a bypass range >= 2^23, so a single byte
// shift restores range >= 2^24.
// ---------------------------------------------------------------------------
struct vcg_rdec {
    const uint8_t *s;
    uint32_t n, pos, code, range;
};
VCG_HD uint32_t vcg_rd_next(vcg_rdec &d) {
    uint32_t b = d.pos < d.n ? (uint32_t)d.s[d.pos] : 0u;   // past-end bytes read as 0 (vc.c:425)
    d.pos++;
    return b;
}
VCG_HD int vcg_rd_bit(vcg_rdec &d, uint32_t &p) {
    uint32_t r0 = (uint32_t)(((uint64_t)d.range * p) >> 12);
    int b;
    if (d.code < r0) { d.range = r0; p += (4096u - p) >> 5; b = 0; }
    else { d.code -= r0; d.range -= r0; p -= p >> 5; b = 1; }
    if (d.range < (1u << 24)) { d.code = (d.code << 8) | vcg_rd_next(d); d.range <<= 8; }
    return b;
}
VCG_HD int vcg_rd_bypass(vcg_rdec &d) {
    d.range >>= 1;
    int b = d.code >= d.range;
    if (b) d.code -= d.range;
    if (d.range < (1u << 24)) { d.code = (d.code << 8) | vcg_rd_next(d); d.range <<= 8; }
    return b;
}

// Decode one atom's range-coder bytes (after the dc byte).  Writes the nonzero
// levels as (scan position << 16 | uint16 level) in increasing position to
// out (capacity 32768) and their count to *cnt.  Returns 0, or 1 when an
// Exp-Golomb prefix exceeds VCG_EG_PREFIX_CAP bits (reported as a format error;
// the reference never returns from such a stream).
VCG_HD int vcg_rc_decode_atom(const uint8_t *s, uint32_t n, const uint32_t *band_start, uint32_t *out,
                              uint32_t *cnt_out) {
    vcg_rdec d;
    d.s = s; d.n = n; d.pos = 0; d.code = 0; d.range = 0xFFFFFFFFu;
    for (int k = 0; k < 5; ++k) d.code = (d.code << 8) | vcg_rd_next(d);      // vc.c:406-409
    uint32_t eob = 0;
    for (int k = 0; k < 16; ++k) eob = (eob << 1) | (uint32_t)vcg_rd_bypass(d);  // vc.c:520-525
    if (eob > 32768u) eob = 32768u;                                            // vc.c:549
    // sig contexts are used band by band in increasing scan position, so one
    // live context suffices: a new band starts from the reset value 2048.
    uint32_t psig = 2048u;
    uint32_t m0 = 2048u, m1 = 2048u, m2 = 2048u, m3 = 2048u, m4 = 2048u, m5 = 2048u,
             m6 = 2048u, m7 = 2048u, m8 = 2048u, m9 = 2048u, m10 = 2048u, m11 = 2048u;
    int band = 0;
    uint32_t nextb = band_start[1];
    uint32_t cnt = 0;
    for (uint32_t pp = 0; pp < eob; ++pp) {
        while (pp >= nextb) { ++band; nextb = band_start[band + 1]; psig = 2048u; }
        if (!vcg_rd_bit(d, psig)) continue;
        // magnitude: unary over mag[0..10], mag[11] escape, EG0 bypass (vc.c:491-506)
        uint32_t m;
        if (!vcg_rd_bit(d, m0)) m = 1;
        else if (!vcg_rd_bit(d, m1)) m = 2;
        else if (!vcg_rd_bit(d, m2)) m = 3;
        else if (!vcg_rd_bit(d, m3)) m = 4;
        else if (!vcg_rd_bit(d, m4)) m = 5;
        else if (!vcg_rd_bit(d, m5)) m = 6;
        else if (!vcg_rd_bit(d, m6)) m = 7;
        else if (!vcg_rd_bit(d, m7)) m = 8;
        else if (!vcg_rd_bit(d, m8)) m = 9;
        else if (!vcg_rd_bit(d, m9)) m = 10;
        else if (!vcg_rd_bit(d, m10)) m = 11;
        else if (!vcg_rd_bit(d, m11)) m = 12;
        else {
            uint32_t nb = 0;
            while (vcg_rd_bypass(d)) {
                if (++nb > VCG_EG_PREFIX_CAP) { *cnt_out = 0; return 1; }
            }
            uint32_t x = 1;
            for (uint32_t k = 0; k < nb; ++k) x = (x << 1) | (uint32_t)vcg_rd_bypass(d);
            m = 11u + (x - 1u) + 1u;
        }
        int neg = vcg_rd_bypass(d);
        uint32_t lv = neg ? 0u - m : m;
        out[cnt++] = (pp << 16) | (lv & 0xFFFFu);
    }
    *cnt_out = cnt;
    return 0;
}

// ---------------------------------------------------------------------------
// Decoder as a one-bin-per-step state machine (the batched kernels' form).
//
// Every lane executes the same instructions each step whatever symbol it is
// in (significance, magnitude unary, Exp-Golomb prefix/suffix, sign), so a
// wave of 64 atoms does not serialise on divergent branches.  Input bytes come
// from a 16-byte little-endian register window; every VCG_WSTEP steps lanes
// whose window holds <= 8 bytes append the next 8, which were loaded (as an
// aligned u64) two refills earlier, so global-load latency hides behind the
// steps in between.  The window never underflows: a step consumes at most one
// byte (single renormalisation per bin, see vcg_rd_bit).  Magnitude contexts
// 1..11 live in `mg` (LDS on the GPU); the one a step may need next is read
// at the start of the step and mag[0] stays in a register.
//
// Reads are whole aligned 8-byte blocks covering [s, s+n): the caller
// guarantees the buffer extends to the next 8-byte boundary past s+n plus 16
// bytes (the device archive copy is padded); bytes outside [s, s+n) are masked
// to zero (past-end bytes read as 0, vc.c:425).
// ---------------------------------------------------------------------------
#define VCG_WSTEP 8
#ifndef VCG_ENTROPY_FSM
#define VCG_ENTROPY_FSM 0
#endif
enum { VCG_S_SIG = 0, VCG_S_MAG = 1, VCG_S_EGP = 2, VCG_S_EGS = 3, VCG_S_SIGN = 4, VCG_S_DONE = 5, VCG_S_BAD = 6 };

// aligned 8-byte block `blk` of the stream whose first byte sits `mis` bytes
// into block 0 and which holds `n` bytes; bytes outside the stream are zero
VCG_HD uint64_t vcg_fetch8(const uint64_t *base, uint32_t blk, uint32_t mis, uint32_t n) {
    const int64_t lo = (int64_t)mis - 8 * (int64_t)blk;          // first valid byte index within the block
    const int64_t hi = (int64_t)mis + n - 8 * (int64_t)blk;      // one past the last
    if (hi <= 0 || lo >= 8) return 0;
    uint64_t v = base[blk];
    if (lo > 0) v &= ~0ull << (8 * lo);
    if (hi < 8) v &= (1ull << (8 * hi)) - 1ull;
    return v;
}

struct vcg_fsm {
    // range decoder
    uint32_t code, range;
    // byte window (little endian, wn valid bytes) and two pending blocks
    uint64_t wlo, whi, p1, p2;
    uint32_t wn, blk;
    // symbol state
    uint32_t st, pos, eob, k, nb, x, m;
    uint32_t psig, pm, pnext, m0;
    uint32_t band, nextb, cnt;
};

VCG_HD void vcg_win_pop(vcg_fsm &f, uint32_t &byte) {
    byte = (uint32_t)f.wlo & 0xFFu;
    f.wlo = (f.wlo >> 8) | (f.whi << 56);
    f.whi >>= 8;
    f.wn--;
}
// append 8 bytes at byte offset wn (wn <= 8)
VCG_HD void vcg_win_push8(vcg_fsm &f, uint64_t v) {
    const uint32_t sh = 8 * f.wn;
    if (sh == 0) { f.wlo |= v; }
    else if (sh == 64) { f.whi |= v; }
    else { f.wlo |= v << sh; f.whi |= v >> (64 - sh); }
    f.wn += 8;
}

VCG_HD void vcg_fsm_init(vcg_fsm &f, const uint64_t *base, uint32_t mis, uint32_t n, const uint32_t *band_start) {
    f.wlo = vcg_fetch8(base, 0, mis, n) >> (8 * mis);
    f.whi = 0;
    f.wn = 8 - mis;
    vcg_win_push8(f, vcg_fetch8(base, 1, mis, n));
    f.p1 = vcg_fetch8(base, 2, mis, n);
    f.p2 = vcg_fetch8(base, 3, mis, n);
    f.blk = 4;
    f.code = 0; f.range = 0xFFFFFFFFu;
    uint32_t b;
    for (int k = 0; k < 5; ++k) { vcg_win_pop(f, b); f.code = (f.code << 8) | b; }   // vc.c:406-409
    // EOB: 16 bypass bins (vc.c:520-525).  The window holds >= 4 bytes here and
    // 16 halvings of a full range renormalise at most 3 times.
    uint32_t eob = 0;
    for (int k = 0; k < 16; ++k) {
        f.range >>= 1;
        uint32_t bit = f.code >= f.range;
        if (bit) f.code -= f.range;
        if (f.range < (1u << 24)) { vcg_win_pop(f, b); f.code = (f.code << 8) | b; f.range <<= 8; }
        eob = (eob << 1) | bit;
    }
    f.eob = eob > 32768u ? 32768u : eob;                                   // vc.c:549
    f.st = f.eob ? VCG_S_SIG : VCG_S_DONE;
    f.pos = 0; f.k = 0; f.nb = 0; f.x = 0; f.m = 0;
    f.psig = 2048u; f.pm = 2048u; f.pnext = 2048u; f.m0 = 2048u;
    f.band = 0; f.nextb = band_start[1]; f.cnt = 0;
}

// refill (every VCG_WSTEP steps): lanes holding <= 8 bytes append the oldest pending block
VCG_HD void vcg_fsm_refill(vcg_fsm &f, const uint64_t *base, uint32_t mis, uint32_t n) {
    if (f.wn <= 8) {
        vcg_win_push8(f, f.p1);
        f.p1 = f.p2;
        f.p2 = vcg_fetch8(base, f.blk, mis, n);
        f.blk++;
    }
}

// one bin of the symbol stream; mg = magnitude contexts (index 1..11 used)
template <class MagMem>
VCG_HD void vcg_fsm_step(vcg_fsm &f, MagMem &mg, const uint32_t *band_start, uint32_t *out) {
    const uint32_t st = f.st;
    if (st >= VCG_S_DONE) return;
    // context read for the magnitude bin after this one (used only if this bin is 1)
    if (st == VCG_S_MAG && f.k + 1 < 12) f.pnext = mg.get(f.k + 1);
    const bool byp = st >= VCG_S_EGP;
    const uint32_t p = st == VCG_S_SIG ? f.psig : f.pm;
    const uint32_t r0 = byp ? (f.range >> 1) : (uint32_t)(((uint64_t)f.range * p) >> 12);
    const uint32_t bit = f.code >= r0;
    if (bit) f.code -= r0;
    f.range = (byp || !bit) ? r0 : f.range - r0;
    const uint32_t pn = bit ? p - (p >> 5) : p + ((4096u - p) >> 5);
    if (f.range < (1u << 24)) { uint32_t b; vcg_win_pop(f, b); f.code = (f.code << 8) | b; f.range <<= 8; }
    if (st == VCG_S_SIG) {
        f.psig = pn;
        if (bit) { f.st = VCG_S_MAG; f.k = 0; f.pm = f.m0; return; }
        f.pos++;
    } else if (st == VCG_S_MAG) {
        if (f.k == 0) f.m0 = pn; else mg.set(f.k, pn);
        if (bit) {
            f.k++;
            if (f.k == 12) { f.st = VCG_S_EGP; f.nb = 0; }
            else f.pm = f.pnext;
        } else { f.m = f.k + 1; f.st = VCG_S_SIGN; }
        return;
    } else if (st == VCG_S_EGP) {
        if (bit) { if (++f.nb > VCG_EG_PREFIX_CAP) f.st = VCG_S_BAD; }
        else { f.x = 1; if (f.nb == 0) { f.m = 12; f.st = VCG_S_SIGN; } else f.st = VCG_S_EGS; }
        return;
    } else if (st == VCG_S_EGS) {
        f.x = (f.x << 1) | bit;
        if (--f.nb == 0) { f.m = 11u + f.x; f.st = VCG_S_SIGN; }     // v + (x-1) + 1 with v = 11
        return;
    } else {                                                          // sign bypass: emit
        const uint32_t lv = bit ? 0u - f.m : f.m;
        out[f.cnt++] = (f.pos << 16) | (lv & 0xFFFFu);
        f.pos++;
        f.st = VCG_S_SIG;
    }
    // advanced past a position (zero significance or emitted coefficient)
    if (f.pos == f.eob) { f.st = VCG_S_DONE; return; }
    if (f.pos == f.nextb) { f.band++; f.nextb = band_start[f.band + 1]; f.psig = 2048u; }
}

// Whole-atom decode with the state machine (CPU-side reference of the kernel's
// per-lane logic).  Returns 0, or 1 for an over-long Exp-Golomb prefix.
struct vcg_mag_array {
    uint32_t v[12];
    VCG_HDM uint32_t get(uint32_t k) const { return v[k]; }
    VCG_HDM void set(uint32_t k, uint32_t p) { v[k] = p; }
};
VCG_HD int vcg_rc_decode_atom_fsm(const uint64_t *base, uint32_t mis, uint32_t n, const uint32_t *band_start,
                                  uint32_t *out, uint32_t *cnt_out) {
    vcg_fsm f;
    vcg_fsm_init(f, base, mis, n, band_start);
    vcg_mag_array mg;
    for (int k = 0; k < 12; ++k) mg.v[k] = 2048u;
    while (f.st < VCG_S_DONE) {
        vcg_fsm_refill(f, base, mis, n);
        for (int s = 0; s < VCG_WSTEP; ++s) vcg_fsm_step(f, mg, band_start, out);
    }
    *cnt_out = f.st == VCG_S_BAD ? 0 : f.cnt;
    return f.st == VCG_S_BAD;
}

// ---------------------------------------------------------------------------
// encoder
// ---------------------------------------------------------------------------
struct vcg_renc {
    uint8_t *out;
    uint32_t pos, cap;
    uint64_t low;
    uint32_t range;
    uint32_t cache;
    uint32_t csize;
    int ovf;
};
VCG_HD void vcg_re_put(vcg_renc &e, uint32_t b) {
    if (e.pos < e.cap) e.out[e.pos] = (uint8_t)b; else e.ovf = 1;
    e.pos++;
}
VCG_HD void vcg_re_shift_low(vcg_renc &e) {                  // vc.c:369-380
    if ((uint32_t)(e.low >> 32) != 0u || e.low < 0xFF000000ull) {
        uint32_t carry = (uint32_t)(e.low >> 32);
        do { vcg_re_put(e, (e.cache + carry) & 0xFFu); e.cache = 0xFFu; } while (--e.csize);
        e.cache = (uint32_t)(e.low >> 24) & 0xFFu;
    }
    e.csize++;
    e.low = (e.low << 8) & 0xFFFFFFFFull;
}
VCG_HD void vcg_re_bit(vcg_renc &e, uint32_t &p, int bit) { // vc.c:381-393
    uint32_t r0 = (uint32_t)(((uint64_t)e.range * p) >> 12);
    if (!bit) { e.range = r0; p += (4096u - p) >> 5; }
    else { e.low += r0; e.range -= r0; p -= p >> 5; }
    while (e.range < (1u << 24)) { vcg_re_shift_low(e); e.range <<= 8; }
}
VCG_HD void vcg_re_bypass(vcg_renc &e, int bit) {           // vc.c:394-398
    e.range >>= 1;
    if (bit) e.low += e.range;
    while (e.range < (1u << 24)) { vcg_re_shift_low(e); e.range <<= 8; }
}
VCG_HD int vcg_nbits(uint32_t t) {                          // floor(log2 t), t >= 1
    int nb = 0;
    while (t > 1) { t >>= 1; nb++; }
    return nb;
}

// Encode levels given in scan order (L[p], p < eob) into out (capacity cap).
// Returns the byte count; *ovf set if cap was exceeded.  vc.c:513-544,399-401.
VCG_HD uint32_t vcg_rc_encode_atom(const int16_t *L, uint32_t eob, const uint32_t *band_start, uint8_t *out,
                                   uint32_t cap, int *ovf) {
    vcg_renc e;
    e.out = out; e.pos = 0; e.cap = cap; e.low = 0; e.range = 0xFFFFFFFFu; e.cache = 0; e.csize = 1; e.ovf = 0;
    for (int b = 15; b >= 0; --b) vcg_re_bypass(e, (eob >> b) & 1);
    uint32_t m0 = 2048u, m1 = 2048u, m2 = 2048u, m3 = 2048u, m4 = 2048u, m5 = 2048u,
             m6 = 2048u, m7 = 2048u, m8 = 2048u, m9 = 2048u, m10 = 2048u, m11 = 2048u;
    uint32_t psig = 2048u;
    int band = 0;
    uint32_t nextb = band_start[1];
    for (uint32_t p = 0; p < eob; ++p) {
        while (p >= nextb) { ++band; nextb = band_start[band + 1]; psig = 2048u; }
        const int v = L[p];
        if (v == 0) { vcg_re_bit(e, psig, 0); continue; }
        vcg_re_bit(e, psig, 1);
        const uint32_t mm = (uint32_t)(v < 0 ? -v : v) - 1u;   // ones before the stop bin (vc.c:471-490)
        // unary: mag[k] gets 1 for k < mm and 0 at k == mm (k <= 10); mm >= 11
        // continues into mag[11]: 0 for mm == 11, else 1 + EG0(mm - 11).
        if (mm == 0) { vcg_re_bit(e, m0, 0); goto sign; }
        vcg_re_bit(e, m0, 1);
        if (mm == 1) { vcg_re_bit(e, m1, 0); goto sign; }
        vcg_re_bit(e, m1, 1);
        if (mm == 2) { vcg_re_bit(e, m2, 0); goto sign; }
        vcg_re_bit(e, m2, 1);
        if (mm == 3) { vcg_re_bit(e, m3, 0); goto sign; }
        vcg_re_bit(e, m3, 1);
        if (mm == 4) { vcg_re_bit(e, m4, 0); goto sign; }
        vcg_re_bit(e, m4, 1);
        if (mm == 5) { vcg_re_bit(e, m5, 0); goto sign; }
        vcg_re_bit(e, m5, 1);
        if (mm == 6) { vcg_re_bit(e, m6, 0); goto sign; }
        vcg_re_bit(e, m6, 1);
        if (mm == 7) { vcg_re_bit(e, m7, 0); goto sign; }
        vcg_re_bit(e, m7, 1);
        if (mm == 8) { vcg_re_bit(e, m8, 0); goto sign; }
        vcg_re_bit(e, m8, 1);
        if (mm == 9) { vcg_re_bit(e, m9, 0); goto sign; }
        vcg_re_bit(e, m9, 1);
        if (mm == 10) { vcg_re_bit(e, m10, 0); goto sign; }
        vcg_re_bit(e, m10, 1);
        if (mm == 11) { vcg_re_bit(e, m11, 0); goto sign; }
        vcg_re_bit(e, m11, 1);
        {
            const uint32_t t = mm - 11u + 1u;                  // EG0 of r = m-12
            const int nb = vcg_nbits(t);
            for (int b = 0; b < nb; ++b) vcg_re_bypass(e, 1);
            vcg_re_bypass(e, 0);
            for (int b = nb - 1; b >= 0; --b) vcg_re_bypass(e, (t >> b) & 1);
        }
    sign:
        vcg_re_bypass(e, v < 0);
    }
    for (int k = 0; k < 5; ++k) vcg_re_shift_low(e);
    *ovf = e.ovf;
    return e.pos;
}

// dead-zone quantizer, evaluated in source order with no contraction
// (step = base_q*g; dz = 0.8f*step; level = (int)((a-dz)/step+1)).
VCG_HD int vcg_quant1(int c, float base_q, float g) {
#pragma clang fp contract(off)
    float step = base_q * g;
    float dz = 0.80f * step;
    float a = (float)(c < 0 ? -c : c);
    int level = 0;
    if (a >= dz) level = (int)((a - dz) / step + 1.0f);
    if (level > 32767) level = 32767;
    return c < 0 ? -level : level;
}
