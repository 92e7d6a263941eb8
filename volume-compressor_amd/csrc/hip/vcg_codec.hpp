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

#if defined(__clang__)
#define VCG_UNPRED(x) __builtin_unpredictable(x)
#else
#define VCG_UNPRED(x) (x)
#endif

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

// Coefficient list entry: scan position << 17 | (2 m - 11 + sign) mod 2^17
// for a coefficient of magnitude m.  The reference keeps the level as an
// int16 (vc.c:559), so m mod 2^16 is all that matters, and 17 bits hold it
// with the sign.
VCG_HD uint32_t vcg_ent_make(uint32_t pos, uint32_t m, uint32_t neg) {
    return (pos << 17) | ((2u * m - 11u + neg) & 0x1FFFFu);
}
VCG_HD uint32_t vcg_ent_pos(uint32_t e) { return e >> 17; }
// the int16 level of an entry, as vc.c:559 stores it
VCG_HD int vcg_ent_level(uint32_t e) {
    const uint32_t t = (e + 11u) & 0x1FFFFu;              // 2 m + sign, mod 2^17
    const uint32_t m = t >> 1;
    return (int)(int16_t)(uint16_t)((t & 1u) ? 0u - m : m);
}

// Decode one atom's range-coder bytes (after the dc byte).  Writes the nonzero
// coefficients as list entries in increasing position to out (capacity 32768)
// and their count to *cnt.  Returns 0, or 1 when an
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
        const uint32_t neg = (uint32_t)vcg_rd_bypass(d);
        out[cnt++] = vcg_ent_make(pp, m, neg);
    }
    *cnt_out = cnt;
    return 0;
}

// ---------------------------------------------------------------------------
// Decoder as a one-bin-per-step state machine (the batched kernels' form).
//
// A wavefront decodes 64 atoms in lockstep, so the step is written as
// selects: every lane runs the same instructions whatever symbol it is in.
// Bin states S, numbered so that most transitions are S + 1:
//    0          significance bin of position pos (context psig, in a register)
//    1..12      magnitude unary bin with context mag[S-1]
//    13 + c     Exp-Golomb prefix bin (bypass) after c ones; 78..127 means the
//               prefix passed VCG_EG_PREFIX_CAP ones
//    255 - c    Exp-Golomb suffix bin (bypass) with c bits still to come, c >= 1
//    255        sign bin (bypass), which emits the coefficient
// Transitions: after bit 1, and after any suffix bin, S + 1.  After bit 0 in
// a magnitude or prefix state, 268 - max(S, 13): the sign bin after a unary
// magnitude or an empty prefix, else the suffix with c bits to come.  After
// bit 0 in state 0, and after the sign bin, state 0 at the next position.
// acc holds the magnitude: S in a unary state (so 12 = 11 + X, X = 1, when
// the prefix starts), then 2 acc - 11 + bit in every suffix bin
// (vc.c:491-506: X = 2X + bit, m = 11 + X) and in the sign bin, where
// 2 m - 11 + sign is the entry's low part.
//
// The probability of the next bin is resolved inside the step.  Context
// memory slot k+1 holds mag[k] (k < 12), slots 13 and 15 hold VCG_P_BYPASS
// and are never written.  A context bin writes its adapted probability to
// slot S, a bypass bin to the sink slot 14 (Ctx::at picks the slot), and the
// step reads the slot after the one it writes, ahead of the bit: that is the next probability after
// every bit 1 that does not start a new position (mag[S] after states 0..11,
// bypass after 12 and the bypass states); after a bit 0 it is psig at a new
// position, else bypass.  A bypass bin splits the range like a context bin
// with p = 2048, but keeps floor(range/2) for both bit values and adapts
// nothing (vc.c:434-441); bit 12 of P marks it and vanishes in P << 20.
//
// A lane that has reached EOB (or failed) is parked: code = 0 and an empty,
// never refilled window make every later bin a zero significance bin, which
// emits nothing.
//
// Band edges, EOB and the prefix cap are tested after each step only in
// blocks where some lane can meet one: a step advances pos by at most one
// position and a prefix state by at most one, so a block of k steps needs no
// test while every lane is more than k positions before its next edge and
// below state 78 - k (vcg_fsm_near).
//
// Input: a 64-bit little-endian window of 64 - c valid bits ending at stream
// byte E, c the bits consumed since the last top-up.  Every 8 steps (a step consumes at most 8 bits: one renormalisation per
// bin, see vcg_rd_bit) the window is topped up to 64 bits.  Refills alternate
// between two 16-byte look-ahead buffers, each loaded at the window end of its
// previous use, two refills (16 steps) before it is needed: E has moved by at
// most 8 bytes since, so the bytes to append lie inside it.  Bytes at or past
// s[n] are masked to zero (past-end bytes read as 0, vc.c:425); loads are
// clamped to s + n, so the buffer must extend 16 bytes past s[n-1].
//
// Output: each step writes its would-be entry into slot k of the lane's ring,
// k the entries emitted since the last flush, and advances k only on an
// emission; every 64 steps the caller copies ring slots [0, 22) to the list
// (at most 22 emissions fit in 64 steps: a coefficient takes at least 3
// bins) and k restarts at 0, so the ring addresses of a flush are constants,
// list stores are unmasked and entries past the count are overwritten by the
// next flush.  Flushing every 64 steps rather than every 16 (a 32-slot ring
// instead of 8) quarters the partly written list lines and the stores that
// carry them; those stores, contending with the transform kernel's traffic,
// held back the entropy waves that share their vector memory counter.  One register holds position << 17 | k << 8 | c: an add of three
// terms per step advances the position, k (an emission is also a new
// position) and the consumed bits c; the top-up clears c and the flush k.
// ---------------------------------------------------------------------------
#ifndef VCG_ENTROPY_FSM
#define VCG_ENTROPY_FSM 1
#endif
#define VCG_REFILL_STEPS 8
#define VCG_FLUSH_EVERY 4                       // loop iterations (of 16 steps) per flush
#define VCG_FLUSH_N 22u
#define VCG_RING 32u
// a flush covers 16 * VCG_FLUSH_EVERY steps, and a coefficient takes at
// least 3 bins; the ring holds the would-be entry past the last emission too,
// its slot index sits in bits 8..16 of the position register, and slot
// offsets OR into the lane's offset (a power of two)
static_assert(VCG_FLUSH_N >= (16u * VCG_FLUSH_EVERY + 2u) / 3u, "a flush must cover every emission of its steps");
static_assert(VCG_RING > VCG_FLUSH_N && (VCG_RING & (VCG_RING - 1u)) == 0u && VCG_RING <= 256u,
              "ring slots: a power of two above the flush's entries");
#define VCG_PARK_E 0x40000000u
#define VCG_P_BYPASS (2048u | 4096u)
#define VCG_S_PFX 13u
#define VCG_S_BAD (VCG_S_PFX + VCG_EG_PREFIX_CAP + 1u)      // 78
#define VCG_S_TAIL 128u
#define VCG_S_SIGN 255u
#define VCG_CTX_SLOTS 16u
#define VCG_POS_MASK 0xFFFE0000u

// mask keeping the stream bytes of the 8 at byte E (those before s[n])
VCG_HD uint64_t vcg_tail_mask(uint32_t E, uint32_t n) {
    const int32_t rem = (int32_t)n - (int32_t)E;
    const uint32_t keep = rem <= 0 ? 0u : rem >= 8 ? 8u : (uint32_t)rem;
    return keep >= 8u ? ~0ull : (1ull << (8u * keep)) - 1ull;
}

// 16 stream bytes starting at byte e (address clamped to s + n)
struct vcg_look {
    uint64_t lo, hi;
    uint32_t e;
};
VCG_HD void vcg_look_load(vcg_look &L, const uint8_t *s, uint32_t n, uint32_t e) {
    uint64_t v[2];
    __builtin_memcpy(v, s + (e < n ? e : n), 16);
    L.lo = v[0]; L.hi = v[1]; L.e = e;
}

struct vcg_fsm {
    uint32_t code, range;
    uint64_t w;                  // window, 64 - c valid bits, ending at stream byte E
    uint32_t E;
    uint32_t P, S, psig, acc;
    uint32_t pc;                 // scan position << 17 (mod 2^32) | k << 8 | c (see above)
    uint32_t lim17;              // next band edge / EOB << 17 (mod 2^32)
    uint32_t eob;
    uint32_t fl;                 // entries flushed to the list
    uint32_t band, bad;
};

// lim17 is odd, so no position ever meets it
VCG_HD void vcg_fsm_park(vcg_fsm &f) {
    f.code = 0; f.w = 0; f.E = VCG_PARK_E; f.S = 0; f.lim17 = 0xFFFFFFFFu; f.P = 2048u; f.psig = 2048u;
}
VCG_HD bool vcg_fsm_parked(const vcg_fsm &f) { return f.E >= VCG_PARK_E; }

// top the window up to 64 bits from L (loaded at L.e, E - L.e <= 8), then
// reload L at the new window end
VCG_HD void vcg_fsm_refill(vcg_fsm &f, vcg_look &L, const uint8_t *s, uint32_t n) {
    const uint32_t o = f.E - L.e < 8u ? f.E - L.e : 8u;   // >= 8 only for a parked lane
    const uint32_t a = 4u * o, b = 32u - a;                 // two shifts each, all < 64
    const uint64_t data = ((L.lo >> a) >> a) | ((L.hi << b) << b);
    const uint32_t c = f.pc & 0xFFu;                        // consumed bits, a multiple of 8, <= 64
    const uint64_t app = c ? (data & vcg_tail_mask(f.E, n)) << (64u - c) : 0ull;
    f.w |= app;
    f.E += c >> 3;
    f.pc &= ~0xFFu;
    vcg_look_load(L, s, n, f.E);
}

// a * b for a, b < 2^24
VCG_HD uint32_t vcg_mul24(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul24(a, b);
#else
    return a * b;
#endif
}

// low `width` bits of x (width 0..31)
VCG_HD uint32_t vcg_low_bits(uint32_t x, uint32_t width) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_ubfe(x, 0u, width);
#else
    return x & ((1u << width) - 1u);
#endif
}

// a + b + c as one v_add3_u32 (the compiler turns b + c of disjoint bits into
// an or first, which costs a second instruction)
VCG_HD uint32_t vcg_add3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t r;
    asm("v_add3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
#else
    return a + b + c;
#endif
}

// renormalise (at most one byte, range >= 2^17 here); returns the bits taken
// from the window, for the caller to add to the consumed count
VCG_HD uint32_t vcg_fsm_renorm_sh(vcg_fsm &f) {
    const uint32_t sh = (uint32_t)__builtin_clz(f.range) & 8u;
    f.code = (f.code << sh) | vcg_low_bits((uint32_t)f.w, sh);
    f.range <<= sh;
    f.w >>= sh;
    return sh;
}
VCG_HD void vcg_fsm_renorm(vcg_fsm &f) { f.pc += vcg_fsm_renorm_sh(f); }

// s: first range-coder byte; n: byte count.  Context memory slots 1..12 must
// hold 2048 and slots 13 and 15 VCG_P_BYPASS on entry.
VCG_HD void vcg_fsm_init(vcg_fsm &f, vcg_look &A, vcg_look &B, const uint8_t *s, uint32_t n,
                         const uint32_t *band_start) {
    f.w = 0; f.pc = 64u; f.E = 0;                     // an empty window: 64 bits consumed
    vcg_look_load(A, s, n, 0);
    vcg_fsm_refill(f, A, s, n);                       // bytes [0, 8): c = 0
    f.code = 0;
    for (int k = 0; k < 5; ++k) { f.code = (f.code << 8) | ((uint32_t)f.w & 0xFFu); f.w >>= 8; f.pc += 8u; }   // vc.c:406-409
    f.range = 0xFFFFFFFFu;
    vcg_fsm_refill(f, A, s, n);                       // A was reloaded at 8: bytes [8, 13), c = 0;
                                                      // the EOB renormalises at most twice
    uint32_t eob = 0;
    for (int k = 0; k < 16; ++k) {                    // vc.c:520-525
        f.range >>= 1;
        const uint32_t bit = f.code >= f.range;
        if (bit) f.code -= f.range;
        vcg_fsm_renorm(f);
        eob = (eob << 1) | bit;
    }
    B = A;                                            // both at the window end (13)
    eob = eob > 32768u ? 32768u : eob;                // vc.c:549
    f.eob = eob;
    f.P = 2048u; f.S = 0; f.psig = 2048u; f.acc = 0; f.pc &= 0xFFu; f.fl = 0;
    f.band = 0; f.bad = 0;
    f.lim17 = (band_start[1] < eob ? band_start[1] : eob) << 17;   // EOB 32768 wraps to 0: met only at the end
    if (eob == 0) vcg_fsm_park(f);
}

// true when the lane may meet a band edge, EOB or the prefix cap within the
// next k steps
VCG_HD bool vcg_fsm_near(const vcg_fsm &f, uint32_t k) {
    return (f.lim17 - (f.pc & VCG_POS_MASK) <= (k << 17)) | (f.S - (VCG_S_BAD - k) < VCG_S_TAIL - VCG_S_BAD + k);
}

// the state after a bin in state S with bit b (the transitions above); a new
// position (after a zero significance bin or the sign bin) is state 0
VCG_HD uint32_t vcg_fsm_next(uint32_t S, uint32_t b) {
    if (S == VCG_S_SIGN) return 0u;
    if (b || S >= VCG_S_TAIL) return S + 1u;
    if (S == 0u) return 0u;
    return 268u - (S > VCG_S_PFX ? S : VCG_S_PFX);
}

// one bin.  Ctx: h = at(S, S < 13) the handle of the slot a bin in state S
// writes (S, or the sink 14), ld_next(h) the next slot's value, st(h, v),
// nxt(S + 256 b) = vcg_fsm_next(S, b) (a table: two byte loads that do not
// wait on the bit replace the transition arithmetic);
// Ring: st(slot << 8, v) on the lane's output ring.  With CHECK, returns nonzero
// when pos reached a band edge or EOB, or the prefix passed its cap (the
// caller then runs vcg_fsm_edge); without, the caller has ruled these out.
template <bool CHECK, class Ctx, class Ring>
VCG_HD uint32_t vcg_fsm_step(vcg_fsm &f, Ctx &ctx, Ring &ring) {
    const uint32_t S = f.S, P = f.P;
    const bool isCtx = S < VCG_S_PFX;
    const uint32_t a = ctx.at(S, isCtx);
    const uint32_t R = ctx.ld_next(a);
    const uint32_t n0 = ctx.nxt(S), n1 = ctx.nxt(S + 256u);
    // binary decision (vc.c:414-441)
    const uint32_t r0 = (uint32_t)(((uint64_t)f.range * (P << 20)) >> 32);
    const bool nbit = f.code < r0, bit = !nbit;
    f.code = bit ? f.code - r0 : f.code;
    f.range = (bit && isCtx) ? f.range - r0 : r0;
    // adaptation (vc.c:420-431): bit 1: P - (P>>5), bit 0: P + ((4096-P)>>5);
    // both equal (t + 31 P) >> 5 with t = 31 for bit 1 and 4096 for bit 0
    const bool isSig = S == 0u;
    const uint32_t pn = ((bit ? 31u : 4096u) + vcg_mul24(31u, P)) >> 5;
    ctx.st(a, pn);
    f.psig = isSig ? pn : f.psig;
    const uint32_t sh = vcg_fsm_renorm_sh(f);
    // symbol bookkeeping
    const bool tail = S >= VCG_S_TAIL;
    const bool sgn = S == VCG_S_SIGN;
    const uint32_t acc2 = 2u * f.acc + (bit ? 0u - 10u : 0u - 11u);
    ring.st(f.pc & ((VCG_RING - 1u) << 8), (f.pc & VCG_POS_MASK) | (acc2 & 0x1FFFFu));
    f.S = bit ? n1 : n0;
    const bool adv = f.S == 0u;                     // a new position: after (sig, 0) or the sign bin
    // one add moves the position (every new position), the ring slot (an
    // emission, which is also a new position) and the consumed bits
    f.pc = vcg_add3(f.pc, sgn ? 0x20100u : adv ? 0x20000u : 0u, sh);
    f.acc = tail ? acc2 : (S < 12u ? S : 12u);
    f.P = adv ? f.psig : (bit ? R : VCG_P_BYPASS);
    if (!CHECK) return 0;
    return ((f.pc & VCG_POS_MASK) == f.lim17) | (f.S - VCG_S_BAD < VCG_S_TAIL - VCG_S_BAD);
}

// band boundary / EOB / Exp-Golomb cap for a lane whose step returned nonzero
VCG_HD void vcg_fsm_edge(vcg_fsm &f, const uint32_t *band_start) {
    const bool bad = f.S - VCG_S_BAD < VCG_S_TAIL - VCG_S_BAD;
    const bool end = bad || (f.pc & VCG_POS_MASK) == f.eob << 17;
    // a new band starts from the reset context (vc.c:552-555) and its first
    // bin is a significance bin; a parked lane runs on the same values
    f.psig = 2048u;
    f.P = 2048u;
    if (end) {
        f.bad |= bad ? 1u : 0u;
        f.code = 0; f.w = 0; f.E = VCG_PARK_E; f.S = 0; f.lim17 = 0xFFFFFFFFu;
    } else {
        f.band++;
        const uint32_t nb = band_start[f.band + 1];
        f.lim17 = (nb < f.eob ? nb : f.eob) << 17;
    }
}

// 8 steps, tested for edges only when `check`
template <class Ctx, class Ring>
VCG_HD void vcg_fsm_block(vcg_fsm &f, Ctx &ctx, Ring &ring, const uint32_t *band_start, bool check) {
    if (check) {
#pragma unroll
        for (int k = 0; k < VCG_REFILL_STEPS; ++k)
            if (vcg_fsm_step<true>(f, ctx, ring)) vcg_fsm_edge(f, band_start);
    } else {
#pragma unroll
        for (int k = 0; k < VCG_REFILL_STEPS; ++k) vcg_fsm_step<false>(f, ctx, ring);
    }
}

// copy ring slots [0, VCG_FLUSH_N) to out + flushed (entries past the count
// are placeholders that a later flush overwrites; out needs VCG_FLUSH_N - 1
// slack entries past the list capacity) and restart the ring.  Returns the
// number of real entries (at most VCG_FLUSH_N).
template <class Ring>
VCG_HD uint32_t vcg_fsm_flush(vcg_fsm &f, Ring &ring, uint32_t *out) {
    const uint32_t pending = (f.pc >> 8) & 0x1FFu;
    uint32_t *o = out + f.fl;
#pragma unroll
    for (uint32_t k = 0; k < VCG_FLUSH_N; ++k) o[k] = ring.ld(k << 8);
    f.fl += pending;
    f.pc &= ~0x1FF00u;
    return pending;
}

// Whole-atom decode with the state machine, driven as the kernel drives it
// (CPU-side reference of the kernel's per-lane logic).  always_check tests
// for edges in every block, as a lane does when another lane of its wave is
// near one.  Returns 0, 1 for an over-long Exp-Golomb prefix, or 2 if a flush
// ever held more than VCG_FLUSH_N entries (never, by the 3-bins-per-
// coefficient bound).
struct vcg_ctx_array {
    uint32_t v[VCG_CTX_SLOTS];
    VCG_HDM uint32_t at(uint32_t S, bool isCtx) const { return isCtx ? S : 14u; }
    VCG_HDM uint32_t ld_next(uint32_t h) const { return v[h + 1u]; }
    VCG_HDM uint32_t ld(uint32_t k) const { return v[k]; }
    VCG_HDM void st(uint32_t h, uint32_t p) { v[h] = p; }
    VCG_HDM uint32_t nxt(uint32_t i) const { return vcg_fsm_next(i & 255u, i >> 8); }
};
struct vcg_ring_array {
    uint32_t v[VCG_RING];
    VCG_HDM uint32_t ld(uint32_t off) const { return v[off >> 8]; }
    VCG_HDM void st(uint32_t off, uint32_t p) { v[off >> 8] = p; }
};
VCG_HD int vcg_rc_decode_atom_fsm(const uint8_t *s, uint32_t n, const uint32_t *band_start, uint32_t *out,
                                  uint32_t *cnt_out, bool always_check) {
    vcg_ctx_array ctx;
    vcg_ring_array ring;
    for (uint32_t k = 0; k < VCG_CTX_SLOTS; ++k) ctx.v[k] = k == 13u || k == 15u ? VCG_P_BYPASS : 2048u;
    for (uint32_t k = 0; k < VCG_RING; ++k) ring.v[k] = 0xDEADBEEFu;
    vcg_fsm f;
    vcg_look A, B;
    vcg_fsm_init(f, A, B, s, n, band_start);
    int over = 0;
    for (uint32_t it = 0; !vcg_fsm_parked(f); ++it) {
        vcg_fsm_refill(f, A, s, n);
        if (it % VCG_FLUSH_EVERY == VCG_FLUSH_EVERY - 1 && vcg_fsm_flush(f, ring, out) > VCG_FLUSH_N) over = 1;
        vcg_fsm_block(f, ctx, ring, band_start, always_check || vcg_fsm_near(f, VCG_REFILL_STEPS));
        vcg_fsm_refill(f, B, s, n);
        vcg_fsm_block(f, ctx, ring, band_start, always_check || vcg_fsm_near(f, VCG_REFILL_STEPS));
    }
    if (vcg_fsm_flush(f, ring, out) > VCG_FLUSH_N) over = 1;
    *cnt_out = f.bad ? 0 : f.fl;
    return over ? 2 : (int)f.bad;
}

// ---------------------------------------------------------------------------
// encoder
// ---------------------------------------------------------------------------
struct vcg_renc {
    uint8_t *out;               // NULL: count bytes only
    uint32_t pos, cap;
    uint64_t low;
    uint32_t range;
    uint32_t cache;
    uint32_t csize;
    int ovf;
    // optional shared byte budget: every 1024 bytes a lane adds 1024 to *kb;
    // once the total passes `budget` the encode stops (stop = 1)
    uint32_t *kb;
    uint32_t budget;
    int stop;
};
VCG_HD uint32_t vcg_atomic_add(uint32_t *p, uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return atomicAdd(p, v);
#else
    const uint32_t o = *p; *p = o + v; return o;
#endif
}
VCG_HD void vcg_re_put(vcg_renc &e, uint32_t b) {
    if (e.pos < e.cap) { if (e.out) e.out[e.pos] = (uint8_t)b; }
    else e.ovf = 1;
    e.pos++;
    if (e.kb && (e.pos & 1023u) == 0u && vcg_atomic_add(e.kb, 1024u) + 1024u > e.budget) e.stop = 1;
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

// Sequential reader of the 32768 scan-order levels of one atom (16-byte
// aligned): 8 levels per 16-byte load, the next block requested one block
// ahead, so a lane's load latency hides behind the 8 positions before it.
struct vcg_lvread {
    struct b16 { uint32_t w[4]; };
    const b16 *blk;
    b16 cur, nxt;
    VCG_HDM void init(const int16_t *L) {
        blk = (const b16 *)L;
        cur = blk[0];
        nxt = blk[1];
    }
    // level at position p; positions must be requested in increasing order, one by one
    VCG_HDM int get(uint32_t p) {
        if ((p & 7u) == 0u && p) {
            cur = nxt;
            const uint32_t b = (p >> 3) + 1u;
            if (b < 4096u) nxt = blk[b];
        }
        const uint32_t q = (p >> 1) & 3u;
        const uint32_t w = q == 0 ? cur.w[0] : q == 1 ? cur.w[1] : q == 2 ? cur.w[2] : cur.w[3];
        return (int)(int16_t)(uint16_t)((p & 1u) ? w >> 16 : w);
    }
};

// Encode levels given in scan order (L[p], p < eob) into out (capacity cap;
// out NULL: count only).  Returns the byte count; *ovf set if cap was
// exceeded.  vc.c:513-544,399-401.  kb/budget: optional shared byte budget
// (see vcg_renc); *stopped set when the encode gave up past it.
VCG_HD uint32_t vcg_rc_encode_atom(const int16_t *L, uint32_t eob, const uint32_t *band_start, uint8_t *out,
                                   uint32_t cap, int *ovf, uint32_t *kb = nullptr, uint32_t budget = 0,
                                   int *stopped = nullptr) {
    vcg_lvread rd;
    rd.init(L);
    vcg_renc e;
    e.out = out; e.pos = 0; e.cap = cap; e.low = 0; e.range = 0xFFFFFFFFu; e.cache = 0; e.csize = 1; e.ovf = 0;
    e.kb = kb; e.budget = budget; e.stop = 0;
    for (int b = 15; b >= 0; --b) vcg_re_bypass(e, (eob >> b) & 1);
    uint32_t m0 = 2048u, m1 = 2048u, m2 = 2048u, m3 = 2048u, m4 = 2048u, m5 = 2048u,
             m6 = 2048u, m7 = 2048u, m8 = 2048u, m9 = 2048u, m10 = 2048u, m11 = 2048u;
    uint32_t psig = 2048u;
    int band = 0;
    uint32_t nextb = band_start[1];
    for (uint32_t p = 0; p < eob; ++p) {
        if (e.stop) break;
        while (p >= nextb) { ++band; nextb = band_start[band + 1]; psig = 2048u; }
        const int v = rd.get(p);
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
    if (stopped) *stopped = e.stop;
    return e.pos;
}

// ---------------------------------------------------------------------------
// Encoder as a one-bin-per-step state machine (the batched encode kernels'
// form), the mirror of the decoder's: the same bin states and transitions,
// each bin's bit taken from the level v being coded, with a = |v|,
// t = a - 11 and T = 2t + sign (vc.c:471-490, 528-544):
//    0..12      significance and unary magnitude bins: a > S
//    13 + c     Exp-Golomb prefix after c ones: t >= 2^(c+1), i.e.
//               a > 10 + (2 << c) (nb = floor(log2 t) ones in all)
//    255 - c    suffix bin with c bits to come: bit c-1 of t = bit c of T
//    255        sign bin: bit 0 of T
// so in a tail state (S >= 128) the bit is bit 255 - S of T: the top bit of
// T << (S mod 32).  Range coder (vc.c:364-401): low + range never exceeds its
// value at the last shift, so low stays below 2^33 and a carry is bit 32;
// after a bin range >= 2^17, so one 8-bit shift renormalises.  A shift with no
// carry and low >= 0xFF000000 only lengthens the pending run (csize);
// otherwise the pending byte (cache + carry) and csize - 1 bytes 0xFF + carry
// go out.  The count-only form keeps no low at all: every shift eventually
// puts out exactly one byte (bytes put + csize = 1 + shifts, and the last of
// the five flush shifts always flushes, leaving csize = 1), so the byte count
// is the number of shifts, which depends on range alone.
//
// Levels reach the lane through a ring of four 8-level blocks (Lv: ld(p) the
// level at scan position p, st(k, blk) block k into ring slot k & 3).  Every
// VCG_EFSM_REFILL steps the ring is brought to blocks k0 .. k0 + 2 with
// k0 = p >> 3 (position p advances at most once per step, so the ring always
// holds p + 1), from a block loaded one refill earlier; the load for the next
// refill is issued unconditionally, so no wait falls on a load in flight.  A
// finished lane records its byte count and keeps stepping with no effect on
// the result: its output capacity drops to zero and it puts nothing more.
// Band edges and EOB are tested after a step only in blocks where some lane
// can meet one (vcg_efsm_near), as in the decoder.
// ---------------------------------------------------------------------------
#define VCG_EFSM_REFILL 8
struct vcg_blk16 { uint32_t w[4]; };
struct vcg_efsm {
    uint64_t low;
    uint32_t range, cache, csize, pos;   // pos: bytes put out
    uint32_t S, P, psig;
    uint32_t p, lim, eob, band;
    int v;
    uint32_t done, len;                  // len: final byte count once done
    uint32_t cap;                        // output capacity (0 once done)
    int ovf;
    uint32_t k0;                         // ring holds blocks k0 .. k0 + 2
    const vcg_blk16 *blk;
    vcg_blk16 pref;                      // block k0 + 3, loaded at the last refill
};

template <bool WRITE>
VCG_HD void vcg_efsm_put(vcg_efsm &f, uint8_t *out, uint32_t b) {
    if (f.done) return;
    if (WRITE) {
        if (f.pos < f.cap) out[f.pos] = (uint8_t)b;
        else f.ovf = 1;
    }
    f.pos++;
}

// general shift of low (vc.c:369-380), for the EOB prefix and the final flush
template <bool WRITE>
VCG_HD void vcg_efsm_shift(vcg_efsm &f, uint8_t *out) {
    const uint32_t hi = (uint32_t)(f.low >> 32), lo = (uint32_t)f.low;
    if (hi != 0u || lo < 0xFF000000u) {
        vcg_efsm_put<WRITE>(f, out, (f.cache + hi) & 0xFFu);
        for (uint32_t k = 1; k < f.csize; ++k) vcg_efsm_put<WRITE>(f, out, (0xFFu + hi) & 0xFFu);
        f.cache = lo >> 24;
        f.csize = 1;
    } else {
        f.csize++;
    }
    f.low = (uint64_t)(lo << 8);
}

// the five-shift flush (vc.c:399-401); the lane is finished afterwards (and
// puts nothing more)
template <bool WRITE>
VCG_HD void vcg_efsm_finish(vcg_efsm &f, uint8_t *out) {
    if (WRITE) for (int k = 0; k < 5; ++k) vcg_efsm_shift<WRITE>(f, out);
    else f.pos += 5;                                  // count: one byte per shift
    f.len = f.pos;
    f.done = 1;
    f.lim = 0xFFFFFFFFu;
}

// a flush's pending run: csize - 1 bytes 0xFF + carry after the cache byte
template <bool WRITE>
VCG_HD void vcg_efsm_run(vcg_efsm &f, uint8_t *out, uint32_t hi) {
    for (uint32_t k = 1; k < f.csize; ++k) vcg_efsm_put<WRITE>(f, out, (0xFFu + hi) & 0xFFu);
}

VCG_HD vcg_blk16 vcg_efsm_load(const vcg_efsm &f, uint32_t k) { return f.blk[k < 4095u ? k : 4095u]; }

// L: scan-order levels (16-byte aligned, 32768 of them), eob <= 32768.
// out: the range-coder bytes (WRITE), capacity cap.  Context memory slots
// 1..12 must hold 2048 and slots 13 and 15 VCG_P_BYPASS on entry.
template <bool WRITE, class Lv>
VCG_HD void vcg_efsm_init(vcg_efsm &f, Lv &lv, const int16_t *L, uint32_t eob, const uint32_t *band_start,
                          uint8_t *out, uint32_t cap) {
    f.low = 0; f.range = 0xFFFFFFFFu; f.cache = 0; f.csize = 1; f.pos = 0;
    f.cap = cap; f.ovf = 0; f.done = 0; f.len = 0;
    for (int b = 15; b >= 0; --b) {                   // EOB, 16 bypass bits (vc.c:513-519)
        f.range >>= 1;
        if ((eob >> b) & 1u) f.low += f.range;
        if (f.range < (1u << 24)) {
            if (WRITE) vcg_efsm_shift<WRITE>(f, out);
            else f.pos++;
            f.range <<= 8;
        }
    }
    f.S = 0; f.P = 2048u; f.psig = 2048u;
    f.p = 0; f.eob = eob; f.band = 0;
    f.lim = band_start[1] < eob ? band_start[1] : eob;
    f.blk = (const vcg_blk16 *)L;
    f.k0 = 0;
    for (uint32_t k = 0; k < 3; ++k) lv.st(k, vcg_efsm_load(f, k));
    f.pref = vcg_efsm_load(f, 3);
    f.v = lv.ld(0);
    if (eob == 0) vcg_efsm_finish<WRITE>(f, out);
}

// every VCG_EFSM_REFILL steps: ring to blocks p>>3 .. (p>>3) + 2
template <class Lv>
VCG_HD void vcg_efsm_refill(vcg_efsm &f, Lv &lv) {
    const uint32_t k = f.p >> 3;                      // k0 or k0 + 1
    if (k != f.k0) lv.st(k + 2u, f.pref);
    f.k0 = k;
    f.pref = vcg_efsm_load(f, k + 3u);
}

// true when the lane may meet a band edge or EOB within the next k steps
VCG_HD bool vcg_efsm_near(const vcg_efsm &f, uint32_t k) { return f.lim - f.p <= k; }

// one bin.  Ctx as for the decoder step.  With CHECK, returns nonzero when p
// reached a band boundary or the EOB (the caller then runs vcg_efsm_edge);
// without, the caller has ruled that out.
template <bool CHECK, bool WRITE, class Ctx, class Lv>
VCG_HD uint32_t vcg_efsm_step(vcg_efsm &f, Ctx &ctx, Lv &lv, uint8_t *out) {
    const uint32_t S = f.S, P = f.P;
    const bool isCtx = S < VCG_S_PFX;
    const uint32_t h = ctx.at(S, isCtx);
    const uint32_t R = ctx.ld_next(h);
    const uint32_t n0 = ctx.nxt(S), n1 = ctx.nxt(S + 256u);   // vcg_fsm_next(S, 0 / 1)
    const int vn = lv.ld(f.p + 1u);                   // the next level, in case p advances
    const int v = f.v;
    const uint32_t a = (uint32_t)(v < 0 ? -v : v);
    // the bit (lanes of a wave sit in different states: selects, no branch)
    const uint32_t thr = isCtx ? S : (2u << ((S + 19u) & 31u)) + 10u;   // prefix: 10 + (2 << (S - 13))
    const uint32_t T = 2u * a - 22u + (v < 0 ? 1u : 0u);
    const bool tail = S >= VCG_S_TAIL;
    const bool bT = (int32_t)(T << (S & 31u)) < 0, bA = a > thr;
    const bool bit = (tail & bT) | (!tail & bA);
    // binary decision (vc.c:381-398); a bypass bin is a context bin with
    // p = 2048 that keeps floor(range/2) for both bits (bit 12 of P)
    const uint32_t r0 = (uint32_t)(((uint64_t)f.range * (P << 20)) >> 32);
    if (WRITE) f.low += bit ? (uint64_t)r0 : 0ull;
    f.range = (bit && isCtx) ? f.range - r0 : r0;
    // adaptation as in the decoder step: (t + 31 P) >> 5, t = 31 or 4096
    const bool isSig = S == 0u;
    const uint32_t pn = ((bit ? 31u : 4096u) + vcg_mul24(31u, P)) >> 5;
    ctx.st(h, pn);
    f.psig = isSig ? pn : f.psig;
    // renormalisation.  A flush puts out the cache byte (plus, rarely, a
    // pending run).  That byte is stored every step, unmasked, at out[pos], or
    // at out[cap] (a spare byte past the capacity) once the lane is finished
    // or full; pos advances only on a flush, so the store of a step without
    // one is overwritten by the next.  out must hold cap + 1 bytes.
    const bool sh = f.range < (1u << 24);
    if (WRITE) {
        const uint32_t hi = (uint32_t)(f.low >> 32), lo = (uint32_t)f.low;
        const bool fl = sh && (hi != 0u || lo < 0xFF000000u);
        const bool put = fl && !f.done;
        out[!f.done && f.pos < f.cap ? f.pos : f.cap] = (uint8_t)(f.cache + hi);
        f.ovf = put && f.pos >= f.cap ? 1 : f.ovf;
        f.pos += put ? 1u : 0u;
        if (fl && f.csize > 1u) vcg_efsm_run<WRITE>(f, out, hi);
        f.cache = fl ? lo >> 24 : f.cache;
        f.csize = fl ? 1u : f.csize + (sh ? 1u : 0u);
        f.low = sh ? (uint64_t)(lo << 8) : f.low;
    } else {
        f.pos += sh ? 1u : 0u;                        // count: one byte per shift
    }
    f.range = sh ? f.range << 8 : f.range;
    // symbol bookkeeping (the decoder's transitions, from its table)
    f.S = bit ? n1 : n0;
    const bool adv = f.S == 0u;                       // a new position
    f.P = adv ? f.psig : (bit ? R : VCG_P_BYPASS);
    f.p += adv ? 1u : 0u;
    f.v = adv ? vn : v;
    if (!CHECK) return 0;
    return f.p == f.lim;
}

// band boundary or EOB for a lane whose step returned nonzero
template <bool WRITE>
VCG_HD void vcg_efsm_edge(vcg_efsm &f, const uint32_t *band_start, uint8_t *out) {
    if (f.p >= f.eob) {
        vcg_efsm_finish<WRITE>(f, out);
    } else {
        // a new band starts from the reset context (vc.c:552-555)
        uint32_t lim = f.lim, band = f.band;
        while (f.p == lim) {
            band++;
            const uint32_t nb = band_start[band + 1];
            lim = nb < f.eob ? nb : f.eob;
        }
        f.lim = lim;
        f.band = band;
    }
    // the next bin is a significance bin (harmless for a finished lane)
    f.psig = 2048u;
    f.P = 2048u;
}

// VCG_EFSM_REFILL steps, tested for edges only when `check`
template <bool WRITE, class Ctx, class Lv>
VCG_HD void vcg_efsm_block(vcg_efsm &f, Ctx &ctx, Lv &lv, const uint32_t *band_start, uint8_t *out, bool check) {
    if (check) {
#pragma unroll
        for (int k = 0; k < VCG_EFSM_REFILL; ++k)
            if (vcg_efsm_step<true, WRITE>(f, ctx, lv, out)) vcg_efsm_edge<WRITE>(f, band_start, out);
    } else {
#pragma unroll
        for (int k = 0; k < VCG_EFSM_REFILL; ++k) vcg_efsm_step<false, WRITE>(f, ctx, lv, out);
    }
}

// the lane's level ring as a plain array (CPU reference)
struct vcg_lv_array {
    int16_t v[32];
    VCG_HDM int ld(uint32_t p) const { return v[p & 31u]; }
    VCG_HDM void st(uint32_t k, const vcg_blk16 &b) {
        for (int q = 0; q < 4; ++q) {
            v[((k & 3u) << 3) + 2 * q] = (int16_t)(uint16_t)b.w[q];
            v[((k & 3u) << 3) + 2 * q + 1] = (int16_t)(uint16_t)(b.w[q] >> 16);
        }
    }
};

// Whole-atom encode with the state machine, driven as the kernels drive it
// (CPU-side reference of the kernels' per-lane logic): a ring refill, then a
// block of VCG_EFSM_REFILL steps, the shared byte budget checked between
// blocks.  always_check tests for edges in every block, as a lane does when
// another lane of its wave is near one.  Same results as vcg_rc_encode_atom
// (out NULL: count only).
VCG_HD uint32_t vcg_rc_encode_atom_fsm(const int16_t *L, uint32_t eob, const uint32_t *band_start, uint8_t *out,
                                       uint32_t cap, int *ovf, uint32_t *kb = nullptr, uint32_t budget = 0,
                                       int *stopped = nullptr, bool always_check = false) {
    vcg_ctx_array ctx;
    for (uint32_t k = 0; k < VCG_CTX_SLOTS; ++k) ctx.v[k] = k == 13u || k == 15u ? VCG_P_BYPASS : 2048u;
    vcg_lv_array lv;
    vcg_efsm f;
    int stop = 0;
    uint32_t kbm = 1024u;
    if (out) {
        vcg_efsm_init<true>(f, lv, L, eob, band_start, out, cap);
        while (!f.done) {
            vcg_efsm_refill(f, lv);
            vcg_efsm_block<true>(f, ctx, lv, band_start, out, always_check || vcg_efsm_near(f, VCG_EFSM_REFILL));
        }
    } else {
        vcg_efsm_init<false>(f, lv, L, eob, band_start, nullptr, cap);
        while (!f.done) {
            vcg_efsm_refill(f, lv);
            vcg_efsm_block<false>(f, ctx, lv, band_start, nullptr,
                                  always_check || vcg_efsm_near(f, VCG_EFSM_REFILL));
            while (kb && !f.done && f.pos >= kbm) {
                kbm += 1024u;
                if (vcg_atomic_add(kb, 1024u) + 1024u > budget) { stop = 1; f.done = 1; f.lim = 0xFFFFFFFFu; }
            }
        }
        if (f.len > cap) f.ovf = 1;
    }
    *ovf = f.ovf;
    if (stopped) *stopped = stop;
    return f.len;
}

// dead-zone quantizer, evaluated in source order with no contraction
// (step = base_q*g; dz = 0.8f*step; level = (int)((a-dz)/step+1)).
// with the step and dead zone of the coefficient's frequency precomputed
// (step = base_q * g, dz = 0.80f * step, as below)
VCG_HD int vcg_quant1s(int c, float step, float dz) {
#pragma clang fp contract(off)
    float a = (float)(c < 0 ? -c : c);
    int level = 0;
    if (a >= dz) level = (int)((a - dz) / step + 1.0f);
    if (level > 32767) level = 32767;
    return c < 0 ? -level : level;
}
VCG_HD int vcg_quant1(int c, float base_q, float g) {
#pragma clang fp contract(off)
    float step = base_q * g;
    float dz = 0.80f * step;
    return vcg_quant1s(c, step, dz);
}
