// CPU build of the device codec code (volume-compressor_amd/csrc/hip/
// vcg_codec.hpp, vcg_xform.hpp): the per-atom range coder and the transform
// kernels' phases, run thread by thread between the kernels' barriers, so the
// exact kernel logic is checked against the oracle on the host.  Test
// infrastructure.
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#include "vcg_xform.hpp"

static const int NT = 512;   // threads per workgroup in the forward kernel
static const int NTI = 512;  // threads per workgroup in the inverse kernel

extern "C" {
uint32_t kc_encode(const int16_t *L, uint32_t eob, const uint32_t *band_start, uint8_t *out, uint32_t cap,
                   int *ovf) {
    return vcg_rc_encode_atom(L, eob, band_start, out, cap, ovf);
}
// the kernels' state-machine encoder (out NULL: count only); levels are
// copied to a 16-byte aligned buffer as the kernels hold them.
// always_check: test for edges in every block
uint32_t kc_encode_fsm(const int16_t *L, uint32_t eob, const uint32_t *band_start, uint8_t *out, uint32_t cap,
                       int *ovf, uint32_t *kb, uint32_t budget, int *stopped, int always_check) {
    int16_t *a = (int16_t *)aligned_alloc(16, 32768 * 2);
    memcpy(a, L, 32768 * 2);
    uint32_t n = vcg_rc_encode_atom_fsm(a, eob, band_start, out, cap, ovf, kb, budget, stopped, always_check != 0);
    free(a);
    return n;
}
int kc_decode(const uint8_t *s, uint32_t n, const uint32_t *band_start, uint32_t *out, uint32_t *cnt) {
    return vcg_rc_decode_atom(s, n, band_start, out, cnt);
}
// the kernels' state-machine decoder; the stream is copied to offset `mis` of
// an 8-aligned, zero-padded buffer (as in the padded device archive).
// always_check: test for edges in every block, as a lane does when another
// lane of its wave is near one
int kc_decode_fsm(const uint8_t *s, uint32_t n, uint32_t mis, const uint32_t *band_start, uint32_t *out,
                  uint32_t *cnt, int always_check) {
    size_t words = (mis + n + 7) / 8 + 4;          // 24+ bytes of slack past the stream
    uint64_t *buf = (uint64_t *)calloc(words, 8);
    memset(buf, 0xA5, words * 8);                 // garbage around the stream must be masked
    memcpy((uint8_t *)buf + mis, s, n);
    int r = vcg_rc_decode_atom_fsm((const uint8_t *)buf + mis, n, band_start, out, cnt, always_check != 0);
    free(buf);
    return r;
}
// One wavefront of k_entropy, its lanes in lockstep as on the GPU: lane k
// decodes the range-coder bytes [offs[k] + 1, offs[k] + lens[k]) of buf, a
// padded archive-like pool where the bytes past a stream belong to the next
// payload.  Every lane, parked or not, runs every refill, flush and block
// until no lane is active (k_entropy's loop), so what a parked lane does
// with the bytes past its stream shows.  out: nl lists of 32768 + 64
// entries; cnt, bad per lane.  variant 1 plants a refill that appends the
// look-ahead bytes unmasked when no active lane is within 16 bytes of its
// stream end (a variant once tried on the device): a parked lane then takes
// in the next payload's bytes and decodes them as bins.
static void wave_refill(vcg_fsm &f, vcg_look &L, const uint8_t *s, uint32_t n, bool mask) {
    if (mask) { vcg_fsm_refill(f, L, s, n); return; }
    const uint32_t o = f.E - L.e < 8u ? f.E - L.e : 8u;
    const uint32_t a = 4u * o, b = 32u - a;
    const uint64_t data = ((L.lo >> a) >> a) | ((L.hi << b) << b);
    const uint32_t c = f.pc & 0xFFu;
    f.w |= c ? data << (64u - c) : 0ull;
    f.E += c >> 3;
    f.pc &= ~0xFFu;
    vcg_look_load(L, s, n, f.E);
}
int kc_decode_wave(const uint8_t *buf, const uint32_t *offs, const uint32_t *lens, int nl,
                   const uint32_t *band_start, uint32_t *out, uint32_t *cnt, int *bad, int variant) {
    if (nl < 1 || nl > 64) return -1;
    const size_t LS = 32768 + 64;
    vcg_ctx_array ctx[64];
    vcg_ring_array ring[64];
    vcg_fsm f[64];
    vcg_look A[64], B[64];
    const uint8_t *s[64];
    uint32_t n[64];
    // a lane of the planted variant can go on emitting past any real list:
    // its flushes then go to a sink and the lane reports bad = 3
    static uint32_t sink[64];
    int over[64] = {0};
    auto lane_out = [&](int k) -> uint32_t * {
        if (f[k].fl + VCG_FLUSH_N > LS) over[k] = 1;
        if (over[k]) { f[k].fl = 0; return sink; }
        return out + k * LS;
    };
    for (int k = 0; k < nl; ++k) {
        for (uint32_t c = 0; c < VCG_CTX_SLOTS; ++c) ctx[k].v[c] = c == 13u || c == 15u ? VCG_P_BYPASS : 2048u;
        for (uint32_t r = 0; r < VCG_RING; ++r) ring[k].v[r] = 0xDEADBEEFu;
        s[k] = buf + offs[k] + 1;
        n[k] = lens[k] - 1;
        vcg_fsm_init(f[k], A[k], B[k], s[k], n[k], band_start);
    }
    auto any_active = [&] { for (int k = 0; k < nl; ++k) if (!vcg_fsm_parked(f[k])) return true; return false; };
    auto any_near = [&] { for (int k = 0; k < nl; ++k) if (vcg_fsm_near(f[k], VCG_REFILL_STEPS)) return true; return false; };
    auto masked = [&] {
        if (variant != 1) return true;
        for (int k = 0; k < nl; ++k)
            if (!vcg_fsm_parked(f[k]) && f[k].E + 16u > n[k]) return true;
        return false;
    };
    for (uint32_t it = 0; any_active(); ++it) {
        bool m = masked();
        for (int k = 0; k < nl; ++k) wave_refill(f[k], A[k], s[k], n[k], m);
        if (it % VCG_FLUSH_EVERY == VCG_FLUSH_EVERY - 1)
            for (int k = 0; k < nl; ++k) vcg_fsm_flush(f[k], ring[k], lane_out(k));
        bool chk = any_near();
        for (int k = 0; k < nl; ++k) vcg_fsm_block(f[k], ctx[k], ring[k], band_start, chk);
        m = masked();
        for (int k = 0; k < nl; ++k) wave_refill(f[k], B[k], s[k], n[k], m);
        chk = any_near();
        for (int k = 0; k < nl; ++k) vcg_fsm_block(f[k], ctx[k], ring[k], band_start, chk);
    }
    for (int k = 0; k < nl; ++k) {
        vcg_fsm_flush(f[k], ring[k], lane_out(k));
        bad[k] = over[k] ? 3 : (int)f[k].bad;
        cnt[k] = bad[k] ? 0 : f[k].fl;
    }
    return 0;
}
void kc_quant(const int32_t *c, size_t n, float q, float g, int32_t *out) {
    for (size_t i = 0; i < n; ++i) out[i] = vcg_quant1(c[i], q, g);
}

static void mfma_cpu(const uint32_t A[64][4], const uint32_t B[64][4], int32_t D[64][16]);

// k_fdct on one atom of an (nx,ny,nz) volume: returns 1 for an all-zero atom.
// The three passes follow the kernel's matrix-core form lane by lane (each
// product emulated as v_mfma_i32_32x32x32_i8 computes it): pass 1 from the
// 64 lanes' 16-voxel rows (vcg_fwd1_row, top bit flipped) into the i16 tile,
// passes 2 and 3 from the tile's byte planes, every block read before any
// block is written (the kernel's barriers); pass 3 into the i32 tile.
int kc_fdct(const uint8_t *vol, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t ax, uint32_t ay, uint32_t az,
            const int32_t cmat[32][32], const uint16_t *scan, const float *gstep, float q, int16_t *lv,
            int32_t *coef_raster, uint32_t *eob, uint32_t *dc_out) {
    int *t = (int *)aligned_alloc(16, VCG_TILE_INTS * sizeof(int));
    uint8_t *tile = (uint8_t *)t;
    static int8_t fa[3][64][16];
    static int32_t rowsum[32];
    static uint32_t raw[32][64][4], F[3][64][4];
    vcg_fwd1_tables(cmat, fa, rowsum);
    memcpy(F, fa, sizeof F);
    uint32_t sum = 0;
    for (int z = 0; z < 32; ++z)
        for (int l = 0; l < 64; ++l) {
            vcg_fwd1_row(vol, nx, ny, nz, ax, ay, az, z, l & 31, l >> 5, raw[z][l]);
            for (int e = 0; e < 16; ++e) sum += (raw[z][l][e >> 2] >> (8 * (e & 3))) & 0xFFu;
            for (int w = 0; w < 4; ++w) raw[z][l][w] ^= 0x80808080u;
        }
    if (sum == 0) { free(t); *eob = 0; *dc_out = 0; return 1; }
    const int dc = (int)((sum + 16384u) / 32768u);
    static int32_t D[64][16];
    // pass 1, scaled by 4: D[y][kx] = (v-128) x Ch, << 8 + 4 corr, + (v-128) x 4 Cl
    for (int z = 0; z < 32; ++z) {
        memset(D, 0, sizeof D);
        mfma_cpu(raw[z], F[0], D);
        for (int l = 0; l < 64; ++l)
            for (int r = 0; r < 16; ++r)
                D[l][r] = (int32_t)(((uint32_t)D[l][r] << 8) + 4u * (uint32_t)vcg_fwd1_corr(dc, rowsum[l & 31]));
        mfma_cpu(raw[z], F[2], D);
        for (int l = 0; l < 64; ++l)
            for (int qq = 0; qq < 4; ++qq) {
                uint32_t w[2];
                vcg_im_pack16h(D[l] + 4 * qq, w);
                memcpy(tile + vcg_im_out_addr(0, z, l & 31, l >> 5, qq), w, 8);
            }
    }
    static uint32_t d[32][64][8], hi[64][4], lo[64][4];
    for (int p = 1; p < 3; ++p) {
        for (int B = 0; B < 32; ++B)
            for (int l = 0; l < 64; ++l) vcg_im_load(tile, p, B, l & 31, l >> 5, d[B][l]);
        for (int B = 0; B < 32; ++B) {
            for (int l = 0; l < 64; ++l) vcg_im_planes(d[B][l], hi[l], lo[l]);
            memset(D, 0, sizeof D);
            // pass 2: data x basis (D[row][ky]); pass 3: basis x data (D[kz][kx])
            if (p == 1) mfma_cpu(hi, F[0], D); else mfma_cpu(F[0], hi, D);
            for (int l = 0; l < 64; ++l) for (int r = 0; r < 16; ++r) D[l][r] = (int32_t)((uint32_t)D[l][r] << 8);
            if (p == 1) { mfma_cpu(hi, F[2], D); mfma_cpu(lo, F[0], D); } else { mfma_cpu(F[2], hi, D); mfma_cpu(F[0], lo, D); }
            // pass 2 scaled by 4 (<< 8, 4 rnd, 4 Cl), pass 3 not (i32 output)
            const uint32_t rnd = p == 1 ? (uint32_t)vcg_im_rnd4(1, 0) : 8192u;
            const int sh = p == 1 ? 8 : 6;
            for (int l = 0; l < 64; ++l) for (int r = 0; r < 16; ++r) D[l][r] = (int32_t)(((uint32_t)D[l][r] << sh) + rnd);
            if (p == 1) mfma_cpu(lo, F[2], D); else mfma_cpu(F[1], lo, D);
            for (int l = 0; l < 64; ++l) {
                if (p == 1) {
                    for (int qq = 0; qq < 4; ++qq) {
                        uint32_t w[2];
                        vcg_im_pack16h(D[l] + 4 * qq, w);
                        memcpy(tile + vcg_im_out_addr(1, B, l & 31, l >> 5, qq), w, 8);
                    }
                } else {
                    for (int r = 0; r < 16; ++r) t[vcg_tidx(vcg_fwd1_k(r, l >> 5), B, l & 31)] = D[l][r] >> 14;
                }
            }
        }
    }
    if (coef_raster)
        for (int k = 0; k < 32768; ++k) coef_raster[k] = t[vcg_tidx(k >> 10, (k >> 5) & 31, k & 31)];
    int last = 0;
    float step[VCG_NFREQ], dz[VCG_NFREQ];
    vcg_quant_tables(gstep, q, step, dz, 0, 1);
    for (int tid = 0; tid < NT; ++tid) { int l = vcg_quant_scan(t, scan, step, dz, lv, tid, NT); if (l > last) last = l; }
    *eob = (uint32_t)last; *dc_out = (uint32_t)dc;
    free(t);
    return 0;
}

// k_idct on one atom's entries: force = 0 picks the path like the kernel,
// 1 forces the 24-bit path, 2 the full 32-bit path; returns the path used.
int kc_idct(const uint32_t *L, uint32_t cnt, int dc, const int32_t *st, const int32_t *bi,
            const int32_t cmat[32][32], const uint16_t *scan, int force, uint8_t *out) {
    int *t = (int *)malloc(VCG_TILE_INTS * sizeof(int));
    for (int tid = 0; tid < NTI; ++tid) vcg_tile_zero(t, tid, NTI);
    const int fmax = cnt ? vcg_freq(scan[vcg_ent_pos(L[cnt - 1])]) : -1;
    uint64_t l1 = 0;
    for (int tid = 0; tid < NTI; ++tid) l1 += vcg_scatter(t, L, cnt, scan, st, bi, tid, NTI);
    const bool fast = force == 1 ? true : force == 2 ? false : l1 <= 8000000ull;
    static uint8_t vox[NTI][64];
    if (fast) {
        for (int tid = 0; tid < NTI; ++tid) vcg_inv_pass_x<true>(t, cmat, fmax, tid, NTI);
        for (int tid = 0; tid < NTI; ++tid) vcg_inv_pass_y<true>(t, cmat, fmax, tid, NTI);
        for (int tid = 0; tid < NTI; ++tid) vcg_inv_pass_z<true>(t, cmat, fmax, dc, tid, vox[tid]);
    } else {
        for (int tid = 0; tid < NTI; ++tid) vcg_inv_pass_x<false>(t, cmat, fmax, tid, NTI);
        for (int tid = 0; tid < NTI; ++tid) vcg_inv_pass_y<false>(t, cmat, fmax, tid, NTI);
        for (int tid = 0; tid < NTI; ++tid) vcg_inv_pass_z<false>(t, cmat, fmax, dc, tid, vox[tid]);
    }
    // the staged raster is exactly the output atom
    for (int tid = 0; tid < NTI; ++tid) vcg_inv_stage(out, vox[tid], tid);
    free(t);
    return fast ? 1 : 2;
}

// v_mfma_i32_32x32x32_i8 over a wave's operands: D[l][r] = C[l][r] + sum
// over both lane halves h2 and bytes e of A[row + 32 h2].e x B[col + 32 h2].e
// (signed bytes), row = 8 (r / 4) + 4 (l / 32) + r % 4, col = l % 32; i32 wrap
static void mfma_cpu(const uint32_t A[64][4], const uint32_t B[64][4], int32_t D[64][16]) {
    for (int l = 0; l < 64; ++l)
        for (int r = 0; r < 16; ++r) {
            const int row = vcg_fwd1_k(r, l >> 5), col = l & 31;
            uint32_t acc = (uint32_t)D[l][r];
            for (int h2 = 0; h2 < 2; ++h2)
                for (int e = 0; e < 16; ++e) {
                    const int8_t a = (int8_t)((A[row + 32 * h2][e >> 2] >> (8 * (e & 3))) & 0xFFu);
                    const int8_t b = (int8_t)((B[col + 32 * h2][e >> 2] >> (8 * (e & 3))) & 0xFFu);
                    acc += (uint32_t)((int)a * (int)b);
                }
            D[l][r] = (int32_t)acc;
        }
}

// k_idct_m's product chain for one block (d[l] = lane l's 16 w), as im_chain:
// scaled by 4, the output in bits 16..31 of o
static void im_chain_cpu(const uint32_t d[64][8], const uint32_t B[3][64][4], int rnd4, int32_t o[64][16]) {
    static uint32_t hi[64][4], lo[64][4];
    for (int l = 0; l < 64; ++l) vcg_im_planes(d[l], hi[l], lo[l]);
    int32_t t[64][16];
    memset(t, 0, sizeof t);
    mfma_cpu(hi, B[0], t);
    for (int l = 0; l < 64; ++l) for (int r = 0; r < 16; ++r) t[l][r] = (int32_t)((uint32_t)t[l][r] << 8);
    mfma_cpu(hi, B[1], t);
    mfma_cpu(lo, B[0], t);
    for (int l = 0; l < 64; ++l) for (int r = 0; r < 16; ++r) t[l][r] = (int32_t)(((uint32_t)t[l][r] << 8) + (uint32_t)rnd4);
    mfma_cpu(lo, B[1], t);
    for (int l = 0; l < 64; ++l) for (int r = 0; r < 16; ++r) o[l][r] = t[l][r];
}

// k_idct_m on one atom's entries, block by block with the kernel's lane
// layout: pass x reads and writes each plane in turn, passes y and z read
// every block before writing any (the kernel's barriers).  Atoms past the
// coefficient bound go to the i32 path like the kernel's status-4 items.
// Returns 3 for the matrix-core path, else kc_idct's path.
int kc_idct_m(const uint32_t *L, uint32_t cnt, int dc, const int32_t *st, const int32_t *bi,
              const int32_t cmat[32][32], const uint16_t *scan, uint8_t *out) {
    static int8_t ib[3][64][16];
    vcg_im_tables(cmat, ib);
    static uint32_t B[3][64][4];
    memcpy(B, ib, sizeof B);
    uint8_t *t = (uint8_t *)aligned_alloc(16, VCG_IM_TILE_BYTES);
    for (int tid = 0; tid < NTI; ++tid) vcg_im_fill((uint32_t *)t, tid, NTI);
    const int fmax = cnt ? vcg_freq(scan[vcg_ent_pos(L[cnt - 1])]) : -1;
    // the kernel's tables: per scan position tile offset | frequency, and
    // per frequency the (step, bias) pair
    static uint32_t imtab[32768];
    for (int q = 0; q < 32768; ++q) imtab[q] = vcg_im_tab(scan[q]);
    int32_t sb[2 * VCG_NFREQ];
    for (int f = 0; f < VCG_NFREQ; ++f) { sb[2 * f] = st[f]; sb[2 * f + 1] = bi[f]; }
    uint64_t ss = 0;
    uint32_t big = 0;
    for (int tid = 0; tid < NTI; ++tid) vcg_im_scatter(t, L, cnt, imtab, sb, tid, NTI, ss, big);
    if (big || ss > (uint64_t)VCG_IM_L2 * VCG_IM_L2) {
        free(t);
        return kc_idct(L, cnt, dc, st, bi, cmat, scan, 0, out);
    }
    static uint32_t d[32][64][8];
    static int32_t o[64][16];
    for (int p = 0; p < 3; ++p) {
        for (int Bk = 0; Bk < 32; ++Bk) {
            if (vcg_im_skip(p, Bk, fmax)) continue;
            for (int l = 0; l < 64; ++l) vcg_im_load(t, p, Bk, l & 31, l >> 5, d[Bk][l]);
            if (p == 0) {                          // in place, plane by plane
                im_chain_cpu(d[Bk], B, vcg_im_rnd4(0, dc), o);
                for (int l = 0; l < 64; ++l)
                    for (int q = 0; q < 4; ++q) {
                        uint32_t w[2];
                        vcg_im_pack16h(o[l] + 4 * q, w);
                        memcpy(t + vcg_im_out_addr(0, Bk, l & 31, l >> 5, q), w, 8);
                    }
            }
        }
        if (p == 0) continue;
        for (int Bk = 0; Bk < 32; ++Bk) {
            if (vcg_im_skip(p, Bk, fmax)) continue;
            im_chain_cpu(d[Bk], B, vcg_im_rnd4(p, dc), o);
            for (int l = 0; l < 64; ++l)
                for (int q = 0; q < 4; ++q) {
                    if (p == 1) {
                        uint32_t w[2];
                        vcg_im_pack16h(o[l] + 4 * q, w);
                        memcpy(t + vcg_im_out_addr(1, Bk, l & 31, l >> 5, q), w, 8);
                    } else {
                        const uint32_t w = vcg_im_pack8h(o[l] + 4 * q);
                        memcpy(t + vcg_im_vox_addr(Bk, l & 31, l >> 5, q), &w, 4);
                    }
                }
        }
    }
    for (int z = 0; z < 32; ++z) memcpy(out + z * 1024, t + z * VCG_IM_ZS, 1024);
    free(t);
    return 3;
}
}
