// vcg_xform.hpp — the per-workgroup phases of the forward and inverse DCT-32
// kernels, written as __host__ __device__ functions of (tid, nthreads) over a
// padded 32x32x(32+1) i32 tile.  The HIP kernels call them between barriers;
// the CPU build (tests/kernel_cpu) runs the same phases thread by thread, so
// the exact kernel logic is checked against the oracle on the host.
//
// Tile layout: element (z, y, x) at (z*32 + y)*33 + x (x stride 1, y stride 33,
// z stride 1056); the +1 pad spreads the y/z-strided line accesses across LDS
// banks.
#pragma once
#include <stdint.h>
#include "vcg_codec.hpp"

#define VCG_PADW 33
#define VCG_TILE_INTS (32 * 32 * VCG_PADW)
#define VCG_SX 1
#define VCG_SY VCG_PADW
#define VCG_SZ (32 * VCG_PADW)

VCG_HD int vcg_tidx(int z, int y, int x) { return (z * 32 + y) * VCG_PADW + x; }
VCG_HD int vcg_freq(uint32_t idx) { return (int)((idx >> 10) + ((idx >> 5) & 31) + (idx & 31)); }

#if defined(__HIP_DEVICE_COMPILE__)
VCG_HD int vcg_mul24(int a, int b) { return __mul24(a, b); }
// even term count covering every lane of the wave, made provably uniform
VCG_HD int vcg_nk(int kmax) {
    int m = kmax;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
    m = m < 0 ? 0 : (m > 31 ? 32 : ((m + 2) & ~1));
    return __builtin_amdgcn_readfirstlane(m);
}
#else
// signed 24-bit multiply: low 32 bits of sext24(a) * sext24(b)
VCG_HD int vcg_mul24(int a, int b) {
    int64_t x = (int64_t)(int32_t)((uint32_t)a << 8) >> 8, y = (int64_t)(int32_t)((uint32_t)b << 8) >> 8;
    return (int)(uint32_t)(uint64_t)(x * y);
}
VCG_HD int vcg_nk(int kmax) { return kmax < 0 ? 0 : (kmax > 31 ? 32 : ((kmax + 2) & ~1)); }
#endif

template <bool FAST>
VCG_HD int vcg_mac(int a, int c) {
    if constexpr (FAST) return vcg_mul24(a, c);
    else return (int)((uint32_t)a * (uint32_t)c);
}

// acc + c*a (i32 wrap): FAST = 24-bit signed operands, one v_mad_i32_i24 with
// the basis constant c read from an SGPR (c is workgroup-uniform)
template <bool FAST>
VCG_HD int vcg_acc(int acc, int c, int a) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (FAST) {
        int r;
        asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "s"(c), "v"(a), "v"(acc));
        return r;
    }
#endif
    return (int)((uint32_t)acc + (uint32_t)vcg_mac<FAST>(c, a));
}

// ===========================================================================
// inverse
// ===========================================================================
// Two inverse lines read from the tile (element k of line i at s_i[k*stride])
// with only the first nk coefficients possibly nonzero (nk a multiple of 8):
//   evn[n] = sum_{k even} C[k][n] v[k],  odd[n] = sum_{k odd} C[k][n] v[k]
//   o[n] = (evn+odd+8192)>>14,  o[31-n] = (evn-odd+8192)>>14     (i32 wrap)
// src/vc/vc.c:178-184.  evn is regrouped exactly (integer sums, mod 2^32) by
// the even-row symmetries of the basis, C[2j][15-n] = (-1)^j C[2j][n],
// C[4m][7-n] = (-1)^m C[4m][n], C[8p][3-n] = (-1)^p C[8p][n] (checked on the
// table at device init): evn from 8x8 + 4x4 + 2x2 + 2x2 products instead of
// 16x16, 344 multiply-adds per line instead of 512.  The two lines share
// every basis load; basis rows are workgroup-uniform (scalar registers).
template <bool FAST>
VCG_HD void vcg_inv_line2(const int32_t (*C)[32], const int *s0, const int *s1, int stride, int nk, int *o0,
                          int *o1) {
    struct row16 { int32_t v[16]; };
    struct row8 { int32_t v[8]; };
    struct row4 { int32_t v[4]; };
    struct row2 { int32_t v[2]; };
    int d0[16], d1[16], q0[8], q1[8];
#pragma unroll
    for (int n = 0; n < 16; ++n) { d0[n] = 0; d1[n] = 0; }
#pragma unroll
    for (int n = 0; n < 8; ++n) { q0[n] = 0; q1[n] = 0; }
    // odd rows k = 1,3,...  and rows k = 2 mod 4, eight coefficients per step
#pragma unroll 1
    for (int kb = 0; kb < nk; kb += 8) {
#pragma unroll
        for (int j = 1; j < 8; j += 2) {
            const int a0 = s0[(kb + j) * stride], a1 = s1[(kb + j) * stride];
            const row16 r = *(const row16 *)C[kb + j];
#pragma unroll
            for (int n = 0; n < 16; ++n) {
                d0[n] = vcg_acc<FAST>(d0[n], r.v[n], a0);
                d1[n] = vcg_acc<FAST>(d1[n], r.v[n], a1);
            }
        }
#pragma unroll
        for (int j = 2; j < 8; j += 4) {
            const int a0 = s0[(kb + j) * stride], a1 = s1[(kb + j) * stride];
            const row8 r = *(const row8 *)C[kb + j];
#pragma unroll
            for (int n = 0; n < 8; ++n) {
                q0[n] = vcg_acc<FAST>(q0[n], r.v[n], a0);
                q1[n] = vcg_acc<FAST>(q1[n], r.v[n], a1);
            }
        }
    }
    // rows k = 4 mod 8 (n < 4), k = 8 mod 16 and k = 0 mod 16 (n < 2)
    int eeo0[4] = {0, 0, 0, 0}, eeo1[4] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 4; k < 32; k += 8) {
        const int a0 = s0[k * stride], a1 = s1[k * stride];
        const row4 r = *(const row4 *)C[k];
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            eeo0[n] = vcg_acc<FAST>(eeo0[n], r.v[n], a0);
            eeo1[n] = vcg_acc<FAST>(eeo1[n], r.v[n], a1);
        }
    }
    int x0[2] = {0, 0}, x1[2] = {0, 0}, y0[2] = {0, 0}, y1[2] = {0, 0};   // x: k = 8, 24; y: k = 0, 16
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
        const int a0 = s0[k * stride], a1 = s1[k * stride];
        const row2 r = *(const row2 *)C[k];
#pragma unroll
        for (int n = 0; n < 2; ++n) {
            if (k & 8) { x0[n] = vcg_acc<FAST>(x0[n], r.v[n], a0); x1[n] = vcg_acc<FAST>(x1[n], r.v[n], a1); }
            else { y0[n] = vcg_acc<FAST>(y0[n], r.v[n], a0); y1[n] = vcg_acc<FAST>(y1[n], r.v[n], a1); }
        }
    }
    // recombine: eee (n<4) -> ee (n<8) -> evn (n<16) -> outputs
    uint32_t e0[16], e1[16];
    {
        uint32_t a[4], b[4], c[8], d[8];
#pragma unroll
        for (int n = 0; n < 2; ++n) {
            a[n] = (uint32_t)y0[n] + (uint32_t)x0[n]; a[3 - n] = (uint32_t)y0[n] - (uint32_t)x0[n];
            b[n] = (uint32_t)y1[n] + (uint32_t)x1[n]; b[3 - n] = (uint32_t)y1[n] - (uint32_t)x1[n];
        }
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            c[n] = a[n] + (uint32_t)eeo0[n]; c[7 - n] = a[n] - (uint32_t)eeo0[n];
            d[n] = b[n] + (uint32_t)eeo1[n]; d[7 - n] = b[n] - (uint32_t)eeo1[n];
        }
#pragma unroll
        for (int n = 0; n < 8; ++n) {
            e0[n] = c[n] + (uint32_t)q0[n]; e0[15 - n] = c[n] - (uint32_t)q0[n];
            e1[n] = d[n] + (uint32_t)q1[n]; e1[15 - n] = d[n] - (uint32_t)q1[n];
        }
    }
#pragma unroll
    for (int n = 0; n < 16; ++n) {
        o0[n] = (int)(e0[n] + (uint32_t)d0[n] + 8192u) >> 14;
        o0[31 - n] = (int)(e0[n] - (uint32_t)d0[n] + 8192u) >> 14;
        o1[n] = (int)(e1[n] + (uint32_t)d1[n] + 8192u) >> 14;
        o1[31 - n] = (int)(e1[n] - (uint32_t)d1[n] + 8192u) >> 14;
    }
}

// term count covering coefficients 0..kmax, rounded up to a multiple of 8
VCG_HD int vcg_nk8(int kmax) { return kmax < 0 ? 0 : (kmax > 23 ? 32 : ((kmax + 8) & ~7)); }


VCG_HD void vcg_tile_zero(int *t, int tid, int nthr) {
    struct i4 { int a, b, c, d; };
    i4 *q = (i4 *)t;
    for (int k = tid; k < VCG_TILE_INTS / 4; k += nthr) q[k] = i4{0, 0, 0, 0};
}

// Dequantize list entry v, coefficient at raster index idx, into the tile
// (src/vc/vc.c:327-338); returns |coef| (for the 24-bit bound).
VCG_HD uint64_t vcg_scatter1(int *t, uint32_t v, uint32_t idx, const int32_t *st, const int32_t *bi) {
    const int lv = vcg_ent_level(v);
    const int f = vcg_freq(idx);
    const uint32_t mag = (uint32_t)(lv < 0 ? -lv : lv);
    int r = (int)(mag * (uint32_t)st[f] + (uint32_t)bi[f] + 128u) >> 8;
    r = lv < 0 ? (int)(0u - (uint32_t)r) : r;
    t[vcg_tidx((int)(idx >> 10), (int)((idx >> 5) & 31), (int)(idx & 31))] = r;
    return (uint64_t)(r < 0 ? -(int64_t)r : (int64_t)r);
}

// Dequantize entries [tid, cnt) step nthr into the tile; returns this
// thread's sum of |coef|.  Entries go in groups of 8 per thread so their list
// and scan-table loads overlap.
VCG_HD uint64_t vcg_scatter(int *t, const uint32_t *L, uint32_t cnt, const uint16_t *scan, const int32_t *st,
                            const int32_t *bi, int tid, int nthr) {
    uint64_t mine = 0;
    const uint32_t un = (uint32_t)nthr;
    for (uint32_t e0 = (uint32_t)tid; e0 < cnt; e0 += 8u * un) {
        uint32_t v[8], idx[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = e0 + q * un < cnt ? L[e0 + q * un] : 0u;
#pragma unroll
        for (int q = 0; q < 8; ++q) idx[q] = e0 + q * un < cnt ? scan[vcg_ent_pos(v[q])] : 0u;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if (e0 + q * un >= cnt) break;
            mine += vcg_scatter1(t, v[q], idx[q], st, bi);
        }
    }
    return mine;
}

// The three passes give each thread the line pairs (line, line + nthr) for
// line = tid, tid + 2*nthr, ... (nthr divides 512).  The pair's term count is
// the larger line's, made wave-uniform on the GPU (vcg_nk) and rounded to 8;
// the extra terms multiply zeros.

// pass 1 (along x): lines (kz, ky); coefficients beyond kx = fmax-kz-ky are zero
template <bool FAST>
VCG_HD void vcg_inv_pass_x(int *t, const int32_t (*C)[32], int fmax, int tid, int nthr) {
    for (int l0 = tid; l0 < 1024; l0 += 2 * nthr) {
        const int l1 = l0 + nthr;
        const int km0 = fmax - (l0 >> 5) - (l0 & 31), km1 = fmax - (l1 >> 5) - (l1 & 31);
        const int nk = vcg_nk8(vcg_nk(km0 > km1 ? km0 : km1) - 1);
        if (nk == 0) continue;
        int *r0 = t + vcg_tidx(l0 >> 5, l0 & 31, 0), *r1 = t + vcg_tidx(l1 >> 5, l1 & 31, 0);
        int o0[32], o1[32];
        vcg_inv_line2<FAST>(C, r0, r1, VCG_SX, nk, o0, o1);
        if (km0 >= 0) {
#pragma unroll
            for (int k = 0; k < 32; ++k) r0[k * VCG_SX] = o0[k];
        }
        if (km1 >= 0) {
#pragma unroll
            for (int k = 0; k < 32; ++k) r1[k * VCG_SX] = o1[k];
        }
    }
}
// pass 2 (along y): lines (kz, x); ky beyond fmax-kz are zero
template <bool FAST>
VCG_HD void vcg_inv_pass_y(int *t, const int32_t (*C)[32], int fmax, int tid, int nthr) {
    for (int l0 = tid; l0 < 1024; l0 += 2 * nthr) {
        const int l1 = l0 + nthr;
        const int km0 = fmax - (l0 >> 5), km1 = fmax - (l1 >> 5);
        const int nk = vcg_nk8(vcg_nk(km0 > km1 ? km0 : km1) - 1);
        if (nk == 0) continue;
        int *c0 = t + vcg_tidx(l0 >> 5, 0, l0 & 31), *c1 = t + vcg_tidx(l1 >> 5, 0, l1 & 31);
        int o0[32], o1[32];
        vcg_inv_line2<FAST>(C, c0, c1, VCG_SY, nk, o0, o1);
        if (km0 >= 0) {
#pragma unroll
            for (int k = 0; k < 32; ++k) c0[k * VCG_SY] = o0[k];
        }
        if (km1 >= 0) {
#pragma unroll
            for (int k = 0; k < 32; ++k) c1[k * VCG_SY] = o1[k];
        }
    }
}
// pass 3 (along z): lines (y, x) -> 32 output voxels each (dc added, clamped).
// For nthr = 512 each thread has one pair; vox[0..31] = line tid (z = 0..31),
// vox[32..63] = line tid + 512.
template <bool FAST>
VCG_HD void vcg_inv_pass_z(const int *t, const int32_t (*C)[32], int fmax, int dc, int tid, uint8_t *vox) {
    const int nk = vcg_nk8(fmax);
    const int l0 = tid, l1 = tid + 512;
    int o0[32], o1[32];
    vcg_inv_line2<FAST>(C, t + vcg_tidx(0, l0 >> 5, l0 & 31), t + vcg_tidx(0, l1 >> 5, l1 & 31), VCG_SZ, nk, o0,
                        o1);
#pragma unroll
    for (int z = 0; z < 32; ++z) {
        const int v0 = o0[z] + dc, v1 = o1[z] + dc;
        vox[z] = (uint8_t)(v0 < 0 ? 0 : v0 > 255 ? 255 : v0);
        vox[32 + z] = (uint8_t)(v1 < 0 ? 0 : v1 > 255 ? 255 : v1);
    }
}
// after a barrier: the pass-3 voxels of thread tid into a raster u8 atom
// (z, y, x) at stage (the tile's memory reused)
VCG_HD void vcg_inv_stage(uint8_t *stage, const uint8_t *vox, int tid) {
#pragma unroll
    for (int z = 0; z < 32; ++z) {
        stage[z * 1024 + tid] = vox[z];
        stage[z * 1024 + 512 + tid] = vox[32 + z];
    }
}

// ===========================================================================
// forward
// ===========================================================================
// ---------------------------------------------------------------------------
// forward transform on the matrix cores (v_mfma_i32_32x32x32_i8)
// ---------------------------------------------------------------------------
// Each forward pass is, per line, out[k] = (sum_n C[k][n] v[n] + 8192) >> 14
// (src/vc/vc.c:142-171, whose even/odd regrouping only reorders the integer
// sum), computed as exact int8 products over the inverse's i16 tile layout
// (vcg_im_chunk, "inverse on the matrix cores"):
//   - pass 1 (x) takes the voxels as bytes: v - dc = (v - 128) + (128 - dc),
//     v - 128 is int8 (the byte with its top bit flipped), and
//     sum_n C[k][n] (v[n] - dc) = sum_n C[k][n] (v[n] - 128) + (128 - dc)
//     rowsum[k]; with C = 64 Ch + Cl two products: T = (v-128)*Ch; T = (T <<
//     6) + corr; T += (v-128)*Cl; out = T >> 14;
//   - passes 2 (y) and 3 (z) take w = v + 128 from the tile in two byte
//     planes and run the inverse's four-product chain with the basis rows
//     C[k][.] in place of the columns.
// Every pass input fits the tile: a pass grows magnitudes by at most
// max_k sum_n |C[k][n]| / 2^14 = 92672 / 16384 < 5.66, so |v| <= 1443 into
// pass 2 and <= 8164 into pass 3 (|voxel - dc| <= 255); pass 3 writes i32
// coefficients.
//
// Per-lane basis operand: lane c + 32 h, byte e = plane of C[c][16h + e]
// (planes Ch, Cl, 4 Cl); as an A operand it supplies row c, as a B operand
// column c.  Products (a wave takes blocks wv, wv + 8, wv + 16, wv + 24):
//   x: block = plane z, A = voxel row y (16 voxels x = 16h.. per lane, one
//      16-byte load), B = basis: D[y][kx] -> tile [z][kx][y];
//   y: blocks and rows as the inverse's pass y (rows (z = 4m + j % 4, kx =
//      8p + j / 4)), A = data, B = basis: D[row][ky] -> tile [ky][kx][z];
//   z: block = plane ky, A = basis, B = data line (ky, kx = col):
//      D[kz][kx] -> i32 tile (vcg_tidx) at (kz, ky, kx).

// output register r of lane half h: row 8 (r / 4) + 4 h + r % 4
// per-lane basis operands (planes Ch, Cl, 4 Cl) and row sums, built once
VCG_HD void vcg_fwd1_tables(const int32_t (*C)[32], int8_t fa[3][64][16], int32_t rowsum[32]) {
    for (int k = 0; k < 32; ++k) {
        int32_t s = 0;
        for (int n = 0; n < 32; ++n) s += C[k][n];
        rowsum[k] = s;
    }
    for (int l = 0; l < 64; ++l)
        for (int e = 0; e < 16; ++e) {
            const int c = C[l & 31][16 * (l >> 5) + e];
            const int lo = ((c + 32) & 63) - 32;
            fa[0][l][e] = (int8_t)((c - lo) / 64);
            fa[1][l][e] = (int8_t)lo;
            fa[2][l][e] = (int8_t)(4 * lo);
        }
}

// the 16 voxels (x = 16h .. 16h+15 of atom row (z, y)) lane (y, h) feeds
// for one z-plane, zero outside the volume, little-endian in w[0..3]
VCG_HD void vcg_fwd1_row(const uint8_t *vol, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t ax, uint32_t ay,
                         uint32_t az, int z, int y, int h, uint32_t w[4]) {
    w[0] = w[1] = w[2] = w[3] = 0;
    const uint32_t vz = az * 32 + (uint32_t)z, vy = ay * 32 + (uint32_t)y, vx = ax * 32 + 16 * (uint32_t)h;
    if (vz >= nz || vy >= ny || vx >= nx) return;
    const uint8_t *p = vol + ((uint64_t)vz * ny + vy) * nx + vx;
    if (vx + 16 <= nx && ((uintptr_t)p & 3) == 0) {
        const uint32_t *q = (const uint32_t *)p;
        w[0] = q[0]; w[1] = q[1]; w[2] = q[2]; w[3] = q[3];
        return;
    }
    const uint32_t m = nx - vx < 16 ? nx - vx : 16;
    for (uint32_t e = 0; e < m; ++e) w[e >> 2] |= (uint32_t)p[e] << (8 * (e & 3));
}

// output register r of lane half h: frequency row k
VCG_HD int vcg_fwd1_k(int r, int h) { return 8 * (r >> 2) + 4 * h + (r & 3); }

// pass-1 rounding term of output column k (the tile stores w = out + 128)
VCG_HD int vcg_fwd1_corr(int dc, int32_t rowsum_k) { return (128 - dc) * rowsum_k + 8192 + (128 << 14); }

// ===========================================================================
// inverse on the matrix cores (i16 tile)
// ===========================================================================
// Each inverse pass is, per line, out[n] = (sum_k C[k][n] v[k] + 8192) >> 14
// (src/vc/vc.c:172-205; the reference's even/odd split only regroups the
// integer sum).  Here a pass is a set of 32x32x32 int8 products
// D[line][n] = sum_k A[line][k] B[k][n], A the data, B the basis, exact:
//   - v = 256 vh + vl with vl in [-128, 127]: the tile holds w = v + 128 as
//     i16, and then vh = byte 1 of w, vl = byte 0 of w xor 0x80;
//   - C = 64 Ch + Cl with Cl = ((C + 32) & 63) - 32 (|C| <= 4091: Ch, Cl
//     and 4 Cl all fit int8);
//   - C v = 64 (256 Ch vh + 4Cl vh + Ch vl) + Cl vl, i.e. the chain
//     T = Ch*vh; T <<= 8; T += 4Cl*vh; T += Ch*vl; T = (T << 6) + rnd;
//     T += Cl*vl; out = T >> 14, every step exact mod 2^32.
// With the atom's coefficient L2 norm at most VCG_IM_L2 (and every
// coefficient at most that), every pass input fits: each 1-D pass has
// sigma_max(C / 2^14) = 1.00008 and adds at most 0.5 of rounding per element
// (L2 <= 0.5 sqrt(32768) = 90.6 per pass), so pass inputs stay below
// 1.00008^2 31000 + 2 x 90.6 = 31186 in magnitude, inside the
// [-32896, 32639] that w = v + 128 covers, and |sum C v| <= 2^14 x 1.00008 x
// 31186 < 2^29, far from i32 overflow.  Atoms past the bound take the i32
// path (vcg_inv_pass_*).
//
// Tile: 1024 lines of 32 i16 (64 bytes, four 16-byte chunks); chunk c of
// line L sits at chunk position c ^ ((L >> 1) & 3) ^ ((L >> 5) & 3), which
// keeps every group of eight lanes of a 16-byte line read on distinct banks
// in all three passes.
//
// v_mfma_i32_32x32x32_i8 lane layout: lane l = j + 32 h supplies A row j
// (k = 16h .. 16h+15, one 32-byte read) and B column j (the same k); output
// register r of lane l is D[8 (r / 4) + 4 h + r % 4][l mod 32].
//
// Passes and their 32 blocks (a block is one product; a wave takes blocks
// wv, wv + 8, wv + 16, wv + 24):
//   x: block = plane kz, rows = ky; in [kz][ky][kx], out [kz][x][ky];
//   y: block = (kz 4m..4m+3, x 8p..8p+7), B = 4m + p, row j = (kz = 4m +
//      j % 4, x = 8p + j / 4); in [kz][x][ky], out [y][x][kz];
//   z: block = plane y, rows = x; in [y][x][kz], out voxels (z, y, x) to the
//      staging raster (z stride VCG_IM_ZS bytes).
// Pass x skips planes kz > fmax and pass y kz groups past fmax: their inputs
// are zero (w = 128), and so are the outputs at the positions they would
// write (lines with kz + ky > fmax).
#define VCG_IM_L2 31000u
#define VCG_IM_TILE_BYTES 65536
#define VCG_IM_ZS 1040                  // staging z stride: 16-byte aligned, banks spread
#define VCG_IM_W0 0x00800080u           // two i16 w = 128 (v = 0)

VCG_HD uint32_t vcg_perm(uint32_t hi, uint32_t lo, uint32_t sel) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_perm(hi, lo, sel);
#else
    // selector bytes 0..7 pick a byte of {hi, lo}, 0x0c gives 0x00, 0x0d 0xff
    const uint64_t v = (uint64_t)hi << 32 | lo;
    uint32_t r = 0;
    for (int b = 0; b < 4; ++b) {
        const uint32_t s = (sel >> (8 * b)) & 0xFFu;
        const uint32_t byte = s < 8 ? (uint32_t)((v >> (8 * s)) & 0xFFu) : s == 0x0c ? 0u : 0xFFu;
        r |= byte << (8 * b);
    }
    return r;
#endif
}

// byte offset of chunk c (8 i16) of tile line L
VCG_HD uint32_t vcg_im_chunk(uint32_t L, uint32_t c) {
    return (L << 6) + ((c ^ ((L >> 1) & 3u) ^ ((L >> 5) & 3u)) << 4);
}

// per-lane B operands: lane l = n + 32 h, byte e = plane of C[16h + e][n];
// plane 0 = Ch, 1 = 4 Cl, 2 = Cl
VCG_HD void vcg_im_tables(const int32_t (*C)[32], int8_t ib[3][64][16]) {
    for (int l = 0; l < 64; ++l)
        for (int e = 0; e < 16; ++e) {
            const int c = C[16 * (l >> 5) + e][l & 31];
            const int lo = ((c + 32) & 63) - 32;
            ib[0][l][e] = (int8_t)((c - lo) / 64);
            ib[1][l][e] = (int8_t)(4 * lo);
            ib[2][l][e] = (int8_t)lo;
        }
}

VCG_HD void vcg_im_fill(uint32_t *t, int tid, int nthr) {
    struct u4 { uint32_t a, b, c, d; };
    u4 *q = (u4 *)t;
    for (int k = tid; k < VCG_IM_TILE_BYTES / 16; k += nthr) q[k] = u4{VCG_IM_W0, VCG_IM_W0, VCG_IM_W0, VCG_IM_W0};
}

// per scan position: the i16-tile byte offset of its raster index idx and,
// in bits 16.., its frequency cz + cy + cx
VCG_HD uint32_t vcg_im_tab(uint32_t idx) {
    const uint32_t L = idx >> 5, x = idx & 31u;
    return (vcg_im_chunk(L, x >> 3) + 2u * (x & 7u)) | (uint32_t)vcg_freq(idx) << 16;
}

// Dequantize list entry v (tab = its position's vcg_im_tab) as vcg_scatter1
// does (sb: per frequency the Q8 step and bias as a pair) and store w = r +
// 128 into the i16 tile; ss += r^2 and big |= r above the bound (then ss is
// not needed).  The magnitude's Q8 product is taken unsigned: one that
// wrapped past 2^31 (pathological q) counts as above the bound, so such an
// atom takes the i32 path, which is exact mod 2^32 like the reference.
VCG_HD void vcg_im_scatter1(uint8_t *t, uint32_t v, uint32_t tab, const int32_t *sb, uint64_t &ss, uint32_t &big) {
    const int lv = vcg_ent_level(v);
    const uint32_t f = tab >> 16;
    const uint32_t mag = (uint32_t)(lv < 0 ? -lv : lv);
    const uint32_t r = (uint32_t)((int)(mag * (uint32_t)sb[2 * f] + (uint32_t)sb[2 * f + 1] + 128u) >> 8);
    big |= r > VCG_IM_L2 ? 1u : 0u;
    ss += (uint64_t)r * r;
    *(int16_t *)(t + (tab & 0xFFFFu)) = (int16_t)(uint16_t)(lv < 0 ? 128u - r : 128u + r);
}

VCG_HD void vcg_im_scatter(uint8_t *t, const uint32_t *L, uint32_t cnt, const uint32_t *imtab, const int32_t *sb,
                           int tid, int nthr, uint64_t &ss, uint32_t &big) {
    const uint32_t un = (uint32_t)nthr;
    for (uint32_t e0 = (uint32_t)tid; e0 < cnt; e0 += 8u * un) {
        uint32_t v[8], tab[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = e0 + q * un < cnt ? L[e0 + q * un] : 0u;
#pragma unroll
        for (int q = 0; q < 8; ++q) tab[q] = e0 + q * un < cnt ? imtab[vcg_ent_pos(v[q])] : 0u;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if (e0 + q * un >= cnt) break;
            vcg_im_scatter1(t, v[q], tab[q], sb, ss, big);
        }
    }
}

// block B of pass p: skipped (all-zero input)?
VCG_HD bool vcg_im_skip(int p, int B, int fmax) {
    return p == 0 ? B > fmax : p == 1 ? 4 * (B >> 2) > fmax : false;
}
// the tile line of A row j of block B in pass p
VCG_HD uint32_t vcg_im_in_line(int p, int B, int j) {
    if (p == 1) return (uint32_t)((4 * (B >> 2) + (j & 3)) * 32 + 8 * (B & 3) + (j >> 2));
    return (uint32_t)(B * 32 + j);
}
// lane (row j, half h) of block B in pass p: its 16 w of A, two 16-byte reads
VCG_HD void vcg_im_load(const uint8_t *t, int p, int B, int j, int h, uint32_t d[8]) {
    const uint32_t L = vcg_im_in_line(p, B, j);
    const uint32_t *a = (const uint32_t *)(t + vcg_im_chunk(L, 2u * (uint32_t)h));
    const uint32_t *b = (const uint32_t *)(t + vcg_im_chunk(L, 2u * (uint32_t)h + 1u));
#pragma unroll
    for (int q = 0; q < 4; ++q) { d[q] = a[q]; d[4 + q] = b[q]; }
}
// the two int8 planes of 16 w (element e of dword e / 2): hi = vh, lo = vl
VCG_HD void vcg_im_planes(const uint32_t d[8], uint32_t hi[4], uint32_t lo[4]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        hi[q] = vcg_perm(d[2 * q + 1], d[2 * q], 0x07050301u);
        lo[q] = vcg_perm(d[2 * q + 1], d[2 * q], 0x06040200u) ^ 0x80808080u;
    }
}
// rounding constant of the chain (before >> 14): passes x, y store w = out
// + 128; pass z adds the atom's dc
VCG_HD int vcg_im_rnd(int p, int dc) { return 8192 + ((p == 2 ? dc : 128) << 14); }

// The chains end scaled by 4: the last product takes 4 Cl in place of Cl and
// the shift and rounding constant before it are 4 times theirs (<< 8,
// vcg_im_rnd4), so a chain leaves T = 4 (sum + rnd) and the output
// (sum + rnd) >> 14 is bits 16..31 of T, taken whole by byte permutes
// instead of a shift per register.  Exact while |sum + rnd| < 2^29 (the
// coefficient bound keeps it below 5.16e8).
VCG_HD int vcg_im_rnd4(int p, int dc) { return (int)(4u * (uint32_t)vcg_im_rnd(p, dc)); }

// pass x / y: lane (n = l, half h) writes its output registers 4q..4q+3
// (four consecutive rows) as 8 bytes; returns the byte offset
VCG_HD uint32_t vcg_im_out_addr(int p, int B, int l, int h, int q) {
    if (p == 0) return vcg_im_chunk((uint32_t)(B * 32 + l), (uint32_t)q) + 8u * (uint32_t)h;
    const int m = B >> 2;
    const uint32_t L = (uint32_t)(l * 32 + 8 * (B & 3) + 2 * q + h);
    return vcg_im_chunk(L, (uint32_t)(m >> 1)) + 8u * (uint32_t)(m & 1);
}
// four outputs of a scaled chain (bits 16..31 of each register) packed
// little-endian as i16
VCG_HD void vcg_im_pack16h(const int t[4], uint32_t w[2]) {
    w[0] = vcg_perm((uint32_t)t[1], (uint32_t)t[0], 0x07060302u);
    w[1] = vcg_perm((uint32_t)t[3], (uint32_t)t[2], 0x07060302u);
}
// pass z: lane (z = l, half h), registers 4q..4q+3 = voxels x = 8q + 4h ..
// +3 of row (z, y = B): clamped, packed, at this staging offset
VCG_HD uint32_t vcg_im_vox_addr(int B, int l, int h, int q) {
    return (uint32_t)(l * VCG_IM_ZS + B * 32 + 8 * q + 4 * h);
}

// four voxels of a scaled chain: clamp(t >> 16, 0, 255) is byte 2 of t
// clamped to [0, 2^24 - 1] (one v_med3 each), gathered by permutes.  (Bytes
// are packed with permutes, not shifts and ors: hipcc fuses "shift, clamp to
// u8, pack two" into v_ashr_pk_u8_i32, which writes only the low 16 bits of
// its destination and leaves the register's stale high half in the word.)
VCG_HD uint32_t vcg_im_pack8h(const int t[4]) {
    uint32_t v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = (uint32_t)(t[i] < 0 ? 0 : t[i] > 0xFFFFFF ? 0xFFFFFF : t[i]);
    const uint32_t p01 = vcg_perm(v[1], v[0], 0x0c0c0602u), p23 = vcg_perm(v[3], v[2], 0x0c0c0602u);
    return vcg_perm(p23, p01, 0x05040100u);
}

#ifndef VCG_NFREQ
#define VCG_NFREQ 94                            // distinct frequencies cz + cy + cx
#endif
// quantize the tile's coefficients in scan order into lv[p]; returns this
// thread's max (p+1) over nonzero levels (the EOB candidate)
VCG_HD int vcg_quant_scan(const int *t, const uint16_t *scan, const float *step, const float *dz, int16_t *lv,
                          int tid, int nthr) {
    int last = 0;
    const int un = nthr;
    for (int p0 = tid; p0 < 32768; p0 += 8 * un) {
        uint32_t idx[8];                       // eight scan-table loads in flight
#pragma unroll
        for (int k = 0; k < 8; ++k) idx[k] = p0 + k * un < 32768 ? scan[p0 + k * un] : 0u;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int p = p0 + k * un;
            if (p >= 32768) break;
            const uint32_t id = idx[k];
            const int f = vcg_freq(id);
            const int l = vcg_quant1s(t[vcg_tidx((int)(id >> 10), (int)((id >> 5) & 31), (int)(id & 31))], step[f], dz[f]);
            lv[p] = (int16_t)l;
            if (l) last = p + 1;
        }
    }
    return last;
}
// the step and dead zone of each frequency at base_q (vcg_quant1's own float
// operations, so every level is unchanged)
VCG_HD void vcg_quant_tables(const float *gstep, float q, float *step, float *dz, int tid, int nthr) {
#pragma clang fp contract(off)
    for (int f = tid; f < VCG_NFREQ; f += nthr) {
        step[f] = q * gstep[f];
        dz[f] = 0.80f * step[f];
    }
}
