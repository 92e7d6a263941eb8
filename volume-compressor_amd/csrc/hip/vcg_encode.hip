// vcg_encode.hip — batched atom encode on gfx950.
//
//   k_downsample : LOD l -> l+1, integer rounded mean of the in-bounds voxels
//                  of each 2x2x2 cell (src/vc/vc.c:1562-1577).
//   k_fdct       : one 512-thread workgroup per atom, atoms of one or many
//                  volumes per launch.  Gathers the 32^3 atom (zero padded),
//                  sums it (zero atom -> no payload, dc = (sum+16384)/32768,
//                  src/vc/vc.c:599-604), runs the three forward passes as
//                  exact int8 MFMA products (pass 1 straight from the voxel
//                  bytes, passes 2 and 3 over an i16 tile in LDS;
//                  src/vc/vc.c:142-171, 212-218) and either stores raster
//                  coefficients (rate-control sample) or quantizes them
//                  (src/vc/vc.c:293-306, strict IEEE in source order) and
//                  stores the levels in scan order plus the EOB.
//   k_quant_groups / k_rc_count_groups : rate-control candidates (a volume's
//                  sample at one base_q), many per launch: re-quantize the
//                  stored coefficients, then count payload bytes per atom,
//                  abandoning a candidate once its total passes its budget.
//   k_rc_encode  : one lane per atom.  Adaptive binary range coder with LZMA
//                  carry and the coefficient context coder
//                  (src/vc/vc.c:360-401, 471-544), payload = [dc][bytes].
//   k_compact / k_assemble : pack payloads, then write the archive image.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <mutex>
#include <sys/mman.h>
#include <thread>
#include <vector>
#include "vcg_dev.hpp"
#include "vcg_xform.hpp"

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

#define FDCT_THREADS 256
#define FDCT_TX 512          // k_fdct: two waves per SIMD, one line pair per thread and pass

struct enc_item {
    uint32_t eob;
    uint32_t dc;
    uint32_t zero;
    uint32_t lod;
};

// One volume's LOD pyramid on device.  Its atoms are numbered LOD-major,
// raster (x fastest) within a LOD; the rate-control sample is an explicit
// list of LOD0 atoms whose raster coefficients and items go to coef/sitems.
struct enc_src {
    const uint8_t *lod[VCG_NLOD];
    uint32_t nx[VCG_NLOD], ny[VCG_NLOD], nz[VCG_NLOD];
    uint32_t acx[VCG_NLOD], acy[VCG_NLOD];
    uint64_t prefix[VCG_NLOD + 1];    // atom index base per LOD
    float q[VCG_NLOD];
    int nlod;
    const uint32_t *list;             // sample atoms (z,y,x)
    int32_t *coef;                    // sample coefficients, 32768 per atom
    enc_item *sitems;                 // sample items
};

__device__ __forceinline__ void enc_atom_of(const enc_src &s, const uint32_t *list, uint64_t gi, int &lod,
                                            uint32_t &ax, uint32_t &ay, uint32_t &az) {
    if (list) {
        lod = 0;
        az = list[3 * gi]; ay = list[3 * gi + 1]; ax = list[3 * gi + 2];
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

// A launch's atoms are a concatenation of segments: flat atoms [base, next
// segment's base) are atoms gi0, gi0 + 1, ... of volume v.
struct enc_seg {
    uint64_t base;
    uint64_t gi0;
    uint32_t v;
    uint32_t pad;
};

// the segment holding flat atom x (segments sorted by base, seg[0].base <= x)
__device__ __forceinline__ uint32_t seg_of(const enc_seg *seg, uint32_t n, uint64_t x) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t m = (lo + hi) / 2;
        if (seg[m].base <= x) lo = m; else hi = m;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// Output voxel o of an ox x oy x oz raster as (x, y, z), in the index type IT.
// The 2x kernels below run grid-stride loops over at most 2^24 threads and
// take IT = uint32_t when the output count plus that stays below 2^32 (64-bit
// division costs several times the 32-bit one).
template <typename IT>
__device__ __forceinline__ void ds_split(IT o, uint32_t ox, uint32_t oy, uint32_t &x, uint32_t &y, uint32_t &z) {
    const IT row = o / ox;
    x = (uint32_t)(o - row * ox);
    z = (uint32_t)(row / oy);
    y = (uint32_t)(row - (IT)z * oy);
}
__host__ __device__ __forceinline__ bool ds_small(uint64_t total) { return total + (1ull << 24) < (1ull << 32); }
// test hook: nonzero makes the launches below take the 64-bit index path
// whatever the output count, so small volumes exercise it
static int g_ds_wide = 0;
extern "C" void vcg_test_wide_index(int on) { g_ds_wide = on != 0; }

// output voxel (x, y, z) of the 2x pyramid level: the rounded mean of its
// in-bounds 2x2x2 cell
__device__ __forceinline__ uint8_t ds_voxel(const uint8_t *__restrict__ in, uint32_t nx, uint32_t ny, uint32_t nz,
                                            uint32_t x, uint32_t y, uint32_t z) {
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
    return (uint8_t)((acc + n / 2) / n);
}
template <typename IT>
__device__ __forceinline__ void downsample_body(const uint8_t *__restrict__ in, uint32_t nx, uint32_t ny,
                                                uint32_t nz, uint8_t *__restrict__ out) {
    const uint32_t ox = (nx + 1) / 2, oy = (ny + 1) / 2, oz = (nz + 1) / 2;
    const IT total = (IT)ox * oy * oz;
    for (IT o = (IT)blockIdx.x * blockDim.x + threadIdx.x; o < total; o += (IT)gridDim.x * blockDim.x) {
        uint32_t x, y, z;
        ds_split<IT>(o, ox, oy, x, y, z);
        out[o] = ds_voxel(in, nx, ny, nz, x, y, z);
    }
}
__global__ void k_downsample(const uint8_t *__restrict__ in, uint32_t nx, uint32_t ny, uint32_t nz,
                             uint8_t *__restrict__ out, int wide) {
    if (!wide && ds_small((uint64_t)((nx + 1) / 2) * ((ny + 1) / 2) * ((nz + 1) / 2)))
        downsample_body<uint32_t>(in, nx, ny, nz, out);
    else
        downsample_body<uint64_t>(in, nx, ny, nz, out);
}

// One pyramid level of many volumes in one launch: job j (input dims nx, ny,
// nz) owns blocks [blk0, next job's blk0), one thread per output voxel;
// output counts below 2^31.
struct ds_job {
    const uint8_t *in;
    uint8_t *out;
    uint32_t nx, ny, nz, blk0;
};
#define DS_JOB_THREADS 256u
__global__ __launch_bounds__(DS_JOB_THREADS) void k_downsample_jobs(const ds_job *__restrict__ J, uint32_t nj) {
    uint32_t lo = 0, hi = nj;
    while (hi - lo > 1) {
        const uint32_t m = (lo + hi) / 2;
        if (J[m].blk0 <= blockIdx.x) lo = m; else hi = m;
    }
    const ds_job j = J[lo];
    const uint32_t ox = (j.nx + 1) / 2, oy = (j.ny + 1) / 2, oz = (j.nz + 1) / 2;
    const uint32_t o = (blockIdx.x - j.blk0) * DS_JOB_THREADS + threadIdx.x;
    if (o >= ox * oy * oz) return;
    uint32_t x, y, z;
    ds_split<uint32_t>(o, ox, oy, x, y, z);
    j.out[o] = ds_voxel(j.in, j.nx, j.ny, j.nz, x, y, z);
}

// The exporter's within-cell 2x downscale (tools/downscale.h:32-93): output
// voxel o from its own 2x2x2 cell, the odd boundary edge-replicated.  method 0
// box (float mean + 0.5, truncated), 1 cbox (the mean pushed by alpha toward
// the first voxel of largest deviation); an all-zero cell gives 0.  Float in
// source order (the file is built without contraction).
__device__ __forceinline__ uint8_t ds_u8(float v) { return (uint8_t)(v < 0.0f ? 0.0f : v > 255.0f ? 255.0f : v + 0.5f); }
__device__ __forceinline__ uint8_t ds_cell(const uint32_t c[8], int method, float alpha) {
    if (!(c[0] | c[1] | c[2] | c[3] | c[4] | c[5] | c[6] | c[7])) return 0;
    const float mean = (float)(c[0] + c[1] + c[2] + c[3] + c[4] + c[5] + c[6] + c[7]) * 0.125f;
    if (method == 0) return ds_u8(mean);
    float best = mean, bestdev = -1.0f;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const float d = (float)c[i] - mean, ad = d < 0.0f ? -d : d;
        if (ad > bestdev) { bestdev = ad; best = (float)c[i]; }
    }
    return ds_u8(mean + alpha * (best - mean));
}
// even dims, nx a multiple of 8 and 8-byte aligned rows: four output voxels
// per thread from one 8-byte load of each of the cell's four input rows
template <typename IT>
__device__ __forceinline__ void ds2x_v_body(const uint8_t *__restrict__ in, uint32_t nx, uint32_t ny, uint32_t nz,
                                            int method, float alpha, uint8_t *__restrict__ out) {
    const uint32_t ox = nx / 2, oy = ny / 2, oz = nz / 2, ox4 = ox / 4;
    const IT total = (IT)ox4 * oy * oz;
    for (IT o = (IT)blockIdx.x * blockDim.x + threadIdx.x; o < total; o += (IT)gridDim.x * blockDim.x) {
        uint32_t x4, y, z;
        ds_split<IT>(o, ox4, oy, x4, y, z);
        const uint64_t b00 = ((uint64_t)(2 * z) * ny + 2 * y) * nx + 8u * x4;
        const uint2 r00 = *(const uint2 *)(in + b00), r01 = *(const uint2 *)(in + b00 + nx);
        const uint2 r10 = *(const uint2 *)(in + b00 + (uint64_t)ny * nx), r11 = *(const uint2 *)(in + b00 + (uint64_t)ny * nx + nx);
        uint32_t packed = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t w00 = k < 2 ? r00.x : r00.y, w01 = k < 2 ? r01.x : r01.y;
            const uint32_t w10 = k < 2 ? r10.x : r10.y, w11 = k < 2 ? r11.x : r11.y;
            const uint32_t sh = 16u * (k & 1);
            const uint32_t c[8] = {(w00 >> sh) & 255u, (w00 >> (sh + 8)) & 255u, (w01 >> sh) & 255u, (w01 >> (sh + 8)) & 255u,
                                   (w10 >> sh) & 255u, (w10 >> (sh + 8)) & 255u, (w11 >> sh) & 255u, (w11 >> (sh + 8)) & 255u};
            packed |= (uint32_t)ds_cell(c, method, alpha) << (8 * k);
        }
        *(uint32_t *)(out + ((uint64_t)z * oy + y) * ox + 4u * x4) = packed;
    }
}
__global__ void k_ds2x_v(const uint8_t *__restrict__ in, uint32_t nx, uint32_t ny, uint32_t nz, int method,
                         float alpha, uint8_t *__restrict__ out, int wide) {
    if (!wide && ds_small((uint64_t)(nx / 8) * (ny / 2) * (nz / 2)))
        ds2x_v_body<uint32_t>(in, nx, ny, nz, method, alpha, out);
    else
        ds2x_v_body<uint64_t>(in, nx, ny, nz, method, alpha, out);
}
template <typename IT>
__device__ __forceinline__ void ds2x_body(const uint8_t *__restrict__ in, uint32_t nx, uint32_t ny, uint32_t nz,
                                          int method, float alpha, uint8_t *__restrict__ out) {
    const uint32_t ox = (nx + 1) / 2, oy = (ny + 1) / 2, oz = (nz + 1) / 2;
    const IT total = (IT)ox * oy * oz;
    for (IT o = (IT)blockIdx.x * blockDim.x + threadIdx.x; o < total; o += (IT)gridDim.x * blockDim.x) {
        uint32_t x, y, z;
        ds_split<IT>(o, ox, oy, x, y, z);
        const uint32_t x0 = 2 * x, x1 = 2 * x + 1 < nx ? 2 * x + 1 : nx - 1;
        const uint32_t y0 = 2 * y, y1 = 2 * y + 1 < ny ? 2 * y + 1 : ny - 1;
        const uint32_t z0 = 2 * z, z1 = 2 * z + 1 < nz ? 2 * z + 1 : nz - 1;
        const uint8_t *r00 = in + ((uint64_t)z0 * ny + y0) * nx, *r01 = in + ((uint64_t)z0 * ny + y1) * nx;
        const uint8_t *r10 = in + ((uint64_t)z1 * ny + y0) * nx, *r11 = in + ((uint64_t)z1 * ny + y1) * nx;
        const uint32_t c[8] = {r00[x0], r00[x1], r01[x0], r01[x1], r10[x0], r10[x1], r11[x0], r11[x1]};
        out[o] = ds_cell(c, method, alpha);
    }
}
__global__ void k_ds2x(const uint8_t *__restrict__ in, uint32_t nx, uint32_t ny, uint32_t nz, int method, float alpha,
                       uint8_t *__restrict__ out, int wide) {
    if (!wide && ds_small((uint64_t)((nx + 1) / 2) * ((ny + 1) / 2) * ((nz + 1) / 2)))
        ds2x_body<uint32_t>(in, nx, ny, nz, method, alpha, out);
    else
        ds2x_body<uint64_t>(in, nx, ny, nz, method, alpha, out);
}

enum { FDCT_STORE_COEF = 0, FDCT_QUANT = 1 };

// Block i handles flat atom base + i of the segments (in FDCT_STORE_COEF
// mode a volume's atoms are its sample list).  FDCT_QUANT: levels to lv_out
// and items to items_out, both indexed by i; FDCT_STORE_COEF: coefficients
// (scan order) and items to the volume's sample arrays, indexed by the
// atom's place in its list.
__global__ __launch_bounds__(FDCT_TX) void k_fdct(const enc_src *__restrict__ srcs,
                                                       const enc_seg *__restrict__ segs, uint32_t nseg,
                                                       uint64_t base, int mode,
                                                       const vcg_dtab *__restrict__ T,
                                                       int16_t *__restrict__ lv_out,
                                                       enc_item *__restrict__ items_out) {
    extern __shared__ int buf[];
    __shared__ uint32_t ssum;
    __shared__ int seob;
    __shared__ float sstep[VCG_NFREQ], sdz[VCG_NFREQ];    // the quantizer's step and dead zone per frequency
    const int tid = threadIdx.x;
    const uint32_t i = blockIdx.x;
    const uint64_t flat = base + i;
    const enc_seg &sg = segs[nseg > 1 ? seg_of(segs, nseg, flat) : 0];
    const enc_src &s = srcs[sg.v];
    const uint64_t gi = sg.gi0 + (flat - sg.base);
    int lod; uint32_t ax, ay, az;
    const bool store = mode == FDCT_STORE_COEF;
    enc_atom_of(s, store ? s.list : nullptr, gi, lod, ax, ay, az);
    enc_item *items = store ? s.sitems + gi : items_out + i;
    if (tid == 0) { ssum = 0; seob = 0; }
    if (!store) vcg_quant_tables(T->gstep, s.q[lod], sstep, sdz, tid, FDCT_TX);
    __syncthreads();
    // pass 1 on the matrix cores (vcg_xform.hpp): wave w takes z-planes w,
    // w+8, w+16, w+24; lane (y, h) loads the 16 voxels x = 16h.. of row (z, y)
    // of each, which also give the atom's sum (and so dc) before any product.
    // In the products below y = lane mod 32 is also the output column (kx in
    // pass 1, ky in pass 2, kx in pass 3).
    const int wv = tid >> 6, ln = tid & 63, y = ln & 31, hh = ln >> 5;
    v4i raw[4];
    uint32_t part = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        uint32_t w[4];
        vcg_fwd1_row(s.lod[lod], s.nx[lod], s.ny[lod], s.nz[lod], ax, ay, az, wv + 8 * b, y, hh, w);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            part = __builtin_amdgcn_sad_u8(w[q], 0u, part);
            raw[b][q] = (int)(w[q] ^ 0x80808080u);     // v - 128 as int8
        }
    }
    if (part) atomicAdd(&ssum, part);
    __syncthreads();
    const uint32_t sum = ssum;
    if (sum == 0) {                          // all-zero atom: KNOWN_ZERO, no payload
        if (tid == 0) { enc_item it = {0, 0, 1, (uint32_t)lod}; *items = it; }
        return;
    }
    const int dc = (int)((sum + 16384u) / 32768u);
    // the forward passes on the matrix cores (vcg_xform.hpp): passes 1 and 2
    // into the i16 tile (the first 64 KiB of buf), pass 3 into the i32 tile
    uint8_t *tile = (uint8_t *)buf;
    const v4i bh = *(const v4i *)T->fa[0][ln], bl = *(const v4i *)T->fa[1][ln], b4l = *(const v4i *)T->fa[2][ln];
    {
        const int corr = vcg_fwd1_corr(dc, T->rowsum[y]);      // output column kx = lane mod 32
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const v16i zero = {};
            // scaled by 4 (vcg_im_pack16h): the second product takes 4 Cl
            v16i t = __builtin_amdgcn_mfma_i32_32x32x32_i8(raw[b], bh, zero, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 16; ++r) t[r] = (int)(((uint32_t)t[r] << 8) + 4u * (uint32_t)corr);
            t = __builtin_amdgcn_mfma_i32_32x32x32_i8(raw[b], b4l, t, 0, 0, 0);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                uint32_t w[2];
                const int o[4] = {t[4 * q], t[4 * q + 1], t[4 * q + 2], t[4 * q + 3]};
                vcg_im_pack16h(o, w);
                *(uint2 *)(tile + vcg_im_out_addr(0, wv + 8 * b, y, hh, q)) = make_uint2(w[0], w[1]);
            }
        }
    }
    __syncthreads();
    {   // pass 2: every wave reads its blocks before any writes
        uint32_t d[4][8];
#pragma unroll
        for (int b = 0; b < 4; ++b) vcg_im_load(tile, 1, wv + 8 * b, y, hh, d[b]);
        __syncthreads();
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            uint32_t hi[4], lo[4];
            vcg_im_planes(d[b], hi, lo);
            const v4i ah = {(int)hi[0], (int)hi[1], (int)hi[2], (int)hi[3]};
            const v4i al = {(int)lo[0], (int)lo[1], (int)lo[2], (int)lo[3]};
            const v16i zero = {};
            v16i t = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah, bh, zero, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 16; ++r) t[r] = (int)((uint32_t)t[r] << 8);
            t = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah, b4l, t, 0, 0, 0);
            t = __builtin_amdgcn_mfma_i32_32x32x32_i8(al, bh, t, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 16; ++r) t[r] = (int)(((uint32_t)t[r] << 8) + (uint32_t)vcg_im_rnd4(1, 0));
            t = __builtin_amdgcn_mfma_i32_32x32x32_i8(al, b4l, t, 0, 0, 0);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                uint32_t w[2];
                const int o[4] = {t[4 * q], t[4 * q + 1], t[4 * q + 2], t[4 * q + 3]};
                vcg_im_pack16h(o, w);
                *(uint2 *)(tile + vcg_im_out_addr(1, wv + 8 * b, y, hh, q)) = make_uint2(w[0], w[1]);
            }
        }
    }
    __syncthreads();
    {   // pass 3: basis as A (rows kz), data line (ky, kx) as B (column kx)
        uint32_t d[4][8];
#pragma unroll
        for (int b = 0; b < 4; ++b) vcg_im_load(tile, 2, wv + 8 * b, y, hh, d[b]);
        __syncthreads();
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            uint32_t hi[4], lo[4];
            vcg_im_planes(d[b], hi, lo);
            const v4i ah = {(int)hi[0], (int)hi[1], (int)hi[2], (int)hi[3]};
            const v4i al = {(int)lo[0], (int)lo[1], (int)lo[2], (int)lo[3]};
            const v16i zero = {};
            v16i t = __builtin_amdgcn_mfma_i32_32x32x32_i8(bh, ah, zero, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 16; ++r) t[r] = (int)((uint32_t)t[r] << 8);
            t = __builtin_amdgcn_mfma_i32_32x32x32_i8(b4l, ah, t, 0, 0, 0);
            t = __builtin_amdgcn_mfma_i32_32x32x32_i8(bh, al, t, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 16; ++r) t[r] = (int)(((uint32_t)t[r] << 6) + 8192u);
            t = __builtin_amdgcn_mfma_i32_32x32x32_i8(bl, al, t, 0, 0, 0);
            const int ky = wv + 8 * b;
#pragma unroll
            for (int r = 0; r < 16; ++r) buf[vcg_tidx(vcg_fwd1_k(r, hh), ky, y)] = t[r] >> 14;
        }
    }
    __syncthreads();
    if (store) {
        int32_t *co = s.coef + gi * VCG_A3;
        for (int p0 = tid; p0 < 32768; p0 += 8 * FDCT_TX) {
            uint32_t idx[8];                 // eight scan-table loads in flight
#pragma unroll
            for (int k = 0; k < 8; ++k) idx[k] = T->scan[p0 + k * FDCT_TX];
#pragma unroll
            for (int k = 0; k < 8; ++k)
                co[p0 + k * FDCT_TX] = buf[vcg_tidx((int)(idx[k] >> 10), (int)((idx[k] >> 5) & 31), (int)(idx[k] & 31))];
        }
        if (tid == 0) { enc_item it = {0, (uint32_t)dc, 0, (uint32_t)lod}; *items = it; }
        return;
    }
    const int last = vcg_quant_scan(buf, T->scan, sstep, sdz, lv_out + (size_t)i * VCG_A3, tid, FDCT_TX);
    if (last) atomicMax(&seob, last);
    __syncthreads();
    if (tid == 0) { enc_item it = {(uint32_t)seob, (uint32_t)dc, 0, (uint32_t)lod}; *items = it; }
}

// Rate-control candidates: group g = (volume, base_q); its n sample atoms
// occupy rows [off, off + n) of the candidate scratch.
struct samp_group {
    const int32_t *coef;         // the volume's cached sample coefficients
    const enc_item *items;
    uint64_t off;
    uint32_t n;
    float q;
    uint32_t budget;             // byte total past which the group cannot meet its target
    uint32_t pad;
};
__device__ __forceinline__ uint32_t group_of(const samp_group *G, uint32_t ng, uint64_t row) {
    uint32_t lo = 0, hi = ng;
    while (hi - lo > 1) {
        const uint32_t m = (lo + hi) / 2;
        if (G[m].off <= row) lo = m; else hi = m;
    }
    return lo;
}

// Rows of a rate-control launch are counted heaviest first: k_quant_groups
// files each row under a bin of its estimated bin count (EOB + 3 x nonzero
// levels, RC_BIN_SHIFT wide), and k_rc_count_groups hands its lanes the rows
// of the heaviest bin first, so that the lanes of a wave finish together.
// The byte totals do not depend on the order (each row's count is exact, and
// a group passes its budget whatever order its rows are counted in).
#define RC_BINS 32u
#define RC_BIN_SHIFT 12u

// quantize every candidate's sample atoms (coefficients in scan order):
// block = one row, four consecutive positions per thread and step.  bins:
// RC_BINS counters (zero on entry); row filed at perm[bin * pcap + k].
__global__ __launch_bounds__(FDCT_THREADS) void k_quant_groups(const samp_group *__restrict__ G, uint32_t ng,
                                                               const vcg_dtab *__restrict__ T,
                                                               int16_t *__restrict__ lv_out,
                                                               uint32_t *__restrict__ eobs,
                                                               uint32_t *__restrict__ bins,
                                                               uint32_t *__restrict__ perm, uint32_t pcap) {
    __shared__ int seob;
    __shared__ uint32_t snz;
    __shared__ float gs[VCG_NFREQ];
    const int tid = threadIdx.x;
    const uint64_t row = blockIdx.x;
    const samp_group &g = G[group_of(G, ng, row)];
    const uint32_t i = (uint32_t)(row - g.off);
    if (g.items[i].zero) {
        if (tid == 0) { eobs[row] = 0; perm[atomicAdd(&bins[0], 1u)] = (uint32_t)row; }
        return;
    }
    if (tid == 0) { seob = 0; snz = 0; }
    if (tid < VCG_NFREQ) gs[tid] = T->gstep[tid];
    __syncthreads();
    const float q = g.q;
    const int4 *co = (const int4 *)(g.coef + (size_t)i * VCG_A3);
    const uint32_t *fq = (const uint32_t *)T->fscan;
    uint2 *lo = (uint2 *)(lv_out + row * VCG_A3);
    int last = 0;
    uint32_t nz = 0;
    for (int k = tid; k < 32768 / 4; k += FDCT_THREADS) {
        const int4 c = co[k];
        const uint32_t f = fq[k];
        const int l0 = vcg_quant1(c.x, q, gs[f & 0xFF]);
        const int l1 = vcg_quant1(c.y, q, gs[(f >> 8) & 0xFF]);
        const int l2 = vcg_quant1(c.z, q, gs[(f >> 16) & 0xFF]);
        const int l3 = vcg_quant1(c.w, q, gs[f >> 24]);
        lo[k] = make_uint2((uint32_t)(uint16_t)l0 | ((uint32_t)(uint16_t)l1 << 16),
                           (uint32_t)(uint16_t)l2 | ((uint32_t)(uint16_t)l3 << 16));
        const int p = 4 * k;
        if (l3) last = p + 4; else if (l2) last = p + 3; else if (l1) last = p + 2; else if (l0) last = p + 1;
        nz += (l0 != 0) + (l1 != 0) + (l2 != 0) + (l3 != 0);
    }
    if (last) atomicMax(&seob, last);
    if (nz) atomicAdd(&snz, nz);
    __syncthreads();
    if (tid == 0) {
        const uint32_t key = (uint32_t)seob + 3u * snz;
        const uint32_t b = key >> RC_BIN_SHIFT < RC_BINS - 1u ? key >> RC_BIN_SHIFT : RC_BINS - 1u;
        perm[b * pcap + atomicAdd(&bins[b], 1u)] = (uint32_t)row;
        eobs[row] = (uint32_t)seob;
    }
}

// per-lane context memory (VCG_CTX_SLOTS slots), slot-major over the 64
// lanes of the wave so a wave's access is conflict-free.  m is the array (a
// fixed LDS address), l4 the lane's byte offset (< 256); handles are byte
// offsets, and sink (the sink slot's) is kept opaque to the compiler so that
// picking a handle stays one select after the shift-or.
struct enc_lds_ctx {
    uint32_t *m;
    const uint8_t *nx;           // the next-state table (vcg_fsm_next)
    uint32_t l4, sink;
    __device__ __forceinline__ uint32_t at(uint32_t S, bool isCtx) const { return isCtx ? (S << 8) | l4 : sink; }
    __device__ __forceinline__ uint32_t nxt(uint32_t i) const { return nx[i]; }
    __device__ __forceinline__ uint32_t ld_next(uint32_t h) const {
        return *(const uint32_t *)((const char *)m + h + 256u);
    }
    __device__ __forceinline__ void st(uint32_t h, uint32_t v) { *(uint32_t *)((char *)m + h) = v; }
};
// (the next-state table is filled by the kernel before its barrier)
__device__ __forceinline__ enc_lds_ctx enc_ctx_init(uint32_t *ctxm, const uint8_t *nx) {
    for (uint32_t k = 1; k < VCG_CTX_SLOTS; ++k)
        ctxm[k * 64u + threadIdx.x] = k == 13u || k == 15u ? VCG_P_BYPASS : 2048u;
    uint32_t sink = (14u << 8) | (threadIdx.x * 4u);
    asm volatile("" : "+v"(sink));
    return enc_lds_ctx{ctxm, nx, threadIdx.x * 4u, sink};
}
// per-lane ring of four 8-level blocks (64 B), lanes 68 B apart so the lanes'
// reads spread over the LDS banks
#define ENC_LV_STRIDE 68u
struct enc_lds_lv {
    uint8_t *b;
    __device__ __forceinline__ int ld(uint32_t p) const { return *(const int16_t *)(b + ((p & 31u) << 1)); }
    __device__ __forceinline__ void st(uint32_t k, const vcg_blk16 &blk) {
        uint32_t *d = (uint32_t *)(b + ((k & 3u) << 4));
        d[0] = blk.w[0]; d[1] = blk.w[1]; d[2] = blk.w[2]; d[3] = blk.w[3];
    }
};

// byte count of every row (lane per row, the state-machine encoder in count
// mode), payload bytes not stored.  A group whose running total passes its
// budget is abandoned (lanes report 0xFFFFFFFE); 0xFFFFFFFF = coder bound
// exceeded.
// A lane reports to its group's running total every RC_REPORT bytes it
// counts.  The total only counts whole reports, so it passes a budget only if
// the group's full total does: the report size changes how soon a group is
// abandoned, never whether.
#define RC_REPORT 256u

// bins, perm: the rows filed by k_quant_groups (bins[b] rows at perm + b * pcap)
__global__ __launch_bounds__(64) void k_rc_count_groups(const samp_group *__restrict__ G, uint32_t ng,
                                                        uint64_t rows, const int16_t *__restrict__ lv,
                                                        const uint32_t *__restrict__ eobs,
                                                        const vcg_dtab *__restrict__ T, uint32_t *__restrict__ kb,
                                                        uint32_t *__restrict__ lens,
                                                        const uint32_t *__restrict__ bins,
                                                        const uint32_t *__restrict__ perm, uint32_t pcap) {
    __shared__ uint32_t ctxm[VCG_CTX_SLOTS * 64];
    __shared__ __align__(16) uint8_t lvm[64 * ENC_LV_STRIDE];
    __shared__ uint32_t bands[9];
    __shared__ uint8_t nxt[512];
    __shared__ uint32_t bpre[RC_BINS];
    if (threadIdx.x < 9) bands[threadIdx.x] = T->band_start[threadIdx.x];
    for (uint32_t t = threadIdx.x; t < 512u; t += 64u) nxt[t] = (uint8_t)vcg_fsm_next(t & 255u, t >> 8);
    if (threadIdx.x == 0) {                               // lanes before bin b (heaviest bin first)
        uint32_t a = 0;
        for (uint32_t b = RC_BINS; b-- > 0;) { bpre[b] = a; a += bins[b]; }
    }
    __syncthreads();
    const uint64_t lane = (uint64_t)blockIdx.x * 64u + threadIdx.x;
    const bool live = lane < rows;
    uint32_t b = RC_BINS - 1u;
    while (b > 0u && lane >= bpre[b - 1u]) --b;          // the lightest bin that starts at or before lane
    const uint64_t row = live ? perm[b * pcap + (uint32_t)(lane - bpre[b])] : 0u;
    const uint64_t r = row;
    const uint32_t gi = group_of(G, ng, r);
    const samp_group &g = G[gi];
    const bool zero = g.items[r - g.off].zero != 0;
    enc_lds_ctx ctx = enc_ctx_init(ctxm, nxt);
    enc_lds_lv lr{lvm + threadIdx.x * ENC_LV_STRIDE};
    vcg_efsm f;
    // lanes past the rows and all-zero atoms run as finished empty encodes
    vcg_efsm_init<false>(f, lr, lv + r * VCG_A3, live && !zero ? eobs[r] : 0u, bands, nullptr, VCG_RC_CAP);
    if (!live || zero) f.done = 1;
    int stop = 0;
    uint32_t kbm = RC_REPORT;
    uint32_t *kg = kb + gi;
    while (__any(!f.done)) {
        vcg_efsm_refill(f, lr);
        // edge tests only where some lane of the wave can meet an edge (a
        // wave-uniform branch)
        vcg_efsm_block<false>(f, ctx, lr, bands, nullptr, __any(vcg_efsm_near(f, VCG_EFSM_REFILL)));
        // every RC_REPORT bytes a lane adds to its group's total; past the
        // budget the group is abandoned
        if (!f.done && f.pos >= kbm) {
            while (!f.done && f.pos >= kbm) {
                kbm += RC_REPORT;
                if (atomicAdd(kg, RC_REPORT) + RC_REPORT > g.budget) { stop = 1; f.done = 1; f.lim = 0xFFFFFFFFu; }
            }
        }
    }
    if (!live) return;
    lens[row] = zero ? 0u : stop ? 0xFFFFFFFEu : f.len > VCG_RC_CAP ? 0xFFFFFFFFu : f.len + 1;
}

// ---------------------------------------------------------------------------
// one lane per atom (the state-machine encoder): payload = [dc][range-coder bytes]
__global__ __launch_bounds__(64) void k_rc_encode(uint32_t n, const int16_t *__restrict__ lv,
                                                  const enc_item *__restrict__ items,
                                                  const vcg_dtab *__restrict__ T,
                                                  uint8_t *__restrict__ pay, uint32_t pay_cap,
                                                  uint32_t *__restrict__ lens) {
    __shared__ uint32_t ctxm[VCG_CTX_SLOTS * 64];
    __shared__ __align__(16) uint8_t lvm[64 * ENC_LV_STRIDE];
    __shared__ uint32_t bands[9];
    __shared__ uint8_t nxt[512];
    if (threadIdx.x < 9) bands[threadIdx.x] = T->band_start[threadIdx.x];
    for (uint32_t t = threadIdx.x; t < 512u; t += 64u) nxt[t] = (uint8_t)vcg_fsm_next(t & 255u, t >> 8);
    __syncthreads();
    const uint32_t i = blockIdx.x * 64u + threadIdx.x;
    const bool live = i < n;
    const uint32_t r = live ? i : 0u;
    const enc_item it = items[r];
    const bool zero = it.zero != 0;
    // lanes past the items and all-zero atoms run as finished empty encodes
    // whose stores all go to byte 1 of their sink: their own (unused) slot, or
    // slot n, the pool's spare one past the last item
    const bool quiet = !live || zero;
    uint8_t *out = pay + (size_t)(live ? i : n) * pay_cap;
    if (!quiet) out[0] = (uint8_t)it.dc;
    // the range-coder bytes start at out + 1; the slot keeps a spare byte past
    // them (the encoder's store sink once a lane is finished or full)
    const uint32_t cap = pay_cap - 2 < VCG_RC_CAP ? pay_cap - 2 : VCG_RC_CAP;
    enc_lds_ctx ctx = enc_ctx_init(ctxm, nxt);
    enc_lds_lv lr{lvm + threadIdx.x * ENC_LV_STRIDE};
    vcg_efsm f;
    vcg_efsm_init<true>(f, lr, lv + (size_t)r * VCG_A3, quiet ? 0u : it.eob, bands, out + 1, quiet ? 0u : cap);
    if (quiet) f.done = 1;
    while (__any(!f.done)) {
        vcg_efsm_refill(f, lr);
        vcg_efsm_block<true>(f, ctx, lr, bands, out + 1, __any(vcg_efsm_near(f, VCG_EFSM_REFILL)));
    }
    if (!live) return;
    lens[i] = zero ? 0u : f.ovf ? 0xFFFFFFFFu : f.len + 1;
}

// copy payload i (len at lens[i]) from batch scratch to the device address dst[i]
__global__ void k_compact(const uint8_t *__restrict__ pay, uint32_t pay_cap, const uint32_t *__restrict__ lens,
                          const uint64_t *__restrict__ dst) {
    const uint32_t i = blockIdx.x;
    const uint32_t len = lens[i];
    if (len == 0 || len == 0xFFFFFFFFu) return;
    const uint8_t *src = pay + (size_t)i * pay_cap;
    uint8_t *d = (uint8_t *)(uintptr_t)dst[i];
    for (uint32_t b = threadIdx.x; b < len; b += blockDim.x) d[b] = src[b];
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

// k_assemble for many archives in one image: atom i's payload (len[i] bytes
// at device address src[i]) to img + dst[i], its 16-byte slot (file offset
// poff[i], or the zero-atom flag) to img + sdst[i] unless that is ~0
__global__ void k_assemble_many(const uint64_t *__restrict__ src, const uint32_t *__restrict__ len_,
                                const uint64_t *__restrict__ dst, const uint64_t *__restrict__ sdst,
                                const uint64_t *__restrict__ poff, uint8_t *__restrict__ img) {
    const uint64_t i = blockIdx.x;
    const uint32_t len = len_[i];
    const uint64_t so = sdst[i];
    const uint8_t *s0 = (const uint8_t *)(uintptr_t)src[i];
    if (threadIdx.x == 0 && so != ~0ull) {
        uint8_t s[16];
        for (int b = 0; b < 16; ++b) s[b] = 0;
        if (len) {
            const uint64_t po = poff[i];
            for (int b = 0; b < 8; ++b) s[b] = (uint8_t)(po >> (8 * b));
            for (int b = 0; b < 4; ++b) s[8 + b] = (uint8_t)(len >> (8 * b));
            s[12] = 1; s[13] = s0[0];
        } else {
            s[12] = 2;
        }
        for (int b = 0; b < 16; ++b) img[so + b] = s[b];
    }
    if (len) {
        uint8_t *d = img + dst[i];
        for (uint32_t b = threadIdx.x; b < len; b += blockDim.x) d[b] = s0[b];
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static const size_t LDS_BYTES = (size_t)VCG_TILE_INTS * 4;

static void enc_setup() {
    static bool done = false;
    if (done) return;
    VCG_CHECK(hipFuncSetAttribute((const void *)k_fdct, hipFuncAttributeMaxDynamicSharedMemorySize, (int)LDS_BYTES));
    done = true;
}

struct vcg_enc {
    enc_src src;
    bool arena = false;                // device memory from the thread's encode arena
    uint64_t total = 0;                // atoms over every LOD
    uint8_t *lod_buf[VCG_NLOD] = {nullptr};
    // rate-control sample
    uint32_t *sample_list = nullptr;
    int32_t *sample_coef = nullptr;
    enc_item *sample_items = nullptr;
    int sample_n = 0;
    // payload pool: atom i's payload at pool + pool_off[i]
    uint8_t *pool = nullptr; uint64_t pool_cap = 0, pool_len = 0;
    std::vector<uint64_t> pool_off;
};

static void *dmalloc(size_t n) {
    void *p = nullptr;
    if (hipMalloc(&p, n ? n : 1) != hipSuccess) { (void)hipGetLastError(); vc__panic("out of device memory (encode)"); }
    return p;
}

// Bump arena for the encoders of one group: a few large chunks instead of a
// dozen allocations and frees per volume (hipFree waits for the device).
struct enc_arena {
    bool open = false;
    std::vector<std::pair<uint8_t *, size_t>> chunks;
    size_t used = 0;                  // bytes used in the last chunk
};
static thread_local enc_arena tl_arena;
#define ARENA_CHUNK (1ull << 30)

extern "C" void vcg_enc_arena_open(void) { tl_arena.open = true; }

static void scratch_release();

// closing the arena also releases the thread's encode scratch, which a
// group of many volumes grows to several GiB
extern "C" void vcg_enc_arena_close(void) {
    enc_arena &A = tl_arena;
    VCG_CHECK(hipStreamSynchronize(vcg_get_ctx()->s));
    for (auto &ch : A.chunks) VCG_CHECK(hipFree(ch.first));
    A.chunks.clear();
    A.used = 0;
    A.open = false;
    scratch_release();
}

static void *enc_alloc(bool arena, size_t n) {
    if (!arena) return dmalloc(n);
    enc_arena &A = tl_arena;
    n = (n + 255) & ~(size_t)255;
    if (A.chunks.empty() || A.used + n > A.chunks.back().second) {
        const size_t cs = n > ARENA_CHUNK ? n : ARENA_CHUNK;
        A.chunks.push_back({(uint8_t *)dmalloc(cs), cs});
        A.used = 0;
    }
    void *p = A.chunks.back().first + A.used;
    A.used += n;
    return p;
}
static void enc_free(bool arena, void *p) {
    if (!arena && p) VCG_CHECK(hipFree(p));
}

extern "C" int vcg_downsample(const void *in, const uint32_t d[3], void *out) {
    vcg_ctx *c = vcg_get_ctx();
    uint64_t total = (uint64_t)((d[0] + 1) / 2) * ((d[1] + 1) / 2) * ((d[2] + 1) / 2);
    uint64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_downsample, dim3((uint32_t)blocks), dim3(256), 0, c->s, (const uint8_t *)in, d[0], d[1],
                       d[2], (uint8_t *)out, g_ds_wide);
    VCG_CHECK(hipGetLastError());
    return 0;
}

extern "C" int vcg_ds2x(const void *in, const uint32_t d[3], int method, float alpha, void *out, int sync) {
    vcg_ctx *c = vcg_get_ctx();
    const bool vec = d[0] % 8 == 0 && d[1] % 2 == 0 && d[2] % 2 == 0 && ((uintptr_t)in & 7) == 0 &&
                     ((uintptr_t)out & 3) == 0;
    const uint64_t total = vec ? (uint64_t)(d[0] / 8) * (d[1] / 2) * (d[2] / 2)
                               : (uint64_t)((d[0] + 1) / 2) * ((d[1] + 1) / 2) * ((d[2] + 1) / 2);
    uint64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (vec)
        hipLaunchKernelGGL(k_ds2x_v, dim3((uint32_t)blocks), dim3(256), 0, c->s, (const uint8_t *)in, d[0], d[1],
                           d[2], method, alpha, (uint8_t *)out, g_ds_wide);
    else
        hipLaunchKernelGGL(k_ds2x, dim3((uint32_t)blocks), dim3(256), 0, c->s, (const uint8_t *)in, d[0], d[1], d[2],
                           method, alpha, (uint8_t *)out, g_ds_wide);
    VCG_CHECK(hipGetLastError());
    if (sync) VCG_CHECK(hipStreamSynchronize(c->s));
    return 0;
}

static uint8_t *desc_area(size_t n);

// Encoders of nv volumes: each volume uploaded (or adopted) and its pyramid
// built, one launch per level across the volumes whose level outputs stay
// below 2^31 voxels (the others, and every volume under the wide-index test
// hook, take a launch of their own).
extern "C" void vcg_enc_begin_many(int nv, const uint8_t *const *vols, int on_dev, const uint32_t (*dims)[3],
                                   const int *nlods, vcg_enc **out) {
    vcg_ctx *c = vcg_get_ctx();
    enc_setup();
    int maxl = 0;
    for (int v = 0; v < nv; ++v) {
        vcg_enc *e = new vcg_enc();
        memset(&e->src, 0, sizeof e->src);
        e->arena = tl_arena.open;
        const int nlod = nlods[v];
        uint32_t x = dims[v][0], y = dims[v][1], z = dims[v][2];
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
        e->total = acc;
        const size_t n0 = (size_t)dims[v][0] * dims[v][1] * dims[v][2];
        if (on_dev) {
            e->src.lod[0] = vols[v];
        } else {
            e->lod_buf[0] = (uint8_t *)enc_alloc(e->arena, n0);
            VCG_CHECK(hipMemcpyAsync(e->lod_buf[0], vols[v], n0, hipMemcpyHostToDevice, c->s));
            e->src.lod[0] = e->lod_buf[0];
        }
        for (int l = 1; l < nlod; ++l) {
            e->lod_buf[l] = (uint8_t *)enc_alloc(e->arena, (size_t)e->src.nx[l] * e->src.ny[l] * e->src.nz[l]);
            e->src.lod[l] = e->lod_buf[l];
        }
        if (nlod > maxl) maxl = nlod;
        out[v] = e;
    }
    // level l from level l - 1: the batched jobs of every level in one upload.
    // The table outlives the call (thread-local), so that the host need not
    // wait for the upload; this thread's next call reuses it only after the
    // encode that follows this one has synchronised its stream.
    static thread_local std::vector<ds_job> jobs;
    jobs.clear();
    std::vector<size_t> lv0(VCG_NLOD + 1, 0);       // level l's jobs: [lv0[l], lv0[l + 1])
    std::vector<uint32_t> lvblocks(VCG_NLOD + 1, 0);
    std::vector<uint8_t> alone((size_t)nv * VCG_NLOD, 0);
    for (int l = 1; l < maxl; ++l) {
        lv0[l] = jobs.size();
        uint64_t blocks = 0;
        for (int v = 0; v < nv; ++v) {
            const vcg_enc *e = out[v];
            if (l >= e->src.nlod) continue;
            const uint64_t total = (uint64_t)e->src.nx[l] * e->src.ny[l] * e->src.nz[l];
            const uint64_t nb = (total + DS_JOB_THREADS - 1) / DS_JOB_THREADS;
            if (g_ds_wide || total >= (1ull << 31) || blocks + nb >= (1ull << 31)) {
                alone[(size_t)v * VCG_NLOD + l] = 1;
                continue;
            }
            jobs.push_back(ds_job{e->src.lod[l - 1], e->lod_buf[l], e->src.nx[l - 1], e->src.ny[l - 1],
                                  e->src.nz[l - 1], (uint32_t)blocks});
            blocks += nb;
        }
        lvblocks[l] = (uint32_t)blocks;
    }
    for (int l = maxl > 1 ? maxl : 1; l <= VCG_NLOD; ++l) lv0[l] = jobs.size();
    ds_job *dj = nullptr;
    if (!jobs.empty()) {
        dj = (ds_job *)desc_area(jobs.size() * sizeof(ds_job));
        VCG_CHECK(hipMemcpyAsync(dj, jobs.data(), jobs.size() * sizeof(ds_job), hipMemcpyHostToDevice, c->s));
    }
    for (int l = 1; l < maxl; ++l) {
        const size_t nj = lv0[l + 1] - lv0[l];
        if (nj) {
            hipLaunchKernelGGL(k_downsample_jobs, dim3(lvblocks[l]), dim3(DS_JOB_THREADS), 0, c->s, dj + lv0[l],
                               (uint32_t)nj);
            VCG_CHECK(hipGetLastError());
        }
        for (int v = 0; v < nv; ++v) {
            if (!alone[(size_t)v * VCG_NLOD + l]) continue;
            const vcg_enc *e = out[v];
            const uint32_t d[3] = {e->src.nx[l - 1], e->src.ny[l - 1], e->src.nz[l - 1]};
            vcg_downsample(e->src.lod[l - 1], d, e->lod_buf[l]);
        }
    }
}

// An encoder over level volumes the caller already built (device memory,
// dims[l] = nx, ny, nz of level l): atom numbering and gathers as for a
// pyramid, nothing computed or freed here.
extern "C" vcg_enc *vcg_enc_begin_levels(const uint8_t *const *lods, const uint32_t (*dims)[3], int nlod) {
    if (nlod < 1 || nlod > VCG_NLOD) return nullptr;
    vcg_get_ctx();
    enc_setup();
    vcg_enc *e = new vcg_enc();
    memset(&e->src, 0, sizeof e->src);
    e->arena = tl_arena.open;
    uint64_t acc = 0;
    for (int l = 0; l < nlod; ++l) {
        const uint32_t x = dims[l][0], y = dims[l][1], z = dims[l][2];
        e->src.nx[l] = x; e->src.ny[l] = y; e->src.nz[l] = z;
        e->src.acx[l] = (x + 31) / 32; e->src.acy[l] = (y + 31) / 32;
        e->src.prefix[l] = acc;
        acc += (uint64_t)((x + 31) / 32) * ((y + 31) / 32) * ((z + 31) / 32);
        e->src.lod[l] = lods[l];
    }
    e->src.prefix[nlod] = acc;
    e->src.nlod = nlod;
    e->total = acc;
    return e;
}

// dst (dd = padded dims, device) = src (sd dims, host or device by pointer)
// with zeros beyond sd; on the calling thread's stream, synchronous
extern "C" int vcg_pad_volume(const uint8_t *src, const uint32_t sd[3], uint8_t *dst, const uint32_t dd[3]) {
    if (sd[0] > dd[0] || sd[1] > dd[1] || sd[2] > dd[2]) return -1;
    vcg_ctx *c = vcg_get_ctx();
    const size_t plane = (size_t)dd[0] * dd[1];
    if (sd[0] != dd[0] || sd[1] != dd[1] || sd[2] != dd[2])
        VCG_CHECK(hipMemsetAsync(dst, 0, plane * dd[2], c->s));
    for (uint32_t z = 0; z < sd[2]; ++z)
        VCG_CHECK(hipMemcpy2DAsync(dst + z * plane, dd[0], src + (size_t)z * sd[0] * sd[1], sd[0], sd[0], sd[1],
                                   hipMemcpyDefault, c->s));
    VCG_CHECK(hipStreamSynchronize(c->s));
    return 0;
}

// scratch shared by the encode launches of this thread
struct enc_scratch {
    int16_t *lv = nullptr; size_t lv_cap = 0;
    enc_item *items = nullptr; size_t items_cap = 0;
    uint8_t *pay = nullptr; size_t pay_cap = 0;
    uint32_t *lens = nullptr; size_t lens_cap = 0;
    uint32_t *aux = nullptr; size_t aux_cap = 0;      // rate-control candidates: eobs, counters
    uint8_t *desc = nullptr; size_t desc_cap = 0;     // per-call descriptors (sources, groups, addresses)
    uint8_t *img = nullptr; size_t img_cap = 0;       // archive image
    // the coarse atoms' launch on its own stream (vcg_enc_all)
    int16_t *clv = nullptr; size_t clv_cap = 0;
    enc_item *citems = nullptr; size_t citems_cap = 0;
    uint8_t *cpay = nullptr; size_t cpay_cap = 0;
    uint32_t *clens = nullptr; size_t clens_cap = 0;
    uint8_t *hstage[2] = {nullptr, nullptr};          // pinned halves of one allocation
    hipEvent_t hev[2] = {nullptr, nullptr};
};
#define HSTAGE_BYTES (32u << 20)

// memcpy into a freshly allocated (pageable) buffer: the first touch of its
// pages dominates, so large copies are split in 1 MiB parts over the caller
// and a pool of host threads (started once, kept until the process ends;
// a batch of hundreds of archives would otherwise start thousands of
// threads).  One job at a time uses the pool; a concurrent caller copies
// alone.  The destination is offered transparent huge pages first (fewer
// first-touch faults where the kernel allows them).  The pool also takes a
// background job that touches the pages of a list of buffers (the archives
// an encode will fill), started before the device work and joined after it,
// so that the faults overlap the encode instead of the copies.
struct copy_pool {
    std::mutex busy, m;
    std::condition_variable go, fin;
    uint64_t gen = 0;
    int active = 0;                       // workers inside the current job (under m)
    // the job: one copy (touch == false) or the pages of tb[i][0, tl[i])
    bool touch = false;
    uint8_t *dst = nullptr;
    const uint8_t *src = nullptr;
    size_t n = 0, parts = 0;
    std::vector<uint8_t *> tb;
    std::vector<size_t> tl, tpre;         // tpre[i]: first part of buffer i
    std::atomic<size_t> next{0}, done{0};
    int nthr = 0;
    void part(size_t k) {
        if (!touch) {
            const size_t o = k << 20, len = n - o < (1u << 20) ? n - o : (1u << 20);
            memcpy(dst + o, src + o, len);
            return;
        }
        const size_t i = (size_t)(std::upper_bound(tpre.begin(), tpre.end(), k) - tpre.begin()) - 1;
        const size_t o = (k - tpre[i]) << 20, len = tl[i] - o < (1u << 20) ? tl[i] - o : (1u << 20);
        volatile uint8_t *b = tb[i] + o;
        for (size_t q = 0; q < len; q += 4096) b[q] = 0;
    }
    // parts are claimed from `next`; the job's fields change only while no
    // worker is inside a job (active == 0, under m)
    void run_parts() {
        for (size_t k; (k = next.fetch_add(1)) < parts;) {
            part(k);
            done.fetch_add(1);
        }
    }
    void worker() {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> l(m);
                go.wait(l, [&] { return gen != seen; });
                seen = gen;
                ++active;
            }
            run_parts();
            {
                std::lock_guard<std::mutex> l(m);
                --active;
            }
            fin.notify_all();
        }
    }
    // with busy held: publish a job whose fields the caller set through f
    template <class F> void start(F f) {
        {
            // a worker that woke for the previous job after it ended leaves
            // first, so the job's fields change only with no worker inside
            std::unique_lock<std::mutex> l(m);
            fin.wait(l, [&] { return active == 0; });
            f();
            done.store(0);
            next.store(0);
            ++gen;
        }
        go.notify_all();
    }
    // with busy held: take parts too, then wait until every part is done and
    // no worker is still inside the job (one that joins late finds no part
    // left, or counts as active until it leaves)
    void join() {
        run_parts();
        std::unique_lock<std::mutex> l(m);
        fin.wait(l, [&] { return done.load() == parts && active == 0; });
    }
};
static copy_pool *pool_get() {
    static copy_pool *p = [] {
        copy_pool *q = new copy_pool;            // never destroyed: its threads live until exit
        const unsigned hw = std::thread::hardware_concurrency();
        q->nthr = hw >= 16 ? 7 : hw >= 4 ? (int)hw / 2 - 1 : 0;
        for (int t = 0; t < q->nthr; ++t) std::thread([q] { q->worker(); }).detach();
        return q;
    }();
    return p;
}
static void thp_advise(uint8_t *dst, size_t n) {
    if (n < (4u << 20)) return;
    const uintptr_t a0 = ((uintptr_t)dst + 4095u) & ~(uintptr_t)4095u;
    const uintptr_t a1 = ((uintptr_t)dst + n) & ~(uintptr_t)4095u;
    if (a1 > a0) (void)madvise((void *)a0, a1 - a0, MADV_HUGEPAGE);
}
// the background touch job of this thread, if one runs (it holds busy)
static thread_local bool tl_touching = false;

static void par_copy(uint8_t *dst, const uint8_t *src, size_t n) {
    thp_advise(dst, n);
    copy_pool *p = n >= (2u << 20) && !tl_touching ? pool_get() : nullptr;
    if (!p || p->nthr == 0 || !p->busy.try_lock()) { memcpy(dst, src, n); return; }
    p->start([&] { p->touch = false; p->dst = dst; p->src = src; p->n = n; p->parts = (n + (1u << 20) - 1) >> 20; });
    p->join();
    p->busy.unlock();
}
extern "C" void vcg_host_prefault_begin(uint8_t *const *bufs, const size_t *lens, int n) {
    if (tl_touching) return;
    copy_pool *p = pool_get();
    if (p->nthr == 0 || !p->busy.try_lock()) return;
    p->start([&] {
        p->touch = true;
        p->tb.clear(); p->tl.clear(); p->tpre.clear();
        size_t parts = 0;
        for (int i = 0; i < n; ++i) {
            if (!bufs[i] || !lens[i]) continue;
            thp_advise(bufs[i], lens[i]);
            p->tb.push_back(bufs[i]); p->tl.push_back(lens[i]); p->tpre.push_back(parts);
            parts += (lens[i] + (1u << 20) - 1) >> 20;
        }
        p->parts = parts;
    });
    tl_touching = true;
}

extern "C" void vcg_host_prefault_wait(void) {
    if (!tl_touching) return;
    copy_pool *p = pool_get();
    p->join();
    p->busy.unlock();
    tl_touching = false;
}
static thread_local enc_scratch tl_es;

static void scratch_release() {
    enc_scratch &S = tl_es;
    void *bufs[] = {S.lv, S.items, S.pay, S.lens, S.aux, S.desc, S.img, S.clv, S.citems, S.cpay, S.clens};
    for (void *b : bufs) if (b) VCG_CHECK(hipFree(b));
    if (S.hstage[0]) VCG_CHECK(hipHostFree(S.hstage[0]));
    for (int k = 0; k < 2; ++k) if (S.hev[k]) VCG_CHECK(hipEventDestroy(S.hev[k]));
    S = enc_scratch();
}

// atoms per encode launch (levels + payload scratch: 192 KiB each)
#define ENC_ROWS 65536u
// payload slot of k_rc_encode: dc byte, the range-coder bytes, one spare byte
#define ENC_SLOT (VCG_PAY_CAP + 1u)
// rate-control rows (candidate x sample atom) per launch: VCG_RC_ROWS
// (levels: 64 KiB each).  A round's rows share one launch up to this size:
// the count lanes are latency-bound, so more waves per SIMD cost little.
static uint32_t g_rc_rows = VCG_RC_ROWS;
extern "C" uint32_t vcg_rc_rows(void) { return g_rc_rows; }
extern "C" void vcg_set_rc_rows(uint32_t rows) {
    g_rc_rows = rows == 0 || rows > VCG_RC_ROWS ? VCG_RC_ROWS : rows < 512 ? 512 : rows;
}

// clear n words with a kernel (no memset node in code that may be captured)
__global__ void k_reset_words(uint32_t *w, uint64_t n) {
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x)
        w[k] = 0;
}

// descriptor scratch: bytes [0, n) of this call's descriptor area
static uint8_t *desc_area(size_t n) {
    vcg_grow((void **)&tl_es.desc, &tl_es.desc_cap, n, 1);
    return tl_es.desc;
}

// sources of volumes es[0..nv) and the segments of the flat atom order
// (device copies at *dsrc / *dseg); returns the flat atom count.  Samples:
// volume-major.  Full encode: coarsest LOD first across every volume, so the
// heaviest atoms (coarse LODs carry the most coefficients) share launches.
static uint64_t upload_sources(vcg_ctx *c, int nv, vcg_enc *const *es, const float (*qs)[VCG_NLOD],
                               bool sample, std::vector<uint8_t> &host, std::vector<enc_seg> &seg,
                               enc_src **dsrc, enc_seg **dseg) {
    seg.clear();
    uint64_t tot = 0;
    if (sample) {
        for (int v = 0; v < nv; ++v) {
            if (!es[v]->sample_n) continue;
            seg.push_back(enc_seg{tot, 0, (uint32_t)v, 0});
            tot += (uint64_t)es[v]->sample_n;
        }
    } else {
        for (int l = VCG_NLOD - 1; l >= 0; --l)
            for (int v = 0; v < nv; ++v) {
                const enc_src &s = es[v]->src;
                if (l >= s.nlod || s.prefix[l + 1] == s.prefix[l]) continue;
                seg.push_back(enc_seg{tot, s.prefix[l], (uint32_t)v, 0});
                tot += s.prefix[l + 1] - s.prefix[l];
            }
    }
    const size_t bo = ((size_t)nv * sizeof(enc_src) + 255) & ~(size_t)255;
    host.assign(bo + seg.size() * sizeof(enc_seg), 0);
    enc_src *hs = (enc_src *)host.data();
    for (int v = 0; v < nv; ++v) {
        enc_src s = es[v]->src;
        if (qs) for (int l = 0; l < VCG_NLOD; ++l) s.q[l] = qs[v][l];
        if (sample) { s.list = es[v]->sample_list; s.coef = es[v]->sample_coef; s.sitems = es[v]->sample_items; }
        hs[v] = s;
    }
    if (!seg.empty()) memcpy(host.data() + bo, seg.data(), seg.size() * sizeof(enc_seg));
    uint8_t *d = desc_area(host.size());
    VCG_CHECK(hipMemcpyAsync(d, host.data(), host.size(), hipMemcpyHostToDevice, c->s));
    *dsrc = (enc_src *)d;
    *dseg = (enc_seg *)(d + bo);
    return tot;
}

// Rate-control samples of nv encoders in one launch: encoder v's ns[v] listed
// LOD0 atoms are transformed and their coefficients kept on device;
// zero_masks[v][i] = 1 for its all-zero atoms.
extern "C" int vcg_enc_samples(int nv, vcg_enc *const *es, const uint32_t *const *lists, const int *ns,
                               uint8_t *const *zero_masks) {
    vcg_ctx *c = vcg_get_ctx();
    if (nv <= 0) return 0;
    // encoders in an arena (a batch) share one list block and one item
    // block, so that the lists go up and the items come back in one copy
    // each; an arena is released whole, so the shared blocks need no owner
    bool shared = nv > 1;
    for (int v = 0; v < nv; ++v) shared = shared && es[v]->arena;
    uint64_t tot = 0;
    for (int v = 0; v < nv; ++v) tot += (uint64_t)(ns[v] > 0 ? ns[v] : 0);
    uint32_t *lblk = nullptr;
    enc_item *iblk = nullptr;
    std::vector<uint32_t> hl;
    if (shared && tot) {
        lblk = (uint32_t *)enc_alloc(true, (size_t)tot * 12);
        iblk = (enc_item *)enc_alloc(true, (size_t)tot * sizeof(enc_item));
        hl.resize((size_t)tot * 3);
    }
    for (int v = 0, k = 0; v < nv; ++v) {
        vcg_enc *e = es[v];
        if (e->sample_coef) vc__panic("encode sample prepared twice");
        const int n = ns[v] > 0 ? ns[v] : 0;
        e->sample_n = n;
        e->sample_coef = (int32_t *)enc_alloc(e->arena, (size_t)n * VCG_A3 * 4);
        if (lblk) {
            e->sample_list = lblk + (size_t)k * 3;
            e->sample_items = iblk + k;
            if (n) memcpy(hl.data() + (size_t)k * 3, lists[v], (size_t)n * 12);
        } else {
            e->sample_list = (uint32_t *)enc_alloc(e->arena, (size_t)n * 12);
            e->sample_items = (enc_item *)enc_alloc(e->arena, (size_t)n * sizeof(enc_item));
            if (n) VCG_CHECK(hipMemcpyAsync(e->sample_list, lists[v], (size_t)n * 12, hipMemcpyHostToDevice, c->s));
        }
        k += n;
    }
    if (lblk) VCG_CHECK(hipMemcpyAsync(lblk, hl.data(), (size_t)tot * 12, hipMemcpyHostToDevice, c->s));
    std::vector<uint8_t> host;
    std::vector<enc_seg> seg;
    enc_src *dsrc; enc_seg *dseg;
    const uint64_t ns_tot = upload_sources(c, nv, es, nullptr, true, host, seg, &dsrc, &dseg);
    if (ns_tot) {
        hipLaunchKernelGGL(k_fdct, dim3((uint32_t)ns_tot), dim3(FDCT_TX), LDS_BYTES, c->s, dsrc, dseg,
                           (uint32_t)seg.size(), (uint64_t)0, (int)FDCT_STORE_COEF, g_dtab,
                           (int16_t *)nullptr, (enc_item *)nullptr);
        VCG_CHECK(hipGetLastError());
    }
    std::vector<enc_item> its(ns_tot);
    if (iblk) {
        if (ns_tot) VCG_CHECK(hipMemcpyAsync(its.data(), iblk, (size_t)ns_tot * sizeof(enc_item), hipMemcpyDeviceToHost, c->s));
    } else {
        for (int v = 0, k = 0; v < nv; k += es[v]->sample_n, ++v)
            if (es[v]->sample_n)
                VCG_CHECK(hipMemcpyAsync(its.data() + k, es[v]->sample_items,
                                         (size_t)es[v]->sample_n * sizeof(enc_item), hipMemcpyDeviceToHost, c->s));
    }
    VCG_CHECK(hipStreamSynchronize(c->s));
    for (int v = 0, k = 0; v < nv; k += es[v]->sample_n, ++v)
        for (int i = 0; i < es[v]->sample_n; ++i) zero_masks[v][i] = (uint8_t)its[k + i].zero;
    return 0;
}

// Sample byte totals of ng rate-control candidates: candidate g is encoder
// es[g]'s sample quantized at base_q qs[g].  bytes[g] = the total (dc bytes
// included); flags[g] bit 0 = coder bound exceeded, bit 1 = the total passed
// budgets[g] and the candidate was abandoned (bytes[g] is then not final).
extern "C" int vcg_enc_candidates(int ng, vcg_enc *const *es, const float *qs, const uint64_t *budgets,
                                  uint64_t *bytes, int *flags) {
    vcg_ctx *c = vcg_get_ctx();
    for (int g = 0; g < ng; ++g) { bytes[g] = 0; flags[g] = 0; }
    if (ng <= 0) return 0;
    enc_scratch &es_ = tl_es;
    // launches of whole groups, at most vcg_rc_rows() rows each
    const uint32_t cap = g_rc_rows;
    std::vector<int> cut{0};
    std::vector<samp_group> G(ng);
    uint64_t rows = 0, tot = 0;
    for (int g = 0; g < ng; ++g) {
        const uint32_t n = (uint32_t)es[g]->sample_n;
        if (n > cap) vc__panic("rate-control sample larger than a launch");
        if (rows + n > cap) { cut.push_back(g); rows = 0; }
        // a candidate's counted total stays below 2^32 (at most 512 atoms x the
        // coder bound), so a budget of 2^32-1 or more never stops an encode
        const uint32_t b32 = budgets[g] >= 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)budgets[g];
        G[g] = samp_group{es[g]->sample_coef, es[g]->sample_items, rows, n, qs[g], b32, 0};
        rows += n; tot += n;
    }
    cut.push_back(ng);
    const size_t maxrows = tot < cap ? (size_t)tot : (size_t)cap;
    vcg_grow((void **)&es_.lv, &es_.lv_cap, maxrows, (size_t)VCG_A3 * 2);
    vcg_grow((void **)&es_.lens, &es_.lens_cap, tot, 4);
    vcg_grow((void **)&es_.aux, &es_.aux_cap, maxrows * (1 + RC_BINS) + (size_t)ng + RC_BINS, 4);
    samp_group *dG = (samp_group *)desc_area((size_t)ng * sizeof(samp_group));
    VCG_CHECK(hipMemcpyAsync(dG, G.data(), (size_t)ng * sizeof(samp_group), hipMemcpyHostToDevice, c->s));
    uint32_t *eobs = es_.aux, *kb = es_.aux + maxrows, *bins = kb + ng, *perm = bins + RC_BINS;
    hipLaunchKernelGGL(k_reset_words, dim3(1), dim3(256), 0, c->s, kb, (uint64_t)ng);
    uint64_t done = 0;
    for (size_t k = 0; k + 1 < cut.size(); ++k) {
        const int g0 = cut[k], g1 = cut[k + 1];
        const uint64_t r = (uint64_t)G[g1 - 1].off + G[g1 - 1].n;
        if (r == 0) continue;
        hipLaunchKernelGGL(k_reset_words, dim3(1), dim3(64), 0, c->s, bins, (uint64_t)RC_BINS);
        hipLaunchKernelGGL(k_quant_groups, dim3((uint32_t)r), dim3(FDCT_THREADS), 0, c->s, dG + g0,
                           (uint32_t)(g1 - g0), g_dtab, es_.lv, eobs, bins, perm, (uint32_t)maxrows);
        VCG_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_rc_count_groups, dim3((uint32_t)((r + 63) / 64)), dim3(64), 0, c->s, dG + g0,
                           (uint32_t)(g1 - g0), r, es_.lv, eobs, g_dtab, kb + g0, es_.lens + done,
                           bins, perm, (uint32_t)maxrows);
        VCG_CHECK(hipGetLastError());
        done += r;
    }
    std::vector<uint32_t> lens(tot);
    if (tot) VCG_CHECK(hipMemcpyAsync(lens.data(), es_.lens, tot * 4, hipMemcpyDeviceToHost, c->s));
    VCG_CHECK(hipStreamSynchronize(c->s));
    for (int g = 0, k = 0; g < ng; ++g)
        for (uint32_t i = 0; i < G[g].n; ++i, ++k) {
            const uint32_t L = lens[k];
            if (L == 0xFFFFFFFFu) flags[g] |= 1;
            else if (L == 0xFFFFFFFEu) flags[g] |= 2;
            else bytes[g] += L;
        }
    return 0;
}

// make room for `need` bytes in e's payload pool (keeps its contents)
static void pool_reserve(vcg_ctx *c, vcg_enc *e, uint64_t need) {
    if (need <= e->pool_cap) return;
    const uint64_t nc = need + need / 2 + 4096;
    uint8_t *np = (uint8_t *)enc_alloc(e->arena, nc);
    if (e->pool_len) VCG_CHECK(hipMemcpyAsync(np, e->pool, e->pool_len, hipMemcpyDeviceToDevice, c->s));
    if (!e->arena) {
        VCG_CHECK(hipStreamSynchronize(c->s));
        enc_free(false, e->pool);
    }
    e->pool = np; e->pool_cap = nc;
}

// Encode every atom of every LOD of nv encoders (encoder v at qs[v][lod]) in
// launches of up to ENC_ROWS atoms taken across volumes, coarsest LOD first.
// lens[v][i]: payload length of encoder v's atom i (0 = all-zero atom);
// ovf[v] = 1 if one of its payloads exceeded the coder bound.
// The coarse atoms (LOD >= ENC_COARSE_LOD, first in the flat order) carry
// the most bins, and the launch that holds them runs as long as the heaviest
// one.  When there are at least ENC_SIDE_MIN atoms and the coarse ones number
// at most ENC_COARSE_MAX, they go in a launch of their own on a second stream
// (with their own scratch), beside the transform and launches of the rest.
#define ENC_COARSE_LOD 2
#define ENC_COARSE_MAX 16384u
#define ENC_SIDE_MIN 8192u

extern "C" int vcg_enc_all(int nv, vcg_enc *const *es, const float (*qs)[VCG_NLOD], uint32_t *const *lens,
                           int *ovf) {
    vcg_ctx *c = vcg_get_ctx();
    enc_scratch &S = tl_es;
    if (nv <= 0) return 0;
    std::vector<uint8_t> host;
    std::vector<enc_seg> seg;
    enc_src *dsrc; enc_seg *dseg;
    const uint64_t tot = upload_sources(c, nv, es, qs, false, host, seg, &dsrc, &dseg);
    for (int v = 0; v < nv; ++v) {
        es[v]->pool_off.assign(es[v]->total, 0);
        es[v]->pool_len = 0;
        ovf[v] = 0;
    }
    seg.push_back(enc_seg{tot, 0, 0, 0});           // sentinel
    uint64_t nc = 0;                                // coarse atoms: flat [0, nc)
    for (int l = VCG_NLOD - 1; l >= ENC_COARSE_LOD; --l)
        for (int v = 0; v < nv; ++v)
            if (l < es[v]->src.nlod) nc += es[v]->src.prefix[l + 1] - es[v]->src.prefix[l];
    const bool side = tot >= ENC_SIDE_MIN && nc > 0 && nc <= ENC_COARSE_MAX && nc < tot;
    const uint64_t b0 = side ? nc : 0;              // the main stream's atoms: flat [b0, tot)
    const uint32_t maxb = tot - b0 < ENC_ROWS ? (uint32_t)(tot - b0) : ENC_ROWS;
    vcg_grow((void **)&S.lv, &S.lv_cap, maxb, (size_t)VCG_A3 * 2);
    vcg_grow((void **)&S.items, &S.items_cap, maxb, sizeof(enc_item));
    vcg_grow((void **)&S.pay, &S.pay_cap, ((size_t)maxb + 1) * ENC_SLOT, 1);   // + k_rc_encode's spare slot
    vcg_grow((void **)&S.lens, &S.lens_cap, maxb, 4);
    const uint32_t maxc = maxb > nc ? maxb : (uint32_t)nc;
    vcg_grow((void **)&S.aux, &S.aux_cap, (size_t)maxc * 2, 4);
    uint64_t *ddst = (uint64_t *)S.aux;       // destination address of each payload
    std::vector<uint32_t> hl(maxc);
    std::vector<uint64_t> dst(maxc), add(nv);
    std::vector<uint32_t> vk(maxc);           // volume of each atom of the launch
    std::vector<uint64_t> ik(maxc);           // its index in that volume
    if (side) {
        vcg_grow((void **)&S.clv, &S.clv_cap, nc, (size_t)VCG_A3 * 2);
        vcg_grow((void **)&S.citems, &S.citems_cap, nc, sizeof(enc_item));
        vcg_grow((void **)&S.cpay, &S.cpay_cap, ((size_t)nc + 1) * ENC_SLOT, 1);
        vcg_grow((void **)&S.clens, &S.clens_cap, nc, 4);
        if (!c->sC) {                               // kept with the thread's context
            VCG_CHECK(hipStreamCreateWithFlags(&c->sC, hipStreamNonBlocking));
            VCG_CHECK(hipEventCreateWithFlags(&c->evCup, hipEventDisableTiming));
            VCG_CHECK(hipEventCreateWithFlags(&c->evCdone, hipEventDisableTiming));
        }
        // after the descriptor upload (and the scratch it may have waited on)
        VCG_CHECK(hipEventRecord(c->evCup, c->s));
        VCG_CHECK(hipStreamWaitEvent(c->sC, c->evCup, 0));
        hipLaunchKernelGGL(k_fdct, dim3((uint32_t)nc), dim3(FDCT_TX), LDS_BYTES, c->sC, dsrc, dseg,
                           (uint32_t)(seg.size() - 1), (uint64_t)0, (int)FDCT_QUANT, g_dtab, S.clv, S.citems);
        VCG_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_rc_encode, dim3((uint32_t)((nc + 63) / 64)), dim3(64), 0, c->sC, (uint32_t)nc, S.clv,
                           S.citems, g_dtab, S.cpay, (uint32_t)ENC_SLOT, S.clens);
        VCG_CHECK(hipGetLastError());
        VCG_CHECK(hipEventRecord(c->evCdone, c->sC));
    }
    // payload lengths of flat atoms [b, b + n) (at dlens, payloads at dpay) to
    // the volumes' pools
    auto place = [&](uint64_t b, uint32_t n, const uint32_t *dlens, const uint8_t *dpay) {
        VCG_CHECK(hipMemcpyAsync(hl.data(), dlens, (size_t)n * 4, hipMemcpyDeviceToHost, c->s));
        VCG_CHECK(hipStreamSynchronize(c->s));
        size_t sg = 0;
        while (b >= seg[sg + 1].base) ++sg;
        for (uint32_t k = 0; k < n; ++k) {
            while (b + k >= seg[sg + 1].base) ++sg;
            const uint32_t v = seg[sg].v;
            const uint64_t i = seg[sg].gi0 + (b + k - seg[sg].base);
            vk[k] = v; ik[k] = i;
            uint32_t L = hl[k];
            if (L == 0xFFFFFFFFu) { ovf[v] = 1; L = 0; }
            lens[v][i] = L;
            es[v]->pool_off[i] = es[v]->pool_len + add[v];
            add[v] += L;
        }
        for (int v = 0; v < nv; ++v)
            if (add[v]) { pool_reserve(c, es[v], es[v]->pool_len + add[v]); }
        for (uint32_t k = 0; k < n; ++k)
            dst[k] = (uint64_t)(uintptr_t)(es[vk[k]]->pool + es[vk[k]]->pool_off[ik[k]]);
        VCG_CHECK(hipMemcpyAsync(ddst, dst.data(), (size_t)n * 8, hipMemcpyHostToDevice, c->s));
        hipLaunchKernelGGL(k_compact, dim3(n), dim3(256), 0, c->s, dpay, (uint32_t)ENC_SLOT, dlens,
                           (const uint64_t *)ddst);
        VCG_CHECK(hipGetLastError());
        for (int v = 0; v < nv; ++v) { es[v]->pool_len += add[v]; add[v] = 0; }
        VCG_CHECK(hipStreamSynchronize(c->s));
    };
    for (uint64_t b = b0; b < tot; b += maxb) {
        const uint32_t n = (uint32_t)(tot - b < maxb ? tot - b : maxb);
        hipLaunchKernelGGL(k_fdct, dim3(n), dim3(FDCT_TX), LDS_BYTES, c->s, dsrc, dseg,
                           (uint32_t)(seg.size() - 1), b, (int)FDCT_QUANT, g_dtab, S.lv, S.items);
        VCG_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_rc_encode, dim3((n + 63) / 64), dim3(64), 0, c->s, n, S.lv, S.items, g_dtab, S.pay,
                           (uint32_t)ENC_SLOT, S.lens);
        VCG_CHECK(hipGetLastError());
        place(b, n, S.lens, S.pay);
    }
    if (side) {
        VCG_CHECK(hipStreamWaitEvent(c->s, c->evCdone, 0));
        place(0, (uint32_t)nc, S.clens, S.cpay);
    }
    return 0;
}

extern "C" int vcg_enc_assemble(vcg_enc *e, const uint8_t *head, size_t head_len, const uint64_t *pay_off,
                                const uint64_t *slot_off, const uint32_t *lens, uint64_t total,
                                uint8_t *out, uint64_t total_len) {
    vcg_ctx *c = vcg_get_ctx();
    enc_scratch &S = tl_es;
    // per-atom arrays in the descriptor area, the image in its own scratch
    const size_t o1 = total * 8, o2 = 2 * o1, o3 = 3 * o1;
    uint8_t *d = desc_area(3 * o1 + total * 4);
    uint64_t *d_pool_off = (uint64_t *)d, *d_pay_off = (uint64_t *)(d + o1), *d_slot_off = (uint64_t *)(d + o2);
    uint32_t *d_lens = (uint32_t *)(d + o3);
    vcg_grow((void **)&S.img, &S.img_cap, total_len, 1);
    uint8_t *img = S.img;
    hipLaunchKernelGGL(k_reset_words, dim3(1024), dim3(256), 0, c->s, (uint32_t *)img, (total_len + 3) / 4);
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
    // to the caller's (pageable) buffer through two pinned halves: the copy
    // of one half out overlaps the transfer into the other
    if (!S.hstage[0]) {
        if (hipHostMalloc((void **)&S.hstage[0], 2 * (size_t)HSTAGE_BYTES, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            vc__panic("out of pinned host memory (encode)");
        }
        S.hstage[1] = S.hstage[0] + HSTAGE_BYTES;
        for (int k = 0; k < 2; ++k)
            if (!S.hev[k]) VCG_CHECK(hipEventCreateWithFlags(&S.hev[k], hipEventDisableTiming));
    }
    const uint64_t nch = (total_len + HSTAGE_BYTES - 1) / HSTAGE_BYTES;
    auto chunk = [&](uint64_t k) {
        const uint64_t off = k * HSTAGE_BYTES, n = total_len - off < HSTAGE_BYTES ? total_len - off : HSTAGE_BYTES;
        VCG_CHECK(hipMemcpyAsync(S.hstage[k & 1], img + off, n, hipMemcpyDeviceToHost, c->s));
        VCG_CHECK(hipEventRecord(S.hev[k & 1], c->s));
    };
    chunk(0);
    for (uint64_t k = 0; k < nch; ++k) {
        if (k + 1 < nch) chunk(k + 1);
        VCG_CHECK(hipEventSynchronize(S.hev[k & 1]));
        const uint64_t off = k * HSTAGE_BYTES, n = total_len - off < HSTAGE_BYTES ? total_len - off : HSTAGE_BYTES;
        par_copy(out + off, S.hstage[k & 1], n);
    }
    return 0;
}

extern "C" int vcg_enc_assemble_many(int n, const vcg_asm_job *J) {
    vcg_ctx *c = vcg_get_ctx();
    enc_scratch &S = tl_es;
    if (n <= 0) return 0;
    // the image: archive k at base[k] (256-byte aligned); the header of each
    // is written on the host, the rest comes back from the image
    std::vector<uint64_t> base(n + 1);
    uint64_t A = 0;
    for (int k = 0; k < n; ++k) {
        base[k + 1] = base[k] + ((J[k].total_len + 255) & ~255ull);
        A += J[k].total_atoms;
    }
    const uint64_t isz = base[n];
    // per-atom arrays in one upload: src, dst, sdst, poff (u64), len (u32)
    std::vector<uint8_t> h((size_t)A * 36);
    uint64_t *hs = (uint64_t *)h.data(), *hd = hs + A, *hsd = hd + A, *hp = hsd + A;
    uint32_t *hl = (uint32_t *)(hp + A);
    for (int k = 0, a = 0; k < n; ++k) {
        const vcg_asm_job &j = J[k];
        for (uint64_t i = 0; i < j.total_atoms; ++i, ++a) {
            hs[a] = (uint64_t)(uintptr_t)(j.e->pool + j.e->pool_off[i]);
            hd[a] = base[k] + j.pay_off[i];
            hsd[a] = j.slot_off[i] ? base[k] + j.slot_off[i] : ~0ull;
            hp[a] = j.pay_off[i];
            hl[a] = j.lens[i];
        }
    }
    uint8_t *d = desc_area(h.size());
    VCG_CHECK(hipMemcpyAsync(d, h.data(), h.size(), hipMemcpyHostToDevice, c->s));
    vcg_grow((void **)&S.img, &S.img_cap, isz, 1);
    uint8_t *img = S.img;
    hipLaunchKernelGGL(k_reset_words, dim3(4096), dim3(256), 0, c->s, (uint32_t *)img, isz / 4);
    const uint64_t *ds = (const uint64_t *)d;
    for (uint64_t b = 0; b < A; b += 1u << 20) {
        const uint64_t m = A - b < (1u << 20) ? A - b : (1u << 20);
        hipLaunchKernelGGL(k_assemble_many, dim3((uint32_t)m), dim3(256), 0, c->s, ds + b, (const uint32_t *)(ds + 4 * A) + b,
                           ds + A + b, ds + 2 * A + b, ds + 3 * A + b, img);
        VCG_CHECK(hipGetLastError());
    }
    for (int k = 0; k < n; ++k) memcpy(J[k].out, J[k].head, J[k].head_len);
    // back through two pinned halves: the copies out of one half overlap the
    // transfer into the other
    if (!S.hstage[0]) {
        if (hipHostMalloc((void **)&S.hstage[0], 2 * (size_t)HSTAGE_BYTES, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            vc__panic("out of pinned host memory (encode)");
        }
        S.hstage[1] = S.hstage[0] + HSTAGE_BYTES;
        for (int k = 0; k < 2; ++k)
            if (!S.hev[k]) VCG_CHECK(hipEventCreateWithFlags(&S.hev[k], hipEventDisableTiming));
    }
    const uint64_t nch = (isz + HSTAGE_BYTES - 1) / HSTAGE_BYTES;
    auto chunk = [&](uint64_t q) {
        const uint64_t off = q * HSTAGE_BYTES, len = isz - off < HSTAGE_BYTES ? isz - off : HSTAGE_BYTES;
        VCG_CHECK(hipMemcpyAsync(S.hstage[q & 1], img + off, len, hipMemcpyDeviceToHost, c->s));
        VCG_CHECK(hipEventRecord(S.hev[q & 1], c->s));
    };
    chunk(0);
    int k0 = 0;                                       // first archive not yet done
    for (uint64_t q = 0; q < nch; ++q) {
        if (q + 1 < nch) chunk(q + 1);
        VCG_CHECK(hipEventSynchronize(S.hev[q & 1]));
        const uint64_t c0 = q * HSTAGE_BYTES, c1 = c0 + (isz - c0 < HSTAGE_BYTES ? isz - c0 : HSTAGE_BYTES);
        for (int k = k0; k < n && base[k] < c1; ++k) {
            // archive k's bytes past its header, within this chunk
            const uint64_t a0 = base[k] + J[k].head_len, a1 = base[k] + J[k].total_len;
            const uint64_t lo = a0 > c0 ? a0 : c0, hi = a1 < c1 ? a1 : c1;
            if (lo < hi) par_copy(J[k].out + (lo - base[k]), S.hstage[q & 1] + (lo - c0), (size_t)(hi - lo));
            if (a1 <= c1) k0 = k + 1;
        }
    }
    return 0;
}

extern "C" void vcg_enc_end(vcg_enc *e) {
    if (!e) return;
    vcg_ctx *c = vcg_get_ctx();
    VCG_CHECK(hipStreamSynchronize(c->s));
    for (int l = 0; l < VCG_NLOD; ++l) enc_free(e->arena, e->lod_buf[l]);
    enc_free(e->arena, e->sample_list);
    enc_free(e->arena, e->sample_coef);
    enc_free(e->arena, e->sample_items);
    enc_free(e->arena, e->pool);
    delete e;
}

extern "C" int vcg_encode_atoms(size_t n, const uint8_t *vox, float q, uint8_t *out, size_t cap, uint32_t *lens) {
    vcg_ctx *c = vcg_get_ctx();
    enc_setup();
    if (n == 0) return 0;
    // the atoms as one LOD0 volume of 32 x 32 x 32n voxels (atom k = z slab k)
    uint8_t *d = (uint8_t *)dmalloc(n * (size_t)VCG_A3);
    VCG_CHECK(hipMemcpyAsync(d, vox, n * (size_t)VCG_A3, hipMemcpyHostToDevice, c->s));
    vcg_enc e;
    memset(&e.src, 0, sizeof e.src);
    e.src.lod[0] = d; e.src.nx[0] = 32; e.src.ny[0] = 32; e.src.nz[0] = (uint32_t)(32 * n);
    e.src.acx[0] = 1; e.src.acy[0] = 1; e.src.prefix[0] = 0; e.src.prefix[1] = n; e.src.nlod = 1;
    e.total = n;
    vcg_enc *ep = &e;
    float qs[1][VCG_NLOD] = {{q}};
    std::vector<uint32_t> L(n);
    uint32_t *lp = L.data();
    int ovf = 0;
    vcg_enc_all(1, &ep, qs, &lp, &ovf);
    int rc = ovf;
    std::vector<uint8_t> pool(e.pool_len);
    if (e.pool_len) VCG_CHECK(hipMemcpyAsync(pool.data(), e.pool, e.pool_len, hipMemcpyDeviceToHost, c->s));
    VCG_CHECK(hipStreamSynchronize(c->s));
    for (size_t k = 0; k < n; ++k) {
        if (L[k] > cap) { rc = 1; lens[k] = 0; continue; }
        lens[k] = L[k];
        if (L[k]) memcpy(out + k * cap, pool.data() + e.pool_off[k], L[k]);
    }
    if (e.pool) VCG_CHECK(hipFree(e.pool));
    VCG_CHECK(hipFree(d));
    return rc;
}
