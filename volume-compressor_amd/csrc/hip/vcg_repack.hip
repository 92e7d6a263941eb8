// vcg_repack.hip — rewrite an archive into the canonical v2 container on the
// device (tools/vc_repack.c).  The archive bytes are already in HBM; the host
// resolves each region's slot array (an L1 probe or a directory read) and does
// the small prefix sums in between.
//   k_repack_size  : one workgroup per region slot array: payload bytes,
//                    present and zero atoms, and the payload-range check
//                    (tools/vc_repack.c:96-121).
//   k_repack_slots : one workgroup per data region: the v2 slot array, output
//                    offsets assigned in slot order (tools/vc_repack.c:131-162).
//   k_repack_copy  : one workgroup per present atom: its payload, verbatim.
#include <hip/hip_runtime.h>
#include <string.h>
#include "vcg_dev.hpp"

const uint8_t *vcg_arc_buf(const vcg_arc *a);
const vcg_geom *vcg_arc_geom_host(const vcg_arc *a);

#define RP_THREADS 1024
#define RP_SLOTS 32768u
#define RP_PER (RP_SLOTS / RP_THREADS)   // slots per thread, consecutive
#define RP_L2 ((uint64_t)RP_SLOTS * 16u)

struct rp_job { uint64_t src, dst; uint32_t len, pad; };

// one 16-byte slot [u64 off][u32 len][u8 flags][u8 dc][u16] at any alignment
__device__ __forceinline__ void rp_read_slot(const uint8_t *p, uint64_t &off, uint32_t &len, uint32_t &flags) {
    if (((uintptr_t)p & 15u) == 0) {
        const uint4 v = *(const uint4 *)p;
        off = (uint64_t)v.x | (uint64_t)v.y << 32; len = v.z; flags = v.w & 0xFFu;
        return;
    }
    uint8_t b[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) b[k] = p[k];
    memcpy(&off, b, 8); memcpy(&len, b + 8, 4); flags = b[12];
}
__device__ __forceinline__ void rp_write_slot(uint8_t *p, uint64_t off, uint32_t len, uint32_t w3) {
    if (((uintptr_t)p & 15u) == 0) {
        *(uint4 *)p = make_uint4((uint32_t)off, (uint32_t)(off >> 32), len, w3);
        return;
    }
    uint8_t b[16];
    memcpy(b, &off, 8); memcpy(b + 8, &len, 4); memcpy(b + 12, &w3, 4);
#pragma unroll
    for (int k = 0; k < 16; ++k) p[k] = b[k];
}

// workgroup-wide exclusive scan of (bytes, count) pairs, thread order
__device__ __forceinline__ void rp_scan(uint64_t &bytes, uint32_t &cnt, uint64_t *sb, uint32_t *sc) {
    const int t = threadIdx.x;
    sb[t] = bytes; sc[t] = cnt;
    __syncthreads();
    for (int o = 1; o < RP_THREADS; o <<= 1) {
        const uint64_t b = t >= o ? sb[t - o] : 0;
        const uint32_t c = t >= o ? sc[t - o] : 0;
        __syncthreads();
        sb[t] += b; sc[t] += c;
        __syncthreads();
    }
    bytes = sb[t] - bytes; cnt = sc[t] - cnt;   // inclusive -> exclusive
}

// per slot array b at blk[b]: sums[3b..3b+2] = payload bytes, present atoms,
// zero atoms; *bad |= 1 for a present payload range outside the archive
__global__ __launch_bounds__(RP_THREADS) void k_repack_size(const uint8_t *__restrict__ arc, uint64_t arc_len,
                                                           const uint64_t *__restrict__ blk,
                                                           uint64_t *__restrict__ sums, uint32_t *bad) {
    __shared__ uint64_t sb[RP_THREADS];
    __shared__ uint32_t sc[RP_THREADS];
    __shared__ uint32_t sz;
    const uint8_t *slots = arc + blk[blockIdx.x];
    if (threadIdx.x == 0) sz = 0;
    uint64_t bytes = 0;
    uint32_t np = 0, nz = 0, oob = 0;
    for (uint32_t k = 0; k < RP_PER; ++k) {
        uint64_t off; uint32_t len, fl;
        rp_read_slot(slots + (uint64_t)(threadIdx.x * RP_PER + k) * 16u, off, len, fl);
        if (fl == 1u) {
            // the payload and its first byte (the slot's dc) lie inside the archive
            oob |= off >= arc_len || len > arc_len - off;
            bytes += len; np++;
        } else if (fl == 2u) {
            nz++;
        }
    }
    if (oob) atomicOr(bad, 1u);
    __syncthreads();
    if (nz) atomicAdd(&sz, nz);
    rp_scan(bytes, np, sb, sc);
    if (threadIdx.x == RP_THREADS - 1) {
        sums[3 * blockIdx.x] = sb[RP_THREADS - 1];
        sums[3 * blockIdx.x + 1] = sc[RP_THREADS - 1];
    }
    __syncthreads();
    if (threadIdx.x == 0) sums[3 * blockIdx.x + 2] = sz;
}

// data region d: slot array from blk[d] rewritten at out + boff[d], payloads
// placed after it in slot order; copy jobs from index jbase[d]
__global__ __launch_bounds__(RP_THREADS) void k_repack_slots(const uint8_t *__restrict__ arc,
                                                            const uint64_t *__restrict__ blk,
                                                            const uint64_t *__restrict__ boff,
                                                            const uint64_t *__restrict__ jbase,
                                                            uint8_t *__restrict__ out, rp_job *__restrict__ jobs) {
    __shared__ uint64_t sb[RP_THREADS];
    __shared__ uint32_t sc[RP_THREADS];
    const uint8_t *slots = arc + blk[blockIdx.x];
    const uint32_t s0 = threadIdx.x * RP_PER;
    uint64_t bytes = 0;
    uint32_t np = 0;
    for (uint32_t k = 0; k < RP_PER; ++k) {
        uint64_t off; uint32_t len, fl;
        rp_read_slot(slots + (uint64_t)(s0 + k) * 16u, off, len, fl);
        if (fl == 1u) { bytes += len; np++; }
    }
    rp_scan(bytes, np, sb, sc);
    const uint64_t base = boff[blockIdx.x];
    uint64_t cur = base + RP_L2 + bytes;
    uint64_t j = jbase[blockIdx.x] + np;
    for (uint32_t k = 0; k < RP_PER; ++k) {
        uint64_t off; uint32_t len, fl;
        rp_read_slot(slots + (uint64_t)(s0 + k) * 16u, off, len, fl);
        uint8_t *o = out + base + (uint64_t)(s0 + k) * 16u;
        if (fl == 1u) {
            rp_write_slot(o, cur, len, 1u | (uint32_t)arc[off] << 8);   // dc = first payload byte
            jobs[j].src = off; jobs[j].dst = cur; jobs[j].len = len; jobs[j].pad = 0;
            cur += len; j++;
        } else {
            rp_write_slot(o, 0, 0, fl == 2u ? 2u : 0u);
        }
    }
}

__global__ __launch_bounds__(256) void k_repack_copy(const uint8_t *__restrict__ arc,
                                                     const rp_job *__restrict__ jobs, uint8_t *__restrict__ out) {
    const rp_job jb = jobs[blockIdx.x];
    const uint8_t *s = arc + jb.src;
    uint8_t *d = out + jb.dst;
    // 16-byte pieces where source and destination share their alignment
    uint32_t head = (uint32_t)((16u - ((uintptr_t)d & 15u)) & 15u);
    if ((((uintptr_t)s ^ (uintptr_t)d) & 15u) != 0 || head >= jb.len) head = jb.len;
    for (uint32_t b = threadIdx.x; b < head; b += blockDim.x) d[b] = s[b];
    const uint32_t nq = (jb.len - head) >> 4;
    const uint4 *sq = (const uint4 *)(s + head);
    uint4 *dq = (uint4 *)(d + head);
    for (uint32_t q = threadIdx.x; q < nq; q += blockDim.x) dq[q] = sq[q];
    for (uint32_t b = head + 16u * nq + threadIdx.x; b < jb.len; b += blockDim.x) d[b] = s[b];
}

// ---------------------------------------------------------------------------
// launchers (on the calling thread's stream; both synchronous)
// ---------------------------------------------------------------------------
extern "C" int vcg_repack_size(vcg_arc *a, size_t nblk, const uint64_t *blk, uint64_t *sums) {
    vcg_ctx *c = vcg_get_ctx();
    if (nblk == 0) return 0;
    uint64_t *dblk = nullptr, *dsum = nullptr;
    uint32_t *dbad = nullptr;
    if (hipMalloc(&dblk, nblk * 8) != hipSuccess || hipMalloc(&dsum, nblk * 24) != hipSuccess ||
        hipMalloc(&dbad, 4) != hipSuccess) {
        (void)hipGetLastError();
        if (dblk) (void)hipFree(dblk);
        if (dsum) (void)hipFree(dsum);
        return -1;
    }
    uint32_t bad = 0;
    VCG_CHECK(hipMemcpyAsync(dblk, blk, nblk * 8, hipMemcpyHostToDevice, c->s));
    VCG_CHECK(hipMemcpyAsync(dbad, &bad, 4, hipMemcpyHostToDevice, c->s));
    hipLaunchKernelGGL(k_repack_size, dim3((uint32_t)nblk), dim3(RP_THREADS), 0, c->s, vcg_arc_buf(a),
                       (uint64_t)vcg_arc_geom_host(a)->len, dblk, dsum, dbad);
    VCG_CHECK(hipGetLastError());
    VCG_CHECK(hipMemcpyAsync(sums, dsum, nblk * 24, hipMemcpyDeviceToHost, c->s));
    VCG_CHECK(hipMemcpyAsync(&bad, dbad, 4, hipMemcpyDeviceToHost, c->s));
    VCG_CHECK(hipStreamSynchronize(c->s));
    VCG_CHECK(hipFree(dblk)); VCG_CHECK(hipFree(dsum)); VCG_CHECK(hipFree(dbad));
    return bad ? 1 : 0;
}

extern "C" int vcg_repack_write(vcg_arc *a, size_t ndata, const uint64_t *blk, const uint64_t *boff,
                                const uint64_t *jbase, uint64_t njobs, uint8_t *dev_out) {
    vcg_ctx *c = vcg_get_ctx();
    if (ndata == 0) return 0;
    uint64_t *d3 = nullptr;
    rp_job *djobs = nullptr;
    if (hipMalloc(&d3, ndata * 24) != hipSuccess ||
        hipMalloc(&djobs, (njobs ? njobs : 1) * sizeof(rp_job)) != hipSuccess) {
        (void)hipGetLastError();
        if (d3) (void)hipFree(d3);
        return -1;
    }
    VCG_CHECK(hipMemcpyAsync(d3, blk, ndata * 8, hipMemcpyHostToDevice, c->s));
    VCG_CHECK(hipMemcpyAsync(d3 + ndata, boff, ndata * 8, hipMemcpyHostToDevice, c->s));
    VCG_CHECK(hipMemcpyAsync(d3 + 2 * ndata, jbase, ndata * 8, hipMemcpyHostToDevice, c->s));
    const uint8_t *arc = vcg_arc_buf(a);
    hipLaunchKernelGGL(k_repack_slots, dim3((uint32_t)ndata), dim3(RP_THREADS), 0, c->s, arc, d3, d3 + ndata,
                       d3 + 2 * ndata, dev_out, djobs);
    VCG_CHECK(hipGetLastError());
    for (uint64_t b = 0; b < njobs; b += 0x7FFFFFFFull) {   // grid x limit
        const uint64_t n = njobs - b < 0x7FFFFFFFull ? njobs - b : 0x7FFFFFFFull;
        hipLaunchKernelGGL(k_repack_copy, dim3((uint32_t)n), dim3(256), 0, c->s, arc, djobs + b, dev_out);
        VCG_CHECK(hipGetLastError());
    }
    VCG_CHECK(hipStreamSynchronize(c->s));
    VCG_CHECK(hipFree(d3)); VCG_CHECK(hipFree(djobs));
    return 0;
}
