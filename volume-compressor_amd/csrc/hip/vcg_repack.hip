// vcg_repack.hip — rewrite an archive into the canonical v2 container on the
// device (tools/vc_repack.c).  The archive bytes are already in HBM; the host
// resolves each region's slot array (an L1 probe or a directory read) and does
// the small prefix sums in between.
//   k_repack_size  : one workgroup per 1024-slot plane of a region's slot
//                    array: payload bytes, present and zero atoms, and the
//                    payload-range check (tools/vc_repack.c:96-121).
//   k_repack_slots : one workgroup per plane of a data region: its v2 slots,
//                    output offsets assigned in slot order from the plane's
//                    base (tools/vc_repack.c:131-162).
//   k_repack_copy  : a wave per present atom at a time: its payload, verbatim.
#include <hip/hip_runtime.h>
#include <string.h>
#include "vcg_dev.hpp"

const uint8_t *vcg_arc_buf(const vcg_arc *a);
const vcg_geom *vcg_arc_geom_host(const vcg_arc *a);

#define RP_THREADS 1024                    // one slot per thread: a chunk is one z plane of a region
#define RP_CHUNKS 32u                      // chunks per region
#define RP_L2 ((uint64_t)32768u * 16u)     // slot array bytes

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

// workgroup-wide exclusive scan of (bytes, count) pairs in thread order; the
// totals land in sb[RP_THREADS - 1], sc[RP_THREADS - 1]
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

// chunk b (region b / 32, plane b % 32) of the slot arrays at blk[]:
// sums[3b .. 3b+2] = payload bytes, present atoms, zero atoms; *bad |= 1 for
// a present payload range (or its first byte) outside the archive
__global__ __launch_bounds__(RP_THREADS) void k_repack_size(const uint8_t *__restrict__ arc, uint64_t arc_len,
                                                           const uint64_t *__restrict__ blk,
                                                           uint64_t *__restrict__ sums, uint32_t *bad) {
    __shared__ uint64_t sb[RP_THREADS];
    __shared__ uint32_t sc[RP_THREADS];
    __shared__ uint32_t sz;
    const uint32_t b = blockIdx.x, r = b / RP_CHUNKS, ch = b % RP_CHUNKS;
    if (threadIdx.x == 0) sz = 0;
    uint64_t off; uint32_t len, fl;
    rp_read_slot(arc + blk[r] + ((uint64_t)ch * RP_THREADS + threadIdx.x) * 16u, off, len, fl);
    uint64_t bytes = 0;
    uint32_t np = 0;
    if (fl == 1u) {
        if (off >= arc_len || len > arc_len - off) atomicOr(bad, 1u);
        bytes = len; np = 1;
    }
    __syncthreads();
    if (fl == 2u) atomicAdd(&sz, 1u);
    rp_scan(bytes, np, sb, sc);
    if (threadIdx.x == RP_THREADS - 1) {
        sums[3 * (uint64_t)b] = sb[RP_THREADS - 1];
        sums[3 * (uint64_t)b + 1] = sc[RP_THREADS - 1];
        sums[3 * (uint64_t)b + 2] = sz;
    }
}

// chunk b of data region b / 32: its 1024 v2 slots at out + boff[r] + plane
// offset, its payloads placed from cdst[b] on in slot order, copy jobs
// numbered from cjob[b]
__global__ __launch_bounds__(RP_THREADS) void k_repack_slots(const uint8_t *__restrict__ arc,
                                                            const uint64_t *__restrict__ blk,
                                                            const uint64_t *__restrict__ boff,
                                                            const uint64_t *__restrict__ cdst,
                                                            const uint64_t *__restrict__ cjob,
                                                            uint8_t *__restrict__ out, rp_job *__restrict__ jobs) {
    __shared__ uint64_t sb[RP_THREADS];
    __shared__ uint32_t sc[RP_THREADS];
    const uint32_t b = blockIdx.x, r = b / RP_CHUNKS, ch = b % RP_CHUNKS;
    const uint64_t s = (uint64_t)ch * RP_THREADS + threadIdx.x;
    uint64_t off; uint32_t len, fl;
    rp_read_slot(arc + blk[r] + s * 16u, off, len, fl);
    uint64_t bytes = fl == 1u ? len : 0;
    uint32_t np = fl == 1u ? 1u : 0u;
    rp_scan(bytes, np, sb, sc);
    uint8_t *o = out + boff[r] + s * 16u;
    if (fl == 1u) {
        const uint64_t cur = cdst[b] + bytes, j = cjob[b] + np;
        rp_write_slot(o, cur, len, 1u | (uint32_t)arc[off] << 8);   // dc = first payload byte
        jobs[j].src = off; jobs[j].dst = cur; jobs[j].len = len; jobs[j].pad = 0;
    } else {
        rp_write_slot(o, 0, 0, fl == 2u ? 2u : 0u);
    }
}

// one wave per payload (a payload is a few KiB at most)
#define RP_COPY_WAVES 4
#define RP_COPY_GRID 4096u                 // 16 workgroups per CU
__device__ __forceinline__ void rp_copy_job(const uint8_t *__restrict__ arc, const rp_job jb, uint32_t lane,
                                            uint8_t *__restrict__ out) {
    const uint8_t *s = arc + jb.src;
    uint8_t *d = out + jb.dst;
    if ((((uintptr_t)s ^ (uintptr_t)d) & 15u) == 0) {
        // 16-byte pieces where source and destination share their alignment
        uint32_t head = (uint32_t)((16u - ((uintptr_t)d & 15u)) & 15u);
        if (head > jb.len) head = jb.len;
        for (uint32_t b = lane; b < head; b += 64u) d[b] = s[b];
        const uint32_t nq = (jb.len - head) >> 4;
        const uint4 *sq = (const uint4 *)(s + head);
        uint4 *dq = (uint4 *)(d + head);
        for (uint32_t q = lane; q < nq; q += 64u) dq[q] = sq[q];
        for (uint32_t b = head + 16u * nq + lane; b < jb.len; b += 64u) d[b] = s[b];
        return;
    }
    // otherwise aligned destination words, each from the two aligned source
    // words it straddles (the archive buffer is padded past its end, so the
    // second word of the last piece stays inside the allocation)
    uint32_t head = (uint32_t)((4u - ((uintptr_t)d & 3u)) & 3u);
    if (head > jb.len) head = jb.len;
    for (uint32_t b = lane; b < head; b += 64u) d[b] = s[b];
    const uint8_t *sp = s + head;
    const uint32_t r = (uint32_t)((uintptr_t)sp & 3u);
    const uint32_t *sw = (const uint32_t *)(sp - r);
    uint32_t *dw = (uint32_t *)(d + head);
    const uint32_t nw = (jb.len - head) >> 2;
    for (uint32_t k = lane; k < nw; k += 64u) {
        const uint32_t lo = sw[k];
        dw[k] = r ? __builtin_amdgcn_alignbyte(sw[k + 1], lo, r) : lo;
    }
    for (uint32_t b = head + 4u * nw + lane; b < jb.len; b += 64u) d[b] = s[b];
}

// waves stride over the jobs: a grid of a few thousand workgroups instead of
// one short-lived wave per payload
__global__ __launch_bounds__(64 * RP_COPY_WAVES) void k_repack_copy(const uint8_t *__restrict__ arc,
                                                                    const rp_job *__restrict__ jobs,
                                                                    uint64_t njobs, uint8_t *__restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * RP_COPY_WAVES;
    const uint32_t lane = threadIdx.x & 63u;
    uint64_t j = (uint64_t)blockIdx.x * RP_COPY_WAVES + (threadIdx.x >> 6);
    if (j >= njobs) return;
    rp_job jb = jobs[j];
    for (;;) {
        // the next job's record is in flight while this payload is copied
        const uint64_t jn = j + stride;
        rp_job nx = jb;
        if (jn < njobs) nx = jobs[jn];
        rp_copy_job(arc, jb, lane, out);
        if (jn >= njobs) break;
        jb = nx;
        j = jn;
    }
}

// ---------------------------------------------------------------------------
// launchers (on the calling thread's stream; both synchronous)
// ---------------------------------------------------------------------------
extern "C" int vcg_repack_size(vcg_arc *a, size_t nblk, const uint64_t *blk, uint64_t *sums) {
    vcg_ctx *c = vcg_get_ctx();
    if (nblk == 0) return 0;
    const size_t nch = nblk * RP_CHUNKS;
    uint64_t *dblk = nullptr, *dsum = nullptr;
    uint32_t *dbad = nullptr;
    if (hipMalloc(&dblk, nblk * 8) != hipSuccess || hipMalloc(&dsum, nch * 24) != hipSuccess ||
        hipMalloc(&dbad, 4) != hipSuccess) {
        (void)hipGetLastError();
        if (dblk) (void)hipFree(dblk);
        if (dsum) (void)hipFree(dsum);
        return -1;
    }
    uint32_t bad = 0;
    VCG_CHECK(hipMemcpyAsync(dblk, blk, nblk * 8, hipMemcpyHostToDevice, c->s));
    VCG_CHECK(hipMemcpyAsync(dbad, &bad, 4, hipMemcpyHostToDevice, c->s));
    for (size_t b = 0; b < nch; b += 0x7FFFFFE0u) {   // grid x limit, whole regions per launch
        const size_t n = nch - b < 0x7FFFFFE0u ? nch - b : 0x7FFFFFE0u;
        hipLaunchKernelGGL(k_repack_size, dim3((uint32_t)n), dim3(RP_THREADS), 0, c->s, vcg_arc_buf(a),
                           (uint64_t)vcg_arc_geom_host(a)->len, dblk + b / RP_CHUNKS, dsum + 3 * b, dbad);
        VCG_CHECK(hipGetLastError());
    }
    VCG_CHECK(hipMemcpyAsync(sums, dsum, nch * 24, hipMemcpyDeviceToHost, c->s));
    VCG_CHECK(hipMemcpyAsync(&bad, dbad, 4, hipMemcpyDeviceToHost, c->s));
    VCG_CHECK(hipStreamSynchronize(c->s));
    VCG_CHECK(hipFree(dblk)); VCG_CHECK(hipFree(dsum)); VCG_CHECK(hipFree(dbad));
    return bad ? 1 : 0;
}

extern "C" int vcg_repack_write(vcg_arc *a, size_t ndata, const uint64_t *blk, const uint64_t *boff,
                                const uint64_t *cdst, const uint64_t *cjob, uint64_t njobs, uint8_t *dev_out) {
    vcg_ctx *c = vcg_get_ctx();
    if (ndata == 0) return 0;
    const size_t nch = ndata * RP_CHUNKS;
    uint64_t *d4 = nullptr;
    rp_job *djobs = nullptr;
    if (hipMalloc(&d4, (2 * ndata + 2 * nch) * 8) != hipSuccess ||
        hipMalloc(&djobs, (njobs ? njobs : 1) * sizeof(rp_job)) != hipSuccess) {
        (void)hipGetLastError();
        if (d4) (void)hipFree(d4);
        return -1;
    }
    uint64_t *dblk = d4, *dboff = d4 + ndata, *dcdst = d4 + 2 * ndata, *dcjob = d4 + 2 * ndata + nch;
    VCG_CHECK(hipMemcpyAsync(dblk, blk, ndata * 8, hipMemcpyHostToDevice, c->s));
    VCG_CHECK(hipMemcpyAsync(dboff, boff, ndata * 8, hipMemcpyHostToDevice, c->s));
    VCG_CHECK(hipMemcpyAsync(dcdst, cdst, nch * 8, hipMemcpyHostToDevice, c->s));
    VCG_CHECK(hipMemcpyAsync(dcjob, cjob, nch * 8, hipMemcpyHostToDevice, c->s));
    const uint8_t *arc = vcg_arc_buf(a);
    for (size_t b = 0; b < nch; b += 0x7FFFFFE0u) {
        const size_t n = nch - b < 0x7FFFFFE0u ? nch - b : 0x7FFFFFE0u;
        hipLaunchKernelGGL(k_repack_slots, dim3((uint32_t)n), dim3(RP_THREADS), 0, c->s, arc, dblk + b / RP_CHUNKS,
                           dboff + b / RP_CHUNKS, dcdst + b, dcjob + b, dev_out, djobs);
        VCG_CHECK(hipGetLastError());
    }
    if (njobs) {
        const uint64_t want = (njobs + RP_COPY_WAVES - 1) / RP_COPY_WAVES;
        hipLaunchKernelGGL(k_repack_copy, dim3((uint32_t)(want < RP_COPY_GRID ? want : RP_COPY_GRID)),
                           dim3(64 * RP_COPY_WAVES), 0, c->s, arc, djobs, njobs, dev_out);
        VCG_CHECK(hipGetLastError());
    }
    VCG_CHECK(hipStreamSynchronize(c->s));
    VCG_CHECK(hipFree(d4)); VCG_CHECK(hipFree(djobs));
    return 0;
}
