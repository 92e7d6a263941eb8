// vcg_decode.hip — batched atom decode on gfx950.
//
// Two kernels per batch:
//   k_entropy : one lane per atom.  Resolves the atom through the device copy
//               of the archive's own index (v1 hash / v2 directory + slot),
//               runs the adaptive binary range decoder and context coder
//               (src/vc/vc.c:403-441, 491-560) and writes the nonzero levels
//               as (scan position, level) entries to HBM scratch.
//   k_idct    : one 512-thread workgroup per atom.  Scatters and dequantizes
//               the entries into a zeroed 32x32x(32+1) i32 LDS tile
//               (src/vc/vc.c:327-338), runs the three inverse partial
//               butterfly passes along x, y, z in place (src/vc/vc.c:172-229),
//               adds dc, clamps and stores u8 voxels to the destination.
#include <hip/hip_runtime.h>
#include <string.h>
#include <stdlib.h>
#include <mutex>
#include "vcg_dev.hpp"
#include "vcg_xform.hpp"

const uint8_t *vcg_arc_buf(const vcg_arc *a);
const vcg_geom *vcg_arc_geom_dev(const vcg_arc *a);
const vcg_geom *vcg_arc_geom_host(const vcg_arc *a);

// ---------------------------------------------------------------------------
// job description
// ---------------------------------------------------------------------------
enum { MODE_BOX = 0, MODE_LIST = 1, MODE_PAYLOAD = 2 };

struct dec_job {
    const uint8_t *arc;         // archive bytes (box/list) or payload pool (payload)
    const vcg_geom *geom;       // device geometry (box/list)
    const int32_t *dqs;         // [8][94] (box/list: geom tables; payload: [94])
    const int32_t *dqb;
    int mode, lod;
    int check;                          // list mode: quads come unchecked (device memory)
    uint32_t ax0, ay0, az0, nax, nay;   // box mode atom range
    uint32_t base, n;                   // item range of this launch
    const int32_t *quads;               // list mode: (lod | placement << 8, ax, ay, az)
    const vcg_place *places;            // list mode: box placements (or null)
    const uint64_t *offs;               // payload mode
    const uint32_t *lens;
    // output placement
    uint8_t *out;
    int32_t *status_out;                // optional: per item 1 = corrupt payload, else 0
    uint32_t bx0, by0, bz0, bx1, by1, bz1;   // box mode: voxel box
};

// Returns false for an unchecked list quad outside the archive (lod not
// below nlod or an atom coordinate outside that LOD's atom grid, the range
// vc_decode_atom rejects, src/vc/vc.c:1406-1409); its lod and coordinates
// then read as 0 so that no table is indexed out of range.
__device__ __forceinline__ bool item_atom(const dec_job &j, uint32_t gi, int &lod, uint32_t &ax,
                                          uint32_t &ay, uint32_t &az) {
    if (j.mode == MODE_BOX) {
        lod = j.lod;
        ax = j.ax0 + gi % j.nax;
        ay = j.ay0 + (gi / j.nax) % j.nay;
        az = j.az0 + gi / (j.nax * j.nay);
    } else if (j.mode == MODE_LIST) {
        const int32_t w = j.quads[4 * gi];
        lod = w & 0xFF; ax = (uint32_t)j.quads[4 * gi + 1];
        ay = (uint32_t)j.quads[4 * gi + 2]; az = (uint32_t)j.quads[4 * gi + 3];
        if (j.check && (w < 0 || w >= j.geom->nlod || ax >= j.geom->acx[w] || ay >= j.geom->acy[w] ||
                        az >= j.geom->acz[w])) {
            lod = 0; ax = ay = az = 0;
            return false;
        }
    } else {
        lod = 0; ax = ay = az = 0;
    }
    return true;
}

// per-lane LDS array (context memory: VCG_CTX_SLOTS slots; output ring: VCG_RING slots),
// slot-major over the 64 lanes of the wave so a wave's access is conflict-free.
// m is the workgroup's array (a fixed LDS address) and l4 the lane's byte
// offset in it: the wave's part (a power of two in size) ORed with lane * 4,
// so a slot's address is one shift-or and the array base goes into the
// instruction's offset field.  Context handles are byte offsets; sink is the
// sink slot's, kept opaque to the compiler so that picking a handle stays one
// select after the shift-or.
struct lds_ctx {
    uint32_t *m;
    const uint8_t *nx;           // the workgroup's next-state table (vcg_fsm_next)
    uint32_t l4, sink;
    __device__ __forceinline__ uint32_t at(uint32_t S, bool isCtx) const { return isCtx ? (S << 8) | l4 : sink; }
    __device__ __forceinline__ uint32_t nxt(uint32_t i) const { return nx[i]; }
    __device__ __forceinline__ uint32_t ld_next(uint32_t h) const {
        return *(const uint32_t *)((const char *)m + h + 256u);
    }
    __device__ __forceinline__ void st(uint32_t h, uint32_t v) { *(uint32_t *)((char *)m + h) = v; }
};
// output ring: slot offsets come pre-scaled (slot << 8 = slot * 64 lanes * 4 B)
struct lds_ring {
    uint32_t *m;
    uint32_t l4;
    __device__ __forceinline__ uint32_t ld(uint32_t off) const {
        return *(const uint32_t *)((const char *)m + (off | l4));
    }
    __device__ __forceinline__ void st(uint32_t off, uint32_t v) { *(uint32_t *)((char *)m + (off | l4)) = v; }
};

#ifndef ENT_WAVES
#define ENT_WAVES 4
#endif
#define ENT_THREADS (64 * ENT_WAVES)

// One lane's atom (item i of the launch): resolve it through the device copy
// of the archive's index, range-decode its payload and write its coefficient
// list and item record.  ctxm / ringm: the workgroup's context and output-ring
// LDS arrays, a wave's part at byte (wave << 12) and wave * VCG_RING * 256, so
// a lane's offset ORs into the slot addressing.
__device__ __forceinline__ void ent_atom(const dec_job &j, uint32_t i, const vcg_dtab *__restrict__ T,
                                         uint32_t *__restrict__ lists, vcg_item *__restrict__ items,
                                         uint32_t *__restrict__ status, uint32_t *ctxm, uint32_t *ringm,
                                         const uint8_t *nxt, const uint32_t *bands, uint32_t lane, uint32_t wave) {
    const uint32_t gi = j.base + i;
    int lod = 0;
    uint64_t off = 0;
    uint32_t len = 0;
    int r;
    if (j.mode == MODE_PAYLOAD) {
        off = j.offs[gi]; len = j.lens[gi];
        r = len >= 1 ? 1 : 0;
    } else {
        uint32_t ax, ay, az;
        r = item_atom(j, gi, lod, ax, ay, az) ? vcg_resolve(j.arc, j.geom, lod, az, ay, ax, &off, &len) : 3;
    }
    vcg_item it;
    it.cnt = 0; it.lod = (uint16_t)lod; it.dc = 0; it.fmax = -1;
    if (r != 1) {
        // 0: zeros; 2: corrupt slot (zeros, status word); 3: out of range (zeros)
        it.status = r == 2 ? 2u : r == 3 ? 3u : 0u;
        items[i] = it;
        if (r == 2) atomicOr(status, 1u);
        return;
    }
    const uint8_t *p = j.arc + off;
    it.dc = p[0];
    uint32_t *out = lists + (size_t)i * VCG_LIST_STRIDE;
#if VCG_ENTROPY_FSM
    // range-coder bytes [p+1, p+len); the device buffers are padded past
    // their end for the 16-byte look-ahead loads
    const uint32_t cl4 = (wave << 12) | (lane * 4u);
    uint32_t sink = (14u << 8) | cl4;
    asm volatile("" : "+v"(sink));
    lds_ctx ctx{ctxm, nxt, cl4, sink};
    // opaque, so that the slot offset and the lane offset stay one and-or
    uint32_t rl4 = (wave * (VCG_RING << 8)) | (lane * 4u);
    asm volatile("" : "+v"(rl4));
    lds_ring ring{ringm, rl4};
    for (uint32_t k = 1; k < VCG_CTX_SLOTS; ++k)
        ctxm[(wave * VCG_CTX_SLOTS + k) * 64u + lane] = k == 13u || k == 15u ? VCG_P_BYPASS : 2048u;
    const uint8_t *s = p + 1;
    const uint32_t n = len - 1;
    vcg_fsm f;
    vcg_look A, B;
    vcg_fsm_init(f, A, B, s, n, bands);
    // the edge test is needed in a block only if some lane of the wave can
    // meet an edge in it (a wave-uniform branch)
    for (uint32_t it = 0; __any(!vcg_fsm_parked(f)); ++it) {
        vcg_fsm_refill(f, A, s, n);
        if (it % VCG_FLUSH_EVERY == VCG_FLUSH_EVERY - 1) vcg_fsm_flush(f, ring, out);
        vcg_fsm_block(f, ctx, ring, bands, __any(vcg_fsm_near(f, VCG_REFILL_STEPS)));
        vcg_fsm_refill(f, B, s, n);
        vcg_fsm_block(f, ctx, ring, bands, __any(vcg_fsm_near(f, VCG_REFILL_STEPS)));
    }
    vcg_fsm_flush(f, ring, out);
    const int bad = (int)f.bad;
    const uint32_t cnt = f.fl;
    // the last entry (highest scan position), read back from the list this
    // lane has just written
    const int fmax = cnt ? vcg_freq(T->scan[vcg_ent_pos(out[cnt - 1u])]) : -1;
#else
    (void)ctxm; (void)ringm; (void)nxt; (void)bands; (void)lane; (void)wave;
    uint32_t cnt = 0;
    const int bad = vcg_rc_decode_atom(p + 1, len - 1, T->band_start, out, &cnt);
    const int fmax = cnt && !bad ? vcg_freq(T->scan[vcg_ent_pos(out[cnt - 1])]) : -1;
#endif
    if (bad) {
        it.status = 2; it.cnt = 0;
        atomicOr(status, 1u);
    } else {
        it.status = 1; it.cnt = cnt; it.fmax = (int16_t)fmax;
    }
    items[i] = it;
}

__global__ __launch_bounds__(ENT_THREADS) void k_entropy(dec_job j, const vcg_dtab *__restrict__ T,
                                                uint32_t *__restrict__ lists,
                                                vcg_item *__restrict__ items,
                                                uint32_t *__restrict__ status) {
    // issue priority over transform waves sharing the SIMD (the serial
    // decode is the critical path; the transform fills its idle issue cycles)
    __builtin_amdgcn_s_setprio(3);
    __shared__ uint32_t ctxw[ENT_WAVES][VCG_CTX_SLOTS * 64];
    __shared__ uint32_t ringw[ENT_WAVES][VCG_RING * 64];
    __shared__ uint8_t nxt[512];
    __shared__ uint32_t bands[9];
    // band table in LDS: a band edge then waits on LDS only, not on the
    // look-ahead loads in flight
    if (threadIdx.x < 9) bands[threadIdx.x] = T->band_start[threadIdx.x];
    for (uint32_t t = threadIdx.x; t < 512u; t += ENT_THREADS) nxt[t] = (uint8_t)vcg_fsm_next(t & 255u, t >> 8);
    __syncthreads();
    const uint32_t i = blockIdx.x * (uint32_t)ENT_THREADS + threadIdx.x;
    if (i >= j.n) return;
    ent_atom(j, i, T, lists, items, status, &ctxw[0][0], &ringw[0][0], nxt, bands, threadIdx.x & 63u,
             threadIdx.x >> 6);
}

// ---------------------------------------------------------------------------
// inverse transform: phases from vcg_xform.hpp separated by barriers
// ---------------------------------------------------------------------------
#define IDCT_THREADS 512
#define IDCT_PRE 10          // list entries per thread loaded ahead (5120 per atom)

// where an item's voxels go: whole atom at out (box == false), or clipped to
// a voxel box whose raster starts at out
struct dst_view {
    uint8_t *out;
    bool box;
    uint32_t bx0, by0, bz0, bx1, by1, bz1;
};
__device__ __forceinline__ dst_view item_dst(const dec_job &j, uint32_t gi) {
    dst_view d;
    if (j.mode == MODE_BOX) {
        d.out = j.out; d.box = true;
        d.bx0 = j.bx0; d.by0 = j.by0; d.bz0 = j.bz0; d.bx1 = j.bx1; d.by1 = j.by1; d.bz1 = j.bz1;
        return d;
    }
    const int p = j.places && j.mode == MODE_LIST ? (j.quads[4 * gi] >> 8) - 1 : -1;
    if (p >= 0) {
        const vcg_place &q = j.places[p];
        d.out = j.out + q.out_off; d.box = true;
        d.bx0 = q.x0; d.by0 = q.y0; d.bz0 = q.z0; d.bx1 = q.x1; d.by1 = q.y1; d.bz1 = q.z1;
        return d;
    }
    d.out = j.out + (uint64_t)gi * VCG_A3; d.box = false;
    d.bx0 = d.by0 = d.bz0 = d.bx1 = d.by1 = d.bz1 = 0;
    return d;
}

// Store the staged raster atom: thread tid owns bytes [64*tid, 64*tid + 64) =
// rows (z, y) and (z, y+1).  Rows that lie wholly inside the box at a 16-byte
// aligned destination go out as 16-byte stores, others byte by byte.
__device__ __forceinline__ void store_atom(const dst_view &d, uint32_t ax, uint32_t ay, uint32_t az,
                                           const uint8_t *stage, int tid, uint32_t zs = 1024) {
    const int z = tid >> 4, y0 = (tid & 15) * 2;
    const uint8_t *row0 = stage + z * zs + y0 * 32;
    const uint4 *src = (const uint4 *)row0;
    if (!d.box) {
        uint4 *dst = (uint4 *)(d.out + 64 * tid);
#pragma unroll
        for (int q = 0; q < 4; ++q) dst[q] = src[q];
        return;
    }
    const uint32_t vz = az * 32 + z, vx = ax * 32;
    const uint64_t ox = d.bx1 - d.bx0, oy = d.by1 - d.by0;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const uint32_t vy = ay * 32 + y0 + r;
        if (vz < d.bz0 || vz >= d.bz1 || vy < d.by0 || vy >= d.by1) continue;
        uint8_t *row = d.out + ((uint64_t)(vz - d.bz0) * oy + (vy - d.by0)) * ox;
        if (vx >= d.bx0 && vx + 32 <= d.bx1 && ((uintptr_t)(row + (vx - d.bx0)) & 15) == 0) {
            uint4 *dst = (uint4 *)(row + (vx - d.bx0));
            dst[0] = src[2 * r];
            dst[1] = src[2 * r + 1];
        } else {
            const uint8_t *b = row0 + 32 * r;
            for (int x = 0; x < 32; ++x)
                if (vx + x >= d.bx0 && vx + x < d.bx1) row[vx + x - d.bx0] = b[x];
        }
    }
}

// Workgroups are dealt to the 8 XCDs round robin (workgroup b to XCD b % 8),
// and each XCD has its own L2.  Item i of a launch goes to workgroup b such
// that every XCD takes a contiguous run of items, in order: atoms that are
// neighbours along x then run side by side on one XCD, and their 32-byte row
// pieces of a raster output meet in that XCD's L2 as whole lines.  A
// bijection on [0, n).
__device__ __forceinline__ uint32_t xcd_item(uint32_t b, uint32_t n) {
    const uint32_t x = b & 7u, k = b >> 3, q = n >> 3, r = n & 7u;
    return x * q + (x < r ? x : r) + k;
}

// One workgroup's item i of the launch: scatter and dequantize its list into
// the LDS tile buf, run the three inverse passes and store the voxels.  Every
// thread of the workgroup calls it with the same i (it holds barriers).
__device__ __forceinline__ void idct_item(const dec_job &j, uint32_t i, const vcg_dtab *__restrict__ T,
                                          const uint32_t *__restrict__ lists, const vcg_item *__restrict__ items,
                                          int *buf, unsigned long long *l1sum, int32_t *dqs) {
    const int tid = threadIdx.x;
    const uint32_t gi = j.base + i;
    int lod;
    uint32_t ax, ay, az;
    item_atom(j, gi, lod, ax, ay, az);
    // loads that do not wait on each other go out together: the item, the
    // dequant tables of the atom's LOD (known from the job, not the item) and
    // the list's first IDCT_PRE entries per thread (a list's allocation spans
    // VCG_LIST_STRIDE entries whatever its count), so the chain of memory
    // round trips before the scatter is two (item and list, then scan table)
    // instead of four.  The tile and the tables stay within 133 KiB of LDS, so
    // a transform workgroup fits on a CU beside four entropy workgroups.
    const uint32_t *L = lists + (size_t)i * VCG_LIST_STRIDE;
    uint32_t pre[IDCT_PRE];
#pragma unroll
    for (int q = 0; q < IDCT_PRE; ++q) pre[q] = L[tid + q * IDCT_THREADS];
    if (tid < VCG_NFREQ) {
        const int o = j.mode == MODE_PAYLOAD ? 0 : lod * VCG_NFREQ;
        dqs[tid] = j.dqs[o + tid];
        dqs[VCG_NFREQ + tid] = j.dqb[o + tid];
    }
    const vcg_item it = items[i];
    uint8_t *stage = (uint8_t *)buf;
    if (tid == 0 && j.status_out) j.status_out[gi] = it.status == 2 ? 1 : it.status == 3 ? 2 : 0;
    // status 4: an atom the matrix-core transform left to this full-range path
    if (it.status != 1 && it.status != 4) {
        for (int v = tid; v < (int)VCG_A3 / 4; v += IDCT_THREADS) ((uint32_t *)stage)[v] = 0;
        __syncthreads();
        store_atom(item_dst(j, gi), ax, ay, az, stage, tid);
        return;
    }
    const uint32_t cnt = it.cnt;
    const int fmax = it.fmax;
    uint32_t sidx[IDCT_PRE];
#pragma unroll
    for (int q = 0; q < IDCT_PRE; ++q)
        sidx[q] = tid + q * IDCT_THREADS < (int)cnt ? T->scan[vcg_ent_pos(pre[q])] : 0u;
    vcg_tile_zero(buf, tid, IDCT_THREADS);
    if (tid == 0) *l1sum = 0;
    __syncthreads();
    const int32_t *st = dqs, *bi = dqs + VCG_NFREQ;
    unsigned long long mine = 0;
#pragma unroll
    for (int q = 0; q < IDCT_PRE; ++q)
        if (tid + q * IDCT_THREADS < (int)cnt) mine += vcg_scatter1(buf, pre[q], sidx[q], st, bi);
    if (cnt > IDCT_PRE * IDCT_THREADS)
        mine += vcg_scatter(buf, L + IDCT_PRE * IDCT_THREADS, cnt - IDCT_PRE * IDCT_THREADS, T->scan, st, bi, tid,
                            IDCT_THREADS);
    if (mine) atomicAdd(l1sum, mine);
    __syncthreads();
    // Every intermediate element is bounded by the atom's L2 norm (each 1-D
    // pass is orthonormal up to Q14 rounding, sigma_max <= 1.001, plus <= 0.5
    // rounding per element), and L2 <= L1.  With L1 <= 8e6 all MAC operands
    // stay below 8.02e6 < 2^23, so the 24-bit multiply path is exact.
    const bool fast = *l1sum <= 8000000ull;
    const int32_t (*C)[32] = T->cmat;
    const int dc = (int)it.dc;
    uint8_t vox[64];
    if (fast) {
        vcg_inv_pass_x<true>(buf, C, fmax, tid, IDCT_THREADS);
        __syncthreads();
        vcg_inv_pass_y<true>(buf, C, fmax, tid, IDCT_THREADS);
        __syncthreads();
        vcg_inv_pass_z<true>(buf, C, fmax, dc, tid, vox);
    } else {
        vcg_inv_pass_x<false>(buf, C, fmax, tid, IDCT_THREADS);
        __syncthreads();
        vcg_inv_pass_y<false>(buf, C, fmax, tid, IDCT_THREADS);
        __syncthreads();
        vcg_inv_pass_z<false>(buf, C, fmax, dc, tid, vox);
    }
    __syncthreads();
    vcg_inv_stage(stage, vox, tid);
    __syncthreads();
    store_atom(item_dst(j, gi), ax, ay, az, stage, tid);
}

__global__ __launch_bounds__(IDCT_THREADS) void k_idct(dec_job j, const vcg_dtab *__restrict__ T,
                                                       const uint32_t *__restrict__ lists,
                                                       const vcg_item *__restrict__ items) {
    extern __shared__ int buf[];
    __shared__ unsigned long long l1sum;
    __shared__ int32_t dqs[2 * VCG_NFREQ];
    idct_item(j, xcd_item(blockIdx.x, gridDim.x), T, lists, items, buf, &l1sum, dqs);
}

// ---------------------------------------------------------------------------
// persistent batch kernels (server batches replayed as hipGraphs)
// ---------------------------------------------------------------------------
// A batch's size and its two claim counters live in device memory (a 16-byte
// header uploaded with the batch's quads), so one captured launch sequence
// serves every batch of up to its capacity.  Entropy waves claim 64 atoms
// (one per lane) at a time until the batch is exhausted; transform
// workgroups claim one atom at a time.  Both loops end once a claim passes n,
// which every wave reaches.
struct vcg_qhdr {
    uint32_t n;        // atoms in the batch
    uint32_t ent;      // next unclaimed atom of the entropy pass
    uint32_t idct;     // next unclaimed atom of the transform
    uint32_t pad;
};

template <int W>
__global__ __launch_bounds__(64 * W) void k_entropy_p(dec_job j, const vcg_dtab *__restrict__ T,
                                                      uint32_t *__restrict__ lists, vcg_item *__restrict__ items,
                                                      uint32_t *__restrict__ status, vcg_qhdr *__restrict__ q) {
    __builtin_amdgcn_s_setprio(3);
    __shared__ uint32_t ctxw[W][VCG_CTX_SLOTS * 64];
    __shared__ uint32_t ringw[W][VCG_RING * 64];
    __shared__ uint8_t nxt[512];
    __shared__ uint32_t bands[9];
    if (threadIdx.x < 9) bands[threadIdx.x] = T->band_start[threadIdx.x];
    for (uint32_t t = threadIdx.x; t < 512u; t += 64u * W) nxt[t] = (uint8_t)vcg_fsm_next(t & 255u, t >> 8);
    __syncthreads();
    const uint32_t n = q->n;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (;;) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&q->ent, 64u);
        base = __shfl(base, 0, 64);
        if (base >= n) break;
        const uint32_t i = base + lane;
        if (i < n) ent_atom(j, i, T, lists, items, status, &ctxw[0][0], &ringw[0][0], nxt, bands, lane, wave);
    }
}

__global__ __launch_bounds__(IDCT_THREADS) void k_idct_p(dec_job j, const vcg_dtab *__restrict__ T,
                                                         const uint32_t *__restrict__ lists,
                                                         const vcg_item *__restrict__ items, vcg_qhdr *__restrict__ q) {
    extern __shared__ int buf[];
    __shared__ unsigned long long l1sum;
    __shared__ int32_t dqs[2 * VCG_NFREQ];
    __shared__ uint32_t cur;
    const uint32_t n = q->n;
    for (;;) {
        __syncthreads();                 // the previous atom is done with the tile and cur
        if (threadIdx.x == 0) cur = atomicAdd(&q->idct, 1u);
        __syncthreads();
        const uint32_t i = cur;
        if (i >= n) break;
        idct_item(j, i, T, lists, items, buf, &l1sum, dqs);
    }
}

// ---------------------------------------------------------------------------
// inverse transform on the matrix cores (vcg_xform.hpp, "inverse on the
// matrix cores"): an i16 tile of 64 KiB, so two transform workgroups share a
// CU, each pass 32 int8 products per atom.  An atom whose coefficients
// exceed the bound is marked status 4 and left to k_idct_ovf (the i32 path),
// launched after it on the same stream.
// ---------------------------------------------------------------------------
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
#define IM_THREADS 512
// registers for four transform waves and one entropy wave per SIMD
#ifndef IM_WAVES
#define IM_WAVES 5
#endif
#define IM_REGS __attribute__((amdgpu_waves_per_eu(IM_WAVES)))
static_assert(IM_THREADS == IDCT_THREADS, "the list prefetch and the store assume 512 threads");

// one block's product chain (data planes hi / lo, basis planes Ch and 4 Cl),
// scaled by 4: o[r] = 4 (sum_k C[k][n] v[k] + rnd), rnd4 = 4 rnd, for output
// register r; the output (sum + rnd) >> 14 is its bits 16..31
__device__ __forceinline__ void im_chain(const uint32_t hi[4], const uint32_t lo[4], v4i bh, v4i b4l, int rnd4,
                                         int o[16]) {
    const v4i ah = {(int)hi[0], (int)hi[1], (int)hi[2], (int)hi[3]};
    const v4i al = {(int)lo[0], (int)lo[1], (int)lo[2], (int)lo[3]};
    const v16i zero = {};
    v16i t = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah, bh, zero, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 16; ++r) t[r] = (int)((uint32_t)t[r] << 8);
    t = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah, b4l, t, 0, 0, 0);
    t = __builtin_amdgcn_mfma_i32_32x32x32_i8(al, bh, t, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 16; ++r) t[r] = (int)(((uint32_t)t[r] << 8) + (uint32_t)rnd4);
    t = __builtin_amdgcn_mfma_i32_32x32x32_i8(al, b4l, t, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 16; ++r) o[r] = t[r];
}

// block B of pass p (x or y): product and the 8-byte writes of this lane
__device__ __forceinline__ void im_block_xy(uint8_t *tile, int p, int B, int jr, int hh, const uint32_t d[8],
                                            v4i bh, v4i b4l) {
    uint32_t hi[4], lo[4];
    vcg_im_planes(d, hi, lo);
    int o[16];
    im_chain(hi, lo, bh, b4l, vcg_im_rnd4(p, 0), o);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        uint32_t w[2];
        vcg_im_pack16h(o + 4 * q, w);
        *(uint2 *)(tile + vcg_im_out_addr(p, B, jr, hh, q)) = make_uint2(w[0], w[1]);
    }
}

// One workgroup's item i: scatter into the i16 tile, the bound, three
// product passes, the store.  Every thread calls it with the same i.
__device__ __forceinline__ void idct_item_m(const dec_job &j, uint32_t i, const vcg_dtab *__restrict__ T,
                                            const uint32_t *__restrict__ lists, vcg_item *__restrict__ items,
                                            uint8_t *tile, unsigned long long *ssum, uint32_t *sbig,
                                            int32_t *dqs) {
    // per-lane addresses derive from tid; an opaque copy keeps the persistent
    // kernel from hoisting them all out of its atom loop (and spilling them)
    int tid = (int)threadIdx.x;
    asm volatile("" : "+v"(tid));
    const uint32_t gi = j.base + i;
    int lod;
    uint32_t ax, ay, az;
    item_atom(j, gi, lod, ax, ay, az);
    // the item, the dequant tables and the list head load together (as idct_item)
    const uint32_t *L = lists + (size_t)i * VCG_LIST_STRIDE;
    uint32_t pre[IDCT_PRE];
#pragma unroll
    for (int q = 0; q < IDCT_PRE; ++q) pre[q] = L[tid + q * IM_THREADS];
    if (tid < VCG_NFREQ) {                   // step and bias of frequency tid as a pair
        const int o = j.mode == MODE_PAYLOAD ? 0 : lod * VCG_NFREQ;
        dqs[2 * tid] = j.dqs[o + tid];
        dqs[2 * tid + 1] = j.dqb[o + tid];
    }
    const vcg_item it = items[i];
    const int ln = tid & 63, wv = tid >> 6, jr = ln & 31, hh = ln >> 5;
    const v4i bh = *(const v4i *)T->ib[0][ln], b4l = *(const v4i *)T->ib[1][ln];
    if (tid == 0 && j.status_out) j.status_out[gi] = it.status == 2 ? 1 : it.status == 3 ? 2 : 0;
    if (it.status != 1) {
        for (int v = tid; v < (int)VCG_A3 / 4; v += IM_THREADS) ((uint32_t *)tile)[v] = 0;
        __syncthreads();
        store_atom(item_dst(j, gi), ax, ay, az, tile, tid);
        return;
    }
    const uint32_t cnt = it.cnt;
    const int fmax = it.fmax, dc = (int)it.dc;
    uint32_t tab[IDCT_PRE];
#pragma unroll
    for (int q = 0; q < IDCT_PRE; ++q)
        tab[q] = tid + q * IM_THREADS < (int)cnt ? T->imtab[vcg_ent_pos(pre[q])] : 0u;
    vcg_im_fill((uint32_t *)tile, tid, IM_THREADS);
    if (tid == 0) { *ssum = 0; *sbig = 0; }
    __syncthreads();
    uint64_t ss = 0;
    uint32_t big = 0;
#pragma unroll
    for (int q = 0; q < IDCT_PRE; ++q)
        if (tid + q * IM_THREADS < (int)cnt) vcg_im_scatter1(tile, pre[q], tab[q], dqs, ss, big);
    if (cnt > IDCT_PRE * IM_THREADS)
        vcg_im_scatter(tile, L + IDCT_PRE * IM_THREADS, cnt - IDCT_PRE * IM_THREADS, T->imtab, dqs, tid, IM_THREADS,
                       ss, big);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor((unsigned long long)ss, o, 64);
    const bool wbig = __ballot(big != 0) != 0;
    if (ln == 0) {
        if (ss) atomicAdd(ssum, (unsigned long long)ss);
        if (wbig) atomicOr(sbig, 1u);
    }
    __syncthreads();
    if (*sbig || *ssum > (unsigned long long)VCG_IM_L2 * VCG_IM_L2) {
        if (tid == 0) items[i].status = 4;
        return;
    }
    // pass x: the wave's planes kz = wv + 8b, in place (it owns them)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int B = wv + 8 * b;
        if (vcg_im_skip(0, B, fmax)) continue;
        uint32_t d[8];
        vcg_im_load(tile, 0, B, jr, hh, d);
        im_block_xy(tile, 0, B, jr, hh, d, bh, b4l);
    }
    __syncthreads();
    // pass y: every wave reads its blocks before any writes
    {
        uint32_t d[4][8];
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (!vcg_im_skip(1, wv + 8 * b, fmax)) vcg_im_load(tile, 1, wv + 8 * b, jr, hh, d[b]);
        __syncthreads();
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (!vcg_im_skip(1, wv + 8 * b, fmax)) im_block_xy(tile, 1, wv + 8 * b, jr, hh, d[b], bh, b4l);
    }
    __syncthreads();
    // pass z: voxels (dc added, clamped) into the staging raster
    {
        uint32_t d[4][8];
#pragma unroll
        for (int b = 0; b < 4; ++b) vcg_im_load(tile, 2, wv + 8 * b, jr, hh, d[b]);
        __syncthreads();
        const int rnd4 = vcg_im_rnd4(2, dc);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            uint32_t hi[4], lo[4];
            vcg_im_planes(d[b], hi, lo);
            int o[16];
            im_chain(hi, lo, bh, b4l, rnd4, o);
#pragma unroll
            for (int q = 0; q < 4; ++q)
                *(uint32_t *)(tile + vcg_im_vox_addr(wv + 8 * b, jr, hh, q)) = vcg_im_pack8h(o + 4 * q);
        }
    }
    __syncthreads();
    store_atom(item_dst(j, gi), ax, ay, az, tile, tid, VCG_IM_ZS);
}

__global__ __launch_bounds__(IM_THREADS) IM_REGS void k_idct_m(dec_job j, const vcg_dtab *__restrict__ T,
                                                       const uint32_t *__restrict__ lists,
                                                       vcg_item *__restrict__ items) {
    extern __shared__ __align__(16) uint8_t tile[];
    __shared__ unsigned long long ssum;
    __shared__ uint32_t sbig;
    __shared__ int32_t dqs[2 * VCG_NFREQ];
    idct_item_m(j, xcd_item(blockIdx.x, gridDim.x), T, lists, items, tile, &ssum, &sbig, dqs);
}

// the items of chunk [c0, c0 + 512) the matrix-core transform marked, through
// the i32 path
__device__ __forceinline__ void ovf_chunk(const dec_job &j, uint32_t c0, uint32_t n, const vcg_dtab *__restrict__ T,
                                          const uint32_t *__restrict__ lists, const vcg_item *__restrict__ items,
                                          int *buf, unsigned long long *l1sum, int32_t *dqs, uint32_t *cnt,
                                          uint32_t *list) {
    __syncthreads();                     // the previous chunk is done with cnt and list
    if (threadIdx.x == 0) *cnt = 0;
    __syncthreads();
    const uint32_t i = c0 + threadIdx.x;
    if (i < n && items[i].status == 4) list[atomicAdd(cnt, 1u)] = i;
    __syncthreads();
    const uint32_t c = *cnt;
    for (uint32_t k = 0; k < c; ++k) {
        idct_item(j, list[k], T, lists, items, buf, l1sum, dqs);
        __syncthreads();
    }
}

__global__ __launch_bounds__(IDCT_THREADS) void k_idct_ovf(dec_job j, const vcg_dtab *__restrict__ T,
                                                           const uint32_t *__restrict__ lists,
                                                           const vcg_item *__restrict__ items) {
    extern __shared__ int buf[];
    __shared__ unsigned long long l1sum;
    __shared__ int32_t dqs[2 * VCG_NFREQ];
    __shared__ uint32_t cnt, list[IDCT_THREADS];
    ovf_chunk(j, blockIdx.x * IDCT_THREADS, j.n, T, lists, items, buf, &l1sum, dqs, &cnt, list);
}

// persistent forms for the server's captured batches
__global__ __launch_bounds__(IM_THREADS) IM_REGS void k_idct_mp(dec_job j, const vcg_dtab *__restrict__ T,
                                                        const uint32_t *__restrict__ lists,
                                                        vcg_item *__restrict__ items, vcg_qhdr *__restrict__ q) {
    extern __shared__ __align__(16) uint8_t tile[];
    __shared__ unsigned long long ssum;
    __shared__ uint32_t sbig;
    __shared__ int32_t dqs[2 * VCG_NFREQ];
    __shared__ uint32_t cur;
    const uint32_t n = q->n;
#pragma clang loop unroll(disable)
    for (;;) {
        __syncthreads();                 // the previous atom is done with the tile and cur
        if (threadIdx.x == 0) cur = atomicAdd(&q->idct, 1u);
        __syncthreads();
        const uint32_t i = __builtin_amdgcn_readfirstlane(cur);   // uniform: scalar addressing
        if (i >= n) break;
        idct_item_m(j, i, T, lists, items, tile, &ssum, &sbig, dqs);
    }
}

__global__ __launch_bounds__(IDCT_THREADS) void k_idct_ovf_p(dec_job j, const vcg_dtab *__restrict__ T,
                                                             const uint32_t *__restrict__ lists,
                                                             const vcg_item *__restrict__ items,
                                                             const vcg_qhdr *__restrict__ q) {
    extern __shared__ int buf[];
    __shared__ unsigned long long l1sum;
    __shared__ int32_t dqs[2 * VCG_NFREQ];
    __shared__ uint32_t cnt, list[IDCT_THREADS];
    const uint32_t n = q->n;
    for (uint32_t c0 = blockIdx.x * IDCT_THREADS; c0 < n; c0 += gridDim.x * IDCT_THREADS)
        ovf_chunk(j, c0, n, T, lists, items, buf, &l1sum, dqs, &cnt, list);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static const size_t LDS_BYTES = (size_t)VCG_TILE_INTS * 4;
static const uint32_t MAX_BATCH = 32768;   // items per launch pair (scratch: 4 GiB)
static const uint64_t VCG_HOST_SLAB = 1ull << 30;   // host-destination staging per slab

static const size_t LDS_M_BYTES = VCG_IM_TILE_BYTES;

// k_entropy's launches carry unused dynamic LDS so that its workgroup holds
// 84 KiB: more than half of a CU's 160 KiB, so a CU takes at most one entropy
// workgroup (one wave per SIMD), and still room for one matrix-core transform
// workgroup (64 KiB tile + tables) beside it.  Entropy launches of
// concurrent decodes then spread over disjoint CUs instead of stacking two
// waves on some SIMDs and leaving others idle (back-to-back LOD0 decode of
// 1024^3 at ratio 10: 5.12 -> 4.33 ms per decode; a lone launch is
// unchanged, it never places two workgroups on a CU).  The server's
// four-wave persistent entropy launches carry the same padding (its
// one- and two-wave shapes do not: they serve small batches).
static const size_t ENT_LDS_STATIC = (size_t)ENT_WAVES * (VCG_CTX_SLOTS + VCG_RING) * 64u * 4u + 512u + 9u * 4u;
static const size_t ENT_LDS_TOTAL = 84u * 1024u;
static const size_t ENT_LDS_PAD = ENT_LDS_TOTAL > ENT_LDS_STATIC ? ENT_LDS_TOTAL - ENT_LDS_STATIC : 0;
static_assert(2 * ENT_LDS_TOTAL > 160u * 1024u, "two entropy workgroups must not fit one CU");
static_assert(ENT_LDS_TOTAL + VCG_IM_TILE_BYTES + 2048u <= 160u * 1024u,
              "an entropy and a transform workgroup must fit one CU");

static void ensure_attr() {
    static std::once_flag once;
    std::call_once(once, [] {
        const void *big[] = {(const void *)k_idct, (const void *)k_idct_p, (const void *)k_idct_ovf,
                             (const void *)k_idct_ovf_p};
        for (const void *k : big)
            VCG_CHECK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)LDS_BYTES));
        const void *mk[] = {(const void *)k_idct_m, (const void *)k_idct_mp};
        for (const void *k : mk)
            VCG_CHECK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)LDS_M_BYTES));
        VCG_CHECK(hipFuncSetAttribute((const void *)k_entropy, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)ENT_LDS_PAD));
        VCG_CHECK(hipFuncSetAttribute((const void *)k_entropy_p<4>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)ENT_LDS_PAD));
    });
}

// The transform of a launch's j.n items on stream s: the matrix-core kernel,
// then the i32 kernel for the items it marked (a scan that ends at once when
// none is).  VCG_IDCT_I32 builds the i32 kernel alone.
static void launch_idct(const dec_job &j, uint32_t *lists, vcg_item *items, hipStream_t s) {
#ifdef VCG_IDCT_I32
    hipLaunchKernelGGL(k_idct, dim3(j.n), dim3(IDCT_THREADS), LDS_BYTES, s, j, g_dtab, lists, items);
#else
    hipLaunchKernelGGL(k_idct_m, dim3(j.n), dim3(IM_THREADS), LDS_M_BYTES, s, j, g_dtab, lists, items);
    hipLaunchKernelGGL(k_idct_ovf, dim3((j.n + IDCT_THREADS - 1) / IDCT_THREADS), dim3(IDCT_THREADS), LDS_BYTES, s,
                       j, g_dtab, lists, items);
#endif
}
// the persistent form: gt transform workgroups, batches of up to cap items
static void launch_idct_p(const dec_job &j, uint32_t gt, size_t cap, uint32_t *lists, vcg_item *items,
                          vcg_qhdr *hq, hipStream_t s) {
#ifdef VCG_IDCT_I32
    hipLaunchKernelGGL(k_idct_p, dim3(gt), dim3(IDCT_THREADS), LDS_BYTES, s, j, g_dtab, lists, items, hq);
#else
    hipLaunchKernelGGL(k_idct_mp, dim3(gt), dim3(IM_THREADS), LDS_M_BYTES, s, j, g_dtab, lists, items, hq);
    const uint32_t go = (uint32_t)((cap + IDCT_THREADS - 1) / IDCT_THREADS);
    hipLaunchKernelGGL(k_idct_ovf_p, dim3(go < 64 ? go : 64), dim3(IDCT_THREADS), LDS_BYTES, s, j, g_dtab, lists,
                       items, hq);
#endif
}

// Launch the entropy / transform pairs for `total` items on the slot streams
// (see vcg_ctx) and join the result into c->s.  inputs_on_s: the job reads
// buffers this call staged on c->s (list quads, payloads), so the entropy pass
// waits for them too; the transform always waits for earlier work on c->s (it
// writes the caller's output).  up_*: an input upload done on the first slot
// stream instead (pinned source).  Returns the scratch slot of the last batch.
// Decoder scratch (coefficient lists and item records) is one pool per
// device, shared by every host thread: VCG_SCRATCH_BLOCKS blocks taken in
// turn, each batch's launches ordered after the previous user of its block
// by that block's last-use event (a wait on the GPU, not on the host).  So
// the scratch stays at most VCG_SCRATCH_BLOCKS x the largest batch however
// many threads decode.  A block grows (after its last user finished) when a
// batch needs more atoms than it holds.
#define VCG_SCRATCH_BLOCKS 3
struct scratch_blk {
    uint32_t *lists = nullptr;
    vcg_item *items = nullptr;
    size_t cap = 0;
    hipEvent_t last = nullptr;
    bool used = false;
};
static std::mutex g_blk_mu;
static scratch_blk g_blk[VCG_SCRATCH_BLOCKS];
static int g_blk_next = 0;

// the next block with room for n atoms (grown if needed), its last user
// waited for on stream sE; returns it with g_blk_mu held
static scratch_blk *blk_take(hipStream_t sE, size_t n) {
    g_blk_mu.lock();
    scratch_blk *b = &g_blk[g_blk_next];
    g_blk_next = (g_blk_next + 1) % VCG_SCRATCH_BLOCKS;
    if (!b->last) VCG_CHECK(hipEventCreateWithFlags(&b->last, hipEventDisableTiming));
    if (b->cap < n) {
        if (b->used) VCG_CHECK(hipEventSynchronize(b->last));
        if (b->lists) VCG_CHECK(hipFree(b->lists));
        if (b->items) VCG_CHECK(hipFree(b->items));
        b->lists = nullptr; b->items = nullptr; b->cap = 0;
        if (hipMalloc((void **)&b->lists, n * (size_t)VCG_LIST_STRIDE * 4) != hipSuccess ||
            hipMalloc((void **)&b->items, n * sizeof(vcg_item)) != hipSuccess) {
            (void)hipGetLastError();
            g_blk_mu.unlock();
            vc__panic("out of device memory (decoder scratch)");
        }
        b->cap = n;
        b->used = false;
    }
    if (b->used) VCG_CHECK(hipStreamWaitEvent(sE, b->last, 0));
    return b;
}
// after the block's last launch on sE: record its use, release the pool
static void blk_give(scratch_blk *b, hipStream_t sE) {
    VCG_CHECK(hipEventRecord(b->last, sE));
    b->used = true;
    g_blk_mu.unlock();
}

static int run_batch(vcg_ctx *c, dec_job j, uint32_t total, bool inputs_on_s, const void *up_src = nullptr,
                     void *up_dst = nullptr, size_t up_bytes = 0) {
    ensure_attr();
    uint32_t maxb = total < MAX_BATCH ? total : MAX_BATCH;
    VCG_CHECK(hipEventRecord(c->evF, c->s));
    const int first = c->slot;
    if (up_bytes) VCG_CHECK(hipMemcpyAsync(up_dst, up_src, up_bytes, hipMemcpyHostToDevice, c->sE[first]));
    for (uint32_t b = 0; b < total; b += maxb) {
        const int sl = c->slot;
        c->slot = (c->slot + 1) % VCG_SLOTS;
        hipStream_t sE = c->sE[sl];
        j.base = b;
        j.n = total - b < maxb ? total - b : maxb;
        // entropy: after the transform that last used this slot (same stream)
        if (inputs_on_s) VCG_CHECK(hipStreamWaitEvent(sE, c->evF, 0));
        if (up_bytes && sl != first) VCG_CHECK(hipStreamWaitEvent(sE, c->evI[first], 0));
        scratch_blk *blk = blk_take(sE, maxb);
        if (c->ev[0]) VCG_CHECK(hipEventRecord((hipEvent_t)c->ev[0], sE));
        hipLaunchKernelGGL(k_entropy, dim3((j.n + ENT_THREADS - 1) / ENT_THREADS), dim3(ENT_THREADS), ENT_LDS_PAD, sE,
                           j, g_dtab, blk->lists, blk->items, c->status_d);
        VCG_CHECK(hipGetLastError());
        if (c->ev[1]) VCG_CHECK(hipEventRecord((hipEvent_t)c->ev[1], sE));
        // transform: after its entropy pass (same stream) and earlier work on c->s
        if (!inputs_on_s) VCG_CHECK(hipStreamWaitEvent(sE, c->evF, 0));
        launch_idct(j, blk->lists, blk->items, sE);
        VCG_CHECK(hipGetLastError());
        blk_give(blk, sE);
        if (c->ev[2]) VCG_CHECK(hipEventRecord((hipEvent_t)c->ev[2], sE));
        VCG_CHECK(hipEventRecord(c->evI[sl], sE));
        c->last_slot = sl;
    }
    // join every slot this call used into c->s
    const uint32_t nb = (total + maxb - 1) / maxb;
    for (uint32_t k = 0; k < nb && k < VCG_SLOTS; ++k)
        VCG_CHECK(hipStreamWaitEvent(c->s, c->evI[(first + k) % VCG_SLOTS], 0));
    return c->last_slot;
}

extern "C" int vcg_decode_box(vcg_arc *a, int lod, const vcg_box *box, uint8_t *out, int out_host,
                              int sync) {
    vcg_ctx *c = vcg_get_ctx();
    const vcg_geom *g = vcg_arc_geom_host(a);
    const vcg_geom *gd = vcg_arc_geom_dev(a);
    dec_job j;
    memset(&j, 0, sizeof j);
    j.arc = vcg_arc_buf(a);
    j.geom = gd;
    j.dqs = &gd->dq_step[0][0];
    j.dqb = &gd->dq_bias[0][0];
    j.mode = MODE_BOX;
    j.lod = lod;
    j.ax0 = box->x0 / 32; j.ay0 = box->y0 / 32; j.az0 = box->z0 / 32;
    j.nax = (box->x1 - 1) / 32 - j.ax0 + 1;
    j.nay = (box->y1 - 1) / 32 - j.ay0 + 1;
    uint32_t naz = (box->z1 - 1) / 32 - j.az0 + 1;
    j.bx0 = box->x0; j.by0 = box->y0; j.bz0 = box->z0;
    j.bx1 = box->x1; j.by1 = box->y1; j.bz1 = box->z1;
    (void)g;
    if (!out_host) {
        j.out = out;
        run_batch(c, j, j.nax * j.nay * naz, false);
        return sync ? vcg_sync() : 0;
    }
    // Host destination: decoded into device staging, then copied.  Boxes of
    // more than VCG_HOST_SLAB bytes go as z-slabs of whole atom layers (the
    // first and last clipped to the box) through two alternating staging
    // buffers, so a thread's staging stays bounded and a slab's copy overlaps
    // the next slab's decode; a slab's raster is the contiguous piece of the
    // box raster at its first z.  The transform of slab k+2 waits for earlier
    // work on c->s (run_batch), which includes slab k's copy out of the same
    // buffer.
    const uint64_t plane = (uint64_t)(box->x1 - box->x0) * (box->y1 - box->y0);
    const size_t bytes = plane * (box->z1 - box->z0);
    uint32_t per = naz;                                     // atom layers per slab
    if (bytes > VCG_HOST_SLAB) {
        per = (uint32_t)(VCG_HOST_SLAB / (plane * 32));
        if (per == 0) per = 1;
    }
    const size_t slab = per >= naz ? bytes : (size_t)per * 32 * plane;
    vcg_grow((void **)&c->stage, &c->stage_cap, per >= naz ? bytes : 2 * slab, 1);
    for (uint32_t k = 0, la = j.az0; la < j.az0 + naz; ++k, la += per) {
        const uint32_t lb = la + per < j.az0 + naz ? la + per : j.az0 + naz;
        dec_job jj = j;
        jj.az0 = la;
        jj.bz0 = la * 32 > box->z0 ? la * 32 : box->z0;
        jj.bz1 = lb * 32 < box->z1 ? lb * 32 : box->z1;
        jj.out = c->stage + (k & 1) * slab;
        run_batch(c, jj, j.nax * j.nay * (lb - la), false);
        VCG_CHECK(hipMemcpyAsync(out + (uint64_t)(jj.bz0 - box->z0) * plane, jj.out, (uint64_t)(jj.bz1 - jj.bz0) * plane,
                                 hipMemcpyDeviceToHost, c->s));
    }
    return vcg_sync();
}

extern "C" int vcg_decode_list(vcg_arc *a, size_t n, const int32_t *quads, uint8_t *out, int32_t *per_item) {
    vcg_ctx *c = vcg_get_ctx();
    const vcg_geom *gd = vcg_arc_geom_dev(a);
    if (n == 0) return 0;
    vcg_grow((void **)&c->quads, &c->quads_cap, n * 4, 4);
    VCG_CHECK(hipMemcpyAsync(c->quads, quads, n * 16, hipMemcpyHostToDevice, c->s));
    vcg_grow((void **)&c->stage, &c->stage_cap, n * (size_t)VCG_A3, 1);
    dec_job j;
    memset(&j, 0, sizeof j);
    j.arc = vcg_arc_buf(a);
    j.geom = gd;
    j.dqs = &gd->dq_step[0][0];
    j.dqb = &gd->dq_bias[0][0];
    j.mode = MODE_LIST;
    j.quads = c->quads;
    j.out = c->stage;
    int32_t *dst = nullptr;
    if (per_item) {
        vcg_grow((void **)&c->stat, &c->stat_cap, n, 4);
        dst = c->stat;
    }
    for (size_t b = 0; b < n; b += MAX_BATCH) {
        uint32_t cnt = (uint32_t)(n - b < MAX_BATCH ? n - b : MAX_BATCH);
        dec_job jj = j;
        jj.quads = c->quads + 4 * b;
        jj.out = c->stage + b * (size_t)VCG_A3;
        jj.status_out = dst ? dst + b : nullptr;
        run_batch(c, jj, cnt, true);
    }
    if (per_item) VCG_CHECK(hipMemcpyAsync(per_item, dst, n * 4, hipMemcpyDeviceToHost, c->s));
    VCG_CHECK(hipMemcpyAsync(out, c->stage, n * (size_t)VCG_A3, hipMemcpyDeviceToHost, c->s));
    return vcg_sync();
}

extern "C" int vcg_decode_payloads(size_t n, const uint8_t *pay, const uint64_t *offs, const uint32_t *lens,
                                   const int32_t dq_step[VCG_NFREQ], const int32_t dq_bias[VCG_NFREQ],
                                   uint8_t *out, int32_t *per_item) {
    vcg_ctx *c = vcg_get_ctx();
    if (n == 0) return 0;
    uint64_t total = 0;
    for (size_t i = 0; i < n; ++i) if (offs[i] + lens[i] > total) total = offs[i] + lens[i];
    uint8_t *dpay = nullptr; uint64_t *doffs = nullptr; uint32_t *dlens = nullptr; int32_t *ddq = nullptr;
    VCG_CHECK(hipMalloc(&dpay, total + 64));
    VCG_CHECK(hipMemsetAsync(dpay + total, 0, 64, c->s));
    VCG_CHECK(hipMalloc(&doffs, n * 8));
    VCG_CHECK(hipMalloc(&dlens, n * 4));
    VCG_CHECK(hipMalloc(&ddq, 2 * VCG_NFREQ * 4));
    VCG_CHECK(hipMemcpyAsync(dpay, pay, total, hipMemcpyHostToDevice, c->s));
    VCG_CHECK(hipMemcpyAsync(doffs, offs, n * 8, hipMemcpyHostToDevice, c->s));
    VCG_CHECK(hipMemcpyAsync(dlens, lens, n * 4, hipMemcpyHostToDevice, c->s));
    VCG_CHECK(hipMemcpyAsync(ddq, dq_step, VCG_NFREQ * 4, hipMemcpyHostToDevice, c->s));
    VCG_CHECK(hipMemcpyAsync(ddq + VCG_NFREQ, dq_bias, VCG_NFREQ * 4, hipMemcpyHostToDevice, c->s));
    vcg_grow((void **)&c->stage, &c->stage_cap, n * (size_t)VCG_A3, 1);
    dec_job j;
    memset(&j, 0, sizeof j);
    j.arc = dpay; j.mode = MODE_PAYLOAD; j.offs = doffs; j.lens = dlens;
    j.dqs = ddq; j.dqb = ddq + VCG_NFREQ;
    j.out = c->stage;
    int32_t *dst = nullptr;
    if (per_item) {
        vcg_grow((void **)&c->stat, &c->stat_cap, n, 4);
        dst = c->stat;
    }
    for (size_t b = 0; b < n; b += MAX_BATCH) {
        uint32_t cnt = (uint32_t)(n - b < MAX_BATCH ? n - b : MAX_BATCH);
        dec_job jj = j;
        jj.base = 0;
        jj.offs = doffs + b; jj.lens = dlens + b;
        jj.out = c->stage + b * (size_t)VCG_A3;
        jj.status_out = dst ? dst + b : nullptr;
        run_batch(c, jj, cnt, true);
    }
    if (per_item) VCG_CHECK(hipMemcpyAsync(per_item, dst, n * 4, hipMemcpyDeviceToHost, c->s));
    VCG_CHECK(hipMemcpyAsync(out, c->stage, n * (size_t)VCG_A3, hipMemcpyDeviceToHost, c->s));
    int rc = vcg_sync();
    VCG_CHECK(hipFree(dpay)); VCG_CHECK(hipFree(doffs)); VCG_CHECK(hipFree(dlens)); VCG_CHECK(hipFree(ddq));
    return rc;
}

// ---------------------------------------------------------------------------
// asynchronous launchers (serving loop): enqueue and return; completion is
// observed through an event recorded on the calling thread's stream
// ---------------------------------------------------------------------------
extern "C" size_t vcg_slot_scratch_bytes(void) { return (size_t)VCG_LIST_STRIDE * 4 + sizeof(vcg_item); }



// ---------------------------------------------------------------------------
// allocation-free list decode (vcx_decode_atoms_noalloc)
// ---------------------------------------------------------------------------
// The decoder's scratch for up to VCG_NOALLOC_MAX atoms lives in one
// file-scope device array: [unused 256 B][items][coefficient lists].  Every
// device holds its own copy, at its own address, until the process ends.  In
// host code `scratch` only names the variable: its device address comes from
// hipGetSymbolAddress, once per device at set-up (vcg_noalloc_setup, called
// from vcg_init with the device current), and only that address is passed to
// kernels.  Nothing in it needs clearing: k_entropy writes every item and
// list entry the transform reads.  Calls that overlap on one device share
// the array and race.
#define NA_ITEMS_OFF 256u
#define NA_LISTS_OFF (NA_ITEMS_OFF + VCG_NOALLOC_MAX * 16u)
#define NA_BYTES (NA_LISTS_OFF + (size_t)VCG_NOALLOC_MAX * VCG_LIST_STRIDE * 4u)
static_assert(sizeof(vcg_item) == 16, "item record size");
static_assert(NA_LISTS_OFF % 256 == 0, "list alignment");
static __device__ __align__(256) uint8_t scratch[NA_BYTES];

#define NA_MAX_DEVICES 64
static uint8_t *g_scratch_dev[NA_MAX_DEVICES];

extern "C" int vcg_noalloc_setup(void) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAX_DEVICES) { (void)hipGetLastError(); return -1; }
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) { (void)hipGetLastError(); return -1; }
    g_scratch_dev[dev] = (uint8_t *)p;
    return 0;
}

extern "C" size_t vcg_noalloc_bytes(void) { return NA_BYTES; }

extern "C" int vcg_list_launch_noalloc(vcg_arc *a, size_t n, const int32_t *quads_dev, uint8_t *out_dev,
                                       int32_t *status_dev, void *stream) {
    if (n == 0 || n > VCG_NOALLOC_MAX) return -1;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAX_DEVICES) { (void)hipGetLastError(); return -1; }
    uint8_t *base = g_scratch_dev[dev];
    if (!base) return -1;
    ensure_attr();
    const vcg_geom *gd = vcg_arc_geom_dev(a);
    dec_job j;
    memset(&j, 0, sizeof j);
    j.arc = vcg_arc_buf(a);
    j.geom = gd;
    j.dqs = &gd->dq_step[0][0];
    j.dqb = &gd->dq_bias[0][0];
    j.mode = MODE_LIST;
    j.check = 1;
    j.quads = quads_dev;
    j.out = out_dev;
    j.status_out = status_dev;
    j.base = 0;
    j.n = (uint32_t)n;
    hipStream_t s = (hipStream_t)stream;
    // the word k_entropy ORs corrupt atoms into sits in the unused head; this
    // path reports per atom and never reads it
    uint32_t *word = (uint32_t *)base;
    vcg_item *items = (vcg_item *)(base + NA_ITEMS_OFF);
    uint32_t *lists = (uint32_t *)(base + NA_LISTS_OFF);
    hipLaunchKernelGGL(k_entropy, dim3((j.n + ENT_THREADS - 1) / ENT_THREADS), dim3(ENT_THREADS), ENT_LDS_PAD, s, j,
                       g_dtab, lists, items, word);
    if (hipGetLastError() != hipSuccess) return -1;
    launch_idct(j, lists, items, s);
    if (hipGetLastError() != hipSuccess) return -1;
    return 0;
}

// ---------------------------------------------------------------------------
// captured persistent batch decode (server slots)
// ---------------------------------------------------------------------------
struct vcg_qgraph {
    hipGraphExec_t exec = nullptr;
    hipStream_t s = nullptr;
    uint8_t *qdev = nullptr;
    size_t cap = 0;
    int slot = 0;
    uint32_t *lists = nullptr;           // the caller's scratch the graph was captured with
    vcg_item *items = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    // the captured launches' arguments (direct launches)
    dec_job j;
    vcg_qcfg cfg;
    uint32_t *status = nullptr;
};

static_assert(sizeof(vcg_qhdr) == 16, "batch header");

extern "C" vcg_qgraph *vcg_qgraph_create(vcg_arc *a, int slot, const vcg_qcfg *cfg, uint8_t *qdev, size_t cap,
                                         uint8_t *out_dev, int32_t *status_dev, void *scratch) {
    if (slot < 0 || slot >= VCG_SLOTS || cap == 0 || cap > MAX_BATCH || !cfg) return nullptr;
    const int W = cfg->ent_waves;
    if ((W != 1 && W != 2 && W != 4) || cfg->ent_groups == 0 || cfg->idct_groups == 0) return nullptr;
    if (!scratch) return nullptr;
    vcg_ctx *c = vcg_get_ctx();
    ensure_attr();
    const vcg_geom *gd = vcg_arc_geom_dev(a);
    dec_job j;
    memset(&j, 0, sizeof j);
    j.arc = vcg_arc_buf(a);
    j.geom = gd;
    j.dqs = &gd->dq_step[0][0];
    j.dqb = &gd->dq_bias[0][0];
    j.mode = MODE_LIST;
    j.quads = (const int32_t *)(qdev + 16);
    j.places = (const vcg_place *)(qdev + 16 + cap * 16);
    j.out = out_dev;
    j.status_out = status_dev;
    vcg_qhdr *hq = (vcg_qhdr *)qdev;
    vcg_qgraph *g = new vcg_qgraph();
    g->s = c->sE[slot]; g->qdev = qdev; g->cap = cap; g->slot = slot;
    // the caller's scratch: cap list strides, then cap item records
    g->lists = (uint32_t *)scratch;
    g->items = (vcg_item *)((uint8_t *)scratch + cap * (size_t)VCG_LIST_STRIDE * 4);
    g->j = j; g->cfg = *cfg; g->status = c->status_d;
    hipGraph_t graph = nullptr;
    VCG_CHECK(hipStreamBeginCapture(g->s, hipStreamCaptureModeThreadLocal));
    const dim3 ge(cfg->ent_groups);
    if (W == 1)
        hipLaunchKernelGGL(k_entropy_p<1>, ge, dim3(64), 0, g->s, j, g_dtab, g->lists, g->items, c->status_d, hq);
    else if (W == 2)
        hipLaunchKernelGGL(k_entropy_p<2>, ge, dim3(128), 0, g->s, j, g_dtab, g->lists, g->items, c->status_d, hq);
    else
        hipLaunchKernelGGL(k_entropy_p<4>, ge, dim3(256), ENT_LDS_PAD, g->s, j, g_dtab, g->lists, g->items,
                           c->status_d, hq);
    launch_idct_p(j, cfg->idct_groups, cap, g->lists, g->items, hq, g->s);
    VCG_CHECK(hipStreamEndCapture(g->s, &graph));
    VCG_CHECK(hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0));
    VCG_CHECK(hipGraphDestroy(graph));
    VCG_CHECK(hipEventCreate(&g->t0));
    VCG_CHECK(hipEventCreate(&g->t1));
    return g;
}

extern "C" int vcg_qgraph_launch(vcg_qgraph *g, const uint8_t *qpin, size_t n, size_t nplace, int direct) {
    if (!g || n == 0 || n > g->cap || nplace > g->cap) return -1;
    // header (n, claim counters at 0) and quads in one copy, placements in another
    VCG_CHECK(hipMemcpyAsync(g->qdev, qpin, 16 + n * 16, hipMemcpyHostToDevice, g->s));
    if (nplace)
        VCG_CHECK(hipMemcpyAsync(g->qdev + 16 + g->cap * 16, qpin + 16 + g->cap * 16, nplace * sizeof(vcg_place),
                                 hipMemcpyHostToDevice, g->s));
    if (direct) {                        // the same two kernels, launched one by one
        vcg_qhdr *hq = (vcg_qhdr *)g->qdev;
        const dim3 ge(g->cfg.ent_groups);
        if (g->cfg.ent_waves == 1)
            hipLaunchKernelGGL(k_entropy_p<1>, ge, dim3(64), 0, g->s, g->j, g_dtab, g->lists, g->items, g->status, hq);
        else if (g->cfg.ent_waves == 2)
            hipLaunchKernelGGL(k_entropy_p<2>, ge, dim3(128), 0, g->s, g->j, g_dtab, g->lists, g->items, g->status, hq);
        else
            hipLaunchKernelGGL(k_entropy_p<4>, ge, dim3(256), ENT_LDS_PAD, g->s, g->j, g_dtab, g->lists, g->items,
                               g->status, hq);
        launch_idct_p(g->j, g->cfg.idct_groups, g->cap, g->lists, g->items, hq, g->s);
        VCG_CHECK(hipGetLastError());
        return 0;
    }
    VCG_CHECK(hipGraphLaunch(g->exec, g->s));
    return 0;
}

extern "C" float vcg_qgraph_time(vcg_qgraph *g, const uint8_t *qpin, size_t n, size_t nplace, int reps) {
    float best = 1e30f;
    for (int r = 0; r < reps; ++r) {
        VCG_CHECK(hipEventRecord(g->t0, g->s));
        if (vcg_qgraph_launch(g, qpin, n, nplace, 0) != 0) return -1.0f;
        VCG_CHECK(hipEventRecord(g->t1, g->s));
        VCG_CHECK(hipEventSynchronize(g->t1));
        float ms = 0;
        VCG_CHECK(hipEventElapsedTime(&ms, g->t0, g->t1));
        if (ms < best) best = ms;
    }
    return best;
}

extern "C" void vcg_qgraph_destroy(vcg_qgraph *g) {
    if (!g) return;
    VCG_CHECK(hipStreamSynchronize(g->s));
    VCG_CHECK(hipGraphExecDestroy(g->exec));
    VCG_CHECK(hipEventDestroy(g->t0));
    VCG_CHECK(hipEventDestroy(g->t1));
    delete g;
}
