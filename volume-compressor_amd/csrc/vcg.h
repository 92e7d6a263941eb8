/* vcg.h — the thin C ABI between the plain-C host library (csrc/host) and the
 * HIP kernels (csrc/hip).  Host C owns the archive format, rate control and
 * layout; the vcg_* functions own device memory, streams and kernel launches.
 * Plain C types only.
 */
#ifndef VCG_H
#define VCG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VCG_A3        32768u
#define VCG_NFREQ     94
#define VCG_NLOD      8
#define VCG_RC_CAP    (32768u * 4u + 64u)   /* range-coder bound, src/vc/vc.c:611 */
#define VCG_PAY_CAP   (VCG_RC_CAP + 1u)     /* + dc byte */

/* Codec tables, computed once on the host (libm) and uploaded. */
typedef struct {
    int32_t  cmat[32][32];      /* Q14 DCT basis CMATA[k][n] */
    float    gstep[VCG_NFREQ];  /* (1+freq)^0.65f */
    uint16_t scan[VCG_A3];      /* scan position -> raster index */
    uint32_t band_start[9];     /* first scan position of each sig band; [8] = 32768 */
} vcg_tables;

/* Archive geometry, parsed and validated on the host. */
typedef struct {
    uint64_t len;                        /* archive bytes */
    uint64_t l1_off, l1_cap;             /* v1 */
    uint64_t dir_off;                    /* v2 */
    uint64_t region_prefix[VCG_NLOD + 1];/* v2 region index base per LOD */
    uint32_t nrx[VCG_NLOD], nry[VCG_NLOD], nrz[VCG_NLOD]; /* region grid */
    uint32_t acx[VCG_NLOD], acy[VCG_NLOD], acz[VCG_NLOD]; /* atom grid */
    uint32_t lnx[VCG_NLOD], lny[VCG_NLOD], lnz[VCG_NLOD]; /* LOD dims */
    int32_t  version, nlod;
    int32_t  dq_step[VCG_NLOD][VCG_NFREQ];  /* Q8 dequant per LOD (host libm) */
    int32_t  dq_bias[VCG_NLOD][VCG_NFREQ];
} vcg_geom;

typedef struct { uint32_t x0, y0, z0, x1, y1, z1; } vcg_box;

/* ---- runtime ---- */
int   vcg_init(const vcg_tables *t);        /* 0 ok, -1 no device */
void *vcg_stream(void);                     /* calling thread's hipStream_t */
void *vcg_alloc(size_t bytes);              /* NULL on failure */
void  vcg_free(void *p);
int   vcg_copy(void *dst, const void *src, size_t bytes, int kind); /* 1 h2d, 2 d2h, 3 d2d */
void  vcg_set_stage_events(void *ev[3]);
int   vcg_copy_kernel(void *dst, const void *src, size_t bytes); /* 16-B aligned, bytes % 16 == 0 */
int   vcg_sync(void);                       /* wait; returns accumulated vc status */
/* asynchronous helpers (serving loop): on the calling thread's stream
 * (slot -1) or on the stream of its scratch slot `slot` (0 .. VCG_SLOTS-1) */
int   vcg_copy_async(void *dst, const void *src, size_t bytes, int kind, int slot); /* kind 0: by pointer */
int   vcg_mem_info(size_t *free_bytes, size_t *total_bytes);   /* hipMemGetInfo; 0 / -1 */
int   vcg_ptr_is_device(const void *p);     /* 1: device memory, 0: host (pinned or pageable) */
void *vcg_event_create(void);
void  vcg_event_destroy(void *e);
void  vcg_event_record(void *e, int slot);
int   vcg_event_done(void *e);              /* 1 when everything before the record finished */
void  vcg_event_sync(void *e);              /* wait until everything before the record finished */
void *vcg_host_alloc(size_t bytes);         /* pinned host memory */
void  vcg_host_free(void *p);

/* ---- device-resident archive ---- */
typedef struct vcg_arc vcg_arc;
vcg_arc *vcg_arc_create(const uint8_t *buf, const vcg_geom *g);   /* copies buf to HBM */
void     vcg_arc_destroy(vcg_arc *a);
/* Streaming archives: an arena of g->len bytes in HBM, head (header and
 * directory, head_len bytes) at its start; region blobs are put into it and
 * rebased there (present slots then address the arena; 0 done, 1 a payload
 * lies in the file outside the blob, -1 bad arguments). */
vcg_arc *vcg_arc_create_arena(const vcg_geom *g, const uint8_t *head, uint64_t head_len);
int      vcg_arc_put(vcg_arc *a, uint64_t off, const void *src, uint64_t n);   /* synchronous; 0 / -1 */
int      vcg_arc_rebase(vcg_arc *a, uint64_t at, uint64_t file_off, uint32_t blob_len, uint64_t file_len);
void     vcg_device_sync(void);                                              /* every stream of the device */

/* Decode every atom overlapping `box` (voxels of LOD `lod`) into out (box
 * raster).  out_host: out is host memory (staged through HBM).  sync: wait
 * and return the status (0 OK / 1 FORMAT); otherwise enqueue and return 0. */
int vcg_decode_box(vcg_arc *a, int lod, const vcg_box *box, uint8_t *out, int out_host, int sync);
/* Decode n whole atoms given as (lod,ax,ay,az) quads into host out (n*32768).
 * Coordinates must already be range-checked. per_item: status per atom. */
int vcg_decode_list(vcg_arc *a, size_t n, const int32_t *quads, uint8_t *out, int32_t *per_item);
/* A box placement for list items: the item's atom is clipped to the voxel
 * box [x0,x1) x [y0,y1) x [z0,z1) of its LOD and written into the box raster
 * at out_dev + out_off. */
typedef struct { uint32_t x0, y0, z0, x1, y1, z1; uint64_t out_off; } vcg_place;
/* Device decoder scratch bytes per atom of a batch's capacity (coefficient
 * list and item record). */
size_t vcg_slot_scratch_bytes(void);
/* Captured persistent batch decode (server slots).  A graph decodes batches
 * of up to cap quads on the stream of scratch slot `slot`: the entropy
 * kernel with cfg->ent_waves (1, 2 or 4) waves per workgroup and
 * cfg->ent_groups workgroups, whose waves claim 64 atoms at a time until the
 * batch is exhausted, then the transform with cfg->idct_groups workgroups
 * claiming one atom at a time.  qdev (device) and the caller's pinned qpin
 * share one layout: a 16-byte header {n, 0, 0, 0}, cap quads (lod |
 * placement << 8, ax, ay, az; a placement p sends the item's atom clipped
 * to box placement p - 1, else whole to out_dev + item * 32768), cap
 * placements (vcg_place); status_dev gets 1 per corrupt atom, else 0.
 * scratch: cap x vcg_slot_scratch_bytes() bytes of device memory the graph
 * keeps using (its batches must not overlap).  launch: upload the header, n
 * quads and nplace placements, then replay the graph, or with direct launch
 * the same two kernels one by one (enqueue only).  time: reps launches, each timed by events on the slot stream
 * (synchronous), the fastest in ms (the autotuner's measurement). */
typedef struct { int32_t ent_waves; uint32_t ent_groups; uint32_t idct_groups; } vcg_qcfg;
typedef struct vcg_qgraph vcg_qgraph;
vcg_qgraph *vcg_qgraph_create(vcg_arc *a, int slot, const vcg_qcfg *cfg, uint8_t *qdev, size_t cap,
                              uint8_t *out_dev, int32_t *status_dev, void *scratch);
int   vcg_qgraph_launch(vcg_qgraph *g, const uint8_t *qpin, size_t n, size_t nplace, int direct);
float vcg_qgraph_time(vcg_qgraph *g, const uint8_t *qpin, size_t n, size_t nplace, int reps);
void  vcg_qgraph_destroy(vcg_qgraph *g);
/* Allocation-free list decode: at most VCG_NOALLOC_MAX quads in device
 * memory, checked on the device (status 2 = out of range, 1 = corrupt, 0 =
 * ok into status_dev when given), whole atoms to out_dev, enqueued on
 * `stream` (a hipStream_t; NULL = the null stream).  Its scratch is a static
 * device array of vcg_noalloc_bytes() bytes whose address vcg_noalloc_setup
 * looks up (vcg_init calls it).  0, or -1 (n out of [1, VCG_NOALLOC_MAX],
 * no set-up on the current device, or a launch error). */
#define VCG_NOALLOC_MAX 64u
int    vcg_noalloc_setup(void);
size_t vcg_noalloc_bytes(void);
int    vcg_list_launch_noalloc(vcg_arc *a, size_t n, const int32_t *quads_dev, uint8_t *out_dev,
                               int32_t *status_dev, void *stream);
/* Container rewrite (vcg_repack.hip), synchronous on the calling thread's
 * stream.  A slot array (32768 slots) is taken as 32 planes of 1024 slots.
 * size: for the nblk slot arrays at archive offsets blk[], per plane p = 32k
 * + z: sums[3p .. 3p+2] = payload bytes, present atoms, zero atoms; returns 0,
 * 1 (a present payload range outside the archive) or -1 (out of device
 * memory).  write: for the ndata data regions, the v2 slot array from blk[k]
 * at dev_out + boff[k]; plane p's payloads placed from cdst[p] on in slot
 * order, its copy jobs numbered from cjob[p] (njobs in all); 0 or -1. */
int vcg_repack_size(vcg_arc *a, size_t nblk, const uint64_t *blk, uint64_t *sums);
int vcg_repack_write(vcg_arc *a, size_t ndata, const uint64_t *blk, const uint64_t *boff, const uint64_t *cdst,
                     const uint64_t *cjob, uint64_t njobs, uint8_t *dev_out);
/* Decode n raw payloads (no archive) at one base_q. */
int vcg_decode_payloads(size_t n, const uint8_t *pay, const uint64_t *offs, const uint32_t *lens,
                        const int32_t dq_step[VCG_NFREQ], const int32_t dq_bias[VCG_NFREQ],
                        uint8_t *out, int32_t *per_item);

/* ---- encode ---- */
typedef struct vcg_enc vcg_enc;
/* Encoders begun between these calls (on this thread) take their device
 * memory from one arena, released by vcg_enc_arena_close after their
 * vcg_enc_end. */
void vcg_enc_arena_open(void);
void vcg_enc_arena_close(void);
/* Encoders of nv LOD0 volumes (volume v: dims[v], nlods[v] levels) into
 * out[v]: each volume uploaded (or adopted) and its pyramid built on device,
 * one launch per level across the volumes. */
void vcg_enc_begin_many(int nv, const uint8_t *const *vols, int vol_on_device, const uint32_t (*dims)[3],
                        const int *nlods, vcg_enc **out);
/* An encoder over nlod level volumes already on the device (dims[l] = nx,
 * ny, nz of level l), encoded by vcg_enc_all like a pyramid. */
vcg_enc *vcg_enc_begin_levels(const uint8_t *const *lods, const uint32_t (*dims)[3], int nlod);
/* dst (device, dims dd) = src (host or device, dims sd <= dd) zero-padded;
 * synchronous.  0 / -1 */
int vcg_pad_volume(const uint8_t *src, const uint32_t sd[3], uint8_t *dst, const uint32_t dd[3]);
/* Rate-control samples of nv encoders in one launch: encoder v's ns[v] listed
 * LOD0 atoms (zyx triples) are transformed and their coefficients kept on
 * device (once per encoder); zero_masks[v][i] = 1 for its all-zero atoms. */
int  vcg_enc_samples(int nv, vcg_enc *const *es, const uint32_t *const *lists_zyx, const int *ns,
                     uint8_t *const *zero_masks);
/* Rate-control sample rows (candidate x atom) one launch takes: VCG_RC_ROWS,
 * or a smaller capacity set for tests (vcx_test_rc_rows). */
#define VCG_RC_ROWS 262144u
uint32_t vcg_rc_rows(void);
void vcg_set_rc_rows(uint32_t rows);
/* Byte totals of ng rate-control candidates in as few launches as fit:
 * candidate g = encoder es[g]'s sample at base_q qs[g].  bytes[g] (dc bytes
 * included), flags[g] bit 0 = coder bound exceeded, bit 1 = the total passed
 * budgets[g] and the candidate was abandoned (bytes[g] not final). */
int  vcg_enc_candidates(int ng, vcg_enc *const *es, const float *qs, const uint64_t *budgets,
                        uint64_t *bytes, int *flags);
/* Encode every atom of every LOD (LOD-major, raster) of nv encoders, encoder
 * v at qs[v][lod], in launches taken across volumes; lens[v][i] = payload
 * length of encoder v's atom i (0 = all-zero atom); ovf[v] = 1 if one of its
 * payloads exceeded the coder bound. */
int  vcg_enc_all(int nv, vcg_enc *const *es, const float (*qs)[VCG_NLOD], uint32_t *const *lens, int *ovf);
/* Assemble the archive image: head (header + directory, head_len bytes) at 0,
 * then for atom i with lens[i] > 0 its payload at pay_off[i] and, for every
 * atom with slot_off[i] != 0, its 16-byte slot at slot_off[i] (flags 1 with
 * off/len/dc for payloads, flags 2 for zero atoms).  Writes total_len bytes
 * to host out. */
int  vcg_enc_assemble(vcg_enc *e, const uint8_t *head, size_t head_len, const uint64_t *pay_off,
                      const uint64_t *slot_off, const uint32_t *lens, uint64_t total_atoms,
                      uint8_t *out, uint64_t total_len);
/* vcg_enc_assemble for many archives at once: one device image for all, one
 * upload of the per-atom offsets, one launch, and one pass of transfers back
 * into the archives' host buffers (out, total_len bytes each). */
typedef struct {
    vcg_enc *e;
    const uint8_t *head;
    size_t head_len;
    const uint64_t *pay_off, *slot_off;
    const uint32_t *lens;
    uint64_t total_atoms;
    uint8_t *out;
    uint64_t total_len;
} vcg_asm_job;
int  vcg_enc_assemble_many(int n, const vcg_asm_job *jobs);
void vcg_enc_end(vcg_enc *e);
/* Touch the pages of n host buffers (bufs[i], lens[i] bytes; NULL entries
 * skipped) on the library's host threads in the background, so that their
 * first-touch faults overlap device work.  The buffers must stay allocated
 * until vcg_host_prefault_wait, which this thread calls before it writes
 * them.  Does nothing when another job holds the threads. */
void vcg_host_prefault_begin(uint8_t *const *bufs, const size_t *lens, int n);
void vcg_host_prefault_wait(void);

/* Batched atom encode without a container (vcx_encode_atoms). */
int vcg_encode_atoms(size_t n, const uint8_t *vox, float base_q, uint8_t *out, size_t cap,
                     uint32_t *lens);
/* Device downsample for a volume already on device (LOD l -> l+1). */
int vcg_downsample(const void *dev_in, const uint32_t dims[3], void *dev_out);
/* the exporter's within-cell downscale (tools/downscale.h), method 0 box,
 * 1 cbox; on the calling thread's stream (sync: wait) */
int vcg_ds2x(const void *dev_in, const uint32_t dims[3], int method, float alpha, void *dev_out, int sync);
void vcg_test_wide_index(int on);   /* test hook: 64-bit index path of the 2x kernels */

/* ---- synthetic generator ---- */
typedef struct { int32_t axis, c0, tu, tv, a1, a2, w1, w2, p1, p2, half, inten; } vcg_sheet;
int vcg_synth(uint64_t seed, const uint32_t dims[3], const vcg_sheet *sheets, int nsheets, void *dev_out);

/* host-side fail-fast (implemented in host C: panic hook + abort) */
void vc__panic(const char *msg);

#ifdef __cplusplus
}
#endif
#endif /* VCG_H */
