/* vc_gpu.h — MI355X extensions to the vc.h API: device-resident decode,
 * batched atom decode, explicit-q encode, stage timing hooks and the seeded
 * synthetic-volume generator used by the benchmarks.
 *
 * None of these has a counterpart in the reference (SuperOptimizer/
 * volume-compressor ships a CPU-only vc.h); they are the batching/serving
 * surface layered beside the drop-in API.  All device pointers are plain
 * `void *` HBM addresses; all work is enqueued on the calling host thread's
 * own HIP stream (vcx_stream), so calls from different host threads overlap.
 */
#ifndef VC_GPU_H
#define VC_GPU_H

#include "vc.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Select/initialise the HIP device (idempotent).  VC_ERR_IO if no usable GPU. */
vc_status vcx_device_init(void);

/* The calling thread's HIP stream (a hipStream_t), created on first use. */
void *vcx_stream(void);

/* Device memory helpers (hipMalloc / hipFree / synchronous copies). */
void     *vcx_dev_alloc(size_t bytes);
void      vcx_dev_free(void *p);
vc_status vcx_h2d(void *dst_dev, const void *src_host, size_t bytes);
vc_status vcx_d2h(void *dst_host, const void *src_dev, size_t bytes);
/* Device-to-device copy by a plain streaming kernel (16-byte aligned
 * pointers, bytes a multiple of 16; else VC_ERR_RANGE): a known-size
 * reference for hardware byte counters. */
vc_status vcx_copy_kernel(void *dst_dev, const void *src_dev, size_t bytes);

/* Decode a whole LOD / a voxel box into device memory (raster of the LOD /
 * of the box).  sync=1: wait and return the status like vc_decode_lod.
 * sync=0: enqueue only and return VC_OK; vcx_sync() later returns the status
 * accumulated since the previous vcx_sync(). */
vc_status vcx_decode_lod_device(vc_archive *a, int lod, void *dev_out, int sync);
vc_status vcx_decode_region_device(vc_archive *a, int lod, vc_box box, void *dev_out, int sync);
vc_status vcx_sync(void);

/* Batched random-access decode: n atoms given as (lod, ax, ay, az) quads;
 * out receives n full 32^3 atoms (host memory, atom i at i*VC_ATOM3), the
 * same bytes vc_decode_atom returns.  per_atom (optional) gets each status.
 * Returns VC_ERR_RANGE if any atom is out of range (none decoded), else the
 * first non-OK per-atom status. */
vc_status vcx_decode_atoms(vc_archive *a, size_t n, const int32_t *lod_ax_ay_az,
                           uint8_t *out, vc_status *per_atom);

/* ---- allocation-free batched atom decode ----
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): vcx_decode_atoms with inputs and outputs in device
 * memory, enqueued on the caller's stream, calling no hipMalloc, hipFree or
 * other allocating routine.  The decoder's scratch is a static device array
 * of VCX_NOALLOC_MAX_ATOMS atoms (about 8.4 MB per GPU, held until the
 * process ends); larger n is VC_ERR_RANGE (loop over chunks on one stream).
 * Every call on a GPU shares that array: calls that overlap on one GPU (on
 * different streams, or from different threads) race and must be ordered by
 * the caller.
 * quads_dev: n (lod, ax, ay, az) quads in device memory, checked on the
 * device; out_dev: n * VC_ATOM3 bytes, atom i at i * VC_ATOM3, the bytes
 * vc_decode_atom returns; status_dev (optional, n int32): 0 OK, 1 corrupt
 * payload (atom zeroed; vc_decode_atom's VC_ERR_FORMAT), 2 quad out of range
 * (atom zeroed; VC_ERR_RANGE).  stream: a hipStream_t (NULL: the null stream)
 * of the library's device.  Returns after enqueueing: VC_OK, or VC_ERR_RANGE
 * for NULL pointers, n > VCX_NOALLOC_MAX_ATOMS, a streaming archive, or an
 * archive without its device copy yet (any earlier device decode of the
 * archive, e.g. one vc_decode_atom, makes it). */
#define VCX_NOALLOC_MAX_ATOMS 64
vc_status vcx_decode_atoms_noalloc(vc_archive *a, size_t n, const int32_t *quads_dev, uint8_t *out_dev,
                                   int32_t *status_dev, void *stream);

/* Stage events: while set, each launch pair of the calling thread's decode
 * calls (vc_decode_*, vcx_decode_*) records events[0] before its entropy
 * kernel, events[1] between the entropy and transform kernels and events[2]
 * after the transform kernel, on the stream those kernels run on (one of the
 * thread's scratch-slot streams, not vcx_stream()).  Server batches (other
 * thread, captured launches) record none.  events are hipEvent_t handles
 * owned by the caller; NULL disables. */
void vcx_set_stage_events(void *events[3]);

/* Whole-volume encode with explicit per-LOD base_q (clamped to [0.05,4096]
 * as vc_set_base_q does) instead of rate control.  Same archive layout as
 * vc_encode. */
vc_status vcx_encode_q(const uint8_t *vol, vc_dims dims, float target_ratio,
                       const float base_q[VC_NLOD], uint8_t **out_archive, size_t *out_len);
/* vc_encode from a volume already resident in device memory. */
vc_status vcx_encode_device(const void *dev_vol, vc_dims dims, float target_ratio,
                            uint8_t **out_archive, size_t *out_len);
/* Batched encode of n independent volumes (all host or all device memory,
 * any dims): archive i is byte-identical to vc_encode(vols[i], dims[i],
 * target_ratios[i]).  Volumes are encoded in groups that share launches
 * (rate-control rounds, transform, quantization and range coding across
 * volumes) while their device footprint fits a fixed budget.  outs[i] gets a
 * malloc buffer (caller frees) and out_lens[i] its length, statuses[i] the
 * status vc_encode would return (outs[i] NULL unless VC_OK).  Returns VC_OK,
 * or the first failing status. */
vc_status vcx_encode_batch(size_t n, const uint8_t *const *vols, int vols_on_device, const vc_dims *dims,
                           const float *target_ratios, uint8_t **outs, size_t *out_lens,
                           vc_status *statuses);
/* Test hook: cap the rate-control rows (candidate x sample atom) per launch
 * (0 = default, 262144; at least 512).  Smaller caps make vcx_encode_batch
 * evaluate fewer bisection levels per launch and split rounds over more
 * launches; archives do not change. */
void      vcx_test_rc_rows(uint32_t rows);
/* The LOD0 base_q that vc_encode's rate control picks for this volume. */
float     vcx_rate_control(const uint8_t *vol, vc_dims dims, float target_ratio);

/* Batched single-atom codec (no container): encode n 32^3 atoms at base_q
 * into payloads ([dc][range-coder bytes], the bytes a v1/v2 slot points at).
 * payload i is written at out + i*cap_per_atom, its length to lens[i]
 * (0 = all-zero atom, stored as KNOWN_ZERO).  VC_ERR_RANGE if a payload
 * would exceed cap_per_atom or the reference coder's 131136-byte bound. */
vc_status vcx_encode_atoms(size_t n, const uint8_t *vox, float base_q,
                           uint8_t *out, size_t cap_per_atom, uint32_t *lens);
/* Decode n payloads (lens[i] bytes at pay + offs[i]) at base_q into n atoms. */
vc_status vcx_decode_payloads(size_t n, const uint8_t *pay, const uint64_t *offs,
                              const uint32_t *lens, float base_q, uint8_t *out,
                              vc_status *per_atom);

/* Seeded synthetic scroll-like volume (background ~N(12,6), nsheets warped
 * sheets of intensity 110..190 and thickness 3..6, 7-tap sigma=1 blur,
 * additive ~N(0,8) noise), integer-exact.  dev_out: nx*ny*nz bytes of device
 * memory. */
vc_status vcx_synth_device(uint64_t seed, vc_dims dims, int nsheets, void *dev_out);

/* ---- the exporter's within-cell 2x downscale ----
 * Replaces ds_downscale2x of tools/downscale.h:48-93 (a header used by the
 * exporter tools/vc_export_stream.c:429-458,618) for device-resident data:
 * dev_out (ceil(nx/2) x ceil(ny/2) x ceil(nz/2) bytes) receives each output
 * voxel from its own 2x2x2 cell of dev_in, the odd boundary edge-replicated,
 * by method VCX_DS_BOX (mean) or VCX_DS_CBOX (the mean pushed by alpha, 0..1,
 * toward the cell's voxel of largest deviation); an all-zero cell gives 0.
 * Float arithmetic in source order without contraction; the reference's
 * own flags (-ffast-math -ffp-contract=fast) give the same bytes on every
 * golden case (checked against builds at the baseline ISA and with FMA).
 * sync: wait; otherwise ordered on vcx_stream().
 * VC_ERR_RANGE for NULL pointers, a zero dimension, an unknown method or a
 * NaN alpha. */
enum { VCX_DS_BOX = 0, VCX_DS_CBOX = 1 };
vc_status vcx_downscale2x_device(const void *dev_in, vc_dims in, int method, float alpha, void *dev_out, int sync);
/* Test hook, no counterpart in the reference: nonzero makes the 2x kernels
 * (this downscale, vc_test_downsample2x and the encoder's pyramid) index
 * their output with 64-bit integers, the path they otherwise take only past
 * about 4.27e9 output voxels, so small volumes exercise it.  Process-wide;
 * 0 restores the default. */
void vcx_test_wide_index(int on);

/* ---- the exporter's pyramid cascade ----
 * Replaces the compute path of tools/vc_export_stream.c (a command-line tool
 * in the reference: zarr/S3 source, pyramid cascade :395-466, tile worker
 * :473-593, coarse tail :595-626, main :658-826) for a u8 volume in host or
 * device memory (src dims, x fastest) whose source chunks (chunk = cx, cy, cz
 * voxels; NULL: 128^3) are all present.  Writes the v1 archive at path:
 * dims rounded up to multiples of 1024 (zero padding); base_q[l] = q0 *
 * q_falloff^l, raised to 0.05 before each level; regions past the source
 * marked zero; per unit (xy tile of `tile` voxels, 0: 1024; z band of 128
 * voxels) in the tool's unit order, LOD0..2 from the within-cell 2x
 * downscale (method VCX_DS_BOX / VCX_DS_CBOX, alpha as vcx_downscale2x_device)
 * appended slab by slab as the tool's cascade fills them, all-zero atoms
 * marked zero, bands past the source's chunk grid stamped zero; the coarser
 * LODs from the archive's own decoded 2x2x2 parents.  The file equals what
 * the reference writer produces driven in that order (the tool run with one
 * compute thread).  The pyramid, the atom encode and the tail's decode and
 * downscale run on the GPU.  VC_ERR_RANGE for bad arguments, VC_ERR_IO if
 * the file cannot be created, VC_ERR_OOM. */
vc_status vcx_export_volume(const uint8_t *vol, vc_dims src, const uint32_t chunk[3], const char *path, float q0,
                            float q_falloff, int method, float alpha, uint32_t tile);

/* ---- streaming archives decoded from HBM ----
 * Extends vc_open_streaming (src/vc/vc.h:108-118; src/vc/vc.c:1325-1338):
 * for a v2 archive, every decode call (vc_decode_atom, vc_decode_region,
 * vc_decode_lod, vcx_decode_atoms and the device variants) first makes the
 * regions it touches resident in a region-blob cache in HBM, fetching each
 * missing region blob with ONE call of the read function
 * (vc_region_blob_range's range), then decodes there.  Results are the bytes
 * the in-memory archive gives.  Regions that cannot be served from their
 * blob (a payload outside it, a failing read) and v1 archives keep the
 * per-atom read path.  vcx_stream_cache sets the cache size in bytes (0:
 * the archive size, at most 8 GiB) and empties the cache; call it when no
 * other thread decodes from the archive.  A call whose blobs exceed the
 * cache uses the per-atom path.  Decode calls on one streaming archive
 * serialize their cache updates and launches; the read function is called
 * with that lock held, so it must not decode from the same archive.
 * vcx_stream_stats: blob reads and cache evictions so far. */
vc_status vcx_stream_cache(vc_archive *a, size_t bytes);
vc_status vcx_stream_stats(const vc_archive *a, uint64_t *blob_reads, uint64_t *evictions);

/* ---- container rewrite into the canonical v2 layout ----
 * Replaces tools/vc_repack.c (main, :184-291; a command-line tool in the
 * reference, no library entry point): the same bytes that tool writes for
 * the same input.  a: a v1 or v2 archive opened from memory (vc_open); its
 * device copy is used.  Regions go coarse-LOD-first, each data region as one
 * blob [512 KiB slot array][payloads in slot order] after the 4 KiB-aligned
 * directory; payloads are copied verbatim.  Errors: VC_ERR_RANGE for NULL
 * arguments or a streaming archive, VC_ERR_FORMAT for a present payload
 * range outside the archive (the tool copies it unchecked), VC_ERR_OOM.
 * vcx_repack returns a malloc'd image (the caller free()s it).
 * vcx_repack_device writes it to dev_out; *out_len receives the image size,
 * and a dev_out that is NULL or smaller (cap) gets VC_ERR_RANGE and nothing
 * written. */
vc_status vcx_repack(vc_archive *a, uint8_t **out, size_t *out_len);
vc_status vcx_repack_device(vc_archive *a, void *dev_out, size_t cap, size_t *out_len);

/* ---- serving: device-resident request loop with continuous batching ----
 * No counterpart in the reference (its vc_decode_atom / vc_decode_region are
 * one call per request); callers are request servers that want many
 * concurrent atom and region requests decoded together.
 *
 * vcx_server_create keeps the archive on the device and starts a worker
 * thread.  device_bytes bounds the device memory of the batches in flight:
 * three batch slots share it at vcx_server_atom_bytes() per atom of slot
 * capacity (output, request staging and decoder scratch, about 160 KiB),
 * all of it allocated before create returns (NULL when it does not fit).
 * device_bytes = 0 admits by free device memory: half of what hipMemGetInfo
 * reports free at create, at most full batches (32768 atoms) in every slot.
 * Requests are admitted in arrival order as slots free up; batches overlap on
 * the device.  A request larger than one slot is not rejected but served in
 * parts over consecutive batches and reported once, when its last part
 * completes: an atom list in chunks of a slot's capacity, a box in z-slabs
 * of whole atom layers (rejected only if one atom layer of the box exceeds a
 * slot).  Submit returns a request id (>= 0) or -1 (invalid request).  Atom
 * requests: quads = (lod, ax, ay, az) x n, out receives n atoms of VC_ATOM3
 * bytes; region requests: out receives the box raster as vc_decode_region
 * writes it.  out may be host or device memory (device memory: the result
 * stays on the device, no PCIe copy) and must stay valid until the request
 * completes; for host memory, pinned memory (vcx_host_alloc) keeps the
 * result copy asynchronous.  A request decodes to the same bytes alone or in
 * any mix.
 * vcx_server_poll returns up to max completions (id, status as vc_status,
 * completion time on CLOCK_MONOTONIC in ns, see vcx_now_ns), waiting up to
 * wait_ms for the first one.  Destroy waits for every submitted request. */
typedef struct vcx_server vcx_server;
vcx_server *vcx_server_create(vc_archive *a, size_t device_bytes);
void        vcx_server_destroy(vcx_server *s);
int64_t     vcx_server_submit_atoms(vcx_server *s, size_t n, const int32_t *quads, uint8_t *out);
int64_t     vcx_server_submit_region(vcx_server *s, int lod, vc_box box, uint8_t *out);
size_t      vcx_server_poll(vcx_server *s, int64_t *ids, int32_t *status, uint64_t *t_done_ns, size_t max,
                            int wait_ms);
size_t      vcx_server_capacity(const vcx_server *s);   /* atoms per batch slot */
/* A completion with its timing, for the breakdown of a request's latency:
 * submitted, admitted into a batch (for a request served in parts, its first
 * batch) and completed, on CLOCK_MONOTONIC in ns (vcx_now_ns); the LOD class
 * of the batch that completed it (0: LOD0 only, 1: a coarser LOD) and that
 * batch's atom count.  vcx_server_poll_served takes completions from the same
 * queue as vcx_server_poll, with these fields.  No counterpart in the
 * reference (it has no server). */
typedef struct {
    int64_t id;
    int32_t status;
    int32_t cls;
    uint64_t t_submit_ns, t_admit_ns, t_done_ns;
    uint64_t batch_atoms;
} vcx_served;
size_t      vcx_server_poll_served(vcx_server *s, vcx_served *out, size_t max, int wait_ms);
/* Batches run as persistent kernels (entropy waves claiming 64 atoms at a
 * time, transform workgroups one atom at a time, the batch size read from
 * device memory), captured once per batch slot, LOD class (LOD0 only / any
 * coarser LOD) and size bucket (up to 64, 512, 4096, 32768 atoms) as a
 * hipGraph.  A batch replays its graph when the server is saturated (8 or
 * more requests still queued), where replays of concurrent slots run one
 * after another; otherwise it launches the same kernels directly, so a short
 * LOD0 batch overlaps a long coarse one.  vcx_server_autotune waits until no batch
 * is in flight, then times every candidate launch shape per class and bucket
 * (entropy workgroups of 1, 2 or 4 waves, enough of them for all atoms at
 * once or for half, 256 / 512 / 1024 transform workgroups or one per atom)
 * on sample batches of the archive's own atoms, keeps the fastest and
 * rebuilds the captured launches; results do not depend on the shape.
 * vcx_server_tuning writes the table (one line per class and bucket: class,
 * bucket, shape, default ms, tuned ms) into buf (NUL-terminated, at most
 * cap bytes) and returns its full length. */
vc_status   vcx_server_autotune(vcx_server *s);
/* Test hook: serve every later batch with one launch shape (ent_waves 1, 2
 * or 4; any positive group counts), waiting for the batches in flight. */
vc_status   vcx_test_server_shape(vcx_server *s, int ent_waves, uint32_t ent_groups, uint32_t idct_groups);
size_t      vcx_server_tuning(vcx_server *s, char *buf, size_t cap);
size_t      vcx_server_atom_bytes(void);                /* device bytes per atom of capacity */
uint64_t    vcx_now_ns(void);
void       *vcx_host_alloc(size_t bytes);               /* pinned host memory */
void        vcx_host_free(void *p);

#ifdef __cplusplus
}
#endif
#endif /* VC_GPU_H */
