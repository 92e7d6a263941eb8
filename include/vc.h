/* vc.h — volume-compressor drop-in C API, MI355X build.
 *
 * The declarations below are the reference's public surface with unchanged
 * signatures, types and constants (SuperOptimizer/volume-compressor
 * src/vc/vc.h), plus the two whole-volume entry points the reference sketches
 * but does not ship (vc_encode, vc_decode_lod: README.md:51-58, plan.txt:206-211).
 * Each declaration cites the reference interface it replaces.
 *
 * Archive format: identical to the reference's v1 (sparse hash index) and v2
 * (region directory) containers; vc_encode writes v2 byte-identically to what
 * the reference's v1 writer followed by tools/vc_repack.c produces for the same
 * atoms and base_q.
 *
 * Execution: vc_encode, vc_decode_lod, vc_decode_atom and vc_decode_region run
 * their codec work in HIP kernels on the current GPU (gfx950).  There is no CPU
 * codec fallback: without a usable HIP device these calls report the failure
 * through the panic hook (vc_set_panic_hook) and abort, like the reference's
 * other fail-fast paths.  Header-only queries (vc_open, vc_lod_dims,
 * vc_atom_coverage, vc_atom_payload_range, vc_archive_version,
 * vc_region_blob_range) run on the host and need no GPU.
 */
#ifndef VC_H
#define VC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/vc/vc.h:15-21 */
typedef enum {
    VC_OK = 0,
    VC_ERR_FORMAT = 1,
    VC_ERR_RANGE = 2,
    VC_ERR_OOM = 3,
    VC_ERR_IO = 4,
} vc_status;

/* src/vc/vc.h:23-25 — logical extent in voxels, x fastest, z slowest */
typedef struct {
    uint32_t nx, ny, nz;
} vc_dims;

/* src/vc/vc.h:27-30 — [x0,x1) x [y0,y1) x [z0,z1) */
typedef struct {
    uint32_t x0, y0, z0;
    uint32_t x1, y1, z1;
} vc_box;

/* src/vc/vc.h:32-35 */
#define VC_ATOM    32u
#define VC_ATOM3   (32u*32u*32u)
#define VC_NLOD    8
#define VC_MAX_DIM (1u<<20)

/* src/vc/vc.h:40 — ABSENT: never written; KNOWN_ZERO: written, all zero */
typedef enum { VC_ABSENT = 0, VC_KNOWN_ZERO = 1, VC_PRESENT = 2 } vc_cover;

/* ===================================================================== */
/* Whole-volume encode (README.md:51-52, plan.txt:206-207)               */
/* ===================================================================== */
/* Encode a u8 volume (raster z,y,x; x fastest) into one in-memory v2 archive
 * holding every LOD of the strict 2x pyramid (ceil halving; the number of LODs
 * follows vc_create's rule, src/vc/vc.c:1001-1006).  LOD l+1 is the integer
 * rounded mean of the in-bounds voxels of each 2x2x2 cell of LOD l
 * (src/vc/vc.c:1562-1577).  base_q is chosen by deterministic rate control so
 * that the LOD0 calibration sample reaches target_ratio (present-atom voxels
 * per payload byte); coarser LODs use q*0.6^l.  *out_archive is malloc'd; the
 * caller free()s it.  Errors: VC_ERR_RANGE for invalid dims, target_ratio < 1
 * or NULL pointers; VC_ERR_OOM when memory runs out. */
vc_status vc_encode(const uint8_t *vol, vc_dims dims, float target_ratio,
                    uint8_t **out_archive, size_t *out_len);

/* ===================================================================== */
/* Writer (src/vc/vc.h:42-97) — appendable v1 archive on disk            */
/* ===================================================================== */
typedef struct vc_writer vc_writer;

vc_writer *vc_create(const char *path, vc_dims lod0_dims, float target_ratio); /* vc.h:51 */
vc_status vc_set_base_q(vc_writer *w, int lod, float q);                          /* vc.h:58 */
vc_status vc_append_atom(vc_writer *w, int lod, uint32_t az, uint32_t ay, uint32_t ax,
                         const uint8_t vox[VC_ATOM3]);                            /* vc.h:61-62 */
vc_status vc_append_box(vc_writer *w, int lod, vc_box voxel_box, const uint8_t *voxels); /* vc.h:68 */
vc_status vc_mark_zero_atom(vc_writer *w, int lod, uint32_t az, uint32_t ay, uint32_t ax); /* vc.h:73 */
vc_status vc_mark_zero_region(vc_writer *w, int lod, uint32_t rz, uint32_t ry, uint32_t rx); /* vc.h:74 */
vc_status vc_writer_decode_atom(vc_writer *w, int lod, uint32_t az, uint32_t ay, uint32_t ax,
                                uint8_t out[VC_ATOM3]);                           /* vc.h:79-80 */
vc_cover  vc_writer_coverage(vc_writer *w, int lod, uint32_t az, uint32_t ay, uint32_t ax); /* vc.h:81 */
vc_status vc_writer_decode_2x2x2(vc_writer *w, int lod, uint32_t az0, uint32_t ay0, uint32_t ax0,
                                 uint8_t out8[8][VC_ATOM3], int *any);            /* vc.h:86-87 */
void      vc_writer_close(vc_writer *w);                                          /* vc.h:90 */
uint32_t  vc_region_atoms(void);                                                  /* vc.h:94 */

/* ===================================================================== */
/* Reader (src/vc/vc.h:99-153)                                           */
/* ===================================================================== */
typedef struct vc_archive vc_archive;

/* Borrows the buffer: it must outlive the handle (vc.h:104-106).  The first
 * GPU decode call copies it into device memory once; later calls decode from
 * that device copy. */
vc_archive *vc_open(const uint8_t *archive, size_t len);                          /* vc.h:106 */

typedef vc_status (*vc_read_fn)(void *ud, uint64_t off, uint32_t len, uint8_t *dst); /* vc.h:112 */
vc_archive *vc_open_streaming(vc_read_fn read, void *ud, uint64_t total_len);     /* vc.h:118 */

void        vc_close(vc_archive *a);                                              /* vc.h:120 */
vc_status   vc_lod_dims(const vc_archive *a, int lod, vc_dims *out);              /* vc.h:123 */

/* Decode every atom of a LOD into out_vol (nx*ny*nz bytes of that LOD, raster),
 * and set *out_dims (may be NULL).  README.md:58.  Same status rules as
 * vc_decode_region over the whole LOD box. */
vc_status vc_decode_lod(vc_archive *a, int lod, uint8_t *out_vol, vc_dims *out_dims);

vc_status vc_decode_atom(vc_archive *a, int lod, int ax, int ay, int az,
                         uint8_t out[VC_ATOM3]);                                  /* vc.h:126-127 */
vc_status vc_decode_region(vc_archive *a, int lod, vc_box box, uint8_t *out);     /* vc.h:130 */
vc_cover  vc_atom_coverage(const vc_archive *a, int lod,
                           uint32_t az, uint32_t ay, uint32_t ax);                 /* vc.h:133-134 */
vc_cover  vc_atom_payload_range(const vc_archive *a, int lod,
                                uint32_t az, uint32_t ay, uint32_t ax,
                                uint64_t *off, uint32_t *len);                     /* vc.h:140-142 */
int       vc_archive_version(const vc_archive *a);                                /* vc.h:146 */
vc_cover  vc_region_blob_range(const vc_archive *a, int lod,
                               uint32_t rz, uint32_t ry, uint32_t rx,
                               uint64_t *off, uint64_t *len);                      /* vc.h:151-153 */

void vc_set_panic_hook(void (*hook)(const char *msg));                            /* vc.h:156 */

#ifdef __cplusplus
}
#endif
#endif /* VC_H */
