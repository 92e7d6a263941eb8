/* vc_host.h — internal declarations of the plain-C host library. */
#ifndef VC_HOST_H
#define VC_HOST_H

#include <pthread.h>
#include <stdint.h>
#include <string.h>

#include "vc.h"
#include "vc_gpu.h"
#include "../vcg.h"

#define VC_A          32u
#define VC_A3         32768u
#define R_ATOMS       32u
#define R3            (R_ATOMS * R_ATOMS * R_ATOMS)
#define ATOM_SLOT     16u
#define L2_BLOCK      ((uint64_t)R3 * ATOM_SLOT)   /* 512 KiB slot array */
#define FILE_HDR      128u
#define VC_MAGIC      0x00314356u                  /* "VC1\0" */
#define AF_ABSENT     0u
#define AF_PRESENT    1u
#define AF_ZERO       2u
#define ZERO_REGION   1ull
#define SLOT_ABSENT   0xFFFFFFFFFFFFFFFFull

/* header field offsets (src/vc/vc.c:674-718) */
#define FH_MAGIC 0
#define FH_VER 4
#define FH_ATOM 8
#define FH_LAYOUT 12
#define FH_RGN 16
#define FH_NLOD 20
#define FH_NX 24
#define FH_NY 28
#define FH_NZ 32
#define FH_TRATIO 36
#define FH_L1CAP 40
#define FH_L1CNT 48
#define FH_L1OFF 56
#define FH_CURSOR 64
#define FH_BASEQ 72
#define FH2_DIROFF 40
#define FH2_DIRCNT 48
#define FH2_TOTLEN 56

static inline uint32_t rd32(const uint8_t *p) { uint32_t v; memcpy(&v, p, 4); return v; }
static inline uint64_t rd64(const uint8_t *p) { uint64_t v; memcpy(&v, p, 8); return v; }
static inline float    rdf(const uint8_t *p) { float v; memcpy(&v, p, 4); return v; }
static inline void     wr32(uint8_t *p, uint32_t v) { memcpy(p, &v, 4); }
static inline void     wr64(uint8_t *p, uint64_t v) { memcpy(p, &v, 8); }
static inline void     wrf(uint8_t *p, float v) { memcpy(p, &v, 4); }

/* tables + device bring-up */
const vcg_tables *vc__tables(void);
void vc__dq(float base_q, int32_t step[VCG_NFREQ], int32_t bias[VCG_NFREQ]);
void vc__require_gpu(void);        /* panics without a usable HIP device */
float vc__clamp_q(float q);

/* geometry */
vc_dims vc__lod_dims(vc_dims d0, int lod);
void    vc__region_grid(vc_dims d0, int lod, uint32_t *nrz, uint32_t *nry, uint32_t *nrx);
int     vc__nlod(vc_dims d0);
int     vc__dims_valid(vc_dims d);
uint64_t vc__mix64(uint64_t x);
uint64_t vc__region_key(int lod, uint32_t rz, uint32_t ry, uint32_t rx);

/* reader */
struct vc_archive {
    const uint8_t *buf; size_t len;     /* buf == NULL: streaming */
    vc_read_fn read; void *ud;
    vc_dims dims0; int nlod; int version;
    uint64_t l1_off, l1_cap, dir_off, region_prefix[VC_NLOD + 1];
    float base_q[VC_NLOD];
    pthread_mutex_t mu;
    vcg_arc *dev;                       /* device copy, created on first GPU decode */
    void *sdev;                         /* streaming v2: region-blob arena in HBM (vc_stream.c) */
    int sdev_failed;
    uint64_t stream_cache;              /* arena bytes requested (0: default) */
};
/* a region's slot array (src/vc/vc.c:1349-1360,1376-1400): AF_PRESENT with
 * its offset in *blk, AF_ZERO (all-zero region) or AF_ABSENT (also for a
 * block failing the corrupt guard) */
int vc__region_block(const vc_archive *a, int lod, uint32_t rz, uint32_t ry, uint32_t rx, uint64_t *blk);
/* resolve (lod,az,ay,ax) -> slot flags, payload range (src/vc/vc.c:1365-1402) */
uint8_t vc__resolve(const vc_archive *a, int lod, uint32_t az, uint32_t ay, uint32_t ax,
                    uint64_t *off, uint32_t *len);
int     vc__rd(const vc_archive *a, uint64_t off, uint32_t len, uint8_t *dst);
vcg_arc *vc__arc_dev(vc_archive *a);   /* NULL on allocation failure */
vc_status vc__region_check(const vc_archive *a, int lod, vc_box b);
int     vc__atom_ok(const vc_archive *a, int lod, int ax, int ay, int az);
/* streaming v2 archives: the arena holding every region the call touches
 * (a box of LOD lod, or n quads), with the arena lock held until
 * vc__stream_end; NULL: decode through the host-resolved path instead */
vcg_arc *vc__stream_begin(vc_archive *a, int lod, const vc_box *box, size_t n, const int32_t *quads);
void     vc__stream_end(vc_archive *a);
void     vc__stream_free(vc_archive *a);
/* host-resolved batched decode of whole atoms (streaming archives, writer views) */
vc_status vc__decode_atoms_host(vc_archive *a, size_t n, const int32_t *quads, uint8_t *out,
                                vc_status *per_atom);

/* writer internals for the exporter cascade (vc_export.c) */
vc_status vc__writer_store(vc_writer *w, int lod, uint32_t az, uint32_t ay, uint32_t ax, const uint8_t *pay,
                           uint32_t len);
vc_status vc__writer_decode_atoms(vc_writer *w, size_t n, const int32_t *quads, uint8_t *out);
int       vc__writer_nlod(const vc_writer *w);
float     vc__writer_q(vc_writer *w, int lod);

#endif
