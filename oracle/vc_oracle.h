/* vc_oracle.h — CPU restatement of the volume-compressor atom pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP library is compared
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product library (libvc.so) never links or calls it.
 *
 * Every function restates a piece of the reference (SuperOptimizer/
 * volume-compressor, src/vc/vc.c unless noted) and cites the lines it follows.
 * Functions with no reference counterpart (the reference has no vc_encode,
 * vc_decode_lod, rate control or synthetic generator) restate the definitions
 * the library pins for them; their parity is "unpinned" by reference output and
 * is instead checked against the reference's primitives they compose (the
 * v1 writer + tools/vc_repack.c layout, vc_test_downsample2x) and against
 * closed forms.
 */
#ifndef VC_ORACLE_H
#define VC_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VCO_A      32
#define VCO_A3     32768
#define VCO_NLOD   8
#define VCO_NFREQ  94          /* L1 frequency cz+cy+cx in 0..93 */
#define VCO_RC_CAP (VCO_A3 * 4 + 64)  /* range-coder scratch, vc.c:611 */

/* status codes, same values as vc.h */
enum { VCO_OK = 0, VCO_ERR_FORMAT = 1, VCO_ERR_RANGE = 2, VCO_ERR_OOM = 3, VCO_ERR_IO = 4 };

/* ---- tables (vc.c:75-88, 266-290, 462-468) ---- */
void        vco_tables(int32_t cmat[32][32], float gstep[VCO_NFREQ],
                       uint16_t scan[VCO_A3], uint8_t band_of_pos[VCO_A3]);
void        vco_dq_tables(float base_q, int32_t dq_step[VCO_NFREQ], int32_t dq_bias[VCO_NFREQ]);

/* ---- atom codec stages (for stage-by-stage GPU checks) ---- */
void vco_fdct(const uint8_t vox[VCO_A3], int32_t dc, int32_t coef[VCO_A3]);      /* vc.c:212-218 */
void vco_quantize(const int32_t coef[VCO_A3], float base_q, int16_t q[VCO_A3]);  /* vc.c:293-306 */
void vco_dequantize(const int16_t q[VCO_A3], float base_q, int32_t coef[VCO_A3]);/* vc.c:327-338 */
/* the quantizer on n coefficients of one frequency (exhaustive checks) */
void vco_quant_levels(float base_q, int freq, const int32_t *coef, size_t n, int32_t *out);
void vco_idct(const int32_t coef[VCO_A3], int32_t dc, uint8_t vox[VCO_A3]);      /* vc.c:219-229 */
/* Range-code the quantized levels (raster) → bytes; returns -1 on overflow of cap. vc.c:528-544,399 */
long vco_entropy_encode(const int16_t q[VCO_A3], uint8_t *out, size_t cap);
/* Decode levels (raster, zero-filled).  Returns VCO_OK, or VCO_ERR_FORMAT when an
 * Exp-Golomb prefix exceeds VCO_EG_PREFIX_CAP bits (the reference never returns
 * from such a payload).  *eob_out gets the clamped EOB.  vc.c:545-560 */
#define VCO_EG_PREFIX_CAP 64
int  vco_entropy_decode(const uint8_t *rc, size_t len, int16_t q[VCO_A3], uint32_t *eob_out);

/* Whole-atom encode: payload = [dc][rc bytes] (vc.c:595-622).  Returns payload
 * length, 0 for an all-zero atom (*is_zero = 1), or -1 on coder overflow. */
long vco_encode_atom(const uint8_t vox[VCO_A3], float base_q, uint8_t *payload,
                     size_t cap, int *is_zero);
/* Whole-atom decode (vc.c:625-633).  Returns VCO_OK / VCO_ERR_FORMAT (vox zeroed). */
int  vco_decode_payload(const uint8_t *pay, uint32_t len, float base_q, uint8_t vox[VCO_A3]);

/* ---- pyramid (vc.c:1562-1577) ---- */
void vco_downsample2x(const uint8_t *in, uint32_t nx, uint32_t ny, uint32_t nz, uint8_t *out);
/* the exporter's within-cell downscale (tools/downscale.h): method 0 box, 1 cbox */
void vco_ds2x(const uint8_t *in, uint32_t nx, uint32_t ny, uint32_t nz, int method, float alpha, uint8_t *out);
int  vco_nlod(uint32_t nx, uint32_t ny, uint32_t nz);                    /* vc.c:1001-1006 */
void vco_lod_dims(uint32_t nx, uint32_t ny, uint32_t nz, int lod,
                  uint32_t *ox, uint32_t *oy, uint32_t *oz);               /* vc.c:733-739 */
void vco_gather_atom(const uint8_t *vol, uint32_t nx, uint32_t ny, uint32_t nz,
                     uint32_t az, uint32_t ay, uint32_t ax, uint8_t atom[VCO_A3]);

/* ---- whole-volume encode (build-defined; composes the reference primitives) ----
 * vco_encode_q: explicit per-LOD base_q (already clamped by caller or not: the
 * function clamps to [0.05,4096] like vc_set_base_q, vc.c:1072-1073), writes a
 * v2 container laid out exactly as tools/vc_repack.c:96-270 would lay out the
 * v1 archive a single-threaded raster append would produce.  *out is malloc'd. */
int  vco_encode_q(const uint8_t *vol, uint32_t nx, uint32_t ny, uint32_t nz,
                  float target_ratio, const float base_q[VCO_NLOD],
                  uint8_t **out, size_t *out_len);
/* rate control: returns q0 (LOD0 base_q) for target_ratio, see vc_encode docs */
float vco_rate_control(const uint8_t *vol, uint32_t nx, uint32_t ny, uint32_t nz,
                       float target_ratio);
void vco_lod_schedule(float q0, int nlod, float base_q[VCO_NLOD]);
int  vco_encode(const uint8_t *vol, uint32_t nx, uint32_t ny, uint32_t nz,
                float target_ratio, uint8_t **out, size_t *out_len);

/* ---- reader (v1 + v2), vc.c:1285-1421, 1535-1555 ---- */
typedef struct vco_archive vco_archive;
vco_archive *vco_open(const uint8_t *buf, size_t len);
void vco_close(vco_archive *a);
int  vco_nlod_of(const vco_archive *a);
int  vco_decode_atom(vco_archive *a, int lod, int ax, int ay, int az, uint8_t out[VCO_A3]);
/* coverage: 0 ABSENT, 1 KNOWN_ZERO, 2 PRESENT (vc.c:1423-1430) */
/* canonical v2 image of any archive (tools/vc_repack.c); *out malloc'd */
int  vco_repack(const uint8_t *buf, size_t len, uint8_t **out, size_t *out_len);
int  vco_coverage(const vco_archive *a, int lod, uint32_t az, uint32_t ay, uint32_t ax,
                  uint64_t *off, uint32_t *len);
/* box = x0,y0,z0,x1,y1,z1 */
int  vco_decode_region(vco_archive *a, int lod, const uint32_t box[6], uint8_t *out);
int  vco_decode_lod(vco_archive *a, int lod, uint8_t *out, uint32_t dims[3]);
/* OpenMP decode of a list of LOD atoms (CPU baseline); returns seconds */
double vco_decode_atoms_omp(vco_archive *a, int lod, const uint32_t *atoms_zyx,
                            size_t n, int nthreads, uint8_t *out /* n*32768 or NULL */);

/* ---- synthetic scroll-like volume (build-defined, integer-exact) ---- */
void vco_synth(uint64_t seed, uint32_t nx, uint32_t ny, uint32_t nz, int nsheets, uint8_t *out);

#ifdef __cplusplus
}
#endif
#endif
