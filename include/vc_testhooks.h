/* vc_testhooks.h — the reference's test-only exports, kept for drop-in
 * callers (tests/test_realdata.c:27 and tools/vc_export_stream.c:108 declare
 * vc_test_downsample2x themselves; the reference builds both under
 * VC_TESTHOOKS, src/vc/vc.c:1560-1591).  Here they run on the GPU. */
#ifndef VC_TESTHOOKS_H
#define VC_TESTHOOKS_H

#include "vc.h"

#ifdef __cplusplus
extern "C" {
#endif

/* 2x box downsample (integer rounded mean of in-bounds voxels), malloc'd
 * result, *out = ceil(in/2) per axis.  src/vc/vc.c:1561-1577 */
uint8_t *vc_test_downsample2x(const uint8_t *vol, vc_dims in, vc_dims *out);

/* Encode one atom at q; returns the payload length (0 for an all-zero atom)
 * and optionally reconstructs it.  src/vc/vc.c:1581-1590 */
uint32_t vc_test_atom_rt(const uint8_t vox[VC_ATOM3], float q, uint8_t *recon);

#ifdef __cplusplus
}
#endif
#endif
