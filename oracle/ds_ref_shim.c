/* ds_ref_shim.c — exports the reference's within-cell downscale
 * (tools/downscale.h, a header of static inline functions) from a shared
 * library built by oracle/ref.mk, so the tests can compare against it.  Test
 * infrastructure only; the header is included from the reference tree. */
#include <stdint.h>
#include "downscale.h"

void ref_ds2x(const uint8_t *in, int nx, int ny, int nz, int method, float alpha, uint8_t *out) {
    int ox, oy, oz;
    ds_downscale2x(in, nx, ny, nz, out, &ox, &oy, &oz, method ? DS_CBOX : DS_BOX, alpha);
}
