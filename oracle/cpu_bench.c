/* cpu_bench.c — CPU baseline timer (test/bench infrastructure, not product).
 *
 * Times the reference library's own vc_decode_atom (oracle/_ref/libvcref_*.so,
 * loaded with dlopen) over a list of atoms with an OpenMP parallel-for
 * (schedule dynamic), the CPU path the reference's README describes
 * (README.md:24; the reference ships no OpenMP loop itself).  The archive is
 * the same bytes the GPU decodes.
 */
#include <dlfcn.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <time.h>

typedef void *(*open_fn)(const uint8_t *, size_t);
typedef void (*close_fn)(void *);
typedef int (*dec_fn)(void *, int, int, int, int, uint8_t *);

/* returns seconds (negative on load failure); *bad counts non-OK statuses;
 * if out != NULL atom i is stored at out + i*32768 */
double cpu_decode_atoms(const char *libpath, const uint8_t *arc, size_t len, int lod,
                        const int32_t *zyx, size_t n, int nthreads, long *bad, uint8_t *out) {
    void *h = dlopen(libpath, RTLD_NOW | RTLD_LOCAL);
    if (!h) return -1.0;
    open_fn vopen = (open_fn)dlsym(h, "vc_open");
    close_fn vclose = (close_fn)dlsym(h, "vc_close");
    dec_fn vdec = (dec_fn)dlsym(h, "vc_decode_atom");
    if (!vopen || !vclose || !vdec) { dlclose(h); return -2.0; }
    void *a = vopen(arc, len);
    if (!a) { dlclose(h); return -3.0; }
    if (nthreads > 0) omp_set_num_threads(nthreads);
    long nbad = 0;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    #pragma omp parallel reduction(+:nbad)
    {
        uint8_t *tmp = (uint8_t *)malloc(32768);
        #pragma omp for schedule(dynamic, 4)
        for (long i = 0; i < (long)n; ++i) {
            uint8_t *dst = out ? out + (size_t)i * 32768 : tmp;
            if (vdec(a, lod, zyx[3 * i + 2], zyx[3 * i + 1], zyx[3 * i], dst) != 0) nbad++;
        }
        free(tmp);
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    vclose(a);
    dlclose(h);
    if (bad) *bad = nbad;
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* Open-loop request serving on the host cores, the CPU counterpart of the
 * GPU request server.  Request r arrives t_arr[r] seconds after the start;
 * it is either kind 0, nat[r] atoms given as (lod, ax, ay, az) quads from
 * quads + 4 * qoff[r], decoded with the reference's vc_decode_atom one after
 * another, or kind 1, the box boxes[7r .. 7r+6] = (lod, x0, y0, z0, x1, y1,
 * z1) decoded with its vc_decode_region into a buffer of the box's size.  A
 * pool of nthreads workers takes requests in arrival order (each waits for
 * its request's arrival) and writes t_done[r] (seconds after the start).
 * Returns the wall seconds until the last completion (negative on load
 * failure); *bad counts non-OK statuses. */
typedef struct { uint32_t x0, y0, z0, x1, y1, z1; } box6;
typedef int (*reg_fn)(void *, int, box6, uint8_t *);

static double now_s(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

double cpu_serve(const char *libpath, const uint8_t *arc, size_t len, size_t nreq, const int32_t *kind,
                 const int64_t *qoff, const int32_t *nat, const int32_t *quads, const int32_t *boxes,
                 const double *t_arr, int nthreads, double *t_done, long *bad) {
    void *h = dlopen(libpath, RTLD_NOW | RTLD_LOCAL);
    if (!h) return -1.0;
    open_fn vopen = (open_fn)dlsym(h, "vc_open");
    close_fn vclose = (close_fn)dlsym(h, "vc_close");
    dec_fn vdec = (dec_fn)dlsym(h, "vc_decode_atom");
    reg_fn vreg = (reg_fn)dlsym(h, "vc_decode_region");
    if (!vopen || !vclose || !vdec || !vreg) { dlclose(h); return -2.0; }
    void *a = vopen(arc, len);
    if (!a) { dlclose(h); return -3.0; }
    if (nthreads > 0) omp_set_num_threads(nthreads);
    long nbad = 0, next = 0;
    const double t0 = now_s();
    #pragma omp parallel reduction(+:nbad)
    {
        uint8_t *tmp = (uint8_t *)malloc(32768);
        uint8_t *rbuf = NULL;
        size_t rcap = 0;
        for (;;) {
            long r;
            #pragma omp atomic capture
            r = next++;
            if (r >= (long)nreq) break;
            while (now_s() - t0 < t_arr[r]) { }
            if (kind[r] == 0) {
                const int32_t *q = quads + 4 * qoff[r];
                for (int k = 0; k < nat[r]; ++k)
                    if (vdec(a, q[4 * k], q[4 * k + 1], q[4 * k + 2], q[4 * k + 3], tmp) != 0) nbad++;
            } else {
                const int32_t *b = boxes + 7 * r;
                box6 bx = {(uint32_t)b[1], (uint32_t)b[2], (uint32_t)b[3], (uint32_t)b[4], (uint32_t)b[5], (uint32_t)b[6]};
                const size_t need = (size_t)(bx.x1 - bx.x0) * (bx.y1 - bx.y0) * (bx.z1 - bx.z0);
                if (need > rcap) { free(rbuf); rbuf = (uint8_t *)malloc(need); rcap = need; }
                if (!rbuf || vreg(a, b[0], bx, rbuf) != 0) nbad++;
            }
            t_done[r] = now_s() - t0;
        }
        free(tmp);
        free(rbuf);
    }
    double end = 0;
    for (size_t r = 0; r < nreq; ++r) if (t_done[r] > end) end = t_done[r];
    vclose(a);
    dlclose(h);
    if (bad) *bad = nbad;
    return end;
}
