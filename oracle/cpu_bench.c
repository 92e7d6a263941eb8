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
