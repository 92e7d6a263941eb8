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

/* The reference writer (vc_create / vc_set_base_q / vc_append_atom /
 * vc_mark_zero_atom / vc_writer_close of the reference build) appending every
 * LOD0 atom of an (nx, ny, nz) host volume at base_q q with an OpenMP
 * parallel-for over atoms (its appends are thread-safe), all-zero atoms
 * marked zero: the host-core cost of the exporter's LOD0 share.  Returns the
 * seconds of the appends (negative on a load or create failure). */
typedef struct { uint32_t nx, ny, nz; } dims3;
typedef void *(*create_fn)(const char *, dims3, float);
typedef int (*setq_fn)(void *, int, float);
typedef int (*app_fn)(void *, int, uint32_t, uint32_t, uint32_t, const uint8_t *);
typedef int (*zero_fn)(void *, int, uint32_t, uint32_t, uint32_t);
typedef void (*wclose_fn)(void *);

double cpu_append_lod0(const char *libpath, const char *path, const uint8_t *vol, uint32_t nx, uint32_t ny,
                       uint32_t nz, float q, int nthreads, long *bad) {
    void *h = dlopen(libpath, RTLD_NOW | RTLD_LOCAL);
    if (!h) return -1.0;
    create_fn vcreate = (create_fn)dlsym(h, "vc_create");
    setq_fn vsetq = (setq_fn)dlsym(h, "vc_set_base_q");
    app_fn vapp = (app_fn)dlsym(h, "vc_append_atom");
    zero_fn vzero = (zero_fn)dlsym(h, "vc_mark_zero_atom");
    wclose_fn vclose = (wclose_fn)dlsym(h, "vc_writer_close");
    if (!vcreate || !vsetq || !vapp || !vzero || !vclose) { dlclose(h); return -2.0; }
    dims3 d = {nx, ny, nz};
    void *w = vcreate(path, d, 1.0f);
    if (!w) { dlclose(h); return -3.0; }
    for (int l = 0; l < 8; ++l) vsetq(w, l, q);
    if (nthreads > 0) omp_set_num_threads(nthreads);
    const uint32_t acx = (nx + 31) / 32, acy = (ny + 31) / 32, acz = (nz + 31) / 32;
    const long n = (long)acx * acy * acz;
    long nbad = 0;
    const double t0 = now_s();
    #pragma omp parallel reduction(+:nbad)
    {
        uint8_t *a = (uint8_t *)malloc(32768);
        #pragma omp for schedule(dynamic, 4)
        for (long i = 0; i < n; ++i) {
            const uint32_t ax = (uint32_t)(i % acx), ay = (uint32_t)((i / acx) % acy), az = (uint32_t)(i / ((long)acx * acy));
            uint64_t any = 0;
            for (uint32_t z = 0; z < 32; ++z)
                for (uint32_t y = 0; y < 32; ++y) {
                    uint8_t *row = a + (z * 32 + y) * 32;
                    const uint32_t vz = az * 32 + z, vy = ay * 32 + y;
                    for (uint32_t x = 0; x < 32; ++x) {
                        const uint32_t vx = ax * 32 + x;
                        row[x] = (vz < nz && vy < ny && vx < nx) ? vol[((uint64_t)vz * ny + vy) * nx + vx] : 0;
                        any |= row[x];
                    }
                }
            if ((any ? vapp(w, 0, az, ay, ax, a) : vzero(w, 0, az, ay, ax)) != 0) nbad++;
        }
        free(a);
    }
    const double t = now_s() - t0;
    vclose(w);
    dlclose(h);
    if (bad) *bad = nbad;
    return t;
}
