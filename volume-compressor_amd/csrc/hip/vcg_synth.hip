// vcg_synth.hip — seeded synthetic scroll-like volume, integer-exact.
//
// base  : background ~N(12,6) (Irwin-Hall sum of four 16-bit uniforms from a
//         per-voxel splitmix64 hash, Q8), then the max with every warped sheet
//         containing the voxel (sheet: tilted plane + two Q14 sines of the
//         in-plane coordinates, half thickness and intensity in Q8);
// blur  : separable 7-tap Q12 Gaussian (sigma = 1), edge replicate, x, y, z;
// noise : + ~N(0,8) (Irwin-Hall, second hash stream), round, clip to u8.
// Sheet parameters are drawn on the host.
#include <hip/hip_runtime.h>
#include "vcg_dev.hpp"

__device__ __forceinline__ uint64_t sm64(uint64_t s) {
    uint64_t z = s + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t vhash(uint64_t seed, uint32_t stream, uint64_t idx) {
    return sm64(seed ^ ((uint64_t)stream << 56) ^ (idx * 0xD1B54A32D192ED03ull));
}
__device__ __forceinline__ int64_t irwin4(uint64_t r) {
    return (int64_t)((r & 0xFFFF) + ((r >> 16) & 0xFFFF) + ((r >> 32) & 0xFFFF) + (r >> 48)) - 131070;
}
// Q14 sine of a 16-bit angle (Bhaskara I rational approximation)
__device__ __forceinline__ int32_t isin16(uint32_t th) {
    int64_t h = th & 32767, y = h * (32768 - h);
    int64_t s = (16 * y * 16384) / (5ll * (1ll << 30) - 4 * y);
    return (th & 32768) ? -(int32_t)s : (int32_t)s;
}

__global__ void k_synth_base(uint64_t seed, uint32_t nx, uint32_t ny, uint32_t nz,
                             const vcg_sheet *__restrict__ sh, int ns, uint16_t *__restrict__ out) {
    const uint64_t N = (uint64_t)nx * ny * nz;
    for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < N;
         idx += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t x = (uint32_t)(idx % nx), y = (uint32_t)((idx / nx) % ny), z = (uint32_t)(idx / ((uint64_t)nx * ny));
        int64_t bg = 12 * 256 + ((irwin4(vhash(seed, 0, idx)) * 174353943ll + (1ll << 31)) >> 32);
        if (bg < 0) bg = 0;
        if (bg > 255 * 256) bg = 255 * 256;
        int32_t v = (int32_t)bg;
        const uint32_t c[3] = {z, y, x}, dm[3] = {nz, ny, nx};
        for (int s = 0; s < ns; ++s) {
            const vcg_sheet p = sh[s];
            const int ua = p.axis == 0 ? 1 : 0, va = p.axis == 2 ? 1 : 2;
            const int64_t u = c[ua], w = c[va];
            int64_t surf = (int64_t)p.c0 + (((int64_t)p.tu * (u * 256 - (int64_t)dm[ua] * 128)) >> 16) +
                           (((int64_t)p.tv * (w * 256 - (int64_t)dm[va] * 128)) >> 16) +
                           (((int64_t)p.a1 * isin16((uint32_t)(p.p1 + u * p.w1) & 0xFFFF)) >> 14) +
                           (((int64_t)p.a2 * isin16((uint32_t)(p.p2 + w * p.w2) & 0xFFFF)) >> 14);
            int64_t d = (int64_t)c[p.axis] * 256 + 128 - surf;
            if (d < 0) d = -d;
            if (d <= p.half && p.inten * 256 > v) v = p.inten * 256;
        }
        out[idx] = (uint16_t)v;
    }
}

__global__ void k_synth_blur(const uint16_t *__restrict__ a, uint16_t *__restrict__ b, uint64_t N,
                             uint64_t stride, uint32_t n) {
    const uint32_t W[7] = {18, 221, 991, 1636, 991, 221, 18};
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (uint64_t)gridDim.x * blockDim.x) {
        int64_t pos = (int64_t)((i / stride) % n);
        uint32_t acc = 2048;
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            int64_t q = pos + k - 3;
            if (q < 0) q = 0;
            if (q >= (int64_t)n) q = n - 1;
            acc += W[k] * a[(int64_t)i + (q - pos) * (int64_t)stride];
        }
        b[i] = (uint16_t)(acc >> 12);
    }
}

__global__ void k_synth_noise(uint64_t seed, const uint16_t *__restrict__ a, uint8_t *__restrict__ out, uint64_t N) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (uint64_t)gridDim.x * blockDim.x) {
        int64_t nq = (irwin4(vhash(seed, 1, i)) * 232471924ll + (1ll << 31)) >> 32;
        int64_t t = (int64_t)a[i] + nq + 128;
        out[i] = (uint8_t)(t < 0 ? 0 : (t >> 8) > 255 ? 255 : (t >> 8));
    }
}

extern "C" int vcg_synth(uint64_t seed, const uint32_t d[3], const vcg_sheet *sheets, int ns, void *dev_out) {
    vcg_ctx *c = vcg_get_ctx();
    const uint64_t N = (uint64_t)d[0] * d[1] * d[2];
    uint16_t *a = nullptr, *b = nullptr;
    vcg_sheet *dsh = nullptr;
    if (hipMalloc(&a, N * 2) != hipSuccess || hipMalloc(&b, N * 2) != hipSuccess ||
        hipMalloc(&dsh, sizeof(vcg_sheet) * (ns > 0 ? ns : 1)) != hipSuccess) {
        (void)hipGetLastError();
        if (a) (void)hipFree(a);
        if (b) (void)hipFree(b);
        return -1;
    }
    if (ns > 0) VCG_CHECK(hipMemcpyAsync(dsh, sheets, sizeof(vcg_sheet) * ns, hipMemcpyHostToDevice, c->s));
    uint64_t blocks = (N + 255) / 256;
    if (blocks > 1u << 20) blocks = 1u << 20;
    hipLaunchKernelGGL(k_synth_base, dim3((uint32_t)blocks), dim3(256), 0, c->s, seed, d[0], d[1], d[2], dsh, ns, a);
    VCG_CHECK(hipGetLastError());
    const uint64_t strides[3] = {1, d[0], (uint64_t)d[0] * d[1]};
    for (int ax = 0; ax < 3; ++ax) {
        hipLaunchKernelGGL(k_synth_blur, dim3((uint32_t)blocks), dim3(256), 0, c->s, a, b, N, strides[ax], d[ax]);
        VCG_CHECK(hipGetLastError());
        uint16_t *t = a; a = b; b = t;
    }
    hipLaunchKernelGGL(k_synth_noise, dim3((uint32_t)blocks), dim3(256), 0, c->s, seed, a, (uint8_t *)dev_out, N);
    VCG_CHECK(hipGetLastError());
    VCG_CHECK(hipStreamSynchronize(c->s));
    VCG_CHECK(hipFree(a)); VCG_CHECK(hipFree(b)); VCG_CHECK(hipFree(dsh));
    return 0;
}
