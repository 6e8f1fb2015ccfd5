// svo_scene.hip -- the 3-D fractal-noise cave scenes (scene kinds 3 and 4) on
// the GPU.
//
// Emits exactly the voxel stream of the host generator (svo_host.cpp, kinds 3
// and 4): solid where 5 octaves of periodic 3-D value noise exceed 0.12 inside
// a slab at the bottom of the world (height size/8 for kind 3, the whole
// height for kind 4), boundary voxels only (a solid
// voxel with a non-solid 6-neighbour; x and z wrap, below the slab counts as
// solid, above as empty), stream order z -> x -> y, colour by height with a
// seeded jitter.  The host evaluates the noise per voxel; here the lattice
// values of each octave are tabulated once and the interpolation runs per
// voxel with the host's operation order (no contraction), so every double is
// bit-identical.  Lattice spacings are powers of two, so x / L == x * (1 / L)
// exactly.
//
// Passes: lattice tables -> solid bit mask (one 64-bit word per 64 rows of a
// column) -> boundary voxels per column -> exclusive scan over columns in
// (z, x) order -> emit.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_scan.hpp>

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/svo.h"
#include "svo_internal.h"

namespace {

constexpr int kOctaves = 5;

__host__ __device__ inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// lattice value in [-1, 1) of the host's vnoise3 corner (lattice(...) * 2 - 1)
__device__ inline double corner(uint64_t seed, int oct, int64_t a, int64_t b, int64_t c)
{
    const uint64_t h = mix64(seed * 0x9E3779B97F4A7C15ull ^ mix64((uint64_t)a * 0xD1B54A32D192ED03ull ^
                                                                  (uint64_t)b * 0xABC98388FB8FAC03ull ^
                                                                  (uint64_t)c * 0x8CB92BA72F3D8DD7ull ^
                                                                  (uint64_t)(oct + 1) * 0x94D049BB133111EBull));
    return (double)(h >> 40) * (1.0 / 16777216.0) * 2.0 - 1.0;
}

struct Octave {
    const double* tab;  // [n][ny][n] corner values, index (a, b, c) = (x, y, z) lattice
    int L, n, ny;
    double inv_l;       // 1 / L (exact: L is a power of two)
};

struct CaveArgs {
    Octave o[kOctaves];
    uint64_t seed;
    int size, slab, words;  // words = 64-bit mask words per column
};

__global__ void k_lattice(double* tab, uint64_t seed, int oct, int n, int ny)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t total = (size_t)n * ny * n;
    if (i >= total) return;
    const int a = (int)(i / ((size_t)ny * n)), b = (int)((i / n) % ny), c = (int)(i % n);
    tab[i] = corner(seed, oct, a, b, c);
}

__device__ inline double fade(double t) { return t * t * (3.0 - 2.0 * t); }

// host vnoise3 with tabulated corners
__device__ inline double vnoise3(const Octave& o, int x, int y, int z)
{
    const double fx = (double)x * o.inv_l, fy = (double)y * o.inv_l, fz = (double)z * o.inv_l;
    const int64_t ix = (int64_t)floor(fx), iy = (int64_t)floor(fy), iz = (int64_t)floor(fz);
    const double tx = fade(fx - (double)ix), ty = fade(fy - (double)iy), tz = fade(fz - (double)iz);
    double r = 0;
    for (int k = 0; k < 8; k++) {
        const int dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
        const double wt = (dx ? tx : 1 - tx) * (dy ? ty : 1 - ty) * (dz ? tz : 1 - tz);
        const int64_t a = (((ix + dx) % o.n) + o.n) % o.n, b = (((iy + dy) % o.n) + o.n) % o.n,
                      c = (((iz + dz) % o.n) + o.n) % o.n;
        r += wt * o.tab[((size_t)a * o.ny + (size_t)b) * o.n + (size_t)c];
    }
    return r;
}

// solid bit of (x, y, z) for 0 <= y < slab
__device__ inline bool solid(const CaveArgs& a, int x, int y, int z)
{
    double v = 0, amp = 1;
    for (int oc = 0; oc < kOctaves; oc++) {
        v += amp * vnoise3(a.o[oc], x, y, z);
        amp *= 0.5;
    }
    return v > 0.12;
}

// one thread per (column, 64-row word): mask[(z * size + x) * words + w]
__global__ __launch_bounds__(256) void k_solid(CaveArgs a, uint64_t* mask)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t total = (size_t)a.size * a.size * a.words;
    if (i >= total) return;
    const int w = (int)(i % a.words);
    const size_t col = i / a.words;
    const int x = (int)(col % a.size), z = (int)(col / a.size);
    uint64_t m = 0;
    for (int j = 0; j < 64; j++) {
        const int y = w * 64 + j;
        if (y < a.slab && solid(a, x, y, z)) m |= 1ull << j;
    }
    mask[i] = m;
}

// boundary bits of word w of column (x, z): solid with a non-solid 6-neighbour
__device__ inline uint64_t shell_word(const CaveArgs& a, const uint64_t* mask, int x, int z, int w)
{
    const int S = a.size, W = a.words;
    const uint64_t* c = mask + ((size_t)z * S + x) * W;
    const uint64_t* xm = mask + ((size_t)z * S + (x + S - 1) % S) * W;
    const uint64_t* xp = mask + ((size_t)z * S + (x + 1) % S) * W;
    const uint64_t* zm = mask + ((size_t)((z + S - 1) % S) * S + x) * W;
    const uint64_t* zp = mask + ((size_t)((z + 1) % S) * S + x) * W;
    const uint64_t me = c[w];
    // y - 1: the previous row (below the slab counts as solid)
    const uint64_t below = (me << 1) | (w > 0 ? c[w - 1] >> 63 : 1ull);
    // y + 1: the next row (rows at or above the slab are empty: zero bits)
    const uint64_t above = (me >> 1) | (w + 1 < W ? c[w + 1] << 63 : 0ull);
    const uint64_t inner = xm[w] & xp[w] & zm[w] & zp[w] & below & above;
    return me & ~inner;
}

__global__ __launch_bounds__(256) void k_shell_count(CaveArgs a, const uint64_t* mask, uint32_t* count)
{
    const size_t col = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (col >= (size_t)a.size * a.size) return;
    const int x = (int)(col % a.size), z = (int)(col / a.size);
    uint32_t n = 0;
    for (int w = 0; w < a.words; w++) n += (uint32_t)__popcll(shell_word(a, mask, x, z, w));
    count[col] = n;
}

__device__ inline int clamp255(int v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }

__global__ __launch_bounds__(256) void k_shell_emit(CaveArgs a, const uint64_t* mask, const uint64_t* offset,
                                                    uint16_t* xyz, uint32_t* rgba)
{
    const size_t col = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (col >= (size_t)a.size * a.size) return;
    const int x = (int)(col % a.size), z = (int)(col / a.size);
    uint64_t o = offset[col];
    for (int w = 0; w < a.words; w++) {
        uint64_t m = shell_word(a, mask, x, z, w);
        while (m) {
            const int j = __ffsll((long long)m) - 1;
            m &= m - 1;
            const int y = w * 64 + j;
            // host jitter(seed, 4, x, y, z, 8) and pack_rgba(luma, 2, luma*2/3, luma/3)
            const uint64_t h = mix64(a.seed ^ mix64(((uint64_t)(uint32_t)x << 40) ^ ((uint64_t)(uint32_t)y << 20) ^
                                                    (uint64_t)(uint32_t)z ^ ((uint64_t)4 << 60)));
            const int jit = (int)(h % 17ull) - 8;
            const int luma = 60 + y * 190 / (a.slab > 1 ? a.slab : 1) + jit;
            uint32_t cx = (uint32_t)((clamp255(luma) & 0xfc) | 2);
            if (cx == 0) cx = 4;
            const uint32_t col32 =
                cx | ((uint32_t)clamp255(luma * 2 / 3) << 8) | ((uint32_t)clamp255(luma / 3) << 16);
            xyz[3 * o] = (uint16_t)x;
            xyz[3 * o + 1] = (uint16_t)y;
            xyz[3 * o + 2] = (uint16_t)z;
            rgba[o] = col32;
            o++;
        }
    }
}

inline unsigned grid_for(size_t n) { return (unsigned)((n + 255) / 256); }

struct DevBufs {
    std::vector<void*> p;
    ~DevBufs()
    {
        for (void* q : p) (void)hipFree(q);
    }
    void* get(size_t bytes)
    {
        void* q = nullptr;
        if (hipMalloc(&q, bytes + 16) != hipSuccess) return nullptr;
        p.push_back(q);
        return q;
    }
};

}  // namespace

#define GCHECK(expr)                                                                                 \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess) return svo_fail(SVO_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

extern "C" int svo_scene_generate_gpu(int device, int kind, uint64_t seed, int depth, uint16_t** xyz_out,
                                      uint32_t** rgba_out, size_t* n_out)
{
    if (!xyz_out || !rgba_out || !n_out) return svo_fail(SVO_ERR_ARG, "null output pointer");
    if (kind != 3 && kind != 4) return svo_fail(SVO_ERR_ARG, "scene kind %d is generated on the host only", kind);
    if (depth < 8 || depth > 13) return svo_fail(SVO_ERR_ARG, "depth %d outside 8..13", depth);
    int ndev = 0;
    GCHECK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return svo_fail(SVO_ERR_ARG, "device %d not present", device);
    GCHECK(hipSetDevice(device));
    const int size = 1 << depth, slab = kind == 4 ? size : size / 8;
    CaveArgs a;
    memset(&a, 0, sizeof(a));
    a.seed = seed;
    a.size = size;
    a.slab = slab;
    a.words = (slab + 63) / 64;
    DevBufs bufs;
    int L = size / 16;
    for (int oc = 0; oc < kOctaves; oc++) {
        Octave& o = a.o[oc];
        o.L = L;
        o.n = size / L;
        o.ny = std::min(o.n, (slab - 1) / L + 2);  // lattice rows the slab reaches (wrapped index 0 included)
        o.inv_l = 1.0 / (double)L;
        const size_t cells = (size_t)o.n * o.ny * o.n;
        double* tab = (double*)bufs.get(cells * 8);
        if (!tab) return svo_fail(SVO_ERR_NOMEM, "lattice table of %zu values", cells);
        hipLaunchKernelGGL(k_lattice, dim3(grid_for(cells)), dim3(256), 0, 0, tab, seed, oc, o.n, o.ny);
        GCHECK(hipGetLastError());
        o.tab = tab;
        L = std::max(2, L / 2);
    }
    const size_t cols = (size_t)size * size;
    uint64_t* mask = (uint64_t*)bufs.get(cols * a.words * 8);
    uint32_t* count = (uint32_t*)bufs.get(cols * 4);
    uint64_t* offset = (uint64_t*)bufs.get((cols + 1) * 8);
    uint64_t* count64 = (uint64_t*)bufs.get((cols + 1) * 8);
    if (!mask || !count || !offset || !count64) return svo_fail(SVO_ERR_NOMEM, "cave masks for depth %d", depth);
    hipLaunchKernelGGL(k_solid, dim3(grid_for(cols * a.words)), dim3(256), 0, 0, a, mask);
    GCHECK(hipGetLastError());
    hipLaunchKernelGGL(k_shell_count, dim3(grid_for(cols)), dim3(256), 0, 0, a, (const uint64_t*)mask, count);
    GCHECK(hipGetLastError());
    // 64-bit exclusive scan of the counts (entry cols = total)
    std::vector<uint32_t> hcount(cols);
    GCHECK(hipMemcpy(hcount.data(), count, cols * 4, hipMemcpyDeviceToHost));
    std::vector<uint64_t> hc64(cols + 1, 0);
    for (size_t i = 0; i < cols; i++) hc64[i] = hcount[i];
    GCHECK(hipMemcpy(count64, hc64.data(), (cols + 1) * 8, hipMemcpyHostToDevice));
    size_t tmp = 0;
    GCHECK(rocprim::exclusive_scan(nullptr, tmp, count64, offset, (uint64_t)0, cols + 1, rocprim::plus<uint64_t>(),
                                   (hipStream_t)0));
    void* d_tmp = bufs.get(tmp);
    if (!d_tmp) return svo_fail(SVO_ERR_NOMEM, "scan scratch");
    GCHECK(rocprim::exclusive_scan(d_tmp, tmp, count64, offset, (uint64_t)0, cols + 1, rocprim::plus<uint64_t>(),
                                   (hipStream_t)0));
    uint64_t total = 0;
    GCHECK(hipMemcpy(&total, offset + cols, 8, hipMemcpyDeviceToHost));
    if (total > 0x7ffffff0ull) return svo_fail(SVO_ERR_ARG, "cave shell of %llu voxels is too large", (unsigned long long)total);
    uint16_t* d_xyz = (uint16_t*)bufs.get(total * 6);
    uint32_t* d_rgba = (uint32_t*)bufs.get(total * 4);
    if (!d_xyz || !d_rgba) return svo_fail(SVO_ERR_NOMEM, "cave voxel stream");
    hipLaunchKernelGGL(k_shell_emit, dim3(grid_for(cols)), dim3(256), 0, 0, a, (const uint64_t*)mask,
                       (const uint64_t*)offset, d_xyz, d_rgba);
    GCHECK(hipGetLastError());
    uint16_t* hx = (uint16_t*)malloc(total * 6 + 6);
    uint32_t* hr = (uint32_t*)malloc(total * 4 + 4);
    if (!hx || !hr) {
        free(hx);
        free(hr);
        return svo_fail(SVO_ERR_NOMEM, "cave voxel stream on the host");
    }
    if (hipMemcpy(hx, d_xyz, total * 6, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hr, d_rgba, total * 4, hipMemcpyDeviceToHost) != hipSuccess) {
        free(hx);
        free(hr);
        return svo_fail(SVO_ERR_HIP, "cave voxel read-back failed");
    }
    *xyz_out = hx;
    *rgba_out = hr;
    *n_out = (size_t)total;
    return SVO_OK;
}
