// svo_cnn.hip -- learned hole filler: a 4-layer 3x3 CNN with hand-written
// BF16 MFMA forward / backward kernels (FP32 accumulate) and Adam, trained on
// (warped frame, full raycast) pairs the pipeline produces itself.  No
// counterpart in the reference (it only has the fixed filter of
// kernel.cl:404-470); the plain warp path is unchanged by this file.
//
// Network: conv3x3 5->32, ReLU, 32->32, ReLU, 32->32, ReLU, 32->3 (linear),
// zero "same" padding at crop borders.  Channels are padded to multiples of
// 16 (input 5 -> 16, output 3 -> 16); padded weights stay zero.
// Input features per pixel of the pre-filter warped frame: RGB of the
// colorized screen word (0 at holes), depth z/(z+64) (0 at holes), hole mask.
//
// Tensors are NHWC bf16 over a batch of crops.  One wave computes an 8x8
// output tile: 4 M-tiles of 16 pixels x (COUT/16) N-tiles with
// v_mfma_f32_16x16x16_bf16 (A = 16 pixels x 16 input channels from LDS,
// B = 16 input x 16 output channels from the weights, K = 9 taps x CIN).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "svo_ctx.h"
#include "svo_internal.h"

namespace {

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef uint16_t bf16;

__device__ __forceinline__ bf16 f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (bf16)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float((uint32_t)h << 16); }
// relu(acc + bias) of four channels as four bf16 (channel 0 in the low half)
__device__ __forceinline__ uint2 relu_bf16x4(f4v acc, f4v bias)
{
    uint32_t h[4];
    for (int i = 0; i < 4; i++) {
        const float v = acc[i] + bias[i];
        h[i] = (uint32_t)f2bf(v > 0.f ? v : 0.f);
    }
    return make_uint2(h[0] | (h[1] << 16), h[2] | (h[3] << 16));
}

constexpr uint32_t HOLE_WORD = 0xffffff00u;
constexpr int TILE = 8;          // output tile edge per wave
constexpr int HALO = TILE + 2;   // input tile edge

// MODE 0: + bias, ReLU, bf16 out.  MODE 1: + bias, fp32 out (last layer).
// MODE 2: no bias, multiplied by (mask_src > 0), bf16 out (backward data).
// Work items (crop, 8x8 tile) are taken grid-stride; the crop count is B, or
// min(*bcount, B) when bcount is given (batches sized on the device).
template <int CIN, int COUT, int MODE>
__global__ __launch_bounds__(64) void k_conv(const bf16* __restrict__ X, const bf16* __restrict__ Wt,
                                             const float* __restrict__ bias, const bf16* __restrict__ mask_src,
                                             void* Yout, int H, int W, int B, const uint32_t* bcount)
{
    constexpr int CS = CIN + 4;  // LDS pixel stride (bank spread)
    __shared__ __align__(16) bf16 xs[HALO * HALO * CS];
    const int lane = threadIdx.x;
    const int tiles_x = W / TILE, tiles_y = H / TILE;
    const int nb = bcount ? min((int)*bcount, B) : B;
    for (int item = blockIdx.x; item < nb * tiles_x * tiles_y; item += gridDim.x) {
    const int b = item / (tiles_x * tiles_y);
    const int t = item % (tiles_x * tiles_y);
    const int y0 = (t / tiles_x) * TILE, x0 = (t % tiles_x) * TILE;
    __syncthreads();  // previous item's LDS reads are done
    // stage the (TILE+2)^2 x CIN input halo, zero outside the crop
    for (int i = lane; i < HALO * HALO * (CIN / 4); i += 64) {
        const int pix = i / (CIN / 4), c4 = (i % (CIN / 4)) * 4;
        const int yy = y0 + pix / HALO - 1, xx = x0 + pix % HALO - 1;
        uint2 v = make_uint2(0u, 0u);
        if (yy >= 0 && yy < H && xx >= 0 && xx < W)
            v = *(const uint2*)(X + (((size_t)b * H + yy) * W + xx) * CIN + c4);
        *(uint2*)(xs + pix * CS + c4) = v;
    }
    __syncthreads();

    constexpr int NT = COUT / 16;
    f4v acc[4][NT];
    for (int m = 0; m < 4; m++)
        for (int n = 0; n < NT; n++) acc[m][n] = f4v{0.f, 0.f, 0.f, 0.f};
    const int row = lane & 15, kq = (lane >> 4) * 4;
    for (int tap = 0; tap < 9; tap++) {
        const int dy = tap / 3, dx = tap % 3;
        for (int kc = 0; kc < CIN / 16; kc++) {
            s4v bfr[NT];
            for (int n = 0; n < NT; n++)
                bfr[n] = *(const s4v*)(Wt + ((size_t)tap * COUT + n * 16 + row) * CIN + kc * 16 + kq);
            for (int m = 0; m < 4; m++) {
                const int p = m * 16 + row, py = p / TILE, px = p % TILE;
                const s4v afr = *(const s4v*)(xs + ((py + dy) * HALO + px + dx) * CS + kc * 16 + kq);
                for (int n = 0; n < NT; n++)
                    acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(afr, bfr[n], acc[m][n], 0, 0, 0);
            }
        }
    }
    // D layout: lane holds pixels 4*(lane/16)+i of the M-tile, channel lane%16
    for (int m = 0; m < 4; m++)
        for (int n = 0; n < NT; n++) {
            const int co = n * 16 + row;
            for (int i = 0; i < 4; i++) {
                const int p = m * 16 + kq + i, py = p / TILE, px = p % TILE;
                const size_t o = (((size_t)b * H + y0 + py) * W + x0 + px) * COUT + co;
                float v = acc[m][n][i];
                if (MODE == 0) {
                    v += bias[co];
                    v = v > 0.f ? v : 0.f;
                    ((bf16*)Yout)[o] = f2bf(v);
                } else if (MODE == 1) {
                    ((float*)Yout)[o] = v + bias[co];
                } else {
                    // ReLU derivative: the forward output (>= 0) is positive
                    ((bf16*)Yout)[o] = mask_src[o] != 0 ? f2bf(v) : (bf16)0;
                }
            }
        }
    }
}

// dW partial of crop b: sum over the crop's pixels p of dZ[p][co] X[p + tap
// - 1][ci].  grid = (crops, 9 taps), DW_SLICES waves per workgroup: wave s
// loops over pixel slice s (a multiple of 16 pixels), K = 16 pixels per MFMA
// (A = dZ^T: co x pixels, B = shifted X: pixels x ci); the waves' sums are
// added in slice order through LDS and written to part[b][tap][co][ci];
// k_sum_parts adds the crops' partials in crop order, so the gradient does
// not depend on scheduling.
constexpr int DW_SLICES = 16;
template <int CIN, int COUT>
__global__ __launch_bounds__(64 * DW_SLICES) void k_conv_dw(const bf16* __restrict__ X, const bf16* __restrict__ dZ,
                                                            float* __restrict__ part, int H, int W)
{
    __shared__ float s_acc[DW_SLICES][COUT * CIN];
    const int b = blockIdx.x, tap = blockIdx.y, sl = threadIdx.x >> 6;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int lane = threadIdx.x & 63, row = lane & 15, kq = (lane >> 4) * 4;
    constexpr int MT = COUT / 16, NT = CIN / 16;
    f4v acc[MT][NT];
    for (int m = 0; m < MT; m++)
        for (int n = 0; n < NT; n++) acc[m][n] = f4v{0.f, 0.f, 0.f, 0.f};
    const int npix = H * W;  // a multiple of 16
    const int len = (npix / 16 + DW_SLICES - 1) / DW_SLICES * 16;
    const int kb = min(npix, sl * len), ke = min(npix, kb + len);
    const bf16* dz = dZ + (size_t)b * npix * COUT;
    const bf16* xb = X + (size_t)b * npix * CIN;
    for (int k0 = kb; k0 < ke; k0 += 16) {
        s4v afr[MT], bfr[NT];
        for (int m = 0; m < MT; m++) {
            short tmp[4];
            for (int j = 0; j < 4; j++) tmp[j] = (short)dz[(size_t)(k0 + kq + j) * COUT + m * 16 + row];
            afr[m] = s4v{tmp[0], tmp[1], tmp[2], tmp[3]};
        }
        for (int n = 0; n < NT; n++) {
            short tmp[4];
            for (int j = 0; j < 4; j++) {
                const int p = k0 + kq + j, y = p / W + dy, x = p % W + dx;
                tmp[j] = (y >= 0 && y < H && x >= 0 && x < W) ? (short)xb[((size_t)y * W + x) * CIN + n * 16 + row] : 0;
            }
            bfr[n] = s4v{tmp[0], tmp[1], tmp[2], tmp[3]};
        }
        for (int m = 0; m < MT; m++)
            for (int n = 0; n < NT; n++)
                acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(afr[m], bfr[n], acc[m][n], 0, 0, 0);
    }
    for (int m = 0; m < MT; m++)
        for (int n = 0; n < NT; n++)
            for (int i = 0; i < 4; i++) {
                const int co = m * 16 + kq + i, ci = n * 16 + row;
                s_acc[sl][co * CIN + ci] = acc[m][n][i];
            }
    __syncthreads();
    float* dw = part + ((size_t)b * 9 + tap) * COUT * CIN;
    for (int e = threadIdx.x; e < COUT * CIN; e += 64 * DW_SLICES) {
        float t = 0.f;
        for (int k = 0; k < DW_SLICES; k++) t += s_acc[k][e];
        dw[e] = t;
    }
}

// db partial of block b: part[b][co] = sum over its DB_PIXELS pixels of
// dZ[p][co] (each thread a fixed stride of pixels, then the threads of a
// channel in order)
constexpr int DB_PIXELS = 256;
template <int COUT>
__global__ __launch_bounds__(256) void k_conv_db(const bf16* __restrict__ dZ, float* __restrict__ part, int npix)
{
    __shared__ float s_sum[256];
    const int co = threadIdx.x % COUT, lane_p = threadIdx.x / COUT;
    const int p0 = blockIdx.x * DB_PIXELS, p1 = min(p0 + DB_PIXELS, npix);
    float s = 0.f;
    for (int p = p0 + lane_p; p < p1; p += 256 / COUT) s += bf2f(dZ[(size_t)p * COUT + co]);
    s_sum[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < COUT) {
        float t = 0.f;
        for (int j = 0; j < 256 / COUT; j++) t += s_sum[j * COUT + threadIdx.x];
        part[(size_t)blockIdx.x * COUT + threadIdx.x] = t;
    }
}

// out[e] for many partials and few elements (one workgroup per element):
// thread t sums part[t], part[t + 256], ... in order, then a fixed pairwise
// tree over the 256 threads
__global__ __launch_bounds__(256) void k_sum_parts_tree(const float* __restrict__ part, int n_parts, int n,
                                                        float* __restrict__ out)
{
    __shared__ float s[256];
    const int e = blockIdx.x, t = threadIdx.x;
    float v = 0.f;
    for (int p = t; p < n_parts; p += 256) v += part[(size_t)p * n + e];
    s[t] = v;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) s[t] += s[t + h];
        __syncthreads();
    }
    if (t == 0) out[e] = s[0];
}

// out[e] = sum over p = 0 .. n_parts-1, in order, of part[p][e]
__global__ __launch_bounds__(256) void k_sum_parts(const float* __restrict__ part, int n_parts, int n,
                                                   float* __restrict__ out)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    float s = 0.f;
    for (int p = 0; p < n_parts; p++) s += part[(size_t)p * n + e];
    out[e] = s;
}

// masked MSE on RGB: loss += sum mask*(pred-t)^2 / (3*count); dZ = d loss/d pred
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ pred, const float* __restrict__ target,
                                              const float* __restrict__ mask, bf16* dZ, float* loss, int npix,
                                              const float* __restrict__ mask_count)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= npix) return;
    const float inv_count = *mask_count > 0.f ? 1.f / (3.f * *mask_count) : 0.f;
    float l = 0.f;
    for (int c = 0; c < 16; c++) {
        float g = 0.f;
        if (c < 3) {
            const float d = pred[(size_t)p * 16 + c] - target[(size_t)p * 3 + c];
            l += mask[p] * d * d;
            g = 2.f * mask[p] * d * inv_count;
        }
        dZ[(size_t)p * 16 + c] = f2bf(g);
    }
    if (l != 0.f) atomicAdd(loss, l * inv_count);
}

// Input features of crop c (origin from `origins`, or tile list -> tile*16-4):
// RGB of the colorized word of the CNN input planes (plane 0 as the gate of a
// learned frame leaves it), camera depth z/(z+64) and the hole mask; zero
// outside the screen and at holes.  With `image`, also the target RGB (full
// raycast) and the loss mask (pixels that are holes in `mask_words`, else in
// the input, inside the screen), counting mask pixels into *mask_count.
__device__ __forceinline__ void rgb_of(uint32_t word, float& r, float& g, float& b)
{
    const int a = (int)word, c = a & 3;
    const float i = (float)(a & (255 - 7));
    float tr = 1.0f, tg = 1.0f, tb = 1.0f;
    if (c == 1) { tg = 0.7f; tb = 0.3f; }
    else if (c == 2) { tr = 1.5f; tg = 0.8f; tb = 0.1f; }
    else if (c == 3) { tr = 0.2f; tg = 0.8f; tb = 0.2f; }
    r = fminf((float)(int)(i * tr), 255.f) * (1.f / 255.f);
    g = fminf((float)(int)(i * tg), 255.f) * (1.f / 255.f);
    b = fminf((float)(int)(i * tb), 255.f) * (1.f / 255.f);
}

struct FeatArgs {
    const uint32_t* words;   // screen words of the CNN input (hole channel)
    const float4* coords;    // coords of the CNN input
    const uint32_t* mask_words;  // training: the loss covers pixels that are holes here (else `words`)
    const uint32_t* image;   // full-raycast image (targets) or null
    int res_x, res_y, crop;
    const int2* origins;     // explicit crop origins, or null
    const uint32_t* tiles;   // tile list (origin = tile*16 - 4), or null
    int tiles_x;
    const uint32_t* count;   // device crop count (tile list), or null
    int max_crops;
    float m0x, m0y, m0z, r20, r21, r22;
    bf16* x;
    float* target;
    float* mask;
    float* mask_count;
};

__global__ __launch_bounds__(256) void k_features(FeatArgs a)
{
    const int nb = a.count ? min((int)*a.count, a.max_crops) : a.max_crops;
    const size_t per = (size_t)a.crop * a.crop;
    float cnt = 0.f;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < (size_t)nb * per; i += (size_t)gridDim.x * 256) {
        const int c = (int)(i / per), q = (int)(i % per);
        int ox, oy;
        if (a.origins) {
            ox = a.origins[c].x;
            oy = a.origins[c].y;
        } else {
            const int t = (int)a.tiles[c];
            ox = (t % a.tiles_x) * 16 - 4;
            oy = (t / a.tiles_x) * 16 - 4;
        }
        const int x = ox + q % a.crop, y = oy + q / a.crop;
        float f[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
        float tr = 0.f, tg = 0.f, tb = 0.f, m = 0.f;
        if (x >= 0 && y >= 0 && x < a.res_x && y < a.res_y) {
            const size_t o = (size_t)y * a.res_x + x;
            const uint32_t w = a.words[o];
            m = (a.mask_words ? a.mask_words[o] : w) == HOLE_WORD ? 1.f : 0.f;
            if (w == HOLE_WORD) {
                f[4] = 1.f;
            } else {
                rgb_of(w, f[0], f[1], f[2]);
                const float4 cc = a.coords[o];
                const float qx = cc.x - a.m0x, qy = cc.y - a.m0y, qz = cc.z - a.m0z;
                const float z = fmaxf((qx * a.r20 + qy * a.r21) + qz * a.r22, 0.f);
                f[3] = z / (z + 64.f);
            }
            if (a.image) {
                const uint32_t px = a.image[o];
                tr = (float)((px >> 16) & 255u) * (1.f / 255.f);
                tg = (float)((px >> 8) & 255u) * (1.f / 255.f);
                tb = (float)(px & 255u) * (1.f / 255.f);
            }
        }
        bf16* xo = a.x + i * 16;
        for (int k = 0; k < 16; k++) xo[k] = k < 5 ? f2bf(f[k]) : (bf16)0;
        if (a.target) {
            a.target[i * 3 + 0] = tr;
            a.target[i * 3 + 1] = tg;
            a.target[i * 3 + 2] = tb;
            a.mask[i] = m;
            cnt += m;
        }
    }
    if (a.target && cnt != 0.f) atomicAdd(a.mask_count, cnt);
}

__device__ __forceinline__ uint32_t rgb_word(float r, float g, float b)
{
    const uint32_t R = (uint32_t)fminf(fmaxf(r * 255.f + 0.5f, 0.f), 255.f);
    const uint32_t G = (uint32_t)fminf(fmaxf(g * 255.f + 0.5f, 0.f), 255.f);
    const uint32_t B = (uint32_t)fminf(fmaxf(b * 255.f + 0.5f, 0.f), 255.f);
    return B | (G << 8) | (R << 16);
}

// CNN RGB of the centre 16x16 of each tile crop -> 256 colorized words per tile
__global__ __launch_bounds__(256) void k_stage_fill(const float* __restrict__ out, const uint32_t* __restrict__ count,
                                                    int max_crops, uint32_t* __restrict__ staged)
{
    const int nb = min((int)*count, max_crops);
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nb * 256; i += gridDim.x * 256) {
        const int c = i / 256, q = i % 256;
        const float* v = out + ((size_t)c * 576 + (size_t)((q >> 4) + 4) * 24 + (q & 15) + 4) * 16;
        staged[i] = rgb_word(v[0], v[1], v[2]);
    }
}

// staged CNN colours -> image at the tile pixels that are still holes in
// `words` (plane 2: the frame after its rays, before the filter)
__global__ __launch_bounds__(256) void k_cnn_apply(const uint32_t* __restrict__ staged,
                                                   const uint32_t* __restrict__ tiles,
                                                   const uint32_t* __restrict__ count, int max_crops,
                                                   const uint32_t* __restrict__ words, uint32_t* image, int res_x,
                                                   int res_y, int tiles_x)
{
    const int nb = min((int)*count, max_crops);
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nb * 256; i += gridDim.x * 256) {
        const int c = i / 256, q = i % 256;
        const int t = (int)tiles[c];
        const int x = (t % tiles_x) * 16 + (q & 15), y = (t / tiles_x) * 16 + (q >> 4);
        if (x >= res_x || y >= res_y) continue;
        const size_t o = (size_t)y * res_x + x;
        if (words[o] == HOLE_WORD) image[o] = staged[i];
    }
}

// ---------------------------------------------------------------- fused tile inference
// One workgroup (4 waves) runs the whole network on one 24x24 crop around a
// 16x16 gate tile, activations in LDS.  Each layer is evaluated only where
// the centre 16x16 output needs it: layer 1 on 22x22, layer 2 on 20x20,
// layer 3 on 18x18, layer 4 on 16x16 (the crop border's zero padding never
// reaches the centre, so this equals the full-crop network there).  An
// M-tile is 16 consecutive output pixels in raster order; layer 1 uses
// v_mfma_f32_16x16x16_bf16 over K = 16 channels per tap, layers 2-4
// v_mfma_f32_16x16x32_bf16 over K = 32 (A lane: pixel l%16, channels
// 8*(l/16)..+7; B lane: output channel l%16, the same 8 input channels).
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

// 32-channel pixels: the 16-byte channel group g of pixel p sits in slot
// g ^ ((p >> 1) & 3).  A 128-bit LDS read is served 8 lanes at a time from
// 128 bytes of banks: 8 consecutive pixels (64-byte stride) of one group
// then land on 8 distinct 16-byte bank slots (pixel parity x slot)
__device__ __forceinline__ int swz(int p, int g) { return p * 32 + ((g ^ ((p >> 1) & 3)) << 3); }
// 16-channel input pixels (layer 1): the 8-byte group k (4 channels) of
// pixel p sits in slot k ^ ((p >> 2) & 3); a 64-bit read is served 16 lanes
// at a time from 128 bytes, and 16 consecutive pixels (32-byte stride) of
// one group then cover 16 distinct 8-byte slots
__device__ __forceinline__ int swz1(int p, int k) { return p * 16 + ((k ^ ((p >> 2) & 3)) << 2); }

// One view's CNN tiles in frame mode: features of the crop at tile*16 - 4
// from the view's CNN input planes, 256 colorized words per tile to `staged`;
// the apply step writes them into `image` where plane 2 still holds holes.
struct TileView {
    const uint32_t* words;
    const float4* coords;
    int res_x, res_y, tiles_x;
    const uint32_t* tiles;
    const uint32_t* count;
    int max_tiles;
    float m0x, m0y, m0z, r20, r21, r22;
    uint32_t* staged;
    const uint32_t* plane2;
    uint32_t* image;
};

struct TileArgs {
    // frame mode (x == null): one view, or a device table of n_views views
    // whose tile lists are concatenated (first[i] = the first tile of view i,
    // first[n_views] = all tiles)
    TileView v;
    const TileView* views;
    const uint32_t* first;
    int n_views;
    // explicit mode: x [n][24][24][16] bf16 -> out [n][16][16][3] float
    const bf16* x;
    float* out;
    int n;
    const bf16* w[4];  // [9][32][16], [9][32][32], [9][32][32], [9][16][32]
    const float* b[4];
};

// first tile of each batched view (after their gates), one thread
__global__ void k_tile_first(const TileView* views, int n, uint32_t* first)
{
    uint32_t acc = 0;
    for (int i = 0; i < n; i++) {
        first[i] = acc;
        acc += (uint32_t)min((int)*views[i].count, views[i].max_tiles);
    }
    first[n] = acc;
}

// k_cnn_apply for every view of a batch (grid y = view)
__global__ __launch_bounds__(256) void k_cnn_apply_views(const TileView* views)
{
    const TileView& V = views[blockIdx.y];
    const int nb = min((int)*V.count, V.max_tiles);
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nb * 256; i += gridDim.x * 256) {
        const int c = i / 256, q = i % 256;
        const int t = (int)V.tiles[c];
        const int x = (t % V.tiles_x) * 16 + (q & 15), y = (t / V.tiles_x) * 16 + (q >> 4);
        if (x >= V.res_x || y >= V.res_y) continue;
        const size_t o = (size_t)y * V.res_x + x;
        if (V.plane2[o] == HOLE_WORD) V.image[o] = V.staged[i];
    }
}

// Both tile layers compute the transposed product D[channel][pixel] (the
// weights are the MFMA's A operand, the pixels its B operand), so a lane
// ends with four consecutive output channels of one pixel and stores them as
// one 8-byte write.

// layers 2-4: in = swizzled 32-channel pixels of a win x win region, Wl =
// swizzled [9][COUT][32] weights
template <int COUT, bool LAST>
__device__ __forceinline__ void tile_layer32(const bf16* in, int win, bf16* out, float* outf, int wout,
                                             const bf16* Wl, const float* __restrict__ bias, int wave, int lane,
                                             int nw)
{
    constexpr int NT = COUT / 16;
    const int npix = wout * wout;
    const int row = lane & 15, kg = lane >> 4;
    // the layer's B fragments stay in registers for all of the wave's M-tiles
    bf8v bfr[9][NT];
    for (int tap = 0; tap < 9; tap++)
        for (int nn = 0; nn < NT; nn++) bfr[tap][nn] = *(const bf8v*)(Wl + swz(tap * COUT + nn * 16 + row, kg));
    // this lane's output channels: nn * 16 + kg * 4 + i
    f4v bv[NT];
    for (int nn = 0; nn < NT; nn++)
        for (int i = 0; i < 4; i++) bv[nn][i] = bias[nn * 16 + kg * 4 + i];
    for (int m0 = wave * 2; m0 * 16 < npix; m0 += 2 * nw) {
        f4v acc[2][NT];
        int q0[2];
        for (int m = 0; m < 2; m++) {
            for (int nn = 0; nn < NT; nn++) acc[m][nn] = f4v{0.f, 0.f, 0.f, 0.f};
            const int p = min((m0 + m) * 16 + row, npix - 1);  // past the end: recomputed, not stored
            q0[m] = (p / wout) * win + p % wout;
        }
        for (int tap = 0; tap < 9; tap++) {
            const int dq = (tap / 3) * win + tap % 3;
            for (int m = 0; m < 2; m++) {
                const bf8v afr = *(const bf8v*)(in + swz(q0[m] + dq, kg));
                for (int nn = 0; nn < NT; nn++)
                    acc[m][nn] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfr[tap][nn], afr, acc[m][nn], 0, 0, 0);
            }
        }
        for (int m = 0; m < 2; m++) {
            const int q = (m0 + m) * 16 + row;
            if (q >= npix) continue;
            for (int nn = 0; nn < NT; nn++) {
                const int co0 = nn * 16 + kg * 4;
                if (LAST) {
                    if (co0 == 0)
                        for (int i = 0; i < 3; i++) outf[q * 3 + i] = acc[m][nn][i] + bv[nn][i];
                } else {
                    *(uint2*)(out + swz(q, co0 >> 3) + (co0 & 7)) = relu_bf16x4(acc[m][nn], bv[nn]);
                }
            }
        }
    }
}

// weights of the four layers as the tile kernel keeps them in LDS: layer 1
// [9][32][16] plain, layers 2-4 [9][COUT][32] with swizzled rows
constexpr int W_OFS1 = 9 * 32 * 16, W_OFS2 = W_OFS1 + 9 * 32 * 32, W_OFS3 = W_OFS2 + 9 * 32 * 32;
constexpr int W_ALL = W_OFS3 + 9 * 16 * 32;

__device__ __forceinline__ void stage_weights(bf16* dst, const bf16* __restrict__ src, int rows, bool swizzled,
                                              int tid, int nthreads)
{
    if (!swizzled) {
        for (int i = tid; i < rows * 16 / 8; i += nthreads) *(uint4*)(dst + i * 8) = *(const uint4*)(src + i * 8);
        return;
    }
    for (int i = tid; i < rows * 4; i += nthreads) {
        const int r = i >> 2, g = i & 3;
        *(uint4*)(dst + swz(r, g)) = *(const uint4*)(src + r * 32 + g * 8);
    }
}

// layer 1: 16 -> 32 on 22x22 from the 24x24 plain 16-channel input
__device__ __forceinline__ void tile_layer1(const bf16* P, bf16* Q, const bf16* Wl, const float* __restrict__ bias,
                                            int wave, int lane, int nw)
{
    const int row = lane & 15, kg = lane >> 4;
    s4v b1[9][2];
    for (int tap = 0; tap < 9; tap++)
        for (int nn = 0; nn < 2; nn++) b1[tap][nn] = *(const s4v*)(Wl + (tap * 32 + nn * 16 + row) * 16 + kg * 4);
    f4v bv[2];
    for (int nn = 0; nn < 2; nn++)
        for (int i = 0; i < 4; i++) bv[nn][i] = bias[nn * 16 + kg * 4 + i];
    for (int m0 = wave * 2; m0 * 16 < 484; m0 += 2 * nw) {
        f4v acc[2][2];
        int q0[2];
        for (int m = 0; m < 2; m++) {
            acc[m][0] = acc[m][1] = f4v{0.f, 0.f, 0.f, 0.f};
            const int p = min((m0 + m) * 16 + row, 483);
            q0[m] = (p / 22) * 24 + p % 22;
        }
        for (int tap = 0; tap < 9; tap++) {
            const int dq = (tap / 3) * 24 + tap % 3;
            for (int m = 0; m < 2; m++) {
                const s4v afr = *(const s4v*)(P + swz1(q0[m] + dq, kg));
                for (int nn = 0; nn < 2; nn++)
                    acc[m][nn] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(b1[tap][nn], afr, acc[m][nn], 0, 0, 0);
            }
        }
        for (int m = 0; m < 2; m++) {
            const int q = (m0 + m) * 16 + row;
            if (q >= 484) continue;
            for (int nn = 0; nn < 2; nn++) {
                const int co0 = nn * 16 + kg * 4;
                *(uint2*)(Q + swz(q, co0 >> 3) + (co0 & 7)) = relu_bf16x4(acc[m][nn], bv[nn]);
            }
        }
    }
}

// NW waves per workgroup.  RESIDENT: the four layers' weights are staged in
// LDS once per workgroup (8 waves, one workgroup per CU); otherwise each
// layer's weights are staged per tile (4 waves, two workgroups per CU).
template <int NW, bool RESIDENT>
__global__ __launch_bounds__(NW * 64) void k_cnn_tiles(TileArgs a)
{
    __shared__ __align__(16) bf16 P[400 * 32];  // input 576 x 16 ch / layer-2 output 400 x 32 ch / layer-4 RGB
    __shared__ __align__(16) bf16 Q[484 * 32];  // layer-1 output 484 x 32 / layer-3 output 324 x 32
    __shared__ __align__(16) bf16 Wl[RESIDENT ? W_ALL : 9 * 32 * 32];
    float* OF = (float*)P;  // 256 x 3 floats, written by layer 4 after layer 3 has consumed P
    constexpr int NT = NW * 64;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int n = a.x ? a.n : a.views ? (int)a.first[a.n_views] : min((int)*a.v.count, a.v.max_tiles);
    if (RESIDENT) {
        stage_weights(Wl, a.w[0], 9 * 32, false, tid, NT);
        stage_weights(Wl + W_OFS1, a.w[1], 9 * 32, true, tid, NT);
        stage_weights(Wl + W_OFS2, a.w[2], 9 * 32, true, tid, NT);
        stage_weights(Wl + W_OFS3, a.w[3], 9 * 16, true, tid, NT);
    }
    for (int c = blockIdx.x; c < n; c += gridDim.x) {
        __syncthreads();  // the previous crop is done with every buffer (and the weights are staged)
        int ox = 0, oy = 0, lc = c;
        const TileView* V = &a.v;
        if (a.views) {  // the last view whose first tile is <= c
            int lo = 0, hi = a.n_views - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if ((int)a.first[mid] <= c) lo = mid;
                else hi = mid - 1;
            }
            V = &a.views[lo];
            lc = c - (int)a.first[lo];
        }
        if (!a.x) {
            const int t = (int)V->tiles[lc];
            ox = (t % V->tiles_x) * 16 - 4;
            oy = (t / V->tiles_x) * 16 - 4;
        }
        for (int q = tid; q < 576; q += NT) {
            uint4 lo = make_uint4(0u, 0u, 0u, 0u), hi = lo;
            if (a.x) {
                const uint4* src = (const uint4*)(a.x + ((size_t)c * 576 + q) * 16);
                lo = src[0];
                hi = src[1];
            } else {
                const int x = ox + q % 24, y = oy + q / 24;
                float f[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
                if (x >= 0 && y >= 0 && x < V->res_x && y < V->res_y) {
                    const size_t o = (size_t)y * V->res_x + x;
                    const uint32_t w = V->words[o];
                    if (w == HOLE_WORD) {
                        f[4] = 1.f;
                    } else {
                        rgb_of(w, f[0], f[1], f[2]);
                        const float4 cc = V->coords[o];
                        const float qx = cc.x - V->m0x, qy = cc.y - V->m0y, qz = cc.z - V->m0z;
                        const float z = fmaxf((qx * V->r20 + qy * V->r21) + qz * V->r22, 0.f);
                        f[3] = z / (z + 64.f);
                    }
                }
                lo.x = (uint32_t)f2bf(f[0]) | ((uint32_t)f2bf(f[1]) << 16);
                lo.y = (uint32_t)f2bf(f[2]) | ((uint32_t)f2bf(f[3]) << 16);
                lo.z = (uint32_t)f2bf(f[4]);
            }
            *(uint2*)(P + swz1(q, 0)) = make_uint2(lo.x, lo.y);
            *(uint2*)(P + swz1(q, 1)) = make_uint2(lo.z, lo.w);
            *(uint2*)(P + swz1(q, 2)) = make_uint2(hi.x, hi.y);
            *(uint2*)(P + swz1(q, 3)) = make_uint2(hi.z, hi.w);
        }
        if (!RESIDENT) stage_weights(Wl, a.w[0], 9 * 32, false, tid, NT);
        __syncthreads();
        tile_layer1(P, Q, Wl, a.b[0], wave, lane, NW);
        __syncthreads();
        if (!RESIDENT) {
            stage_weights(Wl, a.w[1], 9 * 32, true, tid, NT);
            __syncthreads();
        }
        tile_layer32<32, false>(Q, 22, P, nullptr, 20, RESIDENT ? Wl + W_OFS1 : Wl, a.b[1], wave, lane, NW);
        __syncthreads();
        if (!RESIDENT) {
            stage_weights(Wl, a.w[2], 9 * 32, true, tid, NT);
            __syncthreads();
        }
        tile_layer32<32, false>(P, 20, Q, nullptr, 18, RESIDENT ? Wl + W_OFS2 : Wl, a.b[2], wave, lane, NW);
        __syncthreads();
        if (!RESIDENT) {
            stage_weights(Wl, a.w[3], 9 * 16, true, tid, NT);
            __syncthreads();
        }
        tile_layer32<16, true>(Q, 18, nullptr, OF, 16, RESIDENT ? Wl + W_OFS3 : Wl, a.b[3], wave, lane, NW);
        __syncthreads();
        for (int q = tid; q < 256; q += NT) {
            const float r = OF[q * 3], g = OF[q * 3 + 1], b = OF[q * 3 + 2];
            if (a.x) {
                float* o = a.out + ((size_t)c * 256 + q) * 3;
                o[0] = r;
                o[1] = g;
                o[2] = b;
            } else {
                V->staged[(size_t)lc * 256 + q] = rgb_word(r, g, b);
            }
        }
    }
}

// W'[t][ci][co] = W[8-t][co][ci] (backward-data weights), bf16
__global__ void k_flip(const float* __restrict__ W, bf16* Wf, int cin, int cout)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= 9 * cin * cout) return;
    const int t = i / (cin * cout), r = i % (cin * cout), ci = r / cout, co = r % cout;
    Wf[i] = f2bf(W[((size_t)(8 - t) * cout + co) * cin + ci]);
}

__global__ void k_to_bf16(const float* __restrict__ a, bf16* b, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = f2bf(a[i]);
}

// Adam (FP32 master weights): m, v moments, bias-corrected step
__global__ void k_adam(float* w, const float* g, float* m, float* v, int n, float lr, float b1, float b2,
                       float c1, float c2, float eps)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace

// ====================================================================== host
struct CnnLayer {
    int cin, cout;
    size_t nw;          // 9*cin*cout
    float* w = nullptr; // fp32 master [9][cout][cin]
    float* b = nullptr; // fp32 [cout]
    float* gw = nullptr;
    float* gb = nullptr;
    float *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;
    bf16* wbf = nullptr;   // forward weights
    bf16* wflip = nullptr; // backward-data weights [9][cin][cout]
};

struct svo_cnn {
    int device = 0;
    int n_cu = 256;        // compute units of the device (tile kernel grid)
    hipStream_t stream = nullptr;
    CnnLayer L[4];
    int step = 0;
    // activation workspace for the current batch size
    size_t cap_pix = 0;
    bf16* x0 = nullptr;    // [pix][16]
    bf16* y[3] = {};       // [pix][32]
    float* out = nullptr;  // [pix][16]
    bf16* dz16 = nullptr;  // [pix][16]
    bf16* dz32a = nullptr; // [pix][32]
    bf16* dz32b = nullptr; // [pix][32]
    float* target = nullptr;
    float* mask = nullptr;
    float* loss = nullptr;
    float* mask_count = nullptr;
    int2* origins = nullptr;
    int max_origins = 0;
    bool fused = true;     // per-tile fused inference kernel for the frame fill
    int tile_mode = 0;     // fused kernel variant (see launch_tiles)
    bf16* xt = nullptr;    // explicit-mode tile inputs / outputs
    float* ot = nullptr;
    size_t cap_tiles = 0;
    float* part = nullptr;  // gradient partials (per crop and tap, or per block)
    size_t cap_part = 0;
};

#define RCC(x)             \
    do {                   \
        int _r = (x);      \
        if (_r) return _r; \
    } while (0)

#define HIPCHECK(expr)                                                                    \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess)                                                             \
            return svo_fail(SVO_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

static const int kCin[4] = {16, 32, 32, 32}, kCout[4] = {32, 32, 32, 16};

static void free_ws(svo_cnn* n)
{
    (void)hipFree(n->x0);
    for (auto& p : n->y) (void)hipFree(p);
    (void)hipFree(n->out);
    (void)hipFree(n->dz16);
    (void)hipFree(n->dz32a);
    (void)hipFree(n->dz32b);
    (void)hipFree(n->target);
    (void)hipFree(n->mask);
    n->x0 = nullptr;
    for (auto& p : n->y) p = nullptr;
    n->out = nullptr;
    n->dz16 = n->dz32a = n->dz32b = nullptr;
    n->target = n->mask = nullptr;
    n->cap_pix = 0;
}

static int ensure_ws(svo_cnn* n, size_t pix)
{
    if (pix <= n->cap_pix) return SVO_OK;
    free_ws(n);
    bool ok = hipMalloc(&n->x0, pix * 16 * 2) == hipSuccess && hipMalloc(&n->y[0], pix * 32 * 2) == hipSuccess &&
              hipMalloc(&n->y[1], pix * 32 * 2) == hipSuccess && hipMalloc(&n->y[2], pix * 32 * 2) == hipSuccess &&
              hipMalloc(&n->out, pix * 16 * 4) == hipSuccess && hipMalloc(&n->dz16, pix * 16 * 2) == hipSuccess &&
              hipMalloc(&n->dz32a, pix * 32 * 2) == hipSuccess && hipMalloc(&n->dz32b, pix * 32 * 2) == hipSuccess &&
              hipMalloc(&n->target, pix * 3 * 4) == hipSuccess && hipMalloc(&n->mask, pix * 4) == hipSuccess;
    if (!ok) {
        free_ws(n);
        return svo_fail(SVO_ERR_NOMEM, "CNN workspace for %zu pixels", pix);
    }
    n->cap_pix = pix;
    return SVO_OK;
}

static unsigned nb(size_t n) { return (unsigned)((n + 255) / 256); }

static int refresh_bf16(svo_cnn* n)
{
    for (auto& l : n->L) {
        hipLaunchKernelGGL(k_to_bf16, dim3(nb(l.nw)), dim3(256), 0, n->stream, l.w, l.wbf, (int)l.nw);
        hipLaunchKernelGGL(k_flip, dim3(nb(l.nw)), dim3(256), 0, n->stream, l.w, l.wflip, l.cin, l.cout);
    }
    HIPCHECK(hipGetLastError());
    return SVO_OK;
}

extern "C" int svo_cnn_destroy(svo_cnn* n)
{
    if (!n) return SVO_OK;
    (void)hipSetDevice(n->device);
    (void)hipStreamSynchronize(n->stream);
    free_ws(n);
    for (auto& l : n->L) {
        for (float* p : {l.w, l.b, l.gw, l.gb, l.mw, l.vw, l.mb, l.vb}) (void)hipFree(p);
        (void)hipFree(l.wbf);
        (void)hipFree(l.wflip);
    }
    (void)hipFree(n->loss);
    (void)hipFree(n->mask_count);
    (void)hipFree(n->part);
    (void)hipFree(n->origins);
    (void)hipFree(n->xt);
    (void)hipFree(n->ot);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
    return SVO_OK;
}

extern "C" size_t svo_cnn_param_count(void)
{
    size_t s = 0;
    for (int i = 0; i < 4; i++) s += 9 * (size_t)kCin[i] * kCout[i] + kCout[i];
    return s;
}

extern "C" int svo_cnn_create(svo_cnn** out, int device, uint64_t seed)
{
    if (!out) return svo_fail(SVO_ERR_ARG, "null output pointer");
    *out = nullptr;
    HIPCHECK(hipSetDevice(device));
    svo_cnn* n = new svo_cnn();
    n->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        n->n_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&n->loss, 4) != hipSuccess || hipMalloc(&n->mask_count, 4) != hipSuccess) {
        svo_cnn_destroy(n);
        return svo_fail(SVO_ERR_HIP, "CNN set-up failed");
    }
    // He-normal init of the real channels (Box-Muller over a splitmix stream)
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
    auto next = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return ((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
    };
    const int real_in[4] = {5, 32, 32, 32}, real_out[4] = {32, 32, 32, 3};
    for (int li = 0; li < 4; li++) {
        CnnLayer& l = n->L[li];
        l.cin = kCin[li];
        l.cout = kCout[li];
        l.nw = 9 * (size_t)l.cin * l.cout;
        std::vector<float> w(l.nw, 0.f), b(l.cout, 0.f);
        const double sd = sqrt(2.0 / (9.0 * real_in[li])) * (li == 3 ? 0.5 : 1.0);
        for (int t = 0; t < 9; t++)
            for (int co = 0; co < real_out[li]; co++)
                for (int ci = 0; ci < real_in[li]; ci++) {
                    double u1 = std::max(next(), 1e-12), u2 = next();
                    w[((size_t)t * l.cout + co) * l.cin + ci] = (float)(sd * sqrt(-2 * log(u1)) * cos(6.283185307179586 * u2));
                }
        bool ok = true;
        for (float** p : {&l.w, &l.gw, &l.mw, &l.vw}) ok = ok && hipMalloc(p, l.nw * 4) == hipSuccess;
        for (float** p : {&l.b, &l.gb, &l.mb, &l.vb}) ok = ok && hipMalloc(p, l.cout * 4) == hipSuccess;
        ok = ok && hipMalloc(&l.wbf, l.nw * 2) == hipSuccess && hipMalloc(&l.wflip, l.nw * 2) == hipSuccess;
        if (!ok) {
            svo_cnn_destroy(n);
            return svo_fail(SVO_ERR_NOMEM, "CNN weights");
        }
        (void)hipMemcpy(l.w, w.data(), l.nw * 4, hipMemcpyHostToDevice);
        (void)hipMemcpy(l.b, b.data(), l.cout * 4, hipMemcpyHostToDevice);
        for (float* p : {l.mw, l.vw}) (void)hipMemset(p, 0, l.nw * 4);
        for (float* p : {l.mb, l.vb}) (void)hipMemset(p, 0, l.cout * 4);
    }
    int rc = refresh_bf16(n);
    if (rc) {
        svo_cnn_destroy(n);
        return rc;
    }
    HIPCHECK(hipStreamSynchronize(n->stream));
    *out = n;
    return SVO_OK;
}

// flat parameter vector: per layer W [9][cout][cin] then b [cout]
extern "C" int svo_cnn_get_params(svo_cnn* n, float* dst, size_t count)
{
    if (!n || !dst || count != svo_cnn_param_count()) return svo_fail(SVO_ERR_ARG, "need %zu floats", svo_cnn_param_count());
    HIPCHECK(hipStreamSynchronize(n->stream));
    for (auto& l : n->L) {
        HIPCHECK(hipMemcpy(dst, l.w, l.nw * 4, hipMemcpyDeviceToHost));
        dst += l.nw;
        HIPCHECK(hipMemcpy(dst, l.b, l.cout * 4, hipMemcpyDeviceToHost));
        dst += l.cout;
    }
    return SVO_OK;
}

extern "C" int svo_cnn_set_params(svo_cnn* n, const float* src, size_t count)
{
    if (!n || !src || count != svo_cnn_param_count()) return svo_fail(SVO_ERR_ARG, "need %zu floats", svo_cnn_param_count());
    HIPCHECK(hipStreamSynchronize(n->stream));
    for (auto& l : n->L) {
        HIPCHECK(hipMemcpy(l.w, src, l.nw * 4, hipMemcpyHostToDevice));
        src += l.nw;
        HIPCHECK(hipMemcpy(l.b, src, l.cout * 4, hipMemcpyHostToDevice));
        src += l.cout;
    }
    return refresh_bf16(n);
}

static unsigned conv_grid(int B, int H, int W) { return (unsigned)std::min(B * (H / TILE) * (W / TILE), 16384); }

int cnn_forward_dev(svo_cnn* n, int B, int H, int W, const uint32_t* bcount = nullptr)
{
    const unsigned g = conv_grid(B, H, W);
    hipLaunchKernelGGL((k_conv<16, 32, 0>), dim3(g), dim3(64), 0, n->stream, n->x0, n->L[0].wbf, n->L[0].b,
                       (const bf16*)nullptr, (void*)n->y[0], H, W, B, bcount);
    hipLaunchKernelGGL((k_conv<32, 32, 0>), dim3(g), dim3(64), 0, n->stream, n->y[0], n->L[1].wbf, n->L[1].b,
                       (const bf16*)nullptr, (void*)n->y[1], H, W, B, bcount);
    hipLaunchKernelGGL((k_conv<32, 32, 0>), dim3(g), dim3(64), 0, n->stream, n->y[1], n->L[2].wbf, n->L[2].b,
                       (const bf16*)nullptr, (void*)n->y[2], H, W, B, bcount);
    hipLaunchKernelGGL((k_conv<32, 16, 1>), dim3(g), dim3(64), 0, n->stream, n->y[2], n->L[3].wbf, n->L[3].b,
                       (const bf16*)nullptr, (void*)n->out, H, W, B, bcount);
    HIPCHECK(hipGetLastError());
    return SVO_OK;
}

// backward from n->out vs n->target/n->mask; gradients into L[].gw/gb
// (per-crop and per-block partials added in a fixed order)
int cnn_backward_dev(svo_cnn* n, int B, int H, int W)
{
    const size_t pix = (size_t)B * H * W;
    const unsigned g = conv_grid(B, H, W);
    const unsigned gdb = (unsigned)((pix + DB_PIXELS - 1) / DB_PIXELS);
    const size_t need = std::max((size_t)B * 9 * 32 * 32, (size_t)gdb * 32);
    if (need > n->cap_part) {
        (void)hipFree(n->part);
        n->part = nullptr;
        n->cap_part = 0;
        if (hipMalloc(&n->part, need * 4) != hipSuccess) return svo_fail(SVO_ERR_NOMEM, "gradient partials");
        n->cap_part = need;
    }
    HIPCHECK(hipMemsetAsync(n->loss, 0, 4, n->stream));
    hipLaunchKernelGGL(k_loss, dim3(nb(pix)), dim3(256), 0, n->stream, n->out, n->target, n->mask, n->dz16, n->loss,
                       (int)pix, n->mask_count);
    // weight and bias gradients of layer l from its input X and output gradient dZ
    auto dwdb = [&](int l, const bf16* X, const bf16* dZ) {
        CnnLayer& L = n->L[l];
        const dim3 gdw((unsigned)B, 9), bdw(64 * DW_SLICES);
        if (l == 3)
            hipLaunchKernelGGL((k_conv_dw<32, 16>), gdw, bdw, 0, n->stream, X, dZ, n->part, H, W);
        else if (l == 0)
            hipLaunchKernelGGL((k_conv_dw<16, 32>), gdw, bdw, 0, n->stream, X, dZ, n->part, H, W);
        else
            hipLaunchKernelGGL((k_conv_dw<32, 32>), gdw, bdw, 0, n->stream, X, dZ, n->part, H, W);
        hipLaunchKernelGGL(k_sum_parts, dim3(nb(L.nw)), dim3(256), 0, n->stream, (const float*)n->part, B,
                           (int)L.nw, L.gw);
        if (l == 3)
            hipLaunchKernelGGL((k_conv_db<16>), dim3(gdb), dim3(256), 0, n->stream, dZ, n->part, (int)pix);
        else
            hipLaunchKernelGGL((k_conv_db<32>), dim3(gdb), dim3(256), 0, n->stream, dZ, n->part, (int)pix);
        hipLaunchKernelGGL(k_sum_parts_tree, dim3((unsigned)L.cout), dim3(256), 0, n->stream, (const float*)n->part,
                           (int)gdb, L.cout, L.gb);
    };
    // layer 4 (32 -> 16)
    dwdb(3, n->y[2], n->dz16);
    hipLaunchKernelGGL((k_conv<16, 32, 2>), dim3(g), dim3(64), 0, n->stream, n->dz16, n->L[3].wflip,
                       (const float*)nullptr, n->y[2], (void*)n->dz32a, H, W, B, (const uint32_t*)nullptr);
    // layer 3
    dwdb(2, n->y[1], n->dz32a);
    hipLaunchKernelGGL((k_conv<32, 32, 2>), dim3(g), dim3(64), 0, n->stream, n->dz32a, n->L[2].wflip,
                       (const float*)nullptr, n->y[1], (void*)n->dz32b, H, W, B, (const uint32_t*)nullptr);
    // layer 2
    dwdb(1, n->y[0], n->dz32b);
    hipLaunchKernelGGL((k_conv<32, 32, 2>), dim3(g), dim3(64), 0, n->stream, n->dz32b, n->L[1].wflip,
                       (const float*)nullptr, n->y[0], (void*)n->dz32a, H, W, B, (const uint32_t*)nullptr);
    // layer 1 (16 -> 32)
    dwdb(0, n->x0, n->dz32a);
    HIPCHECK(hipGetLastError());
    return SVO_OK;
}

int cnn_adam_dev(svo_cnn* n, float lr)
{
    n->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)n->step), c2 = 1.f - powf(b2, (float)n->step);
    for (auto& l : n->L) {
        hipLaunchKernelGGL(k_adam, dim3(nb(l.nw)), dim3(256), 0, n->stream, l.w, l.gw, l.mw, l.vw, (int)l.nw, lr, b1,
                           b2, c1, c2, 1e-8f);
        hipLaunchKernelGGL(k_adam, dim3(nb(l.cout)), dim3(256), 0, n->stream, l.b, l.gb, l.mb, l.vb, l.cout, lr, b1,
                           b2, c1, c2, 1e-8f);
    }
    HIPCHECK(hipGetLastError());
    return refresh_bf16(n);
}

static int check_shape(int B, int H, int W)
{
    if (B <= 0 || H <= 0 || W <= 0 || H % TILE || W % TILE || (size_t)B * H * W > (1u << 26) || (H * W) % 16)
        return svo_fail(SVO_ERR_ARG, "batch %dx%dx%d: H and W must be positive multiples of 8", B, H, W);
    return SVO_OK;
}

// Raw forward on host data: x = [B][H][W][16] bf16 bits, out = [B][H][W][16] fp32.
extern "C" int svo_cnn_forward(svo_cnn* n, const uint16_t* x, int B, int H, int W, float* out)
{
    if (!n || !x || !out) return svo_fail(SVO_ERR_ARG, "null argument");
    int rc = check_shape(B, H, W);
    if (rc) return rc;
    HIPCHECK(hipSetDevice(n->device));
    const size_t pix = (size_t)B * H * W;
    rc = ensure_ws(n, pix);
    if (rc) return rc;
    HIPCHECK(hipMemcpyAsync(n->x0, x, pix * 16 * 2, hipMemcpyHostToDevice, n->stream));
    rc = cnn_forward_dev(n, B, H, W);
    if (rc) return rc;
    HIPCHECK(hipMemcpyAsync(out, n->out, pix * 16 * 4, hipMemcpyDeviceToHost, n->stream));
    HIPCHECK(hipStreamSynchronize(n->stream));
    return SVO_OK;
}

// Raw loss + gradients (no update): target [B][H][W][3], mask [B][H][W];
// grads written as the flat parameter vector layout; returns the loss.
extern "C" int svo_cnn_gradients(svo_cnn* n, const uint16_t* x, const float* target, const float* mask, int B, int H,
                                 int W, float* grads, size_t count, float* loss)
{
    if (!n || !x || !target || !mask || !grads || !loss) return svo_fail(SVO_ERR_ARG, "null argument");
    if (count != svo_cnn_param_count()) return svo_fail(SVO_ERR_ARG, "need %zu floats", svo_cnn_param_count());
    int rc = check_shape(B, H, W);
    if (rc) return rc;
    HIPCHECK(hipSetDevice(n->device));
    const size_t pix = (size_t)B * H * W;
    rc = ensure_ws(n, pix);
    if (rc) return rc;
    HIPCHECK(hipMemcpyAsync(n->x0, x, pix * 16 * 2, hipMemcpyHostToDevice, n->stream));
    HIPCHECK(hipMemcpyAsync(n->target, target, pix * 3 * 4, hipMemcpyHostToDevice, n->stream));
    HIPCHECK(hipMemcpyAsync(n->mask, mask, pix * 4, hipMemcpyHostToDevice, n->stream));
    float cnt = 0.f;
    for (size_t i = 0; i < pix; i++) cnt += mask[i];
    HIPCHECK(hipMemcpyAsync(n->mask_count, &cnt, 4, hipMemcpyHostToDevice, n->stream));
    if ((rc = cnn_forward_dev(n, B, H, W)) || (rc = cnn_backward_dev(n, B, H, W))) return rc;
    HIPCHECK(hipStreamSynchronize(n->stream));
    for (auto& l : n->L) {
        HIPCHECK(hipMemcpy(grads, l.gw, l.nw * 4, hipMemcpyDeviceToHost));
        grads += l.nw;
        HIPCHECK(hipMemcpy(grads, l.gb, l.cout * 4, hipMemcpyDeviceToHost));
        grads += l.cout;
    }
    HIPCHECK(hipMemcpy(loss, n->loss, 4, hipMemcpyDeviceToHost));
    return SVO_OK;
}

static FeatArgs feat_args(svo_cnn* n, svo_ctx* c, const svo_frame_params* p, int crop)
{
    FeatArgs a;
    memset(&a, 0, sizeof(a));
    a.words = c->d_cnn_words;
    a.coords = c->d_cnn_coords;
    a.res_x = c->res_x;
    a.res_y = c->res_y;
    a.crop = crop;
    a.tiles_x = c->n_tiles_x;
    a.m0x = p->m0[0]; a.m0y = p->m0[1]; a.m0z = p->m0[2];
    a.r20 = p->row2[0]; a.r21 = p->row2[1]; a.r22 = p->row2[2];
    a.x = n->x0;
    return a;
}

// One Adam step on n_crops crop x crop windows of the current frame pair:
// features from the warp context's last learned-frame CNN input (plane 0 as
// its gate left it), targets from the full context's image, loss on the
// pixels still holes after the warp frame's rays.  Both contexts must have
// rendered the same camera.  The CNN's stream waits for both contexts.
extern "C" int svo_cnn_train_step(svo_cnn* n, svo_ctx* warp, svo_ctx* full, const svo_frame_params* p, int n_crops,
                                  int crop, uint64_t seed, float lr, float* loss_out)
{
    if (!n || !warp || !full || !p) return svo_fail(SVO_ERR_ARG, "null argument");
    if (warp->res_x != full->res_x || warp->res_y != full->res_y || !warp->d_screen)
        return svo_fail(SVO_ERR_STATE, "contexts must share one resolution");
    if (!warp->cnn_input)
        return svo_fail(SVO_ERR_STATE, "the warp context has rendered no learned frame since its last resize");
    if (n_crops <= 0 || n_crops > 256 || crop < 16 || crop % 8 || crop > warp->res_x || crop > warp->res_y)
        return svo_fail(SVO_ERR_ARG, "bad crop batch %d x %d", n_crops, crop);
    HIPCHECK(hipSetDevice(n->device));
    int rc = ensure_ws(n, (size_t)n_crops * crop * crop);
    if (rc) return rc;
    if (n->max_origins < n_crops) {
        (void)hipFree(n->origins);
        n->origins = nullptr;
        HIPCHECK(hipMalloc(&n->origins, sizeof(int2) * 256));
        n->max_origins = 256;
    }
    std::vector<int2> org(n_crops);
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 0x1234567ull;
    for (auto& o : org) {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        o.x = (int)((st >> 33) % (uint64_t)(warp->res_x - crop + 1));
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        o.y = (int)((st >> 33) % (uint64_t)(warp->res_y - crop + 1));
    }
    // order after both contexts' queued frames
    hipEvent_t e1, e2;
    HIPCHECK(hipEventCreateWithFlags(&e1, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&e2, hipEventDisableTiming));
    (void)hipEventRecord(e1, warp->stream);
    (void)hipEventRecord(e2, full->stream);
    (void)hipStreamWaitEvent(n->stream, e1, 0);
    (void)hipStreamWaitEvent(n->stream, e2, 0);
    HIPCHECK(hipMemcpyAsync(n->origins, org.data(), sizeof(int2) * n_crops, hipMemcpyHostToDevice, n->stream));
    HIPCHECK(hipMemsetAsync(n->mask_count, 0, 4, n->stream));
    FeatArgs a = feat_args(n, warp, p, crop);
    a.mask_words = warp->d_screen + 2 * warp->N;  // holes left after the frame's rays
    a.image = full->d_image;
    a.origins = n->origins;
    a.max_crops = n_crops;
    a.target = n->target;
    a.mask = n->mask;
    a.mask_count = n->mask_count;
    hipLaunchKernelGGL(k_features, dim3(1024), dim3(256), 0, n->stream, a);
    HIPCHECK(hipGetLastError());
    if ((rc = cnn_forward_dev(n, n_crops, crop, crop)) || (rc = cnn_backward_dev(n, n_crops, crop, crop)) ||
        (rc = cnn_adam_dev(n, lr)))
        return rc;
    // the next frames of either context wait for this step's reads
    (void)hipEventRecord(e1, n->stream);
    (void)hipStreamWaitEvent(warp->stream, e1, 0);
    (void)hipStreamWaitEvent(full->stream, e1, 0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    if (loss_out) {
        HIPCHECK(hipStreamSynchronize(n->stream));
        HIPCHECK(hipMemcpy(loss_out, n->loss, 4, hipMemcpyDeviceToHost));
    }
    return SVO_OK;
}

// Inference on the gate's tile list (24x24 crops around each 16x16 tile of
// the CNN input planes) into the staged colours, queued on `stream`.
static TileArgs tile_args(svo_cnn* n)
{
    TileArgs a;
    memset(&a, 0, sizeof(a));
    for (int i = 0; i < 4; i++) {
        a.w[i] = n->L[i].wbf;
        a.b[i] = n->L[i].b;
    }
    return a;
}

static unsigned tile_grid(int max_tiles, int cus, int per_cu)
{
    return (unsigned)std::max(1, std::min(max_tiles, per_cu * cus));
}

// tile kernel variant: 0 = 4 waves, weights staged per layer (two workgroups
// per CU); 1 = 8 waves, weights resident (one workgroup per CU)
// (grid > 0: that many workgroups, at most one per tile; the workgroups loop
// over the tile list, so any count covers every tile)
static int launch_tiles(const TileArgs& a, int mode, int max_tiles, int cus, hipStream_t stream, int grid)
{
    const unsigned g1 = grid > 0 ? (unsigned)std::max(1, std::min(max_tiles, grid)) : tile_grid(max_tiles, cus, 1);
    const unsigned g2 = grid > 0 ? (unsigned)std::max(1, std::min(max_tiles, grid)) : tile_grid(max_tiles, cus, 2);
    if (mode == 1)
        hipLaunchKernelGGL((k_cnn_tiles<8, true>), dim3(g1), dim3(512), 0, stream, a);
    else
        hipLaunchKernelGGL((k_cnn_tiles<4, false>), dim3(g2), dim3(256), 0, stream, a);
    HIPCHECK(hipGetLastError());
    return SVO_OK;
}

static TileView tile_view(svo_ctx* c, const svo_frame_params* p, const uint32_t* tile_list,
                          const uint32_t* tile_count, int max_tiles)
{
    TileView v;
    memset(&v, 0, sizeof(v));
    v.words = c->d_cnn_words;
    v.coords = c->d_cnn_coords;
    v.res_x = c->res_x;
    v.res_y = c->res_y;
    v.tiles_x = c->n_tiles_x;
    v.tiles = tile_list;
    v.count = tile_count;
    v.max_tiles = max_tiles;
    v.m0x = p->m0[0]; v.m0y = p->m0[1]; v.m0z = p->m0[2];
    v.r20 = p->row2[0]; v.r21 = p->row2[1]; v.r22 = p->row2[2];
    v.staged = c->d_cnn_out;
    v.plane2 = c->d_screen + 2 * c->N;
    v.image = c->d_image;
    return v;
}

int cnn_fill_tiles(svo_cnn* n, svo_ctx* c, const svo_frame_params* p, const uint32_t* tile_list,
                   const uint32_t* tile_count, int max_tiles, hipStream_t stream, int grid)
{
    if (n->fused) {
        TileArgs a = tile_args(n);
        a.v = tile_view(c, p, tile_list, tile_count, max_tiles);
        return launch_tiles(a, n->tile_mode, max_tiles, n->n_cu, stream, grid);
    }
    int rc = ensure_ws(n, (size_t)max_tiles * 576);
    if (rc) return rc;
    hipStream_t keep = n->stream;
    n->stream = stream;  // run on the frame's stream
    FeatArgs a = feat_args(n, c, p, 24);
    a.tiles = tile_list;
    a.count = tile_count;
    a.max_crops = max_tiles;
    hipLaunchKernelGGL(k_features, dim3(2048), dim3(256), 0, n->stream, a);
    rc = cnn_forward_dev(n, max_tiles, 24, 24, tile_count);
    if (!rc) {
        hipLaunchKernelGGL(k_stage_fill, dim3(2048), dim3(256), 0, n->stream, n->out, tile_count, max_tiles,
                           c->d_cnn_out);
        if (hipGetLastError() != hipSuccess) rc = svo_fail(SVO_ERR_HIP, "CNN fill launch failed");
    }
    n->stream = keep;
    return rc;
}

size_t cnn_tile_view_bytes() { return sizeof(TileView); }

// Every view of a batch in one tile launch: the views' gate tile lists are
// concatenated through a table (host staging h_tab, device copy d_tab, both
// n_views entries; first: n_views + 1 words) written here and copied on
// `stream` ahead of the launch.  The per-layer kernels (fused mode off) run
// view by view.
int cnn_fill_tiles_batch(svo_cnn* n, svo_ctx* const* views, const svo_frame_params* params, int n_views,
                         void* h_tab, void* d_tab, uint32_t* first, hipStream_t stream)
{
    TileView* hv = (TileView*)h_tab;
    int max_all = 0;
    for (int i = 0; i < n_views; i++) {
        svo_ctx* c = views[i];
        const int mt = c->n_tiles_x * c->n_tiles_y;
        hv[i] = tile_view(c, &params[i], c->d_gate_tiles, c->d_gate, mt);
        max_all += mt;
    }
    if (!n->fused) {
        for (int i = 0; i < n_views; i++) {
            svo_ctx* c = views[i];
            RCC(cnn_fill_tiles(n, c, &params[i], c->d_gate_tiles, c->d_gate, c->n_tiles_x * c->n_tiles_y, stream));
        }
        HIPCHECK(hipMemcpyAsync(d_tab, h_tab, sizeof(TileView) * n_views, hipMemcpyHostToDevice, stream));
        return SVO_OK;
    }
    HIPCHECK(hipMemcpyAsync(d_tab, h_tab, sizeof(TileView) * n_views, hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_tile_first, dim3(1), dim3(1), 0, stream, (const TileView*)d_tab, n_views, first);
    HIPCHECK(hipGetLastError());
    TileArgs a = tile_args(n);
    a.views = (const TileView*)d_tab;
    a.first = first;
    a.n_views = n_views;
    return launch_tiles(a, n->tile_mode, max_all, n->n_cu, stream, 0);
}

// the staged colours of every view of a batch into their images (the table
// cnn_fill_tiles_batch copied to the device)
int cnn_apply_tiles_batch(const void* d_tab, int n_views, int max_tiles, hipStream_t stream)
{
    const unsigned g = (unsigned)std::max(1, std::min(2048, (max_tiles * 256 + 255) / 256));
    hipLaunchKernelGGL(k_cnn_apply_views, dim3(g, (unsigned)n_views), dim3(256), 0, stream, (const TileView*)d_tab);
    HIPCHECK(hipGetLastError());
    return SVO_OK;
}

int cnn_apply_tiles(svo_ctx* c, const uint32_t* tile_list, const uint32_t* tile_count, int max_tiles,
                    hipStream_t stream)
{
    hipLaunchKernelGGL(k_cnn_apply, dim3(2048), dim3(256), 0, stream, c->d_cnn_out, tile_list, tile_count, max_tiles,
                       c->d_screen + 2 * c->N, c->d_image, c->res_x, c->res_y, c->n_tiles_x);
    HIPCHECK(hipGetLastError());
    return SVO_OK;
}

// The fused tile network on explicit 24x24 crops: x [n][24][24][16] bf16
// features -> out [n][16][16][3] float (the centre 16x16, RGB).
extern "C" int svo_cnn_forward_tiles(svo_cnn* n, const uint16_t* x, int count, float* out)
{
    if (!n || !x || !out || count < 1 || count > (1 << 20)) return svo_fail(SVO_ERR_ARG, "bad tile batch");
    HIPCHECK(hipSetDevice(n->device));
    if ((size_t)count > n->cap_tiles) {
        (void)hipFree(n->xt);
        (void)hipFree(n->ot);
        n->xt = nullptr;
        n->ot = nullptr;
        n->cap_tiles = 0;
        if (hipMalloc(&n->xt, (size_t)count * 576 * 16 * 2) != hipSuccess ||
            hipMalloc(&n->ot, (size_t)count * 256 * 3 * 4) != hipSuccess)
            return svo_fail(SVO_ERR_NOMEM, "tile batch of %d", count);
        n->cap_tiles = (size_t)count;
    }
    HIPCHECK(hipMemcpyAsync(n->xt, x, (size_t)count * 576 * 16 * 2, hipMemcpyHostToDevice, n->stream));
    TileArgs a = tile_args(n);
    a.x = n->xt;
    a.out = n->ot;
    a.n = count;
    RCC(launch_tiles(a, n->tile_mode, count, n->n_cu, n->stream, 0));
    HIPCHECK(hipMemcpyAsync(out, n->ot, (size_t)count * 256 * 3 * 4, hipMemcpyDeviceToHost, n->stream));
    HIPCHECK(hipStreamSynchronize(n->stream));
    return SVO_OK;
}

extern "C" int svo_cnn_set_fused(svo_cnn* n, int mode)
{
    if (!n) return svo_fail(SVO_ERR_ARG, "null network");
    if (mode < 0 || mode > 2) return svo_fail(SVO_ERR_ARG, "fused mode %d outside 0..2", mode);
    n->fused = mode != 0;
    n->tile_mode = mode == 2 ? 1 : 0;
    return SVO_OK;
}
