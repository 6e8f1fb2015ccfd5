// svo_device.hip -- gfx950 kernels and the C-ABI runtime of libsvo.
//
// Kernels (all hand-written for CDNA wave64):
//   k_rays          persistent ray traversal: every wave pulls rays from a
//                   device work queue (hole list then refresh tile), refills
//                   idle lanes with a ballot-ranked atomicAdd, keeps its
//                   per-ray node stack in LDS (one column per lane);
//                   restates kernel.cl:32-214 (fetchChildOffset, fetchColor,
//                   CAST_RAY) with the same float op order.
//   k_proj_scatter  warp of cache planes (kernel.cl:472-592): one 64-bit
//                   atomicMin per source of key = depth24|source index|colour,
//                   so the first source with the smallest depth key wins.
//   k_proj_resolve  writes winners' screen word and coords, clears the key,
//                   fills the rest of plane 0 with the hole word (fuses the
//                   per-frame ocl_memset of plane 0).
//   k_counthole /
//   k_writeids      one wave per 16x16 block, one lane per 2x2 quad, ballot
//                   of the hole quads (kernel.cl:234-274, :298-340).
//   k_sumids        single-workgroup scan, total stays on the device in
//                   ids[0] (kernel.cl:276-296; no host readback).
//   k_copy_plane    plane 0 -> plane k (screen + coords), 16-byte accesses.
//   k_fill_color    fillhole2 on a pre-filter snapshot fused with colorize
//                   (kernel.cl:404-470, :944-974).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "svo_ctx.h"
#include "svo_internal.h"

#define HOLE 0xffffff00u

namespace {

constexpr int RAY_BLOCK = 256;
// Trips per wave-uniform exit check.  Measured on the flythrough (lone
// ray / warp frames/s / full frames/s): 1: 0.66-0.68 ms / 2,360-2,400 /
// 880; 2: 0.63 ms / 2,480-2,520 / 920; 4: 0.61-0.63 ms / 2,520-2,560 / 940;
// 8: 0.60-0.61 ms / 2,530-2,580 / 940-950; 16: the same as 8.
#ifndef SVO_TRIP_UNROLL
#define SVO_TRIP_UNROLL 8
#endif

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ uint32_t bc8(uint32_t v) { return (uint32_t)__popc(v & 0xffu); }

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz)
{
    return (ax * bx + ay * by) + az * bz;
}

__device__ __forceinline__ float len3(float x, float y, float z)
{
    return sqrtf((x * x + y * y) + z * z);
}

// OpenCL sign(): +-1 for nonzero values, the zero itself for +-0
__device__ __forceinline__ float cl_sign(float v)
{
    return v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : v);
}

// Octree reads go through a raw buffer resource (buffer_load_dword with an
// SGPR descriptor): 32-bit offsets, and reads at or past num_records return
// 0 in hardware.  Dword 3 = 0x00020000 is the gfx9-family raw-buffer format.
typedef __amdgpu_buffer_rsrc_t OctBuf;
__device__ __forceinline__ OctBuf oct_buffer(const uint32_t* oct, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc((void*)oct, (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ uint32_t oct_load(OctBuf b, uint32_t i)
{
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(b, (int)(i * 4u), 0, 0);
}

// Octree word i: the first cache_words words (the breadth-first normal
// records of the top levels) come from the LDS copy, the rest from HBM.
__device__ __forceinline__ uint32_t oct_word(OctBuf oct, const uint32_t* cache, uint32_t cache_words, uint32_t i)
{
    return i < cache_words ? cache[i] : oct_load(oct, i);
}

// kernel.cl:64-112
__device__ uint32_t fetch_color(OctBuf oct, const uint32_t* cache, uint32_t cache_words,
                                uint32_t a_node, uint32_t a_node_before, uint32_t a_node_before2,
                                uint32_t local_root, int rek, uint32_t child_test)
{
    uint32_t n = a_node >> 9;
    uint32_t nadd = 8;
    if (rek == 1) {
        n = (a_node_before >> 9) & 15u;
        uint32_t ofs = (a_node_before2 >> 9) + local_root;
        nadd = bc8(a_node_before2 & 255u);
        for (uint32_t i = 1; i < n; i++) nadd += bc8((oct_load(oct, ofs + (i >> 2)) >> ((i & 3) << 3)) & 255u);
        nadd += bc8(a_node_before & ((child_test << 1) - 1));
        return oct_load(oct, ofs + (nadd >> 2)) >> ((nadd & 3) << 3);
    }
    if (a_node & 256u) {
        if (!(a_node_before & 256u)) {
            local_root = n << 6;
            n = 0;
        }
        nadd = local_root;
        if (rek == 2) {
            uint32_t ofs = (a_node_before >> 9) + local_root;
            nadd = 1 + bc8(a_node_before & 255u);
            for (uint32_t i = 1; i < n; i++) nadd += bc8((oct_load(oct, ofs + (i >> 2)) >> ((i & 3) << 3)) & 255u);
            return oct_load(oct, ofs + (nadd >> 2)) >> ((nadd & 3) << 3);
        }
    }
    return oct_word(oct, cache, cache_words, n + nadd);
}

// Per-view frame description shared by every kernel.  Single-view calls pass
// it by value; a batch passes a device table and each block (grid z) or ray
// (ray kernel) picks its view.
struct View {
    uint32_t* screen;          // 4 planes x N words
    float4* coords;            // 4 planes x N float4
    uint32_t* ids;             // [S counts | S offsets | hole list]
    uint32_t* image;           // colorized output, or null
    unsigned long long* keys;  // warp keys (N)
    const uint32_t* list;       // ray pixel list (x | y << 16), or null
    const uint32_t* list_count; // its length, read on the device
    const uint32_t* list2;       // second ray pixel list, or null
    const uint32_t* list2_count; // its length, read on the device
    int res_x, res_y;
    uint32_t N, S;
    float m0x, m0y, m0z;
    float r00, r01, r02, r10, r11, r12, r20, r21, r22;  // proj rows
    float cx0, cx1, cx2, cy0, cy1, cy2, cz0, cz1, cz2;  // ray columns
    float fovx, fovy;
    int plane_lo, n_planes;                // proj source planes
    int add_x, add_y, ext_x, ext_y, tiles_x;  // refresh rectangle, tiles_x blocks per row
    uint32_t n_rect;
    int blk_shift;  // its rays are handed out in blocks of 2^blk_shift x 2^(6-blk_shift) pixels
    uint32_t* ray_image;                   // colorize ray words straight into this, or null
    int edge_ok;  // list pixels may be (res_x, y) or (x, res_y) (quad rays)
    const uint32_t* rect_skip;   // per 8x8 block: nonzero = raycast from the list instead, or null
    uint32_t* block_long;        // per 8x8 block: set to 1 by a ray of >= long_trips trips, or null
    uint32_t long_trips;
    // learned frames: destination of the CNN input copy (plane 0 after the gate)
    uint32_t* cnn_words;
    float4* cnn_coords;
    // learned-fill gate of this view: counters, CNN tile list, gate pixel list
    uint32_t* gate;
    uint32_t* gate_tiles;
    uint32_t* gate_rays;
    int gate_tiles_x, gate_n_tiles;
    float gate_depth;  // > 0: a tile with a reprojected pixel farther than this is raycast whole
    // pretrace of warp frames: a ray launch on a second stream traces part
    // of a frame's rays ahead (into one of a few result sets, each pixel
    // stamped with the frame's id); the frame's own ray launch then copies
    // the stored result for every ray whose pixel carries its id
    const uint32_t* pre_stamp;  // the frame's set: stamps, words, positions (or null)
    const uint32_t* pre_word;
    const float4* pre_xyz;
    uint32_t* stamp_out;        // pretrace launches: stamp per traced pixel (or null)
    uint32_t pre_id;
    // with a pretrace the frame's rays run as two launches: pre_mode 1 traces
    // the rays whose pixel is on neither part's pretrace list (bucket byte
    // kPreNone; part 0's list may be absent) while the pretrace still runs,
    // then pre_mode 2 copies the pretraced results; pre_mode 0 does both
    const uint8_t* pre_bkt0;
    const uint8_t* pre_bkt1;
    int pre_mode;
    uint32_t* trips_out;        // per pixel: loop trips of the last ray traced there (or null)
    float4* snap_out;           // warp frames' resolve: plane-0 positions after the warp, holes as NaN (or null)
};

constexpr uint32_t QUAD_NONE = 0xffffffffu;  // list entry without a ray
constexpr uint8_t kPreNone = 0xff;           // pretrace list passes: the pixel is not traced ahead
constexpr uint32_t NO_BLOCK = 0xffffffffu;

#define VIEW_OF(idx) const View& V = views ? views[(idx)] : v

struct RayArgs {
    View v;                       // single view
    const View* views;            // batch table (device), or null
    const uint32_t* view_first;   // batch: first ray of each view, [n_views + 1]
    int n_views;
    const uint32_t* oct;
    uint32_t root;
    int depth;
    uint32_t* queue;
    int slot;
    int refill;
    int iters;
    unsigned long long* ray_count;  // rays traced, accumulated across launches
    uint32_t* max_trips;            // loop trips of the longest ray, a max across launches
    unsigned long long* merged;     // rays taken from a pretrace, accumulated
    // split hand-out (list order only): the first *split_count rays go to the
    // grid's first split_waves waves, at most 2^split_shift lanes of a wave
    // at a time, from their own counters split_queue[0..1]; the other waves
    // take the rest from queue
    const uint32_t* split_count;    // or null
    uint32_t* split_queue;
    uint32_t split_waves;
    int split_shift;
    uint32_t cache_words;           // octree prefix held in LDS
    uint32_t cache_last;            // max(cache_words, 1) - 1
    uint32_t oct_bytes;             // buffer extent of the octree (words + pad)
    int spread_shift;               // log2 of the spread group (>= lanes_shift: list order)
    int spread_fit;                 // spread only when all rays get a lane at once
    int lanes_shift;                // log2 of the lanes of a wave that take rays (6: all 64)
};

struct Ray {
    float px, py, pz;
    float g0y, g0z, g1x, g1z, g2x, g2y;
    float lx, ly, lz;
    float dist, lodsw;
    int ix, iy, iz;
    int rek, lod, sxyz;
    uint32_t node, before, lroot, test;
    uint32_t pix;
    uint32_t guard;
    uint32_t desync;  // nonzero: a step left stale nodes on the stack (no fused trips from then on)
    int view;
    uint32_t blk;  // 8x8 screen block of the ray (long-ray flags of full frames), or NO_BLOCK
};

__device__ __forceinline__ uint32_t colorize_word(uint32_t word)
{
    // kernel.cl:959-971
    const int a = (int)word;
    const int c = a & 3;
    const float i = (float)(a & (255 - 7));
    float tr = 1.0f, tg = 1.0f, tb = 1.0f;
    if (c == 1) { tr = 1.0f; tg = 0.7f; tb = 0.3f; }
    else if (c == 2) { tr = 1.5f; tg = 0.8f; tb = 0.1f; }
    else if (c == 3) { tr = 0.2f; tg = 0.8f; tb = 0.2f; }
    int r = (int)(i * tr); if (r > 255) r = 255;
    int g = (int)(i * tg); if (g > 255) g = 255;
    int b = (int)(i * tb); if (b > 255) b = 255;
    return (uint32_t)(b + g * 256 + r * 65536);
}

__device__ __forceinline__ uint32_t view_rays(const View& V)
{
    // a pixel list holds at most N entries
    return (V.list ? min(*V.list_count, V.N) : 0u) + (V.list2 ? min(*V.list2_count, V.N) : 0u) + V.n_rect;
}

// Ray k of a view (its pixel lists, then its rectangle in 64-pixel blocks):
// false when it has none; blk = its 8x8 screen block for the long-ray flags
// of full frames (NO_BLOCK when not needed).
__device__ __forceinline__ bool ray_pixel(const View& V, uint32_t k, int& x, int& y, uint32_t& blk)
{
    const uint32_t H = V.list ? min(*V.list_count, V.N) : 0u;
    const uint32_t H2 = V.list2 ? min(*V.list2_count, V.N) : 0u;
    if (k < H + H2) {
        const uint32_t w = (k < H) ? V.list[k] : V.list2[k - H];
        if (w == QUAD_NONE) return false;
        x = (int)(w & 0xffffu);
        y = (int)(w >> 16);
        blk = V.block_long ? (uint32_t)((y >> (6 - V.blk_shift)) * V.tiles_x + (x >> V.blk_shift)) : NO_BLOCK;
    } else {
        k -= H + H2;
        const uint32_t t = k >> 6, w = k & 63u;
        if (V.rect_skip && V.rect_skip[t]) return false;
        blk = t;
        int tx = (int)(t % (uint32_t)V.tiles_x), ty = (int)(t / (uint32_t)V.tiles_x);
        const int bs = V.blk_shift;
        int lx = (tx << bs) + (int)(w & ((1u << bs) - 1u)), ly = (ty << (6 - bs)) + (int)(w >> bs);
        if (lx >= V.ext_x || ly >= V.ext_y) return false;
        x = V.add_x + lx;
        y = V.add_y + ly;
    }
    // quad rays may pick the pixel one past the row or frame (their flat
    // offset then lies in the next row or in plane 1, inside the buffers)
    return V.edge_ok || (x < V.res_x && y < V.res_y);
}

// Ray set-up: kernel.cl:883-908 and CAST_RAY prologue :115-150.  Returns
// RAY_NONE (no ray), RAY_TRACE, or RAY_MERGED (the pixel's result came from
// the frame's pretrace and has been written).
constexpr int RAY_NONE = 0, RAY_TRACE = 1, RAY_MERGED = 2;
__device__ __forceinline__ int init_ray(Ray& r, uint32_t k, const RayArgs& a)
{
    int vi = 0;
    if (a.views) {  // last view whose first ray is <= k
        int lo = 0, hi = a.n_views - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (a.view_first[mid] <= k) lo = mid;
            else hi = mid - 1;
        }
        vi = lo;
        k -= a.view_first[vi];
    }
    const View& V = a.views ? a.views[vi] : a.v;
    int x, y;
    if (!ray_pixel(V, k, x, y, r.blk)) return RAY_NONE;
    r.pix = (uint32_t)(y * V.res_x + x);
    r.view = vi;
    if (V.pre_mode == 1 &&
        (V.pre_bkt1[r.pix] != kPreNone || (V.pre_bkt0 && V.pre_bkt0[r.pix] != kPreNone)))
        return RAY_NONE;  // on a pretrace list: copied by the pre_mode 2 launch
    if (V.pre_mode != 1 && V.pre_stamp && V.pre_stamp[r.pix] == V.pre_id) {
        // traced ahead with this frame's camera: the same word and position
        // finish_ray would write
        const uint32_t word = V.pre_word[r.pix];
        const float4 q = V.pre_xyz[r.pix];
        V.screen[r.pix] = word;
        float* c = (float*)V.coords + 4u * r.pix;
        c[0] = q.x;
        c[1] = q.y;
        c[2] = q.z;
        if (V.ray_image) V.ray_image[r.pix] = colorize_word(word);
        return RAY_MERGED;
    }
    if (V.pre_mode == 2) return RAY_NONE;  // traced by the pre_mode 1 launch

    const int D = a.depth;
    const float SCALE_MAX = (float)(1 << (D + 1));
    float d1x = ((float)(x - V.res_x / 2) + 0.5f) * V.fovx / (float)V.res_y;
    float d1y = ((float)(y - V.res_y / 2) + 0.5f) * V.fovy / (float)V.res_y;
    float dx = dot3(d1x, d1y, 1.0f, V.cx0, V.cx1, V.cx2);
    float dy = dot3(d1x, d1y, 1.0f, V.cy0, V.cy1, V.cy2);
    float dz = dot3(d1x, d1y, 1.0f, V.cz0, V.cz1, V.cz2);
    float px = V.m0x * 16.0f, py = V.m0y * 16.0f, pz = V.m0z * 16.0f;
    const float eps = 0x1p-20f;
    if (fabsf(dx) < eps) dx = cl_sign(dx) * eps;
    if (fabsf(dy) < eps) dy = cl_sign(dy) * eps;
    if (fabsf(dz) < eps) dz = cl_sign(dz) * eps;
    int s = 0;
    if (dx < 0) { s |= 1; px = SCALE_MAX - px; dx = -dx; }
    if (dy < 0) { s |= 2; py = SCALE_MAX - py; dy = -dy; }
    if (dz < 0) { s |= 4; pz = SCALE_MAX - pz; dz = -dz; }
    r.g0y = dy / dx; r.g0z = dz / dx;
    r.g1x = dx / dy; r.g1z = dz / dy;
    r.g2x = dx / dz; r.g2y = dy / dz;
    r.lx = len3(1.0f, r.g0y, r.g0z);
    r.ly = len3(r.g1x, 1.0f, r.g1z);
    r.lz = len3(r.g2x, r.g2y, 1.0f);
    r.px = px; r.py = py; r.pz = pz;
    r.ix = (int)px; r.iy = (int)py; r.iz = (int)pz;
    r.sxyz = s;
    r.rek = D;
    r.lod = 1;
    r.dist = 0.0f;
    r.lodsw = (float)(V.res_x * 1 * 2);
    r.node = a.root;
    r.before = a.root;
    r.lroot = 0;
    r.test = 0;
    r.guard = 0;
    r.desync = 0u;
    return RAY_TRACE;
}

// LDS: node stack, rows 0..D+1 x RAY_BLOCK lanes, then the node cache.
// Row `rek` (1..D-1) holds the ray's node at depth D-rek; rows D and D+1 hold
// the root.  So STK(rek) is always the node being tested and STK(rek+1) its
// parent, and a pop reads both rows without a range check: the reference's
// rows D and D+1 (kernel.cl:635-637) belong to no node of the ray and are
// read only where they hold the root (rek == D) or where the value is not
// consulted (a normal depth-1 node's parent, rek == D-1, depths >= 8).
#define STK(rek) stk[(rek) * RAY_BLOCK + tid]
constexpr int STK_ROWS_EXTRA = 2;  // rows D and D+1

// One trip of the CAST_RAY do-while loop (kernel.cl:152-211).  Returns true
// when the loop exits.  A lone wave in the tail of a launch issues one
// instruction per 4 cycles, so the trip is written for a short path with few
// exec-mask regions: one divergent branch (descend or step), selects inside
// each body, the pop as unconditional stack reads at the new level (no pop
// re-reads the current node and its parent, which the stack also holds), the
// pop test as one compare (floor(log2 m) >= rek  <=>  m >= 2^rek, m > 0), and
// the child fetch of fetchChildOffset (kernel.cl:32-62) as one word read.
template <bool kCache>
__device__ __forceinline__ uint32_t step_ray(Ray& r, const RayArgs& a, OctBuf ob, uint32_t* stk,
                                             const uint32_t* cache, int tid)
{
    const int D = a.depth;
    const int rek = r.rek;
    const int p0x = r.ix >> rek, p0y = r.iy >> rek, p0z = r.iz >> rek;
    const int ni = ((p0x & 1) | ((p0y & 1) << 1) | ((p0z & 1) << 2)) ^ r.sxyz;
    const uint32_t node = r.node;
    const uint32_t test = node & (1u << ni);
    r.test = test;
    uint32_t done;  // 0/1 as a value (not a lane mask): no mask merges at the join
    if (test) {
        // ---- descend (kernel.cl:42-61, :169-178).  Block nodes address their
        // children from local_root, set on entry from a normal parent; at
        // rek == 2 the depth-(D-1) ref is synthesised from the byte node.  At
        // rek == 1 the fetched word is never consumed (the hit exits and
        // fetchColor reads only node_before and the stack).
        const uint32_t n0 = node >> 9;
        const bool blk = (node & 256u) != 0;
        const bool enter = blk && !(r.before & 256u);
        const uint32_t lroot = enter ? (n0 << 6) : r.lroot;
        const uint32_t nadd = (uint32_t)__popc(node & ((test << 1) - 1u));
        const bool synth = blk && rek == 2;
        const uint32_t off = synth ? (nadd >> 2) : nadd;
        const uint32_t idx = (enter ? 0u : n0) + (blk ? lroot + off : (uint32_t)ni);
        uint32_t w;
        if (kCache && idx < a.cache_words)
            w = cache[idx];
        else
            w = oct_load(ob, idx);
        const uint32_t child = synth ? ((w >> ((nadd & 3u) << 3)) & 255u) + 256u + (nadd << 9) : w;
        r.lroot = lroot;
        r.before = node;
        r.node = child;
        done = (uint32_t)(rek <= r.lod);
        // the stack write is unconditional: when the ray stops here (rek-1 is
        // then 0..lod-1) the row is never read again (finish reads row 2 only
        // for a ray at rek 1, and row 0 never)
        STK(rek - 1) = child;
        r.rek = done ? rek : rek - 1;
    } else {
        // ---- DDA step to the nearest cell face (kernel.cl:181-194)
        const float mx = (float)((p0x + 1) << rek) - r.px;
        const float my = (float)((p0y + 1) << rek) - r.py;
        const float mz = (float)((p0z + 1) << rek) - r.pz;
        const float ex = mx * r.lx, ey = my * r.ly, ez = mz * r.lz;
        const bool c1 = ey < ex;
        float dist = c1 ? ey : ex;
        const bool c2 = ez < dist;
        dist = c2 ? ez : dist;
        const float m = c2 ? mz : (c1 ? my : mx);
        const float gx = c2 ? r.g2x : (c1 ? r.g1x : 1.0f);
        const float gy = c2 ? r.g2y : (c1 ? 1.0f : r.g0y);
        const float gz = c2 ? 1.0f : (c1 ? r.g1z : r.g0z);
        r.px += gx * m;
        r.py += gy * m;
        r.pz += gz * m;
        const int nx = (int)r.px, ny = (int)r.py, nz = (int)r.pz;
        r.dist += dist;
        // pop level (kernel.cl:197-209): the XOR of the three axes' changes
        const int DAND = (1 << D) - 1;
        const int x0ry = r.iy ^ ny;
        const uint32_t xm = (uint32_t)((r.ix ^ nx) ^ x0ry ^ (r.iz ^ nz)) & (uint32_t)DAND;
        const uint32_t chg = (uint32_t)((r.ix ^ nx) | x0ry | (r.iz ^ nz)) & ((2u << D) - 1u);  // bits 0..D
        r.ix = nx;
        r.iy = ny;
        r.iz = nz;
        const bool far = r.dist > 400000.0f;
        const bool pop = !far && xm >= (1u << rek);
        const int nr = pop ? 32 - __clz((int)xm) : rek;  // rek_new + 1, or stay
        // LOD switch, evaluated on pops only (kernel.cl:209)
        const bool l = pop && r.dist > r.lodsw;
        r.lodsw = l ? r.lodsw * 2.0f : r.lodsw;
        r.lod += l ? 1 : 0;
        const bool ycross = ((uint32_t)x0ry >> (D + 1) & 1u) != 0u;
        // The reference's next trip, fused into this one.  After a pop to
        // level nr that left every position bit from nr to D unchanged (the
        // XOR level can sit below the highest changed bit, and bit D is
        // masked out of it, so this is checked on the OR of the changes),
        // the cell at level nr holding the ray is the one it came from: the
        // next trip tests that child, finds it set, and descends into it
        // (kernel.cl:160-179), fetching the ref the stack still holds in row
        // nr-1 -- or stops there as a hit when nr <= lod.  local_root already
        // belongs to that path.  Counted as two trips; not fused when the
        // second trip would pass the trip cap.
        // A step that changed a position bit above the level it resumes at
        // (the XOR hid it) leaves the nodes there, and the stack rows below,
        // describing cells the ray has left; the reference then carries on
        // with those stale nodes, and so does this kernel, but no later
        // step of the ray is fused.
        r.desync |= chg >> (nr + 1);
        const bool fuse = pop && r.desync == 0u && !ycross && (chg >> nr) == 0u && r.guard + 2u <= SVO_MAX_RAY_ITERS;
        const bool hit = fuse && nr <= r.lod;
        const int base = fuse ? nr - 1 : nr;
        r.node = STK(base);
        r.before = STK(base + 1);
        r.rek = (fuse && !hit) ? nr - 1 : nr;
        r.guard += fuse ? 1u : 0u;
        // a hit's child_test is consumed by fetchColor at level 1 only, and
        // nr >= 2 here
        done = (uint32_t)far | (uint32_t)ycross | (uint32_t)hit;
    }
    return done;
}

__device__ __forceinline__ void finish_ray(Ray& r, const RayArgs& a, OctBuf ob, uint32_t* stk, const uint32_t* cache,
                                           int tid)
{
    const float SCALE_MAX = (float)(1 << (a.depth + 1));
    float px = r.px, py = r.py, pz = r.pz;
    if (r.sxyz & 1) px = SCALE_MAX - px;
    if (r.sxyz & 2) py = SCALE_MAX - py;
    if (r.sxyz & 4) pz = SCALE_MAX - pz;
    uint32_t b2 = (r.rek == 1) ? STK(2) : 0u;
    uint32_t col = fetch_color(ob, cache, a.cache_words, r.node, r.before, b2, r.lroot, r.rek, r.test);
    uint32_t word = 0xff000000u + col;
    const View& V = a.views ? a.views[r.view] : a.v;
    V.screen[r.pix] = word;
    float* c = (float*)V.coords + 4u * r.pix;
    c[0] = px / 16.0f;
    c[1] = py / 16.0f;
    c[2] = pz / 16.0f;
    if (V.ray_image) V.ray_image[r.pix] = colorize_word(word);
    if (V.block_long && r.blk != NO_BLOCK && r.guard + 1u >= V.long_trips) V.block_long[r.blk] = 1u;
    if (V.stamp_out) V.stamp_out[r.pix] = V.pre_id;
    if (V.trips_out) V.trips_out[r.pix] = r.guard;
}

// lanes 0, 64/K, 2*64/K, ... of a wave (K = 2^ks lanes)
__device__ __forceinline__ unsigned long long lane_pattern(int ks)
{
    unsigned long long m = 0;
    const int step = 64 >> ks;
    for (int l = 0; l < 64; l += step) m |= 1ull << l;
    return m;
}

template <bool kCache>
__global__ __launch_bounds__(RAY_BLOCK) void k_rays(RayArgs a)
{
    extern __shared__ uint32_t smem[];
    uint32_t* stk = smem;
    uint32_t* cache = smem + (a.depth + STK_ROWS_EXTRA) * RAY_BLOCK;
    const int tid = threadIdx.x;
    for (uint32_t i = (uint32_t)tid; i < a.cache_words; i += RAY_BLOCK) cache[i] = a.oct[i];
    STK(a.depth) = a.root;  // no ray writes rows D and D+1
    STK(a.depth + 1) = a.root;
    __syncthreads();
    const OctBuf ob = oct_buffer(a.oct, a.oct_bytes);
    const int lane = __lane_id();
    const uint32_t total = a.views ? a.view_first[a.n_views] : view_rays(a.v);
    // Spread order for the pixel-list rays (the first P rays; in a batch all
    // rays) in groups of G = 2^spread_shift: queue slot k < 64 R (wave w =
    // k / 64, lane l = k % 64) takes ray ((l / G) * R + w) * G + l % G,
    // R = ceil(P / 64); the refresh-rectangle rays that follow keep their
    // 8x8-block order (neighbouring rays there take the same path, so a wave
    // of them stays coherent).  The 64 / G groups of a wave lie R groups apart
    // in the list, so the long hole rays of one screen region (near the
    // horizon) land in different waves; a wave then rarely runs many of them
    // through both branch bodies.  G = 64 is list order.
    // Only K = 2^lanes_shift lanes of a wave take rays (every 64/K-th lane):
    // with K < 64 the same rays occupy more waves, so fewer long rays share
    // one.  The spread mapping then works on K-lane queue waves.
    const int ks = a.lanes_shift;
    const uint32_t K = 1u << ks;
    const unsigned long long elig = ks >= 6 ? ~0ull : lane_pattern(ks);
    const uint32_t P = (a.views && a.n_views > 1) ? total : total - (a.views ? a.views[0].n_rect : a.v.n_rect);
    const uint32_t rows = (P + K - 1u) >> ks;
    const bool spread = !a.split_count && a.spread_shift < ks &&
                        (!a.spread_fit || total <= gridDim.x * (uint32_t)(RAY_BLOCK / 64) * K);
    const uint32_t qtotal = spread ? rows * K + (total - P) : total;
    const uint32_t gmask = (1u << a.spread_shift) - 1u;
    if (blockIdx.x == 0 && tid == 0) a.queue[a.slot ^ 1] = 0u;  // next launch's counter
    // this wave's queue: positions k < qn map to rays lo + k (list order) or
    // through the spread order
    uint32_t* q = a.queue + a.slot;
    uint32_t lo = 0, qn = qtotal;
    unsigned long long weli = elig;
    uint32_t Kw = K;
    if (a.split_count) {
        const uint32_t ns = min(*a.split_count, total);
        const uint32_t wid = blockIdx.x * (uint32_t)(RAY_BLOCK / 64) + (uint32_t)(tid >> 6);
        if (blockIdx.x == 0 && tid == 0) a.split_queue[a.slot ^ 1] = 0u;
        // as many split waves as the split rays need, at most split_waves
        const uint32_t sw = min(a.split_waves, (ns + (1u << a.split_shift) - 1u) >> a.split_shift);
        if (wid < sw) {
            q = a.split_queue + a.slot;
            qn = ns;
            Kw = 1u << a.split_shift;
            weli = a.split_shift >= 6 ? ~0ull : lane_pattern(a.split_shift);
        } else {
            lo = ns;
            qn = total - ns;
        }
    }
    const int refill = min(a.refill, (int)Kw);

    Ray r;
    bool active = false, exhausted = false;
    uint32_t traced = 0;   // lane 0 of each wave: rays started by the wave
    uint32_t merged = 0;   // lane 0 of each wave: pretraced rays the wave copied
    uint32_t longest = 0;  // trips of the longest ray this lane finished
    for (;;) {
        const unsigned long long idle_all = __ballot(!active);
        const unsigned long long idle = idle_all & weli;
        const int n_idle = __popcll(idle);
        if (!exhausted && (n_idle >= refill || n_idle == (int)Kw)) {
            const int leader = __ffsll((long long)idle) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(q, (uint32_t)n_idle);
            base = __shfl(base, leader);
            if (base + (uint32_t)n_idle >= qn) exhausted = true;
            bool took = false;
            if ((idle >> lane) & 1ull) {
                const int rank = __popcll(idle & ((1ull << lane) - 1ull));
                const uint32_t k = base + (uint32_t)rank;
                const uint32_t l = k & (K - 1u);
                uint32_t kr = lo + k;
                bool valid = k < qn;
                if (spread) {
                    if (k < rows * K) {
                        kr = (((l >> a.spread_shift) * rows + (k >> ks)) << a.spread_shift) | (l & gmask);
                        valid = kr < P;
                    } else {
                        kr = P + (k - rows * K);
                    }
                }
                if (valid) {
                    const int st = init_ray(r, kr, a);
                    active = st == RAY_TRACE;
                    took = st == RAY_MERGED;
                }
            }
            traced += (uint32_t)__popcll(__ballot(active)) - (uint32_t)__popcll(~idle_all);
            merged += (uint32_t)__popcll(__ballot(took));
        }
        if (__ballot(active) == 0ull) {
            if (exhausted) break;
            continue;
        }
        // A batch of up to `iters` trips.  Lanes whose ray ends stop (run =
        // false) and the batch ends early once no lane runs; the loop exit is
        // wave-uniform, so a trip carries no per-lane loop bookkeeping.  At
        // most SVO_MAX_RAY_ITERS trips per ray: a ray still running then
        // finishes at its current node.
        uint32_t run = active ? 1u : 0u;
        for (int it = 0; it < a.iters; it += SVO_TRIP_UNROLL) {
#pragma unroll
            for (int u = 0; u < SVO_TRIP_UNROLL; ++u) {
                if (run) {
                    const uint32_t d = step_ray<kCache>(r, a, ob, stk, cache, tid);
                    r.guard++;
                    run = (d | (uint32_t)(r.guard >= SVO_MAX_RAY_ITERS)) ^ 1u;
                }
            }
            if (__ballot(run != 0u) == 0ull) break;
        }
        if (active && !run) {
            finish_ray(r, a, ob, stk, cache, tid);
            longest = max(longest, r.guard);
            active = false;
        }
    }
    for (int o = 32; o > 0; o >>= 1) longest = max(longest, (uint32_t)__shfl_xor((int)longest, o));
    if (lane == 0 && traced) atomicAdd(a.ray_count, (unsigned long long)traced);
    if (lane == 0 && merged) atomicAdd(a.merged, (unsigned long long)merged);
    if (lane == 0 && longest) atomicMax(a.max_trips, longest);
}

// The copy half of a frame's rays with a pretrace: every ray of the view
// whose pixel carries the frame's stamp takes the pretraced word and
// position (what finish_ray would write; the other rays were traced by the
// launch before).  One thread per ray; merged rays counted into *merged.
__global__ __launch_bounds__(256) void k_pre_merge(View v, unsigned long long* merged)
{
    const View& V = v;
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    int x = 0, y = 0;
    uint32_t blk;
    bool took = false;
    if (k < view_rays(V) && ray_pixel(V, k, x, y, blk)) {
        const uint32_t pix = (uint32_t)(y * V.res_x + x);
        if (V.pre_stamp[pix] == V.pre_id) {
            const uint32_t word = V.pre_word[pix];
            const float4 q = V.pre_xyz[pix];
            V.screen[pix] = word;
            float* c = (float*)V.coords + 4u * pix;
            c[0] = q.x;
            c[1] = q.y;
            c[2] = q.z;
            if (V.ray_image) V.ray_image[pix] = colorize_word(word);
            took = true;
        }
    }
    const unsigned long long b = __ballot(took);
    if (__lane_id() == 0 && b) atomicAdd(merged, (unsigned long long)__popcll(b));
}

// ---------------------------------------------------------------- warp
__global__ __launch_bounds__(256) void k_memset(uint32_t* __restrict__ dst, uint32_t val, size_t n)
{
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t stride = (size_t)gridDim.x * 256;
    for (; i < n; i += stride) dst[i] = val;
}

__global__ __launch_bounds__(256) void k_fill_u64(unsigned long long* __restrict__ dst, unsigned long long val,
                                                  size_t n)
{
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t stride = (size_t)gridDim.x * 256;
    for (; i < n; i += stride) dst[i] = val;
}

// kernel.cl:498-572 per source pixel; grid y = source plane, z = view.
__global__ __launch_bounds__(256) void k_proj_scatter(View v, const View* views)
{
    VIEW_OF(blockIdx.z);
    const uint32_t N = V.N;
    const uint32_t pix = blockIdx.x * 256u + threadIdx.x;
    if (pix >= N || (int)blockIdx.y >= V.n_planes) return;
    const uint32_t src = (uint32_t)(V.plane_lo + (int)blockIdx.y) * N + pix;
    const uint32_t col = V.screen[src];
    if (col == HOLE) return;
    const float4 c = V.coords[src];
    const float qx = c.x - V.m0x, qy = c.y - V.m0y, qz = c.z - V.m0z;
    const float hx = dot3(qx, qy, qz, V.r00, V.r01, V.r02);
    const float hy = dot3(qx, qy, qz, V.r10, V.r11, V.r12);
    const float hz = dot3(qx, qy, qz, V.r20, V.r21, V.r22);
    if (hz < 0.05f) { V.screen[src] = HOLE; return; }
    const float sfx = (hx * (float)V.res_y + 0.0f) / hz + (float)V.res_x / 2 - 0.0f;
    const float sfy = (hy * (float)V.res_y + 0.0f) / hz + (float)V.res_y / 2 - 0.0f;
    const int sxi = (int)sfx, syi = (int)sfy;
    if (sxi >= V.res_x - 1 || sxi < 0 || syi >= V.res_y - 1 || syi < 0) { V.screen[src] = HOLE; return; }
    const uint32_t t = (uint32_t)syi * (uint32_t)V.res_x + (uint32_t)sxi;
    const uint32_t depth = (uint32_t)(int)(hz * 1000.0f) & 0xffffffu;
    const unsigned long long key = ((unsigned long long)depth << 40) |
                                   ((unsigned long long)src << 8) | (unsigned long long)(col & 255u);
    if (V.keys[t] > key) atomicMin(&V.keys[t], key);
}

// Resolve: winner words/coords into plane 0 (kernel.cl:571-585 semantics
// against the plane-0 word w0), key reset; clear=1 also writes the hole word
// to every pixel without a winner (the per-frame memset of plane 0).  With
// snap_out (warp frames with predicted holes) the winners' positions, NaN
// elsewhere, also go to the frame's snapshot.
__global__ __launch_bounds__(256) void k_proj_resolve(View v, const View* views, int clear)
{
    VIEW_OF(blockIdx.z);
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= V.N) return;
    const unsigned long long key = V.keys[i];
    if (key != ~0ull) {
        V.keys[i] = ~0ull;
        const uint32_t w0 = clear ? HOLE : V.screen[i];
        const uint32_t depth = (uint32_t)(key >> 40);
        if (depth < (w0 >> 8)) {
            const uint32_t src = (uint32_t)(key >> 8);
            const float4 c = V.coords[src];
            const float qx = c.x - V.m0x, qy = c.y - V.m0y, qz = c.z - V.m0z;
            const float hz = dot3(qx, qy, qz, V.r20, V.r21, V.r22);
            V.screen[i] = (depth << 8) + (uint32_t)(key & 255u);
            V.coords[i] = make_float4(c.x, c.y, c.z, hz);
            if (V.snap_out) V.snap_out[i] = make_float4(c.x, c.y, c.z, hz);
            return;
        }
    }
    if (clear) V.screen[i] = HOLE;
    if (V.snap_out) V.snap_out[i] = make_float4(__int_as_float(0x7fc00000), 0.0f, 0.0f, 0.0f);
}

// ---------------------------------------------------------------- compaction
// one wave per 16x16 block, lane q = 2x2 quad (q&7, q>>3): kernel.cl:262-272
__device__ __forceinline__ bool quad_is_hole(const uint32_t* __restrict__ s, int res_x, int x, int y)
{
    const uint32_t o = (uint32_t)y * (uint32_t)res_x + (uint32_t)x;
    int h = (s[o] == HOLE) + (s[o + 1] == HOLE) + (s[o + 1 + res_x] == HOLE) + (s[o + res_x] == HOLE);
    return h >= 4;
}

__global__ __launch_bounds__(64) void k_counthole(View v, const View* views)
{
    VIEW_OF(blockIdx.z);
    const int bxn = V.res_x / 16;
    const uint32_t b = blockIdx.x;
    if (b >= V.S) return;
    const int x = (int)(b % bxn) * 16 + 2 * (threadIdx.x & 7), y = (int)(b / bxn) * 16 + 2 * (threadIdx.x >> 3);
    const unsigned long long m = __ballot(quad_is_hole(V.screen, V.res_x, x, y));
    if (threadIdx.x == 0) V.ids[b] = 4u * (uint32_t)__popcll(m);
}

__global__ __launch_bounds__(64) void k_writeids(View v, const View* views)
{
    VIEW_OF(blockIdx.z);
    const int bxn = V.res_x / 16;
    const uint32_t S = V.S;
    const uint32_t b = blockIdx.x;
    if (b >= S) return;
    const int lane = threadIdx.x;
    const int x = (int)(b % bxn) * 16 + 2 * (lane & 7), y = (int)(b / bxn) * 16 + 2 * (lane >> 3);
    const bool hole = quad_is_hole(V.screen, V.res_x, x, y);
    const unsigned long long m = __ballot(hole);
    if (!hole) return;
    const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    const uint32_t dst = 2u * S + V.ids[S + b] + 4u * rank;
    const uint32_t val = (uint32_t)x + ((uint32_t)y << 16);
    V.ids[dst + 0] = val;
    V.ids[dst + 1] = val + 1u;
    V.ids[dst + 2] = val + 1u + (1u << 16);
    V.ids[dst + 3] = val + (1u << 16);
}

// exclusive scan of ids[0..S) into ids[S..2S), total -> ids[0]; one block per view
// One workgroup per view; each thread scans 8 consecutive block counts per
// round (loads issued together), then the workgroup scans the thread totals.
// With `first` (a single view in the frame graph), thread 0 also writes the
// view's ray range [0, rays) and clears the next ray launch's queue counter.
__global__ __launch_bounds__(1024) void k_sumids(View v, const View* views, uint32_t* first, uint32_t* zero)
{
    VIEW_OF(blockIdx.z);
    uint32_t* ids = V.ids;
    const uint32_t S = V.S;
    __shared__ uint32_t wsum[16];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < S; base += 8192) {
        const uint32_t i0 = base + (uint32_t)tid * 8u;
        uint32_t x[8], pre[8];
        for (int k = 0; k < 8; k++) x[k] = (i0 + (uint32_t)k < S) ? ids[i0 + k] : 0u;
        uint32_t tot = 0;
        for (int k = 0; k < 8; k++) {
            pre[k] = tot;  // exclusive within the thread
            tot += x[k];
        }
        uint32_t incl = tot;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t t = __shfl_up(incl, d);
            if (lane >= d) incl += t;
        }
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        if (w == 0) {
            uint32_t q = (lane < 16) ? wsum[lane] : 0u;
            for (int d = 1; d < 16; d <<= 1) {
                const uint32_t t = __shfl_up(q, d);
                if (lane >= d) q += t;
            }
            if (lane < 16) wsum[lane] = q;  // inclusive wave totals
        }
        __syncthreads();
        const uint32_t tpre = carry + ((w > 0) ? wsum[w - 1] : 0u) + incl - tot;
        for (int k = 0; k < 8; k++)
            if (i0 + (uint32_t)k < S) ids[S + i0 + k] = tpre + pre[k];
        carry += wsum[15];
        __syncthreads();
    }
    if (tid == 0) {
        ids[0] = carry;
        if (first) {
            first[0] = 0u;
            first[1] = view_rays(V);
        }
        if (zero) *zero = 0u;
    }
}

// first ray of each batched view (after their compaction), one thread
__global__ void k_view_first(const View* views, int n, uint32_t* first, uint32_t* zero)
{
    if (zero) *zero = 0u;  // the ray-queue counter of the launch that follows
    uint32_t acc = 0;
    for (int i = 0; i < n; i++) {
        first[i] = acc;
        acc += view_rays(views[i]);
    }
    first[n] = acc;
}

// ---------------------------------------------------------------- copy / filter / colorize
__global__ __launch_bounds__(256) void k_copy_plane(View v, const View* views, uint32_t dst_plane)
{
    VIEW_OF(blockIdx.z);
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= V.N) return;
    const size_t d = (size_t)dst_plane * V.N + i;
    V.screen[d] = V.screen[i];
    V.coords[d] = V.coords[i];
}

// learned frames: the CNN's input, plane 0 as the gate leaves it (the rays
// that follow write plane 0 while the CNN runs)
__global__ __launch_bounds__(256) void k_cnn_input(View v, const View* views)
{
    VIEW_OF(blockIdx.z);
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= V.N) return;
    V.cnn_words[i] = V.screen[i];
    V.cnn_coords[i] = V.coords[i];
}

// kernel.cl:342-401 raycast_fillhole (disabled in the reference, raycast.h:
// 205-220): a plane-0 pixel 3 <= x < res_x-4, 3 <= y < res_y-4 becomes a hole
// when its nearest neighbour among (i, j) in {-1,0}^2 (camera depth =
// coordinate w; the first strictly nearer in i-outer, j-inner order) is
// nearer by more than min(z0, z1) * Z_CONTINUITY * 4 (Z_CONTINUITY = 0.00625)
// and moves differently (motion buffers mx/my, or none = zero motion).  A
// pixel without a nearer neighbour stays (the reference first reads the
// motion at an uninitialised offset there, then returns).  Each pixel reads
// only its own screen word, so the in-place update is order-independent.
__global__ __launch_bounds__(256) void k_fillhole_punch(View v, const int32_t* __restrict__ mx,
                                                        const int32_t* __restrict__ my)
{
    const View& V = v;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= V.N) return;
    const int x = (int)(i % (uint32_t)V.res_x), y = (int)(i / (uint32_t)V.res_x);
    if (x < 3 || y < 3 || x >= V.res_x - 4 || y >= V.res_y - 4) return;
    if (V.screen[i] == HOLE) return;
    const float z0 = V.coords[i].w;
    float z1 = z0;
    int xi = 0, yj = 0;
    for (int a = -1; a < 1; a++)
        for (int b = -1; b < 1; b++) {
            const float z = V.coords[(long)i + (long)b * V.res_x + a].w;
            if (z < z1) {
                xi = a;
                yj = b;
                z1 = z;
            }
        }
    if (z0 <= z1) return;
    const long oij = (long)i + (long)yj * V.res_x + xi;
    const int dx0 = mx ? mx[i] : 0, dy0 = my ? my[i] : 0;
    const int dx = mx ? mx[oij] : 0, dy = my ? my[oij] : 0;
    if (fabsf(z1 - z0) > fminf(z0, z1) * 0.00625f * 4.0f && (dx != dx0 || dy != dy0)) V.screen[i] = HOLE;
}

__device__ __forceinline__ uint32_t rs(const uint32_t* __restrict__ s, long o, long limit)
{
    return o < limit ? s[o] : HOLE;
}

// kernel.cl:423-469 for one hole pixel, neighbours from the snapshot
__device__ uint32_t fill_pixel(const uint32_t* __restrict__ src, long ofs, int res_x, long limit)
{
    uint32_t col = HOLE;
    const uint32_t c1 = rs(src, ofs + 1, limit), c2 = rs(src, ofs - 1, limit);
    const uint32_t c3 = rs(src, ofs + res_x, limit), c4 = rs(src, ofs - res_x, limit);
    if (c1 != HOLE && c2 != HOLE && c3 != HOLE && c4 != HOLE)
        return (c1 & 3u) + ((((c1 & 0xfcu) + (c2 & 0xfcu) + (c3 & 0xfcu) + (c4 & 0xfcu)) >> 2) & 0xfcu);
    if (c1 != HOLE && c2 != HOLE) return (c1 & 3u) + ((((c1 & 0xfcu) + (c2 & 0xfcu)) >> 1) & 0xfcu);
    if (c3 != HOLE && c4 != HOLE) return (c3 & 3u) + ((((c3 & 0xfcu) + (c4 & 0xfcu)) >> 1) & 0xfcu);
    // 2x2: (0,0) is the hole itself, then (1,0), (0,1), (1,1)
    col = rs(src, ofs + 1, limit);
    if (col == HOLE) col = rs(src, ofs + res_x, limit);
    if (col == HOLE) col = rs(src, ofs + 1 + res_x, limit);
    if (col == HOLE) {
        // 5x5, x offset outer, y offset inner; first non-hole wins
        for (int i = -2; i < 3 && col == HOLE; i++)
            for (int j = -2; j < 3 && col == HOLE; j++) col = rs(src, ofs + i + (long)j * res_x, limit);
    }
    return col;
}

// snapshot (plane 2, or `snap` when given) -> plane 0 for holes inside the
// filter window; colorize of the filtered word into V.image when do_color.
__global__ __launch_bounds__(256) void k_fill_color(View v, const View* views, const uint32_t* __restrict__ snap_in,
                                                    int do_fill, int do_color)
{
    VIEW_OF(blockIdx.z);
    const uint32_t N = V.N;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= N) return;
    const uint32_t* snap = snap_in ? snap_in : V.screen + 2 * (size_t)N;
    const int x = (int)(i % (uint32_t)V.res_x), y = (int)(i / (uint32_t)V.res_x);
    uint32_t w = snap[i];
    if (do_fill && w == HOLE && x >= 2 && y >= 2 && x < V.res_x - 1 && y < V.res_y - 1) {
        w = fill_pixel(snap, (long)i, V.res_x, (long)N);
        V.screen[i] = w;
    }
    if (do_color) V.image[i] = colorize_word(w);
}

// ---------------------------------------------------------------- learned-fill gate
// raycast_fine (kernel.cl:696-843, disabled in raycast.h:236): one ray per
// 2x2 quad at (2 gx + add_x, 2 gy + add_y) over the rounded-up (res/4) grid
// (ocl.h:188-198).  The quad's first hole pixel in the order (0,0), (1,0),
// (0,1), (1,1) is traced; without one, pixel ((frame>>2)&1, (frame>>3)&1).
// The reads are flat offsets, so a quad on the last column or row reads and
// may pick the first pixel of the next row or of plane 1 (kernel.cl:727-736).
// The pick is made for all quads before any ray writes (the reference's
// threads read only their own quad except at those edges).
__global__ __launch_bounds__(256) void k_quad_select(View v, int add_x, int add_y, int frame, int gx_n, int gy_n,
                                                     uint32_t* list, uint32_t* count)
{
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g == 0) *count = (uint32_t)(gx_n * gy_n);
    if (g >= gx_n * gy_n) return;
    int idx = 2 * (g % gx_n) + add_x, idy = 2 * (g / gx_n) + add_y;
    if (idx >= v.res_x || idy >= v.res_y) {
        list[g] = QUAD_NONE;
        return;
    }
    bool found = false;
    for (int i = 0; i < 4 && !found; i++) {
        const size_t o = (size_t)(idy + (i >> 1)) * v.res_x + idx + (i & 1);
        if (v.screen[o] == HOLE) {
            idx += i & 1;
            idy += i >> 1;
            found = true;
        }
    }
    if (!found) {
        idx += (frame >> 2) & 1;
        idy += (frame >> 3) & 1;
    }
    list[g] = (uint32_t)idx | ((uint32_t)idy << 16);
}

// Long blocks of a full frame go first.  A ray block (8x8 by default) whose previous full
// frame had a ray of at least long_trips trips (flag in `prev`) has its
// on-screen pixels written to the front list (count in pixels, at most the
// screen) and the rectangle pass skips it; the flags for this frame (`cur`)
// are cleared.  In a flythrough such blocks sit at the horizon, half-way
// through raster order, and their rays alone take a large part of the frame;
// started first, they overlap the rest.  Rays are independent, so the
// results do not depend on the order.  count[0] is this frame's list length
// (zero on entry); count[1], the next frame's, is cleared here.
__global__ __launch_bounds__(256) void k_long_blocks(const uint32_t* __restrict__ prev, uint32_t* __restrict__ cur,
                                                     uint32_t* __restrict__ list, uint32_t* __restrict__ count,
                                                     uint32_t* __restrict__ next_count, uint32_t n, uint32_t tiles_x,
                                                     uint32_t bshift,
                                                     uint32_t res_x, uint32_t res_y)
{
    __shared__ uint32_t s_wave[4], s_base;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint32_t i = blockIdx.x * 256u + (uint32_t)tid;
    if (i == 0) *next_count = 0u;
    const bool hv = i < n && prev && prev[i] != 0u;
    if (i < n) cur[i] = 0u;
    const uint32_t bw = 1u << bshift, bh = 64u >> bshift;
    const uint32_t x0 = (i % tiles_x) * bw, y0 = (i / tiles_x) * bh;
    const uint32_t w = hv ? min(bw, res_x - x0) : 0u, h = hv ? min(bh, res_y - y0) : 0u;
    const uint32_t mine = w * h;
    uint32_t incl = mine;  // inclusive scan of pixel counts over the wave
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(incl, d);
        if (lane >= d) incl += t;
    }
    if (lane == 63) s_wave[wv] = incl;
    __syncthreads();
    if (tid == 0) {
        const uint32_t tot = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
        s_base = tot ? atomicAdd(count, tot) : 0u;  // one atomic per workgroup
    }
    __syncthreads();
    if (!hv) return;
    uint32_t base = s_base + incl - mine;
    for (int k = 0; k < wv; k++) base += s_wave[k];
    for (uint32_t j = 0; j < mine; j++) list[base + j] = (x0 + j % w) | ((y0 + j / w) << 16);
}

// gate counters: [0] CNN tiles, [GATE_RAYS] gate rays (a separate 64-byte line)
constexpr int GATE_RAYS = 16;
__global__ void k_gate_reset(View v, const View* views)
{
    VIEW_OF(blockIdx.y);
    V.gate[0] = 0u;
    V.gate[GATE_RAYS] = 0u;
}

// After compaction and before the rays, each 16x16 tile counts its residual
// holes: hole pixels that neither the hole list (all-hole 2x2 quads of the
// full 16x16 blocks) nor the refresh rectangle will trace.  0 < residual <=
// max_cnn -> the tile goes to the CNN; more -> its residual pixels join the
// frame's ray launch as a second list.  With a depth limit (gate_depth > 0)
// a tile also counts its far pixels: reprojected pixels (outside the refresh
// rectangle) whose camera depth (coordinate w, written by the warp) exceeds
// the limit -- where a pixel spans several voxels, the warped sample is one
// of them and its colour rarely the one the pixel's ray meets.  A tile with
// a far pixel sends its residual holes and its far pixels to the list.  A
// workgroup handles GATE_TILES tiles (each wave two, lane = 4 pixels) and
// appends its CNN tiles and gate pixels with one atomic per list, so the
// counters see few, uncontended atomics.
constexpr int GATE_TILES = 8;
__global__ __launch_bounds__(256) void k_gate(View v, const View* views, int max_cnn)
{
    VIEW_OF(blockIdx.y);
    __shared__ uint32_t s_n[GATE_TILES], s_proj[GATE_TILES], s_base[2];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int res_x = V.res_x, res_y = V.res_y;
    const int tiles_x = V.gate_tiles_x, n_tiles = V.gate_n_tiles;
    uint32_t* gate = V.gate;
    uint32_t* tiles = V.gate_tiles;
    uint32_t* rays = V.gate_rays;
    const uint32_t max_tiles = (uint32_t)n_tiles;
    const int fx = (res_x / 16) * 16, fy = (res_y / 16) * 16;
    const bool depth_gate = V.gate_depth > 0.0f;
    // per pixel: residual hole; far reprojected pixel outside the rectangle
    unsigned long long m[2][4], mp[2][4];
    for (int h = 0; h < 2; h++) {
        const int lt = wave * 2 + h, t = blockIdx.x * GATE_TILES + lt;
        uint32_t n = 0, nf = 0;
        for (int j = 0; j < 4; j++) {
            const int i = lane + 64 * j;
            const int x = (t % tiles_x) * 16 + (i & 15), y = (t / tiles_x) * 16 + (i >> 4);
            const bool in = t < n_tiles && x < res_x && y < res_y;
            const size_t o = (size_t)y * res_x + x;
            const uint32_t w = in ? V.screen[o] : HOLE;
            const bool in_rect = x >= V.add_x && x < V.add_x + V.ext_x && y >= V.add_y && y < V.add_y + V.ext_y;
            bool hole = in && w == HOLE && !in_rect;
            if (hole && x < fx && y < fy && quad_is_hole(V.screen, res_x, x & ~1, y & ~1))
                hole = false;  // in the hole list
            const bool far = depth_gate && in && w != HOLE && !in_rect && V.coords[o].w > V.gate_depth;
            m[h][j] = __ballot(hole);
            mp[h][j] = __ballot(far);
            n += (uint32_t)__popcll(m[h][j]);
            nf += (uint32_t)__popcll(mp[h][j]);
        }
        if (lane == 0) {
            s_n[lt] = n;
            s_proj[lt] = nf;  // a tile with far pixels sends them (and all its residual holes)
        }
    }
    __syncthreads();
    // pixels a tile sends back to raycasting; 0 with residual holes = a CNN tile
    auto tile_rays = [&](int k) -> uint32_t {
        return s_proj[k] ? s_n[k] + s_proj[k] : ((int)s_n[k] > max_cnn ? s_n[k] : 0u);
    };
    if (threadIdx.x == 0) {
        uint32_t nc = 0, nr = 0;
        for (int k = 0; k < GATE_TILES; k++) {
            const uint32_t tr = tile_rays(k);
            nr += tr;
            nc += (!tr && s_n[k]) ? 1u : 0u;
        }
        s_base[0] = nc ? atomicAdd(&gate[0], nc) : 0u;
        s_base[1] = nr ? atomicAdd(&gate[GATE_RAYS], nr) : 0u;
    }
    __syncthreads();
    for (int h = 0; h < 2; h++) {
        const int lt = wave * 2 + h, t = blockIdx.x * GATE_TILES + lt;
        const uint32_t tr = tile_rays(lt);
        if (!tr && !s_n[lt]) continue;
        // this tile's slots: tiles of lower index in the block come first
        uint32_t oc = s_base[0], orr = s_base[1];
        for (int k = 0; k < lt; k++) {
            const uint32_t tk = tile_rays(k);
            orr += tk;
            oc += (!tk && s_n[k]) ? 1u : 0u;
        }
        if (!tr) {  // a CNN tile
            if (lane == 0 && oc < max_tiles) tiles[oc] = (uint32_t)t;
            continue;
        }
        for (int j = 0; j < 4; j++) {
            const unsigned long long mm = m[h][j] | mp[h][j];  // mp: far pixels (none without the depth rule)
            if ((mm >> lane) & 1ull) {
                const int i = lane + 64 * j;
                const int x = (t % tiles_x) * 16 + (i & 15), y = (t / tiles_x) * 16 + (i >> 4);
                rays[orr + (uint32_t)__popcll(mm & ((1ull << lane) - 1ull))] = (uint32_t)x | ((uint32_t)y << 16);
            }
            orr += (uint32_t)__popcll(mm);
        }
    }
}

}  // namespace

// ====================================================================== runtime
#define RC(x)              \
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

static void free_pre(svo_ctx* c);

static int free_frame(svo_ctx* c)
{
    if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);  // no copy reads a freed image
    if (c->d_image && c->d_image != c->image_buf[0] && c->d_image != c->image_buf[1])
        (void)hipFree(c->d_image);  // a resize that failed before recording it
    (void)hipFree(c->d_screen);
    (void)hipFree(c->d_coords);
    (void)hipFree(c->d_ids);
    (void)hipFree(c->image_buf[0]);
    (void)hipFree(c->image_buf[1]);
    c->image_buf[0] = c->image_buf[1] = nullptr;
    c->image_copy_pending[0] = c->image_copy_pending[1] = false;
    c->image_flip = false;
    (void)hipFree(c->d_keys);
    (void)hipFree(c->d_snap);
    c->d_screen = nullptr;
    c->d_coords = nullptr;
    c->d_ids = nullptr;
    c->d_image = nullptr;
    c->d_keys = nullptr;
    c->d_snap = nullptr;
    (void)hipFree(c->d_gate);
    (void)hipFree(c->d_gate_tiles);
    (void)hipFree(c->d_gate_rays);
    c->d_gate = c->d_gate_tiles = c->d_gate_rays = nullptr;
    (void)hipFree(c->d_motion);
    c->d_motion = nullptr;
    (void)hipFree(c->d_cnn_words);
    (void)hipFree(c->d_cnn_coords);
    (void)hipFree(c->d_cnn_out);
    c->d_cnn_words = nullptr;
    c->d_cnn_coords = nullptr;
    c->d_cnn_out = nullptr;
    c->cnn_input = false;
    (void)hipFree(c->d_block_long);
    (void)hipFree(c->d_long_list);
    (void)hipFree(c->d_long_count);
    c->d_block_long = c->d_long_list = c->d_long_count = nullptr;
    c->long_valid = false;
    free_pre(c);
    (void)hipFree(c->d_known);
    c->d_known = nullptr;
    c->res_x = c->res_y = 0;
    c->N = c->S = 0;
    return SVO_OK;
}

// the captured frame graph holds buffer pointers and launch knobs
static void drop_graph(svo_ctx* c)
{
    for (auto& g : c->graph)
        if (g) (void)hipGraphExecDestroy(g);
    for (auto& g : c->graph) g = nullptr;
}

// pretrace pool streams for `sets` sets with every fourth stream left out
static int pool_size(int sets) { return std::min(kPrePool, sets + (sets + 2) / 3); }

extern "C" int svo_create(svo_ctx** out, int device)
{
    if (!out) return svo_fail(SVO_ERR_ARG, "null context pointer");
    *out = nullptr;
    int n = 0;
    HIPCHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return svo_fail(SVO_ERR_ARG, "device %d not present (%d devices)", device, n);
    HIPCHECK(hipSetDevice(device));
    svo_ctx* c = new svo_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        c->n_cu = prop.multiProcessorCount;
    c->cnn_overlap_wgs = std::max(1, c->n_cu / 2);
    bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess &&
              hipStreamCreateWithFlags(&c->cnn_stream, hipStreamNonBlocking) == hipSuccess;
    // the default pretrace's stream pool right behind the context's own two
    // streams, so the queue the default queue_skip leaves out is the slow one
    for (int i = 0; ok && i < pool_size(c->pre_sets); i++)
        ok = hipStreamCreateWithFlags(&c->pre_pool[i], hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&c->cnn_fork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&c->cnn_join, hipEventDisableTiming) == hipSuccess &&
              hipMalloc(&c->d_queue, 128) == hipSuccess && hipMemset(c->d_queue, 0, 128) == hipSuccess;
    for (int i = 0; ok && i <= ST_N; i++) ok = hipEventCreate(&c->ev[i]) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&c->ev_frame_done, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; ok && i < kPreMaxSets; i++)
        ok = hipEventCreateWithFlags(&c->ev_pre[2 * i], hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&c->ev_pre[2 * i + 1], hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&c->ev_list[2 * i], hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&c->ev_list[2 * i + 1], hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&c->ev_main[i], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        svo_destroy(c);
        return svo_fail(SVO_ERR_HIP, "context set-up failed on device %d", device);
    }
    *out = c;
    return SVO_OK;
}

extern "C" int svo_destroy(svo_ctx* c)
{
    if (!c) return SVO_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->cnn_stream) (void)hipStreamSynchronize(c->cnn_stream);
    for (auto& st : c->pre_streams)
        if (st) (void)hipStreamSynchronize(st);
    for (auto& st : c->pre_pool)
        if (st) (void)hipStreamSynchronize(st);
    free_frame(c);
    for (auto& e : c->image_copied)
        if (e) (void)hipEventDestroy(e);
    if (c->image_ready) (void)hipEventDestroy(c->image_ready);
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    drop_graph(c);
    (void)hipFree(c->d_view);
    (void)hipHostFree(c->h_view);
    (void)hipFree(c->d_first);
    for (auto& e : c->view_done)
        if (e) (void)hipEventDestroy(e);
    (void)hipFree(c->d_oct);
    (void)hipFree(c->d_queue);
    for (int i = 0; i <= ST_N; i++)
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    if (c->cnn_fork) (void)hipEventDestroy(c->cnn_fork);
    if (c->cnn_join) (void)hipEventDestroy(c->cnn_join);
    if (c->cnn_stream) (void)hipStreamDestroy(c->cnn_stream);
    for (auto& e : c->ev_pre)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : c->ev_list)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : c->ev_main)
        if (e) (void)hipEventDestroy(e);
    for (auto& st : c->pre_streams)
        if (st) (void)hipStreamDestroy(st);
    for (auto& st : c->pre_pool)
        if (st) (void)hipStreamDestroy(st);
    for (auto& e : c->end_ring)
        if (e) (void)hipEventDestroy(e);
    if (c->ev_frame_done) (void)hipEventDestroy(c->ev_frame_done);
    for (auto& e : c->ev_snap)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SVO_OK;
}

// Pathological fetchColor loops (kernel.cl:104-106, reached only by rays that
// stop at a byte node without a hit) may read up to 2^21 words past a valid
// offset; the device copy carries that many zero words after the array.
static const size_t kOctPad = (1u << 21) + 256;

extern "C" int svo_upload_octree(svo_ctx* c, const uint32_t* words, size_t n_words, uint32_t root_ref, int depth)
{
    if (!c || !words) return svo_fail(SVO_ERR_ARG, "null argument");
    if (depth < 8 || depth > 13) return svo_fail(SVO_ERR_ARG, "depth %d outside 8..13", depth);
    if (n_words < 10 || n_words > (1ull << 29)) return svo_fail(SVO_ERR_ARG, "word count %zu out of range", n_words);
    std::vector<uint32_t> dev;
    uint32_t new_root = 0, normal_words = 0;
    RC(svo_device_layout(words, n_words, root_ref, depth, dev, &new_root, &normal_words));
    HIPCHECK(hipSetDevice(c->device));
    HIPCHECK(hipStreamSynchronize(c->stream));
    drop_graph(c);
    c->long_valid = false;  // the next full frame only collects long-ray flags
    (void)hipFree(c->d_oct);
    c->d_oct = nullptr;
    HIPCHECK(hipMalloc(&c->d_oct, (n_words + kOctPad) * 4));
    HIPCHECK(hipMemcpy(c->d_oct, dev.data(), n_words * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemset(c->d_oct + n_words, 0, kOctPad * 4));
    c->n_words = n_words;
    c->root = new_root;
    c->depth = depth;
    c->normal_words = normal_words;
    return SVO_OK;
}

// the most ray blocks a full screen has over the block shapes 8x8 .. 64x1
static size_t max_ray_blocks(int res_x, int res_y)
{
    size_t m = 0;
    for (int s = 3; s <= 6; s++)
        m = std::max(m, (size_t)((res_x + (1 << s) - 1) >> s) * (size_t)((res_y + (64 >> s) - 1) / (64 >> s)));
    return m;
}

extern "C" int svo_resize(svo_ctx* c, int res_x, int res_y)
{
    if (!c) return svo_fail(SVO_ERR_ARG, "null context");
    if (res_x < 16 || res_y < 16 || res_x > 8192 || res_y > 8192)
        return svo_fail(SVO_ERR_ARG, "resolution %dx%d outside 16..8192", res_x, res_y);
    HIPCHECK(hipSetDevice(c->device));
    HIPCHECK(hipStreamSynchronize(c->stream));
    drop_graph(c);
    free_frame(c);
    const size_t N = (size_t)res_x * res_y, S = (size_t)(res_x / 16) * (res_y / 16);
    const size_t ids_words = 2 * S + N + 64;
    if (hipMalloc(&c->d_screen, 4 * N * 4) != hipSuccess || hipMalloc(&c->d_coords, 4 * N * 16) != hipSuccess ||
        hipMalloc(&c->d_ids, ids_words * 4) != hipSuccess || hipMalloc(&c->d_image, N * 4) != hipSuccess ||
        hipMalloc(&c->d_keys, N * 8) != hipSuccess || hipMalloc(&c->d_snap, N * 4) != hipSuccess ||
        hipMalloc(&c->d_gate, 128) != hipSuccess ||
        hipMalloc(&c->d_gate_tiles, ((size_t)(res_x + 15) / 16) * ((res_y + 15) / 16) * 4) != hipSuccess ||
        hipMalloc(&c->d_gate_rays, N * 4) != hipSuccess ||
        hipMalloc(&c->d_block_long, max_ray_blocks(res_x, res_y) * 8) != hipSuccess ||
        hipMalloc(&c->d_long_list, N * 4) != hipSuccess ||
        hipMalloc(&c->d_long_count, 512) != hipSuccess ||
        hipMalloc(&c->d_known, (N + res_x + 64) * 4) != hipSuccess) {
        free_frame(c);
        return svo_fail(SVO_ERR_NOMEM, "frame buffers for %dx%d", res_x, res_y);
    }
    HIPCHECK(hipMemset(c->d_screen, 0, 4 * N * 4));
    HIPCHECK(hipMemset(c->d_coords, 0, 4 * N * 16));
    HIPCHECK(hipMemset(c->d_ids, 0, ids_words * 4));
    HIPCHECK(hipMemset(c->d_image, 0, N * 4));
    c->image_buf[0] = c->d_image;
    HIPCHECK(hipMemset(c->d_keys, 0xff, N * 8));
    HIPCHECK(hipMemset(c->d_snap, 0, N * 4));
    HIPCHECK(hipMemset(c->d_gate, 0, 128));
    HIPCHECK(hipMemset(c->d_long_count, 0, 512));
    HIPCHECK(hipMemset(c->d_known, 0, (N + res_x + 64) * 4));
    c->n_tiles_x = (res_x + 15) / 16;
    c->n_tiles_y = (res_y + 15) / 16;
    HIPCHECK(hipDeviceSynchronize());
    c->res_x = res_x;
    c->res_y = res_y;
    c->N = N;
    c->S = S;
    c->plane1_holes = false;
    return SVO_OK;
}

static int need_frame(svo_ctx* c)
{
    if (!c) return svo_fail(SVO_ERR_ARG, "null context");
    if (!c->d_screen) return svo_fail(SVO_ERR_STATE, "svo_resize has not been called");
    return SVO_OK;
}
static int need_octree(svo_ctx* c)
{
    int rc = need_frame(c);
    if (rc) return rc;
    if (!c->d_oct) return svo_fail(SVO_ERR_STATE, "no octree uploaded");
    return SVO_OK;
}

static inline void mark(svo_ctx* c, int st)
{
    if (c->timing == 1) {
        (void)hipEventRecord(c->ev[st], c->stream);
        c->stage_used[st] = true;
    }
}

static void frame_begin(svo_ctx* c)
{
    for (int i = 0; i < ST_N; i++) c->stage_used[i] = false;
    c->frame_timed = c->timing == 1 || c->timing == 2;
    if (c->timing == 2) (void)hipEventRecord(c->ev[0], c->stream);
}
static int frame_end(svo_ctx* c)
{
    if (c->timing == 3) {
        if (c->end_count < kEndRing) HIPCHECK(hipEventRecord(c->end_ring[c->end_count++], c->stream));
        return SVO_OK;
    }
    if (c->timing) HIPCHECK(hipEventRecord(c->ev[ST_N], c->stream));
    return SVO_OK;
}

// After svo_read_image_async the next call that rewrites the whole image (a
// frame, a batch frame, svo_colorize) renders into the other image buffer;
// that buffer's own last copy (two frames back) must be complete.  Calls
// that write part of the image or none keep the current buffer.
static void take_image(svo_ctx* c)
{
    if (!c->image_flip || !c->image_buf[1]) return;
    c->image_flip = false;
    const int s = c->d_image == c->image_buf[1] ? 0 : 1;
    if (c->image_copy_pending[s]) {
        (void)hipEventSynchronize(c->image_copied[s]);
        c->image_copy_pending[s] = false;
    }
    c->d_image = c->image_buf[s];
}

static void view_camera(View& v, const svo_frame_params* p);

static View make_view(svo_ctx* c, const svo_frame_params* p)
{
    View v;
    memset(&v, 0, sizeof(v));
    v.screen = c->d_screen;
    v.coords = c->d_coords;
    v.ids = c->d_ids;
    v.image = c->d_image;
    v.keys = c->d_keys;
    v.res_x = c->res_x;
    v.res_y = c->res_y;
    v.N = (uint32_t)c->N;
    v.S = (uint32_t)c->S;
    v.tiles_x = 1;
    v.blk_shift = c->blk_shift;
    v.gate = c->d_gate;
    v.gate_tiles = c->d_gate_tiles;
    v.gate_rays = c->d_gate_rays;
    v.gate_tiles_x = c->n_tiles_x;
    v.gate_n_tiles = c->n_tiles_x * c->n_tiles_y;
    v.gate_depth = c->gate_depth;
    // quad rays may write one row past the frame: d_known holds N + res_x + 64
    v.trips_out = c->d_known;
    if (c->pre_active) {
        const size_t o = (size_t)c->pre_set * c->N;
        v.pre_stamp = c->d_pre_stamp + o;
        v.pre_word = c->d_pre_word + o;
        v.pre_xyz = c->d_pre_xyz + o;
        v.pre_id = c->pre_id;
    }
    view_camera(v, p);
    return v;
}

static void view_camera(View& v, const svo_frame_params* p)
{
    if (p) {
        v.m0x = p->m0[0]; v.m0y = p->m0[1]; v.m0z = p->m0[2];
        v.r00 = p->row0[0]; v.r01 = p->row0[1]; v.r02 = p->row0[2];
        v.r10 = p->row1[0]; v.r11 = p->row1[1]; v.r12 = p->row1[2];
        v.r20 = p->row2[0]; v.r21 = p->row2[1]; v.r22 = p->row2[2];
        v.cx0 = p->col0[0]; v.cx1 = p->col0[1]; v.cx2 = p->col0[2];
        v.cy0 = p->col1[0]; v.cy1 = p->col1[1]; v.cy2 = p->col1[2];
        v.cz0 = p->col2[0]; v.cz1 = p->col2[1]; v.cz2 = p->col2[2];
        v.fovx = p->fovx;
        v.fovy = p->fovy;
    }
}

static void set_rays(View& v, const uint32_t* list, const uint32_t* list_count, int add_x, int add_y, int ext_x,
                     int ext_y, uint32_t* ray_image)
{
    v.list = list;
    v.list_count = list_count;
    v.add_x = add_x;
    v.add_y = add_y;
    v.ext_x = std::max(0, ext_x);
    v.ext_y = std::max(0, ext_y);
    const int bw = 1 << v.blk_shift, bh = 64 >> v.blk_shift;
    v.tiles_x = std::max(1, (v.ext_x + bw - 1) / bw);
    v.n_rect = (uint32_t)(((v.ext_x + bw - 1) / bw) * ((v.ext_y + bh - 1) / bh) * 64);
    v.ray_image = ray_image;
}

static int launch_memset(svo_ctx* c, uint32_t* dst, uint32_t val, size_t n)
{
    if (n == 0) return SVO_OK;
    size_t blocks = std::min<size_t>((n + 255) / 256, (size_t)c->n_cu * 16);
    hipLaunchKernelGGL(k_memset, dim3((unsigned)blocks), dim3(256), 0, c->stream, dst, val, n);
    HIPCHECK(hipGetLastError());
    return SVO_OK;
}

// The warp-stage launchers take the view by value (views == null) or a
// device table of n views (grid z); maxN / maxS size the grid for the largest.
struct Launch {
    hipStream_t stream;
    const View* views;
    int n;
    uint32_t maxN, maxS;
};
static Launch single(svo_ctx* c) { return Launch{c->stream, nullptr, 1, (uint32_t)c->N, (uint32_t)c->S}; }

static int launch_scatter(const Launch& L, const View& v, int max_planes)
{
    if (max_planes <= 0) return SVO_OK;
    hipLaunchKernelGGL(k_proj_scatter, dim3((L.maxN + 255) / 256, (unsigned)max_planes, (unsigned)L.n), dim3(256), 0,
                       L.stream, v, L.views);
    HIPCHECK(hipGetLastError());
    return SVO_OK;
}
static int launch_resolve(const Launch& L, const View& v, int clear)
{
    hipLaunchKernelGGL(k_proj_resolve, dim3((L.maxN + 255) / 256, 1, (unsigned)L.n), dim3(256), 0, L.stream, v,
                       L.views, clear);
    HIPCHECK(hipGetLastError());
    return SVO_OK;
}
static int launch_scatter_resolve(const Launch& L, const View& v, int max_planes, int clear)
{
    RC(launch_scatter(L, v, max_planes));
    return launch_resolve(L, v, clear);
}

static int launch_count(const Launch& L, const View& v)
{
    if (L.maxS == 0) return SVO_OK;
    hipLaunchKernelGGL(k_counthole, dim3(L.maxS, 1, (unsigned)L.n), dim3(64), 0, L.stream, v, L.views);
    HIPCHECK(hipGetLastError());
    return SVO_OK;
}
static int launch_sum(const Launch& L, const View& v, uint32_t* first = nullptr, uint32_t* zero = nullptr)
{
    hipLaunchKernelGGL(k_sumids, dim3(1, 1, (unsigned)L.n), dim3(1024), 0, L.stream, v, L.views, first, zero);
    HIPCHECK(hipGetLastError());
    return SVO_OK;
}
static int launch_write(const Launch& L, const View& v)
{
    if (L.maxS == 0) return SVO_OK;
    hipLaunchKernelGGL(k_writeids, dim3(L.maxS, 1, (unsigned)L.n), dim3(64), 0, L.stream, v, L.views);
    HIPCHECK(hipGetLastError());
    return SVO_OK;
}
static int launch_copy(const Launch& L, const View& v, int dst_plane)
{
    hipLaunchKernelGGL(k_copy_plane, dim3((L.maxN + 255) / 256, 1, (unsigned)L.n), dim3(256), 0, L.stream, v,
                       L.views, (uint32_t)dst_plane);
    HIPCHECK(hipGetLastError());
    return SVO_OK;
}
static int launch_fill(const Launch& L, const View& v, const uint32_t* snap, int do_fill, int do_color)
{
    hipLaunchKernelGGL(k_fill_color, dim3((L.maxN + 255) / 256, 1, (unsigned)L.n), dim3(256), 0, L.stream, v,
                       L.views, snap, do_fill, do_color);
    HIPCHECK(hipGetLastError());
    return SVO_OK;
}

// persistent ray launch over one view (views == null) or a batch table
// counters: [0..1] rays traced (u64), [2] longest ray (trips), [4..5] rays
// merged from a pretrace (u64); default the context's frame counters
// spread_shift: the hand-out spread (default the context's; 6 = list order);
// wg_per_cu: persistent workgroups per CU (default the context's); split:
// the split hand-out (list order) with its counters split_queue[0..1] --
// half of the grid's waves take the first *split_count rays, 2^split_shift
// lanes at a time
struct Split {
    const uint32_t* count = nullptr;
    uint32_t* queue = nullptr;
    int shift = 6;
};
static int launch_rays_v(svo_ctx* c, hipStream_t stream, const View& v, const View* views,
                         const uint32_t* view_first, int n_views, uint32_t* queue, int& seq,
                         uint32_t* counters = nullptr, int spread_shift = -1, int wg_per_cu = 0,
                         const Split& split = Split())
{
    RayArgs a;
    memset(&a, 0, sizeof(a));
    a.v = v;
    a.views = views;
    a.view_first = view_first;
    a.n_views = n_views;
    a.oct = c->d_oct;
    a.root = c->root;
    a.depth = c->depth;
    a.queue = queue;
    a.slot = seq & 1;
    seq++;
    a.refill = c->refill;
    a.iters = c->iters;
    a.spread_shift = spread_shift >= 0 ? spread_shift : c->spread_shift;
    a.spread_fit = c->spread_fit;
    a.lanes_shift = c->lanes_shift;
    if (!counters) counters = c->d_queue + 8;
    a.ray_count = (unsigned long long*)counters;
    a.max_trips = counters + 2;
    a.merged = (unsigned long long*)(counters + 4);
    a.cache_words = std::min(c->cache_req, c->normal_words);
    a.cache_last = a.cache_words ? a.cache_words - 1 : 0;
    a.oct_bytes = (uint32_t)std::min<size_t>((c->n_words + kOctPad) * 4, 0xfffffffcull);
    const unsigned grid = (unsigned)(c->n_cu * (wg_per_cu > 0 ? wg_per_cu : c->wg_per_cu));
    if (split.count) {
        a.split_count = split.count;
        a.split_queue = split.queue;
        a.split_shift = split.shift;
        a.split_waves = grid * (unsigned)(RAY_BLOCK / 64) / 2u;
    }
    const size_t lds = ((size_t)(c->depth + STK_ROWS_EXTRA) * RAY_BLOCK + std::max(a.cache_words, 1u)) * 4;
    if (a.cache_words)
        hipLaunchKernelGGL(k_rays<true>, dim3(grid), dim3(RAY_BLOCK), lds, stream, a);
    else
        hipLaunchKernelGGL(k_rays<false>, dim3(grid), dim3(RAY_BLOCK), lds, stream, a);
    HIPCHECK(hipGetLastError());
    return SVO_OK;
}

static int launch_rays(svo_ctx* c, const svo_frame_params* p, const uint32_t* list, const uint32_t* list_count,
                       int add_x, int add_y, int ext_x, int ext_y, uint32_t* ray_image)
{
    View v = make_view(c, p);
    set_rays(v, list, list_count, add_x, add_y, ext_x, ext_y, ray_image);
    return launch_rays_v(c, c->stream, v, nullptr, nullptr, 1, c->d_queue, c->launch_seq);
}

static float4* snap_out(const svo_ctx* c);
static void snap_end(svo_ctx* c, const svo_frame_params* p);

// the warp stages of raycast_draw up to the hole compaction
static int warp_compact(svo_ctx* c, const svo_frame_params* p)
{
    const Launch L = single(c);
    mark(c, ST_CLEAR);
    if (p->frame < 2) {  // raycast.h:150-154
        RC(launch_memset(c, c->d_screen, HOLE, 4 * c->N));
        c->plane1_holes = true;
    }
    // planes 1 and 2 (raycast.h:176-198); plane 1 is skipped while the
    // library knows it holds only hole words (then it contributes nothing)
    View v = make_view(c, p);
    v.plane_lo = c->plane1_holes ? 2 : 1;
    v.n_planes = 3 - v.plane_lo;
    v.snap_out = snap_out(c);
    mark(c, ST_SCATTER);
    RC(launch_scatter(L, v, v.n_planes));
    mark(c, ST_RESOLVE);
    RC(launch_resolve(L, v, 1));
    snap_end(c, p);  // later frames' pretraces may read the snapshot from here on
    v.snap_out = nullptr;
    mark(c, ST_COUNT);
    RC(launch_count(L, v));
    mark(c, ST_SCAN);
    RC(launch_sum(L, v));
    mark(c, ST_WRITE);
    return launch_write(L, v);
}

// The frame's hole (and gate) and refresh-rectangle rays.  With a
// pretrace, two launches: once the pretrace's pixel lists are final, the
// rays on none of them are traced (overlapping the pretrace's own rays);
// once the pretrace has completed, its results are copied for the rest.
static int frame_rays(svo_ctx* c, View v)
{
    if (!c->pre_active) return launch_rays_v(c, c->stream, v, nullptr, nullptr, 1, c->d_queue, c->launch_seq);
    const int s = c->pre_set;
    HIPCHECK(hipStreamWaitEvent(c->stream, c->ev_list[2 * s + 1], 0));
    if (c->pre_early) HIPCHECK(hipStreamWaitEvent(c->stream, c->ev_list[2 * s], 0));
    v.pre_bkt1 = c->d_pre_bkt + (size_t)(2 * s + 1) * c->N;
    v.pre_bkt0 = c->pre_early ? c->d_pre_bkt + (size_t)(2 * s) * c->N : nullptr;
    v.pre_mode = 1;
    RC(launch_rays_v(c, c->stream, v, nullptr, nullptr, 1, c->d_queue, c->launch_seq));
    HIPCHECK(hipStreamWaitEvent(c->stream, c->ev_pre[2 * s + 1], 0));
    if (c->pre_early) HIPCHECK(hipStreamWaitEvent(c->stream, c->ev_pre[2 * s], 0));
    v.pre_mode = 2;
    // one thread per possible ray: both lists hold at most N pixels each
    const size_t n = (v.list ? c->N : 0) + (v.list2 ? c->N : 0) + v.n_rect;
    hipLaunchKernelGGL(k_pre_merge, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, v,
                       (unsigned long long*)(c->d_queue + 12));
    HIPCHECK(hipGetLastError());
    return SVO_OK;
}

// the warp stages of raycast_draw up to and including the hole and tile rays
static int warp_front(svo_ctx* c, const svo_frame_params* p)
{
    RC(warp_compact(c, p));
    int ax, ay, ex, ey;
    svo_tile_schedule(c->res_x, c->res_y, p->frame, &ax, &ay, &ex, &ey);
    mark(c, ST_RAYS);
    View v = make_view(c, p);
    set_rays(v, c->d_ids + 2 * c->S, c->d_ids, ax, ay, ex, ey, nullptr);
    return frame_rays(c, v);
}

// Capture the warp frame's launches (frames >= 2: no clear) into one graph.
// Every kernel reads the frame's view from c->d_view; the ray launch uses its
// own queue counters (d_queue[2..3]), reset by k_sumids inside the graph.
static int build_graph(svo_ctx* c)
{
    if (!c->d_view) {
        bool ok = hipMalloc(&c->d_view, sizeof(View)) == hipSuccess &&
                  hipHostMalloc(&c->h_view, kViewSlots * sizeof(View)) == hipSuccess &&
                  hipMalloc(&c->d_first, 8) == hipSuccess;
        for (int i = 0; ok && i < kViewSlots; i++)
            ok = hipEventCreateWithFlags(&c->view_done[i], hipEventDisableTiming) == hipSuccess &&
                 hipEventRecord(c->view_done[i], c->stream) == hipSuccess;
        if (!ok) return svo_fail(SVO_ERR_HIP, "frame graph buffers");
    }
    const View* dv = (const View*)c->d_view;
    const View ph = make_view(c, nullptr);  // by-value placeholder, unused when a table is given
    const Launch L{c->stream, dv, 1, (uint32_t)c->N, (uint32_t)c->S};
    HIPCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int seq = 0;
    int rc = launch_scatter_resolve(L, ph, 2, 1);
    if (!rc) rc = launch_count(L, ph);
    if (!rc) rc = launch_sum(L, ph, c->d_first, c->d_queue + 2);  // also the ray range and queue counter
    if (!rc) rc = launch_write(L, ph);
    if (!rc) rc = launch_rays_v(c, c->stream, ph, dv, c->d_first, 1, c->d_queue + 2, seq);
    if (!rc) rc = launch_copy(L, ph, 2);
    if (!rc) rc = launch_fill(L, ph, nullptr, 1, 1);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (!rc && e != hipSuccess) rc = svo_fail(SVO_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
    for (int i = 0; !rc && i < kGraphExecs; i++)
        if (hipGraphInstantiate(&c->graph[i], g, nullptr, nullptr, 0) != hipSuccess) {
            c->graph[i] = nullptr;
            rc = svo_fail(SVO_ERR_HIP, "graph instantiation failed");
        }
    if (rc) drop_graph(c);
    if (g) (void)hipGraphDestroy(g);
    return rc;
}

static int frame_warp_graph(svo_ctx* c, const svo_frame_params* p, float4* snap)
{
    if (!c->graph[0]) RC(build_graph(c));
    View v = make_view(c, p);
    v.snap_out = snap;
    v.plane_lo = c->plane1_holes ? 2 : 1;
    v.n_planes = 3 - v.plane_lo;
    int ax, ay, ex, ey;
    svo_tile_schedule(c->res_x, c->res_y, p->frame, &ax, &ay, &ex, &ey);
    set_rays(v, c->d_ids + 2 * c->S, c->d_ids, ax, ay, ex, ey, nullptr);
    const int slot = c->graph_frames % kViewSlots;
    View* hv = (View*)c->h_view + slot;
    HIPCHECK(hipEventSynchronize(c->view_done[slot]));  // its previous copy has been consumed
    *hv = v;
    HIPCHECK(hipMemcpyAsync(c->d_view, hv, sizeof(View), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipEventRecord(c->view_done[slot], c->stream));
    HIPCHECK(hipGraphLaunch(c->graph[c->graph_frames % kGraphExecs], c->stream));
    c->graph_frames++;
    return SVO_OK;
}

// ---------------------------------------------------------------- pretrace
// Per set: pixel list of N + kPreHdr words ([0] length, [1] split count,
// [8..31] rays per bucket, [32..55] bucket cursors, pixel words x | y << 16
// from kPreHdr), one word per 8x8 block, one byte per pixel (its sort
// bucket), four ray-queue counters.
constexpr int kPreHdr = 64;
constexpr int kPreBuckets = 24;  // bucket b: blocks whose longest known ray has floor(log2(trips)) = b

static size_t pre_blocks(const svo_ctx* c) { return (size_t)((c->res_x + 7) / 8) * (size_t)((c->res_y + 7) / 8); }

static void free_pre(svo_ctx* c)
{
    (void)hipFree(c->d_pre_stamp);
    (void)hipFree(c->d_pre_word);
    (void)hipFree(c->d_pre_xyz);
    (void)hipFree(c->d_pre_list);
    (void)hipFree(c->d_pre_blk);
    (void)hipFree(c->d_pre_queue);
    (void)hipFree(c->d_pre_bkt);
    c->d_pre_stamp = c->d_pre_word = c->d_pre_list = c->d_pre_blk = c->d_pre_queue = nullptr;
    c->d_pre_bkt = nullptr;
    c->d_pre_xyz = nullptr;
    c->pre_alloc_sets = 0;
    c->pre_frames = 0;
    for (auto& m : c->main_recorded) m = false;
    (void)hipFree(c->d_snap_xyz);
    (void)hipFree(c->d_pred_keys);
    (void)hipFree(c->d_pred_xyz);
    (void)hipFree(c->d_pred_cov);
    (void)hipFree(c->d_pred_deep);
    c->d_snap_xyz = nullptr;
    c->d_pred_keys = nullptr;
    c->d_pred_xyz = nullptr;
    c->d_pred_cov = nullptr;
    c->d_pred_deep = nullptr;
    c->pred_alloc_sets = 0;
    for (int i = 0; i < kSnapSlots; i++) c->snap_frame[i] = c->cam_frame[i] = -1;
    c->snap_pending = -1;
}

// the predicted-hole buffers: snapshots and per-set projection scratch
static int ensure_pred(svo_ctx* c)
{
    if (c->pred_alloc_sets == c->pre_alloc_sets && c->d_snap_xyz) return SVO_OK;
    const size_t K = (size_t)c->pre_alloc_sets, N = c->N;
    for (int i = 0; i < kSnapSlots; i++)
        if (!c->ev_snap[i]) HIPCHECK(hipEventCreateWithFlags(&c->ev_snap[i], hipEventDisableTiming));
    if (hipMalloc(&c->d_snap_xyz, kSnapSlots * N * 16) != hipSuccess ||
        hipMalloc(&c->d_pred_keys, K * N * 8) != hipSuccess || hipMalloc(&c->d_pred_xyz, K * N * 16) != hipSuccess ||
        hipMalloc(&c->d_pred_cov, K * N) != hipSuccess || hipMalloc(&c->d_pred_deep, K * N) != hipSuccess) {
        free_pre(c);
        return svo_fail(SVO_ERR_NOMEM, "predicted-hole buffers for %zu sets of %zu pixels", K, N);
    }
    HIPCHECK(hipMemset(c->d_pred_keys, 0xff, K * N * 8));
    HIPCHECK(hipMemset(c->d_pred_cov, 0, K * N));
    HIPCHECK(hipMemset(c->d_pred_deep, 0, K * N));
    HIPCHECK(hipDeviceSynchronize());
    c->pred_alloc_sets = (int)K;
    for (int i = 0; i < kSnapSlots; i++) c->snap_frame[i] = -1;
    return SVO_OK;
}

static int ensure_pre(svo_ctx* c)
{
    if (c->pre_alloc_sets == c->pre_sets) return SVO_OK;
    HIPCHECK(hipStreamSynchronize(c->stream));  // no frame reads the old sets
    free_pre(c);
    const size_t K = (size_t)c->pre_sets, n = K * c->N;
    // the part-1 stream pool, in index order (enough for the sets with every
    // fourth stream left out); part 0's streams are made when first used
    for (int i = 0; i < pool_size(c->pre_sets); i++)
        if (!c->pre_pool[i]) HIPCHECK(hipStreamCreateWithFlags(&c->pre_pool[i], hipStreamNonBlocking));
    if (hipMalloc(&c->d_pre_stamp, n * 4) != hipSuccess || hipMalloc(&c->d_pre_word, n * 4) != hipSuccess ||
        hipMalloc(&c->d_pre_xyz, n * 16) != hipSuccess ||
        hipMalloc(&c->d_pre_list, 2 * K * (c->N + kPreHdr) * 4) != hipSuccess ||
        hipMalloc(&c->d_pre_blk, 2 * K * pre_blocks(c) * 4) != hipSuccess ||
        hipMalloc(&c->d_pre_queue, 2 * K * 16) != hipSuccess || hipMalloc(&c->d_pre_bkt, 2 * n) != hipSuccess) {
        free_pre(c);
        return svo_fail(SVO_ERR_NOMEM, "%d pretrace sets of %zu pixels", c->pre_sets, c->N);
    }
    HIPCHECK(hipMemset(c->d_pre_stamp, 0, n * 4));  // stamps start at 1
    HIPCHECK(hipMemset(c->d_pre_queue, 0, 2 * K * 16));
    HIPCHECK(hipDeviceSynchronize());
    c->pre_alloc_sets = c->pre_sets;
    for (auto& q : c->pre_seq) q = 0;
    return SVO_OK;
}

// the snapshot slot the frame being issued writes (its projection's
// resolve), or null
static float4* snap_out(const svo_ctx* c)
{
    return c->snap_pending >= 0 ? c->d_snap_xyz + (size_t)c->snap_pending * c->N : nullptr;
}

// A warp or learned frame being issued: record its camera; with predicted
// holes on, point its copy at snapshot slot frame % kSnapSlots (the slot's
// last reader, a pretrace of a frame before this one, has completed: the
// frames in between waited for their pretraces).  Returns the snapshot
// buffer or null.
static float4* snap_begin(svo_ctx* c, const svo_frame_params* p)
{
    c->snap_pending = -1;
    if (c->pre_predict <= 0 || c->pre_sets <= 0 || p->frame < 0) return nullptr;
    if (ensure_pre(c) || ensure_pred(c)) {
        c->pre_predict = 0;  // out of memory: long blocks instead (svo_last_error says why)
        return nullptr;
    }
    const int slot = p->frame % kSnapSlots;
    c->cam_hist[slot] = *p;
    c->cam_frame[slot] = p->frame;
    c->snap_frame[slot] = -1;  // being rewritten
    c->snap_pending = slot;
    return snap_out(c);
}

// after the launch that writes the frame's snapshot is queued: a pretrace
// may read it once this event has completed (no-op when already recorded)
static void snap_end(svo_ctx* c, const svo_frame_params* p)
{
    const int slot = c->snap_pending;
    c->snap_pending = -1;
    if (slot < 0) return;
    if (hipEventRecord(c->ev_snap[slot], c->stream) == hipSuccess) c->snap_frame[slot] = p->frame;
}

__device__ __forceinline__ uint32_t trip_bucket(uint32_t t) { return t ? 31u - (uint32_t)__clz(t) : 0u; }

// Pretrace list, pass 1.  A workgroup of 256 covers 8 x 4 screen blocks of
// 8x8 pixels (one column of 8 pixels per thread, coalesced rows).  A block
// is pretraced whole when the longest known ray among its pixels (the
// `known` map: trips of the last ray traced at each pixel, read while other
// frames' rays update it -- any value read only changes which rays are
// traced ahead) has at least `trips` trips; otherwise only its pixels inside
// the refresh rectangle are.  blk[b] = (whole << 31) | pixels (0: none);
// bkt[pixel] = k = floor(log2(trips)) of the pixel's own known ray (the
// value pass 2 sorts by; the known map may change in between), and hdr[8 +
// k] counts the pretraced pixels of bucket k.
__global__ __launch_bounds__(256) void k_pre_scan(const uint32_t* __restrict__ known, int res_x, int res_y, int ax,
                                                  int ay, int ex, int ey, uint32_t trips, uint32_t* __restrict__ blk,
                                                  uint8_t* __restrict__ bkt, uint32_t* __restrict__ hdr)
{
    __shared__ uint32_t hist[kPreBuckets];
    const int t = threadIdx.x;
    if (t < kPreBuckets) hist[t] = 0;
    __syncthreads();
    const int x = blockIdx.x * 64 + (t & 63);
    const int by = blockIdx.y * 4 + (t >> 6);  // block row
    const int y0 = by * 8;
    const bool col_ok = x < res_x && y0 < res_y;
    const int h = col_ok ? min(8, res_y - y0) : 0;
    uint32_t kn[8];
    uint32_t m = 0, in_tile = 0;
    for (int j = 0; j < 8; j++) {
        kn[j] = j < h ? known[(size_t)(y0 + j) * res_x + x] : 0u;
        m = max(m, kn[j]);
        in_tile += (j < h && x >= ax && x < ax + ex && y0 + j >= ay && y0 + j < ay + ey) ? 1u : 0u;
    }
    uint32_t cnt = (uint32_t)h;
    for (int o = 1; o < 8; o <<= 1) {  // over the block's 8 columns
        m = max(m, (uint32_t)__shfl_xor((int)m, o));
        in_tile += (uint32_t)__shfl_xor((int)in_tile, o);
        cnt += (uint32_t)__shfl_xor((int)cnt, o);
    }
    const bool whole = m >= trips && cnt > 0;
    for (int j = 0; j < h; j++)
        if (whole || (x >= ax && x < ax + ex && y0 + j >= ay && y0 + j < ay + ey)) {
            const uint32_t b = trip_bucket(kn[j]);
            bkt[(size_t)(y0 + j) * res_x + x] = (uint8_t)b;
            atomicAdd(&hist[b], 1u);
        } else {
            bkt[(size_t)(y0 + j) * res_x + x] = kPreNone;
        }
    const int nbx = (res_x + 7) >> 3;
    const int bx = x >> 3;
    if ((t & 7) == 0 && bx < nbx && y0 < res_y) {
        const uint32_t n = whole ? cnt : in_tile;
        blk[(size_t)by * nbx + bx] = n ? (((uint32_t)whole << 31) | n) : 0u;
    }
    __syncthreads();
    if (t < kPreBuckets && hist[t]) atomicAdd(&hdr[8 + t], hist[t]);
}

// Pretrace list, pass 2: the pixels in descending order of their own known
// trips' bucket (the longest rays first, so the ray launch hands them out
// first); within a bucket in no defined order.  hdr[0] = list length,
// hdr[1] = the pixels of buckets >= split_bucket (0 when split_bucket is
// outside 0..23).
__global__ __launch_bounds__(256) void k_pre_emit(const uint8_t* __restrict__ bkt, int res_x, int res_y, int ax,
                                                  int ay, int ex, int ey, int split_bucket, int by_bkt,
                                                  const uint32_t* __restrict__ blk, uint32_t* __restrict__ hdr,
                                                  uint32_t* __restrict__ list)
{
    __shared__ uint32_t first[kPreBuckets];  // this workgroup's first list position per bucket
    __shared__ uint32_t hist[kPreBuckets];
    const int t = threadIdx.x;
    if (t < kPreBuckets) hist[t] = 0;
    __syncthreads();
    const int x = blockIdx.x * 64 + (t & 63);
    const int by = blockIdx.y * 4 + (t >> 6);
    const int y0 = by * 8;
    const int nbx = (res_x + 7) >> 3;
    const int bx = x >> 3;
    const bool col_ok = x < res_x && y0 < res_y && bx < nbx;
    const uint32_t w = col_ok ? blk[(size_t)by * nbx + bx] : 0u;
    const bool whole = (w >> 31) != 0u;
    const int h = (col_ok && (w & 0xffffu)) ? min(8, res_y - y0) : 0;
    uint32_t bk[8];
    uint32_t take = 0;  // bit j: pixel (x, y0 + j) is pretraced
    for (int j = 0; j < h; j++) {
        const int y = y0 + j;
        if (by_bkt) {  // k_pre_scan_pred: the pixel's bucket, or kPreNone
            bk[j] = bkt[(size_t)y * res_x + x];
            if (bk[j] != kPreNone) {
                take |= 1u << j;
                atomicAdd(&hist[bk[j]], 1u);
            }
        } else if (whole || (x >= ax && x < ax + ex && y >= ay && y < ay + ey)) {
            take |= 1u << j;
            bk[j] = bkt[(size_t)y * res_x + x];
            atomicAdd(&hist[bk[j]], 1u);
        }
    }
    __syncthreads();
    if (t < kPreBuckets) {
        uint32_t before = 0;  // pixels of the longer buckets
        for (int b = kPreBuckets - 1; b > t; b--) before += hdr[8 + b];
        first[t] = hist[t] ? before + atomicAdd(&hdr[32 + t], hist[t]) : 0u;
        hist[t] = 0;
        if (blockIdx.x == 0 && blockIdx.y == 0) {
            if (t == 0) hdr[0] = before + hdr[8];  // list length
            if (t == split_bucket) hdr[1] = before + hdr[8 + t];
        }
    }
    __syncthreads();
    for (int j = 0; j < h; j++)
        if ((take >> j) & 1u) list[first[bk[j]] + atomicAdd(&hist[bk[j]], 1u)] = (uint32_t)x | ((uint32_t)(y0 + j) << 16);
}

// ---- predicted holes (pre_predict > 0).  Camera of one projection: the
// reference's raycast_proj arithmetic (kernel.cl:543-562), op order kept.
struct PredCam {
    float m0x, m0y, m0z;
    float r00, r01, r02, r10, r11, r12, r20, r21, r22;
};

static PredCam pred_cam(const svo_frame_params* p)
{
    PredCam k;
    k.m0x = p->m0[0]; k.m0y = p->m0[1]; k.m0z = p->m0[2];
    k.r00 = p->row0[0]; k.r01 = p->row0[1]; k.r02 = p->row0[2];
    k.r10 = p->row1[0]; k.r11 = p->row1[1]; k.r12 = p->row1[2];
    k.r20 = p->row2[0]; k.r21 = p->row2[1]; k.r22 = p->row2[2];
    return k;
}

// target pixel of position q under camera k (-1: behind, too near, a NaN
// hole marker, or off-screen incl. the last column and row), and its camera
// depth hz (world units, the coordinate w the warp writes)
__device__ __forceinline__ int pred_target(float4 q, const PredCam& k, int res_x, int res_y, float* hz_out)
{
    const float qx = q.x - k.m0x, qy = q.y - k.m0y, qz = q.z - k.m0z;
    const float hz = dot3(qx, qy, qz, k.r20, k.r21, k.r22);
    if (!(hz >= 0.05f)) return -1;
    const float hx = dot3(qx, qy, qz, k.r00, k.r01, k.r02);
    const float hy = dot3(qx, qy, qz, k.r10, k.r11, k.r12);
    const float sfx = (hx * (float)res_y + 0.0f) / hz + (float)res_x / 2 - 0.0f;
    const float sfy = (hy * (float)res_y + 0.0f) / hz + (float)res_y / 2 - 0.0f;
    if (!(sfx >= 0.0f && sfy >= 0.0f && sfx < (float)(res_x - 1) && sfy < (float)(res_y - 1))) return -1;
    *hz_out = hz;
    return (int)sfy * res_x + (int)sfx;
}

// nearest sample per target pixel of an intermediate camera: key = depth << 32 | source
// (hdr != null: block 0 also clears the pretrace list header, which the
// list passes after this kernel fill -- one launch fewer in the chain)
__global__ __launch_bounds__(256) void k_pred_scatter(const float4* __restrict__ src, PredCam k, int res_x, int res_y,
                                                      uint32_t N, unsigned long long* __restrict__ keys,
                                                      uint32_t* __restrict__ hdr)
{
    if (hdr && blockIdx.x == 0 && threadIdx.x < (unsigned)kPreHdr) hdr[threadIdx.x] = 0u;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= N) return;
    float hz;
    const int t = pred_target(src[i], k, res_x, res_y, &hz);
    if (t < 0) return;
    const uint32_t z = (uint32_t)(int)(hz * 1000.0f) & 0xffffffu;
    const unsigned long long key = ((unsigned long long)z << 32) | i;
    if (keys[t] > key) atomicMin(&keys[t], key);
}

// a sample reaches pixel t of the final camera: coverage byte, and with a
// depth limit (learned frames' gate) a second byte where it lies beyond it
__device__ __forceinline__ void pred_mark(float4 q, const PredCam& fin, int res_x, int res_y, float deep_z,
                                          uint8_t* cov, uint8_t* deep)
{
    float hz;
    const int t = pred_target(q, fin, res_x, res_y, &hz);
    if (t < 0) return;
    cov[t] = 1;
    if (deep_z > 0.0f && hz > deep_z) deep[t] = 1;
}

// the winners of k_pred_scatter (keys back to all ones): their positions
// into dst (NaN where none), or, with dst == null, marked under the final
// camera `fin`
__global__ __launch_bounds__(256) void k_pred_resolve(unsigned long long* __restrict__ keys,
                                                      const float4* __restrict__ src, float4* __restrict__ dst,
                                                      PredCam fin, int res_x, int res_y, uint32_t N, float deep_z,
                                                      uint8_t* __restrict__ cov, uint8_t* __restrict__ deep)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= N) return;
    const unsigned long long key = keys[i];
    if (key == ~0ull) {
        if (dst) dst[i] = make_float4(__int_as_float(0x7fc00000), 0.0f, 0.0f, 0.0f);
        return;
    }
    keys[i] = ~0ull;
    const float4 q = src[(uint32_t)key];
    if (dst) {
        dst[i] = q;
        return;
    }
    pred_mark(q, fin, res_x, res_y, deep_z, cov, deep);
}

// every sample of src marked under the final camera
__global__ __launch_bounds__(256) void k_pred_cover(const float4* __restrict__ src, PredCam fin, int res_x, int res_y,
                                                    uint32_t N, float deep_z, uint8_t* __restrict__ cov,
                                                    uint8_t* __restrict__ deep, uint32_t* __restrict__ hdr)
{
    if (hdr && blockIdx.x == 0 && threadIdx.x < (unsigned)kPreHdr) hdr[threadIdx.x] = 0u;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= N) return;
    pred_mark(src[i], fin, res_x, res_y, deep_z, cov, deep);
}

// Pretrace list, pass 1 of the predicted-hole mode (layout of k_pre_scan: a
// thread per column of 8 pixels of an 8x8 block; a workgroup covers 64x32
// pixels, i.e. 4x2 whole 16x16 gate tiles).  A pixel is traced ahead when it
// lies in the refresh rectangle, or in a 2x2 quad (even x and y, inside the
// whole 16x16 blocks the hole list covers, kernel.cl:262-272) that no
// predicted sample reaches, in a block whose longest known ray has at least
// `trips` trips -- unless the block's longest known ray has at least `whole`
// (> 0) trips: part 0 traces those blocks whole.  Learned frames (max_cnn >=
// 0) also predict the gate (k_gate): a tile's residual holes are the pixels
// outside the rectangle no sample reaches and no predicted hole quad holds;
// a tile with more than max_cnn of them has them traced ahead, and with the
// depth rule (deep != null) a tile where samples beyond the depth limit land
// outside the rectangle has them and its residual holes traced ahead.  bkt[pixel] = its bucket, or
// kPreNone; the coverage (and depth) bytes are cleared for the next use.
__global__ __launch_bounds__(256) void k_pre_scan_pred(const uint32_t* __restrict__ known, uint8_t* __restrict__ cov,
                                                       uint8_t* __restrict__ deep, int max_cnn, int res_x, int res_y,
                                                       int qx_end, int qy_end, int ax, int ay, int ex, int ey,
                                                       uint32_t trips, uint32_t whole, uint32_t* __restrict__ blk,
                                                       uint8_t* __restrict__ bkt, uint32_t* __restrict__ hdr)
{
    __shared__ uint32_t hist[kPreBuckets];
    __shared__ uint32_t s_res[8], s_far[8];  // per gate tile of the workgroup
    const int t = threadIdx.x;
    if (t < kPreBuckets) hist[t] = 0;
    if (t < 8) s_res[t] = s_far[t] = 0;
    __syncthreads();
    const int x = blockIdx.x * 64 + (t & 63);
    const int by = blockIdx.y * 4 + (t >> 6);
    const int y0 = by * 8;
    const int gt = ((t & 63) >> 4) + 4 * ((t >> 6) >> 1);  // this column's gate tile in the workgroup
    const bool col_ok = x < res_x && y0 < res_y;
    const int h = col_ok ? min(8, res_y - y0) : 0;
    uint32_t kn[8];
    uint32_t m = 0, cv = 0, dp = 0;  // bit j: pixel (x, y0 + j) is reached / reached beyond the depth limit
    for (int j = 0; j < 8; j++) {
        const size_t o = (size_t)(y0 + j) * res_x + x;
        kn[j] = j < h ? known[o] : 0u;
        cv |= (j < h && cov[o]) ? 1u << j : 0u;
        dp |= (deep && j < h && deep[o]) ? 1u << j : 0u;
        m = max(m, kn[j]);
    }
    const uint32_t ncv = (uint32_t)__shfl_xor((int)cv, 1);  // the quad's other column (x ^ 1)
    for (int o = 1; o < 8; o <<= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o));
    const bool long_blk = m >= trips;
    const bool part0 = whole > 0u && m >= whole;  // traced whole by part 0
    const bool qx_ok = (x | 1) < qx_end;
    uint32_t quad = 0, res = 0, far = 0;  // bit j: predicted hole quad / residual hole / deep sample
    for (int j = 0; j < h; j++) {
        const int y = y0 + j;
        const uint32_t qm = 3u << (j & 6);  // the quad's two rows
        const bool in_rect = x >= ax && x < ax + ex && y >= ay && y < ay + ey;
        const bool hq = qx_ok && (y | 1) < qy_end && ((cv | ncv) & qm) == 0u;
        const bool reached = (cv >> j) & 1u;
        quad |= hq ? 1u << j : 0u;
        res |= (!reached && !hq && !in_rect) ? 1u << j : 0u;
        far |= (reached && ((dp >> j) & 1u) && !in_rect) ? 1u << j : 0u;
    }
    if (max_cnn >= 0) {
        if (res) atomicAdd(&s_res[gt], (uint32_t)__popc(res));
        if (far) s_far[gt] = 1u;
    }
    __syncthreads();
    const bool gate_far = max_cnn >= 0 && s_far[gt] != 0u;
    const bool gate_res = max_cnn >= 0 && (gate_far || (int)s_res[gt] > max_cnn);
    uint32_t n = 0;
    for (int j = 0; j < h; j++) {
        const int y = y0 + j;
        const size_t o = (size_t)y * res_x + x;
        const bool in_rect = x >= ax && x < ax + ex && y >= ay && y < ay + ey;
        const bool take = !part0 && (in_rect || (((quad >> j) & 1u) && long_blk) ||
                                     (gate_res && ((res >> j) & 1u)) || (gate_far && ((far >> j) & 1u)));
        uint8_t b = kPreNone;
        if (take) {
            b = (uint8_t)trip_bucket(kn[j]);
            atomicAdd(&hist[b], 1u);
            n++;
        }
        bkt[o] = b;
        if ((cv >> j) & 1u) cov[o] = 0;
        if ((dp >> j) & 1u) deep[o] = 0;
    }
    for (int o = 1; o < 8; o <<= 1) n += (uint32_t)__shfl_xor((int)n, o);
    const int nbx = (res_x + 7) >> 3;
    const int bx = x >> 3;
    if ((t & 7) == 0 && bx < nbx && y0 < res_y) blk[(size_t)by * nbx + bx] = n;
    __syncthreads();
    if (t < kPreBuckets && hist[t]) atomicAdd(&hdr[8 + t], hist[t]);
}

// Pretrace a warp frame on the set's own stream: the refresh rectangle and
// the long blocks (k_pre_scan), longest known rays first, with the frame's
// camera, into set s, stamped with the frame's id.  Successive frames use
// successive sets and streams, so their pretraces overlap one another and
// the frames before them; a set is reused after the frame that read it last
// has completed, and the frame's own work on the main stream waits for its
// pretrace.  A primary ray depends only on the camera, the pixel and the
// octree (kernel.cl:594-694, :846-942), so the copied results are those the
// frame's own rays would compute.
// whether frame p's pretrace can predict its holes: the snapshot of frame
// g-k is stored (its event recorded) and the cameras of frames g-k+1 .. g-1
// are known
static bool predict_ready(const svo_ctx* c, const svo_frame_params* p)
{
    const int k = c->pre_predict, g = p->frame;
    if (k <= 0 || !c->d_snap_xyz || g - k < 0) return false;
    if (c->snap_frame[(g - k) % kSnapSlots] != g - k) return false;
    for (int j = 1; j < k; j++)
        if (c->cam_frame[(g - k + j) % kSnapSlots] != g - k + j) return false;
    return true;
}

// the pool stream of set s's part 1: the s-th index left after every i with
// i % 4 == pre_qskip - 1 is taken out (pre_qskip > 0)
static int pool_index(const svo_ctx* c, int s)
{
    int i = 0;
    for (int k = 0;; i++) {
        if (c->pre_qskip > 0 && i % 4 == c->pre_qskip - 1) continue;
        if (k++ == s) break;
    }
    return i;
}

// one part of set s's pretrace: list passes (long blocks, the longest
// blocks only, or predicted holes), then the ray launch into the set's
// results, stamped `id`; records ev_pre[q]
enum PrePart { PRE_LONG, PRE_WHOLE, PRE_PREDICT };
static int pretrace_part(svo_ctx* c, const svo_frame_params* p, int s, int part, PrePart kind, uint32_t id,
                         int gate_max)
{
    const int q = 2 * s + part;
    const size_t N = c->N;
    hipStream_t* sp = part ? &c->pre_pool[pool_index(c, s)] : &c->pre_streams[q];  // the set's stream
    if (!*sp) HIPCHECK(hipStreamCreateWithFlags(sp, hipStreamNonBlocking));
    hipStream_t st = *sp;
    if (c->timing == 2 && part == 1) (void)hipEventRecord(c->ev[0], st);  // the frame starts with its pretrace
    if (c->main_recorded[s]) HIPCHECK(hipStreamWaitEvent(st, c->ev_main[s], 0));
    int ax, ay, ex, ey;
    svo_tile_schedule(c->res_x, c->res_y, p->frame, &ax, &ay, &ex, &ey);
    if (kind == PRE_WHOLE) ax = ay = ex = ey = 0;  // the refresh rectangle is part 1's
    uint32_t* hdr = c->d_pre_list + (size_t)q * (N + kPreHdr);
    uint32_t* list = hdr + kPreHdr;
    uint32_t* blk = c->d_pre_blk + (size_t)q * pre_blocks(c);
    if (kind != PRE_PREDICT) {  // the prediction's first kernel clears the header itself
        hipLaunchKernelGGL(k_memset, dim3(1), dim3(256), 0, st, hdr, 0u, (size_t)kPreHdr);
        HIPCHECK(hipGetLastError());
    }
    const dim3 grid((unsigned)((c->res_x + 63) / 64), (unsigned)(((c->res_y + 7) / 8 + 3) / 4));
    uint8_t* bkt = c->d_pre_bkt + (size_t)q * N;
    if (kind == PRE_PREDICT) {
        // frame g's hole quads as predicted from frame g-k's final plane 0:
        // nearest sample per pixel through the cameras g-k+1 .. g-1, then
        // coverage under camera g (no sample lost by a later z-test is
        // counted, so the quads left cover nearly every hole quad of g)
        const int k = c->pre_predict, g = p->frame;
        const int src_slot = (g - k) % kSnapSlots;
        HIPCHECK(hipStreamWaitEvent(st, c->ev_snap[src_slot], 0));
        const unsigned nb = (unsigned)((N + 255) / 256);
        unsigned long long* keys = c->d_pred_keys + (size_t)s * N;
        float4* inter = c->d_pred_xyz + (size_t)s * N;
        uint8_t* cov = c->d_pred_cov + (size_t)s * N;
        uint8_t* deep = c->d_pred_deep + (size_t)s * N;
        // learned frames with the gate's depth rule: mark where a sample beyond the limit lands
        const float deep_z = gate_max >= 0 && c->gate_depth > 0.0f ? c->gate_depth : 0.0f;
        const float4* cur = c->d_snap_xyz + (size_t)src_slot * N;
        const PredCam fin = pred_cam(p);
        if (k == 1) {
            hipLaunchKernelGGL(k_pred_cover, dim3(nb), dim3(256), 0, st, cur, fin, c->res_x, c->res_y, (uint32_t)N,
                               deep_z, cov, deep, hdr);
        }
        for (int j = 1; j < k; j++) {
            const PredCam mid = pred_cam(&c->cam_hist[(g - k + j) % kSnapSlots]);
            hipLaunchKernelGGL(k_pred_scatter, dim3(nb), dim3(256), 0, st, cur, mid, c->res_x, c->res_y,
                               (uint32_t)N, keys, j == 1 ? hdr : nullptr);
            const bool last = j == k - 1;
            // (k == 3: the intermediate positions go to `inter`, which the
            // next scatter reads; the first scatter read the snapshot)
            hipLaunchKernelGGL(k_pred_resolve, dim3(nb), dim3(256), 0, st, keys, cur, last ? nullptr : inter, fin,
                               c->res_x, c->res_y, (uint32_t)N, deep_z, cov, deep);
            cur = inter;
        }
        HIPCHECK(hipGetLastError());
        // blocks part 0 traces whole are left to it
        const uint32_t whole = c->pre_early ? c->pre_whole_trips : 0u;
        hipLaunchKernelGGL(k_pre_scan_pred, grid, dim3(256), 0, st, (const uint32_t*)c->d_known, cov,
                           deep_z > 0.0f ? deep : nullptr, gate_max, c->res_x, c->res_y, (int)(c->res_x / 16 * 16),
                           (int)(c->res_y / 16 * 16), ax, ay, ex, ey, c->pre_trips, whole, blk, bkt, hdr);
    } else {
        const uint32_t trips = kind == PRE_WHOLE ? c->pre_whole_trips : c->pre_trips;
        hipLaunchKernelGGL(k_pre_scan, grid, dim3(256), 0, st, (const uint32_t*)c->d_known, c->res_x, c->res_y, ax,
                           ay, ex, ey, trips, blk, bkt, hdr);
    }
    HIPCHECK(hipGetLastError());
    const int split_bucket = c->pre_split_trips > 0 ? 31 - __builtin_clz((unsigned)c->pre_split_trips) : -1;
    hipLaunchKernelGGL(k_pre_emit, grid, dim3(256), 0, st, (const uint8_t*)bkt, c->res_x, c->res_y, ax, ay, ex, ey,
                       split_bucket, kind == PRE_PREDICT ? 1 : 0, (const uint32_t*)blk, hdr, list);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventRecord(c->ev_list[q], st));  // bucket bytes final: the frame's first ray launch may read them
    View v = make_view(c, p);  // no pretrace fields: pre_active is clear between frames
    set_rays(v, list, hdr, 0, 0, 0, 0, nullptr);
    v.screen = c->d_pre_word + s * N;
    v.coords = c->d_pre_xyz + s * N;
    v.stamp_out = c->d_pre_stamp + s * N;
    v.pre_id = id;
    // list order (the longest rays first, neighbours together), no spread;
    // the longest go to waves of their own, a few lanes each
    Split split;
    if (split_bucket >= 0) {
        split.count = hdr + 1;
        split.queue = c->d_pre_queue + 4 * q + 2;
        split.shift = c->pre_split_shift;
    }
    RC(launch_rays_v(c, st, v, nullptr, nullptr, 1, c->d_pre_queue + 4 * q, c->pre_seq[q], c->d_queue + 16, 6,
                     c->pre_wg, split));
    HIPCHECK(hipEventRecord(c->ev_pre[q], st));  // the frame's ray launch waits for it
    return SVO_OK;
}

// Pretrace a warp frame into set s (one part, or two with predicted holes
// and a whole-block threshold: part 0 starts at once, part 1 after the
// snapshot it projects).  Parts write different pixels, except where the
// known trips change between their list passes: a pixel traced by both gets
// the same result twice.
static int pretrace(svo_ctx* c, const svo_frame_params* p, int gate_max)
{
    RC(ensure_pre(c));
    const int s = (int)(c->pre_frames % (uint32_t)c->pre_sets);
    const uint32_t id = c->pre_frames + 1u;
    const bool pred = predict_ready(c, p);
    c->pre_early = pred && c->pre_whole_trips > 0;
    if (c->pre_early) RC(pretrace_part(c, p, s, 0, PRE_WHOLE, id, gate_max));
    RC(pretrace_part(c, p, s, 1, pred ? PRE_PREDICT : PRE_LONG, id, gate_max));
    c->pre_frames++;
    c->pre_active = true;
    c->pre_set = s;
    c->pre_id = id;
    return SVO_OK;
}

// after the frame's work is queued: its set may be reused once it completes
static void pretrace_done(svo_ctx* c)
{
    if (hipEventRecord(c->ev_frame_done, c->stream) == hipSuccess) c->frame_done_recorded = true;
    if (!c->pre_active) return;
    c->pre_active = false;
    if (hipEventRecord(c->ev_main[c->pre_set], c->stream) == hipSuccess) c->main_recorded[c->pre_set] = true;
}

// whether a warp or learned frame gets a pretrace: frames >= 2 outside stage
// timing; with only_queued, only while the frame before it is still on the
// GPU (frames issued ahead of it) -- a frame the GPU starts right away has
// no earlier frame to overlap
static bool want_pretrace(svo_ctx* c, const svo_frame_params* p)
{
    if (c->pre_sets <= 0 || c->timing == 1 || p->frame < 2) return false;
    if (!c->pre_only_queued) return true;
    return c->frame_done_recorded && hipEventQuery(c->ev_frame_done) == hipErrorNotReady;
}


extern "C" int svo_frame_warp(svo_ctx* c, const svo_frame_params* p)
{
    RC(need_octree(c));
    if (!p) return svo_fail(SVO_ERR_ARG, "null frame params");
    HIPCHECK(hipSetDevice(c->device));
    take_image(c);
    // a pretrace frame runs direct launches: a graph launch here waits on the
    // host for the GPU, which would keep the host from issuing the next
    // frames' pretraces ahead
    const bool pre = want_pretrace(c, p);
    const bool graph = c->use_graph && c->timing != 1 && p->frame >= 2 && !pre;
    if (graph && !c->graph[0]) RC(build_graph(c));
    float4* snap = snap_begin(c, p);
    frame_begin(c);
    if (pre) RC(pretrace(c, p, -1));  // warp frames: no gate
    int rc;
    if (graph) {
        rc = frame_warp_graph(c, p, snap);
    } else {
        rc = warp_front(c, p);
        const Launch L = single(c);
        const View v = make_view(c, p);
        if (!rc) mark(c, ST_COPY);
        if (!rc) rc = launch_copy(L, v, 2);
        if (!rc) mark(c, ST_FILTER);
        if (!rc) rc = launch_fill(L, v, nullptr, 1, 1);
    }
    if (!rc) snap_end(c, p);
    pretrace_done(c);
    RC(rc);
    return frame_end(c);
}

// the CNN input planes and the staged CNN colours of a learned frame
static int ensure_cnn_buffers(svo_ctx* c)
{
    if (c->d_cnn_words) return SVO_OK;
    const size_t tiles = (size_t)c->n_tiles_x * c->n_tiles_y;
    if (hipMalloc(&c->d_cnn_words, c->N * 4) != hipSuccess || hipMalloc(&c->d_cnn_coords, c->N * 16) != hipSuccess ||
        hipMalloc(&c->d_cnn_out, tiles * 256 * 4) != hipSuccess) {
        (void)hipFree(c->d_cnn_words);
        (void)hipFree(c->d_cnn_coords);
        (void)hipFree(c->d_cnn_out);
        c->d_cnn_words = nullptr;
        c->d_cnn_coords = nullptr;
        c->d_cnn_out = nullptr;
        return svo_fail(SVO_ERR_NOMEM, "learned-frame buffers at %dx%d", c->res_x, c->res_y);
    }
    return SVO_OK;
}

// Warp frame with the learned fill: after hole compaction a per-tile gate
// sends tiles with more than max_cnn_holes residual holes back to
// raycasting (in the frame's ray launch) and the rest to the CNN, whose RGB
// replaces the filtered colour of their hole pixels in the image.  Screen
// words keep the reference filter's result.  cnn == NULL keeps the filter's
// colours (gate only).
static int learned_body(svo_ctx* c, svo_cnn* cnn, const svo_frame_params* p, int max_cnn_holes);

extern "C" int svo_frame_warp_learned(svo_ctx* c, svo_cnn* cnn, const svo_frame_params* p, int max_cnn_holes)
{
    RC(need_octree(c));
    if (!p) return svo_fail(SVO_ERR_ARG, "null frame params");
    if (max_cnn_holes < 0 || max_cnn_holes > 256) return svo_fail(SVO_ERR_ARG, "max_cnn_holes outside 0..256");
    HIPCHECK(hipSetDevice(c->device));
    RC(ensure_cnn_buffers(c));
    take_image(c);
    const bool pre = want_pretrace(c, p);
    snap_begin(c, p);
    frame_begin(c);
    if (pre) RC(pretrace(c, p, max_cnn_holes));
    const int rc = learned_body(c, cnn, p, max_cnn_holes);
    if (!rc) snap_end(c, p);
    pretrace_done(c);
    RC(rc);
    return frame_end(c);
}

// the learned frame after its pretrace (if any) is issued
static int learned_body(svo_ctx* c, svo_cnn* cnn, const svo_frame_params* p, int max_cnn_holes)
{
    RC(warp_compact(c, p));
    int ax, ay, ex, ey;
    svo_tile_schedule(c->res_x, c->res_y, p->frame, &ax, &ay, &ex, &ey);
    View v = make_view(c, p);  // with the pretrace's set when the frame has one
    set_rays(v, c->d_ids + 2 * c->S, c->d_ids, ax, ay, ex, ey, nullptr);
    v.list2 = c->d_gate_rays;
    v.list2_count = c->d_gate + GATE_RAYS;
    v.cnn_words = c->d_cnn_words;
    v.cnn_coords = c->d_cnn_coords;
    const int tiles = c->n_tiles_x * c->n_tiles_y;
    mark(c, ST_GATE);
    hipLaunchKernelGGL(k_gate_reset, dim3(1), dim3(1), 0, c->stream, v, (const View*)nullptr);
    hipLaunchKernelGGL(k_gate, dim3((unsigned)((tiles + GATE_TILES - 1) / GATE_TILES)), dim3(256), 0, c->stream, v,
                       (const View*)nullptr, max_cnn_holes);
    hipLaunchKernelGGL(k_cnn_input, dim3((unsigned)((c->N + 255) / 256)), dim3(256), 0, c->stream, v,
                       (const View*)nullptr);
    HIPCHECK(hipGetLastError());
    c->cnn_input = true;
    // The CNN reads only its input planes and writes only d_cnn_out, so it
    // runs on the second stream alongside the ray launch, which is
    // tail-bound by a few long rays.  Per-stage timing runs it after the
    // fill on the frame's stream instead, so its stage time is its own.
    const bool overlap = cnn && c->timing != 1;
    if (overlap) {
        HIPCHECK(hipEventRecord(c->cnn_fork, c->stream));
        HIPCHECK(hipStreamWaitEvent(c->cnn_stream, c->cnn_fork, 0));
        RC(cnn_fill_tiles(cnn, c, p, c->d_gate_tiles, c->d_gate, tiles, c->cnn_stream, c->cnn_overlap_wgs));
        HIPCHECK(hipEventRecord(c->cnn_join, c->cnn_stream));
    }
    // hole list, gate rays and the refresh tile in one ray launch
    mark(c, ST_RAYS);
    RC(frame_rays(c, v));
    const Launch L = single(c);
    mark(c, ST_COPY);
    RC(launch_copy(L, v, 2));
    mark(c, ST_FILTER);
    RC(launch_fill(L, v, nullptr, 1, 1));
    if (cnn) {
        mark(c, ST_CNN);
        if (overlap)
            HIPCHECK(hipStreamWaitEvent(c->stream, c->cnn_join, 0));
        else
            RC(cnn_fill_tiles(cnn, c, p, c->d_gate_tiles, c->d_gate, tiles, c->stream));
        mark(c, ST_CNN_APPLY);
        RC(cnn_apply_tiles(c, c->d_gate_tiles, c->d_gate, tiles, c->stream));
    }
    return SVO_OK;
}

extern "C" int svo_frame_full(svo_ctx* c, const svo_frame_params* p)
{
    RC(need_octree(c));
    if (!p) return svo_fail(SVO_ERR_ARG, "null frame params");
    HIPCHECK(hipSetDevice(c->device));
    take_image(c);
    frame_begin(c);
    View v = make_view(c, p);
    set_rays(v, nullptr, nullptr, 0, 0, c->res_x, c->res_y, c->d_image);
    mark(c, ST_RAYS);
    if (c->long_trips) {
        // the first full frame after a resize, upload or setting change only collects flags
        const uint32_t nb = v.n_rect >> 6;  // the screen's ray blocks
        uint32_t* cur = c->d_block_long + (size_t)c->long_parity * nb;
        const uint32_t* prev = c->long_valid ? c->d_block_long + (size_t)(c->long_parity ^ 1) * nb : nullptr;
        uint32_t* count = c->d_long_count + 64 * c->long_parity;  // zero: cleared by the previous full frame
        uint32_t* next_count = c->d_long_count + 64 * (c->long_parity ^ 1);
        hipLaunchKernelGGL(k_long_blocks, dim3((nb + 255) / 256), dim3(256), 0, c->stream, prev, cur, c->d_long_list,
                           count, next_count, nb, (uint32_t)v.tiles_x, (uint32_t)v.blk_shift, (uint32_t)c->res_x,
                           (uint32_t)c->res_y);
        HIPCHECK(hipGetLastError());
        v.list = c->d_long_list;
        v.list_count = count;
        v.rect_skip = prev;
        v.block_long = cur;
        v.long_trips = (uint32_t)c->long_trips;
        c->long_parity ^= 1;
        c->long_valid = true;
    }
    RC(launch_rays_v(c, c->stream, v, nullptr, nullptr, 1, c->d_queue, c->launch_seq));
    return frame_end(c);
}

extern "C" int svo_set_ray_block(svo_ctx* c, int width)
{
    if (!c) return svo_fail(SVO_ERR_ARG, "null context");
    int shift = 3;
    while (shift <= 6 && (1 << shift) != width) shift++;
    if (shift > 6) return svo_fail(SVO_ERR_ARG, "ray block width %d is not 8, 16, 32 or 64", width);
    if (c->blk_shift != shift) {
        c->blk_shift = shift;
        c->long_valid = false;  // the long-block flags index blocks of the old shape
    }
    return SVO_OK;
}

extern "C" int svo_set_pretrace(svo_ctx* c, const svo_pretrace_config* cfg)
{
    if (!c || !cfg) return svo_fail(SVO_ERR_ARG, "null argument");
    const int sets = cfg->sets, long_trips = cfg->long_trips, wg_per_cu = cfg->wg_per_cu;
    if (wg_per_cu < 0 || wg_per_cu > 32) return svo_fail(SVO_ERR_ARG, "wg_per_cu %d outside 0..32", wg_per_cu);
    if (cfg->split_trips < 0 || cfg->split_trips > SVO_MAX_RAY_ITERS)
        return svo_fail(SVO_ERR_ARG, "split_trips %d outside 0..%d", cfg->split_trips, (int)SVO_MAX_RAY_ITERS);
    const int sl = cfg->split_lanes;
    if (sl < 1 || sl > 64 || (sl & (sl - 1)))
        return svo_fail(SVO_ERR_ARG, "split_lanes %d is not a power of two in 1..64", sl);
    if (sets < 0 || sets == 1 || sets > kPreMaxSets)
        return svo_fail(SVO_ERR_ARG, "pretrace sets %d: 0 (off) or 2..%d", sets, kPreMaxSets);
    if (long_trips < 1 || long_trips > SVO_MAX_RAY_ITERS)
        return svo_fail(SVO_ERR_ARG, "long_trips %d outside 1..%d", long_trips, (int)SVO_MAX_RAY_ITERS);
    if (cfg->predict < 0 || cfg->predict > kSnapSlots - 1)
        return svo_fail(SVO_ERR_ARG, "predict %d outside 0..%d", cfg->predict, kSnapSlots - 1);
    if (cfg->whole_trips < 0 || cfg->whole_trips > SVO_MAX_RAY_ITERS)
        return svo_fail(SVO_ERR_ARG, "whole_trips %d outside 0..%d", cfg->whole_trips, (int)SVO_MAX_RAY_ITERS);
    if (cfg->queue_skip < 0 || cfg->queue_skip > 4)
        return svo_fail(SVO_ERR_ARG, "queue_skip %d outside 0..4", cfg->queue_skip);
    HIPCHECK(hipSetDevice(c->device));
    if (cfg->queue_skip != c->pre_qskip) {
        // the sets move to other streams: nothing may still run on the old ones
        HIPCHECK(hipDeviceSynchronize());
        c->pre_qskip = cfg->queue_skip;
    }
    if (sets != c->pre_sets) {
        HIPCHECK(hipStreamSynchronize(c->stream));
        free_pre(c);
        c->pre_sets = sets;
    }
    c->pre_trips = (uint32_t)long_trips;
    c->pre_wg = wg_per_cu;
    c->pre_split_trips = cfg->split_trips;
    c->pre_split_shift = 31 - __builtin_clz((unsigned)sl);
    c->pre_only_queued = cfg->only_queued != 0;
    c->pre_predict = cfg->predict;
    c->pre_whole_trips = (uint32_t)cfg->whole_trips;
    return SVO_OK;
}

extern "C" int svo_get_pretrace(svo_ctx* c, svo_pretrace_config* cfg)
{
    if (!c || !cfg) return svo_fail(SVO_ERR_ARG, "null argument");
    cfg->sets = c->pre_sets;
    cfg->long_trips = (int)c->pre_trips;
    cfg->wg_per_cu = c->pre_wg;
    cfg->split_trips = c->pre_split_trips;
    cfg->split_lanes = 1 << c->pre_split_shift;
    cfg->only_queued = c->pre_only_queued ? 1 : 0;
    cfg->predict = c->pre_predict;
    cfg->whole_trips = (int)c->pre_whole_trips;
    cfg->queue_skip = c->pre_qskip;
    return SVO_OK;
}

extern "C" int svo_pretrace_stats(svo_ctx* c, uint64_t* rays, uint64_t* merged, uint32_t* max_trips, int reset)
{
    if (!c || !rays || !merged || !max_trips) return svo_fail(SVO_ERR_ARG, "null argument");
    HIPCHECK(hipSetDevice(c->device));
    HIPCHECK(hipStreamSynchronize(c->stream));
    uint32_t w[10];
    HIPCHECK(hipMemcpy(w, c->d_queue + 12, 40, hipMemcpyDeviceToHost));
    memcpy(merged, w, 8);
    memcpy(rays, w + 4, 8);
    *max_trips = w[6];
    if (reset) HIPCHECK(hipMemset(c->d_queue + 12, 0, 40));
    return SVO_OK;
}

extern "C" int svo_set_long_first(svo_ctx* c, int long_trips)
{
    if (!c) return svo_fail(SVO_ERR_ARG, "null context");
    if (long_trips < 0 || long_trips > SVO_MAX_RAY_ITERS)
        return svo_fail(SVO_ERR_ARG, "long_trips %d outside 0..%d", long_trips, (int)SVO_MAX_RAY_ITERS);
    if (c->long_trips != long_trips) {
        c->long_trips = long_trips;
        c->long_valid = false;  // flags are not kept while off and mean another threshold
    }
    return SVO_OK;
}

extern "C" int svo_set_gate_depth(svo_ctx* c, float max_depth)
{
    if (!c) return svo_fail(SVO_ERR_ARG, "null context");
    if (!(max_depth >= 0.0f) || max_depth > 1e30f) return svo_fail(SVO_ERR_ARG, "gate depth must be >= 0");
    c->gate_depth = max_depth;
    return SVO_OK;
}

extern "C" int svo_set_cnn_overlap(svo_ctx* c, int workgroups)
{
    if (!c) return svo_fail(SVO_ERR_ARG, "null context");
    if (workgroups < 0 || workgroups > 65536)
        return svo_fail(SVO_ERR_ARG, "CNN workgroups %d outside 0..65536", workgroups);
    c->cnn_overlap_wgs = workgroups;
    return SVO_OK;
}

extern "C" int svo_memset(svo_ctx* c, size_t byte_ofs, uint32_t val, size_t bytes)
{
    RC(need_frame(c));
    if ((byte_ofs | bytes) & 3) return svo_fail(SVO_ERR_ARG, "memset offset/size must be multiples of 4");
    if (byte_ofs + bytes > 16 * c->N) return svo_fail(SVO_ERR_ARG, "memset past the screen buffer");
    HIPCHECK(hipSetDevice(c->device));
    RC(launch_memset(c, c->d_screen + byte_ofs / 4, val, bytes / 4));
    const size_t p1b = 4 * c->N, p1e = 8 * c->N;
    if (byte_ofs <= p1b && byte_ofs + bytes >= p1e) c->plane1_holes = (val == HOLE);
    else if (byte_ofs < p1e && byte_ofs + bytes > p1b && val != HOLE) c->plane1_holes = false;
    return SVO_OK;
}

extern "C" int svo_proj(svo_ctx* c, int plane, const svo_frame_params* p)
{
    RC(need_frame(c));
    if (!p || plane < 1 || plane > 3) return svo_fail(SVO_ERR_ARG, "plane %d outside 1..3", plane);
    HIPCHECK(hipSetDevice(c->device));
    View v = make_view(c, p);
    v.plane_lo = plane;
    v.n_planes = 1;
    return launch_scatter_resolve(single(c), v, 1, 0);
}

extern "C" int svo_counthole(svo_ctx* c)
{
    RC(need_frame(c));
    HIPCHECK(hipSetDevice(c->device));
    return launch_count(single(c), make_view(c, nullptr));
}
extern "C" int svo_sumids(svo_ctx* c)
{
    RC(need_frame(c));
    HIPCHECK(hipSetDevice(c->device));
    return launch_sum(single(c), make_view(c, nullptr));
}
extern "C" int svo_writeids(svo_ctx* c)
{
    RC(need_frame(c));
    HIPCHECK(hipSetDevice(c->device));
    return launch_write(single(c), make_view(c, nullptr));
}
extern "C" int svo_raycast_holes(svo_ctx* c, const svo_frame_params* p)
{
    RC(need_octree(c));
    if (!p) return svo_fail(SVO_ERR_ARG, "null frame params");
    HIPCHECK(hipSetDevice(c->device));
    return launch_rays(c, p, c->d_ids + 2 * c->S, c->d_ids, 0, 0, 0, 0, nullptr);
}
extern "C" int svo_raycast_rect(svo_ctx* c, const svo_frame_params* p, int add_x, int add_y, int ext_x, int ext_y)
{
    RC(need_octree(c));
    if (!p) return svo_fail(SVO_ERR_ARG, "null frame params");
    if (add_x < 0 || add_y < 0 || ext_x < 0 || ext_y < 0 || ext_x > 65536 || ext_y > 65536)
        return svo_fail(SVO_ERR_ARG, "bad rectangle");
    HIPCHECK(hipSetDevice(c->device));
    return launch_rays(c, p, nullptr, nullptr, add_x, add_y, ext_x, ext_y, nullptr);
}
extern "C" int svo_raycast_quads(svo_ctx* c, const svo_frame_params* p, int add_x, int add_y)
{
    RC(need_octree(c));
    if (!p) return svo_fail(SVO_ERR_ARG, "null frame params");
    if (add_x < 0 || add_y < 0 || add_x > 65535 || add_y > 65535) return svo_fail(SVO_ERR_ARG, "bad quad offset");
    const int gx_n = ((c->res_x / 4 + 15) / 16) * 16, gy_n = ((c->res_y / 4 + 15) / 16) * 16;
    if ((size_t)gx_n * gy_n > c->N) return svo_fail(SVO_ERR_ARG, "quad grid larger than the frame");
    HIPCHECK(hipSetDevice(c->device));
    View v = make_view(c, p);
    hipLaunchKernelGGL(k_quad_select, dim3((unsigned)((gx_n * gy_n + 255) / 256)), dim3(256), 0, c->stream, v, add_x,
                       add_y, p->frame, gx_n, gy_n, c->d_gate_rays, c->d_gate + GATE_RAYS);
    HIPCHECK(hipGetLastError());
    set_rays(v, c->d_gate_rays, c->d_gate + GATE_RAYS, 0, 0, 0, 0, nullptr);
    v.edge_ok = 1;
    return launch_rays_v(c, c->stream, v, nullptr, nullptr, 1, c->d_queue, c->launch_seq);
}

extern "C" int svo_fillhole_punch(svo_ctx* c, const int32_t* motion_x, const int32_t* motion_y)
{
    RC(need_frame(c));
    HIPCHECK(hipSetDevice(c->device));
    if ((motion_x || motion_y) && !c->d_motion && hipMalloc(&c->d_motion, c->N * 8) != hipSuccess) {
        c->d_motion = nullptr;
        return svo_fail(SVO_ERR_NOMEM, "motion buffers at %dx%d", c->res_x, c->res_y);
    }
    if (motion_x || motion_y) {  // the caller's arrays are read before returning
        HIPCHECK(hipStreamSynchronize(c->stream));
        if (motion_x) HIPCHECK(hipMemcpy(c->d_motion, motion_x, c->N * 4, hipMemcpyHostToDevice));
        if (motion_y) HIPCHECK(hipMemcpy(c->d_motion + c->N, motion_y, c->N * 4, hipMemcpyHostToDevice));
    }
    hipLaunchKernelGGL(k_fillhole_punch, dim3((unsigned)((c->N + 255) / 256)), dim3(256), 0, c->stream,
                       make_view(c, nullptr), motion_x ? c->d_motion : nullptr,
                       motion_y ? c->d_motion + c->N : nullptr);
    HIPCHECK(hipGetLastError());
    return SVO_OK;
}

extern "C" int svo_copy_cache(svo_ctx* c, int dst_plane)
{
    RC(need_frame(c));
    if (dst_plane < 1 || dst_plane > 3) return svo_fail(SVO_ERR_ARG, "plane %d outside 1..3", dst_plane);
    HIPCHECK(hipSetDevice(c->device));
    if (dst_plane == 1) c->plane1_holes = false;
    return launch_copy(single(c), make_view(c, nullptr), dst_plane);
}
extern "C" int svo_fillhole(svo_ctx* c)
{
    RC(need_frame(c));
    HIPCHECK(hipSetDevice(c->device));
    HIPCHECK(hipMemcpyAsync(c->d_snap, c->d_screen, c->N * 4, hipMemcpyDeviceToDevice, c->stream));
    return launch_fill(single(c), make_view(c, nullptr), c->d_snap, 1, 0);
}
extern "C" int svo_colorize(svo_ctx* c)
{
    RC(need_frame(c));
    HIPCHECK(hipSetDevice(c->device));
    take_image(c);
    return launch_fill(single(c), make_view(c, nullptr), c->d_screen, 0, 1);
}

// ---------------------------------------------------------------- batched views
// Many views (contexts) of one scene run each warp stage as one launch over a
// device view table (grid z = view) and all their hole + tile rays in ONE
// persistent ray launch, so the long-ray tails of the views overlap.
struct svo_batch {
    std::vector<svo_ctx*> views;
    std::vector<size_t> view_n;  // each view's pixel count and octree at creation
    std::vector<uint32_t*> view_oct;
    size_t n_words = 0;
    uint32_t root = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t cnn_stream = nullptr;  // the views' CNN fills, alongside the ray launch
    hipEvent_t cnn_fork = nullptr, cnn_join = nullptr;
    View* d_views = nullptr;     // 2 slots x n
    View* h_views = nullptr;     // pinned staging, 2 slots x n
    hipEvent_t slot_done[2] = {};
    std::vector<hipEvent_t> pre;  // per view: its stream's queued work
    uint32_t* d_first = nullptr;
    uint32_t* d_queue = nullptr;
    // learned frames: the CNN tile table of all views (2 slots, like the views)
    char* d_cnn_tab = nullptr;
    char* h_cnn_tab = nullptr;
    uint32_t* d_cnn_first = nullptr;
    int seq = 0;
    int frame_no = 0;
    uint32_t maxN = 0, maxS = 0;
};

extern "C" int svo_batch_destroy(svo_batch* b)
{
    if (!b) return SVO_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->cnn_stream) (void)hipStreamSynchronize(b->cnn_stream);
    (void)hipFree(b->d_views);
    (void)hipHostFree(b->h_views);
    (void)hipFree(b->d_first);
    (void)hipFree(b->d_queue);
    (void)hipFree(b->d_cnn_tab);
    (void)hipHostFree(b->h_cnn_tab);
    (void)hipFree(b->d_cnn_first);
    for (auto& e : b->slot_done)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : b->pre)
        if (e) (void)hipEventDestroy(e);
    if (b->cnn_fork) (void)hipEventDestroy(b->cnn_fork);
    if (b->cnn_join) (void)hipEventDestroy(b->cnn_join);
    if (b->cnn_stream) (void)hipStreamDestroy(b->cnn_stream);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return SVO_OK;
}

extern "C" int svo_batch_create(svo_batch** out, svo_ctx* const* views, int n_views)
{
    if (!out || !views || n_views < 1 || n_views > 256) return svo_fail(SVO_ERR_ARG, "1..256 views required");
    *out = nullptr;
    svo_ctx* v0 = views[0];
    for (int i = 0; i < n_views; i++) {
        svo_ctx* c = views[i];
        if (!c || !c->d_screen || !c->d_oct) return svo_fail(SVO_ERR_STATE, "view %d has no octree or frame", i);
        if (c->device != v0->device || c->n_words != v0->n_words || c->root != v0->root || c->depth != v0->depth)
            return svo_fail(SVO_ERR_ARG, "view %d does not hold the same octree on the same device", i);
        for (int j = 0; j < i; j++)
            if (views[j] == c) return svo_fail(SVO_ERR_ARG, "view %d repeats view %d", i, j);
    }
    HIPCHECK(hipSetDevice(v0->device));
    svo_batch* b = new svo_batch();
    b->device = v0->device;
    b->views.assign(views, views + n_views);
    for (auto* c : b->views) {
        b->view_n.push_back(c->N);
        b->view_oct.push_back(c->d_oct);
    }
    b->n_words = v0->n_words;
    b->root = v0->root;
    for (auto* c : b->views) {
        b->maxN = std::max(b->maxN, (uint32_t)c->N);
        b->maxS = std::max(b->maxS, (uint32_t)c->S);
    }
    bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess &&
              hipStreamCreateWithFlags(&b->cnn_stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&b->cnn_fork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&b->cnn_join, hipEventDisableTiming) == hipSuccess &&
              hipMalloc(&b->d_views, sizeof(View) * 2 * n_views) == hipSuccess &&
              hipHostMalloc(&b->h_views, sizeof(View) * 2 * n_views) == hipSuccess &&
              hipMalloc(&b->d_first, 4 * (n_views + 1)) == hipSuccess && hipMalloc(&b->d_queue, 64) == hipSuccess &&
              hipMalloc(&b->d_cnn_tab, cnn_tile_view_bytes() * 2 * n_views) == hipSuccess &&
              hipHostMalloc(&b->h_cnn_tab, cnn_tile_view_bytes() * 2 * n_views) == hipSuccess &&
              hipMalloc(&b->d_cnn_first, 4 * (n_views + 1)) == hipSuccess &&
              hipMemset(b->d_queue, 0, 64) == hipSuccess &&
              hipEventCreateWithFlags(&b->slot_done[0], hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&b->slot_done[1], hipEventDisableTiming) == hipSuccess &&
              hipEventRecord(b->slot_done[0], b->stream) == hipSuccess &&
              hipEventRecord(b->slot_done[1], b->stream) == hipSuccess;
    b->pre.assign(n_views, nullptr);
    for (int i = 0; ok && i < n_views; i++)
        ok = hipEventCreateWithFlags(&b->pre[i], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        svo_batch_destroy(b);
        return svo_fail(SVO_ERR_HIP, "batch set-up failed");
    }
    *out = b;
    return SVO_OK;
}

// svo_frame_warp for every view of the batch (params[i] for view i), with all
// stages batched over the views.  Results per view are identical to
// svo_frame_warp on that view alone.
static int batch_frame(svo_batch* b, const svo_frame_params* params, svo_cnn* cnn, int max_cnn_holes, bool learned);

extern "C" int svo_batch_frame_warp(svo_batch* b, const svo_frame_params* params)
{
    return batch_frame(b, params, nullptr, 0, false);
}

// svo_frame_warp_learned for every view of the batch: the gate of all views
// in one launch, their gate pixels in the shared ray launch, the CNN per view.
extern "C" int svo_batch_frame_warp_learned(svo_batch* b, svo_cnn* cnn, const svo_frame_params* params,
                                            int max_cnn_holes)
{
    if (max_cnn_holes < 0 || max_cnn_holes > 256) return svo_fail(SVO_ERR_ARG, "max_cnn_holes outside 0..256");
    return batch_frame(b, params, cnn, max_cnn_holes, true);
}

static int batch_frame(svo_batch* b, const svo_frame_params* params, svo_cnn* cnn, int max_cnn_holes, bool learned)
{
    if (!b || !params) return svo_fail(SVO_ERR_ARG, "null argument");
    for (size_t i = 0; i < b->views.size(); i++)
        if (b->views[i]->N != b->view_n[i] || b->views[i]->d_oct != b->view_oct[i] || !b->views[i]->d_screen ||
            b->views[i]->n_words != b->n_words || b->views[i]->root != b->root)
            return svo_fail(SVO_ERR_STATE, "view %zu was resized or given a new octree after svo_batch_create", i);
    const int n = (int)b->views.size();
    svo_ctx* c0 = b->views[0];
    HIPCHECK(hipSetDevice(b->device));
    if (learned)
        for (auto* c : b->views) RC(ensure_cnn_buffers(c));
    // order after the views' own queued work
    for (int i = 0; i < n; i++) {
        HIPCHECK(hipEventRecord(b->pre[i], b->views[i]->stream));
        HIPCHECK(hipStreamWaitEvent(b->stream, b->pre[i], 0));
    }
    const int slot = b->frame_no & 1;
    HIPCHECK(hipEventSynchronize(b->slot_done[slot]));  // staging slot free again
    View* hv = b->h_views + slot * n;
    View* dv = b->d_views + slot * n;
    for (int i = 0; i < n; i++) {
        svo_ctx* c = b->views[i];
        const svo_frame_params* p = &params[i];
        if (p->frame < 2) {
            hipLaunchKernelGGL(k_memset, dim3((unsigned)std::min<size_t>((4 * c->N + 255) / 256, 4096)), dim3(256),
                               0, b->stream, c->d_screen, HOLE, 4 * c->N);
            c->plane1_holes = true;
        }
        take_image(c);
        View v = make_view(c, p);
        v.plane_lo = c->plane1_holes ? 2 : 1;
        v.n_planes = 3 - v.plane_lo;
        int ax, ay, ex, ey;
        svo_tile_schedule(c->res_x, c->res_y, p->frame, &ax, &ay, &ex, &ey);
        set_rays(v, c->d_ids + 2 * c->S, c->d_ids, ax, ay, ex, ey, nullptr);
        if (learned) {
            v.list2 = c->d_gate_rays;
            v.list2_count = c->d_gate + GATE_RAYS;
            v.cnn_words = c->d_cnn_words;
            v.cnn_coords = c->d_cnn_coords;
        }
        hv[i] = v;
    }
    HIPCHECK(hipMemcpyAsync(dv, hv, sizeof(View) * n, hipMemcpyHostToDevice, b->stream));
    const Launch L{b->stream, dv, n, b->maxN, b->maxS};
    const View dummy = hv[0];
    RC(launch_scatter_resolve(L, dummy, 2, 1));
    RC(launch_count(L, dummy));
    RC(launch_sum(L, dummy));
    RC(launch_write(L, dummy));
    if (learned) {
        uint32_t max_gate_tiles = 0;
        for (auto* c : b->views) max_gate_tiles = std::max(max_gate_tiles, (uint32_t)(c->n_tiles_x * c->n_tiles_y));
        hipLaunchKernelGGL(k_gate_reset, dim3(1, (unsigned)n), dim3(1), 0, b->stream, dummy, (const View*)dv);
        hipLaunchKernelGGL(k_gate, dim3((max_gate_tiles + GATE_TILES - 1) / GATE_TILES, (unsigned)n), dim3(256), 0,
                           b->stream, dummy, (const View*)dv, max_cnn_holes);
        hipLaunchKernelGGL(k_cnn_input, dim3((b->maxN + 255) / 256, 1, (unsigned)n), dim3(256), 0, b->stream, dummy,
                           (const View*)dv);
        HIPCHECK(hipGetLastError());
        for (auto* c : b->views) c->cnn_input = true;
        if (cnn) {  // all views' CNN tiles in one launch on the batch's second stream, alongside the rays
            HIPCHECK(hipEventRecord(b->cnn_fork, b->stream));
            HIPCHECK(hipStreamWaitEvent(b->cnn_stream, b->cnn_fork, 0));
            const size_t tb = cnn_tile_view_bytes() * n;
            RC(cnn_fill_tiles_batch(cnn, b->views.data(), params, n, b->h_cnn_tab + slot * tb,
                                    b->d_cnn_tab + slot * tb, b->d_cnn_first, b->cnn_stream));
            HIPCHECK(hipEventRecord(b->cnn_join, b->cnn_stream));
        }
    }
    hipLaunchKernelGGL(k_view_first, dim3(1), dim3(1), 0, b->stream, (const View*)dv, n, b->d_first,
                       (uint32_t*)nullptr);
    HIPCHECK(hipGetLastError());
    RC(launch_rays_v(c0, b->stream, dummy, dv, b->d_first, n, b->d_queue, b->seq));
    RC(launch_copy(L, dummy, 2));
    RC(launch_fill(L, dummy, nullptr, 1, 1));
    if (learned && cnn) {
        HIPCHECK(hipStreamWaitEvent(b->stream, b->cnn_join, 0));
        int max_tiles = 0;
        for (auto* c : b->views) max_tiles = std::max(max_tiles, c->n_tiles_x * c->n_tiles_y);
        RC(cnn_apply_tiles_batch(b->d_cnn_tab + slot * cnn_tile_view_bytes() * n, n, max_tiles, b->stream));
    }
    HIPCHECK(hipEventRecord(b->slot_done[slot], b->stream));
    for (auto* c : b->views) HIPCHECK(hipStreamWaitEvent(c->stream, b->slot_done[slot], 0));
    b->frame_no++;
    return SVO_OK;
}

extern "C" int svo_batch_sync(svo_batch* b)
{
    if (!b) return svo_fail(SVO_ERR_ARG, "null batch");
    HIPCHECK(hipSetDevice(b->device));
    HIPCHECK(hipStreamSynchronize(b->stream));
    return SVO_OK;
}

static int buf_ptr(svo_ctx* c, int buf, size_t byte_ofs, size_t bytes, char** p)
{
    size_t cap = 0;
    char* base = nullptr;
    switch (buf) {
    case SVO_BUF_SCREEN: base = (char*)c->d_screen; cap = 16 * c->N; break;
    case SVO_BUF_COORDS: base = (char*)c->d_coords; cap = 64 * c->N; break;
    case SVO_BUF_IDS: base = (char*)c->d_ids; cap = 4 * (2 * c->S + c->N); break;
    case SVO_BUF_IMAGE: base = (char*)c->d_image; cap = 4 * c->N; break;
    case SVO_BUF_OCTREE: base = (char*)c->d_oct; cap = 4 * c->n_words; break;
    default: return svo_fail(SVO_ERR_ARG, "unknown buffer %d", buf);
    }
    if (!base) return svo_fail(SVO_ERR_STATE, "buffer %d not allocated", buf);
    if (byte_ofs > cap || bytes > cap - byte_ofs) return svo_fail(SVO_ERR_ARG, "range outside buffer %d", buf);
    *p = base + byte_ofs;
    return SVO_OK;
}

extern "C" int svo_read(svo_ctx* c, int buf, size_t byte_ofs, size_t bytes, void* dst)
{
    if (!c || (!dst && bytes)) return svo_fail(SVO_ERR_ARG, "null argument");
    char* p;
    RC(buf_ptr(c, buf, byte_ofs, bytes, &p));
    HIPCHECK(hipSetDevice(c->device));
    HIPCHECK(hipMemcpyAsync(dst, p, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return SVO_OK;
}

extern "C" int svo_host_alloc(size_t bytes, void** out)
{
    if (!out || bytes == 0) return svo_fail(SVO_ERR_ARG, "bad host allocation arguments");
    *out = nullptr;
    HIPCHECK(hipHostMalloc(out, bytes, hipHostMallocDefault));
    return SVO_OK;
}

extern "C" int svo_host_free(void* p)
{
    if (p) HIPCHECK(hipHostFree(p));
    return SVO_OK;
}

extern "C" int svo_read_image_async(svo_ctx* c, uint32_t* dst)
{
    RC(need_frame(c));
    if (!dst) return svo_fail(SVO_ERR_ARG, "null destination");
    HIPCHECK(hipSetDevice(c->device));
    if (!c->copy_stream) {
        bool ok = hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking) == hipSuccess &&
                  hipEventCreateWithFlags(&c->image_ready, hipEventDisableTiming) == hipSuccess &&
                  hipEventCreateWithFlags(&c->image_copied[0], hipEventDisableTiming) == hipSuccess &&
                  hipEventCreateWithFlags(&c->image_copied[1], hipEventDisableTiming) == hipSuccess;
        if (!ok) return svo_fail(SVO_ERR_HIP, "image copy stream set-up failed");
    }
    if (!c->image_buf[1]) {
        if (hipMalloc(&c->image_buf[1], 4 * c->N) != hipSuccess) {
            c->image_buf[1] = nullptr;
            return svo_fail(SVO_ERR_NOMEM, "second image buffer");
        }
        HIPCHECK(hipMemset(c->image_buf[1], 0, 4 * c->N));
    }
    const int s = c->d_image == c->image_buf[1] ? 1 : 0;
    HIPCHECK(hipEventRecord(c->image_ready, c->stream));
    HIPCHECK(hipStreamWaitEvent(c->copy_stream, c->image_ready, 0));
    HIPCHECK(hipMemcpyAsync(dst, c->d_image, 4 * c->N, hipMemcpyDeviceToHost, c->copy_stream));
    HIPCHECK(hipEventRecord(c->image_copied[s], c->copy_stream));
    c->image_copy_pending[s] = true;
    c->image_flip = true;
    return SVO_OK;
}

extern "C" int svo_write(svo_ctx* c, int buf, size_t byte_ofs, size_t bytes, const void* src)
{
    if (!c || (!src && bytes)) return svo_fail(SVO_ERR_ARG, "null argument");
    if (buf == SVO_BUF_OCTREE) return svo_fail(SVO_ERR_ARG, "use svo_upload_octree for the octree");
    char* p;
    RC(buf_ptr(c, buf, byte_ofs, bytes, &p));
    HIPCHECK(hipSetDevice(c->device));
    HIPCHECK(hipStreamSynchronize(c->stream));
    HIPCHECK(hipMemcpy(p, src, bytes, hipMemcpyHostToDevice));
    if (buf == SVO_BUF_SCREEN && byte_ofs < 8 * c->N && byte_ofs + bytes > 4 * c->N) c->plane1_holes = false;
    return SVO_OK;
}

extern "C" int svo_sync(svo_ctx* c)
{
    if (!c) return svo_fail(SVO_ERR_ARG, "null context");
    HIPCHECK(hipSetDevice(c->device));
    HIPCHECK(hipStreamSynchronize(c->stream));
    if (c->copy_stream) {
        HIPCHECK(hipStreamSynchronize(c->copy_stream));
        c->image_copy_pending[0] = c->image_copy_pending[1] = false;
    }
    if (c->frame_timed) {
        c->frame_timed = false;
        int last = ST_N, first = -1;
        for (int i = ST_N - 1; i >= 0; i--) {
            c->stage_ms[i] = 0.0f;
            if (!c->stage_used[i]) continue;
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, c->ev[i], c->ev[last]);
            c->stage_ms[i] = ms;
            last = i;
            first = i;
        }
        c->frame_ms = 0.0f;
        if (first >= 0 || c->timing == 2)
            (void)hipEventElapsedTime(&c->frame_ms, c->ev[first >= 0 ? first : 0], c->ev[ST_N]);
    }
    return SVO_OK;
}

extern "C" int svo_gate_counts(svo_ctx* c, uint32_t* cnn_tiles, uint32_t* gate_rays)
{
    RC(need_frame(c));
    if (!cnn_tiles || !gate_rays) return svo_fail(SVO_ERR_ARG, "null argument");
    uint32_t g[2];
    HIPCHECK(hipSetDevice(c->device));
    HIPCHECK(hipStreamSynchronize(c->stream));
    HIPCHECK(hipMemcpy(&g[0], c->d_gate, 4, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(&g[1], c->d_gate + GATE_RAYS, 4, hipMemcpyDeviceToHost));
    *cnn_tiles = g[0];
    *gate_rays = g[1];
    return SVO_OK;
}

extern "C" int svo_hole_count(svo_ctx* c, uint32_t* holes)
{
    if (!holes) return svo_fail(SVO_ERR_ARG, "null argument");
    return svo_read(c, SVO_BUF_IDS, 0, 4, holes);
}

extern "C" int svo_ray_count(svo_ctx* c, uint64_t* rays, int reset)
{
    if (!rays) return svo_fail(SVO_ERR_ARG, "null argument");
    uint32_t longest = 0;
    return svo_ray_stats(c, rays, &longest, reset);
}

extern "C" int svo_ray_stats(svo_ctx* c, uint64_t* rays, uint32_t* max_trips, int reset)
{
    if (!c || !rays || !max_trips) return svo_fail(SVO_ERR_ARG, "null argument");
    HIPCHECK(hipSetDevice(c->device));
    HIPCHECK(hipStreamSynchronize(c->stream));
    uint32_t w[3];
    HIPCHECK(hipMemcpy(w, c->d_queue + 8, 12, hipMemcpyDeviceToHost));
    memcpy(rays, w, 8);
    *max_trips = w[2];
    if (reset) HIPCHECK(hipMemset(c->d_queue + 8, 0, 12));
    return SVO_OK;
}

extern "C" int svo_enable_timing(svo_ctx* c, int mode)
{
    if (!c) return svo_fail(SVO_ERR_ARG, "null context");
    if (mode < 0 || mode > 3) return svo_fail(SVO_ERR_ARG, "timing mode %d outside 0..3", mode);
    if (mode == 3) {
        HIPCHECK(hipSetDevice(c->device));
        for (auto& e : c->end_ring)
            if (!e) HIPCHECK(hipEventCreate(&e));
        c->end_count = 0;
    }
    c->timing = mode;
    return SVO_OK;
}

extern "C" int svo_frame_intervals(svo_ctx* c, float* ms, int max_n, int* n)
{
    if (!c || !n || (max_n > 0 && !ms) || max_n < 0) return svo_fail(SVO_ERR_ARG, "bad argument");
    HIPCHECK(hipSetDevice(c->device));
    HIPCHECK(hipStreamSynchronize(c->stream));
    const int k = std::min(std::max(c->end_count - 1, 0), max_n);
    for (int i = 0; i < k; i++) HIPCHECK(hipEventElapsedTime(&ms[i], c->end_ring[i], c->end_ring[i + 1]));
    *n = k;
    return SVO_OK;
}

extern "C" int svo_frame_time(svo_ctx* c, float* ms)
{
    if (!c || !ms) return svo_fail(SVO_ERR_ARG, "null argument");
    *ms = c->frame_ms;
    return SVO_OK;
}

extern "C" int svo_set_ray_spread(svo_ctx* c, int group, int only_if_fits)
{
    if (!c) return svo_fail(SVO_ERR_ARG, "null context");
    if (group < 1 || group > 64 || (group & (group - 1)))
        return svo_fail(SVO_ERR_ARG, "spread group %d is not a power of two in 1..64", group);
    HIPCHECK(hipSetDevice(c->device));
    HIPCHECK(hipStreamSynchronize(c->stream));
    drop_graph(c);
    int shift = 0;
    while ((1 << shift) < group) shift++;
    c->spread_shift = shift;
    c->spread_fit = only_if_fits != 0;
    return SVO_OK;
}

extern "C" int svo_set_ray_lanes(svo_ctx* c, int lanes)
{
    if (!c) return svo_fail(SVO_ERR_ARG, "null context");
    if (lanes < 1 || lanes > 64 || (lanes & (lanes - 1)))
        return svo_fail(SVO_ERR_ARG, "lanes %d is not a power of two in 1..64", lanes);
    HIPCHECK(hipSetDevice(c->device));
    HIPCHECK(hipStreamSynchronize(c->stream));
    drop_graph(c);
    int shift = 0;
    while ((1 << shift) < lanes) shift++;
    c->lanes_shift = shift;
    return SVO_OK;
}

extern "C" int svo_set_graph(svo_ctx* c, int enable)
{
    if (!c) return svo_fail(SVO_ERR_ARG, "null context");
    HIPCHECK(hipSetDevice(c->device));
    HIPCHECK(hipStreamSynchronize(c->stream));
    c->use_graph = enable != 0;
    if (!c->use_graph) drop_graph(c);
    return SVO_OK;
}

extern "C" int svo_stage_times(svo_ctx* c, float* ms, int n)
{
    if (!c || (!ms && n)) return svo_fail(SVO_ERR_ARG, "null argument");
    int k = std::min(n, (int)ST_N);
    for (int i = 0; i < k; i++) ms[i] = c->stage_ms[i];
    return k;
}

extern "C" const char* svo_stage_name(int stage)
{
    return (stage >= 0 && stage < ST_N) ? kStageNames[stage] : "";
}

extern "C" int svo_set_tuning(svo_ctx* c, int refill_threshold, int wg_per_cu, int iters_per_refill,
                              int lds_cache_words)
{
    if (!c) return svo_fail(SVO_ERR_ARG, "null context");
    if (refill_threshold == 0 || refill_threshold > 64 || wg_per_cu == 0 || wg_per_cu > 32 ||
        iters_per_refill == 0 || iters_per_refill > 4096 || lds_cache_words > 12288)
        return svo_fail(SVO_ERR_ARG, "tuning value out of range");
    HIPCHECK(hipSetDevice(c->device));
    HIPCHECK(hipStreamSynchronize(c->stream));
    drop_graph(c);
    if (refill_threshold > 0) c->refill = refill_threshold;
    if (wg_per_cu > 0) c->wg_per_cu = wg_per_cu;
    if (iters_per_refill > 0) c->iters = iters_per_refill;
    if (lds_cache_words >= 0) c->cache_req = (uint32_t)lds_cache_words;
    return SVO_OK;
}

// ---------------------------------------------------------------- warp without allocation
// svo_proj_noalloc keeps its 64-bit warp keys (depth | source | colour, one
// per target pixel) in this file-scope device array instead of a context's
// buffer: every GPU the library runs on holds its own copy (16 MiB) until the
// process ends.  In host code `scratch` names a host-side stand-in, not a
// device address; the address comes from hipGetSymbolAddress once per device
// (with that device current) and only that pointer reaches kernels.
static __device__ __align__(256) unsigned long long scratch[SVO_NOALLOC_MAX_PIXELS];

static std::mutex g_scratch_mu;
static unsigned long long* g_scratch_dev[64];

static int scratch_for_device(int device, unsigned long long** out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return svo_fail(SVO_ERR_HIP, "key array of device %d: %s", device, hipGetErrorString(e));
        g_scratch_dev[device] = (unsigned long long*)ptr;
    }
    *out = g_scratch_dev[device];
    return SVO_OK;
}

extern "C" int svo_proj_noalloc(int device, uint32_t* screen, float* coords, int res_x, int res_y, int plane,
                                const svo_frame_params* p, void* stream)
{
    if (!screen || !coords || !p) return svo_fail(SVO_ERR_ARG, "null argument");
    if (plane < 1 || plane > 3) return svo_fail(SVO_ERR_ARG, "plane %d outside 1..3", plane);
    if (res_x < 16 || res_y < 16 || res_x > 8192 || res_y > 8192)
        return svo_fail(SVO_ERR_ARG, "resolution %dx%d outside 16..8192", res_x, res_y);
    const size_t N = (size_t)res_x * res_y;
    if (N > SVO_NOALLOC_MAX_PIXELS)
        return svo_fail(SVO_ERR_ARG, "%dx%d is %zu pixels, more than the %u svo_proj_noalloc takes", res_x, res_y, N,
                        (unsigned)SVO_NOALLOC_MAX_PIXELS);
    if (((uintptr_t)coords & 15) || ((uintptr_t)screen & 3))
        return svo_fail(SVO_ERR_ARG, "screen must be 4-byte and coords 16-byte aligned");
    int n = 0;
    HIPCHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n || device >= 64) return svo_fail(SVO_ERR_ARG, "device %d not present", device);
    HIPCHECK(hipSetDevice(device));
    unsigned long long* keys = nullptr;
    RC(scratch_for_device(device, &keys));
    View v;
    memset(&v, 0, sizeof(v));
    v.screen = screen;
    v.coords = (float4*)coords;
    v.keys = keys;
    v.res_x = res_x;
    v.res_y = res_y;
    v.N = (uint32_t)N;
    v.S = (uint32_t)((res_x / 16) * (res_y / 16));
    v.tiles_x = 1;
    view_camera(v, p);
    v.plane_lo = plane;
    v.n_planes = 1;
    const hipStream_t st = (hipStream_t)stream;
    // the keys start at all ones; the resolve resets every key it reads, but
    // each call sets them again and so depends on no earlier call
    const unsigned blocks = (unsigned)std::min<size_t>((N + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_u64, dim3(blocks), dim3(256), 0, st, keys, ~0ull, N);
    HIPCHECK(hipGetLastError());
    const Launch L{st, nullptr, 1, (uint32_t)N, v.S};
    return launch_scatter_resolve(L, v, 1, 0);
}
