/*
 * svo.h -- C ABI of the MI355X (gfx950) SVO raycast + image-warp frame
 * pipeline (libsvo.so).
 *
 * The reference drives its per-frame kernels through an OpenCL host shim
 * (src/ocl.h) from src/raycast.h.  This header is the drop-in replacement for
 * that boundary: plain C, plain pointers and sizes, one context per GPU,
 * status codes instead of error_stop()/exit().
 *
 * Conventions
 *   - every call returns SVO_OK (0) or a negative SVO_ERR_*; the message is
 *     in svo_last_error() (per thread).
 *   - the library owns device memory; the caller owns host memory.
 *   - one context = one device + one HIP stream; calls on a context are not
 *     reentrant; launches are asynchronous unless the entry point says it
 *     blocks (svo_read, svo_sync, svo_hole_count).
 *   - buffer layouts are the reference's (src/raycast.h:79-85,:169-173,
 *     :267-270): screen = 4 planes x N uint32 words, coords = 4 planes x N
 *     float4, ids = [S counts | S offsets | H packed x|(y<<16)], with
 *     N = res_x*res_y, S = (res_x/16)*(res_y/16), plane stride N.
 */
#ifndef SVO_H
#define SVO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SVO_OK 0
#define SVO_ERR_ARG (-1)    /* bad argument / size out of range        */
#define SVO_ERR_HIP (-2)    /* HIP runtime error                       */
#define SVO_ERR_STATE (-3)  /* call order (no octree / no resize)       */
#define SVO_ERR_NOMEM (-4)  /* allocation failed                       */
#define SVO_ERR_IO (-5)     /* file I/O                                */

#define SVO_HOLE 0xffffff00u /* empty screen word (kernel.cl:267)       */

typedef struct svo_ctx svo_ctx;

/* Per-frame arguments of the reference kernels (kernel.cl:479-482,
 * :609-614, :858-863; host side raycast.h:159-162, :322-326).  row* are the
 * vectors raycast_proj receives (rows of the camera matrix), col* the ones the
 * ray kernels receive (after m.transpose()).  Only .xyz are used. */
typedef struct svo_frame_params {
    float m0[4];
    float row0[4], row1[4], row2[4];
    float col0[4], col1[4], col2[4];
    float fovx, fovy;
    int32_t frame;
    int32_t reserved;
} svo_frame_params;

enum svo_buffer {
    SVO_BUF_SCREEN = 0, /* 4 planes x N uint32                     */
    SVO_BUF_COORDS = 1, /* 4 planes x N float4                     */
    SVO_BUF_IDS = 2,    /* 2S + N uint32                           */
    SVO_BUF_IMAGE = 3,  /* N uint32, b + g<<8 + r<<16 (PBO format) */
    SVO_BUF_OCTREE = 4  /* n_words uint32                          */
};

/* ---- errors ---------------------------------------------------------- */
/* CL_CHECK / error_stop (ocl.h:29-41, error.h:2-11): message of the last
 * failing call on this thread. */
const char* svo_last_error(void);

/* ---- context ---------------------------------------------------------- */
/* ocl_init (ocl.h:57-146): pick the device, create the queue. */
int svo_create(svo_ctx** out, int device);
/* ocl_exit (ocl.h:168-174) + raycast_exit (raycast.h:511-516). */
int svo_destroy(svo_ctx* ctx);

/* ocl_malloc(octree, COPY_HOST_PTR) (raycast.h:68, ocl.h:200-213): copy the
 * compact word array (octree.h:232-293 layout) and its root ref
 * (octree_root_normal) to the device.  depth = OCTREE_DEPTH (8..13). */
int svo_upload_octree(svo_ctx* ctx, const uint32_t* words, size_t n_words,
                      uint32_t root_ref, int depth);

/* raycast_init buffer allocation (raycast.h:77-85, :168-172, :267-270) for
 * one resolution; all planes start zeroed. */
int svo_resize(svo_ctx* ctx, int res_x, int res_y);

/* ---- whole frames ------------------------------------------------------ */
/* raycast_draw (raycast.h:147-438): memset, proj planes 1+2, hole compaction,
 * hole rays, 8x4 tile refresh (schedule from frame), copy plane 0 -> 2,
 * fillhole2, colorize.  No device->host readback inside the frame. */
int svo_frame_warp(svo_ctx* ctx, const svo_frame_params* p);
/* raycast_fine_2 over the whole screen with add = 0, then colorize. */
int svo_frame_full(svo_ctx* ctx, const svo_frame_params* p);

/* ---- per-stage entry points (reference kernel names) ------------------- */
/* ocl_memset(mem_screenbuffer, byte_ofs, val, bytes) (ocl.h:299-309) */
int svo_memset(svo_ctx* ctx, size_t byte_ofs, uint32_t val, size_t bytes);
/* raycast_proj with ofs_add = plane*N (kernel.cl:472-592, raycast.h:176-198);
 * plane in 1..3.  Deterministic: the first source (row-major) with the
 * smallest depth key wins a target pixel. */
int svo_proj(svo_ctx* ctx, int plane, const svo_frame_params* p);
/* svo_proj without any allocation, on the caller's device buffers: screen
 * (4 planes x N uint32) and coords (4 planes x N float4, 16-byte aligned) of
 * one res_x x res_y frame on `device`, launched on `stream` (a hipStream_t,
 * NULL = the null stream) and not waited for.  Same results as svo_proj on a
 * context holding those planes.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc); it calls no hipMalloc, hipFree or
 * allocating library routine.  Its warp keys live in a static 16 MiB device
 * array of the library (one per GPU, held until the process ends), so
 * N = res_x * res_y may not exceed 2^21 pixels (SVO_ERR_ARG beyond); calls
 * that overlap on one GPU share that array and race. */
int svo_proj_noalloc(int device, uint32_t* screen, float* coords, int res_x, int res_y, int plane,
                     const svo_frame_params* p, void* stream);
/* raycast_counthole (kernel.cl:234-274) */
int svo_counthole(svo_ctx* ctx);
/* raycast_sumids (kernel.cl:276-296); the total stays on the device in
 * ids[0] -- read it with svo_hole_count() if the host needs it. */
int svo_sumids(svo_ctx* ctx);
/* raycast_writeids (kernel.cl:298-340) */
int svo_writeids(svo_ctx* ctx);
/* raycast_holes (kernel.cl:594-694): one ray per id; the count is read on
 * the device from ids[0] (no idbuf_size argument needed). */
int svo_raycast_holes(svo_ctx* ctx, const svo_frame_params* p);
/* raycast_fine_2 (kernel.cl:846-942): rays for pixels (add_x+i, add_y+j),
 * i < ext_x, j < ext_y, clipped to the screen. */
int svo_raycast_rect(svo_ctx* ctx, const svo_frame_params* p, int add_x,
                     int add_y, int ext_x, int ext_y);
/* raycast_fine (kernel.cl:696-843; launched by raycast.h:236-259 under an
 * if(0)): one ray per 2x2 quad at (2 gx + add_x, 2 gy + add_y) over the
 * roundup16(res/4) grid; the quad's first hole pixel is traced, else pixel
 * ((frame>>2)&1, (frame>>3)&1) of the quad.  The reference host used
 * add_x = (res_x/2)*(frame&1), add_y = (res_y/2)*((frame>>1)&1).  Picks use
 * flat offsets, so quads on the last column/row may read and write the first
 * pixel of the next row or of plane 1, as in the reference; all picks are made
 * before any ray writes.  Uses the learned-fill gate lists as scratch. */
int svo_raycast_quads(svo_ctx* ctx, const svo_frame_params* p, int add_x, int add_y);
/* raycast_fillhole (kernel.cl:342-401; its call in raycast.h:205-220 is
 * disabled): on plane 0, a pixel at a depth step (its nearest neighbour among
 * the 2x2 up-left block is nearer by more than 2.5%) whose motion differs
 * from that neighbour's becomes a hole, to be traced by the next hole pass.
 * motion_x / motion_y: N ints each (the reference's mem_x / mem_y, which no
 * reference kernel writes, so its own result is undefined); NULL = zero
 * motion, under which no pixel changes.  Copied before the call returns. */
int svo_fillhole_punch(svo_ctx* ctx, const int32_t* motion_x, const int32_t* motion_y);
/* ocl_memcpy plane 0 -> dst_plane of screen and coords (raycast.h:394-405) */
int svo_copy_cache(svo_ctx* ctx, int dst_plane);
/* raycast_fillhole2 (kernel.cl:404-470), defined on a pre-filter snapshot of
 * plane 0 (deterministic; the reference filters in place). */
int svo_fillhole(svo_ctx* ctx);
/* raycast_colorize (kernel.cl:944-974) into SVO_BUF_IMAGE */
int svo_colorize(svo_ctx* ctx);

/* ---- transfers ------------------------------------------------------- */
/* ocl_copy_to_host (ocl.h:215-221): blocking read of `bytes` at byte_ofs. */
int svo_read(svo_ctx* ctx, int buf, size_t byte_ofs, size_t bytes, void* dst);
/* Non-blocking copy of the last frame's colorized image (SVO_BUF_IMAGE,
 * N uint32) into dst, on a copy stream of its own: the next call that
 * rewrites the whole image (a frame, a batch frame or svo_colorize) renders
 * into a second image buffer meanwhile, so the copy overlaps it (the
 * reference uploads the image to a GL pixel buffer instead, raycast.h:
 * 425-455); calls that write part of the image or none keep the buffer.
 * dst should be pinned host memory (svo_host_alloc) for the copy to run
 * asynchronously; it holds the image once svo_sync returns.  A buffer is
 * reused two frames later, after its copy has completed.  For a view of a
 * batch, call svo_batch_sync first. */
int svo_read_image_async(svo_ctx* ctx, uint32_t* dst);
/* Pinned (page-locked) host memory for svo_read_image_async destinations,
 * and its release.  No reference counterpart (ocl_copy_to_host, ocl.h:215,
 * reads into ordinary memory, blocking). */
int svo_host_alloc(size_t bytes, void** out);
int svo_host_free(void* p);
/* blocking write (host -> device) into a frame buffer */
int svo_write(svo_ctx* ctx, int buf, size_t byte_ofs, size_t bytes, const void* src);
/* ocl_end_all_kernels (ocl.h:253-265): wait for all queued work. */
int svo_sync(svo_ctx* ctx);
/* blocking read of ids[0] (raycast.h:298) */
int svo_hole_count(svo_ctx* ctx, uint32_t* holes);

/* ---- host-side helpers (no device work) ------------------------------ */
/* Camera vectors of raycast.h:113-162/:322-326 for yaw/pitch/roll (radians)
 * and a world position; pos is wrapped like raycast.h:136-145.
 * pos_wrapped may be NULL. */
int svo_camera(float yaw, float pitch, float roll, const float pos_world[3],
               int depth, int frame, svo_frame_params* out, float pos_wrapped[3]);
/* 8x4 tile schedule (raycast.h:363-366) and the rounded-up extents of the
 * raycast_fine_2 launch (ocl.h:188-198,:234-235). */
int svo_tile_schedule(int res_x, int res_y, int frame, int* add_x, int* add_y,
                      int* ext_x, int* ext_y);

/* set_voxel x n then convert_tree_blocks (octree.h:32-293, raycast.h:13-39):
 * voxels in stream order (xyz = 3 x uint16 each, rgba = uchar4 packed x in the
 * low byte); emits the reference word array.  *words is malloc'ed; free it
 * with svo_free. */
int svo_build_octree(const uint16_t* xyz, const uint32_t* rgba, size_t n,
                     int depth, uint32_t** words, size_t* n_words,
                     uint32_t* root_ref);
void svo_free(void* p);
/* svo_build_octree on a GPU (device index): same words and root ref, built
 * with a Morton sort and per-level scans instead of the host recursion.  No
 * counterpart in the reference (its builder is the host code above). */
int svo_build_octree_gpu(int device, const uint16_t* xyz, const uint32_t* rgba, size_t n,
                         int depth, uint32_t** words, size_t* n_words, uint32_t* root_ref);

/* Save / load a built octree (word array + root ref + depth) as a small
 * binary file, so large scenes are built once. */
int svo_octree_save(const char* path, const uint32_t* words, size_t n_words,
                    uint32_t root_ref, int depth);
int svo_octree_load(const char* path, uint32_t** words, size_t* n_words,
                    uint32_t* root_ref, int* depth);

/* RLE4 volume files (RLE4::load, src/octree/Rle4.cpp:8-165): map 0 of the
 * file as a voxel stream for svo_build_octree, in the reference's set_voxel
 * order (slice, x, rows upward), voxel y = sy-1-row, colours by the rule of
 * Rle4.cpp:127-160 (palette != 0: palette bits + intensity; 0: inverted
 * luminance).  Voxels are offset by add_*, and every one must lie inside the
 * depth-`depth` cube.  Truncated or inconsistent files fail with SVO_ERR_IO
 * (the reference reads past them).  *xyz / *rgba are malloc'ed (svo_free). */
int svo_rle4_load(const char* path, int palette, int add_x, int add_y, int add_z, int depth,
                  uint16_t** xyz, uint32_t** rgba, size_t* n);
/* Write a one-map RLE4 file from a dense volume: vol[(z*sx + x)*sy + row]
 * is the 16-bit colour word of file row `row` of column (x, z), or < 0 for
 * empty. */
int svo_rle4_save(const char* path, int sx, int sy, int sz, const int32_t* vol);

/* Write a colorized image (SVO_BUF_IMAGE words, b + g<<8 + r<<16, the PBO
 * layout of raycast.h:425-455) as a binary PPM, rows top to bottom. */
int svo_image_save_ppm(const char* path, const uint32_t* image, int w, int h);
/* The same image as an 8-bit RGB PNG (uncompressed deflate blocks).  Both
 * writers replace the reference's display path (the PBO upload,
 * raycast.h:425-472, shown through GL), which has no file output. */
int svo_image_save_png(const char* path, const uint32_t* image, int w, int h);

/* Seeded synthetic scenes (voxel streams for svo_build_octree).
 *   kind 1: value-noise heightfield shell (2 octaves)
 *   kind 2: 4-octave fractal terrain shell + n_spheres sphere shells
 *   kind 3: 3-D fractal-noise cave in a slab of 1/8 of the world's height,
 *           boundary voxels only
 *   kind 4: the same cave over the whole height of the world
 *   kind 5: kind 2 without the per-voxel colour jitter (colour by height)
 * *xyz / *rgba are malloc'ed (svo_free). */
int svo_scene_generate(int kind, uint64_t seed, int depth, int n_spheres,
                       uint16_t** xyz, uint32_t** rgba, size_t* n);
/* svo_scene_generate for kinds 3 and 4 (the caves) on a GPU: the same voxel
 * stream, bit for bit, in seconds at depth 12 (the host needs hours there). */
int svo_scene_generate_gpu(int device, int kind, uint64_t seed, int depth,
                           uint16_t** xyz, uint32_t** rgba, size_t* n);

/* Seeded Catmull-Rom flythrough: 6 floats per frame
 * (yaw, pitch, roll, pos_x, pos_y, pos_z), world units (1 voxel = 1/8).
 * max_voxels_per_frame / max_yaw_deg_per_frame bound the motion. */
int svo_camera_path(uint64_t seed, int depth, int n_frames,
                    float max_voxels_per_frame, float max_yaw_deg_per_frame,
                    float height_voxels, float pitch, float* out);

/* ---- measurement -------------------------------------------------------- */
/* GPU timing with HIP events on the context's stream (off by default).
 *   mode 0  off
 *   mode 1  per stage: frames run as direct launches with an event before
 *           each stage; after svo_sync, svo_stage_times fills up to n floats
 *           (ms) in stage order and returns the number written (names via
 *           svo_stage_name)
 *   mode 2  per frame: an event pair around each frame (graph launches
 *           included; a warp frame with a pretrace starts at its pretrace)
 *   mode 3  frame completions: an event at the end of each of the first 512
 *           frames after the mode is set; svo_frame_intervals returns the
 *           times between successive completions (frames issued back to
 *           back: the frame period the GPU sustains)
 * In modes 1 and 2, svo_frame_time gives the last frame's GPU time (ms)
 * once svo_sync has returned.  Stage timing (mode 1) runs warp frames
 * without a pretrace. */
int svo_enable_timing(svo_ctx* ctx, int mode);
int svo_frame_time(svo_ctx* ctx, float* ms);
/* Timing mode 3: up to max_n intervals (ms) between the completions of
 * successive frames since the mode was set; *n receives the count
 * (blocking). */
int svo_frame_intervals(svo_ctx* ctx, float* ms, int max_n, int* n);
/* svo_frame_warp runs frames >= 2 that have no pretrace (svo_set_pretrace)
 * as one captured HIP graph per context (launch sequence fixed, the frame's
 * camera and tile in a device-side view record written before each launch;
 * a few instances of the graph launched in turn).  On by default; 0 makes
 * every frame use direct launches.  The graph is re-captured after
 * svo_resize, svo_upload_octree and svo_set_tuning.  Results are identical
 * either way. */
int svo_set_graph(svo_ctx* ctx, int enable);
/* Rays traced by the ray kernels since the last reset (blocking). */
int svo_ray_count(svo_ctx* ctx, uint64_t* rays, int reset);
/* The same count and the loop trips of the longest of those rays (one trip
 * is one pass of the reference's traversal loop, kernel.cl:152-211), both
 * since the last reset of either (blocking).  No counterpart in the
 * reference. */
int svo_ray_stats(svo_ctx* ctx, uint64_t* rays, uint32_t* max_trips, int reset);
int svo_stage_times(svo_ctx* ctx, float* ms, int n);
const char* svo_stage_name(int stage);

/* Ray-kernel tuning (the autotuner's knobs); a negative value keeps the
 * current setting.
 *   refill_threshold  idle lanes (1..64) that trigger a work-queue refill
 *   wg_per_cu         persistent workgroups per compute unit (1..32)
 *   iters_per_refill  traversal steps between refill checks (1..4096)
 *   lds_cache_words   octree words (top levels, breadth-first) kept in LDS
 *                     per workgroup (0..12288; 0 disables the cache) */
int svo_set_tuning(svo_ctx* ctx, int refill_threshold, int wg_per_cu,
                   int iters_per_refill, int lds_cache_words);
/* Ray hand-out order of the ray kernel.  The 64 lanes of a wave take 64/group
 * groups of `group` consecutive rays that lie 1/64 of the ray list apart, so
 * the long rays of one screen region do not share a wave; group 64 is list
 * order.  only_if_fits != 0 spreads only launches whose rays all get a lane
 * at once (warp frames) and keeps list order for larger ones (full frames).
 * Default: group 1, only_if_fits 1.  Results do not depend on it. */
int svo_set_ray_spread(svo_ctx* ctx, int group, int only_if_fits);
/* Lanes per wave that take rays (power of two, 1..64; default 64).  Fewer
 * lanes spread the same rays over more waves, so fewer long rays share one
 * (a wave runs both branch bodies while its rays disagree).  The spread order
 * then places the rays of a wave 1/lanes of the list apart.  Results do not
 * depend on it. */
int svo_set_ray_lanes(svo_ctx* ctx, int lanes);
/* Shape of the ray blocks in which rectangle rays (refresh tile, full
 * frame) are handed out, one block per wave: width 8 (8x8, default), 16
 * (16x4), 32 (32x2) or 64 (64x1) pixels.  The reference's raycast_fine_2
 * runs 16x16 work-groups (raycast.h:366); here the shape only orders the
 * hand-out, so results do not depend on it. */
int svo_set_ray_block(svo_ctx* ctx, int width);
/* Full frames (svo_frame_full): the 8x8 screen blocks in which the previous
 * full frame had a ray of at least long_trips traversal trips are raycast
 * first (as a pixel list ahead of the raster-order blocks), so the long
 * horizon rays overlap the rest of the frame; 0 = raster order only.  The
 * first full frame after a resize, an upload or a change of this setting
 * only collects the flags.  The reference launches raycast_fine_2 as a fixed
 * 16x16-workgroup grid over its rectangle (raycast.h:361-366), so no
 * counterpart.  Default 1024; results do not depend on it. */
int svo_set_long_first(svo_ctx* ctx, int long_trips);
/* Warp and learned frames (svo_frame_warp, svo_frame_warp_learned; frames
 * >= 2, timing mode 0, 2 or 3): each frame's
 * refresh rectangle and every pixel of the 8x8 blocks whose last traced rays
 * include one of at least long_trips loop trips are traced ahead, into one
 * of `sets` result sets; each set has a stream of its own and is reused every
 * `sets` frames, once the frame that read it has completed.  The frame's own
 * ray launch waits for its pretrace and copies the stored result of each of
 * its rays that was traced ahead, so the long rays of a frame overlap the
 * frames before it.  The pretrace hands out its rays longest first (by the
 * trips of the last ray traced at each pixel); with split_trips > 0 the
 * pixels whose last ray had at least that many trips go to up to half of
 * its waves, split_lanes lanes per wave.  Warp frames with a pretrace run as
 * direct launches (svo_set_graph applies to the others).  With only_queued
 * a frame is pretraced only while the warp frame before it is still on the
 * GPU (the caller issues frames ahead): a frame the GPU starts at once has
 * no earlier frame to overlap, and its own rays finish sooner without one.
 * Defaults: 3 sets, long_trips 1, 1 workgroup per CU, no split,
 * only_queued 1, predict 2, whole_trips 0, queue_skip 3.  A primary ray depends
 * only on the camera, the pixel and the octree (kernel.cl:594-694,
 * :846-942): results do not depend on these settings.  HIP maps streams onto
 * at most GPU_MAX_HW_QUEUES hardware queues (4 unless the environment sets
 * it before the first HIP call); pretraces whose streams share one run one
 * after the other.
 * predict = k (1..3) replaces the long blocks by predicted holes: every warp
 * or learned frame also stores its plane-0 positions after its projection
 * (before its rays; holes marked) in a ring of 4 snapshots, and frame g's
 * pretrace, once frame g-k's projection has run, projects frame g-k's
 * snapshot through the cameras of frames g-k+1 .. g-1 (keeping the nearest
 * sample per pixel, as the warp does) and then g, and traces the 2x2 quads
 * no sample reaches -- the hole quads frame g's warp can leave -- in the 8x8
 * blocks whose last traced rays include one of at least long_trips trips,
 * with the refresh rectangle; for a learned frame (svo_frame_warp_learned)
 * also the gate's rays as the prediction sees them: the residual holes of
 * tiles with more than max_cnn_holes of them, and with the depth rule
 * (svo_set_gate_depth) the pixels where a sample beyond the limit lands and
 * the residual holes of their tiles.
 * A frame with a pretrace traces its rays on
 * none of the pretrace's lists as soon as the lists are made, and copies
 * the pretraced results once the pretrace has completed.  A
 * frame whose frames g-k .. g-1 did not all pass through svo_frame_warp or
 * svo_frame_warp_learned (consecutive frame counters) falls back to the long
 * blocks.  With predict > 0 and whole_trips > 0, the 8x8 blocks whose last
 * traced rays include one of at least whole_trips trips are traced whole in
 * a second launch of the set that starts as soon as the frame is issued
 * (their rays are the longest; a prediction would start them too late), and
 * the predicted holes skip them.  The sets run on a pool of streams, taken
 * in index order: the context makes the pool for its default set count
 * right after its own two streams in svo_create, and more streams when the
 * set count grows;
 * with queue_skip = k > 0 the pool streams at index k-1, k+3, k+7, ... are
 * left out.  HIP gives each new stream the next of GPU_MAX_HW_QUEUES
 * hardware queues, and on MI355X one hardware queue in four runs the
 * pretrace far slower than the others (the flythrough at 1,860 against
 * 2,500 frames/s for the same work, with 3 sets on the other queues): the
 * autotuner searches queue_skip.  The slow queues have been those four,
 * eight, ... after the main stream's (the main stream's pipe, with queues
 * dealt to four pipes in turn); with the pool right behind the context's
 * two streams that is pool index 2, 6, 10, which the default, 3, leaves
 * out.  The reference traces each
 * frame's rays in
 * the frame (raycast.h:296-387), so no counterpart. */
typedef struct {
    int sets;         /* 0 = off, 2..8 */
    int long_trips;   /* >= 1 */
    int wg_per_cu;    /* persistent workgroups per CU, 1..32; 0 = svo_set_tuning's */
    int split_trips;  /* 0 = no split */
    int split_lanes;  /* power of two, 1..64 */
    int only_queued;  /* 1: pretrace a frame only while the warp frame before it is still running */
    int predict;      /* 0: long blocks whole; 1..3: holes predicted from the frame that many frames back */
    int whole_trips;  /* predict > 0: blocks traced whole at once (0 = none) */
    int queue_skip;   /* 0: the pool's streams in order; k = 1..4: leave out those at index k-1 mod 4 */
} svo_pretrace_config;
int svo_set_pretrace(svo_ctx* ctx, const svo_pretrace_config* cfg);
int svo_get_pretrace(svo_ctx* ctx, svo_pretrace_config* cfg);
/* Since the last reset (blocking): rays traced ahead, the loop trips of the
 * longest of them, and the frame rays that took a result traced ahead
 * (svo_ray_stats counts only the frame rays traced in the frame). */
int svo_pretrace_stats(svo_ctx* ctx, uint64_t* rays, uint64_t* merged, uint32_t* max_trips, int reset);
/* Learned frames (svo_frame_warp_learned): the number of workgroups the
 * fused CNN tile kernel runs with while it shares the GPU with the ray
 * launch (0 = two per CU, its grid when it runs alone).  Fewer workgroups
 * take longer but leave the ray launch's SIMDs freer; the default is half
 * the CUs.  Batched learned frames and timing mode 1 use the full grid.
 * No counterpart in the reference (it has no CNN); results do not depend
 * on it. */
int svo_set_cnn_overlap(svo_ctx* ctx, int workgroups);

/* ---- learned hole filler (no counterpart in the reference) ------------
 * A 4-layer 3x3 CNN (5 -> 32 -> 32 -> 32 -> 3 channels, ReLU) run with BF16
 * MFMA kernels accumulating in FP32, trained with Adam on FP32 master
 * weights.  Parameters are a flat float vector: per layer the weights
 * [9 taps][cout][cin] (channels padded to 16/32/32/16 in, 32/32/32/16 out;
 * padded entries stay zero) followed by the bias [cout]. */
typedef struct svo_cnn svo_cnn;
int svo_cnn_create(svo_cnn** out, int device, uint64_t seed);
int svo_cnn_destroy(svo_cnn* cnn);
size_t svo_cnn_param_count(void);
int svo_cnn_get_params(svo_cnn* cnn, float* dst, size_t count);
int svo_cnn_set_params(svo_cnn* cnn, const float* src, size_t count);
/* x: [batch][h][w][16] bf16 bit patterns (h, w multiples of 8), zero
 * "same" padding at crop borders; out: [batch][h][w][16] float (channels
 * 0..2 = RGB in 0..1). */
int svo_cnn_forward(svo_cnn* cnn, const uint16_t* x, int batch, int h, int w, float* out);
/* loss = sum mask*(out-target)^2 / (3*sum mask) over RGB; writes the
 * gradient of every parameter (same layout as the parameter vector). */
int svo_cnn_gradients(svo_cnn* cnn, const uint16_t* x, const float* target, const float* mask,
                      int batch, int h, int w, float* grads, size_t count, float* loss);

/* One Adam step (learning rate lr) on n_crops random crop x crop windows
 * (origins drawn from seed) of the frame pair the two contexts currently
 * hold: features from the CNN input of `warp`'s last svo_frame_warp_learned
 * (plane 0 as the gate left it), targets from `full`'s image, loss on the
 * pixels still holes after that frame's rays.  Both contexts must have
 * rendered the same camera at one resolution; SVO_ERR_STATE if `warp` has
 * rendered no learned frame since its last resize.  loss_out (optional)
 * makes the call block and returns the step's loss. */
int svo_cnn_train_step(svo_cnn* cnn, svo_ctx* warp, svo_ctx* full, const svo_frame_params* p,
                       int n_crops, int crop, uint64_t seed, float lr, float* loss_out);
/* The frame fill's inference kernel: one workgroup runs all four layers on
 * a 24x24 crop with activations in LDS and returns the centre 16x16 RGB.
 * x: [count][24][24][16] bf16 features; out: [count][16][16][3] float.
 * svo_cnn_set_fused(cnn, mode): 1 (default) the fused kernel with 4 waves
 * per tile staging each layer's weights (two workgroups per CU); 2 the fused
 * kernel with 8 waves and all weights resident in LDS (one workgroup per
 * CU, about 10% slower on MI355X); 0 the per-layer kernels of
 * svo_cnn_forward (same network, bf16 rounding differences). */
int svo_cnn_forward_tiles(svo_cnn* cnn, const uint16_t* x, int count, float* out);
int svo_cnn_set_fused(svo_cnn* cnn, int mode);
/* svo_frame_warp with the learned fill and per-tile confidence gate.  After
 * hole compaction each 16x16 tile counts its residual holes: hole pixels that
 * neither the hole list nor the refresh tile will trace.  A tile with more
 * than max_cnn_holes of them has them raycast in the same ray launch as the
 * hole and tile rays; a tile with fewer is filled by the CNN (image only;
 * screen words keep fillhole2's result).  The CNN's input is plane 0 as the
 * gate leaves it (copied before any ray), so it runs on a second stream
 * alongside the ray launch; its colours replace the fill's at the tile
 * pixels still holes after the rays.  Timing mode 1 runs it after the fill
 * instead (same results).  cnn may be NULL: the gate alone, fillhole2
 * colours everywhere. */
int svo_frame_warp_learned(svo_ctx* ctx, svo_cnn* cnn, const svo_frame_params* p, int max_cnn_holes);
/* The gate's confidence rule for reprojected pixels (learned frames; no
 * counterpart in the reference): a reprojected pixel outside the refresh
 * rectangle whose camera depth exceeds max_depth world units (where a pixel
 * spans several voxels, so the warped sample's colour is rarely the one the
 * pixel's ray meets) is raycast in the frame's ray launch, and so are the
 * residual holes of its 16x16 tile (the CNN gets no such tile).  0
 * (default) = off.  Trades frames/s for PSNR against full raycasting. */
int svo_set_gate_depth(svo_ctx* ctx, float max_depth);
/* Gate outcome of the last svo_frame_warp_learned (blocking): tiles sent to
 * the CNN and pixels sent back to raycasting (residual holes and, with the
 * depth rule, the reprojected pixels of whole tiles). */
int svo_gate_counts(svo_ctx* ctx, uint32_t* cnn_tiles, uint32_t* gate_rays);

/* ---- batched views (no counterpart in the reference) -------------------
 * A batch runs the warp frame of many contexts that hold the same octree on
 * one device: every warp stage is one launch over all views and the hole and
 * tile rays of all views share one persistent ray launch, so the long-ray
 * tails of different views overlap.  Each view's outputs equal those of
 * svo_frame_warp on that context.  The views keep their own resolution and
 * frame counter; the contexts must not be resized, given a new octree or
 * destroyed while the batch exists.  The ray kernel uses the first view's tuning and counts
 * rays on the first view. */
typedef struct svo_batch svo_batch;
int svo_batch_create(svo_batch** out, svo_ctx* const* views, int n_views);
int svo_batch_destroy(svo_batch* batch);
/* params[i] drives view i; ordered after each view's queued work, and each
 * view's later work is ordered after the batch. */
int svo_batch_frame_warp(svo_batch* batch, const svo_frame_params* params);
int svo_batch_sync(svo_batch* batch);
/* svo_frame_warp_learned for every view (same cnn and max_cnn_holes for all):
 * the gate of all views in one launch, their gate pixels in the shared ray
 * launch, the CNN of each view on the batch's second stream alongside it.
 * Per view identical to svo_frame_warp_learned on that context. */
int svo_batch_frame_warp_learned(svo_batch* batch, svo_cnn* cnn, const svo_frame_params* params,
                                 int max_cnn_holes);

#ifdef __cplusplus
}
#endif
#endif /* SVO_H */
