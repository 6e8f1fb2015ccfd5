// svo_ctx.h -- the per-device context shared by the ray/warp runtime
// (svo_device.hip) and the learned hole filler (svo_cnn.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "svo_internal.h"

enum Stage {
    ST_CLEAR, ST_SCATTER, ST_RESOLVE, ST_COUNT, ST_SCAN, ST_WRITE, ST_GATE, ST_RAYS, ST_COPY, ST_FILTER, ST_CNN,
    ST_CNN_APPLY, ST_N
};
static const char* const kStageNames[ST_N] = {"clear",    "proj_scatter", "proj_resolve",  "counthole",
                                              "sumids",   "writeids",     "gate",          "rays",
                                              "copy",     "fill_colorize", "cnn",          "cnn_apply"};

constexpr int kPreMaxSets = 8;  // pretrace result sets at most
constexpr int kPrePool = 12;    // pretrace part-1 streams at most (8 sets, every fourth left out)
constexpr int kViewSlots = 8;   // pinned staging slots of the graph frame's view record
// instances of the warp frame's graph, launched in turn: a relaunch of one
// instance waits on the host for its previous launch, so successive frames
// use different instances and the host can run frames ahead
constexpr int kGraphExecs = 4;
constexpr int kEndRing = 512;  // timing mode 3: completion events of the first frames
constexpr int kSnapSlots = 4;  // predicted-hole pretrace: final plane-0 snapshots of recent frames

struct svo_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int n_cu = 256;
    // octree
    uint32_t* d_oct = nullptr;  // validated copy, normal records in breadth-first order
    size_t n_words = 0;
    uint32_t normal_words = 0;  // size of the renumbered normal region
    uint32_t root = 0;
    int depth = 0;
    // frame buffers
    int res_x = 0, res_y = 0;
    size_t N = 0, S = 0;
    uint32_t* d_screen = nullptr;
    float4* d_coords = nullptr;
    uint32_t* d_ids = nullptr;
    uint32_t* d_image = nullptr;  // the frame's image: image_buf[0] or image_buf[1]
    // svo_read_image_async: a second image buffer, so the next frame renders
    // into one while the other is copied to the host on copy_stream
    uint32_t* image_buf[2] = {nullptr, nullptr};
    hipStream_t copy_stream = nullptr;
    hipEvent_t image_ready = nullptr;
    hipEvent_t image_copied[2] = {nullptr, nullptr};
    bool image_copy_pending[2] = {false, false};
    bool image_flip = false;  // the next frame switches buffers
    unsigned long long* d_keys = nullptr;
    uint32_t* d_snap = nullptr;
    int32_t* d_motion = nullptr;  // svo_fillhole_punch: motion x (N) then y (N)
    // [0..1] ray-queue counters, [2..3] the graph's, [4..5] the pretrace's;
    // frame ray launches: [8..9] rays (u64), [10] longest ray (trips), [12..13]
    // rays merged from a pretrace (u64); pretrace launches: [16..17] rays,
    // [18] longest ray
    uint32_t* d_queue = nullptr;
    // learned-fill gate: [0] CNN tile count, [16] gate ray count
    uint32_t* d_gate = nullptr;
    uint32_t* d_gate_tiles = nullptr;  // tile indices sent to the CNN
    uint32_t* d_gate_rays = nullptr;   // x | y << 16 of holes sent back to raycasting
    int n_tiles_x = 0, n_tiles_y = 0;  // 16x16 gate tiles (partial ones included)
    // learned frames: the CNN's input is plane 0 as the gate leaves it (copied
    // before any ray), so the CNN runs on cnn_stream alongside the ray launch;
    // its colours wait in d_cnn_out until the fill has run
    uint32_t* d_cnn_words = nullptr;   // N screen words
    float4* d_cnn_coords = nullptr;    // N coords
    uint32_t* d_cnn_out = nullptr;     // 256 colorized words per gate tile
    bool cnn_input = false;            // d_cnn_* hold the last learned frame's input
    hipStream_t cnn_stream = nullptr;
    hipEvent_t cnn_fork = nullptr, cnn_join = nullptr;
    int launch_seq = 0;
    bool plane1_holes = false;
    // tuning
    int refill = 16, wg_per_cu = 4, iters = 64;
    uint32_t cache_req = 0;  // requested LDS node-cache words
    int spread_shift = 0;    // ray hand-out: log2 of the spread group (6 = list order)
    int spread_fit = 1;      // spread only launches whose rays all fit in the resident lanes
    int lanes_shift = 6;     // log2 of the lanes per wave that take rays
    int blk_shift = 3;       // rectangle rays in blocks of 2^blk_shift x 2^(6-blk_shift) (8x8)
    // full frames: 8x8 blocks whose previous full frame had a ray of at least
    // long_trips trips are raycast first (0 = raster order only)
    int long_trips = 1024;
    // learned frames: workgroups of the CNN tile kernel while it runs
    // alongside the ray launch (0 = two per CU); set to half the CUs at creation
    int cnn_overlap_wgs = 0;
    // learned frames: tiles with a reprojected pixel deeper than this (world
    // units) are raycast whole (0 = off)
    float gate_depth = 0.0f;
    uint32_t* d_block_long = nullptr;  // two flag arrays (previous / current frame)
    uint32_t* d_long_list = nullptr;   // pixel words (x | y << 16) of the long blocks
    uint32_t* d_long_count = nullptr;
    int long_parity = 0;
    bool long_valid = false;
    // pretrace of warp frames (svo_set_pretrace): the refresh rectangle and
    // the 8x8 blocks whose last rays ran at least pre_trips trips are traced
    // into one of pre_sets result sets (each on a stream of its own) while
    // earlier frames still run; the frame's own ray launch copies those of
    // its rays
    int pre_sets = 3;              // 0 = off
    uint32_t pre_trips = 1;
    int pre_wg = 1;                // workgroups per CU of the pretrace launch (0: wg_per_cu)
    int pre_split_trips = 0;       // blocks with a known ray this long go to the split waves (0: no split)
    int pre_split_shift = 3;       // log2 of the lanes per split wave
    bool pre_only_queued = true;   // pretrace a frame only while the one before it is still running
    hipEvent_t ev_frame_done = nullptr;  // the last warp frame's end
    bool frame_done_recorded = false;
    // set s runs part 1 (the long blocks, or the predicted holes) on a
    // stream of pre_pool and, with predicted holes and pre_whole_trips > 0,
    // part 0 (the longest blocks whole, issued without waiting for a
    // snapshot) on pre_streams[2s]; both write the set's results.  The pool
    // streams are made in index order, the default count right after the
    // context's own two streams (HIP gives each new stream the next hardware
    // queue); the sets take them in order, leaving out every index i with
    // i % 4 == pre_qskip - 1 when pre_qskip > 0
    hipStream_t pre_streams[2 * kPreMaxSets] = {};  // part 0, created on first use
    hipStream_t pre_pool[kPrePool] = {};
    int pre_qskip = 3;  // pool index 2 is four queues after the main stream's: its pipe
    uint32_t* d_known = nullptr;      // N: loop trips of the last ray traced per pixel
    uint32_t* d_pre_stamp = nullptr;  // pre_sets x N
    uint32_t* d_pre_word = nullptr;   // pre_sets x N
    float4* d_pre_xyz = nullptr;      // pre_sets x N
    uint32_t* d_pre_list = nullptr;   // per set and part: list header and pixel words (x | y << 16)
    uint32_t* d_pre_blk = nullptr;    // per set and part: one word per 8x8 block
    uint32_t* d_pre_queue = nullptr;  // per set and part: four ray-queue counters
    uint8_t* d_pre_bkt = nullptr;     // per set and part: one byte per pixel
    int pre_alloc_sets = 0;           // sets the buffers hold (0: not set up)
    hipEvent_t ev_pre[2 * kPreMaxSets] = {};  // a set's pretrace part is complete
    hipEvent_t ev_list[2 * kPreMaxSets] = {};  // ... its pixel list (bucket bytes) is final
    hipEvent_t ev_main[kPreMaxSets] = {};  // a set's frame is complete
    bool main_recorded[kPreMaxSets] = {};
    uint32_t pre_frames = 0;  // frames pretraced since the sets were cleared; the stamp is this count
    int pre_seq[2 * kPreMaxSets] = {};  // ray-queue slot parity per set and part
    bool pre_active = false;  // the frame being issued has a pretrace in set pre_set, stamp pre_id
    bool pre_early = false;   // ... with a part 0
    // predicted-hole pretrace (pre_predict = k > 0): each warp/learned frame's
    // projection (resolve) stores its plane-0 positions (holes as NaN) in
    // snapshot slot frame % kSnapSlots; frame g's pretrace projects snapshot
    // g-k through the cameras of frames g-k+1 .. g and traces the hole quads
    // no sample reaches (and the refresh rectangle)
    int pre_predict = 2;
    uint32_t pre_whole_trips = 0;           // part 0's block threshold (0: no part 0)
    float4* d_snap_xyz = nullptr;           // kSnapSlots x N
    int snap_frame[kSnapSlots] = {-1, -1, -1, -1};  // frame counter each slot holds (-1: none)
    hipEvent_t ev_snap[kSnapSlots] = {};    // the slot's frame has written it
    svo_frame_params cam_hist[kSnapSlots] = {};  // cameras of recent frames, by frame counter
    int cam_frame[kSnapSlots] = {-1, -1, -1, -1};
    int snap_pending = -1;                  // slot written by the frame being issued (-1: none)
    unsigned long long* d_pred_keys = nullptr;  // per set: N nearest-sample keys (all ones between uses)
    float4* d_pred_xyz = nullptr;               // per set: N positions of an intermediate camera
    uint8_t* d_pred_cov = nullptr;              // per set: N coverage bytes (zero between uses)
    uint8_t* d_pred_deep = nullptr;             // per set: N bytes, a sample beyond the gate's depth limit (zero between uses)
    int pred_alloc_sets = 0;
    int pre_set = 0;
    uint32_t pre_id = 0;
    // per-frame graph of the warp frame (frames >= 2): its kernels read the
    // frame's view from d_view, refreshed from a pinned staging ring
    bool use_graph = true;
    hipGraphExec_t graph[kGraphExecs] = {};
    void* d_view = nullptr;  // one view descriptor (device)
    void* h_view = nullptr;  // kViewSlots view descriptors (pinned host)
    hipEvent_t view_done[kViewSlots] = {};
    uint32_t* d_first = nullptr;  // [0] = 0, [1] = rays of the frame
    int graph_frames = 0;
    // timing: 0 off, 1 per stage (direct launches), 2 per frame, 3 frame
    // completion times (end_ring)
    int timing = 0;
    hipEvent_t end_ring[kEndRing] = {};
    int end_count = 0;
    hipEvent_t ev[ST_N + 1] = {};
    float stage_ms[ST_N] = {};
    bool stage_used[ST_N] = {};
    bool frame_timed = false;
    float frame_ms = 0.0f;
};


// learned-fill hooks implemented in svo_cnn.hip: the CNN on the gate's tiles
// (input d_cnn_words/d_cnn_coords, output d_cnn_out), then its colours into
// the image at the tiles' pixels that are still holes in plane 2
// (grid: workgroups of the fused tile kernel, 0 = two per CU)
int cnn_fill_tiles(svo_cnn* n, svo_ctx* c, const svo_frame_params* p, const uint32_t* tile_list,
                   const uint32_t* tile_count, int max_tiles, hipStream_t stream, int grid = 0);
int cnn_apply_tiles(svo_ctx* c, const uint32_t* tile_list, const uint32_t* tile_count, int max_tiles,
                    hipStream_t stream);
// a batch's CNN: one tile launch over every view through a table of
// cnn_tile_view_bytes() per view (host staging h_tab, device copy d_tab;
// first: n_views + 1 words), then one apply launch over the same table
size_t cnn_tile_view_bytes();
int cnn_fill_tiles_batch(svo_cnn* n, svo_ctx* const* views, const svo_frame_params* params, int n_views,
                         void* h_tab, void* d_tab, uint32_t* first, hipStream_t stream);
int cnn_apply_tiles_batch(const void* d_tab, int n_views, int max_tiles, hipStream_t stream);
