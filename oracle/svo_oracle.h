/*
 * svo_oracle.h -- CPU restatement of the reference SVO raycast + image-warp
 * frame pipeline.  TEST INFRASTRUCTURE ONLY: the product library never links
 * or calls this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, as the checker / CPU baseline.
 *
 * Reference files restated (paths relative to the reference checkout):
 *   kernel/kernel.cl            device kernels (traversal, warp, compaction,
 *                               filter, colorize)
 *   src/raycast.h:93-437        per-frame launch sequence
 *   src/octree/octree.h:32-293  pointer octree + compact word format
 *   ext/mathlib/_matrix44.h     camera matrix (rotate_x/y/z, transpose, m*v)
 *
 * Arithmetic semantics (the reference GPU build used -cl-fast-relaxed-math,
 * so its own results were implementation-defined; these are the defined
 * semantics):  IEEE binary32, round-to-nearest, no FMA contraction,
 * correctly rounded '/' and sqrt, length(v) = sqrtf((x*x + y*y) + z*z),
 * dot in the same order, float->int truncation toward zero.
 *
 * Parity pins (what checks each function; see tests/):
 *   orc_build_octree    word-for-word vs the reference octree.h compiled on
 *                       the host (oracle/_ref) + hand-derived 1/8-voxel arrays
 *                       (like the reference, its 24-bit node handles wrap
 *                       above 2^24 pointer nodes, octree.h:74-76)
 *   orc_trace_pixel     dense-grid FP64 first-hit DDA on small scenes
 *                       (hit voxel + colour byte); LOD hits against the
 *                       octree's node colours (creator voxel, first leaf);
 *                       pop levels against the geometric rule (one above
 *                       the highest changed position bit) via the step log;
 *                       the whole step/pop sequence, the LOD stop level,
 *                       the exit kind, exit node and colour against an
 *                       independent FP64 model of the traversal, with
 *                       planted LOD and exit-node mistakes failing it
 *                       (orc_set_mutation)
 *   orc_frame_warp_gated the gate's depth rule against full raycasts
 *   orc_proj            zero-motion invariant + literal sequential semantics
 *   orc_counthole/...   closed forms + brute force
 *   orc_fillhole2_*     closed-form per-pixel rules
 *   orc_fillhole_punch  closed-form per-pixel rules
 *   orc_colorize        exhaustive over the low byte
 */
#ifndef SVO_ORACLE_H
#define SVO_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_HOLE 0xffffff00u

typedef struct { float x, y, z, w; } orc_f4;

/* Per-frame camera / kernel arguments (raycast.h:159-162, :322-326). */
typedef struct {
    orc_f4 m0;               /* camera position, world units (w = 1) */
    orc_f4 row0, row1, row2; /* rows of m  -> raycast_proj a_mx/a_my/a_mz */
    orc_f4 col0, col1, col2; /* columns of m -> ray kernels a_mx/a_my/a_mz */
    float fovx, fovy;
    int32_t frame;
    int32_t pad;
} orc_params;

/* ---- camera (raycast.h:113-162, _matrix44.h:334-344,529-580,863-888) ---- */
void orc_camera(float yaw, float pitch, float roll, const float pos_world[3],
                int depth, int frame, orc_params* out, float pos_wrapped[3]);

/* ---- octree builder (octree.h:32-89 set_voxel, :95-293 conversion) ---- */
/* xyz: 3 x uint16 per voxel, rgba: uchar4 as little-endian u32 (x = low byte).
 * Returns 0 on success; *words_out is malloc'ed (free with orc_free). */
int orc_build_octree(const uint16_t* xyz, const uint32_t* rgba, size_t n,
                     int depth, uint32_t** words_out, size_t* n_words,
                     uint32_t* root_ref);
void orc_free(void* p);

/* ---- ray traversal (kernel.cl:23-214, :594-694, :846-942) ---- */
typedef struct {
    uint64_t rays, iters, descends, steps, pops, fetch_words;
    uint64_t desc_by_rek[16];   /* descends by level (rekursion) */
    /* the last traced ray: its level (rekursion) and lod at loop exit,
     * whether it exited by a hit (node_test set), its integer position in
     * the octant-folded frame and the fold (sign_xyz); not accumulated */
    int64_t last_rek, last_lod, last_hit;
    int64_t last_ix, last_iy, last_iz, last_sign;
    uint64_t max_iters;         /* loop trips of the longest ray (accumulated as a max) */
} orc_stats;

/* One primary ray for pixel (idx,idy); returns the screen word
 * (0xff000000 + colour) and writes xyz = raypos/16. */
uint32_t orc_trace_pixel(const uint32_t* oct, size_t n_words, uint32_t root,
                         int depth, int res_x, int res_y, const orc_params* p,
                         int idx, int idy, float xyz[3], orc_stats* st);
/* The same ray, logging every DDA step (kernel.cl:181-208): per step 9 ints,
 * the integer position before and after the step (folded frame), the level
 * of the step, the XOR level + 1 (floor(log2((dx^dy^dz) & (2^D-1))) + 1, 0
 * when no bit changed; the step pops to it when it is >= the level), and 1
 * if the step ends the ray by distance (no pop then).  At most max_pops
 * steps are logged.  st (optional) receives the ray's exit state. */
uint32_t orc_trace_pixel_pops(const uint32_t* oct, size_t n_words, uint32_t root,
                              int depth, int res_x, int res_y, const orc_params* p,
                              int idx, int idy, float xyz[3], int32_t* pops, int max_pops,
                              int* n_pops, orc_stats* st);
/* Test hook: plant a mistake in orc_trace_pixel* so the tests can show their
 * checks catch it.  0 = the reference's rules (default); 1 = the LOD switch
 * evaluated on every loop trip instead of on pops (kernel.cl:209); 2 = on
 * descends instead; 3 = a ray that exits without a hit returns its parent
 * node's colour.  Process-global: only for single-threaded trace calls. */
void orc_set_mutation(int m);

/* ---- frame state (raycast.h:79-85, :169-173, :267-270) ---- */
typedef struct {
    int res_x, res_y, depth;
    const uint32_t* oct; size_t n_words; uint32_t root;
    uint32_t* screen;   /* 4 planes x N words            */
    float*    coords;   /* 4 planes x N float4 (x,y,z,w) */
    uint32_t* ids;      /* 2S + N words                  */
    uint32_t* image;    /* N words BGRX                  */
    int nthreads;       /* ray kernels only; results are order-independent */
    orc_stats st;
} orc_state;

int  orc_state_init(orc_state* s, int res_x, int res_y, const uint32_t* oct,
                    size_t n_words, uint32_t root, int depth);
void orc_state_free(orc_state* s);

/* individual kernels, reference argument meaning */
void orc_memset_screen(orc_state* s, size_t word_ofs, uint32_t val, size_t n_words);
void orc_proj(orc_state* s, int plane, const orc_params* p);      /* kernel.cl:472-592 */
void orc_counthole(orc_state* s);                                 /* kernel.cl:234-274 */
void orc_sumids(orc_state* s);                                    /* kernel.cl:276-296 */
void orc_writeids(orc_state* s);                                  /* kernel.cl:298-340 */
void orc_raycast_holes(orc_state* s, const orc_params* p);        /* kernel.cl:594-694 */
void orc_raycast_rect(orc_state* s, const orc_params* p, int add_x, int add_y,
                      int ext_x, int ext_y);                      /* kernel.cl:846-942 */
void orc_copy_plane(orc_state* s, int dst_plane);                 /* raycast.h:394-405 */
void orc_fillhole2_snapshot(orc_state* s, const uint32_t* snapshot); /* kernel.cl:404-470 */
void orc_fillhole2_inplace(orc_state* s);                         /* serial 16x16-group order */
void orc_colorize(orc_state* s);                                  /* kernel.cl:944-974 */

void orc_raycast_quads(orc_state* s, const orc_params* p, int add_x, int add_y); /* kernel.cl:696-843 */
/* kernel.cl:342-401 raycast_fillhole (disabled in the reference): plane-0
 * pixels at a depth discontinuity whose motion differs from the nearer
 * neighbour's become holes.  motion_x/y: N ints each (the reference's
 * mem_x/mem_y, which no kernel writes), or NULL for zero motion. */
void orc_fillhole_punch(orc_state* s, const int32_t* motion_x, const int32_t* motion_y);
void orc_tile_schedule(int res_x, int res_y, int frame, int* add_x, int* add_y,
                       int* ext_x, int* ext_y);                   /* raycast.h:363-366 */

/* whole frames: warp (raycast.h:147-438, filter on a pre-filter snapshot when
 * inplace_filter == 0) and full raycast (raycast_fine_2 over the screen). */
void orc_frame_warp(orc_state* s, const orc_params* p, int inplace_filter);
void orc_frame_full(orc_state* s, const orc_params* p);
/* gated warp frame (learned-fill gate, no reference counterpart): tiles with
 * more than max_holes residual holes -- hole pixels that neither the hole
 * list nor the refresh rectangle traces -- have them raycast with the frame's
 * hole and tile rays; with max_depth > 0 a tile holding a reprojected pixel
 * (outside the rectangle) of camera depth above max_depth has its residual
 * holes and every such far reprojected pixel raycast */
void orc_frame_warp_gated(orc_state* s, const orc_params* p, int max_holes, float max_depth);

/* exhaustive-check helpers */
uint32_t orc_bitcount8(uint32_t b);
uint32_t orc_colorize_word(uint32_t a);

#ifdef __cplusplus
}
#endif
#endif
