/*
 * svo_oracle.c -- CPU restatement of the reference SVO raycast + image-warp
 * pipeline.  TEST INFRASTRUCTURE: used only by tests/, smoke() and the
 * cpu_baseline leg of bench.py, as the checker.  Never linked into the
 * product library.  Build: oracle/Makefile (gcc -O2 -ffp-contract=off).
 *
 * Each function cites the reference lines it restates.  Semantics notes are
 * in svo_oracle.h.
 *
 * How each part is pinned (tests/): the octree builder word for word against
 * the reference's own octree.h (compiled from the checkout) and committed
 * fixtures; bitcount and colorize exhaustively; the tile schedule and the
 * hole compaction by closed forms and an independent formulation; fillhole2
 * per rule; the projection by its zero-motion invariant and an independent
 * key-min formulation; the camera by matrix identities.  orc_trace_pixel:
 * full-depth hits against an FP64 voxel march, LOD hits against the octree's
 * node colours, and the whole step/pop sequence, the level a LOD hit stops
 * at, the exit kind (hit, distance, y crossing), the exit node and its colour
 * against an independent FP64 model of the traversal (largest-empty-cell
 * skipping with the LOD rule applied at cell exits); planted mistakes in the
 * LOD rule and the exit node fail those checks (orc_set_mutation).
 */
#include "svo_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SCREEN_HOLE ORC_HOLE

void orc_free(void* p) { free(p); }

/* kernel.cl:23-29 get_bitcount: SWAR popcount of the low 8 bits. */
uint32_t orc_bitcount8(uint32_t b)
{
    b = (b & 0x55u) + (b >> 1 & 0x55u);
    b = (b & 0x33u) + (b >> 2 & 0x33u);
    b = (b + (b >> 4)) & 0x0fu;
    return b;
}

/* ------------------------------------------------------------------------
 * Camera: raycast.h:113-162 and :225-229/:322-326 with the nebula mathlib
 * (_matrix44.h:529-580 rotate_x/y/z, :334-344 transpose, :875-888 m*v4;
 * nmath.h:33-34 n_sin/n_cos = float(sin(double))).
 * ---------------------------------------------------------------------- */
typedef struct { float m[4][4]; } mat44;

static void mat_ident(mat44* a)
{
    memset(a, 0, sizeof(*a));
    for (int i = 0; i < 4; i++) a->m[i][i] = 1.0f;
}
static void mat_rotate_x(mat44* a, float ang)
{
    float c = (float)cos((double)ang), s = (float)sin((double)ang);
    for (int i = 0; i < 4; i++) {
        float mi1 = a->m[i][1], mi2 = a->m[i][2];
        a->m[i][1] = mi1 * c + mi2 * -s;
        a->m[i][2] = mi1 * s + mi2 * c;
    }
}
static void mat_rotate_y(mat44* a, float ang)
{
    float c = (float)cos((double)ang), s = (float)sin((double)ang);
    for (int i = 0; i < 4; i++) {
        float mi0 = a->m[i][0], mi2 = a->m[i][2];
        a->m[i][0] = mi0 * c + mi2 * s;
        a->m[i][2] = mi0 * -s + mi2 * c;
    }
}
static void mat_rotate_z(mat44* a, float ang)
{
    float c = (float)cos((double)ang), s = (float)sin((double)ang);
    for (int i = 0; i < 4; i++) {
        float mi0 = a->m[i][0], mi1 = a->m[i][1];
        a->m[i][0] = mi0 * c + mi1 * -s;
        a->m[i][1] = mi0 * s + mi1 * c;
    }
}
static void mat_transpose(mat44* a)
{
    for (int i = 0; i < 4; i++)
        for (int j = i + 1; j < 4; j++) {
            float t = a->m[i][j]; a->m[i][j] = a->m[j][i]; a->m[j][i] = t;
        }
}
/* _matrix44.h:875-888: v' = (M11*x + M21*y + M31*z + M41*w, ...) */
static orc_f4 mat_mul4(const mat44* a, float x, float y, float z, float w)
{
    orc_f4 r;
    r.x = a->m[0][0] * x + a->m[1][0] * y + a->m[2][0] * z + a->m[3][0] * w;
    r.y = a->m[0][1] * x + a->m[1][1] * y + a->m[2][1] * z + a->m[3][1] * w;
    r.z = a->m[0][2] * x + a->m[1][2] * y + a->m[2][2] * z + a->m[3][2] * w;
    r.w = a->m[0][3] * x + a->m[1][3] * y + a->m[2][3] * z + a->m[3][3] * w;
    return r;
}

void orc_camera(float yaw, float pitch, float roll, const float pos_world[3],
                int depth, int frame, orc_params* out, float pos_wrapped[3])
{
    /* raycast.h:136-145: wrap pos*16 into [0, 2*2^D) per axis */
    float dim2 = (float)((1 << depth) * 2);
    float p[3];
    for (int i = 0; i < 3; i++) {
        p[i] = pos_world[i] * 16.0f;
        while (p[i] < 0) p[i] += dim2;
        while (p[i] >= dim2) p[i] -= dim2;
    }
    float inv16 = 1.0f / 16.0f; /* _vector3.h:328-333 operator/ */
    for (int i = 0; i < 3; i++) pos_wrapped[i] = p[i] * inv16;

    mat44 m;
    mat_ident(&m);
    mat_rotate_z(&m, roll);
    mat_rotate_x(&m, pitch);
    mat_rotate_y(&m, yaw);
    out->m0.x = pos_wrapped[0]; out->m0.y = pos_wrapped[1];
    out->m0.z = pos_wrapped[2]; out->m0.w = 1.0f; /* _vector4.h:120-125 */
    out->row0 = mat_mul4(&m, 1, 0, 0, 0);  /* raycast.h:160-162 */
    out->row1 = mat_mul4(&m, 0, 1, 0, 0);
    out->row2 = mat_mul4(&m, 0, 0, 1, 0);
    mat_transpose(&m);                       /* raycast.h:322-325 */
    out->col0 = mat_mul4(&m, 1, 0, 0, 0);
    out->col1 = mat_mul4(&m, 0, 1, 0, 0);
    out->col2 = mat_mul4(&m, 0, 0, 1, 0);
    out->fovx = 1.0f; out->fovy = 1.0f;     /* raycast.h:109-110 */
    out->frame = frame;
    out->pad = 0;
}

/* ------------------------------------------------------------------------
 * Octree builder.  Restates octree.h:32-89 (set_voxel into a pointer octree
 * whose node handle is (index<<8)|childmask), :95-191 (convert_tree: compact
 * and byte nodes inside a block), :232-293 (convert_tree_blocks: 10-word
 * normal nodes, 64-word aligned block roots) and raycast.h:13-17,38-39
 * (zero root node, normal region pre-sized to 2^21 words).
 * ---------------------------------------------------------------------- */
typedef struct { uint32_t child[8]; uint32_t color; } pnode;

typedef struct {
    pnode* nodes; size_t n, cap;
    uint32_t root_handle;       /* octree_root */
    int depth;
    uint32_t* out; size_t out_n, out_cap;
    uint32_t normal_ofs;        /* octree_normal_offset */
    int iteration;              /* _iteration */
    int err;
} obuild;

static uint32_t ob_push(obuild* b, const pnode* nd)
{
    if (b->n == b->cap) {
        b->cap = b->cap ? b->cap * 2 : 1024;
        b->nodes = (pnode*)realloc(b->nodes, b->cap * sizeof(pnode));
    }
    b->nodes[b->n] = *nd;
    return (uint32_t)(b->n++);
}

static void ob_set_voxel(obuild* b, uint32_t x, uint32_t y, uint32_t z, uint32_t color)
{
    /* octree.h:32-89 */
    uint32_t offset = 0, parent_offset = 0, parent_childbit = 0, childbit = 0;
    pnode fresh; memset(&fresh, 0, sizeof(fresh)); fresh.color = color;
    int D = b->depth;
    for (int bit = D - 1;; bit--) {
        uint32_t bx = (x >> bit) & 1, by = (y >> bit) & 1, bz = (z >> bit) & 1;
        parent_childbit = childbit;
        childbit = bx + by * 2 + bz * 4;
        uint32_t child_offset = b->nodes[offset >> 8].child[childbit];
        if (bit < D - 1)
            b->nodes[parent_offset >> 8].child[parent_childbit] |= 1u << childbit;
        else
            b->root_handle |= 1u << childbit;
        if (bit == 0) {
            b->nodes[offset >> 8].child[childbit] = color;
            break;
        }
        parent_offset = offset;
        if (child_offset == 0) {
            uint32_t insert_ofs = (uint32_t)b->n << 8;
            b->nodes[offset >> 8].child[childbit] = insert_ofs;
            ob_push(b, &fresh);
            offset = insert_ofs;
        } else {
            offset = child_offset;
        }
    }
}

static void ob_out_resize(obuild* b, size_t n)
{
    if (n > b->out_cap) {
        size_t c = b->out_cap ? b->out_cap : 1 << 22;
        while (c < n) c *= 2;
        b->out = (uint32_t*)realloc(b->out, c * sizeof(uint32_t));
        b->out_cap = c;
    }
    if (n > b->out_n) memset(b->out + b->out_n, 0, (n - b->out_n) * sizeof(uint32_t));
    b->out_n = n;
}
static void ob_out_push(obuild* b, uint32_t w)
{
    ob_out_resize(b, b->out_n + 1);
    b->out[b->out_n - 1] = w;
}

/* octree.h:95-191 */
static uint32_t ob_convert_tree(obuild* b, uint32_t offset, uint32_t local_root)
{
    b->iteration++;
    int D = b->depth;
    if (b->iteration == D - 1) {
        /* byte node: colour byte, child masks, leaf colour bytes */
        uint32_t cur = (uint32_t)(((b->out_n - local_root) << 9) + 256 + (offset & 0xff));
        pnode nd = b->nodes[offset >> 8];
        uint8_t blk[1 + 8 + 64];
        int size = 0;
        blk[size++] = (uint8_t)nd.color;
        for (int j = 0; j < 8; j++)
            if (nd.child[j] > 0) blk[size++] = (uint8_t)nd.child[j];
        for (int j = 0; j < 8; j++) {
            if (nd.child[j] > 0) {
                pnode c = b->nodes[nd.child[j] >> 8];
                for (int i = 0; i < 8; i++)
                    if (c.child[i] > 0) blk[size++] = (uint8_t)c.child[i];
            }
        }
        uint32_t dw = 0; int i03 = 0;
        for (int i = 0; i < size; i++) {
            i03 = i & 3;
            dw |= (uint32_t)blk[i] << (i03 * 8);
            if (i03 == 3) { ob_out_push(b, dw); dw = 0; }
        }
        if (i03 < 3) ob_out_push(b, dw);
        b->iteration--;
        return cur;
    }
    pnode nd = b->nodes[offset >> 8];
    for (int j = 0; j < 8; j++)
        if (nd.child[j] > 0) nd.child[j] = ob_convert_tree(b, nd.child[j], local_root);
    uint32_t cur = (uint32_t)(((b->out_n - local_root) << 9) + 256 + (offset & 0xff));
    ob_out_push(b, nd.color);
    for (int j = 0; j < 8; j++)
        if (nd.child[j] > 0) ob_out_push(b, nd.child[j]);
    b->iteration--;
    return cur;
}

/* octree.h:232-293 */
static uint32_t ob_convert_blocks(obuild* b, uint32_t offset)
{
    b->iteration++;
    int D = b->depth;
    pnode nd = b->nodes[offset >> 8];
    for (int j = 0; j < 8; j++) {
        uint32_t childofs = nd.child[j];
        if (childofs > 0) {
            if (b->iteration < D - 6) {
                nd.child[j] = ob_convert_blocks(b, childofs);
            } else {
                uint32_t size64 = (uint32_t)(((b->out_n + 63) >> 6) << 6);
                uint32_t local_root = size64;
                uint32_t num_entries = orc_bitcount8(childofs & 255) + 1;
                ob_out_resize(b, size64 + num_entries);
                nd.child[j] = ((local_root >> 6) << 9) | (1u << 8) | (childofs & 255);
                uint32_t subtree_root = ob_convert_tree(b, childofs, local_root) >> 9;
                for (uint32_t i = 0; i < num_entries; i++)
                    b->out[local_root + i] = b->out[local_root + subtree_root + i];
                ob_out_resize(b, size64 + subtree_root);
                if (((uint64_t)local_root >> 6) >= (1u << 23)) b->err = 1;
            }
        }
    }
    if (b->normal_ofs + 10 > (1u << 21)) { b->err = 2; b->iteration--; return 0; }
    b->out[b->normal_ofs + 8] = nd.color;
    b->out[b->normal_ofs + 9] = nd.color;
    for (int j = 0; j < 8; j++) b->out[b->normal_ofs + j] = nd.child[j];
    b->normal_ofs += 10;
    b->iteration--;
    return ((b->normal_ofs - 10) << 9) | (0u << 8) | (offset & 255);
}

int orc_build_octree(const uint16_t* xyz, const uint32_t* rgba, size_t n,
                     int depth, uint32_t** words_out, size_t* n_words,
                     uint32_t* root_ref)
{
    if (depth < 8 || depth > 13) return -1;
    obuild b; memset(&b, 0, sizeof(b));
    b.depth = depth;
    pnode zero; memset(&zero, 0, sizeof(zero));
    ob_push(&b, &zero);                       /* raycast.h:15-17 */
    for (size_t i = 0; i < n; i++)
        ob_set_voxel(&b, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], rgba[i]);
    ob_out_resize(&b, 2097152);                /* raycast.h:38 */
    uint32_t root = ob_convert_blocks(&b, b.root_handle);
    free(b.nodes);
    if (b.err) { free(b.out); return -2; }
    *words_out = b.out; *n_words = b.out_n; *root_ref = root;
    return 0;
}

/* ------------------------------------------------------------------------
 * Traversal.  kernel.cl:32-62 fetchChildOffset, :64-112 fetchColor,
 * :114-214 CAST_RAY, ray set-up :639-663 / :883-908, output :686-693.
 * Reads past the end of the word array return 0 (defined here; the
 * reference reads unspecified memory there).
 * ---------------------------------------------------------------------- */
typedef struct { const uint32_t* w; size_t n; orc_stats* st; } octv;

static inline uint32_t rd(const octv* o, uint32_t i)
{
    if (o->st) o->st->fetch_words++;
    return (size_t)i < o->n ? o->w[i] : 0u;
}

static uint32_t fetch_child(const octv* o, uint32_t a_node, uint32_t a_node_before,
                            uint32_t* local_root, uint32_t a_child,
                            uint32_t child_test, int rek)
{
    uint32_t n = a_node >> 9;
    uint32_t nadd = a_child;
    if (a_node & 256) {
        if (!(a_node_before & 256)) { *local_root = n << 6; n = 0; }
        n += *local_root;
        nadd = orc_bitcount8(a_node & ((child_test << 1) - 1));
        if (rek == 2)
            return ((rd(o, n + (nadd >> 2)) >> ((nadd & 3) << 3)) & 255) + 256 + (nadd << 9);
    }
    return rd(o, n + nadd);
}

static uint32_t fetch_color(const octv* o, uint32_t a_node, uint32_t a_node_before,
                            uint32_t a_node_before2, uint32_t* local_root, int rek,
                            uint32_t child_test)
{
    uint32_t n = a_node >> 9;
    uint32_t nadd = 8;
    if (rek == 1) {
        n = (a_node_before >> 9) & 15;
        uint32_t ofs = (a_node_before2 >> 9) + *local_root;
        nadd = 0 + orc_bitcount8(a_node_before2 & 255);
        for (uint32_t i = 1; i < n; i++)
            nadd += orc_bitcount8((rd(o, ofs + (i >> 2)) >> ((i & 3) << 3)) & 255);
        nadd += orc_bitcount8(a_node_before & ((child_test << 1) - 1));
        return rd(o, ofs + (nadd >> 2)) >> ((nadd & 3) << 3);
    }
    if (a_node & 256) {
        if (!(a_node_before & 256)) { *local_root = n << 6; n = 0; }
        nadd = *local_root;
        if (rek == 2) {
            uint32_t ofs = (a_node_before >> 9) + *local_root;
            nadd = 1 + orc_bitcount8(a_node_before & 255);
            for (uint32_t i = 1; i < n; i++)
                nadd += orc_bitcount8((rd(o, ofs + (i >> 2)) >> ((i & 3) << 3)) & 255);
            return rd(o, ofs + (nadd >> 2)) >> ((nadd & 3) << 3);
        }
    }
    return rd(o, n + nadd);
}

/* OpenCL sign(): +-1 for nonzero, the zero itself for +-0 */
static inline float cl_sign(float v)
{
    if (v > 0.0f) return 1.0f;
    if (v < 0.0f) return -1.0f;
    return v;
}
static inline float len3(float x, float y, float z) { return sqrtf((x * x + y * y) + z * z); }
static inline float dot3(float ax, float ay, float az, float bx, float by, float bz)
{
    return (ax * bx + ay * by) + az * bz;
}
/* (int)log2(convert_float(v)) for v in [0, 2^D): floor(log2 v) for v > 0
 * (exact: every such v is an integer below 2^24); v == 0 gives -inf whose
 * int conversion is negative on both x86 and gfx9, and any negative value
 * takes the same branch (kernel.cl:204), so -1 stands for it. */
static inline int ilog2_or_neg(uint32_t v) { return v ? 31 - __builtin_clz(v) : -1; }

static uint32_t trace_impl(const uint32_t* oct, size_t n_words, uint32_t root,
                           int depth, int res_x, int res_y, const orc_params* p,
                           int idx, int idy, float xyz[3], orc_stats* st,
                           int32_t* pops, int max_pops, int* n_pops);

uint32_t orc_trace_pixel(const uint32_t* oct, size_t n_words, uint32_t root,
                         int depth, int res_x, int res_y, const orc_params* p,
                         int idx, int idy, float xyz[3], orc_stats* st)
{
    return trace_impl(oct, n_words, root, depth, res_x, res_y, p, idx, idy, xyz, st, NULL, 0, NULL);
}

uint32_t orc_trace_pixel_pops(const uint32_t* oct, size_t n_words, uint32_t root,
                              int depth, int res_x, int res_y, const orc_params* p,
                              int idx, int idy, float xyz[3], int32_t* pops, int max_pops,
                              int* n_pops, orc_stats* st)
{
    *n_pops = 0;
    return trace_impl(oct, n_words, root, depth, res_x, res_y, p, idx, idy, xyz, st, pops, max_pops, n_pops);
}

/* Planted mistakes for the tests that show the traversal checks catch them
 * (0 = the reference's rules; see svo_oracle.h).  Not thread-safe: set it
 * only around single-threaded trace calls. */
static int g_mutation = 0;
void orc_set_mutation(int m) { g_mutation = m; }

static uint32_t trace_impl(const uint32_t* oct, size_t n_words, uint32_t root,
                           int depth, int res_x, int res_y, const orc_params* p,
                           int idx, int idy, float xyz[3], orc_stats* st,
                           int32_t* pops, int max_pops, int* n_pops)
{
    const int D = depth;
    const int DAND = (1 << D) - 1;
    const int SCALE_MAX = 1 << (D + 1);
    octv o = { oct, n_words, st };

    /* kernel.cl:883-891 */
    float d1x = ((float)(idx - res_x / 2) + 0.5f) * p->fovx / (float)res_y;
    float d1y = ((float)(idy - res_y / 2) + 0.5f) * p->fovy / (float)res_y;
    float d1z = 1.0f;
    float dx = dot3(d1x, d1y, d1z, p->col0.x, p->col0.y, p->col0.z);
    float dy = dot3(d1x, d1y, d1z, p->col1.x, p->col1.y, p->col1.z);
    float dz = dot3(d1x, d1y, d1z, p->col2.x, p->col2.y, p->col2.z);
    float distance = 0.0f;
    float px = p->m0.x * 16.0f, py = p->m0.y * 16.0f, pz = p->m0.z * 16.0f;
    int lod = 1;
    float lodswitch = (float)(res_x * 1 * 2);

    /* CAST_RAY, kernel.cl:114-214 */
    int sign_x = 0, sign_y = 0, sign_z = 0;
    int rek = D;
    const float eps = 0x1p-20f;
    if (fabsf(dx) < eps) dx = cl_sign(dx) * eps;
    if (fabsf(dy) < eps) dy = cl_sign(dy) * eps;
    if (fabsf(dz) < eps) dz = cl_sign(dz) * eps;
    if (dx < 0) { sign_x = 1; px = (float)SCALE_MAX - px; dx = -dx; }
    if (dy < 0) { sign_y = 1; py = (float)SCALE_MAX - py; dy = -dy; }
    if (dz < 0) { sign_z = 1; pz = (float)SCALE_MAX - pz; dz = -dz; }
    float g0x = 1, g0y = dy / dx, g0z = dz / dx;
    float g1x = dx / dy, g1y = 1, g1z = dz / dy;
    float g2x = dx / dz, g2y = dy / dz, g2z = 1;
    float lx = len3(g0x, g0y, g0z), ly = len3(g1x, g1y, g1z), lz = len3(g2x, g2y, g2z);
    int ix = (int)px, iy = (int)py, iz = (int)pz;
    uint32_t nodeid = root, nodeid_before = root, local_root = 0, node_test = 0;
    int x0ry = 0;
    int sign_xyz = sign_x | (sign_y << 1) | (sign_z << 2);
    uint32_t stack[16];
    memset(stack, 0, sizeof(stack));

    if (st) st->rays++;
    uint32_t guard = 0;
    do {
        /* iteration bound shared with the GPU path: a ray still running
         * after 2^22 loop trips finishes as a miss at its current node */
        if (++guard > (1u << 22)) break;
        if (st) st->iters++;
        if (g_mutation == 1 && distance > lodswitch) { lodswitch *= 2; ++lod; }  /* planted: every trip */
        int p0x = ix >> rek, p0y = iy >> rek, p0z = iz >> rek;
        int node_index = ((p0x & 1) | ((p0y & 1) << 1) | ((p0z & 1) << 2)) ^ sign_xyz;
        node_test = nodeid & (1u << node_index);
        if (node_test) {
            uint32_t tmp = nodeid;
            if (st) { st->descends++; st->desc_by_rek[rek & 15]++; }
            nodeid = fetch_child(&o, nodeid, nodeid_before, &local_root,
                                 (uint32_t)node_index, node_test, rek);
            nodeid_before = tmp;
            if (g_mutation == 2 && distance > lodswitch) { lodswitch *= 2; ++lod; }  /* planted: on descends */
            if (rek <= lod) break;
            rek--;
            stack[rek] = nodeid;
            continue;
        }
        if (st) st->steps++;
        float mx = (float)(int)((uint32_t)(p0x + 1) << rek) - px;
        float my = (float)(int)((uint32_t)(p0y + 1) << rek) - py;
        float mz = (float)(int)((uint32_t)(p0z + 1) << rek) - pz;
        int bx = (int)px, by = (int)py, bz = (int)pz;
        float idist = mx * lx, idy_ = my * ly, idz = mz * lz;
        float sx = g0x * mx, sy = g0y * mx, sz = g0z * mx;
        if (idy_ < idist) { idist = idy_; sx = g1x * my; sy = g1y * my; sz = g1z * my; }
        if (idz < idist) { idist = idz; sx = g2x * mz; sy = g2y * mz; sz = g2z * mz; }
        px += sx; py += sy; pz += sz;
        ix = (int)px; iy = (int)py; iz = (int)pz;
        distance += idist;
        int x0rx = bx ^ ix;
        x0ry = by ^ iy;
        int x0rz = bz ^ iz;
        int x0r = x0rx ^ x0ry ^ x0rz;
        int rek_new = ilog2_or_neg((uint32_t)(x0r & DAND));
        if (pops && *n_pops < max_pops) {
            /* every step: integer cells before and after, the level, the XOR
             * level + 1 (0: no bit changed), whether it ends the ray by distance */
            int32_t* e = pops + 9 * (*n_pops)++;
            e[0] = bx; e[1] = by; e[2] = bz; e[3] = ix; e[4] = iy; e[5] = iz;
            e[6] = rek; e[7] = rek_new + 1; e[8] = distance > 400000.0f;
        }
        if (distance > 400000.0f) break;
        if (rek_new < rek) continue;
        if (st) st->pops++;
        rek = rek_new + 1;
        nodeid = stack[rek];
        nodeid_before = stack[rek + 1];
        if (rek == D) nodeid_before = nodeid = root;
        if (g_mutation != 1 && g_mutation != 2 && distance > lodswitch) { lodswitch *= 2; ++lod; }
    } while (!(x0ry & (2 * DAND + 2)));
    if (sign_x) px = (float)SCALE_MAX - px;
    if (sign_y) py = (float)SCALE_MAX - py;
    if (sign_z) pz = (float)SCALE_MAX - pz;

    if (g_mutation == 3 && !node_test) nodeid = nodeid_before;  /* planted: exit colour from the parent */
    uint32_t col = fetch_color(&o, nodeid, nodeid_before, stack[rek + 1], &local_root,
                               rek, node_test);
    xyz[0] = px / 16.0f; xyz[1] = py / 16.0f; xyz[2] = pz / 16.0f;
    if (st) {
        st->last_rek = rek; st->last_lod = lod; st->last_hit = node_test != 0;
        st->last_ix = ix; st->last_iy = iy; st->last_iz = iz; st->last_sign = sign_xyz;
        const uint64_t trips = guard > (1u << 22) ? (1u << 22) : guard;
        if (trips > st->max_iters) st->max_iters = trips;
    }
    return 0xff000000u + col;
}

/* ------------------------------------------------------------------------
 * Frame state and per-frame kernels.
 * ---------------------------------------------------------------------- */
int orc_state_init(orc_state* s, int res_x, int res_y, const uint32_t* oct,
                   size_t n_words, uint32_t root, int depth)
{
    memset(s, 0, sizeof(*s));
    s->res_x = res_x; s->res_y = res_y; s->depth = depth;
    s->oct = oct; s->n_words = n_words; s->root = root;
    size_t N = (size_t)res_x * res_y;
    size_t S = (size_t)(res_x / 16) * (res_y / 16);
    s->screen = (uint32_t*)calloc(4 * N, 4);
    s->coords = (float*)calloc(16 * N, 4);
    s->ids = (uint32_t*)calloc(2 * S + N + 1, 4);
    s->image = (uint32_t*)calloc(N, 4);
    s->nthreads = 1;
    return (s->screen && s->coords && s->ids && s->image) ? 0 : -1;
}

void orc_state_free(orc_state* s)
{
    free(s->screen); free(s->coords); free(s->ids); free(s->image);
    memset(s, 0, sizeof(*s));
}

/* kernel.cl:5 memset, issued by ocl_memset (ocl.h:299-309) */
void orc_memset_screen(orc_state* s, size_t word_ofs, uint32_t val, size_t n_words)
{
    for (size_t i = 0; i < n_words; i++) s->screen[word_ofs + i] = val;
}

/* kernel.cl:472-592, executed sequentially in row-major source order (the
 * defined order: the first source with the strictly smallest depth key wins
 * a target pixel). */
void orc_proj(orc_state* s, int plane, const orc_params* p)
{
    const int res_x = s->res_x, res_y = s->res_y;
    const size_t N = (size_t)res_x * res_y;
    const size_t ofs_add = (size_t)plane * N;
    uint32_t* scr = s->screen;
    float* bb = s->coords;
    for (int idy = 0; idy < res_y; idy++)
        for (int idx = 0; idx < res_x; idx++) {
            size_t ofs = (size_t)idy * res_x + idx + ofs_add;
            size_t srcofs = ofs;
            uint32_t col = scr[ofs];
            if (col == SCREEN_HOLE) continue;
            float cx = bb[ofs * 4 + 0], cy = bb[ofs * 4 + 1], cz = bb[ofs * 4 + 2];
            float qx = cx - p->m0.x, qy = cy - p->m0.y, qz = cz - p->m0.z;
            float hx = dot3(qx, qy, qz, p->row0.x, p->row0.y, p->row0.z);
            float hy = dot3(qx, qy, qz, p->row1.x, p->row1.y, p->row1.z);
            float hz = dot3(qx, qy, qz, p->row2.x, p->row2.y, p->row2.z);
            if (hz < 0.05f) { scr[srcofs] = SCREEN_HOLE; continue; }
            float sfx = (hx * (float)res_y + 0.0f) / hz + (float)res_x / 2 - 0.0f;
            float sfy = (hy * (float)res_y + 0.0f) / hz + (float)res_y / 2 - 0.0f;
            int sxi = (int)sfx, syi = (int)sfy;
            if (sxi >= res_x - 1 || sxi < 0 || syi >= res_y - 1 || syi < 0) {
                scr[srcofs] = SCREEN_HOLE; continue;
            }
            size_t t = (size_t)syi * res_x + sxi;
            uint32_t sz = (uint32_t)(int)(hz * 1000.0f) << 8;
            uint32_t val = sz + (col & 255);
            if ((scr[t] & 0xffffff00u) <= sz) continue;
            uint32_t old = scr[t];
            if (val < old) scr[t] = val;             /* atom_min */
            if (old < val) continue;
            if (scr[t] != val) continue;
            bb[t * 4 + 0] = cx; bb[t * 4 + 1] = cy; bb[t * 4 + 2] = cz;
            bb[t * 4 + 3] = hz;
        }
}

/* kernel.cl:234-274 */
void orc_counthole(orc_state* s)
{
    const int res_x = s->res_x, res_y = s->res_y;
    const int bx = res_x / 16, by = res_y / 16;
    for (int idy = 0; idy < by; idy++)
        for (int idx = 0; idx < bx; idx++) {
            int count = 0, x = idx * 16, y = idy * 16;
            int ofs = x + y * res_x;
            int dx = (res_x - x < 16 ? res_x - x : 16) / 2;
            int dy = (res_y - y < 16 ? res_y - y : 16) / 2;
            for (int j = 0; j < dy; j++)
                for (int i = 0; i < dx; i++) {
                    int o = ofs + i * 2 + j * 2 * res_x, a = 0;
                    if (s->screen[o] == SCREEN_HOLE) a++;
                    if (s->screen[o + 1] == SCREEN_HOLE) a++;
                    if (s->screen[o + 1 + res_x] == SCREEN_HOLE) a++;
                    if (s->screen[o + res_x] == SCREEN_HOLE) a++;
                    if (a >= 4) count += 4;
                }
            s->ids[idx + idy * bx] = (uint32_t)count;
        }
}

/* kernel.cl:276-296 */
void orc_sumids(orc_state* s)
{
    int size = (s->res_x / 16) * (s->res_y / 16);
    uint32_t ofs = 0;
    for (int i = 0; i < size; i++) {
        uint32_t len = s->ids[i];
        s->ids[i + size] = ofs;
        ofs += len;
    }
    s->ids[0] = ofs;
}

/* kernel.cl:298-340 */
void orc_writeids(orc_state* s)
{
    const int res_x = s->res_x, res_y = s->res_y;
    const int bx = res_x / 16, by = res_y / 16;
    const int size = bx * by;
    for (int idy = 0; idy < by; idy++)
        for (int idx = 0; idx < bx; idx++) {
            int x = idx * 16, y = idy * 16;
            int ofs = x + y * res_x;
            uint32_t dst = s->ids[idx + idy * bx + size] + (uint32_t)size * 2;
            int dx = (res_x - x < 16 ? res_x - x : 16) / 2;
            int dy = (res_y - y < 16 ? res_y - y : 16) / 2;
            for (int j = 0; j < dy; j++)
                for (int i = 0; i < dx; i++) {
                    int a = 0, o = ofs + i * 2 + j * 2 * res_x;
                    uint32_t val = (uint32_t)(x + i * 2 + ((j * 2 + y) << 16));
                    if (s->screen[o] == SCREEN_HOLE) a++;
                    if (s->screen[o + 1] == SCREEN_HOLE) a++;
                    if (s->screen[o + 1 + res_x] == SCREEN_HOLE) a++;
                    if (s->screen[o + res_x] == SCREEN_HOLE) a++;
                    if (a >= 4) {
                        s->ids[dst++] = val;
                        s->ids[dst++] = val + 1;
                        s->ids[dst++] = val + 1 + (1u << 16);
                        s->ids[dst++] = val + (1u << 16);
                    }
                }
        }
}

static void write_ray(orc_state* s, size_t ofs, uint32_t word, const float xyz[3])
{
    s->screen[ofs] = word;
    s->coords[ofs * 4 + 0] = xyz[0];
    s->coords[ofs * 4 + 1] = xyz[1];
    s->coords[ofs * 4 + 2] = xyz[2];
}

static void add_stats(orc_stats* a, const orc_stats* b)
{
    a->rays += b->rays; a->iters += b->iters; a->descends += b->descends;
    a->steps += b->steps; a->pops += b->pops; a->fetch_words += b->fetch_words;
    for (int i = 0; i < 16; i++) a->desc_by_rek[i] += b->desc_by_rek[i];
    if (b->max_iters > a->max_iters) a->max_iters = b->max_iters;
}

/* kernel.cl:594-694: one ray per id in ids[2S .. 2S+ids[0]) */
void orc_raycast_holes(orc_state* s, const orc_params* p)
{
    const int res_x = s->res_x, res_y = s->res_y;
    const size_t S = (size_t)(res_x / 16) * (res_y / 16);
    const long H = (long)s->ids[0];
    orc_stats tot; memset(&tot, 0, sizeof(tot));
#pragma omp parallel num_threads(s->nthreads)
    {
        orc_stats loc; memset(&loc, 0, sizeof(loc));
#pragma omp for schedule(dynamic, 256)
        for (long id = 0; id < H; id++) {
            uint32_t idxy = s->ids[id + S * 2];
            int idx = (int)(idxy & 0xffff), idy = (int)(idxy >> 16);
            if (idx >= res_x || idy >= res_y) continue;
            float xyz[3];
            uint32_t w = orc_trace_pixel(s->oct, s->n_words, s->root, s->depth,
                                         res_x, res_y, p, idx, idy, xyz, &loc);
            write_ray(s, (size_t)idy * res_x + idx, w, xyz);
        }
#pragma omp critical
        add_stats(&tot, &loc);
    }
    add_stats(&s->st, &tot);
}

/* kernel.cl:846-942: one ray per pixel (add_x+i, add_y+j), i<ext_x, j<ext_y,
 * clipped to the screen (ext = the rounded-up global work size). */
void orc_raycast_rect(orc_state* s, const orc_params* p, int add_x, int add_y,
                      int ext_x, int ext_y)
{
    const int res_x = s->res_x, res_y = s->res_y;
    orc_stats tot; memset(&tot, 0, sizeof(tot));
    long total = (long)ext_x * ext_y;
#pragma omp parallel num_threads(s->nthreads)
    {
        orc_stats loc; memset(&loc, 0, sizeof(loc));
#pragma omp for schedule(dynamic, 256)
        for (long k = 0; k < total; k++) {
            int idx = (int)(k % ext_x) + add_x, idy = (int)(k / ext_x) + add_y;
            if (idx >= res_x || idy >= res_y) continue;
            float xyz[3];
            uint32_t w = orc_trace_pixel(s->oct, s->n_words, s->root, s->depth,
                                         res_x, res_y, p, idx, idy, xyz, &loc);
            write_ray(s, (size_t)idy * res_x + idx, w, xyz);
        }
#pragma omp critical
        add_stats(&tot, &loc);
    }
    add_stats(&s->st, &tot);
}

/* raycast_fine (kernel.cl:696-843; its launch at raycast.h:236-259 is
 * disabled): one ray per 2x2 quad at (2 gx + add_x, 2 gy + add_y) for gx <
 * roundup16(res_x/4), gy < roundup16(res_y/4) (ocl.h:188-198 global-size
 * rounding), skipped when that corner is off the screen (:722).  The traced
 * pixel is the quad's first hole in the order (0,0), (1,0), (0,1), (1,1)
 * (:726-737), else ((frame>>2)&1, (frame>>3)&1) (:738-742); the reads and the
 * write use flat offsets y*res_x + x (:727, :804), so at the last column or
 * row they reach the next row or plane 1.  All picks are made before any ray
 * writes (defined order; the reference's work-items race only at those
 * edges). */
void orc_raycast_quads(orc_state* s, const orc_params* p, int add_x, int add_y)
{
    const int res_x = s->res_x, res_y = s->res_y;
    const int gx_n = ((res_x / 4 + 15) / 16) * 16, gy_n = ((res_y / 4 + 15) / 16) * 16;
    const long nq = (long)gx_n * gy_n;
    int* pick = (int*)malloc(sizeof(int) * 2 * (size_t)nq);
    for (long g = 0; g < nq; g++) {
        int idx = 2 * (int)(g % gx_n) + add_x, idy = 2 * (int)(g / gx_n) + add_y;
        if (idx >= res_x || idy >= res_y) {
            pick[2 * g] = -1;
            continue;
        }
        int found = 0;
        for (int i = 0; i < 4 && !found; i++) {
            size_t o = (size_t)(idy + (i >> 1)) * res_x + idx + (i & 1);
            if (s->screen[o] == SCREEN_HOLE) {
                idx += i & 1;
                idy += i >> 1;
                found = 1;
            }
        }
        if (!found) {
            idx += (p->frame >> 2) & 1;
            idy += (p->frame >> 3) & 1;
        }
        pick[2 * g] = idx;
        pick[2 * g + 1] = idy;
    }
    orc_stats tot; memset(&tot, 0, sizeof(tot));
#pragma omp parallel num_threads(s->nthreads)
    {
        orc_stats loc; memset(&loc, 0, sizeof(loc));
#pragma omp for schedule(dynamic, 64)
        for (long g = 0; g < nq; g++) {
            if (pick[2 * g] < 0) continue;
            int idx = pick[2 * g], idy = pick[2 * g + 1];
            float xyz[3];
            uint32_t w = orc_trace_pixel(s->oct, s->n_words, s->root, s->depth, res_x, res_y, p, idx, idy, xyz, &loc);
            write_ray(s, (size_t)idy * res_x + idx, w, xyz);
        }
#pragma omp critical
        add_stats(&tot, &loc);
    }
    add_stats(&s->st, &tot);
    free(pick);
}

/* raycast.h:394-405 (ocl_memcpy of the screen words then the float4 coords) */
void orc_copy_plane(orc_state* s, int dst_plane)
{
    size_t N = (size_t)s->res_x * s->res_y;
    memcpy(s->screen + dst_plane * N, s->screen, N * 4);
    memcpy(s->coords + dst_plane * N * 4, s->coords, N * 16);
}

/* kernel.cl:404-470 for one pixel.  src holds the words the filter reads;
 * reads at linear offsets >= limit (only the 5x5 scan of the last two rows
 * reaches them) return the hole word. */
static inline uint32_t rs(const uint32_t* src, long o, long limit)
{
    return o < limit ? src[o] : SCREEN_HOLE;
}

static uint32_t fill_pixel(const uint32_t* src, long ofs, int res_x, long limit)
{
    uint32_t col = rs(src, ofs, limit);
    int col1 = (int)rs(src, ofs + 1, limit);
    int col2 = (int)rs(src, ofs - 1, limit);
    int col3 = (int)rs(src, ofs + res_x, limit);
    int col4 = (int)rs(src, ofs - res_x, limit);
    if ((uint32_t)col1 != SCREEN_HOLE && (uint32_t)col2 != SCREEN_HOLE &&
        (uint32_t)col3 != SCREEN_HOLE && (uint32_t)col4 != SCREEN_HOLE)
        return (uint32_t)((col1 & 3) +
                          ((((col1 & 0xfc) + (col2 & 0xfc) + (col3 & 0xfc) + (col4 & 0xfc)) >> 2) & 0xfc));
    if ((uint32_t)col1 != SCREEN_HOLE && (uint32_t)col2 != SCREEN_HOLE)
        return (uint32_t)((col1 & 3) + ((((col1 & 0xfc) + (col2 & 0xfc)) >> 1) & 0xfc));
    if ((uint32_t)col3 != SCREEN_HOLE && (uint32_t)col4 != SCREEN_HOLE)
        return (uint32_t)((col3 & 3) + ((((col3 & 0xfc) + (col4 & 0xfc)) >> 1) & 0xfc));
    for (int i = 0; i < 4; i++) {
        long o = ofs + (i & 1) + ((i >> 1) & 1) * res_x;
        col = rs(src, o, limit);
        if (col != SCREEN_HOLE) break;
    }
    if (col == SCREEN_HOLE) {
        /* x offset outer, y offset inner; the reference's break leaves only
         * the inner loop, so later columns break at once */
        for (int i = -2; i < 3; i++)
            for (int j = -2; j < 3; j++) {
                long o = ofs + i + (long)j * res_x;
                if (col != SCREEN_HOLE) break;
                col = rs(src, o, limit);
            }
    }
    return col;
}

/* kernel.cl:342-401 raycast_fillhole.  Per pixel 3 <= x < res_x-4,
 * 3 <= y < res_y-4 of plane 0 that is not a hole: z1 = the smallest camera
 * depth (coordinate w) among the pixel and its neighbours (i, j) in
 * {-1, 0}^2, the first strictly smaller in i-outer, j-inner order
 * (:376-381); if none is smaller the pixel stays (the reference reads the
 * motion buffers at an uninitialised offset first, :384-388, then returns);
 * else it becomes a hole when |z1 - z0| > min(z0, z1) * Z_CONTINUITY * 4
 * (Z_CONTINUITY = 0.00625, single precision) and the pixel's motion differs
 * from that neighbour's (:390-392).  Each pixel reads only its own screen
 * word, so the in-place update has no order dependence. */
void orc_fillhole_punch(orc_state* s, const int32_t* mx, const int32_t* my)
{
    const int res_x = s->res_x, res_y = s->res_y;
    for (int y = 3; y < res_y - 4; y++)
        for (int x = 3; x < res_x - 4; x++) {
            const long of = (long)y * res_x + x;
            if (s->screen[of] == SCREEN_HOLE) continue;
            const float z0 = s->coords[of * 4 + 3];
            float z1 = z0;
            int xi = 0, yj = 0;
            for (int i = -1; i < 1; i++)
                for (int j = -1; j < 1; j++) {
                    const float z = s->coords[(of + (long)j * res_x + i) * 4 + 3];
                    if (z < z1) { xi = i; yj = j; z1 = z; }
                }
            if (z0 <= z1) continue;
            const long oij = of + (long)yj * res_x + xi;
            const int dx0 = mx ? mx[of] : 0, dy0 = my ? my[of] : 0;
            const int dx = mx ? mx[oij] : 0, dy = my ? my[oij] : 0;
            const float zmin = z0 < z1 ? z0 : z1;
            if (fabsf(z1 - z0) > zmin * 0.00625f * 4.0f && (abs(dx - dx0) > 0 || abs(dy - dy0) > 0))
                s->screen[of] = SCREEN_HOLE;
        }
}

/* Filter defined on a pre-filter snapshot of plane 0 (every hole reads
 * unfiltered neighbours). */
void orc_fillhole2_snapshot(orc_state* s, const uint32_t* snap)
{
    const int res_x = s->res_x, res_y = s->res_y;
    const long N = (long)res_x * res_y;
    for (int idy = 2; idy < res_y - 1; idy++)
        for (int idx = 2; idx < res_x - 1; idx++) {
            long ofs = (long)idy * res_x + idx;
            if (snap[ofs] != SCREEN_HOLE) continue;
            s->screen[ofs] = fill_pixel(snap, ofs, res_x, N);
        }
}

/* Filter in place, work-items serial in OpenCL order: 16x16 groups in
 * row-major group order, items x-fastest inside a group.  Reads below the
 * last row land in plane 1, as in the reference's single buffer. */
void orc_fillhole2_inplace(orc_state* s)
{
    const int res_x = s->res_x, res_y = s->res_y;
    const long lim = 4L * res_x * res_y;
    int gx = (res_x + 15) / 16, gy = (res_y + 15) / 16;
    for (int by = 0; by < gy; by++)
        for (int bx = 0; bx < gx; bx++)
            for (int ty = 0; ty < 16; ty++)
                for (int tx = 0; tx < 16; tx++) {
                    int idx = bx * 16 + tx, idy = by * 16 + ty;
                    if (idx >= res_x - 1 || idy >= res_y - 1 || idx <= 1 || idy <= 1) continue;
                    long ofs = (long)idy * res_x + idx;
                    if (s->screen[ofs] != SCREEN_HOLE) continue;
                    s->screen[ofs] = fill_pixel(s->screen, ofs, res_x, lim);
                }
}

/* kernel.cl:944-974 */
uint32_t orc_colorize_word(uint32_t word)
{
    static const float tab[4][3] = {
        {1.0f, 1.0f, 1.0f}, {1.0f, 0.7f, 0.3f}, {1.5f, 0.8f, 0.1f}, {0.2f, 0.8f, 0.2f}};
    int a = (int)word;
    int c = a & 3;
    float i = (float)(a & (255 - 7));
    int r = (int)(i * tab[c][0]); if (r > 255) r = 255;
    int g = (int)(i * tab[c][1]); if (g > 255) g = 255;
    int b = (int)(i * tab[c][2]); if (b > 255) b = 255;
    return (uint32_t)(b + g * 256 + r * 65536);
}

void orc_colorize(orc_state* s)
{
    size_t N = (size_t)s->res_x * s->res_y;
    for (size_t i = 0; i < N; i++) s->image[i] = orc_colorize_word(s->screen[i]);
}

/* raycast.h:363-366 with the 16-rounding of ocl_begin (ocl.h:188-198,234) */
void orc_tile_schedule(int res_x, int res_y, int frame, int* add_x, int* add_y,
                       int* ext_x, int* ext_y)
{
    *add_x = (res_x / 8) * (frame & 7);
    *add_y = (res_y / 4) * ((frame >> 3) & 3);
    int gx = res_x / 8, gy = res_y / 4;
    *ext_x = (gx % 16) ? gx + 16 - gx % 16 : gx;
    *ext_y = (gy % 16) ? gy + 16 - gy % 16 : gy;
}

/* raycast.h:147-438 */
void orc_frame_warp(orc_state* s, const orc_params* p, int inplace_filter)
{
    size_t N = (size_t)s->res_x * s->res_y;
    if (p->frame < 2) orc_memset_screen(s, 0, SCREEN_HOLE, 4 * N);
    orc_memset_screen(s, 0, SCREEN_HOLE, N);
    orc_proj(s, 1, p);
    orc_proj(s, 2, p);
    orc_counthole(s);
    orc_sumids(s);
    orc_writeids(s);
    if (s->ids[0] > 0) orc_raycast_holes(s, p);
    int ax, ay, ex, ey;
    orc_tile_schedule(s->res_x, s->res_y, p->frame, &ax, &ay, &ex, &ey);
    orc_raycast_rect(s, p, ax, ay, ex, ey);
    orc_copy_plane(s, 2);
    if (inplace_filter) orc_fillhole2_inplace(s);
    else orc_fillhole2_snapshot(s, s->screen + 2 * N);
    orc_colorize(s);
}

/* Gated warp frame (no reference counterpart).  After hole compaction, the
 * residual holes of a 16x16 tile (partial tiles included) are its hole pixels
 * that neither the hole list (all-hole 2x2 quads of full 16x16 blocks) nor the
 * refresh rectangle will trace; a tile with more than max_holes of them has
 * those pixels raycast along with the hole and tile rays.  Rays write only
 * their own pixel, so their order does not matter.  The CNN part of the
 * learned fill is not restated here. */
void orc_frame_warp_gated(orc_state* s, const orc_params* p, int max_holes, float max_depth)
{
    const int rx = s->res_x, ry = s->res_y;
    size_t N = (size_t)rx * ry;
    if (p->frame < 2) orc_memset_screen(s, 0, SCREEN_HOLE, 4 * N);
    orc_memset_screen(s, 0, SCREEN_HOLE, N);
    orc_proj(s, 1, p);
    orc_proj(s, 2, p);
    orc_counthole(s);
    orc_sumids(s);
    orc_writeids(s);
    int ax, ay, ex, ey;
    orc_tile_schedule(rx, ry, p->frame, &ax, &ay, &ex, &ey);
    /* residual holes: hole pixels that neither the hole list (all-hole 2x2
     * quads of the full 16x16 blocks) nor the refresh rectangle traces */
    const int fx = (rx / 16) * 16, fy = (ry / 16) * 16;
    uint8_t* resid = (uint8_t*)calloc(N, 1);
    for (int y = 0; y < ry; y++)
        for (int x = 0; x < rx; x++) {
            size_t o = (size_t)y * rx + x;
            if (s->screen[o] != SCREEN_HOLE) continue;
            if (x < fx && y < fy) {
                size_t q = (size_t)(y & ~1) * rx + (x & ~1);
                if (s->screen[q] == SCREEN_HOLE && s->screen[q + 1] == SCREEN_HOLE && s->screen[q + rx] == SCREEN_HOLE &&
                    s->screen[q + rx + 1] == SCREEN_HOLE)
                    continue;
            }
            if (x >= ax && x < ax + ex && y >= ay && y < ay + ey) continue;
            resid[o] = 1;
        }
    /* the depth rule: reprojected pixels outside the refresh rectangle, and
     * whether one lies deeper than max_depth (camera depth = coordinate w) */
    uint8_t* proj = (uint8_t*)calloc(N, 1);
    for (int y = 0; max_depth > 0.0f && y < ry; y++)
        for (int x = 0; x < rx; x++) {
            size_t o = (size_t)y * rx + x;
            if (s->screen[o] == SCREEN_HOLE) continue;
            if (x >= ax && x < ax + ex && y >= ay && y < ay + ey) continue;
            proj[o] = s->coords[4 * o + 3] > max_depth ? 2 : 1;
        }
    int tx = (rx + 15) / 16, ty = (ry + 15) / 16;
    uint32_t* list = (uint32_t*)malloc(N * 4);
    size_t n = 0;
    for (int t = 0; t < tx * ty; t++) {
        int x0 = (t % tx) * 16, y0 = (t / tx) * 16, h = 0, far = 0;
        for (int j = 0; j < 16; j++)
            for (int i = 0; i < 16; i++)
                if (x0 + i < rx && y0 + j < ry) {
                    size_t o = (size_t)(y0 + j) * rx + x0 + i;
                    h += resid[o];
                    far += proj[o] == 2;
                }
        if (!far && h <= max_holes) continue;  /* CNN tile (or nothing to do) */
        for (int j = 0; j < 16; j++)
            for (int i = 0; i < 16; i++) {
                if (x0 + i >= rx || y0 + j >= ry) continue;
                size_t o = (size_t)(y0 + j) * rx + x0 + i;
                if (resid[o] || proj[o] == 2)
                    list[n++] = (uint32_t)(x0 + i) | ((uint32_t)(y0 + j) << 16);
            }
    }
    free(proj);
    free(resid);
    if (s->ids[0] > 0) orc_raycast_holes(s, p);
    orc_raycast_rect(s, p, ax, ay, ex, ey);
    for (size_t k = 0; k < n; k++) {
        int x = (int)(list[k] & 0xffff), y = (int)(list[k] >> 16);
        float xyz[3];
        uint32_t w = orc_trace_pixel(s->oct, s->n_words, s->root, s->depth, rx, ry, p, x, y, xyz, &s->st);
        write_ray(s, (size_t)y * rx + x, w, xyz);
    }
    free(list);
    orc_copy_plane(s, 2);
    orc_fillhole2_snapshot(s, s->screen + 2 * N);
    orc_colorize(s);
}

void orc_frame_full(orc_state* s, const orc_params* p)
{
    orc_raycast_rect(s, p, 0, 0, s->res_x, s->res_y);
    orc_colorize(s);
}
