// svo_host.cpp -- host side of libsvo: octree builder emitting the reference
// compact word layout, seeded synthetic scenes, camera math and flythrough
// paths, octree save/load, thread-local error reporting.
//
// Builder design (own): the voxel stream is radix-sorted by Morton key with
// the stream index kept stable, leaves are de-duplicated (last write wins, as
// set_voxel overwrites the leaf slot, octree.h:67-70), every node's colour is
// the colour of the earliest voxel in its subtree (the voxel that created it,
// octree.h:46-48,:75-81), blocks (the depth D-6 subtrees) are emitted in
// parallel into private buffers with block-relative refs, then the 10-word
// normal region is written in post-order and blocks are placed at 64-word
// aligned offsets in DFS order (octree.h:232-293).
#include "svo_internal.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

// ---------------------------------------------------------------- errors
static thread_local std::string g_err;

int svo_fail(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

extern "C" const char* svo_last_error(void) { return g_err.c_str(); }
extern "C" void svo_free(void* p) { free(p); }

static int popc8(uint32_t v) { return __builtin_popcount(v & 0xffu); }

int svo_host_threads()
{
    unsigned n = std::thread::hardware_concurrency();
    if (n == 0) n = 4;
    if (n > 16) n = 16;
    return (int)n;
}

template <class F>
static void parallel_for(size_t n, F&& f)
{
    int nt = svo_host_threads();
    if (n < 2048 || nt <= 1) {
        for (size_t i = 0; i < n; i++) f(i);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::thread> th;
    const size_t chunk = 256;
    for (int t = 0; t < nt; t++)
        th.emplace_back([&]() {
            for (;;) {
                size_t b = next.fetch_add(chunk);
                if (b >= n) break;
                size_t e = std::min(n, b + chunk);
                for (size_t i = b; i < e; i++) f(i);
            }
        });
    for (auto& t : th) t.join();
}

// ---------------------------------------------------------------- builder
namespace {

struct Leaf {
    uint64_t key;      // Morton key, 3 bits per level, root level first
    uint32_t color;    // last write
    uint32_t first;    // smallest stream index at this position
};

struct Builder {
    int D;
    std::vector<Leaf> leaves;
    const uint32_t* rgba;

    int digit(uint64_t key, int d) const { return (int)((key >> (3 * (D - 1 - d))) & 7); }

    // children of the node at depth d spanning leaves [b,e): split points
    int split(size_t b, size_t e, int d, size_t cb[8], size_t ce[8]) const
    {
        int mask = 0;
        size_t i = b;
        for (int j = 0; j < 8; j++) {
            size_t s = i;
            while (i < e && digit(leaves[i].key, d) == j) i++;
            cb[j] = s;
            ce[j] = i;
            if (i > s) mask |= 1 << j;
        }
        return mask;
    }

    // byte node at depth D-2 (octree.h:106-157); returns its min stream index
    uint32_t emit_bytes(std::vector<uint32_t>& out, size_t b, size_t e, uint32_t* ref)
    {
        const int d = D - 2;
        size_t cb[8], ce[8];
        int mask = split(b, e, d, cb, ce);
        uint32_t first = UINT32_MAX;
        for (size_t i = b; i < e; i++) first = std::min(first, leaves[i].first);
        *ref = (uint32_t)((out.size() << 9) + 256 + (uint32_t)mask);
        uint8_t blk[1 + 8 + 64];
        int size = 0;
        blk[size++] = (uint8_t)rgba[first];
        size_t gb[8][8], ge[8][8];
        int cmask[8];
        for (int j = 0; j < 8; j++) {
            if (!(mask >> j & 1)) continue;
            cmask[j] = split(cb[j], ce[j], d + 1, gb[j], ge[j]);
            blk[size++] = (uint8_t)cmask[j];
        }
        for (int j = 0; j < 8; j++) {
            if (!(mask >> j & 1)) continue;
            for (int i = 0; i < 8; i++) {
                if (!(cmask[j] >> i & 1)) continue;
                uint32_t c = leaves[gb[j][i]].color;
                if (c > 0) blk[size++] = (uint8_t)c;  // octree.h:133-135
            }
        }
        uint32_t dw = 0;
        int i03 = 0;
        for (int i = 0; i < size; i++) {
            i03 = i & 3;
            dw |= (uint32_t)blk[i] << (i03 * 8);
            if (i03 == 3) { out.push_back(dw); dw = 0; }
        }
        if (i03 < 3) out.push_back(dw);
        return first;
    }

    // compact node at depth d in [D-6, D-3] (octree.h:166-190)
    uint32_t emit_compact(std::vector<uint32_t>& out, size_t b, size_t e, int d, uint32_t* ref)
    {
        size_t cb[8], ce[8];
        int mask = split(b, e, d, cb, ce);
        uint32_t refs[8] = {0};
        uint32_t first = UINT32_MAX;
        for (int j = 0; j < 8; j++) {
            if (!(mask >> j & 1)) continue;
            uint32_t f = (d + 1 == D - 2) ? emit_bytes(out, cb[j], ce[j], &refs[j])
                                          : emit_compact(out, cb[j], ce[j], d + 1, &refs[j]);
            first = std::min(first, f);
        }
        *ref = (uint32_t)((out.size() << 9) + 256 + (uint32_t)mask);
        out.push_back(rgba[first]);
        for (int j = 0; j < 8; j++)
            if (mask >> j & 1) out.push_back(refs[j]);
        return first;
    }

    // block rooted at depth D-6 (octree.h:252-276), block-relative words
    uint32_t emit_block(std::vector<uint32_t>& out, size_t b, size_t e, int* mask_out)
    {
        size_t cb[8], ce[8];
        int mask = split(b, e, D - 6, cb, ce);
        int num_entries = popc8((uint32_t)mask) + 1;
        out.assign((size_t)num_entries, 0u);
        uint32_t sub;
        uint32_t first = emit_compact(out, b, e, D - 6, &sub);
        uint32_t sr = sub >> 9;
        for (int i = 0; i < num_entries; i++) out[i] = out[sr + i];
        out.resize(sr);
        *mask_out = mask;
        return first;
    }
};

struct BlockOut {
    std::vector<uint32_t> words;
    uint32_t first;
    int mask;
    size_t pos;
};

}  // namespace

static void radix_sort_pairs(std::vector<uint64_t>& key, std::vector<uint32_t>& idx, int bits)
{
    size_t n = key.size();
    std::vector<uint64_t> k2(n);
    std::vector<uint32_t> i2(n);
    const int R = 11;
    for (int shift = 0; shift < bits; shift += R) {
        std::vector<size_t> cnt((size_t)1 << R, 0);
        for (size_t i = 0; i < n; i++) cnt[(key[i] >> shift) & ((1u << R) - 1)]++;
        size_t s = 0;
        for (auto& c : cnt) { size_t t = c; c = s; s += t; }
        for (size_t i = 0; i < n; i++) {
            size_t p = cnt[(key[i] >> shift) & ((1u << R) - 1)]++;
            k2[p] = key[i];
            i2[p] = idx[i];
        }
        key.swap(k2);
        idx.swap(i2);
    }
}

extern "C" int svo_build_octree(const uint16_t* xyz, const uint32_t* rgba, size_t n, int depth,
                                uint32_t** words_out, size_t* n_words, uint32_t* root_ref)
{
    if (!words_out || !n_words || !root_ref) return svo_fail(SVO_ERR_ARG, "null output pointer");
    if (depth < 8 || depth > 13) return svo_fail(SVO_ERR_ARG, "depth %d outside 8..13", depth);
    if (n > 0xfffffff0ull) return svo_fail(SVO_ERR_ARG, "too many voxels");
    const int D = depth;
    const uint32_t lim = 1u << D;

    // Morton keys, stream index as the stable payload
    std::vector<uint64_t> key(n);
    std::vector<uint32_t> idx(n);
    for (size_t i = 0; i < n; i++) {
        uint32_t x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        if (x >= lim || y >= lim || z >= lim)
            return svo_fail(SVO_ERR_ARG, "voxel %zu (%u,%u,%u) outside 2^%d", i, x, y, z, D);
        uint64_t k = 0;
        for (int bit = D - 1; bit >= 0; bit--)
            k = (k << 3) | (((x >> bit) & 1) | (((y >> bit) & 1) << 1) | (((z >> bit) & 1) << 2));
        key[i] = k;
        idx[i] = (uint32_t)i;
    }
    radix_sort_pairs(key, idx, 3 * D);

    Builder B;
    B.D = D;
    B.rgba = rgba;
    B.leaves.reserve(n);
    for (size_t i = 0; i < n;) {
        size_t j = i;
        while (j + 1 < n && key[j + 1] == key[i]) j++;
        Leaf L;
        L.key = key[i];
        L.color = rgba[idx[j]];  // last write at this position
        L.first = idx[i];        // earliest voxel at this position
        B.leaves.push_back(L);
        i = j + 1;
    }
    std::vector<uint64_t>().swap(key);
    std::vector<uint32_t>().swap(idx);

    // blocks = groups of leaves sharing the top 3*(D-6) key bits
    const int bshift = 3 * 6;
    std::vector<size_t> bstart;
    for (size_t i = 0; i < B.leaves.size(); i++)
        if (i == 0 || (B.leaves[i].key >> bshift) != (B.leaves[i - 1].key >> bshift)) bstart.push_back(i);
    size_t nb = bstart.size();
    bstart.push_back(B.leaves.size());
    std::vector<BlockOut> blocks(nb);
    parallel_for(nb, [&](size_t b) {
        blocks[b].first = B.emit_block(blocks[b].words, bstart[b], bstart[b + 1], &blocks[b].mask);
    });

    // block placement: 64-aligned, DFS order, after the 2^21-word normal region
    size_t pos = 2097152;
    for (size_t b = 0; b < nb; b++) {
        pos = (pos + 63) & ~(size_t)63;
        blocks[b].pos = pos;
        pos += blocks[b].words.size();
        if ((pos >> 6) >= (1u << 23)) return svo_fail(SVO_ERR_ARG, "octree exceeds 2^29 words");
    }
    size_t total = (nb ? pos : 2097152);
    uint32_t* out = (uint32_t*)calloc(total, 4);
    if (!out) return svo_fail(SVO_ERR_NOMEM, "octree output %zu words", total);
    parallel_for(nb, [&](size_t b) {
        memcpy(out + blocks[b].pos, blocks[b].words.data(), blocks[b].words.size() * 4);
    });

    // normal region, post-order (octree.h:232-293); root colour is 0
    uint32_t normal_ofs = 0;
    size_t next_block = 0;
    bool overflow = false;
    struct Rec {
        static uint32_t go(Builder& B, std::vector<BlockOut>& blocks, size_t& next_block,
                           uint32_t* out, uint32_t& normal_ofs, bool& overflow, size_t b,
                           size_t e, int d, uint32_t* first_out)
        {
            size_t cb[8], ce[8];
            int mask = B.split(b, e, d, cb, ce);
            uint32_t refs[8] = {0};
            uint32_t first = UINT32_MAX;
            for (int j = 0; j < 8; j++) {
                if (!(mask >> j & 1)) continue;
                uint32_t f;
                if (d + 1 <= B.D - 7) {
                    refs[j] = go(B, blocks, next_block, out, normal_ofs, overflow, cb[j], ce[j],
                                 d + 1, &f);
                } else {
                    BlockOut& bo = blocks[next_block++];
                    refs[j] = (uint32_t)(((bo.pos >> 6) << 9) | (1u << 8) | (uint32_t)bo.mask);
                    f = bo.first;
                }
                first = std::min(first, f);
            }
            *first_out = first;
            if (normal_ofs + 10 > 2097152u) { overflow = true; return 0; }
            uint32_t color = (d == 0) ? 0u : B.rgba[first];
            for (int j = 0; j < 8; j++) out[normal_ofs + j] = refs[j];
            out[normal_ofs + 8] = color;
            out[normal_ofs + 9] = color;
            normal_ofs += 10;
            return ((normal_ofs - 10) << 9) | (uint32_t)mask;
        }
    };
    uint32_t root = 0, first;
    if (!B.leaves.empty())
        root = Rec::go(B, blocks, next_block, out, normal_ofs, overflow, 0, B.leaves.size(), 0, &first);
    else {
        // empty scene: a root record with no children at offset 0
        root = 0;
        normal_ofs = 10;
    }
    if (overflow) {
        free(out);
        return svo_fail(SVO_ERR_ARG, "normal region exceeds 2^21 words");
    }
    *words_out = out;
    *n_words = total;
    *root_ref = root;
    return SVO_OK;
}

// ---------------------------------------------------------------- device layout
// Validates every ref of a compact word array (octree.h:232-293 layout) and
// renumbers its 10-word normal records in breadth-first order (root at word
// 0, then depth 1, ...).  Normal-node refs are only ever used to address
// their own records (kernel.cl:42,:61,:74,:111); the words a ray returns are
// colour words and block-relative refs, which do not move, so every output is
// unchanged while the top levels become a prefix of the array that the ray
// kernel keeps in LDS.
namespace {
struct Validator {
    const uint32_t* w;
    size_t n;
    int D;
    bool ok = true;
    std::string why;
    bool in(size_t i) { return i < n; }
    void fail(const char* m, size_t at)
    {
        if (ok) {
            char b[160];
            snprintf(b, sizeof(b), "%s (word %zu)", m, at);
            why = b;
        }
        ok = false;
    }
    // compact node at depth d (D-6 .. D-3) at absolute word a, or byte node
    void node(size_t local_root, uint32_t ref, int d)
    {
        if (!ok) return;
        size_t a = local_root + (ref >> 9);
        int mask = (int)(ref & 255u);
        if (!(ref & 256u)) { fail("block-internal ref without the block bit", a); return; }
        if (d == D - 2) {  // byte node: colour, masks, leaf colours
            size_t nb = 1 + (size_t)popc8((uint32_t)mask);
            if (!in(a + (nb + 3) / 4)) { fail("byte node past the array", a); return; }
            size_t leaves = 0;
            for (size_t i = 1; i < nb; i++) leaves += popc8((w[a + (i >> 2)] >> ((i & 3) * 8)) & 255u);
            if (!in(a + (nb + leaves + 3) / 4 - 1)) fail("byte node leaves past the array", a);
            return;
        }
        int k = popc8((uint32_t)mask);
        if (!in(a + (size_t)k)) { fail("compact node past the array", a); return; }
        for (int j = 0; j < k && ok; j++) node(local_root, w[a + 1 + j], d + 1);
    }
};
}  // namespace

int svo_device_layout(const uint32_t* words, size_t n_words, uint32_t root, int depth,
                      std::vector<uint32_t>& out, uint32_t* new_root, uint32_t* normal_words)
{
    const size_t nrec_max = std::min<size_t>(n_words, 2097152) / 10;
    std::vector<int32_t> newidx(nrec_max + 1, -1);
    std::vector<uint32_t> order;
    if ((root & 256u) || (root >> 9) % 10 != 0 || (root >> 9) / 10 >= nrec_max)
        return svo_fail(SVO_ERR_ARG, "root ref 0x%x is not a normal record", root);
    Validator V{words, n_words, depth};
    order.push_back(root >> 9);
    newidx[(root >> 9) / 10] = 0;
    std::vector<int> level(1, 0);
    for (size_t q = 0; q < order.size(); q++) {
        uint32_t o = order[q];
        for (int j = 0; j < 8; j++) {
            uint32_t v = words[o + j];
            if (!v) continue;
            if (v & 256u) {
                size_t lr = (size_t)(v >> 9) << 6;
                if (level[q] + 1 != depth - 6)
                    return svo_fail(SVO_ERR_ARG, "block ref at depth %d (expected %d)", level[q] + 1, depth - 6);
                if (lr < (size_t)10 * nrec_max && lr < 2097152)
                    return svo_fail(SVO_ERR_ARG, "block at word %zu overlaps the normal region", lr);
                // block root record sits at relative offset 0
                V.node(lr, (v & 255u) | 256u, depth - 6);
                if (!V.ok) return svo_fail(SVO_ERR_ARG, "invalid octree: %s", V.why.c_str());
                continue;
            }
            uint32_t c = v >> 9;
            if (c % 10 != 0 || c / 10 >= nrec_max)
                return svo_fail(SVO_ERR_ARG, "normal ref 0x%x out of range", v);
            if (level[q] + 1 > depth - 7)
                return svo_fail(SVO_ERR_ARG, "normal node below depth %d", depth - 7);
            if (newidx[c / 10] < 0) {
                newidx[c / 10] = (int32_t)order.size();
                order.push_back(c);
                level.push_back(level[q] + 1);
            }
        }
    }
    out.assign(words, words + n_words);
    for (size_t q = 0; q < order.size(); q++) {
        const uint32_t* src = words + order[q];
        uint32_t* dst = out.data() + 10 * q;
        for (int j = 0; j < 8; j++) {
            uint32_t v = src[j];
            if (v && !(v & 256u)) v = ((uint32_t)(10 * newidx[(v >> 9) / 10]) << 9) | (v & 255u);
            dst[j] = v;
        }
        dst[8] = src[8];
        dst[9] = src[9];
    }
    *new_root = root & 255u;  // record 0
    *normal_words = (uint32_t)(10 * order.size());
    return SVO_OK;
}

// ---------------------------------------------------------------- save/load
static const char kMagic[8] = {'S', 'V', 'O', 'W', 'O', 'R', 'D', '1'};

extern "C" int svo_octree_save(const char* path, const uint32_t* words, size_t n_words,
                               uint32_t root_ref, int depth)
{
    if (!path || (!words && n_words)) return svo_fail(SVO_ERR_ARG, "null argument");
    FILE* f = fopen(path, "wb");
    if (!f) return svo_fail(SVO_ERR_IO, "cannot open %s for writing", path);
    uint64_t hdr[3] = {(uint64_t)n_words, (uint64_t)root_ref, (uint64_t)depth};
    bool ok = fwrite(kMagic, 1, 8, f) == 8 && fwrite(hdr, 8, 3, f) == 3 &&
              fwrite(words, 4, n_words, f) == n_words;
    ok = (fclose(f) == 0) && ok;
    return ok ? SVO_OK : svo_fail(SVO_ERR_IO, "short write to %s", path);
}

extern "C" int svo_octree_load(const char* path, uint32_t** words, size_t* n_words,
                               uint32_t* root_ref, int* depth)
{
    if (!path || !words || !n_words || !root_ref || !depth) return svo_fail(SVO_ERR_ARG, "null argument");
    FILE* f = fopen(path, "rb");
    if (!f) return svo_fail(SVO_ERR_IO, "cannot open %s", path);
    char m[8];
    uint64_t hdr[3];
    if (fread(m, 1, 8, f) != 8 || memcmp(m, kMagic, 8) != 0 || fread(hdr, 8, 3, f) != 3) {
        fclose(f);
        return svo_fail(SVO_ERR_IO, "%s is not an SVO word file", path);
    }
    if (hdr[0] > (1ull << 29) || hdr[2] < 8 || hdr[2] > 13) {
        fclose(f);
        return svo_fail(SVO_ERR_IO, "%s: bad header", path);
    }
    uint32_t* w = (uint32_t*)malloc(hdr[0] * 4 + 4);
    if (!w) { fclose(f); return svo_fail(SVO_ERR_NOMEM, "load %zu words", (size_t)hdr[0]); }
    if (fread(w, 4, hdr[0], f) != hdr[0]) {
        fclose(f);
        free(w);
        return svo_fail(SVO_ERR_IO, "%s: truncated", path);
    }
    fclose(f);
    *words = w;
    *n_words = (size_t)hdr[0];
    *root_ref = (uint32_t)hdr[1];
    *depth = (int)hdr[2];
    return SVO_OK;
}

// ---------------------------------------------------------------- camera
// raycast.h:113-162, :322-326 with the nebula mathlib semantics
// (_matrix44.h:529-580 rotations, :334-344 transpose, :875-888 m*v;
// nmath.h:33-34 sin/cos evaluated in double then rounded to float).
namespace {
struct M44 { float m[4][4]; };
void rot_x(M44& a, float ang)
{
    float c = (float)cos((double)ang), s = (float)sin((double)ang);
    for (int i = 0; i < 4; i++) {
        float a1 = a.m[i][1], a2 = a.m[i][2];
        a.m[i][1] = a1 * c + a2 * -s;
        a.m[i][2] = a1 * s + a2 * c;
    }
}
void rot_y(M44& a, float ang)
{
    float c = (float)cos((double)ang), s = (float)sin((double)ang);
    for (int i = 0; i < 4; i++) {
        float a0 = a.m[i][0], a2 = a.m[i][2];
        a.m[i][0] = a0 * c + a2 * s;
        a.m[i][2] = a0 * -s + a2 * c;
    }
}
void rot_z(M44& a, float ang)
{
    float c = (float)cos((double)ang), s = (float)sin((double)ang);
    for (int i = 0; i < 4; i++) {
        float a0 = a.m[i][0], a1 = a.m[i][1];
        a.m[i][0] = a0 * c + a1 * -s;
        a.m[i][1] = a0 * s + a1 * c;
    }
}
void mul4(const M44& a, float x, float y, float z, float w, float o[4])
{
    for (int j = 0; j < 4; j++)
        o[j] = a.m[0][j] * x + a.m[1][j] * y + a.m[2][j] * z + a.m[3][j] * w;
}
}  // namespace

extern "C" int svo_camera(float yaw, float pitch, float roll, const float pos_world[3], int depth,
                          int frame, svo_frame_params* out, float pos_wrapped[3])
{
    if (!pos_world || !out) return svo_fail(SVO_ERR_ARG, "null argument");
    if (depth < 1 || depth > 20) return svo_fail(SVO_ERR_ARG, "bad depth");
    const float dim2 = (float)((1 << depth) * 2);
    float p[3];
    for (int i = 0; i < 3; i++) {
        if (!std::isfinite(pos_world[i])) return svo_fail(SVO_ERR_ARG, "non-finite position");
        p[i] = pos_world[i] * 16.0f;
        while (p[i] < 0) p[i] += dim2;
        while (p[i] >= dim2) p[i] -= dim2;
    }
    const float inv16 = 1.0f / 16.0f;
    float pw[3] = {p[0] * inv16, p[1] * inv16, p[2] * inv16};
    if (pos_wrapped)
        for (int i = 0; i < 3; i++) pos_wrapped[i] = pw[i];
    M44 m;
    memset(&m, 0, sizeof(m));
    for (int i = 0; i < 4; i++) m.m[i][i] = 1.0f;
    rot_z(m, roll);
    rot_x(m, pitch);
    rot_y(m, yaw);
    out->m0[0] = pw[0]; out->m0[1] = pw[1]; out->m0[2] = pw[2]; out->m0[3] = 1.0f;
    mul4(m, 1, 0, 0, 0, out->row0);
    mul4(m, 0, 1, 0, 0, out->row1);
    mul4(m, 0, 0, 1, 0, out->row2);
    M44 t;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) t.m[i][j] = m.m[j][i];
    mul4(t, 1, 0, 0, 0, out->col0);
    mul4(t, 0, 1, 0, 0, out->col1);
    mul4(t, 0, 0, 1, 0, out->col2);
    out->fovx = 1.0f;
    out->fovy = 1.0f;
    out->frame = frame;
    out->reserved = 0;
    return SVO_OK;
}

extern "C" int svo_tile_schedule(int res_x, int res_y, int frame, int* add_x, int* add_y,
                                 int* ext_x, int* ext_y)
{
    if (res_x <= 0 || res_y <= 0 || !add_x || !add_y || !ext_x || !ext_y)
        return svo_fail(SVO_ERR_ARG, "bad tile schedule arguments");
    *add_x = (res_x / 8) * (frame & 7);
    *add_y = (res_y / 4) * ((frame >> 3) & 3);
    *ext_x = (res_x / 8 + 15) / 16 * 16;
    *ext_y = (res_y / 4 + 15) / 16 * 16;
    return SVO_OK;
}

// ---------------------------------------------------------------- scenes
namespace {
inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
// PCG32 (O'Neill, pcg32_random_r) for sequential draws
struct Pcg32 {
    uint64_t state, inc;
    explicit Pcg32(uint64_t seed, uint64_t seq = 54u)
    {
        state = 0;
        inc = (seq << 1u) | 1u;
        next();
        state += seed;
        next();
    }
    uint32_t next()
    {
        uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31));
    }
    double uniform() { return next() * (1.0 / 4294967296.0); }
};
inline double lattice(uint64_t seed, int oct, int64_t a, int64_t b, int64_t c)
{
    uint64_t h = mix64(seed * 0x9E3779B97F4A7C15ull ^ mix64((uint64_t)a * 0xD1B54A32D192ED03ull ^
                                                            (uint64_t)b * 0xABC98388FB8FAC03ull ^
                                                            (uint64_t)c * 0x8CB92BA72F3D8DD7ull ^
                                                            (uint64_t)(oct + 1) * 0x94D049BB133111EBull));
    return (double)(h >> 40) * (1.0 / 16777216.0);
}
inline double fade(double t) { return t * t * (3.0 - 2.0 * t); }
// periodic 2-D value noise, lattice spacing L voxels, period P voxels
double vnoise2(uint64_t seed, int oct, double x, double z, int L, int P)
{
    int n = P / L;
    double fx = x / L, fz = z / L;
    int64_t ix = (int64_t)floor(fx), iz = (int64_t)floor(fz);
    double tx = fade(fx - ix), tz = fade(fz - iz);
    auto w = [&](int64_t a, int64_t b) {
        return lattice(seed, oct, ((a % n) + n) % n, ((b % n) + n) % n, 0);
    };
    double v00 = w(ix, iz), v10 = w(ix + 1, iz), v01 = w(ix, iz + 1), v11 = w(ix + 1, iz + 1);
    return (v00 * (1 - tx) + v10 * tx) * (1 - tz) + (v01 * (1 - tx) + v11 * tx) * tz;
}
double vnoise3(uint64_t seed, int oct, double x, double y, double z, int L, int P)
{
    int n = P / L;
    double fx = x / L, fy = y / L, fz = z / L;
    int64_t ix = (int64_t)floor(fx), iy = (int64_t)floor(fy), iz = (int64_t)floor(fz);
    double tx = fade(fx - ix), ty = fade(fy - iy), tz = fade(fz - iz);
    auto w = [&](int64_t a, int64_t b, int64_t c) {
        return lattice(seed, oct, ((a % n) + n) % n, ((b % n) + n) % n, ((c % n) + n) % n) * 2.0 - 1.0;
    };
    double r = 0;
    for (int k = 0; k < 8; k++) {
        int dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
        double wt = (dx ? tx : 1 - tx) * (dy ? ty : 1 - ty) * (dz ? tz : 1 - tz);
        r += wt * w(ix + dx, iy + dy, iz + dz);
    }
    return r;
}
inline int clamp255(int v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }
inline int jitter(uint64_t seed, int salt, int x, int y, int z, int amp)
{
    uint64_t h = mix64(seed ^ mix64(((uint64_t)(uint32_t)x << 40) ^ ((uint64_t)(uint32_t)y << 20) ^
                                    (uint64_t)(uint32_t)z ^ ((uint64_t)salt << 60)));
    return (int)(h % (uint64_t)(2 * amp + 1)) - amp;
}
// reference colour byte: bits 0-1 palette, bits 2-7 intensity (Rle4.cpp:145-152)
inline uint32_t pack_rgba(int luma, int palette, int g, int b)
{
    uint32_t x = (uint32_t)((clamp255(luma) & 0xfc) | (palette & 3));
    if (x == 0) x = 4;  // never emit an all-zero colour word (octree.h:122 tests > 0)
    return x | ((uint32_t)clamp255(g) << 8) | ((uint32_t)clamp255(b) << 16);
}

struct VoxOut {
    std::vector<uint16_t> xyz;
    std::vector<uint32_t> rgba;
    void add(int x, int y, int z, uint32_t c)
    {
        xyz.push_back((uint16_t)x);
        xyz.push_back((uint16_t)y);
        xyz.push_back((uint16_t)z);
        rgba.push_back(c);
    }
};

int emit_voxels(VoxOut& vo, uint16_t** xyz, uint32_t** rgba, size_t* n)
{
    const size_t nv = vo.rgba.size();
    uint16_t* a = (uint16_t*)malloc(nv * 6 + 6);
    uint32_t* c = (uint32_t*)malloc(nv * 4 + 4);
    if (!a || !c) {
        free(a);
        free(c);
        return svo_fail(SVO_ERR_NOMEM, "voxel stream of %zu voxels", nv);
    }
    if (nv) {
        memcpy(a, vo.xyz.data(), nv * 6);
        memcpy(c, vo.rgba.data(), nv * 4);
    }
    *xyz = a;
    *rgba = c;
    *n = nv;
    return SVO_OK;
}

// heightfield shell: columns in z -> x -> y order (like Rle4.cpp:95-123)
void heightfield(VoxOut& vo, const std::vector<int>& H, int size, uint64_t seed, int hmin, int hmax, int jamp)
{
    auto h = [&](int x, int z) { return H[(size_t)((z + size) % size) * size + (x + size) % size]; };
    for (int z = 0; z < size; z++)
        for (int x = 0; x < size; x++) {
            int top = h(x, z);
            int lo = std::min(top, std::min(std::min(h(x - 1, z), h(x + 1, z)), std::min(h(x, z - 1), h(x, z + 1)))) - 1;
            if (lo < 0) lo = 0;
            for (int y = lo; y <= top; y++) {
                int luma = 40 + (y - hmin) * 215 / std::max(1, hmax - hmin) + jitter(seed, 1, x, y, z, jamp);
                int g = luma * 3 / 4 + jitter(seed, 2, x, y, z, jamp);
                int b = luma / 2 + jitter(seed, 3, x, y, z, jamp);
                vo.add(x, y, z, pack_rgba(luma, 1, g, b));
            }
        }
}
}  // namespace

extern "C" int svo_scene_generate(int kind, uint64_t seed, int depth, int n_spheres, uint16_t** xyz,
                                  uint32_t** rgba, size_t* n)
{
    if (!xyz || !rgba || !n) return svo_fail(SVO_ERR_ARG, "null output pointer");
    if (depth < 8 || depth > 13) return svo_fail(SVO_ERR_ARG, "depth %d outside 8..13", depth);
    const int size = 1 << depth;
    VoxOut vo;
    if (kind == 1 || kind == 2 || kind == 5) {
        // kind 5: kind 2's terrain and spheres with no per-voxel colour jitter
        const int jamp = kind == 5 ? 0 : 8;
        if (kind == 5) kind = 2;
        int octaves = (kind == 1) ? 2 : 4;
        int L0 = (kind == 1) ? size / 4 : size / 8;
        int hmin = (kind == 1) ? size / 8 : size / 16;
        int hspan = (kind == 1) ? size * 3 / 4 : size / 2;
        std::vector<int> H((size_t)size * size);
        parallel_for((size_t)size, [&](size_t zi) {
            int z = (int)zi;
            for (int x = 0; x < size; x++) {
                double v = 0, amp = 1, norm = 0;
                int L = L0;
                for (int o = 0; o < octaves; o++) {
                    v += amp * vnoise2(seed, o, x, z, L, size);
                    norm += amp;
                    amp *= 0.5;
                    L /= 2;
                }
                v /= norm;
                H[(size_t)z * size + x] = hmin + (int)floor(v * hspan);
            }
        });
        heightfield(vo, H, size, seed, hmin, hmin + hspan, jamp);
        if (kind == 2) {
            Pcg32 rng(seed, 7);
            for (int s = 0; s < n_spheres; s++) {
                int r = 4 + (int)(rng.uniform() * 61.0);  // U[4,64]
                int cx = (int)(rng.uniform() * size), cz = (int)(rng.uniform() * size);
                int ground = H[(size_t)cz * size + cx];
                int cy = ground + r + (int)(rng.uniform() * (size / 16));
                if (cy + r >= size) cy = size - 1 - r;
                int pal = 1 + (int)(rng.next() % 3);
                int cr = (int)(rng.next() & 255), cg = (int)(rng.next() & 255), cb = (int)(rng.next() & 255);
                uint32_t col = pack_rgba(cr, pal, cg, cb);
                long r2 = (long)r * r;
                // boundary voxels of the solid ball (a 6-neighbour lies outside)
                for (int dz = -r; dz <= r; dz++)
                    for (int dx = -r; dx <= r; dx++)
                        for (int dy = -r; dy <= r; dy++) {
                            long d2 = (long)dx * dx + (long)dy * dy + (long)dz * dz;
                            if (d2 > r2) continue;
                            auto out = [&](int a, int b, int c) { return (long)a * a + (long)b * b + (long)c * c > r2; };
                            if (!(out(dx + 1, dy, dz) || out(dx - 1, dy, dz) || out(dx, dy + 1, dz) ||
                                  out(dx, dy - 1, dz) || out(dx, dy, dz + 1) || out(dx, dy, dz - 1)))
                                continue;
                            int x = ((cx + dx) % size + size) % size, z = ((cz + dz) % size + size) % size;
                            int y = cy + dy;
                            if (y < 0 || y >= size) continue;
                            vo.add(x, y, z, col);
                        }
            }
        }
    } else if (kind == 3 || kind == 4) {
        // cave: solid where 5-octave 3-D noise > threshold inside a slab at
        // the bottom of the world (height size/8 for kind 3, the whole height
        // for kind 4); boundary voxels only.
        int slab = kind == 4 ? size : size / 8;
        auto solid = [&](int x, int y, int z) {
            if (y < 0) return true;
            if (y >= slab) return false;
            double v = 0, amp = 1;
            int L = size / 16;
            for (int o = 0; o < 5; o++) {
                v += amp * vnoise3(seed, o, x, y, z, L, size);
                amp *= 0.5;
                L = std::max(2, L / 2);
            }
            return v > 0.12;
        };
        // evaluate per z-slice in parallel, keep stream order z -> x -> y
        std::vector<VoxOut> per(size);
        parallel_for((size_t)size, [&](size_t zi) {
            int z = (int)zi;
            for (int x = 0; x < size; x++)
                for (int y = 0; y < slab; y++) {
                    if (!solid(x, y, z)) continue;
                    int xm = (x + size - 1) % size, xp = (x + 1) % size;
                    int zm = (z + size - 1) % size, zp = (z + 1) % size;
                    if (solid(xm, y, z) && solid(xp, y, z) && solid(x, y - 1, z) && solid(x, y + 1, z) &&
                        solid(x, y, zm) && solid(x, y, zp))
                        continue;
                    int luma = 60 + y * 190 / std::max(1, slab) + jitter(seed, 4, x, y, z, 8);
                    per[z].add(x, y, z, pack_rgba(luma, 2, luma * 2 / 3, luma / 3));
                }
        });
        for (auto& p : per) {
            vo.xyz.insert(vo.xyz.end(), p.xyz.begin(), p.xyz.end());
            vo.rgba.insert(vo.rgba.end(), p.rgba.begin(), p.rgba.end());
        }
    } else {
        return svo_fail(SVO_ERR_ARG, "unknown scene kind %d", kind);
    }
    return emit_voxels(vo, xyz, rgba, n);
}

// ---------------------------------------------------------------- RLE4 volumes
// File layout read by RLE4::load (src/octree/Rle4.cpp:8-165): int32 map
// count, then per map int32 sx, sy, sz, slab count and that many uint16
// slabs.  The slabs hold one record per column, x fastest then z
// (Rle4.cpp:60-90): [n_runs][n_colours][n_runs run words][n_colours colour
// words].  A run word is skip = bits 0-9, length = bits 10-15: the run starts
// `skip` rows after the previous run's end (Rle4.cpp:121-124).  File row r
// is voxel y = sy-1-r.  Only map 0 becomes voxels (Rle4.cpp:111).
namespace {
struct Rle4Map {
    int sx = 0, sy = 0, sz = 0;
    std::vector<uint16_t> slabs;
};

int rle4_read(const char* path, std::vector<Rle4Map>& maps)
{
    FILE* f = fopen(path, "rb");
    if (!f) return svo_fail(SVO_ERR_IO, "cannot open %s", path);
    int32_t nmaps = 0;
    int rc = SVO_OK;
    if (fread(&nmaps, 4, 1, f) != 1 || nmaps < 1 || nmaps > 16) {
        rc = svo_fail(SVO_ERR_IO, "%s: bad map count", path);
    }
    for (int m = 0; rc == SVO_OK && m < nmaps; m++) {
        int32_t hdr[4];
        if (fread(hdr, 4, 4, f) != 4 || hdr[0] < 1 || hdr[1] < 1 || hdr[2] < 1 || hdr[3] < 2 || hdr[0] > 65536 ||
            hdr[1] > 65536 || hdr[2] > 65536) {
            rc = svo_fail(SVO_ERR_IO, "%s: bad header of map %d", path, m);
            break;
        }
        Rle4Map mp;
        mp.sx = hdr[0];
        mp.sy = hdr[1];
        mp.sz = hdr[2];
        mp.slabs.resize((size_t)hdr[3]);
        if (fread(mp.slabs.data(), 2, mp.slabs.size(), f) != mp.slabs.size()) {
            rc = svo_fail(SVO_ERR_IO, "%s: map %d truncated", path, m);
            break;
        }
        maps.push_back(std::move(mp));
    }
    fclose(f);
    return rc;
}

// tcol -> uchar4 of Rle4.cpp:127-160, packed x | y<<8 | z<<16 (w = 0)
uint32_t rle4_color(uint16_t tcol, int palette)
{
    const int lum = tcol & 255;
    int inten = lum * lum / 255;
    inten = std::min(255, std::max(2, inten));
    uint32_t x = (uint32_t)(((tcol >> 8) & 3) + ((inten >> 3) << 3));
    if (!palette) x = 1u + (uint32_t)((255 - lum) & 0xfc);
    const uint32_t y = (uint32_t)((tcol >> 5) << 3) & 255u;
    const uint32_t z = (uint32_t)((tcol >> 10) << 3) & 255u;
    return x | (y << 8) | (z << 16);
}
}  // namespace

extern "C" int svo_rle4_load(const char* path, int palette, int add_x, int add_y, int add_z, int depth,
                             uint16_t** xyz, uint32_t** rgba, size_t* n)
{
    if (!path || !xyz || !rgba || !n) return svo_fail(SVO_ERR_ARG, "null argument");
    if (depth < 1 || depth > 16) return svo_fail(SVO_ERR_ARG, "depth %d outside 1..16", depth);
    std::vector<Rle4Map> maps;
    const int rrc = rle4_read(path, maps);
    if (rrc) return rrc;
    const long size = 1L << depth;
    // column offsets of every map (the walk of Rle4.cpp:66-90)
    std::vector<size_t> col0;
    for (size_t m = 0; m < maps.size(); m++) {
        const Rle4Map& mp = maps[m];
        size_t ofs = 0;
        const size_t cols = (size_t)mp.sx * mp.sz;
        if (m == 0) col0.resize(cols);
        for (size_t c = 0; c < cols; c++) {
            if (ofs + 2 > mp.slabs.size()) return svo_fail(SVO_ERR_IO, "%s: map %zu column %zu past the slabs", path, m, c);
            if (m == 0) col0[c] = ofs;
            ofs += (size_t)mp.slabs[ofs] + mp.slabs[ofs + 1] + 2;
            if (ofs > mp.slabs.size()) return svo_fail(SVO_ERR_IO, "%s: map %zu column %zu past the slabs", path, m, c);
        }
    }
    const Rle4Map& mp = maps[0];
    VoxOut vo;
    for (int z = 0; z < mp.sz; z++)  // slice outer, x inner, rows upward (Rle4.cpp:113-121)
        for (int x = 0; x < mp.sx; x++) {
            const size_t ofs = col0[(size_t)z * mp.sx + x];
            const uint32_t runs = mp.slabs[ofs], cols = mp.slabs[ofs + 1];
            const uint16_t* run = mp.slabs.data() + ofs + 2;
            const uint16_t* col = run + runs;
            uint32_t used = 0;
            uint32_t r0 = 0;
            for (uint32_t k = 0; k < runs; k++) {
                r0 += run[k] & 1023u;
                const uint32_t r1 = r0 + (run[k] >> 10);
                for (; r0 < r1; r0++, used++) {
                    if (used >= cols)
                        return svo_fail(SVO_ERR_IO, "%s: column (%d,%d) has more voxels than colours", path, x, z);
                    if (r0 >= (uint32_t)mp.sy) continue;  // Rle4.cpp:126
                    const long vx = (long)x + add_x, vy = (long)mp.sy - 1 - (long)r0 + add_y, vz = (long)z + add_z;
                    if (vx < 0 || vy < 0 || vz < 0 || vx >= size || vy >= size || vz >= size)
                        return svo_fail(SVO_ERR_ARG, "voxel (%ld,%ld,%ld) outside the depth-%d cube", vx, vy, vz,
                                        depth);
                    vo.add((int)vx, (int)vy, (int)vz, rle4_color(col[used], palette));
                }
            }
        }
    return emit_voxels(vo, xyz, rgba, n);
}

extern "C" int svo_rle4_save(const char* path, int sx, int sy, int sz, const int32_t* vol)
{
    if (!path || !vol) return svo_fail(SVO_ERR_ARG, "null argument");
    if (sx < 1 || sy < 1 || sz < 1 || sx > 65536 || sy > 65536 || sz > 65536)
        return svo_fail(SVO_ERR_ARG, "bad volume size");
    std::vector<uint16_t> slabs;
    std::vector<uint16_t> runs, cols;
    for (int z = 0; z < sz; z++)
        for (int x = 0; x < sx; x++) {
            runs.clear();
            cols.clear();
            const int32_t* c = vol + ((size_t)z * sx + x) * sy;
            uint32_t gap = 0;
            for (int r = 0; r < sy;) {
                if (c[r] < 0) {
                    gap++;
                    r++;
                    continue;
                }
                int len = 0;
                while (r + len < sy && c[r + len] >= 0 && len < 63) len++;
                while (gap > 1023) {  // empty run words carry long gaps
                    runs.push_back(1023);
                    gap -= 1023;
                }
                runs.push_back((uint16_t)(gap | ((uint32_t)len << 10)));
                for (int k = 0; k < len; k++) cols.push_back((uint16_t)c[r + k]);
                r += len;
                gap = 0;
            }
            if (runs.size() > 65535 || cols.size() > 65535) return svo_fail(SVO_ERR_ARG, "column too long");
            slabs.push_back((uint16_t)runs.size());
            slabs.push_back((uint16_t)cols.size());
            slabs.insert(slabs.end(), runs.begin(), runs.end());
            slabs.insert(slabs.end(), cols.begin(), cols.end());
        }
    if (slabs.size() > 0x7fffffff) return svo_fail(SVO_ERR_ARG, "volume too large");
    FILE* f = fopen(path, "wb");
    if (!f) return svo_fail(SVO_ERR_IO, "cannot create %s", path);
    const int32_t hdr[5] = {1, sx, sy, sz, (int32_t)slabs.size()};
    bool ok = fwrite(hdr, 4, 5, f) == 5 && fwrite(slabs.data(), 2, slabs.size(), f) == slabs.size();
    ok = (fclose(f) == 0) && ok;
    return ok ? SVO_OK : svo_fail(SVO_ERR_IO, "write to %s failed", path);
}

// ---------------------------------------------------------------- paths
extern "C" int svo_camera_path(uint64_t seed, int depth, int n_frames, float max_voxels_per_frame,
                               float max_yaw_deg_per_frame, float height_voxels, float pitch, float* out)
{
    if (!out || n_frames <= 0 || depth < 1 || depth > 20) return svo_fail(SVO_ERR_ARG, "bad path arguments");
    const int seg = 16;
    const int K = n_frames / seg + 4;
    const double size = (double)(1 << depth);
    Pcg32 rng(seed, 11);
    std::vector<double> px(K), pz(K), hd(K);
    px[0] = rng.uniform() * size;
    pz[0] = rng.uniform() * size;
    hd[0] = rng.uniform() * 6.283185307179586;
    const double ymax = max_yaw_deg_per_frame * 3.141592653589793 / 180.0;
    for (int k = 1; k < K; k++) {
        hd[k] = hd[k - 1] + (rng.uniform() * 2 - 1) * ymax * seg * 0.6;
        double sp = max_voxels_per_frame * (0.45 + 0.3 * rng.uniform());
        px[k] = px[k - 1] + sp * seg * sin(hd[k]);
        pz[k] = pz[k - 1] + sp * seg * cos(hd[k]);
    }
    auto cr = [](double a, double b, double c, double d, double t) {
        double t2 = t * t, t3 = t2 * t;
        return 0.5 * ((2 * b) + (-a + c) * t + (2 * a - 5 * b + 4 * c - d) * t2 + (-a + 3 * b - 3 * c + d) * t3);
    };
    for (int f = 0; f < n_frames; f++) {
        int s = f / seg + 1;
        double t = (double)(f % seg) / seg;
        double x = cr(px[s - 1], px[s], px[s + 1], px[s + 2], t);
        double z = cr(pz[s - 1], pz[s], pz[s + 1], pz[s + 2], t);
        double yaw = cr(hd[s - 1], hd[s], hd[s + 1], hd[s + 2], t);
        float* o = out + 6 * (size_t)f;
        o[0] = (float)yaw;
        o[1] = pitch;
        o[2] = 0.0f;
        o[3] = (float)(x / 8.0);
        o[4] = (float)(height_voxels / 8.0);
        o[5] = (float)(z / 8.0);
    }
    return SVO_OK;
}

// ---------------------------------------------------------------- image output
// The colorized words are the PBO layout b + g<<8 + r<<16 that raycast.h
// uploads with glTexSubImage2D(GL_BGRA) (raycast.h:425-455); a binary PPM
// stores r, g, b bytes per pixel, rows top to bottom.
// PNG (RFC 2083 / ISO 15948): 8-bit RGB, filter 0 on every row, the zlib
// stream (RFC 1950) made of stored deflate blocks (RFC 1951 BTYPE 00), so no
// compressor is needed.  CRC-32 is the reflected polynomial 0xEDB88320 over
// chunk type + data; Adler-32 is taken modulo 65521 over the raw rows.
namespace {
uint32_t crc32_update(uint32_t c, const unsigned char* p, size_t n)
{
    static uint32_t table[256];
    static bool ready = false;
    if (!ready) {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t v = i;
            for (int k = 0; k < 8; k++) v = (v & 1u) ? 0xEDB88320u ^ (v >> 1) : v >> 1;
            table[i] = v;
        }
        ready = true;
    }
    for (size_t i = 0; i < n; i++) c = table[(c ^ p[i]) & 255u] ^ (c >> 8);
    return c;
}
void put_be32(std::vector<unsigned char>& o, uint32_t v)
{
    for (int s = 24; s >= 0; s -= 8) o.push_back((unsigned char)(v >> s));
}
void put_chunk(std::vector<unsigned char>& o, const char* type, const std::vector<unsigned char>& data)
{
    put_be32(o, (uint32_t)data.size());
    const size_t at = o.size();
    o.insert(o.end(), type, type + 4);
    o.insert(o.end(), data.begin(), data.end());
    put_be32(o, crc32_update(0xffffffffu, o.data() + at, 4 + data.size()) ^ 0xffffffffu);
}
}  // namespace

extern "C" int svo_image_save_png(const char* path, const uint32_t* image, int w, int h)
{
    if (!path || !image || w < 1 || h < 1 || w > (1 << 20) || h > (1 << 20))
        return svo_fail(SVO_ERR_ARG, "bad image arguments");
    const size_t row = (size_t)w * 3 + 1;
    std::vector<unsigned char> raw(row * h);
    for (int y = 0; y < h; y++) {
        unsigned char* r = raw.data() + row * y;
        r[0] = 0;  // filter type None
        for (int x = 0; x < w; x++) {
            const uint32_t v = image[(size_t)y * w + x];
            r[1 + 3 * x] = (unsigned char)(v >> 16);
            r[2 + 3 * x] = (unsigned char)(v >> 8);
            r[3 + 3 * x] = (unsigned char)v;
        }
    }
    std::vector<unsigned char> z;
    z.reserve(raw.size() + raw.size() / 65535 * 5 + 16);
    z.push_back(0x78);  // CMF: deflate, 32 KiB window
    z.push_back(0x01);  // FLG: (0x78 * 256 + 0x01) % 31 == 0, no dictionary
    uint32_t a = 1, b = 0;
    for (size_t ofs = 0; ofs < raw.size() || ofs == 0;) {
        const size_t len = std::min<size_t>(65535, raw.size() - ofs);
        const bool last = ofs + len == raw.size();
        z.push_back(last ? 1 : 0);  // BFINAL, BTYPE 00
        z.push_back((unsigned char)len);
        z.push_back((unsigned char)(len >> 8));
        z.push_back((unsigned char)~len);
        z.push_back((unsigned char)(~len >> 8));
        z.insert(z.end(), raw.begin() + ofs, raw.begin() + ofs + len);
        for (size_t i = ofs; i < ofs + len; i++) {
            a = (a + raw[i]) % 65521u;
            b = (b + a) % 65521u;
        }
        ofs += len;
        if (last) break;
    }
    put_be32(z, (b << 16) | a);
    std::vector<unsigned char> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<unsigned char> ihdr;
    put_be32(ihdr, (uint32_t)w);
    put_be32(ihdr, (uint32_t)h);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8 bits, RGB, deflate, filter set 0, no interlace
    put_chunk(out, "IHDR", ihdr);
    put_chunk(out, "IDAT", z);
    put_chunk(out, "IEND", {});
    FILE* f = fopen(path, "wb");
    if (!f) return svo_fail(SVO_ERR_IO, "cannot create %s", path);
    bool ok = fwrite(out.data(), 1, out.size(), f) == out.size();
    ok = (fclose(f) == 0) && ok;
    return ok ? SVO_OK : svo_fail(SVO_ERR_IO, "write to %s failed", path);
}

extern "C" int svo_image_save_ppm(const char* path, const uint32_t* image, int w, int h)
{
    if (!path || !image || w < 1 || h < 1) return svo_fail(SVO_ERR_ARG, "bad image arguments");
    std::vector<unsigned char> rgb((size_t)w * h * 3);
    for (size_t i = 0; i < (size_t)w * h; i++) {
        rgb[3 * i + 0] = (unsigned char)(image[i] >> 16);
        rgb[3 * i + 1] = (unsigned char)(image[i] >> 8);
        rgb[3 * i + 2] = (unsigned char)image[i];
    }
    FILE* f = fopen(path, "wb");
    if (!f) return svo_fail(SVO_ERR_IO, "cannot create %s", path);
    bool ok = fprintf(f, "P6\n%d %d\n255\n", w, h) > 0 && fwrite(rgb.data(), 1, rgb.size(), f) == rgb.size();
    ok = (fclose(f) == 0) && ok;
    return ok ? SVO_OK : svo_fail(SVO_ERR_IO, "write to %s failed", path);
}
