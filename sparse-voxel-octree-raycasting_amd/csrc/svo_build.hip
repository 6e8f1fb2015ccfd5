// svo_build.hip -- octree builder on the GPU.
//
// Emits the same compact word array as the host builder (svo_host.cpp), i.e.
// the layout of set_voxel + convert_tree_blocks (src/octree/octree.h:32-293):
// 10-word normal records for depths <= D-7 in post-order from word 0, blocks
// rooted at depth D-6 placed 64-word aligned from word 2^21 in depth-first
// order, compact records [colour][child refs] for depths D-6..D-3 in
// post-order inside a block with the block root's record moved to the front,
// byte records at depth D-2, and synthesized depth-(D-1) refs.
//
// Data-parallel construction:
//   1. Morton keys (root digit first) and a stable radix sort of (key, stream
//      index); equal keys keep stream order.
//   2. Unique leaves: colour of the last voxel at a position, stream index of
//      the first (the creator rule of octree.h:46-48, :75-81).
//   3. Per depth d, nodes are the runs of equal key prefixes; node ids come
//      from a scan of run-start flags.  Child masks and creator indices are
//      reduced bottom-up with one atomic per child node.
//   4. Post-order = order by (last leaf of the node's range, depth deepest
//      first): every leaf carries the chain of records that end at it, so the
//      block-relative offsets are one exclusive scan of per-leaf record words,
//      and the normal-record indices one scan of per-leaf normal-node counts.
//   5. One thread per record writes it.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/svo.h"
#include "svo_internal.h"

namespace {

constexpr int kMaxLevels = 16;
constexpr uint32_t kNormalWords = 2097152u;  // normal region (raycast.h:38)

struct Levels {
    uint32_t* nid[kMaxLevels];     // per leaf: id of its depth-d node
    uint32_t* nstart[kMaxLevels];  // per node: first leaf
    uint32_t* mask[kMaxLevels];    // per node: child mask
    uint32_t* first[kMaxLevels];   // per node: smallest stream index below it
    uint32_t* ofs[kMaxLevels];     // per node: record offset (block-relative, or normal word offset)
    uint32_t count[kMaxLevels];    // nodes at depth d
};

__device__ __forceinline__ bool starts(const uint64_t* key, uint32_t j, int shift)
{
    return j == 0 || (key[j] >> shift) != (key[j - 1] >> shift);
}
__device__ __forceinline__ bool ends(const uint64_t* key, uint32_t j, uint32_t L, int shift)
{
    return j + 1 == L || (key[j] >> shift) != (key[j + 1] >> shift);
}

__global__ void k_morton(const uint16_t* __restrict__ xyz, uint32_t n, int D, uint64_t* __restrict__ key,
                         uint32_t* __restrict__ idx, uint32_t* __restrict__ err)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const uint32_t x = xyz[3 * (size_t)i], y = xyz[3 * (size_t)i + 1], z = xyz[3 * (size_t)i + 2];
    const uint32_t lim = 1u << D;
    if (x >= lim || y >= lim || z >= lim) atomicOr(err, 1u);
    uint64_t k = 0;
    for (int b = D - 1; b >= 0; b--)
        k = (k << 3) | (uint64_t)(((x >> b) & 1u) | (((y >> b) & 1u) << 1) | (((z >> b) & 1u) << 2));
    key[i] = k;
    idx[i] = i;
}

__global__ void k_heads(const uint64_t* __restrict__ key, uint32_t n, uint32_t* __restrict__ flag)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) flag[i] = (i == 0 || key[i] != key[i - 1]) ? 1u : 0u;
}

// lid = inclusive scan of the head flags
__global__ void k_leaves(const uint64_t* __restrict__ key, const uint32_t* __restrict__ idx,
                         const uint32_t* __restrict__ rgba, const uint32_t* __restrict__ lid, uint32_t n,
                         uint64_t* __restrict__ leaf_key, uint32_t* __restrict__ leaf_first,
                         uint32_t* __restrict__ leaf_color)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = lid[i] - 1u;
    if (i == 0 || key[i] != key[i - 1]) {
        leaf_key[j] = key[i];
        leaf_first[j] = idx[i];
    }
    if (i + 1 == n || key[i] != key[i + 1]) leaf_color[j] = rgba[idx[i]];
}

__global__ void k_level_flags(const uint64_t* __restrict__ key, uint32_t L, int shift, uint32_t* __restrict__ flag)
{
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j < L) flag[j] = starts(key, j, shift) ? 1u : 0u;
}

// nid holds the inclusive scan of the start flags; make it 0-based and record
// each node's first leaf
__global__ void k_level_start(const uint64_t* __restrict__ key, uint32_t L, int shift, uint32_t* __restrict__ nid,
                              uint32_t* __restrict__ nstart)
{
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j >= L) return;
    const uint32_t id = nid[j] - 1u;
    nid[j] = id;
    if (starts(key, j, shift)) nstart[id] = j;
}

// child (depth d+1, or leaf when d+1 == D) -> parent mask bit and creator index
__global__ void k_level_link(const uint64_t* __restrict__ key, const uint32_t* __restrict__ leaf_first, uint32_t L,
                             int D, int d, Levels lv)
{
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j >= L) return;
    const int cd = d + 1;
    if (cd < D && !starts(key, j, 3 * (D - cd))) return;
    const uint32_t parent = lv.nid[d][j];
    const uint32_t digit = (uint32_t)(key[j] >> (3 * (D - 1 - d))) & 7u;
    const uint32_t cf = (cd == D) ? leaf_first[j] : lv.first[cd][lv.nid[cd][j]];
    atomicOr(&lv.mask[d][parent], 1u << digit);
    atomicMin(&lv.first[d][parent], cf);
}

__global__ void k_bytecount(const uint32_t* __restrict__ leaf_color, uint32_t L, const uint32_t* __restrict__ nid,
                            uint32_t* __restrict__ nbytes)
{
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j < L && leaf_color[j] != 0u) atomicAdd(&nbytes[nid[j]], 1u);  // octree.h:133-135
}

__device__ __forceinline__ uint32_t rec_words(const Levels& lv, const uint32_t* nbytes, int D, int d, uint32_t id)
{
    const uint32_t kids = (uint32_t)__popc(lv.mask[d][id]);
    if (d == D - 2) return (1u + kids + nbytes[id] + 3u) >> 2;
    return 1u + kids;
}

// per leaf: words of the non-root block records ending at it (W) and the
// number of normal records ending at it (C); entry L of both is 0
__global__ void k_chain(const uint64_t* __restrict__ key, uint32_t L, int D, Levels lv,
                        const uint32_t* __restrict__ nbytes, uint64_t* __restrict__ W, uint32_t* __restrict__ C)
{
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j > L) return;
    uint64_t w = 0;
    uint32_t c = 0;
    if (j < L) {
        for (int d = D - 2; d > D - 6; d--) {
            if (!ends(key, j, L, 3 * (D - d))) break;
            w += rec_words(lv, nbytes, D, d, lv.nid[d][j]);
        }
        for (int d = D - 7; d >= 0; d--) {
            if (!ends(key, j, L, 3 * (D - d))) break;
            c++;
        }
    }
    W[j] = w;
    C[j] = c;
}

// sW, sC: exclusive scans of W and C (L + 1 entries)
__global__ void k_offsets(const uint64_t* __restrict__ key, uint32_t L, int D, Levels lv,
                          const uint32_t* __restrict__ nbytes, const uint64_t* __restrict__ sW,
                          const uint32_t* __restrict__ sC)
{
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j >= L) return;
    const int db = D - 6;
    const uint32_t b = lv.nid[db][j];
    uint32_t run = (uint32_t)(sW[j] - sW[lv.nstart[db][b]]) + 1u + (uint32_t)__popc(lv.mask[db][b]);
    for (int d = D - 2; d > db; d--) {
        if (!ends(key, j, L, 3 * (D - d))) break;
        const uint32_t id = lv.nid[d][j];
        lv.ofs[d][id] = run;
        run += rec_words(lv, nbytes, D, d, id);
    }
    uint32_t k = sC[j];
    for (int d = D - 7; d >= 0; d--) {
        if (!ends(key, j, L, 3 * (D - d))) break;
        lv.ofs[d][lv.nid[d][j]] = 10u * k;
        k++;
    }
}

// block sizes: root record + the block's non-root records; size64 rounded to 64
__global__ void k_block_size(uint32_t L, int D, Levels lv, const uint64_t* __restrict__ sW,
                             uint64_t* __restrict__ size, uint64_t* __restrict__ size64)
{
    const int db = D - 6;
    const uint32_t b = blockIdx.x * 256u + threadIdx.x;
    if (b >= lv.count[db]) return;
    const uint32_t s = lv.nstart[db][b];
    const uint32_t e = (b + 1 < lv.count[db]) ? lv.nstart[db][b + 1] : L;
    const uint64_t w = sW[e] - sW[s] + 1u + (uint64_t)__popc(lv.mask[db][b]);
    size[b] = w;
    size64[b] = (w + 63u) & ~(uint64_t)63u;
}

// compact records, depths D-6..D-3 (octree.h:166-190): [colour][refs]
__global__ void k_emit_compact(int D, int d, Levels lv, const uint32_t* __restrict__ rgba,
                               const uint64_t* __restrict__ bpos, uint32_t* __restrict__ out)
{
    const uint32_t id = blockIdx.x * 256u + threadIdx.x;
    if (id >= lv.count[d]) return;
    const uint32_t s = lv.nstart[d][id];
    const uint64_t base = bpos[lv.nid[D - 6][s]] + lv.ofs[d][id];
    out[base] = rgba[lv.first[d][id]];
    const uint32_t c0 = lv.nid[d + 1][s];
    const int kids = __popc(lv.mask[d][id]);
    for (int k = 0; k < kids; k++) {
        const uint32_t c = c0 + (uint32_t)k;
        out[base + 1 + k] = (lv.ofs[d + 1][c] << 9) | 256u | lv.mask[d + 1][c];
    }
}

// byte records, depth D-2 (octree.h:106-157)
__global__ void k_emit_bytes(uint32_t L, int D, Levels lv, const uint32_t* __restrict__ rgba,
                             const uint32_t* __restrict__ leaf_color, const uint64_t* __restrict__ bpos,
                             uint32_t* __restrict__ out)
{
    const int d = D - 2;
    const uint32_t id = blockIdx.x * 256u + threadIdx.x;
    if (id >= lv.count[d]) return;
    const uint32_t s = lv.nstart[d][id];
    uint32_t* dst = out + bpos[lv.nid[D - 6][s]] + lv.ofs[d][id];
    uint32_t word = rgba[lv.first[d][id]] & 255u, nb = 1;
    auto put = [&](uint32_t byte) {
        word |= byte << (8 * (nb & 3u));
        nb++;
        if ((nb & 3u) == 0) {
            *dst++ = word;
            word = 0;
        }
    };
    const uint32_t c0 = lv.nid[d + 1][s];
    const uint32_t kids = (uint32_t)__popc(lv.mask[d][id]);
    const uint32_t m1 = lv.count[d + 1];
    for (uint32_t k = 0; k < kids; k++) put(lv.mask[d + 1][c0 + k]);
    for (uint32_t k = 0; k < kids; k++) {
        const uint32_t c = c0 + k;
        const uint32_t le = (c + 1 < m1) ? lv.nstart[d + 1][c + 1] : L;
        for (uint32_t j = lv.nstart[d + 1][c]; j < le; j++)
            if (leaf_color[j] != 0u) put(leaf_color[j] & 255u);
    }
    if (nb & 3u) *dst = word;
}

// normal records, depths <= D-7 (octree.h:280-288)
__global__ void k_emit_normal(int D, int d, Levels lv, const uint32_t* __restrict__ rgba,
                              const uint64_t* __restrict__ bpos, uint32_t* __restrict__ out)
{
    const uint32_t id = blockIdx.x * 256u + threadIdx.x;
    if (id >= lv.count[d]) return;
    uint32_t* dst = out + lv.ofs[d][id];
    const uint32_t mask = lv.mask[d][id];
    const uint32_t c0 = lv.nid[d + 1][lv.nstart[d][id]];
    uint32_t rank = 0;
    for (int k = 0; k < 8; k++) {
        uint32_t ref = 0;
        if (mask >> k & 1u) {
            const uint32_t c = c0 + rank++;
            if (d + 1 <= D - 7)
                ref = (lv.ofs[d + 1][c] << 9) | lv.mask[d + 1][c];
            else
                ref = ((uint32_t)(bpos[c] >> 6) << 9) | 256u | lv.mask[d + 1][c];
        }
        dst[k] = ref;
    }
    const uint32_t color = (d == 0) ? 0u : rgba[lv.first[d][id]];
    dst[8] = color;
    dst[9] = color;
}

inline unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }

// device allocations of one build, freed on every exit path
struct Pool {
    std::vector<void*> ptrs;
    ~Pool()
    {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    T* get(size_t count)
    {
        void* p = nullptr;
        if (hipMalloc(&p, count * sizeof(T) + 16) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return (T*)p;
    }
};

struct Stream {
    hipStream_t s = nullptr;
    ~Stream()
    {
        if (s) (void)hipStreamDestroy(s);
    }
};

}  // namespace

#define BCHECK(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) return svo_fail(SVO_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)
#define BALLOC(ptr, T, count)                                                             \
    do {                                                                                  \
        ptr = pool.get<T>(count);                                                         \
        if (!ptr) return svo_fail(SVO_ERR_NOMEM, "octree build: %zu bytes", (size_t)(count) * sizeof(T)); \
    } while (0)

extern "C" int svo_build_octree_gpu(int device, const uint16_t* xyz, const uint32_t* rgba, size_t n, int depth,
                                    uint32_t** words_out, size_t* n_words, uint32_t* root_ref)
{
    if (!words_out || !n_words || !root_ref) return svo_fail(SVO_ERR_ARG, "null output pointer");
    if (depth < 8 || depth > 13) return svo_fail(SVO_ERR_ARG, "depth %d outside 8..13", depth);
    if (n > 0x7ffffff0ull) return svo_fail(SVO_ERR_ARG, "too many voxels for the GPU builder");
    if (n && (!xyz || !rgba)) return svo_fail(SVO_ERR_ARG, "null voxel stream");
    const int D = depth;
    if (n == 0) {  // empty scene: the zeroed normal region, root ref 0
        uint32_t* out = (uint32_t*)calloc(kNormalWords, 4);
        if (!out) return svo_fail(SVO_ERR_NOMEM, "octree output");
        *words_out = out;
        *n_words = kNormalWords;
        *root_ref = 0;
        return SVO_OK;
    }
    int ndev = 0;
    BCHECK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return svo_fail(SVO_ERR_ARG, "device %d not present", device);
    BCHECK(hipSetDevice(device));
    Stream st;
    BCHECK(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
    hipStream_t s = st.s;
    Pool pool;
    const uint32_t N = (uint32_t)n;

    uint16_t* d_xyz;
    uint32_t *d_rgba, *d_idx, *d_idx2, *d_err, *d_flag;
    uint64_t *d_key, *d_key2;
    BALLOC(d_xyz, uint16_t, 3 * n);
    BALLOC(d_rgba, uint32_t, n);
    BALLOC(d_key, uint64_t, n);
    BALLOC(d_key2, uint64_t, n);
    BALLOC(d_idx, uint32_t, n);
    BALLOC(d_idx2, uint32_t, n);
    BALLOC(d_flag, uint32_t, n + 1);
    BALLOC(d_err, uint32_t, 1);
    BCHECK(hipMemcpyAsync(d_xyz, xyz, 6 * n, hipMemcpyHostToDevice, s));
    BCHECK(hipMemcpyAsync(d_rgba, rgba, 4 * n, hipMemcpyHostToDevice, s));
    BCHECK(hipMemsetAsync(d_err, 0, 4, s));
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(n)), dim3(256), 0, s, d_xyz, N, D, d_key, d_idx, d_err);
    BCHECK(hipGetLastError());

    // scratch for the library sort / scans, sized for the largest call
    size_t tmp_sort = 0, tmp_scan = 0;
    BCHECK(rocprim::radix_sort_pairs(nullptr, tmp_sort, d_key, d_key2, d_idx, d_idx2, N, 0u, (unsigned)(3 * D), s));
    BCHECK(rocprim::inclusive_scan(nullptr, tmp_scan, d_flag, d_flag, N + 1, rocprim::plus<uint32_t>(), s));
    size_t tmp_ex32 = 0, tmp_ex64 = 0;
    BCHECK(rocprim::exclusive_scan(nullptr, tmp_ex32, d_flag, d_flag, 0u, N + 1, rocprim::plus<uint32_t>(), s));
    BCHECK(rocprim::exclusive_scan(nullptr, tmp_ex64, d_key, d_key2, (uint64_t)0, N + 1, rocprim::plus<uint64_t>(),
                                   s));
    size_t tmp_bytes = std::max(std::max(tmp_sort, tmp_scan), std::max(tmp_ex32, tmp_ex64));
    void* d_tmp;
    BALLOC(d_tmp, char, tmp_bytes);
    BCHECK(rocprim::radix_sort_pairs(d_tmp, tmp_sort, d_key, d_key2, d_idx, d_idx2, N, 0u, (unsigned)(3 * D), s));
    uint32_t herr = 0;
    BCHECK(hipMemcpyAsync(&herr, d_err, 4, hipMemcpyDeviceToHost, s));
    BCHECK(hipStreamSynchronize(s));
    if (herr) return svo_fail(SVO_ERR_ARG, "a voxel lies outside 2^%d", D);

    // unique leaves
    uint32_t* d_lid;
    BALLOC(d_lid, uint32_t, n);
    hipLaunchKernelGGL(k_heads, dim3(blocks_for(n)), dim3(256), 0, s, d_key2, N, d_flag);
    BCHECK(hipGetLastError());
    size_t tb = tmp_bytes;
    BCHECK(rocprim::inclusive_scan(d_tmp, tb, d_flag, d_lid, N, rocprim::plus<uint32_t>(), s));
    uint32_t L = 0;
    BCHECK(hipMemcpyAsync(&L, d_lid + (N - 1), 4, hipMemcpyDeviceToHost, s));
    BCHECK(hipStreamSynchronize(s));
    uint64_t* d_lkey;
    uint32_t *d_lfirst, *d_lcolor;
    BALLOC(d_lkey, uint64_t, L);
    BALLOC(d_lfirst, uint32_t, L);
    BALLOC(d_lcolor, uint32_t, L);
    hipLaunchKernelGGL(k_leaves, dim3(blocks_for(n)), dim3(256), 0, s, d_key2, d_idx2, d_rgba, d_lid, N, d_lkey,
                       d_lfirst, d_lcolor);
    BCHECK(hipGetLastError());

    // nodes per depth
    Levels lv;
    memset(&lv, 0, sizeof(lv));
    for (int d = 0; d < D; d++) {
        const int shift = 3 * (D - d);
        BALLOC(lv.nid[d], uint32_t, L);
        hipLaunchKernelGGL(k_level_flags, dim3(blocks_for(L)), dim3(256), 0, s, d_lkey, L, shift, d_flag);
        BCHECK(hipGetLastError());
        tb = tmp_bytes;
        BCHECK(rocprim::inclusive_scan(d_tmp, tb, d_flag, lv.nid[d], L, rocprim::plus<uint32_t>(), s));
        uint32_t cnt = 0;
        BCHECK(hipMemcpyAsync(&cnt, lv.nid[d] + (L - 1), 4, hipMemcpyDeviceToHost, s));
        BCHECK(hipStreamSynchronize(s));
        lv.count[d] = cnt;
        BALLOC(lv.nstart[d], uint32_t, cnt);
        BALLOC(lv.mask[d], uint32_t, cnt);
        BALLOC(lv.first[d], uint32_t, cnt);
        BALLOC(lv.ofs[d], uint32_t, cnt);
        BCHECK(hipMemsetAsync(lv.mask[d], 0, 4 * (size_t)cnt, s));
        BCHECK(hipMemsetAsync(lv.first[d], 0xff, 4 * (size_t)cnt, s));
        BCHECK(hipMemsetAsync(lv.ofs[d], 0, 4 * (size_t)cnt, s));
        hipLaunchKernelGGL(k_level_start, dim3(blocks_for(L)), dim3(256), 0, s, d_lkey, L, shift, lv.nid[d],
                           lv.nstart[d]);
        BCHECK(hipGetLastError());
    }
    for (int d = D - 1; d >= 0; d--) {
        hipLaunchKernelGGL(k_level_link, dim3(blocks_for(L)), dim3(256), 0, s, d_lkey, d_lfirst, L, D, d, lv);
        BCHECK(hipGetLastError());
    }
    uint32_t* d_nbytes;
    BALLOC(d_nbytes, uint32_t, lv.count[D - 2]);
    BCHECK(hipMemsetAsync(d_nbytes, 0, 4 * (size_t)lv.count[D - 2], s));
    hipLaunchKernelGGL(k_bytecount, dim3(blocks_for(L)), dim3(256), 0, s, d_lcolor, L, lv.nid[D - 2], d_nbytes);
    BCHECK(hipGetLastError());

    // post-order offsets
    uint64_t *d_W, *d_sW;
    uint32_t *d_C, *d_sC;
    BALLOC(d_W, uint64_t, (size_t)L + 1);
    BALLOC(d_C, uint32_t, (size_t)L + 1);
    BALLOC(d_sW, uint64_t, (size_t)L + 1);
    BALLOC(d_sC, uint32_t, (size_t)L + 1);
    hipLaunchKernelGGL(k_chain, dim3(blocks_for((size_t)L + 1)), dim3(256), 0, s, d_lkey, L, D, lv, d_nbytes, d_W,
                       d_C);
    BCHECK(hipGetLastError());
    tb = tmp_bytes;
    BCHECK(rocprim::exclusive_scan(d_tmp, tb, d_W, d_sW, (uint64_t)0, L + 1, rocprim::plus<uint64_t>(), s));
    tb = tmp_bytes;
    BCHECK(rocprim::exclusive_scan(d_tmp, tb, d_C, d_sC, 0u, L + 1, rocprim::plus<uint32_t>(), s));
    hipLaunchKernelGGL(k_offsets, dim3(blocks_for(L)), dim3(256), 0, s, d_lkey, L, D, lv, d_nbytes, d_sW, d_sC);
    BCHECK(hipGetLastError());
    const uint32_t nblk = lv.count[D - 6];
    uint64_t *d_bsize, *d_bsize64, *d_bpos;
    BALLOC(d_bsize, uint64_t, nblk);
    BALLOC(d_bsize64, uint64_t, nblk);
    BALLOC(d_bpos, uint64_t, nblk);
    hipLaunchKernelGGL(k_block_size, dim3(blocks_for(nblk)), dim3(256), 0, s, L, D, lv, d_sW, d_bsize, d_bsize64);
    BCHECK(hipGetLastError());
    tb = tmp_bytes;
    BCHECK(rocprim::exclusive_scan(d_tmp, tb, d_bsize64, d_bpos, (uint64_t)kNormalWords, nblk,
                                   rocprim::plus<uint64_t>(), s));
    uint64_t last_pos = 0, last_size = 0;
    uint32_t n_normal = 0;
    BCHECK(hipMemcpyAsync(&last_pos, d_bpos + (nblk - 1), 8, hipMemcpyDeviceToHost, s));
    BCHECK(hipMemcpyAsync(&last_size, d_bsize + (nblk - 1), 8, hipMemcpyDeviceToHost, s));
    BCHECK(hipMemcpyAsync(&n_normal, d_sC + L, 4, hipMemcpyDeviceToHost, s));
    BCHECK(hipStreamSynchronize(s));
    const uint64_t total = last_pos + last_size;
    if ((total >> 6) >= (1ull << 23)) return svo_fail(SVO_ERR_ARG, "octree exceeds 2^29 words");
    if ((uint64_t)n_normal * 10u > kNormalWords) return svo_fail(SVO_ERR_ARG, "normal region exceeds 2^21 words");

    uint32_t* d_out;
    BALLOC(d_out, uint32_t, total);
    BCHECK(hipMemsetAsync(d_out, 0, 4 * total, s));
    for (int d = D - 6; d <= D - 3; d++) {
        hipLaunchKernelGGL(k_emit_compact, dim3(blocks_for(lv.count[d])), dim3(256), 0, s, D, d, lv, d_rgba, d_bpos,
                           d_out);
        BCHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_emit_bytes, dim3(blocks_for(lv.count[D - 2])), dim3(256), 0, s, L, D, lv, d_rgba, d_lcolor,
                       d_bpos, d_out);
    BCHECK(hipGetLastError());
    for (int d = 0; d <= D - 7; d++) {
        hipLaunchKernelGGL(k_emit_normal, dim3(blocks_for(lv.count[d])), dim3(256), 0, s, D, d, lv, d_rgba, d_bpos,
                           d_out);
        BCHECK(hipGetLastError());
    }
    uint32_t root_ofs = 0, root_mask = 0;
    BCHECK(hipMemcpyAsync(&root_ofs, lv.ofs[0], 4, hipMemcpyDeviceToHost, s));
    BCHECK(hipMemcpyAsync(&root_mask, lv.mask[0], 4, hipMemcpyDeviceToHost, s));
    uint32_t* out = (uint32_t*)malloc(4 * total);
    if (!out) return svo_fail(SVO_ERR_NOMEM, "octree output %llu words", (unsigned long long)total);
    const hipError_t e = hipMemcpyAsync(out, d_out, 4 * total, hipMemcpyDeviceToHost, s);
    const hipError_t e2 = hipStreamSynchronize(s);
    if (e != hipSuccess || e2 != hipSuccess) {
        free(out);
        return svo_fail(SVO_ERR_HIP, "octree read-back failed");
    }
    *words_out = out;
    *n_words = (size_t)total;
    *root_ref = (root_ofs << 9) | root_mask;
    return SVO_OK;
}
