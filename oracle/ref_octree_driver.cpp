// ref_octree_driver.cpp -- host harness that compiles the REFERENCE octree
// builder (src/octree/octree.h, included from the reference checkout at build
// time, never copied) and exposes it through a C entry point, so tests can
// compare the oracle's and the product's word arrays with the reference's.
// TEST INFRASTRUCTURE ONLY.  Build recipe: oracle/Makefile -> oracle/_ref/.
//
// core.h (included by octree.h via Rle4.h) #defines uint/uchar/ushort as
// multi-word types, which breaks functional casts like uint(x) in octree.h;
// defining them first as one-word typedef names makes core.h skip its own.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

typedef unsigned int svo_ref_uint;
typedef unsigned char svo_ref_uchar;
typedef unsigned short svo_ref_ushort;
#define uint svo_ref_uint
#define uchar svo_ref_uchar
#define ushort svo_ref_ushort

#include "octree.h"

extern "C" int ref_build_octree(const uint16_t* xyz, const uint32_t* rgba, size_t n,
                                int depth, uint32_t** words_out, size_t* n_words,
                                uint32_t* root_ref)
{
    // reset the reference's globals (octree.h:24-30, :91, :201-206)
    octree_array.clear();
    octree_array_compact.clear();
    octree_array_normal.clear();
    color_block.clear();
    octree_root = 0;
    octree_root_normal = 0;
    octree_size = (uint)depth;
    num_voxels = 0;
    _iteration = 0;
    lastblockoffset = 0;
    blockcount = 0;
    octree_normal_offset = 0;

    // raycast.h:13-17 then the voxel stream, raycast.h:38-39
    OctreeNodeNew node;
    memset(&node, 0, sizeof(OctreeNodeNew));
    octree_array.push_back(node);
    for (size_t i = 0; i < n; i++) {
        uchar4 c;
        memcpy(&c, &rgba[i], 4);
        set_voxel(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], c);
    }
    octree_array_compact.resize(2097152);
    octree_root_normal = convert_tree_blocks(octree_root);

    size_t nw = octree_array_compact.size();
    uint32_t* out = (uint32_t*)malloc(nw * 4);
    if (!out) return -1;
    memcpy(out, octree_array_compact.data(), nw * 4);
    *words_out = out;
    *n_words = nw;
    *root_ref = octree_root_normal;
    std::vector<OctreeNodeNew>().swap(octree_array);
    std::vector<uint>().swap(octree_array_compact);
    return 0;
}

extern "C" void ref_free(void* p) { free(p); }
