/* A headless host loop written against the C ABI alone: the depth-8
 * heightfield flythrough (scene seed 1001, path seed 2001) at 512x256, warp
 * frames along the seeded camera path, the last frame's image saved as PNG.
 *
 *     render_flythrough OUT.png [frames]
 *
 * The calls follow the reference's raycast_init / raycast_draw order
 * (raycast.h:13-85 build and upload, :93-510 per frame): build the octree,
 * create the context, upload, size the buffers, then one svo_frame_warp per
 * frame and a blocking read of the colorized image. */
#include <stdio.h>
#include <stdlib.h>

#include "svo.h"

#define CHECK(call)                                                        \
    do {                                                                   \
        if ((call) != 0) {                                                 \
            fprintf(stderr, "%s failed: %s\n", #call, svo_last_error()); \
            return 1;                                                      \
        }                                                                  \
    } while (0)

int main(int argc, char** argv)
{
    enum { DEPTH = 8, RES_X = 512, RES_Y = 256 };
    int frames, f;
    uint16_t* xyz = NULL;
    uint32_t *rgba = NULL, *words = NULL, *image = NULL, root = 0;
    size_t n_voxels = 0, n_words = 0;
    float* path = NULL;
    svo_ctx* ctx = NULL;

    if (argc < 2) {
        fprintf(stderr, "usage: %s OUT.png [frames 1..64]\n", argv[0]);
        return 2;
    }
    frames = argc > 2 ? atoi(argv[2]) : 16;
    if (frames < 1 || frames > 64) {
        fprintf(stderr, "frames must be 1..64\n");
        return 2;
    }
    CHECK(svo_scene_generate(1, 1001, DEPTH, 0, &xyz, &rgba, &n_voxels));
    CHECK(svo_build_octree(xyz, rgba, n_voxels, DEPTH, &words, &n_words, &root));
    svo_free(xyz);
    svo_free(rgba);
    path = (float*)malloc(sizeof(float) * 6 * (size_t)frames);
    image = (uint32_t*)malloc(sizeof(uint32_t) * RES_X * RES_Y);
    if (!path || !image) {
        fprintf(stderr, "out of host memory\n");
        return 1;
    }
    CHECK(svo_camera_path(2001, DEPTH, frames, 2.0f, 0.5f, 240.0f, 0.35f, path));

    CHECK(svo_create(&ctx, 0));
    CHECK(svo_upload_octree(ctx, words, n_words, root, DEPTH));
    CHECK(svo_resize(ctx, RES_X, RES_Y));
    for (f = 0; f < frames; f++) {
        const float* row = path + 6 * f;  /* yaw, pitch, roll, position */
        svo_frame_params p;
        CHECK(svo_camera(row[0], row[1], row[2], row + 3, DEPTH, f, &p, NULL));
        CHECK(svo_frame_warp(ctx, &p));
    }
    CHECK(svo_read(ctx, SVO_BUF_IMAGE, 0, sizeof(uint32_t) * RES_X * RES_Y, image));
    CHECK(svo_image_save_png(argv[1], image, RES_X, RES_Y));
    CHECK(svo_destroy(ctx));
    printf("%d warp frames of %zu voxels (%zu octree words) -> %s\n", frames, n_voxels, n_words, argv[1]);
    svo_free(words);
    free(path);
    free(image);
    return 0;
}
