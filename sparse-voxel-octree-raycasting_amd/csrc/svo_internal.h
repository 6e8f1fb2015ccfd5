// svo_internal.h -- shared declarations between the host (.cpp) and device
// (.hip) halves of libsvo.
#pragma once
#include <cstdint>

#include "../../include/svo.h"

// Records the message for svo_last_error() and returns code.
int svo_fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

// Host worker threads used by the builder and scene generators.
int svo_host_threads();

#include <vector>
// Validated device copy of a compact word array with the normal region in
// breadth-first order; *normal_words = size of the renumbered region.
int svo_device_layout(const uint32_t* words, size_t n_words, uint32_t root, int depth,
                      std::vector<uint32_t>& out, uint32_t* new_root, uint32_t* normal_words);

// Hard bound on traversal iterations per ray (a ray that has not exited
// after this many loop trips is finished as a miss at its current node).
// Far above what any ray of a valid scene needs: every step advances the
// ray to a cell face and rays stop at 400000 ray units.
#define SVO_MAX_RAY_ITERS (1u << 22)

// Pixels svo_proj_noalloc takes at most (its key array: 8 bytes per pixel,
// 16 MiB per device): 1920x1024 and 2048x1024 fit.
#define SVO_NOALLOC_MAX_PIXELS (1u << 21)
