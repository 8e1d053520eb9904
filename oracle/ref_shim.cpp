/*
 * ref_shim.cpp -- C entry points around the two self-contained headers of
 * the reference's hot path, compiled from the reference tree where it lies
 * (oracle/Makefile, target ref).  Output: oracle/_ref/libwv_ref.so.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle's restatement
 * of Distance::compare (include/weavess/distance.h:12-36) and GenRandom
 * (include/weavess/util.h:13-27).  The rest of the path (index.h and the
 * component sources) includes Boost headers that this image lacks, so it is
 * not built.
 */
#include <cstdint>
#include <random>
#include "weavess/distance.h"
#include "weavess/util.h"

extern "C" {

float wvr_dist(const float *a, const float *b, uint32_t d) {
    weavess::Distance dist;
    return dist.compare(a, b, d);
}

/* GenRandom driven by std::mt19937(seed), as the reference seeds it. */
void wvr_gen_random(uint32_t seed, uint32_t *addr, uint32_t size, uint32_t N) {
    std::mt19937 rng(seed);
    weavess::GenRandom(rng, addr, size, N);
}

/* The first n outputs of std::mt19937(seed). */
void wvr_mt19937(uint32_t seed, uint32_t *out, uint32_t n) {
    std::mt19937 rng(seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = (uint32_t)rng();
}
}
