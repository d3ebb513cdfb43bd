/* lq_driver.c -- times the reference tree's only CPU proximity code,
 * OpenSteer's locality-query bin lattice (Samples/OpensteerDemo/src/lq.c,
 * interface Samples/OpensteerDemo/include/OpenSteer/lq.h), on a flock, for a
 * reference-code CPU baseline of the neighbour search.  Test and measurement
 * infrastructure: linked with lq.c from the reference sources into
 * oracle/_ref/ by oracle/build_ref.sh.
 *
 * lq answers uncapped sphere queries (d^2 < r^2, strict) over per-bin linked
 * lists with no canonical order, so its neighbour sets are not BehaveRT's
 * capped lists; it is timed, not compared bit for bit. */
#include <stdint.h>
#include <stdlib.h>

#include <omp.h>

#include "OpenSteer/lq.h"

typedef struct {
    uint64_t count;
} lq_count_state;

static void lq_count(void* client, float distance_squared, void* state) {
    (void)client;
    (void)distance_squared;
    ((lq_count_state*)state)->count += 1;
}

/* pos: n float4 (x, y, z, w).  Builds the database (lqCreateDatabase over the
 * box origin + size with div bins per axis, lqUpdateForNewLocation per agent,
 * serial: the bins are linked lists), then runs one
 * lqMapOverAllObjectsInLocality for every stride-th agent (agents 0, stride,
 * 2 stride, ...) on `threads` OpenMP threads (the queries only read the
 * lattice).  Returns 0, or -1 when out of memory; neighbours (summed over the
 * queried agents) excludes each agent itself. */
int lq_flock_baseline(const float* pos, uint32_t n, const float origin[3], const float size[3], const int div[3],
                      float radius, uint32_t stride, int threads, double* build_s, double* query_s,
                      uint64_t* neighbours) {
    lqClientProxy* proxies = (lqClientProxy*)malloc(sizeof(lqClientProxy) * (size_t)(n ? n : 1));
    if (!proxies) return -1;
    const double t0 = omp_get_wtime();
    lqDB* db = lqCreateDatabase(origin[0], origin[1], origin[2], size[0], size[1], size[2], div[0], div[1], div[2]);
    for (uint32_t i = 0; i < n; ++i) {
        lqInitClientProxy(&proxies[i], (void*)(uintptr_t)i);
        lqUpdateForNewLocation(db, &proxies[i], pos[4 * (size_t)i], pos[4 * (size_t)i + 1], pos[4 * (size_t)i + 2]);
    }
    const double t1 = omp_get_wtime();
    uint64_t total = 0;
    if (stride == 0) stride = 1;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 256) reduction(+ : total)
    for (int64_t i = 0; i < (int64_t)n; i += stride) {
        lq_count_state st = {0};
        lqMapOverAllObjectsInLocality(db, pos[4 * i], pos[4 * i + 1], pos[4 * i + 2], radius, lq_count, &st);
        total += st.count - 1; /* the agent itself: distance 0 < radius^2 */
    }
    const double t2 = omp_get_wtime();
    lqDeleteDatabase(db);
    free(proxies);
    *build_s = t1 - t0;
    *query_s = t2 - t1;
    *neighbours = total;
    return 0;
}
