/*
 * A C caller of libselas_gpu built against include/selas_gpu.h alone (no HIP headers): one scene,
 * committed once, queried from several host threads at the same time, each thread on a stream of
 * its own, the way Selas's path tracer calls RayPick / OcclusionRay from its worker threads
 * (Source/Applications/Selas/Source/PathTracer.cpp:325-344, one TBB task per worker).
 *
 * The scene is a Gaussian-cluster triangle soup; the rays are pinhole camera rays plus cosine bounce
 * rays leaving the soup. The main thread first answers every query on one stream (closest hit and
 * occlusion); then each of the threads copies the rays into buffers of its own, answers the same
 * queries `rounds` times on its own stream and compares every result with the serial one bit for
 * bit (traversal is a function of the triangle set and the ray alone). Prints one line per thread
 * and a summary; exit status 0 when every comparison matched, 1 on a mismatch, 2 on a library error.
 *
 * usage: capi_threads [threads=8] [rounds=4] [triangles_per_cluster=1024]
 */
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "selas_gpu.h"

#define CHECK(call)                                                                        \
    do {                                                                                   \
        int rc_ = (call);                                                                  \
        if (rc_ != SGPU_OK) {                                                              \
            fprintf(stderr, "%s:%d: %s -> %d (%s)\n", __FILE__, __LINE__, #call, rc_,      \
                    sgpu_last_error());                                                    \
            exit(2);                                                                       \
        }                                                                                  \
    } while (0)

enum { N_RAY_FIELDS = 8, N_HIT_FIELDS = 5 };

typedef struct {
    sgpu_context ctx;
    sgpu_scene scene;
    uint32_t n;                      /* rays */
    float* host_rays[N_RAY_FIELDS];  /* org xyz, dir xyz, tnear, tfar */
    uint32_t* want_hits[N_HIT_FIELDS]; /* t, u, v (as bits), prim, geom of the serial pass */
    uint32_t* want_occ;
    int rounds;
} Shared;

typedef struct {
    const Shared* sh;
    int index;
    uint64_t mismatches;
    uint64_t hits;
} Worker;

/* Device copies of one ray set and its outputs. */
typedef struct {
    void* rays[N_RAY_FIELDS];
    void* hits[N_HIT_FIELDS];
    void* occ;
} DeviceSet;

static void device_set_alloc(sgpu_context ctx, uint32_t n, DeviceSet* d)
{
    for (int k = 0; k < N_RAY_FIELDS; ++k) CHECK(sgpu_malloc(ctx, (size_t)n * 4, &d->rays[k]));
    for (int k = 0; k < N_HIT_FIELDS; ++k) CHECK(sgpu_malloc(ctx, (size_t)n * 4, &d->hits[k]));
    CHECK(sgpu_malloc(ctx, (size_t)n * 4, &d->occ));
}

static void device_set_free(sgpu_context ctx, DeviceSet* d)
{
    for (int k = 0; k < N_RAY_FIELDS; ++k) CHECK(sgpu_free(ctx, d->rays[k]));
    for (int k = 0; k < N_HIT_FIELDS; ++k) CHECK(sgpu_free(ctx, d->hits[k]));
    CHECK(sgpu_free(ctx, d->occ));
}

/* Uploads the rays, answers both queries on `stream` and downloads the results. */
static void query(const Shared* sh, DeviceSet* d, void* stream, uint32_t* hits_out[N_HIT_FIELDS], uint32_t* occ_out)
{
    const size_t bytes = (size_t)sh->n * 4;
    for (int k = 0; k < N_RAY_FIELDS; ++k) CHECK(sgpu_memcpy_htod(d->rays[k], sh->host_rays[k], bytes, stream));
    sgpu_rays_soa rays = {(const float*)d->rays[0], (const float*)d->rays[1], (const float*)d->rays[2],
                          (const float*)d->rays[3], (const float*)d->rays[4], (const float*)d->rays[5],
                          (const float*)d->rays[6], (const float*)d->rays[7]};
    sgpu_hits_soa hits;
    memset(&hits, 0, sizeof(hits));
    hits.t = (float*)d->hits[0];
    hits.u = (float*)d->hits[1];
    hits.v = (float*)d->hits[2];
    hits.prim_id = (uint32_t*)d->hits[3];
    hits.geom_id = (uint32_t*)d->hits[4];
    CHECK(sgpu_intersect(sh->scene, &rays, &hits, sh->n, stream));
    CHECK(sgpu_occluded(sh->scene, &rays, (uint32_t*)d->occ, sh->n, stream));
    for (int k = 0; k < N_HIT_FIELDS; ++k) CHECK(sgpu_memcpy_dtoh(hits_out[k], d->hits[k], bytes, stream));
    CHECK(sgpu_memcpy_dtoh(occ_out, d->occ, bytes, stream));
    CHECK(sgpu_stream_synchronize(stream));
}

static void* worker_main(void* arg)
{
    Worker* w = (Worker*)arg;
    const Shared* sh = w->sh;
    void* stream = NULL;
    CHECK(sgpu_stream_create(&stream));
    DeviceSet d;
    device_set_alloc(sh->ctx, sh->n, &d);
    uint32_t* got[N_HIT_FIELDS];
    for (int k = 0; k < N_HIT_FIELDS; ++k) got[k] = (uint32_t*)malloc((size_t)sh->n * 4);
    uint32_t* occ = (uint32_t*)malloc((size_t)sh->n * 4);
    for (int r = 0; r < sh->rounds; ++r) {
        for (int k = 0; k < N_HIT_FIELDS; ++k) memset(got[k], 0xA5, (size_t)sh->n * 4);
        memset(occ, 0xA5, (size_t)sh->n * 4);
        query(sh, &d, stream, got, occ);
        for (uint32_t i = 0; i < sh->n; ++i) {
            int bad = occ[i] != sh->want_occ[i];
            for (int k = 0; k < N_HIT_FIELDS; ++k) bad |= got[k][i] != sh->want_hits[k][i];
            w->mismatches += (uint64_t)bad;
            if (r == 0) w->hits += got[3][i] != SGPU_INVALID_ID;
        }
    }
    for (int k = 0; k < N_HIT_FIELDS; ++k) free(got[k]);
    free(occ);
    device_set_free(sh->ctx, &d);
    CHECK(sgpu_stream_destroy(stream));
    return NULL;
}

int main(int argc, char** argv)
{
    const int n_threads = argc > 1 ? atoi(argv[1]) : 8;
    const int rounds = argc > 2 ? atoi(argv[2]) : 4;
    const uint32_t per_cluster = argc > 3 ? (uint32_t)atoi(argv[3]) : 1024u;
    const uint32_t clusters = 64, side = 128, bounce = 32768;
    if (n_threads < 1 || n_threads > 64 || rounds < 1 || per_cluster < 1) {
        fprintf(stderr, "usage: capi_threads [threads 1..64] [rounds >= 1] [triangles_per_cluster >= 1]\n");
        return 2;
    }
    Shared sh;
    memset(&sh, 0, sizeof(sh));
    sh.rounds = rounds;
    CHECK(sgpu_context_create(0, &sh.ctx));

    /* the scene: an indexed mesh whose vertices are the soup's corners */
    const uint32_t n_tris = clusters * per_cluster;
    float* soup = (float*)malloc((size_t)n_tris * 9 * sizeof(float));
    uint32_t* idx = (uint32_t*)malloc((size_t)n_tris * 3 * sizeof(uint32_t));
    CHECK(sgpu_gen_cluster_soup(5, clusters, per_cluster, 0.05f, 2e-3f, 3e-2f, soup));
    for (uint32_t i = 0; i < 3 * n_tris; ++i) idx[i] = i;
    uint32_t geom = 0;
    CHECK(sgpu_scene_create(sh.ctx, &sh.scene));
    CHECK(sgpu_scene_add_triangles(sh.scene, soup, 3 * n_tris, idx, n_tris, NULL, &geom));
    CHECK(sgpu_scene_commit(sh.scene, NULL));

    /* the rays: camera rays toward the unit box, then bounce rays leaving the soup */
    const uint32_t n_cam = side * side;
    sh.n = n_cam + bounce;
    for (int k = 0; k < N_RAY_FIELDS; ++k) sh.host_rays[k] = (float*)malloc((size_t)sh.n * 4);
    const float pos[3] = {0.5f, 0.5f, -1.2f}, at[3] = {0.5f, 0.5f, 0.5f}, up[3] = {0.0f, 1.0f, 0.0f};
    float** r = sh.host_rays;
    CHECK(sgpu_gen_camera_rays(pos, at, up, 50.0f, side, side, 1, r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7]));
    CHECK(sgpu_gen_bounce_rays(9, soup, n_tris, bounce, r[0] + n_cam, r[1] + n_cam, r[2] + n_cam, r[3] + n_cam,
                               r[4] + n_cam, r[5] + n_cam, r[6] + n_cam, r[7] + n_cam));

    /* the serial answer, on the null stream */
    for (int k = 0; k < N_HIT_FIELDS; ++k) sh.want_hits[k] = (uint32_t*)malloc((size_t)sh.n * 4);
    sh.want_occ = (uint32_t*)malloc((size_t)sh.n * 4);
    DeviceSet d;
    device_set_alloc(sh.ctx, sh.n, &d);
    query(&sh, &d, NULL, sh.want_hits, sh.want_occ);
    device_set_free(sh.ctx, &d);
    uint64_t serial_hits = 0, serial_occ = 0;
    for (uint32_t i = 0; i < sh.n; ++i) {
        serial_hits += sh.want_hits[3][i] != SGPU_INVALID_ID;
        serial_occ += sh.want_occ[i] != 0;
    }
    printf("scene: %u triangles; %u rays; serial pass: %llu hits, %llu occluded\n", n_tris, sh.n,
           (unsigned long long)serial_hits, (unsigned long long)serial_occ);
    fflush(stdout);

    /* the same queries from n_threads threads at once, each on its own stream */
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)n_threads);
    Worker* ws = (Worker*)calloc((size_t)n_threads, sizeof(Worker));
    for (int t = 0; t < n_threads; ++t) {
        ws[t].sh = &sh;
        ws[t].index = t;
        if (pthread_create(&th[t], NULL, worker_main, &ws[t]) != 0) {
            fprintf(stderr, "pthread_create failed\n");
            return 2;
        }
    }
    uint64_t total_bad = 0;
    for (int t = 0; t < n_threads; ++t) {
        pthread_join(th[t], NULL);
        printf("thread %d: %d rounds x %u rays, %llu hits, %llu mismatches\n", t, rounds, sh.n,
               (unsigned long long)ws[t].hits, (unsigned long long)ws[t].mismatches);
        total_bad += ws[t].mismatches;
    }
    CHECK(sgpu_scene_check(sh.scene));
    CHECK(sgpu_scene_release(sh.scene));
    CHECK(sgpu_context_destroy(sh.ctx));
    printf("%s: %d threads, %llu mismatches\n", total_bad == 0 ? "ok" : "FAILED", n_threads,
           (unsigned long long)total_bad);
    for (int k = 0; k < N_RAY_FIELDS; ++k) free(sh.host_rays[k]);
    for (int k = 0; k < N_HIT_FIELDS; ++k) free(sh.want_hits[k]);
    free(sh.want_occ);
    free(soup);
    free(idx);
    free(th);
    free(ws);
    return total_bad == 0 ? 0 : 1;
}
