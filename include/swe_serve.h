/*
 * libswe_serve.so -- continuous-batching forecast server over libswe_hip.
 *
 * No counterpart in the reference (which runs one forecast per process,
 * swe_code/main/main.c:294-392).  Requests (mesh, initial states of one or
 * more ensemble members, step count) are queued as they arrive; meshes stay
 * resident on the device; members are packed into batched contexts of the
 * same mesh (one weight read serves every member of a batch); each batch
 * runs on its own stream so batches overlap on the device; a member joins a
 * batch at the next chunk boundary when a slot is free and the HBM budget
 * allows, and leaves it as soon as its steps are done.  Every member's
 * arithmetic is independent of the other members and of chunking, so a
 * request's result is the same alone and in any mix.
 *
 * Threading: one host thread drives a server (submit / pump / completed).
 * Host buffers passed to swe_srv_add_mesh and swe_srv_submit are borrowed:
 * they must stay valid until the server is destroyed (mesh data) or the
 * request completes (states).  Return codes are SWE_HIP_* (swe_hip.h).
 */
#ifndef SWE_SERVE_H
#define SWE_SERVE_H

#include "swe_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct swe_srv swe_srv;

#define SWE_SRV_EIO (-5)       /* a request's result file could not be written */

typedef struct swe_srv_config {
    long long hbm_cap_bytes;   /* admission budget: static data + state buffers */
    int slots_per_batch;       /* member slots of one batched context (even)    */
    int member_block;          /* lane-group size of the ensemble kernel        */
    int chunk_steps;           /* RK steps per launch chunk (admission period)  */
    int max_batches_per_mesh;
    unsigned flags;            /* SWE_HIP_EXACT                                 */
    int device;
    int tiled;                 /* 1: batches of meshes with 31-point stencils use
                                  the member-lane LDS-tile kernel over a Hilbert-
                                  ordered copy (member_block 32, slots_per_batch
                                  a multiple of 32); 0: the member_block kernel */
} swe_srv_config;

typedef struct swe_srv_stats {
    long long member_steps_done;
    long long requests_done, requests_pending;
    long long hbm_in_use, hbm_peak;
    int batches, meshes_resident;
    long long chunks_launched;
    long long evictions;       /* idle batches destroyed to make room          */
} swe_srv_stats;

void swe_srv_default_config(swe_srv_config *cfg);
int swe_srv_create(const swe_srv_config *cfg, swe_srv **out);
void swe_srv_destroy(swe_srv *srv);

/* Registers a mesh (static data in the host layout of swe_hip_static);
 * its device copy is made at the first admission and then kept. */
int swe_srv_add_mesh(swe_srv *srv, const swe_hip_static *s, int *mesh_id);
/* The same with the host -> file node mapping (the RCM mapping of the host's
 * init pipeline, n_nodes real nodes; NULL = identity) that result files use. */
int swe_srv_add_mesh_ex(swe_srv *srv, const swe_hip_static *s, const int *mapping, int n_nodes,
                        int *mesh_id);

/* Queues a request: init[m] / result[m] are host-layout states of member m
 * (patch_size*4 doubles each).  Returns its id in *req_id. */
int swe_srv_submit(swe_srv *srv, int mesh_id, int members, const double *const *init,
                   double *const *result, int nsteps, long long tag, int *req_id);
/* The same, and when result_path is not NULL the server also writes the
 * request's result file when it completes (in swe_srv_pump): the history
 * format of swe_run (the reference's USE_HIST output, runtime_params.c:24-27)
 * -- int32 N, then per member an int32 step (= nsteps), f64 time and
 * f64 H[N][4] (u, v, w, h) in the file node order of the mesh's mapping.
 * A write failure makes swe_srv_pump return SWE_SRV_EIO. */
int swe_srv_submit_ex(swe_srv *srv, int mesh_id, int members, const double *const *init,
                      double *const *result, int nsteps, long long tag, const char *result_path,
                      int *req_id);

/* One non-blocking scheduling pass: retire members of batches whose chunk
 * finished, admit queued members, launch the next chunks. *progress is set
 * to the number of actions taken (0: nothing could be done right now). */
int swe_srv_pump(swe_srv *srv, int *progress);

/* Blocks until some in-flight chunk finishes (or returns at once if none). */
int swe_srv_wait(swe_srv *srv);

/* Drains up to max completion records: tag, latency (submit -> last member
 * done) and queueing delay (submit -> first member admitted), seconds. */
int swe_srv_completed(swe_srv *srv, long long *tags, double *latency_s, double *queue_s,
                      int max, int *count);

int swe_srv_get_stats(const swe_srv *srv, swe_srv_stats *st);

#ifdef __cplusplus
}
#endif
#endif
