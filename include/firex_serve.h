/*
 * firex_serve.h — serving many independent Fire-X plumes on one MI355X.
 *
 * An addition: the reference runs one engine per page.  A server keeps many
 * plumes of different grid sizes resident in HBM, admits queued requests as
 * memory (an HBM byte budget) and slots free up, and advances every admitted
 * plume concurrently: each plume's steps go to one of a pool of HIP streams,
 * one step in flight per plume by default, with the part of each step after
 * `simulate` replayed from a captured hipGraph.  With batch_plumes, small
 * plumes (whose launches alone fill a fraction of the device) share streams
 * and step together, each Jacobi launch of their step covering all of them.  Requests that arrive while
 * others run join at the next pump (continuous batching).  Each plume's
 * kernels touch only its own buffers, so a plume's fields are bit-identical
 * whether it runs alone or in the mix.
 *
 * One server must not be used from two threads at once.
 */
#ifndef FIREX_SERVE_H
#define FIREX_SERVE_H

#include <stddef.h>
#include <stdint.h>

#include "firex.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct firex_server firex_server;

typedef struct firex_server_config {
  int32_t device;
  uint64_t memory_budget_bytes;  /* resident plume bytes must stay within this (0 = 3/4 of free HBM) */
  int32_t max_active;            /* resident plume slots                                     */
  int32_t streams;               /* HIP streams the plumes are spread over (1..8)            */
  int32_t plan;                  /* FIREX_PLAN_* for every plume                              */
  int32_t capture_graphs;        /* 1: replay the post-simulate part of a step as a hipGraph */
  int32_t max_inflight_steps;    /* steps issued but not finished per plume (>= 1; 0 = 1)    */
  int32_t jacobi_autotune;       /* 1: autotune the Jacobi plan per grid shape (cached)      */
  int32_t batch_plumes;          /* 1: small plumes (at most batch_max_cells cells) go to the
                                    first batch_streams streams, and those on one stream that
                                    share iteration counts, pressure storage and dt step
                                    together: each Jacobi launch of their step is one ring
                                    launch over all their tiles (16-wave ring plan of depth
                                    4 on 8-plane z chunks; grids the ring kernel cannot run
                                    step alone).  Larger plumes step alone on the other
                                    streams.                                                  */
  int32_t batch_max_cells;       /* 0 = 4 Mi cells                                            */
  int32_t batch_streams;         /* 0 = half of the streams (at least 1)                      */
  int32_t reserved[1];
} firex_server_config;

typedef struct firex_request {
  int32_t dims[3];
  int32_t pressure_iterations;
  int32_t correction_iterations;
  int32_t steps;                 /* fixed steps of dt seconds                                 */
  double dt;
  int32_t species_storage;       /* FIREX_STORAGE_*                                           */
  int32_t emitter_count;         /* 0: the reference emitter; else seeded emitters            */
  firex_emitter emitters[8];
  float jitter;
  float velocity_noise;
  uint64_t seed;
  int32_t keep_result;           /* 1: stay resident after the last step until released       */
  int32_t reserved[5];
} firex_request;

enum {
  FIREX_REQUEST_QUEUED = 0,
  FIREX_REQUEST_ACTIVE = 1,
  FIREX_REQUEST_DONE = 2,
  FIREX_REQUEST_FAILED = 3,
  FIREX_REQUEST_RELEASED = 4
};

typedef struct firex_request_status {
  int32_t state;                 /* FIREX_REQUEST_*                                           */
  int32_t steps_done;
  double submit_time;            /* seconds on the server clock                               */
  double admit_time;
  double finish_time;
  double latency_ms_p50;         /* per step: issue on the host to completion on the device   */
  double latency_ms_p99;
  double latency_ms_max;
  uint64_t device_bytes;
} firex_request_status;

typedef struct firex_server_stats {
  int32_t queued, active, done;
  int32_t failed;
  uint64_t resident_bytes;
  uint64_t memory_budget_bytes;
  int64_t steps_completed;
  double cell_updates;           /* sum over completed steps of the plume's cell count       */
  double elapsed_seconds;        /* server clock                                              */
  double latency_ms_p50;         /* over every completed step                                 */
  double latency_ms_p99;
  double latency_ms_max;
  int64_t grouped_steps;         /* batch_plumes: plume steps issued inside groups of two or
                                    more (solver steps that shared their Jacobi launches)     */
  int64_t group_steps;           /* batch_plumes: group steps issued                          */
} firex_server_stats;

int firex_server_create(const firex_server_config* config, firex_server** out);
void firex_server_destroy(firex_server* server);
/* Queue a request; *id identifies it afterwards.  Fails (FIREX_ERR_BUDGET)
 * only when the request could never fit the budget. */
int firex_server_submit(firex_server* server, const firex_request* request, int64_t* id);
/* Admit what fits, issue steps up to the in-flight limit, retire finished
 * steps and plumes.  Never blocks on the device. *busy receives 1 while
 * requests are queued or active (may be NULL). */
int firex_server_pump(firex_server* server, int32_t* busy);
/* Pump until every request is done (blocks). */
int firex_server_drain(firex_server* server);
int firex_server_request_status(firex_server* server, int64_t id, firex_request_status* out);
int firex_server_get_stats(firex_server* server, firex_server_stats* out);
/* Fields of a finished request kept resident with keep_result = 1. */
int firex_server_read_field(firex_server* server, int64_t id, int32_t field, void* dst, size_t bytes, int32_t dtype);
/* Free a finished request's device memory. */
int firex_server_release(firex_server* server, int64_t id);
/* Seconds on the server clock (the clock submit/admit/finish times use). */
double firex_server_clock(firex_server* server);

#ifdef __cplusplus
}
#endif

#endif /* FIREX_SERVE_H */
