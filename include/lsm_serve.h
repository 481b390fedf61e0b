/* lsm_serve.h - device-resident serving of chemistry steps from concurrent CFD cases.
 *
 * No counterpart in the reference, where every MPI rank integrates its own cells in
 * a serial loop (solvers/laminarSMOKE/unsteady/chemistry_DI.H:92).  Here several CFD
 * cases (each with its own mechanism handle and mesh) submit chemistry steps to one
 * GPU:
 *   - each case's state [ns+1][n_cells] lives in HBM while the case is resident;
 *   - admission is by an HBM budget over everything the cases pin on the device:
 *     every case's handle (its tables and kernel workspaces, allocated by lsm_create,
 *     and the buffers the handle grows, re-read after each step) always, and the
 *     state of each resident case (allocated by the server).  A step whose case is
 *     not resident waits until its state fits, idle resident cases being moved to
 *     pinned host memory (least recently used first) to make room;
 *   - steps of one case run in submission order on that case's stream; steps of
 *     different cases overlap on the device (up to max_inflight at once): as the
 *     workgroups of one step drain, those of the next admitted step fill the CUs;
 *   - every step after a case's first visits its cells in the cost order of the
 *     case's previous step (lsm_cost_order).
 * A step's result does not depend on what else runs: cells are integrated
 * independently (bit for bit with the FP64 kernels).
 * Not thread-safe: one host thread drives a server. */
#ifndef LSM_SERVE_H
#define LSM_SERVE_H

#include <stddef.h>
#include <stdint.h>

#include "laminarsmoke_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lsm_server lsm_server;

/* hbm_budget_bytes: device bytes the cases' handles and resident states may hold;
 * max_inflight: steps on the device at once (1-64). */
int lsm_server_create(size_t hbm_budget_bytes, int32_t max_inflight, lsm_server **out);
int lsm_server_destroy(lsm_server *s);

/* Registers a case: its handle (mechanism + ODE parameters, owned by the caller,
 * used by this case only, must outlive the server), its cell count and initial
 * state y0 (host, SoA [ns+1][n_cells], copied).  Fails with LSM_E_NOMEM if the
 * handles of all cases and this case's state exceed the budget. */
int lsm_server_add_case(lsm_server *s, lsm_handle *h, int64_t n_cells, const double *y0, int32_t *case_id);

/* Queues one chemistry step of every cell of a case from t0 (the case's time at the
 * start of the step, t - dt of its CFD step) over dt, at the pressure of the case's
 * handle (lsm_reactor_options::pressure); returns its ticket.
 * y (may be NULL): new fields for the case (host, SoA [ns+1][n_cells], copied at
 * once), e.g. the state after the CFD solver's transport step; NULL continues
 * from the case's current state. */
int lsm_server_submit(lsm_server *s, int32_t case_id, double t0, double dt, const double *y, uint64_t *ticket);

/* Non-blocking progress: retires finished steps, admits queued ones.  Writes up to
 * `capacity` finished tickets with their latency (submission to observed
 * completion, host clock), device time, and the device timestamps of the step's
 * start and end (s after an event recorded when the server was created; steps of
 * different cases whose [start, end] intervals intersect were on the device
 * together); *n_done = number written.  Any output array may be NULL. */
int lsm_server_poll(lsm_server *s, uint64_t *tickets, double *latency_s, double *device_s, double *start_s,
                    double *end_s, int32_t capacity, int32_t *n_done);

/* Blocks until every queued step has finished (results stay available to poll). */
int lsm_server_drain(lsm_server *s);

/* Copies a case's current state to y (host, SoA [ns+1][n_cells]); waits for the
 * case's running step. */
int lsm_server_read_state(lsm_server *s, int32_t case_id, double *y);

/* {resident state bytes, peak of handle + resident state bytes, evictions, steps
 * admitted, steps finished, cases, bytes held by the cases' handles, budget}. */
int lsm_server_stats(lsm_server *s, int64_t *out7);

#ifdef __cplusplus
}
#endif

#endif /* LSM_SERVE_H */
