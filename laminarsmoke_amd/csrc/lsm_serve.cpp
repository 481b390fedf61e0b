// lsm_serve.cpp - device-resident serving of chemistry steps (include/lsm_serve.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <vector>

#include "../../include/lsm_serve.h"
#include "lsm_internal.hpp"

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Pending {
    uint64_t ticket;
    double t0, dt;
    double t_submit;
    std::vector<double> y;       // new fields shipped with the request (empty: continue from the state)
};

struct Case {
    lsm_handle *h = nullptr;
    int64_t n = 0;
    int ne = 0;
    size_t state_bytes = 0;      // device bytes while resident (state + per-call arrays)
    size_t handle_bytes = 0;     // the handle's own tables and workspaces (always held)
    double *h_y = nullptr;       // pinned host copy while not resident
    double *d_y = nullptr, *d_dt = nullptr;
    int32_t *d_status = nullptr, *d_cnt = nullptr, *d_order = nullptr;
    bool resident = false, in_flight = false, have_order = false;
    double last_used = 0.;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::deque<Pending> queue;
    Pending running{};
};

struct Done {
    uint64_t ticket;
    double latency, device;
    double start, end;           // device timestamps of the step, s after the server's epoch event
};

}  // namespace

struct lsm_server {
    size_t budget = 0;                        // device bytes for the cases' handles and resident states
    size_t used = 0;                          // resident case states (server allocations)
    size_t handle_bytes = 0;                  // the cases' handles (allocated by lsm_create), always held
    size_t peak = 0;                          // peak of handle_bytes + used
    hipEvent_t epoch = nullptr;               // time origin of the step timestamps
    int max_inflight = 1, inflight = 0;
    uint64_t next_ticket = 1;
    int64_t evictions = 0, admitted = 0, finished = 0;
    std::vector<Case> cases;
    std::deque<Done> done;
};

namespace {

void free_device(Case &c) {
    if (c.d_y) (void)hipFree(c.d_y);
    if (c.d_dt) (void)hipFree(c.d_dt);
    if (c.d_status) (void)hipFree(c.d_status);
    if (c.d_cnt) (void)hipFree(c.d_cnt);
    if (c.d_order) (void)hipFree(c.d_order);
    c.d_y = c.d_dt = nullptr;
    c.d_status = c.d_cnt = c.d_order = nullptr;
}

// Moves an idle resident case to pinned host memory.
int evict(lsm_server *s, Case &c) {
    if (hipMemcpyAsync(c.h_y, c.d_y, sizeof(double) * c.ne * c.n, hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
        hipStreamSynchronize(c.stream) != hipSuccess)
        return lsm_fail(LSM_E_HIP, "eviction copy failed");
    free_device(c);
    c.resident = false;
    c.have_order = false;
    s->used -= c.state_bytes;
    s->evictions++;
    return LSM_OK;
}

// Makes room for `need` bytes by evicting idle resident cases, least recently used
// first.  Returns false when the handles and the running cases alone leave too little room.
bool make_room(lsm_server *s, size_t need, int *rc) {
    *rc = LSM_OK;
    while (s->handle_bytes + s->used + need > s->budget) {
        int victim = -1;
        for (size_t i = 0; i < s->cases.size(); i++) {
            const Case &c = s->cases[i];
            if (c.resident && !c.in_flight && c.queue.empty() &&
                (victim < 0 || c.last_used < s->cases[victim].last_used))
                victim = (int)i;
        }
        if (victim < 0)   // only cases with queued steps left: take the least recently used of those
            for (size_t i = 0; i < s->cases.size(); i++) {
                const Case &c = s->cases[i];
                if (c.resident && !c.in_flight && (victim < 0 || c.last_used < s->cases[victim].last_used))
                    victim = (int)i;
            }
        if (victim < 0) return false;
        *rc = evict(s, s->cases[victim]);
        if (*rc) return false;
    }
    return true;
}

int make_resident(lsm_server *s, Case &c) {
    if (hipMalloc(&c.d_y, sizeof(double) * c.ne * c.n) != hipSuccess ||
        hipMalloc(&c.d_dt, sizeof(double) * c.n) != hipSuccess ||
        hipMalloc(&c.d_status, sizeof(int32_t) * c.n) != hipSuccess ||
        hipMalloc(&c.d_cnt, sizeof(int32_t) * LSM_NCOUNTERS * c.n) != hipSuccess ||
        hipMalloc(&c.d_order, sizeof(int32_t) * c.n) != hipSuccess) {
        free_device(c);
        return lsm_fail(LSM_E_NOMEM, "case state allocation failed");
    }
    if (hipMemcpyAsync(c.d_y, c.h_y, sizeof(double) * c.ne * c.n, hipMemcpyHostToDevice, c.stream) != hipSuccess)
        return lsm_fail(LSM_E_HIP, "case upload failed");
    c.resident = true;
    s->used += c.state_bytes;
    s->peak = std::max(s->peak, s->handle_bytes + s->used);
    return LSM_OK;
}

int launch_step(lsm_server *s, Case &c) {
    c.running = std::move(c.queue.front());
    c.queue.pop_front();
    if (!c.running.y.empty()) {      // fields from the CFD side (e.g. after transport): through the pinned copy
        if (hipStreamSynchronize(c.stream) != hipSuccess) return lsm_fail(LSM_E_HIP, "case stream failed");
        memcpy(c.h_y, c.running.y.data(), sizeof(double) * c.ne * c.n);
        std::vector<double>().swap(c.running.y);
        if (hipMemcpyAsync(c.d_y, c.h_y, sizeof(double) * c.ne * c.n, hipMemcpyHostToDevice, c.stream) != hipSuccess)
            return lsm_fail(LSM_E_HIP, "request upload failed");
    }
    if (lsm::launch_fill(c.d_dt, c.n, c.running.dt, c.stream)) return lsm_fail(LSM_E_HIP, "dt fill failed");
    if (hipEventRecord(c.ev0, c.stream) != hipSuccess) return lsm_fail(LSM_E_HIP, "event record failed");
    const double P = lsm_handle_pressure(c.h);
    int rc = c.have_order
                 ? lsm_integrate_batch_device_ordered(c.h, c.n, c.running.t0, P, c.d_dt, c.d_y, c.d_y, c.d_status,
                                                      nullptr, c.d_cnt, c.d_order, c.stream)
                 : lsm_integrate_batch_device(c.h, c.n, c.running.t0, P, c.d_dt, c.d_y, c.d_y, c.d_status, nullptr,
                                              c.d_cnt, c.stream);
    if (rc) return rc;
    // visiting order of the next step: this step's per-cell cost
    rc = lsm_cost_order(c.h, c.n, c.d_cnt, c.d_order, c.stream);
    if (rc) return rc;
    if (hipEventRecord(c.ev1, c.stream) != hipSuccess) return lsm_fail(LSM_E_HIP, "event record failed");
    c.have_order = true;
    c.in_flight = true;
    s->inflight++;
    s->admitted++;
    return LSM_OK;
}

// Admits queued steps, oldest submission first, while slots and memory allow.
int admit(lsm_server *s) {
    for (;;) {
        if (s->inflight >= s->max_inflight) return LSM_OK;
        int pick = -1;
        for (size_t i = 0; i < s->cases.size(); i++) {
            const Case &c = s->cases[i];
            if (c.in_flight || c.queue.empty()) continue;
            if (pick < 0 || c.queue.front().t_submit < s->cases[pick].queue.front().t_submit) pick = (int)i;
        }
        if (pick < 0) return LSM_OK;
        Case &c = s->cases[pick];
        if (!c.resident) {
            int rc = LSM_OK;
            if (!make_room(s, c.state_bytes, &rc)) return rc;   // wait for a running step to finish
            rc = make_resident(s, c);
            if (rc) return rc;
        }
        c.last_used = now_s();
        int rc = launch_step(s, c);
        if (rc) return rc;
    }
}

int retire(lsm_server *s) {
    for (Case &c : s->cases) {
        if (!c.in_flight) continue;
        const hipError_t q = hipEventQuery(c.ev1);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) return lsm_fail(LSM_E_HIP, "chemistry step failed");
        float ms = 0.f, m0 = 0.f, m1 = 0.f;
        (void)hipEventElapsedTime(&ms, c.ev0, c.ev1);
        (void)hipEventElapsedTime(&m0, s->epoch, c.ev0);
        (void)hipEventElapsedTime(&m1, s->epoch, c.ev1);
        const double t = now_s();
        s->done.push_back(Done{c.running.ticket, t - c.running.t_submit, ms * 1e-3, m0 * 1e-3, m1 * 1e-3});
        // the handle may have grown buffers during the step (plain-kernel workspace)
        const size_t hb = lsm_handle_device_bytes(c.h);
        s->handle_bytes = s->handle_bytes - c.handle_bytes + hb;
        c.handle_bytes = hb;
        s->peak = std::max(s->peak, s->handle_bytes + s->used);
        c.in_flight = false;
        c.last_used = t;
        s->inflight--;
        s->finished++;
    }
    return LSM_OK;
}

}  // namespace

extern "C" {

int lsm_server_create(size_t hbm_budget_bytes, int32_t max_inflight, lsm_server **out) {
    if (!out) return lsm_fail(LSM_E_INVALID, "null output");
    *out = nullptr;
    if (max_inflight < 1 || max_inflight > 64) return lsm_fail(LSM_E_INVALID, "max_inflight must be 1-64");
    lsm_server *s = new lsm_server();
    s->budget = hbm_budget_bytes;
    s->max_inflight = max_inflight;
    if (hipEventCreate(&s->epoch) != hipSuccess || hipEventRecord(s->epoch, nullptr) != hipSuccess ||
        hipEventSynchronize(s->epoch) != hipSuccess) {
        if (s->epoch) (void)hipEventDestroy(s->epoch);
        delete s;
        return lsm_fail(LSM_E_HIP, "server epoch event");
    }
    *out = s;
    return LSM_OK;
}

int lsm_server_destroy(lsm_server *s) {
    if (!s) return LSM_OK;
    for (Case &c : s->cases) {
        if (c.stream) (void)hipStreamSynchronize(c.stream);
        free_device(c);
        if (c.h_y) (void)hipHostFree(c.h_y);
        if (c.ev0) (void)hipEventDestroy(c.ev0);
        if (c.ev1) (void)hipEventDestroy(c.ev1);
        if (c.stream) (void)hipStreamDestroy(c.stream);
    }
    if (s->epoch) (void)hipEventDestroy(s->epoch);
    delete s;
    return LSM_OK;
}

int lsm_server_add_case(lsm_server *s, lsm_handle *h, int64_t n_cells, const double *y0, int32_t *case_id) {
    if (!s || !h || !y0 || !case_id) return lsm_fail(LSM_E_INVALID, "null argument");
    if (n_cells <= 0 || n_cells > (int64_t)0x7fffffff) return lsm_fail(LSM_E_INVALID, "cell count out of range");
    for (const Case &c : s->cases)
        if (c.h == h) return lsm_fail(LSM_E_INVALID, "handle already serves another case");
    Case c;
    c.h = h;
    c.n = n_cells;
    c.ne = lsm_handle_neq(h);
    c.state_bytes = (size_t)n_cells * (sizeof(double) * (c.ne + 1) + sizeof(int32_t) * (LSM_NCOUNTERS + 2));
    c.handle_bytes = lsm_handle_device_bytes(h);
    if (s->handle_bytes + c.handle_bytes + c.state_bytes > s->budget)
        return lsm_fail(LSM_E_NOMEM, "the cases' handles and this case's state exceed the HBM budget");
    if (hipHostMalloc((void **)&c.h_y, sizeof(double) * c.ne * n_cells, hipHostMallocDefault) != hipSuccess)
        return lsm_fail(LSM_E_NOMEM, "pinned host state");
    memcpy(c.h_y, y0, sizeof(double) * c.ne * n_cells);
    if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c.ev0) != hipSuccess || hipEventCreate(&c.ev1) != hipSuccess) {
        (void)hipHostFree(c.h_y);
        return lsm_fail(LSM_E_HIP, "stream / event creation failed");
    }
    s->handle_bytes += c.handle_bytes;
    s->peak = std::max(s->peak, s->handle_bytes + s->used);
    *case_id = (int32_t)s->cases.size();
    s->cases.push_back(c);
    return LSM_OK;
}

int lsm_server_submit(lsm_server *s, int32_t case_id, double t0, double dt, const double *y, uint64_t *ticket) {
    if (!s || !ticket) return lsm_fail(LSM_E_INVALID, "null argument");
    if (case_id < 0 || case_id >= (int32_t)s->cases.size()) return lsm_fail(LSM_E_INVALID, "unknown case");
    if (!(dt > 0.)) return lsm_fail(LSM_E_INVALID, "dt must be positive");
    if (!(t0 == t0) || t0 - t0 != 0.) return lsm_fail(LSM_E_INVALID, "t0 must be finite");
    Case &c = s->cases[case_id];
    *ticket = s->next_ticket++;
    Pending p{*ticket, t0, dt, now_s(), {}};
    if (y) p.y.assign(y, y + (size_t)c.ne * c.n);
    c.queue.push_back(std::move(p));
    return admit(s);
}

int lsm_server_poll(lsm_server *s, uint64_t *tickets, double *latency_s, double *device_s, double *start_s,
                    double *end_s, int32_t capacity, int32_t *n_done) {
    if (!s) return lsm_fail(LSM_E_INVALID, "null server");
    int rc = retire(s);
    if (rc) return rc;
    rc = admit(s);
    if (rc) return rc;
    int32_t k = 0;
    while (k < capacity && !s->done.empty()) {
        const Done d = s->done.front();
        s->done.pop_front();
        if (tickets) tickets[k] = d.ticket;
        if (latency_s) latency_s[k] = d.latency;
        if (device_s) device_s[k] = d.device;
        if (start_s) start_s[k] = d.start;
        if (end_s) end_s[k] = d.end;
        k++;
    }
    if (n_done) *n_done = k;
    return LSM_OK;
}

int lsm_server_drain(lsm_server *s) {
    if (!s) return lsm_fail(LSM_E_INVALID, "null server");
    for (;;) {
        int rc = retire(s);
        if (rc) return rc;
        rc = admit(s);
        if (rc) return rc;
        bool queued = false;
        for (const Case &c : s->cases) queued = queued || !c.queue.empty();
        if (s->inflight == 0) {
            if (!queued) return LSM_OK;
            return lsm_fail(LSM_E_NOMEM, "queued steps do not fit in the HBM budget");
        }
        for (const Case &c : s->cases)
            if (c.in_flight) {
                if (hipEventSynchronize(c.ev1) != hipSuccess) return lsm_fail(LSM_E_HIP, "chemistry step failed");
                break;
            }
    }
}

int lsm_server_read_state(lsm_server *s, int32_t case_id, double *y) {
    if (!s || !y) return lsm_fail(LSM_E_INVALID, "null argument");
    if (case_id < 0 || case_id >= (int32_t)s->cases.size()) return lsm_fail(LSM_E_INVALID, "unknown case");
    Case &c = s->cases[case_id];
    if (c.resident) {
        if (hipMemcpyAsync(y, c.d_y, sizeof(double) * c.ne * c.n, hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
            hipStreamSynchronize(c.stream) != hipSuccess)
            return lsm_fail(LSM_E_HIP, "state copy failed");
    } else {
        memcpy(y, c.h_y, sizeof(double) * c.ne * c.n);
    }
    return LSM_OK;
}

int lsm_server_stats(lsm_server *s, int64_t *out) {   // out[8]
    if (!s || !out) return lsm_fail(LSM_E_INVALID, "null argument");
    out[0] = (int64_t)s->used;
    out[1] = (int64_t)s->peak;
    out[2] = s->evictions;
    out[3] = s->admitted;
    out[4] = s->finished;
    out[5] = (int64_t)s->cases.size();
    out[6] = (int64_t)s->handle_bytes;
    out[7] = (int64_t)s->budget;
    return LSM_OK;
}

}  // extern "C"
