// mel_serve.cpp - a request loop that keeps many independent MELODY 2D jobs
// on one GPU at once (continuous batching).
//
// The reference runs one model per process (Code/main.cpp: load, then the
// Euler_step loop of Code/main.cpp:141-238 until Tend); a batch of models is
// a series of such processes.  Here a job is one model container with a step
// count and an arrival time.  Every job gets its own context (its own device state and
// HIP stream), so the work of all admitted jobs overlaps on the device while
// each job's results stay exactly what the job gives alone: a context's
// launches depend only on its own state, and a run of steps gives the same
// state however it is cut into calls.
//
// One host thread drives the loop.  Each pass
//   1. retires the slices (runs of `steps_per_slice` steps) whose event has
//      completed, and marks jobs done when their last slice has;
//   2. admits arrived jobs in arrival order while their device footprint fits
//      the HBM budget (a job that does not fit waits, and so do the jobs
//      behind it, so large jobs are not starved; a job larger than the whole
//      budget fails); done jobs that the caller did not ask to keep give
//      their memory back only then, because freeing device memory waits for
//      the whole device.  Loading (parsing, upload, the first contact
//      detection) runs on this thread, while the running jobs' queued slices
//      keep the device busy;
//   3. keeps every running job `SLICE_DEPTH` slices ahead of the device.
// A job's footprint is known exactly once it is loaded; before that it is
// estimated from its container's size by the largest device/file ratio seen
// so far (the first estimate is DEFAULT_RATIO).
#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <string>
#include <thread>
#include <vector>

#include "melody2d.h"

namespace {

constexpr int SLICE_DEPTH = 2;          // slices queued per running job
constexpr double DEFAULT_RATIO = 4.0;   // device bytes per container byte, before any load

struct Slice {
    hipEvent_t ev;
    long long steps;
};

struct Job {
    std::string path;
    mel_precision precision = MEL_FP64;
    long long steps = 0, issued = 0, done = 0;
    int keep = 0;
    double arrival = 0., admitted = -1., finished = -1.;
    size_t file_bytes = 0, device_bytes = 0;
    int n_nodes = 0;
    mel_ctx* ctx = nullptr;
    std::deque<Slice> inflight;
    int state = MEL_JOB_WAITING;
    mel_status status = MEL_OK;
    std::string err;
};

}  // namespace

struct mel_server {
    int device = 0;
    size_t budget = 0, used = 0;
    long long slice = 0;
    double ratio = DEFAULT_RATIO;
    int graph_replay = -1;              // of the jobs loaded from now on (-1: the context's own choice)
    std::vector<Job> jobs;
    std::vector<size_t> order;          // job indices by arrival (stable)
    size_t next_admit = 0;              // first entry of `order` not yet admitted or failed
    std::vector<hipEvent_t> events;     // free slice events
    bool started = false;
    std::chrono::steady_clock::time_point t0;
    std::string err;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    hipError_t event(hipEvent_t* e) {
        if (!events.empty()) { *e = events.back(); events.pop_back(); return hipSuccess; }
        return hipEventCreateWithFlags(e, hipEventDisableTiming);
    }
    void fail(Job& j, mel_status st, const std::string& msg) {
        j.state = MEL_JOB_FAILED;
        j.status = st;
        j.err = msg;
        j.finished = now();
        // the job's queued work ends before its stream goes; its slice
        // events are destroyed, not reused (their last record was on that
        // stream)
        if (j.ctx) (void)mel_synchronize(j.ctx);
        for (Slice& s : j.inflight) (void)hipEventDestroy(s.ev);
        j.inflight.clear();
        release(j);
    }
    void release(Job& j) {
        if (j.ctx) {
            mel_destroy(j.ctx);
            j.ctx = nullptr;
            used -= std::min(used, j.device_bytes);
        }
    }
    // device footprint assumed before the job is loaded
    size_t estimate(const Job& j) const { return (size_t)(ratio * (double)j.file_bytes); }
};

namespace {

std::string ctx_error(mel_ctx* c) {
    const char* m = mel_last_error(c);
    return m ? std::string(m) : std::string();
}

// 1. completed slices
bool retire(mel_server& s) {
    bool moved = false;
    for (Job& j : s.jobs) {
        if (j.state != MEL_JOB_RUNNING) continue;
        while (!j.inflight.empty()) {
            hipError_t e = hipEventQuery(j.inflight.front().ev);
            if (e == hipErrorNotReady) break;
            if (e != hipSuccess) {
                s.fail(j, MEL_E_HIP, std::string("slice event: ") + hipGetErrorString(e));
                break;
            }
            j.done += j.inflight.front().steps;
            s.events.push_back(j.inflight.front().ev);
            j.inflight.pop_front();
            moved = true;
        }
        if (j.state == MEL_JOB_RUNNING && j.done == j.steps && j.inflight.empty()) {
            // flags raised by the last contact updates (non-finite state,
            // overflowed buffers) surface here
            mel_status st = mel_synchronize(j.ctx);
            if (st != MEL_OK) { s.fail(j, st, ctx_error(j.ctx)); continue; }
            j.state = MEL_JOB_DONE;
            j.finished = s.now();
        }
    }
    return moved;
}

// done jobs that the caller does not keep give their memory back
void reclaim(mel_server& s) {
    for (Job& j : s.jobs)
        if (j.state == MEL_JOB_DONE && !j.keep) s.release(j);
}

// 2. admission by HBM footprint, in arrival order
bool admit(mel_server& s) {
    bool moved = false;
    bool reclaimed = false;
    while (s.next_admit < s.order.size()) {
        Job& j = s.jobs[s.order[s.next_admit]];
        if (j.arrival > s.now()) break;
        const size_t need = s.estimate(j);
        if (s.used + need > s.budget) {
            if (!reclaimed) {
                reclaim(s);
                reclaimed = true;
                continue;
            }
            // wait for running jobs; with nothing else held, load the job
            // and judge its real footprint
            if (s.used > 0) break;
        }
        mel_ctx* c = nullptr;
        const double t_admit = s.now();
        mel_status st = mel_load_binary(j.path.c_str(), j.precision, s.device, &c);
        s.next_admit++;
        moved = true;
        j.admitted = t_admit;
        if (st != MEL_OK) { s.fail(j, st, ctx_error(nullptr)); continue; }
        j.ctx = c;
        if (s.graph_replay >= 0) mel_set_graph_replay(c, s.graph_replay);
        size_t bytes = 0;
        mel_get_device_bytes(c, &bytes);
        mel_info_t info;
        st = mel_get_info(c, &info);
        if (j.file_bytes) s.ratio = std::max(s.ratio, (double)bytes / (double)j.file_bytes);
        if (st == MEL_OK && bytes <= s.budget && s.used > 0 && s.used + bytes > s.budget) {
            // the estimate was low and other jobs hold the memory: the job
            // goes back to the head of the queue until they have finished
            // (its next estimate is at least its real footprint)
            mel_destroy(c);
            j.ctx = nullptr;
            j.admitted = -1.;
            s.next_admit--;
            break;
        }
        j.device_bytes = bytes;
        s.used += bytes;
        if (st != MEL_OK) { s.fail(j, st, ctx_error(c)); continue; }
        if (bytes > s.budget) {
            s.fail(j, MEL_E_CAPACITY, "job footprint " + std::to_string(bytes) + " B exceeds the HBM budget");
            continue;
        }
        j.n_nodes = info.n_nodes;
        j.state = MEL_JOB_RUNNING;
    }
    return moved;
}

// 3. keep each running job SLICE_DEPTH slices ahead
bool issue(mel_server& s) {
    bool moved = false;
    for (Job& j : s.jobs) {
        while (j.state == MEL_JOB_RUNNING && (int)j.inflight.size() < SLICE_DEPTH && j.issued < j.steps) {
            const long long n = std::min(s.slice, j.steps - j.issued);
            mel_status st = mel_euler_steps(j.ctx, (int)n);
            if (st != MEL_OK) { s.fail(j, st, ctx_error(j.ctx)); break; }
            void* stream = nullptr;
            mel_get_stream(j.ctx, &stream);
            Slice sl{nullptr, n};
            hipError_t e = s.event(&sl.ev);
            if (e == hipSuccess) e = hipEventRecord(sl.ev, (hipStream_t)stream);
            if (e != hipSuccess) {
                if (sl.ev) s.events.push_back(sl.ev);
                s.fail(j, MEL_E_HIP, std::string("slice event: ") + hipGetErrorString(e));
                break;
            }
            j.inflight.push_back(sl);
            j.issued += n;
            moved = true;
        }
    }
    return moved;
}

bool finished(const Job& j) {
    return j.state == MEL_JOB_DONE || j.state == MEL_JOB_FAILED || j.state == MEL_JOB_RELEASED;
}

}  // namespace

extern "C" {

mel_status mel_server_create(int hip_device, size_t hbm_budget, long long steps_per_slice, mel_server** out) {
    if (!out || hbm_budget == 0 || steps_per_slice <= 0 || steps_per_slice > (1LL << 30)) return MEL_E_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || hip_device < 0 || hip_device >= n) return MEL_E_ARG;
    mel_server* s = new mel_server;
    s->device = hip_device;
    s->budget = hbm_budget;
    s->slice = steps_per_slice;
    *out = s;
    return MEL_OK;
}

mel_status mel_server_submit(mel_server* s, const char* binary_path, mel_precision precision, long long nsteps,
                             double arrival_s, int keep, int* job) {
    if (!s || !binary_path || !job || nsteps < 0 || !(arrival_s >= 0.) ||
        (precision != MEL_FP64 && precision != MEL_FP32_MIXED))
        return MEL_E_ARG;
    struct stat st;
    if (stat(binary_path, &st) != 0) {
        s->err = std::string("cannot stat ") + binary_path;
        return MEL_E_IO;
    }
    Job j;
    j.path = binary_path;
    j.precision = precision;
    j.steps = nsteps;
    j.arrival = arrival_s;
    j.keep = keep != 0;
    j.file_bytes = (size_t)st.st_size;
    *job = (int)s->jobs.size();
    s->jobs.push_back(std::move(j));
    // keep the not-yet-admitted part of the arrival order sorted (stable)
    const size_t id = s->jobs.size() - 1;
    auto it = std::upper_bound(s->order.begin() + s->next_admit, s->order.end(), id, [&](size_t a, size_t b) {
        return s->jobs[a].arrival < s->jobs[b].arrival;
    });
    s->order.insert(it, id);
    return MEL_OK;
}

mel_status mel_server_run(mel_server* s, double timeout_s, int* n_unfinished) {
    if (!s || !(timeout_s >= 0.)) return MEL_E_ARG;
    if (hipSetDevice(s->device) != hipSuccess) return MEL_E_HIP;
    if (!s->started) {
        s->t0 = std::chrono::steady_clock::now();
        s->started = true;
    }
    const double end = s->now() + timeout_s;
    int left = 0;
    for (;;) {
        bool moved = retire(*s);
        moved |= admit(*s);
        moved |= issue(*s);
        left = 0;
        for (const Job& j : s->jobs) left += !finished(j);
        if (left == 0 || s->now() >= end) break;
        if (!moved) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    if (n_unfinished) *n_unfinished = left;
    return MEL_OK;
}

mel_status mel_server_job(const mel_server* s, int job, mel_job_info_t* info) {
    if (!s || !info || job < 0 || job >= (int)s->jobs.size()) return MEL_E_ARG;
    const Job& j = s->jobs[job];
    info->state = j.state;
    info->status = j.status;
    info->steps = j.steps;
    info->steps_done = j.done;
    info->arrival = j.arrival;
    info->admitted = j.admitted;
    info->finished = j.finished;
    info->device_bytes = j.device_bytes;
    info->n_nodes = j.n_nodes;
    return MEL_OK;
}

mel_status mel_server_job_context(mel_server* s, int job, mel_ctx** ctx) {
    if (!s || !ctx || job < 0 || job >= (int)s->jobs.size()) return MEL_E_ARG;
    const Job& j = s->jobs[job];
    if (j.state != MEL_JOB_DONE || !j.ctx) return MEL_E_ARG;
    *ctx = j.ctx;
    return MEL_OK;
}

mel_status mel_server_release(mel_server* s, int job) {
    if (!s || job < 0 || job >= (int)s->jobs.size()) return MEL_E_ARG;
    Job& j = s->jobs[job];
    if (j.state != MEL_JOB_DONE) return MEL_E_ARG;
    s->release(j);
    j.state = MEL_JOB_RELEASED;
    return MEL_OK;
}

mel_status mel_server_set_graph_replay(mel_server* s, int on) {
    if (!s) return MEL_E_ARG;
    s->graph_replay = on != 0;
    return MEL_OK;
}

mel_status mel_server_usage(const mel_server* s, size_t* used, size_t* budget) {
    if (!s || !used || !budget) return MEL_E_ARG;
    *used = s->used;
    *budget = s->budget;
    return MEL_OK;
}

const char* mel_server_job_error(const mel_server* s, int job) {
    if (!s) return "no server";
    if (job < 0 || job >= (int)s->jobs.size()) return s->err.c_str();
    return s->jobs[job].err.c_str();
}

void mel_server_destroy(mel_server* s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    for (Job& j : s->jobs) {
        if (j.ctx) (void)mel_synchronize(j.ctx);
        for (Slice& sl : j.inflight) s->events.push_back(sl.ev);
        j.inflight.clear();
        s->release(j);
    }
    for (hipEvent_t e : s->events) (void)hipEventDestroy(e);
    delete s;
}

}  // extern "C"
