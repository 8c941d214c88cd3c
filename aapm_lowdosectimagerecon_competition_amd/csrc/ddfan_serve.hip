// ddfan_serve.hip -- reconstruction request loop with continuous batching.
//
// A submission that finds no request waiting, a free slot and room in the HBM
// budget is admitted at once and its inputs are copied by the submitting
// thread straight to the device; any other is copied to host memory and
// queued.  A worker thread owns the device: it admits waiting requests in
// arrival order while a slot is free and their device footprint fits in the
// HBM budget (ddf::Admission), creates their plans and uploads the queued
// requests' inputs, and runs every admitted
// reconstruction together, one batched OS-SART iteration (launched directly, or
// replayed as a captured hipGraph with DDFAN_SERVER_GRAPH)
// at a time.  Requests join and leave at iteration boundaries, so a new one
// starts while others are mid-run, and each request's arithmetic is the same
// launch sequence it would get alone.  The worker keeps at most two batch
// iterations in flight and finishes requests whose last iteration completed
// (result copied back on a second stream while the next iteration runs).  A
// third thread frees finished requests' device memory (freeing waits for the
// GPU work in flight, which the worker must not) and gives their bytes back to
// the budget.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "ddfan_guard.h"
#include "ddfan_plan.h"
#include "ddfan_sched.h"
#include "ddfan_serve_hooks.h"

using namespace ddf;

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// memcpy split over up to 8 threads for large buffers.  A fresh destination
// is faulted in by the copying threads, so the first touch of its pages runs
// in parallel too (a 434 MB sinogram: 87 ms single-threaded into a new vector).
// A thread that cannot be started leaves its share to the calling thread.
void parallel_copy(void* dst, const void* src, size_t bytes) {
    constexpr size_t kMinPerThread = 16u << 20;
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>({8, hw, bytes / kMinPerThread});
    if (nt <= 1) {
        if (bytes) memcpy(dst, src, bytes);
        return;
    }
    const size_t step = ((bytes + nt - 1) / nt + 4095) & ~(size_t)4095;
    std::vector<std::thread> th;
    th.reserve(nt);
    for (size_t i = 0; i < nt; ++i) {
        const size_t a = i * step;
        if (a >= bytes) break;
        const size_t n = std::min(step, bytes - a);
        auto part = [=] { memcpy((char*)dst + a, (const char*)src + a, n); };
        try {
            th.emplace_back(part);
        } catch (...) {
            part();
        }
    }
    for (auto& t : th) t.join();
}

// Host float buffer without value-initialisation (filled by a copy).
struct HostFloats {
    std::unique_ptr<float[]> p;
    size_t n = 0;
    bool empty() const { return n == 0; }
    size_t size() const { return n; }
    float* data() { return p.get(); }
    const float* data() const { return p.get(); }
    bool alloc(size_t count) {
        p.reset(count ? new (std::nothrow) float[count] : nullptr);
        n = p ? count : 0;
        return count == 0 || p != nullptr;
    }
    bool assign_copy(const float* src, size_t count) {
        if (!alloc(count)) return false;
        parallel_copy(p.get(), src, count * sizeof(float));
        return true;
    }
    void release() {
        p.reset();
        n = 0;
    }
};

// Device state of a finished request, freed by the server's reaper thread.
struct Dead {
    ddfan_plan* plan = nullptr;
    void* meas = nullptr;
    void* x = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    size_t bytes = 0;
};

struct Request {
    long long id = 0;
    // inputs (host copies)
    ddfan_geom G{};
    std::vector<float> xds, yds, hangs;
    HostFloats meas, x0;
    std::vector<uint8_t> mask;
    bool preloaded = false;  // meas (and x0, if given) already copied to d_meas / d_x
    bool x0Given = false;
    int PN = 0, numIter = 0, numOS = 1, fista = 0;
    size_t bytes = 0;
    // device state while admitted
    ddfan_plan* plan = nullptr;
    DBuf d_meas, d_x;
    OsJob job;
    int left = 0;                    // iterations not yet enqueued
    int steps = 1;                   // iterations of the current batch step
    hipEvent_t uploaded = nullptr;   // recorded on the copy stream after each input upload
    hipEvent_t lastIter = nullptr;   // recorded after its last iteration
    // outcome
    int status = DDFAN_OK;
    std::string err;
    bool done = false;
    bool claimed = false;        // a wait has taken the request (it is collected once)
    float* dst = nullptr;        // the waiting caller's output, registered before completion
    bool inDst = false;          // the result was copied straight into dst
    HostFloats result;
    Clock::time_point tSubmit, tAdmit, tDone;

    ~Request() { release_device(); }
    // Moves the device state out (the request keeps its host side).
    Dead take_device() {
        Dead d;
        d.plan = plan;
        plan = nullptr;
        d.meas = d_meas.detach();
        d.x = d_x.detach();
        d.ev[0] = uploaded;
        d.ev[1] = lastIter;
        uploaded = lastIter = nullptr;
        d.bytes = bytes;
        return d;
    }
    void release_device() {
        if (uploaded) (void)hipEventDestroy(uploaded);
        uploaded = nullptr;
        if (lastIter) (void)hipEventDestroy(lastIter);
        lastIter = nullptr;
        d_meas.release();
        d_x.release();
        if (plan) ddfan_plan_destroy(plan);
        plan = nullptr;
    }
};

}  // namespace

struct ddfan_server {
    int device = 0, mode = 0;
    BatchShape shape;
    hipStream_t stream = nullptr, copyStream = nullptr;
    hipStream_t upStream = nullptr;  // inputs of requests admitted at submission
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    ddf::Admission adm;
    std::map<long long, std::unique_ptr<Request>> reqs;  // every request not yet collected
    std::vector<Request*> ready;  // admitted at submission, inputs on the device
    std::deque<Dead> dead;        // finished, device memory not yet freed
    std::condition_variable cvDead;
    long long nextId = 1;
    bool stop = false, reaperStop = false;
    ddf::ServerHooks hooks;  // how the worker checks GPU completion (ddfan_serve_hooks.h)
    bool graphs_ = false;    // DDFAN_SERVER_GRAPH: replay each membership's iteration as a hipGraph
    const Clock::time_point t0 = Clock::now();  // the admission policy's clock
    double maxWaitS = 10.0;                      // DDFAN_SERVER_MAX_WAIT_MS
    double clock_s(Clock::time_point t) const {
        return std::chrono::duration<double>(t - t0).count();
    }
    // With graphs on, the worker's stream capture and graph instantiation, the
    // reaper's frees (hipFree waits for all device work, the capturing stream
    // included) and the submit path's allocations and copies hold this, so no
    // other library thread calls the runtime while a capture is in progress.
    std::mutex captureMu;
    std::unique_lock<std::mutex> capture_guard() {
        return graphs_ ? std::unique_lock<std::mutex>(captureMu) : std::unique_lock<std::mutex>();
    }
    std::thread worker, reaper;
    // statistics
    long long completed = 0, failed = 0;
    double busyVoxelViews = 0.0;

    ddfan_server(size_t budget, int slots) : adm(budget, slots) {}
    void loop();
    void reap();
};

namespace {

// Iterations of one request per batch step of the request loop, at most.
constexpr int kMaxSteps = 8;

// Work of one OS-SART iteration of a request: voxel-views per operator.
double request_work(const Request& R) {
    return (double)R.G.XN * R.G.YN * R.G.SLN * (double)R.PN;
}

// Device bytes one request holds while admitted: its plan with OS-SART state,
// the measured sinogram and the iterate.
int request_bytes(const ddfan_geom* G, int PN, int mode, int numOS, size_t* bytes) {
    size_t b = 0;
    TRY(ddfan_plan_footprint(G, PN, mode, 0, numOS, &b));
    b += sizeof(float) * (size_t)PN * G->DNU * G->SLN;
    b += sizeof(float) * (size_t)G->XN * G->YN * G->SLN;
    *bytes = b;
    return DDFAN_OK;
}

// Creates the plan, uploads the initial iterate on `cst`, enqueues the OS-SART
// set-up on `st` after it, then uploads the sinogram on `cst` while the set-up
// runs; `st` waits for that upload before anything enqueued after this call.
// The host inputs stay alive until `uploaded` has completed.
int admit_request(Request& R, int mode, hipStream_t st, hipStream_t cst) {
    ddfan_geom G = R.G;
    G.xds = R.xds.data();
    G.yds = R.yds.data();
    TRY(ddfan_plan_create(&G, R.hangs.data(), R.PN, R.mask.empty() ? nullptr : R.mask.data(), mode,
                          0, &R.plan));
    int tuned = 0;  // recorded launch shapes for this size of subset, if any
    TRY(ddfan_plan_apply_tuned(R.plan, subset_views(R.PN, R.numOS, 0, nullptr), &tuned));
    const size_t nv = (size_t)G.XN * G.YN * G.SLN;
    if (!R.preloaded) {
        HIPCALL(R.d_meas.reserve(R.meas.size() * sizeof(float)));
        HIPCALL(R.d_x.reserve(nv * sizeof(float)));
    }
    HIPCALL(hipEventCreateWithFlags(&R.uploaded, hipEventDisableTiming));
    if (!R.x0Given) {
        HIPCALL(launch_fill(R.d_x.as<float>(), 0.0f, nv, st));
    } else if (!R.preloaded) {
        HIPCALL(hipMemcpyAsync(R.d_x.p, R.x0.data(), nv * sizeof(float), hipMemcpyHostToDevice, cst));
        HIPCALL(hipEventRecord(R.uploaded, cst));
        HIPCALL(hipStreamWaitEvent(st, R.uploaded, 0));  // the FISTA copy reads it
    }
    R.job.P = R.plan;
    R.job.meas = R.d_meas.as<float>();
    R.job.x = R.d_x.as<float>();
    R.job.numIter = R.numIter;
    R.job.numOS = R.numOS;
    R.job.fista = R.fista;
    TRY(ossart_prepare(R.job, st));  // weights only: the sinogram is not read yet
    if (!R.preloaded) {
        HIPCALL(hipMemcpyAsync(R.d_meas.p, R.meas.data(), R.meas.size() * sizeof(float),
                               hipMemcpyHostToDevice, cst));
        HIPCALL(hipEventRecord(R.uploaded, cst));
        HIPCALL(hipStreamWaitEvent(st, R.uploaded, 0));
    }
    HIPCALL(hipEventCreateWithFlags(&R.lastIter, hipEventDisableTiming));
    R.left = R.numIter;
    return DDFAN_OK;
}

// Copies a request admitted at submission from the caller's buffers to its
// device buffers on `up` (the submitting thread, with the server's device
// current) and waits for the copies.
int upload_inputs(Request& R, const float* meas, size_t nm, const float* x0, size_t nv,
                  int device, hipStream_t up) {
    int prev = -1;
    HIPCALL(hipGetDevice(&prev));
    struct Restore {
        int d;
        ~Restore() { (void)hipSetDevice(d); }
    } restore{prev};
    HIPCALL(hipSetDevice(device));
    HIPCALL(R.d_meas.reserve(nm * sizeof(float)));
    HIPCALL(R.d_x.reserve(nv * sizeof(float)));
    HIPCALL(hipMemcpyAsync(R.d_meas.p, meas, nm * sizeof(float), hipMemcpyHostToDevice, up));
    if (x0) HIPCALL(hipMemcpyAsync(R.d_x.p, x0, nv * sizeof(float), hipMemcpyHostToDevice, up));
    HIPCALL(hipStreamSynchronize(up));
    R.preloaded = true;
    return DDFAN_OK;
}

}  // namespace

void ddfan_server::loop() {
    (void)hipSetDevice(device);
    std::vector<Request*> running;   // admitted, iterations left to enqueue
    std::vector<Request*> finishing; // last iteration enqueued
    const bool useGraphs = graphs_;  // fixed before the first submission
    IterGraph graphs[3];             // current graph + up to two with launches in flight
    std::vector<OsJob*> jobs;        // the running requests' jobs, in launch order
    hipEvent_t iterEv[2] = {nullptr, nullptr};
    (void)hipEventCreateWithFlags(&iterEv[0], hipEventDisableTiming);
    (void)hipEventCreateWithFlags(&iterEv[1], hipEventDisableTiming);
    int graphIdx[2] = {-1, -1};  // graph slot each in-flight iteration used
    int cur = -1;                // slot of the graph matching `running`
    long long iterNo = 0;
    bool inFlight[2] = {false, false};

    // Completes a request and hands its device memory to the reaper; its
    // waiter may destroy the request as soon as it is marked done.
    auto finish = [&](Request* R, int status, const std::string& err) {
        Dead d = R->take_device();
        std::lock_guard<std::mutex> lk(mu);
        R->status = status;
        R->err = err;
        R->done = true;
        R->tDone = Clock::now();
        if (status == DDFAN_OK)
            ++completed;
        else
            ++failed;
        dead.push_back(d);
        cvDone.notify_all();
        cvDead.notify_one();
    };
    auto fail_all_running = [&](int status, const std::string& err) {
        for (Request* R : running) finish(R, status, err);
        for (Request* R : finishing) finish(R, status, err);
        running.clear();
        finishing.clear();
    };
    // A failed completion check means the device reported an error for work
    // in flight: every admitted request fails with it (none of their results
    // can be trusted) and the next iteration is captured afresh.
    auto device_failed = [&](hipError_t e, const char* what) {
        fail_all_running(DDFAN_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
        inFlight[0] = inFlight[1] = false;
        graphIdx[0] = graphIdx[1] = -1;
        cur = -1;
    };
    // Copies back and completes the requests whose last iteration has run:
    // into the output of a caller already waiting for it, else into a host
    // buffer of the request.  Returns false after a device error.
    auto collect = [&]() -> bool {
        for (size_t i = 0; i < finishing.size();) {
            Request* R = finishing[i];
            const hipError_t q = hooks.eventQuery(R->lastIter);
            if (q == hipErrorNotReady) {
                ++i;
                continue;
            }
            if (q != hipSuccess) {
                device_failed(q, "checking a request's last iteration");
                return false;
            }
            const size_t nv = (size_t)R->G.XN * R->G.YN * R->G.SLN;
            int st = DDFAN_OK;
            std::string err;
            float* dst = nullptr;
            {
                // A registered output stays valid: its caller blocks until done.
                std::lock_guard<std::mutex> lk(mu);
                dst = R->dst;
                R->inDst = dst != nullptr;
            }
            if (dst) {
                if (hipMemcpyAsync(dst, R->d_x.p, nv * sizeof(float), hipMemcpyDeviceToHost,
                                   copyStream) != hipSuccess ||
                    hooks.streamSync(copyStream) != hipSuccess) {
                    st = DDFAN_EDEVICE;
                    err = "result copy failed";
                }
            } else if (!R->result.alloc(nv)) {
                st = DDFAN_ENOMEM;
                err = "host allocation of the result failed";
            } else if (hipMemcpyAsync(R->result.data(), R->d_x.p, nv * sizeof(float),
                                      hipMemcpyDeviceToHost, copyStream) != hipSuccess ||
                       hooks.streamSync(copyStream) != hipSuccess) {
                st = DDFAN_EDEVICE;
                err = "result copy failed";
            }
            finish(R, st, err);
            finishing.erase(finishing.begin() + i);
        }
        return true;
    };

    for (;;) {
        // admission
        std::vector<Request*> admitted;
        {
            std::unique_lock<std::mutex> lk(mu);
            for (;;) {
                for (Request* R : ready) admitted.push_back(R);
                ready.clear();
                for (long long id : adm.admit(clock_s(Clock::now()))) {
                    Request* R = reqs[id].get();
                    R->tAdmit = Clock::now();
                    admitted.push_back(R);
                }
                const bool idle = running.empty() && finishing.empty() && admitted.empty();
                if (!idle) break;
                if (stop && adm.waiting() == 0) break;
                cvWork.wait(lk);
            }
            if (stop && running.empty() && finishing.empty() && admitted.empty() &&
                adm.waiting() == 0)
                break;
        }
        bool changed = false;
        for (Request* R : admitted) {
            const int rc = admit_request(*R, mode, stream, copyStream);
            if (rc != DDFAN_OK) {
                // the request's set-up and upload may still be in flight
                (void)hipStreamSynchronize(stream);
                (void)hipStreamSynchronize(copyStream);
                finish(R, rc, ddfan_last_error());
                continue;
            }
            if (R->left == 0) {  // zero iterations: the result is the initial iterate
                (void)hipEventRecord(R->lastIter, stream);
                finishing.push_back(R);
            } else {
                running.push_back(R);
                changed = true;
            }
        }
        if (running.empty()) {
            const hipError_t e = hooks.streamSync(stream);
            if (e != hipSuccess)
                device_failed(e, "waiting for the last iterations");
            else
                collect();
            continue;
        }
        // one batch iteration of every running request
        const int slot = (int)(iterNo & 1);
        if (inFlight[slot]) {  // keep at most two iterations in flight
            const hipError_t e = hooks.eventSync(iterEv[slot]);
            inFlight[slot] = false;
            graphIdx[slot] = -1;
            if (e != hipSuccess) {
                device_failed(e, "waiting for a batch iteration");
                continue;
            }
        }
        if (!useGraphs) {
            // Direct launches of the same kernels: no stream capture is ever in
            // progress on the worker while other threads allocate, copy or free.
            // One batch step runs `steps` iterations of each request: the
            // largest running request one, a smaller one about as many as fit
            // in the same work (at most kMaxSteps and what it has left), each
            // step's iterations as batch iterations of the requests still due.
            double wmax = 0.0;
            for (Request* R : running) wmax = std::max(wmax, request_work(*R));
            int smax = 1;
            for (Request* R : running) {
                const double r = wmax / std::max(request_work(*R), 1.0);
                R->steps = std::min(R->left, (int)std::min<double>(kMaxSteps, std::max(1.0, r)));
                smax = std::max(smax, R->steps);
            }
            hipError_t e = hipSuccess;
            for (int k = 0; k < smax && e == hipSuccess; ++k) {
                jobs.clear();
                for (Request* R : running)
                    if (R->steps > k) jobs.push_back(&R->job);
                e = enqueue_batch_iteration(jobs.data(), (int)jobs.size(), mode, shape, stream,
                                            nullptr);
            }
            if (e != hipSuccess) {
                (void)hipStreamSynchronize(stream);
                fail_all_running(DDFAN_EDEVICE, std::string("batch iteration launch: ") +
                                                    hipGetErrorString(e));
                continue;
            }
        } else {
            for (Request* R : running) R->steps = 1;
            if (changed || cur < 0) {
                int g = 0;  // a slot that no in-flight iteration uses
                while (g == graphIdx[0] || g == graphIdx[1]) ++g;
                jobs.clear();
                for (Request* R : running) jobs.push_back(&R->job);
                int rc;
                {
                    auto cg = capture_guard();
                    rc = graphs[g].capture(jobs.data(), (int)jobs.size(), mode, shape, stream);
                }
                if (rc != DDFAN_OK) {
                    (void)hipStreamSynchronize(stream);
                    fail_all_running(rc, ddfan_last_error());
                    cur = -1;
                    continue;
                }
                cur = g;
            }
            if (graphs[cur].launch(stream) != DDFAN_OK) {
                (void)hipStreamSynchronize(stream);
                fail_all_running(DDFAN_EDEVICE, ddfan_last_error());
                cur = -1;
                continue;
            }
        }
        (void)hipEventRecord(iterEv[slot], stream);
        inFlight[slot] = true;
        graphIdx[slot] = cur;
        ++iterNo;
        {
            std::lock_guard<std::mutex> lk(mu);
            for (Request* R : running) busyVoxelViews += 2.0 * request_work(*R) * R->steps;
        }
        for (size_t i = 0; i < running.size();) {
            Request* R = running[i];
            R->left -= R->steps;
            if (R->left == 0) {
                (void)hipEventRecord(R->lastIter, stream);
                finishing.push_back(R);
                running.erase(running.begin() + i);
                cur = -1;  // membership changes: capture again next time
            } else {
                ++i;
            }
        }
        // host copies of inputs already on the device are no longer needed
        for (auto* list : {&running, &finishing})
            for (Request* R : *list)
                if (!R->meas.empty() && hipEventQuery(R->uploaded) == hipSuccess) {
                    R->meas.release();
                    R->x0.release();
                }
        // wait for the previous iteration, then complete finished requests
        const int prev = slot ^ 1;
        if (inFlight[prev]) {
            const hipError_t e = hooks.eventSync(iterEv[prev]);
            inFlight[prev] = false;
            graphIdx[prev] = -1;
            if (e != hipSuccess) {
                device_failed(e, "waiting for a batch iteration");
                continue;
            }
        }
        collect();
    }
    const hipError_t e = hooks.streamSync(stream);
    if (e != hipSuccess)
        device_failed(e, "waiting for the last iterations");
    else
        collect();
    for (auto& g : graphs) g.reset();
    (void)hipEventDestroy(iterEv[0]);
    (void)hipEventDestroy(iterEv[1]);
}

// Frees finished requests' device memory, then returns their bytes to the
// budget (the worker admits waiting requests against it).
void ddfan_server::reap() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        cvDead.wait(lk, [&] { return !dead.empty() || reaperStop; });
        if (dead.empty()) break;  // stopping, nothing left
        Dead d = dead.front();
        dead.pop_front();
        lk.unlock();
        {
            auto cg = capture_guard();
            if (d.plan) (void)ddfan_plan_destroy(d.plan);
            if (d.meas) (void)hipFree(d.meas);
            if (d.x) (void)hipFree(d.x);
            for (hipEvent_t e : d.ev)
                if (e) (void)hipEventDestroy(e);
        }
        lk.lock();
        adm.release(d.bytes);
        cvWork.notify_all();
    }
}

void ddf::set_server_hooks(ddfan_server* S, const ServerHooks& h) {
    std::lock_guard<std::mutex> lk(S->mu);
    S->hooks = h;
}

extern "C" {

int ddfan_server_create(size_t hbm_budget, int max_slots, int sat_mode, ddfan_server** out) {
    return ddf::guarded([&]() -> int {
        CHECK(out, DDFAN_EINVAL, "out is NULL");
        *out = nullptr;
        CHECK(max_slots >= 1, DDFAN_EINVAL, "max_slots must be >= 1");
        CHECK(sat_mode >= 0 && sat_mode <= 2, DDFAN_EINVAL, "sat_mode must be 0, 1 or 2");
        int dev;
        TRY(check_device(&dev));
        if (hbm_budget == 0) {
            size_t fb = 0, tb = 0;
            HIPCALL(hipMemGetInfo(&fb, &tb));
            hbm_budget = fb / 10 * 8;  // 80% of what is free now
        }
        ddfan_server* S = new (std::nothrow) ddfan_server(hbm_budget, max_slots);
        CHECK(S, DDFAN_ENOMEM, "host allocation failed");
        S->device = dev;
        S->mode = sat_mode;
        hipError_t e = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&S->copyStream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&S->upStream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            if (S->stream) (void)hipStreamDestroy(S->stream);
            if (S->copyStream) (void)hipStreamDestroy(S->copyStream);
            delete S;
            HIPCALL(e);
        }
        try {
            S->reaper = std::thread([S] { S->reap(); });
            S->worker = std::thread([S] { S->loop(); });
        } catch (...) {
            if (S->reaper.joinable()) {
                {
                    std::lock_guard<std::mutex> lk(S->mu);
                    S->reaperStop = true;
                    S->cvDead.notify_all();
                }
                S->reaper.join();
            }
            (void)hipStreamDestroy(S->stream);
            (void)hipStreamDestroy(S->copyStream);
            (void)hipStreamDestroy(S->upStream);
            delete S;
            return fail(DDFAN_ENOMEM, "cannot start the worker threads");
        }
        *out = S;
        return DDFAN_OK;
    });
}

int ddfan_server_set_param(ddfan_server* S, int param, long long value) {
    return ddf::guarded([&]() -> int {
        CHECK(S, DDFAN_EINVAL, "server is NULL");
        std::lock_guard<std::mutex> lk(S->mu);
        CHECK(S->nextId == 1, DDFAN_EINVAL, "server parameters are set before the first submission");
        switch (param) {
            case DDFAN_SERVER_GRAPH:
                CHECK(value == 0 || value == 1, DDFAN_EINVAL, "DDFAN_SERVER_GRAPH takes 0 or 1");
                S->graphs_ = value != 0;
                return DDFAN_OK;
            case DDFAN_SERVER_POLICY:
                CHECK(value == DDFAN_POLICY_FIFO || value == DDFAN_POLICY_SHORTEST, DDFAN_EINVAL,
                      "DDFAN_SERVER_POLICY takes DDFAN_POLICY_FIFO or DDFAN_POLICY_SHORTEST");
                S->adm.set_policy((int)value, S->maxWaitS);
                return DDFAN_OK;
            case DDFAN_SERVER_MAX_WAIT_MS:
                CHECK(value >= 0, DDFAN_EINVAL, "DDFAN_SERVER_MAX_WAIT_MS must be >= 0");
                S->maxWaitS = (double)value * 1e-3;
                S->adm.set_policy(S->adm.policy(), S->maxWaitS);
                return DDFAN_OK;
            default:
                return fail(DDFAN_EINVAL, "unknown server parameter %d", param);
        }
    });
}

int ddfan_server_submit(ddfan_server* S, const ddfan_geom* G, const float* hangs, int PN,
                        const unsigned char* mask, const float* meas, const float* x0,
                        int numIter, int numOS, int useFISTA, long long* id) {
    return ddf::guarded([&]() -> int {
        const Clock::time_point arrival = Clock::now();
        CHECK(S && G && hangs && meas && id, DDFAN_EINVAL, "NULL argument");
        CHECK(G->XN >= 1 && G->YN >= 1 && G->SLN >= 1 && G->DNU >= 2 && G->xds && G->yds,
              DDFAN_EINVAL, "bad geometry");
        CHECK(PN >= 1 && numIter >= 0 && numOS >= 1 && numOS <= PN, DDFAN_EINVAL,
              "bad view or iteration counts (PN %d numIter %d numOS %d)", PN, numIter, numOS);
        size_t bytes = 0;
        TRY(request_bytes(G, PN, S->mode, numOS, &bytes));
        auto R = std::make_unique<Request>();
        R->G = *G;
        R->xds.assign(G->xds, G->xds + G->DNU);
        R->yds.assign(G->yds, G->yds + G->DNU);
        R->hangs.assign(hangs, hangs + (size_t)PN * G->SLN);
        if (mask) R->mask.assign(mask, mask + (size_t)G->XN * G->YN);
        R->PN = PN;
        R->numIter = numIter;
        R->numOS = numOS;
        R->fista = useFISTA ? 1 : 0;
        R->x0Given = x0 != nullptr;
        R->bytes = bytes;
        R->tSubmit = arrival;
        const size_t nm = (size_t)PN * G->DNU * G->SLN, nv = (size_t)G->XN * G->YN * G->SLN;
        bool now = false;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            CHECK(!S->stop, DDFAN_EINVAL, "server is shutting down");
            CHECK(bytes <= S->adm.budget(), DDFAN_EBUSY,
                  "request needs %zu device bytes, more than the whole budget of %zu", bytes,
                  S->adm.budget());
            R->id = S->nextId++;
            now = S->adm.admit_now(bytes);
        }
        if (now) {
            // Admitted: copy the inputs from the caller's buffers to the device.
            R->tAdmit = Clock::now();
            int rc;
            {
                auto cg = S->capture_guard();
                rc = upload_inputs(*R, meas, nm, x0, nv, S->device, S->upStream);
            }
            std::lock_guard<std::mutex> lk(S->mu);
            if (rc != DDFAN_OK) {
                S->adm.release(bytes);
                S->cvWork.notify_all();
                return rc;
            }
            const long long rid = R->id;
            *id = rid;
            S->ready.push_back(R.get());
            S->reqs[rid] = std::move(R);
            S->cvWork.notify_all();
            return DDFAN_OK;
        }
        CHECK(R->meas.assign_copy(meas, nm) && (!x0 || R->x0.assign_copy(x0, nv)), DDFAN_ENOMEM,
              "host copy of the request's inputs failed");
        std::lock_guard<std::mutex> lk(S->mu);
        CHECK(!S->stop, DDFAN_EINVAL, "server is shutting down");
        // work of the whole request (voxel-views of its iterations) orders
        // shortest-first admission
        const double work = request_work(*R) * std::max(numIter, 1);
        const bool queued = S->adm.submit(R->id, bytes, work, S->clock_s(arrival));
        CHECK(queued, DDFAN_EBUSY, "request rejected by the HBM budget");
        const long long rid = R->id;
        *id = rid;
        S->reqs[rid] = std::move(R);
        S->cvWork.notify_all();
        return DDFAN_OK;
    });
}

int ddfan_server_poll(ddfan_server* S, long long id, int* done) {
    return ddf::guarded([&]() -> int {
        CHECK(S && done, DDFAN_EINVAL, "NULL argument");
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->reqs.find(id);
        CHECK(it != S->reqs.end(), DDFAN_EINVAL, "unknown request %lld", id);
        *done = it->second->done ? 1 : 0;
        return DDFAN_OK;
    });
}

int ddfan_server_wait(ddfan_server* S, long long id, float* x, double* queue_ms,
                      double* latency_ms) {
    return ddf::guarded([&]() -> int {
        CHECK(S, DDFAN_EINVAL, "server is NULL");
        std::unique_ptr<Request> R;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            auto it = S->reqs.find(id);
            CHECK(it != S->reqs.end(), DDFAN_EINVAL, "unknown request %lld", id);
            Request* r = it->second.get();
            CHECK(!r->claimed, DDFAN_EINVAL, "request %lld is already being waited for", id);
            r->claimed = true;
            if (!r->done) r->dst = x;  // the loop copies the result straight into x
            S->cvDone.wait(lk, [&] { return r->done; });
            R = std::move(it->second);
            S->reqs.erase(it);
        }
        if (queue_ms) *queue_ms = R->status == DDFAN_OK ? ms_between(R->tSubmit, R->tAdmit) : 0.0;
        if (latency_ms) *latency_ms = ms_between(R->tSubmit, R->tDone);
        if (R->status != DDFAN_OK)
            return fail(R->status, "request %lld failed: %s", id, R->err.c_str());
        if (x && !R->inDst) parallel_copy(x, R->result.data(), R->result.size() * sizeof(float));
        return DDFAN_OK;
    });
}

int ddfan_server_stats(ddfan_server* S, long long* completed, long long* failed, int* running,
                       long long* waiting, size_t* used_bytes, double* voxel_views) {
    return ddf::guarded([&]() -> int {
        CHECK(S, DDFAN_EINVAL, "server is NULL");
        std::lock_guard<std::mutex> lk(S->mu);
        if (completed) *completed = S->completed;
        if (failed) *failed = S->failed;
        if (running) *running = S->adm.active();
        if (waiting) *waiting = (long long)S->adm.waiting();
        if (used_bytes) *used_bytes = S->adm.used();
        if (voxel_views) *voxel_views = S->busyVoxelViews;
        return DDFAN_OK;
    });
}

int ddfan_server_destroy(ddfan_server* S) {
    return ddf::guarded([&]() -> int {
        if (!S) return DDFAN_OK;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->stop = true;
            S->cvWork.notify_all();
        }
        if (S->worker.joinable()) S->worker.join();
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->reaperStop = true;
            S->cvDead.notify_all();
        }
        if (S->reaper.joinable()) S->reaper.join();  // frees everything finished
        (void)hipSetDevice(S->device);
        S->reqs.clear();
        (void)hipStreamDestroy(S->stream);
        (void)hipStreamDestroy(S->copyStream);
        (void)hipStreamDestroy(S->upStream);
        delete S;
        return DDFAN_OK;
    });
}

}  // extern "C"
