// A device error reported to the request loop while requests are in flight:
// every outstanding request must complete with DDFAN_EDEVICE (no waiter left
// blocked, destroy returns), and the loop must serve the next request
// normally.  The error status comes from replaced completion checks
// (ddf::set_server_hooks); the GPU itself runs normally throughout.
#include <signal.h>
#include <stdio.h>
#include <unistd.h>

#include <atomic>
#include <vector>

#include "ddfan.h"
#include "ddfan_serve_hooks.h"

namespace {

std::atomic<int> g_fail{0};
std::atomic<long> g_checks{0};

hipError_t query(hipEvent_t e) {
    ++g_checks;
    return g_fail ? hipErrorLaunchFailure : hipEventQuery(e);
}
hipError_t event_sync(hipEvent_t e) {
    ++g_checks;
    return g_fail ? hipErrorLaunchFailure : hipEventSynchronize(e);
}
hipError_t stream_sync(hipStream_t s) {
    ++g_checks;
    return g_fail ? hipErrorLaunchFailure : hipStreamSynchronize(s);
}

#define EXPECT(c)                                                        \
    do {                                                                 \
        if (!(c)) {                                                      \
            fprintf(stderr, "FAILED %s:%d: %s (%s)\n", __FILE__, __LINE__, #c, \
                    ddfan_last_error());                                 \
            return 1;                                                    \
        }                                                                \
    } while (0)

struct Job {
    int XN = 48, YN = 48, SLN = 2, DNU = 96, PN = 48;
    std::vector<float> xds, yds, hangs, meas, x;
    ddfan_geom g{};
    Job() {
        xds.resize(DNU);
        yds.resize(DNU);
        ddfan_detector_arc(DNU, 48.5f, 4.8f, 1085.6f, 595.0f, xds.data(), yds.data());
        for (int v = 0; v < PN; ++v)
            for (int z = 0; z < SLN; ++z) hangs.push_back(6.2831853f * v / PN);
        meas.assign((size_t)PN * DNU * SLN, 1.0f);
        x.assign((size_t)XN * YN * SLN, -1.0f);
        g = ddfan_geom{0.0f, 595.0f, DNU, xds.data(), yds.data(), 48.5f, 0.0f, 0.0f, XN, YN, SLN, 8.0f};
    }
};

}  // namespace

int main() {
    alarm(120);  // a waiter left blocked ends the test as a failure, not a hang
    int n = 0;
    if (ddfan_device_count(&n) != DDFAN_OK || n < 1) {
        printf("serve-fault skipped: no GPU\n");
        return 77;
    }
    ddfan_server* S = nullptr;
    EXPECT(ddfan_server_create(0, 2, DDFAN_SAT_F32, &S) == DDFAN_OK);
    ddf::ServerHooks h;
    h.eventQuery = query;
    h.eventSync = event_sync;
    h.streamSync = stream_sync;
    ddf::set_server_hooks(S, h);

    // two requests in flight when the error arrives
    Job a, b;
    long long ia = 0, ib = 0;
    g_fail = 1;
    EXPECT(ddfan_server_submit(S, &a.g, a.hangs.data(), a.PN, nullptr, a.meas.data(), nullptr,
                               4, 4, 1, &ia) == DDFAN_OK);
    EXPECT(ddfan_server_submit(S, &b.g, b.hangs.data(), b.PN, nullptr, b.meas.data(), nullptr,
                               1, 2, 0, &ib) == DDFAN_OK);
    double q = 0, lat = 0;
    EXPECT(ddfan_server_wait(S, ia, a.x.data(), &q, &lat) == DDFAN_EDEVICE);
    EXPECT(ddfan_server_wait(S, ib, b.x.data(), &q, &lat) == DDFAN_EDEVICE);
    EXPECT(g_checks > 0);
    long long done = 0, failed = 0;
    EXPECT(ddfan_server_stats(S, &done, &failed, nullptr, nullptr, nullptr, nullptr) == DDFAN_OK);
    EXPECT(failed == 2 && done == 0);

    // the device recovers: the next request runs normally
    g_fail = 0;
    Job c;
    long long ic = 0;
    EXPECT(ddfan_server_submit(S, &c.g, c.hangs.data(), c.PN, nullptr, c.meas.data(), nullptr, 2,
                               4, 1, &ic) == DDFAN_OK);
    EXPECT(ddfan_server_wait(S, ic, c.x.data(), &q, &lat) == DDFAN_OK);
    bool changed = false;
    for (float v : c.x) changed = changed || v != -1.0f;
    EXPECT(changed);

    // an error with a request outstanding at destroy: destroy still returns
    Job d;
    long long id = 0;
    g_fail = 1;
    EXPECT(ddfan_server_submit(S, &d.g, d.hangs.data(), d.PN, nullptr, d.meas.data(), nullptr, 3,
                               4, 1, &id) == DDFAN_OK);
    EXPECT(ddfan_server_destroy(S) == DDFAN_OK);
    printf("serve-fault ok\n");
    return 0;
}
