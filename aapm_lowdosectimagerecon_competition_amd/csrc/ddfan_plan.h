// ddfan_plan.h -- library-private plan state and helpers shared by the C ABI
// (ddfan_api.hip), the OS-SART engine and batch (ddfan_ossart.hip) and the
// request loop (ddfan_serve.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "ddfan.h"
#include "ddfan_host.h"
#include "ddfan_internal.h"

namespace ddf {

// Records a message for ddfan_last_error() of the calling thread; returns code.
int fail(int code, const char* fmt, ...);

#define HIPCALL(x)                                                                           \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess)                                                                \
            return ::ddf::fail(e_ == hipErrorOutOfMemory ? DDFAN_ENOMEM : DDFAN_EDEVICE,     \
                               "%s: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__,        \
                               __LINE__);                                                    \
    } while (0)

#define CHECK(cond, code, ...)                              \
    do {                                                    \
        if (!(cond)) return ::ddf::fail(code, __VA_ARGS__); \
    } while (0)

#define TRY(x)                           \
    do {                                 \
        int rc_ = (x);                   \
        if (rc_ != DDFAN_OK) return rc_; \
    } while (0)

// Owning device allocation.
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    // Gives up ownership: the caller frees the returned pointer.
    void* detach() {
        void* q = p;
        p = nullptr;
        bytes = 0;
        return q;
    }
    hipError_t reserve(size_t b) {
        if (p && b <= bytes) return hipSuccess;
        release();
        hipError_t e = hipMalloc(&p, b ? b : 16);
        if (e == hipSuccess)
            bytes = b;
        else
            p = nullptr;
        return e;
    }
    template <class T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

int check_device(int* dev);
int upload(DBuf& b, const void* src, size_t bytes, hipStream_t st);

// Argument checks and kernel scalars shared by the entry points.
int check_subset(int PN, int first, int stride, int count);
int check_start_idx(const int* startIdx, int nStart, int SLN);
// The plan's projector launch shape fits its geometry (DDFAN_EINVAL if not).
int check_proj_shape(const ddfan_plan* P);
int validate_geom(const ddfan_geom* G);
void fill_geom(GeomDev& g, const ddfan_geom* G, int SLN);
// SAT precision of the drop-in entry points (ddfan_set_default_sat_mode).
int default_sat_mode();

// Tables carry 256 bytes of slack: the staged projector and backprojector copy
// rows in 16-byte pieces, which may end one 8-byte f32 entry past a row.
inline size_t sat_vol_bytes(int mode, int nzc, int ZT, int XN, int YN) {
    return sat_elem_bytes(mode) * (size_t)nzc * ZT * (size_t)(XN + 1) * YN + 256;
}
inline size_t sat_sino_bytes(int mode, int nzc, int ZT, int DNU, int views) {
    return sat_elem_bytes(mode) * (size_t)nzc * ZT * (size_t)(DNU + 1) * views + 256;
}

}  // namespace ddf

struct ddfan_plan {
    int device = 0;
    hipStream_t stream = nullptr;
    ddf::GeomDev g{};
    int PN = 0, mode = 0;
    int ZTmax = 1;         // widest chunk of consecutive slices with identical angles
    int ZT = 1, nzc = 1;   // projector slice chunk and chunk count
    int ZTb = 1, nzcb = 1; // backprojector slice chunk and chunk count
    int slabTile = 0;      // projector slabs per pass (0: all slabs in one pass)
    int viewsPerWave = 4;  // projector views per wave (1, 2 or 4); tasks per wave (3-6, 9, 12) at stage >= 2
    int backStage = 0;     // backprojector views per LDS-staged group (0: per-lane gathers)
    int projStage = 0;     // 0: per-lane gathers, 1: staged runs, 2/3: whole rows (4/2 slices)
    int backPass = 0;      // backprojector views per pass of a plain call (0: one pass)
    std::vector<float> h_hangs, h_cs;
    ddf::DBuf cs, src, bx, by, xds, yds, mask;  // src: source positions (source_table)
    ddf::DBuf satX, satY, satP;
    // OS-SART state
    // cs0: slice 0's (cos, sin) pairs of every view, then its source positions
    ddf::DBuf rbuf, rowSum, colSum, xprev, momTab, counter, cs0, onesV, onesP, satX0, satY0, satP0;
    bool timing = false;
    struct Ev {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Ev> events;

    ~ddfan_plan() {
        if (stream) (void)hipStreamSynchronize(stream);
        clear_events();
        if (stream) (void)hipStreamDestroy(stream);
    }
    void clear_events() {
        for (auto& e : events) {
            (void)hipEventDestroy(e.a);
            (void)hipEventDestroy(e.b);
        }
        events.clear();
    }
    // Brackets the launches f() enqueues on `st` with events when timing is on.
    template <class F>
    hipError_t timed(int kind, F&& f, hipStream_t st = nullptr) {
        if (!timing) return f();
        if (!st) st = stream;
        hipEvent_t a, b;
        hipError_t e = hipEventCreate(&a);
        if (e != hipSuccess) return e;
        e = hipEventCreate(&b);
        if (e != hipSuccess) {
            (void)hipEventDestroy(a);
            return e;
        }
        (void)hipEventRecord(a, st);
        e = f();
        (void)hipEventRecord(b, st);
        events.push_back({kind, a, b});
        return e;
    }
};

namespace ddf {

// Volume summed-area tables sized for the projector's current slice chunking.
int ensure_vol_sat(ddfan_plan* P);

// One OS-SART(+FISTA) reconstruction (OSSART_AAPM.m:49-117) on a plan: the
// plan's tables and OS-SART buffers, the measured sinogram and the iterate.
struct OsJob {
    ddfan_plan* P = nullptr;
    const float* meas = nullptr;  // [PN][DNU][SLN]
    float* x = nullptr;           // [YN][XN][SLN], updated in place
    int numIter = 0, numOS = 1, fista = 0;
    std::vector<int> cnt;         // views per subset
    int maxc = 0;
    std::vector<float> h_cs0, h_mom;  // host staging of uploads in flight
};

// Allocates the plan's OS-SART buffers and enqueues on `st` the subset
// weights (slice-0 geometry), the FISTA momentum table, the iteration counter
// reset and the FISTA copy of the initial iterate.
int ossart_prepare(OsJob& J, hipStream_t st);

// Launch grouping of a batched iteration.  A reconstruction whose two volume
// tables take at least soloBytes runs its four kernels back to back; the others
// share launches per kernel kind (up to kBatch per launch) while a launch holds
// at most projFill projector waves / backFill backprojector blocks.  Measured
// on MI355X, 10 iterations of 48-view subsets: 16 jobs of 256^2 x 8-16 slices
// (34 MB of tables each) 1.33x faster than one after another when they share
// launches freely and 0.99x under a 2048-wave / 512-block limit; a mix of 16
// jobs of 256^2-768^2 x 8-123 slices (150-800 MB of tables) 0.70x when packed
// freely, 0.97x with its kernels run back to back.
constexpr long long kSoloTableBytes = 64LL << 20;
struct BatchShape {
    long long projFill = kNoFillLimit;
    long long backFill = kNoFillLimit;
    long long soloBytes = kSoloTableBytes;
};

// Enqueues one OS-SART iteration of every job on `st`.  The iteration is
// S = max(numOS) rounds; job j runs its subset k in round floor(k*S/numOS_j),
// so a job's subsets keep their order whatever else shares the launches.
// Each round is one batched launch per kernel kind; the iteration ends with
// every job's counter increment.  `timingPlan`, when not null, brackets the
// launches of each kind with its events.
hipError_t enqueue_batch_iteration(OsJob* const* jobs, int n, int mode, const BatchShape& shape,
                                   hipStream_t st, ddfan_plan* timingPlan);

// One batch iteration captured as a hipGraph (stream capture of
// enqueue_batch_iteration); replaying it runs one more iteration of every job,
// since the FISTA momentum is read through each job's device counter.
struct IterGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    IterGraph() = default;
    IterGraph(const IterGraph&) = delete;
    IterGraph& operator=(const IterGraph&) = delete;
    ~IterGraph() { reset(); }
    // Only after every launch of this graph has completed.
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    int capture(OsJob* const* jobs, int n, int mode, const BatchShape& shape, hipStream_t st);
    int launch(hipStream_t st);
};

}  // namespace ddf
