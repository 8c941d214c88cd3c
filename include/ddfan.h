/*
 * ddfan.h -- C ABI of the MI355X distance-driven fan-beam reconstruction path.
 *
 * Drop-in replacements for the reference's native entry points
 * (Angus1996/AAPM_LowDoseCTImageRecon_competition, AAPM_script/):
 *
 *   ddfan_proj            <- DD2ProjMultiGPUs_CPP17  MultiSlicesProjGPU.cu:486-494
 *                            (MEX argument list MultiSlicesProjGPU.cu:550-575)
 *   ddfan_back            <- DD2BackMultiGPUs         MultiSlicesBackGPU.cu:349-359
 *                            (MEX argument list MultiSlicesBackGPU.cu:409-436)
 *   ddfan_helical_to_fan  <- HelicalToFan             HelicalToFanFunc_mex.cpp:142-158
 *                            (MEX argument list HelicalToFanFunc_mex.cpp:190-232)
 *
 * plus the device-resident plan API that the OS-SART driver
 * (OSSART_AAPM.m:49-117), the batched launch and the request loop use.
 *
 * Conventions (same buffers and layouts as the reference):
 *   volume   vol[(y*XN + x)*SLN + z]          (MATLAB single(SLN, XN, YN))
 *   sinogram prj[(v*DNU + u)*SLN + z]         (MATLAB single(SLN, DNU, PN))
 *   angles   hangs[v*SLN + z]                 (MATLAB single(SLN, PN)), radians
 *   mask     mask[y*XN + x], uint8; only XN*YN bytes are read (the reference's
 *            [mask, mask'] second half is never read).  NULL means all ones.
 *   raw helical proj[(view*DetWidth + ch)*DetHeight + row]
 *   rebinned Proj[(s*DefTimes + v)*DetWidth + ch], Views[s*DefTimes + v]
 *
 * Every entry point returns an int status (DDFAN_OK = 0, negative on error)
 * instead of calling exit(); ddfan_last_error() gives a message for the
 * calling thread.  Host-buffer entry points are synchronous, re-entrant and
 * thread-safe; they never reset the device.  Output buffers are caller-owned.
 */
#ifndef DDFAN_H
#define DDFAN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DDFAN_OK 0
#define DDFAN_EINVAL (-1)  /* invalid argument or shape */
#define DDFAN_ENOMEM (-2)  /* device or host allocation failed */
#define DDFAN_EBOUNDS (-3) /* input would be read out of bounds */
#define DDFAN_EDEVICE (-4) /* HIP runtime error */
#define DDFAN_ENODEV (-5)  /* no usable GPU */
#define DDFAN_EBUSY (-6)   /* request rejected by the HBM budget */
#define DDFAN_EIO (-7)     /* file missing, unreadable or not in the expected format */

/* Summed-area-table precision. */
#define DDFAN_SAT_F32 0      /* FP32 sequential prefix: the reference's arithmetic */
#define DDFAN_SAT_TWOFLOAT 1 /* compensated FP32 (TwoSum) prefix, hi/lo storage */
#define DDFAN_SAT_F64 2      /* FP64 prefix */

const char* ddfan_status_string(int status);
const char* ddfan_last_error(void);
int ddfan_version(void);
/* Number of visible GPUs (the reference's gpuDeviceCount_mex.cu:7-9). */
int ddfan_device_count(int* count);

/* Precision used by the drop-in entry points below (default DDFAN_SAT_F32). */
int ddfan_set_default_sat_mode(int sat_mode);

/* ---------------------------------------------------------------- drop-ins */

/* DD projector.  Same argument order and meaning as DD2ProjMultiGPUs_CPP17;
 * nStart is the length of startIdx (the reference reads one entry per GPU).
 * The slice split is validated (startIdx[0] == 0, non-decreasing, < SLN) and
 * all shards run on the current device.  hvol is not modified; the mask
 * multiplies a private copy (MultiSlicesProjGPU.cu:202-209,215).  startIdx may
 * be NULL (nStart 0). */
int ddfan_proj(float x0, float y0, int DNU, const float* xds, const float* yds,
               float imgXCenter, float imgYCenter, const float* hangs, int PN, int XN, int YN,
               int SLN, const float* hvol, float* hprj, float dx, const unsigned char* mask,
               const int* startIdx, int nStart);

/* DD backprojector.  Same argument order as DD2BackMultiGPUs.  Masked voxels
 * are written as 0 (the reference leaves them uninitialised). */
int ddfan_back(float x0, float y0, int DNU, const float* xds, const float* yds, float detCntIdx,
               float imgXCenter, float imgYCenter, const float* hangs, int PN, int XN, int YN,
               int SLN, float* hvol, const float* hprj, float dx, const unsigned char* mask,
               const int* startIdx, int nStart);

/* Helical-to-fan single-slice rebinning.  Same arguments as HelicalToFan plus
 * TotalViews, the number of raw views in proj: a slice whose rotation window
 * ends past TotalViews returns DDFAN_EBOUNDS (the reference reads out of
 * bounds there). */
int ddfan_helical_to_fan(float* Proj, float* Views, const float* proj, long long TotalViews,
                         const float* zPos, int SLN, float SD, float SO, float BVAngle,
                         int DetWidth, int DetHeight, float PerDetW, float PerDetH, int DefTimes,
                         float DetCenterW, float SpiralPitchFactor);

/* DD projector without device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under stream capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Arguments as
 * ddfan_proj, except that d_vol ([YN][XN][SLN]) and d_prj ([PN][DNU][SLN]) are
 * device buffers on the current device; the small host arrays (xds, yds,
 * hangs, mask) are copied in.  The working tables live in a fixed device array
 * of 512 MiB per GPU (held until the process ends): a projection whose tables
 * need more (e.g. beyond 512x512x64 slices with the default f32 tables)
 * returns DDFAN_EINVAL.  Calls on one GPU share that array and are serialised;
 * each returns after its projection has completed.  Results are bit-identical
 * to ddfan_proj. */
int ddfan_proj_noalloc(float x0, float y0, int DNU, const float* xds, const float* yds,
                       float imgXCenter, float imgYCenter, const float* hangs, int PN, int XN,
                       int YN, int SLN, const float* d_vol, float* d_prj, float dx,
                       const unsigned char* mask, const int* startIdx, int nStart);

/* ------------------------------------------------------------ geometry tables */

/* Equiangular arc detector centres, DD2MutiSlices.m:61-76 (MATLAB single
 * arithmetic, platform libm). */
int ddfan_detector_arc(int DNU, float detCntIdx, float detCellWidth, float sdd, float sid,
                       float* xds, float* yds);
/* Circular reconstruction mask, reconscript.m:28-38 (mask[y*XN + x]). */
int ddfan_recon_mask(int XN, int YN, unsigned char* mask);
/* OS-SART subset k (0-based) of OSSART_AAPM.m:60: views k, k+numOS, ...
 * Writes *count entries to out (out may be NULL to query the count). */
int ddfan_subset_views(int numViews, int numOS, int k, int* out, int* count);
/* Rebin per-slice first raw view, HelicalToFanFunc_mex.cpp:71 (1-based view). */
int ddfan_rebin_starts(const float* zPos, int SLN, float SD, float SO, int DetHeight,
                       float PerDetW, float PerDetH, int DefTimes, float SpiralPitchFactor,
                       int* start);

/* --------------------------------------------------------- device memory */
int ddfan_dev_alloc(size_t bytes, void** ptr);
int ddfan_dev_free(void* ptr);
int ddfan_h2d(void* dst, const void* src, size_t bytes);
int ddfan_d2h(void* dst, const void* src, size_t bytes);
int ddfan_d2d(void* dst, const void* src, size_t bytes);
int ddfan_dev_mem_info(size_t* free_bytes, size_t* total_bytes);

/* ------------------------------------------------------------ plan API */

/* Geometry of one reconstruction, the fields DD2MutiSlices.m:49-76 derives
 * from conf.  The plan copies xds/yds. */
typedef struct ddfan_geom {
    float x0, y0;  /* source position at angle 0, (0, sid) in the reference */
    int DNU;
    const float* xds;
    const float* yds;
    float detCntIdx;
    float imgXCenter, imgYCenter;
    int XN, YN, SLN;
    float dx;
} ddfan_geom;

typedef struct ddfan_plan ddfan_plan;

/* Creates a plan on the current device for PN views with angles
 * hangs[v*SLN + z] (copied).  slice_chunk: 0 picks the widest chunk of
 * consecutive slices whose angles agree in every view (1, 2, 4, 8 or 16);
 * a positive value caps it. */
int ddfan_plan_create(const ddfan_geom* geom, const float* hangs, int PN,
                      const unsigned char* mask, int sat_mode, int slice_chunk,
                      ddfan_plan** out);
int ddfan_plan_destroy(ddfan_plan* plan);
/* slice_chunk: widest chunk of consecutive slices with identical angles. */
int ddfan_plan_query(const ddfan_plan* plan, int* slice_chunk, int* n_chunks, int* sat_mode);
/* Tunable launch parameters (the autotuner sets them; defaults are heuristic).
 * DDFAN_PARAM_SLAB_TILE: projector slabs per pass (0 = one pass over all
 * slabs); smaller passes keep the table band each pass reads inside L2. */
#define DDFAN_PARAM_SLAB_TILE 1
/* DDFAN_PARAM_VIEWS_PER_WAVE: consecutive views one projector wave handles
 * (1, 2 or 4); they share table rows through L1.  With DDFAN_PARAM_PROJ_STAGE
 * 2 or 3 it counts the whole-row projector's tasks (view x 63 detectors) per
 * wave instead: 9 or 12 in blocks of 4 waves, 5 or 6 in blocks of 8, 3 or 4 in
 * blocks of 12. */
#define DDFAN_PARAM_VIEWS_PER_WAVE 2
/* DDFAN_PARAM_PROJ_CHUNK / DDFAN_PARAM_BACK_CHUNK: slices per chunk (a power
 * of two up to the widest chunk whose angles agree) of the projector and of
 * the backprojector. */
#define DDFAN_PARAM_PROJ_CHUNK 3
#define DDFAN_PARAM_BACK_CHUNK 4
/* DDFAN_PARAM_BACK_STAGE: views per group whose detector windows the
 * backprojector copies into LDS before summing (0 = every lane reads the
 * table in global memory; 1, 2 or 4, with value x back chunk <= 32).  Results
 * are bit-identical for every value.  Narrowing the back chunk halves it as
 * needed. */
#define DDFAN_PARAM_BACK_STAGE 5
/* DDFAN_PARAM_PROJ_STAGE: 1 = the projector copies each slab's table run
 * into LDS for four view groups at a time; 0 = every lane reads the table in
 * global memory; 2 / 3 = whole table rows in LDS for the views of a block
 * (4 / 2 slices per block; for spaced views such as OS-SART subsets; needs a
 * projector chunk that is a multiple of 4 / 2 and images up to 775 voxels
 * wide, else the projector call fails with DDFAN_EINVAL).  Results are
 * bit-identical for every value. */
#define DDFAN_PARAM_PROJ_STAGE 6
/* DDFAN_PARAM_BACK_PASS: a plain backprojection of more views than this runs
 * as launches of at most this many views, each continuing the sums the
 * previous one stored (0 = one launch).  Results are bit-identical. */
#define DDFAN_PARAM_BACK_PASS 7
/* DDFAN_PARAM_BACK_MID_RANGE (read-only): 1 when the geometry keeps every
 * voxel-source distance and view-axis offset of the backprojector's footprint
 * inside the range where its correctly rounded square root and quotients run
 * as their short sequences with no per-lane check (any source farther from the
 * grid centre than sqrt(2) times its largest half-extent); 0: each wave checks
 * its lanes.  Results are the same either way. */
#define DDFAN_PARAM_BACK_MID_RANGE 8
int ddfan_plan_set_param(ddfan_plan* plan, int param, int value);
int ddfan_plan_get_param(const ddfan_plan* plan, int param, int* value);
/* Autotuner: times the projector over slice chunks x views per wave x LDS
 * staging, then slab tilings, and the backprojector over slice chunks x
 * staged view groups, for launches of `views`
 * views (0: the plan's PN) spaced PN / views apart, as in an OS-SART subset,
 * with this plan's shape, sets the fastest on the
 * plan and records it in a process-wide table (key: SAT mode, XN, YN, SLN,
 * DNU, widest uniform chunk, views).  flags DDFAN_TUNE_USE_CACHE takes a
 * recorded entry without timing.  out_params (may be NULL, else 6 ints)
 * receives {proj chunk, views per wave, slab tile, back chunk, back stage,
 * proj stage};
 * out_ms (may be NULL)
 * the projector and backprojector milliseconds per call.  No parameter
 * changes results beyond the projector's slab-pass summation grouping. */
#define DDFAN_TUNE_USE_CACHE 1
int ddfan_plan_autotune(ddfan_plan* plan, int views, int reps, int flags, int* out_params,
                        double* out_ms);
/* Applies a recorded entry for launches of `views` views; *found = 0 if none. */
int ddfan_plan_apply_tuned(ddfan_plan* plan, int views, int* found);
/* The table as text, one entry per line (load merges; malformed files are
 * rejected whole). */
int ddfan_tune_save(const char* path);
int ddfan_tune_load(const char* path, int* entries);
int ddfan_tune_clear(void);

/* Device bytes a plan of this shape holds for full-view operators plus an
 * OS-SART run with numOS subsets (0: operators only). */
int ddfan_plan_footprint(const ddfan_geom* geom, int PN, int sat_mode, int slice_chunk, int numOS,
                         size_t* bytes);

/* Device-resident operators, enqueued on the plan's stream.  The view subset
 * is first, first+stride, ... (count views) of the plan's angle table;
 * (0, 1, PN) is the reference's full call.  d_prj holds the subset's views
 * only, in subset order. */
int ddfan_plan_proj(ddfan_plan* plan, const float* d_vol, float* d_prj, int first, int stride,
                    int count);
int ddfan_plan_back(ddfan_plan* plan, const float* d_prj, float* d_vol, int first, int stride,
                    int count);
int ddfan_plan_sync(ddfan_plan* plan);

/* The summed-area tables the operators read (for tests and diagnostics; the
 * reference builds them inside its calls, MultiSlicesProjGPU.cu:53-97 and
 * MultiSlicesBackGPU.cu:34-66).  Entry k of a row holds the prefix S[k] (sum of
 * samples 0..k-1) and the sample v[k] (v[N] = 0): float2 (S, v) for
 * DDFAN_SAT_F32, floats (S_hi, S_lo, v, 0) with S = S_hi + S_lo for
 * DDFAN_SAT_TWOFLOAT, double2 (S, v) for DDFAN_SAT_F64.
 * ddfan_plan_vol_tables builds the projector's two volume tables from d_vol
 * (masked by the plan's mask) and copies them to d_satX / d_satY:
 * E[chunk][row][zt][k], slice z = chunk*ZT + zt (ZT = the projector's slice
 * chunk, DDFAN_PARAM_PROJ_CHUNK; slices past SLN are zero), rows y with k
 * along x (k = 0..XN) in d_satX, rows x with k along y (k = 0..YN) in d_satY.
 * ddfan_plan_sino_table builds the backprojector's table of a sinogram of
 * `count` views d_prj[(v*DNU + u)*SLN + z] into d_out: E[chunk][v][zt][k],
 * k = 0..DNU, ZT = the backprojector's slice chunk (DDFAN_PARAM_BACK_CHUNK).
 * With NULL output buffers they only report the bytes. */
int ddfan_plan_vol_tables(ddfan_plan* plan, const float* d_vol, void* d_satX, void* d_satY,
                          size_t* bytes_x, size_t* bytes_y);
int ddfan_plan_sino_table(ddfan_plan* plan, const float* d_prj, int count, void* d_out,
                          size_t* bytes);

/* OS-SART(+FISTA) of OSSART_AAPM.m:49-117 on device-resident data:
 * d_meas[(v*DNU + u)*SLN + z] over the plan's PN views, d_x in/out.
 * flags: DDFAN_OSSART_GRAPH replays one captured iteration as a hipGraph. */
#define DDFAN_OSSART_GRAPH 1
int ddfan_ossart(ddfan_plan* plan, const float* d_meas, float* d_x, int numIter, int numOS,
                 int useFISTA, int flags);

/* ------------------------------------------------------------ batched OS-SART
 *
 * Many independent reconstructions of uneven size (volume, slices, views,
 * subsets, iterations) run together: every kernel launch of an iteration
 * carries up to 8 reconstructions, and each iteration of the batch is one
 * captured hipGraph.  The reference has no batched form; per reconstruction
 * the arithmetic is that of ddfan_ossart (bit-identical results).  Each plan
 * holds one reconstruction's state, so a plan is added once; all plans share
 * the batch's SAT mode and device; each keeps its own launch parameters. */
typedef struct ddfan_batch ddfan_batch;
int ddfan_batch_create(int sat_mode, ddfan_batch** out);
int ddfan_batch_add(ddfan_batch* batch, ddfan_plan* plan, const float* d_meas, float* d_x,
                    int numIter, int numOS, int useFISTA, int* index);
int ddfan_batch_size(const ddfan_batch* batch, int* n);
/* Launch grouping.  A reconstruction whose two volume tables take at least
 * DDFAN_BATCH_SOLO_BYTES bytes (default 64 MiB) runs its kernels back to back;
 * the others share launches per kernel kind, up to 8 per launch, while a
 * projector (backprojector) launch holds at most DDFAN_BATCH_PROJ_FILL 64-lane
 * waves (DDFAN_BATCH_BACK_FILL 256-thread blocks; default: no limit).  Grouping
 * never changes results. */
#define DDFAN_BATCH_PROJ_FILL 1
#define DDFAN_BATCH_BACK_FILL 2
#define DDFAN_BATCH_SOLO_BYTES 3
/* DDFAN_BATCH_STREAMS (0..8): the reconstructions are split into that many
 * groups of about equal work, each group running its own batch iterations on
 * its own stream, so the kernels of different groups share the GPU.  0
 * (default): 2 groups. */
#define DDFAN_BATCH_STREAMS 4
int ddfan_batch_set_param(ddfan_batch* batch, int param, long long value);
/* Throughput launch shapes for the reconstructions that share launches (volume
 * tables below DDFAN_BATCH_SOLO_BYTES).  A plan's own autotuning times it
 * alone, where a small reconstruction leaves most of the GPU idle; here one
 * projector shape (slice chunk, views or tasks per wave, staging) and one
 * backprojector shape (slice chunk, staged views) for all of them (a plan
 * whose slices allow only narrower chunks takes its widest) are timed on
 * their shared launches (median of reps, on two subsets of each) and the
 * fastest is applied to their plans when it beats the shapes in place.
 * Iterates are only read, slab tiles are kept, results do not change.
 * out_ms (4 doubles, may be NULL): projector and backprojector launch time per
 * subset round before, then after. */
int ddfan_batch_tune(ddfan_batch* batch, int reps, double* out_ms);
/* Runs every reconstruction to completion (flags: DDFAN_OSSART_GRAPH). */
int ddfan_batch_run(ddfan_batch* batch, int flags);
int ddfan_batch_destroy(ddfan_batch* batch);

/* ------------------------------------------------------------ request loop
 *
 * A worker thread serves reconstruction requests on the current device.
 * Requests are admitted (in arrival order by default, see DDFAN_SERVER_POLICY)
 * while one of max_slots is free and their device footprint fits in what is
 * left of hbm_budget bytes (0: 80% of
 * the free device memory); a request larger than the whole budget is
 * rejected at submission (DDFAN_EBUSY).  Admitted requests keep their
 * sinogram and iterate in device memory and run together as a batch
 * (ddfan_batch_run's iteration), joining and leaving at iteration boundaries
 * while others run; in each step of the batch the largest running request
 * runs one iteration and a smaller one up to 8, about as many as fit in the
 * same work, so small requests do not wait for a large one's iterations.
 * Every request's result is that of running it alone.  The reference has no
 * serving layer. */
typedef struct ddfan_server ddfan_server;
int ddfan_server_create(size_t hbm_budget, int max_slots, int sat_mode, ddfan_server** out);
/* Copies the request's host inputs (meas [PN][DNU][SLN], x0 [YN][XN][SLN] or
 * NULL for zeros, mask or NULL) and queues it; *id names it.  With no request
 * waiting, a free slot and room in the budget it is admitted at once and the
 * inputs go straight to the device; else they are copied to host memory.
 * Either way the caller's buffers are not read after the call returns. */
int ddfan_server_submit(ddfan_server* server, const ddfan_geom* geom, const float* hangs, int PN,
                        const unsigned char* mask, const float* meas, const float* x0,
                        int numIter, int numOS, int useFISTA, long long* id);
int ddfan_server_poll(ddfan_server* server, long long id, int* done);
/* Server parameters, set before the first submission.  DDFAN_SERVER_GRAPH:
 * 1 captures each batch iteration of a request set as a hipGraph and replays
 * it while the set is unchanged (every request then runs one iteration per
 * step); 0 (default) launches every step's kernels directly from the worker,
 * with the smaller requests running several iterations per step.  The worker
 * keeps at most two steps in flight either way.  Results are the same bit
 * for bit. */
#define DDFAN_SERVER_GRAPH 1
/* DDFAN_SERVER_POLICY: the order in which queued requests are admitted.
 * DDFAN_POLICY_FIFO (default): arrival order, a request that does not fit is
 * not overtaken.  DDFAN_POLICY_SHORTEST: the queued request with the least
 * work (voxels x views x iterations) that fits goes first, unless a request
 * has waited DDFAN_SERVER_MAX_WAIT_MS (default 10000) or longer: then the
 * longest-waiting of those goes first and is not overtaken.  A submission that
 * finds nobody queued is admitted at once under both.  Results do not depend
 * on the policy. */
#define DDFAN_SERVER_POLICY 2
#define DDFAN_SERVER_MAX_WAIT_MS 3
#define DDFAN_POLICY_FIFO 0
#define DDFAN_POLICY_SHORTEST 1
int ddfan_server_set_param(ddfan_server* server, int param, long long value);
/* Blocks until the request is done, copies its reconstruction to x (may be
 * NULL) and returns the milliseconds from submission to admission and to
 * completion.  Called before the request completes, the loop copies the result
 * from the device straight into x.  A request is collected once: a second wait
 * on it, concurrent or later, returns DDFAN_EINVAL; a failed request returns
 * its status. */
int ddfan_server_wait(ddfan_server* server, long long id, float* x, double* queue_ms,
                      double* latency_ms);
/* Counts so far; voxel_views sums 2*XN*YN*SLN*PN over every request iteration run.
 * used_bytes counts admitted requests until their device memory has been freed,
 * which happens on a thread of the server shortly after each completes. */
int ddfan_server_stats(ddfan_server* server, long long* completed, long long* failed,
                       int* running, long long* waiting, size_t* used_bytes,
                       double* voxel_views);
/* Finishes every queued request, then stops the worker and frees the server.
 * Submit, poll, wait and stats may be called from any number of threads
 * concurrently; destroy must be the last call and must not overlap any other
 * call on the same server. */
int ddfan_server_destroy(ddfan_server* server);

/* The reference's plain form on host buffers: every subset calls the
 * drop-in projector and backprojector (allocate, copy in, compute, copy out,
 * free) and the ratio/normalise/update run on the host (OSSART_AAPM.m:49-117
 * with the calls of DD2MutiSlices.m:106-115). */
int ddfan_ossart_host_plain(float x0, float y0, int DNU, const float* xds, const float* yds,
                            float detCntIdx, float imgXCenter, float imgYCenter,
                            const float* hangs, int numViews, int XN, int YN, int SLN,
                            const float* meas, float* x, float dx, const unsigned char* mask,
                            int numIter, int numOS, int useFISTA, int sat_mode);

/* Kernel timing on the plan's stream: when enabled every kernel launch is
 * bracketed by HIP events; ddfan_plan_timing_read synchronises and returns
 * the total milliseconds and launch count of one kernel kind since the last
 * reset.  kind: 0 volume SAT, 1 projector, 2 sinogram SAT, 3 backprojector. */
int ddfan_plan_timing(ddfan_plan* plan, int enable);
int ddfan_plan_timing_read(ddfan_plan* plan, int kind, double* total_ms, int* launches);

/* Integer index traces computed by the same device functions as the
 * operators (for bit-exactness checks).  Projector: d_trace int32
 * [count][DNU][SLN][max(XN,YN)][2] texel indices (L, R) per slab;
 * backprojector: [YN][XN][SLN][count][2] detector texel indices. */
int ddfan_plan_trace_proj(ddfan_plan* plan, int first, int stride, int count, int32_t* d_trace);
int ddfan_plan_trace_back(ddfan_plan* plan, int first, int stride, int count, int32_t* d_trace);

/* Device-resident rebin.  raw_kind 0: d_raw float line integrals; 1: d_raw
 * uint16 counts, converted to log(count_scale / max(count, 1)) on the fly;
 * 2: d_raw uint16 line integrals times count_scale (1: used unscaled, as
 * readProj.m:47 does).  raw_order 0: raw[(view*DetWidth+ch)*DetHeight+row]
 * (HelicalToFan's input); 1: the projection frames in file order, as the
 * DICOM-CT-PD files store them (DetWidth rows of DetHeight samples each,
 * raw[(f*DetWidth+ch)*DetHeight+row]), which readProj.m:43-53 reads into the
 * same per-frame order and flips to the reverse view order
 * (view = TotalViews-1-f) -- the gather reads them in place.
 * layout 0: reference layout (Proj[(s*DefTimes+v)*W+ch], Views[s*DefTimes+v]);
 * layout 1: sinogram layout (prj[(v*W+ch)*SLN+s], hangs[v*SLN+s]).
 * d_cut (optional) [SLN][DefTimes] receives 1 for rows zeroed as cut views;
 * d_idv (optional, layout-independent) int32 [SLN][DefTimes][W] source rows. */
int ddfan_helical_to_fan_dev(float* d_Proj, float* d_Views, const void* d_raw, int raw_kind,
                             int raw_order, float count_scale, int layout, long long TotalViews,
                             const float* zPos, int SLN, float SD, float SO, float BVAngle,
                             int DetWidth, int DetHeight, float PerDetW, float PerDetH,
                             int DefTimes, float DetCenterW, float SpiralPitchFactor,
                             unsigned char* d_cut, int32_t* d_idv);

/* ---- DICOM-CT-PD projection files and CT image headers (what readProj.m,
 * CollectReconCfg.m and CollectImageCfg.m read through MATLAB's dicominfo /
 * dicomread and the cov_* byte converters).  Little-endian files, explicit or
 * implicit VR; private elements are matched by tag number and decoded from
 * their bytes as the reference's converters do.  Host-only (no GPU). */

/* Projection header, CollectReconCfg.m:5-92.  The fields marked (opt) are
 * NaN (or empty) when the file lacks them; the others are required. */
typedef struct ddfan_ctpd_header {
    double HUCalibrationFactor;              /* (0018,0061) decimal string (opt) */
    double DataCollectionDiameter;           /* (0018,0090) decimal string (opt) */
    double SpiralPitchFactor;                /* (0018,9311) float64 */
    int NumberofDetectorRows;                /* (7029,1010) uint16 */
    int NumberofDetectorColumns;             /* (7029,1011) uint16 */
    float DetectorElementTransverseSpacing;  /* (7029,1002) float32, mm */
    float DetectorElementAxialSpacing;       /* (7029,1006) float32, mm */
    char DetectorShape[17];                  /* (7029,100B) text (opt) */
    float DetectorFocalCenterAngularPosition; /* (7031,1001) float32 */
    float DetectorFocalCenterAxialPosition;   /* (7031,1002) float32 (opt) */
    float DetectorFocalCenterRadialDistance;  /* (7031,1003) float32, mm */
    float DetectorCentralElementX;            /* (7031,1033) first float32, 1-based */
    float DetectorCentralElementY;            /* (7031,1033) second float32, 1-based */
    float ConstantRadialDistance;             /* (7031,1031) float32, mm */
    float SourceAngularPositionShift;         /* (7033,100B) float32 (opt) */
    float SourceAxialPositionShift;           /* (7033,100C) float32 (opt) */
    float SourceRadialDistanceShift;          /* (7033,100D) float32 (opt) */
    int NumberofSourceAngularSteps;           /* (7033,1013) uint16 */
    int Rows, Columns;                        /* (0028,0010/0011): the stored frame (0 if absent) */
    int PhotonStatisticsCount;                /* float32 values in (7033,1065) */
} ddfan_ctpd_header;
/* Reads one file's header; photon_stats (capacity floats, may be NULL with
 * capacity 0) receives the first values of (7033,1065). */
int ddfan_ctpd_read_header(const char* path, ddfan_ctpd_header* out, float* photon_stats,
                           int capacity);

/* CT image header, CollectImageCfg.m:3-31 (MATLAB's dicominfo names).  Width,
 * Height and ReconstructionDiameter are required; absent numbers are NaN
 * (BitDepth and the pixel value bounds 0). */
typedef struct ddfan_ct_image_header {
    int Width, Height;                       /* (0028,0011) Columns, (0028,0010) Rows */
    int BitDepth;                            /* (0028,0101) */
    double SliceThickness;                   /* (0018,0050) */
    double KVP;                              /* (0018,0060) */
    double DataCollectionDiameter;           /* (0018,0090) */
    double ReconstructionDiameter;           /* (0018,1100) */
    double SingleCollimationWidth;           /* (0018,9306) */
    double TotalCollimationWidth;            /* (0018,9307) */
    double TableSpeed;                       /* (0018,9309) */
    double TableFeedPerRotation;             /* (0018,9310) */
    double SpiralPitchFactor;                /* (0018,9311) */
    double DataCollectionCenterPatient[3];   /* (0018,9313) */
    double ReconstructionTargetCenterPatient[3]; /* (0018,9318) */
    double PixelSpacing[2];                  /* (0028,0030) */
    int SmallestImagePixelValue;             /* (0028,0106) */
    int LargestImagePixelValue;              /* (0028,0107) */
    double WindowCenter[2], WindowWidth[2];  /* (0028,1050/1051), first two values */
    double RescaleIntercept, RescaleSlope;   /* (0028,1052/1053) */
    double PixelSize;                        /* ReconstructionDiameter / Width */
} ddfan_ct_image_header;
int ddfan_ct_image_read_header(const char* path, ddfan_ct_image_header* out);

/* The 16-bit pixel data of one file (rows x cols samples, row-major, one
 * frame, uncompressed) into out.  A file of another size is DDFAN_EIO. */
int ddfan_dicom_read_u16(const char* path, int rows, int cols, uint16_t* out);
/* The frames of nfiles files, in the given order, into out[f*rows*cols ...]
 * with `threads` host threads (0: one per hardware thread, at most 64).  For
 * DICOM-CT-PD projections rows = detector columns and cols = detector rows,
 * so out is the rebin's raw input with the views reversed (raw_order 1 of
 * ddfan_helical_to_fan_dev; readProj.m:41-53). */
int ddfan_ctpd_read_frames(const char* const* paths, int nfiles, int rows, int cols,
                           uint16_t* out, int threads);

#ifdef __cplusplus
}
#endif

#endif /* DDFAN_H */
