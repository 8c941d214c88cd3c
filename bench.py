#!/usr/bin/env python3
"""Benchmark of the DD fan-beam path on one MI355X.

Step = one full-view DD projector call plus one full-view DD backprojector
call on the 512x512x64-slice, 736-channel, 2304-view workload (0.98 mm
pixels, arc detector, sid 595 mm, sdd 1085.6 mm), with the volume and
sinogram resident in HBM.  value = voxel-views/s over the timed steps
(2 * 512*512*64*2304 voxel-views per step).

Also reported on the same line:
  roofline      the dominant kernel against the bound it runs at: VALU issue
                (wave-level VALU instructions per launch from the committed
                rocprofv3 counters x 64 lanes / its average launch duration,
                HIP events on the plan's stream, against 256 CUs x 4 SIMDs x
                16 lanes x 2.4 GHz), with its measured HBM traffic per launch
                (committed rocprofv3 FETCH_SIZE/WRITE_SIZE) over the same
                duration and the per-view-sweep byte model beside it
  kernels       average kernel durations
  cpu_baseline  the FP32 CPU restatement on the host cores: a stated sample of
                the workload's views (median of 3 after a warm-up) and the
                single-slice 256x256, 1152-view OS-SART workload end to end
  ossart        device-resident fused + hipGraph OS-SART against the
                reference's plain per-subset form on the same workload
  batch, serve  the uneven 16-job mix batched, and served as a Poisson stream
                at a stated fraction of the GPU's capacity
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DD proj+back voxel-views/s & %HBM peak vs plain; served recon p50/p99 latency"
HBM_PEAK_GBS = 8000.0
# FP32 VALU issue of MI355X: 256 CUs x 4 SIMDs x 16 lanes at 2.4 GHz (a 64-lane
# instruction occupies a SIMD for 4 cycles), in lane-instructions per second
VALU_PEAK_LANE_OPS = 256 * 4 * 16 * 2.4e9
COPY_BYTES = 256 << 20

# workload: 512x512x64 slices @ 0.98 mm, 736 channels x 2304 views
XN = YN = 512
SLN = 64
DX = 0.98
DNU = 736
PN = 2304
SID, SDD, STEP = 595.0, 1085.6, 1.1845e-3
DETC = 368.5


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# Table precision the committed counter summaries were recorded with; runs with
# other tables report no counter-derived figures.
PROFILED_SAT = "twofloat"


def traffic_from_profiles(kernel: str, sat: str = PROFILED_SAT):
    """HBM bytes per launch measured by rocprofv3 --pmc (committed summary)."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if sat != PROFILED_SAT or not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 2 and parts[0] == kernel:  # "key bytes" lines
            try:
                return float(parts[1])
            except ValueError:
                return None
    return None


def valu_from_profiles(kernel: str, sat: str = PROFILED_SAT):
    """Wave-level VALU instructions per launch measured by rocprofv3 --pmc
    SQ_INSTS_VALU (committed summary, "valu key count" lines)."""
    path = os.path.join(ROOT, "profiles", f"counters_512x512x64_2304v_{sat}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "valu" and parts[1] == kernel:
            try:
                return float(parts[2])
            except ValueError:
                return None
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def median_of_3(fn):
    """fn() once to warm up, then the median of three timed runs (seconds)."""
    fn()
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[1], ts


def cpu_nthreads():
    from oracle import oracle as O
    return O.nthreads_default()


def cpu_baseline(g, hangs, vol, mask, nsample=288):
    """FP32 restatement (projector + backprojector) of the 512x512x64,
    2304-view workload on `nsample` evenly spaced views, median of 3 after a
    warm-up, scaled to voxel-views/s."""
    import numpy as np

    from oracle import oracle as O
    # built for this host's CPU unless something already loaded the prebuilt one
    march = O.use_host_tuned_build() if O._lib is None else "x86-64-v3"
    nth = cpu_nthreads()
    gd = g.as_dict()
    sel = np.linspace(0, PN - 1, nsample).astype(np.int64)
    h = np.ascontiguousarray(hangs[sel])

    def run():
        p = O.proj(gd, h, vol, mask, fp64=False, nthreads=nth)
        O.back(gd, h, p, mask, XN, YN, fp64=False, nthreads=nth)

    med, ts = median_of_3(run)
    return {"value": 2.0 * XN * YN * SLN * nsample / med, "unit": "voxel-views/s",
            "cores": nth, "kind": "port", "cpu_model": cpu_model(),
            "sample": f"{nsample} of {PN} views (evenly spaced) of the 512x512x64 workload, FP32 "
                      f"projector + backprojector restatement (oracle/, gcc -O3 -march={march}, "
                      f"no fast-math), "
                      f"median of 3 after a warm-up: {med:.2f} s",
            "build": f"gcc -O3 -march={march} -ffp-contract=off", "runs_s": ts}


def bench_config0(dd, nth):
    """The single-slice CPU-sized workload end to end: 256x256 over 500 mm,
    736 channels, 1152 views, Shepp-Logan (water 0.0195/mm) plus 20 random
    ellipses, Poisson noise at I0 = 2.5e5; 10 OS-SART iterations x 24 subsets
    with FISTA on the FP32 CPU restatement (median of 3 after a warm-up), on
    the GPU device-resident (two-float tables, hipGraph) and in the
    reference's plain per-subset form; GPU vs FP64 oracle relative L2."""
    import numpy as np

    from aapm_lowdosectimagerecon_competition_amd import phantoms
    from oracle import oracle as O
    XN0, PN0, it, nos = 256, 1152, 10, 24
    dx0 = 500.0 / XN0
    conf = dd.Conf.make(XN=XN0, YN=XN0, SLN=1, dx=dx0, sid=SID, sdd=SDD, DNU=DNU,
                        detCellWidth=float(np.float32(2 * SDD * np.tan(STEP / 2))),
                        detCntIdx=DETC, PN=PN0)
    g0 = dd.detector_geometry(conf)
    h0 = phantoms.uniform_angles(PN0, 1)
    m0 = dd.recon_mask(XN0, XN0)
    img = phantoms.shepp_logan_2d(XN0, XN0, dx0, 250.0, seed=0)[:, :, None]
    clean = O.proj(g0.as_dict(), h0, img, np.ones((XN0, XN0), np.uint8))
    meas = phantoms.poisson_sinogram(clean, dx0, 2.5e5, seed=2)
    x0 = np.zeros((XN0, XN0, 1), np.float32)
    gd = g0.as_dict()
    cpu_s, cpu_runs = median_of_3(
        lambda: O.ossart(gd, h0, meas, m0, x0, it, nos, 1, fp64=False, nthreads=nth))
    ref = O.ossart(gd, h0, meas, m0, x0, it, nos, 1, fp64=True, nthreads=nth)
    with dd.Plan(g0, h0, m0, sat_mode="twofloat") as plan, \
            dd.DeviceBuffer.from_array(meas) as d_meas, dd.DeviceBuffer.from_array(x0) as d_x:
        def gpu():
            d_x.upload(x0)
            plan.ossart(d_meas, d_x, it, nos, 1, graph=True)
            plan.sync()
        gpu_s, _ = median_of_3(gpu)
        xg = d_x.to_numpy(x0.shape)
    tc = time.perf_counter()
    dd.ossart_aapm(np.asfortranarray(meas.T), np.asfortranarray(h0.T),
                   np.zeros((1, XN0, XN0), np.float32, order="F"), conf, it, nos,
                   np.asfortranarray(m0.T), 1, mode="plain", sat_mode="f32")
    plain_s = time.perf_counter() - tc
    vv = 2.0 * XN0 * XN0 * PN0 * it
    return {"workload": "256x256x1 @1.95 mm, 736 channels, 1152 views, 10 it x 24 subsets, "
                        "FISTA, Poisson I0 2.5e5",
            "cpu_s": cpu_s, "cpu_runs_s": cpu_runs, "cpu_voxel_views_per_s": vv / cpu_s,
            "gpu_device_s": gpu_s, "gpu_plain_s": plain_s,
            "gpu_device_speedup_vs_cpu": cpu_s / gpu_s,
            "rel_l2_gpu_vs_fp64_oracle": float(np.linalg.norm(xg - ref) / np.linalg.norm(ref))}


# ---------------------------------------------------------------- batched and served workloads
def mix_specs(n, seed):
    from aapm_lowdosectimagerecon_competition_amd import phantoms
    return phantoms.mix_specs(n, seed)


class DeviceJob:
    """One synthetic reconstruction resident on the GPU: plan, noisy sinogram
    and iterate (library device buffers), and the host copy of the sinogram."""

    def __init__(self, dd, spec, sat, seed, iters, keep_plan=True, tune=True):
        import numpy as np

        from aapm_lowdosectimagerecon_competition_amd import phantoms
        XN, SLN, PN = spec["XN"], spec["SLN"], spec["PN"]
        dx = 500.0 / XN
        conf = dd.Conf.make(XN=XN, YN=XN, SLN=SLN, dx=dx, sid=SID, sdd=SDD, DNU=DNU,
                            detCellWidth=float(np.float32(2 * SDD * np.tan(STEP / 2))),
                            detCntIdx=DETC, PN=PN)
        self.g = dd.detector_geometry(conf)
        self.hangs = phantoms.uniform_angles(PN, SLN)
        self.mask = dd.recon_mask(XN, XN)
        self.PN, self.iters, self.shape = PN, iters, (XN, XN, SLN)
        self.nos = max(1, PN // 48)  # subsets of 48 views
        self.vv = float(XN) * XN * SLN * PN  # voxel-views per operator call
        self.plan = dd.Plan(self.g, self.hangs, self.mask, sat_mode=sat)
        vol = phantoms.ellipsoid_phantom(XN, XN, SLN, dx, n=20, seed=seed)
        if tune:  # launch shapes for its OS-SART subsets, recorded for the server too
            self.plan.autotune(views=(PN + self.nos - 1) // self.nos, reps=2)
        clean = np.empty((PN, DNU, SLN), np.float32)
        with dd.DeviceBuffer.from_array(vol) as dv, dd.DeviceBuffer(clean.nbytes) as dp:
            self.plan.proj(dv, dp)
            self.plan.sync()
            dp.download(clean)
        self.meas_host = phantoms.gaussian_count_sinogram(clean, dx, spec["I0"], seed=seed + 1)
        del clean, vol
        self.meas = dd.DeviceBuffer.from_array(self.meas_host)
        self.x = dd.DeviceBuffer(4 * XN * XN * SLN)
        self.zero = np.zeros(self.shape, np.float32)
        if not keep_plan:
            self.plan.close()

    def reset(self):
        self.x.upload(self.zero)

    def result(self):
        return self.x.to_numpy(self.shape)

    def close(self):
        self.plan.close()
        self.meas.free()
        self.x.free()


def small_specs(n, seed):
    """Small reconstructions (256x256 over 500 mm, 8-16 slices, 720 views) that
    leave most of the GPU idle when run one at a time."""
    import numpy as np
    rng = np.random.default_rng(seed)
    return [{"XN": 256, "SLN": int(rng.integers(8, 17)), "PN": 720,
             "I0": float(np.exp(rng.uniform(np.log(5e4), np.log(5e5))))} for _ in range(n)]


def bench_batch(dd, sat, n=16, iters=10, seed=3, specs=None, streams=(0, 1, 2, 4), untuned=True):
    """Uneven reconstructions: one at a time (device OS-SART, graph, each plan's
    own autotuned shapes) against one batch (shared launches, graph per batch
    iteration), with the jobs in the automatic grouping (0) and in 1, 2 or 4
    groups side by side on their own streams; results must agree bit for bit.
    With `untuned`, the batch first runs in one stream with the shapes in
    place; then it tunes its shared launches for throughput (ddfan_batch_tune, outside the
    timed runs).  `batch_s` is the automatic grouping's time."""
    import numpy as np
    log(f"[bench] batch: generating {n} jobs")
    specs = specs or mix_specs(n, seed)
    jobs = [DeviceJob(dd, sp, sat, 1000 + i, iters) for i, sp in enumerate(specs)]
    vv = sum(2.0 * j.vv * j.iters for j in jobs)
    alone_s, ref = [], []
    log("[bench] batch: each job alone")
    for j in jobs:
        j.reset()
        t0 = time.perf_counter()
        j.plan.ossart(j.meas, j.x, j.iters, j.nos, 1, graph=True)
        alone_s.append(time.perf_counter() - t0)
        ref.append(j.result())
    runs = {}
    same = True
    tuned = None
    for ns in (("untuned",) if untuned else ()) + tuple(streams):
        for j in jobs:
            j.reset()
        log(f"[bench] batch: one batch, {ns} stream(s)")
        with dd.Batch(sat) as B:
            B.set_streams(1 if ns == "untuned" else ns)
            for j in jobs:
                B.add(j.plan, j.meas, j.x, j.iters, j.nos, 1)
            if ns != "untuned" and tuned is None:
                t0 = time.perf_counter()
                tuned = B.tune(2)
                tuned["tune_s"] = time.perf_counter() - t0
                log(f"[bench] batch: shared launches tuned {tuned}")
            t0 = time.perf_counter()
            B.run(graph=True)
            runs[ns] = time.perf_counter() - t0
        same = same and all(np.array_equal(j.result(), r) for j, r in zip(jobs, ref))
    out = {"jobs": n, "iters": iters, "subset_views": 48,
           "sizes": [f"{sp['XN']}x{sp['XN']}x{sp['SLN']}x{sp['PN']}" for sp in specs],
           "sequential_s": sum(alone_s), "batch_s": runs[0],
           "batch_s_by_streams": {("auto" if k == 0 else str(k)): v for k, v in runs.items()
                                  if k != "untuned"},
           "batch_s_untuned_one_stream": runs.get("untuned"),
           "shared_launch_tuning": tuned,
           "alone_s": alone_s,
           "speedup_vs_sequential": sum(alone_s) / runs[0],
           "voxel_views_per_s": vv / runs[0], "bitwise_equal_to_alone": bool(same),
           "noise": "normal approximation of Poisson counts, I0 log-uniform in [5e4, 5e5]"}
    return jobs, alone_s, out


class BigRequest:
    """One 768x768x256, 2304-view request of the served stream: its inputs,
    its time alone (device-resident OS-SART, 10 iterations) and that result."""

    def __init__(self, dd, sat, iters=10):
        self.job = DeviceJob(dd, {"XN": 768, "SLN": 256, "PN": 2304, "I0": 1e5}, sat, 7, iters)
        j = self.job
        j.reset()
        t0 = time.perf_counter()
        j.plan.ossart(j.meas, j.x, iters, j.nos, 1, graph=True)
        j.plan.sync()
        self.alone_s = time.perf_counter() - t0
        self.ref = j.result()
        self.g, self.hangs, self.mask = j.g, j.hangs, j.mask
        self.meas_host, self.nos = j.meas_host, j.nos

    def close(self):
        self.job.close()


def big_request(dd, sat):
    log("[bench] serve: generating a 768x768x256, 2304-view request")
    return BigRequest(dd, sat)


def bench_serve(dd, jobs, alone_s, sat, load=0.8, nreq=40, iters=10, seed=4, slots=4, big=None,
                graphs=False, policy="fifo", max_wait_ms=None):
    """Poisson arrivals of `nreq` requests drawn from the batch mix plus one
    768x768x256, 2304-view request (`big`), each `iters` OS-SART iterations,
    at `load` times the GPU's capacity for mix requests (capacity = 1 / the
    mean time of a mix request run alone).  Latency is submission to result
    on the host; the offered load counts every request's time alone."""
    import numpy as np
    rng = np.random.default_rng(seed)
    mean_alone = float(np.mean(alone_s))
    rate = load / mean_alone
    picks = [int(rng.integers(len(jobs))) for _ in range(nreq)]
    reqs = [jobs[i] for i in picks]
    solo = [alone_s[i] for i in picks]
    ibig = None
    if big is not None:
        ibig = nreq // 3
        reqs.insert(ibig, big)
        solo.insert(ibig, big.alone_s)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(reqs)))
    log(f"[bench] serve: {len(reqs)} requests at {rate:.2f}/s (load {load}), graphs {graphs}")
    with dd.Server(hbm_budget=0, max_slots=slots, sat_mode=sat, graphs=graphs, policy=policy,
                   max_wait_ms=max_wait_ms) as S:
        t0 = time.perf_counter()
        ids = []
        for i, j in enumerate(reqs):
            wait = arrivals[i] - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            ids.append(S.submit(j.g, j.hangs, j.meas_host, iters, j.nos, 1, mask=j.mask))
        res = [S.wait(r) for r in ids]
        t1 = time.perf_counter()
        st = S.stats()
    lat = np.array([r.latency_ms for r in res])
    q = np.array([r.queue_ms for r in res])
    mix = np.array([k != ibig for k in range(len(res))])
    out = {"requests": len(reqs), "load": load, "rate_per_s": rate,
           "capacity_req_per_s": 1.0 / mean_alone,
           "offered_load": float(sum(solo) / arrivals[-1]), "iters": iters, "slots": slots,
           "policy": policy, "max_wait_ms": max_wait_ms,
           "graphs": graphs, "p50_ms": float(np.percentile(lat, 50)),
           "p99_ms": float(np.percentile(lat, 99)),
           "mix_p50_ms": float(np.percentile(lat[mix], 50)),
           "mix_p99_ms": float(np.percentile(lat[mix], 99)),
           "mean_queue_ms": float(q.mean()), "wall_s": t1 - t0,
           "voxel_views_per_s": st["voxel_views"] / (t1 - t0), "failed": st["failed"]}
    if big is not None:
        out["large_request"] = {"shape": "768x768x256x2304", "latency_ms": float(lat[ibig]),
                                "alone_s": big.alone_s,
                                "equals_alone": bool(np.array_equal(res[ibig].x, big.ref))}
    return out


# ---------------------------------------------------------------- helical workload
def bench_helical(dd, sat, iters, nos=48, SLN2=200, rotations=20, seed=5):
    """Helical scan (pitch 0.8, 64 rows x 1.09472 mm at the detector, 736
    channels, 1152 views per rotation) as uint16 counts of a water cylinder
    with a documented gain of 0.5 at I0 = 1e5, rebinned on the device to
    200 fan-beam slices in the sinogram layout (counts -> line integrals fused
    into the gather), then OS-SART on 512x512x200 slices with the rebinned
    per-slice angles inside the scanned field of view."""
    import numpy as np

    from aapm_lowdosectimagerecon_competition_amd import phantoms, rebin
    W, H, DefTimes, pitch = DNU, 64, 1152, 0.8
    SD, SO = SDD, SID
    PerDetW = float(np.float32(2 * SD * np.tan(STEP / 2)))
    PerDetH = 0.6 * SD / SO
    TV = DefTimes * rotations
    I0, gain = 1e5, 0.5
    log(f"[bench] helical: {TV} raw views x {W} x {H} uint16")
    raw = phantoms.helical_cylinder_counts(W, H, DefTimes, rotations, SO, STEP, DETC, I0=I0,
                                           gain=gain, seed=seed)
    h = pitch * PerDetH * H * SO / SD
    deltaZ = h / DefTimes
    zPos = np.linspace(h, deltaZ * TV - h, SLN2).astype(np.float32)
    d_raw = dd.DeviceBuffer.from_array(raw)
    del raw
    d_prj = dd.DeviceBuffer(4 * DefTimes * W * SLN2)
    d_views = dd.DeviceBuffer(4 * DefTimes * SLN2)

    def rebin_call():
        rebin.helical_to_fan_device(d_raw, 1, gain * I0, 1, TV, zPos, SD, SO, 0.0, W, H, PerDetW,
                                    PerDetH, DefTimes, DETC, pitch, d_prj, d_views)

    rebin_call()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        rebin_call()
    rebin_s = (time.perf_counter() - t0) / reps
    samples = DefTimes * W * SLN2
    out = {"raw": f"{TV}x{W}x{H} uint16 (gain {gain}, I0 {I0:g})",
           "rebinned": f"{SLN2} slices x {DefTimes} views x {W} channels",
           "rebin_ms_per_call": rebin_s * 1e3,
           "rebin_GBs_algorithmic": samples * 8 / rebin_s / 1e9}
    log(f"[bench] helical rebin {out}")
    if iters <= 0:  # rebin only
        d_raw.free()
        d_prj.free()
        d_views.free()
        return out
    views = d_views.to_numpy((DefTimes, SLN2))
    XN2 = 512
    conf = dd.Conf.make(XN=XN2, YN=XN2, SLN=SLN2, dx=500.0 / XN2, sid=SID, sdd=SDD, DNU=DNU,
                        detCellWidth=PerDetW, detCntIdx=DETC, PN=DefTimes)
    g = dd.detector_geometry(conf)
    # Mask = the scanned field of view (radius SO*sin(half fan angle)).  The
    # reference's mask (radius 1.3x the half-width, reconscript.m:28-38) keeps
    # image corners no ray of the fan crosses; with 24-view subsets OS-SART
    # diverges there within a few iterations.
    fov = SO * np.sin((DNU / 2) * STEP)
    yy, xx = np.mgrid[0:XN2, 0:XN2]
    mask = (np.hypot(xx - (XN2 - 1) / 2, yy - (XN2 - 1) / 2) * g.dx < 0.98 * fov).astype(np.uint8)
    with dd.Plan(g, views, mask, sat_mode=sat) as plan, \
            dd.DeviceBuffer(4 * XN2 * XN2 * SLN2) as d_x:
        plan.autotune(views=DefTimes // nos, reps=2)
        d_x.upload(np.zeros((XN2, XN2, SLN2), np.float32))
        plan_chunk = plan.slice_chunk
        log(f"[bench] helical: OS-SART {iters} it x {nos} subsets, slice chunk {plan_chunk}")
        t0 = time.perf_counter()
        plan.ossart(d_prj, d_x, iters, nos, 1, graph=True)
        os_s = time.perf_counter() - t0
        x = d_x.to_numpy((XN2, XN2, SLN2))
    # The same reconstruction with the backprojector's channels aligned to the
    # projector's (DetectorGeometry.matched_backprojector; not the reference's
    # arithmetic, whose one-channel offset keeps 24-view subsets from converging).
    with dd.Plan(g.matched_backprojector(), views, mask, sat_mode=sat) as plan, \
            dd.DeviceBuffer(4 * XN2 * XN2 * SLN2) as d_x:
        plan.apply_tuned(views=DefTimes // nos)
        d_x.upload(np.zeros((XN2, XN2, SLN2), np.float32))
        t0 = time.perf_counter()
        plan.ossart(d_prj, d_x, iters, nos, 1, graph=True)
        om_s = time.perf_counter() - t0
        xm = d_x.to_numpy((XN2, XN2, SLN2))
    matched = {"ossart_s": om_s,
               "recon_center_mu_per_mm": float(xm[XN2 // 2 - 8:XN2 // 2 + 8,
                                                  XN2 // 2 - 8:XN2 // 2 + 8].mean()) / g.dx,
               "recon_max_abs": float(np.abs(xm).max()),
               "recon_p99_abs": float(np.percentile(np.abs(xm), 99))}
    vv = 2.0 * XN2 * XN2 * SLN2 * DefTimes * iters
    out.update({"ossart_iters": iters, "ossart_subsets": nos, "ossart_s": os_s,
                "ossart_voxel_views_per_s": vv / os_s, "slice_chunk": plan_chunk,
                "recon_center_mu_per_mm": float(x[XN2 // 2 - 8:XN2 // 2 + 8,
                                                  XN2 // 2 - 8:XN2 // 2 + 8].mean()) / g.dx,
                "recon_max_abs": float(np.abs(x).max()),
                "recon_p99_abs": float(np.percentile(np.abs(x), 99)),
                "recon_nonfinite": int(np.count_nonzero(~np.isfinite(x))),
                "mask": f"scanned field of view, radius {0.98 * fov:.1f} mm",
                "matched_backprojector": matched})
    d_raw.free()
    d_prj.free()
    d_views.free()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--sat", default="twofloat", choices=["f32", "twofloat", "f64"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ossart", action="store_true")
    ap.add_argument("--ossart-iters", type=int, default=10)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched and served mix")
    ap.add_argument("--no-tune", action="store_true", help="default launch shapes, no autotuning")
    ap.add_argument("--helical-iters", type=int, default=20,
                    help="OS-SART iterations of the helical workload (0 skips it)")
    ap.add_argument("--shapes", help="launch-shape table to apply instead of autotuning "
                                     "(written by --save-shapes)")
    ap.add_argument("--save-shapes", help="write the autotuned launch-shape table here")
    ap.add_argument("--rebin-pass", action="store_true",
                    help="only the helical workload's device rebin (23040 uint16 raw views -> "
                         "200 slices x 1152 views, 6 calls); prints no JSON (for rocprofv3 runs)")
    ap.add_argument("--profile-pass", action="store_true",
                    help="one projector + backprojector call, then one 256 MiB device-to-device "
                         "copy whose counter readings calibrate FETCH_SIZE / WRITE_SIZE; "
                         "prints no JSON (for rocprofv3 runs)")
    ap.add_argument("--ossart-pass", action="store_true",
                    help="only the device-resident OS-SART (--ossart-iters iterations x 24 "
                         "subsets, hipGraph per iteration, after one warm-up iteration) on the "
                         "projection of the phantom; prints no JSON (for rocprofv3 runs)")
    ap.add_argument("--helical-pass", action="store_true",
                    help="only the helical workload (device rebin, then --helical-iters OS-SART "
                         "iterations with the reference's and with the aligned backprojector); "
                         "prints no JSON (for rocprofv3 runs)")
    ap.add_argument("--batch-pass", action="store_true",
                    help="only the batch of 16 small reconstructions (256x256 x 8-16 slices, 720 "
                         "views, 48-view subsets): shared launches tuned, then 2 iterations in "
                         "one stream with direct launches; prints no JSON (for rocprofv3 runs)")
    ap.add_argument("--ossart-eager", action="store_true",
                    help="with --ossart-pass: launch the iterations eagerly, without hipGraphs")
    ap.add_argument("--serve-load", type=float, default=0.8,
                    help="offered load of the served stream, as a fraction of the GPU's capacity "
                         "(1 / the mean time of a mix request run alone)")
    ap.add_argument("--serve-requests", type=int, default=40)
    ap.add_argument("--serve-policy", default="shortest", choices=["shortest", "fifo"],
                    help="admission order of the served stream (the other policy is measured "
                         "on the same arrivals beside it)")
    ap.add_argument("--serve-max-wait-ms", type=int, default=4000,
                    help="age limit of shortest-first admission")
    ap.add_argument("--serve-graphs", action="store_true",
                    help="the request loop replays each request set's iteration as a hipGraph")
    ap.add_argument("--serve-sweep", action="store_true",
                    help="also serve 100-request streams at 0.5, 0.8 and 1.0 of capacity")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the path shards by slice with no exchange); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import aapm_lowdosectimagerecon_competition_amd as dd
    from aapm_lowdosectimagerecon_competition_amd import phantoms

    if dd.device_count() < 1:
        raise RuntimeError("no GPU visible")
    if args.batch_pass:
        specs = small_specs(16, 7)
        jobs = [DeviceJob(dd, sp, args.sat, 1000 + i, 2) for i, sp in enumerate(specs)]
        with dd.Batch(args.sat) as B:
            B.set_streams(1)
            for j in jobs:
                j.reset()
                B.add(j.plan, j.meas, j.x, j.iters, j.nos, 1)
            tuned = B.tune(2)
            ta = time.perf_counter()
            B.run(graph=False)
            tb = time.perf_counter()
        for j in jobs:
            j.close()
        log(f"[bench] batch pass done: tuned {tuned}, 2 iterations {tb - ta:.4f} s")
        return
    if args.rebin_pass:
        bench_helical(dd, args.sat, 0)
        log("[bench] rebin pass done")
        return
    if args.helical_pass:
        h = bench_helical(dd, args.sat, args.helical_iters)
        log(f"[bench] helical pass done: OS-SART {h.get('ossart_s', 0.0):.3f} s")
        return
    conf = dd.Conf.make(XN=XN, YN=YN, SLN=SLN, dx=DX, sid=SID, sdd=SDD, DNU=DNU,
                        detCellWidth=float(np.float32(2 * SDD * np.tan(STEP / 2))),
                        detCntIdx=DETC, PN=PN)
    g = dd.detector_geometry(conf)
    hangs = phantoms.uniform_angles(PN, SLN)
    vol = phantoms.ellipsoid_phantom(XN, YN, SLN, DX, seed=1)
    mask = dd.recon_mask(XN, YN)
    vv = XN * YN * SLN * PN

    log(f"[bench] plan sat={args.sat}")
    plan = dd.Plan(g, hangs, mask, sat_mode=args.sat)
    if args.shapes:
        dd.tune_load(args.shapes)

    def shape_for(views):
        """Launch shape for calls of `views` views: from --shapes, else autotuned."""
        if args.no_tune:
            return None
        if args.shapes:
            if not plan.apply_tuned(views=views):
                raise RuntimeError(f"{args.shapes} has no entry for {views}-view calls")
            return {"from": args.shapes}
        return plan.autotune(views=views, reps=2)

    tuned = shape_for(PN)
    log(f"[bench] launch shape {tuned}")
    d_vol = dd.DeviceBuffer.from_array(vol)
    d_prj = dd.DeviceBuffer(4 * PN * DNU * SLN)
    d_bp = dd.DeviceBuffer(4 * XN * YN * SLN)
    if args.profile_pass:
        plan.proj(d_vol, d_prj)
        plan.back(d_prj, d_bp)
        plan.sync()
        with dd.DeviceBuffer(COPY_BYTES) as a, dd.DeviceBuffer(COPY_BYTES) as b:
            b.copy_from(a)  # hipMemcpy device-to-device: returns when the copy is done
        log("[bench] profile pass done")
        return
    if args.ossart_pass:
        nos, it = 24, args.ossart_iters
        plan.proj(d_vol, d_prj)
        shape_for(PN // nos)
        with dd.DeviceBuffer(4 * XN * YN * SLN) as d_x:
            graph = not args.ossart_eager
            plan.ossart(d_prj, d_x, 1, nos, 1, graph=graph)  # warm-up iteration
            d_x.upload(np.zeros((YN, XN, SLN), np.float32))
            ta = time.perf_counter()
            plan.ossart(d_prj, d_x, it, nos, 1, graph=graph)
            tb = time.perf_counter()
        log(f"[bench] ossart pass done: {it} it x {nos} subsets {tb - ta:.4f} s")
        return
    log(f"[bench] warmup {args.warmup}")
    for _ in range(args.warmup):
        plan.proj(d_vol, d_prj)
        plan.back(d_prj, d_bp)
    plan.sync()
    plan.timing(True)
    log(f"[bench] timed {args.steps} steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.proj(d_vol, d_prj)
        plan.back(d_prj, d_bp)
    plan.sync()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    kt = {}
    for kind, name in ((0, "sat_vol"), (1, "proj"), (2, "sat_sino"), (3, "back")):
        tot, n = plan.timing_read(kind)
        kt[name] = tot / max(n, 1)
    plan.timing(False)
    log(f"[bench] ms/step {ms_step:.3f} kernels {kt}")

    wide = args.sat != "f32"
    # per-view-sweep byte model (one volume-side operand per voxel-view; LDS
    # and the caches serve most of it, so it can exceed the HBM peak)
    model_bytes = {"ddfan_proj": (8 if wide else 4) * vv + 4 * SLN * DNU * PN + 8 * SLN * PN,
                   "ddfan_back": 4 * vv + 4 * SLN * (DNU + 1) * PN + 4 * XN * YN * SLN}
    kms = {"ddfan_proj": kt["proj"], "ddfan_back": kt["back"]}
    per_kernel = {}
    for k, ms in kms.items():
        sec = ms * 1e-3
        wi = valu_from_profiles(k, args.sat)
        tb = traffic_from_profiles(k, args.sat)
        per_kernel[k] = {
            "ms": ms,
            "valu_lane_ops_per_s": None if wi is None else wi * 64 / sec,
            "valu_frac": None if wi is None else wi * 64 / sec / VALU_PEAK_LANE_OPS,
            "hbm_GBs": None if tb is None else tb / sec / 1e9,
            "hbm_frac": None if tb is None else tb / sec / 1e9 / HBM_PEAK_GBS,
            "model_GBs": model_bytes[k] / sec / 1e9,
            "model_frac": model_bytes[k] / sec / 1e9 / HBM_PEAK_GBS}
    dom = max(kms, key=kms.get)
    d = per_kernel[dom]
    if d["valu_frac"] is not None and d["valu_frac"] > 1.0:
        log(f"[bench] WARNING: VALU fraction {d['valu_frac']:.3f} > 1: the committed counters do "
            "not match this build's kernels")
    roof = {"bound": "valu",
            "achieved": None if d["valu_lane_ops_per_s"] is None else d["valu_lane_ops_per_s"] / 1e12,
            "peak": VALU_PEAK_LANE_OPS / 1e12, "unit": "T lane-instructions/s (FP32 VALU issue)",
            "frac": d["valu_frac"], "traffic": traffic_from_profiles(dom, args.sat), "kernel": dom,
            "hbm_GBs": d["hbm_GBs"], "hbm_frac": d["hbm_frac"],
            "model_GBs": d["model_GBs"], "model_frac": d["model_frac"],
            "model_bytes_per_launch": model_bytes[dom],
            "step_hbm_GBs": (None if any(traffic_from_profiles(k, args.sat) is None for k in
                                         ("ddfan_proj", "ddfan_back", "ddfan_sat_sino",
                                          "ddfan_sat_vol"))
                             else sum(traffic_from_profiles(k, args.sat) for k in
                                      ("ddfan_proj", "ddfan_back", "ddfan_sat_sino",
                                       "ddfan_sat_vol")) / (ms_step * 1e-3) / 1e9),
            "per_kernel": per_kernel,
            "counters": (f"profiles/counters_512x512x64_2304v_{PROFILED_SAT}.txt, "
                         "profiles/traffic.txt" if args.sat == PROFILED_SAT else
                         f"none: the committed counters are of {PROFILED_SAT} tables")}

    out = {"metric": METRIC, "value": 2.0 * vv / (ms_step * 1e-3), "unit": "voxel-views/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded random-ellipsoid phantom, circular FOV mask)",
           "config": {"workload": "DD projector + backprojector, full view set, device-resident: "
                                  "512x512x64 @0.98 mm, 736-channel arc, 2304 views",
                      "XN": XN, "YN": YN, "SLN": SLN, "DNU": DNU, "PN": PN, "sat": args.sat,
                      "slice_chunk": plan.slice_chunk,
                      "launch_shape": {"proj_chunk": plan.proj_chunk,
                                       "views_per_wave": plan.views_per_wave,
                                       "slab_tile": plan.slab_tile,
                                       "back_chunk": plan.back_chunk,
                                       "back_stage": plan.back_stage,
                                       "proj_stage": plan.proj_stage,
                                       "autotuned": tuned is not None}},
           "roofline": roof, "kernels_ms": kt}

    # the drop-in entry points on host buffers (plan, copies in and out, free
    # per call, as the reference's MEX calls do): PCIe-inclusive, not `value`
    from aapm_lowdosectimagerecon_competition_amd import _native as N
    hprj = np.empty((PN, DNU, SLN), np.float32)
    hbp = np.empty((YN, XN, SLN), np.float32)
    t0 = time.perf_counter()
    N.call("ddfan_proj", g.x0, g.y0, DNU, N.ptr(g.xds), N.ptr(g.yds), g.imgXCenter, g.imgYCenter,
           N.ptr(hangs), PN, XN, YN, SLN, N.ptr(vol), N.ptr(hprj), g.dx, N.ptr(mask), None, 0)
    t1 = time.perf_counter()
    N.call("ddfan_back", g.x0, g.y0, DNU, N.ptr(g.xds), N.ptr(g.yds), g.detCntIdx, g.imgXCenter,
           g.imgYCenter, N.ptr(hangs), PN, XN, YN, SLN, N.ptr(hbp), N.ptr(hprj), g.dx,
           N.ptr(mask), None, 0)
    t2 = time.perf_counter()
    out["dropin_host_buffers"] = {"proj_ms": (t1 - t0) * 1e3, "back_ms": (t2 - t1) * 1e3,
                                  "voxel_views_per_s": 2.0 * vv / (t2 - t0),
                                  "sat": "f32 (drop-in default)"}
    log(f"[bench] drop-in host buffers {out['dropin_host_buffers']}")

    if not args.no_ossart:
        # device-resident fused + graph OS-SART vs the plain per-subset host form
        nos, it = 24, args.ossart_iters
        meas = np.empty((PN, DNU, SLN), np.float32)
        d_prj.download(meas)
        x0 = np.zeros((YN, XN, SLN), np.float32)
        shape_for(PN // nos)
        log(f"[bench] ossart device {it} it x {nos} subsets")
        d_meas = dd.DeviceBuffer.from_array(meas)
        d_x = dd.DeviceBuffer.from_array(x0)
        plan.ossart(d_meas, d_x, 1, nos, 1, graph=True)  # warm
        d_x.upload(x0)
        ta = time.perf_counter()
        plan.ossart(d_meas, d_x, it, nos, 1, graph=True)
        tb = time.perf_counter()
        xd = d_x.to_numpy((YN, XN, SLN))
        log(f"[bench] ossart plain {it} it x {nos} subsets")
        P_m = np.asfortranarray(meas.T)
        V_m = np.asfortranarray(hangs.T)
        init = np.zeros((SLN, XN, YN), np.float32, order="F")
        tc = time.perf_counter()
        xp = dd.ossart_aapm(P_m, V_m, init, conf, it, nos, np.asfortranarray(mask.T), 1,
                            mode="plain", sat_mode="f32")
        td = time.perf_counter()
        rd = float(np.linalg.norm(xd - xp.T) / max(np.linalg.norm(xp), 1e-30))
        # the same job through the request loop, alone: submit (inputs copied
        # from the host) -> result on the host, like the plain form's host buffers
        log(f"[bench] ossart served {it} it x {nos} subsets")
        with dd.Server(hbm_budget=0, max_slots=4, sat_mode=args.sat) as srv:
            te = time.perf_counter()
            res = srv.wait(srv.submit(g, hangs, meas, it, nos, 1, mask=mask, x0=x0))
            tf = time.perf_counter()
        out["ossart"] = {"iters": it, "subsets": nos, "device_s": tb - ta, "plain_s": td - tc,
                         "served_s": tf - te,
                         "speedup_vs_plain": (td - tc) / (tb - ta),
                         "speedup_served_vs_plain": (td - tc) / (tf - te),
                         "device_voxel_views_per_s": 2.0 * vv * it / (tb - ta),
                         "rel_diff_device_twofloat_vs_plain_f32": rd,
                         "served_equals_device": bool(np.array_equal(res.x, xd)),
                         "rel_diff_served_vs_device": float(np.linalg.norm(res.x - xd) /
                                                            max(np.linalg.norm(xd), 1e-30))}
        d_meas.free()
        d_x.free()

    if not args.no_batch:
        d_vol.free()
        d_prj.free()
        d_bp.free()
        plan.close()
        small, _, out["batch_small"] = bench_batch(dd, args.sat, specs=small_specs(16, 7))
        log(f"[bench] batch of small jobs {out['batch_small']}")
        for j in small:
            j.close()
        del small
        jobs, alone_s, out["batch"] = bench_batch(dd, args.sat, untuned=False)
        log(f"[bench] batch {out['batch']}")
        big = big_request(dd, args.sat)
        mw = args.serve_max_wait_ms
        out["serve"] = bench_serve(dd, jobs, alone_s, args.sat, load=args.serve_load,
                                   nreq=args.serve_requests, big=big, graphs=args.serve_graphs,
                                   policy=args.serve_policy,
                                   max_wait_ms=mw if args.serve_policy == "shortest" else None)
        log(f"[bench] serve {out['serve']}")
        other = "fifo" if args.serve_policy == "shortest" else "shortest"
        o = bench_serve(dd, jobs, alone_s, args.sat, load=args.serve_load,
                        nreq=args.serve_requests, big=big, graphs=args.serve_graphs, policy=other,
                        max_wait_ms=mw if other == "shortest" else None)
        out["serve"]["other_policy"] = {k: o[k] for k in ("policy", "max_wait_ms", "p50_ms", "p99_ms",
                                                          "mean_queue_ms", "voxel_views_per_s")}
        out["serve"]["other_policy"]["large_request_ms"] = o["large_request"]["latency_ms"]
        log(f"[bench] serve, {other} admission {out['serve']['other_policy']}")
        if args.serve_sweep:
            out["serve_sweep"] = [bench_serve(dd, jobs, alone_s, args.sat, load=ld, nreq=100,
                                              big=big, graphs=args.serve_graphs, seed=10 + i,
                                              policy=pol, max_wait_ms=mw if pol == "shortest" else None)
                                  for i, ld in enumerate((0.5, 0.8, 1.0))
                                  for pol in ("shortest", "fifo")]
            log(f"[bench] serve sweep {out['serve_sweep']}")
        big.close()
        for j in jobs:
            j.close()
        del jobs

    if args.helical_iters > 0 and not args.no_batch:
        out["helical"] = bench_helical(dd, args.sat, args.helical_iters)
        log(f"[bench] helical {out['helical']}")

    if args.save_shapes:
        dd.tune_save(args.save_shapes)

    if not args.no_cpu:
        log("[bench] cpu baseline")
        out["cpu_baseline"] = cpu_baseline(g, hangs, vol, mask)
        log(f"[bench] cpu baseline {out['cpu_baseline']}")
        log("[bench] config 0 end to end (CPU restatement, GPU)")
        out["cpu_baseline"]["config0_ossart"] = bench_config0(dd, out["cpu_baseline"]["cores"])
        log(f"[bench] config 0 {out['cpu_baseline']['config0_ossart']}")
    d_vol.free()
    d_prj.free()
    d_bp.free()
    plan.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
