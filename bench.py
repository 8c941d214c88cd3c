#!/usr/bin/env python3
"""Benchmark of the DD fan-beam path on one MI355X.

Step = one full-view DD projector call plus one full-view DD backprojector
call on the 512x512x64-slice, 736-channel, 2304-view workload (0.98 mm
pixels, arc detector, sid 595 mm, sdd 1085.6 mm), with the volume and
sinogram resident in HBM.  value = voxel-views/s over the timed steps
(2 * 512*512*64*2304 voxel-views per step).

Also reported on the same line:
  roofline      the dominant kernel's algorithmic bytes / its average
                launch duration (HIP events on the plan's stream)
  kernels       average kernel durations
  cpu_baseline  the FP32 CPU restatement on a bounded sample of views
  ossart        device-resident fused + hipGraph OS-SART against the
                reference's plain per-subset form on the same workload
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


def traffic_from_profiles(kernel: str):
    """HBM bytes per launch measured by rocprofv3 --pmc (committed summary)."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 2 and parts[0] == kernel:  # "key bytes" lines
            try:
                return float(parts[1])
            except ValueError:
                return None
    return None


def cpu_baseline(g, hangs, vol, mask, target_s=20.0):
    """FP32 restatement (projector + backprojector) on a bounded sample of
    the workload's views, scaled to voxel-views/s."""
    import numpy as np

    from oracle import oracle as O
    nth = int(os.environ.get("OMP_NUM_THREADS", "0")) or O.nthreads_default()
    nth = max(1, min(nth, O.nthreads_default()))
    gd = g.as_dict()

    def run(nv):
        sel = np.linspace(0, PN - 1, nv).astype(np.int64)
        h = np.ascontiguousarray(hangs[sel])
        t0 = time.perf_counter()
        p = O.proj(gd, h, vol, mask, fp64=False, nthreads=nth)
        O.back(gd, h, p, mask, XN, YN, fp64=False, nthreads=nth)
        return time.perf_counter() - t0

    run(2)  # thread pool and caches warm
    per_view = run(8) / 8
    nv = int(max(8, min(PN, round(target_s / max(per_view, 1e-4)))))
    dt = run(nv)
    vv = 2.0 * XN * YN * SLN * nv
    return {"value": vv / dt, "unit": "voxel-views/s", "cores": nth, "kind": "port",
            "sample": f"{nv} of {PN} views (evenly spaced) of the 512x512x64 workload, "
                      f"FP32 projector + backprojector restatement, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--sat", default="twofloat", choices=["f32", "twofloat", "f64"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ossart", action="store_true")
    ap.add_argument("--ossart-iters", type=int, default=2)
    ap.add_argument("--profile-pass", action="store_true",
                    help="one projector + backprojector call, then one 256 MiB device-to-device "
                         "copy whose counter readings calibrate FETCH_SIZE / WRITE_SIZE; "
                         "prints no JSON (for rocprofv3 runs)")
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
    proj_bytes = (8 if wide else 4) * vv + 4 * SLN * DNU * PN + 8 * SLN * PN
    back_bytes = 4 * vv + 4 * SLN * (DNU + 1) * PN + 4 * XN * YN * SLN
    if kt["proj"] >= kt["back"]:
        dom, dbytes, dms = "ddfan_proj", proj_bytes, kt["proj"]
    else:
        dom, dbytes, dms = "ddfan_back", back_bytes, kt["back"]
    achieved = dbytes / (dms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profiles(dom),
            "kernel": dom, "bytes_per_launch": dbytes,
            "proj_GBs": proj_bytes / (kt["proj"] * 1e-3) / 1e9,
            "back_GBs": back_bytes / (kt["back"] * 1e-3) / 1e9}

    out = {"metric": METRIC, "value": 2.0 * vv / (ms_step * 1e-3), "unit": "voxel-views/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded random-ellipsoid phantom, circular FOV mask)",
           "config": {"workload": "DD projector + backprojector, full view set, device-resident: "
                                  "512x512x64 @0.98 mm, 736-channel arc, 2304 views",
                      "XN": XN, "YN": YN, "SLN": SLN, "DNU": DNU, "PN": PN, "sat": args.sat,
                      "slice_chunk": plan.slice_chunk},
           "roofline": roof, "kernels_ms": kt}

    if not args.no_ossart:
        # device-resident fused + graph OS-SART vs the plain per-subset host form
        nos, it = 24, args.ossart_iters
        meas = np.empty((PN, DNU, SLN), np.float32)
        d_prj.download(meas)
        x0 = np.zeros((YN, XN, SLN), np.float32)
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
        out["ossart"] = {"iters": it, "subsets": nos, "device_s": tb - ta, "plain_s": td - tc,
                         "speedup_vs_plain": (td - tc) / (tb - ta),
                         "device_voxel_views_per_s": 2.0 * vv * it / (tb - ta),
                         "rel_diff_device_twofloat_vs_plain_f32": rd}
        d_meas.free()
        d_x.free()

    if not args.no_cpu:
        log("[bench] cpu baseline")
        out["cpu_baseline"] = cpu_baseline(g, hangs, vol, mask)
    d_vol.free()
    d_prj.free()
    d_bp.free()
    plan.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
