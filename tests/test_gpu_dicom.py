"""DICOM-CT-PD scan to rebinned sinogram on the GPU: the projection files as
stored (read_proj_frames: the files' frames into one host array, no
reordering) rebinned on the device with the frames read in place
(ORDER_FRAMES_REVERSED, uint16 line integrals unscaled) equal, bit for bit,
the reference's route (readProj.m -> HelicalToFan_routine.m: read_proj and
the drop-in rebin) and the FP32 restatement; slice positions and geometry
come from the file headers (reconscript.m:18-22, HelicalToFan_routine.m)."""
import numpy as np
import pytest

import dicom_writer as DW

pytestmark = pytest.mark.gpu

W, H, DT, TURNS = 32, 8, 48, 6


def test_dicom_scan_rebinned_on_device(tmp_path, oracle):
    import aapm_lowdosectimagerecon_competition_amd as dd
    from aapm_lowdosectimagerecon_competition_amd import rebin as R
    geom = dict(pitch=0.8, PerDetW=1.2859 * 23.0, PerDetH=1.09472, angle=17.25, SO=595.0,
                SD=1085.6, cx=16.5, cy=4.5, DefTimes=DT)
    rng = np.random.default_rng(21)
    TV = DT * TURNS
    stored = rng.integers(0, 6000, (TV, W, H), dtype=np.uint16)  # files in order
    scan = tmp_path / "scan"
    scan.mkdir()
    for f in range(TV):
        DW.write(scan / f"L000_proj.{f + 1:05d}.dcm", DW.ctpd_elements(geom, stored[f]),
                 ts=DW.IMPLICIT if f % 2 else DW.EXPLICIT)
    print(f"\n[dicom] {TV} projection files of {W}x{H}", flush=True)
    frames, cfg = dd.read_proj_frames(scan)
    assert np.array_equal(frames, stored)
    hc = cfg.helical()
    SLN = 7
    z = R.slice_positions(hc, cfg.NumOfDataViews, SLN).astype(np.float32)
    BV = np.float32(hc.DetectorFocalCenterAngularPosition / 180.0 * np.pi)  # routine.m:22
    with dd.DeviceBuffer.from_array(frames) as d_raw, \
            dd.DeviceBuffer(4 * SLN * DT * W) as d_out, dd.DeviceBuffer(4 * SLN * DT) as d_v:
        R.helical_to_fan_device(d_raw, R.RAW_LINE_INTEGRALS_U16, 1.0, 0, TV, z,
                                hc.ConstantRadialDistance, hc.DetectorFocalCenterRadialDistance,
                                BV, W, H, hc.DetectorElementTransverseSpacing,
                                hc.DetectorElementAxialSpacing, DT, hc.DetectorCentralElementX,
                                hc.SpiralPitchFactor, d_out, d_v,
                                raw_order=R.ORDER_FRAMES_REVERSED)
        dev_p = d_out.to_numpy((SLN, DT, W))
        dev_v = d_v.to_numpy((SLN, DT))
    proj, cfg2 = dd.read_proj(scan)  # readProj.m
    Proj, Views = dd.helical_to_fan_routine(proj, cfg2.helical(), z)  # (SLN, W, DT), (SLN, DT)
    assert np.array_equal(dev_p, np.transpose(Proj, (0, 2, 1)))
    assert np.array_equal(dev_v, np.asarray(Views))
    raw = np.ascontiguousarray(np.asarray(proj).T)  # [view, ch, row]
    o = oracle.helical_to_fan(raw, z, hc.ConstantRadialDistance,
                              hc.DetectorFocalCenterRadialDistance, BV,
                              hc.DetectorElementTransverseSpacing, hc.DetectorElementAxialSpacing,
                              DT, hc.DetectorCentralElementX, hc.SpiralPitchFactor, fp64=False)
    assert np.array_equal(dev_p, o["Proj"])
    assert np.array_equal(dev_v, o["Views"])
    print(f"[dicom] {SLN} slices bit-identical; cut views per slice "
          f"{int(o['cut'].sum(axis=1).min())}..{int(o['cut'].sum(axis=1).max())}", flush=True)
