"""Edge cases of the operators and the rebin on the GPU: empty view sets,
one slice, non-square images without a mask, slice counts that leave the
last slice chunk partly empty, fewer views than a projector wave groups, and
the widest detector the rebin accepts.  Values against the FP64 oracle
(relative L2 <= 5e-6, two-float tables) or exact where the result is."""
import numpy as np
import pytest

from conftest import rel_l2
from helpers import small_conf

pytestmark = pytest.mark.gpu


def _geom(XN, YN, SLN, PN, seed=3, mask=True):
    import aapm_lowdosectimagerecon_competition_amd as dd
    from aapm_lowdosectimagerecon_competition_amd import phantoms
    conf = small_conf(XN=XN, YN=YN, SLN=SLN)
    g = dd.detector_geometry(conf)
    hangs = phantoms.uniform_angles(PN, SLN)
    vol = phantoms.ellipsoid_phantom(XN, YN, SLN, conf.recon.dx, seed=seed)
    m = dd.recon_mask(XN, YN) if mask else None
    return dd, conf, g, hangs, vol, m


def test_zero_views():
    from aapm_lowdosectimagerecon_competition_amd import _native as N
    dd, conf, g, hangs, vol, mask = _geom(32, 24, 2, 8)
    print("\n[edges] zero views", flush=True)
    out = np.full(4, 7.0, np.float32)  # untouched: no views, no sinogram
    N.call("ddfan_proj", g.x0, g.y0, g.DNU, N.ptr(g.xds), N.ptr(g.yds), g.imgXCenter, g.imgYCenter,
           N.ptr(hangs), 0, g.XN, g.YN, g.SLN, N.ptr(vol), N.ptr(out), g.dx, N.ptr(mask), None, 0)
    assert np.all(out == 7.0)
    bp = np.full(vol.shape, 7.0, np.float32)
    N.call("ddfan_back", g.x0, g.y0, g.DNU, N.ptr(g.xds), N.ptr(g.yds), g.detCntIdx, g.imgXCenter,
           g.imgYCenter, N.ptr(hangs), 0, g.XN, g.YN, g.SLN, N.ptr(bp), N.ptr(out), g.dx,
           N.ptr(mask), None, 0)
    assert np.all(bp == 0.0)  # the backprojection of no views
    with dd.Plan(g, hangs, mask, sat_mode="twofloat") as plan, \
            dd.DeviceBuffer.from_array(vol) as dv, dd.DeviceBuffer(16) as dp:
        plan.proj(dv, dp, 0, 1, 0)
        plan.back(dp, dv, 0, 1, 0)
        plan.sync()
        assert np.all(dv.to_numpy(vol.shape) == 0.0)


@pytest.mark.parametrize("SLN,limit", [(1, 0), (9, 0), (17, 0), (17, 2), (30, 0)])
def test_slice_counts_and_nonsquare_without_mask(oracle, SLN, limit):
    """SLN 1 (one chunk of one slice), 9 (one chunk of 16, 7 slices of it
    padding), 17 (two chunks of 16, the second one slice; or, with chunks
    limited to 2, the last one half empty), 30 (chunks of 16, last one 14
    slices); 37x53 image, no mask."""
    dd, conf, g, hangs, vol, _ = _geom(37, 53, SLN, 20, mask=False)
    print(f"\n[edges] SLN {SLN} (chunk limit {limit}), 37x53, no mask", flush=True)
    with dd.Plan(g, hangs, None, sat_mode="twofloat", slice_chunk=limit) as plan:
        zt = plan.slice_chunk
        prj = plan.proj_host(vol)
        bp = plan.back_host(prj)
    ones = np.ones((53, 37), np.uint8)
    ref_p = oracle.proj(g.as_dict(), hangs, vol, ones)
    ref_b = oracle.back(g.as_dict(), hangs, prj, ones, 37, 53)
    ep, eb = rel_l2(prj, ref_p), rel_l2(bp, ref_b)
    print(f"[edges] chunk {zt}: projector {ep:.2e}, backprojector {eb:.2e}", flush=True)
    assert {(1, 0): 1, (9, 0): 16, (17, 0): 16, (17, 2): 2, (30, 0): 16}[(SLN, limit)] == zt
    assert ep <= 5e-6 and eb <= 5e-6


@pytest.mark.parametrize("sat", ["f32", "twofloat"])
@pytest.mark.parametrize("SLN", [13, 15])
def test_results_do_not_depend_on_slice_chunk(sat, SLN):
    """Fewer than 16 slices, padded to 16 by some chunks and not by others
    (13 slices: 13, 14 or 16 table slices; 15: 15 or 16): every projector chunk
    and every backprojector chunk gives the same bits, since the volume tables'
    builder follows the volume's own slice count."""
    dd, conf, g, hangs, vol, mask = _geom(48, 40, SLN, 36)
    print(f"\n[edges] chunk independence, SLN {SLN}, {sat} tables", flush=True)
    with dd.Plan(g, hangs, mask, sat_mode=sat) as plan:
        assert plan.slice_chunk == 16
        prj, bp = {}, {}
        for z in (1, 2, 4, 8, 16):
            plan.proj_chunk = z
            plan.proj_stage = 0
            prj[z] = plan.proj_host(vol)
        ref = prj[1]
        for z in (1, 2, 4, 8, 16):
            plan.back_chunk = z
            plan.back_stage = 0
            bp[z] = plan.back_host(ref)
    for z in (2, 4, 8, 16):
        assert np.array_equal(prj[z], prj[1]), ("projector", z)
        assert np.array_equal(bp[z], bp[1]), ("backprojector", z)


def test_fewer_views_than_a_wave_groups(oracle):
    dd, conf, g, hangs, vol, mask = _geom(40, 40, 4, 3)
    print("\n[edges] 3 views, 4 views per wave; single-view subsets", flush=True)
    with dd.Plan(g, hangs, mask, sat_mode="twofloat") as plan:
        plan.views_per_wave = 4
        prj = plan.proj_host(vol)
        one = plan.proj_host(vol, first=2, stride=1, count=1)
    ref = oracle.proj(g.as_dict(), hangs, vol, mask)
    assert rel_l2(prj, ref) <= 5e-6
    assert np.array_equal(one[0], prj[2])


def test_rebin_widest_detector(oracle):
    """2048 channels (the widest the gather handles: 8 per thread) match the
    FP32 restatement bit for bit; 2049 are refused."""
    import aapm_lowdosectimagerecon_competition_amd as dd
    from aapm_lowdosectimagerecon_competition_amd import rebin as R
    SD, SO, H, DT, turns, pitch = 1085.6, 595.0, 4, 16, 4, 0.8
    W = 2048
    PerDetW, PerDetH = 0.3, 1.09472
    rng = np.random.default_rng(2)
    raw = rng.uniform(0, 3, (DT * turns, W, H)).astype(np.float32)
    p = oracle.rebin_params(SD, SO, H, PerDetW, PerDetH, DT, pitch)
    z = np.linspace(float(p["h"]), float(p["deltaZ"]) * DT * turns - float(p["h"]), 3)
    z = z.astype(np.float32)
    print("\n[edges] rebin with 2048 channels", flush=True)
    P, V = R.helical_to_fan_func(np.asfortranarray(np.transpose(raw, (2, 1, 0))), z, 3, SD, SO,
                                 0.0, W, H, PerDetW, PerDetH, DT, W / 2 + 0.5, pitch)
    ref = oracle.helical_to_fan(raw, z, SD, SO, 0.0, PerDetW, PerDetH, DT, W / 2 + 0.5, pitch,
                                fp64=False)
    assert np.array_equal(np.transpose(P, (2, 1, 0)), ref["Proj"])
    raw2 = np.zeros((DT * turns, W + 1, H), np.float32)
    with pytest.raises(dd.DDFanError):
        R.helical_to_fan_func(np.asfortranarray(np.transpose(raw2, (2, 1, 0))), z, 3, SD, SO, 0.0,
                              W + 1, H, PerDetW, PerDetH, DT, W / 2 + 0.5, pitch)
