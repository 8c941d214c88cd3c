"""LDS-staged projector against the per-lane gather kernel: every slice chunk
x views per wave x slab pass shape gives bit-identical sinograms (tolerance
0), in all three table precisions, including footprints wider than the
128-entry staged run (lanes read global memory), images narrower than one
run, a non-square image (views of both slab orientations in one block), view
counts that leave idle waves in the last block, strided subsets, per-slice
angles and the OS-SART ratio epilogue."""
import numpy as np
import pytest

from conftest import rel_l2
from helpers import per_slice_angles, small_conf

pytestmark = pytest.mark.gpu

# 'coarse': the test scanner; 'wide': 8-channel bins over 1 mm voxels, so a
# block's 64 boundary rays cross ~360 entries of a row (most lanes outside the
# staged run); 'narrow': rows of 41 entries; 'oblong': 72 x 40 voxels.
CASES = [
    ("coarse", dict(XN=64, YN=64, SLN=8), 37),
    ("wide", dict(XN=160, YN=160, SLN=4, dx=1.0, DNU=64, bin=8), 23),
    ("narrow", dict(XN=40, YN=40, SLN=4, DNU=40, bin=8), 19),
    ("oblong", dict(XN=72, YN=40, SLN=8), 33),
]


def _setup(kw, PN, perslice=False, seed=8):
    import aapm_lowdosectimagerecon_competition_amd as dd
    from aapm_lowdosectimagerecon_competition_amd import phantoms
    conf = small_conf(**kw)
    g = dd.detector_geometry(conf)
    hangs = per_slice_angles(PN, g.SLN) if perslice else phantoms.uniform_angles(PN, g.SLN)
    vol = phantoms.ellipsoid_phantom(g.XN, g.YN, g.SLN, conf.recon.dx, seed=seed)
    return dd, g, hangs, vol, dd.recon_mask(g.XN, g.YN)


@pytest.mark.parametrize("sat", ["f32", "twofloat", "f64"])
@pytest.mark.parametrize("name,kw,PN", CASES, ids=[c[0] for c in CASES])
def test_staged_projector_equals_gather(sat, name, kw, PN):
    dd, g, hangs, vol, mask = _setup(kw, PN)
    print(f"\n[pstage] {name} {sat}: {g.XN}x{g.YN}x{g.SLN}, DNU {g.DNU}, {PN} views", flush=True)
    ran = 0
    with dd.Plan(g, hangs, mask, sat_mode=sat) as plan:
        widest = plan.slice_chunk
        for zt in (8, 4, 2, 1):
            if zt > widest:
                continue
            plan.proj_chunk = zt
            for V in (4, 2, 1):
                plan.views_per_wave = V
                for tile in (0, 16):
                    plan.slab_tile = tile
                    out = {}
                    for st in (0, 1):
                        plan.proj_stage = st
                        out[st] = (plan.proj_host(vol),
                                   plan.proj_host(vol, first=2, stride=3, count=(PN - 2 + 2) // 3))
                    assert np.array_equal(out[1][0], out[0][0]), (zt, V, tile)
                    assert np.array_equal(out[1][1], out[0][1]), (zt, V, tile)
                    ran += 1
    print(f"[pstage] {ran} chunk x views-per-wave x pass shapes bit-identical", flush=True)
    assert ran >= 6


def test_staged_projector_per_slice_angles_and_oracle(oracle):
    dd, g, hangs, vol, mask = _setup(dict(XN=48, YN=48, SLN=5), 21, perslice=True)
    print("\n[pstage] per-slice angles, chunk 1", flush=True)
    with dd.Plan(g, hangs, mask, sat_mode="twofloat") as plan:
        assert plan.slice_chunk == 1
        plan.proj_stage = 1
        staged = plan.proj_host(vol)
        plan.proj_stage = 0
        gather = plan.proj_host(vol)
    assert np.array_equal(staged, gather)
    e = rel_l2(staged, oracle.proj(g.as_dict(), hangs, vol, mask))
    print(f"[pstage] rel-L2 vs FP64 oracle {e:.3e}", flush=True)
    assert e <= 5e-6


def test_staged_projector_in_ossart():
    """The ratio epilogue (b - Ax) / rowSum on the last slab pass: OS-SART with
    the staged projector equals the per-lane one bit for bit."""
    dd, g, hangs, vol, mask = _setup(dict(XN=64, YN=64, SLN=8), 48)
    with dd.Plan(g, hangs, mask, sat_mode="twofloat") as plan:
        meas = plan.proj_host(vol)
    x0 = np.zeros_like(vol)
    res = {}
    print("\n[pstage] OS-SART 3 iterations x 4 subsets, staged vs per-lane projector", flush=True)
    for st in (0, 1):
        with dd.Plan(g, hangs, mask, sat_mode="twofloat") as plan:
            plan.proj_stage = st
            plan.slab_tile = 16  # ratio on the last of four passes
            res[st] = plan.ossart_host(meas, x0, 3, 4, use_fista=1)
    assert np.array_equal(res[0], res[1])


def test_proj_stage_parameter():
    dd, g, hangs, vol, mask = _setup(dict(XN=32, YN=32, SLN=8), 8)
    from aapm_lowdosectimagerecon_competition_amd import _native as N
    with dd.Plan(g, hangs, mask, sat_mode="twofloat") as plan:
        assert plan.proj_chunk == 8 and plan.proj_stage == 1  # the defaults at chunk 8
        for bad in (4, -1):
            with pytest.raises(N.DDFanError):
                plan.proj_stage = bad
        plan.proj_stage = 0
        assert plan.proj_stage == 0
    with dd.Plan(g, hangs, mask, sat_mode="f32") as plan:
        assert plan.proj_chunk == 8 and plan.proj_stage == 1
