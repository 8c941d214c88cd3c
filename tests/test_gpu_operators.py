"""GPU parity of the DD operators against the FP64 restatement.

Tolerances (relative L2 against the FP64 oracle, FP32 device arithmetic):
  projector      f32 SAT <= 2e-5, two-float / f64 SAT <= 5e-6
  backprojector  f32 SAT <= 2e-5, two-float / f64 SAT <= 5e-6
Integer texel indices of both operators are compared bit-exactly.
"""
import numpy as np
import pytest

from conftest import rel_l2
from helpers import per_slice_angles, small_conf

pytestmark = pytest.mark.gpu

TOL = {"f32": 2e-5, "twofloat": 5e-6, "f64": 5e-6}


def _setup(SLN=8, PN=64, angles="uniform", XN=64, YN=64, seed=11):
    import aapm_lowdosectimagerecon_competition_amd as dd
    from aapm_lowdosectimagerecon_competition_amd import phantoms
    conf = small_conf(XN=XN, YN=YN, SLN=SLN)
    g = dd.detector_geometry(conf)
    hangs = phantoms.uniform_angles(PN, SLN) if angles == "uniform" else per_slice_angles(PN, SLN)
    vol = phantoms.ellipsoid_phantom(XN, YN, SLN, conf.recon.dx, seed=seed)
    mask = dd.recon_mask(XN, YN)
    return dd, g, hangs, vol, mask


@pytest.mark.parametrize("angles", ["uniform", "perslice"])
@pytest.mark.parametrize("sat", ["f32", "twofloat", "f64"])
def test_projector_matches_oracle(oracle, sat, angles):
    dd, g, hangs, vol, mask = _setup(angles=angles)
    print(f"\n[proj] sat={sat} angles={angles} vol={vol.shape} PN={hangs.shape[0]}", flush=True)
    with dd.Plan(g, hangs, mask, sat_mode=sat) as plan:
        assert plan.slice_chunk == (8 if angles == "uniform" else 1)
        got = plan.proj_host(vol)
    ref = oracle.proj(g.as_dict(), hangs, vol, mask)
    e = rel_l2(got, ref)
    print(f"[proj] sat={sat} angles={angles} rel-L2 vs FP64 oracle = {e:.3e}", flush=True)
    assert e <= TOL[sat]


@pytest.mark.parametrize("angles", ["uniform", "perslice"])
@pytest.mark.parametrize("sat", ["f32", "twofloat", "f64"])
def test_backprojector_matches_oracle(oracle, sat, angles):
    dd, g, hangs, vol, mask = _setup(angles=angles)
    prj = oracle.proj(g.as_dict(), hangs, vol, mask).astype(np.float32)
    print(f"\n[back] sat={sat} angles={angles} prj={prj.shape}", flush=True)
    with dd.Plan(g, hangs, mask, sat_mode=sat) as plan:
        got = plan.back_host(prj)
    ref = oracle.back(g.as_dict(), hangs, prj, mask, g.XN, g.YN)
    e = rel_l2(got, ref)
    print(f"[back] sat={sat} angles={angles} rel-L2 vs FP64 oracle = {e:.3e}", flush=True)
    assert e <= TOL[sat]
    assert np.all(got[mask == 0] == 0.0)


@pytest.mark.parametrize("angles", ["uniform", "perslice"])
def test_projector_texel_indices_bit_exact(oracle, angles):
    dd, g, hangs, vol, mask = _setup(SLN=2, PN=24, angles=angles, XN=40, YN=36)
    print(f"\n[proj-trace] angles={angles}", flush=True)
    with dd.Plan(g, hangs, mask) as plan:
        got = plan.trace_proj()
    _, ref = oracle.proj(g.as_dict(), hangs, vol, mask, trace=True)
    assert got.shape == ref.shape
    ndiff = int(np.count_nonzero(got != ref))
    print(f"[proj-trace] {got.size} indices, {ndiff} differ", flush=True)
    assert ndiff == 0


@pytest.mark.parametrize("angles", ["uniform", "perslice"])
def test_backprojector_detector_indices_bit_exact(oracle, angles):
    dd, g, hangs, vol, mask = _setup(SLN=2, PN=24, angles=angles, XN=40, YN=36)
    print(f"\n[back-trace] angles={angles}", flush=True)
    prj = np.zeros((hangs.shape[0], g.DNU, g.SLN), np.float32)
    with dd.Plan(g, hangs, mask) as plan:
        got = plan.trace_back()
    _, ref = oracle.back(g.as_dict(), hangs, prj, mask, g.XN, g.YN, trace=True)
    m = mask.astype(bool)
    ndiff = int(np.count_nonzero(got[m] != ref[m]))
    print(f"[back-trace] {got[m].size} indices, {ndiff} differ", flush=True)
    assert ndiff == 0


@pytest.mark.parametrize("dx,mid", [(4.0, 1), (14.0, 0)])
def test_backprojector_footprint_range_paths(oracle, dx, mid):
    """With 4 mm voxels the 64x64 grid lies well inside the source circle and
    the footprint's short correctly rounded sequences run unchecked; with
    14 mm voxels the grid's corners reach past the source (sqrt(2) x 441 mm >
    595 mm), so every wave checks its lanes' operands.  Both give the oracle's
    detector indices bit for bit and its backprojection within tolerance, in
    the staged and the per-lane-gather kernel."""
    import aapm_lowdosectimagerecon_competition_amd as dd
    from aapm_lowdosectimagerecon_competition_amd import phantoms
    conf = small_conf(XN=64, YN=64, SLN=2, dx=dx)
    g = dd.detector_geometry(conf)
    hangs = phantoms.uniform_angles(24, 2)
    vol = phantoms.ellipsoid_phantom(64, 64, 2, dx, seed=4)
    mask = dd.recon_mask(64, 64)
    prj = oracle.proj(g.as_dict(), hangs, vol, mask).astype(np.float32)
    ref = oracle.back(g.as_dict(), hangs, prj, mask, g.XN, g.YN)
    _, ref_tr = oracle.back(g.as_dict(), hangs, np.zeros_like(prj), mask, g.XN, g.YN, trace=True)
    m = mask.astype(bool)
    with dd.Plan(g, hangs, mask, sat_mode="twofloat") as plan:
        assert plan.back_mid_range == mid
        got_tr = plan.trace_back()
        ndiff = int(np.count_nonzero(got_tr[m] != ref_tr[m]))
        outs = []
        for stage in (plan.back_stage, 0):
            plan.back_stage = stage
            outs.append(plan.back_host(prj))
    errs = [rel_l2(o, ref) for o in outs]
    print(f"\n[back-range] dx={dx} mid={mid}: {ndiff} of {got_tr[m].size} indices differ, "
          f"rel-L2 staged {errs[0]:.3e} gather {errs[1]:.3e}", flush=True)
    assert ndiff == 0
    assert max(errs) <= TOL["twofloat"]
    assert np.array_equal(outs[0], outs[1])


def test_view_subsets_match_full_call():
    dd, g, hangs, vol, mask = _setup(PN=48)
    print("\n[subset] strided subsets vs full call", flush=True)
    with dd.Plan(g, hangs, mask) as plan:
        full = plan.proj_host(vol)
        for k in range(3):
            sub = plan.proj_host(vol, first=k, stride=3, count=16)
            assert np.array_equal(sub, full[k::3])
        fb = plan.back_host(full[1::4], first=1, stride=4, count=12)
    sep = dd.Plan(g, np.ascontiguousarray(hangs[1::4]), mask)
    ref = sep.back_host(np.ascontiguousarray(full[1::4]))
    sep.close()
    assert np.array_equal(fb, ref)


def test_dropin_entry_points_match_plan():
    import aapm_lowdosectimagerecon_competition_amd as dd
    from aapm_lowdosectimagerecon_competition_amd import phantoms
    conf = small_conf(SLN=4)
    g = dd.detector_geometry(conf)
    PN = 32
    hangs = phantoms.uniform_angles(PN, 4)
    vol = phantoms.ellipsoid_phantom(64, 64, 4, 4.0, seed=4)
    mask = dd.recon_mask(64, 64)
    print("\n[drop-in] dd2_multi_slices Proj/Back vs plan", flush=True)
    # MATLAB-shaped column-major views of the same buffers
    vol_m = np.asfortranarray(vol.T)          # (SLN, XN, YN)
    hangs_m = np.asfortranarray(hangs.T)      # (SLN, PN)
    mask_m = np.asfortranarray(mask.T)        # (XN, YN)
    p_m = dd.dd2_multi_slices("Proj", conf, vol_m, hangs_m, None, mask_m)
    b_m = dd.dd2_multi_slices("Back", conf, p_m, hangs_m, list(range(1, PN + 1, 2)), mask_m)
    with dd.Plan(g, hangs, mask) as plan:
        p = plan.proj_host(vol)
        b = plan.back_host(np.ascontiguousarray(p[0::2]), first=0, stride=2, count=PN // 2)
    assert np.array_equal(p_m.T, p)
    assert np.array_equal(b_m.T, b)


@pytest.mark.parametrize("angles", ["uniform", "perslice"])
def test_projector_slab_passes_match_single_pass(angles):
    """Splitting the slab loop into passes (ragged last pass) changes only the
    summation grouping: results agree to FP32 rounding, OS-SART ratios too."""
    dd, g, hangs, vol, mask = _setup(angles=angles)
    print(f"\n[proj-passes] angles={angles}", flush=True)
    with dd.Plan(g, hangs, mask, sat_mode="twofloat") as plan:
        plan.slab_tile = 0
        one = plan.proj_host(vol)
        plan.slab_tile = 7
        assert plan.slab_tile == 7
        many = plan.proj_host(vol)
    assert rel_l2(many, one) < 1e-6


@pytest.mark.parametrize("sat", ["f32", "twofloat"])
def test_projector_views_per_wave_bit_identical(sat):
    """Grouping 2 or 4 views per wave only shares table rows through L1; each
    view's arithmetic is unchanged, so results are bit-identical (including a
    view count that is not a multiple of the group and a strided subset)."""
    dd, g, hangs, vol, mask = _setup(PN=45)
    print(f"\n[proj-vpw] sat={sat}", flush=True)
    with dd.Plan(g, hangs, mask, sat_mode=sat) as plan:
        outs = {}
        for v in (1, 2, 4):
            plan.views_per_wave = v
            outs[v] = (plan.proj_host(vol), plan.proj_host(vol, first=2, stride=3, count=14))
    for v in (2, 4):
        assert np.array_equal(outs[v][0], outs[1][0])
        assert np.array_equal(outs[v][1], outs[1][1])
