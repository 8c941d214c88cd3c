"""The kernels' table fetch against the clamp-to-edge texel fetch.

The oracle samples a summed-area table like the reference's texture fetch
(MultiSlicesProjGPU.cu:236-315): texel i = floor(c - 0.5) limited to [-1, N],
weight t = (c - 0.5) - i, and t = 0 where the clamp reads one texel twice.
The kernels (fetch_pos in ddfan_kernels.hip) instead clamp the texel into
[0, N] and the weight into [0, 1], and read one paired entry (S[i0], v[i0])
with v[N] = 0.  This pins, in float32 on the CPU, that both give the same
entry and the same weight wherever the weight matters (i0 < N), for
coordinates below, inside, at and beyond the table, NaN and infinities.
"""
import numpy as np

F32 = np.float32


def texel_fetch(c, N):
    """Clamp-to-edge fetch of the oracle: (entry, weight)."""
    cm = F32(c) - F32(0.5)
    f = np.floor(cm)
    f = np.where(f >= F32(-1.0), f, F32(-1.0))  # NaN -> -1
    f = np.where(f > F32(N), F32(N), f)
    i = f.astype(np.int64)
    i0 = np.clip(i, 0, N)
    t = np.where((i < 0) | (i >= N), F32(0.0), cm - i.astype(F32))
    return i0, t.astype(F32)


def kernel_fetch(c, N):
    """fetch_pos: i0 = clamp(floor(c - 0.5), 0, N), t = clamp(c - 0.5 - i0, 0, 1)."""
    cm = F32(c) - F32(0.5)
    fc = np.fmin(np.fmax(np.floor(cm), F32(0.0)), F32(N))  # fmax(NaN, 0) = 0
    i0 = fc.astype(np.int64)
    t = np.fmin(np.fmax(cm - fc, F32(0.0)), F32(1.0))
    return i0, t.astype(F32)


def test_fetch_matches_clamp_to_edge_texel():
    N = 736
    rng = np.random.default_rng(11)
    c = np.concatenate([
        rng.uniform(-50.0, N + 50.0, 200000),
        np.arange(-3.0, N + 4.0, 0.25),                      # texel centres and edges
        np.nextafter(np.arange(-2, N + 3, dtype=F32) + F32(0.5), F32(-np.inf)),
        np.nextafter(np.arange(-2, N + 3, dtype=F32) + F32(0.5), F32(np.inf)),
        [0.0, 0.5, 1.0, N + 0.5, N + 1.0, 1e30, -1e30, np.inf, -np.inf, np.nan],
    ]).astype(F32)
    with np.errstate(invalid="ignore"):
        i_ref, t_ref = texel_fetch(c, N)
        i_k, t_k = kernel_fetch(c, N)
    assert np.array_equal(i_ref, i_k)
    assert np.all((t_k >= 0) & (t_k <= 1))          # finite: t * v[N] = 0 exactly
    live = i_k < N                                  # v[i0] != 0 only there
    assert np.array_equal(t_ref[live], t_k[live])
    # entries and sampled values agree on a random table with v[N] = 0
    S = np.cumsum(rng.uniform(0, 0.03, N + 1)).astype(F32)
    v = np.append(np.diff(S), F32(0.0)).astype(F32)
    T_ref = (S[i_ref].astype(np.float64) + t_ref.astype(np.float64) * v[i_ref]).astype(F32)
    T_k = (S[i_k].astype(np.float64) + t_k.astype(np.float64) * v[i_k]).astype(F32)
    assert np.array_equal(T_ref, T_k)
