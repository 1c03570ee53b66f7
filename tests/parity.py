"""Checks of the HIP pass against the FP64 oracle, shared by the GPU tests.

Bars: integer work (counts, profiles, centre choice) bit-exact except where
the oracle flags a near-tie (a decision within TOL_TIE relative of flipping
under the float32 rounding of the split-precision pass); float results
within the tolerances written below.

Full-size catalogues are checked through sub-boxes: the particles whose
cells lie in [lo, lo + sub) per axis, with lo + sub <= nx (no wrap), moved
to the sub-box frame by x*nx - lo.  For a power-of-two nx that subtraction
is exact in float32, so every FP64 difference of positions the oracle forms
in the sub-box equals the one it would form in the full box, and a halo whose
reach stays inside the sub-box gets the oracle's full-box answer.
"""
import numpy as np

TOL_TIE = 1e-5


def host(t):
    return t.cpu().numpy()


def check_sod(o, r, nd, min_same=0.99, nx=None):
    """Batched outputs o against oracle outputs r (same halos, same seeds).
    Returns the mask of halos with the same centre.  nx (cells per side)
    bounds the centre-of-mass shift of a boundary swap (below)."""
    assert np.all(o["status"] & 1), "every halo closes"
    same = o["centre"] == r["centre"]
    robust = r["robust_centre"] == 1
    assert np.all(same | ~robust), np.where(~same & robust)
    assert same.mean() >= min_same, same.mean()
    np.testing.assert_array_equal(o["centre_id"][same], r["centre_id"][same])
    for d in range(nd):
        cnt = o["count"][:, d] == r["count"][:, d]
        rob = r["margin"][:, d] > TOL_TIE
        assert np.all(cnt | ~same | ~rob), np.where(same & rob & ~cnt)
        # discrete R is the FP64 distance of the boundary particle, the
        # (count)-th closest: bit-exact unless another particle lies within
        # TOL_TIE of it (oracle rgap), when the split-float32 ranking may
        # take that one; R then keeps the 1e-6 relative bar (float32
        # distances are good to about 1e-6 cells)
        ok = same & cnt
        rg = r["rgap"][:, d] > TOL_TIE
        np.testing.assert_allclose(o["radius"][ok & rg, d],
                                   r["radius"][ok & rg, d], rtol=1e-12)
        np.testing.assert_allclose(o["radius"][ok & ~rg, d],
                                   r["radius"][ok & ~rg, d], rtol=1e-6)
        # interpolated R and M are continuous across near-ties
        np.testing.assert_allclose(o["rint"][same, d], r["rint"][same, d],
                                   rtol=1e-6)
        np.testing.assert_allclose(o["mint"][same, d], r["mint"][same, d],
                                   rtol=3e-6)
    ok = same & (o["count"][:, 0] == r["count"][:, 0])
    c2 = o["count2"] == r["count2"]
    assert np.all(c2 | ~ok | ~(r["margin2"] > TOL_TIE))
    pm = (o["profile"] == r["profile"]).all(axis=1)
    assert np.all(pm | ~ok | ~(r["profile_margin"] > TOL_TIE))
    np.testing.assert_allclose(o["com_vel"][ok], r["com_vel"][ok], rtol=1e-9,
                               atol=1e-9)
    dp = np.abs(o["com_pos"][ok] - r["com_pos"][ok])
    dp = np.minimum(dp, 1 - dp)
    # the same particle set: the same centre of mass (sums of exact float
    # offsets); a swap across the boundary (rgap below TOL_TIE) moves it by
    # at most 2R/n cells
    rg0 = r["rgap"][ok, 0] > TOL_TIE
    assert np.all(dp[rg0] < 1e-12), dp[rg0].max()
    if nx is not None:
        lim = 2.0 * r["radius"][ok, 0] / np.maximum(r["count"][ok, 0], 1) / nx
        assert np.all(dp[~rg0] <= lim[~rg0][:, None] + 1e-12)
    # acceptance (mass and centre window, main file:2118-2120)
    acc = (o["status"] & 2) != 0
    np.testing.assert_array_equal(acc[ok], r["accepted"][ok] == 1)
    # the boundary listing: particles within tol_abs of R[0].  The batched
    # pass ranks by split float32 distances, so a particle this close to R
    # may fall on either side; the count of those listed can differ by the
    # ones whose float32 and FP64 distances straddle tol_abs, a handful.
    nb_o, nb_r = o["nboundary"][ok], r["nboundary"][ok]
    assert np.all(np.abs(nb_o - nb_r) <= 2 + nb_r // 8), (nb_o, nb_r)
    return same


def subbox(pos, nx, lo, sub, vel=None):
    """Particles of the sub-box [lo, lo + sub) (cells, per axis) of a
    device catalogue: (positions in sub-box units, float32; their full-box
    indices, ascending; velocities or None)."""
    import torch
    assert all(0 <= l and l + sub <= nx for l in lo), (lo, sub, nx)
    assert sub & (sub - 1) == 0, "sub-box side must be a power of two"
    inside = None
    for d in range(3):
        xs = pos[:, d] * nx
        xs = torch.where(xs >= nx, xs - nx, xs)  # x == 1.0 is cell 0
        c = torch.floor(xs)
        m = (c >= lo[d]) & (c < lo[d] + sub)
        inside = m if inside is None else inside & m
        del xs, c
    idx = torch.nonzero(inside).squeeze(1)
    sp = pos[idx] * nx
    sp = torch.where(sp >= nx, sp - nx, sp)
    sp = (sp - torch.tensor(lo, dtype=torch.float32, device=pos.device)) / sub
    sv = host(vel[idx]) if vel is not None else None
    return host(sp).astype(np.float32), host(idx).astype(np.int64), sv


def oracle_subbox(orc, spos, sidx, sub, seeds, nmem, mpart, sod_kw, vel=None,
                  centres=None, tol_abs=None):
    """The oracle's FoF-mode SOD of the halos with full-box seeds `seeds`
    (all inside the sub-box) on the sub-box particles, with the full box's
    particle mass.  Results are translated to full-box indices (centre,
    centre_id = the generator's id = index); com_pos stays in sub-box
    units (see com_to_full).  tol_abs is the full box's boundary listing
    width in cells (the pass's default is 1e-7 nx of the full box)."""
    pos_in = np.searchsorted(sidx, seeds)
    assert np.all(sidx[pos_in] == seeds)
    cen_in = None
    if centres is not None:
        cen_in = np.searchsorted(sidx, centres)
        assert np.all(sidx[cen_in] == centres)
    cs, perm = orc.cell_sort(spos, sub, 8)
    fp = orc.fof_params(sub, len(spos), **sod_kw)
    fp.mpart = mpart
    if tol_abs is not None:  # the full box's boundary width, in cells
        fp.tol_abs = tol_abs
    r = orc.sod_fof(spos, cs, perm, 8, pos_in, nmem, fp, vel=vel, ids=sidx,
                    centres=cen_in)
    r["centre_sub"] = r["centre"].copy()
    r["centre"] = sidx[r["centre"]]
    return r


def com_to_full(com_sub, lo, sub, nx):
    """Centre of mass from sub-box units to full-box units."""
    return ((com_sub * sub + np.asarray(lo, dtype=np.float64)) / nx) % 1.0


def reach_inside(r, spos, sub):
    """Halos whose final reach (plus the centre cell) stays inside the
    sub-box, so that the oracle's periodic wrap of the sub-box never
    brought in particles from its far side."""
    cc = np.floor(spos[r["centre_sub"]] * sub)
    lo_room = cc.min(axis=1)
    hi_room = (sub - 1 - cc).min(axis=1)
    w = r["w"]
    return (w + 1 <= lo_room) & (w + 1 <= hi_room)


def s_fp64(spos, sub, cand, rs=1.0, eps=1.0 / 64.0):
    """S(c) = sum over particles within rs of 1/sqrt(d2 + eps^2) in FP64
    (main file:1837-1858) for sub-box particle indices `cand`, by brute
    force over the particles of the cube of half-width rs around each."""
    xs = spos.astype(np.float64) * sub
    out = []
    for c in cand:
        d = xs - xs[c]
        near = np.all(np.abs(d) <= rs, axis=1)
        dd = d[near]
        d2 = (dd * dd).sum(axis=1)
        out.append(np.sum(1.0 / np.sqrt(d2[d2 <= rs * rs] + eps * eps)))
    return np.array(out)
