"""psod_cic_deposit_noalloc (no reference counterpart): the deposit for
callers that cannot allocate, with its tile list in a fixed device array.
Called through the C ABI and compared with psod_cic_deposit on two calls in
a row, the second on a smaller catalogue: the grids must be bit-identical
(both are exact 64-bit fixed-point sums), and the call must not grow the
context's device scratch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _catalogue(ctx, nx, nclump, seed):
    import psod_amd as P
    g = P.gen_params(nx ** 3, nclump, nlo=20, nhi=5000, seed=seed)
    pos, _, lab, _, _ = ctx.generate(g, with_vel=False)
    return ctx.cell_sort(pos, nx, labels=lab)


def _deposit(L, ctx, f, cat, ng, shift):
    import torch
    grid = torch.full((ng, ng, ng), float("nan"), dtype=torch.float32,
                      device="cuda:0")
    torch.cuda.synchronize()
    rc = f(ctx.h, cat["npart"], cat["nx"],
           C.c_void_p(cat["cell_start"].data_ptr()),
           C.c_void_p(cat["rec"].data_ptr()), ng, shift,
           C.c_void_p(grid.data_ptr()))
    assert rc == 0, L.psod_last_error()
    assert L.psod_synchronize(ctx.h) == 0
    return grid.cpu().numpy()


@pytest.mark.parametrize("order", ["noalloc_first", "alloc_first"])
def test_noalloc_deposit_matches_deposit(ctx, order):
    from psod_amd import _lib
    L = _lib.load()
    # a larger catalogue, then a smaller one (and another grid shape, so the
    # second call's tile list is shorter than what the array still holds)
    for nx, ng, shift, seed in ((64, 64, 0, 3), (32, 64, 5, 4)):
        cat = _catalogue(ctx, nx, 60, seed)
        calls = [("noalloc", L.psod_cic_deposit_noalloc),
                 ("alloc", L.psod_cic_deposit)]
        if order == "alloc_first":
            calls.reverse()
        out = {}
        for name, f in calls:
            before = L.psod_scratch_bytes(ctx.h)
            out[name] = _deposit(L, ctx, f, cat, ng, shift)
            if name == "noalloc":
                assert L.psod_scratch_bytes(ctx.h) == before
        assert np.all(np.isfinite(out["noalloc"]))
        np.testing.assert_array_equal(out["noalloc"], out["alloc"])
        # mass conservation: rho/rho_bar averages to 1 over the nodes
        assert abs(out["noalloc"].astype(np.float64).mean() - 1.0) < 1e-6


def test_noalloc_rejects_bad_arguments(ctx):
    from psod_amd import _lib
    L = _lib.load()
    rc = L.psod_cic_deposit_noalloc(ctx.h, 10, 12, None, None, 12, 0, None)
    assert rc == _lib.E_ARG
    assert b"nx=12" in L.psod_last_error()
