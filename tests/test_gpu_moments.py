"""Plain FP64 scene moments against float64 numpy sums, over band counts
that exercise every split of the work: whole 16-band blocks on the matrix
cores, one or two leftover bands on the vector lanes, a zero-padded last
block, and ragged pixel counts over one or many workgroups. Every product of
two floats is exact in double, so the only difference is the order of the
sums: each entry must lie within 1e-12 of the sum of |products| behind it.
(The golden fixtures pin the same moments against the reference at 16, 113
and 224 bands.)"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BANDS = [1, 2, 3, 15, 16, 17, 18, 19, 31, 32, 33, 34, 35, 97, 98, 113, 127, 128, 129, 130, 131,
         224, 225, 226, 239, 241, 242, 255, 256]


@pytest.fixture(scope="module")
def H():
    import torch
    import hsi_mi355x
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return hsi_mi355x


@pytest.fixture(scope="module")
def ctx(H):
    c = H.Context(0)
    yield c
    c.close()


def check(H, ctx, bands, h, w, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(0.3, 0.5, (bands, h * w)).astype(np.float32)
    cube = H.to_device_cube(x.reshape(bands, h, w), h, w)
    ctx.stats_accumulate(cube, h, w, bands)
    got = ctx.moments_read(bands)
    xd = x.astype(np.float64)
    assert got["count"][0] == h * w
    lower = np.tril_indices(bands)
    want_o = (xd @ xd.T)[lower]
    bound_o = (np.abs(xd) @ np.abs(xd).T)[lower] * 1e-12
    err_o = np.abs(got["sum_outer"][lower] - want_o)
    assert np.all(err_o <= bound_o), f"{bands} bands: worst {np.max(err_o / bound_o):.3g} of bound"
    err_s = np.abs(got["sum"] - xd.sum(axis=1))
    assert np.all(err_s <= np.abs(xd).sum(axis=1) * 1e-12), f"{bands} bands: sums"


@pytest.mark.parametrize("bands", BANDS)
def test_moments_match_float64_sums(H, ctx, bands):
    print(f"moments {bands} bands 37x29", flush=True)
    check(H, ctx, bands, 37, 29, 1000 + bands)


@pytest.mark.parametrize("bands,h,w", [(113, 517, 509), (224, 263, 255), (18, 1024, 1021)])
def test_moments_over_many_workgroups(H, ctx, bands, h, w):
    print(f"moments {bands} bands {h}x{w}", flush=True)
    check(H, ctx, bands, h, w, 2000 + bands)
