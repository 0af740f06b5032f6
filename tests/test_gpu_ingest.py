"""Device ingest helpers against numpy: FP32 -> FP16 narrowing (the
reference's launch_narrow_to_half, sam.cu:311-316) bit for bit against
numpy's round-to-nearest-even astype(float16), and the BIP -> BSQ transpose
(launch_transpose_bip_to_bsq, sam.cu:281-309) element for element."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


@pytest.fixture(scope="module")
def H():
    import hsi_mi355x
    return hsi_mi355x


def test_narrow_to_half_matches_numpy(H, torch):
    rng = np.random.default_rng(3)
    n = 8 * 1000 + 5  # not a multiple of the kernel's eight elements per lane
    x = (rng.normal(0, 1, n) * np.exp(rng.uniform(-20, 12, n))).astype(np.float32)
    # ties at FP16 precision, the overflow edge, subnormals, signed zero, inf, NaN
    special = np.array([1 + 2 ** -11, 1 + 3 * 2 ** -11, 65504, 65519.99, 65520, -65520, 2 ** -24,
                        2 ** -25, 3 * 2 ** -26, 1e-9, -0.0, np.inf, -np.inf, np.nan], np.float32)
    x[:special.size] = special
    x[-special.size:] = special
    src = torch.from_numpy(x).cuda()
    dst = torch.full((n,), 7.0, dtype=torch.float16, device="cuda")
    H.narrow_to_half(src, dst)
    torch.cuda.synchronize()
    got = dst.cpu().numpy()
    with np.errstate(over="ignore"):
        want = x.astype(np.float16)
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint16), want[~nan].view(np.uint16))


def test_narrow_to_half_partial_count_and_alignment(H, torch):
    x = torch.arange(64, dtype=torch.float32, device="cuda") + 0.25
    dst = torch.full((64,), -1.0, dtype=torch.float16, device="cuda")
    H.narrow_to_half(x, dst, count=13)
    torch.cuda.synchronize()
    got = dst.cpu().numpy()
    assert np.array_equal(got[:13], (np.arange(13) + 0.25).astype(np.float16))
    assert np.all(got[13:] == -1.0)  # nothing past `count` is written
    with pytest.raises(H.HsiError):
        H.narrow_to_half(x[1:], dst, count=8)  # 4 bytes past a 16-byte boundary


@pytest.mark.parametrize("h,w,bands", [(37, 29, 70), (16, 64, 64), (1, 1, 3), (300, 301, 113)])
def test_transpose_bip_to_bsq_matches_numpy(H, torch, h, w, bands):
    rng = np.random.default_rng(h * 1000 + bands)
    p = h * w
    bip = rng.normal(size=(p, bands)).astype(np.float32)
    stride = H.plane_stride(h, w)
    src = torch.from_numpy(bip).cuda()
    dst = torch.full((bands, stride), -9.0, dtype=torch.float32, device="cuda")
    H.transpose_bip_to_bsq(src, dst, h, w, bands)
    torch.cuda.synchronize()
    got = dst.cpu().numpy()
    assert np.array_equal(got[:, :p], bip.T)
    assert np.all(got[:, p:] == -9.0)  # the plane padding is left alone
