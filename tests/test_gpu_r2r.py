"""GPU r2r transforms vs the oracle's FFTW-semantics transforms (FP64 and FP32)."""
import itertools
import os

import numpy as np
import pytest

import smaiso_amd as S
import smaiso_oracle as O
from helpers import rel_err

pytestmark = pytest.mark.gpu

PLANS = {
    "CT": O.CT, "ICT": O.ICT,
    "STx": O.ST[0], "STy": O.ST[1], "STz": O.ST[2],
    "ISTx": O.IST[0], "ISTy": O.IST[1], "ISTz": O.IST[2],
}

# FFT path (16..512, radix 2/3/4/8/16) and the direct O(N^2) path (others)
SHAPES = [(8, 6, 4), (5, 7, 3), (16, 16, 16), (32, 48, 96), (64, 64, 64), (96, 96, 96),
          (128, 128, 64), (12, 10, 2), (2, 3, 2)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("plan", list(PLANS))
def test_plan_fp64(shape, plan):
    rng = np.random.default_rng(hash((shape, plan)) % 2**32)
    x = rng.standard_normal(shape)
    got = S.r2r3d(x, PLANS[plan], S.FP64)
    ref = O.r2r3d(x, PLANS[plan])
    mx, l2 = rel_err(got, ref)
    assert mx < 1e-12 and l2 < 1e-12, (mx, l2)


@pytest.mark.parametrize("shape", [(64, 64, 64), (96, 96, 96), (128, 128, 64), (5, 7, 3)],
                         ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("plan", list(PLANS))
def test_plan_fp32(shape, plan):
    rng = np.random.default_rng(7)
    x = rng.standard_normal(shape)
    got = S.r2r3d(x, PLANS[plan], S.FP32_MIXED)
    ref = O.r2r3d(x, PLANS[plan])
    mx, l2 = rel_err(got, ref)
    assert mx < 2e-5 and l2 < 2e-6, (mx, l2)


@pytest.mark.parametrize("n", [16, 32, 48, 64, 96, 128, 192, 256, 384, 512, 1024])
@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_every_fft_length_each_axis(n, kind, axis):
    """1-D transform along one axis (identity-free: kinds on the other axes act
    on length-2 extents and are undone by the oracle comparison)."""
    shape = [3, 2, 5]
    shape[axis] = n
    rng = np.random.default_rng(n * 10 + kind)
    x = rng.standard_normal(shape)
    kinds = [kind, kind, kind]
    got = S.r2r3d(x, kinds, S.FP64)
    ref = O.r2r3d(x, kinds)
    mx, l2 = rel_err(got, ref)
    assert mx < 1e-12, (mx, l2)


def test_golden_fixtures():
    """Committed vectors from the FFTW manual's defining sums (tests/golden)."""
    path = os.path.join(os.path.dirname(__file__), "golden", "r2r_golden.npz")
    g = np.load(path)
    for key in g.files:
        if not key.startswith("in_"):
            continue
        name = key[3:]
        plan, shape = name.split("@")
        kinds = PLANS[plan]
        got = S.r2r3d(g[key], kinds, S.FP64)
        mx, l2 = rel_err(got, g["out_" + name])
        assert mx < 1e-12, (name, mx)


def test_round_trip_256():
    """Size-independent property at the bench size: ICT(CT(x)) = 8 N^3 x, and
    the DST pair likewise."""
    n = 256
    rng = np.random.default_rng(3)
    x = rng.standard_normal((n, n, n))
    for fwd, inv in (("CT", "ICT"), ("STx", "ISTx"), ("STz", "ISTz")):
        y = S.r2r3d(S.r2r3d(x, PLANS[fwd]), PLANS[inv]) * (0.125 / n ** 3)
        mx, l2 = rel_err(y, x)
        assert mx < 1e-12, (fwd, mx)


def test_device_api_in_place():
    from hipmem import DeviceArray
    rng = np.random.default_rng(11)
    x = rng.standard_normal((64, 32, 48))
    d = DeviceArray(x)
    S.r2r3d_device(d.ptr, d.ptr, x.shape, O.ST[1], S.FP64)
    mx, _ = rel_err(d.get(), O.r2r3d(x, O.ST[1]))
    assert mx < 1e-12
    xf = x.astype(np.float32)
    a, b = DeviceArray(xf), DeviceArray.empty(x.shape, np.float32)
    S.r2r3d_device(a.ptr, b.ptr, x.shape, O.CT, S.FP32_MIXED)
    mx, _ = rel_err(b.get().astype(np.float64), O.r2r3d(xf.astype(np.float64), O.CT))
    assert mx < 2e-5


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("dtype", [S.FP64, S.FP32_MIXED])
@pytest.mark.parametrize("shape", [(64, 96, 128), (256, 32, 48), (16, 512, 32)])
def test_tile_variants(variant, dtype, shape):
    """Every autotuner tile variant gives the transform (all four kinds on
    the three axes)."""
    from hipmem import DeviceArray
    rng = np.random.default_rng(variant + 10 * dtype)
    x = rng.standard_normal(shape)
    npdt = np.float64 if dtype == S.FP64 else np.float32
    tol = 1e-12 if dtype == S.FP64 else 2e-5
    for kinds in (O.CT, O.ICT, O.ST[0], O.IST[2]):
        xin = x.astype(npdt)
        a, b = DeviceArray(xin), DeviceArray.empty(shape, npdt)
        try:
            S.r2r3d_tiles(a.ptr, b.ptr, shape, kinds, dtype, variant)
        except S.SmaisoError as e:
            assert e.code == S.E_INVALID and "not built" in str(e)
            continue
        mx, _ = rel_err(b.get().astype(np.float64), O.r2r3d(xin.astype(np.float64), kinds))
        assert mx < tol, (kinds, mx)


def test_hbm_copy_rate():
    """smaiso_hbm_copy_rate: a plain copy kernel's HBM rate, and its argument
    checks."""
    import ctypes as C
    L = S.lib()
    L.smaiso_hbm_copy_rate.argtypes = [C.c_int64, C.c_int32, C.POINTER(C.c_double)]
    g = C.c_double()
    assert L.smaiso_hbm_copy_rate(256 << 20, 3, C.byref(g)) == 0
    assert 1000.0 < g.value < 20000.0, g.value
    assert L.smaiso_hbm_copy_rate(1000, 3, C.byref(g)) == S.E_INVALID
    assert L.smaiso_hbm_copy_rate(256 << 20, 0, C.byref(g)) == S.E_INVALID
