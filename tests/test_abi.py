"""C ABI checks that need no GPU: exports, parameter validation, host IC."""
import ctypes as C
import math
import subprocess

import numpy as np
import pytest

import smaiso_amd as S
import smaiso_oracle as O


def test_library_exports_every_declared_symbol():
    L = S.lib()
    declared = S.header_symbols()
    assert len(declared) >= 19
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", S.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(declared) <= exported, set(declared) - exported


def test_abi_version_and_defaults():
    L = S.lib()
    assert L.smaiso_abi_version() == 1
    cp = S.CParams()
    L.smaiso_params_default(C.byref(cp))
    # the reference's constants, smaiso-wcomp.cpp:206-283
    assert (cp.nx, cp.ny, cp.nz) == (512, 512, 512)
    assert cp.dt == 0.001 and cp.divv_switch == 50
    assert math.isclose(cp.dx, 2 * math.pi / 8, rel_tol=0, abs_tol=0)
    assert cp.epsilon == -0.005 and cp.rho0 == 0.01
    assert (cp.alpha, cp.beta, cp.gamma, cp.q0, cp.nu, cp.lambda_, cp.zeta) == (1, 2, 1, 1, 1, 1, 1)


@pytest.mark.parametrize("bad,msg", [
    (dict(nx=1), "extent"), (dict(nz=0), "extent"), (dict(dt=0.0), "dt"), (dict(dt=-1e-3), "dt"),
    (dict(nu=0.0), "nu"), (dict(nu=-1.0), "nu"), (dict(lam=-3.0, nu=1.0), "lambda"),
    (dict(dx=0.0), "dx"), (dict(precision=7), "precision"), (dict(dt=float("nan")), "non-finite"),
    (dict(epsilon=-100.0), "Amp"), (dict(divv_switch=-1), "divv_switch"), (dict(slab_n0=-2), "slab"),
])
def test_create_rejects_bad_parameters(bad, msg):
    p = S.Params(nx=8, ny=8, nz=8)
    for k, v in bad.items():
        setattr(p, k, v)
    rc, text = S.create_error(p)
    assert rc == S.E_INVALID, (rc, text)
    assert msg in text


def test_r2r_rejects_bad_arguments():
    x = np.zeros((4, 4, 4))
    with pytest.raises(S.SmaisoError) as e:
        S.r2r3d(x, [0, 5, 0])
    assert e.value.code == S.E_INVALID
    with pytest.raises(S.SmaisoError) as e:
        S.r2r3d(x, [0, 0, 0], dtype=3)
    assert e.value.code == S.E_INVALID


@pytest.mark.parametrize("und,seed", [(0.0, 1), (0.5, 2), (1.0, 77)])
def test_interface_ic_matches_oracle(und, seed):
    p = S.Params(nx=12, ny=10, nz=32, rho0=0.8)
    ic = S.InterfaceIC(rho_iso=0.8, undulation=und, seed=seed)
    psi, rho = S.make_interface_ic(p, ic)
    op = O.Params(nx=12, ny=10, nz=32, rho0=0.8)
    psi2, rho2 = O.make_interface_ic(op, O.InterfaceIC(rho_iso=0.8, undulation=und, seed=seed))
    assert np.abs(psi - psi2).max() < 1e-14
    assert np.abs(rho - rho2).max() < 1e-14


def test_interface_ic_explicit_height():
    p = S.Params(nx=6, ny=5, nz=16)
    h = np.random.default_rng(0).uniform(4, 8, (6, 5))
    psi, rho = S.make_interface_ic(p, S.InterfaceIC(noise=0.0), height=h)
    op = O.Params(nx=6, ny=5, nz=16)
    psi2, rho2 = O.make_interface_ic(op, O.InterfaceIC(noise=0.0, height=h))
    assert np.abs(psi - psi2).max() < 1e-14 and np.abs(rho - rho2).max() < 1e-14


def test_splitmix64_known_answers():
    """First outputs of splitmix64 for seed 0 (Vigna's reference generator):
    0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F."""
    u = O.splitmix64_uniform(0, 3)
    want = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert np.array_equal(u, np.array([(w >> 11) * 2.0 ** -53 for w in want]))
    p = S.Params(nx=1 + 1, ny=2, nz=2)
    # the library's generator on the same stream: psi = noise*(2u-1) with amplitude 0
    psi, _ = S.make_interface_ic(S.Params(nx=2, ny=2, nz=2), S.InterfaceIC(amplitude=0.0, noise=1.0, seed=0))
    u8 = O.splitmix64_uniform(0, 8)
    assert np.array_equal(psi.ravel(), 2.0 * u8 - 1.0)


def test_ensemble_create_rejects_bad_members():
    good = S.Params(nx=8, ny=8, nz=8)
    bad = S.Params(nx=8, ny=8, nz=8, nu=-1.0)
    with pytest.raises(S.SmaisoError) as e:
        S.Ensemble([good, bad])
    assert e.value.code == S.E_INVALID
    assert "member 1" in str(e.value) and "nu" in str(e.value)
    with pytest.raises(S.SmaisoError) as e:
        S.Ensemble([])
    assert e.value.code == S.E_INVALID
    L = S.lib()
    assert L.smaiso_ensemble_step_count(None, 0) == -1
    assert L.smaiso_ensemble_groups(None) == -1
    assert L.smaiso_set_graph(None, 1) == S.E_INVALID
    assert L.smaiso_ensemble_step(None, 1) == S.E_INVALID


@pytest.mark.parametrize("explicit", [False, True])
def test_interface_ic_vjp_matches_oracle(explicit):
    import smaiso_adjoint as A
    p = S.Params(nx=6, ny=5, nz=12)
    op = O.Params(nx=6, ny=5, nz=12)
    rng = np.random.default_rng(3)
    dpsi, drho = rng.standard_normal(p.shape), rng.standard_normal(p.shape)
    ic = S.InterfaceIC(amplitude=0.8, width=1.5, rho_iso=0.85, rho_sm=1.0, undulation=0.7, noise=0.0)
    x = np.arange(6) * op.dx
    y = np.arange(5) * op.dy
    h = op.nz * op.dz / 2 + 0.7 * np.cos(2 * np.pi * x / (6 * op.dx))[:, None] * np.cos(
        2 * np.pi * y / (5 * op.dy))[None, :]
    if explicit:
        h = h + 0.3 * rng.standard_normal(h.shape)
    got = S.interface_ic_vjp(p, ic, dpsi, drho, height=h if explicit else None)
    oic = O.InterfaceIC(amplitude=0.8, width=1.5, rho_iso=0.85, rho_sm=1.0, noise=0.0, height=h)
    want = A.interface_ic_vjp(op, oic, h, dpsi, drho)
    assert np.allclose(got, want, rtol=1e-13, atol=1e-13)


def test_read_psi_text_rank_files(tmp_path):
    """The reference's restart files: each rank's slab, one value per line
    (written here with C++ ostream-like %g formatting), read back in rank order."""
    rng = np.random.default_rng(4)
    psi = rng.standard_normal((4, 3, 5))
    slabs = [psi[:2], psi[2:]]
    paths = []
    for r, sl in enumerate(slabs):
        p = tmp_path / f"psi{r}.dat"
        p.write_text("".join("%g\n" % v for v in sl.reshape(-1)))
        paths.append(p)
    got = S.read_psi_text(paths, psi.size)
    want = np.array([float("%g" % v) for v in psi.reshape(-1)])
    assert np.array_equal(got, want)
    with pytest.raises(S.SmaisoError) as e:
        S.read_psi_text(paths[:1], psi.size)
    assert e.value.code == S.E_INVALID and "values" in str(e.value)
    with pytest.raises(S.SmaisoError) as e:
        S.read_psi_text([tmp_path / "missing.dat"], 1)
    assert e.value.code == S.E_INVALID
