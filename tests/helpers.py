"""Shared helpers for the parity tests (tests only)."""
import numpy as np

import smaiso_amd as S
import smaiso_oracle as O


def rel_err(a, b, scale=None, l2=None):
    """(max-abs error, L2 error) relative to the reference field's scale.

    For one component of a vector field pass the vector's scale (max-abs and
    L2 norm over all components): a component that is ~0 by symmetry has no
    scale of its own."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if scale is None:
        scale = np.abs(b).max()
    if l2 is None:
        l2 = np.linalg.norm(b)
    scale = max(scale, 1e-300)
    return np.abs(a - b).max() / scale, np.linalg.norm(a - b) / max(l2, 1e-300)


# components of a vector share its scale; the solenoidal and irrotational
# parts are measured against the whole velocity (either can vanish by symmetry)
VECTOR_GROUPS = [("gpx", "gpy", "gpz"), ("rdx", "rdy", "rdz"),
                 ("vx", "vy", "vz", "vsolx", "vsoly", "vsolz", "virrx", "virry", "virrz")]


def group_scales(ref: dict):
    """Scale (max-abs, L2) per field; vector components share their vector's."""
    out = {k: (np.abs(v).max(), np.linalg.norm(v)) for k, v in ref.items()}
    for g in VECTOR_GROUPS:
        if all(k in ref for k in g):
            mx = max(np.abs(ref[k]).max() for k in g)
            l2 = float(np.sqrt(sum(np.linalg.norm(ref[k]) ** 2 for k in g)))
            for k in g:
                out[k] = (mx, l2)
    return out


def params_pair(**kw):
    """The same parameters for the GPU library and the oracle."""
    sp = S.Params(**kw)
    ok = {k: v for k, v in kw.items() if k not in ("precision",)}
    if "rho_m" in ok and ok["rho_m"] <= 0:
        ok["rho_m"] = None
    if "lam" in ok:
        pass
    op = O.Params(**ok)
    return sp, op


def config1(n=64, steps_dt=0.01, rho_m=1.0, precision=S.FP64, **kw):
    """Config 1: n^3, dx=pi/4, dt=0.01, rho 1.0/0.9 tanh profile, divvSwitch=0."""
    base = dict(nx=n, ny=n, nz=n, dt=steps_dt, rho0=0.9, rho_m=rho_m, divv_switch=0,
                precision=precision)
    base.update(kw)
    return params_pair(**base)
