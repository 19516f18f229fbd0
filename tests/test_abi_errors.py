"""Error returns of gl_plan_create for malformed descriptors.  Every case here
is rejected by the host-side validation before any device call, so the tests
run without a GPU: the return code, the thread-local message, and a null plan."""
import ctypes
import os

import pytest

from gigalens_amd import _native as nat

pytestmark = pytest.mark.skipif(not os.path.exists(nat.LIB_PATH), reason="library not built")


def _comp(kind, role, ncol, first=0):
    c = nat.ComponentDesc()
    c.kind, c.role = kind, role
    for j in range(nat.GL_MAX_COMPONENT_PARAMS):
        c.param_col[j] = first + j if j < ncol else -1
        c.param_const[j] = 0.0
    c.niter = 50 if kind == nat.GL_EPL else 0
    return c


def _desc(comps, n_params=None, N=60, ss=2):
    arr = (nat.ComponentDesc * len(comps))(*comps)
    d = nat.PlanDesc()
    d.num_pix, d.supersample = N, ss
    d.n_components = len(comps)
    d.components = arr
    d.n_params = n_params if n_params is not None else 22
    d.background_rms, d.exp_time = 0.2, 100.0
    d._keep = arr
    return d


def _create(desc):
    L = nat.lib()
    out = ctypes.c_void_p(12345)
    rc = L.gl_plan_create(ctypes.byref(desc) if desc is not None else None, ctypes.byref(out))
    msg = L.gl_last_error().decode()
    return rc, out.value, msg


def _default_comps():
    return [_comp(nat.GL_EPL, nat.GL_ROLE_LENS, 6, 0), _comp(nat.GL_SHEAR, nat.GL_ROLE_LENS, 2, 6),
            _comp(nat.GL_SERSIC_ELLIPSE, nat.GL_ROLE_LENS_LIGHT, 7, 8),
            _comp(nat.GL_SERSIC_ELLIPSE, nat.GL_ROLE_SOURCE, 7, 15)]


@pytest.mark.parametrize("case,expect", [
    ("null", "null argument"),
    ("num_pix", "num_pix and supersample must be positive"),
    ("supersample", "num_pix and supersample must be positive"),
    ("n_params", "n_params must be in"),
    ("no_components", "n_components must be in"),
    ("too_many_components", "n_components must be in"),
    ("role_order", "components must be ordered"),
    ("light_as_lens", "light profile given as a lens"),
    ("mass_as_light", "mass profile given as a light component"),
    ("param_col", "param_col out of range"),
    ("unknown_kind", "profile kind 99"),
    ("epl_niter", "EPL niter must be in"),
])
def test_plan_create_rejects(case, expect):
    comps = _default_comps()
    n_params, N, ss = 22, 60, 2
    if case == "num_pix":
        N = 0
    elif case == "supersample":
        ss = -1
    elif case == "n_params":
        n_params = nat.GL_MAX_PARAMS + 1
    elif case == "no_components":
        comps = []
    elif case == "too_many_components":
        comps = comps + [_comp(nat.GL_SERSIC_ELLIPSE, nat.GL_ROLE_SOURCE, 7, 15)] * 5
    elif case == "role_order":
        comps = [comps[2], comps[0], comps[1], comps[3]]
    elif case == "light_as_lens":
        comps[1] = _comp(nat.GL_SERSIC, nat.GL_ROLE_LENS, 5, 6)
    elif case == "mass_as_light":
        comps[3] = _comp(nat.GL_SIS, nat.GL_ROLE_SOURCE, 3, 15)
    elif case == "param_col":
        comps[3] = _comp(nat.GL_SERSIC_ELLIPSE, nat.GL_ROLE_SOURCE, 7, 20)  # columns 20..26 of 22
    elif case == "unknown_kind":
        comps[3] = _comp(99, nat.GL_ROLE_SOURCE, 3, 15)
    elif case == "epl_niter":
        comps[0].niter = nat.GL_MAX_EPL_ITER + 1
    if case == "null":
        rc, out, msg = _create(None)
    elif case == "no_components":
        d = nat.PlanDesc()
        d.num_pix, d.supersample, d.n_components, d.n_params = N, ss, 0, n_params
        rc, out, msg = _create(d)
    else:
        rc, out, msg = _create(_desc(comps, n_params, N, ss))
    assert rc != 0, case
    assert not out, "no plan is returned on error"
    assert expect in msg, (case, msg)
