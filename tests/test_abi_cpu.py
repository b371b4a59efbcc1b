"""CPU checks of the C-ABI library: it loads, exports every declared entry
point, and its host-only numerics (wavelet, parabola) match the oracle.
No call here touches a GPU."""
import os
import re

import numpy as np
import pytest

import fwi2d_amd as F
from fwi2d_amd import abi
import oracle_bind as ob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    names = []
    for hdr in os.listdir(os.path.join(ROOT, "include")):
        if not hdr.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(fwi_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = F.lib()
    names = declared_entry_points()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    bound = {s[0] for s in abi.SIGNATURES + abi.NN_SIGNATURES + abi.UPDATE_SIGNATURES}
    assert set(names) == bound, set(names) ^ bound


@pytest.mark.parametrize("nt,dt,fmax", [(2001, 1e-3, 45.0), (4001, 8e-4, 30.0), (2000, 1e-3, 50.0)])
def test_host_wavelet_matches_oracle(nt, dt, fmax):
    w, tlag = F.ricker_wavelet(nt, dt, fmax)
    wo, tlo = ob.wavelet(nt, dt, fmax)
    assert tlag == tlo
    assert np.abs(w - wo).max() <= 1e-7 * np.abs(wo).max()


def test_host_parabola_step():
    a = np.array([0.0, 0.02, 0.05])
    f = 3.0 * (a - 0.031) ** 2 + 1.0
    assert abs(F.parabola_step(a, f) - 0.031) < 1e-12
    for aa, ff in [(a, [1.0, 2.0, 3.5]),                       # collinear: den = 0
                   ([0.0, 0.0, 0.1], [1.0, 1.0, 1.0])]:        # repeated abscissa
        with pytest.raises(F.FwiError) as e:
            F.parabola_step(aa, ff)
        assert e.value.code == abi.FWI_ERR_NUMERIC


def test_guarded_step_rules():
    a = np.array([0.0, 0.02, 0.05])
    m = np.full(50, 1.0, np.float32)
    g = np.linspace(-1.0, 1.0, 50).astype(np.float32)
    # convex, vertex inside (0, 4 alpha_2]: the vertex itself (rule 0)
    f = 3.0 * (a - 0.031) ** 2 + 1.0
    s, r = F.guarded_step(a, f, m, g)
    assert r == 0 and s == F.parabola_step(a, f)
    # convex but the vertex beyond 4 alpha_2: best trial (rule 1)
    f = 3.0 * (a - 0.5) ** 2
    s, r = F.guarded_step(a, f, m, g)
    assert (s, r) == (0.05, 1)
    # concave (the vertex is a maximum): best trial below f0 (rule 1)
    f = np.array([1.0, 0.9, 0.5])
    s, r = F.guarded_step(a, f, m, g)
    assert (s, r) == (0.05, 1)
    # concave and nothing decreases: no step (rule 2)
    f = np.array([1.0, 1.5, 1.6])
    assert F.guarded_step(a, f, m, g) == (0.0, 2)
    # the vertex would make some m non-positive: fall back
    f = 3.0 * (a - 0.04) ** 2 + 1.0
    m2 = m.copy()
    m2[-1] = 0.03                                   # m - 0.04 * 1 < 0 at the last cell
    s, r = F.guarded_step(a, f, m2, g)
    assert r == 1 and s == 0.02 and np.all(m2 - np.float32(s) * g > 0)
    with pytest.raises(F.FwiError):
        F.guarded_step([0.0, 0.05, 0.02], [1.0, 0.9, 0.8], m, g)   # alphas not increasing
    # a trial that could not be modeled (+inf) is no decrease
    s, r = F.guarded_step(a, [1.0, 0.9, np.inf], m, g)
    assert (s, r) == (0.02, 1)
    assert F.guarded_step(a, [1.0, np.inf, np.inf], m, g) == (0.0, 2)
    # m_min bounds the update from below, not just at zero
    f = 3.0 * (a - 0.03) ** 2 + 1.0
    s, r = F.guarded_step(a, f, m, g, m_min=0.975)    # the vertex would reach 0.97
    assert (s, r) == (0.02, 1)
    assert np.all(m - np.float32(s) * g >= 0.975)


def test_wavelet_rejects_bad_arguments():
    with pytest.raises(F.FwiError) as e:
        F.ricker_wavelet(1, 1e-3, 30.0)
    assert e.value.code == abi.FWI_ERR_ARG
