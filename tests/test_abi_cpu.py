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
    bound = {s[0] for s in abi.SIGNATURES + abi.NN_SIGNATURES}
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


def test_wavelet_rejects_bad_arguments():
    with pytest.raises(F.FwiError) as e:
        F.ricker_wavelet(1, 1e-3, 30.0)
    assert e.value.code == abi.FWI_ERR_ARG
