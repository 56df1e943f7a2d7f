"""Zero- and infinite-frequency limits (SYBO = 1, ImplementSubs.f90:45-70) of
the oracle's full chain, pinned by results that do not come from the oracle.

A floating hemisphere of radius a in deep water, added mass normalised by
(2/3) pi rho a^3, the displaced mass of the hemisphere (a whole sphere in
unbounded fluid has added mass (2/3) pi rho a^3, half its displaced mass):
- zero frequency (V = 0, Green function 1/r + 1/r1, rigid free surface): surge
  of the hemisphere is half of the double-body sphere's translation, 0.5
  exactly (reflection symmetry about z = 0); heave 0.8310 (Hulme 1982, J. Fluid Mech. 121);
- infinite frequency (V = -1, Green function 1/r - 1/r1, phi = 0 on z = 0):
  heave is half of the sphere's translation, 0.5 exactly (the translating
  sphere's potential is odd in z); surge 0.2732 (Hulme 1982).
The panel method converges as O(h^2); the values are Richardson-extrapolated
from two meshes (24 x 6 and 48 x 12 panels). A wrong image sign (1/r1 added
instead of subtracted) moves them by tens of percent. Damping and excitation
are zero at zero frequency (W1 = 0 multiplies both, PotentWavForceMulti.f90:
108,162); at infinite frequency in deep water the damping is Re(i rho W1 phi)
with phi real, zero, and the excitation vanishes with exp(1e20 z)
(PatclVelct.f90:97-108).
"""
import numpy as np
import pytest

import meshes
from oracle import Oracle


@pytest.fixture(scope="module")
def limits():
    o = Oracle()
    out = {}
    for nt, nphi in ((24, 6), (48, 12)):
        p = meshes.Problem([meshes.hemisphere(1.0, nt, nphi)], h=-1.0, omegas=[1.0], beta=np.array([0.0, 1.0]))
        norm = 2.0 / 3.0 * np.pi * p.rho
        out[nt] = {w: o.solve_limit(p, w) for w in (0, 1)}
        out[nt]["norm"] = norm
    return out


def _extrapolated(limits, which, i):
    coarse = limits[24][which]["amas"][i, i] / limits[24]["norm"]
    fine = limits[48][which]["amas"][i, i] / limits[48]["norm"]
    return fine + (fine - coarse) / 3.0         # h ratio 2, O(h^2)


@pytest.mark.parametrize("which,mode,value,tol", [
    (0, 0, 0.5, 2e-3),      # zero frequency, surge: exact
    (1, 2, 0.5, 2e-3),      # infinite frequency, heave: exact
    (0, 2, 0.8310, 5e-3),   # zero frequency, heave: Hulme 1982
    (1, 0, 0.2732, 5e-3),   # infinite frequency, surge: Hulme 1982
])
def test_hemisphere_limit_added_mass(limits, which, mode, value, tol):
    got = _extrapolated(limits, which, mode)
    assert abs(got - value) < tol, (which, mode, got, value)


def test_limit_damping_and_excitation_vanish(limits):
    for nt in (24, 48):
        for which in (0, 1):
            r = limits[nt][which]
            assert np.all(np.isfinite(r["amas"]))
            assert np.max(np.abs(r["bdmp"])) == 0.0, (nt, which)
            assert np.max(np.abs(r["exfc"])) == 0.0, (nt, which)
