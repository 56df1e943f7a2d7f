"""Field-point potential, pressure and elevation on the GPU (hm_field_points:
CalPotentialMulti, CalPressureMulti, CalElevationMulti,
PressureElevationMulti.f90:55-348) against the oracle's restatement
(hmo_field_points) on the same problem and frequency, after the plain FP64
solve.

Tolerance: infinite depth 1e-12 of each column's largest value. Finite depth
2e-5: the points' pairs with the hull fall in FINGREEN3D's series regimes, whose
series and Wynn tables stop on absolute 1e-6 tests (FinGreen3D.f90:341, :787)
and move by up to ~1e-6 under last-bit changes of the inputs (the element
kernels are held to 2e-6 absolute for the same reason, tests/test_gpu_kernels.py);
summed over the hull's 192 panels that is what is seen here (4e-6). The on-hull
check pins the free term: at a panel centroid the
representation reproduces the solved potential of that panel (it is the same
collocation equation); with the fluid's 4 pi instead of 2 pi it would be half."""
import numpy as np
import pytest

import hm_bem
import meshes

pytestmark = pytest.mark.gpu


def _problem(h):
    bs = [meshes.cylinder(3.0, 4.0, 16, 4, 2, center=(12.0 * i, 0.0)) for i in range(2)]
    return meshes.Problem(bs, h=h, omegas=[0.9], beta=np.array([0.0, np.pi / 3]))


@pytest.mark.parametrize("h", [40.0, -1.0], ids=["finite", "deep"])
def test_field_points_match_oracle(oracle, h):
    p = _problem(h)
    j = 5
    xfp = np.array([p.xyz_p[j],              # on a hull panel
                    [6.0, 0.5, -1.0],        # in the fluid between the bodies
                    [20.0, 15.0, 0.0],       # on the free surface
                    [6.0, -3.0, -12.0]])     # below the bodies
    ctx = hm_bem.Context()
    try:
        ctx.setup(p)
        ctx.solve_frequency(0.9, hm_bem.HM_FP64)
        got = ctx.field_points(xfp)
        x = ctx.get_potentials()
    finally:
        ctx.close()
    ref = oracle.field_points(p, 0.9, xfp)
    assert ref["info"] == 0
    errs = {}
    for k in ("pot", "prs", "elv"):
        # per column, at least 1e-3 of the largest value: the yaw columns of a
        # vertical cylinder are rounding noise (~1e-15 of the largest value, its
        # yaw normals vanish), whose pattern depends on how the normals were
        # rounded (body frames or global coordinates)
        scale = np.maximum(np.max(np.abs(ref[k]), axis=0), 1e-3 * np.max(np.abs(ref[k])))
        errs[k] = float(np.max(np.abs(got[k] - ref[k]) / scale))
    print("h", h, "field-point errors against the oracle", errs, flush=True)
    bar = 2e-5 if h > 0 else 1e-12
    for k, e in errs.items():
        assert e < bar, (k, errs)
    # pressure and elevation from the potential (i w rho phi, i w / g phi)
    np.testing.assert_allclose(got["prs"], 1j * 0.9 * p.rho * got["pot"], rtol=1e-14, atol=0)
    np.testing.assert_allclose(got["elv"], 1j * 0.9 / p.g * got["pot"], rtol=1e-14, atol=0)
    # on the hull: the solved potential of that panel (radiation columns; the
    # diffraction columns add the incident potential). At finite depth the
    # field-point rows take FINGREEN3D per lane and the solve's matrix the
    # lane-group series: two evaluations within the Wynn band above, so the
    # equation is reproduced to ~1e-5 (a wrong free term would be off by 2x)
    m = 6 * p.nbody
    np.testing.assert_allclose(got["pot"][0, :m], x[j, :m], rtol=1e-4 if h > 0 else 1e-9,
                               atol=1e-12 * np.max(np.abs(x[:, :m])))


def test_field_points_need_a_solution():
    p = _problem(40.0)
    ctx = hm_bem.Context()
    try:
        ctx.setup(p)
        with pytest.raises(hm_bem.HMError, match="before hm_solve"):
            ctx.field_points(np.zeros((1, 3)))
    finally:
        ctx.close()
