"""Pins of the oracle's geometry preprocessing (src/Geometry_preprocessing.cpp:29-403) by
independent constructions:

* the smoothed solid indicator is four Jacobi passes of the D3Q27 filter with weights
  8/27, 2/27, 1/54, 1/216 (src/Geometry_preprocessing.cpp:187-206); that filter is the
  tensor product of the 1-D filter (1, 4, 1)/6, so numpy applies (1, 4, 1)/6 along x, y and
  z in turn and must agree to rounding;
* the solid normals are the reference's ISO8 gradient (src/Geometry_preprocessing.cpp:
  229-386, weights includes/Fluid_multiphase.h:35) of that field.  Here the stencil is built
  by enumerating every integer offset e with |e_i| <= 2, e_a > 0 along the derivative axis a
  and |e|^2 in {1, 2, 3, 4, 5, 6, 8}, each difference f(x + e) - f(x - e) weighted by
  w(|e|^2) with coefficient 1.  That is the reference's form: it does not multiply the
  offsets with e_a = 2 by 2 as the textbook E8 gradient does, so its gain on a linear field
  is 37/42, not 1 (pinned below, so that nobody "corrects" it);
* normalisation s_n = nw / (|nw| + FLT_EPSILON) (src/Geometry_preprocessing.cpp:377-383),
  checked at a flat wall against exact rational arithmetic.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
from helpers import small_drainage
from mflbm import cases, config

G = 10  # ghost layers of the reference's temporary arrays (src/Geometry_preprocessing.cpp:12-20)
FLT_EPS = float(np.finfo(np.float32).eps)
ISO8_W = {1: Fraction(4, 45), 2: Fraction(1, 21), 3: Fraction(2, 105), 4: Fraction(5, 504),
          5: Fraction(1, 315), 6: Fraction(1, 630), 8: Fraction(1, 5040)}


def iso8_offsets(axis):
    """(offset, |e|^2) of every term of the reference-form ISO8 derivative along axis
    (0 = x, 1 = y, 2 = z; offsets as (ex, ey, ez))."""
    out = []
    r = range(-2, 3)
    for ex in r:
        for ey in r:
            for ez in r:
                e = (ex, ey, ez)
                m = ex * ex + ey * ey + ez * ez
                if e[axis] > 0 and m in ISO8_W:
                    out.append((e, m))
    return out


def test_iso8_enumeration_has_the_reference_group_sizes_and_gain():
    """1, 4, 4, 1, 8, 12, 4 terms per weight (src/Geometry_preprocessing.cpp:232-384), gain
    on f = x of sum 2 e_x w = 37/42, and the textbook isotropy sum_e w e_x^2 = 1 over
    all 2*34 offsets (the factor the reference leaves out)."""
    for axis in range(3):
        offs = iso8_offsets(axis)
        sizes = {m: sum(1 for _, mm in offs if mm == m) for m in ISO8_W}
        assert sizes == {1: 1, 2: 4, 3: 4, 4: 1, 5: 8, 6: 12, 8: 4}
        # f = x_axis: each difference f(x + e) - f(x - e) is 2 e_axis
        gain = sum(ISO8_W[m] * 2 * e[axis] for e, m in offs)
        assert gain == Fraction(37, 42)
        iso = sum(ISO8_W[m] * 2 * e[axis] * e[axis] for e, m in offs)
        assert iso == 1


def padded_indicator(geo, p):
    """Solid indicator with the domain walls (src/Misc.cpp:67-78) and G ghost layers filled
    z first, then y, then x, each by periodic wrap or by copying the boundary plane
    (src/Geometry_preprocessing.cpp:56-123).  Array axes (z, y, x)."""
    g = (np.asarray(geo) != 0).astype(np.float64)
    if p["wall_z_min"]:
        g[0] = 1
    if p["wall_z_max"]:
        g[-1] = 1
    if p["wall_x_min"]:
        g[:, :, 0] = 1
    if p["wall_x_max"]:
        g[:, :, -1] = 1
    if p["wall_y_min"]:
        g[:, 0, :] = 1
    if p["wall_y_max"]:
        g[:, -1, :] = 1
    for axis, per in ((0, p["kper"]), (1, p["jper"]), (2, p["iper"])):
        w = [(0, 0)] * 3
        w[axis] = (G, G)
        g = np.pad(g, w, mode="wrap" if per else "edge")
    return g


def smooth_separable(a, passes=4):
    """passes x (1,4,1)/6 along x, y, z on every node but the outermost layer, which keeps
    its value (the reference sweeps [2-G, n+G-1])."""
    a = a.copy()
    for _ in range(passes):
        b = a.copy()
        b[:, :, 1:-1] = (a[:, :, :-2] + 4.0 * a[:, :, 1:-1] + a[:, :, 2:]) / 6.0
        c = b.copy()
        c[:, 1:-1, :] = (b[:, :-2, :] + 4.0 * b[:, 1:-1, :] + b[:, 2:, :]) / 6.0
        d = c.copy()
        d[1:-1, :, :] = (c[:-2, :, :] + 4.0 * c[1:-1, :, :] + c[2:, :, :]) / 6.0
        a[1:-1, 1:-1, 1:-1] = d[1:-1, 1:-1, 1:-1]
    return a


def s4_window(a, shape):
    """[-3, n+4]^3 of a padded (G ghost layers) array."""
    nz, ny, nx = shape
    return a[G - 4:G + nz + 4, G - 4:G + ny + 4, G - 4:G + nx + 4]


def iso8_normals(sm, sel):
    """Reference-form ISO8 gradient of the padded field sm at the nodes sel (mask over the
    s4 window), normalised with FLT_EPSILON."""
    nz, ny, nx = (s - 2 * G for s in sm.shape)
    nw = np.zeros((3,) + sel.shape)
    z0 = G - 4
    for axis in range(3):
        acc = np.zeros(sel.shape)
        for (ex, ey, ez), m in iso8_offsets(axis):
            plus = sm[z0 + ez:z0 + ez + nz + 8, z0 + ey:z0 + ey + ny + 8, z0 + ex:z0 + ex + nx + 8]
            minus = sm[z0 - ez:z0 - ez + nz + 8, z0 - ey:z0 - ey + ny + 8, z0 - ex:z0 - ex + nx + 8]
            acc += float(ISO8_W[m]) * (plus - minus)
        nw[axis] = np.where(sel, acc, 0.0)
    inv = 1.0 / (np.sqrt((nw ** 2).sum(0)) + FLT_EPS)
    return np.where(sel, nw * inv, 0.0)


CASES = {
    "drainage_open_z_periodic_y": lambda: small_drainage(n=(14, 16, 18), seed=3, jper=1),
    "drainage_walled_y": lambda: small_drainage(n=(16, 14, 20), seed=5, jper=0, radius=3.0),
    "periodic_yz_obstacles": lambda: (cases.geometry.sphere_pack((18, 16, 20), 3.0, 0.7, 11, periodic=(True, True, False)),
                                      config.derive(cases.periodic_control(n=(20, 16, 18)))),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_smoothing_is_the_separable_141_filter(case):
    geo, p = CASES[case]()
    o = O.Oracle(p, geo)
    ref = s4_window(smooth_separable(padded_indicator(geo, p)), geo.shape)
    got = o.array("smooth")
    np.testing.assert_allclose(got, ref, rtol=0, atol=2e-15)


@pytest.mark.parametrize("case", sorted(CASES))
def test_solid_normals_are_the_enumerated_iso8_gradient(case):
    geo, p = CASES[case]()
    o = O.Oracle(p, geo)
    sm = smooth_separable(padded_indicator(geo, p))
    sel = o.int_array("walls_type") == -1
    assert sel.sum() > 100
    ref = iso8_normals(sm, sel)
    got = np.stack([o.array(c) for c in ("s_nx", "s_ny", "s_nz")])
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13)


def test_smoothing_conserves_the_solid_volume_of_a_periodic_domain():
    """The filter's weights sum to 1 (8/27 + 6 * 2/27 + 12/54 + 8/216), so on a domain
    periodic in x, y and z the smoothed indicator summed over one period equals the solid
    volume.  (The reference requires x walls; the oracle itself takes any iper.)"""
    n = (16, 14, 12)
    p = config.derive(cases.periodic_control(n=n))
    p.update(iper=1, wall_x_min=0, wall_x_max=0)  # past the control-file check, on the oracle only
    geo = cases.geometry.sphere_pack((12, 14, 16), 3.0, 0.6, 4, periodic=(True, True, True))
    o = O.Oracle(p, geo)
    sm = o.array("smooth")[4:-4, 4:-4, 4:-4]  # [1, n]^3
    assert abs(sm.sum() - float(geo.sum())) < 1e-10
    assert sm.min() >= 0.0 and sm.max() <= 1.0 + 1e-15


def test_flat_wall_normal_length_in_exact_arithmetic():
    """Solid for k >= 9 of a z-periodic 16-plane domain: the smoothed field depends on z
    alone, so nw = (0, 0, sum_e w(|e|^2) (s(k+e_z) - s(k-e_z))) with s the periodic 1-D
    smoothing of the indicator, and s_nz = nw_z / (|nw_z| + FLT_EPSILON).  Both are computed
    here in exact rationals (the oracle rounds a few times: 1e-15)."""
    nz = 16
    p = config.derive(cases.periodic_control(n=(24, 16, nz)))
    geo = np.zeros((nz, 16, 24), np.int8)
    geo[8:] = 1
    o = O.Oracle(p, geo)
    c = [Fraction(1 if k >= 9 else 0) for k in range(1, nz + 1)]  # k = 1 .. nz
    for _ in range(4):
        c = [(c[(k - 1) % nz] + 4 * c[k] + c[(k + 1) % nz]) / 6 for k in range(nz)]

    def s(k):  # 1-based, periodic
        return c[(k - 1) % nz]

    eps = Fraction(FLT_EPS)
    sm = o.array("smooth")
    sn = np.stack([o.array(a) for a in ("s_nx", "s_ny", "s_nz")])
    wt = o.int_array("walls_type")
    for k in (8, 1):  # the fluid-boundary planes below and above the solid
        nwz = sum(ISO8_W[m] * (s(k + e[2]) - s(k - e[2])) for e, m in iso8_offsets(2))
        expect = nwz / (abs(nwz) + eps)
        xs = np.s_[8 + 3:17 + 3]  # x in [8, 17]: out of reach of the x walls
        assert (wt[k + 3, 4:-4, xs] == -1).all()
        np.testing.assert_allclose(sm[k + 3, 4:-4, xs], float(s(k)), rtol=0, atol=1e-15)
        np.testing.assert_allclose(sn[2][k + 3, 4:-4, xs], float(expect), rtol=0, atol=1e-15)
        assert np.abs(sn[0][k + 3, 4:-4, xs]).max() < 1e-15 and np.abs(sn[1][k + 3, 4:-4, xs]).max() < 1e-15
    assert float(sum(ISO8_W[m] * (s(8 + e[2]) - s(8 - e[2])) for e, m in iso8_offsets(2))) > 0  # into the solid
