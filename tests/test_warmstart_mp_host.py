"""Host side of the multiphase warm start (mflbm.warmstart_mp): the drainage cases, the
refined cases' pressure scaling and the network features."""
import numpy as np

from mflbm import warmstart_mp as W
from mflbm.warmstart import coarsen


def test_drainage_case_is_seeded_with_empty_end_layers():
    g1, p1, m1 = W.drainage_case(3, 32)
    g2, p2, m2 = W.drainage_case(3, 32)
    assert np.array_equal(g1, g2) and m1 == m2
    assert not g1[:W.BUFFER].any() and not g1[-W.BUFFER:].any()
    assert g1[W.BUFFER:-W.BUFFER].any()
    assert p1["n_exclude_inlet"] == W.BUFFER and p1["interface_z0"] == W.BUFFER
    assert 0.30 <= m1["porosity"] <= 0.40 and 4.0 <= m1["radius"] <= 7.0
    g3, _, _ = W.drainage_case(4, 32)
    assert not np.array_equal(g1, g3)


def test_fine_case_halves_the_pressure_drop():
    """rho_in - rho_out = F nz / 3 (src/Init_multiphase.cpp:196-214); the refined case keeps
    the ratio of that drop to the capillary entry pressures, which halve with the pore
    sizes: half the drop over twice the planes."""
    gc, pc, mc = W.drainage_case(5, 32)
    gf, pf, mf = W.fine_case(5, 32)
    assert np.array_equal(coarsen(gf), gc) and gf.shape == (64, 64, 64)
    dc = pc["rho_in"] - pc["rho_out"]
    df = pf["rho_in"] - pf["rho_out"]
    assert abs(df / dc - 0.5) < 1e-12
    assert pf["n_exclude_inlet"] == 2 * W.BUFFER and pf["interface_z0"] == 2 * W.BUFFER
    assert mf["radius"] == 2 * mc["radius"]


def test_features():
    g, _, _ = W.drainage_case(1, 32)
    x = W.features(g)
    solid = W.solid_of(g)
    assert x.shape == g.shape + (8,) and x.dtype == np.float32
    assert np.array_equal(x[..., 0] > 0.5, solid)
    assert np.all(x[..., 1][solid] <= 0) and np.all(x[..., 1][~solid] > 0)
    assert np.all(np.diff(x[:, 0, 0, 2]) > 0) and 0 < x[0, 0, 0, 2] < x[-1, 0, 0, 2] < 1
    assert np.all(x[:W.BUFFER, ..., 3] == 1) and np.all(x[W.BUFFER:, ..., 3] == 0)
    assert not x[..., 4:].any()


def test_field_agreement():
    g, _, _ = W.drainage_case(1, 32)
    pore = ~W.solid_of(g)
    a = np.where(pore, 1.0, -1.0)
    assert W.field_agreement(a, a, g) == 1.0
    assert W.field_agreement(a, -a, g) == 0.0


def test_sharpening_keeps_the_outputs_saturation_estimate():
    g, _, _ = W.drainage_case(2, 32)
    pore = ~W.solid_of(g)
    rng = np.random.default_rng(0)
    y = rng.uniform(-1.2, 0.6, size=g.shape)
    phi = W.PhaseWarmStart.sharpen(y, g)
    assert set(np.unique(phi)) <= {-1.0, 1.0} and np.all(phi[~pore] == -1.0)
    s_est = np.mean((1 + np.clip(y[pore], -1, 1)) / 2)
    assert abs((phi[pore] > 0).mean() - s_est) < 1.0 / pore.sum() + 1e-12
    # fluid 1 goes to the highest outputs
    assert y[pore][phi[pore] > 0].min() >= y[pore][phi[pore] < 0].max()
