"""Independent numpy formulation of the two-colour D3Q19 MRT step, written from the
textbook definitions rather than from the reference code:

* moment matrix of d'Humieres et al., Phil. Trans. R. Soc. A 360 (2002) 437, built from
  its polynomial definitions (rho, e, eps, j, q, 3p_xx, 3pi_xx, p_ww, pi_ww, p_xy, p_yz,
  p_xz, m_x, m_y, m_z);
* equilibrium moments with the (omega_eps, omega_epsj, omega_xx) = (0, -475/63, 0) choice;
* Guo et al. (PRE 65, 046308, 2002) forcing projected on the moments,
  F_i = w_i [3 (e_i - u) . F + 9 (e_i . u)(e_i . F)], with u = sum e_i f_i + F/2;
* Latva-Kokko & Rothman (PRE 71, 056702, 2005) recolouring,
  g1_i = (rho1/rho) f*_i + beta (rho1 rho2 / rho) w_i (e_i . n) / |e_i|.

It is used only to pin the oracle's hand-ordered arithmetic.
"""
import numpy as np

E = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
              [1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1], [1, 0, -1], [-1, 0, -1],
              [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1]], dtype=np.float64)
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
OPP = np.array([int(np.where((E == -e).all(axis=1))[0][0]) for e in E])


def moment_matrix():
    ex, ey, ez = E[:, 0], E[:, 1], E[:, 2]
    e2 = ex ** 2 + ey ** 2 + ez ** 2
    rows = [
        np.ones(19),
        19 * e2 - 30,
        (21 * e2 ** 2 - 53 * e2 + 24) / 2,
        ex, (5 * e2 - 9) * ex,
        ey, (5 * e2 - 9) * ey,
        ez, (5 * e2 - 9) * ez,
        3 * ex ** 2 - e2, (3 * e2 - 5) * (3 * ex ** 2 - e2),
        ey ** 2 - ez ** 2, (3 * e2 - 5) * (ey ** 2 - ez ** 2),
        ex * ey, ey * ez, ex * ez,
        (ey ** 2 - ez ** 2) * ex, (ez ** 2 - ex ** 2) * ey, (ex ** 2 - ey ** 2) * ez,
    ]
    return np.array(rows)


M = moment_matrix()
MINV = np.linalg.inv(M)


def model_rates(model, omega):
    """(s_e, s_e2, s_pi, s_q, s_t) of the reference's four collision models (its `mrt` macro):
    1 bounce-back optimised, 2 original, 3 single relaxation time, 4 advection optimised."""
    if model == 1:
        sq = 8 * (2 - omega) / (8 - omega)
        return omega, omega, omega, sq, sq
    if model == 3:
        return omega, omega, omega, omega, omega
    if model == 4:
        return omega, omega, omega, (6 - 3 * omega) / (3 - omega), omega
    return 1.19, 1.4, 1.4, 1.2, 1.98


def collide(g1, g2, n, curv, cnorm, gamma, force_z, nui1, nui2, beta, model=2):
    """Post-collision (o1, o2) and phi for populations g1, g2 of shape (19, ...)."""
    g1 = np.asarray(g1, dtype=np.float64)
    g2 = np.asarray(g2, dtype=np.float64)
    f = g1 + g2
    rho1, rho2 = g1.sum(0), g2.sum(0)
    rho = rho1 + rho2
    phi = (rho1 - rho2) / rho
    n = np.asarray(n, dtype=np.float64)
    fmag = 0.5 * gamma * curv * cnorm
    F = np.stack([fmag * n[0], fmag * n[1], fmag * n[2] + force_z])
    j = np.tensordot(E.T, f, axes=(1, 0)) + 0.5 * F  # "velocity" of the incompressible form
    u2 = (j ** 2).sum(0)
    omega = 1.0 / (6.0 / ((1 + phi) * nui1 + (1 - phi) * nui2) + 0.5)
    S = np.zeros((19,) + f.shape[1:])
    se, se2, spi, sq, st = model_rates(model, omega)
    for r, v in ((1, se), (2, se2), (4, sq), (6, sq), (8, sq), (10, spi), (12, spi), (16, st), (17, st), (18, st)):
        S[r] = v
    for r in (9, 11, 13, 14, 15):
        S[r] = omega
    m = np.tensordot(M, f, axes=(1, 0))
    ux, uy, uz = j
    meq = np.zeros_like(m)
    meq[0] = rho
    meq[1] = -11 * rho + 19 * u2
    meq[2] = 0.0 * rho - 475.0 / 63.0 * u2
    meq[3], meq[5], meq[7] = ux, uy, uz
    meq[4], meq[6], meq[8] = -2 / 3 * ux, -2 / 3 * uy, -2 / 3 * uz
    meq[9] = 3 * ux ** 2 - u2
    meq[11] = uy ** 2 - uz ** 2
    meq[13], meq[14], meq[15] = ux * uy, uy * uz, ux * uz
    # Guo forcing distribution and its moments
    eu = np.tensordot(E, j, axes=(1, 0))
    eF = np.tensordot(E, F, axes=(1, 0))
    uF = (j * F).sum(0)
    Fi = W.reshape((19,) + (1,) * (f.ndim - 1)) * (3 * (eF - uF) + 9 * eu * eF)
    mF = np.tensordot(M, Fi, axes=(1, 0))
    mpost = m - S * (m - meq) + (1 - S / 2) * mF
    fpost = np.tensordot(MINV, mpost, axes=(1, 0))
    en = np.tensordot(E, n, axes=(1, 0))
    elen = np.sqrt((E ** 2).sum(1))
    elen[0] = 1.0
    coef = (W / elen).reshape((19,) + (1,) * (f.ndim - 1))
    o1 = (rho1 / rho) * fpost + beta * (rho1 * rho2 / rho) * coef * en
    o1[0] = (rho1 / rho) * fpost[0]
    o2 = fpost - o1
    return phi, o1, o2


def equilibrium_moments_state(rho):
    """Populations whose moments are the MRT equilibrium at rest (rho, -11 rho, 0, ...)."""
    m = np.zeros(19)
    m[0] = rho
    m[1] = -11 * rho
    return MINV @ m
