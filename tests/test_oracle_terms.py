"""Pins the FP64 oracle's term energies and gradients (both formula families).

Checks that do not restate the implementation: closed-form anchor values, harmonic
second derivatives, central finite differences, and translation/rotation invariance.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from mlxmolkit_archive_amd.mmff_params import MMFFParams

MDYNE = 143.9325
ABC = 143.9325 * (np.pi / 180) ** 2  # 0.043844 kcal/mol/deg^2 per mdyn A/rad^2
FAMS = (oracle.FAMILY_TG, oracle.FAMILY_MEGA)


def _p(n, **kw):
    p = MMFFParams(n_atoms=n)
    for k, v in kw.items():
        dt = np.int32 if ("idx" in k or k in ("angle_is_linear", "ele_is_1_4")) else np.float32
        setattr(p, k, np.asarray(v, dtype=dt))
    return p


def E(p, x, fam, diel=1, mask=oracle.T_ALL):
    return oracle.energy_grad(p, np.asarray(x, float), family=fam, diel=diel, mask=mask)[0]


def G(p, x, fam, diel=1):
    return oracle.energy_grad(p, np.asarray(x, float), family=fam, diel=diel)[1]


def bond_p(kb=5.0, r0=1.5):
    return _p(2, bond_idx1=[0], bond_idx2=[1], bond_kb=[kb], bond_r0=[r0])


@pytest.mark.parametrize("fam", FAMS)
def test_bond_closed_forms(fam):
    kb, r0 = 4.5, float(np.float32(1.51))  # parameters are stored as float32
    p = bond_p(kb, r0)
    at = lambda r: np.array([0, 0, 0, r, 0, 0.0])
    assert abs(E(p, at(r0), fam)) < 1e-9
    assert np.abs(G(p, at(r0), fam)).max() < 1e-8
    h = 1e-4  # harmonic limit: E''(r0) = MDYNE * kb
    d2 = (E(p, at(r0 + h), fam) - 2 * E(p, at(r0), fam) + E(p, at(r0 - h), fam)) / h ** 2
    assert d2 == pytest.approx(MDYNE * kb, rel=1e-4)
    # MMFF cubic stretch: cs = -2 A^-1 -> third derivative at r0 equals 3*cs*MDYNE*kb
    d3 = (E(p, at(r0 + 2 * h), fam) - 2 * E(p, at(r0 + h), fam) + 2 * E(p, at(r0 - h), fam)
          - E(p, at(r0 - 2 * h), fam)) / (2 * h ** 3)
    assert d3 == pytest.approx(3 * -2.0 * MDYNE * kb, rel=2e-3)


def angle_geom(theta_deg, r1=1.1, r2=1.3):
    t = np.deg2rad(theta_deg)
    return np.array([r1, 0, 0, 0, 0, 0, r2 * np.cos(t), r2 * np.sin(t), 0.0])


@pytest.mark.parametrize("fam", FAMS)
def test_angle_closed_forms(fam):
    ka, t0 = 0.7, 112.0
    p = _p(3, angle_idx1=[0], angle_idx2=[1], angle_idx3=[2], angle_ka=[ka], angle_theta0=[t0],
           angle_is_linear=[0])
    assert abs(E(p, angle_geom(t0), fam)) < 1e-9
    h = 1e-3  # in degrees: harmonic limit E''(theta0) = 0.043844 * ka per deg^2
    d2 = (E(p, angle_geom(t0 + h), fam) - 2 * E(p, angle_geom(t0), fam) + E(p, angle_geom(t0 - h), fam)) / h ** 2
    assert d2 == pytest.approx(ABC * ka, rel=1e-4)
    # cubic bend: E(t0+D) - E(t0-D) = ABC*ka*cb*D^3, cb = -0.4*pi/180 (TG) or -0.007 (MEGA)
    D = 5.0
    cb = -0.4 * np.pi / 180 if fam == oracle.FAMILY_TG else -0.007
    asym = E(p, angle_geom(t0 + D), fam) - E(p, angle_geom(t0 - D), fam)
    assert asym == pytest.approx(ABC * ka * cb * D ** 3, rel=1e-3)


def test_linear_angle_branch():
    """TG family: E = 143.9325 ka (1 + cos theta) for linear centres (mmff_bfgs_header.metal:82-84)."""
    ka = float(np.float32(0.3))
    p = _p(3, angle_idx1=[0], angle_idx2=[1], angle_idx3=[2], angle_ka=[ka], angle_theta0=[180.0],
           angle_is_linear=[1])
    assert abs(E(p, angle_geom(180.0), oracle.FAMILY_TG)) < 1e-9
    assert E(p, angle_geom(90.0), oracle.FAMILY_TG) == pytest.approx(MDYNE * ka, rel=1e-12)
    assert E(p, angle_geom(0.0 + 1e-9), oracle.FAMILY_TG) == pytest.approx(2 * MDYNE * ka, rel=1e-9)


def torsion_geom(phi_deg):
    t = np.deg2rad(phi_deg)
    # 1 at +x, 2 at origin, 3 on +z, 4 rotated by phi about the 2-3 axis (phi = 0 is cis)
    return np.array([1.0, 0, 0, 0, 0, 0, 0, 0, 1.5, np.cos(t), np.sin(t), 1.5])


@pytest.mark.parametrize("fam", FAMS)
def test_torsion_closed_forms(fam):
    V1, V2, V3 = (float(np.float32(v)) for v in (0.8, 2.1, 0.4))
    p = _p(4, torsion_idx1=[0], torsion_idx2=[1], torsion_idx3=[2], torsion_idx4=[3], torsion_V1=[V1],
           torsion_V2=[V2], torsion_V3=[V3])
    assert E(p, torsion_geom(0.0), fam) == pytest.approx(V1 + V3, abs=1e-7)
    assert E(p, torsion_geom(180.0), fam) == pytest.approx(0.0, abs=1e-7)
    assert E(p, torsion_geom(90.0), fam) == pytest.approx(0.5 * (V1 + 2 * V2 + V3), abs=1e-7)
    assert E(p, torsion_geom(60.0), fam) == pytest.approx(E(p, torsion_geom(-60.0), fam), abs=1e-9)


@pytest.mark.parametrize("fam", FAMS)
def test_oop_planar_zero_and_perpendicular(fam):
    koop = float(np.float32(0.05))
    abc = ABC if fam == oracle.FAMILY_TG else 0.043844  # the mega family uses the rounded constant
    p = _p(4, oop_idx1=[0], oop_idx2=[1], oop_idx3=[2], oop_idx4=[3], oop_koop=[koop])
    planar = np.array([1.0, 0, 0, 0, 0, 0, -0.5, 0.9, 0, -0.5, -0.9, 0.0])
    assert abs(E(p, planar, fam)) < 1e-12
    # atom l 60 degrees out of the (i, j, k) plane (away from chi = 90, where asin is ill-conditioned)
    tilt = np.array([1.0, 0, 0, 0, 0, 0, -0.5, 0.9, 0, -np.cos(np.pi / 3), 0, np.sin(np.pi / 3)])
    assert E(p, tilt, fam) == pytest.approx(0.5 * abc * koop * 60.0 ** 2, rel=1e-9)


@pytest.mark.parametrize("fam", FAMS)
def test_vdw_at_rstar_is_minus_eps(fam):
    R, eps = float(np.float32(3.6)), float(np.float32(0.08))
    p = _p(2, vdw_idx1=[0], vdw_idx2=[1], vdw_R_star=[R], vdw_eps=[eps])
    x = np.array([0, 0, 0, R, 0, 0.0])
    assert E(p, x, fam) == pytest.approx(-eps, rel=1e-6)
    far = E(p, np.array([0, 0, 0, 40.0, 0, 0]), fam)
    assert -1e-6 < far < 0.0  # weak attractive r^-7 tail


@pytest.mark.parametrize("fam", FAMS)
def test_coulomb_buffered(fam):
    qq, r = -0.12, 3.0
    for is14, sc in ((0, 1.0), (1, 0.75)):
        p = _p(2, ele_idx1=[0], ele_idx2=[1], ele_charge_term=[qq], ele_is_1_4=[is14])
        x = np.array([0, 0, 0, r, 0, 0.0])
        assert E(p, x, fam) == pytest.approx(332.0716 * np.float32(qq) * sc / (r + 0.05), rel=1e-6)
    p = _p(2, ele_idx1=[0], ele_idx2=[1], ele_charge_term=[qq], ele_is_1_4=[0])
    assert E(p, np.array([0, 0, 0, r, 0, 0.0]), oracle.FAMILY_TG, diel=2) == pytest.approx(
        332.0716 * np.float32(qq) / (r + 0.05) ** 2, rel=1e-6)


@pytest.mark.parametrize("fam", FAMS)
def test_strbend_zero_when_either_factor_zero(fam):
    p = _p(3, strbend_idx1=[0], strbend_idx2=[1], strbend_idx3=[2], strbend_kba_ijk=[0.3], strbend_kba_kji=[0.2],
           strbend_r0_ij=[1.1], strbend_r0_kj=[1.3], strbend_theta0=[110.0])
    assert abs(E(p, angle_geom(110.0, 1.4, 1.0), fam)) < 1e-7   # d theta = 0
    assert abs(E(p, angle_geom(95.0, 1.1, 1.3), fam)) < 1e-7    # d r = 0
    # product form: E = 2.51210 * dtheta * (kij dr_ij + kkj dr_kj)
    v = E(p, angle_geom(115.0, 1.2, 1.3), fam)
    assert v == pytest.approx(MDYNE * np.pi / 180 * 5.0 * (np.float32(0.3) * 0.1), rel=1e-4)


def _fd(p, x, fam, h=1e-5):
    g = np.zeros_like(x)
    for i in range(len(x)):
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        g[i] = (E(p, xp, fam) - E(p, xm, fam)) / (2 * h)
    return g


@pytest.mark.parametrize("fam", FAMS)
def test_gradient_matches_central_differences(fam, small_wl):
    """Every term's analytic gradient vs FD of the FP64 energy on generated molecules."""
    wl = small_wl
    off = 0
    for p in wl.params[:4]:
        x = wl.x[off:off + 3 * p.n_atoms].astype(np.float64)
        off += 3 * p.n_atoms * wl.conf_counts[0]
        for mask in (1, 2, 4, 8, 16, 32, 64):
            ga = oracle.energy_grad(p, x, family=fam, mask=mask)[1]
            gn = np.zeros_like(x)
            for i in range(len(x)):
                xp, xm = x.copy(), x.copy()
                xp[i] += 1e-5
                xm[i] -= 1e-5
                gn[i] = (oracle.energy_grad(p, xp, family=fam, mask=mask)[0]
                         - oracle.energy_grad(p, xm, family=fam, mask=mask)[0]) / 2e-5
            scale = max(np.abs(gn).max(), 1e-3)
            assert np.abs(ga - gn).max() / scale < 2e-5, (fam, mask)


@pytest.mark.parametrize("fam", FAMS)
def test_translation_and_rotation_invariance(fam, small_wl):
    wl = small_wl
    p = wl.params[1]
    x = wl.x[3 * wl.params[0].n_atoms * wl.conf_counts[0]:][:3 * p.n_atoms].astype(np.float64)
    e0, g = oracle.energy_grad(p, x, family=fam)
    X = x.reshape(-1, 3)
    G3 = g.reshape(-1, 3)
    assert np.abs(G3.sum(0)).max() < 1e-8 * max(1.0, np.abs(G3).max())       # net force
    assert np.abs(np.cross(X, G3).sum(0)).max() < 1e-7 * max(1.0, np.abs(G3).max() * np.abs(X).max())  # torque
    rng = np.random.default_rng(0)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    xr = (X @ q.T + np.array([3.0, -1.0, 2.0])).reshape(-1)
    assert oracle.energy_grad(p, xr, family=fam)[0] == pytest.approx(e0, rel=1e-10, abs=1e-9)


def test_families_agree_near_equilibrium(small_wl):
    """TG vs MEGA total energies differ only through the cubic-bend constant and guards
    (restates tests/test_mmff_minimize_nk.py:25-40 at <= 1e-3 kcal/mol)."""
    wl = small_wl
    ob = oracle.OracleBatch(wl.params[:3], [1, 1, 1])
    x0 = np.concatenate([wl.x[sum(3 * q.n_atoms * 2 for q in wl.params[:k]):][:3 * wl.params[k].n_atoms]
                         for k in range(3)]).astype(np.float64)
    xm, _, st, _ = oracle.minimize_nk(ob, x0, max_iters=2000, grad_tol=1e-6)
    e_tg, _, et_tg = oracle.energy_grad_batch(ob, xm, family=oracle.FAMILY_TG)
    e_mg, _, et_mg = oracle.energy_grad_batch(ob, xm, family=oracle.FAMILY_MEGA)
    assert np.abs(e_tg - e_mg).max() < 1e-2
    # every term but the angle bend is identical up to the guards; out-of-plane differs by
    # the constant (0.043844 in the mega family vs 143.9325*(pi/180)^2 = 0.04384384...)
    for k in (0, 2, 4, 5, 6):
        assert np.abs(et_tg[:, k] - et_mg[:, k]).max() < 1e-6
    assert np.allclose(et_mg[:, 3], et_tg[:, 3] * (0.043844 / ABC), rtol=1e-9, atol=1e-12)
