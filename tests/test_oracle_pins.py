"""Pins of the oracle functions that the closed-form tests of test_oracle_physics.py leave open.

Each test checks the oracle (CPU, FP64) against what the mathematics fixes, chosen so that a
plausible slip (dropped factor, wrong sign, wrong index or exponent, swapped operands) fails:
  - FlatWall (Wall.hpp:68-79): minus the LJ pair law at the wall distance, zero at 2^(1/6),
    constant below the 0.82 clamp, none beyond rcut, directed by the wall's side;
  - the two-disc bond (MD.cu:255-265, 547-559): minus the gradient of the FENE / push-spring
    energy in r^2, including the r2max = max(r2max, sigma_max) rule;
  - the Dividing state of binaryCellForce (MD.cu:560-568): half the contraction, the growth
    activity terms one at a time (k1 constant, k0 times the FENE value at Rmin^2, k2 r^4);
  - EA (MD.cuh:296-304): a linear spring in the overlap with the lower-UID disc's modulus;
  - cross adhesion (MD.cu:345-360): attractive terms between species scale by each cell's own
    factor, repulsive ones and same-species ones do not;
  - pressure friction (CIL.cu:921): special cases in which the mobility halves;
  - force shift (MD.cuh:100-104, 200-202): a constant offset of the force magnitude that
    vanishes at the cut-off;
  - active noise (CIL.cu:803-814, MD.cu:573-578): draw order (all radial, then all angular)
    and the component-wise force it produces;
  - linear growth: rest-length ramp and split geometry;
  - drain (CIL.cu:945-966): removed cells were outside, and the cells still outside are the
    ones swapped into a freed slot during the call;
  - calculDivisions (CIL.cu:369-455): switch timing, division and failure geometry;
  - updateDiameters (CIL.cu:970-990): a linear ramp that reaches the maximum one step before
    the switch time;
  - StochasticHeun (CIL.cu:1028-1049): the bond relaxes by (1 - a/2)^2 per step, motility
    moves a disc by dt m / zeta, and the noise variance per step is D dt.
"""
import math

import numpy as np
import pytest

from cil_cuda_amd import abi, configs
from cil_cuda_amd.colony import Colony
from oracle import oracle as O


def _single_disc(dim=2, L=40.0, periodic=1):
    p = configs.base_params(dim, 1, 1)
    configs.set_law(p, abi.LAW_NONE, 1.0)
    box = configs.make_box([L] * dim, [periodic] * dim)
    col = Colony(dim, 1, [1], ncells=[1])
    return p, box, col


# ------------------------------------------------------------------------------------------
# FlatWall
# ------------------------------------------------------------------------------------------
def _wall_force(x, sigma=1.2, eps=1.7, x0=3.0, positive=1, rcut=2 ** (1 / 6)):
    p, box, col = _single_disc(L=40.0, periodic=0)
    p.nwalls = 1
    w = p.walls[0]
    w.kind, w.axis, w.x0, w.rcut, w.epsilon, w.positive = 0, 0, x0, rcut, eps, positive
    s = col.species[0]
    s["pos"][0, 0] = [x, 20.0]
    s["sigma"][0, 0] = sigma
    s["vel"][0, 0] = [0.0, 1.0]  # no zero-velocity kick
    O.Oracle(p, box, traversal=1).forces(col, 0)
    return s["frc"][0, 0].copy()


def test_flat_wall_is_minus_lj_law_at_wall_distance():
    sigma, eps = 1.2, 1.7
    q = configs.base_params(2, 1, 1)
    q.species[0].eps_well[0] = 1.0
    configs.set_law(q, abi.LAW_LJ, 100.0)
    for d in (0.83, 0.9, 1.0, 1.05, 1.1):  # wall distance in units of sigma, above the clamp
        f = _wall_force(3.0 + d * sigma, sigma, eps)
        hit, f0 = O.pair_f0(q, 2, 0, 0, sigma, sigma, (d * sigma) ** 2)
        assert hit
        # LJParams' law is 24 eps (s^6 - 2 s^12) / r^2; the wall's is eps (2 s^12 - s^6) / r^2
        assert f[0] == pytest.approx(-eps * f0 / 24.0, rel=1e-12)
        assert f[1] == 0.0
        assert f[0] > 0 or d > 2 ** (1 / 6)


def test_flat_wall_zero_at_minimum_clamped_and_cut():
    sigma = 1.2
    assert abs(_wall_force(3.0 + 2 ** (1 / 6) * sigma * (1 - 1e-12), sigma)[0]) < 1e-9
    f_clamp = _wall_force(3.0 + 0.82 * sigma, sigma)[0]
    for d in (0.3, 0.5, 0.7, 0.8199):
        assert _wall_force(3.0 + d * sigma, sigma)[0] == pytest.approx(f_clamp, rel=1e-13)
    assert _wall_force(3.0 + 1.2 * sigma, sigma)[0] == 0.0
    assert f_clamp > 0
    # the wall on the other side pushes the other way, with the same magnitude
    assert _wall_force(3.0 - 0.9 * sigma, sigma, positive=0)[0] == pytest.approx(
        -_wall_force(3.0 + 0.9 * sigma, sigma)[0], rel=1e-15)


# ------------------------------------------------------------------------------------------
# Two-disc bond: FENE + push spring, and the Dividing state
# ------------------------------------------------------------------------------------------
def _bond(r, kappa=3.0, kappa_push=5.0, ratio=0.4, r2max=2.0, sig=(1.0, 1.0), state=abi.CS_CRAWLING,
          div_r2max=3.0, r2min=1.0, k0=(0.0, 0.0), k1=(0.0, 0.0), k2=(0.0, 0.0), angle=0.3, cil=1):
    p = configs.base_params(2, 1, 2)
    sp = p.species[0]
    sp.spring = abi.SPRING_FENE
    sp.kappa, sp.kappa_push, sp.spring_ratio, sp.r2max, sp.div_r2max = kappa, kappa_push, ratio, r2max, div_r2max
    sp.division_r2min = r2min
    for e in range(2):
        sp.growthk0[e], sp.growthk1[e], sp.growthk2[e] = k0[e], k1[e], k2[e]
    sp.cil = cil
    configs.set_law(p, abi.LAW_NONE, 1.0)
    box = configs.make_box([40.0, 40.0], [1, 1])
    col = Colony(2, 2, [1], ncells=[1])
    s = col.species[0]
    u = np.array([math.cos(angle), math.sin(angle)])
    s["pos"][0, 0] = [20.0, 20.0]
    s["pos"][0, 1] = np.array([20.0, 20.0]) + r * u
    s["sigma"][0] = sig
    s["state"][0] = state
    O.Oracle(p, box, traversal=1).forces(col, 0)
    return s["frc"][0].copy(), r * u


def _bond_energy(u, kappa, kappa_push, ratio, R2):
    # energy as a function of u = r^2 whose u-derivative is the reference's spring factor:
    # push region u <= R2 rho: kp (R2 rho - u)^2 / 2; FENE region: kappa Rf artanh(x / Rf)
    # with x = u - R2 rho, Rf = R2 (1 - rho)
    a = R2 * ratio
    if u <= a:
        return 0.5 * kappa_push * (a - u) ** 2
    Rf = R2 * (1 - ratio)
    return kappa * Rf * math.atanh((u - a) / Rf)


@pytest.mark.parametrize("sig,R2", [((1.0, 1.0), 2.0), ((1.1, 2.6), 2.6)])
def test_bond_is_minus_gradient_of_fene_push_energy(sig, R2):
    kappa, kp, ratio = 3.0, 5.0, 0.4
    for r in (0.3, 0.6, 0.85, 0.95, 1.2, 1.35):
        if r * r >= R2 - 1e-3:
            continue
        f, rij = _bond(r, kappa, kp, ratio, 2.0, sig)
        h = 1e-6
        dU = (_bond_energy((r + h) ** 2, kappa, kp, ratio, R2) - _bond_energy((r - h) ** 2, kappa, kp, ratio, R2)) / (2 * h)
        # force on the front disc: -dU/dr along the axis (the reference's contraction factor 2
        # of a crawling cell is the chain-rule factor d(r^2)/dr = 2 r)
        expect = -dU * rij / r
        assert np.allclose(f[1], expect, rtol=2e-6, atol=1e-9), (r, f[1], expect)
        assert np.array_equal(f[0], -f[1])


def test_dividing_bond_uses_division_rmax_and_half_contraction():
    kappa, kp, ratio, R2div = 3.0, 5.0, 0.4, 3.0
    for r in (0.5, 1.1, 1.5):
        f, rij = _bond(r, kappa, kp, ratio, 2.0, state=abi.CS_DIVIDING, div_r2max=R2div)
        h = 1e-6
        dU = (_bond_energy((r + h) ** 2, kappa, kp, ratio, R2div) - _bond_energy((r - h) ** 2, kappa, kp, ratio, R2div)) / (2 * h)
        # contraction factor 1: half the gradient of the bond energy
        assert np.allclose(f[1], -0.5 * dU * rij / r, rtol=2e-6, atol=1e-9)


def test_dividing_growth_activity_terms():
    r, kappa = 0.9, 3.0
    # growthk1 alone: constant activity along the unscaled axis, opposite on the two discs
    f, rij = _bond(r, kappa=0.0, kappa_push=0.0, state=abi.CS_DIVIDING, k1=(0.7, 0.25), cil=0)
    assert np.allclose(f[0], -0.7 * rij, rtol=1e-14) and np.allclose(f[1], 0.25 * rij, rtol=1e-14)
    # growthk0 alone: -k0 times FENE(kappa, R2, Rmin^2) = k0 kappa / (1 - (Rmin^2/R2)^2) = 4/3 k0 kappa
    # at Rmin^2 = R2 / 2; the bond itself is switched off by a zero push constant inside rho R2
    R2 = 3.0
    f, rij = _bond(r, kappa=kappa, kappa_push=0.0, ratio=0.5, state=abi.CS_DIVIDING, div_r2max=R2, r2min=R2 / 2,
                   k0=(0.3, 0.6))
    assert r * r <= 0.5 * R2
    assert np.allclose(f[1], 4.0 / 3.0 * 0.6 * kappa * rij, rtol=1e-13)
    assert np.allclose(f[0], -4.0 / 3.0 * 0.3 * kappa * rij, rtol=1e-13)
    # growthk2 alone: k2 r^4, so the front force grows as r^5 with the disc distance
    fa, ra = _bond(0.4, kappa=0.0, kappa_push=0.0, state=abi.CS_DIVIDING, k2=(0.0, 1.0))
    fb, rb = _bond(0.8, kappa=0.0, kappa_push=0.0, state=abi.CS_DIVIDING, k2=(0.0, 1.0))
    assert np.linalg.norm(fb[1]) / np.linalg.norm(fa[1]) == pytest.approx(32.0, rel=1e-12)
    assert np.linalg.norm(fa[1]) == pytest.approx(0.4 ** 5, rel=1e-12)


# ------------------------------------------------------------------------------------------
# EA law and cross adhesion
# ------------------------------------------------------------------------------------------
def test_ea_is_linear_in_overlap_with_lower_uid_modulus():
    p = configs.base_params(2, 2, 1)
    p.species[0].eps_core[0] = 40.0
    p.species[1].eps_core[0] = 90.0
    configs.set_law(p, abi.LAW_EA, 1.0)
    for si, sj in ((1.0, 1.0), (0.8, 1.3)):
        for h in (1e-3, 0.05, 0.2):
            r = 0.5 * (si + sj) - h
            for ti, E in ((0, 40.0), (1, 90.0)):
                hit, f0 = O.pair_f0(p, 2, ti, 1 - ti, si, sj, r * r)
                assert hit
                assert -f0 / h == pytest.approx(E, rel=1e-9)
        assert not O.pair_f0(p, 2, 0, 1, si, sj, (0.5 * (si + sj) * (1 + 1e-12)) ** 2)[0]


def _two_cell_sc(r, nspec, ca=(0.3, 0.8)):
    p = configs.base_params(2, nspec, 1)
    for s in range(nspec):
        sp = p.species[s]
        sp.eps_core[0], sp.eps_well[0], sp.sig_well[0] = 2.0, 0.5, 0.4
        sp.cross_adhesion = ca[s]
    configs.set_law(p, abi.LAW_SC, 1.0)
    box = configs.make_box([30.0, 30.0], [1, 1])
    col = Colony(2, 1, [1] * nspec, ncells=[1] * nspec) if nspec == 2 else Colony(2, 1, [2], ncells=[2])
    pts = ([10.0, 10.0], [10.0 + r * 0.6, 10.0 + r * 0.8])
    for k in range(2):
        sp_, c = (k, 0) if nspec == 2 else (0, k)
        col.species[sp_]["pos"][c, 0] = pts[k]
        col.species[sp_]["vel"][c, 0] = [1.0, 0.0]
    O.Oracle(p, box, traversal=1).forces(col, 0)
    if nspec == 2:
        return col.species[0]["frc"][0, 0].copy(), col.species[1]["frc"][0, 0].copy()
    return col.species[0]["frc"][0, 0].copy(), col.species[0]["frc"][1, 0].copy()


def test_cross_adhesion_scales_attraction_between_species_only():
    ca = (0.3, 0.8)
    # r > sigma: attractive well (f0 > 0); each cell's force scales by its own factor
    fi1, fj1 = _two_cell_sc(1.2, 1)
    fi2, fj2 = _two_cell_sc(1.2, 2, ca)
    assert np.dot(fi1, [0.6, 0.8]) > 0  # attraction pulls i towards j
    assert np.allclose(fi2, ca[0] * fi1, rtol=1e-14) and np.allclose(fj2, ca[1] * fj1, rtol=1e-14)
    assert np.allclose(fi1, -fj1, rtol=1e-14)
    # r < sigma: repulsive core (f0 < 0): no scaling
    fi1, fj1 = _two_cell_sc(0.8, 1)
    fi2, fj2 = _two_cell_sc(0.8, 2, ca)
    assert np.dot(fi1, [0.6, 0.8]) < 0
    assert np.array_equal(fi1, fi2) and np.array_equal(fj1, fj2)


# ------------------------------------------------------------------------------------------
# Pressure friction
# ------------------------------------------------------------------------------------------
def _disp_with_pressure_friction(zeta_p, beta, p0=2.0):
    """Two overlapping Hertz discs: one Euler step, the displacement of disc 0."""
    p = configs.base_params(2, 1, 1)
    sp = p.species[0]
    sp.eps_core[0], sp.sig_well[0] = 1000.0, 0.5
    sp.zeta[0], sp.zeta_p[0], sp.zeta_p_beta[0], sp.cycle_p0 = 1.5, zeta_p, beta, p0
    configs.set_law(p, abi.LAW_HZ, 1.0)
    p.pressure = 1
    p.dt = 1e-5
    box = configs.make_box([30.0, 30.0], [1, 1])
    col = Colony(2, 1, [2], ncells=[2])
    s = col.species[0]
    s["pos"][0, 0] = [10.0, 10.0]
    s["pos"][1, 0] = [10.9, 10.0]
    s["vel"][:, 0] = [0.0, 1.0]
    orc = O.Oracle(p, box, traversal=1)
    orc.forces(col, 0)
    pr = s["pressure"][0]
    x0 = s["pos"][0, 0].copy()
    orc.step(col, 1, 0)
    return s["pos"][0, 0] - x0, pr


def test_pressure_friction_halves_mobility_in_special_cases():
    d0, pr = _disp_with_pressure_friction(0.0, 1.0)
    assert pr > 0
    p0 = 2.0
    # beta 1: zeta_p = zeta p0 / p doubles the friction
    d1, _ = _disp_with_pressure_friction(1.5 * p0 / pr, 1.0, p0)
    assert np.allclose(d1, 0.5 * d0, rtol=1e-12)
    # beta 2 with p0 = p / 2: (p / p0)^2 = 4, zeta_p = zeta / 4
    d2, _ = _disp_with_pressure_friction(1.5 / 4, 2.0, pr / 2)
    assert np.allclose(d2, 0.5 * d0, rtol=1e-12)
    # beta 0: (p / p0)^0 = 1, zeta_p = zeta
    d3, _ = _disp_with_pressure_friction(1.5, 0.0, p0)
    assert np.allclose(d3, 0.5 * d0, rtol=1e-12)


# ------------------------------------------------------------------------------------------
# Force shift
# ------------------------------------------------------------------------------------------
@pytest.mark.parametrize("law", [abi.LAW_LJ, abi.LAW_SC])
def test_force_shift_is_constant_offset_vanishing_at_cutoff(law):
    p = configs.base_params(3, 1, 1)
    sp = p.species[0]
    sp.eps_well[0], sp.eps_core[0], sp.sig_well[0] = 1.3, 2.0, 0.4
    configs.set_law(p, law, 2.5 ** 2 if law == abi.LAW_LJ else 1.0)
    q = abi.Params.from_buffer_copy(p)
    q.force_shifted = 1
    si, sj = 0.9, 1.2
    sb = 0.5 * (si + sj)
    rc = 2.5 * sb if law == abi.LAW_LJ else sb + 0.4
    _, frc_c = O.pair_f0(p, 3, 0, 0, si, sj, rc * rc)
    Fc = frc_c * rc  # unshifted force magnitude at the cut-off (f0 multiplies r_ij)
    lo = 0.9 * sb if law == abi.LAW_LJ else sb - 0.35
    for r in np.linspace(lo, rc, 9):
        _, f = O.pair_f0(p, 3, 0, 0, si, sj, r * r)
        _, g = O.pair_f0(q, 3, 0, 0, si, sj, r * r)
        assert g * r - f * r == pytest.approx(-Fc, rel=1e-10, abs=1e-12)
    _, g = O.pair_f0(q, 3, 0, 0, si, sj, rc * rc)
    assert abs(g) < 1e-12


# ------------------------------------------------------------------------------------------
# Active noise
# ------------------------------------------------------------------------------------------
def _active_noise_colony(angles, d_r=0.7, d_a=0.4, r=0.8):
    n = len(angles)
    p = configs.base_params(2, 1, 2)
    sp = p.species[0]
    sp.spring = abi.SPRING_HARMONIC
    sp.k_harm = 0.0
    sp.d_noise_radial, sp.d_noise_angular = d_r, d_a
    p.active_noise = 1
    p.rng = abi.RNG_EXTERNAL
    configs.set_law(p, abi.LAW_NONE, 1.0)
    box = configs.make_box([60.0, 60.0], [1, 1])
    col = Colony(2, 2, [n], ncells=[n])
    s = col.species[0]
    for c, a in enumerate(angles):
        s["pos"][c, 0] = [5.0 + 5 * c, 20.0]
        s["pos"][c, 1] = s["pos"][c, 0] + r * np.array([math.cos(a), math.sin(a)])
        s["rest_len"][c] = r
    return p, box, col


def test_active_noise_draw_order_and_force():
    angles = [0.0, math.pi / 2, math.pi / 4]
    p, box, col = _active_noise_colony(angles)
    n, r, d_r, d_a = len(angles), 0.8, 0.7, 0.4
    z = np.array([[1.0, 2.0, 3.0, -1.0, -2.0, -3.0]])  # radial draws of cells 0..2, then angular
    orc = O.Oracle(p, box, traversal=1)
    orc.set_external_noise(z)
    orc.forces(col, 0)
    f = col.species[0]["frc"][:n]
    for c, a in enumerate(angles):
        Wr, Wt = d_r * z[0, c], d_a * z[0, n + c]
        rij = r * np.array([math.cos(a), math.sin(a)])
        # rndForce (MD.cu:573-578) multiplies rij component-wise
        rnd = 0.5 * np.array([Wr * rij[0] / r - Wt * rij[1], Wr * rij[1] / r + Wt * rij[0]])
        assert np.allclose(f[c, 1], rnd * rij, rtol=1e-14, atol=1e-15)
        assert np.allclose(f[c, 0], -rnd * rij, rtol=1e-14, atol=1e-15)
    # an axis along x or y gets no force from the angular draw
    assert f[0, 1, 1] == 0.0 and f[1, 1, 0] == 0.0
    assert f[0, 1, 0] == pytest.approx(0.5 * d_r * 1.0 * r, rel=1e-14)


def test_active_noise_variance_counter_draws():
    n = 4000
    p, box, col = _active_noise_colony([0.0] * 1)
    p.rng = abi.RNG_COUNTER
    col = Colony(2, 2, [n], ncells=[n])
    s = col.species[0]
    rng = np.random.default_rng(3)
    s["pos"][:, 0] = rng.uniform(5, 55, (n, 2))
    s["pos"][:, 1] = s["pos"][:, 0] + [0.8, 0.0]
    s["rest_len"][:] = 0.8
    O.Oracle(p, box, traversal=1).forces(col, 0)
    fx = s["frc"][:, 1, 0]  # 0.5 Wr r along x
    assert np.var(fx) == pytest.approx((0.5 * 0.7 * 0.8) ** 2, rel=0.06)
    assert np.all(s["frc"][:, 1, 1] == 0.0)


# ------------------------------------------------------------------------------------------
# Linear growth and drain
# ------------------------------------------------------------------------------------------
def test_linear_growth_ramp_and_split_geometry():
    p, box, col = configs.channel_2d_hz(width=40.0, length=200.0, capacity=8, dt=1e-3, tau=0.01, jitter=0.0,
                                        k_harm=0.0, kT=0.0)
    configs.set_law(p, abi.LAW_NONE, 1.0)
    p.nwalls = 0
    s = col.species[0]
    u = np.array([0.6, 0.8])
    s["pos"][0, 1] = s["pos"][0, 0] + 0.5 * u
    orc = O.Oracle(p, box, traversal=1)
    orc.forces(col, 0)
    b0, f0 = s["pos"][0, 0].copy(), s["pos"][0, 1].copy()
    # rate 100 per time unit, dt 1e-3: l grows 0.1 per step from 0.5, reaching 2.0 at step 15
    ts = orc.step(col, 14, 0)
    assert col.ncells() == [1] and s["rest_len"][0] == pytest.approx(0.5 + 14 * 0.1, rel=1e-12)
    assert np.allclose(s["pos"][0, 0], b0) and np.allclose(s["pos"][0, 1], f0)
    orc.step(col, 1, ts)
    assert col.ncells() == [2]
    assert s["rest_len"][0] == s["rest_len"][1] == 0.5
    # back daughter keeps the back disc, front daughter the front disc; each spans l_birth along u
    assert np.allclose(s["pos"][0, 0], b0, atol=1e-14) and np.allclose(s["pos"][0, 1], b0 + 0.5 * u, atol=1e-14)
    assert np.allclose(s["pos"][1, 1], f0, atol=1e-14) and np.allclose(s["pos"][1, 0], f0 - 0.5 * u, atol=1e-14)
    assert s["t0"][0] == s["t0"][1] == 15


def test_drain_removal_set_properties():
    rng = np.random.default_rng(8)
    n = 300
    p = configs.base_params(2, 1, 1)
    configs.set_law(p, abi.LAW_NONE, 1.0)
    p.drain = 1
    box = configs.make_box([10.0, 10.0], [0, 0])
    col = Colony(2, 1, [n], ncells=[n])
    s = col.species[0]
    s["pos"][:, 0] = rng.uniform(-2.0, 12.0, (n, 2))
    s["t0"][:] = np.arange(n)  # cell identity
    inside = lambda x: np.all((x >= 0.0) & (x < 10.0), axis=-1)
    before_out = set(np.flatnonzero(~inside(s["pos"][:, 0])).tolist())
    removed = O.Oracle(p, box).check_positions(col)
    m = col.ncells()[0]
    ids = s["t0"][:m]
    assert removed == n - m
    gone = set(range(n)) - set(ids.tolist())
    assert gone <= before_out  # only cells outside were removed
    still_out = set(ids[~inside(s["pos"][:m, 0])].tolist())
    # the skip after a swap: every survivor outside the box is one that moved into a freed slot
    moved = {int(i) for k, i in enumerate(ids) if i != k}
    assert still_out <= moved
    assert len(still_out) > 0 and len(gone) > 0
    # repeated calls remove them (each call may again skip cells swapped into freed slots)
    orc = O.Oracle(p, box)
    for _ in range(len(still_out) + 1):
        if orc.check_positions(col) == 0:
            break
    assert np.all(inside(s["pos"][:col.ncells()[0], 0]))
    assert col.ncells()[0] == n - len(before_out)


# ------------------------------------------------------------------------------------------
# calculDivisions, updateDiameters, StochasticHeun
# ------------------------------------------------------------------------------------------
def _binary_cell(state, tswitch, r=1.0, r2min=1.2, r0=0.3, swelling=0):
    p = configs.base_params(2, 1, 2)
    sp = p.species[0]
    sp.spring = abi.SPRING_HARMONIC
    sp.k_harm = 0.0
    sp.sigma[0], sp.sigma[1] = 1.0, 1.4
    sp.division_r2min, sp.division_r0 = r2min, r0
    sp.division_swelling = swelling
    sp.division_tau.mu, sp.division_tau.sigma = 0.05, 0.0
    sp.migration_tau.mu, sp.migration_tau.sigma = 0.02, 0.0
    p.division = abi.DIV_BINARY
    p.dt = 1e-3
    configs.set_law(p, abi.LAW_NONE, 1.0)
    box = configs.make_box([30.0, 30.0], [1, 1])
    col = Colony(2, 2, [4], ncells=[1])
    s = col.species[0]
    s["pos"][0, 0] = [10.0, 10.0]
    s["pos"][0, 1] = [10.0 + 0.6 * r, 10.0 + 0.8 * r]
    s["sigma"][0] = [0.8, 0.9]
    s["state"][0] = state
    s["tswitch"][0] = tswitch
    s["rest_len"][0] = r
    return p, box, col


def test_binary_division_cycle_timing_and_geometry():
    # crawling until tswitch = 3, then dividing for tau_division / dt = 50 steps
    p, box, col = _binary_cell(abi.CS_CRAWLING, 3, r=1.0, r2min=1.2)
    s = col.species[0]
    orc = O.Oracle(p, box, traversal=1)
    orc.forces(col, 0)
    ts = orc.step(col, 2, 0)
    assert s["state"][0] == abi.CS_CRAWLING
    ts = orc.step(col, 1, ts)
    assert s["state"][0] == abi.CS_DIVIDING and s["tswitch"][0] == 53
    # r^2 = 1 < Rmin^2 = 1.2: no division; at tswitch it fails: back disc to the midpoint,
    # discs re-placed +-R0 around it, Crawling, next switch 20 steps later, species diameters
    mid = 0.5 * (s["pos"][0, 0] + s["pos"][0, 1])
    ts = orc.step(col, 50, ts)
    assert ts == 53 and col.ncells() == [1]
    assert s["state"][0] == abi.CS_CRAWLING and s["tswitch"][0] == 73
    assert np.allclose(0.5 * (s["pos"][0, 0] + s["pos"][0, 1]), mid, atol=1e-13)
    assert np.linalg.norm(s["pos"][0, 1] - s["pos"][0, 0]) == pytest.approx(0.6, rel=1e-13)
    assert np.array_equal(s["sigma"][0], [1.0, 1.4])


def test_binary_division_success_geometry():
    p, box, col = _binary_cell(abi.CS_DIVIDING, 40, r=1.2, r2min=1.2)
    s = col.species[0]
    b, f = s["pos"][0, 0].copy(), s["pos"][0, 1].copy()
    orc = O.Oracle(p, box, traversal=1)
    orc.forces(col, 0)
    orc.step(col, 1, 0)
    assert col.ncells() == [2]
    # mother: discs +-R0 around its old front disc; daughter: +-R0 around the old back disc
    assert np.allclose(0.5 * (s["pos"][0, 0] + s["pos"][0, 1]), f, atol=1e-13)
    assert np.allclose(0.5 * (s["pos"][1, 0] + s["pos"][1, 1]), b, atol=1e-13)
    for c in range(2):
        assert np.linalg.norm(s["pos"][c, 1] - s["pos"][c, 0]) == pytest.approx(0.6, rel=1e-13)
        assert s["state"][c] == abi.CS_CRAWLING and s["t0"][c] == 1
        assert np.array_equal(s["sigma"][c], [1.0, 1.4])
    assert s["tswitch"][1] == 1 + 20 and s["tswitch"][0] == 40 + 20


def test_swelling_is_linear_ramp_to_max_before_switch_time():
    p, box, col = _binary_cell(abi.CS_DIVIDING, 41, r=0.5, r2min=100.0, swelling=1)
    s = col.species[0]
    orc = O.Oracle(p, box, traversal=1)
    orc.forces(col, 0)
    ts = 0
    for k in range(1, 41):
        ts = orc.step(col, 1, ts)
        # the deficit shrinks by (T - t - 1) / (T - t) at step t: a linear ramp
        # sigma_e(t) = max - (max - sigma_e(0)) (T - 1 - t) / (T - 1), max 1.4, T = 41, which
        # reaches the maximum one step before the switch time
        for e, s0 in enumerate((0.8, 0.9)):
            assert s["sigma"][0, e] == pytest.approx(1.4 - (1.4 - s0) * (40 - k) / 40, rel=1e-12, abs=1e-14)
    assert s["state"][0] == abi.CS_DIVIDING


def _heun(p):
    q = abi.Params.from_buffer_copy(p)
    q.integrator = abi.INT_HEUN
    return q


def test_heun_bond_relaxation_factor():
    # harmonic bond e = r - l: one Heun step multiplies e by (1 - a/2)^2, a = 2 k dt / zeta
    # (updatePositions uses 1/zeta, no area factor)
    p = configs.base_params(2, 1, 2)
    sp = p.species[0]
    sp.spring, sp.k_harm, sp.zeta[0], sp.zeta[1] = abi.SPRING_HARMONIC, 10.0, 2.0, 2.0
    configs.set_law(p, abi.LAW_NONE, 1.0)
    p.dt = 1e-2
    box = configs.make_box([30.0, 30.0], [1, 1])
    col = Colony(2, 2, [1], ncells=[1])
    s = col.species[0]
    s["pos"][0, 0] = [10.0, 10.0]
    s["pos"][0, 1] = [10.9, 10.0]
    s["rest_len"][0] = 0.5
    orc = O.Oracle(_heun(p), box, traversal=1)
    orc.forces(col, 0)
    orc.step(col, 20, 0)
    a = 2 * 10.0 * 1e-2 / 2.0
    e = s["pos"][0, 1, 0] - s["pos"][0, 0, 0] - 0.5
    assert e == pytest.approx(0.4 * (1 - a / 2) ** 40, rel=1e-11)
    assert np.all(s["vel"][0] == 0.0)


def test_heun_motility_and_noise_variance():
    n = 4000
    p = configs.base_params(2, 1, 1)
    sp = p.species[0]
    sp.zeta[0], sp.motility[0] = 2.0, 0.3
    D = 0.05
    p.dt = 1e-3
    sp.d_noise = math.sqrt(2 * D) / math.sqrt(p.dt)
    p.noise = abi.NOISE_CELL_WIENER
    configs.set_law(p, abi.LAW_NONE, 1.0)
    box = configs.make_box([1000.0, 1000.0], [1, 1])
    col = Colony(2, 1, [n], ncells=[n])
    s = col.species[0]
    rng = np.random.default_rng(1)
    s["pos"][:, 0] = rng.uniform(100, 900, (n, 2))
    s["vel"][:, 0] = [1.0, 0.0]
    x0 = s["pos"][:, 0].copy()
    orc = O.Oracle(_heun(p), box, traversal=1)
    orc.forces(col, 0)
    orc.step(col, 20, 0)
    dx = s["pos"][:, 0] - x0
    # motility moves every disc by dt m / zeta per step along its (unchanged) velocity;
    # the noise force zeta W enters both half steps with independent draws: variance D dt
    assert np.mean(dx[:, 0]) == pytest.approx(20 * 1e-3 * 0.3 / 2.0, abs=4 * math.sqrt(20 * D * 1e-3 / n))
    assert np.var(dx[:, 1]) == pytest.approx(20 * D * 1e-3, rel=0.06)
