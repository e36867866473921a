"""Pins of the oracle's physics restatement against closed forms and invariants.

Each check is chosen so that a plausible mistake (dropped term, wrong sign, wrong index,
transposed operand) fails: pair forces against the derivative of the textbook potential,
Hertz against (4/3) E* sqrt(R*) h^(3/2), link-list pairs against brute force, the Euler
update against exact solutions, and the noise against its variance.
"""
import math

import numpy as np
import pytest

from cil_cuda_amd import abi, configs
from cil_cuda_amd.colony import Colony
from oracle import oracle as O


def _lj_energy(r, s, eps):
    x = (s / r) ** 6
    return 4 * eps * (x * x - x)


def test_lj_force_is_minus_gradient_of_textbook_potential():
    p = configs.base_params(3, 1, 1)
    p.species[0].eps_well[0] = 1.7
    configs.set_law(p, abi.LAW_LJ, 2.5 ** 2)
    for si, sj in ((1.0, 1.0), (0.9, 1.2)):
        s = 0.5 * (si + sj)
        for r in np.linspace(0.85 * s, 2.45 * s, 13):
            hit, f0 = O.pair_f0(p, 3, 0, 0, si, sj, r * r)
            assert hit
            h = 1e-6 * r
            dphi = (_lj_energy(r + h, s, 1.7) - _lj_energy(r - h, s, 1.7)) / (2 * h)
            assert f0 == pytest.approx(dphi / r, rel=1e-6, abs=1e-9)
        # potential minimum at 2^(1/6) sigma: zero force
        hit, f0 = O.pair_f0(p, 3, 0, 0, si, sj, (2 ** (1 / 6) * s) ** 2)
        assert abs(f0) < 1e-12
        # cut-off predicate r^2 <= RCUT^2 sigma_ij^2 (MD.cuh:100-101)
        assert O.pair_f0(p, 3, 0, 0, si, sj, (2.5 * s) ** 2 * (1 - 1e-12))[0]
        assert not O.pair_f0(p, 3, 0, 0, si, sj, (2.5 * s) ** 2 * (1 + 1e-12))[0]


def _sc_energy(r, sig, xi, e0, ew):
    # soft-core polynomial potential: cubic smoothstep wells of width xi around sigma
    y = r - sig
    t = lambda z: z * z * (3 * xi - 2 * z) / xi ** 3
    if y <= -xi:
        return e0
    if y < 0:
        return (e0 + ew) * t(-y) - ew
    if y < xi:
        return ew * t(y) - ew
    return 0.0


def test_sc_force_is_minus_gradient_of_potential():
    p = configs.base_params(2, 1, 1)
    sp = p.species[0]
    sp.eps_core[0] = 2.0
    sp.eps_well[0] = 0.3
    sp.sig_well[0] = 0.4
    configs.set_law(p, abi.LAW_SC, 1.0)
    si = sj = 1.0
    for r in np.linspace(0.62, 1.38, 31):
        if abs(r - 1.0) < 1e-3:
            continue
        hit, f0 = O.pair_f0(p, 2, 0, 0, si, sj, r * r)
        assert hit
        h = 1e-7
        dphi = (_sc_energy(r + h, 1.0, 0.4, 2.0, 0.3) - _sc_energy(r - h, 1.0, 0.4, 2.0, 0.3)) / (2 * h)
        assert f0 == pytest.approx(dphi / r, rel=1e-5, abs=1e-7)
    # outside the well (r > sigma + xi) and inside the core (r < sigma - xi): no interaction
    assert not O.pair_f0(p, 2, 0, 0, si, sj, 1.41 ** 2)[0]
    assert not O.pair_f0(p, 2, 0, 0, si, sj, 0.59 ** 2)[0]


def test_hertz_matches_textbook_contact_law():
    p = configs.base_params(2, 1, 1)
    sp = p.species[0]
    sp.eps_core[0] = 1000.0
    sp.sig_well[0] = 0.3
    configs.set_law(p, abi.LAW_HZ, 1.0)
    for si, sj in ((1.0, 1.0), (0.8, 1.3)):
        Ri, Rj = si / 2, sj / 2
        Estar = 1.0 / (2 * (1 - 0.3 ** 2) / 1000.0)
        Rstar = 1.0 / (1 / Ri + 1 / Rj)
        for h in (1e-4, 1e-2, 0.1, 0.3):
            r = 0.5 * (si + sj) - h
            hit, f0 = O.pair_f0(p, 2, 0, 0, si, sj, r * r)
            assert hit
            textbook = (4.0 / 3.0) * Estar * math.sqrt(Rstar) * h ** 1.5
            assert -f0 == pytest.approx(textbook, rel=1e-12)
        assert not O.pair_f0(p, 2, 0, 0, si, sj, (0.5 * (si + sj) + 1e-9) ** 2)[0]


def _brute_pairs(col, params, box):
    """Independent numpy enumeration of interacting element pairs (UID order)."""
    D = col.dim
    pos = col.flat("pos").reshape(-1, D)
    sig = col.flat("sigma").reshape(-1)
    nE = col.nelem
    spec = np.concatenate([np.full(sp["ncells"] * nE, k) for k, sp in enumerate(col.species)])
    typ = spec * nE + np.tile(np.arange(nE), len(pos) // nE)
    cell = np.repeat(np.arange(len(pos) // nE), nE)
    L = np.array([box.length[d] for d in range(D)])
    per = np.array([box.periodic[d] for d in range(D)], dtype=bool)
    out = []
    for i in range(len(pos)):
        dx = pos[i + 1:] - pos[i]
        for d in range(D):
            if per[d]:
                n = dx[:, d] / L[d]
                dx[:, d] -= np.where(n >= 0, np.floor(n + 0.5), -np.floor(-n + 0.5)) * L[d]
        r2 = np.sum(dx * dx, axis=1)
        js = np.arange(i + 1, len(pos))
        for j, rr in zip(js, r2):
            if cell[j] == cell[i]:
                continue
            law = params.law[typ[i]][typ[j]]
            sb = 0.5 * (sig[i] + sig[j])
            rc2 = params.rcut2[typ[i]][typ[j]]
            if law == abi.LAW_LJ:
                hit = rr <= sb * sb * rc2
            elif law == abi.LAW_SC:
                si_ = params.species[spec[i]].sig_well[typ[i] % nE]
                sj_ = params.species[spec[j]].sig_well[typ[j] % nE]
                xi = 0.5 * (si_ + sj_)
                hit = rr > rc2 * (sb - xi) ** 2 and rr <= rc2 * (sb + xi) ** 2
            elif law in (abi.LAW_HZ, abi.LAW_EA):
                hit = math.sqrt(rr) <= (sig[i] + sig[j]) / 2
            else:
                hit = False
            if hit:
                out.append((i, j))
    return np.array(out, dtype=np.int32).reshape(-1, 2)


def _random_colony(dim, nelem, n, L, periodic, law, seed, nspec=1):
    rng = np.random.default_rng(seed)
    p = configs.base_params(dim, nspec, nelem)
    for s in range(nspec):
        for e in range(nelem):
            p.species[s].sig_well[e] = 0.5 + 0.2 * s
            p.species[s].eps_core[e] = 10.0
            p.species[s].eps_well[e] = 1.0
    configs.set_law(p, law, 2.0 if law == abi.LAW_LJ else 1.0)
    box = configs.make_box([L] * dim, [periodic] * dim)
    per = n // nspec
    col = Colony(dim, nelem, [per] * nspec, ncells=[per] * nspec)
    for sp in col.species:
        c = rng.uniform(0, L, (per, dim))
        sp["pos"][:, 0] = c
        if nelem == 2:
            u = rng.normal(size=(per, dim))
            u /= np.linalg.norm(u, axis=1, keepdims=True)
            sp["pos"][:, 1] = np.mod(c + 0.6 * u, L) if periodic else c + 0.6 * u
        sp["sigma"][:] = rng.uniform(0.8, 1.25, (per, nelem))
    return p, box, col


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("periodic", [1, 0])
@pytest.mark.parametrize("nelem", [1, 2])
@pytest.mark.parametrize("law", [abi.LAW_LJ, abi.LAW_SC, abi.LAW_HZ])
def test_linklist_pairs_equal_brute_force(dim, periodic, nelem, law):
    n = 160 if dim == 2 else 220
    L = 12.0 if dim == 2 else 7.5
    p, box, col = _random_colony(dim, nelem, n, L, periodic, law, seed=dim * 10 + nelem + law, nspec=2)
    orc = O.Oracle(p, box)
    ll = orc.pairs(col, mode=0)
    allp = orc.pairs(col, mode=1)
    ns, width = orc.linklist_info()
    bf = _brute_pairs(col, p, box)
    assert len(bf) > 0
    assert np.array_equal(allp, bf)
    assert np.array_equal(ll, bf)


def test_linklist_is_used_on_large_box():
    p, box, col = _random_colony(2, 1, 400, 30.0, 1, abi.LAW_SC, seed=5)
    orc = O.Oracle(p, box)
    ll = orc.pairs(col, mode=0)
    ns, width = orc.linklist_info()
    assert ns[0] >= 3 and ns[1] >= 3
    assert np.array_equal(ll, _brute_pairs(col, p, box))


def test_euler_update_moves_single_disc_by_mobility_times_motility():
    # single-disc cell, motility m along its velocity, no neighbours: dx = dt*m/(zeta*pi*(sigma/2)^2)
    p = configs.base_params(2, 1, 1)
    sp = p.species[0]
    sp.motility[0] = 0.3
    sp.zeta[0] = 2.0
    configs.set_law(p, abi.LAW_SC, 1.0)
    p.dt = 1e-3
    box = configs.make_box([20.0, 20.0], [1, 1])
    col = Colony(2, 1, [1], ncells=[1])
    s = col.species[0]
    s["pos"][0, 0] = [5.0, 5.0]
    s["sigma"][0, 0] = 1.2
    s["vel"][0, 0] = [3.0, 4.0]
    orc = O.Oracle(p, box)
    orc.forces(col, 0)
    assert np.allclose(s["frc"][0, 0], [0.18, 0.24])
    ts = orc.step(col, 10, 0)
    assert ts == 10
    mu = 1.0 / (2.0 * math.pi * 0.6 ** 2)
    disp = 10 * 1e-3 * mu * 0.3
    assert np.allclose(s["pos"][0, 0], [5.0 + 0.6 * disp, 5.0 + 0.8 * disp], rtol=0, atol=1e-13)
    assert np.allclose(s["vel"][0, 0], np.array([0.6, 0.8]) * 0.3 * mu, rtol=1e-13)


def test_harmonic_bond_relaxes_geometrically():
    # one two-disc cell, no motility or noise: (r - l) shrinks by (1 - 2 k mu dt) per Euler step
    p = configs.base_params(2, 1, 2)
    sp = p.species[0]
    sp.spring = abi.SPRING_HARMONIC
    sp.k_harm = 10.0
    configs.set_law(p, abi.LAW_SC, 1.0)
    p.dt = 1e-3
    box = configs.make_box([30.0, 30.0], [1, 1])
    col = Colony(2, 2, [1], ncells=[1])
    s = col.species[0]
    s["pos"][0, 0] = [10.0, 10.0]
    s["pos"][0, 1] = [10.9, 10.0]
    s["sigma"][0] = [1.0, 1.0]
    s["rest_len"][0] = 0.5
    orc = O.Oracle(p, box)
    orc.forces(col, 0)
    orc.step(col, 50, 0)
    mu = 1.0 / (math.pi * 0.25)
    expect = 0.4 * (1 - 2 * 10.0 * mu * 1e-3) ** 50
    r = s["pos"][0, 1, 0] - s["pos"][0, 0, 0]
    assert r - 0.5 == pytest.approx(expect, rel=1e-10)
    # centre of mass fixed
    assert 0.5 * (s["pos"][0, 1, 0] + s["pos"][0, 0, 0]) == pytest.approx(10.45, abs=1e-13)


def test_fdt_noise_variance():
    # free discs (no contacts, no motility): per-component displacement variance 2 kT mu dt
    n = 4000
    p = configs.base_params(2, 1, 1)
    sp = p.species[0]
    sp.kT = 0.5
    configs.set_law(p, abi.LAW_NONE, 1.0)
    p.noise = abi.NOISE_FDT
    p.dt = 1e-3
    box = configs.make_box([1000.0, 1000.0], [1, 1])
    col = Colony(2, 1, [n], ncells=[n])
    rng = np.random.default_rng(0)
    col.species[0]["pos"][:, 0] = rng.uniform(100, 900, (n, 2))
    x0 = col.species[0]["pos"][:, 0].copy()
    orc = O.Oracle(p, box, traversal=1)
    orc.forces(col, 0)
    orc.step(col, 20, 0)
    dx = col.species[0]["pos"][:, 0] - x0
    mu = 1.0 / (math.pi * 0.25)
    var = 2 * 0.5 * mu * 1e-3 * 20
    assert np.var(dx) == pytest.approx(var, rel=0.05)
    assert abs(np.mean(dx)) < 4 * math.sqrt(var / dx.size)


def test_pair_forces_conserve_momentum():
    p, box, col = _random_colony(3, 1, 200, 7.0, 1, abi.LAW_LJ, seed=11, nspec=2)
    orc = O.Oracle(p, box, traversal=1)
    orc.forces(col, 0)
    f = col.flat("frc").reshape(-1, 3)
    scale = np.sum(np.abs(f))
    assert np.abs(f.sum(axis=0)).max() < 1e-12 * scale


def test_com_of_two_disc_cell_across_boundary():
    p = configs.base_params(2, 1, 2)
    box = configs.make_box([10.0, 10.0], [1, 1])
    col = Colony(2, 2, [1], ncells=[1])
    s = col.species[0]
    s["pos"][0, 0] = [9.8, 5.0]
    s["pos"][0, 1] = [0.4, 5.0]
    orc = O.Oracle(p, box)
    com = orc.coms(col)
    assert com[0] == pytest.approx([0.1, 5.0], abs=1e-12)


@pytest.mark.parametrize("dim,nelem", [(2, 2), (3, 1)])
def test_threaded_gather_matches_serial_traversal(dim, nelem):
    p, box, col = _random_colony(dim, nelem, 600 if dim == 2 else 1500, 30.0 if dim == 2 else 14.0, 1, abi.LAW_SC,
                                 seed=21, nspec=2)
    p.pressure = 1
    a, b = col.copy(), col.copy()
    O.Oracle(p, box, traversal=0, threads=1).forces(a, 0)
    orc = O.Oracle(p, box, traversal=0, threads=4)
    orc.forces(b, 0)
    assert orc.linklist_info()[0][0] >= 3
    fa, fb = a.flat("frc"), b.flat("frc")
    assert np.max(np.abs(fa - fb)) <= 1e-12 * np.max(np.abs(fa))
    assert np.allclose(a.flat("pressure"), b.flat("pressure"), rtol=1e-12, atol=1e-12)
