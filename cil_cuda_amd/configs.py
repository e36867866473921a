"""Workload builders: parameters and seeded synthetic initial colonies.

Each builder returns (params, box, colony).  Parameter names follow the reference's
CellParams after readSpecs' derivations (Input.cpp:6-293); where a workload needs something
the reference lacks (harmonic bond, FDT noise, per-pair law matrix, linear growth) the
extension is named in include/cil_hip.h.

  plain_2d_sc      2D periodic, two-disc cells, soft-core + harmonic bond + motility, FDT
                   noise (the 2000-cell, L=64 workload and smaller test variants)
  channel_2d_hz    2D growing colony in a 40-wide walled channel, Hertz contacts, linear
                   growth and division, drain at the open ends
  bulk_3d_mixed    3D periodic bulk of single discs, lognormal diameters at packing fraction
                   0.6, LJ (2.5 sigma) between species-0 pairs and soft core otherwise
  theta_2d_hz      single-disc Hertz colony with the reference's theta cell cycle
  binary_2d_fene   two-disc cells with the reference's FENE / push-spring bond, soft-core
                   contacts, per-cell Wiener noise and the reference's two-disc division cycle
                   (calculDivisions: Crawling <-> Dividing, growth activity, swelling), for the
                   Euler or the stochastic Heun integrator
"""
from __future__ import annotations

import math

import numpy as np

from . import abi
from .colony import Colony


def base_params(dim: int, nspecies: int, nelem: int) -> abi.Params:
    p = abi.Params()
    p.abi_version = abi.ABI_VERSION
    p.dim = dim
    p.nspecies = nspecies
    for s in range(abi.MAX_SPECIES):
        sp = p.species[s]
        sp.nelem = nelem
        for e in range(abi.MAX_ELEMS):
            sp.sigma[e] = 1.0
            sp.zeta[e] = 1.0
        sp.cross_adhesion = 1.0
        sp.cycle_p0 = 1.0
    for i in range(abi.MAX_TYPES):
        for j in range(abi.MAX_TYPES):
            p.rcut2[i][j] = 1.0
            p.law[i][j] = abi.LAW_NONE
    p.dt = 1e-4
    p.seed = 1
    p.precision = abi.PREC_FP64
    return p


def set_law(p: abi.Params, law: int, rcut2: float = 1.0, types=None):
    nt = p.nspecies * p.species[0].nelem
    rng = range(nt) if types is None else types
    for i in rng:
        for j in rng:
            p.law[i][j] = law
            p.rcut2[i][j] = rcut2


def make_box(lengths, periodic) -> abi.Box:
    b = abi.Box()
    for d in range(3):
        b.length[d] = lengths[d] if d < len(lengths) else 1.0
        b.periodic[d] = int(periodic[d]) if d < len(periodic) else 0
    return b


def _min_image(dx, L):
    return dx - L * np.round(dx / L)


def poisson_centres(n, L, dmin, rng, dim=2, max_tries=200):
    """Uniform random centres in a periodic box, rejecting any closer than dmin."""
    L = np.asarray(L, dtype=float)
    cell = dmin
    ng = np.maximum((L / cell).astype(int), 1)
    grid = {}
    pts = np.zeros((n, dim))
    k = 0
    tries = 0
    while k < n:
        x = rng.uniform(0, 1, dim) * L
        tries += 1
        if tries > max_tries * n:
            raise RuntimeError("rejection sampling failed; lower the density")
        g = tuple((x / (L / ng)).astype(int) % ng)
        ok = True
        for off in np.ndindex(*([3] * dim)):
            gg = tuple((np.array(g) + np.array(off) - 1) % ng)
            for j in grid.get(gg, ()):
                if np.sum(_min_image(pts[j] - x, L) ** 2) < dmin * dmin:
                    ok = False
                    break
            if not ok:
                break
        if ok:
            pts[k] = x
            grid.setdefault(g, []).append(k)
            k += 1
    return pts


# ------------------------------------------------------------------------------------------
# 2D periodic soft-core colony of two-disc cells (plain method, CPU-runnable workload)
# ------------------------------------------------------------------------------------------
def plain_2d_sc(n_cells=2000, L=64.0, seed=1, kT=0.01, dt=1e-4, motility=0.1, k_harm=10.0, r0=0.5,
                precision=abi.PREC_FP64, noise=abi.NOISE_FDT, capacity=None):
    rng = np.random.default_rng(seed)
    p = base_params(2, 1, 2)
    sp = p.species[0]
    sp.spring = abi.SPRING_HARMONIC
    sp.k_harm = k_harm
    sp.motility[0] = 0.0
    sp.motility[1] = motility
    for e in range(2):
        sp.eps_core[e] = 1.0      # EPSILON_CORE
        sp.eps_well[e] = 0.0      # EPSILON_RATIO 0: repulsive only
        sp.sig_well[e] = 1.0      # SIGMA_RATIO 1 x SIGMA_CORE 1
    sp.kT = kT
    set_law(p, abi.LAW_SC, 1.0)
    p.dt = dt
    p.noise = noise
    p.seed = seed
    p.precision = precision
    box = make_box([L, L], [1, 1])
    cap = capacity or n_cells
    col = Colony(2, 2, [cap], ncells=[n_cells])
    c = poisson_centres(n_cells, [L, L], 0.8, rng)
    ang = rng.uniform(0, 2 * math.pi, n_cells)
    u = np.stack([np.cos(ang), np.sin(ang)], axis=1)
    s = col.species[0]
    s["pos"][:n_cells, 0] = np.mod(c - 0.5 * r0 * u, L)
    s["pos"][:n_cells, 1] = np.mod(c + 0.5 * r0 * u, L)
    s["sigma"][:n_cells] = rng.uniform(0.9, 1.1, (n_cells, 2))
    s["rest_len"][:n_cells] = r0
    s["state"][:n_cells] = abi.CS_CRAWLING
    return p, box, col


# ------------------------------------------------------------------------------------------
# 3D periodic mixed LJ / soft-core bulk of single discs
# ------------------------------------------------------------------------------------------
def bulk_3d_box_length(n, phi=0.6, cv=0.1):
    m3 = (1.0 + cv * cv) ** 3  # <d^3> of a lognormal with mean 1 and coefficient of variation cv
    return (n * math.pi * m3 / (6.0 * phi)) ** (1.0 / 3.0)


def bulk_3d_mixed(n=1 << 20, phi=0.6, seed=1, kT=0.05, dt=1e-4, precision=abi.PREC_FP32, sort_spatially=True):
    rng = np.random.default_rng(seed)
    L = bulk_3d_box_length(n, phi)
    p = base_params(3, 2, 1)
    for s in range(2):
        sp = p.species[s]
        sp.eps_core[0] = 1.0     # soft core E0
        sp.eps_well[0] = 1.0 if s == 0 else 0.0   # species 0: LJ epsilon; species 1: no SC well
        sp.sig_well[0] = 1.0     # SIGMA_RATIO 1
        sp.kT = kT
        sp.motility[0] = 0.0
    # law matrix: LJ between two species-0 discs (RCUT 2.5 sigma_ij), soft core otherwise
    set_law(p, abi.LAW_SC, 1.0)
    p.law[0][0] = abi.LAW_LJ
    p.rcut2[0][0] = 2.5 ** 2
    # the SC entries involving species 0 use its eps_well as Ew = sqrt(Ew_i Ew_j) -> 0 with species 1
    p.dt = dt
    p.noise = abi.NOISE_FDT
    p.seed = seed
    p.precision = precision
    box = make_box([L, L, L], [1, 1, 1])
    mu_ln = -0.5 * math.log(1.01)
    s_ln = math.sqrt(math.log(1.01))
    x = rng.uniform(0, L, (n, 3))
    sig = np.exp(rng.normal(mu_ln, s_ln, n))
    spec = (rng.uniform(0, 1, n) < 0.5).astype(int)
    if sort_spatially:
        # order particles along a coarse spatial grid so memory order follows space
        g = np.floor(x / (L / 32)).astype(np.int64)
        key = (g[:, 0] * 32 + g[:, 1]) * 32 + g[:, 2]
        o = np.argsort(key, kind="stable")
        x, sig, spec = x[o], sig[o], spec[o]
    n0 = int(np.sum(spec == 0))
    col = Colony(3, 1, [n0, n - n0], ncells=[n0, n - n0])
    for s, m in ((0, spec == 0), (1, spec == 1)):
        sp = col.species[s]
        sp["pos"][:, 0] = x[m]
        sp["sigma"][:, 0] = sig[m]
        sp["state"][:] = abi.CS_UNDEFINED
    return p, box, col


def relaxation_params(p: abi.Params, dt=1e-3, max_disp=0.05) -> abi.Params:
    """Relaxation of random initial positions: the production laws, a 10x larger time step,
    no noise, and every displacement capped at `max_disp` sigma per step (the restated laws all
    vanish or diverge at full overlap, so an uncapped push-off cannot separate coincident
    discs).  Used for the 100 relaxation steps of the 3D bulk."""
    q = abi.Params.from_buffer_copy(p)
    q.dt = dt
    q.noise = abi.NOISE_NONE
    q.max_disp = max_disp
    return q


def relaxed_bulk_3d(n, seed=1, steps=100, **kw):
    """The 3D mixed bulk of `n` particles relaxed on the GPU with the capped relaxation above
    (`steps` steps); returns production (params, box, colony)."""
    from .colony import HipMD

    p, box, col = bulk_3d_mixed(n=n, seed=seed, **kw)
    q = relaxation_params(p)
    md = HipMD(q, box, [sp["capacity"] for sp in col.species])
    try:
        md.upload(col)
        md.forces(want=False)
        md.step(steps, 0)
        md.download(col)
    finally:
        md.close()
    return p, box, col


# ------------------------------------------------------------------------------------------
# 2D channel: growing two-disc colony with Hertz contacts, walls and drain
# ------------------------------------------------------------------------------------------
def channel_2d_hz(width=40.0, length=4096.0, seed=1, kT=0.001, dt=5e-5, tau=1.0, E=1000.0, nu=0.5,
                  k_harm=100.0, motility=0.0, capacity=65536, l_birth=0.5, l_div=2.0, jitter=0.05,
                  precision=abi.PREC_FP64):
    p = base_params(2, 1, 2)
    sp = p.species[0]
    sp.spring = abi.SPRING_HARMONIC
    sp.k_harm = k_harm
    sp.motility[1] = motility
    for e in range(2):
        sp.eps_core[e] = E       # ELASTIC_MODULUS
        sp.sig_well[e] = nu      # POISSON_RATIO
        sp.eps_well[e] = 0.0     # no adhesion
    sp.kT = kT
    sp.growth_rate = 1.0 / tau
    sp.l_div = l_div
    sp.l_birth = l_birth
    sp.sigma_jitter = jitter
    set_law(p, abi.LAW_HZ, 1.0)
    p.dt = dt
    p.noise = abi.NOISE_FDT
    p.seed = seed
    p.division = abi.DIV_LINEAR_GROWTH
    p.drain = 1
    p.precision = precision
    rc = 2.0 ** (1.0 / 6.0)
    p.nwalls = 2
    for k, (x0, pos) in enumerate(((0.0, 1), (width, 0))):
        w = p.walls[k]
        w.kind = 0
        w.axis = 0
        w.x0 = x0
        w.rcut = rc
        w.epsilon = 1.0
        w.positive = pos
    box = make_box([width, length], [0, 0])
    col = Colony(2, 2, [capacity], ncells=[1])
    s = col.species[0]
    cx, cy = width / 2, length / 2
    s["pos"][0, 0] = [cx, cy - 0.5 * l_birth]
    s["pos"][0, 1] = [cx, cy + 0.5 * l_birth]
    s["sigma"][0] = 1.0
    s["rest_len"][0] = l_birth
    s["state"][0] = abi.CS_CRAWLING
    return p, box, col


# ------------------------------------------------------------------------------------------
# theta cell cycle (reference division model) on a small single-disc Hertz colony
# ------------------------------------------------------------------------------------------
def theta_2d_hz(n_cells=16, L=40.0, seed=3, dt=1e-3, capacity=256, precision=abi.PREC_FP64):
    rng = np.random.default_rng(seed)
    p = base_params(2, 1, 1)
    sp = p.species[0]
    sp.eps_core[0] = 1000.0
    sp.sig_well[0] = 0.5
    sp.eps_well[0] = 0.0
    sp.motility[0] = 0.1
    sp.cycle_p0 = 1.0
    sp.cycle_tau_r = 0.5 / dt
    sp.cycle_tau_v = 0.5 / dt
    sp.cycle_tau_p = 0.1 / dt
    sp.cycle_v_s = 1.0
    sp.cycle_v_1 = 2.0
    sp.cycle_beta_vs = 0.1
    sp.cycle_rmaxed_inv = 1.0
    sp.cycle_rho_ratio = 1.0
    sp.apop_pressure = 1e9
    sp.division_tau.mu = 0.02
    sp.division_tau.sigma = 0.002
    sp.migration_tau.mu = 0.0
    sp.migration_tau.sigma = 0.0
    set_law(p, abi.LAW_HZ, 1.0)
    p.dt = dt
    p.noise = abi.NOISE_NONE
    p.seed = seed
    p.division = abi.DIV_THETA_CYCLE
    p.pressure = 1
    p.precision = precision
    box = make_box([L, L], [1, 1])
    col = Colony(2, 1, [capacity], ncells=[n_cells])
    s = col.species[0]
    c = poisson_centres(n_cells, [L, L], 1.5, rng)
    s["pos"][:n_cells, 0] = c
    s["sigma"][:n_cells, 0] = 1.0
    s["sigma_t"][:n_cells] = 1.0
    s["state"][:n_cells] = abi.CS_DIVIDING
    s["rho"][:n_cells] = 1.0
    s["theta"][:n_cells] = rng.uniform(0, 1, n_cells)
    s["ang_vel"][:n_cells] = 1.0 / np.maximum(np.rint(rng.normal(0.02, 0.002, n_cells) / dt), 1)
    return p, box, col


# ------------------------------------------------------------------------------------------
# two-disc division cycle (calculDivisions, CIL.cu:369-455) with the reference's FENE bond
# ------------------------------------------------------------------------------------------
def binary_2d_fene(n_cells=150, L=20.0, seed=9, dt=2e-3, capacity=1024, integrator=abi.INT_EULER,
                   swelling=1, D=0.005, precision=abi.PREC_FP64):
    """Parameter names follow Input.cpp:117-245: FENE KAPPA / KAPPA_PUSH / SPRING_RATIO / RMAX
    (r2max = RMAX^2) and RMAX_DIVISION, DIVISION R0 / Rmin / TAU_DIVISION / TAU_MIGRATION /
    SWELLING, GROWTHK0..2 (the Dividing-state activity).  Crawling cells pull at their front disc
    (rss-scaled motility); a Dividing cell's growth activity (growthk1 > kappa) stretches its
    bond until |r_1 - r_0| >= Rmin and it divides into two cells of discs +-R0 apart."""
    rng = np.random.default_rng(seed)
    p = base_params(2, 1, 2)
    sp = p.species[0]
    sp.sigma[0], sp.sigma[1] = 0.9, 1.0
    sp.spring = abi.SPRING_FENE
    sp.kappa, sp.kappa_push, sp.spring_ratio = 1.0, 5.0, 0.3
    sp.r2max, sp.div_r2max = 2.0, 2.5
    sp.rss = 0.6
    sp.motility[0], sp.motility[1] = 0.02, 0.05
    for e in range(2):
        sp.eps_core[e] = 1.0     # SC EPSILON_CORE
        sp.eps_well[e] = 0.0     # no attraction
        sp.sig_well[e] = 0.3     # SIGMA_RATIO * SIGMA_CORE
        sp.growthk1[e] = 2.0
    sp.division_r0 = 0.3
    sp.division_r2min = 1.2 ** 2
    sp.division_swelling = swelling
    sp.division_tau.mu, sp.division_tau.sigma = 1.0, 0.2
    sp.migration_tau.mu, sp.migration_tau.sigma = 0.2, 0.05
    sp.d_noise = math.sqrt(2 * D) / math.sqrt(dt)
    set_law(p, abi.LAW_SC, 1.0)
    p.dt = dt
    p.noise = abi.NOISE_CELL_WIENER if D > 0 else abi.NOISE_NONE
    p.seed = seed
    p.division = abi.DIV_BINARY
    p.integrator = integrator
    p.precision = precision
    box = make_box([L, L], [1, 1])
    col = Colony(2, 2, [capacity], ncells=[n_cells])
    s = col.species[0]
    c = poisson_centres(n_cells, [L, L], 1.2, rng)
    ang = rng.uniform(0, 2 * math.pi, n_cells)
    u = np.stack([np.cos(ang), np.sin(ang)], axis=1)
    s["pos"][:n_cells, 0] = np.mod(c - 0.3 * u, L)
    s["pos"][:n_cells, 1] = np.mod(c + 0.3 * u, L)
    s["sigma"][:n_cells] = [0.85, 0.95]
    s["state"][:n_cells] = abi.CS_CRAWLING
    s["tswitch"][:n_cells] = rng.integers(1, 150, n_cells)
    return p, box, col
