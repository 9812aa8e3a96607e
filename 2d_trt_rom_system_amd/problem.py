"""Problem descriptions for the full-order TRT iteration.

A :class:`Problem` holds everything the reference reads from its input deck
(MainSrc/INPUTS.f08): grid, quadrature, group structure, material map, time
step, boundary conditions and solver options, already converted to the
reference's code units (eV, sh, cm, erg/comp_unit).

Unit conventions (MainSrc/main.f08:18-22, MainSrc/INPUTS.f08:89-90,121,346):
    temperature/energies  eV          (1 keV = 1e3 eV)
    time                  shake       (1 ns = 0.1 sh)
    c                     299.792458 cm/sh
    heat capacity         erg/comp_unit per cm^3 per eV; with comp_unit = 1e13,
                          0.5 GJ/(cm^3 keV) = 0.5 code units
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace
from typing import List, Optional

import numpy as np

C_LIGHT = 2.99792458e2        # cm/sh (main.f08:19)
KAP0 = 27.0e9                 # eV^3/cm (main.f08:22)
COMP_UNIT = 1.0e13            # erg per code unit (INPUTS.f08:346)

# ---------------------------------------------------------------------------
# Level-symmetric quadrature (LQn), Lewis & Miller, "Computational Methods of
# Neutron Transport", Table 4-1: first direction cosine mu_1 and the point
# weights of each weight class (octant weights sum to 1).  Class membership is
# fixed by the multiset of level indices (i, j, k), i + j + k = N/2 + 2.
# ---------------------------------------------------------------------------
_LQN = {
    4: (0.3500212, {(1, 1, 2): 0.3333333}),
    8: (0.2182179, {(1, 1, 4): 0.1209877, (1, 2, 3): 0.0907407, (2, 2, 2): 0.0925926}),
    12: (0.1672126, {(1, 1, 6): 0.0707626, (1, 2, 5): 0.0558811, (1, 3, 4): 0.0373377,
                     (2, 2, 4): 0.0502819, (2, 3, 3): 0.0258513}),
    16: (0.1389568, {(1, 1, 8): 0.0489872, (1, 2, 7): 0.0413296, (1, 3, 6): 0.0212326,
                     (1, 4, 5): 0.0256207, (2, 2, 6): 0.0360486, (2, 3, 5): 0.0144589,
                     (2, 4, 4): 0.0344958, (3, 3, 4): 0.0085179}),
}


# ---------------------------------------------------------------------------
# Named quadrature sets of the reference's input deck (INPUT_QUAD,
# MainSrc/INPUTS.f08:1579-1674): first-quadrant direction cosines Omega_x and
# weights of the Abu-Shumays sets abu36 and abu20 (36 and 20 directions per
# quadrant) and of the 2-direction '1d_test' set.  In each set Omega_y is the
# Omega_x list reversed and the weights read the same backwards.  The other
# quadrants mirror the first (:1659-1668) and the weights are renormalised to
# 4*pi (:1669), as level_symmetric() does.
# ---------------------------------------------------------------------------
_NAMED_QUAD = {
    "abu36": (
        [9.717784813336e-1, 7.656319455497e-1, 4.445439440056e-1, 9.701698603928e-1, 7.633693960835e-1,
         1.483114568272e-1, 9.622473153642e-1, 7.524467626583e-1, 9.410672772109e-1, 4.288508824476e-1,
         7.241384940891e-1, 8.997294996538e-1, 6.711819639118e-1, 1.293388490485e-1, 8.335743322378e-1,
         3.670788892962e-1, 5.909368760506e-1, 7.417637460141e-1, 6.279014865859e-1, 4.869502395267e-1,
         2.519357740235e-1, 4.996274255819e-1, 7.447663982495e-2, 3.685670882907e-1, 3.673697853806e-1,
         2.488983098158e-1, 1.196054590036e-1, 2.426233944222e-1, 1.419535016004e-1, 1.362124657777e-1,
         1.670387759191e-2, 5.995074957044e-2, 5.695764868253e-2, 2.447911451942e-2, 1.160393058611e-2,
         1.096881837272e-2],
        [8.454511187252e-3, 8.352354145856e-3, 1.460888798152e-2, 1.913728513580e-2, 1.873220073879e-2,
         6.404244616724e-3, 2.863542971348e-2, 2.759429759588e-2, 3.648716160597e-2, 2.995376809966e-2,
         3.442681426024e-2, 4.244873302980e-2, 3.901232700510e-2, 1.162080754372e-2, 4.642823955812e-2,
         3.798783310581e-2, 4.130171453748e-2, 4.841339013884e-2, 4.841339013884e-2, 4.130171453748e-2,
         3.798783310581e-2, 4.642823955812e-2, 1.162080754372e-2, 3.901232700510e-2, 4.244873302980e-2,
         3.442681426024e-2, 2.995376809966e-2, 3.648716160597e-2, 2.759429759588e-2, 2.863542971348e-2,
         6.404244616724e-3, 1.873220073879e-2, 1.913728513580e-2, 1.460888798152e-2, 8.352354145856e-3,
         8.454511187252e-3]),
    "abu20": (
        [9.713274064903e-1, 7.645615896150e-1, 4.424202396002e-1, 9.586898685237e-1, 7.375714298063e-1,
         1.409476441875e-1, 9.028558915298e-1, 3.858240341629e-1, 6.313311043797e-1, 7.770210099715e-1,
         5.837054752370e-1, 4.332989313333e-1, 2.221674140412e-1, 3.596178122512e-1, 4.908227124734e-2,
         2.057068622698e-1, 1.593344524838e-1, 4.982847370367e-2, 4.210110375297e-2, 3.157215799340e-2],
        [2.419260514149e-2, 2.998205782366e-2, 2.932993043666e-2, 5.213067212540e-2, 6.147460425028e-2,
         1.802505216045e-2, 7.185542471164e-2, 5.322055875020e-2, 7.796304620960e-2, 8.182604839076e-2,
         8.182604839076e-2, 7.796304620960e-2, 5.322055875020e-2, 7.185542471164e-2, 1.802505216045e-2,
         6.147460425028e-2, 5.213067212540e-2, 2.932993043666e-2, 2.998205782366e-2, 2.419260514149e-2]),
    "1d_test": ([0.92388, 0.382683], [0.5, 0.5]),
}


def named_quadrature(name: str):
    """(Omega_x, Omega_y, quad_weight) of a named set of the input deck (quad key)."""
    if name not in _NAMED_QUAD:
        raise ValueError(f"unrecognized quadrature {name!r} (abu36, abu20, 1d_test)")
    ox, w = (np.array(v) for v in _NAMED_QUAD[name])
    oy = ox[::-1].copy()
    Ox = np.concatenate([ox, -ox, -ox, ox])
    Oy = np.concatenate([oy, oy, -oy, -oy])
    W = np.concatenate([w, w, w, w])
    return Ox, Oy, 4.0 * math.pi * W / W.sum()


# Group structures of the input deck (INPUT_ENERGY, MainSrc/INPUTS.f08:1517-1577): edges in eV.
GROUP_STRUCTURES = {
    "JCP": [0.0, 7.075e2, 1.415e3, 2.123e3, 2.830e3, 3.538e3, 4.245e3, 5.129e3, 6.014e3, 6.898e3, 7.783e3, 8.667e3,
            9.551e3, 1.044e4, 1.132e4, 1.220e4, 1.309e4, 1e10],
    "NE795": [0.0, 1.9047, 3.6278, 6.9097, 13.161, 25.067, 47.744, 90.937, 173.21, 329.90, 628.35, 1196.8, 2279.5,
              4341.7, 8269.5, 15751.0, 3e4, 1e10],
    "ONE_GROUP": [0.0, 1e10],
}


def level_symmetric(order: int):
    """Return (Omega_x, Omega_y, quad_weight) for a 2D level-symmetric S_N set.

    The N(N+2)/2 directions of the upper hemisphere are laid out in four
    contiguous quadrant blocks in the reference's order (INPUTS.f08:1658-1668):
    q1 (+,+), q2 (-,+), q3 (-,-), q4 (+,-); the weights are renormalised to sum
    to 4*pi (INPUTS.f08:1669).
    """
    if order not in _LQN:
        raise ValueError(f"unsupported level-symmetric order S{order}")
    mu1, classes = _LQN[order]
    n = order // 2
    d = 2.0 * (1.0 - 3.0 * mu1 * mu1) / (order - 2) if order > 2 else 0.0
    mu = [math.sqrt(mu1 * mu1 + i * d) for i in range(n)]
    ox, oy, w = [], [], []
    for i in range(1, n + 1):
        for j in range(1, n + 2 - i):
            k = n + 2 - i - j
            if k < 1:
                continue
            ox.append(mu[i - 1])
            oy.append(mu[j - 1])
            w.append(classes[tuple(sorted((i, j, k)))])
    ox = np.array(ox)
    oy = np.array(oy)
    w = np.array(w)
    Ox = np.concatenate([ox, -ox, -ox, ox])
    Oy = np.concatenate([oy, oy, -oy, -oy])
    W = np.concatenate([w, w, w, w])
    W = 4.0 * math.pi * W / W.sum()
    return Ox, Oy, W


def log_group_edges(ng: int, lo_kev: float, hi_kev: float) -> np.ndarray:
    """ng log-spaced groups between lo and hi (keV); edges returned in eV."""
    return 1.0e3 * np.logspace(math.log10(lo_kev), math.log10(hi_kev), ng + 1)


# ---------------------------------------------------------------------------
# Project RNG: SplitMix64, shared by every producer of seeded problem data so
# that maps and request streams are identical wherever they are generated.
# ---------------------------------------------------------------------------
class SplitMix64:
    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """Uniform double in (0, 1)."""
        return ((self.next_u64() >> 11) + 0.5) * (1.0 / 9007199254740992.0)

    def normal(self) -> float:
        u1, u2 = self.uniform(), self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def lognormal(self, mu: float, sigma: float) -> float:
        return math.exp(mu + sigma * self.normal())

    def choice(self, seq):
        return seq[self.next_u64() % len(seq)]


def region_map(nx: int, ny: int, nrx: int = 8, nry: int = 8):
    """Material id of each cell for an nrx x nry piecewise-constant region map.

    Returned in Fortran order (i fastest), 0-based ids, region = rx + nrx*ry.
    """
    i = np.arange(nx)
    j = np.arange(ny)
    rx = (i * nrx) // nx
    ry = (j * nry) // ny
    mat = rx[:, None] + nrx * ry[None, :]
    return np.asfortranarray(mat.astype(np.int32))


@dataclass
class Problem:
    nx: int
    ny: int
    order: int                       # level-symmetric S_N order
    ng: int
    nu: np.ndarray                   # group edges [eV], ng+1
    xlen: float = 1.0
    ylen: float = 1.0
    dt: float = 2.0e-4               # sh
    nsteps: int = 10
    theta: float = 1.0
    cv: float = 0.5
    tini: float = 1.0                # eV
    bc_type: tuple = (0, 0, 0, 0)    # left, bottom, right, top
    bcT: tuple = (1000.0, 0.0, 0.0, 0.0)
    mat: Optional[np.ndarray] = None
    kmult: Optional[np.ndarray] = None
    opacity_model: int = 1           # 0 = kapB_calc Planck mean, 1 = group midpoint
    c: float = C_LIGHT
    comp_unit: float = COMP_UNIT
    kap0: float = KAP0
    maxit_rte: int = 25
    maxit_mloqd: int = 500
    maxit_gloqd: int = 100
    conv_type: int = 1
    conv_ho: float = 1e-8
    conv_lo: float = 1e-9
    conv_gr1: float = 1e-10
    conv_gr2: float = 1e-15
    chi: float = 0.7
    line_src: float = 2.0
    use_line_search: int = 1
    use_safety_search: int = 0
    kapE_dT_flag: int = 1
    E_bound_low: float = -1.7976931348623157e308
    T_bound_low: float = -1.7976931348623157e308
    threads: int = 1
    # options beyond the reference's defaults
    grey_emission: int = 0           # 0 = analytic grey emission (GLOQD_SOLVES.f08:439-440);
                                     # 1 = sum_g 4 pi kappa_g B_g, consistent with the group sources
    no_grey: int = 0                 # 1 = use_grey 0: MEB_SOLVE instead of the grey Newton (AL:767-770)
    name: str = "custom"
    quadrature: Optional[str] = None  # a named set of the input deck (abu36, abu20, 1d_test) instead of LQn

    def __post_init__(self):
        self.nu = np.ascontiguousarray(self.nu, dtype=np.float64)
        if self.mat is None:
            self.mat = np.zeros((self.nx, self.ny), dtype=np.int32, order="F")
        if self.kmult is None:
            self.kmult = np.ones(1)
        self.mat = np.asfortranarray(self.mat, dtype=np.int32)
        self.kmult = np.ascontiguousarray(self.kmult, dtype=np.float64)
        ox, oy, w = named_quadrature(self.quadrature) if self.quadrature else level_symmetric(self.order)
        self.ox, self.oy, self.w = ox, oy, w
        self.dx = np.full(self.nx, self.xlen / self.nx)
        self.dy = np.full(self.ny, self.ylen / self.ny)

    @property
    def nm(self) -> int:
        return len(self.w)

    @property
    def nmat(self) -> int:
        return len(self.kmult)

    def updates_per_sweep(self) -> int:
        return self.nx * self.ny * self.nm * self.ng

    def with_(self, **kw) -> "Problem":
        q = replace(self, **kw)
        return q


def _conv_cascade(p: Problem, conv_ho: float) -> Problem:
    # INPUTS.f08:588-594: conv_lo = conv_ho/10, conv_gr1 = conv_ho/100
    p.conv_ho, p.conv_lo, p.conv_gr1 = conv_ho, conv_ho / 10.0, conv_ho / 100.0
    return p


def cfg0() -> Problem:
    """Reference CPU case: 20x20 cells, 4 groups 1e-2..1e1 keV, S4, 10 steps."""
    p = Problem(nx=20, ny=20, order=4, ng=4, nu=log_group_edges(4, 1e-2, 1e1), nsteps=10, name="cfg0")
    return _conv_cascade(p, 1e-8)


def cfg1(seed: int = 20240611, nx: int = 256) -> Problem:
    """256x256 cells, 32 groups 1e-3..1e2 keV, S16, 50 steps, 8x8 lognormal(0,1.5) map."""
    rng = SplitMix64(seed)
    kmult = np.array([rng.lognormal(0.0, 1.5) for _ in range(64)])
    p = Problem(nx=nx, ny=nx, order=16, ng=32, nu=log_group_edges(32, 1e-3, 1e2), nsteps=50,
                mat=region_map(nx, nx), kmult=kmult, name="cfg1")
    return _conv_cascade(p, 1e-8)


def mapped(nx: int = 48, order: int = 8, ng: int = 16, sigma: float = 1.0, seed: int = 20240611,
           nsteps: int = 10) -> Problem:
    """The 256x256 problem's physics at a smaller size with a milder multiplier map:
    an 8x8-region lognormal(0, sigma) map, groups 1e-3..1e2 keV.  With sigma = 1.5
    (multipliers 0.04..53) the grey Newton iteration of EGP_FV_NEWT diverges within
    the first step (NaN in the state, also in the oracle); sigma = 1 stays finite."""
    rng = SplitMix64(seed)
    kmult = np.array([rng.lognormal(0.0, sigma) for _ in range(64)])
    p = Problem(nx=nx, ny=nx, order=order, ng=ng, nu=log_group_edges(ng, 1e-3, 1e2), nsteps=nsteps,
                mat=region_map(nx, nx), kmult=kmult, name=f"mapped{nx}")
    return _conv_cascade(p, 1e-8)


CFG3_GRIDS = (64, 128, 256, 384, 512)
CFG3_GROUPS = (16, 32, 64)
CFG3_ORDERS = (8, 12, 16)


def cfg3_problem(base_seed: int, idx: int, nsteps: int = 20) -> Problem:
    """One problem of the batched distribution (grid, groups, S_N, map and T_b drawn per problem)."""
    rng = SplitMix64((base_seed * 0x9E3779B97F4A7C15 + idx * 0xD1B54A32D192ED03) & SplitMix64.MASK)
    n = rng.choice(CFG3_GRIDS)
    ng = rng.choice(CFG3_GROUPS)
    order = rng.choice(CFG3_ORDERS)
    kmult = np.array([rng.lognormal(0.0, 1.5) for _ in range(64)])
    tb = 500.0 + 1000.0 * rng.uniform()
    p = Problem(nx=n, ny=n, order=order, ng=ng, nu=log_group_edges(ng, 1e-3, 1e2), nsteps=nsteps,
                mat=region_map(n, n), kmult=kmult, bcT=(tb, 0.0, 0.0, 0.0), name=f"cfg3[{idx}]")
    return _conv_cascade(p, 1e-8)


def small(nx: int = 12, ny: int = 10, order: int = 4, ng: int = 3, nsteps: int = 2, seed: int = 7,
          opacity_model: int = 1) -> Problem:
    """A small non-square problem with a random multiplier map, for parity tests."""
    rng = SplitMix64(seed)
    kmult = np.array([rng.lognormal(0.0, 1.0) for _ in range(4)])
    p = Problem(nx=nx, ny=ny, order=order, ng=ng, nu=log_group_edges(ng, 1e-2, 1e1), nsteps=nsteps,
                mat=region_map(nx, ny, 2, 2), kmult=kmult, opacity_model=opacity_model, name="small")
    return _conv_cascade(p, 1e-8)


def by_name(name: str) -> Problem:
    if name == "cfg0":
        return cfg0()
    if name == "cfg1":
        return cfg1()
    if name == "small":
        return small()
    raise ValueError(f"unknown problem {name}")
