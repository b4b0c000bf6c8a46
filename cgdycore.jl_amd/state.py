"""Seeded baroclinic-jet initial state on the global nodes.

The prognostic array ``U`` has the reference's layout ``(Nz, NumG, NumV)``
(``Examples/testNHSphere.jl:194-198``: 1 = rho, 2 = u, 3 = v, 4 = w at the
top interface of each layer, 5 = rho*theta), held as numpy ``(5, NumG, Nz)``.

State definition (fixed for every workload of this build):

* T(z) = 288 - 0.005 z, p = 1e5 (T/288)^(g/(Rd 0.005)), rho = p/(Rd T)
* theta = T (p0/p)^kappa + noise, noise ~ U(-1, 1) K (amplitude 1 K)
* u = 35 sin^2(2 lat) taper(z), taper = 1 below 15 km and
  cos^2(pi/2 (z-15 km)/(H-15 km)) above, v = 0, w = 0
* z is the layer midpoint (Iz - 1/2) H/nz, lat is the latitude of the node
  (taken from its first occurrence in (IF, ID) order)
* the noise is a counter-based generator: SplitMix64 of
  ``(seed << 40) + (node << 12) + level`` (0-based node and level), mapped to
  [-1, 1) with 53 random bits, so every global node and level draws the same
  value on every device and in every run.
"""
from __future__ import annotations

import numpy as np

from .discretization import Discretization, Phys

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 output function applied to counters ``x`` (uint64)."""
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        z = z ^ (z >> np.uint64(31))
    return z


def noise(seed: int, num_g: int, nz: int) -> np.ndarray:
    """U(-1, 1) noise of shape (NumG, Nz) keyed by (seed, node, level)."""
    if not (0 <= seed < (1 << 23) and num_g < (1 << 28) and nz < (1 << 12)):
        raise ValueError("noise counter fields out of range")
    g = np.arange(num_g, dtype=np.uint64)[:, None]
    iz = np.arange(nz, dtype=np.uint64)[None, :]
    ctr = (np.uint64(seed) << np.uint64(40)) + (g << np.uint64(12)) + iz
    bits = splitmix64(ctr) >> np.uint64(11)
    u01 = bits.astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * u01 - 1.0


def node_latitude(disc: Discretization) -> np.ndarray:
    """Latitude of every global node (first occurrence in (IF, ID) order)."""
    X = disc.X  # (NF, Nz, 3, 2, N^2)
    nf, nn = disc.glob.shape
    idx = disc.glob.reshape(-1) - 1
    first = np.full(disc.num_g, -1, dtype=np.int64)
    order = np.arange(idx.size)[::-1]
    first[idx[::-1]] = order  # last assignment wins -> smallest (IF, ID)
    e, i = np.divmod(first, nn)
    x = X[e, 0, 0, 0, i]
    y = X[e, 0, 1, 0, i]
    zc = X[e, 0, 2, 0, i]
    return np.arcsin(zc / np.sqrt(x * x + y * y + zc * zc))


def taper(zm: np.ndarray, H: float, z_top_jet: float = 15000.0) -> np.ndarray:
    s = np.clip((zm - z_top_jet) / (H - z_top_jet), 0.0, 1.0)
    return np.where(zm <= z_top_jet, 1.0, np.cos(0.5 * np.pi * s) ** 2)


def jet_state(disc: Discretization, seed: int = 0, theta_amp: float = 1.0,
              phys: Phys = Phys()) -> np.ndarray:
    """Seeded jet state ``U`` as numpy (5, NumG, Nz), float64."""
    nz, num_g = disc.nz, disc.num_g
    zm = (np.arange(nz) + 0.5) * (disc.H / nz)
    T0, lapse, ps = 288.0, 0.005, 1.0e5
    Rd, g, p0, kappa = phys.Rd, phys.Grav, phys.p0, phys.kappa
    T = T0 - lapse * zm
    p = ps * (T / T0) ** (g / (Rd * lapse))
    rho = p / (Rd * T)
    theta = T * (p0 / p) ** kappa
    lat = node_latitude(disc)
    U = np.zeros((5, num_g, nz))
    th = theta[None, :] + theta_amp * noise(seed, num_g, nz)
    U[0] = rho[None, :]
    U[1] = 35.0 * np.sin(2.0 * lat)[:, None] ** 2 * taper(zm, disc.H)[None, :]
    U[4] = rho[None, :] * th
    return U


def ensemble_members(n: int = 32, nes=(8, 12, 16)):
    """The ensemble workload's members: member m uses ne = nes[m mod len(nes)]
    and theta-noise seed m.  Returns (ne list, seed list)."""
    return [nes[m % len(nes)] for m in range(n)], list(range(n))


def ensemble_state(discs, seeds, theta_amp: float = 1.0, phys: Phys = Phys()) -> np.ndarray:
    """Members' seeded jet states concatenated along the node axis:
    numpy (5, sum NumG, Nz), member m's nodes after those of members 0..m-1."""
    if len(discs) != len(seeds):
        raise ValueError("one seed per member expected")
    return np.concatenate([jet_state(d, seed=s, theta_amp=theta_amp, phys=phys) for d, s in zip(discs, seeds)],
                          axis=1)
