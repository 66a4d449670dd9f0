"""Seeded synthetic configurations used by the benchmark and the tests.

BCC/FCC crystals of n x n x n cubic cells with Gaussian displacements, random close packings at a
given number density with a hard minimum separation, Maxwell-Boltzmann velocities with zero net
momentum, and seeded linear SNAP coefficients.  All arrays are float64, positions (N, 3) in Angstrom.
"""
import numpy as np

BOLTZ_METAL = 8.617343e-5           # eV/K            (src/C++/Main/setunits.cpp:51-58, metal units)
MVV2E_METAL = 1.0364269e-4          # amu*(A/ps)^2 -> eV
FTM2V_METAL = 1.0 / 1.0364269e-4    # eV/A/amu -> A/ps^2

_BASIS = {
    "bcc": np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.5]]),
    "fcc": np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]]),
}


def crystal(kind, ncell, a, sigma=0.0, seed=0, wrap=True):
    """Atoms of an ncell^3 supercell of the cubic ``kind`` lattice with constant ``a``; Gaussian
    displacement of std ``sigma`` per component.  Returns (x (N,3), box (3,))."""
    basis = _BASIS[kind]
    g = np.arange(ncell)
    cells = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    x = ((cells[:, None, :] + basis[None, :, :]).reshape(-1, 3) * a).astype(np.float64)
    box = np.full(3, ncell * a, dtype=np.float64)
    if sigma > 0:
        rng = np.random.default_rng(seed)
        x = x + rng.normal(0.0, sigma, size=x.shape)
    if wrap:
        x = np.mod(x, box)
    return np.ascontiguousarray(x), box


def random_packing(natoms, density, rmin, seed=0, max_tries=200000):
    """``natoms`` points in a periodic cube of number density ``density`` (atoms/A^3), sequential
    random insertion with minimum image separation ``rmin``.  Returns (x, box)."""
    L = (natoms / density) ** (1.0 / 3.0)
    box = np.full(3, L)
    rng = np.random.default_rng(seed)
    pts = np.zeros((natoms, 3))
    k, tries = 0, 0
    while k < natoms:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("random_packing: could not place atoms at this density/separation")
        p = rng.uniform(0.0, L, size=3)
        if k:
            d = pts[:k] - p
            d -= L * np.round(d / L)
            if np.min(np.einsum("ij,ij->i", d, d)) < rmin * rmin:
                continue
        pts[k] = p
        k += 1
    return pts, box


def maxwell_boltzmann(natoms, mass, temp, seed=0):
    """Velocities (A/ps) at temperature ``temp`` (K) for atoms of ``mass`` (amu), zero net momentum,
    rescaled to the exact kinetic temperature with 3N-3 degrees of freedom."""
    rng = np.random.default_rng(seed)
    v = rng.normal(0.0, 1.0, size=(natoms, 3))
    v -= v.mean(axis=0)
    ke = 0.5 * mass * np.sum(v * v) * MVV2E_METAL
    dof = 3 * natoms - 3
    t_now = 2.0 * ke / (dof * BOLTZ_METAL)
    if t_now > 0:
        v *= np.sqrt(temp / t_now)
    return np.ascontiguousarray(v)


def snap_coefficients(ncoeff, beta0=-11.7, sigma=0.05, seed=0):
    """Linear SNAP coefficients [beta0, beta_1..beta_ncoeff], beta_k ~ N(0, sigma)."""
    rng = np.random.default_rng(seed)
    c = np.empty(ncoeff + 1)
    c[0] = beta0
    c[1:] = rng.normal(0.0, sigma, size=ncoeff)
    return c
