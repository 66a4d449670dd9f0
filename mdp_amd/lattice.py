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
    """Atoms of an ncell^3 (or n1 x n2 x n3 for a 3-tuple ``ncell``) supercell of the cubic ``kind``
    lattice with constant ``a``; Gaussian displacement of std ``sigma`` per component.
    Returns (x (N,3), box (3,))."""
    basis = _BASIS[kind]
    nc = np.broadcast_to(np.asarray(ncell, dtype=np.int64), (3,))
    g = [np.arange(k) for k in nc]
    cells = np.stack(np.meshgrid(*g, indexing="ij"), axis=-1).reshape(-1, 3)
    x = ((cells[:, None, :] + basis[None, :, :]).reshape(-1, 3) * a).astype(np.float64)
    box = (nc * a).astype(np.float64)
    if sigma > 0:
        rng = np.random.default_rng(seed)
        x = x + rng.normal(0.0, sigma, size=x.shape)
    if wrap:
        x = np.mod(x, box)
    return np.ascontiguousarray(x), box


def random_packing(natoms, density, rmin, seed=0, max_rounds=100000):
    """``natoms`` points in a periodic cube of number density ``density`` (atoms/A^3) by random
    sequential insertion with minimum image separation ``rmin``: uniform candidates are drawn in
    batches, each is rejected if it lies within ``rmin`` of an accepted point.  Returns (x, box)."""
    from scipy.spatial import cKDTree
    L = (natoms / density) ** (1.0 / 3.0)
    box = np.full(3, L)
    rng = np.random.default_rng(seed)
    pts = np.zeros((natoms, 3))
    r2 = rmin * rmin
    k = 0
    for _ in range(max_rounds):
        if k == natoms:
            return pts, box
        cand = rng.uniform(0.0, L, size=(2048, 3))
        if k:
            d, _ = cKDTree(pts[:k], boxsize=L).query(cand, k=1)
            cand = cand[d >= rmin]
        start = k
        for p in cand:
            if k == natoms:
                break
            if k > start:
                dd = pts[start:k] - p
                dd -= L * np.round(dd / L)
                if np.min(np.einsum("ij,ij->i", dd, dd)) < r2:
                    continue
            pts[k] = p
            k += 1
    if k < natoms:
        raise RuntimeError("random_packing: could not place atoms at this density/separation")
    return pts, box


def _supercell_for(natoms, per_cell):
    """Near-cubic n1 <= n2 <= n3 (n3 <= 2 n1) whose atom count per_cell*n1*n2*n3 is closest to natoms."""
    best = None
    top = int(np.ceil((natoms / per_cell) ** (1.0 / 3.0))) + 2
    for n1 in range(1, top + 1):
        for n2 in range(n1, 2 * n1 + 1):
            for n3 in range(n2, 2 * n1 + 1):
                d = abs(per_cell * n1 * n2 * n3 - natoms)
                if best is None or d < best[0]:
                    best = (d, (n1, n2, n3))
    return best[1]


def fitting_batch(nconfig=512, nmin=16, nmax=2000, sigma=0.1, seed=20241203):
    """Independent periodic configurations for batched descriptor fitting: atom counts drawn
    log-uniformly in [nmin, nmax]; configurations cycle BCC (a ~ U[3.2, 3.45]), FCC (a ~ U[4.0, 4.3])
    and random packing (0.0554 atoms/A^3, minimum separation 2.2 A); every one gets Gaussian
    displacements of std ``sigma``.  Crystal counts are rounded to the nearest near-cubic supercell.
    Returns a list of (x (N,3), box (3,), kind)."""
    rng = np.random.default_rng(seed)
    out = []
    for c in range(nconfig):
        kind = ("bcc", "fcc", "liquid")[c % 3]
        target = float(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        sub = int(rng.integers(0, 2**31 - 1))
        if kind == "liquid":
            x, box = random_packing(max(2, int(round(target))), 0.0554, 2.2, seed=sub)
        else:
            a = rng.uniform(3.2, 3.45) if kind == "bcc" else rng.uniform(4.0, 4.3)
            x, box = crystal(kind, _supercell_for(target, len(_BASIS[kind])), a, wrap=False)
        d = np.random.default_rng(sub + 1).normal(0.0, sigma, size=x.shape)
        x = np.mod(x + d, box)
        out.append((np.ascontiguousarray(x), box, kind))
    return out


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
