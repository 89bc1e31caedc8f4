"""Seeded synthetic inputs: rigid water clusters and even-tempered bases.

The workloads are water clusters (H2O)n: O atoms uniform in a cube, rejection
sampled to a minimum O-O distance, each H2O rigid (O-H 0.96 A, HOH 104.5 deg)
in a uniformly random orientation.  Bases are uncontracted even-tempered
shells, alpha_k = alpha0 * beta**k, alpha0 ~ U[0.1, 0.3] drawn per
(element, l), contraction coefficient 1.0; `contracted_shells` builds
segmented contractions over the same kind of primitive set.

Writers for the reference's input formats are provided so the same system can
be fed to the reference driver: XYZ (sys_definitions.f90:588-723) and the
EMSL/GAMESS basis text (basis_sets.f90:445-656).
"""
from __future__ import annotations

import math

import numpy as np

from .basis import BOHR_ANGSTROM, AOBasis, build_basis

L_LETTERS = "SPDFGH"

# shell counts per element and l
BASIS_SPECS = {
    "small": {"O": (6, 4, 2), "H": (3, 1)},          # O 6s4p2d, H 3s1p (l <= 2)
    "large": {"O": (8, 5, 3, 1), "H": (4, 2, 1)},    # O 8s5p3d1f, H 4s2p1d (l <= 3)
    # g and h shells (the reference's AUTO2E_MAXL = 5): O 4s3p2d1f1g1h, H 3s2p1d1f1g
    "gh": {"O": (4, 3, 2, 1, 1, 1), "H": (3, 2, 1, 1, 1)},
}


def water_cluster(n: int, box_angstrom: float, seed: int, min_oo: float = 2.7,
                  roh: float = 0.96, angle_deg: float = 104.5):
    """Return (symbols, coords_angstrom (3n, 3)) for a seeded (H2O)n."""
    rng = np.random.default_rng(seed)
    oxy = []
    tries = 0
    while len(oxy) < n:
        tries += 1
        if tries > 200000:
            raise RuntimeError("cannot place %d waters in a %.1f A box" % (n, box_angstrom))
        p = rng.uniform(0.0, box_angstrom, 3)
        if all(np.linalg.norm(p - q) >= min_oo for q in oxy):
            oxy.append(p)
    half = math.radians(angle_deg) / 2.0
    h1_local = np.array([roh * math.sin(half), 0.0, roh * math.cos(half)])
    h2_local = np.array([-roh * math.sin(half), 0.0, roh * math.cos(half)])
    symbols, coords = [], []
    for o in oxy:
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        symbols += ["O", "H", "H"]
        coords += [o, o + R @ h1_local, o + R @ h2_local]
    return symbols, np.array(coords)


def even_tempered_shells(spec: dict, seed: int, beta: float = 2.5):
    """{element: [(l, [alpha], [1.0]), ...]} with s shells first, then p, ...;
    exponents ascending within each l.  alpha0 is drawn per (element, l) in
    sorted element order."""
    rng = np.random.default_rng(seed)
    shells = {}
    for el in sorted(spec):
        lst = []
        for l, count in enumerate(spec[el]):
            a0 = rng.uniform(0.1, 0.3)
            for k in range(count):
                lst.append((l, [a0 * beta ** k], [1.0]))
        shells[el] = lst
    return shells


def contracted_shells(spec: dict, seed: int, beta: float = 2.5):
    """Segmented contractions over an even-tempered primitive set: per
    (element, l) with `count` shells, the first shell contracts the three
    tightest primitives, the second the next two, the rest are single
    primitives; coefficients ~ U[0.2, 0.8].  Exercises the primitive-pair
    loops that uncontracted shells leave at one iteration."""
    rng = np.random.default_rng(seed)
    shells = {}
    for el in sorted(spec):
        lst = []
        for l, count in enumerate(spec[el]):
            a0 = rng.uniform(0.1, 0.3)
            nprims = [3, 2] + [1] * max(0, count - 2)
            nprims = nprims[:count]
            alphas = [a0 * beta ** k for k in range(sum(nprims))][::-1]
            pos = 0
            for npr in nprims:
                ex = alphas[pos:pos + npr]
                pos += npr
                lst.append((l, ex, list(rng.uniform(0.2, 0.8, npr))))
        shells[el] = lst
    return shells


def water_system(n: int, box_angstrom: float, basis: str, seed: int,
                 spher_ao: bool = True, contracted: bool = False) -> AOBasis:
    symbols, xyz = water_cluster(n, box_angstrom, seed)
    make = contracted_shells if contracted else even_tempered_shells
    shells = make(BASIS_SPECS[basis], seed + 1000003)
    return build_basis(symbols, xyz / BOHR_ANGSTROM, shells, spher_ao=spher_ao)


# Named workloads (sizes follow the literal shell specifications).
def config_small_cluster(seed: int = 1) -> AOBasis:
    """(H2O)4, 6 A box, O 6s4p2d / H 3s1p: 160 spherical AOs."""
    return water_system(4, 6.0, "small", seed)


def config_large_cluster(seed: int = 2) -> AOBasis:
    """(H2O)16, 10 A box, O 8s5p3d1f / H 4s2p1d: 1200 spherical AOs."""
    return water_system(16, 10.0, "large", seed)


def batch_systems(nmol: int = 64, seed: int = 3, nmin: int = 1, nmax: int = 12,
                  basis: str = "small") -> list:
    """Independent molecules for the batched form: (H2O)n with n uniform in
    nmin..nmax (seeded), each in a cube of 6 A * (n / 4)^(1/3) (the 6 A box
    of the (H2O)4 workload at the same water density), each with its own
    seeded geometry and the same basis specification."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(nmin, nmax + 1, nmol)
    return [water_system(int(n), 6.0 * (n / 4.0) ** (1.0 / 3.0), basis, seed=seed * 1000 + k)
            for k, n in enumerate(ns)]


def config_gradient_cluster(seed: int = 5) -> AOBasis:
    """(H2O)8 in a 6 A * 2^(1/3) box (the water density of the (H2O)4
    workload), O 6s4p2d / H 3s1p: 320 spherical AOs, 24 atoms."""
    return water_system(8, 6.0 * 2.0 ** (1.0 / 3.0), "small", seed)


def atom_centred_points(basis: AOBasis, radii=(0.35, 0.9, 1.8, 3.0), ndir: int = 8,
                        seed: int = 0) -> np.ndarray:
    """Synthetic THC grid points (npts, 3), bohr: around every atom, ndir
    directions of a Fibonacci sphere (randomly rotated per atom and radius)
    at each radius.  Stands in for the caller's pruned molecular grid."""
    rng = np.random.default_rng(seed)
    k = np.arange(ndir) + 0.5
    z = 1.0 - 2.0 * k / ndir
    phi = math.pi * (1.0 + 5.0 ** 0.5) * k
    rxy = np.sqrt(1.0 - z * z)
    dirs = np.stack([rxy * np.cos(phi), rxy * np.sin(phi), z], axis=1)
    pts = []
    for A in basis.atom_coords:
        for r in radii:
            q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
            pts.append(A + r * dirs @ q.T)
    return np.concatenate(pts, axis=0)


def occupied_density(nao: int, nocc: int, seed: int) -> np.ndarray:
    """D = C_occ C_occ^T from a random orthonormal C (QR of an N(0, 1)
    nao x nao matrix, seeded), C_occ its first nocc columns."""
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((nao, nao)))
    c = q[:, :nocc]
    return c @ c.T


def write_xyz(path: str, symbols, coords_angstrom) -> None:
    """XYZ block as read by sys_Read_XYZ (sys_definitions.f90:588-723)."""
    with open(path, "w") as f:
        f.write("XYZ\n%d\n" % len(symbols))
        for s, c in zip(symbols, coords_angstrom):
            f.write("%s %.12f %.12f %.12f\n" % (s, c[0], c[1], c[2]))
        f.write("END\n")


ELEMENT_NAMES = {"H": "HYDROGEN", "O": "OXYGEN", "C": "CARBON", "N": "NITROGEN"}


def write_emsl(path: str, element_shells: dict) -> None:
    """GAMESS-US/EMSL text basis (basis_sets.f90:445-656): element long name,
    then "L nprim" lines followed by "k exponent coefficient"."""
    with open(path, "w") as f:
        f.write("$DATA\n")
        for el, shells in element_shells.items():
            f.write("%s\n" % ELEMENT_NAMES[el])
            for l, exps, coefs in shells:
                f.write("%s %d\n" % (L_LETTERS[l], len(exps)))
                for k, (e, c) in enumerate(zip(exps, coefs), 1):
                    f.write("%d %.15E %.15E\n" % (k, e, c))
            f.write("\n")
        f.write("$END\n")
