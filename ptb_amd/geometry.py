"""Deterministic synthetic geometries for the density-build workloads.

Three families are generated, all in bohr (Angstrom / 0.52917726, the
conversion PTB applies when it reads an xyz file, source/main.f90:277 ->
source/rdcoord.f:55):

* water boxes: O atoms on an a x b x c cubic grid (3.1 A spacing) with a
  uniform +-jitter, each water rigid (O-H 0.957 A, H-O-H 104.5 deg) with a
  uniformly random orientation (uniform unit quaternion), numpy
  ``default_rng(seed)``;
* all-trans linear alkanes CnH2n+2 (C-C 1.54 A, C-H 1.09 A, tetrahedral
  109.47 deg), zigzag along x, no noise;
* the mixed batch of 256 molecules (128 cropped water clusters of 4-40 H2O
  and 128 alkanes with n in 3..40), sizes drawn from one seed.

The AO layout follows the vDZP shell pattern PTB uses for these elements
(H: s,s,p -> 5 functions; C and O: s,s,p,p,d -> 13 functions), so ndim and
the AO -> atom map ``aoat`` have the reference's shapes (ndim 1472 for 64
H2O, 11776 for 512 H2O, 9210 for C400H802).  Valence electron counts are
H 1, C 4, O 6 (source/com.f90:52).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

ANGSTROM_TO_BOHR = 1.0 / 0.52917726

# element -> (nuclear charge, valence electrons, AO count)
ELEMENTS = {"H": (1, 1, 5), "C": (6, 4, 13), "O": (8, 6, 13)}
# element -> shell (1-based, within the atom) of each of its AOs:
# H s, s, p(3); C and O s, s, p(3), p(3), d(5)
SHELLS = {"H": [1, 2, 3, 3, 3], "C": [1, 2, 3, 3, 3, 4, 4, 4, 5, 5, 5, 5, 5]}
SHELLS["O"] = SHELLS["C"]


@dataclass
class Molecule:
    """A geometry plus the AO bookkeeping the density build needs."""

    symbols: list
    xyz: np.ndarray            # (n, 3) bohr, C order; Fortran sees xyz(3, n)
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return len(self.symbols)

    @property
    def nel(self) -> int:
        return int(sum(ELEMENTS[s][1] for s in self.symbols))

    @property
    def ndim(self) -> int:
        return int(sum(ELEMENTS[s][2] for s in self.symbols))

    @property
    def aoat(self) -> np.ndarray:
        """1-based atom index of every AO (the reference's ``aoat``)."""
        counts = [ELEMENTS[s][2] for s in self.symbols]
        return np.repeat(np.arange(1, self.n + 1, dtype=np.int64), counts)

    @property
    def shell2ao(self) -> np.ndarray:
        """Shell (1-based, within its atom) of every AO (the reference's
        ``shell2ao``, used by the population analysis)."""
        return np.concatenate([SHELLS[s] for s in self.symbols]).astype(np.int64)

    @property
    def znum(self) -> np.ndarray:
        return np.array([ELEMENTS[s][0] for s in self.symbols], dtype=np.int64)


def _random_rotations(rng: np.random.Generator, k: int) -> np.ndarray:
    """k uniformly distributed rotation matrices from unit quaternions."""
    q = rng.normal(size=(k, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    r = np.empty((k, 3, 3))
    r[:, 0, 0] = 1 - 2 * (y * y + z * z)
    r[:, 0, 1] = 2 * (x * y - z * w)
    r[:, 0, 2] = 2 * (x * z + y * w)
    r[:, 1, 0] = 2 * (x * y + z * w)
    r[:, 1, 1] = 1 - 2 * (x * x + z * z)
    r[:, 1, 2] = 2 * (y * z - x * w)
    r[:, 2, 0] = 2 * (x * z - y * w)
    r[:, 2, 1] = 2 * (y * z + x * w)
    r[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return r


def _water_frame() -> np.ndarray:
    """Rigid H2O in its own frame (Angstrom): O at the origin."""
    roh = 0.957
    half = np.deg2rad(104.5) / 2.0
    return np.array([
        [0.0, 0.0, 0.0],
        [roh * np.sin(half), 0.0, roh * np.cos(half)],
        [-roh * np.sin(half), 0.0, roh * np.cos(half)],
    ])


def water_box(shape=(4, 4, 4), spacing=3.1, jitter=0.2, seed=0,
              keep=None) -> Molecule:
    """Water molecules on a jittered cubic grid.

    ``shape`` is the grid (a, b, c); ``keep`` optionally limits the box to
    the first ``keep`` grid sites in C order (the cropped-grid clusters of
    the mixed batch).  All random numbers come from ``default_rng(seed)``:
    first the (a*b*c, 3) jitter, then the (a*b*c, 4) quaternions.
    """
    a, b, c = shape
    rng = np.random.default_rng(seed)
    grid = np.stack(np.meshgrid(np.arange(a), np.arange(b), np.arange(c),
                                indexing="ij"), axis=-1).reshape(-1, 3)
    oxy = grid * spacing + rng.uniform(-jitter, jitter, size=grid.shape)
    rot = _random_rotations(rng, len(grid))
    if keep is not None:
        oxy, rot = oxy[:keep], rot[:keep]
    frame = _water_frame()
    pos = oxy[:, None, :] + np.einsum("kij,aj->kai", rot, frame)
    symbols = ["O", "H", "H"] * len(oxy)
    return Molecule(symbols, pos.reshape(-1, 3) * ANGSTROM_TO_BOHR,
                    name=f"h2o_{len(oxy)}",
                    meta={"shape": list(shape), "spacing": spacing,
                          "jitter": jitter, "seed": seed})


def alkane(nc: int) -> Molecule:
    """All-trans CnH2n+2 zigzag along x (C-C 1.54 A, C-H 1.09 A)."""
    rcc, rch = 1.54, 1.09
    theta = np.deg2rad(109.47)
    dx = rcc * np.sin(theta / 2.0)
    dy = rcc * np.cos(theta / 2.0)
    carbons = np.array([[i * dx, (dy / 2.0) * (1 if i % 2 == 0 else -1), 0.0]
                        for i in range(nc)])

    def virtual(i):
        # where carbon i would sit if the zigzag went on
        return np.array([i * dx, (dy / 2.0) * (1 if i % 2 == 0 else -1), 0.0])

    symbols, pos = [], []
    half = theta / 2.0
    for i in range(nc):
        ci = carbons[i]
        left = carbons[i - 1] if i > 0 else virtual(i - 1)
        right = carbons[i + 1] if i < nc - 1 else virtual(i + 1)
        u1 = (left - ci) / np.linalg.norm(left - ci)
        u2 = (right - ci) / np.linalg.norm(right - ci)
        bis = u1 + u2
        bis /= np.linalg.norm(bis)
        perp = np.cross(u1, u2)
        perp /= np.linalg.norm(perp)
        hs = [ci + rch * (-bis * np.cos(half) + s * perp * np.sin(half))
              for s in (1.0, -1.0)]
        if nc == 1:
            hs += [ci + rch * u1, ci + rch * u2]
        elif i == 0:
            hs.append(ci + rch * u1)      # trans H at the left chain end
        elif i == nc - 1:
            hs.append(ci + rch * u2)      # trans H at the right chain end
        symbols.append("C")
        pos.append(ci)
        for h in hs:
            symbols.append("H")
            pos.append(h)
    return Molecule(symbols, np.array(pos) * ANGSTROM_TO_BOHR,
                    name=f"C{nc}H{2 * nc + 2}", meta={"nc": nc})


def mixed_batch(count=256, seed=2):
    """The batch of independent molecules: half water clusters, half alkanes.

    Water cluster k has ``nw`` molecules, uniform in 4..40, cut from the
    first ``nw`` sites of a 4x4x3 (48-site) grid with its own seed; alkane k
    has n carbons uniform in 3..40.  All sizes come from ``default_rng(seed)``.
    """
    rng = np.random.default_rng(seed)
    half = count // 2
    nws = rng.integers(4, 41, size=half)
    ncs = rng.integers(3, 41, size=count - half)
    wseeds = rng.integers(0, 2**31 - 1, size=half)
    mols = []
    for nw, ws in zip(nws, wseeds):
        mols.append(water_box((4, 4, 3), 3.1, 0.2, int(ws), keep=int(nw)))
    for nc in ncs:
        mols.append(alkane(int(nc)))
    return mols


def config_molecule(name: str) -> Molecule:
    """The named single-system workloads (sizes as in the workload table)."""
    if name == "h2o_64":
        return water_box((4, 4, 4), 3.1, 0.2, 0)
    if name == "h2o_512":
        return water_box((8, 8, 8), 3.1, 0.2, 1)
    if name == "h2o_2048":
        return water_box((16, 16, 8), 3.1, 0.25, 3)
    if name == "c400h802":
        return alkane(400)
    raise KeyError(name)
