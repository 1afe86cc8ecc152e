"""Synthetic water-cluster geometries for the benchmark and batch workloads.

(H2O)_n: monomers at the reference README geometry (O 0,-0.075791844,0;
H +-0.866811829,0.601435779,0 A), each rotated uniformly at random about its
O (Shoemake quaternion from three uniforms) and placed uniformly in a cube,
rejecting placements with an O-O distance below `min_oo`.  Optionally the
O-H lengths are perturbed by N(0, 0.02 A) and the H-O-H angle by N(0, 2 deg)
before rotation.  Atoms are emitted molecule by molecule as O, H, H, in a
neutral-singlet geometry string for ``Molecule``.
"""
import numpy as np


def water_cluster(n, box, seed, min_oo=2.7, perturb=False):
    """(H2O)_n: monomers at the README geometry, each rotated uniformly
    (Shoemake quaternion) about its O, O placed uniformly in [0, box]^3 A
    with all O-O distances >= min_oo (rejection sampling).  With `perturb`,
    O-H lengths get N(0, 0.02 A) and the H-O-H angle N(0, 2 deg) first."""
    rng = np.random.default_rng(seed)
    O = np.array([0.0, -0.075791844, 0.0])
    H1 = np.array([0.866811829, 0.601435779, 0.0])
    H2v = np.array([-0.866811829, 0.601435779, 0.0])
    mono0 = np.array([O, H1, H2v]) - O
    placed = []
    out = []
    while len(placed) < n:
        mono = mono0.copy()
        if perturb:
            r1 = np.linalg.norm(mono[1]) + rng.normal(0.0, 0.02)
            r2 = np.linalg.norm(mono[2]) + rng.normal(0.0, 0.02)
            ang0 = np.arccos(np.dot(mono[1], mono[2]) / np.linalg.norm(mono[1]) / np.linalg.norm(mono[2]))
            ang = ang0 + np.deg2rad(rng.normal(0.0, 2.0))
            bis = (mono[1] / np.linalg.norm(mono[1]) + mono[2] / np.linalg.norm(mono[2]))
            bis /= np.linalg.norm(bis)
            perp = np.cross(bis, [0.0, 0.0, 1.0])
            perp /= np.linalg.norm(perp)
            mono[1] = r1 * (np.cos(ang / 2) * bis + np.sin(ang / 2) * perp)
            mono[2] = r2 * (np.cos(ang / 2) * bis - np.sin(ang / 2) * perp)
        u1, u2, u3 = rng.random(3)
        q = np.array([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                      np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)])
        x, y, z, w = q
        Rm = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                       [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                       [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        pos = rng.random(3) * box
        if any(np.linalg.norm(pos - p) < min_oo for p in placed):
            continue
        placed.append(pos)
        for sym, r in zip("OHH", mono):
            out.append((sym,) + tuple(pos + Rm @ r))
    return "\n0 1\n" + "\n".join("%s %.12f %.12f %.12f" % a for a in out) + "\n"
