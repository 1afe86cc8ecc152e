"""Atom-pair learned initial density (no reference counterpart).

The guess keeps the SAD atomic blocks (mmd.guess) and fills the off-diagonal
atom-pair blocks, which SAD leaves at zero, from a network per pair of
elements:

* Pair (i, j) of atoms closer than R_CUT gets the bond-axis frame: z along
  r_j - r_i, x and y completing a right-handed frame.  Cartesian p functions
  transform as vectors, so the block in that frame is B = U_i^T P_ij U_j
  (U: identity on s functions, the frame's axes on every p triple).
* Input: 32 radial features of the distance, Gaussians exp(-g (r - mu_k)^2)
  on an even grid mu_k in [R_MIN, R_MAX], times the cosine cutoff
  (1 + cos(pi r / R_CUT)) / 2.
* Output: the parts of B that are unchanged by a rotation about the bond
  axis: (s|s), (s|p_z), (p_z|s), (p_z|p_z) and the mean of (p_x|p_x) and
  (p_y|p_y) for every pair of shells; the rest of B is predicted as zero.
* Network: three hidden layers of 128 SiLU units and a linear output layer,
  one per ordered element pair (O-O, O-H, H-O, H-H), in libmmd's HIP
  perceptron (mmd_mlp_* in include/mmd.h).

Both orders of a pair are predicted and averaged, P_ij = (B_ij-part +
(B_ji-part)^T) / 2 rotated back, so that the guess is symmetric.  The radial
features alone cannot see a pair's surroundings; mmd.learned's monomer model
can (field of the other monomers, monomer geometry) and also corrects the
atomic blocks.  make_pair_guess.py trains this model and compares the two.
"""
import os
import threading

import numpy as np

NRBF = 32
R_MIN, R_MAX, R_CUT = 0.5, 9.0, 9.5          # bohr
_MU = np.linspace(R_MIN, R_MAX, NRBF)
_GAMMA = 0.5 / (_MU[1] - _MU[0]) ** 2
HIDDEN = (128, 128, 128)
ELEMENTS = (8, 1)
DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_MODEL = os.path.join(DATA_DIR, "pair_guess_631g.npz")


def radial_features(r):
    """(n, NRBF) features of distances r (bohr)."""
    r = np.asarray(r, dtype=float)[:, None]
    fc = 0.5 * (1.0 + np.cos(np.pi * np.minimum(r, R_CUT) / R_CUT))
    return np.exp(-_GAMMA * (r - _MU[None, :]) ** 2) * fc


def atom_layout(lmn):
    """s-function indices and first indices of the p triples of one atom's
    functions (lmn rows in basis order)."""
    L = np.asarray(lmn).sum(axis=1)
    if np.any(L > 1):
        raise ValueError("the pair guess supports s and p functions only")
    s_idx, p_idx = [], []
    k = 0
    while k < len(L):
        if L[k] == 0:
            s_idx.append(k)
            k += 1
        else:
            p_idx.append(k)
            k += 3
    return s_idx, p_idx


def bond_frames(d):
    """Bond-axis frames R (n, 3, 3), axes as columns, z along d (n, 3)."""
    z = d / np.linalg.norm(d, axis=1)[:, None]
    a = np.where(np.abs(z[:, :1]) < 0.9, np.array([[1.0, 0.0, 0.0]]), np.array([[0.0, 1.0, 0.0]]))
    x = a - np.sum(a * z, axis=1)[:, None] * z
    x /= np.linalg.norm(x, axis=1)[:, None]
    y = np.cross(z, x)
    return np.stack([x, y, z], axis=2)


class PairShape:
    """Invariant slots of a block between two atoms' function layouts: a list
    of (kind, row, col) with kind 0 a plain element B[row, col] and kind 1 the
    pi pair of p triples starting at (row, col)."""

    def __init__(self, lay_i, lay_j, ni, nj):
        si, pi = lay_i
        sj, pj = lay_j
        self.ni, self.nj = ni, nj
        self.slots = ([(0, a, b) for a in si for b in sj] + [(0, a, b + 2) for a in si for b in pj]
                      + [(0, a + 2, b) for a in pi for b in sj] + [(0, a + 2, b + 2) for a in pi for b in pj]
                      + [(1, a, b) for a in pi for b in pj])
        self.p_i, self.p_j = pi, pj

    @property
    def size(self):
        return len(self.slots)

    def invariants(self, B):
        """(n, size) invariants of blocks B (n, ni, nj) in the bond frame."""
        out = np.empty((B.shape[0], self.size))
        for k, (kind, a, b) in enumerate(self.slots):
            out[:, k] = B[:, a, b] if kind == 0 else 0.5 * (B[:, a, b] + B[:, a + 1, b + 1])
        return out

    def blocks(self, inv):
        """Blocks (n, ni, nj) in the bond frame from invariants (n, size)."""
        B = np.zeros((inv.shape[0], self.ni, self.nj))
        for k, (kind, a, b) in enumerate(self.slots):
            if kind == 0:
                B[:, a, b] = inv[:, k]
            else:
                B[:, a, b] = inv[:, k]
                B[:, a + 1, b + 1] = inv[:, k]
        return B


def _rotations(p_idx, n, R):
    """U (m, n, n): identity on s functions, R on every p triple."""
    U = np.repeat(np.eye(n)[None], len(R), axis=0)
    for k in p_idx:
        U[:, k:k + 3, k:k + 3] = R
    return U


def _pairs(Z, xyz):
    """Ordered atom pairs (i, j), i != j, closer than R_CUT: (i, j, r, d)."""
    d = xyz[None, :, :] - xyz[:, None, :]          # d[i, j] = r_j - r_i
    r = np.linalg.norm(d, axis=2)
    i, j = np.nonzero((r < R_CUT) & ~np.eye(len(Z), dtype=bool))
    return i, j, r[i, j], d[i, j]


class PairModel:
    """One mmd.learned.Model per ordered element pair (SiLU hidden layers)."""

    def __init__(self, nets, layouts):
        self.nets = nets            # {(Zi, Zj): Model}
        self.layouts = layouts      # {Z: (n functions, (s indices, p starts))}

    def shape(self, zi, zj):
        ni, li = self.layouts[zi]
        nj, lj = self.layouts[zj]
        return PairShape(li, lj, ni, nj)

    def save(self, path):
        d = {"pairs": np.array(sorted(self.nets), dtype=np.int32)}
        for (zi, zj), m in self.nets.items():
            key = "%d_%d_" % (zi, zj)
            d[key + "x_mean"], d[key + "x_std"], d[key + "y_std"] = m.x_mean, m.x_std, m.y_std
            d[key + "nlayer"] = len(m.layers)
            for k, (W, b) in enumerate(m.layers):
                d[key + "W%d" % k], d[key + "b%d" % k] = W, b
        for z, (n, (s_idx, p_idx)) in self.layouts.items():
            d["layout_%d_s" % z] = np.array(s_idx, dtype=np.int32)
            d["layout_%d_p" % z] = np.array(p_idx, dtype=np.int32)
            d["layout_%d_n" % z] = n
        np.savez(path, **d)

    @classmethod
    def load(cls, path):
        from mmd.learned import Model
        d = np.load(path, allow_pickle=False)
        nets = {}
        for zi, zj in d["pairs"]:
            key = "%d_%d_" % (zi, zj)
            layers = [(d[key + "W%d" % k], d[key + "b%d" % k]) for k in range(int(d[key + "nlayer"]))]
            m = Model(layers, d[key + "x_mean"], d[key + "x_std"], float(d[key + "y_std"]))
            m.activation = "silu"
            nets[(int(zi), int(zj))] = m
        layouts = {}
        for z in ELEMENTS:
            if "layout_%d_n" % z in d:
                layouts[z] = (int(d["layout_%d_n" % z]),
                              ([int(x) for x in d["layout_%d_s" % z]], [int(x) for x in d["layout_%d_p" % z]]))
        return cls(nets, layouts)


def element_layouts(basis):
    """{Z: (n functions, layout)} of the water elements in ``basis``."""
    from mmd.batch import _element_template
    out = {}
    for z in ELEMENTS:
        lmn = _element_template(basis, z)[0]
        out[z] = (len(lmn), atom_layout(lmn))
    return out


def pair_blocks_frame(Z, xyz, layouts):
    """Per ordered element pair: atom indices (i, j), radial features and
    bond frames of every pair of one molecule within R_CUT."""
    i, j, r, d = _pairs(np.asarray(Z), np.asarray(xyz, dtype=float))
    R = bond_frames(d) if len(i) else np.zeros((0, 3, 3))
    F = radial_features(r) if len(i) else np.zeros((0, NRBF))
    out = {}
    Z = np.asarray(Z)
    for zi in ELEMENTS:
        for zj in ELEMENTS:
            m = (Z[i] == zi) & (Z[j] == zj)
            if m.any():
                out[(zi, zj)] = (i[m], j[m], F[m], R[m])
    return out


def _offsets(Z, layouts):
    o = np.zeros(len(Z) + 1, dtype=int)
    for k, z in enumerate(Z):
        o[k + 1] = o[k] + layouts[int(z)][0]
    return o


def training_pairs(Z, xyz, P, layouts):
    """{(Zi, Zj): (features, invariant targets)} of one converged density P
    (the SAD start has zero off-diagonal blocks, so the target is P's block)."""
    off = _offsets(Z, layouts)
    out = {}
    for key, (i, j, F, R) in pair_blocks_frame(Z, xyz, layouts).items():
        ni, li = layouts[key[0]]
        nj, lj = layouts[key[1]]
        sh = PairShape(li, lj, ni, nj)
        blk = np.stack([P[off[a]:off[a] + ni, off[b]:off[b] + nj] for a, b in zip(i, j)])
        Ui = _rotations(li[1], ni, R)
        Uj = _rotations(lj[1], nj, R)
        B = np.transpose(Ui, (0, 2, 1)) @ blk @ Uj
        out[key] = (F, sh.invariants(B))
    return out


_LOCK = threading.Lock()
_MODELS = {}


def default_model():
    """The trained model shipped in mmd/data (loaded once per process)."""
    m = _MODELS.get(DEFAULT_MODEL)
    if m is None:
        m = PairModel.load(DEFAULT_MODEL)
        _MODELS[DEFAULT_MODEL] = m
    return m


def predict_offdiagonal(molecules, model, ctx=None):
    """Off-diagonal atom-pair blocks of many molecules at once: per molecule
    a list of (i, j, block) with i < j (block of rows of atom i)."""
    from mmd import _native
    per = []
    feats = {k: [] for k in model.nets}
    for Z, xyz in molecules:
        pb = pair_blocks_frame(Z, xyz, model.layouts)
        per.append(pb)
        for k, v in pb.items():
            feats[k].append(v[2])
    preds = {}
    for k, lst in feats.items():
        if not lst:
            continue
        m = model.nets[k]
        X = ((np.concatenate(lst) - m.x_mean) / m.x_std).astype(np.float32)
        with _LOCK:   # the networks' device buffers are shared
            preds[k] = _native.mlp_forward(ctx, m, X).astype(float) * float(m.y_std)
    used = {k: 0 for k in preds}
    out = []
    for (Z, xyz), pb in zip(molecules, per):
        acc = {}
        for k, (i, j, F, R) in pb.items():
            n = len(i)
            inv = preds[k][used[k]:used[k] + n]
            used[k] += n
            ni, li = model.layouts[k[0]]
            nj, lj = model.layouts[k[1]]
            B = PairShape(li, lj, ni, nj).blocks(inv)
            blk = _rotations(li[1], ni, R) @ B @ np.transpose(_rotations(lj[1], nj, R), (0, 2, 1))
            for a, b, x in zip(i, j, blk):
                if a < b:
                    acc.setdefault((a, b), []).append(x)
                else:
                    acc.setdefault((b, a), []).append(x.T)
        out.append([(a, b, sum(v) / len(v)) for (a, b), v in sorted(acc.items())])
    return out


def batch_pair_densities(molecules, basis, model=None, ctx=None, element_density=None):
    """Initial half densities of many water clusters: SAD atomic blocks plus
    the predicted atom-pair blocks.  molecules: per cluster (Z, xyz bohr,
    lmn of its functions, electron count), as for
    mmd.learned.batch_learned_densities."""
    from mmd.learned import _element_density
    model = model or default_model()
    element_density = element_density or _element_density
    offd = predict_offdiagonal([(m[0], m[1]) for m in molecules], model, ctx)
    out = []
    for (Z, xyz, lmn, nelec), blocks in zip(molecules, offd):
        Z = np.asarray(Z)
        off = _offsets(Z, model.layouts)
        P = np.zeros((off[-1], off[-1]))
        for k, z in enumerate(Z):
            b = element_density(int(z), basis)
            P[off[k]:off[k + 1], off[k]:off[k + 1]] = b
        if Z.sum() != nelec:
            P *= nelec / Z.sum()
        for a, b, x in blocks:
            P[off[a]:off[a + 1], off[b]:off[b + 1]] = x
            P[off[b]:off[b + 1], off[a]:off[a + 1]] = x.T
        out.append(P)
    return out


def pair_density(mol, model=None):
    """batch_pair_densities() of one Molecule (RHF(guess='pair'))."""
    Z = np.array([a.charge for a in mol.atoms])
    if np.any((Z != 8) & (Z != 1)):
        raise ValueError("the pair guess handles water clusters (O and H)")
    xyz = np.array([a.origin for a in mol.atoms], dtype=float)
    return batch_pair_densities([(Z, xyz, None, mol.nelec)], mol.basis_name, model, getattr(mol, "native", None))[0]


def train(data, layouts, epochs=60, batch=512, lr=1e-3, seed=0, device=0):
    """Fit one network per ordered element pair to {(Zi, Zj): (X, Y)};
    returns the PairModel and the per-pair loss histories."""
    from mmd.learned import train_model
    nets, losses = {}, {}
    for k, (X, Y) in sorted(data.items()):
        m, l = train_model(X, Y, hidden=HIDDEN, epochs=epochs, batch=batch, lr=lr, seed=seed, device=device,
                           activation="silu")
        m.activation = "silu"
        nets[k] = m
        losses[k] = l
    return PairModel(nets, layouts), losses
