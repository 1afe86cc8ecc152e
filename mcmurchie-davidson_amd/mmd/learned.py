"""Learned initial density for water clusters (no reference counterpart).

The guess is the SAD density (mmd.guess) with every water monomer's diagonal
block corrected by a small multilayer perceptron:

* Each monomer (atoms O, H, H in input order) gets a local frame: origin at
  O, x along the H-O-H bisector, y in the molecular plane towards H1, z =
  x cross y.  Cartesian p functions transform as vectors, so a monomer's
  density block in that frame is P_loc = U^T P U with U = diag(1 | R | ...)
  (R: local axes as columns).
* Input features (15 per monomer): the two O-H lengths and the H-O-H angle,
  and the electrostatic potential and local-frame field of the other
  monomers' point charges (O -0.834, H +0.417; TIP3P) at O, H1 and H2.
* Output: the upper triangle of P_loc(converged) - P_loc(SAD) (one value per
  unique element of the monomer block).

Inter-monomer blocks stay zero, as in SAD.  The network's forward and
backward passes and its training run in libmmd's HIP kernels
(mmd_mlp_* in include/mmd.h); ``train_model`` drives them on converged
densities that the GPU SCF itself produces (``make_dataset``).
"""
import os
import threading

import numpy as np

from mmd.guess import sad_density

Q_O, Q_H = -0.834, 0.417
DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_MODEL = os.path.join(DATA_DIR, "learned_guess_631g.npz")


def monomers(mol):
    """Per water: (atom indices (O, H1, H2), basis-function indices)."""
    atoms = mol.atoms
    if len(atoms) % 3:
        raise ValueError("the learned guess handles water clusters (O, H, H per monomer)")
    nbf_atom = []
    for a in atoms:
        nbf_atom.append(int(round(a.mask.sum())))
    start = np.concatenate([[0], np.cumsum(nbf_atom)])
    out = []
    for m in range(len(atoms) // 3):
        ia = (3 * m, 3 * m + 1, 3 * m + 2)
        if [atoms[i].charge for i in ia] != [8, 1, 1]:
            raise ValueError("the learned guess handles water clusters (O, H, H per monomer)")
        bf = np.arange(start[ia[0]], start[ia[2] + 1])
        out.append((ia, bf))
    return out


def frame(mol, ia):
    O, H1, H2 = (np.asarray(mol.atoms[i].origin, dtype=float) for i in ia)
    u1 = (H1 - O) / np.linalg.norm(H1 - O)
    u2 = (H2 - O) / np.linalg.norm(H2 - O)
    x = u1 + u2
    x /= np.linalg.norm(x)
    y = u1 - np.dot(u1, x) * x
    y /= np.linalg.norm(y)
    z = np.cross(x, y)
    return O, np.stack([x, y, z], axis=1)


def block_rotation(mol, bf, R):
    """U with P_loc = U^T P_block U: identity on s functions, R on every
    (x, y, z) triple of p functions."""
    n = len(bf)
    U = np.eye(n)
    k = 0
    while k < n:
        shell = tuple(mol.bfs[bf[k]].shell)
        if sum(shell) == 0:
            k += 1
        elif sum(shell) == 1:
            U[k:k + 3, k:k + 3] = R
            k += 3
        else:
            raise ValueError("the learned guess supports s and p functions only")
    return U


def features(mol):
    """(n_monomer, 15) inputs of the network, in the order of monomers(mol)."""
    mons = monomers(mol)
    xyz = np.array([a.origin for a in mol.atoms], dtype=float)
    q = np.array([Q_O if a.charge == 8 else Q_H for a in mol.atoms])
    F = np.zeros((len(mons), 15))
    for m, (ia, _) in enumerate(mons):
        O, R = frame(mol, ia)
        r1 = np.linalg.norm(xyz[ia[1]] - O)
        r2 = np.linalg.norm(xyz[ia[2]] - O)
        th = np.arccos(np.clip(np.dot(xyz[ia[1]] - O, xyz[ia[2]] - O) / (r1 * r2), -1.0, 1.0))
        F[m, :3] = (r1, r2, th)
        others = np.array([i for i in range(len(xyz)) if i not in ia], dtype=int)
        for k, i in enumerate(ia):
            if len(others) == 0:
                continue
            d = xyz[i] - xyz[others]
            r = np.linalg.norm(d, axis=1)
            F[m, 3 + k] = np.sum(q[others] / r)
            E = np.sum((q[others] / r ** 3)[:, None] * d, axis=0)
            F[m, 6 + 3 * k:9 + 3 * k] = R.T @ E
    return F


def local_blocks(mol, P):
    """Upper triangles of every monomer's local-frame diagonal block of P."""
    out = []
    for ia, bf in monomers(mol):
        _, R = frame(mol, ia)
        U = block_rotation(mol, bf, R)
        B = U.T @ P[np.ix_(bf, bf)] @ U
        iu = np.triu_indices(len(bf))
        out.append(B[iu])
    return np.array(out)


def assemble(mol, P_base, corrections):
    """P_base with each monomer block's local-frame correction added."""
    P = np.array(P_base, dtype=float, copy=True)
    for (ia, bf), c in zip(monomers(mol), corrections):
        _, R = frame(mol, ia)
        U = block_rotation(mol, bf, R)
        n = len(bf)
        C = np.zeros((n, n))
        iu = np.triu_indices(n)
        C[iu] = c
        C = C + C.T - np.diag(np.diag(C))
        P[np.ix_(bf, bf)] += U @ C @ U.T
    return P


def training_pair(mol, P_converged):
    """(features, targets) of one converged molecule: target = converged
    minus SAD local blocks."""
    P_sad = sad_density(mol)
    return features(mol), local_blocks(mol, np.real(P_converged)) - local_blocks(mol, P_sad)


class Model:
    """Weights and input/output normalisation of the perceptron.  ``layers``
    is a list of (W, b) with W of shape (fan_in, fan_out); hidden layers use
    tanh, the output layer is linear."""

    def __init__(self, layers, x_mean, x_std, y_std):
        self.layers = [(np.asarray(W, dtype=np.float32), np.asarray(b, dtype=np.float32)) for W, b in layers]
        self.x_mean = np.asarray(x_mean, dtype=np.float32)
        self.x_std = np.asarray(x_std, dtype=np.float32)
        self.y_std = np.float32(y_std)

    @property
    def dims(self):
        return [self.layers[0][0].shape[0]] + [W.shape[1] for W, _ in self.layers]

    def flat(self):
        return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in self.layers]).astype(np.float32)

    def set_flat(self, v):
        k = 0
        layers = []
        for W, b in self.layers:
            nW, nb = W.size, b.size
            layers.append((v[k:k + nW].reshape(W.shape).astype(np.float32), v[k + nW:k + nW + nb].astype(np.float32)))
            k += nW + nb
        self.layers = layers
        self._gpu = None

    def save(self, path):
        d = {"x_mean": self.x_mean, "x_std": self.x_std, "y_std": self.y_std, "nlayer": len(self.layers)}
        for i, (W, b) in enumerate(self.layers):
            d["W%d" % i], d["b%d" % i] = W, b
        np.savez(path, **d)

    @classmethod
    def load(cls, path):
        d = np.load(path, allow_pickle=False)
        layers = [(d["W%d" % i], d["b%d" % i]) for i in range(int(d["nlayer"]))]
        return cls(layers, d["x_mean"], d["x_std"], float(d["y_std"]))

    @classmethod
    def init(cls, dims, x_mean, x_std, y_std, seed=0):
        rng = np.random.default_rng(seed)
        layers = []
        for a, b in zip(dims[:-1], dims[1:]):
            layers.append((rng.standard_normal((a, b)) * np.sqrt(1.0 / a), np.zeros(b)))
        return cls(layers, x_mean, x_std, y_std)


def predict(mol, model, ctx=None):
    """Local-frame corrections for every monomer (forward pass on the GPU)."""
    from mmd import _native
    X = ((features(mol) - model.x_mean) / model.x_std).astype(np.float32)
    ctx = ctx or mol.native
    Y = _native.mlp_forward(ctx, model, X)
    return Y.astype(float) * float(model.y_std)


_MODELS = {}


def default_model():
    """The trained model shipped in mmd/data (loaded once per process)."""
    m = _MODELS.get(DEFAULT_MODEL)
    if m is None:
        m = Model.load(DEFAULT_MODEL)
        _MODELS[DEFAULT_MODEL] = m
    return m


def learned_density(mol, model=None):
    """SAD density with every monomer block corrected by the perceptron."""
    model = model or default_model()
    return assemble(mol, sad_density(mol), predict(mol, model))


_LOCK = threading.Lock()


def _element_density(Z, basis):
    """SAD half density of a free atom of element Z in ``basis`` (cached)."""
    from mmd.guess import atomic_density

    class _Atom:
        charge = int(Z)

    class _Mol:
        basis_name = basis

    with _LOCK:
        return atomic_density(_Mol, _Atom)


def _features_stack(Z, X):
    """batch_features() of M clusters of the same composition at once: Z
    (n_atom,), X (M, n_atom, 3) bohr.  Returns features (M, n_monomer, 15)
    and local frames (M, n_monomer, 3, 3)."""
    Z = np.asarray(Z)
    if len(Z) % 3 or np.any(Z.reshape(-1, 3) != (8, 1, 1)):
        raise ValueError("the learned guess handles water clusters (O, H, H per monomer)")
    X = np.asarray(X, dtype=float)
    M, na = X.shape[0], len(Z)
    nm = na // 3
    O, H1, H2 = X[:, 0::3], X[:, 1::3], X[:, 2::3]
    v1, v2 = H1 - O, H2 - O
    r1 = np.linalg.norm(v1, axis=2)
    r2 = np.linalg.norm(v2, axis=2)
    u1, u2 = v1 / r1[..., None], v2 / r2[..., None]
    x = u1 + u2
    x /= np.linalg.norm(x, axis=2)[..., None]
    y = u1 - np.sum(u1 * x, axis=2)[..., None] * x
    y /= np.linalg.norm(y, axis=2)[..., None]
    z = np.cross(x, y)
    R = np.stack([x, y, z], axis=3)
    F = np.zeros((M, nm, 15))
    F[..., 0], F[..., 1] = r1, r2
    F[..., 2] = np.arccos(np.clip(np.sum(v1 * v2, axis=2) / (r1 * r2), -1.0, 1.0))
    if nm > 1:
        q = np.where(Z == 8, Q_O, Q_H)
        d = X[:, :, None, :] - X[:, None, :, :]
        r = np.linalg.norm(d, axis=3)
        mon = np.arange(na) // 3
        other = (mon[:, None] != mon[None, :])[None]
        rinv = np.where(other, 1.0 / np.where(other, r, 1.0), 0.0)
        pot = (q[None, None, :] * rinv).sum(axis=2)
        E = np.einsum("Mij,Mijx->Mix", q[None, None, :] * rinv ** 3, d)
        F[..., 3:6] = pot.reshape(M, nm, 3)
        El = np.einsum("Mnxa,Mnkx->Mnka", R, E.reshape(M, nm, 3, 3))
        F[..., 6:15] = El.reshape(M, nm, 9)
    return F, R


def batch_features(Z, xyz):
    """features() of one water cluster from arrays: atomic numbers Z and
    coordinates xyz (bohr) in O, H, H order per monomer.  Returns
    (features (n_monomer, 15), local frames R (n_monomer, 3, 3))."""
    F, R = _features_stack(Z, np.asarray(xyz, dtype=float)[None])
    return F[0], R[0]


def _monomer_rotation(lmn, R):
    """block_rotation() for every monomer at once: U (n_monomer, n, n)."""
    n = len(lmn)
    L = np.asarray(lmn).sum(axis=1)
    if np.any(L > 1):
        raise ValueError("the learned guess supports s and p functions only")
    U = np.repeat(np.eye(n)[None], len(R), axis=0)
    k = 0
    while k < n:
        if L[k] == 1:
            U[:, k:k + 3, k:k + 3] = R
            k += 3
        else:
            k += 1
    return U


def batch_learned_densities(molecules, basis, model=None, ctx=None, element_density=None):
    """learned_density() of many water clusters with one forward pass.

    molecules: per cluster (Z, xyz bohr, lmn of its functions, electron
    count).  Returns the half densities, one (n, n) array each.  Clusters of
    the same composition are handled together (features, frames and the
    assembly as stacked array operations).  element_density(Z, basis) gives
    the free-atom block (default: the cached SAD atomic density)."""
    from mmd import _native
    model = model or default_model()
    element_density = element_density or _element_density
    kinds = {}
    for k, (Z, xyz, lmn, nelec) in enumerate(molecules):
        kinds.setdefault((tuple(int(z) for z in Z), int(nelec)), []).append(k)
    stacks = []
    for (Zt, nelec), ks in kinds.items():
        F, R = _features_stack(np.array(Zt), np.stack([np.asarray(molecules[k][1], dtype=float) for k in ks]))
        stacks.append((Zt, nelec, ks, F, R))
    X = np.concatenate([st[3].reshape(-1, 15) for st in stacks])
    X = ((X - model.x_mean) / model.x_std).astype(np.float32)
    with _LOCK:   # the network's device buffers are shared
        Y = _native.mlp_forward(ctx, model, X).astype(float) * float(model.y_std)
    out = [None] * len(molecules)
    row = 0
    for Zt, nelec, ks, F, R in stacks:
        Z = np.array(Zt)
        M, nm = F.shape[0], F.shape[1]
        blocks = [element_density(int(z), basis) for z in Z]
        sizes = [b.shape[0] for b in blocks]
        n = sum(sizes)
        P0 = np.zeros((n, n))
        o = 0
        for b in blocks:
            P0[o:o + len(b), o:o + len(b)] = b
            o += len(b)
        if Z.sum() != nelec:   # sad_density's scaling of a charged cluster
            P0 *= nelec / Z.sum()
        nb = sizes[0] + sizes[1] + sizes[2]   # functions of one monomer
        iu = np.triu_indices(nb)
        C = np.zeros((M * nm, nb, nb))
        C[:, iu[0], iu[1]] = Y[row:row + M * nm]
        row += M * nm
        C = C + np.transpose(C, (0, 2, 1)) - C * np.eye(nb)[None]
        lmn = np.asarray(molecules[ks[0]][2])[:nb]
        U = _monomer_rotation(lmn, R.reshape(M * nm, 3, 3))
        D = (U @ C @ np.transpose(U, (0, 2, 1))).reshape(M, nm, nb, nb)
        P = np.repeat(P0[None], M, axis=0)
        for m in range(nm):
            P[:, m * nb:(m + 1) * nb, m * nb:(m + 1) * nb] += D[:, m]
        for j, k in enumerate(ks):
            out[k] = P[j]
    return out


def make_dataset(geometries, basis="6-31g", workers=8):
    """Features and targets of every monomer of every geometry, from
    converged direct RHF densities computed by the batch driver."""
    from mmd.batch import run_batch
    from mmd.molecule import Molecule
    res = run_batch(geometries, basis, workers=workers, keep_density=True, conver=1e-8)
    X, Y = [], []
    for g, r in zip(geometries, res):
        if not r["converged"]:
            continue
        mol = Molecule(g, basis)
        f, t = training_pair(mol, r["P"])
        X.append(f)
        Y.append(t)
    return np.concatenate(X), np.concatenate(Y)


def train_model(X, Y, hidden=(64, 64), epochs=200, batch=512, lr=2e-3, seed=0, device=0, activation="tanh"):
    """Fit the perceptron to (X, Y) with the GPU trainer; returns the Model
    and the per-epoch training loss (normalised units).  activation: of the
    hidden layers, 'tanh' or 'silu'."""
    from mmd import _native
    x_mean, x_std = X.mean(axis=0), X.std(axis=0) + 1e-8
    y_std = float(Y.std()) + 1e-12
    dims = [X.shape[1]] + list(hidden) + [Y.shape[1]]
    model = Model.init(dims, x_mean, x_std, y_std, seed=seed)
    model.activation = activation
    net = _native.MLP(dims, device, activation=activation)
    net.set_params(model.flat())
    Xn = ((X - x_mean) / x_std).astype(np.float32)
    Yn = (Y / y_std).astype(np.float32)
    losses = net.train(Xn, Yn, epochs=epochs, batch=batch, lr=lr, seed=seed)
    model.set_flat(net.get_params())
    return model, losses
