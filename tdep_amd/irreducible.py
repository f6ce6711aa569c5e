"""Symmetry-adapted coefficients of the third-order force constants.

TDEP parametrises Phi3 by irreducible coefficients x: every triplet tensor is a fixed linear
image of x (type_forcemap_returntensors.f90:294-335 expands them; the coefficients come
from extract_forceconstants, outfile.irrifc_thirdorder).  The force map that fixes TDEP's
own choice of basis is outside this path, so this module builds a basis of the same space
from the IFC list itself: the tensors m[t, i, j, k] (the layout of the engine's fc3 list)
that are

  * invariant under the space group: the image of triplet t under (R, tau) is a triplet t'
    of the list with m_t' = (R x R x R) m_t;
  * invariant under permutations of the three atoms of a triplet (the same tensor seen
    from another atom of the cluster);
  * optionally, obey the acoustic sum rule over the third atom,
    sum_(a3, R3) m(a1, a2 R2, a3 R3)[i, j, k] = 0 for every (a1, a2 R2, i, j, k).

`fc3_basis` returns an orthonormal B (27 ntrip x ncoef) with m = B x for every tensor list in
that space; gradients with respect to x are B^T dL/dm, which torch forms by itself when B x
is fed to the autograd functions of tdep_amd.autograd.  A tensor list is projected onto
the space by B B^T.
"""
import itertools

import numpy as np


def _frac_key(Linv, v):
    f = Linv @ v
    r = np.round(f)
    if np.abs(f - r).max() > 1e-6:
        raise ValueError("lattice vector off the lattice")
    return tuple(int(x) for x in r)


def _perm_matrix(p):
    """27 x 27 matrix of the permuted tensor: new[x0, x1, x2] = old[y] with y[p[s]] = x[s]."""
    P = np.zeros((27, 27))
    for x in itertools.product(range(3), repeat=3):
        y = [0, 0, 0]
        for s_ in range(3):
            y[p[s_]] = x[s_]
        P[9 * x[0] + 3 * x[1] + x[2], 9 * y[0] + 3 * y[1] + y[2]] = 1.0
    return P


def _null_space(A, tol):
    n = A.shape[1]
    if A.shape[0] == 0:
        return np.eye(n)
    # only the right singular vectors are needed (no U of a tall constraint stack)
    _, s, vt = np.linalg.svd(A, full_matrices=A.shape[0] < n)
    s = np.concatenate([s, np.zeros(max(0, vt.shape[0] - len(s)))])
    rank = int(np.sum(s > tol * max(1.0, s[0])))
    return vt[rank:].T


def fc3_basis(latvec, rcart, rot, tr, fmap, fc3, asr=True, tol=1e-8):
    """Orthonormal basis B [27 * ntrip, ncoef] of the symmetric third-order tensor lists.

    latvec: (3, 3) columns are lattice vectors; rcart: (na, 3) Cartesian positions;
    rot, tr, fmap: space-group operations (Cartesian rotation, translation, atom map with
    R r_a + tau = r_fmap[a] + lattice vector); fc3: dict with a1, a2, a3 (ntrip), lv2, lv3
    (ntrip, 3) Cartesian lattice vectors.  Everything in one length unit."""
    a1, a2, a3 = (np.asarray(fc3[k], int) for k in ("a1", "a2", "a3"))
    lv2, lv3 = np.asarray(fc3["lv2"], float), np.asarray(fc3["lv3"], float)
    nt = len(a1)
    Linv = np.linalg.inv(latvec)
    key = {}
    for t in range(nt):
        k = (a1[t], a2[t], a3[t], _frac_key(Linv, lv2[t]), _frac_key(Linv, lv3[t]))
        if k in key:
            raise ValueError("repeated triplet in the IFC list")
        key[k] = t
    # generators: (image triplet, 27 x 27 tensor map) for every triplet
    gens = []
    for R, tau, fm in zip(rot, tr, fmap):
        Lsh = np.array([R @ rcart[a] + tau - rcart[fm[a]] for a in range(len(rcart))])
        M = np.kron(np.kron(R, R), R)
        img = np.empty(nt, int)
        for t in range(nt):
            b1, b2, b3 = fm[a1[t]], fm[a2[t]], fm[a3[t]]
            l2 = Lsh[a2[t]] + R @ lv2[t] - Lsh[a1[t]]
            l3 = Lsh[a3[t]] + R @ lv3[t] - Lsh[a1[t]]
            k = (b1, b2, b3, _frac_key(Linv, l2), _frac_key(Linv, l3))
            if k not in key:
                raise ValueError("IFC list not closed under the space group")
            img[t] = key[k]
        gens.append((img, M))
    for p in itertools.permutations(range(3)):
        if p == (0, 1, 2):
            continue
        P = _perm_matrix(p)
        img = np.empty(nt, int)
        for t in range(nt):
            at = (a1[t], a2[t], a3[t])
            lv = (np.zeros(3), lv2[t], lv3[t])
            s0 = lv[p[0]]
            k = (at[p[0]], at[p[1]], at[p[2]], _frac_key(Linv, lv[p[1]] - s0), _frac_key(Linv, lv[p[2]] - s0))
            if k not in key:
                raise ValueError("IFC list not closed under permutations of the triplet")
            img[t] = key[k]
        gens.append((img, P))
    # orbits: transport map T_u (m_u = T_u m_rep) by breadth-first search; every further
    # generator edge inside an orbit is a constraint on the representative
    seen = np.full(nt, -1)
    cols = []
    for r in range(nt):
        if seen[r] >= 0:
            continue
        T = {r: np.eye(27)}
        seen[r] = r
        queue = [r]
        cons = []
        while queue:
            u = queue.pop()
            for img, G in gens:
                v = img[u]
                Tv = G @ T[u]
                if v in T:
                    d = T[v] - Tv
                    if np.abs(d).max() > tol:
                        cons.append(d)
                else:
                    T[v] = Tv
                    seen[v] = r
                    queue.append(v)
        V = _null_space(np.vstack(cons) if cons else np.zeros((0, 27)), tol)
        for c in range(V.shape[1]):
            col = np.zeros(27 * nt)
            for u, Tu in T.items():
                col[27 * u:27 * u + 27] = Tu @ V[:, c]
            cols.append(col)
    B = np.array(cols).T if cols else np.zeros((27 * nt, 0))
    if asr and B.shape[1]:
        # sum over the third atom at fixed (a1, a2, lv2), every component
        groups = {}
        for t in range(nt):
            groups.setdefault((a1[t], a2[t], _frac_key(Linv, lv2[t])), []).append(t)
        CB = np.zeros((27 * len(groups), B.shape[1]))
        for g, ts in enumerate(groups.values()):
            for t in ts:
                CB[27 * g:27 * g + 27] += B[27 * t:27 * t + 27]
        N = _null_space(CB, 1e-10)
        B = B @ N
    # orthonormal columns
    if B.shape[1]:
        q, rr = np.linalg.qr(B)
        keep = np.abs(np.diag(rr)) > 1e-10 * max(1.0, np.abs(np.diag(rr)).max())
        B = q[:, keep]
    return B


def engine_basis(engine, asr=True):
    """fc3_basis for the structure, symmetry and IFC list an Engine holds (Bohr)."""
    p = engine.structure
    sy = engine.symmetry()
    rc = p.rfrac @ p.latvec.T
    return fc3_basis(p.latvec, rc, sy["rot"], sy["tr"], sy["fmap"], engine.fc3, asr=asr)
