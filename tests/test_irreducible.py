"""Symmetry-adapted third-order coefficients (tdep_amd/irreducible.py) on the CPU.

The basis is pinned independently of its construction: random members of the span obey
every space-group and permutation relation and the acoustic sum rule explicitly, the
generated Morse force constants (symmetric by construction) lie in the span, and for the
FCC cell the dimension equals that of the null space of all constraints stacked and solved
by brute force."""
import itertools

import numpy as np
import pytest

from tdep_amd import irreducible as IR


def _engine(inputs_dir, name, grid=(2, 2, 2)):
    from tdep_amd.engine import Engine
    return Engine.from_directory(inputs_dir(name), grid, device=-1)


def _relations(e):
    """Explicit (t, t', 27x27 map) relations m_t' = G m_t of the list."""
    p, sy, fc3 = e.structure, e.symmetry(), e.fc3
    L = p.latvec
    Li = np.linalg.inv(L)
    rc = p.rfrac @ L.T
    key = {}
    for t in range(len(fc3["a1"])):
        k = (fc3["a1"][t], fc3["a2"][t], fc3["a3"][t], tuple(np.round(Li @ fc3["lv2"][t]).astype(int)),
             tuple(np.round(Li @ fc3["lv3"][t]).astype(int)))
        key[k] = t
    rel = []
    for R, tau, fm in zip(sy["rot"], sy["tr"], sy["fmap"]):
        M = np.einsum("ia,jb,kc->ijkabc", R, R, R).reshape(27, 27)
        for t, k in enumerate(key):
            a, b, c, l2, l3 = k
            sh = [R @ rc[x] + tau - rc[fm[x]] for x in (a, b, c)]
            n2 = tuple(np.round(Li @ (sh[1] + R @ (L @ np.array(l2)) - sh[0])).astype(int))
            n3 = tuple(np.round(Li @ (sh[2] + R @ (L @ np.array(l3)) - sh[0])).astype(int))
            rel.append((key[k], key[(fm[a], fm[b], fm[c], n2, n3)], M))
    for perm in itertools.permutations(range(3)):
        P = np.zeros((27, 27))
        for x in itertools.product(range(3), repeat=3):
            y = [0, 0, 0]
            for s in range(3):
                y[perm[s]] = x[s]
            P[9 * x[0] + 3 * x[1] + x[2], 9 * y[0] + 3 * y[1] + y[2]] = 1
        for k, t in key.items():
            at, lv = k[:3], (np.zeros(3, int), np.array(k[3]), np.array(k[4]))
            s0 = lv[perm[0]]
            k2 = (at[perm[0]], at[perm[1]], at[perm[2]], tuple(lv[perm[1]] - s0), tuple(lv[perm[2]] - s0))
            rel.append((t, key[k2], P))
    return rel, key


@pytest.mark.parametrize("name", ["fcc_al", "rocksalt"])
def test_span_obeys_every_relation_and_the_sum_rule(inputs_dir, name):
    e = _engine(inputs_dir, name)
    B = IR.engine_basis(e, asr=True)
    assert np.allclose(B.T @ B, np.eye(B.shape[1]), atol=1e-12)
    rel, key = _relations(e)
    x = np.random.default_rng(1).standard_normal(B.shape[1])
    m = (B @ x).reshape(-1, 27)
    worst = max(np.abs(m[t2] - G @ m[t1]).max() for t1, t2, G in rel)
    assert worst < 1e-12 * np.abs(m).max()
    groups = {}
    for k, t in key.items():
        groups.setdefault((k[0], k[1], k[3]), []).append(t)
    assert max(np.abs(m[ts].sum(axis=0)).max() for ts in groups.values()) < 1e-12 * np.abs(m).max()
    # the generated (symmetric, sum-rule obeying) force constants lie in the span
    m0 = e.fc3["m"].reshape(-1)
    assert np.linalg.norm(m0 - B @ (B.T @ m0)) < 1e-12 * np.linalg.norm(m0)


def test_dimension_equals_brute_force_null_space(inputs_dir):
    e = _engine(inputs_dir, "fcc_al")
    rel, key = _relations(e)
    nt = len(key)
    # A^T A of the stacked constraints, block by block: a relation m_t2 - G m_t1 = 0
    # contributes I at (t2, t2), G^T G at (t1, t1) and -G, -G^T off the diagonal; a sum-rule
    # group contributes I at every pair of its triplets
    AtA = np.zeros((27 * nt, 27 * nt))
    blk = lambda i, j: (slice(27 * i, 27 * i + 27), slice(27 * j, 27 * j + 27))
    I = np.eye(27)
    for t1, t2, G in rel:
        AtA[blk(t2, t2)] += I
        AtA[blk(t1, t1)] += G.T @ G
        AtA[blk(t2, t1)] -= G
        AtA[blk(t1, t2)] -= G.T
    groups = {}
    for k, t in key.items():
        groups.setdefault((k[0], k[1], k[3]), []).append(t)
    for ts in groups.values():
        for t in ts:
            for u in ts:
                AtA[blk(t, u)] += I
    w = np.linalg.eigvalsh(AtA)
    dim = int(np.sum(w < 1e-9 * w.max()))
    assert dim == IR.engine_basis(e, asr=True).shape[1]
