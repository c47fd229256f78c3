"""Stencil tables and the element-type map against the reference's own data files.

Golden fixtures (tests/golden/make_golden.py): the reference's Stencils/*.inp tables
(printed with 15 decimals, so |file - exact| <= 5e-16) and the hanging masks of
include/oda/cnumEtypes.h.  Both the product generator (closed-form moments) and the oracle
(the Matlab recipe with 2-point Gauss quadrature) are checked against them, against each
other and against mathematical invariants.
"""
import json
import os

import numpy as np
import pytest

import dendro_amd as da
import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(os.path.join(GOLD, "stencils_ref.npz")))


@pytest.fixture(scope="module")
def masks():
    with open(os.path.join(GOLD, "etype_masks.json")) as f:
        return {int(k): v for k, v in json.load(f).items()}


@pytest.fixture(scope="module", params=["product", "oracle"])
def tables(request):
    return da.stencils() if request.param == "product" else oracle.stencils()


def test_tables_match_reference_files(tables, gold):
    for name in ("lap2", "mass2"):
        assert np.abs(tables[name] - gold[name]).max() <= 1e-15, name
    for name in ("rmat2", "rmat1"):
        assert np.abs(tables[name] - gold[name]).max() <= 2e-16, name


def test_product_matches_oracle():
    p, o = da.stencils(), oracle.stencils()
    for k in p:
        assert np.abs(p[k] - o[k]).max() <= 1e-15, k


def test_etype_map_matches_reference(masks):
    for c in range(8):
        for m in range(256):
            if m in masks[c]:
                assert da.etype(c, m) == masks[c].index(m)
                assert oracle.etype(c, m) == masks[c].index(m)
            else:
                with pytest.raises(da.DendroError):
                    da.etype(c, m)
                assert oracle.etype(c, m) < 0


def test_etype_masks_are_the_18_balanced_configurations(masks):
    """For child c, vertex c (parent corner) and 7-c (parent centre) never hang; a hanging
    face centre (c^3, c^5, c^6) needs both adjacent edge midpoints hanging: 18 subsets."""
    for c in range(8):
        valid = []
        for m in range(256):
            if m & (1 << c) or m & (1 << (7 - c)):
                continue
            ok = True
            for f, (e1, e2) in ((3, (1, 2)), (5, (1, 4)), (6, (2, 4))):
                if m & (1 << (c ^ f)) and not (m & (1 << (c ^ e1)) and m & (1 << (c ^ e2))):
                    ok = False
            if ok:
                valid.append(m)
        assert sorted(valid) == sorted(masks[c])


def test_invariants(tables):
    L, M, R2, R1 = tables["lap2"], tables["mass2"], tables["rmat2"], tables["rmat1"]
    assert np.abs(L - L.transpose(0, 1, 3, 2)).max() <= 1e-15
    assert np.abs(M - M.transpose(0, 1, 3, 2)).max() <= 1e-15
    assert np.abs(L.sum(axis=3)).max() <= 1e-14          # Laplacian kills constants
    assert np.abs(M.sum(axis=(2, 3)) - 8.0).max() <= 1e-13  # volume of [-1,1]^3
    for c in range(8):
        assert np.abs(R2[c, 0] - np.eye(8)).max() <= 1e-15
        assert np.abs(L[c, 0] - L[0, 0]).max() == 0.0 or np.abs(L[c, 0] - L[0, 0]).max() <= 1e-15
    assert np.abs(R1.sum(axis=3) - 1.0).max() <= 1e-14   # partition of unity at fine vertices
    assert np.abs(R2.sum(axis=2) - 1.0).max() <= 1e-14


def test_regular_element_is_textbook_q1(tables):
    """eType 0: the trilinear Q1 element on [-1,1]^3, i.e. Kronecker products of the 1D
    stiffness [[1/2,-1/2],[-1/2,1/2]] and mass [[2/3,1/3],[1/3,2/3]] matrices."""
    k1 = np.array([[0.5, -0.5], [-0.5, 0.5]])
    m1 = np.array([[2.0, 1.0], [1.0, 2.0]]) / 3.0
    # vertex k = 4z + 2y + x  ->  kron(z, y, x)
    lap = np.kron(np.kron(m1, m1), k1) + np.kron(np.kron(m1, k1), m1) + np.kron(np.kron(k1, m1), m1)
    mass = np.kron(np.kron(m1, m1), m1)
    assert np.abs(tables["lap2"][0, 0] - lap).max() <= 1e-15
    assert np.abs(tables["mass2"][0, 0] - mass).max() <= 1e-15


def test_child_mirror_symmetry(tables, masks):
    """K[c][t][k][j] = K[0][t'][k^c][j^c] (t': child-0 type of the mirrored mask), which lets
    the device keep only the 18 child-0 stencils."""
    for c in range(8):
        for t in range(18):
            H = masks[c][t]
            Hp = sum(1 << (k ^ c) for k in range(8) if H >> k & 1)
            tp = masks[0].index(Hp)
            perm = [k ^ c for k in range(8)]
            for name in ("lap2", "mass2"):
                K = tables[name]
                assert np.abs(K[c, t] - K[0, tp][np.ix_(perm, perm)]).max() <= 1e-15
