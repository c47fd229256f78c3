"""The hierarchy blob (dgh_damg_save / dgh_damg_load): a saved hierarchy loads back with every
array bit-identical, and damaged or foreign files are refused (no GPU)."""
import os

import numpy as np
import pytest

import dendro_amd as da


def _arrays(damg):
    out = []
    for l in range(damg.nlevels):
        L = damg.level(l)
        out += [L.nlist, L.child, L.hnmask, L.lev, L.anchor, damg.boundary(l), *damg.octants(l)]
        if l:
            out += list(damg.transfer(l))
            out += list(damg.coarsen_coefficients(np.arange(damg.level(damg.nlevels - 1).n_elem, dtype=float) + 1.0,
                                                  None, l - 1)[:1])
    return out


def test_round_trip_bit_identical(tmp_path, hier_small):
    d = hier_small.damg
    path = tmp_path / "mesh.dghb"
    d.save(path)
    e = da.DAMG.load(path)
    assert e.nlevels == d.nlevels
    for l in range(d.nlevels):
        a, b = d.info(l), e.info(l)
        assert (a.n_elem, a.n_node, a.n_oct, a.max_depth, a.n_hanging) == (b.n_elem, b.n_node, b.n_oct,
                                                                          b.max_depth, b.n_hanging)
    for x, y in zip(_arrays(d), _arrays(e)):
        assert x.dtype == y.dtype and np.array_equal(x, y)


def test_damaged_files_refused(tmp_path, hier_tiny):
    path = tmp_path / "mesh.dghb"
    hier_tiny.damg.save(path)
    raw = bytearray(path.read_bytes())
    for name, blob in [("truncated", raw[: len(raw) // 2]), ("flipped", raw[:100] + bytes([raw[100] ^ 1]) + raw[101:]),
                       ("version", raw[:8] + bytes([9, 0, 0, 0]) + raw[12:]), ("foreign", b"not a mesh" * 10)]:
        bad = tmp_path / f"{name}.dghb"
        bad.write_bytes(bytes(blob))
        with pytest.raises(da.DendroError):
            da.DAMG.load(bad)
    with pytest.raises(da.DendroError):
        da.DAMG.load(tmp_path / "missing.dghb")
