"""Coherent backscattering in the CPU restatement (NextEventEstimation.cuh:1212-1235,1385-1391,
1432-1454): the reversed path v -> x1 -> ... -> xK -> l of every path, field output.

Pinned by the reciprocity of the terms: when every view is the exact negation of an
illumination (reflection geometry), the view connected at x1 is the illumination connected at
x1 and the illumination connected at x_k is the view connected at x_k, so the reversed-path
terms equal the direct ones index for index, and the rendered field is S + (M + M^T) / sqrt(2)
with M the direct multiple-scattering field."""
import numpy as np
import pytest

import gpu_cases as G
import oracle_lib as O


def _paths(cfg, n):
    d = O.paths(cfg, 3, 0, n)
    return d


def test_reversed_terms_equal_direct_terms_under_negated_views():
    cfg = G.cbs_backscatter(paths=192)
    n = cfg["simulation"]["batchSize"]
    d = _paths(cfg, n)
    ar, br = O.cbs_terms_from_vertices(cfg, 3, d["k"], d["off"], d["pos"], d["mat"], d["rho"], d["xt"], d["len"])
    multi = d["k"] >= 2
    assert multi.sum() > n // 4
    # Ar[p, t, v] == A[p, t, l = v] and Br[p, t, l] == B[p, t, v = l], exactly
    assert np.array_equal(ar, d["a"]) and np.array_equal(br, d["b"])
    assert np.all(ar[~multi] == 0) and np.all(br[~multi] == 0)


def test_rendered_field_is_symmetrised_multiple_scattering():
    cfg = G.cbs_backscatter(paths=160)
    n = cfg["simulation"]["batchSize"]
    d = _paths(cfg, n)
    S = d["s"].sum(axis=0)                                  # [t][v][l]
    M = np.einsum("ptl,ptv->tvl", d["a"], d["b"])            # direct multiple scattering
    expect = S + (M + np.transpose(M, (0, 2, 1))) / np.sqrt(2.0)
    u, it, _ = O.render(cfg, 3)                              # [l][v][t]
    got = np.transpose(u, (2, 1, 0))
    assert np.linalg.norm(got - expect) / np.linalg.norm(expect) < 1e-12
    # and it differs from the plain render by exactly the reversed part
    plain = dict(cfg, tracer={"type": "nee"})
    u0, _, _ = O.render(plain, 3)
    got0 = np.transpose(u0, (2, 1, 0))
    assert np.linalg.norm(got0 - (S + M)) / np.linalg.norm(S + M) < 1e-12


def test_point_source_reversed_terms_follow_the_three_point_function():
    # point sources and sensors on the same face; the reversed terms are checked against a
    # direct evaluation from the vertices: Ar_v = g_v(x1) f(n(x1 - s_v) . n(x2 - x1)) w2
    cfg = G.cbs_backscatter(paths=128, point=True)
    n = cfg["simulation"]["batchSize"]
    d = _paths(cfg, n)
    ar, br = O.cbs_terms_from_vertices(cfg, 3, d["k"], d["off"], d["pos"], d["mat"], d["rho"], d["xt"], d["len"])
    # swapping the roles of the two sources turns reversed terms into direct ones
    swapped = dict(cfg, illumination=cfg["view"], view=cfg["illumination"])
    ds = _paths(swapped, n)
    # same vertices are not guaranteed (first vertices follow the illumination), so compare on
    # the vertices of the original paths through the term function of the swapped problem
    a2, b2, _ = O.terms_from_vertices(swapped, 3, d["k"], d["off"], d["pos"], d["mat"], d["rho"], d["xt"], d["len"])
    multi = d["k"] >= 2
    # Br (illumination connected at x_k>=2) == B of the swapped problem (its view is our illumination)
    assert np.allclose(br, b2, rtol=0, atol=1e-12 * np.abs(b2).max())
    # Ar differs from the swapped A only by the multiple-scattering weight w2, which is built from
    # the sampling source: the ratio is real and one number per path
    r = ar[multi] / np.where(np.abs(a2[multi]) > 0, a2[multi], 1.0)
    per_path = r.reshape(r.shape[0], -1)
    mask = np.abs(a2[multi]).reshape(per_path.shape) > 1e-6 * np.abs(a2).max()
    for row, mk in zip(per_path, mask):
        if mk.any():
            vals = row[mk]
            assert np.allclose(vals.imag, 0, atol=1e-9) and np.allclose(vals.real, vals.real[0], rtol=1e-9)
    del ds


def test_cbs_with_correlation_is_not_supported():
    cfg = G.cbs_backscatter(paths=64)
    cfg["tracer"]["output"] = "correlation"
    with pytest.raises(RuntimeError):
        O.render(cfg, 3)
