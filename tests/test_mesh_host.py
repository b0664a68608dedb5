"""Pin the oracle's marching cubes (CPU only) by the properties of the surface it must
produce: every case cuts exactly its straddling edges with triangles facing from inside
to outside; closed surfaces come out closed and consistently oriented with the right
Euler characteristic (sphere 2, torus 0); enclosed volume and area match the shapes'
closed forms; every vertex is the zero crossing of the grid's trilinear interpolant; and
the library's case table equals the oracle's."""

import ctypes
import math

import numpy as np
import pytest

from oracle import ref_cpu as R


def _pos(k):
    return np.array([(k >> 2) & 1, (k >> 1) & 1, k & 1], float)


def test_case_table_cuts_straddling_edges_facing_outwards():
    table = R.mc_case_table()
    assert table[0] == [] and table[255] == [] and max(len(t) for t in table) == 5
    for c in range(256):
        inside = [(c >> k) & 1 for k in range(8)]
        straddle = {e for e in range(12) if inside[R._mc_edge_corners(e)[0]] != inside[R._mc_edge_corners(e)[1]]}
        used = {e for tri in table[c] for e in tri}
        assert used == straddle, c
        dots = []
        for tri in table[c]:
            P = [sum(_pos(k) for k in R._mc_edge_corners(e)) / 2 for e in tri]
            n = np.cross(P[1] - P[0], P[2] - P[0])
            # no triangle faces towards the inside corners of its three edges
            cor = [k for e in tri for k in R._mc_edge_corners(e)]
            li = np.mean([_pos(k) for k in cor if inside[k]], 0)
            lo = np.mean([_pos(k) for k in cor if not inside[k]], 0)
            dots.append(float(np.dot(n, lo - li)))
        if table[c]:
            assert min(dots) >= 0 and max(dots) > 0, (c, dots)


def _mesh_stats(v, f):
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    directed = {tuple(x) for x in e.tolist()}
    closed = len(directed) == len(e) and all((b, a) in directed for a, b in directed)
    und = {tuple(sorted(x)) for x in e.tolist()}
    euler = v.shape[0] - len(und) + f.shape[0]
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    vol = float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0)
    area = float(np.linalg.norm(np.cross(b - a, c - a), axis=1).sum() / 2.0)
    return closed, euler, vol, area


def _lattice(res, lo=-1.0, hi=1.0):
    ax = np.linspace(lo, hi, res)
    return np.meshgrid(ax, ax, ax, indexing="ij")


def test_sphere_and_torus():
    X, Y, Z = _lattice(48)
    r = 0.6
    sphere = (np.sqrt(X * X + Y * Y + Z * Z) - r).astype(np.float32)
    v, f = R.extract_mesh_marching_cubes(sphere, [-1, -1, -1], [1, 1, 1])
    closed, euler, vol, area = _mesh_stats(v.astype(np.float64), f)
    assert closed and euler == 2
    assert abs(vol / (4 / 3 * math.pi * r ** 3) - 1) < 0.01 and abs(area / (4 * math.pi * r * r) - 1) < 0.02
    Rm, rm = 0.55, 0.2
    torus = (np.sqrt((np.sqrt(X * X + Y * Y) - Rm) ** 2 + Z * Z) - rm).astype(np.float32)
    v, f = R.extract_mesh_marching_cubes(torus, [-1, -1, -1], [1, 1, 1])
    closed, euler, vol, area = _mesh_stats(v.astype(np.float64), f)
    assert closed and euler == 0
    assert abs(vol / (2 * math.pi ** 2 * Rm * rm * rm) - 1) < 0.03


def test_random_fields_close_and_vertices_sit_on_the_crossing():
    rng = np.random.default_rng(0)
    for seed in range(3):
        vol = rng.normal(size=(14, 13, 15)).astype(np.float32)
        vol[0], vol[-1], vol[:, 0], vol[:, -1], vol[:, :, 0], vol[:, :, -1] = 1, 1, 1, 1, 1, 1
        v, f = R.marching_cubes(vol, 0.1)
        closed, _, _, _ = _mesh_stats(v, f)
        assert closed and f.shape[0] > 0
        # the trilinear interpolant along a grid edge is linear: the vertex is its crossing
        i0 = np.floor(v).astype(int)
        frac = v - i0
        axis = np.argmax(frac > 0, axis=1)
        a = vol[i0[:, 0], i0[:, 1], i0[:, 2]].astype(np.float64)
        step = np.eye(3, dtype=int)[axis]
        j = i0 + step
        b = vol[j[:, 0], j[:, 1], j[:, 2]].astype(np.float64)
        t = frac[np.arange(len(v)), axis]
        assert np.allclose(a + t * (b - a), 0.1, atol=1e-6)


def test_library_table_equals_the_oracle():
    from voxneus_amd import _native as N
    tri = (ctypes.c_int8 * (256 * 16))()
    nt = (ctypes.c_uint8 * 256)()
    N.check(N.lib().vxn_mc_case_table(tri, nt))
    tri = np.frombuffer(tri, np.int8).reshape(256, 16)
    nt = np.frombuffer(nt, np.uint8)
    for c, tris in enumerate(R.mc_case_table()):
        assert int(nt[c]) == len(tris)
        assert tri[c][:3 * nt[c]].reshape(-1, 3).tolist() == [list(t) for t in tris]
        assert (tri[c][3 * nt[c]:] == -1).all()


@pytest.mark.parametrize("sigma", [0.1, 0.5, 1.0, 1.5, 2.3, 7.9])
def test_gaussian_weights_are_scipys(sigma):
    """The oracle's and the library's weights are scipy's own (its private kernel builder,
    read here only as the checker)."""
    from scipy.ndimage import _filters
    from voxneus_amd.functional import gaussian_weights
    w, r = R.gaussian_weights(sigma)
    assert r == int(4.0 * sigma + 0.5)
    ref = _filters._gaussian_kernel1d(sigma, 0, r)
    assert np.array_equal(w, ref)
    lw, lr = gaussian_weights(sigma)
    assert lr == r and np.array_equal(lw, ref[::-1]) and np.array_equal(lw, lw[::-1])


def _separable(vol, sigma):
    """The device kernel's algorithm in numpy: per axis (x, y, z), a double-precision
    correlation with edge replication, the centre term first and then the symmetric pairs
    from the outermost inwards, rounded to float32 after each axis."""
    w, r = R.gaussian_weights(sigma)
    out = vol.astype(np.float32)
    for ax in range(3):
        a = np.moveaxis(out.astype(np.float64), ax, -1)
        n = a.shape[-1]
        pad = np.concatenate([np.repeat(a[..., :1], r, -1), a, np.repeat(a[..., -1:], r, -1)], -1)
        acc = pad[..., r:r + n] * w[r]
        for j in range(r, 0, -1):
            acc = acc + (pad[..., r - j:r - j + n] + pad[..., r + j:r + j + n]) * w[r - j]
        out = np.moveaxis(acc.astype(np.float32), -1, ax)
    return out


@pytest.mark.parametrize("sigma", [0.7, 1.0, 1.5, 2.3])
def test_separable_double_accumulation_is_scipys_filter(sigma):
    rng = np.random.default_rng(int(sigma * 10))
    for shape in ((23, 19, 31), (1, 7, 12), (40, 3, 2)):
        vol = rng.normal(size=shape).astype(np.float32)
        assert np.array_equal(_separable(vol, sigma), R.apply_gaussian_filter_to_grid(vol, sigma))


def test_gaussian_filter_closed_forms():
    """Constants stay constant; away from the edges a linear field is unchanged (the
    kernel is symmetric and sums to 1); at the edges replication bends it."""
    c = np.full((12, 13, 14), 0.375, np.float32)
    assert np.array_equal(R.apply_gaussian_filter_to_grid(c, 1.3), c)
    X = np.broadcast_to(np.arange(30, dtype=np.float32)[:, None, None], (30, 8, 8)).copy()
    out = R.apply_gaussian_filter_to_grid(X, 1.0)
    assert np.abs(out[4:-4] - X[4:-4]).max() < 1e-5
    assert out[0, 0, 0] > X[0, 0, 0]
