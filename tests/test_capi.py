"""The C-ABI library loads and exports every entry point include/voxneus_hip.h declares;
argument validation rejects bad calls on the host (no GPU needed: the checks run before
any launch)."""

import ctypes
import os
import re

import pytest

from voxneus_amd import _native as N

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(REPO, "include", "voxneus_hip.h")).read()
    return sorted(set(re.findall(r"\b(vxn_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(N.exported_symbols())


def test_library_exports_every_declared_symbol():
    L = N.lib()
    for name in declared_symbols():
        assert hasattr(L, name), name
    assert L.vxn_version().decode().startswith("voxneus_hip")


def test_struct_layouts_match_header_field_order():
    text = open(os.path.join(REPO, "include", "voxneus_hip.h")).read()
    for struct, cls in (("vxn_sampler_args", N.SamplerArgs), ("vxn_render_args", N.RenderArgs),
                        ("vxn_scene_desc", N.SceneDesc), ("vxn_serve_render_args", N.ServeRenderArgs),
                        ("vxn_composite_args", N.CompositeArgs)):
        body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (struct, struct), text, re.S).group(1)
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        names = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            decl = re.sub(r"\[[^\]]*\]", "", decl)
            for part in decl.split(","):
                names.append(re.findall(r"[A-Za-z_][A-Za-z0-9_]*", part)[-1])
        assert names == [f[0] for f in cls._fields_], struct


def test_host_validation_errors():
    L = N.lib()
    box = N.box6((-1, -1, -1), (1, 1, 1))
    rc = L.vxn_grid_query_fwd(None, 8, 8, 8, box, None, 1, None, None, None, None, None, 0, None)
    assert rc == -1 and b"grid" in L.vxn_last_error()
    rc = L.vxn_grid_query_fwd(ctypes.c_void_p(16), 1, 8, 8, box, ctypes.c_void_p(16), 1, None, None, None, None, None,
                              0, None)
    assert rc == -2
    bad = N.box6((1, -1, -1), (-1, 1, 1))
    rc = L.vxn_grid_query_fwd(ctypes.c_void_p(16), 8, 8, 8, bad, ctypes.c_void_p(16), 1, None, None, None, None, None,
                              0, None)
    assert rc == -3
    # gradient_mode outside {0 interpolated, 1 analytical}: a value error before any launch
    rc = L.vxn_grid_query_fwd(ctypes.c_void_p(16), 8, 8, 8, box, ctypes.c_void_p(16), 1, None, None, None, None, None,
                              2, None)
    assert rc == -3 and b"gradient_mode" in L.vxn_last_error()
    rc = L.vxn_lattice_query(ctypes.c_void_p(16), 8, 8, 8, box, 4, 4, 4, box, ctypes.c_void_p(16),
                             ctypes.c_void_p(16), -1, None)
    assert rc == -3 and b"gradient_mode" in L.vxn_last_error()
    res = (ctypes.c_int32 * 16)(*([16] * 16))
    rc = L.vxn_hash_fwd(ctypes.c_void_p(16), 0, 16, 15, 4, res, None, ctypes.c_void_p(16), 1, ctypes.c_void_p(16),
                        None)
    assert rc == -4 and b"features" in L.vxn_last_error()
    a = N.RenderArgs()
    a.grid, a.X, a.Y, a.Z, a.box = 16, 8, 8, 8, box
    a.n_rays, a.n_samples = 4, 200
    assert L.vxn_render_forward(ctypes.byref(a), None) == -4
    a.n_samples, a.n_levels, a.n_features = 128, 16, 2
    a.mlp_mode = 0
    assert L.vxn_render_forward(ctypes.byref(a), None) == -1  # null rays
    nb = ctypes.c_int64(0)
    assert L.vxn_render_workspace_bytes(ctypes.byref(a), ctypes.byref(nb)) == 0 and nb.value > 0
    s = N.SamplerArgs()
    s.grid, s.X, s.Y, s.Z, s.box = 16, 8, 8, 8, box
    s.n_rays, s.n_samples, s.n_importance, s.up_sample_steps = 4, 64, 128, 4
    assert L.vxn_upsample(ctypes.byref(s), None) == -4  # 64 + 4*32 > 128
    eps = N.f3((0.1, 0.1, 0.1))
    assert L.vxn_reg_eik_curv(ctypes.c_void_p(16), 8, 8, 8, eps, ctypes.c_void_p(16), ctypes.c_void_p(16), 0.1, 0.1,
                              ctypes.c_void_p(16), ctypes.c_void_p(16), None, 0, None) == -5
    assert L.vxn_reg_eik_curv(ctypes.c_void_p(16), 8, 8, 8, N.f3((0.1, 0.0, 0.1)), ctypes.c_void_p(16),
                              ctypes.c_void_p(16), 0.1, 0.1, ctypes.c_void_p(16), ctypes.c_void_p(16), None, 0,
                              None) == -3


def test_serving_validation_errors():
    L = N.lib()
    nb = ctypes.c_int64(0)
    assert L.vxn_serve_scene_bytes(ctypes.byref(nb)) == 0 and nb.value > 0
    d = N.SceneDesc()
    buf = (ctypes.c_uint8 * nb.value)()
    assert L.vxn_serve_pack_scene(ctypes.byref(d), buf, nb.value) == -1  # null grid
    d.grid, d.X, d.Y, d.Z = 16, 8, 8, 8
    d.box = (ctypes.c_float * 6)(-1, -1, -1, 1, 1, 1)
    d.table, d.mlp, d.n_levels, d.n_features, d.log2_T = 16, 16, 8, 2, 15
    assert L.vxn_serve_pack_scene(ctypes.byref(d), buf, nb.value) == -4  # 8 levels: unsupported
    d.n_levels = 16
    for l in range(16):
        d.level_res[l] = 16 + l
    assert L.vxn_serve_pack_scene(ctypes.byref(d), buf, nb.value - 1) == -5  # record too small
    assert L.vxn_serve_pack_scene(ctypes.byref(d), buf, nb.value) == 0
    d.table = 24  # an FP32 table needs 8-byte alignment, a BF16 one 16-byte
    assert L.vxn_serve_pack_scene(ctypes.byref(d), buf, nb.value) == 0
    d.table_dtype = N.VXN_DTYPE_BF16
    assert L.vxn_serve_pack_scene(ctypes.byref(d), buf, nb.value) == -3
    d.table, d.table_dtype = 20, N.VXN_DTYPE_F32
    assert L.vxn_serve_pack_scene(ctypes.byref(d), buf, nb.value) == -3
    d.table = 16
    a = N.ServeRenderArgs()
    a.capacity, a.n_samples, a.n_importance, a.up_sample_steps = 4, 64, 64, 4
    assert L.vxn_serve_render_workspace_bytes(ctypes.byref(a), ctypes.byref(nb)) == 0
    assert nb.value >= 3 * 4 * 128 * 4
    assert L.vxn_serve_render(ctypes.byref(a), None) == -1  # null pointers
    a.n_samples = 100
    assert L.vxn_serve_render(ctypes.byref(a), None) == -4  # 164 samples per ray
    assert L.vxn_serve_move(None, None, 0, 4, 4, 1, None, None, 3, None, None, 0, None, None, None) == -3
    assert L.vxn_serve_move(None, None, 8, 4, 4, 1, None, None, 3, None, None, 0, None, None, None) == -1
    assert L.vxn_serve_query(None, 1, None, None, 4, None, None, None, None) == -1



def test_noalloc_regulariser_refuses_grids_past_its_scratch():
    """The entry point without a workspace holds partial sums for grids of up to 2^27
    vertices; a larger grid is refused on the host before anything is launched."""
    L = N.lib()
    p = ctypes.c_void_p(16)
    eps = N.f3((2.0 / 599,) * 3)
    rc = L.vxn_reg_eik_curv_vr_noalloc(p, 600, 600, 600, eps, N.box6((-1, -1, -1), (1, 1, 1)), 1.2, p, 0.1, 5e-4, p,
                                       p, p, None)
    assert rc == -2 and b"at most" in L.vxn_last_error()
