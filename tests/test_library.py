"""CPU checks of the C ABI library and the host-side helpers (no GPU needed):
libsfp.so loads and exports every function include/*.h declares, the ctypes
structures match the C layouts, and the scene generators are deterministic."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import sfp
from sfp import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def _declared(header):
    text = open(os.path.join(INCLUDE, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sfp_\w+)\s*\(", text)))


@pytest.mark.parametrize("header", sorted(sfp.EXPORTS))
def test_library_exports_every_declared_function(header):
    lib = sfp.lib()
    names = _declared(header)
    assert names, header
    assert sorted(sfp.EXPORTS[header]) == names
    for n in names:
        assert hasattr(lib, n), n


def test_headers_cover_all_exports():
    out = subprocess.run(["nm", "-D", "--defined-only", sfp.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = sorted(set(re.findall(r" T (sfp_\w+)", out)))
    declared = sorted(set(sum((_declared(h) for h in os.listdir(INCLUDE) if h.endswith(".h")), [])))
    assert exported == declared


def test_struct_layouts_match_c():
    src = r'''
#include <stddef.h>
#include <stdio.h>
#include "sfp.h"
#include "sfp_profile.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(sfp_params), offsetof(sfp_params, max_iters),
         offsetof(sfp_params, cheby_ratio), offsetof(sfp_params, stall_ratio), offsetof(sfp_params, verbose),
         offsetof(sfp_params, coarse_cheby_ratio), offsetof(sfp_params, net_cheby_ratio),
         offsetof(sfp_params, net_coarse_cheby_ratio), offsetof(sfp_params, magnetic_bc),
         offsetof(sfp_params, pressure_coarse_cheby_ratio),
         sizeof(sfp_stats), offsetof(sfp_stats, rel_residual), offsetof(sfp_stats, unknowns),
         sizeof(sfp_kernel_profile), offsetof(sfp_kernel_profile, bytes));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-std=c99", "-Wall", "-I", INCLUDE, c, "-o", exe], check=True)
        vals = [int(v) for v in subprocess.run([exe], capture_output=True, text=True,
                                               check=True).stdout.split()]
    P, S, K = sfp.Params, sfp.Stats, sfp.KernelProfile
    assert vals == [C.sizeof(P), P.max_iters.offset, P.cheby_ratio.offset, P.stall_ratio.offset,
                    P.verbose.offset, P.coarse_cheby_ratio.offset, P.net_cheby_ratio.offset,
                    P.net_coarse_cheby_ratio.offset, P.magnetic_bc.offset, P.pressure_coarse_cheby_ratio.offset,
                    C.sizeof(S), S.rel_residual.offset,
                    S.unknowns.offset,
                    C.sizeof(K), K.bytes.offset]


def test_default_params_and_status_strings():
    p = sfp.default_params()
    assert p.tol == 1e-8 and p.max_iters == 100 and p.precond == sfp.PRECOND_MG
    assert p.mg_coarsest == 4 and p.cheby_degree == 3
    assert p.coarse_cheby_degree == 5 and p.coarse_cheby_ratio == 6.0
    L = sfp.lib()
    for st in range(7):
        assert L.sfp_status_string(st)
    assert L.sfp_status_string(99) == b"unknown status"


def test_create_rejects_bad_arguments_before_touching_a_device():
    L = sfp.lib()
    h = C.c_void_p()
    org = (C.c_double * 3)(0, 0, 0)
    assert L.sfp_create(C.byref(h), 0, 4, 4, 0.1, C.cast(org, sfp._dp), 0) == sfp.ERR_ARGUMENT
    assert L.sfp_create(C.byref(h), 4, 4, 4, -1.0, C.cast(org, sfp._dp), 0) == sfp.ERR_ARGUMENT
    assert L.sfp_create(C.byref(h), 2048, 2048, 1024, 0.1, C.cast(org, sfp._dp), 0) == sfp.ERR_ARGUMENT
    assert b"int" in L.sfp_last_error(None)
    assert L.sfp_set_params(None, None) == sfp.ERR_ARGUMENT
    assert L.sfp_pressure_project(None, None, None, None, None, None, 0.0, None, None) == sfp.ERR_ARGUMENT


def test_create_without_gpu_fails_loudly():
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(sfp.SfpError) as e:
        sfp.Context(8, 1 / 8)
    assert e.value.status == sfp.ERR_NO_DEVICE


def test_splitmix64_known_answers():
    # SplitMix64 (Steele, Lea, Flood 2014) from state 0 and 1
    got = scenes.splitmix64(np.array([0, 1], dtype=np.uint64))
    assert int(got[0]) == 0xE220A8397B1DCDAF
    assert int(got[1]) == 0x910A2DEC89025CC1


def test_scenes_are_deterministic_and_counter_based():
    a = scenes.uniform(3, 5, 100)
    b = scenes.uniform(3, 5, 60, start=40)
    assert np.array_equal(a[40:], b)
    assert not np.array_equal(a, scenes.uniform(4, 5, 100))
    assert np.all((a >= 0) & (a < 1))
    s1 = scenes.pressure_scene(12, seed=1)
    s2 = scenes.pressure_scene(12, seed=1)
    for k in range(3):
        assert np.array_equal(s1["velocity"][k], s2["velocity"][k])
    liquid = (s1["phi"] <= 0).mean()
    assert 0.25 < liquid < 0.40


def test_geometry_matches_staggered_grid_construction():
    g = scenes.Geometry((20, 18, 22), 1.0 / 22, center=(0.41, 0.52, 0.47))
    for a, (c, r) in enumerate(zip((0.41, 0.52, 0.47), (20, 18, 22))):
        assert g.origin[a] == c - (float(r) * (1.0 / 22)) / 2
    X, Y, Z = g.cell_centers()
    assert X[0, 0, 0] == g.origin[0] + (1.0 / 22) / 2
    assert g.face_shape(1) == (20, 19, 22)


def test_trained_weights_load_and_checksum():
    from sfp import learn
    params, meta = learn.load_weights()
    assert params.dtype == np.float32 and params.size == meta["count"] == 321745
    assert np.all(np.isfinite(params))
    assert meta["n"] == 128 and meta["samples"] > 0
    sm = learn.smoother(meta)
    assert set(sm) == {"net_" + k for k in learn.SMOOTHER_KEYS} and all(v > 0 for v in sm.values())
    lp = learn.learned_params(meta)
    assert lp["precond"] == sfp.PRECOND_LEARNED and 0 < lp["stall_ratio"] <= 1
    # the learned cycle's smoother is a separate set of parameters: the plain
    # cycle keeps the library's defaults for the fall-back
    p = sfp.default_params()
    assert p.net_cheby_degree == 0 and p.net_coarse_cheby_ratio == 0.0
    assert p.magnetic_bc == sfp.MAGNETIC_BC_REFERENCE
    assert (p.pressure_coarse_cheby_degree, p.pressure_coarse_cheby_ratio) == (3, 4.0)
    assert (p.coarse_cheby_degree, p.coarse_cheby_ratio) == (5, 6.0)


def test_apply_full_matches_csr():
    from sfp import learn
    from helpers import csr_from_full_grid
    rng = np.random.default_rng(3)
    shape = (5, 4, 6)
    diag = rng.uniform(4, 6, shape)
    cpl = [rng.uniform(0, 1, shape) for _ in range(3)]
    for a in range(3):
        sl = [slice(None)] * 3
        sl[a] = 0
        cpl[a][tuple(sl)] = 0.0
    x = rng.normal(size=shape)
    A = csr_from_full_grid(diag, cpl)
    np.testing.assert_allclose(learn.apply_full(diag, cpl, x).ravel(), A @ x.ravel(), rtol=1e-14, atol=1e-14)


def _unbound_context(n):
    """A Context object that never reached sfp_create: any library call on it
    fails, so a ValueError proves the check ran before the call."""
    ctx = object.__new__(sfp.Context)
    ctx.n = tuple(n)
    ctx._h = C.c_void_p()
    return ctx


def test_wrapper_rejects_arrays_that_do_not_fit_the_grid():
    ctx = _unbound_context((6, 5, 4))
    good = scenes.random_pressure_scene((6, 5, 4), seed=1)
    bad_cells = scenes.random_pressure_scene((6, 5, 5), seed=1)
    wrong_phi = dict(good, phi=good["phi"][:, :, :3])
    wrong_face = dict(good, velocity=[good["velocity"][0], good["velocity"][2], good["velocity"][1]])
    for sc in (bad_cells, wrong_phi, wrong_face, dict(good, fraction=good["fraction"][:2])):
        with pytest.raises(ValueError):
            ctx.pressure_project(sc)
    msc = dict(good, hext=[np.zeros(s) for s in ((7, 5, 4), (6, 6, 4), (6, 5, 4))], chi=2.0, band=0.1)
    with pytest.raises(ValueError):
        ctx.magnetic_solve(msc)
    with pytest.raises(ValueError):
        ctx.magnetic_solve(dict(msc, phi=np.zeros((6, 5, 3))))
    with pytest.raises(ValueError):
        ctx.vertex_pressures(np.zeros((4, 2)), np.zeros((4, 2)), 1.0)
    with pytest.raises(ValueError):
        ctx.vertex_pressures(np.zeros((4, 3)), np.zeros((3, 3)), 1.0)
    faces = [np.zeros(ctx._face_shape(a)) for a in range(3)]
    with pytest.raises(ValueError):
        ctx.extrapolate_faces(faces, [np.zeros((6, 5, 4), np.uint8)] * 3)
    with pytest.raises(ValueError):
        ctx.extrapolate_velocity(faces, np.zeros((5, 5, 4)), faces)
    with pytest.raises(ValueError):
        ctx.collider_enforce(faces, faces, faces[::-1], faces)
    for hook in (ctx.apply64, ctx.precondition, ctx.residual):
        with pytest.raises(ValueError):
            hook(0, np.zeros((6, 5, 5)))
    # arrays that fit get past the checks and reach the (unbound) library
    with pytest.raises(sfp.SfpError):
        ctx.pressure_project(good)


def test_noalloc_extrapolation_rejects_before_touching_a_device():
    """The entry point without allocation checks its arguments, including
    that the grid fits its fixed 64 MiB scratch, before any HIP call."""
    L = sfp.lib()
    dummy = (C.c_double * 1)()
    p3 = sfp.P3(*[C.cast(dummy, sfp._dp)] * 3)
    d = C.cast(dummy, sfp._dp)
    f = L.sfp_extrapolate_velocity_noalloc
    # 330^3: each face grid ~36 M faces, two bytes per face do not fit
    assert f(330, 330, 330, p3, d, p3, 6, None) == sfp.ERR_ARGUMENT
    assert b"64 MiB" in L.sfp_last_error(None)
    assert f(0, 4, 4, p3, d, p3, 6, None) == sfp.ERR_ARGUMENT
    assert f(4, 4, 4, p3, d, p3, -1, None) == sfp.ERR_ARGUMENT
    assert f(4, 4, 4, None, d, p3, 6, None) == sfp.ERR_ARGUMENT
    assert f(4, 4, 4, p3, None, p3, 6, None) == sfp.ERR_ARGUMENT
    assert f(2 ** 30, 2 ** 30, 2 ** 30, p3, d, p3, 6, None) == sfp.ERR_ARGUMENT
