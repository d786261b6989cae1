"""CPU-side checks of the C ABI: the library loads, exports every symbol that
include/lp2d.h declares, and rejects bad arguments before touching a device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "lp2d.h")


def header_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(lp2d_[a-z0-9_]+)\s*\(", src)))


@pytest.fixture(scope="module")
def lib():
    from lightprop2d_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("liblp2d.so not built")
    return _lib.load()


def test_header_declares_the_bound_symbols():
    from lightprop2d_amd import _lib
    assert header_functions() == sorted(_lib.EXPORTED)


def test_every_declared_symbol_is_exported(lib):
    for name in header_functions():
        assert hasattr(lib, name), name


def test_host_only_queries(lib):
    assert lib.lp2d_version() >= 10000
    assert lib.lp2d_max_n(0) == 8192 and lib.lp2d_max_n(1) == 8192 and lib.lp2d_max_n(7) == 0
    buf = C.create_string_buffer(128)
    for n in [2, 16, 128, 1024, 8192]:
        for dt in (0, 1):
            for kind in (0, 1):
                cnt = lib.lp2d_variant_count(n, dt, kind)
                assert cnt >= 1
                for v in range(cnt):
                    assert lib.lp2d_variant_name(n, dt, kind, v, buf, 128) > 0
    assert lib.lp2d_variant_count(3, 0, 0) == 0
    assert lib.lp2d_set_variant(1024, 0, 1, 99) == -1
    # chained-plane row passes (kind 3): complex64 sizes whose plane is block-tiled
    for n in (256, 512, 1024, 4096):
        cnt = lib.lp2d_variant_count(n, 0, 3)
        assert 2 <= cnt <= 4
        for v in range(cnt):
            assert lib.lp2d_variant_name(n, 0, 3, v, buf, 128) > 0
            assert buf.value.startswith(b"chain E=")
    assert lib.lp2d_variant_count(1024, 1, 3) == 0
    assert lib.lp2d_variant_count(128, 0, 3) == 0


def test_argument_errors_before_any_device_work(lib):
    from lightprop2d_amd import _lib
    p = C.c_void_p(16)
    assert lib.lp2d_fft2(3, 1, 0, p, p, 0, None) == _lib.LP2D_E_SIZE
    assert b"power of two" in lib.lp2d_last_error()
    assert lib.lp2d_fft2(16384, 1, 0, p, p, 0, None) == _lib.LP2D_E_SIZE
    assert lib.lp2d_fft2(64, 1, 9, p, p, 0, None) == _lib.LP2D_E_DTYPE
    assert lib.lp2d_fft2(64, 0, 0, p, p, 0, None) == _lib.LP2D_E_SIZE
    assert lib.lp2d_fft2(64, 1, 0, None, p, 0, None) == _lib.LP2D_E_ARG
    o = _lib.lp2d_optics(1.0, 1e-4, 0.1, 0.0, 0.0, 0, 0)
    assert lib.lp2d_propagate(64, 1, 0, p, p, None, 0, None, None, None) == _lib.LP2D_E_ARG
    bad = _lib.lp2d_optics(-1.0, 1e-4, 0.1, 0.0, 0.0, 0, 0)
    assert lib.lp2d_propagate(64, 1, 0, p, p, None, 0, C.byref(bad), None, None) == _lib.LP2D_E_ARG
    assert lib.lp2d_propagate(64, 1, 0, p, p, None, 0, C.byref(o), p, None) == _lib.LP2D_E_ARG
    assert lib.lp2d_propagate_adjoint(64, 1, 0, p, None, p, 0, C.byref(o), p, None, 0,
                                      None) == _lib.LP2D_E_ARG
    assert lib.lp2d_propagate_adjoint(64, 1, 0, p, p, None, 0, C.byref(o), C.c_void_p(32), None, 6,
                                      None) == _lib.LP2D_E_ARG
    # the workspace must not alias grad_out / field_in / grad_in
    q = C.c_void_p(48)
    assert lib.lp2d_propagate_adjoint_ws(64, 1, 0, p, None, None, 0, C.byref(o), q, None, 0, q,
                                         None) == _lib.LP2D_E_ARG
    assert lib.lp2d_propagate_adjoint_ws(64, 1, 0, p, None, None, 0, C.byref(o), q, None, 0, p,
                                         None) == _lib.LP2D_E_ARG
    assert b"work must not alias" in lib.lp2d_last_error()
    m = (_lib.lp2d_mask * 1)(_lib.lp2d_mask(16, 42, 0, 0, 0, 0))
    assert lib.lp2d_spatial_apply(64, 1, 0, p, m, 1, None, None, 0, None, None) == _lib.LP2D_E_DTYPE
    assert lib.lp2d_spectral_apply(64, 1, 0, p, m, 3, None, None, None) == _lib.LP2D_E_ARG
    assert lib.lp2d_intensity(-1, 0, p, p, None) == _lib.LP2D_E_SIZE
    d = (_lib.lp2d_field_desc * 1)(_lib.lp2d_field_desc(48, 0, 16, 16, 1.0, 1e-4, 0.1, 0.0))
    h = C.c_void_p()
    assert lib.lp2d_batch_create(d, 1, 0, C.byref(h)) == _lib.LP2D_E_SIZE
    assert lib.lp2d_batch_create(d, 1, 7, C.byref(h)) == _lib.LP2D_E_DTYPE
    d[0] = _lib.lp2d_field_desc(64, 0, None, 16, 1.0, 1e-4, 0.1, 0.0)
    assert lib.lp2d_batch_create(d, 1, 0, C.byref(h)) == _lib.LP2D_E_ARG


def test_python_layer_rejects_cpu_tensors():
    import torch
    from lightprop2d_amd import propagation as P
    from lightprop2d_amd._lib import Lp2dError, LIB_PATH
    if not os.path.exists(LIB_PATH):
        pytest.skip("liblp2d.so not built")
    with pytest.raises(Lp2dError):
        P.fft2(torch.zeros((8, 8), dtype=torch.complex64))
    with pytest.raises(Lp2dError):
        P.propagate(torch.zeros((8, 8), dtype=torch.complex64), 1e-3, 1e-2, 5e-5)


def test_argument_errors_of_the_generator_resample_and_mode_entry_points(lib):
    from lightprop2d_amd import _lib
    p = C.c_void_p(16)
    prm = _lib.lp2d_profile_params(1, 1, 1, 0.6, 0, 0, 1, 1, 1, 1, 0)
    assert lib.lp2d_profile(99, 8, 8, p, p, C.byref(prm), None, p, 3, None) == _lib.LP2D_E_ARG
    assert lib.lp2d_profile(1, 0, 8, p, p, C.byref(prm), None, p, 3, None) == _lib.LP2D_E_SIZE
    assert lib.lp2d_profile(1, 8, 8, None, p, C.byref(prm), None, p, 3, None) == _lib.LP2D_E_ARG
    # wrong output dtype for the kind (round hole -> bool)
    assert lib.lp2d_profile(2, 8, 8, p, p, C.byref(prm), None, p, 3, None) == _lib.LP2D_E_DTYPE
    # random kinds need the host draw
    assert lib.lp2d_profile(5, 8, 8, p, p, C.byref(prm), None, p, 3, None) == _lib.LP2D_E_ARG
    assert lib.lp2d_expand(64, 63, 1, 0, p, p, None) == _lib.LP2D_E_SIZE
    assert lib.lp2d_expand(64, 129, 1, 0, p, p, None) == _lib.LP2D_E_SIZE
    assert lib.lp2d_expand(64, 128, 1, 5, p, p, None) == _lib.LP2D_E_DTYPE
    assert lib.lp2d_coarse(64, 3, 1, 0, p, p, None) == _lib.LP2D_E_SIZE
    assert lib.lp2d_coarse(64, 2, 1, 0, None, p, None) == _lib.LP2D_E_ARG
    assert lib.lp2d_modes_project(64, 65, 3, 1, 0, p, p, p, None) == _lib.LP2D_E_SIZE
    assert lib.lp2d_modes_project(64, 63, 3, 1, 0, p, p, p, None) == _lib.LP2D_E_SIZE
    # matrix-core projection: workspace size and aliasing are checked before launch
    need = lib.lp2d_modes_project_ws_size(512, 256, 64, 32)
    assert need > 0 and need % 16 == 0
    assert lib.lp2d_modes_project_ws_size(64, 65, 3, 1) == 0
    w = C.c_void_p(64)
    assert lib.lp2d_modes_project_ws(512, 256, 64, 32, 0, p, C.c_void_p(32), C.c_void_p(48), w,
                                     need - 16, None) == _lib.LP2D_E_ARG
    assert b"workspace too small" in lib.lp2d_last_error()
    assert lib.lp2d_modes_project_ws(512, 256, 64, 32, 0, p, C.c_void_p(32), w, w, need,
                                     None) == _lib.LP2D_E_ARG
    assert lib.lp2d_modes_project_ws(64, 63, 3, 1, 0, p, p, p, w, 1 << 20, None) == _lib.LP2D_E_SIZE
    assert lib.lp2d_modes_construct(64, 32, 0, 1, 0, p, p, p, None) == _lib.LP2D_E_SIZE
    assert lib.lp2d_modes_construct(64, 32, 3, 1, 2, p, p, p, None) == _lib.LP2D_E_DTYPE


def test_argument_errors_of_the_interpolation_entry_point(lib):
    from lightprop2d_amd import _lib
    p, q = C.c_void_p(16), C.c_void_p(1 << 40)

    def desc(**kw):
        d = dict(win_off=0, win_len=64, src_i0=-32, dst_i0=-32, dst_sign=1, grid_i0=-32,
                 src_dx=1e-4, dst_dx=1e-4, dst_scale=1.0, scale=1.0, phase_coef=0.0,
                 grid_dl=1e-4)
        d.update(kw)
        return C.byref(_lib.lp2d_interp_desc(**d))

    f = lib.lp2d_interp_linear
    assert f(64, 64, 1, 0, p, q, None, None) == _lib.LP2D_E_ARG
    assert f(64, 0, 1, 0, p, q, desc(), None) == _lib.LP2D_E_SIZE
    assert f(64, 64, 1, 0, p, q, desc(win_len=1), None) == _lib.LP2D_E_SIZE
    assert f(64, 64, 1, 0, p, q, desc(win_off=1), None) == _lib.LP2D_E_SIZE
    assert b"window" in lib.lp2d_last_error()
    assert f(64, 64, 1, 5, p, q, desc(), None) == _lib.LP2D_E_DTYPE
    assert f(64, 64, 1, 0, None, q, desc(), None) == _lib.LP2D_E_ARG
    assert f(64, 64, 1, 0, p, q, desc(dst_sign=0), None) == _lib.LP2D_E_ARG
    assert f(64, 64, 1, 0, p, q, desc(src_dx=0.0), None) == _lib.LP2D_E_ARG
    assert f(64, 64, 1, 0, p, q, desc(scale=float("nan")), None) == _lib.LP2D_E_ARG
    assert f(64, 64, 1, 0, p, q, desc(phase_coef=float("inf")), None) == _lib.LP2D_E_ARG
    assert f(64, 64, 1, 0, p, C.c_void_p(16 + 8 * 64), desc(), None) == _lib.LP2D_E_ARG
    assert b"overlap" in lib.lp2d_last_error()


def test_interpolation_grids_are_the_reference_coordinates():
    """The descriptors of crop / lens_image reproduce the reference's coordinate arrays
    bit for bit: the window's knots old_X and the new grid X (beam2d.py:419-434), and
    the sorted queries X * (-l1/l2) (beam2d.py:525-534)."""
    import numpy as np
    from lightprop2d_amd import propagation as P

    def knots(d):
        return (np.arange(d.win_len) + d.src_i0) * d.src_dx

    def queries(d, m):
        return ((d.dst_sign * np.arange(m) + d.dst_i0) * d.dst_dx) * d.dst_scale

    for n, area, frac, npts in ((256, 200e-4, 0.6, 0), (512, 400e-4, 0.5, 128), (64, 1.0, 0.3, 50)):
        d, m = P.crop_desc(n, area, frac * area, npts)
        X = np.arange(-n // 2, n // 2, 1) * (area / n)
        old = X[np.abs(X) < frac * area / 2]
        h, Nc = len(old) // 2, n // 2
        assert np.array_equal(knots(d), X[Nc - h:Nc + h])
        assert d.win_off == Nc - h and d.win_len == 2 * h
        mm = npts or n
        Xn = np.arange(-mm // 2, mm // 2, 1) * (frac * area / mm)
        assert m == mm and np.array_equal(queries(d, m), Xn)
    with pytest.raises(ValueError):
        P.crop_desc(64, 1.0, 1.0 / 64)
    for n, l1, l2 in ((256, 1.0, 2.0), (128, -1.5, 1.0), (64, 0.0, 1.0), (64, 3.0, 0.7)):
        area, wl = 200e-4, 632e-7
        d = P.lens_image_desc(n, area, wl, l1, l2)
        X = np.arange(-n // 2, n // 2, 1) * (area / n)
        assert np.array_equal(knots(d), X)
        assert np.array_equal(queries(d, n), np.sort(X * (-l1 / l2), kind="mergesort"))
        assert d.scale == -l1 / l2 and d.grid_i0 == -n // 2 and d.grid_dl == area / n


def test_caller_buffers_must_match_the_input():
    """Output / workspace buffers of another size, dtype or layout than the input are
    rejected before any launch (the kernels access batch * n * n elements of the
    input's dtype through them)."""
    import torch
    from lightprop2d_amd import propagation as P
    from lightprop2d_amd._lib import Lp2dError
    ref = torch.zeros((2, 8, 8), dtype=torch.complex64)
    P._check_like(torch.zeros((2, 8, 8), dtype=torch.complex64), ref, "out")
    for bad in (torch.zeros((1, 8, 8), dtype=torch.complex64),
                torch.zeros((2, 8, 8), dtype=torch.complex128),
                torch.zeros((2, 4, 4), dtype=torch.complex64),
                torch.zeros((2, 8, 16), dtype=torch.complex64)[..., ::2]):
        with pytest.raises(Lp2dError):
            P._check_like(bad, ref, "out")
    with pytest.raises(TypeError):
        P._check_like(np.zeros((2, 8, 8)), ref, "out")


def test_uneven_batch_rejects_overlapping_destinations():
    """Destinations may share memory only with their own source, at the same address."""
    from lightprop2d_amd.propagation import check_disjoint_fields
    from lightprop2d_amd._lib import Lp2dError
    MB = 1 << 20
    srcs = [(0, MB), (MB, MB), (2 * MB, MB)]
    check_disjoint_fields(srcs, srcs)                                   # in place
    check_disjoint_fields(srcs, [(10 * MB, MB), (11 * MB, MB), (12 * MB, MB)])
    check_disjoint_fields([(0, MB)] * 3, [(10 * MB, MB), (11 * MB, MB), (12 * MB, MB)])
    with pytest.raises(Lp2dError):      # two fields write the same plane
        check_disjoint_fields(srcs, [(10 * MB, MB), (10 * MB, MB), (12 * MB, MB)])
    with pytest.raises(Lp2dError):      # a field writes another's source
        check_disjoint_fields(srcs, [(0, MB), (0, MB), (2 * MB, MB)])
    with pytest.raises(Lp2dError):      # a shared source overwritten by one of its readers
        check_disjoint_fields([(0, MB)] * 3, [(0, MB), (11 * MB, MB), (12 * MB, MB)])
    with pytest.raises(Lp2dError):      # a partial overlap with its own source
        check_disjoint_fields([(0, MB)], [(MB // 2, MB)])
    with pytest.raises(Lp2dError):      # destinations overlapping in part
        check_disjoint_fields(srcs, [(10 * MB, MB), (10 * MB + 64, MB), (12 * MB, MB)])


def test_noalloc_adjoint_rejects_sizes_beyond_its_scratch(lib):
    """The allocation-free adjoint checks the scratch its call needs before any device
    work: a shared screen over 300 fields of 128^2 (whole-field kernel, one gradient
    plane per field: 300 x 64 KiB) exceeds the 16 MiB array."""
    from lightprop2d_amd import _lib
    p = C.c_void_p(16)
    o = _lib.lp2d_optics(1.0, 1e-4, 0.1, 0.0, 0.0, 0, 0)
    rc = lib.lp2d_propagate_adjoint_noalloc(128, 300, 0, p, C.c_void_p(32), C.c_void_p(48), 0,
                                            C.byref(o), C.c_void_p(64), C.c_void_p(80), 0, None)
    assert rc == _lib.LP2D_E_SIZE
    assert b"scratch" in lib.lp2d_last_error()
    assert lib.lp2d_propagate_adjoint_noalloc(64, 1, 0, p, None, p, 0, C.byref(o), p, None, 0,
                                              None) == _lib.LP2D_E_ARG
    assert lib.lp2d_propagate_adjoint_noalloc(64, 1, 0, p, p, None, 0, C.byref(o),
                                              C.c_void_p(32), p, 0, None) == _lib.LP2D_E_ARG


def test_stack_adjoint_checks_arguments_and_workspace(lib):
    """lp2d_propagate_stack_adjoint validates its plane list and workspace before any
    device work; the workspace query is host-only: per-group gradient planes of a
    shared screen over 4 fields at 512^2, none for one field, -1 for a size that is
    not a power of two."""
    from lightprop2d_amd import _lib
    need = lib.lp2d_propagate_stack_adjoint_ws_size(512, 4, _lib.LP2D_C64)
    assert need > 0 and need % (512 * 512 * 4) == 0
    assert lib.lp2d_propagate_stack_adjoint_ws_size(512, 1, _lib.LP2D_C64) == 0
    assert lib.lp2d_propagate_stack_adjoint_ws_size(500, 4, _lib.LP2D_C64) == -1
    o = _lib.lp2d_optics(1.0, 1e-4, 0.1, 0.0, 0.0, 0, 0)
    g, gin, scr, gs, work = (C.c_void_p(a) for a in (16, 32, 48, 64, 80))
    planes = (C.c_void_p * 2)(96, 112)

    def call(pl, k=2, out=gin, wk=work, wb=need):
        return lib.lp2d_propagate_stack_adjoint(512, 4, _lib.LP2D_C64, g, pl, scr, k,
                                                C.byref(o), out, gs, 0, wk, wb, None)

    assert call(planes, k=0) == _lib.LP2D_E_ARG
    assert call(None) == _lib.LP2D_E_ARG
    assert call((C.c_void_p * 2)(96, None)) == _lib.LP2D_E_ARG
    assert call(planes, out=C.c_void_p(112)) == _lib.LP2D_E_ARG      # grad_in = u_1
    assert call(planes, wb=need - 4) == _lib.LP2D_E_SIZE
    assert call(planes, wk=None) == _lib.LP2D_E_SIZE
    assert call(planes, wk=gin) == _lib.LP2D_E_ARG                   # work = grad_in
    assert call(planes, wk=C.c_void_p(96)) == _lib.LP2D_E_ARG        # work = u_0
    assert call(planes, wk=scr) == _lib.LP2D_E_ARG                   # work = screens


def test_plan_argument_errors(lib):
    """lp2d_plan_create validates its geometry before touching a device; every plan
    call rejects a null plan, and the workspace query answers -1 for one."""
    from lightprop2d_amd import _lib
    h = C.c_void_p(1234)
    assert lib.lp2d_plan_create(500, 1, _lib.LP2D_C64, C.byref(h)) == _lib.LP2D_E_SIZE
    assert h.value is None                       # *out cleared on failure
    assert lib.lp2d_plan_create(512, 0, _lib.LP2D_C64, C.byref(h)) == _lib.LP2D_E_SIZE
    assert lib.lp2d_plan_create(512, 1, 7, C.byref(h)) == _lib.LP2D_E_DTYPE
    assert lib.lp2d_plan_create(512, 1, _lib.LP2D_C64, None) == _lib.LP2D_E_ARG
    assert lib.lp2d_plan_workspace_size(None, _lib.LP2D_OP_ADJOINT) == -1
    assert lib.lp2d_plan_destroy(None) == _lib.LP2D_OK
    p = C.c_void_p(16)
    o = _lib.lp2d_optics(1.0, 1e-4, 0.1, 0.0, 0.0, 0, 0)
    assert lib.lp2d_plan_fft2(None, p, p, 0, None) == _lib.LP2D_E_ARG
    assert lib.lp2d_plan_propagate(None, p, p, None, 0, C.byref(o), None, None) == _lib.LP2D_E_ARG
    assert lib.lp2d_plan_propagate_adjoint(None, p, p, None, 0, C.byref(o), p, None, 0, None, 0,
                                           None) == _lib.LP2D_E_ARG
    assert lib.lp2d_plan_propagate_stack(None, p, p, p, 1, C.byref(o), None,
                                         None) == _lib.LP2D_E_ARG
    assert lib.lp2d_plan_propagate_stack_adjoint(None, p, None, p, 1, C.byref(o), p, None, 0,
                                                 None, 0, None) == _lib.LP2D_E_ARG


def test_phase_grad_argument_errors(lib):
    """lp2d_phase_grad rejects bad geometry, null pointers, a phase stride other than 0
    or n*n and an output aliasing a field, before any launch."""
    from lightprop2d_amd import _lib
    p, q, r = C.c_void_p(16), C.c_void_p(32), C.c_void_p(48)
    assert lib.lp2d_phase_grad(100, 1, 0, p, q, None, 0, r, 1, None) == _lib.LP2D_E_SIZE
    assert lib.lp2d_phase_grad(64, 1, 0, None, q, None, 0, r, 1, None) == _lib.LP2D_E_ARG
    assert lib.lp2d_phase_grad(64, 1, 0, p, q, None, 0, None, 1, None) == _lib.LP2D_E_ARG
    assert lib.lp2d_phase_grad(64, 2, 0, p, q, p, 7, r, 1, None) == _lib.LP2D_E_ARG
    assert lib.lp2d_phase_grad(64, 1, 0, p, q, None, 0, p, 1, None) == _lib.LP2D_E_ARG


def test_header_compiles_as_c_and_cpp(tmp_path):
    """include/lp2d.h is a C header: a C99 translation unit (and a C++11 one) that
    includes it and names a few of its types and entry points compiles cleanly with
    warnings as errors."""
    import shutil
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "use_lp2d.c"
    src.write_text('#include "lp2d.h"\n'
                   "int main(void) {\n"
                   "  lp2d_plan* p = 0;\n"
                   "  lp2d_optics o = {1.0, 1e-4, 0.1, 0.0, 0.0, 0, 0};\n"
                   "  (void)o;\n"
                   "  return (int)lp2d_plan_workspace_size(p, LP2D_OP_ADJOINT) + lp2d_version();\n"
                   "}\n")
    for cc, args in (("gcc", ["-std=c99", "-pedantic"]), ("g++", ["-std=c++11", "-x", "c++"])):
        if shutil.which(cc) is None:
            pytest.skip(f"{cc} not available")
        r = subprocess.run([cc, *args, "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                            "-I" + os.path.join(root, "include"), str(src)],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
