"""The C-ABI library loads and exports every symbol include/kielflow_amd.h
declares; host-side entry points (no GPU work) agree with the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kielflow_amd.h")


def header_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import kielflow_amd as kf
    lib = kf.load()
    syms = header_symbols()
    assert len(syms) > 30
    for s in syms:
        assert hasattr(lib, s), s
    # every declared symbol has a ctypes signature in the binding
    from kielflow_amd.api import SIGNATURES
    missing = [s for s in syms if s not in SIGNATURES]
    assert not missing, missing


def test_axis_values_match_oracle(oracle):
    import kielflow_amd as kf
    from kielflow_amd.api import Axis
    lib = kf.load()
    cases = [(kf.EQST, 32, 0.0), (kf.SIST, 32, 16 * (1 - 1 / 3) / (2 * np.pi)), (kf.HYST, 24, 3.0)]
    for t, n, m in cases:
        a = Axis(t, n, -8.0, 16.0, m)
        oa = oracle.axis(t, n, -8.0, 16.0, m)
        for i in (-2, -1, 0, 1, 7, n // 2, n - 1, n, n + 1):
            v, k, d = C.c_double(), C.c_double(), C.c_double()
            assert lib.kf_axis_values(C.byref(a), i, C.byref(v), C.byref(k), C.byref(d)) == 0
            assert v.value == oracle.val_at(oa, i)
            assert k.value == oracle.k_val_at(oa, i)
            if 0 <= i < n:
                assert d.value == oracle.dS(oa, i)


def test_setup_host_matches_oracle(oracle):
    import kielflow_amd as kf
    axes = kf.cube(16)
    ln_n, nd, mask = kf.setup_host(axes, seed=7, amp=1e-3, Q=-0.1, a=2.0)
    g = oracle.Grid.cube(16)
    z = oracle.normals(7, 16 ** 3).reshape(g.rshape)
    assert np.allclose(ln_n, np.log(1 + 1e-3 * z), rtol=1e-15, atol=0)  # numpy log vs libm: 1 ulp
    assert np.array_equal(nd, g.gauss(-0.1, 2.0))
    assert np.array_equal(mask, g.mask(2.0))


def test_errors_are_reported_not_thrown():
    import kielflow_amd as kf
    lib = kf.load()
    out = C.c_void_p()
    code = lib.kf_ctx_create(-1, C.byref(out))
    assert code != 0
    assert len(lib.kf_last_error()) > 0
    assert lib.kf_rhs_solve(None, 0.0, None, None, None, None) == kf.api.KF_ERR_ARG
    assert lib.kf_poisson_set_tolerance(None, 1e-10) == kf.api.KF_ERR_ARG


def test_product_path_does_not_import_the_oracle():
    for root, _, files in os.walk(os.path.join(ROOT, "kielflow_amd")):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                text = open(os.path.join(root, f)).read()
                assert "oracle" not in text.replace("Oracle", ""), f


class _NoCallLib:
    """Stands in for the library: any call is a failure of the argument checks."""

    def __getattr__(self, name):
        def f(*a):
            raise AssertionError(f"{name} was called with unchecked arguments")
        return f


class _Fake:
    def __init__(self, shape, dtype):
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(shape)) * self.dtype.itemsize
        self.ptr = C.c_void_p(0x1000)


def _objects(n=8):
    from kielflow_amd import api
    ctx = type("Ctx", (), {})()
    ctx.lib = _NoCallLib()
    g = api.Grid.__new__(api.Grid)
    g.ctx, g.h = ctx, None
    g.n = (n, n, n)
    g.rshape, g.cshape = g.n, (n, n, n // 2 + 1)
    ps = api.Poisson.__new__(api.Poisson)
    ps.grid, ps.h = g, None
    r = api.Rhs.__new__(api.Rhs)
    r.grid, r.h = g, None
    rk = api.Rk4.__new__(api.Rk4)
    rk.rhs, rk.h = r, None
    fft = api.FFT.__new__(api.FFT)
    fft.grid, fft.h, fft.precision = g, None, api.KF_FP64
    return g, ps, r, rk, fft


def test_wrong_fields_raise_before_any_library_call():
    """A field of another grid, another dtype or a freed one is rejected on
    the host; the library, which trusts raw pointers, is never called."""
    g, ps, r, rk, fft = _objects(8)
    other = _Fake((16, 16, 16), np.float64)
    f32 = _Fake(g.rshape, np.float32)
    good_c = [_Fake(g.cshape, np.complex128) for _ in range(4)]
    with pytest.raises(ValueError):
        r.set_fields(other, None, None)
    with pytest.raises(ValueError):
        r.set_fields(f32, None, None)
    with pytest.raises(ValueError):
        ps.solve(_Fake(g.rshape, np.float64), other)
    with pytest.raises(ValueError):
        r.solve(0.0, good_c[:3] + [_Fake(g.cshape, np.complex64)])
    with pytest.raises(ValueError):
        rk.step(good_c[:3])
    with pytest.raises(ValueError):
        rk.step(good_c[:3] + [_Fake((8, 8, 8), np.complex128)])
    with pytest.raises(ValueError):
        fft.r2c(_Fake(g.rshape, np.float32), good_c[0])
    with pytest.raises(ValueError):
        g.deriv(0, good_c[0], _Fake(g.rshape, np.float64))
    freed = _Fake(g.rshape, np.float64)
    freed.ptr = None
    with pytest.raises(ValueError):
        r.set_fields(None, freed, None)
    with pytest.raises(TypeError):
        r.set_fields(np.zeros(g.rshape), None, None)


def test_rhs_keeps_the_fields_whose_pointers_the_library_holds():
    from kielflow_amd import api
    g, ps, r, rk, fft = _objects(8)
    calls = []
    g.ctx.lib = type("L", (), {"kf_rhs_set_fields": lambda self, *a: calls.append(a) or 0})()
    phi, nd, mask = (_Fake(g.rshape, np.float64) for _ in range(3))
    r.set_fields(phi, nd, mask)
    assert len(calls) == 1
    assert r._phi is phi
    r.set_fields(None, None, None)   # a null keeps what the library keeps
    assert r._phi is phi


def test_adaptor_defaults_follow_the_shipped_switches(tmp_path):
    """rhs_standard.hpp:22-30 ships _RHS_SVISCOSITY on and _RHS_PENALIZATION
    commented out; the adaptor's defaults are the same, and the adaptor
    header compiles on its own."""
    import subprocess
    src = tmp_path / "defaults.cpp"
    src.write_text('#include "kielflow_compat.hpp"\n'
                   "static_assert(rhs_standard::default_spectral_viscosity, \"SV on\");\n"
                   "static_assert(!rhs_standard::default_penalization, \"penalization off\");\n"
                   "int main() { return 0; }\n")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-I",
                        os.path.join(ROOT, "kielflow_amd", "host"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
