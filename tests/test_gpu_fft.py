"""GPU 3D R2C/C2R (OP_FFT / OP_iFFT) against the oracle FFT."""
import numpy as np
import pytest

from helpers import rel_l2

pytestmark = pytest.mark.gpu

SHAPES = [(4, 6, 8), (8, 8, 8), (12, 6, 12), (16, 24, 32), (32, 32, 32), (48, 48, 48),
          (24, 96, 48), (64, 64, 64), (128, 32, 16)]


def _grid(ctx, shape):
    import kielflow_amd as kf
    return kf.Grid(ctx, [kf.axis(kf.EQST, n, -8.0, 16.0) for n in shape])


@pytest.mark.parametrize("shape", SHAPES)
def test_r2c_c2r_fp64_vs_oracle(ctx, oracle, shape):
    import kielflow_amd as kf
    print("fft fp64", shape, flush=True)
    g = _grid(ctx, shape)
    og = oracle.Grid(*[oracle.axis(oracle.EQST, n, -8.0, 16.0) for n in shape])
    f = kf.FFT(g, kf.KF_FP64)
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape)
    dx = ctx.to_device(x)
    dX = g.cfield()
    f.r2c(dx, dX)
    X = dX.get()
    assert rel_l2(X, og.r2c(x)) < 2e-15
    # C2R of a non-Hermitian spectrum: the kz = 0 / Nz/2 projection rule
    Y = rng.standard_normal(og.cshape) + 1j * rng.standard_normal(og.cshape)
    dY = ctx.to_device(Y)
    dy = g.rfield()
    f.c2r(dY, dy)
    assert rel_l2(dy.get(), og.c2r(Y)) < 2e-15
    assert np.array_equal(dY.get(), Y)  # input preserved
    f.close()
    g.close()


@pytest.mark.parametrize("shape", [(8, 8, 8), (32, 32, 32), (48, 48, 48), (64, 64, 64)])
def test_r2c_c2r_fp32_vs_oracle(ctx, oracle, shape):
    import kielflow_amd as kf
    print("fft fp32", shape, flush=True)
    g = _grid(ctx, shape)
    og = oracle.Grid(*[oracle.axis(oracle.EQST, n, -8.0, 16.0) for n in shape])
    f = kf.FFT(g, kf.KF_FP32)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(shape)
    dx = ctx.to_device(x.astype(np.float32))
    dX = g.cfield(np.complex64)
    f.r2c(dx, dX)
    assert rel_l2(dX.get(), og.r2c(x.astype(np.float32).astype(np.float64))) < 2e-6
    dy = g.rfield(np.float32)
    f.c2r(dX, dy)
    assert rel_l2(dy.get(), x) < 2e-6
    f.close()
    g.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_roundtrip_and_parseval_256(ctx, prec):
    """Full-size properties: c2r(r2c(x)) = x and Parseval's identity."""
    import kielflow_amd as kf
    print("fft 256 roundtrip", prec, flush=True)
    shape = (256, 256, 256)
    g = _grid(ctx, shape)
    p = kf.KF_FP64 if prec == "fp64" else kf.KF_FP32
    rt = np.float64 if prec == "fp64" else np.float32
    ct = np.complex128 if prec == "fp64" else np.complex64
    f = kf.FFT(g, p)
    rng = np.random.default_rng(11)
    x = rng.standard_normal(shape).astype(rt)
    dx = ctx.to_device(x)
    dX = g.cfield(ct)
    f.r2c(dx, dX)
    X = dX.get().astype(np.complex128)
    w = np.full(X.shape[2], 2.0)
    w[0] = 1.0
    w[-1] = 1.0
    energy = (np.abs(X) ** 2 * w).sum() / x.size
    tol = 1e-12 if prec == "fp64" else 1e-5
    assert abs(energy - (x.astype(np.float64) ** 2).sum()) / energy < tol
    dy = g.rfield(rt)
    f.c2r(dX, dy)
    assert rel_l2(dy.get(), x) < (1e-14 if prec == "fp64" else 2e-6)
    f.close()
    g.close()
