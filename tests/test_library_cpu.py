"""CPU-side checks of the HIP library: it loads, exports every declared symbol,
its host code matches the oracle, and its errors are status codes (no GPU calls)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import fe as ofe, objective as oobj

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "maple_hip.h")


@pytest.fixture(scope="module")
def lib():
    from maple_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built (run python -m maple_amd.build)")
    return _lib.load()


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(maple_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol(lib):
    from maple_amd import _lib
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), f"{s} declared in maple_hip.h but not exported"
    assert set(syms) == set(_lib.SIGNATURES), "ctypes signatures and the header disagree"


def test_host_element_matrices_match_oracle(lib):
    from maple_amd.cells import hex8_matrices
    for r in (1, 8, 64):
        Ke24, Fe24, X024, Ke8, Fe8, X08 = hex8_matrices(r)
        Ko, Fo, Xo = ofe.hexahedron(r, ofe.isotropic_elastic_tensor(1.0, 0.3))
        assert np.abs(Ke24 - Ko).max() <= 1e-14 * np.abs(Ko).max()
        assert np.abs(Fe24 - Fo).max() <= 1e-14 * np.abs(Fo).max()
        assert np.abs(X024 - Xo).max() <= 1e-12 * np.abs(Xo).max()
        Ks, Fs, Xs = ofe.hexahedron_scalar(r)
        assert np.abs(Ke8 - Ks).max() <= 1e-14 * np.abs(Ks).max()
        assert np.abs(Fe8 - Fs).max() <= 1e-14 * np.abs(Fs).max()
        assert np.abs(X08 - Xs).max() <= 1e-12 * np.abs(Xs).max()


def test_errors_without_gpu(lib):
    from maple_amd import _lib
    if lib.maple_device_count() > 0:
        pytest.skip("a GPU is visible")
    h = C.c_void_p()
    dims = (C.c_int * 3)(8, 8, 8)
    rc = lib.maple_ctx_create(0, 1, dims, 3, 0, C.byref(h))
    assert rc == _lib.MAPLE_EDEVICE
    assert b"device" in lib.maple_last_error()
    assert not h.value


def test_argument_validation(lib):
    from maple_amd import _lib
    h = C.c_void_p()
    bad = (C.c_int * 3)(12, 12, 12)   # 12 is not divisible by 2^3
    assert lib.maple_ctx_create(0, 1, bad, 4, 0, C.byref(h)) == _lib.MAPLE_EINVAL
    assert b"multiple" in lib.maple_last_error()
    assert lib.maple_ctx_create(0, 0, bad, 2, 0, C.byref(h)) == _lib.MAPLE_EINVAL
    assert lib.maple_hex8_matrices(0, 1.0, 0.3, 1.0, None, None, None, None, None, None) == _lib.MAPLE_EINVAL
    assert lib.maple_ctx_destroy(None) == 0


def test_product_objective_matches_oracle():
    import torch
    from maple_amd import objective as po
    rng = np.random.default_rng(0)
    A = rng.standard_normal((6, 6))
    CH = ofe.isotropic_elastic_tensor(1.0, 0.3) * 0.3 + 0.01 * (A @ A.T)
    kap = np.diag([0.2, 0.3, 0.4]) + 0.01
    w = {"stiffness": 0.4, "shear": 0.3, "thermal": 0.2, "bulk": 0.1}
    J, G6, G3 = po.evaluate_np(w, CH, kap)
    Jo, G6o, G3o = oobj.evaluate(w, CH, kap)
    assert abs(J - Jo) < 1e-14 and np.abs(G6 - G6o).max() < 1e-12 and np.abs(G3 - G3o).max() < 1e-14
    # torch form with autograd gives the same gradient
    t = torch.tensor(CH, requires_grad=True)
    k = torch.tensor(kap, requires_grad=True)
    Jt = po.objective_from_spec(obj="stiffness:0.4,shear:0.3,thermal:0.2,bulk:0.1", CH=t, kappa=k)
    Jt.backward()
    assert abs(Jt.item() - Jo) < 1e-12
    assert np.abs(t.grad.numpy() - G6o).max() < 1e-10
    assert np.abs(k.grad.numpy() - G3o).max() < 1e-12


def test_objective_grammar():
    from maple_amd.objective import parse_objective_spec, objective_requires
    assert parse_objective_spec("stiffness") == {"stiffness": 1.0}
    assert parse_objective_spec("0.4*stiffness+0.6*thermal") == {"stiffness": 0.4, "thermal": 0.6}
    assert parse_objective_spec("shear=0.3+thermal=0.7") == {"shear": 0.3, "thermal": 0.7}
    assert parse_objective_spec("E:0.5, kappa:0.5") == {"stiffness": 0.5, "thermal": 0.5}
    assert parse_objective_spec(None) == {"shear": 1.0}
    assert objective_requires("bulk") == (True, False)
    # the reference's other forms (objective.py:93-157): name*w, repeated terms add up,
    # a lone name with surrounding spaces, scientific weights
    assert parse_objective_spec("stiffness*2,g:1") == {"stiffness": 2.0, "shear": 1.0}
    assert parse_objective_spec("e+e") == {"stiffness": 2.0}
    assert parse_objective_spec("  Thermal ") == {"thermal": 1.0}
    assert parse_objective_spec("1e-3*shear") == {"shear": 1e-3}
    # rejected as the reference rejects them: unknown names, two numbers or two names
    # around '*', a bare weight, an exponent with '+' (the separator splits it)
    for bad in ("strength", "2*3", "e*g", "0.5", "stiffness:abc", "stiffness:1e+3", ","):
        with pytest.raises(ValueError):
            parse_objective_spec(bad)


def test_workload_designs_match_oracle():
    from maple_amd import optim
    from oracle import driver
    assert np.array_equal(optim.x0_sphere_void(16), driver.x0_sphere_void(16))
    assert np.array_equal(optim.x0_uniform_noise(16), driver.x0_uniform_noise(16))


def test_warmstart_network_layout_matches_oracle(lib):
    """Parameter count and initialisation of the warm-start network agree between the
    library, its Python wrapper and the FP64 restatement; the layout helpers invert."""
    import torch
    from maple_amd import learned
    from oracle import unet as ounet
    assert lib.maple_unet_num_params() == learned.num_params() == ounet.num_params()
    assert [l[0] for l in learned.LAYERS] == [l[0] for l in ounet.LAYERS]
    np.testing.assert_array_equal(learned.init_params(3), ounet.init_params(3).astype(np.float32))
    ue = torch.randn(6, 3, 64, dtype=torch.float64)
    ut = torch.randn(3, 1, 64, dtype=torch.float64)
    ch = learned.fields_to_channels(ue, ut)
    assert ch.shape == (64, 21)
    e, t = learned.channels_to_fields(ch)
    np.testing.assert_allclose(e.numpy(), (ue - ue.mean(dim=2, keepdim=True)).numpy(), atol=1e-12)
    np.testing.assert_allclose(t.numpy(), (ut - ut.mean(dim=2, keepdim=True)).numpy(), atol=1e-12)
    # the feature of a cell twice the grid is the 2x2x2 mean of rho^(1/3)
    rho = torch.rand(8 ** 3, dtype=torch.float64)
    f = learned.density_feature(rho, 8, 4).numpy()
    ref = (rho.numpy() ** (1 / 3)).reshape(4, 2, 4, 2, 4, 2).mean(axis=(1, 3, 5)).ravel()
    np.testing.assert_allclose(f, ref, rtol=1e-6)


def test_tuner_choice_rule():
    """The tuner keeps the fastest variant; under concurrent timing (which scatters by a
    few per cent) the library's default unless a candidate beats it by the margin."""
    from maple_amd import tune
    m = tune.CONCURRENT_MARGIN
    t = {5: 1.00, 8: 1.02, 9: 0.99}
    assert tune.choose(t, 8, concurrent=False) == 9
    assert tune.choose(t, 8, concurrent=True) == 8                 # 9 is within the margin of 8
    t = {5: 1.00, 8: 1.02, 9: 1.02 * (1 - m) - 1e-3}
    assert tune.choose(t, 8, concurrent=True) == 9                 # clearly faster: taken
    assert tune.choose({0: 2.0, 1: 1.0}, None, concurrent=True) == 1   # no default: the fastest
    assert tune.choose({0: 2.0, 1: 1.0}, 7, concurrent=True) == 1      # default not timed: the fastest
    assert set(tune.DEFAULTS) == set(tune.KERNELS)
    for name, v in tune.DEFAULTS.items():
        assert v in tune.KERNELS[name][1]
