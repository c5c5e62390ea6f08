"""The kernel-variant tuner (maple_amd.tune.autotune).

It must set, for every tunable kernel, one of that kernel's variants and the one its
own timings rank fastest; the solve with the tuned variants must give the results of
the default ones (every variant computes the same operator: parity tests pin each one
against the oracle); a second call on a batch of the same shape reuses the choice; the
concurrent form (streams > 1) keeps the library's default unless a candidate beats it
by the margin.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0) if torch.cuda.is_available() else None


def _rho(n, seed=0):
    rng = np.random.default_rng(seed)
    x = np.clip(0.35 + 0.3 * rng.standard_normal(n), 1e-3, 1.0)
    return torch.from_numpy(1e-9 + (1 - 1e-9) * x ** 3).to(DEV)


@pytest.mark.parametrize("precision,dims", [("mixed", [16, 32]), ("fp64", [32])])
def test_autotune_choice_and_results(precision, dims):
    import maple_amd
    from maple_amd import tune as mt
    print(f"[tune] {precision} {dims}", flush=True)
    cells = maple_amd.CellBatch(dims, nlevels=3, device=DEV, precision=precision)
    rho = _rho(cells.nelem)
    CH0, k0, _ = cells.homogenize(rho, rtol=1e-9, maxit=2000)
    CH0, k0 = CH0.clone(), k0.clone()
    res = mt.autotune(cells, reps=2, force=True)
    assert set(res) == set(mt.KERNELS)
    for name, (v, times) in res.items():
        assert v in mt.KERNELS[name][1]
        assert set(times) == set(mt.KERNELS[name][1])
        assert times[v] == min(times.values())
    CH1, k1, _ = cells.homogenize(rho, rtol=1e-9, maxit=2000)
    assert float(torch.linalg.norm(CH1 - CH0) / torch.linalg.norm(CH0)) <= 1e-7
    assert float(torch.linalg.norm(k1 - k0) / torch.linalg.norm(k0)) <= 1e-7
    again = mt.autotune(cells)
    assert {k: v for k, (v, _) in again.items()} == {k: v for k, (v, _) in res.items()}
    assert all(t is None for _, t in again.values())
    conc = mt.autotune(cells, reps=2, force=True, streams=3)
    for name, (v, times) in conc.items():
        fastest = min(times.values())
        assert v in mt.KERNELS[name][1]
        # concurrent timings: the default stays unless a candidate beats it by the margin
        d = mt.DEFAULTS[name]
        if times[d] <= fastest / (1.0 - mt.CONCURRENT_MARGIN):
            assert v == d, (name, v, times)
        else:
            assert times[v] == fastest, (name, v, times)
    CH2, _, _ = cells.homogenize(rho, rtol=1e-9, maxit=2000)
    assert float(torch.linalg.norm(CH2 - CH0) / torch.linalg.norm(CH0)) <= 1e-7
    cells.close()
