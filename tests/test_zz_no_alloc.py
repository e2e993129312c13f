"""bd_inverse_cholesky_static_dev, the allocation-free border factor (its
scratch is a static device array), against bd_inverse_cholesky_dev on the
same inputs: the same T bits, verdicts and pivot ratios. Two calls in a row,
the second at a smaller capacity (the shared array keeps the first call's
contents), a rejected border, and a capacity above the static limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATIC_CAP = 640


def _factor(name, S, T, cap, frm, to):
    import torch
    from baandi_amd import lib as L
    Sd = torch.from_numpy(np.asfortranarray(S).ravel(order="F").copy()).cuda()
    Td = torch.from_numpy(np.asfortranarray(T).ravel(order="F").copy()).cuda()
    acc = C.c_int32(-1)
    ratio = C.c_double(0.0)
    L.check(getattr(L.load(), name)(L.ptr(Sd), L.ptr(Td), cap, frm, to, C.byref(acc), C.byref(ratio)))
    return bool(acc.value), ratio.value, Td.cpu().numpy().reshape(cap, cap).T


def _overlap(rng, cap, rows, dependent=None):
    V = rng.standard_normal((rows, cap)) * rng.uniform(0.5, 2.0, cap)[None, :]
    if dependent is not None:
        V[:, dependent] = V[:, 7] + 1e-9 * V[:, dependent + 1]
    return V.T @ V


@pytest.mark.parametrize("steps", [
    # (cap, [(from, to), ...]): a wide border in blocks and extensions at
    # cap 600, then a smaller capacity reusing the same static array.
    (600, [(0, 288), (288, 500), (500, 600)]),
    (96, [(0, 40), (40, 96)]),
])
def test_static_factor_equals_allocating_factor(steps):
    rng = np.random.default_rng(steps[0])
    cap, borders = steps
    S = _overlap(rng, cap, 3 * cap)
    Ta = np.zeros((cap, cap))
    Tb = np.zeros((cap, cap))
    for frm, to in borders:
        oka, ra, Ta = _factor("bd_inverse_cholesky_dev", S, Ta, cap, frm, to)
        okb, rb, Tb = _factor("bd_inverse_cholesky_static_dev", S, Tb, cap, frm, to)
        ident = np.abs(Tb[:to, :to] @ S[:to, :to] @ Tb[:to, :to].T - np.eye(to)).max()
        print(f"[no-alloc factor] cap={cap} {frm}->{to}: accepted {oka}/{okb}, ratio {ra:.3e}/{rb:.3e}, "
              f"|T S T^T - I| {ident:.2e}, bitwise {np.array_equal(Ta, Tb)}", flush=True)
        assert oka and okb
        assert ra == rb
        assert np.array_equal(Ta, Tb)
        assert ident < 1e-11


def test_static_factor_two_calls_second_smaller():
    """Two calls in a row on the shared array, the second with a smaller
    input; each equals the allocating entry point bit for bit."""
    rng = np.random.default_rng(11)
    for cap in (STATIC_CAP, 200):
        S = _overlap(rng, cap, 2 * cap)
        oka, ra, Ta = _factor("bd_inverse_cholesky_dev", S, np.zeros((cap, cap)), cap, 0, cap)
        okb, rb, Tb = _factor("bd_inverse_cholesky_static_dev", S, np.zeros((cap, cap)), cap, 0, cap)
        print(f"[no-alloc factor] cap={cap} 0->{cap}: accepted {oka}/{okb}, bitwise {np.array_equal(Ta, Tb)}",
              flush=True)
        assert oka and okb and ra == rb and np.array_equal(Ta, Tb)


def test_static_factor_rejects_dependent_border():
    rng = np.random.default_rng(12)
    cap = 320
    S = _overlap(rng, cap, 1000, dependent=250)
    oka, ra, _ = _factor("bd_inverse_cholesky_dev", S, np.zeros((cap, cap)), cap, 0, 300)
    okb, rb, _ = _factor("bd_inverse_cholesky_static_dev", S, np.zeros((cap, cap)), cap, 0, 300)
    print(f"[no-alloc factor] dependent border: accepted {oka}/{okb}, ratio {ra:.3e}/{rb:.3e}", flush=True)
    assert not oka and not okb
    assert ra == rb and ra < np.finfo(float).eps / 1e-12


def test_static_factor_refuses_capacity_above_limit():
    from baandi_amd import lib as L
    cap = STATIC_CAP + 1
    S = np.eye(cap)
    with pytest.raises(L.BaandiError) as e:
        _factor("bd_inverse_cholesky_static_dev", S, np.zeros((cap, cap)), cap, 0, 4)
    assert e.value.status == L.BD_ERR_INVALID_PARAMETER
    assert "640" in str(e.value)
