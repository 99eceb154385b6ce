"""felis_run_noalloc (no device allocation) against felis_run through the
C-ABI, two calls in a row on one device, the second with a smaller input, so
the second reads a scratch array the first left behind.

Batched FP32/FP64: the deposit's float atomics sum in a run-dependent order,
so the per-period field energies agree to 1e-6 relative, bunching and twiss
to 1e-6 of their scale (the two runs of felis_run itself differ by as
much). Plain FP64: 1e-12. A run larger than the array is refused with
FELIS_ERR_ARG and leaves the context usable.
"""
import ctypes as C

import numpy as np
import pytest

import fel_cases as fc

pytestmark = pytest.mark.gpu

fl = pytest.importorskip("felis_amd._lib")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _call(fn, ctx, tab, n):
    d = ctx.desc
    e = np.zeros((n, 2 * d.n_harm))
    b = np.zeros((n, 2 * d.n_harm + 1))
    t = np.zeros((n, 7 * d.n_species))
    arr = (fl.Lattice * len(tab))(*tab)
    rc = fn(ctx.h, arr, 0, n, 0, None, _dp(e), _dp(b), _dp(t))
    return rc, e, b, t


def _pair(case, prec, n):
    """The same state run by felis_run and felis_run_noalloc."""
    L = fl.load()
    tab = [fc.to_lib_lattice(case.lattice(current_flag=0))] + [fc.to_lib_lattice(case.lattice())] * (2 * n)
    ref_ctx = fc.gpu_context(case, prec)
    rc, e0, b0, t0 = _call(L.felis_run, ref_ctx, tab, n)
    assert rc == 0
    u0 = ref_ctx.download_field(fl.UX)
    ref_ctx.close()
    ctx = fc.gpu_context(case, prec)
    rc, e1, b1, t1 = _call(L.felis_run_noalloc, ctx, tab, n)
    assert rc == 0, L.felis_last_error(ctx.h)
    u1 = ctx.download_field(fl.UX)
    ctx.close()
    return (e0, b0, t0, u0), (e1, b1, t1, u1)


def _check(ref, got, tol):
    for a, b in zip(ref, got):
        scale = np.max(np.abs(a))
        assert scale > 0
        assert np.max(np.abs(a - b)) <= tol * scale


@pytest.mark.parametrize("prec,tol", [pytest.param(1, 1e-6, id="mixed"), pytest.param(0, 1e-12, id="fp64")])
def test_noalloc_matches_run_twice_second_smaller(prec, tol):
    big = fc.make_case(slice_n=12, harmonics=(1, 3, 5), ppslice=512, nx=33, ny=31, uneven=True)
    small = fc.make_case(slice_n=4, harmonics=(1,), ppslice=128, nx=21, ny=21, seed=4)
    for case, n in ((big, 14), (small, 6)):
        ref, got = _pair(case, prec, n)
        _check(ref, got, tol)


def test_noalloc_refuses_a_run_past_its_array():
    L = fl.load()
    case = fc.make_case(slice_n=4, harmonics=(1,), ppslice=128, nx=21, ny=21)
    ctx = fc.gpu_context(case, 1)
    n = 200000   # 200k periods of output rows exceed the 16 MiB array
    tab = [fc.to_lib_lattice(case.lattice())] * (2 * n + 1)
    arr = (fl.Lattice * len(tab))(*tab)
    rc = L.felis_run_noalloc(ctx.h, arr, 0, n, 0, None, None, None, None)
    assert rc == fl.ERR_ARG
    assert b"scratch" in L.felis_last_error(ctx.h)
    rc, e, _, _ = _call(L.felis_run_noalloc, ctx, tab[:9], 4)
    assert rc == 0 and np.all(e[:, 0] > 0)
    ctx.close()
