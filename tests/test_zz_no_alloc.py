"""vgpu_filter_noalloc (the filter whose scratch is a fixed device array)
against vgpu_filter and the closed form of src/filter.c:13-39: selection
vectors bit-exact, on two calls in a row, the second with a smaller input,
with and without an incoming selection vector, on the fast-path comparison
and on a program run by the device interpreter."""
import ctypes as C

import numpy as np
import pytest

from helpers import dbl_with_specials, rng, with_na

pytestmark = pytest.mark.gpu


def _filter(nv, fn, prog, cols, n_rows, sel_in, n_sel, cap):
    lib = nv.vgpu()
    out = lib.vgpu_malloc(max(cap, 1) * 4)
    assert out, lib.vgpu_last_error().decode()
    try:
        n = C.c_int64(-1)
        rc = getattr(lib, fn)(C.byref(prog), nv.col_array(cols), len(cols), n_rows, sel_in, n_sel, out, C.byref(n))
        assert rc == 0, f"{fn}: {lib.vgpu_last_error().decode()}"
        sel = np.zeros(max(n.value, 1), np.int32)
        if n.value:
            assert lib.vgpu_memcpy_d2h(sel.ctypes.data, out, n.value * 4) == 0
        return sel[:n.value]
    finally:
        lib.vgpu_free(out)


def _programs(nv):
    fast = nv.cmp_const_prog(0, nv.VEC_DOUBLE, "GT", 0.0)
    # (x > -1 & k % 7 != 3) | is.na(k), through the interpreter
    vm = nv.compile_prog([
        ("COL", nv.VEC_DOUBLE, 0, 0, 0, 0),
        ("CONST", nv.VEC_DOUBLE, 0, 0, 0, -1.0),
        ("GT", nv.VEC_DOUBLE, 0, 1, 0, 0),
        ("COL", nv.VEC_INT64, 0, 0, 1, 0),
        ("CONST", nv.VEC_INT64, 0, 0, 0, 7),
        ("MOD", nv.VEC_INT64, 3, 4, 0, 0),
        ("CONST", nv.VEC_INT64, 0, 0, 0, 3),
        ("NE", nv.VEC_INT64, 5, 6, 0, 0),
        ("AND", nv.VEC_BOOL, 2, 7, 0, 0),
        ("ISNA", nv.VEC_BOOL, 3, 0, 0, 0),
        ("OR", nv.VEC_BOOL, 8, 9, 0, 0),
    ])
    return {"fast": fast, "vm": vm}


def _expected(name, x, xv, k, kv, rows):
    """The closed form: ascending rows whose predicate is valid and true
    (NaN > t is false; R's three-valued & and |)."""
    xs, xvs, ks, kvs = x[rows], xv[rows], k[rows], kv[rows]
    if name == "fast":
        keep = xvs & (xs > 0.0)
    else:
        a_val, a = xvs, xs > -1.0
        b_val, b = kvs, np.where(kvs, np.fmod(ks, 7), 0) != 3   # k >= 0: C's % equals fmod
        isna = ~kvs
        and_false = (a_val & ~a) | (b_val & ~b)
        and_true = a_val & a & b_val & b
        and_valid = and_true | and_false
        keep = (and_valid & and_true) | isna
    return rows[keep].astype(np.int32)


@pytest.mark.parametrize("pname", ["fast", "vm"])
def test_noalloc_matches_filter(gpu, pname):
    nv = gpu
    prog = _programs(nv)[pname]
    for call, (n, seed) in enumerate([(3_000_017, 11), (70_001, 12)]):
        r = rng(seed)
        x = dbl_with_specials(r, n)
        xv = with_na(rng(seed + 100), n, 0.05)
        k = r.integers(0, 1000, n)
        kv = with_na(rng(seed + 200), n, 0.05)
        cx, ck = nv.DeviceColumn(x, xv), nv.DeviceColumn(k, kv)
        sel_host = np.flatnonzero(r.random(n) < 0.6).astype(np.int32)
        dsel = nv.vgpu().vgpu_malloc(max(sel_host.nbytes, 8))
        try:
            assert nv.vgpu().vgpu_memcpy_h2d(dsel, sel_host.ctypes.data, sel_host.nbytes) == 0
            for sel_in, n_sel, rows in [(None, 0, np.arange(n)), (dsel, len(sel_host), sel_host.astype(np.int64))]:
                cap = n if sel_in is None else n_sel
                got = _filter(nv, "vgpu_filter_noalloc", prog, [cx, ck], n, sel_in, n_sel, cap)
                ref = _filter(nv, "vgpu_filter", prog, [cx, ck], n, sel_in, n_sel, cap)
                assert np.array_equal(got, ref), f"call {call}: no-alloc filter differs from vgpu_filter"
                exp = _expected(pname, x, xv, k, kv, rows)
                assert np.array_equal(got, exp), f"call {call}: selection differs from the closed form"
        finally:
            nv.vgpu().vgpu_free(dsel)
            cx.free()
            ck.free()


def test_noalloc_empty_and_limit(gpu):
    nv = gpu
    lib = nv.vgpu()
    prog = _programs(nv)["fast"]
    x = nv.DeviceColumn(np.array([1.0, -1.0, 2.0]))
    try:
        assert len(_filter(nv, "vgpu_filter_noalloc", prog, [x], 0, None, 0, 1)) == 0
        assert list(_filter(nv, "vgpu_filter_noalloc", prog, [x], 3, None, 0, 3)) == [0, 2]
        # more logical rows than the fixed scratch holds: refused on the host
        # before anything is enqueued
        n = C.c_int64(-1)
        too_many = (1 << 26) + 1
        rc = lib.vgpu_filter_noalloc(C.byref(prog), nv.col_array([x]), 1, too_many, None, 0, None, C.byref(n))
        assert rc < 0
        assert "exceed the maximum" in lib.vgpu_last_error().decode()
    finally:
        x.free()
