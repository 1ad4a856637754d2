"""cave_step_run_static (ticks on caller-owned device buffers, no device
allocation, block partials in a static device array) against cave_step_run on
a handle and against the CPU oracle: per-tick statistics and cells, bit-exact,
through the C-ABI.  Each kernel form runs two calls in a row, the second on a
smaller grid, so the static scratch keeps the first call's contents when the
second starts.  Device memory for the test's own buffers comes from the HIP
runtime the library itself loaded (hipMalloc through ctypes).
"""
import ctypes as C

import numpy as np
import pytest

import cave_amd as cv
import np_model as NP
import oracle as O

FE, FEV = cv.FACES | cv.EDGES, cv.FACES | cv.EDGES | cv.CORNERS


class Hip:
    """The HIP runtime functions the test needs, from the library the
    native engine already loaded."""

    def __init__(self):
        cv.lib()
        L = C.CDLL("libamdhip64.so.7")
        vp, sz = C.c_void_p, C.c_size_t
        L.hipMalloc.argtypes = [C.POINTER(vp), sz]
        L.hipFree.argtypes = [vp]
        L.hipMemcpy.argtypes = [vp, vp, sz, C.c_int]
        L.hipMemset.argtypes = [vp, C.c_int, sz]
        L.hipStreamCreate.argtypes = [C.POINTER(vp)]
        L.hipStreamSynchronize.argtypes = [vp]
        L.hipStreamDestroy.argtypes = [vp]
        for f in (L.hipMalloc, L.hipFree, L.hipMemcpy, L.hipMemset, L.hipStreamCreate, L.hipStreamSynchronize,
                  L.hipStreamDestroy, L.hipDeviceSynchronize):
            f.restype = C.c_int
        self.L = L

    def ok(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def malloc(self, n):
        p = C.c_void_p()
        self.ok(self.L.hipMalloc(C.byref(p), n))
        return p.value

    def h2d(self, dst, a: np.ndarray):
        self.ok(self.L.hipMemcpy(dst, a.ctypes.data, a.nbytes, 1))  # hipMemcpyHostToDevice

    def d2h(self, a: np.ndarray, src):
        self.ok(self.L.hipMemcpy(a.ctypes.data, src, a.nbytes, 2))  # hipMemcpyDeviceToHost


def random_grid(dims, M, seed):
    X, Y, Z = dims
    rng = np.random.default_rng(seed)
    return NP.pack(rng.integers(0, M + 1, size=(Z, Y, X)).astype(np.uint8))


def check_pair(hip, stream, kernel, handle_kernel, cases, ticks):
    for dims, M, flags, seed in cases:
        X, Y, Z = dims
        print(f"static step {X}x{Y}x{Z} M={M} flags={flags} kernel={kernel} ticks={ticks}", flush=True)
        rng = np.random.default_rng(seed)
        n = O.lib().orc_neighbor_count(flags)
        birth, surv = int(rng.integers(1, 1 << n)), int(rng.integers(1, 1 << n))
        packed = random_grid(dims, M, seed)
        # the existing entry point: a handle's cave_step_run
        with cv.Step(dims, M, flags, birth, surv, initial_cells=packed, kernel=handle_kernel) as h:
            want_rows, _ = h.run(ticks)
            want_cells = h.cells()
        # the oracle
        ref = packed
        w0, w1 = ((M - 1) & 0xF) << 60 | birth, flags << 60 | surv
        for t in range(ticks):
            ref, st = O.step(ref, X, Y, Z, w0, w1)
            assert want_rows[t].as_tuple() == st.as_tuple(), f"handle vs oracle, tick {t}"
        assert np.array_equal(want_cells, ref)
        # the static entry point on device buffers
        a, b = hip.malloc(packed.nbytes), hip.malloc(packed.nbytes)
        pt = hip.malloc(48 * ticks)
        try:
            hip.h2d(a, packed)
            hip.ok(hip.L.hipMemset(b, 0xA5, packed.nbytes))  # whatever b held must not matter
            hip.ok(hip.L.hipMemset(pt, 0xFF, 48 * ticks))
            cv.step_run_static(dims, M, flags, birth, surv, a, b, ticks, pt, stream, kernel)
            hip.ok(hip.L.hipStreamSynchronize(stream))
            got_cells = np.zeros_like(packed)
            hip.d2h(got_cells, a if ticks % 2 == 0 else b)
            rows = (cv.TickStats * ticks)()
            hip.ok(hip.L.hipMemcpy(C.addressof(rows), pt, 48 * ticks, 2))
        finally:
            for p in (a, b, pt):
                hip.L.hipFree(p)
        for t in range(ticks):
            assert rows[t].as_tuple() == want_rows[t].as_tuple(), f"static vs handle, tick {t}"
        assert np.array_equal(got_cells, want_cells), "static vs handle, cells"


# two calls in a row per kernel form, the second on a smaller grid
STREAM_CASES = [((64, 40, 30), 5, FEV, 1), ((32, 9, 7), 2, FE, 2)]
GATHER_CASES = [((70, 33, 21), 4, FEV, 3), ((37, 6, 5), 3, FE, 4)]


@pytest.fixture(scope="module")
def hip_stream():
    hip = Hip()
    s = C.c_void_p()
    hip.ok(hip.L.hipStreamCreate(C.byref(s)))
    yield hip, s.value
    hip.L.hipStreamDestroy(s.value)


@pytest.mark.gpu
@pytest.mark.parametrize("ticks", [1, 6, 7])
def test_static_stream_nibble_matches_handle(hip_stream, ticks):
    hip, s = hip_stream
    check_pair(hip, s, cv.KERNEL_STREAM_NIBBLE, cv.KERNEL_STREAM_NIBBLE, STREAM_CASES, ticks)


@pytest.mark.gpu
@pytest.mark.parametrize("ticks", [1, 6, 7])
def test_static_gather_matches_handle(hip_stream, ticks):
    hip, s = hip_stream
    check_pair(hip, s, cv.KERNEL_GATHER, cv.KERNEL_GATHER, GATHER_CASES, ticks)


@pytest.mark.gpu
def test_static_auto_matches_handle(hip_stream):
    """AUTO takes the streaming kernel where the reference layout fits it
    (ceil(X/8) a multiple of 4) and the gather kernel elsewhere."""
    hip, s = hip_stream
    check_pair(hip, s, cv.KERNEL_AUTO, cv.KERNEL_STREAM_NIBBLE, STREAM_CASES, 5)
    check_pair(hip, s, cv.KERNEL_AUTO, cv.KERNEL_GATHER, GATHER_CASES, 5)


@pytest.mark.gpu
def test_static_null_stream_and_no_stats(hip_stream):
    """The null stream, and no per-tick statistics."""
    hip, _ = hip_stream
    dims, M, flags = (48, 20, 11), 3, FEV
    X, Y, Z = dims
    packed = random_grid(dims, M, 9)
    birth, surv = 0b11100, 0b111000
    a, b = hip.malloc(packed.nbytes), hip.malloc(packed.nbytes)
    try:
        hip.h2d(a, packed)
        cv.step_run_static(dims, M, flags, birth, surv, a, b, 3, 0, 0, cv.KERNEL_AUTO)
        hip.ok(hip.L.hipDeviceSynchronize())
        got = np.zeros_like(packed)
        hip.d2h(got, b)
    finally:
        hip.L.hipFree(a)
        hip.L.hipFree(b)
    ref = packed
    for _ in range(3):
        ref, _ = O.step(ref, X, Y, Z, ((M - 1) << 60) | birth, (flags << 60) | surv)
    assert np.array_equal(got, ref)


# ---- argument checks: host-side, before any HIP call (run on the CPU) ----
def test_static_rejects_stream_nibble_where_the_layout_does_not_fit():
    with pytest.raises(cv.CaveError) as e:
        cv.step_run_static((49, 8, 8), 1, FE, 1, 1, 1 << 40, 2 << 40, 1, kernel=cv.KERNEL_STREAM_NIBBLE)
    assert e.value.code == cv.CAVE_E_UNSUPPORTED


@pytest.mark.parametrize("dims", [(0, 4, 4), (4, 65537, 4), (65536, 65536, 2)])
def test_static_rejects_bad_dims(dims):
    with pytest.raises(cv.CaveError) as e:
        cv.step_run_static(dims, 1, FE, 1, 1, 1 << 40, 2 << 40, 1)
    assert e.value.code == cv.CAVE_E_INVALID


def test_static_rejects_null_buffers_and_erd():
    with pytest.raises(cv.CaveError):
        cv.step_run_static((8, 8, 8), 1, FE, 1, 1, 0, 0, 1)
    p = cv.encode_params(1, FE, 1, 1, shape=cv.SHAPE_ERD)
    d = (C.c_uint32 * 3)(8, 8, 8)
    assert cv.lib().cave_step_run_static(C.byref(p), d, 0, 1 << 40, 2 << 40, 1, None, None) == cv.CAVE_E_UNSUPPORTED
