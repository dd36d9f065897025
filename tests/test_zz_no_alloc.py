"""bp_neighbor_grid_noalloc (the neighbour grid with its scratch in a fixed
per-GPU array, no allocation) against bp_neighbor_grid, bit for bit, through
the C ABI: two calls in a row, the second smaller, so that the first call's
scratch contents are still in place when the second runs."""
import ctypes as C

import numpy as np
import pytest

import brasspan as bp

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def case(msg):
    print(f"[no-alloc] {msg}", flush=True)


class DevBuf:
    """A device buffer owned by the test (the caller allocates; the entry point
    under test does not)."""

    def __init__(self, hip, nbytes):
        self.hip = hip
        self.nbytes = max(int(nbytes), 4)
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.nbytes)) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.ptr, C.c_void_p(a.ctypes.data), C.c_size_t(a.nbytes), H2D) == 0

    def fill(self, byte):
        assert self.hip.hipMemset(self.ptr, byte, C.c_size_t(self.nbytes)) == 0

    def download(self, count, dtype=np.int32):
        out = np.empty(count, dtype)
        if count:
            assert self.hip.hipMemcpy(C.c_void_p(out.ctypes.data), self.ptr, C.c_size_t(out.nbytes), D2H) == 0
        return out

    def free(self):
        if self.ptr.value:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def grid_noalloc(hip, xs, cell_size, grid_dims, bufs):
    """One call of bp_neighbor_grid_noalloc on the default stream, results
    copied back."""
    lib = bp.load_library()
    n = len(xs)
    cells = int(np.prod(grid_dims))
    x4 = np.zeros((n, 4), np.float32)
    x4[:, :3] = xs
    x4[:, 3] = np.nan  # w is not read
    bufs["x"].upload(x4)
    for name in ("keys", "perm", "start", "end"):
        bufs[name].fill(0x5A)  # every output element must be written
    cs = np.asarray(cell_size, np.float32)
    dims = np.asarray(grid_dims, np.int32)
    st = lib.bp_neighbor_grid_noalloc(bufs["x"].ptr, n, bp._fp(cs), bp._ip(dims), bufs["keys"].ptr,
                                      bufs["perm"].ptr, bufs["start"].ptr, bufs["end"].ptr, None)
    bp._check(st)
    assert hip.hipDeviceSynchronize() == 0
    return (bufs["keys"].download(n), bufs["perm"].download(n), bufs["start"].download(cells),
            bufs["end"].download(cells))


def test_noalloc_grid_matches_allocating_grid_twice():
    hip = bp.hip_runtime()
    nmax, cmax = 1 << 20, 1 << 20
    bufs = {"x": DevBuf(hip, 16 * nmax), "keys": DevBuf(hip, 4 * nmax), "perm": DevBuf(hip, 4 * nmax),
            "start": DevBuf(hip, 4 * cmax), "end": DevBuf(hip, 4 * cmax)}
    try:
        # 2^20 positions (the limit; 4096-key sort tiles) on 2^20 cells: three
        # 8-bit digit passes, so the sorted pairs end in the scratch array and
        # are copied out; then 100k positions on 2^14 cells: two passes, the
        # result in the caller's buffers, over scratch the first call filled.
        calls = [(1 << 20, (128, 128, 64), (0.05, 0.05, 0.1), 3), (100_000, (64, 32, 8), (0.1, 0.2, 0.4), 4)]
        for n, dims, cs, seed in calls:
            rng = np.random.default_rng(seed)
            xs = rng.uniform(-2.0, 5.0, size=(n, 3)).astype(np.float32)
            xs[: n // 10] = xs[n // 10: 2 * (n // 10)]  # duplicated cells stress stability
            case(f"n={n} dims={dims}")
            got = grid_noalloc(hip, xs, cs, dims, bufs)
            want = bp.neighbor_grid(xs, cs, dims)
            for a, b, name in zip(got, want, ("keys", "perm", "cell_start", "cell_end")):
                assert np.array_equal(a, b), name
        case("empty input and optional outputs")
        lib = bp.load_library()
        bufs["start"].fill(0x5A)
        cs = np.asarray((0.1, 0.1, 0.1), np.float32)
        dims = np.asarray((4, 8, 16), np.int32)
        bp._check(lib.bp_neighbor_grid_noalloc(None, 0, bp._fp(cs), bp._ip(dims), None, None, bufs["start"].ptr,
                                               None, None))
        assert hip.hipDeviceSynchronize() == 0
        assert np.array_equal(bufs["start"].download(512), np.full(512, -1, np.int32))
        case("beyond the scratch limit")
        with pytest.raises(bp.BrassPanError) as e:
            bp._check(lib.bp_neighbor_grid_noalloc(bufs["x"].ptr, (1 << 20) + 1, bp._fp(cs), bp._ip(dims),
                                                   bufs["keys"].ptr, bufs["perm"].ptr, None, None, None))
        assert e.value.status == 2  # BP_ERR_CAPACITY
    finally:
        for b in bufs.values():
            b.free()
