"""The allocation-free sort entry point (bh_hip_sort_pairs_noalloc) against
bh_hip_sort_pairs, through the C ABI: bit-exact keys and values, two calls in
a row on one GPU (the second smaller, so it runs over the first call's
scratch contents), even and odd pass counts, in place and out of place, and
the size limit rejected before any device work.

Device buffers and the stream come from the HIP runtime the library itself
is bound to (libamdhip64 by its soname, through ctypes), so the test works
whatever else the process has initialised."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_KEYS = 1 << 21
H2D, D2H = 1, 2


def say(*a):
    print(*a, flush=True)


class Hip:
    """Minimal ctypes view of the HIP runtime libbh_hip.so uses."""

    def __init__(self):
        import blitzar_amd as B
        B.lib()  # loads libamdhip64 as the library's dependency
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.rt.hipStreamDestroy.argtypes = [C.c_void_p]
        self.held = []

    def _ok(self, rc, what):
        assert rc == 0, f"{what} failed: hip error {rc}"

    def alloc(self, nbytes):
        p = C.c_void_p()
        self._ok(self.rt.hipMalloc(C.byref(p), max(nbytes, 1)), "hipMalloc")
        self.held.append(p.value)
        return p.value

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        self._ok(self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        return p

    def get(self, p, n, dtype):
        out = np.empty(n, dtype)
        self._ok(self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def stream(self):
        s = C.c_void_p()
        self._ok(self.rt.hipStreamCreate(C.byref(s)), "hipStreamCreate")
        return s.value

    def close(self, streams=()):
        for s in streams:
            self.rt.hipStreamDestroy(s)
        for p in self.held:
            self.rt.hipFree(p)
        self.held = []


def _sort_noalloc(lib, dk, dv, n, bits, dko, dvo, stream):
    err = C.create_string_buffer(512)
    rc = lib.bh_hip_sort_pairs_noalloc(dk, dv, n, bits, dko, dvo, stream, err, 512)
    assert rc == 0, err.value.decode(errors="replace")


def _case(seed, n, bits, dup):
    rng = np.random.default_rng(seed)
    hi = (1 << bits) - 1 if bits < 64 else (1 << 64) - 1
    if dup:  # many equal keys: stability is visible in the values
        keys = rng.integers(0, 97, n, dtype=np.uint64) * np.uint64(hi // 97)
    else:
        keys = rng.integers(0, hi, n, dtype=np.uint64, endpoint=True)
    # bits above key_bits must not influence the order
    if bits < 64:
        keys = (keys & np.uint64(hi)) | (
            rng.integers(0, 1 << (64 - bits), n, dtype=np.uint64) << np.uint64(bits))
    vals = rng.permutation(n).astype(np.uint32)
    return keys, vals


@pytest.mark.parametrize("bits,in_place", [(63, False), (21, False), (30, True), (21, True)])
def test_noalloc_matches_sort_pairs_two_calls(bits, in_place):
    import blitzar_amd as B
    lib = B.lib()
    hip = Hip()
    st = hip.stream()
    try:
        for call, (n, dup) in enumerate(((1_500_000, False), (70_001, True))):
            say(f"case noalloc sort bits={bits} in_place={in_place} call={call} n={n}")
            keys, vals = _case(100 * bits + call, n, bits, dup)
            ref_k, ref_v = B.sort_pairs(keys, vals, bits)
            dk, dv = hip.put(keys), hip.put(vals)
            if in_place:
                dko, dvo = dk, dv
            else:
                dko, dvo = hip.alloc(keys.nbytes), hip.alloc(vals.nbytes)
            _sort_noalloc(lib, dk, dv, n, bits, dko, dvo, st)
            assert np.array_equal(hip.get(dko, n, np.uint64), ref_k)
            assert np.array_equal(hip.get(dvo, n, np.uint32), ref_v)
            if not in_place:  # inputs untouched
                assert np.array_equal(hip.get(dk, n, np.uint64), keys)
                assert np.array_equal(hip.get(dv, n, np.uint32), vals)
    finally:
        hip.close([st])


def test_noalloc_small_and_limit():
    import blitzar_amd as B
    lib = B.lib()
    hip = Hip()
    try:
        for n in (1, 2, 4097):
            say(f"case noalloc sort small n={n}")
            keys, vals = _case(7 + n, n, 63, False)
            dk, dv = hip.put(keys), hip.put(vals)
            dko, dvo = hip.alloc(keys.nbytes), hip.alloc(vals.nbytes)
            _sort_noalloc(lib, dk, dv, n, 63, dko, dvo, None)  # the null stream
            ref_k, ref_v = B.sort_pairs(keys, vals, 63)
            assert np.array_equal(hip.get(dko, n, np.uint64), ref_k)
            assert np.array_equal(hip.get(dvo, n, np.uint32), ref_v)
        # beyond the fixed scratch: rejected on the host, nothing launched
        err = C.create_string_buffer(256)
        d = hip.alloc(64)
        rc = lib.bh_hip_sort_pairs_noalloc(d, d, MAX_KEYS + 1, 63, d, d, None, err, 256)
        assert rc == 1 and b"maximum" in err.value
    finally:
        hip.close()
