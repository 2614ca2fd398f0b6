"""se_grids_step_noalloc (the batched step over caller-owned grids that
allocates nothing) against se_batch_run on se_worlds, bit for bit.

Two calls in a row through the C ABI, the second with fewer and smaller grids
and fewer modifications than the first, so that stale contents of the shared
scratch array would show: grids that fit one LDS tile and grids that span
several, different step counts, modifications on the first step, and a grid
whose next frame is the clearing frame.  The grids are device buffers the test
allocates itself through the HIP runtime the library uses (torch carries a
HIP runtime of its own and is not used here), and the call is enqueued on a
stream the test creates.
"""
import ctypes

import numpy as np
import pytest

import sandengine_amd as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FILL = ([0] + list(range(3, 15)), [40] + [5] * 12)  # twelve materials
H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def hip():
    S.lib()  # the library's HIP runtime is loaded first; the same soname resolves to it
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return h


def ok(rc):
    assert rc == 0, f"HIP call failed with {rc}"


class DevGrid:
    """Two device buffers of pitch*h u32 cells (zeroed), owned by the test."""

    def __init__(self, H, w, h, pitch):
        self.H, self.w, self.h, self.pitch = H, w, h, pitch
        self.nbytes = 4 * pitch * h
        self.p = ctypes.c_void_p()
        ok(H.hipMalloc(ctypes.byref(self.p), 2 * self.nbytes))
        z = np.zeros(2 * pitch * h, np.uint32)
        ok(H.hipMemcpy(self.p, z.ctypes.data, 2 * self.nbytes, H2D))

    def ptr(self, i):
        return self.p.value + i * self.nbytes

    def upload(self, i, g):
        a = np.zeros((self.h, self.pitch), np.uint32)
        a[:, :self.w] = g
        ok(self.H.hipMemcpy(self.ptr(i), a.ctypes.data, self.nbytes, H2D))

    def download(self, i):
        a = np.zeros((self.h, self.pitch), np.uint32)
        ok(self.H.hipMemcpy(a.ctypes.data, self.ptr(i), self.nbytes, D2H))
        return a[:, :self.w]

    def free(self):
        ok(self.H.hipFree(self.p))


def run_case(ctx, rs, specs, seed):
    H = hip()
    stream = ctypes.c_void_p()
    ok(H.hipStreamCreate(ctypes.byref(stream)))
    try:
        _run_case(ctx, rs, specs, seed, H, stream)
    finally:
        ok(H.hipStreamDestroy(stream))


def _run_case(ctx, rs, specs, seed, H, stream):
    mods_all, grids, bufs, refs = [], [], [], []
    for j, (w, h, steps, frame, nm) in enumerate(specs):
        pitch = (w + 63) // 64 * 64
        g0 = O.fill(seed * 100 + j, w, h, *FILL)
        rng = np.random.RandomState(seed * 100 + j)
        mods = [(int(rng.randint(0, w)), int(rng.randint(0, h)), int(rng.randint(0, 2)), int(rng.randint(1, 24)),
                 int(rng.randint(0, 15))) for _ in range(nm)]
        # the existing entry point: a world stepped by se_batch_run
        wd = S.World(ctx, rs, w, h)
        wd.frame = frame
        wd.upload(g0)
        for m in mods:
            wd.push_modification(*m)
        refs.append((wd, steps))
        # the caller-owned pair of grids
        t = DevGrid(H, w, h, pitch)
        t.upload(0, g0)
        bufs.append(t)
        grids.append(dict(grid0=t.ptr(0), grid1=t.ptr(1), w=w, h=h, pitch=pitch, cur=0, frame=frame,
                          nsteps=steps, mod_first=len(mods_all), mod_count=nm))
        mods_all += mods
    b = S.Batch(ctx)
    for wd, steps in refs:
        b.add(wd, steps)
    b.run(mode=2)
    out = S.grids_step_noalloc(ctx, rs, grids, mods_all, stream.value)
    ok(H.hipStreamSynchronize(stream))
    for j, ((wd, steps), (cur, frame), t) in enumerate(zip(refs, out, bufs)):
        got = t.download(cur)
        t.free()
        assert frame == wd.frame, (j, frame, wd.frame)
        want = wd.download()
        bad = int((got != want).sum())
        assert bad == 0, f"grid {j} {specs[j]}: {bad} cells differ from se_batch_run"
        wd.close()
    b.close()


def test_noalloc_equals_batch_run_twice():
    print("\n[gpu] no-alloc batched step vs se_batch_run: two calls, the second smaller", flush=True)
    ctx = S.Context(0)
    rs = S.RuleSet(ctx, S.bundled_ruleset("twelve_materials"))
    # (w, h, steps, frame before, modifications)
    first = [(60, 50, 37, 1, 6), (112, 112, 23, 1, 0), (113, 97, 29, 5, 4), (300, 260, 41, 1, 9),
             (250, 64, 9, 1, 3), (90, 300, 17, 0, 5), (200, 200, 0, 1, 2)]
    second = [(40, 30, 11, 3, 1), (130, 120, 19, 1, 0), (64, 200, 7, 1, 2)]
    run_case(ctx, rs, first, 1)
    run_case(ctx, rs, second, 2)
    ctx.close()


def test_noalloc_rejects_sizes_beyond_its_limits():
    ctx = S.Context(0)
    rs = S.RuleSet(ctx, S.bundled_ruleset("four_materials"))
    g = dict(grid0=1 << 20, grid1=1 << 21, w=8, h=8, pitch=64, cur=0, frame=1, nsteps=1)
    with pytest.raises(S.SandengineError, match="SE_NOALLOC_MAX_GRIDS"):
        S.grids_step_noalloc(ctx, rs, [g] * 2049)
    with pytest.raises(S.SandengineError, match="SE_NOALLOC_MAX_MODS"):
        S.grids_step_noalloc(ctx, rs, [g], [(1, 1, 0, 1, 3)] * 2049)
    with pytest.raises(S.SandengineError, match="modification range"):
        S.grids_step_noalloc(ctx, rs, [dict(g, mod_first=0, mod_count=2)], [(1, 1, 0, 1, 3)])
    ctx.close()
