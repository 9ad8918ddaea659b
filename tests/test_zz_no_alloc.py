"""oka_trace_chronicle_noalloc, the variant of oka_trace_chronicle that
allocates nothing (its scratch is a static device array in the library),
against oka_trace_chronicle on the same device inputs: two calls in a row, the
second on fewer realizations, for a lattice whose bitmap lives in LDS and for
one whose tiles spill into the scratch's HBM region (and a flat HBM bitmap)."""
import ctypes as C

import numpy as np
import pytest

from onekapy_amd import _lib, sites, start_points
from onekapy_amd.sites import fixed_field
from helpers import oracle_fit_draw, small_site

pytestmark = pytest.mark.gpu


class HipMemory:
    """Device buffers from the HIP runtime liboka is linked against (through
    ctypes, so that no second runtime is initialised in the process)."""

    H2D, D2H = 1, 2

    def __init__(self):
        _lib.context()   # loads liboka and, with it, its HIP runtime
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, self.H2D) == 0
        return p.value

    def get(self, ptr, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, self.D2H) == 0
        return out

    def close(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


class DeviceSite:
    """A site's realization arrays, start points and wells in device memory."""

    def __init__(self, mem, s, d, coef, pf, umbra, npaths):
        self.mem = mem
        self.s, self.pf, self.umbra, self.P = s, pf, umbra, npaths
        f64 = lambda a: mem.put(np.asarray(a, dtype=np.float64))
        self.wells = f64(s.wells_xy())
        self.starts = f64(start_points(s.xt, s.yt, s.rt, npaths))
        self.K, self.n, self.b = f64(d["K"]), f64(d["n"]), f64(d["b"])
        self.q, self.coef = f64(d["q"]), f64(coef)
        self.site = _lib.make_site(s.wells_xy(), s.xt, s.yt, wells_ptr=self.wells)
        self.grid = _lib.make_grid(pf.xmin, pf.ymin, pf.xmax, pf.ymax, pf.deltax, pf.deltay,
                                   pf.nrows, pf.ncols)

    def run(self, fn, nreal, precision, bitmap_mode=_lib.BITMAP_AUTO, tile_slots=0):
        mem = self.mem
        ntr = nreal * self.P
        shape = (self.pf.nrows, self.pf.ncols)
        counts = mem.put(np.zeros(shape, np.int32))
        bbox = mem.put(np.zeros((ntr, 4)))
        status = mem.put(np.full(ntr, -7, np.int32))
        steps = mem.put(np.zeros(ntr, np.int64))
        opts = _lib.make_trace_opts(self.P, self.s.duration, self.s.tol, self.s.maxstep, self.umbra,
                                    self.starts, precision, True, bitmap_mode=bitmap_mode,
                                    tile_slots=tile_slots)
        st = _lib.TraceStats_()
        ctx = _lib.context()
        ctx.check(fn(ctx.h, C.byref(self.site), nreal, self.K, self.n, self.b, self.q, self.coef,
                     C.byref(opts), C.byref(self.grid), 0, None, 1, counts, bbox, status, steps,
                     C.byref(st), _lib.PTR_DEVICE))
        return (mem.get(counts, shape, np.int32), mem.get(bbox, (ntr, 4), np.float64),
                mem.get(status, ntr, np.int32), mem.get(steps, ntr, np.int64), st)


@pytest.fixture
def mem(gpu_ctx):
    m = HipMemory()
    yield m
    m.close()


STAT_KEYS = ("attempted", "accepted", "escaped", "visited", "candidates")


def _same(a, b):
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(x, y)
    for k in STAT_KEYS:
        assert getattr(a[4], k) == getattr(b[4], k), k
    assert list(a[4].status_count) == list(b[4].status_count)


def _two_calls(ds, nreals, precision, **kw):
    lib = _lib.context().lib
    for nreal in nreals:
        print(f"noalloc: nreal {nreal} precision {precision} {kw}", flush=True)
        want = ds.run(lib.oka_trace_chronicle, nreal, precision, **kw)
        got = ds.run(lib.oka_trace_chronicle_noalloc, nreal, precision, **kw)
        assert want[0].sum() > 0 and (want[2] == 0).all()
        _same(got, want)


@pytest.mark.parametrize("precision", [_lib.FP64_EXACT, _lib.FP32], ids=["fp64", "fp32"])
def test_noalloc_matches_chronicle_lds_bitmap(mem, precision):
    """The configuration-1 site (400^2 nodes: flat LDS bitmap), 12 then 5
    realizations of 1000 particles."""
    s = sites.config1(12)
    d = sites.draw(s, seed=21)
    _, _, coef = oracle_fit_draw(s, d)
    ds = DeviceSite(mem, s, d, coef, s.fixed_field(), s.umbra, s.npaths)
    _two_calls(ds, (12, 5), precision)


MODES = [("tiled_overflow", _lib.BITMAP_TILED, 2), ("hbm_flat", _lib.BITMAP_HBM_FLAT, 0)]


@pytest.mark.parametrize("name,mode,slots", MODES, ids=[m[0] for m in MODES])
@pytest.mark.parametrize("precision", [_lib.FP64_EXACT, _lib.FP32], ids=["fp64", "fp32"])
def test_noalloc_matches_chronicle_hbm_bitmaps(mem, precision, name, mode, slots):
    """A 1100 x 1200 lattice at 5 m with a 10 m umbra: tiles beyond 2 LDS slots
    (or the whole flat bitmap) live in the scratch's HBM region; 6 then 3
    realizations of 40 particles."""
    s, d = small_site(nreal=6, npaths=40)
    _, _, coef = oracle_fit_draw(s, d)
    pf = fixed_field(s.xt, s.yt, 5.0, 1100, 1200)
    ds = DeviceSite(mem, s, d, coef, pf, 10.0, 40)
    _two_calls(ds, (6, 3), precision, bitmap_mode=mode, tile_slots=slots)


def test_noalloc_rejects_what_it_cannot_run_without_allocating(mem):
    """Host arrays, the pool schedule and a bitmap larger than the scratch are
    refused on the host, before any trace kernel is launched."""
    s, d = small_site(nreal=2, npaths=8)
    _, _, coef = oracle_fit_draw(s, d)
    ds = DeviceSite(mem, s, d, coef, s.fixed_field(), s.umbra, 8)
    ctx = _lib.context()
    lib = ctx.lib
    st = _lib.TraceStats_()
    counts = mem.put(np.zeros(16, np.int32))
    opts = _lib.make_trace_opts(8, s.duration, s.tol, s.maxstep, s.umbra, ds.starts)
    host = np.zeros((ds.pf.nrows, ds.pf.ncols), np.int32)
    args = (C.byref(ds.site), 2, ds.K, ds.n, ds.b, ds.q, ds.coef)
    rc = lib.oka_trace_chronicle_noalloc(ctx.h, *args, C.byref(opts), C.byref(ds.grid), 0, None,
                                         1, host.ctypes.data, None, None, None, C.byref(st), 0)
    assert rc == _lib.ERR_ARG and b"OKA_PTR_DEVICE" in lib.oka_last_error(ctx.h)
    pool = _lib.make_trace_opts(8, s.duration, s.tol, s.maxstep, s.umbra, ds.starts, _lib.FP32,
                                schedule=_lib.SCHED_POOL)
    rc = lib.oka_trace_chronicle_noalloc(ctx.h, *args, C.byref(pool), C.byref(ds.grid), 0, None,
                                         1, counts, None, None, None, C.byref(st),
                                         _lib.PTR_DEVICE)
    assert rc == _lib.ERR_ARG
    # a 30000^2 lattice as one flat HBM bitmap needs 112 MB per workgroup
    huge = _lib.make_grid(s.xt - 15000.0, s.yt - 15000.0, s.xt + 14999.0, s.yt + 14999.0, 1.0,
                          1.0, 30000, 30000)
    flat = _lib.make_trace_opts(8, s.duration, s.tol, s.maxstep, s.umbra, ds.starts,
                                bitmap_mode=_lib.BITMAP_HBM_FLAT)
    rc = lib.oka_trace_chronicle_noalloc(ctx.h, *args, C.byref(flat), C.byref(huge), 0, None, 1,
                                         counts, None, None, None, C.byref(st), _lib.PTR_DEVICE)
    assert rc == _lib.ERR_CAPACITY
    ctx.synchronize()
    # the scratch is intact: a regular call right after still matches
    _two_calls(ds, (2,), _lib.FP64_EXACT)
