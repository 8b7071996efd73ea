"""sgpu_eval_confs_noalloc: the allocation-free variant of sgpu_eval_confs.

Inputs and outputs are device arrays the test allocates through the HIP
runtime the library itself links (/opt/rocm's libamdhip64); the call must
give exactly what sgpu_eval_confs gives for the same confs (same kernels,
same precision), on two calls in a row with the second one smaller, on the
receptor maps and on the direct receptor term, in FP64 and FP32."""
import ctypes as C

import numpy as np
import pytest

import smina_gpu as sg
from common import gd_c0, heavy_types, ligand, random_confs, receptor

pytestmark = pytest.mark.gpu

I32P = C.POINTER(C.c_int)
F64P = C.POINTER(C.c_double)


class EvalRequestDev(C.Structure):
    _fields_ = [("batch_id", C.c_int), ("grid_id", C.c_int), ("receptor_id", C.c_int),
                ("gd", C.POINTER(sg.GridDims)), ("slope", C.c_double), ("v", C.c_double * 3), ("nconfs", C.c_int),
                ("model", C.c_void_p), ("conf", C.c_void_p), ("conf_stride", C.c_int), ("e", C.c_void_p),
                ("g", C.c_void_p), ("change_stride", C.c_int), ("coords", C.c_void_p), ("coord_stride", C.c_int),
                ("status", C.c_void_p), ("stream", C.c_void_p)]


class Hip:
    """device buffers through the HIP runtime libsgpu.so links"""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), C.c_size_t(max(nbytes, 8))) == 0
        self.bufs.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        return p

    def get(self, p, shape, dtype):
        out = np.zeros(shape, dtype=dtype)
        assert self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


@pytest.fixture(scope="module")
def setup():
    lib = sg.lib
    lib.sgpu_eval_confs_noalloc.restype = C.c_int
    lib.sgpu_eval_confs_noalloc.argtypes = [C.c_void_p, C.POINTER(EvalRequestDev)]
    rec = receptor()
    ligs = [ligand(1, 24, 6), ligand(2, 38, 10), ligand(3, 12, 2)]
    flats = [sg.Flat(l) for l in ligs]
    gd = gd_c0()
    ctx = sg.Context(0)
    n, smooth, fast = sg.build_tables(sg.vina_scoring())
    ctx.upload_tables(n, smooth, fast)
    rid = ctx.upload_receptor(rec.xyz, rec.type)
    types = heavy_types(*ligs)
    out = {"ctx": ctx, "rid": rid, "gd": gd, "flats": flats, "hip": Hip(), "lib": lib}
    for prec in (sg.FP64, sg.FP32):
        out[("grid", prec)] = ctx.build_grids(rid, gd, types, prec)
        out[("batch", prec)] = ctx.upload_models(flats, prec)
    yield out
    out["hip"].free()
    ctx.close()


def _confs(flats, gd, n, seed):
    rng = np.random.default_rng(seed)
    models = rng.integers(0, len(flats), n).astype(np.int32)
    cs = max(f.conf_size for f in flats)
    confs = np.zeros((n, cs))
    for i, m in enumerate(models):
        c = random_confs(flats[m], gd, 1, seed0=seed * 100 + i)[0]
        confs[i, :len(c)] = c
    return models, confs


def _noalloc(S, prec, backend, models, confs, status=True):
    ctx, hip, flats = S["ctx"], S["hip"], S["flats"]
    n = len(models)
    gs = max(f.change_size for f in flats)
    xs = 3 * max(f.natoms for f in flats)
    q = EvalRequestDev()
    q.batch_id = S[("batch", prec)]
    q.grid_id = S[("grid", prec)] if backend == "maps" else -1
    q.receptor_id = S["rid"]
    gd = S["gd"]
    q.gd = C.pointer(gd)
    q.slope = 1e6
    for k in range(3):
        q.v[k] = 10.0 if backend == "maps" else 1000.0
    q.nconfs = n
    q.model = hip.put(models.astype(np.int32))
    q.conf = hip.put(confs)
    q.conf_stride = confs.shape[1]
    q.e = hip.alloc(8 * n)
    q.g = hip.alloc(8 * n * gs)
    q.change_stride = gs
    q.coords = hip.alloc(8 * n * xs)
    q.coord_stride = xs
    st = hip.alloc(4) if status else None
    q.status = st
    q.stream = None
    rc = S["lib"].sgpu_eval_confs_noalloc(ctx.h, C.byref(q))
    if rc != 0:
        return rc, None
    if status:
        # the call returned once enqueued: wait for the context's stream
        assert hip.rt.hipDeviceSynchronize() == 0
    e = hip.get(q.e, (n,), np.float64)
    g = hip.get(q.g, (n, gs), np.float64)
    x = hip.get(q.coords, (n, xs), np.float64)
    word = int(hip.get(st, (1,), np.int32)[0]) if status else 0
    hip.free()
    return 0, (e, g, x, word)


def _reference(S, prec, backend, models, confs):
    flats = S["flats"]
    gs = max(f.change_size for f in flats)
    na = max(f.natoms for f in flats)
    v = (10.0,) * 3 if backend == "maps" else (1000.0,) * 3
    kw = dict(grid_id=S[("grid", prec)]) if backend == "maps" else dict(receptor_id=S["rid"], gd=S["gd"])
    return S["ctx"].eval_confs(S[("batch", prec)], models, confs, gs, na, v, **kw)


@pytest.mark.parametrize("backend", ["maps", "direct"])
@pytest.mark.parametrize("prec", [sg.FP64, sg.FP32], ids=["fp64", "fp32"])
def test_noalloc_equals_eval_confs(setup, prec, backend):
    S = setup
    flats = S["flats"]
    for call, (n, seed) in enumerate(((40, 11), (7, 12))):  # the second call smaller
        models, confs = _confs(flats, S["gd"], n, seed)
        e0, g0, x0 = _reference(S, prec, backend, models, confs)
        print(f"noalloc {backend} prec {prec} call {call}: {n} confs", flush=True)
        rc, out = _noalloc(S, prec, backend, models, confs, status=(call == 0))
        assert rc == 0, S["ctx"].check(rc)
        e, g, x, word = out
        assert word == 0
        assert np.array_equal(e, e0), (call, np.max(np.abs(e - e0)))
        for i, m in enumerate(models):
            f = flats[m]
            assert np.array_equal(g[i, :f.change_size], g0[i, :f.change_size]), (call, i)
            assert np.array_equal(x[i, :3 * f.natoms], x0[i, :3 * f.natoms]), (call, i)
            assert not g[i, f.change_size:].any() and not x[i, 3 * f.natoms:].any()


def test_noalloc_refuses_bad_requests(setup):
    S = setup
    flats = S["flats"]
    models, confs = _confs(flats, S["gd"], 4, 21)
    # a model index past the batch: the kernels clamp it and report it
    bad = models.copy()
    bad[2] = 17
    rc, _ = _noalloc(S, sg.FP64, "maps", bad, confs, status=False)
    assert rc == sg.E_INVALID and "model index" in sg.lib.sgpu_last_error(S["ctx"].h).decode()
    rc, out = _noalloc(S, sg.FP64, "maps", bad, confs, status=True)
    assert rc == 0 and out[3] == 2
    # a request past the scratch size is refused before any launch
    big = 200000
    models = np.zeros(big, dtype=np.int32)
    confs = np.zeros((big, confs.shape[1]))
    rc, _ = _noalloc(S, sg.FP64, "maps", models, confs, status=False)
    assert rc == sg.E_INVALID and "scratch" in sg.lib.sgpu_last_error(S["ctx"].h).decode()
