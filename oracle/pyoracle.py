"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle (liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
as the checker.  Model descriptions, scoring and boxes are passed as the
same C structs the product takes (include/sgpu.h, include/sgpu_model.h).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

F64P = C.POINTER(C.c_double)
I32P = C.POINTER(C.c_int)
U32P = C.POINTER(C.c_uint32)
I64P = C.POINTER(C.c_int64)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "orc_create": (vp, [vp, vp, vp]),
        "orc_destroy": (None, [vp]),
        "orc_set_user_grid": (None, [vp, vp, F64P, C.c_double]),
        "orc_error": (C.c_char_p, [vp]),
        "orc_table_n": (C.c_int, [vp]),
        "orc_tables": (None, [vp, F64P, F64P]),
        "orc_eval_terms": (C.c_double, [vp, C.c_int, C.c_int, C.c_double]),
        "orc_eval_components": (None, [vp, C.c_int, C.c_int, C.c_double, F64P]),
        "orc_conf_independent": (C.c_double, [vp, C.c_double, C.c_double]),
        "orc_build_grids": (C.c_int, [vp, I32P, C.c_int]),
        "orc_grid_values": (C.c_int, [vp, C.c_int, F64P]),
        "orc_table_ncomp": (C.c_int, [vp]),
        "orc_qtables": (None, [vp, F64P, F64P]),
        "orc_charge_grid_values": (C.c_int, [vp, C.c_int, F64P]),
        "orc_add_model": (C.c_int, [vp, vp]),
        "orc_model_info": (None, [vp, C.c_int, I32P]),
        "orc_initial_conf": (None, [vp, C.c_int, F64P]),
        "orc_model_pairs": (None, [vp, C.c_int, I32P, I32P, I32P]),
        "orc_num_tors": (C.c_double, [vp, C.c_int]),
        "orc_eval_deriv": (C.c_double, [vp, C.c_int, C.c_int, C.c_double, F64P, F64P, F64P, F64P]),
        "orc_eval": (C.c_double, [vp, C.c_int, C.c_double, F64P, F64P]),
        "orc_eval_intramolecular": (C.c_double, [vp, C.c_int, F64P, F64P]),
        "orc_bfgs": (C.c_double, [vp, C.c_int, C.c_int, C.c_double, F64P, C.c_int, C.c_int, C.c_int, F64P, I64P]),
        "orc_bfgs_quadratic": (C.c_double, [C.c_int, F64P, F64P, F64P, C.c_int, C.c_int, I64P]),
        "orc_randomize": (None, [vp, C.c_int, C.c_uint32, F64P, U32P]),
        "orc_use_splines": (C.c_int, [vp, C.c_double, F64P, F64P]),
        "orc_visited_check": (C.c_int, [C.c_int, C.c_int, F64P, F64P, F64P, F64P, C.c_double, F64P]),
        "orc_mc_chain": (C.c_int, [vp, C.c_int, vp, C.c_int, C.c_int, F64P, U32P, C.c_int, F64P, F64P, F64P,
                                   I64P]),
        "orc_container_add": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, F64P, F64P, I32P, C.c_double,
                                        F64P, C.c_int]),
        "orc_refine": (C.c_double, [vp, C.c_int, F64P, C.c_int, F64P, F64P]),
        "orc_refine_opts": (C.c_double, [vp, C.c_int, F64P, C.c_int, C.c_int, F64P, F64P]),
        "orc_dock": (C.c_int, [vp, C.c_int, vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double,
                               C.c_int, C.c_int, F64P, F64P, I64P, C.c_double, C.c_double, C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def _p(a, t):
    return a.ctypes.data_as(t)


def visited_check(points, energies, grads, x, f, g):
    """QuickVina 2's history fed `points` in order, then whether (x, f, g)
    is interesting"""
    points = np.ascontiguousarray(points, dtype=np.float64)
    energies = np.ascontiguousarray(energies, dtype=np.float64)
    grads = np.ascontiguousarray(grads, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    g = np.ascontiguousarray(g, dtype=np.float64)
    n = x.size
    return bool(lib.orc_visited_check(n, len(energies), _p(points, F64P), _p(energies, F64P), _p(grads, F64P),
                                      _p(x, F64P), float(f), _p(g, F64P)))


def container_add(poses, pose, max_size=20, min_rmsd=1.0):
    """add_to_output_container: poses is a list of (tag, e, coords[natoms][3]),
    pose one more; returns the new list of (tag, e, coords)"""
    natoms = len(pose[2])
    n = len(poses)
    e = np.zeros(max_size + 1)
    xyz = np.zeros((max_size + 1, natoms, 3))
    tag = np.zeros(max_size + 1, np.int32)
    for i, (t, en, x) in enumerate(poses):
        tag[i], e[i], xyz[i] = t, en, x
    tx = np.ascontiguousarray(pose[2], dtype=np.float64)
    k = lib.orc_container_add(n, natoms, max_size, min_rmsd, _p(e, F64P), _p(xyz, F64P), _p(tag, I32P), pose[1],
                              _p(tx, F64P), pose[0])
    return [(int(tag[i]), float(e[i]), xyz[i].copy()) for i in range(k)]


def bfgs_quadratic(A, b, x0, maxiters, accurate=False):
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.ascontiguousarray(x0, dtype=np.float64).copy()
    ev = np.zeros(1, np.int64)
    f = lib.orc_bfgs_quadratic(len(b), _p(A, F64P), _p(b, F64P), _p(x, F64P), maxiters, int(accurate), _p(ev, I64P))
    return f, x, int(ev[0])


class Oracle:
    """One scoring function + receptor + box, with models added to it."""

    def __init__(self, scoring, receptor_xyz, receptor_types, gd, receptor_charges=None):
        # scoring, gd: the product binding's ctypes structs (same C layout)
        self._xyz = np.ascontiguousarray(receptor_xyz, dtype=np.float64).reshape(-1)
        self._types = np.ascontiguousarray(receptor_types, dtype=np.int32)
        self._q = None if receptor_charges is None else np.ascontiguousarray(receptor_charges, dtype=np.float64)

        class Rec(C.Structure):
            _fields_ = [("natoms", C.c_int), ("xyz", F64P), ("type", I32P), ("charge", F64P)]

        self._rec = Rec(len(self._types), _p(self._xyz, F64P), _p(self._types, I32P),
                        None if self._q is None else _p(self._q, F64P))
        self._sf = scoring
        self._gd = gd
        self.h = C.c_void_p(lib.orc_create(C.addressof(scoring), C.addressof(self._rec), C.addressof(gd)))
        self.n = lib.orc_table_n(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_destroy(self.h)
            self.h = None

    def use_splines(self, factor):
        """switch to precalculate_splines(sf, factor); returns the tables:
        coef [378, nseg, 5] (x, a, b, c, d) and cut [378]"""
        nseg = int(factor * 8.0)
        coef = np.zeros((378, nseg, 5))
        cut = np.zeros(378)
        k = lib.orc_use_splines(self.h, factor, _p(coef, F64P), _p(cut, F64P))
        if k < 0:
            raise RuntimeError(self.error())
        assert k == nseg
        return coef, cut

    def error(self):
        return lib.orc_error(self.h).decode()

    def tables(self):
        smooth = np.zeros(378 * self.n * 2)
        fast = np.zeros(378 * self.n)
        lib.orc_tables(self.h, _p(smooth, F64P), _p(fast, F64P))
        return smooth, fast

    def qtables(self):
        """(ncomp, qsmooth [378, n, 3, 2], qfast [378, n, 3]); None arrays
        without charge-dependent terms"""
        k = lib.orc_table_ncomp(self.h)
        if k != 4:
            return k, None, None
        qs = np.zeros(378 * self.n * 6)
        qf = np.zeros(378 * self.n * 3)
        lib.orc_qtables(self.h, _p(qs, F64P), _p(qf, F64P))
        return k, qs, qf

    def charge_grid(self, t):
        gd = self._gd
        g = np.zeros((gd.n[0] + 1) * (gd.n[1] + 1) * (gd.n[2] + 1))
        if lib.orc_charge_grid_values(self.h, t, _p(g, F64P)) != 0:
            raise RuntimeError("no charge map for this type")
        return g

    def conf_independent(self, num_tors, e):
        return lib.orc_conf_independent(self.h, num_tors, e)

    def eval_terms(self, t1, t2, r):
        return lib.orc_eval_terms(self.h, t1, t2, r)

    def eval_components(self, t1, t2, r):
        """weighted_terms::eval_fast's four components at distance r"""
        out = np.zeros(4)
        lib.orc_eval_components(self.h, t1, t2, r, _p(out, F64P))
        return out

    def set_user_grid(self, ugd, values, map_slope=1e6):
        """--user_grid: values ((n+1)^3, x fastest) as grid::init stores
        them; call before build_grids (the maps carry it)"""
        self._ug = np.ascontiguousarray(values, dtype=np.float64)
        self._ugd = ugd
        lib.orc_set_user_grid(self.h, C.addressof(ugd), _p(self._ug, F64P), map_slope)

    def build_grids(self, types):
        t = np.ascontiguousarray(types, dtype=np.int32)
        if lib.orc_build_grids(self.h, _p(t, I32P), len(t)) != 0:
            raise RuntimeError(self.error())

    def grid(self, t):
        gd = self._gd
        out = np.zeros((gd.n[0] + 1) * (gd.n[1] + 1) * (gd.n[2] + 1))
        if lib.orc_grid_values(self.h, t, _p(out, F64P)) != 0:
            raise RuntimeError("no grid for type %d" % t)
        return out

    def add_model(self, desc_ptr):
        m = lib.orc_add_model(self.h, C.cast(desc_ptr, C.c_void_p))
        if m < 0:
            raise RuntimeError(self.error())
        return m

    def info(self, m):
        a = np.zeros(10, np.int32)
        lib.orc_model_info(self.h, m, _p(a, I32P))
        keys = ["natoms", "all_atoms", "nframes", "npairs", "nother", "conf_size", "change_size", "nheavy",
                "nligands", "ntrees"]
        return dict(zip(keys, a.tolist()))

    def pairs(self, m):
        n = self.info(m)["npairs"]
        a, b, lst = (np.zeros(n, np.int32) for _ in range(3))
        lib.orc_model_pairs(self.h, m, _p(a, I32P), _p(b, I32P), _p(lst, I32P))
        return a, b, lst

    def num_tors(self, m):
        return lib.orc_num_tors(self.h, m)

    def initial_conf(self, m):
        c = np.zeros(self.info(m)["conf_size"])
        lib.orc_initial_conf(self.h, m, _p(c, F64P))
        return c

    def eval_deriv(self, m, conf, v, backend=0, slope=1e6):
        info = self.info(m)
        conf = np.ascontiguousarray(conf, dtype=np.float64)
        g = np.zeros(info["change_size"])
        x = np.zeros(3 * info["natoms"])
        vv = np.ascontiguousarray(v, dtype=np.float64)
        e = lib.orc_eval_deriv(self.h, m, backend, slope, _p(vv, F64P), _p(conf, F64P), _p(g, F64P), _p(x, F64P))
        if np.isnan(e):
            raise RuntimeError(self.error())
        return e, g, x

    def eval(self, m, conf, v=(1000.0, 1000.0, 1000.0), slope=1e6):
        conf = np.ascontiguousarray(conf, dtype=np.float64)
        vv = np.ascontiguousarray(v, dtype=np.float64)
        return lib.orc_eval(self.h, m, slope, _p(vv, F64P), _p(conf, F64P))

    def eval_intramolecular(self, m, conf, v=(1000.0, 1000.0, 1000.0)):
        conf = np.ascontiguousarray(conf, dtype=np.float64)
        vv = np.ascontiguousarray(v, dtype=np.float64)
        return lib.orc_eval_intramolecular(self.h, m, _p(vv, F64P), _p(conf, F64P))

    def bfgs(self, m, conf, v, maxiters, backend=0, slope=1e6, accurate=False):
        conf = np.ascontiguousarray(conf, dtype=np.float64).copy()
        vv = np.ascontiguousarray(v, dtype=np.float64)
        ev = np.zeros(1, np.int64)
        f = lib.orc_bfgs(self.h, m, backend, slope, _p(vv, F64P), maxiters, int(accurate), 0, _p(conf, F64P),
                         _p(ev, I64P))
        return f, conf, int(ev[0])

    def randomize(self, m, seed):
        c = np.zeros(self.info(m)["conf_size"])
        s = np.zeros(625, np.uint32)
        lib.orc_randomize(self.h, m, seed, _p(c, F64P), _p(s, U32P))
        return c, s

    def mc_chain(self, m, params, steps, maxiters, conf0, rng_state, max_out=20):
        info = self.info(m)
        conf0 = np.ascontiguousarray(conf0, dtype=np.float64)
        rng_state = np.ascontiguousarray(rng_state, dtype=np.uint32)
        e = np.zeros(max_out)
        c = np.zeros((max_out, info["conf_size"]))
        x = np.zeros((max_out, 3 * info["nheavy"]))
        ev = np.zeros(1, np.int64)
        k = lib.orc_mc_chain(self.h, m, C.addressof(params), steps, maxiters, _p(conf0, F64P), _p(rng_state, U32P),
                             max_out, _p(e, F64P), _p(c, F64P), _p(x, F64P), _p(ev, I64P))
        if k < 0:
            raise RuntimeError(self.error())
        return e[:k], c[:k], x[:k], int(ev[0])

    def refine(self, m, conf, maxiters, v=(1000.0, 1000.0, 1000.0), accurate=False):
        info = self.info(m)
        conf = np.ascontiguousarray(conf, dtype=np.float64).copy()
        vv = np.ascontiguousarray(v, dtype=np.float64)
        x = np.zeros(3 * info["nheavy"])
        e = lib.orc_refine_opts(self.h, m, _p(vv, F64P), maxiters, int(accurate), _p(conf, F64P), _p(x, F64P))
        return e, conf, x

    def dock(self, m, params, seed=1, exhaustiveness=8, num_modes=9, energy_range=3.0, threads=1, max_out=32,
             step_fraction=1.0, split=1, force_cap=0.0, min_rmsd_filter=1.0, minimize_iters=0):
        """force_cap: the refinement's and final scoring's cap (0: the
        chains' params.authentic_v); min_rmsd_filter: remove_redundant's;
        minimize_iters: every BFGS's iterations (0: (25 + N) / 3)"""
        info = self.info(m)
        e = np.zeros(max_out)
        c = np.zeros((max_out, info["conf_size"]))
        ev = np.zeros(1, np.int64)
        k = lib.orc_dock(self.h, m, C.addressof(params), seed, exhaustiveness, num_modes, energy_range, threads,
                         step_fraction, split, max_out, _p(e, F64P), _p(c, F64P), _p(ev, I64P), force_cap,
                         min_rmsd_filter, minimize_iters)
        if k < 0:
            raise RuntimeError(self.error())
        return e[:k], c[:k], int(ev[0])
