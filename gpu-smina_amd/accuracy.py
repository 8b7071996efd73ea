"""The reference's FP32 accuracy rule for device evaluations against FP64
(src/gpu/tests/gpu_eval_test.cpp:55-92, :186-233): discontinuity flags
computed from FP64 coordinates, and the rule itself.  Used by the parity
tests and by bench.py --single."""
import numpy as np


def discontinuity_flags(flat, gd, rec, xs, backend, cutoff_sqr=64.0):
    """the reference's discontinuity flags (src/gpu/tests/gpu_eval_test.cpp:
    186-233): a conf is flagged when, in its FP64 coordinates xs[k], a
    movable heavy atom sits within 1e-4 grid units of a map cell face (maps
    only), or a receptor heavy atom lies within 2e-4 A^2 of r^2 = cutoff of
    the box-clamped atom, or an interacting pair lies within 2e-4 A^2 of the
    cutoff.  Beside the reference's flags, an atom within 1e-4 grid units of
    the box edge (where the out-of-box penalty starts) is flagged on the
    maps as well."""
    v = flat.view
    types = np.ctypeslib.as_array(v.atom_type, (v.natoms + v.ninflex,)).copy()
    fixed = np.ctypeslib.as_array(v.atom_init, (3 * (v.natoms + v.ninflex),)).reshape(-1, 3)[v.natoms:] \
        if v.ninflex else np.zeros((0, 3))
    heavy = np.nonzero((types[:v.natoms] > 1) & (types[:v.natoms] < 32))[0]
    begin = np.array([gd.begin[i] for i in range(3)])
    end = np.array([gd.end[i] for i in range(3)])
    nint = np.array([gd.n[i] for i in range(3)], dtype=np.float64)
    factor = nint / (end - begin)
    rxyz = rec.xyz[rec.type > 1]
    flags = []
    for x in xs:
        x = np.asarray(x).reshape(-1, 3)
        allx = np.vstack([x, fixed])
        xh = x[heavy]
        flag = False
        if backend == 0:
            sd = (xh - begin) * factor
            frac = sd - np.floor(sd)
            inside = (sd > 0) & (sd < nint)
            flag |= bool(np.any(inside & (np.minimum(frac, 1 - frac) < 1e-4)))
            flag |= bool(np.any((np.abs(sd) < 1e-4) | (np.abs(sd - nint) < 1e-4)))
        adj = np.clip(xh, begin, end)
        d2 = np.sum((adj[:, None, :] - rxyz[None, :, :]) ** 2, axis=2)
        flag |= bool(np.any(np.abs(d2 - cutoff_sqr) < 2e-4))
        if flat.npairs:
            pd = np.sum((allx[flat.pair_a] - allx[flat.pair_b]) ** 2, axis=1)
            flag |= bool(np.any(np.abs(pd - cutoff_sqr) < 2e-4))
        flags.append(flag)
    return np.array(flags)


def fp32_rule(e, g, ref, flags, tol=1e-4):
    """the reference's FP32 rule (src/gpu/tests/gpu_eval_test.cpp:55-92):
    at least 99% of evaluations within tol relative (energy over max(1, |E|),
    gradient over max(1, max |g|)), and every evaluation not flagged at a
    discontinuity within 10 x tol.  Returns (fraction within, max error of
    the unflagged, the number flagged, the worst unflagged index)."""
    errs = []
    for k, (eo, go, _) in enumerate(ref):
        rel_e = abs(e[k] - eo) / max(1.0, abs(eo))
        gs = max(1.0, float(np.max(np.abs(go)))) if len(go) else 1.0
        rel_g = float(np.max(np.abs(g[k][:len(go)] - go))) / gs if len(go) else 0.0
        errs.append(max(rel_e, rel_g))
    errs = np.array(errs)
    within = float(np.mean(errs <= tol))
    unflagged = errs[~flags]
    worst = int(np.flatnonzero(~flags)[np.argmax(unflagged)]) if len(unflagged) else -1
    return within, float(unflagged.max()) if len(unflagged) else 0.0, int(flags.sum()), worst
