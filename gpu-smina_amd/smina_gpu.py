"""Python bindings (ctypes) of libsgpu.so: the C ABI of include/sgpu.h and the
host entry points of include/sgpu_host.h.

The library is built in-tree (``make -C gpu-smina_amd``); importing this
module fails loudly when it is missing, so nothing here can silently run
without the HIP kernels.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsgpu.so")

FP32, FP64, MIXED = 0, 1, 2  # sgpu_precision
OK = 0
E_INVALID, E_HIP, E_OOM, E_ATOM_OVERLAP, E_NO_DEVICE, E_NOT_FOUND = 1, 2, 3, 4, 5, 6  # sgpu_status
STATUS = {0: "OK", 1: "INVALID", 2: "HIP", 3: "OOM", 4: "ATOM_OVERLAP", 5: "NO_DEVICE", 6: "NOT_FOUND"}

I32P = C.POINTER(C.c_int)
F64P = C.POINTER(C.c_double)
U32P = C.POINTER(C.c_uint32)
I64P = C.POINTER(C.c_int64)


class Tables(C.Structure):
    _fields_ = [("n", C.c_int), ("factor", C.c_double), ("cutoff_sqr", C.c_double),
                ("smooth", F64P), ("fast", F64P), ("ncomp", C.c_int), ("qsmooth", F64P), ("qfast", F64P)]


class Splines(C.Structure):
    _fields_ = [("nseg", C.c_int), ("fraction", C.c_double), ("cutoff_sqr", C.c_double),
                ("coef", F64P), ("cut", F64P)]


class Receptor(C.Structure):
    _fields_ = [("natoms", C.c_int), ("xyz", F64P), ("type", I32P), ("charge", F64P)]


class GridDims(C.Structure):
    _fields_ = [("begin", C.c_double * 3), ("end", C.c_double * 3), ("n", C.c_int * 3)]


class FlatModel(C.Structure):
    _fields_ = [("natoms", C.c_int), ("ninflex", C.c_int), ("nframes", C.c_int), ("nligands", C.c_int),
                ("ntrees", C.c_int), ("npairs", C.c_int), ("nother", C.c_int), ("conf_size", C.c_int),
                ("change_size", C.c_int), ("num_tors", C.c_int),
                ("atom_type", I32P), ("atom_frame", I32P), ("atom_local", F64P), ("atom_init", F64P),
                ("frame_parent", I32P), ("frame_kind", I32P), ("frame_begin", I32P), ("frame_end", I32P),
                ("frame_tree", I32P), ("frame_conf", I32P), ("frame_change", I32P),
                ("frame_origin", F64P), ("frame_axis", F64P),
                ("pair_a", I32P), ("pair_b", I32P), ("pair_tri", I32P), ("atom_charge", F64P)]


class McParams(C.Structure):
    _fields_ = [("temperature", C.c_double), ("mutation_amplitude", C.c_double), ("min_rmsd", C.c_double),
                ("num_saved_mins", C.c_int), ("hunt_cap", C.c_double * 3), ("authentic_v", C.c_double * 3),
                ("cache_slope", C.c_double), ("accurate_line_search", C.c_int), ("early_term", C.c_int),
                ("early_stop_checkpoints", C.c_int), ("early_stop_first", C.c_int),
                ("early_stop_fraction", C.c_double), ("early_stop_eps", C.c_double), ("visited", C.c_int)]


class Chains(C.Structure):
    _fields_ = [("nchains", C.c_int), ("model", I32P), ("conf", F64P), ("conf_stride", C.c_int),
                ("rng", U32P), ("num_steps", I32P), ("maxiters", I32P), ("unit", I32P)]


class ChainResults(C.Structure):
    _fields_ = [("max_poses", C.c_int), ("conf_stride", C.c_int), ("coord_stride", C.c_int),
                ("count", I32P), ("e", F64P), ("conf", F64P), ("coords", F64P), ("evals", I64P)]


class EvalRequest(C.Structure):
    _fields_ = [("batch_id", C.c_int), ("grid_id", C.c_int), ("receptor_id", C.c_int),
                ("gd", C.POINTER(GridDims)), ("slope", C.c_double), ("v", C.c_double * 3), ("nconfs", C.c_int),
                ("model", I32P), ("conf", F64P), ("conf_stride", C.c_int), ("e", F64P), ("g", F64P),
                ("change_stride", C.c_int), ("coords", F64P), ("coord_stride", C.c_int)]


class RefineRequest(C.Structure):
    _fields_ = [("batch_id", C.c_int), ("receptor_id", C.c_int), ("gd", C.POINTER(GridDims)),
                ("v", C.c_double * 3), ("accurate_line_search", C.c_int), ("early_term", C.c_int),
                ("nposes", C.c_int), ("model", I32P), ("conf", F64P), ("conf_stride", C.c_int),
                ("maxiters", I32P), ("e", F64P), ("coords", F64P), ("coord_stride", C.c_int),
                ("boxes", F64P), ("tries", I32P)]


class BfgsRequest(C.Structure):
    _fields_ = [("batch_id", C.c_int), ("grid_id", C.c_int), ("receptor_id", C.c_int), ("gd", C.POINTER(GridDims)),
                ("slope", C.c_double), ("v", C.c_double * 3), ("accurate_line_search", C.c_int),
                ("early_term", C.c_int), ("nconfs", C.c_int), ("model", I32P), ("conf", F64P),
                ("conf_stride", C.c_int), ("maxiters", I32P), ("e", F64P), ("evals", I64P)]


class ScoreRequest(C.Structure):
    _fields_ = [("batch_id", C.c_int), ("receptor_id", C.c_int), ("gd", C.POINTER(GridDims)),
                ("slope", C.c_double), ("v", C.c_double * 3), ("mode", C.c_int), ("nconfs", C.c_int),
                ("model", I32P), ("conf", F64P), ("conf_stride", C.c_int), ("e", F64P)]


class ModelDesc(C.Structure):
    _fields_ = [("natoms", C.c_int), ("nmovable", C.c_int), ("type", I32P), ("xyz", F64P),
                ("nbonds", C.c_int), ("bonds", I32P), ("nframes", C.c_int),
                ("frame_parent", I32P), ("frame_begin", I32P), ("frame_end", I32P),
                ("frame_origin_atom", I32P), ("frame_axis_atom", I32P),
                ("nligands", C.c_int), ("ntrees", C.c_int), ("tree_frame", I32P), ("charge", F64P)]


class Term(C.Structure):
    _fields_ = [("kind", C.c_int), ("p", C.c_double * 5), ("weight", C.c_double)]


# sgpu_term_kind and sgpu_ci_kind (include/sgpu_model.h)
(TERM_GAUSS, TERM_REPULSION, TERM_HYDROPHOBIC, TERM_NON_HYDROPHOBIC, TERM_NON_DIR_H_BOND,
 TERM_NON_DIR_ANTI_H_BOND_QUADRATIC, TERM_DONOR_DONOR_QUADRATIC, TERM_ACCEPTOR_ACCEPTOR_QUADRATIC, TERM_VDW,
 TERM_NON_DIR_H_BOND_LJ, TERM_ELECTROSTATIC, TERM_AD4_SOLVATION) = range(12)
(CI_NUM_TORS_DIV, CI_NUM_TORS_ADD, CI_NUM_TORS_SQR, CI_NUM_TORS_SQRT, CI_NUM_TORS_DIV_SIMPLE, CI_CONSTANT) = range(6)


class CITerm(C.Structure):
    _fields_ = [("kind", C.c_int), ("weight", C.c_double)]


class Scoring(C.Structure):
    _fields_ = [("nterms", C.c_int), ("terms", C.POINTER(Term)), ("nci", C.c_int), ("ci", C.POINTER(CITerm)),
                ("atom_params", C.c_int)]


PROGRESS_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_double)


class TuneReport(C.Structure):
    _fields_ = [("nclasses", C.c_int), ("class_index", C.c_int * 4), ("ncand", C.c_int * 4),
                ("cand_waves", (C.c_int * 8) * 4), ("cand_persist", (C.c_int * 8) * 4),
                ("cand_ms", (C.c_double * 8) * 4), ("chosen_waves", C.c_int * 4), ("chosen_persist", C.c_int * 4)]


class DockSettings(C.Structure):
    _fields_ = [("device", C.c_int), ("precision", C.c_int), ("exhaustiveness", C.c_int), ("seed", C.c_int),
                ("num_modes", C.c_int), ("energy_range", C.c_double), ("hydrophobic_scale", C.c_double),
                ("ligands_per_launch", C.c_int), ("step_fraction", C.c_double), ("progress_chunks", C.c_int),
                ("progress", PROGRESS_FN), ("progress_user", C.c_void_p), ("scoring", C.c_char_p),
                ("chain_split", C.c_int), ("early_stop", C.c_int), ("fast", C.c_int), ("skip_visited", C.c_int),
                ("approximation", C.c_int), ("approx_factor", C.c_double),
                ("user_grid", C.c_void_p), ("force_cap", C.c_double), ("accurate_line", C.c_int),
                ("min_rmsd_filter", C.c_double), ("minimize_iters", C.c_int), ("minimize_early_term", C.c_int),
                ("custom_scoring", C.c_char_p)]


class UserGrid(C.Structure):
    _fields_ = [("gd", GridDims), ("data", F64P), ("map_slope", C.c_double)]


class ServerSettings(C.Structure):
    _fields_ = [("device", C.c_int), ("workers", C.c_int), ("max_batch", C.c_int), ("budget_gb", C.c_double),
                ("precision", C.c_int), ("seed", C.c_int), ("num_modes", C.c_int), ("energy_range", C.c_double),
                ("hydrophobic_scale", C.c_double), ("scoring", C.c_char_p), ("progress_chunks", C.c_int),
                ("progress", PROGRESS_FN), ("progress_user", C.c_void_p), ("fast", C.c_int),
                ("step_fraction", C.c_double)]


class ServeResult(C.Structure):
    _fields_ = [("nposes", C.c_int), ("e", C.c_double * 32), ("evals", C.c_int64), ("latency_s", C.c_double),
                ("queued_s", C.c_double), ("batch_size", C.c_int), ("fp64_error", C.c_double)]


class MinimizeResult(C.Structure):
    _fields_ = [("e", C.c_double), ("rmsd", C.c_double), ("box_begin", C.c_double * 3), ("box_end", C.c_double * 3)]


class DockResult(C.Structure):
    _fields_ = [("nposes", C.c_int), ("e", C.c_double * 32), ("evals", C.c_int64), ("seconds", C.c_double),
                ("coords", F64P), ("coords_capacity", C.c_int), ("ncoords", C.c_int),
                ("confs", F64P), ("confs_capacity", C.c_int), ("conf_size", C.c_int),
                ("fp64_error", C.c_double), ("e_search", C.c_double), ("e_fp64", C.c_double)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libsgpu.so is not built (run `make -C gpu-smina_amd` or __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "sgpu_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "sgpu_destroy": (None, [vp]),
        "sgpu_last_error": (C.c_char_p, [vp]),
        "sgpu_device_info": (C.c_int, [vp, C.c_char_p, C.c_int, I32P]),
        "sgpu_mc_capacity": (C.c_int, [vp, C.c_int, I32P]),
        "sgpu_upload_tables": (C.c_int, [vp, C.POINTER(Tables)]),
        "sgpu_upload_receptor": (C.c_int, [vp, C.POINTER(Receptor), I32P]),
        "sgpu_build_grids": (C.c_int, [vp, C.c_int, C.POINTER(GridDims), I32P, C.c_int, C.c_int, I32P]),
        "sgpu_download_grid": (C.c_int, [vp, C.c_int, C.c_int, F64P]),
        "sgpu_download_charge_grid": (C.c_int, [vp, C.c_int, C.c_int, F64P]),
        "sgpu_release_grid": (C.c_int, [vp, C.c_int]),
        "sgpu_upload_models": (C.c_int, [vp, C.POINTER(FlatModel), C.c_int, C.c_int, I32P]),
        "sgpu_release_models": (C.c_int, [vp, C.c_int]),
        "sgpu_eval_confs": (C.c_int, [vp, C.POINTER(EvalRequest)]),
        "sgpu_run_mc": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(McParams), C.POINTER(Chains),
                                  C.POINTER(ChainResults)]),
        "sgpu_run_refine": (C.c_int, [vp, C.POINTER(RefineRequest)]),
        "sgpu_run_bfgs": (C.c_int, [vp, C.POINTER(BfgsRequest)]),
        "sgpu_tune_mc": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(McParams), C.POINTER(Chains), C.c_int,
                                   C.POINTER(TuneReport)]),
        "sgpu_docking_tune": (C.c_int, [vp, C.POINTER(vp), C.c_int, C.c_int, C.POINTER(TuneReport)]),
        "sgpu_set_user_grid": (C.c_int, [vp, vp]),
        "sgpu_read_user_grid": (C.c_int, [C.c_char_p, C.c_double, C.POINTER(GridDims), F64P, C.c_int, C.c_char_p,
                                          C.c_int]),
        "sgpu_score": (C.c_int, [vp, C.POINTER(ScoreRequest)]),
        "sgpu_last_kernel_ms": (C.c_int, [vp, C.c_char_p, F64P]),
        "sgpu_set_progress": (C.c_int, [vp, C.c_int, PROGRESS_FN, vp]),
        "sgpu_mol_free": (None, [vp]),
        "sgpu_mol_desc": (C.POINTER(ModelDesc), [vp]),
        "sgpu_mol_set_charges": (C.c_int, [vp, F64P]),
        "sgpu_build_qtables": (C.c_int, [C.POINTER(Scoring), C.c_double, F64P, F64P]),
        "sgpu_synth_receptor": (vp, [C.c_uint32, C.c_int, C.c_double]),
        "sgpu_synth_ligand": (vp, [C.c_uint32, C.c_int, C.c_int, F64P]),
        "sgpu_read_pdbqt_ligand": (vp, [C.c_char_p, C.c_char_p, C.c_int]),
        "sgpu_read_pdbqt_receptor": (vp, [C.c_char_p, C.c_char_p, C.c_int]),
        "sgpu_read_pdbqt_flex": (vp, [vp, vp, C.c_char_p, C.c_char_p, C.c_int]),
        "sgpu_write_pdbqt_poses": (C.c_int, [vp, C.c_int, F64P, C.c_int, F64P, F64P, C.c_char_p, C.c_int]),
        "sgpu_write_sdf_poses": (C.c_int, [vp, C.c_int, F64P, C.c_int, F64P, F64P, C.POINTER(Scoring), vp,
                                           C.c_char_p, C.c_int]),
        "sgpu_atom_terms": (C.c_int, [C.POINTER(Scoring), vp, vp, F64P, C.c_char_p, C.c_int]),
        "sgpu_synth_flex": (vp, [vp, vp, C.c_uint32, C.c_int, C.c_int, C.c_int, F64P, C.POINTER(vp)]),
        "sgpu_vina_scoring": (None, [C.POINTER(Term), C.POINTER(CITerm), C.POINTER(Scoring), C.c_double]),
        "sgpu_conf_independent": (C.c_double, [C.POINTER(Scoring), C.c_double, C.c_double]),
        "sgpu_builtin_scoring": (C.c_int, [C.c_char_p, C.c_double, C.POINTER(Term), C.POINTER(CITerm),
                                           C.POINTER(Scoring)]),
        "sgpu_parse_custom_scoring": (C.c_int, [C.c_char_p, C.c_double, C.POINTER(Term), C.c_int, C.POINTER(CITerm),
                                                C.c_int, C.POINTER(Scoring), C.c_char_p, C.c_int]),
        "sgpu_build_tables": (C.c_int, [C.POINTER(Scoring), C.c_double, F64P, F64P]),
        "sgpu_build_splines": (C.c_int, [C.POINTER(Scoring), C.c_double, F64P, F64P]),
        "sgpu_upload_splines": (C.c_int, [vp, C.POINTER(Splines)]),
        "sgpu_flatten": (vp, [C.POINTER(ModelDesc)]),
        "sgpu_flat_get": (C.POINTER(FlatModel), [vp]),
        "sgpu_flat_free": (None, [vp]),
        "sgpu_initial_conf": (None, [vp, F64P]),
        "sgpu_make_chain": (None, [vp, C.c_uint32, C.POINTER(GridDims), F64P, U32P]),
        "sgpu_docking_create": (vp, [C.POINTER(DockSettings), vp, C.POINTER(GridDims)]),
        "sgpu_docking_free": (None, [vp]),
        "sgpu_docking_error": (C.c_char_p, [vp]),
        "sgpu_dock": (C.c_int, [vp, C.POINTER(vp), C.c_int, C.POINTER(DockResult)]),
        "sgpu_autobox": (C.c_int, [C.POINTER(vp), C.c_int, C.c_double, C.c_double, F64P, C.c_void_p]),
        "sgpu_minimize": (C.c_int, [vp, C.POINTER(vp), C.c_int, C.c_int, C.c_int, C.c_double,
                                    C.POINTER(MinimizeResult)]),
        "sgpu_server_create": (vp, [C.POINTER(ServerSettings)]),
        "sgpu_server_free": (None, [vp]),
        "sgpu_server_error": (C.c_char_p, [vp]),
        "sgpu_server_add_receptor": (C.c_int, [vp, vp, C.POINTER(GridDims), I32P, C.c_int]),
        "sgpu_server_submit": (C.c_int, [vp, C.c_int, vp, C.c_int]),
        "sgpu_server_wait": (C.c_int, [vp, C.c_int, C.POINTER(ServeResult)]),
        "sgpu_server_memory": (C.c_int, [vp, F64P, F64P, F64P]),
        "sgpu_device_memory": (C.c_int, [vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
        "sgpu_docking_prepare": (C.c_int, [vp, C.POINTER(vp), C.c_int]),
        "sgpu_docking_ctx": (vp, [vp]),
        "sgpu_docking_receptor_id": (C.c_int, [vp]),
        "sgpu_docking_full_searches": (C.c_int, [vp]),
        "sgpu_full_search_cutoff": (C.c_double, [F64P, C.c_int, C.c_double, C.c_double]),
        "sgpu_search_plan": (C.c_int, [C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_double, I32P, F64P,
                                       I32P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()

EXPORTED = ["sgpu_create", "sgpu_destroy", "sgpu_last_error", "sgpu_device_info", "sgpu_upload_tables",
            "sgpu_upload_receptor", "sgpu_build_grids", "sgpu_download_grid", "sgpu_download_charge_grid",
            "sgpu_release_grid",
            "sgpu_upload_models", "sgpu_release_models", "sgpu_eval_confs", "sgpu_run_mc", "sgpu_run_refine",
            "sgpu_run_bfgs", "sgpu_score", "sgpu_last_kernel_ms", "sgpu_set_progress", "sgpu_eval_confs_noalloc"]


def ptr(a, t):
    return a.ctypes.data_as(t)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ------------------------------------------------------------------ molecules
class Mol:
    """A library-owned molecule (sgpu_mol) with numpy views of its description."""

    def __init__(self, handle):
        if not handle:
            raise RuntimeError("molecule generation failed")
        self.h = C.c_void_p(handle)
        self.desc_ptr = lib.sgpu_mol_desc(self.h)
        d = self.desc_ptr.contents
        self.natoms = d.natoms
        self.nmovable = d.nmovable
        self.type = np.ctypeslib.as_array(d.type, (d.natoms,)).copy()
        self.xyz = np.ctypeslib.as_array(d.xyz, (d.natoms * 3,)).copy().reshape(-1, 3)
        self.nframes = d.nframes
        self.bonds = np.ctypeslib.as_array(d.bonds, (2 * d.nbonds,)).copy().reshape(-1, 2) if d.nbonds else \
            np.zeros((0, 2), np.int32)
        self.frames = [dict(parent=d.frame_parent[f], begin=d.frame_begin[f], end=d.frame_end[f],
                            origin=d.frame_origin_atom[f], axis=d.frame_axis_atom[f]) for f in range(d.nframes)]
        self.charge = np.ctypeslib.as_array(d.charge, (d.natoms,)).copy() if d.charge else np.zeros(d.natoms)

    def set_charges(self, q):
        """the atoms' partial charges (sgpu_mol_set_charges)"""
        q = f64(q)
        assert len(q) == self.natoms
        lib.sgpu_mol_set_charges(self.h, ptr(q, F64P))
        self.desc_ptr = lib.sgpu_mol_desc(self.h)
        self.charge = q.copy()

    def __del__(self):
        if getattr(self, "h", None):
            lib.sgpu_mol_free(self.h)
            self.h = None

    @staticmethod
    def receptor(seed, nheavy=1800, pocket=8.0):
        return Mol(lib.sgpu_synth_receptor(seed, nheavy, pocket))

    @staticmethod
    def ligand(seed, nheavy=24, ntors=6, center=(0.0, 0.0, 0.0)):
        c = f64(center)
        return Mol(lib.sgpu_synth_ligand(seed, nheavy, ntors, ptr(c, F64P)))

    @staticmethod
    def pdbqt_ligand(text):
        """a ligand read from PDBQT text (sgpu_read_pdbqt_ligand)"""
        err = C.create_string_buffer(512)
        h = lib.sgpu_read_pdbqt_ligand(text.encode(), err, 512)
        if not h:
            raise ValueError(err.value.decode())
        return Mol(h)

    @staticmethod
    def pdbqt_receptor(text):
        """a rigid receptor read from PDBQT text (sgpu_read_pdbqt_receptor)"""
        err = C.create_string_buffer(512)
        h = lib.sgpu_read_pdbqt_receptor(text.encode(), err, 512)
        if not h:
            raise ValueError(err.value.decode())
        return Mol(h)

    @staticmethod
    def pdbqt_flex(ligand, rigid, text):
        """the ligand plus the side chains of a flexible-residue PDBQT file
        (sgpu_read_pdbqt_flex)"""
        err = C.create_string_buffer(512)
        h = lib.sgpu_read_pdbqt_flex(ligand.h, rigid.h, text.encode(), err, 512)
        if not h:
            raise ValueError(err.value.decode())
        return Mol(h)

    def write_pdbqt(self, confs, energies, rmsds=None):
        """the poses as PDBQT MODEL blocks (sgpu_write_pdbqt_poses)"""
        confs = np.atleast_2d(f64(confs))
        e = f64(energies)
        r = None if rmsds is None else f64(rmsds)
        n = lib.sgpu_write_pdbqt_poses(self.h, len(e), ptr(confs, F64P), confs.shape[1], ptr(e, F64P),
                                       None if r is None else ptr(r, F64P), None, 0)
        if n < 0:
            raise ValueError("cannot write these poses")
        buf = C.create_string_buffer(n + 1)
        lib.sgpu_write_pdbqt_poses(self.h, len(e), ptr(confs, F64P), confs.shape[1], ptr(e, F64P),
                                   None if r is None else ptr(r, F64P), buf, n + 1)
        return buf.value.decode()

    def write_sdf(self, confs, energies, rmsds=None, atom_terms=None):
        """the poses as SDF (sgpu_write_sdf_poses); atom_terms: (scoring,
        receptor) to add each pose's atomic_interaction_terms item"""
        confs = np.atleast_2d(f64(confs))
        e = f64(energies)
        r = None if rmsds is None else f64(rmsds)
        sf, rec = (C.byref(atom_terms[0]), atom_terms[1].h) if atom_terms else (None, None)
        args = (self.h, len(e), ptr(confs, F64P), confs.shape[1], ptr(e, F64P), None if r is None else ptr(r, F64P),
                sf, rec)
        n = lib.sgpu_write_sdf_poses(*args, None, 0)
        if n < 0:
            raise ValueError("cannot write these poses")
        buf = C.create_string_buffer(n + 1)
        lib.sgpu_write_sdf_poses(*args, buf, n + 1)
        return buf.value.decode()

    def atom_terms(self, scoring, receptor, conf):
        """--atom_terms: the per-atom term table of this pose (sgpu_atom_terms)"""
        c = f64(conf)
        n = lib.sgpu_atom_terms(C.byref(scoring), receptor.h, self.h, ptr(c, F64P), None, 0)
        if n < 0:
            raise ValueError("cannot compute atom terms")
        buf = C.create_string_buffer(n + 1)
        lib.sgpu_atom_terms(C.byref(scoring), receptor.h, self.h, ptr(c, F64P), buf, n + 1)
        return buf.value.decode()

    @staticmethod
    def flex(receptor, ligand, seed, nflex=4, min_chi=2, max_chi=4, center=(0.0, 0.0, 0.0)):
        """(model, rigid): `ligand` plus `nflex` flexible side chains cut from
        `receptor`, and the receptor without them (sgpu_synth_flex)."""
        c = f64(center)
        rigid = C.c_void_p()
        h = lib.sgpu_synth_flex(receptor.h, ligand.h, seed, nflex, min_chi, max_chi, ptr(c, F64P), C.byref(rigid))
        if not h:
            raise RuntimeError("receptor has too few chain ends for the side chains")
        return Mol(h), Mol(rigid.value)


def builtin_scoring(name="vina", hydrophobic_scale=1.0):
    """a built-in scoring function by smina's name (sgpu_builtin_scoring)"""
    terms, ci, sf = (Term * 8)(), (CITerm * 4)(), Scoring()
    rc = lib.sgpu_builtin_scoring(name.encode(), hydrophobic_scale, terms, ci, C.byref(sf))
    if rc != 0:
        raise ValueError(f"scoring {name!r}: unknown")
    sf._keep = (terms, ci)  # keep alive
    return sf


def custom_scoring(text, scale=1.0):
    """--custom_scoring from the file's text (sgpu_parse_custom_scoring);
    ValueError names the line of an unknown or unsupported term"""
    n = max(1, text.count("\n") + 1)
    terms, ci, sf = (Term * n)(), (CITerm * n)(), Scoring()
    err = C.create_string_buffer(512)
    if lib.sgpu_parse_custom_scoring(text.encode(), scale, terms, n, ci, n, C.byref(sf), err, 512) != 0:
        raise ValueError(err.value.decode())
    sf._keep = (terms, ci)
    return sf


def vina_scoring(hydrophobic_scale=1.0):
    return builtin_scoring("vina", hydrophobic_scale)


def build_tables(sf, factor=32.0):
    n = lib.sgpu_build_tables(C.byref(sf), factor, None, None)
    smooth = np.zeros(378 * n * 2)
    fast = np.zeros(378 * n)
    lib.sgpu_build_tables(C.byref(sf), factor, ptr(smooth, F64P), ptr(fast, F64P))
    return n, smooth, fast


def build_qtables(sf, factor=32.0):
    """(ncomp, qsmooth, qfast): the charge-dependent components of the
    tables (None arrays when the scoring has no charge-dependent terms)"""
    n = lib.sgpu_build_tables(C.byref(sf), factor, None, None)
    k = lib.sgpu_build_qtables(C.byref(sf), factor, None, None)
    if k != 4:
        return k, None, None
    qs = np.zeros(378 * n * 6)
    qf = np.zeros(378 * n * 3)
    lib.sgpu_build_qtables(C.byref(sf), factor, ptr(qs, F64P), ptr(qf, F64P))
    return k, qs, qf


def search_plan(tasks, capacity, chain_split=0, step_fraction=0.0, early_stop=1, budget=0.0):
    """the GPU-style search plan (sgpu_search_plan): (split, fraction,
    early-stopping checkpoints)"""
    split, frac, ck = C.c_int(), C.c_double(), C.c_int()
    if lib.sgpu_search_plan(chain_split, step_fraction, early_stop, tasks, capacity, budget, C.byref(split),
                            C.byref(frac), C.byref(ck)) != 0:
        raise ValueError("bad search plan request")
    return split.value, frac.value, ck.value


def full_search_cutoff(history, fast_top=0.1, margin=0.3):
    """two-pass screening's cutoff (sgpu_full_search_cutoff)"""
    h = f64(history)
    return lib.sgpu_full_search_cutoff(ptr(h, F64P), len(h), fast_top, margin)


def build_splines(sf, factor=10.0):
    """precalculate_splines(sf, factor): nseg, coef [378, nseg, 5], cut [378]"""
    nseg = lib.sgpu_build_splines(C.byref(sf), factor, None, None)
    if nseg < 0:
        raise ValueError("bad spline factor")
    coef = np.zeros((378, nseg, 5))
    cut = np.zeros(378)
    lib.sgpu_build_splines(C.byref(sf), factor, ptr(coef, F64P), ptr(cut, F64P))
    return nseg, coef, cut


def grid_dims(center, size, spacing=0.375):
    """smina's box: n = ceil(size / spacing) intervals, the span widened to
    n * spacing around the center (src/main/main.cpp:1370-1382)."""
    gd = GridDims()
    for i in range(3):
        n = int(np.ceil(size[i] / spacing))
        real = spacing * n
        gd.begin[i] = center[i] - real / 2
        gd.end[i] = gd.begin[i] + real
        gd.n[i] = n
    return gd


def autobox(mols, add=4.0, spacing=0.375, size=None):
    """--autobox_ligand / --autobox_add (sgpu_autobox): the box around every
    atom of mols, each side grown by add; size components > 0 replace the
    grown extent (--size_x/y/z)"""
    arr = (C.c_void_p * len(mols))(*[m.h.value for m in mols])
    sz = f64(size if size is not None else (0.0, 0.0, 0.0))
    gd = GridDims()
    if lib.sgpu_autobox(arr, len(mols), add, spacing, ptr(sz, F64P), C.addressof(gd)) != 0:
        raise ValueError("autobox: no atoms or a bad argument")
    return gd


def tune_dict(rep):
    """an sgpu_mc_tune_report as a list of per-class dicts"""
    out = []
    for i in range(rep.nclasses):
        out.append({"class": rep.class_index[i], "chosen": {"waves": rep.chosen_waves[i],
                                                             "persistent": bool(rep.chosen_persist[i])},
                    "candidates": [{"waves": rep.cand_waves[i][j], "persistent": bool(rep.cand_persist[i][j]),
                                    "ms": rep.cand_ms[i][j]} for j in range(rep.ncand[i])]})
    return out


def read_user_grid(text, scaling=1.0):
    """a --user_grid map file (sgpu_read_user_grid): (GridDims of its box,
    the (n0+1)(n1+1)(n2+1) values grid::init stores, x fastest)"""
    gd = GridDims()
    err = C.create_string_buffer(512)
    n = lib.sgpu_read_user_grid(text.encode(), scaling, C.byref(gd), None, 0, err, 512)
    if n < 0:
        raise ValueError(err.value.decode())
    vals = np.zeros(n)
    lib.sgpu_read_user_grid(text.encode(), scaling, C.byref(gd), ptr(vals, F64P), n, err, 512)
    return gd, vals


class Flat:
    """A flattened model (sgpu_flat) for the device layer."""

    def __init__(self, mol):
        self.h = C.c_void_p(lib.sgpu_flatten(mol.desc_ptr))
        if not self.h:
            raise RuntimeError("flatten failed")
        self.mol = mol
        self.view = lib.sgpu_flat_get(self.h).contents
        v = self.view
        self.conf_size = v.conf_size
        self.change_size = v.change_size
        self.natoms = v.natoms
        self.npairs = v.npairs
        self.nother = v.nother
        # torsions of the search's degrees of freedom (ligand and flexible)
        self.ntors_total = v.change_size - 6 * v.nligands
        self.pair_a = np.ctypeslib.as_array(v.pair_a, (v.npairs,)).copy() if v.npairs else np.zeros(0, np.int32)
        self.pair_b = np.ctypeslib.as_array(v.pair_b, (v.npairs,)).copy() if v.npairs else np.zeros(0, np.int32)
        t = np.ctypeslib.as_array(v.atom_type, (v.natoms,))
        self.nheavy = int(np.sum((t != 0) & (t != 1)))

    def __del__(self):
        if getattr(self, "h", None):
            lib.sgpu_flat_free(self.h)
            self.h = None

    def initial_conf(self):
        c = np.zeros(self.conf_size)
        lib.sgpu_initial_conf(self.h, ptr(c, F64P))
        return c

    def make_chain(self, seed, gd):
        c = np.zeros(self.conf_size)
        s = np.zeros(625, np.uint32)
        lib.sgpu_make_chain(self.h, seed, C.byref(gd), ptr(c, F64P), ptr(s, U32P))
        return c, s


# ------------------------------------------------------------------ device
class SgpuError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"sgpu {STATUS.get(status, status)}: {msg}")
        self.status = status


class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        st = lib.sgpu_create(device, C.byref(h))
        if st != OK:
            raise SgpuError(st, "sgpu_create failed")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib.sgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def check(self, st):
        if st != OK:
            raise SgpuError(st, lib.sgpu_last_error(self.h).decode())

    def device_info(self):
        name = C.create_string_buffer(256)
        cus = C.c_int()
        self.check(lib.sgpu_device_info(self.h, name, 256, C.byref(cus)))
        return name.value.decode(), cus.value

    def upload_tables(self, n, smooth, fast, factor=32.0, cutoff_sqr=64.0, qsmooth=None, qfast=None):
        """the linear tables; with charge-dependent terms also their three
        other components (qsmooth [378, n, 3, 2], qfast [378, n, 3])"""
        self._tables = (f64(smooth), f64(fast), None if qsmooth is None else f64(qsmooth),
                        None if qfast is None else f64(qfast))
        t = Tables(n, factor, cutoff_sqr, ptr(self._tables[0], F64P), ptr(self._tables[1], F64P),
                   1 if qsmooth is None else 4, None if qsmooth is None else ptr(self._tables[2], F64P),
                   None if qfast is None else ptr(self._tables[3], F64P))
        self.check(lib.sgpu_upload_tables(self.h, C.byref(t)))

    def upload_splines(self, nseg, coef, cut, cutoff=8.0):
        self._splines = (f64(coef).reshape(-1), f64(cut))
        s = Splines(nseg, cutoff / nseg, cutoff * cutoff, ptr(self._splines[0], F64P), ptr(self._splines[1], F64P))
        self.check(lib.sgpu_upload_splines(self.h, C.byref(s)))

    def set_user_grid(self, ugd=None, values=None, map_slope=1e6):
        """sgpu_set_user_grid: the --user_grid map (None: remove it)"""
        if ugd is None:
            self.check(lib.sgpu_set_user_grid(self.h, None))
            return
        vals = f64(values)
        ug = UserGrid(ugd, ptr(vals, F64P), map_slope)
        self.check(lib.sgpu_set_user_grid(self.h, C.byref(ug)))

    def upload_receptor(self, xyz, types, charges=None):
        xyz = f64(xyz).reshape(-1)
        types = i32(types)
        q = None if charges is None else f64(charges)
        r = Receptor(len(types), ptr(xyz, F64P), ptr(types, I32P), None if q is None else ptr(q, F64P))
        rid = C.c_int()
        self.check(lib.sgpu_upload_receptor(self.h, C.byref(r), C.byref(rid)))
        return rid.value

    def build_grids(self, rid, gd, types, precision):
        t = i32(types)
        gid = C.c_int()
        self.check(lib.sgpu_build_grids(self.h, rid, C.byref(gd), ptr(t, I32P), len(t), precision, C.byref(gid)))
        return gid.value

    def download_grid(self, gid, t, gd, charge=False):
        """type t's map (charge=True: its charge map), x fastest"""
        out = np.zeros((gd.n[0] + 1) * (gd.n[1] + 1) * (gd.n[2] + 1))
        fn = lib.sgpu_download_charge_grid if charge else lib.sgpu_download_grid
        self.check(fn(self.h, gid, t, ptr(out, F64P)))
        return out

    def upload_models(self, flats, precision):
        arr = (FlatModel * len(flats))(*[f.view for f in flats])
        bid = C.c_int()
        self.check(lib.sgpu_upload_models(self.h, arr, len(flats), precision, C.byref(bid)))
        return bid.value

    def release_models(self, bid):
        self.check(lib.sgpu_release_models(self.h, bid))

    def eval_confs(self, bid, models, confs, change_size, natoms, v, grid_id=-1, receptor_id=0, gd=None, slope=1e6):
        models = i32(models)
        confs = f64(confs)
        n = len(models)
        e = np.zeros(n)
        g = np.zeros((n, change_size))
        x = np.zeros((n, 3 * natoms))
        q = EvalRequest()
        q.batch_id, q.grid_id, q.receptor_id = bid, grid_id, receptor_id
        q.gd = C.pointer(gd) if gd is not None else None
        q.slope = slope
        for k in range(3):
            q.v[k] = v[k]
        q.nconfs = n
        q.model, q.conf, q.conf_stride = ptr(models, I32P), ptr(confs, F64P), confs.shape[1]
        q.e, q.g, q.change_stride = ptr(e, F64P), ptr(g, F64P), change_size
        q.coords, q.coord_stride = ptr(x, F64P), 3 * natoms
        self.check(lib.sgpu_eval_confs(self.h, C.byref(q)))
        return e, g, x

    def run_mc(self, bid, gid, params, models, confs, rngs, steps, maxiters, max_poses=20, coord_stride=None,
               units=None):
        """units: each chain's search unit within its model (early stopping)"""
        models = i32(models)
        confs = f64(confs)
        rngs = np.ascontiguousarray(rngs, dtype=np.uint32)
        steps = i32(steps)
        maxiters = i32(maxiters)
        units = None if units is None else i32(units)
        n = len(models)
        cs = confs.shape[1]
        xs = coord_stride or 3 * 256
        ch = Chains(n, ptr(models, I32P), ptr(confs, F64P), cs, ptr(rngs, U32P), ptr(steps, I32P),
                    ptr(maxiters, I32P), ptr(units, I32P) if units is not None else None)
        cnt = np.zeros(n, np.int32)
        e = np.zeros((n, max_poses))
        c = np.zeros((n, max_poses, cs))
        x = np.zeros((n, max_poses, xs))
        ev = np.zeros(n, np.int64)
        res = ChainResults(max_poses, cs, xs, ptr(cnt, I32P), ptr(e, F64P), ptr(c, F64P), ptr(x, F64P),
                           ptr(ev, I64P))
        self.check(lib.sgpu_run_mc(self.h, bid, gid, C.byref(params), C.byref(ch), C.byref(res)))
        return cnt, e, c, x, ev

    def tune_mc(self, bid, gid, params, models, confs, rngs, maxiters, probe_steps=8):
        """sgpu_tune_mc on these chains; returns tune_dict's list"""
        models = i32(models)
        confs = f64(confs)
        rngs = np.ascontiguousarray(rngs, dtype=np.uint32)
        steps = np.full(len(models), probe_steps, np.int32)
        maxiters = i32(maxiters)
        ch = Chains(len(models), ptr(models, I32P), ptr(confs, F64P), confs.shape[1], ptr(rngs, U32P),
                    ptr(steps, I32P), ptr(maxiters, I32P), None)
        rep = TuneReport()
        self.check(lib.sgpu_tune_mc(self.h, bid, gid, C.byref(params), C.byref(ch), probe_steps, C.byref(rep)))
        return tune_dict(rep)

    def refine(self, bid, rid, gd, models, confs, maxiters, v=(1000.0, 1000.0, 1000.0), coord_stride=3 * 256,
               boxes=None, accurate=False, tries=False):
        """refine_structure of each pose; boxes: optional [n, 6] per-pose
        (begin, end) inside gd; tries: also return the box-slope increases
        each refinement took (its last slope is 10 x 10^tries)"""
        models = i32(models)
        confs = f64(confs).copy()
        maxiters = i32(maxiters)
        boxes = None if boxes is None else f64(boxes)
        n = len(models)
        e = np.zeros(n)
        x = np.zeros((n, coord_stride))
        q = RefineRequest()
        q.batch_id, q.receptor_id, q.gd = bid, rid, C.pointer(gd)
        for k in range(3):
            q.v[k] = v[k]
        q.nposes = n
        q.model, q.conf, q.conf_stride = ptr(models, I32P), ptr(confs, F64P), confs.shape[1]
        q.maxiters, q.e, q.coords, q.coord_stride = ptr(maxiters, I32P), ptr(e, F64P), ptr(x, F64P), coord_stride
        q.accurate_line_search = int(accurate)
        q.boxes = ptr(boxes, F64P) if boxes is not None else None
        nt = np.zeros(n, np.int32)
        q.tries = ptr(nt, I32P) if tries else None
        self.check(lib.sgpu_run_refine(self.h, C.byref(q)))
        return (e, confs, x, nt) if tries else (e, confs, x)

    def bfgs(self, bid, models, confs, maxiters, v, grid_id=-1, receptor_id=0, gd=None, slope=1e6, accurate=False):
        """sgpu_run_bfgs: one BFGS minimisation per conf on the maps (grid_id)
        or the direct receptor term (receptor_id, gd); returns (e, confs,
        evals)"""
        models = i32(models)
        confs = f64(confs).copy()
        maxiters = i32(maxiters)
        n = len(models)
        e = np.zeros(n)
        ev = np.zeros(n, np.int64)
        q = BfgsRequest()
        q.batch_id, q.grid_id, q.receptor_id = bid, grid_id, receptor_id
        q.gd = C.pointer(gd) if gd is not None else None
        q.slope = slope
        for k in range(3):
            q.v[k] = v[k]
        q.accurate_line_search = int(accurate)
        q.nconfs = n
        q.model, q.conf, q.conf_stride = ptr(models, I32P), ptr(confs, F64P), confs.shape[1]
        q.maxiters, q.e, q.evals = ptr(maxiters, I32P), ptr(e, F64P), ptr(ev, I64P)
        self.check(lib.sgpu_run_bfgs(self.h, C.byref(q)))
        return e, confs, ev

    def score(self, bid, rid, gd, models, confs, mode, v=(1000.0, 1000.0, 1000.0), slope=1e6):
        models = i32(models)
        confs = f64(confs)
        n = len(models)
        e = np.zeros(n)
        q = ScoreRequest()
        q.batch_id, q.receptor_id, q.gd, q.slope = bid, rid, C.pointer(gd), slope
        for k in range(3):
            q.v[k] = v[k]
        q.mode, q.nconfs = mode, n
        q.model, q.conf, q.conf_stride, q.e = ptr(models, I32P), ptr(confs, F64P), confs.shape[1], ptr(e, F64P)
        self.check(lib.sgpu_score(self.h, C.byref(q)))
        return e

    def last_ms(self, kind):
        ms = C.c_double()
        self.check(lib.sgpu_last_kernel_ms(self.h, kind.encode(), C.byref(ms)))
        return ms.value

    def set_progress(self, chunks, fn=None):
        """run_mc in `chunks` launches, fn(fraction) after each"""
        self._progress = PROGRESS_FN(lambda user, frac: fn(frac)) if fn else PROGRESS_FN()
        self.check(lib.sgpu_set_progress(self.h, chunks, self._progress, None))


def mc_params(hunt=10.0, authentic=1000.0, num_saved=20, visited=False):
    p = McParams()
    p.visited = int(visited)
    p.temperature = 1.2
    p.mutation_amplitude = 2.0
    p.min_rmsd = 1.0
    p.num_saved_mins = num_saved
    for k in range(3):
        p.hunt_cap[k] = hunt
        p.authentic_v[k] = authentic
    p.cache_slope = 1e6
    p.accurate_line_search = 0
    p.early_term = 0
    return p


def search_params(natoms, ntors_total, nligands=1):
    """setup_search (src/main/main.cpp:513-529): steps and BFGS iterations."""
    heuristic = natoms + 10 * (ntors_total + 6 * nligands)
    return int(70 * 3 * (50 + heuristic) // 2), (25 + natoms) // 3


class Docking:
    """GpuDocking driver: search, merge, refine, FP64 rescoring."""

    def __init__(self, receptor, gd, fp64=False, exhaustiveness=8, seed=1, num_modes=9, energy_range=3.0,
                 hydrophobic_scale=1.0, ligands_per_launch=0, device=0, step_fraction=1.0, precision=None,
                 progress=None, progress_chunks=1, scoring="vina", chain_split=0, early_stop=False, fast=False,
                 skip_visited=False, approximation="linear", approx_factor=0.0, user_grid=None, force_cap=1000.0,
                 accurate_line=False, min_rmsd_filter=1.0, minimize_iters=0, minimize_early_term=False,
                 custom_scoring=None):
        """precision: FP32, FP64 or MIXED (default: FP64 if fp64 else FP32);
        progress(fraction): called after each of progress_chunks launches
        of the Monte Carlo search; chain_split, early_stop, skip_visited,
        fast: the GPU-style search options --gpu_split, --gpu_early_stop,
        --gpu_qvina, --fast; approximation "linear" or "spline" (--minimize's
        default, with factor 10 unless approx_factor is given); user_grid:
        (GridDims, values) of --user_grid (read_user_grid); force_cap:
        --force_cap (authentic_v of the refinement and the final scoring;
        the chains keep monte_carlo's 1000); accurate_line: --accurate_line
        (every BFGS with the accurate line search); min_rmsd_filter:
        --min_rmsd_filter; minimize_iters, minimize_early_term: every BFGS's
        iterations (0: (25 + movable atoms) / 3) and early termination;
        custom_scoring: the text of a --custom_scoring file, in place of
        scoring"""
        prec = precision if precision is not None else (FP64 if fp64 else FP32)
        # keep the callback alive as long as the docking object
        self._progress = PROGRESS_FN(lambda user, frac: progress(frac)) if progress else PROGRESS_FN()
        s = DockSettings(device, prec, exhaustiveness, seed, num_modes, energy_range, hydrophobic_scale,
                         ligands_per_launch, step_fraction, progress_chunks, self._progress, None,
                         scoring.encode(), chain_split, int(early_stop), int(fast), int(skip_visited),
                         {"linear": 0, "spline": 1}[approximation], approx_factor)
        s.force_cap = force_cap
        s.accurate_line = int(accurate_line)
        s.min_rmsd_filter = min_rmsd_filter
        s.minimize_iters = minimize_iters
        s.minimize_early_term = int(minimize_early_term)
        self._custom = custom_scoring.encode() if custom_scoring else None
        s.custom_scoring = self._custom
        ug = None
        if user_grid is not None:
            ugd, vals = user_grid
            vals = f64(vals)
            ug = UserGrid(ugd, ptr(vals, F64P), 1e6)
            s.user_grid = C.cast(C.pointer(ug), C.c_void_p)
        self.gd = gd
        self.h = C.c_void_p(lib.sgpu_docking_create(C.byref(s), receptor.h, C.byref(gd)))
        err = lib.sgpu_docking_error(self.h)
        if err:
            msg = err.decode()
            lib.sgpu_docking_free(self.h)
            self.h = None
            raise RuntimeError(msg)

    def __del__(self):
        if getattr(self, "h", None):
            lib.sgpu_docking_free(self.h)
            self.h = None

    def prepare(self, ligands):
        arr = (C.c_void_p * len(ligands))(*[m.h.value for m in ligands])
        if lib.sgpu_docking_prepare(self.h, arr, len(ligands)) != 0:
            raise RuntimeError(lib.sgpu_docking_error(self.h).decode())

    def kernel_ms(self, kind):
        ms = C.c_double()
        ctx = lib.sgpu_docking_ctx(self.h)
        if lib.sgpu_last_kernel_ms(ctx, kind.encode(), C.byref(ms)) != OK:
            raise RuntimeError(lib.sgpu_last_error(ctx).decode())
        return ms.value

    def dock(self, ligands, top_coords=False, confs=False):
        """per ligand: energies of the modes, evaluations, seconds; with
        top_coords, the top pose's heavy-atom coordinates ([H, 3]); with
        confs, every mode's conf ([modes, conf_size])"""
        arr = (C.c_void_p * len(ligands))(*[m.h.value for m in ligands])
        out = (DockResult * len(ligands))()
        bufs, cbufs = [], []
        if top_coords:
            for r in out:
                b = np.zeros(3 * 256)
                bufs.append(b)
                r.coords = ptr(b, F64P)
                r.coords_capacity = b.size
        if confs:
            for r in out:
                b = np.zeros(32 * 512)
                cbufs.append(b)
                r.confs = ptr(b, F64P)
                r.confs_capacity = b.size
        if lib.sgpu_dock(self.h, arr, len(ligands), out) != 0:
            raise RuntimeError(lib.sgpu_docking_error(self.h).decode())
        # fp64_error: |e_search - e_fp64| of the top pose (FP32 and mixed
        # searches; None for FP64 or a ligand without poses)
        res = [dict(e=list(r.e[:r.nposes]), evals=r.evals, seconds=r.seconds,
                    fp64_error=r.fp64_error if r.fp64_error >= 0 else None, e_search=r.e_search, e_fp64=r.e_fp64)
               for r in out]
        for k, b in enumerate(bufs):
            res[k]["coords"] = b[:out[k].ncoords].reshape(-1, 3).copy()
        for k, b in enumerate(cbufs):
            cs = out[k].conf_size
            res[k]["confs"] = b[:out[k].nposes * cs].reshape(-1, cs).copy() if cs else np.zeros((0, 0))
        return res

    def tune(self, ligands, probe_steps=512):
        """the launch-plan autotuner on the chains a docking of these
        ligands would run: per size class, every candidate plan's probe time
        and the plan kept for later dockings"""
        arr = (C.c_void_p * len(ligands))(*[m.h.value for m in ligands])
        rep = TuneReport()
        if lib.sgpu_docking_tune(self.h, arr, len(ligands), probe_steps, C.byref(rep)) != 0:
            raise RuntimeError(lib.sgpu_docking_error(self.h).decode())
        return tune_dict(rep)

    def full_searches(self):
        """ligands of the last dock() that got the full search"""
        return lib.sgpu_docking_full_searches(self.h)

    def minimize(self, ligands, maxiters=0, accurate=True, force_cap=10.0):
        """--minimize: FP64 refine_structure of each parsed pose in its own box,
        every BFGS under force_cap (--minimize's 10; --local_only keeps
        --force_cap's 1000)"""
        arr = (C.c_void_p * len(ligands))(*[m.h.value for m in ligands])
        out = (MinimizeResult * len(ligands))()
        if lib.sgpu_minimize(self.h, arr, len(ligands), maxiters, int(accurate), force_cap, out) != 0:
            raise RuntimeError(lib.sgpu_docking_error(self.h).decode())
        return [dict(e=r.e, rmsd=r.rmsd, box=(tuple(r.box_begin), tuple(r.box_end))) for r in out]


class Server:
    """The serving path: receptors resident on the device, docking jobs
    admitted as they arrive and batched continuously over `workers`
    concurrent launches (sgpu_server_*)."""

    def __init__(self, workers=2, max_batch=64, budget_gb=0.0, fp64=False, seed=1, num_modes=9,
                 energy_range=3.0, hydrophobic_scale=1.0, device=0, precision=None, scoring="vina", progress=None,
                 progress_chunks=1, fast=False, step_fraction=1.0):
        """progress(fraction): after each of progress_chunks parts of every
        launch's Monte Carlo search, on the worker's thread; budget_gb: the
        device-wide memory account (0: 75% of the device's memory)"""
        prec = precision if precision is not None else (FP64 if fp64 else FP32)
        self._progress = PROGRESS_FN(lambda user, frac: progress(frac)) if progress else PROGRESS_FN()
        s = ServerSettings(device, workers, max_batch, budget_gb, prec, seed, num_modes, energy_range,
                           hydrophobic_scale, scoring.encode(), progress_chunks, self._progress, None, int(fast),
                           step_fraction)
        self.h = C.c_void_p(lib.sgpu_server_create(C.byref(s)))
        err = lib.sgpu_server_error(self.h)
        if err:
            msg = err.decode()
            lib.sgpu_server_free(self.h)
            self.h = None
            raise RuntimeError(msg)
        self._keep = {}

    def __del__(self):
        if getattr(self, "h", None):
            lib.sgpu_server_free(self.h)
            self.h = None

    def _check(self, v):
        if v < 0:
            raise RuntimeError(lib.sgpu_server_error(self.h).decode())
        return v

    def add_receptor(self, receptor, gd, types):
        t = i32(types)
        self._keep[("rec", id(receptor))] = (receptor, gd)
        return self._check(lib.sgpu_server_add_receptor(self.h, receptor.h, C.byref(gd), ptr(t, I32P), len(t)))

    def submit(self, receptor_index, ligand, exhaustiveness=8):
        job = self._check(lib.sgpu_server_submit(self.h, receptor_index, ligand.h, exhaustiveness))
        self._keep[job] = ligand
        return job

    def memory(self):
        """the device-wide memory account: budget, resident maps (once per
        device), peak bytes of the launches in flight"""
        v = np.zeros(3)
        self._check(lib.sgpu_server_memory(self.h, ptr(v[0:], F64P), ptr(v[1:], F64P), ptr(v[2:], F64P)))
        return {"budget": v[0], "resident_maps": v[1], "peak_inflight": v[2]}

    def wait(self, job):
        r = ServeResult()
        self._check(lib.sgpu_server_wait(self.h, job, C.byref(r)))
        self._keep.pop(job, None)
        return dict(e=list(r.e[:r.nposes]), evals=r.evals, latency_s=r.latency_s, queued_s=r.queued_s,
                    batch_size=r.batch_size, fp64_error=r.fp64_error if r.fp64_error >= 0 else None)
