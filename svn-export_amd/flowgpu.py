"""Python binding of the MI355X Darcy / transport C ABI (include/flow123d_gpu.h).

Mirrors the reference's equation interface: ``DarcyFlowMH_Steady``
(update_solution, get_solution_vector, solution_precision, water balance)
and ``ConvectionTransport`` driven as TransportOperatorSplitting
(set_velocity_field, update_solution, concentrations, mass balance).
Every call goes through libflow123d_gpu.so; there is no CPU fallback: a
missing library or device raises.
"""
import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libflow123d_gpu.so")
_LIB = None

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class FlowError(RuntimeError):
    pass


class SolverOpts(C.Structure):
    _fields_ = [("r_tol", C.c_double), ("a_tol", C.c_double), ("max_it", C.c_int), ("precision", C.c_int),
                ("norm", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("residual", C.c_double),
                ("rhs_norm", C.c_double), ("ms_assembly", C.c_double), ("ms_solve", C.c_double),
                ("ms_resolve", C.c_double), ("ms_total", C.c_double), ("ms_spmv_avg", C.c_double),
                ("ms_update_avg", C.c_double), ("kernel_samples", C.c_int), ("reliable_updates", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchInfo(C.Structure):
    _fields_ = [("n_real", C.c_int), ("chunks", C.c_int), ("reliable_updates", C.c_int), ("fp64_finish", C.c_int),
                ("max_iterations", C.c_int), ("n_converged", C.c_int), ("total_iterations", C.c_longlong),
                ("ms_setup", C.c_double), ("ms_solve", C.c_double), ("ms_resolve", C.c_double),
                ("ms_total", C.c_double), ("ms_op_avg", C.c_double), ("ms_update_avg", C.c_double),
                ("grid", C.c_int), ("xcd_map", C.c_int), ("chunk", C.c_int), ("compactions", C.c_int),
                ("handovers", C.c_int), ("handover_iterations", C.c_longlong), ("ms_handover", C.c_double),
                ("ms_coarse_avg", C.c_double), ("ms_coarse_setup", C.c_double), ("ms_queue_empty", C.c_double),
                ("ms_narrow16", C.c_double), ("ms_narrow4", C.c_double), ("ms_loop_end", C.c_double),
                ("ms_handover_end", C.c_double), ("tile_rows", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# symbol -> (restype, argtypes)
SIGNATURES = {
    "fl_last_error": (C.c_char_p, []),
    "fl_mesh_read_gmsh": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "fl_mesh_from_arrays": (C.c_int, [C.c_int, _f64p, C.c_int, C.c_void_p, _i32p, _i32p, _i32p, C.c_int,
                                      C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "fl_mesh_free": (None, [C.c_void_p]),
    "fl_mesh_counts": (C.c_int, [C.c_void_p, _i32p]),
    "fl_mesh_topology": (C.c_int, [C.c_void_p] + [_i32p] * 13),
    "fl_mesh_regions": (C.c_int, [C.c_void_p, _i32p, _i32p, _i32p, C.c_char_p, C.c_int]),
    "fl_mesh_element_ids": (C.c_int, [C.c_void_p, C.c_int, _i32p]),
    "fl_element_data_open": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int, C.c_int,
                                       C.POINTER(C.c_void_p)]),
    "fl_element_data_set_time": (C.c_int, [C.c_void_p, C.c_double, C.POINTER(C.c_int)]),
    "fl_element_data_values": (C.c_int, [C.c_void_p, _f64p]),
    "fl_element_data_info": (C.c_int, [C.c_void_p, _i32p, C.c_char_p, C.c_int]),
    "fl_element_data_free": (None, [C.c_void_p]),
    "fl_msh_write_head": (C.c_int, [C.c_void_p, C.c_char_p]),
    "fl_msh_write_element_data": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_double, C.c_int, C.c_int,
                                            C.c_int, _f64p, C.c_int]),
    "fl_old_bcd_read_flow": (C.c_int, [C.c_void_p, C.c_char_p, _i32p, _f64p, _f64p, _f64p, C.POINTER(C.c_void_p)]),
    "fl_old_bcd_read_transport": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, _f64p]),
    "fl_old_bcd_free": (None, [C.c_void_p]),
    "fl_write_water_balance": (C.c_int, [C.c_void_p, C.c_char_p, C.c_double, _f64p, _f64p, _f64p, _f64p]),
    "fl_write_mass_balance": (C.c_int, [C.c_void_p, C.c_char_p, C.c_double, C.c_double, C.c_int, C.c_void_p,
                                        _f64p, _f64p, _f64p, _f64p, _f64p, _f64p]),
    "fl_darcy_write_water_balance": (C.c_int, [C.c_void_p, C.c_char_p, C.c_double]),
    "fl_darcy_element_output": (C.c_int, [C.c_void_p, _f64p, _f64p, _f64p]),
    "fl_darcy_node_pressure": (C.c_int, [C.c_void_p, _f64p]),
    "fl_transport_set_dual_porosity_sorption": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _f64p, C.c_void_p,
                                                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "fl_transport_get_conc_phase": (C.c_int, [C.c_void_p, C.c_int, _f64p]),
    "fl_transport_set_bc_conc_function": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fl_transport_set_reaction_sorption": (C.c_int, [C.c_void_p, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p,
                                                     C.c_void_p, C.c_double, C.c_int]),
    "fl_device_stream_bandwidth": (C.c_int, [C.c_size_t, C.c_int, _f64p]),
    "fl_transport_set_reaction_sorption_elementwise": (C.c_int, [C.c_void_p, C.c_int, _i32p, _i32p, _f64p, _f64p,
                                                                 _f64p, _f64p, C.c_void_p, C.c_double]),
    "fl_transport_get_reaction_sorbed": (C.c_int, [C.c_void_p, _f64p]),
    "fl_msh_write_node_data": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_double, C.c_int, C.c_int,
                                         C.c_int, _f64p, C.c_int]),
    "fl_transport_write_mass_balance": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p]),
    "fl_darcy_create": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "fl_darcy_free": (None, [C.c_void_p]),
    "fl_darcy_set_data": (C.c_int, [C.c_void_p, _f64p, _f64p, C.c_void_p, _f64p, _f64p, _i32p, _f64p, _f64p,
                                    _f64p]),
    "fl_darcy_set_conductivity": (C.c_int, [C.c_void_p, _f64p]),
    "fl_darcy_update_solution": (C.c_int, [C.c_void_p, C.POINTER(SolverOpts), C.POINTER(SolveInfo)]),
    "fl_darcy_update_solution_noalloc": (C.c_int, [C.c_void_p, C.POINTER(SolverOpts), C.POINTER(SolveInfo)]),
    "fl_darcy_set_unsteady": (C.c_int, [C.c_void_p, C.c_int, _f64p, _f64p]),
    "fl_darcy_unsteady_step": (C.c_int, [C.c_void_p, C.c_double, C.POINTER(SolverOpts), C.POINTER(SolveInfo)]),
    "fl_darcy_set_time": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double]),
    "fl_darcy_time": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "fl_darcy_set_time_upper_constraint": (C.c_int, [C.c_void_p, C.c_double]),
    "fl_darcy_planned_time": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "fl_darcy_solved_time": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "fl_darcy_choose_next_time": (C.c_int, [C.c_void_p]),
    "fl_transport_time_governor": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                             C.POINTER(C.c_int)]),
    "fl_transport_set_time_upper_constraint": (C.c_int, [C.c_void_p, C.c_double]),
    "fl_transport_planned_time": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "fl_transport_solved_time": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "fl_transport_choose_next_time": (C.c_int, [C.c_void_p]),
    "fl_transport_update_solution_step": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "fl_darcy_get_solution": (C.c_int, [C.c_void_p, _f64p]),
    "fl_darcy_device_solution": (C.c_void_p, [C.c_void_p]),
    "fl_darcy_solution_precision": (C.c_double, [C.c_void_p, C.c_double, C.c_double]),
    "fl_darcy_element_balance": (C.c_int, [C.c_void_p, _f64p]),
    "fl_darcy_water_balance": (C.c_int, [C.c_void_p, _f64p, _f64p, _f64p]),
    "fl_darcy_edge_nnz": (C.c_int, [C.c_void_p]),
    "fl_darcy_edge_system": (C.c_int, [C.c_void_p, _i32p, _i32p, _f64p, _f64p]),
    "fl_darcy_batch_create": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    "fl_darcy_batch_create_slots": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "fl_darcy_batch_free": (None, [C.c_void_p]),
    "fl_darcy_batch_set_conductivity": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "fl_darcy_batch_update_solution": (C.c_int, [C.c_void_p, C.POINTER(SolverOpts), C.POINTER(BatchInfo), _i32p,
                                                 _i32p, _f64p]),
    "fl_darcy_batch_get_solution": (C.c_int, [C.c_void_p, C.c_int, _f64p]),
    "fl_darcy_batch_diagnostics": (C.c_int, [C.c_void_p, _i32p, _i32p]),
    "fl_darcy_batch_device_solution": (C.c_void_p, [C.c_void_p, C.POINTER(C.c_int)]),
    "fl_darcy_batch_set_launch": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "fl_darcy_batch_autotune": (C.c_int, [C.c_void_p]),
    "fl_darcy_batch_set_handover": (C.c_int, [C.c_void_p, C.c_int]),
    "fl_darcy_batch_set_narrow": (C.c_int, [C.c_void_p, C.c_int]),
    "fl_darcy_batch_set_coarse": (C.c_int, [C.c_void_p, C.c_int]),
    "fl_darcy_batch_set_coarse_basis": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fl_darcy_batch_aggregates": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fl_darcy_batch_lambda": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fl_coarse_invert": (C.c_int, [_f64p, C.c_int, C.c_int, C.c_void_p, _i32p]),
    "fl_gnn_create": (C.c_int, [C.c_void_p, C.c_uint, C.POINTER(C.c_void_p)]),
    "fl_gnn_free": (None, [C.c_void_p]),
    "fl_gnn_n_params": (C.c_int, [C.c_void_p]),
    "fl_gnn_get_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fl_gnn_set_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fl_gnn_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fl_gnn_loss_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]),
    "fl_gnn_zero_grad": (C.c_int, [C.c_void_p]),
    "fl_gnn_get_grad": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fl_gnn_adam_step": (C.c_int, [C.c_void_p, C.c_double]),
    "fl_gnn_train": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double,
                               C.c_void_p]),
    "fl_gnn_infer": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "fl_gnn_features": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fl_transport_create": (C.c_int, [C.c_void_p, C.c_int, _f64p, _f64p, _f64p, _f64p, C.c_double, C.c_double,
                                      C.POINTER(C.c_void_p)]),
    "fl_transport_free": (None, [C.c_void_p]),
    "fl_transport_set_velocity": (C.c_int, [C.c_void_p, _f64p]),
    "fl_transport_set_velocity_from_darcy": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fl_transport_prepare": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "fl_transport_update_solution": (C.c_int, [C.c_void_p, C.c_double, C.POINTER(C.c_int)]),
    "fl_transport_get_conc": (C.c_int, [C.c_void_p, _f64p]),
    "fl_transport_time": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]),
    "fl_transport_mass_balance": (C.c_int, [C.c_void_p] + [_f64p] * 6),
    "fl_transport_set_sources": (C.c_int, [C.c_void_p, _f64p, _f64p, _f64p]),
    "fl_transport_set_bc_conc": (C.c_int, [C.c_void_p, _f64p]),
    "fl_transport_set_pade_reactions": (C.c_int, [C.c_void_p, C.c_int, _i32p, _i32p, _i32p, _f64p, _f64p, C.c_int,
                                                  C.c_int]),
    "fl_transport_set_linear_reactions": (C.c_int, [C.c_void_p, C.c_int, _i32p, _i32p, _i32p, _f64p, _f64p]),
    "fl_transport_nnz": (C.c_int, [C.c_void_p]),
    "fl_transport_matrix": (C.c_int, [C.c_void_p, _i32p, _i32p, _f64p]),
    "fl_transport_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
}


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise FlowError("libflow123d_gpu.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


# status codes (include/flow123d_gpu.h)
NOT_CONVERGED = 1
E_SINGULAR_ELEMENT = -10
E_BAD_FIELD = -11
E_DIVERGED = -12
E_CALLBACK = -13
E_SORPTION = -14


def _check(rc):
    """Raises on a negative status; a positive one (a solver stopped at its
    iteration limit, the result still valid) is returned to the caller."""
    if rc < 0:
        err = FlowError("flow123d_gpu error %d: %s" % (rc, lib().fl_last_error().decode()))
        err.status = rc
        raise err
    return rc


def stream_bandwidth(nbytes=1 << 30, reps=10):
    """Achievable HBM bandwidth (GB/s) of the current device: copy (read + write
    counted), read only and write only over buffers of nbytes each."""
    g = np.zeros(3)
    _check(lib().fl_device_stream_bandwidth(int(nbytes), int(reps), g))
    return {"copy": float(g[0]), "read": float(g[1]), "write": float(g[2])}


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    if n is not None and a.size != n:
        raise ValueError("expected %d values, got %d" % (n, a.size))
    return a


class Mesh:
    """Mesh + topology (Mesh, msh_gmshreader.cc / mesh.cc) built by the library's host module."""

    def __init__(self, path=None, arrays=None):
        L = lib()
        h = C.c_void_p()
        if path is not None:
            _check(L.fl_mesh_read_gmsh(os.fsencode(path), C.byref(h)))
        else:
            a = arrays
            xyz = _f64(a["nodes"])
            en = np.ascontiguousarray(a["elem_nodes"], np.int32).reshape(-1)
            et = np.ascontiguousarray(a["elem_type"], np.int32)
            er = np.ascontiguousarray(a["elem_region"], np.int32)
            phys = a.get("physical", [])
            pid = (C.c_int * max(len(phys), 1))(*[p[1] for p in phys])
            pnm = (C.c_char_p * max(len(phys), 1))(*[p[2].encode() for p in phys])
            _check(L.fl_mesh_from_arrays(len(xyz) // 3, xyz, len(et), None, et, er, en, len(phys),
                                         C.cast(pid, C.c_void_p), C.cast(pnm, C.c_void_p), C.byref(h)))
        self.h = h
        c = np.zeros(9, np.int32)
        L.fl_mesh_counts(h, c)
        (self.n_nodes, self.n_elements, self.n_bc_elements, self.n_sides, self.n_edges, self.n_boundaries,
         self.n_vb, self.n_regions, self.n_file_bc) = [int(v) for v in c]
        E, F, S, V, B, NB = self.n_elements, self.n_edges, self.n_sides, self.n_vb, self.n_boundaries, \
            self.n_bc_elements
        self.elem_dim = np.zeros(E, np.int32)
        self.elem_id = np.zeros(E, np.int32)
        self.elem_region = np.zeros(E, np.int32)
        self.elem_nodes = np.zeros(4 * E, np.int32)
        self.elem_edge = np.zeros(4 * E, np.int32)
        self.elem_bnd = np.zeros(4 * E, np.int32)
        self.edge_ptr = np.zeros(F + 1, np.int32)
        self.edge_side = np.zeros(max(S, 1), np.int32)
        vb_elem = np.zeros(max(V, 1), np.int32)
        vb_edge = np.zeros(max(V, 1), np.int32)
        bnd_edge = np.zeros(max(B, 1), np.int32)
        bnd_bcele = np.zeros(max(B, 1), np.int32)
        bc_region = np.zeros(max(NB, 1), np.int32)
        _check(L.fl_mesh_topology(h, self.elem_dim, self.elem_id, self.elem_region, self.elem_nodes,
                                  self.elem_edge, self.elem_bnd, self.edge_ptr, self.edge_side, vb_elem, vb_edge,
                                  bnd_edge, bnd_bcele, bc_region))
        self.vb_elem, self.vb_edge = vb_elem[:V], vb_edge[:V]
        self.bnd_edge, self.bnd_bcele, self.bc_region = bnd_edge[:B], bnd_bcele[:B], bc_region[:NB]
        R = self.n_regions
        ids, isb, sub = np.zeros(R, np.int32), np.zeros(R, np.int32), np.zeros(R, np.int32)
        buf = C.create_string_buffer(1 << 16)
        _check(L.fl_mesh_regions(h, ids, isb, sub, buf, len(buf)))
        labels = buf.value.decode().split("\n")[:R]
        self.regions = [dict(id=int(i), boundary=bool(b), sub_idx=int(s), label=l)
                        for i, b, s, l in zip(ids, isb, sub, labels)]
        self.side_off = np.concatenate([[0], np.cumsum(self.elem_dim + 1)]).astype(np.int64)

    @property
    def n_boundary_regions(self):
        return sum(1 for r in self.regions if r["boundary"])

    @property
    def n_bulk_regions(self):
        return sum(1 for r in self.regions if not r["boundary"])

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.fl_mesh_free(self.h)
            self.h = None


def coarse_invert(a):
    """FP32 inverses of a stack of SPD matrices a[np][n][n] by the batched
    Cholesky + inverse kernels of the two-level preconditioner's coarse space;
    returns (inv[np][n][n] float32, info[np]: 1 where not positive definite)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim == 2:
        a = a[None]
    npm, n = a.shape[0], a.shape[1]
    inv = np.zeros((npm, n, n), np.float32)
    info = np.zeros(npm, np.int32)
    _check(lib().fl_coarse_invert(a, n, npm, inv.ctypes.data_as(C.c_void_p), info))
    return inv, info


def element_ids(mesh, boundary=False):
    """Mesh::elements_id_maps (mesh.cc:781-812): gmsh ids of the bulk elements, or of
    the boundary elements with -1 for the implicit ones."""
    n = mesh.n_bc_elements if boundary else mesh.n_elements
    out = np.zeros(max(n, 1), np.int32)
    _check(lib().fl_mesh_element_ids(mesh.h, 1 if boundary else 0, out))
    return out[:n]


class FieldElementwise:
    """Element-wise field read from gmsh `$ElementData` sections
    (FieldElementwise, field_elementwise_impl.hh:67-156, over
    GmshMeshReader::read_element_data, msh_gmshreader.cc:306-397).

    ``set_time(t)`` applies every section of the field with time <= t not yet
    applied and returns whether the values changed; ``value()`` returns them,
    shape (n_entities,) for one component, else (n_entities, n_comp)."""

    def __init__(self, mesh, gmsh_file, field_name, n_comp=1, boundary=False):
        self.mesh = mesh
        self.n_comp = int(n_comp)
        h = C.c_void_p()
        _check(lib().fl_element_data_open(mesh.h, os.fsencode(gmsh_file), field_name.encode(), self.n_comp,
                                          1 if boundary else 0, C.byref(h)))
        self.h = h
        self.n_entities = mesh.n_bc_elements if boundary else mesh.n_elements

    def set_time(self, t):
        u = C.c_int()
        _check(lib().fl_element_data_set_time(self.h, float(t), C.byref(u)))
        return bool(u.value)

    def value(self):
        out = np.zeros(max(self.n_entities * self.n_comp, 1))
        _check(lib().fl_element_data_values(self.h, out))
        out = out[:self.n_entities * self.n_comp]
        return out if self.n_comp == 1 else out.reshape(self.n_entities, self.n_comp)

    def info(self):
        """(rows stored by the last section applied, warnings of the last set_time)."""
        i = np.zeros(3, np.int32)
        buf = C.create_string_buffer(1 << 16)
        _check(lib().fl_element_data_info(self.h, i, buf, len(buf)))
        return int(i[2]), [w for w in buf.value.decode().split("\n") if w]

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.fl_element_data_free(self.h)
            self.h = None


class OldBcdInput:
    """Legacy boundary files (OldBcdInput, old_bcd.cc:44-250).  ``read_flow(path)``
    returns per boundary element (bc_type, bc_pressure, bc_flux, bc_sigma);
    ``read_transport(path, n_subst)`` returns bc_conc[n_subst][NBC]."""

    def __init__(self, mesh):
        self.mesh, self.h = mesh, None

    def read_flow(self, path):
        NB = self.mesh.n_bc_elements
        t = np.zeros(max(NB, 1), np.int32)
        p, f, s = np.zeros(max(NB, 1)), np.zeros(max(NB, 1)), np.zeros(max(NB, 1))
        h = C.c_void_p()
        _check(lib().fl_old_bcd_read_flow(self.mesh.h, os.fsencode(path), t, p, f, s, C.byref(h)))
        self._free()
        self.h = h
        return t[:NB], p[:NB], f[:NB], s[:NB]

    def read_transport(self, path, n_subst=1):
        if self.h is None:
            raise FlowError("read_flow must come first: transport rows refer to the flow file's rows")
        NB = self.mesh.n_bc_elements
        c = np.zeros(max(n_subst * NB, 1))
        _check(lib().fl_old_bcd_read_transport(self.h, os.fsencode(path), int(n_subst), c))
        return c[:n_subst * NB].reshape(n_subst, NB)

    def _free(self):
        if self.h and _LIB is not None:
            _LIB.fl_old_bcd_free(self.h)
        self.h = None

    def __del__(self):
        self._free()


def _names(names):
    arr = (C.c_char_p * len(names))(*[n.encode() for n in names])
    return arr, C.cast(arr, C.c_void_p)


def write_water_balance(mesh, path, time, bal, plus, minus, src):
    """One DarcyFlowMHOutput::water_balance block (darcy_flow_mh_output.cc:671-773) from
    per-region arrays (boundary regions, then bulk regions, in the mesh's region order)."""
    nb, nk = max(mesh.n_boundary_regions, 1), max(mesh.n_bulk_regions, 1)
    pad = lambda a, n: np.resize(np.ascontiguousarray(a, np.float64).reshape(-1), n) if np.size(a) else np.zeros(n)
    _check(lib().fl_write_water_balance(mesh.h, os.fsencode(path), float(time), pad(bal, nb), pad(plus, nb),
                                        pad(minus, nb), pad(src, nk)))


def write_mass_balance(mesh, path, time, initial_time, names, mb):
    """One MassBalance::output block (mass_balance.cc:242-354) from a mass_balance()
    dictionary; names = substance names."""
    ns = len(names)
    keep, ptr = _names(names)
    f = lambda k, n: np.resize(np.ascontiguousarray(mb[k], np.float64).reshape(-1), max(n, 1))
    nb, nk = ns * mesh.n_boundary_regions, ns * mesh.n_bulk_regions
    _check(lib().fl_write_mass_balance(mesh.h, os.fsencode(path), float(time), float(initial_time), ns, ptr,
                                       f("bal", nb), f("plus", nb), f("minus", nb), f("mass", nk), f("src", nk),
                                       f("totals", 4 * ns)))


class OutputMSH:
    """gmsh 2.x ASCII output of per-element results (OutputMSH, output_msh.cc:53-98,
    100-188, 383-467): the mesh head once, then one `$ElementData` section per field
    and output time.  ``precision`` = significant digits (the reference's 6 by default)."""

    def __init__(self, mesh, path, precision=6):
        self.mesh, self.path, self.precision = mesh, os.fsencode(path), int(precision)
        self.header_written = False
        self.current_step = 0

    def write_head(self):
        _check(lib().fl_msh_write_head(self.mesh.h, self.path))
        self.header_written = True

    def write_elem_data(self, name, units, values, time, step):
        v = np.ascontiguousarray(values, np.float64)
        n = v.shape[0]
        ncomp = 1 if v.ndim == 1 else int(np.prod(v.shape[1:]))
        _check(lib().fl_msh_write_element_data(self.path, name.encode(), units.encode(), float(time), int(step),
                                               ncomp, n, v.reshape(-1) if n else np.zeros(1), self.precision))

    def write_node_data(self, name, units, values, time, step):
        v = np.ascontiguousarray(values, np.float64)
        n = v.shape[0]
        ncomp = 1 if v.ndim == 1 else int(np.prod(v.shape[1:]))
        _check(lib().fl_msh_write_node_data(self.path, name.encode(), units.encode(), float(time), int(step),
                                            ncomp, n, v.reshape(-1) if n else np.zeros(1), self.precision))

    def write_data(self, time, fields):
        """OutputMSH::write_data(time) (output_msh.cc:440-462): `fields` is a list of
        (name, units, values[E] or values[E, n_comp]) in registration order."""
        if not self.header_written:
            self.write_head()
        for name, units, v in fields:
            self.write_elem_data(name, units, v, time, self.current_step)
        self.current_step += 1


class _TimeCoupling:
    """Equation time interface (equation.hh:113-146) over the handle's
    TimeGovernor: time() -> (t, dt, is_end), planned_time, solved_time,
    choose_next_time, set_time_upper_constraint."""
    _pfx = None

    def time(self):
        t, dt, e = C.c_double(), C.c_double(), C.c_int()
        _check(getattr(lib(), self._pfx + ("time_governor" if self._pfx == "fl_transport_" else "time"))(
            self.h, C.byref(t), C.byref(dt), C.byref(e)))
        return t.value, dt.value, bool(e.value)

    def planned_time(self):
        t = C.c_double()
        _check(getattr(lib(), self._pfx + "planned_time")(self.h, C.byref(t)))
        return t.value

    def solved_time(self):
        t = C.c_double()
        _check(getattr(lib(), self._pfx + "solved_time")(self.h, C.byref(t)))
        return t.value

    def choose_next_time(self):
        _check(getattr(lib(), self._pfx + "choose_next_time")(self.h))

    def set_time_upper_constraint(self, dt):
        _check(getattr(lib(), self._pfx + "set_time_upper_constraint")(self.h, float(dt)))


class DarcyFlowMH_Steady(_TimeCoupling):
    """GPU replacement of DarcyFlowMH_Steady (darcy_flow_mh.cc:249-505)."""
    _pfx = "fl_darcy_"

    def __init__(self, mesh):
        self.mesh = mesh
        h = C.c_void_p()
        _check(lib().fl_darcy_create(mesh.h, C.byref(h)))
        self.h = h
        self.S, self.E, self.F = mesh.n_sides, mesh.n_elements, mesh.n_edges

    def set_data(self, K, cs, sigma, f, bc_type, bc_pressure, bc_flux, bc_sigma, aniso=None):
        E, NB = self.E, self.mesh.n_bc_elements
        self._keep = [_f64(K, E), _f64(cs, E), _f64(sigma, E), _f64(f, E), _f64(bc_pressure, NB),
                      _f64(bc_flux, NB), _f64(bc_sigma, NB)]
        bct = np.ascontiguousarray(bc_type, np.int32)
        an = None if aniso is None else _f64(aniso, 9 * E)
        self._an = an
        K, cs, sigma, f, bp, bf, bs = self._keep
        _check(lib().fl_darcy_set_data(self.h, K, cs, None if an is None else an.ctypes.data, sigma, f, bct, bp,
                                       bf, bs))

    def set_conductivity(self, K):
        _check(lib().fl_darcy_set_conductivity(self.h, _f64(K, self.E)))

    def update_solution(self, r_tol=1e-7, a_tol=1e-9, max_it=10000, precision=0, norm=0):
        """precision 0: FP64 CG; 1: FP32 CG with FP64 reliable updates.
        norm 0: stop test on ||r||_2; 1: on the natural Jacobi norm sqrt(r^T D^-1 r)."""
        o = SolverOpts(r_tol, a_tol, max_it, precision, norm)
        info = SolveInfo()
        st = _check(lib().fl_darcy_update_solution(self.h, C.byref(o), C.byref(info)))
        self.info = info.as_dict()
        self.info["status"] = st
        return self.info

    def set_unsteady(self, lumped, storativity, init_pressure):
        """Unsteady MH (lumped False) or LMH (lumped True) from the data already set;
        the solution becomes the initial state (darcy_flow_mh.cc:1601-1749)."""
        E = self.mesh.n_elements
        self._unst = [_f64(storativity, E), _f64(init_pressure, E)]
        _check(lib().fl_darcy_set_unsteady(self.h, 1 if lumped else 0, *self._unst))

    def set_time(self, end_time, init_dt, save_step=0.0):
        """Unsteady governor with a fixed step and output marks every save_step
        (after set_unsteady): every update_solution is then one step of it."""
        _check(lib().fl_darcy_set_time(self.h, float(end_time), float(init_dt), float(save_step)))

    def unsteady_step(self, dt, r_tol=1e-7, a_tol=1e-9, max_it=10000, precision=0, norm=0):
        """One implicit Euler step of length dt (darcy_flow_mh.cc:375-433, 1660-1827)."""
        o = SolverOpts(r_tol, a_tol, max_it, precision, norm)
        info = SolveInfo()
        st = _check(lib().fl_darcy_unsteady_step(self.h, float(dt), C.byref(o), C.byref(info)))
        self.info = info.as_dict()
        self.info["status"] = st
        return self.info

    def get_solution_vector(self):
        out = np.zeros(self.S + self.E + self.F)
        _check(lib().fl_darcy_get_solution(self.h, out))
        return out

    def solution_precision(self, r_tol, a_tol):
        return lib().fl_darcy_solution_precision(self.h, r_tol, a_tol)

    def element_balance(self):
        out = np.zeros(self.E)
        _check(lib().fl_darcy_element_balance(self.h, out))
        return out

    def water_balance(self):
        n = max(self.mesh.n_boundary_regions, 1)
        b, p, m = np.zeros(n), np.zeros(n), np.zeros(n)
        _check(lib().fl_darcy_water_balance(self.h, b, p, m))
        k = self.mesh.n_boundary_regions
        return b[:k], p[:k], m[:k]

    def element_output(self):
        """(pressure[E], piezo_head[E], velocity[E, 3]) of the current solution
        (DarcyFlowMHOutput::make_element_scalar / make_element_vector)."""
        E = self.mesh.n_elements
        p, h, v = np.zeros(E), np.zeros(E), np.zeros(3 * E)
        _check(lib().fl_darcy_element_output(self.h, p, h, v))
        return p, h, v.reshape(E, 3)

    def node_pressure(self):
        """Pressure interpolated to the nodes (DarcyFlowMHOutput::make_node_scalar_param)."""
        out = np.zeros(max(self.mesh.n_nodes, 1))
        _check(lib().fl_darcy_node_pressure(self.h, out))
        return out[:self.mesh.n_nodes]

    def write_water_balance(self, path, time=float("inf")):
        """DarcyFlowMHOutput::water_balance block for this solution (steady: time inf)."""
        _check(lib().fl_darcy_write_water_balance(self.h, os.fsencode(path), float(time)))

    def edge_system(self):
        nnz = lib().fl_darcy_edge_nnz(self.h)
        ptr, col, val, rhs = np.zeros(self.F + 1, np.int32), np.zeros(nnz, np.int32), np.zeros(nnz), \
            np.zeros(self.F)
        _check(lib().fl_darcy_edge_system(self.h, ptr, col, val, rhs))
        return ptr, col, val, rhs

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.fl_darcy_free(self.h)
            self.h = None


class DarcyBatch:
    """n_real conductivity realizations of one DarcyFlowMH_Steady problem solved together (Monte Carlo).

    Shares mesh, cross-sections, sources and boundary data with `darcy` (set_data must have been called)."""

    def __init__(self, darcy, n_real, slots=0):
        self.darcy = darcy
        self.n_real = int(n_real)
        h = C.c_void_p()
        _check(lib().fl_darcy_batch_create_slots(darcy.h, self.n_real, int(slots), C.byref(h)))
        self.h = h
        self.S, self.E, self.F = darcy.S, darcy.E, darcy.F

    def set_conductivity(self, K):
        """K: [n_real, E] host array, or a device pointer (int) to n_real*E doubles."""
        if isinstance(K, int):
            _check(lib().fl_darcy_batch_set_conductivity(self.h, C.c_void_p(K), 1))
        else:
            Kh = _f64(K, self.n_real * self.E)
            _check(lib().fl_darcy_batch_set_conductivity(self.h, Kh.ctypes.data, 0))

    def update_solution(self, r_tol=1e-7, a_tol=1e-9, max_it=10000, precision=0, norm=0):
        """norm as DarcyFlowMH_Steady.update_solution"""
        o = SolverOpts(r_tol, a_tol, max_it, precision, norm)
        info = BatchInfo()
        it = np.zeros(self.n_real, np.int32)
        conv = np.zeros(self.n_real, np.int32)
        res = np.zeros(self.n_real)
        st = _check(lib().fl_darcy_batch_update_solution(self.h, C.byref(o), C.byref(info), it, conv, res))
        self.info = info.as_dict()
        self.info.update(iterations=it, converged=conv, residual=res, status=st)
        return self.info

    def get_solution_vector(self, r):
        out = np.zeros(self.S + self.E + self.F)
        _check(lib().fl_darcy_batch_get_solution(self.h, int(r), out))
        return out

    def diagnostics(self):
        st, ru = np.zeros(self.n_real, np.int32), np.zeros(self.n_real, np.int32)
        _check(lib().fl_darcy_batch_diagnostics(self.h, st, ru))
        return st, ru

    def device_solution(self):
        n = C.c_int()
        p = lib().fl_darcy_batch_device_solution(self.h, C.byref(n))
        return p, n.value

    def set_launch(self, grid, xcd_map, chunk):
        _check(lib().fl_darcy_batch_set_launch(self.h, grid, xcd_map, chunk))

    def autotune(self):
        _check(lib().fl_darcy_batch_autotune(self.h))

    def set_handover(self, n):
        _check(lib().fl_darcy_batch_set_handover(self.h, int(n)))

    def set_narrow(self, on):
        """Two-level tail in the batch with 16 / 4 lanes per row (default) instead of the hand-over."""
        _check(lib().fl_darcy_batch_set_narrow(self.h, 1 if on else 0))

    def set_coarse(self, ncol):
        """Two-level preconditioner: 0 Jacobi, 1 constant per aggregate, 2 constant + basis phi."""
        _check(lib().fl_darcy_batch_set_coarse(self.h, int(ncol)))

    def set_coarse_basis(self, phi_ptr):
        """phi: device pointer (int) to n_real x F float32 values in edge order."""
        _check(lib().fl_darcy_batch_set_coarse_basis(self.h, C.c_void_p(phi_ptr)))

    def lambda_device(self, out_ptr):
        """Edge traces of all realizations into a device array [n_real, F] (float64, edge order)."""
        _check(lib().fl_darcy_batch_lambda(self.h, C.c_void_p(out_ptr)))

    def aggregates(self):
        out = np.zeros(self.F, np.int32)
        n = lib().fl_darcy_batch_aggregates(self.h, out.ctypes.data)
        if n < 0:
            _check(n)
        return n, out

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.fl_darcy_batch_free(self.h)
            self.h = None


class LearnedCoarseBasis:
    """Graph network producing the second coarse column phi of a DarcyBatch's
    two-level preconditioner (trained on the GPU on solver-generated pairs).
    Device arrays are passed as integer pointers (e.g. torch tensor.data_ptr())."""

    def __init__(self, batch, seed=0):
        self.batch = batch
        h = C.c_void_p()
        _check(lib().fl_gnn_create(batch.h, int(seed), C.byref(h)))
        self.h = h
        self.n_params = lib().fl_gnn_n_params(h)
        self.E, self.F = batch.E, batch.F

    def get_params(self):
        out = np.zeros(self.n_params, np.float32)
        _check(lib().fl_gnn_get_params(self.h, out.ctypes.data))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check(lib().fl_gnn_set_params(self.h, p.ctypes.data))

    def forward(self, K_ptr, phi_ptr=0):
        _check(lib().fl_gnn_forward(self.h, C.c_void_p(K_ptr), C.c_void_p(phi_ptr or None)))

    def loss_grad(self, lam_ptr):
        L = C.c_double()
        _check(lib().fl_gnn_loss_grad(self.h, C.c_void_p(lam_ptr), C.byref(L)))
        return L.value

    def zero_grad(self):
        _check(lib().fl_gnn_zero_grad(self.h))

    def get_grad(self):
        out = np.zeros(self.n_params, np.float32)
        _check(lib().fl_gnn_get_grad(self.h, out.ctypes.data))
        return out

    def adam_step(self, lr):
        _check(lib().fl_gnn_adam_step(self.h, float(lr)))

    def train(self, K_ptr, lam_ptr, n, epochs, batch, lr):
        losses = np.zeros(epochs)
        _check(lib().fl_gnn_train(self.h, C.c_void_p(K_ptr), C.c_void_p(lam_ptr), int(n), int(epochs), int(batch),
                                  float(lr), losses.ctypes.data))
        return losses

    def infer(self, K_ptr, R, phi_ptr):
        _check(lib().fl_gnn_infer(self.h, C.c_void_p(K_ptr), int(R), C.c_void_p(phi_ptr)))

    def features(self):
        xe = np.zeros((self.E, 16), np.float32)
        xf = np.zeros((self.F, 16), np.float32)
        _check(lib().fl_gnn_features(self.h, xe.ctypes.data, xf.ctypes.data))
        return xe, xf

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.fl_gnn_free(self.h)
            self.h = None


BC_CONC_FN = C.CFUNCTYPE(C.c_int, C.c_double, C.POINTER(C.c_double), C.c_void_p)


class ConvectionTransport(_TimeCoupling):
    """GPU replacement of ConvectionTransport + TransportOperatorSplitting (transport.cc)."""
    _pfx = "fl_transport_"

    def __init__(self, mesh, por, cs, bc_conc, init_conc, end_time, save_step):
        self.mesh = mesh
        NB, E = mesh.n_bc_elements, mesh.n_elements
        bcc = np.ascontiguousarray(bc_conc, np.float64).reshape(-1, NB)
        self.nsub = bcc.shape[0]
        self._keep = [_f64(por, E), _f64(cs, E), _f64(bcc), _f64(init_conc, self.nsub * E)]
        h = C.c_void_p()
        _check(lib().fl_transport_create(mesh.h, self.nsub, self._keep[0], self._keep[1], self._keep[2],
                                         self._keep[3], end_time, save_step, C.byref(h)))
        self.h = h

    def set_velocity_field(self, side_flux=None, darcy=None):
        if darcy is not None:
            _check(lib().fl_transport_set_velocity_from_darcy(self.h, darcy.h))
        else:
            _check(lib().fl_transport_set_velocity(self.h, _f64(side_flux, self.mesh.n_sides)))

    def prepare(self):
        """Builds the transport matrix now and returns the CFL time step."""
        c = C.c_double()
        _check(lib().fl_transport_prepare(self.h, C.byref(c)))
        return c.value

    def update_solution(self, t=None):
        """Convection up to target time t; without t the operator-splitting
        step of its own governor (TransportOperatorSplitting::update_solution)."""
        s = C.c_int()
        self._bc_exc = None
        try:
            if t is None:
                _check(lib().fl_transport_update_solution_step(self.h, C.byref(s)))
            else:
                _check(lib().fl_transport_update_solution(self.h, t, C.byref(s)))
        except FlowError as err:
            if self._bc_exc is not None:
                raise self._bc_exc from err
            raise
        return s.value

    def times(self):
        t, dt, cfl = C.c_double(), C.c_double(), C.c_double()
        lib().fl_transport_time(self.h, C.byref(t), C.byref(dt), C.byref(cfl))
        return t.value, dt.value, cfl.value

    def conc(self):
        out = np.zeros(self.nsub * self.mesh.n_elements)
        _check(lib().fl_transport_get_conc(self.h, out))
        return out.reshape(self.nsub, -1)

    def matrix(self):
        nnz = lib().fl_transport_nnz(self.h)
        ptr, col, val = np.zeros(self.mesh.n_elements + 1, np.int32), np.zeros(nnz, np.int32), np.zeros(nnz)
        _check(lib().fl_transport_matrix(self.h, ptr, col, val))
        return ptr, col, val

    def mass_balance(self):
        ns, nbr, nblk = self.nsub, self.mesh.n_boundary_regions, self.mesh.n_bulk_regions
        bal, plus, minus = np.zeros(ns * max(nbr, 1)), np.zeros(ns * max(nbr, 1)), np.zeros(ns * max(nbr, 1))
        mass, src, tot = np.zeros(ns * max(nblk, 1)), np.zeros(ns * max(nblk, 1)), np.zeros(4 * ns)
        _check(lib().fl_transport_mass_balance(self.h, bal, plus, minus, mass, src, tot))
        return dict(bal=bal[:ns * nbr].reshape(ns, nbr), plus=plus[:ns * nbr].reshape(ns, nbr),
                    minus=minus[:ns * nbr].reshape(ns, nbr), mass=mass[:ns * nblk].reshape(ns, nblk),
                    src=src[:ns * nblk].reshape(ns, nblk), totals=tot.reshape(ns, 4))

    def set_bc_conc(self, bc_conc):
        """New boundary concentrations [n_subst][NBC] from the next step on (transport.cc:598-627)."""
        self._bcc_next = _f64(bc_conc, self.nsub * self.mesh.n_bc_elements)
        _check(lib().fl_transport_set_bc_conc(self.h, self._bcc_next))

    def set_pade_reactions(self, decays, names=None, nom_pol_deg=2, den_pol_deg=2):
        """Pade_approximant (pade_approximant.cc:97-290) with the decays of set_linear_reactions."""
        self._decay_arrays(decays, names)
        _check(lib().fl_transport_set_pade_reactions(self.h, self._n_decays, *self._react, int(nom_pol_deg),
                                                     int(den_pol_deg)))

    def set_linear_reactions(self, decays, names=None):
        """Linear_reaction decays (linear_reaction.cc:200-258): a list of dicts with
        parent, products, optional branch_ratios (default one product, ratio 1.0) and
        half_life or kinetic (half_life = ln 2 / kinetic).  Substances are given by
        index or, with `names`, by name."""
        self._decay_arrays(decays, names)
        _check(lib().fl_transport_set_linear_reactions(self.h, self._n_decays, *self._react))

    def _decay_arrays(self, decays, names):
        idx = (lambda x: names.index(x) if isinstance(x, str) else int(x))
        parent, pptr, prod, hl, br = [], [0], [], [], []
        for d in decays:
            parent.append(idx(d["parent"]))
            ps = [idx(p) for p in d["products"]]
            ratios = d.get("branch_ratios", [1.0])
            if len(ratios) != len(ps):
                raise FlowError("Number of branches %d has to match number of products %d" % (len(ratios), len(ps)))
            prod += ps
            br += list(ratios)
            pptr.append(len(prod))
            hl.append(d["half_life"] if "half_life" in d else math.log(2) / d["kinetic"])
        self._n_decays = len(parent)
        self._react = [np.array(parent or [0], np.int32), np.array(pptr, np.int32), np.array(prod or [0], np.int32),
                       np.array(hl or [1.0], np.float64), np.array(br or [0.0], np.float64)]

    def write_mass_balance(self, path, names):
        """MassBalance::output block at the current time (mass_balance.cc:242-354)."""
        keep, ptr = _names(names)
        _check(lib().fl_transport_write_mass_balance(self.h, os.fsencode(path), ptr))

    def set_sources(self, density, sigma, conc):
        """sources_density, sources_sigma, sources_conc, each [n_subst][E] (transport.cc:501-544)."""
        n = self.nsub * self.mesh.n_elements
        self._src = [_f64(density, n), _f64(sigma, n), _f64(conc, n)]
        _check(lib().fl_transport_set_sources(self.h, *self._src))

    def set_dual_porosity_sorption(self, dual_porosity, sorption, por_imm, alpha=None, phi=None, sorp_type=None,
                                   sorp_coef0=None, sorp_coef1=None):
        """dual_porosity / sorption_enable (transport.cc:1026-1160): por_imm, phi [E];
        alpha, sorp_type (0 none, 1 linear, 2 freundlich, 3 langmuir), sorp_coef0/1
        [n_subst][E]; scalars and per-substance columns broadcast."""
        E, ns = self.mesh.n_elements, self.nsub

        def arr(a, dt, shape):
            return None if a is None else np.ascontiguousarray(np.broadcast_to(a, shape), dt).reshape(-1)

        self._phases = [arr(por_imm, np.float64, (E,)), arr(alpha, np.float64, (ns, E)), arr(phi, np.float64, (E,)),
                        arr(sorp_type, np.int32, (ns, E)), arr(sorp_coef0, np.float64, (ns, E)),
                        arr(sorp_coef1, np.float64, (ns, E))]
        ptr = [None if a is None else a.ctypes.data for a in self._phases[1:]]
        _check(lib().fl_transport_set_dual_porosity_sorption(self.h, int(bool(dual_porosity)), int(bool(sorption)),
                                                             self._phases[0], *ptr))

    def set_bc_conc_function(self, fn):
        """Boundary concentrations fn(t) -> [n_subst][NBC] at every step's start time
        (a time-dependent bc_conc formula); None: keep the last data fixed."""
        shape = (self.nsub, self.mesh.n_bc_elements)
        if fn is None:
            self._bc_fn = None
            _check(lib().fl_transport_set_bc_conc_function(self.h, None, None))
            return

        def cb(t, out, user):
            # an exception cannot cross the C frames: keep it, stop the update
            # (non-zero return) and re-raise it when the native call returns
            try:
                np.ctypeslib.as_array(out, shape=shape)[:] = fn(t)
                return 0
            except BaseException as e:  # noqa: B902
                self._bc_exc = e
                return 1

        self._bc_exc = None
        self._bc_fn = BC_CONC_FN(cb)
        _check(lib().fl_transport_set_bc_conc_function(self.h, C.cast(self._bc_fn, C.c_void_p), None))

    def set_reaction_sorption(self, subst, types, mult, second, rock_density, molar_mass, solubility=None,
                              solvent_dens=1.0, substeps=1000):
        """The reaction module's sorption: species on substances subst[n]; types
        (0 none, 1 linear, 2 freundlich, 3 langmuir), mult, second [n_bulk_regions][n];
        rock_density [n_bulk_regions]; molar_mass, solubility [n]."""
        n = len(subst)
        nreg = self.mesh.n_bulk_regions
        self._rs = [np.ascontiguousarray(subst, np.int32),
                    np.ascontiguousarray(np.broadcast_to(types, (nreg, n)), np.int32).reshape(-1),
                    np.ascontiguousarray(np.broadcast_to(mult, (nreg, n)), np.float64).reshape(-1),
                    np.ascontiguousarray(np.broadcast_to(second, (nreg, n)), np.float64).reshape(-1),
                    np.ascontiguousarray(np.broadcast_to(rock_density, (nreg,)), np.float64),
                    np.ascontiguousarray(np.broadcast_to(molar_mass, (n,)), np.float64),
                    None if solubility is None else np.ascontiguousarray(np.broadcast_to(solubility, (n,)), np.float64)]
        sol = None if self._rs[6] is None else self._rs[6].ctypes.data
        _check(lib().fl_transport_set_reaction_sorption(self.h, n, *self._rs[:6], sol, float(solvent_dens),
                                                        int(substeps)))

    def set_reaction_sorption_elementwise(self, subst, types, mult, second, rock_density, molar_mass,
                                          solubility=None, solvent_dens=1.0):
        """The same with element-wise parameter fields: types, mult, second [n][E];
        rock_density [E]; every element's isotherm is solved directly."""
        n, E = len(subst), self.mesh.n_elements
        self._rs = [np.ascontiguousarray(subst, np.int32),
                    np.ascontiguousarray(np.broadcast_to(types, (n, E)), np.int32).reshape(-1),
                    np.ascontiguousarray(np.broadcast_to(mult, (n, E)), np.float64).reshape(-1),
                    np.ascontiguousarray(np.broadcast_to(second, (n, E)), np.float64).reshape(-1),
                    np.ascontiguousarray(np.broadcast_to(rock_density, (E,)), np.float64),
                    np.ascontiguousarray(np.broadcast_to(molar_mass, (n,)), np.float64),
                    None if solubility is None else np.ascontiguousarray(np.broadcast_to(solubility, (n,)), np.float64)]
        sol = None if self._rs[6] is None else self._rs[6].ctypes.data
        _check(lib().fl_transport_set_reaction_sorption_elementwise(self.h, n, *self._rs[:6], sol,
                                                                    float(solvent_dens)))

    def reaction_sorbed(self):
        out = np.zeros(len(self._rs[0]) * self.mesh.n_elements)
        _check(lib().fl_transport_get_reaction_sorbed(self.h, out))
        return out.reshape(len(self._rs[0]), -1)

    def conc_phase(self, phase):
        """get_concentration_matrix()[phase]: 0 mobile, 1 immobile, 2 mobile sorbed, 3 immobile sorbed."""
        out = np.zeros(self.nsub * self.mesh.n_elements)
        _check(lib().fl_transport_get_conc_phase(self.h, int(phase), out))
        return out.reshape(self.nsub, -1)

    def stats(self):
        ms, st = C.c_double(), C.c_int()
        lib().fl_transport_stats(self.h, C.byref(ms), C.byref(st))
        return ms.value, st.value

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.fl_transport_free(self.h)
            self.h = None
