"""ctypes binding of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline,
never by the product path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(path)
        L.orc_last_error.restype = C.c_char_p
        L.orc_mesh_load.restype = C.c_void_p
        L.orc_mesh_load.argtypes = [C.c_char_p]
        L.orc_mesh_free.argtypes = [C.c_void_p]
        L.orc_mesh_from_arrays.restype = C.c_void_p
        L.orc_mesh_from_arrays.argtypes = [C.c_int, _f64p, C.c_int, _i32p, _i32p, _i32p, C.c_int, C.c_void_p,
                                           C.c_void_p]
        L.orc_mesh_counts.argtypes = [C.c_void_p, _i32p]
        L.orc_mesh_topology.argtypes = [C.c_void_p] + [_i32p] * 13
        L.orc_mesh_nodes.argtypes = [C.c_void_p, _f64p]
        L.orc_mesh_geometry.argtypes = [C.c_void_p, _f64p, _f64p, _f64p, _f64p]
        L.orc_mesh_regions.argtypes = [C.c_void_p, _i32p, _i32p, _i32p, C.c_char_p, C.c_int]
        L.orc_local_matrix.restype = C.c_double
        L.orc_local_matrix.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_void_p, _f64p, _f64p]
        L.orc_darcy_create.restype = C.c_void_p
        L.orc_darcy_create.argtypes = [C.c_void_p, _f64p, _f64p, C.c_void_p, _f64p, _f64p, _i32p, _f64p, _f64p, _f64p]
        L.orc_darcy_free.argtypes = [C.c_void_p]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_darcy_solve.restype = C.c_int
        L.orc_darcy_solve.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int]
        L.orc_darcy_size.argtypes = [C.c_void_p, _i32p]
        L.orc_darcy_solution.argtypes = [C.c_void_p, _f64p]
        L.orc_darcy_stats.argtypes = [C.c_void_p, _f64p]
        L.orc_darcy_edge_system.argtypes = [C.c_void_p, _i32p, _i32p, _f64p, _f64p]
        L.orc_darcy_mh_residual.restype = C.c_double
        L.orc_darcy_mh_residual.argtypes = [C.c_void_p]
        L.orc_darcy_water_balance.argtypes = [C.c_void_p, _f64p, _f64p, _f64p]
        L.orc_darcy_element_balance.argtypes = [C.c_void_p, _f64p]
        L.orc_transport_create.restype = C.c_void_p
        L.orc_transport_create.argtypes = [C.c_void_p, _f64p, _f64p, _f64p, C.c_int, _f64p, _f64p, C.c_double, C.c_double]
        L.orc_transport_free.argtypes = [C.c_void_p]
        L.orc_transport_update.restype = C.c_int
        L.orc_transport_update.argtypes = [C.c_void_p, C.c_double]
        for n in ("orc_transport_time", "orc_transport_dt", "orc_transport_cfl", "orc_transport_step_seconds"):
            getattr(L, n).restype = C.c_double
            getattr(L, n).argtypes = [C.c_void_p]
        L.orc_transport_conc.argtypes = [C.c_void_p, _f64p]
        L.orc_transport_matrix.restype = C.c_int
        L.orc_transport_matrix.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_transport_mass_balance.argtypes = [C.c_void_p] + [_f64p] * 6
        L.orc_transport_set_sources.argtypes = [C.c_void_p, _f64p, _f64p, _f64p]
        L.orc_transport_set_bc_conc.argtypes = [C.c_void_p, _f64p]
        L.orc_darcy_element_output.argtypes = [C.c_void_p, _f64p, _f64p, _f64p]
        L.orc_darcy_node_pressure.argtypes = [C.c_void_p, _f64p]
        L.orc_transport_set_phases.argtypes = [C.c_void_p, C.c_int, C.c_int, _f64p, C.c_void_p, C.c_void_p,
                                               C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_transport_conc_phase.argtypes = [C.c_void_p, C.c_int, _f64p]
        L.orc_transport_set_bc_function.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_transport_set_reaction_sorption.argtypes = [C.c_void_p, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p,
                                                          _f64p, _f64p, C.c_double, C.c_int]
        L.orc_transport_set_reaction_sorption_el.argtypes = [C.c_void_p, C.c_int, _i32p, _i32p, _f64p, _f64p,
                                                             _f64p, _f64p, _f64p, C.c_double]
        L.orc_transport_reaction_sorbed.argtypes = [C.c_void_p, _f64p]
        L.orc_transport_set_velocity.argtypes = [C.c_void_p, _f64p]
        L.orc_transport_set_pade.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_darcy_unsteady_setup.argtypes = [C.c_void_p, C.c_int, _f64p, _f64p, C.c_double]
        L.orc_darcy_unsteady_step.restype = C.c_int
        L.orc_darcy_unsteady_step.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double, C.c_double, C.c_int]
        L.orc_transport_set_reactions.argtypes = [C.c_void_p, C.c_int, _i32p, _i32p, _i32p, _f64p, _f64p]
        _LIB = L
    return _LIB


class Mesh:
    """Mesh read and topologised by the oracle (msh_gmshreader.cc, mesh.cc)."""

    def __init__(self, path=None, arrays=None):
        L = lib()
        if path is not None:
            self.h = L.orc_mesh_load(os.fsencode(path))
        else:
            a = arrays
            xyz = np.ascontiguousarray(a["nodes"], np.float64).reshape(-1)
            et = np.ascontiguousarray(a["elem_type"], np.int32)
            er = np.ascontiguousarray(a["elem_region"], np.int32)
            en = np.ascontiguousarray(a["elem_nodes"], np.int32).reshape(-1)
            phys = a.get("physical", [])
            pid = (C.c_int * max(len(phys), 1))(*[p[1] for p in phys])
            pnm = (C.c_char_p * max(len(phys), 1))(*[p[2].encode() for p in phys])
            self.h = L.orc_mesh_from_arrays(len(xyz) // 3, xyz, len(et), et, er, en, len(phys),
                                            C.cast(pid, C.c_void_p), C.cast(pnm, C.c_void_p))
        if not self.h:
            raise RuntimeError("oracle mesh load failed: %s" % L.orc_last_error().decode())
        c = np.zeros(9, np.int32)
        L.orc_mesh_counts(self.h, c)
        (self.n_nodes, self.n_elements, self.n_bc_elements, self.n_sides, self.n_edges,
         self.n_boundaries, self.n_vb, self.n_regions, self.n_file_bc) = [int(v) for v in c]
        E, F, S, V, B, NB = self.n_elements, self.n_edges, self.n_sides, self.n_vb, self.n_boundaries, self.n_bc_elements
        self.elem_dim = np.zeros(E, np.int32)
        self.elem_id = np.zeros(E, np.int32)
        self.elem_region = np.zeros(E, np.int32)
        self.elem_nodes = np.zeros(4 * E, np.int32)
        self.elem_edge = np.zeros(4 * E, np.int32)
        self.elem_bnd = np.zeros(4 * E, np.int32)
        self.edge_ptr = np.zeros(F + 1, np.int32)
        self.edge_side = np.zeros(max(S, 1), np.int32)
        self.vb_elem = np.zeros(max(V, 1), np.int32)
        self.vb_edge = np.zeros(max(V, 1), np.int32)
        self.bnd_edge = np.zeros(max(B, 1), np.int32)
        self.bnd_bcele = np.zeros(max(B, 1), np.int32)
        self.bc_region = np.zeros(max(NB, 1), np.int32)
        L.orc_mesh_topology(self.h, self.elem_dim, self.elem_id, self.elem_region, self.elem_nodes, self.elem_edge,
                            self.elem_bnd, self.edge_ptr, self.edge_side, self.vb_elem, self.vb_edge,
                            self.bnd_edge, self.bnd_bcele, self.bc_region)
        self.vb_elem, self.vb_edge = self.vb_elem[:V], self.vb_edge[:V]
        self.bnd_edge, self.bnd_bcele, self.bc_region = self.bnd_edge[:B], self.bnd_bcele[:B], self.bc_region[:NB]
        self.nodes = np.zeros(3 * self.n_nodes)
        L.orc_mesh_nodes(self.h, self.nodes)
        self.measure = np.zeros(E)
        self.centre = np.zeros(3 * E)
        self.bc_centre = np.zeros(3 * max(NB, 1))
        self.bc_measure = np.zeros(max(NB, 1))
        L.orc_mesh_geometry(self.h, self.measure, self.centre, self.bc_centre, self.bc_measure)
        R = self.n_regions
        ids, isb, sub = np.zeros(R, np.int32), np.zeros(R, np.int32), np.zeros(R, np.int32)
        buf = C.create_string_buffer(1 << 16)
        L.orc_mesh_regions(self.h, ids, isb, sub, buf, len(buf))
        labels = buf.value.decode().split("\n")[:R]
        self.regions = [dict(id=int(i), boundary=bool(b), sub_idx=int(s), label=l)
                        for i, b, s, l in zip(ids, isb, sub, labels)]

    def local_matrix(self, e, K=1.0, cs=1.0, aniso=None):
        d = int(self.elem_dim[e]) + 1
        loc, inv = np.zeros(16), np.zeros(16)
        a = None if aniso is None else np.ascontiguousarray(aniso, np.float64)
        det = lib().orc_local_matrix(self.h, e, K, cs, None if a is None else a.ctypes.data, loc, inv)
        return loc[:d * d].reshape(d, d), inv[:d * d].reshape(d, d), det

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mesh_free(self.h)
            self.h = None


def set_threads(n):
    """OpenMP threads of the Jacobi-CG (CPU baseline timing only)."""
    lib().orc_set_threads(int(n))


class Darcy:
    """DarcyFlowMH_Steady restated (darcy_flow_mh.cc); BC arrays are per bc element."""

    def __init__(self, mesh, K, cs, sigma, f, bc_type, bc_p, bc_flux, bc_sigma, aniso=None):
        self.mesh = mesh
        L = lib()
        self._keep = [np.ascontiguousarray(a, np.float64) for a in (K, cs, sigma, f, bc_p, bc_flux, bc_sigma)]
        self._bct = np.ascontiguousarray(bc_type, np.int32)
        self._an = None if aniso is None else np.ascontiguousarray(aniso, np.float64).reshape(-1)
        K, cs, sigma, f, bc_p, bc_flux, bc_sigma = self._keep
        self.h = L.orc_darcy_create(mesh.h, K, cs, None if self._an is None else self._an.ctypes.data, sigma, f,
                                    self._bct, bc_p, bc_flux, bc_sigma)

    def solve(self, method=0, rtol=1e-14, atol=0.0, max_it=100000):
        """method 0: Jacobi-CG (threaded when set_threads(n > 1)), 1: dense LU of
        the edge system, 2: dense LU of the whole MH system, 3: IC(0)-CG."""
        it = lib().orc_darcy_solve(self.h, method, rtol, atol, max_it)
        if it < 0:
            raise RuntimeError(lib().orc_last_error().decode())
        sz = np.zeros(5, np.int32)
        lib().orc_darcy_size(self.h, sz)
        self.S, self.E, self.F, self.N, self.nnz = [int(v) for v in sz]
        self.solution = np.zeros(self.N)
        lib().orc_darcy_solution(self.h, self.solution)
        st = np.zeros(7)
        lib().orc_darcy_stats(self.h, st)
        self.stats = dict(iters=int(st[0]), resid=st[1], bnorm=st[2], t_assembly=st[3], t_schur=st[4],
                          t_solve=st[5], t_resolve=st[6])
        return it

    def element_output(self):
        E = self.mesh.n_elements
        p, h, v = np.zeros(E), np.zeros(E), np.zeros(3 * E)
        lib().orc_darcy_element_output(self.h, p, h, v)
        return p, h, v.reshape(E, 3)

    def node_pressure(self):
        out = np.zeros(max(self.mesh.n_nodes, 1))
        lib().orc_darcy_node_pressure(self.h, out)
        return out[:self.mesh.n_nodes]

    def set_unsteady(self, lumped, storativity, init_pressure, dt):
        self._unst = [np.ascontiguousarray(storativity, np.float64), np.ascontiguousarray(init_pressure, np.float64)]
        lib().orc_darcy_unsteady_setup(self.h, int(lumped), self._unst[0], self._unst[1], float(dt))
        sz = np.zeros(5, np.int32)
        lib().orc_darcy_size(self.h, sz)
        self.S, self.E, self.F, self.N, self.nnz = [int(v) for v in sz]
        self.solution = np.zeros(self.N)
        lib().orc_darcy_solution(self.h, self.solution)

    def unsteady_step(self, dt, method=0, rtol=1e-14, atol=0.0, max_it=100000):
        it = lib().orc_darcy_unsteady_step(self.h, float(dt), method, rtol, atol, max_it)
        self.solution = np.zeros(self.N)
        lib().orc_darcy_solution(self.h, self.solution)
        return it

    def edge_system(self):
        ptr = np.zeros(self.F + 1, np.int32)
        col = np.zeros(self.nnz, np.int32)
        val = np.zeros(self.nnz)
        rhs = np.zeros(self.F)
        lib().orc_darcy_edge_system(self.h, ptr, col, val, rhs)
        return ptr, col, val, rhs

    def mh_residual(self):
        return lib().orc_darcy_mh_residual(self.h)

    def water_balance(self):
        nbr = sum(1 for r in self.mesh.regions if r["boundary"])
        b, p, m = np.zeros(max(nbr, 1)), np.zeros(max(nbr, 1)), np.zeros(max(nbr, 1))
        lib().orc_darcy_water_balance(self.h, b, p, m)
        return b[:nbr], p[:nbr], m[:nbr]

    def element_balance(self):
        out = np.zeros(self.mesh.n_elements)
        lib().orc_darcy_element_balance(self.h, out)
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_darcy_free(self.h)
            self.h = None


BC_FN = C.CFUNCTYPE(None, C.c_double, C.POINTER(C.c_double), C.c_void_p)


class Transport:
    """ConvectionTransport + TransportOperatorSplitting loop restated (transport.cc)."""

    def __init__(self, mesh, side_flux, cs, por, bc_conc, init_conc, end_time, save_step):
        self.mesh = mesh
        self.nsub = int(np.asarray(bc_conc).reshape(-1, mesh.n_bc_elements).shape[0])
        self._keep = [np.ascontiguousarray(a, np.float64).reshape(-1) for a in (side_flux, cs, por, bc_conc, init_conc)]
        sf, cs, por, bcc, ic = self._keep
        self.h = lib().orc_transport_create(mesh.h, sf, cs, por, self.nsub, bcc, ic, end_time, save_step)

    def update(self, t):
        n = lib().orc_transport_update(self.h, t)
        if n == -1:
            raise ValueError("sorption: the isotherms cannot be made (zero scale_sorbed or a table without a root)")
        if n == -2:
            raise ValueError("sorption: negative total mass or an isotherm equation without a root in its bracket")
        return n

    @property
    def time(self):
        return lib().orc_transport_time(self.h)

    @property
    def cfl(self):
        return lib().orc_transport_cfl(self.h)

    @property
    def dt(self):
        return lib().orc_transport_dt(self.h)

    def conc(self):
        out = np.zeros(self.nsub * self.mesh.n_elements)
        lib().orc_transport_conc(self.h, out)
        return out.reshape(self.nsub, -1)

    def matrix(self):
        n = lib().orc_transport_matrix(self.h, None, None, None)
        ptr = np.zeros(self.mesh.n_elements + 1, np.int32)
        col = np.zeros(n, np.int32)
        val = np.zeros(n)
        lib().orc_transport_matrix(self.h, ptr.ctypes.data, col.ctypes.data, val.ctypes.data)
        return ptr, col, val

    def mass_balance(self):
        nbr = sum(1 for r in self.mesh.regions if r["boundary"])
        nblk = sum(1 for r in self.mesh.regions if not r["boundary"])
        ns = self.nsub
        bal, plus, minus = np.zeros(ns * max(nbr, 1)), np.zeros(ns * max(nbr, 1)), np.zeros(ns * max(nbr, 1))
        mass, src, tot = np.zeros(ns * max(nblk, 1)), np.zeros(ns * max(nblk, 1)), np.zeros(4 * ns)
        lib().orc_transport_mass_balance(self.h, bal, plus, minus, mass, src, tot)
        return dict(bal=bal[:ns * nbr].reshape(ns, nbr), plus=plus[:ns * nbr].reshape(ns, nbr),
                    minus=minus[:ns * nbr].reshape(ns, nbr), mass=mass[:ns * nblk].reshape(ns, nblk),
                    src=src[:ns * nblk].reshape(ns, nblk), totals=tot.reshape(ns, 4))

    def set_sources(self, density, sigma, conc):
        self._src = [np.ascontiguousarray(a, np.float64).reshape(-1) for a in (density, sigma, conc)]
        lib().orc_transport_set_sources(self.h, *self._src)

    def set_velocity(self, side_flux):
        self._keep[0] = np.ascontiguousarray(side_flux, np.float64).reshape(-1).copy()
        lib().orc_transport_set_velocity(self.h, self._keep[0])

    def set_bc_conc(self, bc_conc):
        self._bcc = np.ascontiguousarray(bc_conc, np.float64).reshape(-1)
        lib().orc_transport_set_bc_conc(self.h, self._bcc)

    def set_reactions(self, parent, products, half_life, branch):
        """parent[n]; products: list of lists of substance indices; branch: same shape."""
        pptr = np.concatenate([[0], np.cumsum([len(p) for p in products])]).astype(np.int32)
        self._react = [np.ascontiguousarray(parent, np.int32), pptr,
                       np.ascontiguousarray([q for p in products for q in p], np.int32),
                       np.ascontiguousarray(half_life, np.float64),
                       np.ascontiguousarray([b for p in branch for b in p], np.float64)]
        lib().orc_transport_set_reactions(self.h, len(parent), *self._react)

    def set_phases(self, dual, sorption, por_imm, alpha=None, phi=None, sorp_type=None, sorp_coef0=None,
                   sorp_coef1=None):
        """Dual porosity / sorption: por_imm, phi [E]; alpha, sorp_type (0 none,
        1 linear, 2 freundlich, 3 langmuir), sorp_coef0/1 [nsub][E]."""
        E = self.mesh.n_elements

        def arr(a, dt, shape):
            return None if a is None else np.ascontiguousarray(np.broadcast_to(a, shape), dt).reshape(-1)

        self._phases = [arr(por_imm, np.float64, (E,)), arr(alpha, np.float64, (self.nsub, E)),
                        arr(phi, np.float64, (E,)), arr(sorp_type, np.int32, (self.nsub, E)),
                        arr(sorp_coef0, np.float64, (self.nsub, E)), arr(sorp_coef1, np.float64, (self.nsub, E))]
        ptr = [None if a is None else a.ctypes.data for a in self._phases[1:]]
        lib().orc_transport_set_phases(self.h, int(dual), int(sorption), self._phases[0], *ptr)

    def conc_phase(self, phase):
        out = np.zeros(self.nsub * self.mesh.n_elements)
        lib().orc_transport_conc_phase(self.h, int(phase), out)
        return out.reshape(self.nsub, -1)

    def set_bc_function(self, fn):
        """Boundary concentrations fn(t) -> [nsub][NBC], evaluated at every step's start."""
        shape = (self.nsub, self.mesh.n_bc_elements)

        def cb(t, out, user):
            np.ctypeslib.as_array(out, shape=shape)[:] = fn(t)

        self._bc_fn = BC_FN(cb)
        lib().orc_transport_set_bc_function(self.h, C.cast(self._bc_fn, C.c_void_p), None)

    def set_reaction_sorption(self, subst, types, mult, second, rock_density, molar_mass, solubility=None,
                              solvent_dens=1.0, substeps=1000):
        """Sorption of the reaction module: species on substances subst[n]; types, mult,
        second [n_bulk_regions][n]; rock_density [n_bulk_regions]; molar_mass, solubility [n]."""
        n = len(subst)
        nreg = max(r["sub_idx"] for r in self.mesh.regions if not r["boundary"]) + 1
        self._rs = [np.ascontiguousarray(subst, np.int32),
                    np.ascontiguousarray(np.broadcast_to(types, (nreg, n)), np.int32).reshape(-1),
                    np.ascontiguousarray(np.broadcast_to(mult, (nreg, n)), np.float64).reshape(-1),
                    np.ascontiguousarray(np.broadcast_to(second, (nreg, n)), np.float64).reshape(-1),
                    np.ascontiguousarray(np.broadcast_to(rock_density, (nreg,)), np.float64),
                    np.ascontiguousarray(np.broadcast_to(molar_mass, (n,)), np.float64),
                    np.ascontiguousarray(np.broadcast_to(-1.0 if solubility is None else solubility, (n,)),
                                         np.float64)]
        lib().orc_transport_set_reaction_sorption(self.h, n, *self._rs, float(solvent_dens), int(substeps))

    def set_reaction_sorption_elementwise(self, subst, types, mult, second, rock_density, molar_mass,
                                          solubility=None, solvent_dens=1.0):
        """The same with element-wise parameter fields: types, mult, second [n][E];
        rock_density [E] (every element's isotherm solved directly)."""
        n, E = len(subst), self.mesh.n_elements
        self._rs = [np.ascontiguousarray(subst, np.int32),
                    np.ascontiguousarray(np.broadcast_to(types, (n, E)), np.int32).reshape(-1),
                    np.ascontiguousarray(np.broadcast_to(mult, (n, E)), np.float64).reshape(-1),
                    np.ascontiguousarray(np.broadcast_to(second, (n, E)), np.float64).reshape(-1),
                    np.ascontiguousarray(np.broadcast_to(rock_density, (E,)), np.float64),
                    np.ascontiguousarray(np.broadcast_to(molar_mass, (n,)), np.float64),
                    np.ascontiguousarray(np.broadcast_to(-1.0 if solubility is None else solubility, (n,)),
                                         np.float64)]
        lib().orc_transport_set_reaction_sorption_el(self.h, n, *self._rs, float(solvent_dens))

    def reaction_sorbed(self):
        out = np.zeros(len(self._rs[0]) * self.mesh.n_elements)
        lib().orc_transport_reaction_sorbed(self.h, out)
        return out.reshape(len(self._rs[0]), -1)

    def set_pade(self, nom_deg, den_deg):
        """Use the Pade approximant of the given degrees for the decays already set."""
        lib().orc_transport_set_pade(self.h, int(nom_deg), int(den_deg))

    def step_seconds(self):
        return lib().orc_transport_step_seconds(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_transport_free(self.h)
            self.h = None
