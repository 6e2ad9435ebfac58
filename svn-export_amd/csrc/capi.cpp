// C ABI over the host mesh module and the HIP solvers (include/flow123d_gpu.h).
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <string>

#include "../../include/flow123d_gpu.h"
#include "batched.hh"
#include "common.hh"
#include "darcy.hh"
#include "gnn.hh"
#include "mesh.hh"
#include "msh_data.hh"
#include "old_bcd.hh"
#include "balance_out.hh"
#include "transport.hh"

struct fl_mesh {
    fl::HostMesh m;
};
struct fl_darcy {
    std::unique_ptr<fl::DarcyGPU> g;
    fl_mesh *mesh;
    fl::TimeGovernorR tg;  // steady until fl_darcy_set_time
};
struct fl_darcy_batch {
    std::unique_ptr<fl::DarcyBatchGPU> g;
};
struct fl_gnn {
    std::unique_ptr<fl::GnnGPU> g;
};
struct fl_transport {
    std::unique_ptr<fl::TransportGPU> g;
    fl_mesh *mesh;
};
struct fl_old_bcd {
    fl::OldBcd b;
    const fl_mesh *mesh;
};
struct fl_element_data {
    fl::ElementDataReader reader;
    std::vector<int> ids;
    std::vector<double> data;
    std::string field;
    int n_comp = 1;
};

static thread_local std::string g_err;

template <class F>
static int guard(F &&f) {
    try {
        return f();
    } catch (const fl::Error &e) {
        g_err = e.what();
        return e.code;
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}

static int check_device() {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) throw fl::Error(-6, "no HIP device available (the GPU path has no CPU fallback)");
    return 0;
}

extern "C" {

const char *fl_last_error(void) { return g_err.c_str(); }

int fl_mesh_read_gmsh(const char *path, fl_mesh **out) {
    return guard([&] {
        auto *mh = new fl_mesh();
        std::string err = fl::read_gmsh(path, mh->m);
        if (err.empty()) err = fl::build_topology(mh->m);
        if (!err.empty()) {
            delete mh;
            throw fl::Error(-1, err);
        }
        *out = mh;
        return 0;
    });
}

int fl_mesh_from_arrays(int n_nodes, const double *xyz, int n_elem, const int *elem_id, const int *gmsh_type,
                        const int *region_id, const int *nodes4, int n_phys, const int *phys_id,
                        const char *const *phys_name, fl_mesh **out) {
    return guard([&] {
        auto *mh = new fl_mesh();
        std::string err = fl::mesh_from_arrays(mh->m, n_nodes, xyz, n_elem, elem_id, gmsh_type, region_id, nodes4,
                                               n_phys, phys_id, phys_name);
        if (err.empty()) err = fl::build_topology(mh->m);
        if (!err.empty()) {
            delete mh;
            throw fl::Error(-1, err);
        }
        *out = mh;
        return 0;
    });
}

void fl_mesh_free(fl_mesh *m) { delete m; }

int fl_mesh_counts(const fl_mesh *mh, int *c) {
    const fl::HostMesh &m = mh->m;
    c[0] = m.n_nodes(); c[1] = m.n_elements(); c[2] = m.n_bc_elements(); c[3] = m.n_sides(); c[4] = m.n_edges();
    c[5] = m.n_boundaries(); c[6] = m.n_vb(); c[7] = (int)m.regions.size(); c[8] = m.n_file_bc;
    return 0;
}

int fl_mesh_topology(const fl_mesh *mh, int *elem_dim, int *elem_id, int *elem_region, int *elem_nodes,
                     int *elem_edge, int *elem_bnd, int *edge_ptr, int *edge_side, int *vb_elem, int *vb_edge,
                     int *bnd_edge, int *bnd_bcele, int *bc_region) {
    const fl::HostMesh &m = mh->m;
    const int E = m.n_elements();
    for (int e = 0; e < E; e++) {
        elem_dim[e] = m.el_dim[e];
        elem_id[e] = m.el_id[e];
        elem_region[e] = m.regions[m.el_region[e]].id;
        for (int k = 0; k < 4; k++) {
            elem_nodes[4 * e + k] = m.el_nodes[4 * e + k];
            bool has = k <= m.el_dim[e];
            elem_edge[4 * e + k] = has ? m.side_edge[m.side_off[e] + k] : -1;
            elem_bnd[4 * e + k] = has ? m.side_bnd[m.side_off[e] + k] : -1;
        }
    }
    std::memcpy(edge_ptr, m.edge_ptr.data(), m.edge_ptr.size() * sizeof(int));
    std::memcpy(edge_side, m.edge_side.data(), m.edge_side.size() * sizeof(int));
    std::memcpy(vb_elem, m.vb_elem.data(), m.vb_elem.size() * sizeof(int));
    std::memcpy(vb_edge, m.vb_edge.data(), m.vb_edge.size() * sizeof(int));
    std::memcpy(bnd_edge, m.bnd_edge.data(), m.bnd_edge.size() * sizeof(int));
    std::memcpy(bnd_bcele, m.bnd_bcele.data(), m.bnd_bcele.size() * sizeof(int));
    for (int b = 0; b < m.n_bc_elements(); b++) bc_region[b] = m.regions[m.bc_region[b]].id;
    return 0;
}

int fl_mesh_regions(const fl_mesh *mh, int *ids, int *is_boundary, int *sub_idx, char *buf, int buflen) {
    const fl::HostMesh &m = mh->m;
    std::string all;
    for (size_t r = 0; r < m.regions.size(); r++) {
        ids[r] = m.regions[r].id;
        is_boundary[r] = m.regions[r].boundary;
        sub_idx[r] = m.regions[r].sub_idx;
        all += m.regions[r].label + "\n";
    }
    std::snprintf(buf, buflen, "%s", all.c_str());
    return 0;
}

int fl_mesh_element_ids(const fl_mesh *mh, int boundary, int *ids) {
    return guard([&] {
        std::vector<int> v;
        std::string err = fl::element_id_map(mh->m, boundary != 0, v);
        if (!err.empty()) throw fl::Error(-1, err);
        if (!v.empty()) std::memcpy(ids, v.data(), v.size() * sizeof(int));
        return 0;
    });
}

int fl_element_data_open(const fl_mesh *m, const char *gmsh_file, const char *field_name, int n_comp, int boundary,
                         fl_element_data **out) {
    return guard([&] {
        if (!m || !gmsh_file || !field_name || !out) throw fl::Error(-3, "null argument");
        if (n_comp < 1) throw fl::Error(-3, "n_comp must be positive");
        std::unique_ptr<fl_element_data> f(new fl_element_data());
        std::string err = fl::element_id_map(m->m, boundary != 0, f->ids);
        if (err.empty()) err = f->reader.open(gmsh_file);
        if (!err.empty()) throw fl::Error(-1, err);
        f->field = field_name;
        f->n_comp = n_comp;
        f->data.assign(f->ids.size() * (size_t)n_comp, 0.0);
        *out = f.release();
        return 0;
    });
}

int fl_element_data_set_time(fl_element_data *f, double time, int *updated) {
    return guard([&] {
        if (updated) *updated = 0;
        if (time == std::numeric_limits<double>::infinity()) return 0;
        fl::MshDataHeader search;
        search.actual = false;
        search.field_name = f->field;
        search.n_components = (unsigned)f->n_comp;
        search.n_entities = (unsigned)f->ids.size();
        search.time = time;
        std::string err = f->reader.read(search, f->data.data(), f->ids);
        if (!err.empty()) throw fl::Error(-1, err);
        if (updated) *updated = search.actual ? 1 : 0;
        return 0;
    });
}

int fl_element_data_values(const fl_element_data *f, double *values) {
    if (!f->data.empty()) std::memcpy(values, f->data.data(), f->data.size() * sizeof(double));
    return 0;
}

int fl_element_data_info(const fl_element_data *f, int *info, char *buf, int buflen) {
    info[0] = (int)f->ids.size();
    info[1] = f->n_comp;
    info[2] = f->reader.n_read;
    if (buf && buflen > 0) {
        std::string all;
        for (const auto &w : f->reader.warnings) all += w + "\n";
        std::snprintf(buf, (size_t)buflen, "%s", all.c_str());
    }
    return 0;
}

void fl_element_data_free(fl_element_data *f) { delete f; }

int fl_msh_write_head(const fl_mesh *m, const char *path) {
    return guard([&] {
        std::string err = fl::msh_write_head(m->m, path);
        if (!err.empty()) throw fl::Error(-1, err);
        return 0;
    });
}

int fl_msh_write_element_data(const char *path, const char *name, const char *units, double time, int step,
                              int n_comp, int n_values, const double *values, int precision) {
    return guard([&] {
        std::string err =
            fl::msh_write_element_data(path, name, units, time, step, n_comp, n_values, values, precision);
        if (!err.empty()) throw fl::Error(-1, err);
        return 0;
    });
}

int fl_old_bcd_read_flow(const fl_mesh *m, const char *path, int *bc_type, double *bc_pressure, double *bc_flux,
                         double *bc_sigma, fl_old_bcd **out) {
    return guard([&] {
        std::unique_ptr<fl_old_bcd> h(new fl_old_bcd());
        h->mesh = m;
        std::string err = h->b.read_flow(m->m, path, bc_type, bc_pressure, bc_flux, bc_sigma);
        if (!err.empty()) throw fl::Error(-1, err);
        *out = h.release();
        return 0;
    });
}

int fl_old_bcd_read_transport(const fl_old_bcd *h, const char *path, int n_subst, double *bc_conc) {
    return guard([&] {
        if (n_subst < 1) throw fl::Error(-3, "n_subst must be positive");
        std::string err = h->b.read_transport(h->mesh->m, path, n_subst, bc_conc);
        if (!err.empty()) throw fl::Error(-1, err);
        return 0;
    });
}

void fl_old_bcd_free(fl_old_bcd *h) { delete h; }

int fl_write_water_balance(const fl_mesh *m, const char *path, double time, const double *bal, const double *plus,
                           const double *minus, const double *src) {
    return guard([&] {
        std::vector<fl::RegionRow> bnd, blk;
        fl::region_sets(m->m, bnd, blk);
        std::string err = fl::write_water_balance(path, time, bnd, bal, plus, minus, blk, src);
        if (!err.empty()) throw fl::Error(-1, err);
        return 0;
    });
}

int fl_write_mass_balance(const fl_mesh *m, const char *path, double time, double initial_time, int n_subst,
                          const char *const *subst_names, const double *bal, const double *plus, const double *minus,
                          const double *mass, const double *src, const double *totals) {
    return guard([&] {
        if (n_subst < 1) throw fl::Error(-3, "n_subst must be positive");
        std::vector<fl::RegionRow> bnd, blk;
        fl::region_sets(m->m, bnd, blk);
        std::vector<std::string> names(n_subst);
        std::vector<double> a(n_subst), b(n_subst), c(n_subst);
        for (int k = 0; k < n_subst; k++) {
            names[k] = subst_names && subst_names[k] ? subst_names[k] : "substance_" + std::to_string(k);
            a[k] = totals[4 * k];
            c[k] = totals[4 * k + 1];
            b[k] = totals[4 * k + 2];
        }
        std::string err = fl::write_mass_balance(path, time, initial_time, n_subst, names, bnd, bal, plus, minus,
                                                 blk, mass, src, a.data(), b.data(), c.data());
        if (!err.empty()) throw fl::Error(-1, err);
        return 0;
    });
}

int fl_darcy_create(fl_mesh *m, fl_darcy **out) {
    return guard([&] {
        check_device();
        auto *d = new fl_darcy();
        d->mesh = m;
        d->tg.init_steady();
        try {
            d->g.reset(new fl::DarcyGPU(m->m));
        } catch (...) {
            delete d;
            throw;
        }
        *out = d;
        return 0;
    });
}

void fl_darcy_free(fl_darcy *d) { delete d; }

int fl_darcy_set_data(fl_darcy *d, const double *K, const double *cs, const double *aniso, const double *sigma,
                      const double *f, const int *bc_type, const double *bc_pressure, const double *bc_flux,
                      const double *bc_robin_sigma) {
    return guard([&] {
        d->g->set_data(K, cs, aniso, sigma, f, bc_type, bc_pressure, bc_flux, bc_robin_sigma);
        return 0;
    });
}

int fl_darcy_set_conductivity(fl_darcy *d, const double *K) {
    return guard([&] {
        d->g->set_conductivity(K);
        return 0;
    });
}

static void fill_info(const fl::SolveStats &st, fl_solve_info *info);

// 0 converged; a positive warning when the iteration limit stopped CG (the
// solution is still the last iterate); an error when the residual is not finite
static int solve_status(const fl::SolveStats &st) {
    if (st.converged) return 0;
    char buf[160];
    if (!std::isfinite(st.residual)) {
        snprintf(buf, sizeof buf, "CG diverged: residual %g after %d iterations", st.residual, st.iterations);
        throw fl::Error(fl::kErrDiverged, buf);
    }
    snprintf(buf, sizeof buf, "CG stopped at the iteration limit: %d iterations, residual %.6g of rhs norm %.6g",
             st.iterations, st.residual, st.bnorm);
    g_err = buf;
    return (int)fl::kStatusNotConverged;
}

int fl_darcy_update_solution(fl_darcy *d, const fl_solver_opts *o, fl_solve_info *info) {
    return guard([&] {
        double rtol = o ? o->r_tol : 1e-7, atol = o ? o->a_tol : 1e-9;
        int max_it = o ? o->max_it : 10000;  // solve.cc:69-84 defaults
        int prec = o ? o->precision : 0;
        const int norm = o ? o->norm : 0;
        if (norm != 0 && norm != 1) throw fl::Error(-8, "norm must be 0 (2-norm) or 1 (natural Jacobi norm)");
        if (prec != 0 && prec != 1) throw fl::Error(-8, "precision must be 0 (FP64) or 1 (FP32 + FP64 reliable updates)");
        fl::SolveStats st;
        if (!d->tg.steady) {
            // unsteady governor (fl_darcy_set_time): one step of its dt, nothing
            // at the end time (DarcyFlowMH_Steady::update_solution, darcy_flow_mh.cc:375-382)
            if (d->tg.is_end()) {
                st.converged = 1;
                fill_info(st, info);
                return 0;
            }
            d->tg.next_time();
            st = d->g->unsteady_step(d->tg.time_step, rtol, atol, max_it, prec, norm);
        } else {
            // steady: solved again on every call (new data), the governor moves
            // to t = inf, dt = 0 as after the reference's one steady solve
            st = d->g->solve(rtol, atol, max_it, prec, norm);
            d->tg.next_time();
        }
        fill_info(st, info);
        return solve_status(st);
    });
}

int fl_darcy_set_time(fl_darcy *d, double end_time, double init_dt, double save_step) {
    return guard([&] {
        if (!(init_dt > 0.0) || !(end_time >= init_dt) || !std::isfinite(end_time))
            throw fl::Error(-1, "time: need 0 < init_dt <= end_time < inf");
        if (!d->g->is_unsteady()) throw fl::Error(-3, "an unsteady time governor needs fl_darcy_set_unsteady first");
        d->tg.init_fixed(end_time, init_dt, save_step);
        d->tg.fix_dt_until_mark();  // DarcyFlowMH_Unsteady constructor (darcy_flow_mh.cc:1606-1608)
        return 0;
    });
}

int fl_darcy_time(fl_darcy *d, double *time, double *dt, int *is_end) {
    if (time) *time = d->tg.time;
    if (dt) *dt = d->tg.time_step;
    if (is_end) *is_end = d->tg.is_end() ? 1 : 0;
    return 0;
}

int fl_darcy_set_time_upper_constraint(fl_darcy *d, double dt) {
    return guard([&] {
        if (!(dt > 0.0)) throw fl::Error(-1, "time step constraint must be positive");
        (void)d->tg.set_upper_constraint(dt);
        return 0;
    });
}

int fl_darcy_planned_time(fl_darcy *d, double *time) {
    if (!time) return -1;
    *time = d->tg.estimate_time();
    return 0;
}

int fl_darcy_solved_time(fl_darcy *d, double *time) {
    if (!time) return -1;
    *time = d->tg.time;
    return 0;
}

int fl_darcy_choose_next_time(fl_darcy *d) {
    (void)d->tg.fix_dt_until_mark();
    return 0;
}

int fl_darcy_update_solution_noalloc(fl_darcy *d, const fl_solver_opts *o, fl_solve_info *info) {
    return guard([&] {
        double rtol = o ? o->r_tol : 1e-7, atol = o ? o->a_tol : 1e-9;
        int max_it = o ? o->max_it : 10000;
        const int prec = o ? o->precision : 0;
        const int norm = o ? o->norm : 0;
        if (prec != 0) throw fl::Error(-8, "the no-alloc solve runs FP64 CG only (precision 0)");
        if (norm != 0 && norm != 1) throw fl::Error(-8, "norm must be 0 (2-norm) or 1 (natural Jacobi norm)");
        fl::SolveStats st = d->g->solve_noalloc(rtol, atol, max_it, norm);
        fill_info(st, info);
        return solve_status(st);
    });
}

int fl_darcy_set_unsteady(fl_darcy *d, int lumped, const double *storativity, const double *init_pressure) {
    return guard([&] {
        d->g->set_unsteady(lumped, storativity, init_pressure);
        return 0;
    });
}

int fl_darcy_unsteady_step(fl_darcy *d, double dt, const fl_solver_opts *o, fl_solve_info *info) {
    return guard([&] {
        double rtol = o ? o->r_tol : 1e-7, atol = o ? o->a_tol : 1e-9;
        int max_it = o ? o->max_it : 10000;
        int prec = o ? o->precision : 0;
        const int norm = o ? o->norm : 0;
        if (norm != 0 && norm != 1) throw fl::Error(-8, "norm must be 0 (2-norm) or 1 (natural Jacobi norm)");
        if (prec != 0 && prec != 1) throw fl::Error(-8, "precision must be 0 (FP64) or 1 (FP32 + FP64 reliable updates)");
        fl::SolveStats st = d->g->unsteady_step(dt, rtol, atol, max_it, prec, norm);
        fill_info(st, info);
        return solve_status(st);
    });
}

static void fill_info(const fl::SolveStats &st, fl_solve_info *info) {
    {
        if (info) {
            info->iterations = st.iterations;
            info->converged = st.converged;
            info->residual = st.residual;
            info->rhs_norm = st.bnorm;
            info->ms_assembly = st.ms_assembly;
            info->ms_solve = st.ms_cg;
            info->ms_resolve = st.ms_resolve;
            info->ms_total = st.ms_total;
            info->ms_spmv_avg = st.ms_spmv_avg;
            info->ms_update_avg = st.ms_update_avg;
            info->kernel_samples = st.kernel_samples;
            info->reliable_updates = st.reliable_updates;
        }
    }
}

int fl_darcy_get_solution(fl_darcy *d, double *out) {
    return guard([&] {
        d->g->get_solution(out);
        return 0;
    });
}

const double *fl_darcy_device_solution(fl_darcy *d) { return d->g->device_solution(); }

double fl_darcy_solution_precision(fl_darcy *d, double r_tol, double a_tol) {
    double v = -1.0;
    guard([&] {
        v = d->g->solution_precision(r_tol, a_tol);
        return 0;
    });
    return v;
}

int fl_darcy_element_balance(fl_darcy *d, double *out) {
    return guard([&] {
        d->g->element_balance(out);
        return 0;
    });
}

int fl_darcy_water_balance(fl_darcy *d, double *b, double *p, double *m) {
    return guard([&] {
        d->g->water_balance(b, p, m);
        return 0;
    });
}

int fl_darcy_write_water_balance(fl_darcy *d, const char *path, double time) {
    return guard([&] {
        const fl::HostMesh &m = d->mesh->m;
        std::vector<double> b(m.n_boundary_regions + 1), p(b.size()), mi(b.size()), src(m.n_bulk_regions + 1);
        d->g->water_balance(b.data(), p.data(), mi.data());
        d->g->source_balance(src.data());
        std::vector<fl::RegionRow> bnd, blk;
        fl::region_sets(m, bnd, blk);
        std::string err = fl::write_water_balance(path, time, bnd, b.data(), p.data(), mi.data(), blk, src.data());
        if (!err.empty()) throw fl::Error(-1, err);
        return 0;
    });
}

int fl_darcy_element_output(fl_darcy *d, double *pressure, double *piezo_head, double *velocity) {
    return guard([&] {
        d->g->element_output(pressure, piezo_head, velocity);
        return 0;
    });
}

int fl_darcy_node_pressure(fl_darcy *d, double *pressure_nodes) {
    return guard([&] {
        d->g->node_pressure(pressure_nodes);
        return 0;
    });
}

int fl_msh_write_node_data(const char *path, const char *name, const char *units, double time, int step,
                           int n_comp, int n_values, const double *values, int precision) {
    return guard([&] {
        std::string err =
            fl::msh_write_node_data(path, name, units, time, step, n_comp, n_values, values, precision);
        if (!err.empty()) throw fl::Error(-1, err);
        return 0;
    });
}

int fl_darcy_edge_nnz(fl_darcy *d) { return d->g->nnz(); }

int fl_darcy_edge_system(fl_darcy *d, int *ptr, int *col, double *val, double *rhs) {
    return guard([&] {
        d->g->get_edge_system(ptr, col, val, rhs);
        return 0;
    });
}

int fl_darcy_batch_create(fl_darcy *base, int n_real, fl_darcy_batch **out) {
    return fl_darcy_batch_create_slots(base, n_real, 0, out);
}

int fl_darcy_batch_create_slots(fl_darcy *base, int n_real, int n_slots, fl_darcy_batch **out) {
    return guard([&] {
        check_device();
        if (!base || !out) throw fl::Error(-1, "null handle");
        auto *b = new fl_darcy_batch();
        try {
            b->g.reset(new fl::DarcyBatchGPU(*base->g, n_real, n_slots));
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
        return 0;
    });
}

void fl_darcy_batch_free(fl_darcy_batch *b) { delete b; }

int fl_darcy_batch_set_conductivity(fl_darcy_batch *b, const double *K, int device_ptr) {
    return guard([&] {
        b->g->set_conductivity(K, device_ptr != 0);
        return 0;
    });
}

int fl_darcy_batch_update_solution(fl_darcy_batch *b, const fl_solver_opts *o, fl_batch_info *info,
                                   int *iterations, int *converged, double *residual) {
    return guard([&] {
        double rtol = o ? o->r_tol : 1e-7, atol = o ? o->a_tol : 1e-9;
        int max_it = o ? o->max_it : 10000;
        int prec = o ? o->precision : 0;
        const int norm = o ? o->norm : 0;
        if (norm != 0 && norm != 1) throw fl::Error(-8, "norm must be 0 (2-norm) or 1 (natural Jacobi norm)");
        if (prec != 0 && prec != 1) throw fl::Error(-8, "precision must be 0 (FP64) or 1 (FP32 + FP64 reliable updates)");
        fl::BatchStats st = b->g->solve(rtol, atol, max_it, prec, norm);
        const int R = b->g->n_real();
        int nconv = 0;
        for (int r = 0; r < R; r++) {
            nconv += st.converged[r];
            if (iterations) iterations[r] = st.iterations[r];
            if (converged) converged[r] = st.converged[r];
            if (residual) residual[r] = st.residual[r];
        }
        if (info) {
            info->n_real = R;
            info->chunks = st.chunks;
            info->reliable_updates = st.reliable_updates;
            info->fp64_finish = st.fp64_finish;
            info->max_iterations = st.max_iterations;
            info->n_converged = nconv;
            info->total_iterations = st.total_iterations;
            info->ms_setup = st.ms_setup;
            info->ms_solve = st.ms_cg;
            info->ms_resolve = st.ms_resolve;
            info->ms_total = st.ms_total;
            info->ms_op_avg = st.ms_op_avg;
            info->ms_update_avg = st.ms_update_avg;
            info->grid = st.grid;
            info->xcd_map = st.xcd_map;
            info->chunk = st.chunk;
            info->compactions = st.compactions;
            info->handovers = st.handovers;
            info->handover_iterations = st.handover_iterations;
            info->ms_handover = st.ms_handover;
            info->ms_coarse_avg = st.ms_coarse_avg;
            info->ms_coarse_setup = st.ms_coarse_setup;
            info->ms_queue_empty = st.ms_queue_empty;
            info->ms_narrow16 = st.ms_narrow16;
            info->ms_narrow4 = st.ms_narrow4;
            info->ms_loop_end = st.ms_loop_end;
            info->ms_handover_end = st.ms_handover_end;
            info->tile_rows = st.tile_rows;
        }
        if (nconv == R) return 0;
        int ndiv = 0;
        for (int r = 0; r < R; r++) ndiv += !st.converged[r] && !std::isfinite(st.residual[r]);
        g_err = std::to_string(R - nconv) + " of " + std::to_string(R) + " realizations did not converge (" +
                std::to_string(ndiv) + " with a residual that is not finite)";
        return (int)fl::kStatusNotConverged;
    });
}

int fl_darcy_batch_diagnostics(fl_darcy_batch *b, int *status, int *reliable_updates) {
    return guard([&] {
        b->g->diagnostics(status, reliable_updates);
        return 0;
    });
}

int fl_darcy_batch_get_solution(fl_darcy_batch *b, int r, double *out) {
    return guard([&] {
        b->g->get_solution(r, out);
        return 0;
    });
}

const double *fl_darcy_batch_device_solution(fl_darcy_batch *b, int *n_padded) {
    if (n_padded) *n_padded = b->g->padded();
    return b->g->device_solution();
}

int fl_darcy_batch_set_launch(fl_darcy_batch *b, int grid, int xcd_map, int chunk) {
    return guard([&] {
        b->g->set_launch(grid, xcd_map, chunk);
        return 0;
    });
}

int fl_darcy_batch_set_handover(fl_darcy_batch *b, int n) {
    return guard([&] {
        b->g->set_handover(n);
        return 0;
    });
}

int fl_darcy_batch_set_narrow(fl_darcy_batch *b, int on) {
    return guard([&] {
        b->g->set_narrow(on != 0);
        return 0;
    });
}

int fl_darcy_batch_set_coarse(fl_darcy_batch *b, int ncol) {
    return guard([&] {
        b->g->set_coarse(ncol);
        return 0;
    });
}

int fl_darcy_batch_set_coarse_basis(fl_darcy_batch *b, const float *phi) {
    return guard([&] {
        if (!phi) throw fl::Error(-1, "null basis");
        b->g->set_coarse_basis(phi);
        return 0;
    });
}

int fl_darcy_batch_aggregates(fl_darcy_batch *b, int *edge_agg) {
    return guard([&] {
        const auto &pe = b->g->row_edges();
        const auto &ag = b->g->row_aggregate();
        if (edge_agg)
            for (size_t r = 0; r < pe.size(); r++) edge_agg[pe[r]] = ag[r];
        return b->g->aggregates();
    });
}

int fl_coarse_invert(const double *a, int n, int np, float *inv, int *info) {
    return guard([&] {
        if (!a || !inv || !info) throw fl::Error(-1, "null argument");
        fl::coarse_invert_host(a, n, np, inv, info);
        return 0;
    });
}

int fl_darcy_batch_lambda(fl_darcy_batch *b, double *out) {
    return guard([&] {
        if (!out) throw fl::Error(-1, "null output");
        b->g->lambda_device(out);
        return 0;
    });
}

int fl_gnn_create(fl_darcy_batch *b, unsigned seed, fl_gnn **out) {
    return guard([&] {
        check_device();
        if (!b || !out) throw fl::Error(-1, "null handle");
        auto *g = new fl_gnn();
        try {
            g->g.reset(new fl::GnnGPU(*b->g, seed));
        } catch (...) {
            delete g;
            throw;
        }
        *out = g;
        return 0;
    });
}

void fl_gnn_free(fl_gnn *g) { delete g; }

int fl_gnn_n_params(fl_gnn *g) { return g ? g->g->n_params() : -1; }

int fl_gnn_get_params(fl_gnn *g, float *out) {
    return guard([&] {
        g->g->get_params(out);
        return 0;
    });
}

int fl_gnn_set_params(fl_gnn *g, const float *in) {
    return guard([&] {
        g->g->set_params(in);
        return 0;
    });
}

int fl_gnn_forward(fl_gnn *g, const double *K, float *phi) {
    return guard([&] {
        g->g->forward(K, phi);
        return 0;
    });
}

int fl_gnn_loss_grad(fl_gnn *g, const double *lambda, double *loss) {
    return guard([&] {
        double L = g->g->loss_backward(lambda);
        if (loss) *loss = L;
        return 0;
    });
}

int fl_gnn_zero_grad(fl_gnn *g) {
    return guard([&] {
        g->g->zero_grad();
        return 0;
    });
}

int fl_gnn_get_grad(fl_gnn *g, float *out) {
    return guard([&] {
        g->g->get_grad(out);
        return 0;
    });
}

int fl_gnn_adam_step(fl_gnn *g, double lr) {
    return guard([&] {
        g->g->adam_step(lr);
        return 0;
    });
}

int fl_gnn_train(fl_gnn *g, const double *K, const double *lambda, int n, int epochs, int batch, double lr,
                 double *epoch_loss) {
    return guard([&] {
        if (n < 1 || epochs < 1 || batch < 1) throw fl::Error(-1, "bad training sizes");
        const size_t E = g->g->n_elements(), F = g->g->n_edges();
        for (int ep = 0; ep < epochs; ep++) {
            double sum = 0.0;
            g->g->zero_grad();
            for (int i = 0; i < n; i++) {
                g->g->forward(K + (size_t)i * E, nullptr);
                sum += g->g->loss_backward(lambda + (size_t)i * F);
                if ((i + 1) % batch == 0 || i + 1 == n) {
                    g->g->adam_step(lr);
                    g->g->zero_grad();
                }
            }
            if (epoch_loss) epoch_loss[ep] = sum / n;
        }
        return 0;
    });
}

int fl_gnn_infer(fl_gnn *g, const double *K, int R, float *phi) {
    return guard([&] {
        g->g->infer(K, R, phi);
        return 0;
    });
}

int fl_gnn_features(fl_gnn *g, float *xe, float *xf) {
    return guard([&] {
        g->g->get_features(xe, xf);
        return 0;
    });
}

int fl_darcy_batch_autotune(fl_darcy_batch *b) {
    return guard([&] {
        b->g->autotune();
        return 0;
    });
}

int fl_transport_create(fl_mesh *m, int n_subst, const double *por, const double *cs, const double *bc_conc,
                        const double *init_conc, double end_time, double save_step, fl_transport **out) {
    return guard([&] {
        check_device();
        if (n_subst < 1 || n_subst > fl::TransportGPU::kMaxSubst) throw fl::Error(-5, "n_subst must be in 1..8");
        auto *t = new fl_transport();
        t->mesh = m;
        try {
            t->g.reset(new fl::TransportGPU(m->m, n_subst, por, cs, bc_conc, init_conc, end_time, save_step));
        } catch (...) {
            delete t;
            throw;
        }
        *out = t;
        return 0;
    });
}

void fl_transport_free(fl_transport *t) { delete t; }

int fl_transport_set_velocity(fl_transport *t, const double *side_flux) {
    return guard([&] {
        t->g->set_velocity_host(side_flux);
        return 0;
    });
}

int fl_transport_set_velocity_from_darcy(fl_transport *t, fl_darcy *d) {
    return guard([&] {
        FL_HIP(hipStreamSynchronize(d->g->stream()));
        t->g->set_velocity_device(d->g->device_solution());
        return 0;
    });
}

int fl_transport_prepare(fl_transport *t, double *cfl) {
    return guard([&] {
        double c = t->g->prepare();
        if (cfl) *cfl = c;
        return 0;
    });
}

int fl_transport_update_solution(fl_transport *t, double target, int *steps) {
    return guard([&] {
        int s = t->g->update_solution(target);
        if (steps) *steps = s;
        return 0;
    });
}

int fl_transport_get_conc(fl_transport *t, double *out) {
    return guard([&] {
        t->g->get_conc(out);
        return 0;
    });
}

int fl_transport_time(fl_transport *t, double *time, double *dt, double *cfl) {
    if (time) *time = t->g->time();
    if (dt) *dt = t->g->dt();
    if (cfl) *cfl = t->g->cfl();
    return 0;
}

int fl_transport_time_governor(fl_transport *t, double *time, double *dt, int *is_end) {
    const fl::TimeGovernorR &g = t->g->tos();
    if (time) *time = g.time;
    if (dt) *dt = g.time_step;
    if (is_end) *is_end = g.is_end() ? 1 : 0;
    return 0;
}

int fl_transport_set_time_upper_constraint(fl_transport *t, double dt) {
    return guard([&] {
        if (!(dt > 0.0)) throw fl::Error(-1, "time step constraint must be positive");
        (void)t->g->tos().set_upper_constraint(dt);
        return 0;
    });
}

int fl_transport_planned_time(fl_transport *t, double *time) {
    if (!time) return -1;
    *time = t->g->tos().estimate_time();
    return 0;
}

int fl_transport_solved_time(fl_transport *t, double *time) {
    if (!time) return -1;
    *time = t->g->tos().time;
    return 0;
}

int fl_transport_choose_next_time(fl_transport *t) {
    (void)t->g->tos().fix_dt_until_mark();
    return 0;
}

int fl_transport_update_solution_step(fl_transport *t, int *steps) {
    return guard([&] {
        const int n = t->g->update_solution_step();
        if (steps) *steps = n;
        return 0;
    });
}

int fl_transport_mass_balance(fl_transport *t, double *bal, double *plus, double *minus, double *mass, double *src,
                              double *totals) {
    return guard([&] {
        t->g->mass_balance(bal, plus, minus, mass, src, totals);
        return 0;
    });
}

int fl_transport_write_mass_balance(fl_transport *t, const char *path, const char *const *subst_names) {
    const fl::HostMesh &m = t->mesh->m;
    const int ns = t->g->n_subst(), nbr = m.n_boundary_regions, nblk = m.n_bulk_regions;
    std::vector<double> bal((size_t)ns * nbr + 1), plus(bal.size()), minus(bal.size()),
        mass((size_t)ns * nblk + 1), src(mass.size()), tot((size_t)4 * ns);
    int rc = guard([&] {
        t->g->mass_balance(bal.data(), plus.data(), minus.data(), mass.data(), src.data(), tot.data());
        return 0;
    });
    if (rc) return rc;
    return fl_write_mass_balance(t->mesh, path, t->g->time(), t->g->balance_initial_time(), ns, subst_names,
                                 bal.data(), plus.data(), minus.data(), mass.data(), src.data(), tot.data());
}

int fl_transport_set_pade_reactions(fl_transport *t, int n_decays, const int *parent, const int *product_ptr,
                                    const int *products, const double *half_life, const double *branch_ratio,
                                    int nom_pol_deg, int den_pol_deg) {
    return guard([&] {
        t->g->set_pade_reactions(n_decays, parent, product_ptr, products, half_life, branch_ratio, nom_pol_deg,
                                 den_pol_deg);
        return 0;
    });
}

int fl_transport_set_bc_conc(fl_transport *t, const double *bc_conc) {
    return guard([&] {
        t->g->set_bc_conc(bc_conc);
        return 0;
    });
}

int fl_transport_set_linear_reactions(fl_transport *t, int n_decays, const int *parent, const int *product_ptr,
                                      const int *products, const double *half_life, const double *branch_ratio) {
    return guard([&] {
        t->g->set_linear_reactions(n_decays, parent, product_ptr, products, half_life, branch_ratio);
        return 0;
    });
}

int fl_transport_set_sources(fl_transport *t, const double *density, const double *sigma, const double *conc) {
    return guard([&] {
        t->g->set_sources(density, sigma, conc);
        return 0;
    });
}

int fl_transport_set_dual_porosity_sorption(fl_transport *t, int dual_porosity, int sorption, const double *por_imm,
                                            const double *alpha, const double *phi, const int *sorp_type,
                                            const double *sorp_coef0, const double *sorp_coef1) {
    return guard([&] {
        t->g->set_phases(dual_porosity != 0, sorption != 0, por_imm, alpha, phi, sorp_type, sorp_coef0, sorp_coef1);
        return 0;
    });
}

int fl_transport_set_reaction_sorption(fl_transport *t, int n, const int *subst, const int *sorption_type,
                                       const double *mult_coef, const double *second_param,
                                       const double *rock_density, const double *molar_mass,
                                       const double *solubility, double solvent_dens, int substeps) {
    return guard([&] {
        t->g->set_reaction_sorption(n, subst, sorption_type, mult_coef, second_param, rock_density, molar_mass,
                                    solubility, solvent_dens, substeps, false);
        return 0;
    });
}

int fl_transport_set_reaction_sorption_elementwise(fl_transport *t, int n, const int *subst, const int *sorption_type,
                                                   const double *mult_coef, const double *second_param,
                                                   const double *rock_density, const double *molar_mass,
                                                   const double *solubility, double solvent_dens) {
    return guard([&] {
        t->g->set_reaction_sorption(n, subst, sorption_type, mult_coef, second_param, rock_density, molar_mass,
                                    solubility, solvent_dens, 1, true);
        return 0;
    });
}

int fl_transport_get_reaction_sorbed(fl_transport *t, double *out) {
    return guard([&] {
        t->g->get_reaction_sorbed(out);
        return 0;
    });
}

int fl_transport_set_bc_conc_function(fl_transport *t, fl_bc_conc_fn fn, void *user) {
    return guard([&] {
        t->g->set_bc_function(fn, user);
        return 0;
    });
}

int fl_transport_get_conc_phase(fl_transport *t, int phase, double *out) {
    return guard([&] {
        t->g->get_conc_phase(phase, out);
        return 0;
    });
}

int fl_transport_nnz(fl_transport *t) { return t->g->nnz(); }

int fl_transport_matrix(fl_transport *t, int *ptr, int *col, double *val) {
    return guard([&] {
        t->g->get_matrix(ptr, col, val);
        return 0;
    });
}

int fl_transport_stats(fl_transport *t, double *ms, int *steps) {
    if (ms) *ms = t->g->stats().ms_steps;
    if (steps) *steps = t->g->stats().steps;
    return 0;
}

int fl_device_stream_bandwidth(size_t bytes, int reps, double *gbs) {
    return guard([&] {
        fl::stream_bandwidth(bytes, reps, gbs);
        return 0;
    });
}

}  // extern "C"
