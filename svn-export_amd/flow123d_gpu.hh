// C++ adapter over the C ABI (include/flow123d_gpu.h) with the method set of
// the reference's equation classes, so a Flow123d call site maps one to one:
//
//   DarcyFlowMH_Steady  (src/flow/darcy_flow_mh.hh:242-331)
//     update_solution(), get_solution_vector(double*&, unsigned&),
//     solution_precision(), side_flux / side_scalar / element_scalar
//     (MH_DofHandler, src/flow/mh_dofhandler.hh:35-43)
//   ConvectionTransport + TransportOperatorSplitting
//   (src/transport/transport.h:81-290, transport_operator_splitting.hh:37-144)
//     set_velocity_field(), update_solution(), concentration matrix,
//     mass balance.
//
// Errors: the reference terminates through xprintf(Err) (system.hh:74-109);
// here a failed call throws fl::GpuError carrying fl_last_error().
#pragma once
#include <stdexcept>
#include <string>
#include <vector>

#include "../include/flow123d_gpu.h"

namespace fl {

struct GpuError : std::runtime_error {
    explicit GpuError(const std::string &m) : std::runtime_error(m) {}
};

inline void check(int rc, const char *what) {
    if (rc != 0) throw GpuError(std::string(what) + ": " + fl_last_error());
}

class GpuMesh {
   public:
    explicit GpuMesh(const std::string &gmsh_path) { check(fl_mesh_read_gmsh(gmsh_path.c_str(), &m_), "mesh"); }
    ~GpuMesh() { fl_mesh_free(m_); }
    GpuMesh(const GpuMesh &) = delete;
    GpuMesh &operator=(const GpuMesh &) = delete;
    fl_mesh *handle() const { return m_; }
    // n_nodes, E, NBC, S, F, boundaries, vb, regions, file bc
    std::vector<int> counts() const {
        std::vector<int> c(9);
        check(fl_mesh_counts(m_, c.data()), "counts");
        return c;
    }
    unsigned n_elements() const { return counts()[1]; }
    unsigned n_bc_elements() const { return counts()[2]; }
    unsigned n_sides() const { return counts()[3]; }
    unsigned n_edges() const { return counts()[4]; }

   private:
    fl_mesh *m_ = nullptr;
};

// Field values as the reference reads them from EqData (darcy_flow_mh.cc:152-183).
struct DarcyData {
    std::vector<double> conductivity, cross_section, sigma, water_source_density;
    std::vector<double> anisotropy;  // 9 per element, empty = identity
    std::vector<int> bc_type;        // per boundary element: 0 none, 1 dirichlet, 2 neumann, 3 robin
    std::vector<double> bc_pressure, bc_flux, bc_robin_sigma;
};

class DarcyFlowMH_Steady {
   public:
    DarcyFlowMH_Steady(GpuMesh &mesh, const DarcyData &data, fl_solver_opts solver = {1e-7, 1e-9, 10000, 0})
        : mesh_(mesh), solver_(solver) {
        check(fl_darcy_create(mesh.handle(), &d_), "darcy create");
        check(fl_darcy_set_data(d_, data.conductivity.data(), data.cross_section.data(),
                                data.anisotropy.empty() ? nullptr : data.anisotropy.data(), data.sigma.data(),
                                data.water_source_density.data(), data.bc_type.data(), data.bc_pressure.data(),
                                data.bc_flux.data(), data.bc_robin_sigma.data()),
              "darcy data");
        auto c = mesh.counts();
        n_sides_ = c[3];
        n_elem_ = c[1];
        solution_.resize((size_t)c[3] + c[1] + c[4]);
    }
    ~DarcyFlowMH_Steady() { fl_darcy_free(d_); }

    // darcy_flow_mh.cc:375-433
    void update_solution() {
        check(fl_darcy_update_solution(d_, &solver_, &last_), "update_solution");
        check(fl_darcy_get_solution(d_, solution_.data()), "get_solution");
    }
    // darcy_flow_mh.cc:485-500: borrowed pointer, valid until the next update_solution
    void get_solution_vector(double *&vec, unsigned &vec_size) {
        vec = solution_.data();
        vec_size = (unsigned)solution_.size();
    }
    double solution_precision() const { return fl_darcy_solution_precision(d_, solver_.r_tol, solver_.a_tol); }
    // MH_DofHandler accessors (mh_dofhandler.cc:37-50) by index
    double side_flux(unsigned global_side) const { return solution_[global_side]; }
    double element_scalar(unsigned elem) const { return solution_[n_sides_ + elem]; }
    double edge_scalar(unsigned edge) const { return solution_[(size_t)n_sides_ + n_elem_ + edge]; }
    const fl_solve_info &last_solve() const { return last_; }
    fl_darcy *handle() const { return d_; }

   private:
    GpuMesh &mesh_;
    fl_darcy *d_ = nullptr;
    fl_solver_opts solver_;
    fl_solve_info last_{};
    std::vector<double> solution_;
    unsigned n_sides_ = 0, n_elem_ = 0;
};

class ConvectionTransport {
   public:
    ConvectionTransport(GpuMesh &mesh, int n_subst, const std::vector<double> &por_m,
                        const std::vector<double> &cross_section, const std::vector<double> &bc_conc,
                        const std::vector<double> &init_conc, double end_time, double save_step)
        : n_subst_(n_subst), n_elem_(mesh.n_elements()) {
        check(fl_transport_create(mesh.handle(), n_subst, por_m.data(), cross_section.data(), bc_conc.data(),
                                  init_conc.data(), end_time, save_step, &t_),
              "transport create");
    }
    ~ConvectionTransport() { fl_transport_free(t_); }
    // TransportBase::set_velocity_field(const MH_DofHandler&)
    void set_velocity_field(DarcyFlowMH_Steady &darcy) {
        check(fl_transport_set_velocity_from_darcy(t_, darcy.handle()), "set_velocity_field");
    }
    // TransportOperatorSplitting::update_solution up to target_time
    int update_solution(double target_time) {
        int steps = 0;
        check(fl_transport_update_solution(t_, target_time, &steps), "transport update");
        return steps;
    }
    // get_concentration_matrix()[MOBILE][substance][element]
    std::vector<double> concentrations() const {
        std::vector<double> c((size_t)n_subst_ * n_elem_);
        check(fl_transport_get_conc(t_, c.data()), "conc");
        return c;
    }
    double time() const {
        double t = 0;
        fl_transport_time(t_, &t, nullptr, nullptr);
        return t;
    }

   private:
    fl_transport *t_ = nullptr;
    int n_subst_;
    unsigned n_elem_;
};

}  // namespace fl
