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
//     mass balance; sources, linear / Pade reactions, time-dependent
//     boundary data.
//   DarcyFlowMH_Unsteady / DarcyFlowLMH_Unsteady (darcy_flow_mh.cc:1601-1827)
//   FieldElementwise (src/fields/field_elementwise_impl.hh:67-156)
//   OldBcdInput (src/flow/old_bcd.cc:44-250)
//   OutputMSH (src/io/output_msh.cc:53-467), water / mass balance files
//     (darcy_flow_mh_output.cc:671-773, mass_balance.cc:242-354)
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
    if (rc < 0) throw GpuError(std::string(what) + ": " + fl_last_error());
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
    unsigned n_nodes() const { return counts()[0]; }
    unsigned n_elements() const { return counts()[1]; }
    unsigned n_bc_elements() const { return counts()[2]; }
    unsigned n_sides() const { return counts()[3]; }
    unsigned n_edges() const { return counts()[4]; }
    // bulk regions (RegionDB::bulk_size): the regions without the boundary flag
    unsigned n_bulk_regions() const {
        const int nr = counts()[7];
        std::vector<int> ids(nr), bnd(nr), sub(nr);
        std::vector<char> buf(64 * (size_t)(nr + 1));
        check(fl_mesh_regions(m_, ids.data(), bnd.data(), sub.data(), buf.data(), (int)buf.size()), "regions");
        unsigned n = 0;
        for (int b : bnd) n += b == 0;
        return n;
    }

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
    DarcyFlowMH_Steady(GpuMesh &mesh, const DarcyData &data, fl_solver_opts solver = {1e-7, 1e-9, 10000, 0, 0})
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
    // like the reference (solve.cc:454-458) an iteration limit is reported, not
    // fatal: last_solve().converged tells, the solution is the last iterate
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
    // DarcyFlowMHOutput::water_balance: one block appended to path
    void write_water_balance(const std::string &path, double time) {
        check(fl_darcy_write_water_balance(d_, path.c_str(), time), "water balance");
    }
    // DarcyFlowMHOutput::output fields (darcy_flow_mh_output.cc:261-325, 424-541):
    // node pressure [n_nodes], element pressure and piezometric head [E], centre velocity [3E]
    void output_fields(std::vector<double> &pressure_nodes, std::vector<double> &pressure,
                       std::vector<double> &piezo_head, std::vector<double> &velocity) {
        pressure_nodes.assign(mesh_.n_nodes(), 0.0);
        pressure.assign(n_elem_, 0.0);
        piezo_head.assign(n_elem_, 0.0);
        velocity.assign(3 * (size_t)n_elem_, 0.0);
        check(fl_darcy_node_pressure(d_, pressure_nodes.data()), "node pressure");
        check(fl_darcy_element_output(d_, pressure.data(), piezo_head.data(), velocity.data()), "element output");
    }

   protected:
    GpuMesh &mesh_;
    fl_darcy *d_ = nullptr;
    fl_solver_opts solver_;
    fl_solve_info last_{};
    std::vector<double> solution_;
    unsigned n_sides_ = 0, n_elem_ = 0;
};

// DarcyFlowMH_Unsteady (lumped = false) / DarcyFlowLMH_Unsteady (lumped = true):
// update_solution() advances by one implicit Euler step of dt.
class DarcyFlowUnsteady : public DarcyFlowMH_Steady {
   public:
    DarcyFlowUnsteady(GpuMesh &mesh, const DarcyData &data, bool lumped, const std::vector<double> &storativity,
                      const std::vector<double> &init_pressure, double dt,
                      fl_solver_opts solver = {1e-7, 1e-9, 10000, 0, 0})
        : DarcyFlowMH_Steady(mesh, data, solver), dt_(dt) {
        check(fl_darcy_set_unsteady(d_, lumped ? 1 : 0, storativity.data(), init_pressure.data()), "set_unsteady");
        check(fl_darcy_get_solution(d_, solution_.data()), "get_solution");
    }
    void set_dt(double dt) { dt_ = dt; }
    void update_solution() {
        check(fl_darcy_unsteady_step(d_, dt_, &solver_, &last_), "unsteady step");
        check(fl_darcy_get_solution(d_, solution_.data()), "get_solution");
        time_ += dt_;
    }
    double time() const { return time_; }

   private:
    double dt_, time_ = 0.0;
};

// FieldElementwise: values of one gmsh $ElementData field on the bulk or boundary elements.
class FieldElementwise {
   public:
    FieldElementwise(GpuMesh &mesh, const std::string &gmsh_file, const std::string &field_name, int n_comp = 1,
                     bool boundary = false) {
        check(fl_element_data_open(mesh.handle(), gmsh_file.c_str(), field_name.c_str(), n_comp, boundary ? 1 : 0,
                                   &f_),
              "element data");
        int info[3];
        check(fl_element_data_info(f_, info, nullptr, 0), "element data info");
        n_ = (size_t)info[0] * info[1];
    }
    ~FieldElementwise() { fl_element_data_free(f_); }
    FieldElementwise(const FieldElementwise &) = delete;
    FieldElementwise &operator=(const FieldElementwise &) = delete;
    bool set_time(double t) {
        int updated = 0;
        check(fl_element_data_set_time(f_, t, &updated), "element data set_time");
        return updated != 0;
    }
    std::vector<double> value() const {
        std::vector<double> v(n_);
        check(fl_element_data_values(f_, v.data()), "element data values");
        return v;
    }

   private:
    fl_element_data *f_ = nullptr;
    size_t n_ = 0;
};

// OldBcdInput: legacy .fbc (into DarcyData) and .tbc (bc_conc[n_subst][NBC]) files.
class OldBcdInput {
   public:
    explicit OldBcdInput(GpuMesh &mesh) : mesh_(mesh) {}
    ~OldBcdInput() {
        if (b_) fl_old_bcd_free(b_);
    }
    OldBcdInput(const OldBcdInput &) = delete;
    OldBcdInput &operator=(const OldBcdInput &) = delete;
    void read_flow(const std::string &path, DarcyData &d) {
        const unsigned nb = mesh_.n_bc_elements();
        d.bc_type.assign(nb, 0);
        d.bc_pressure.assign(nb, 0.0);
        d.bc_flux.assign(nb, 0.0);
        d.bc_robin_sigma.assign(nb, 0.0);
        if (b_) fl_old_bcd_free(b_), b_ = nullptr;
        check(fl_old_bcd_read_flow(mesh_.handle(), path.c_str(), d.bc_type.data(), d.bc_pressure.data(),
                                   d.bc_flux.data(), d.bc_robin_sigma.data(), &b_),
              "flow bcd");
    }
    std::vector<double> read_transport(const std::string &path, int n_subst) {
        std::vector<double> c((size_t)n_subst * mesh_.n_bc_elements());
        check(fl_old_bcd_read_transport(b_, path.c_str(), n_subst, c.data()), "transport bcd");
        return c;
    }

   private:
    GpuMesh &mesh_;
    fl_old_bcd *b_ = nullptr;
};

// OutputMSH: the mesh head once, then $NodeData / $ElementData sections (values at
// `precision` significant digits; the reference's stream default is 6).
class OutputMSH {
   public:
    OutputMSH(GpuMesh &mesh, const std::string &path, int precision = 6)
        : mesh_(mesh), path_(path), precision_(precision) {}
    void write_elem_data(const std::string &name, const std::string &units, const double *values, int n_comp,
                         int n_values, double time) {
        head();
        check(fl_msh_write_element_data(path_.c_str(), name.c_str(), units.c_str(), time, step_, n_comp, n_values,
                                        values, precision_),
              "msh element data");
    }
    void write_node_data(const std::string &name, const std::string &units, const double *values, int n_comp,
                         int n_values, double time) {
        head();
        check(fl_msh_write_node_data(path_.c_str(), name.c_str(), units.c_str(), time, step_, n_comp, n_values,
                                     values, precision_),
              "msh node data");
    }
    void next_step() { step_++; }

   private:
    void head() {
        if (!head_) {
            check(fl_msh_write_head(mesh_.handle(), path_.c_str()), "msh head");
            head_ = true;
        }
    }
    GpuMesh &mesh_;
    std::string path_;
    int precision_, step_ = 0;
    bool head_ = false;
};

class ConvectionTransport {
   public:
    ConvectionTransport(GpuMesh &mesh, int n_subst, const std::vector<double> &por_m,
                        const std::vector<double> &cross_section, const std::vector<double> &bc_conc,
                        const std::vector<double> &init_conc, double end_time, double save_step)
        : n_subst_(n_subst), n_elem_(mesh.n_elements()), n_bulk_(mesh.n_bulk_regions()) {
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
    // ConvectionTransport sources_density / sources_sigma / sources_conc, each [n_subst][E]
    void set_sources(const std::vector<double> &density, const std::vector<double> &sigma,
                     const std::vector<double> &conc) {
        check(fl_transport_set_sources(t_, density.data(), sigma.data(), conc.data()), "sources");
    }
    // dual_porosity / sorption_enable (transport.cc:1026-1160): por_imm, phi [E];
    // alpha, sorp_type, sorp_coef0, sorp_coef1 [n_subst][E] (empty where unused)
    void set_dual_porosity_sorption(bool dual_porosity, bool sorption, const std::vector<double> &por_imm,
                                    const std::vector<double> &alpha, const std::vector<double> &phi,
                                    const std::vector<int> &sorp_type, const std::vector<double> &sorp_coef0,
                                    const std::vector<double> &sorp_coef1) {
        auto p = [](const auto &v) { return v.empty() ? nullptr : v.data(); };
        check(fl_transport_set_dual_porosity_sorption(t_, dual_porosity, sorption, por_imm.data(), p(alpha), p(phi),
                                                      p(sorp_type), p(sorp_coef0), p(sorp_coef1)),
              "dual porosity / sorption");
    }
    // the reaction module's sorption (Sorption, sorption.cc:71-285): species on substances subst;
    // type, mult_coef, second_param [n_bulk_regions][n]; rock_density [n_bulk_regions];
    // molar_mass, solubility [n] (empty solubility: no limit)
    void set_reaction_sorption(const std::vector<int> &subst, const std::vector<int> &type,
                               const std::vector<double> &mult_coef, const std::vector<double> &second_param,
                               const std::vector<double> &rock_density, const std::vector<double> &molar_mass,
                               const std::vector<double> &solubility, double solvent_dens = 1.0,
                               int substeps = 1000) {
        const size_t n = subst.size(), nr = n_bulk_;
        if (type.size() != nr * n || mult_coef.size() != nr * n || second_param.size() != nr * n ||
            rock_density.size() != nr || molar_mass.size() != n || (!solubility.empty() && solubility.size() != n))
            throw std::invalid_argument("set_reaction_sorption: type, mult_coef, second_param need n_bulk_regions x "
                                        "species values, rock_density n_bulk_regions, molar_mass and solubility "
                                        "(or empty) one per species");
        check(fl_transport_set_reaction_sorption(t_, (int)subst.size(), subst.data(), type.data(), mult_coef.data(),
                                                 second_param.data(), rock_density.data(), molar_mass.data(),
                                                 solubility.empty() ? nullptr : solubility.data(), solvent_dens,
                                                 substeps),
              "reaction sorption");
    }
    // the same with element-wise parameter fields: type, mult_coef, second_param [n][E],
    // rock_density [E]; every element's isotherm solved directly (sorption.cc:236-260)
    void set_reaction_sorption_elementwise(const std::vector<int> &subst, const std::vector<int> &type,
                                           const std::vector<double> &mult_coef,
                                           const std::vector<double> &second_param,
                                           const std::vector<double> &rock_density,
                                           const std::vector<double> &molar_mass,
                                           const std::vector<double> &solubility, double solvent_dens = 1.0) {
        const size_t n = subst.size(), ne = n_elem_;
        if (type.size() != n * ne || mult_coef.size() != n * ne || second_param.size() != n * ne ||
            rock_density.size() != ne || molar_mass.size() != n || (!solubility.empty() && solubility.size() != n))
            throw std::invalid_argument("set_reaction_sorption_elementwise: type, mult_coef, second_param need "
                                        "species x elements values, rock_density one per element, molar_mass and "
                                        "solubility (or empty) one per species");
        check(fl_transport_set_reaction_sorption_elementwise(t_, (int)n, subst.data(), type.data(), mult_coef.data(),
                                                             second_param.data(), rock_density.data(),
                                                             molar_mass.data(),
                                                             solubility.empty() ? nullptr : solubility.data(),
                                                             solvent_dens),
              "reaction sorption");
    }
    // bc_conc as a function of time: fn(t, bc_conc[n_subst][NBC], user) at every step's start
    // (a non-zero return stops the update before that step)
    void set_bc_conc_function(fl_bc_conc_fn fn, void *user) {
        check(fl_transport_set_bc_conc_function(t_, fn, user), "bc function");
    }
    // get_concentration_matrix()[phase]: 0 mobile, 1 immobile, 2 mobile sorbed, 3 immobile sorbed
    std::vector<double> concentrations(int phase) const {
        std::vector<double> c((size_t)n_subst_ * n_elem_);
        check(fl_transport_get_conc_phase(t_, phase, c.data()), "conc phase");
        return c;
    }
    // Linear_reaction / Pade_approximant decays: parent[i], products in CSR form
    void set_reactions(const std::vector<int> &parent, const std::vector<int> &product_ptr,
                       const std::vector<int> &products, const std::vector<double> &half_life,
                       const std::vector<double> &branch_ratio, int pade_nom = -1, int pade_den = -1) {
        if (pade_nom < 0)
            check(fl_transport_set_linear_reactions(t_, (int)parent.size(), parent.data(), product_ptr.data(),
                                                    products.data(), half_life.data(), branch_ratio.data()),
                  "reactions");
        else
            check(fl_transport_set_pade_reactions(t_, (int)parent.size(), parent.data(), product_ptr.data(),
                                                  products.data(), half_life.data(), branch_ratio.data(), pade_nom,
                                                  pade_den),
                  "reactions");
    }
    // new boundary data from the next step on
    void set_bc_conc(const std::vector<double> &bc_conc) {
        check(fl_transport_set_bc_conc(t_, bc_conc.data()), "bc_conc");
    }
    // MassBalance::output: one block appended to path
    void write_mass_balance(const std::string &path, const std::vector<std::string> &names) {
        std::vector<const char *> p;
        for (const auto &n : names) p.push_back(n.c_str());
        check(fl_transport_write_mass_balance(t_, path.c_str(), p.data()), "mass balance");
    }

   private:
    fl_transport *t_ = nullptr;
    int n_subst_;
    unsigned n_elem_, n_bulk_;
};

}  // namespace fl
