// GPU explicit upwind transport (ConvectionTransport, src/transport/transport.cc)
// driven by a restated TimeGovernor (src/coupling/time_governor.cc:261-355).
//
//  * k_tm_build: one thread per element row computes its upwind entries
//    (create_transport_matrix_mpi, transport.cc:788-943) in the order the
//    reference inserts them; the sparsity pattern (column-sorted rows, as a
//    PETSc AIJ row) is built on the host once per mesh.
//  * k_tm_scale: tm := dt*A + I, or (tm - I)*dt'/dt + I on a time-step change
//    (transport.cc:628-648).
//  * k_tm_compact: the scaled matrix without its exact zeros (an upwind row
//    keeps only its inflow entries), compacted per group of 64 rows.
//  * k_step_c: c_new = tm * c_old + bcvcorr + sources for all substances at
//    once, one lane per row; concentrations are substance-interleaved [E][m] so
//    one gathered column feeds every substance and the matrix row is read once
//    per step.  The sum runs in the row's column order with no fused
//    multiply-add, which is the sequential MatMultAdd order
//    (transport.cc:668-683).  Each XCD works on one contiguous eighth of the
//    rows so the gathered columns stay in its L2.  The same kernel adds each
//    block's boundary mass flux and source term, times the step length, to the
//    block's running totals of the cumulative mass balance
//    (MassBalance::calculate, mass_balance.cc:79-239); they are folded over the
//    blocks in a fixed order at output times.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "mesh.hh"
#include "time_governor.hh"

namespace fl {

struct TransportStats {
    int steps = 0;
    double ms_steps = 0, ms_build = 0;
};

// Device view of the transport problem, passed to kernels by value.
struct TmDev {
    int E, nsub, NBC, has_src;
    const int *ptr, *col, *kind, *s1, *s2, *diag;
    const int *side_off, *side_edge, *side_bnd, *bnd_bcele, *edge_ptr, *edge_side;
    const int *el_vb_ptr, *el_vb, *vb_hside;
    const int *bl_ptr, *bl_side, *bl_bcele;  // representative boundary sides per element
    const double *meas, *cs, *por, *flux, *edge_flow, *bc_conc;
    const double *src_density, *src_sigma, *src_conc;  // [E][m] interleaved (when has_src)
    const double *react;                               // [m][m] reaction matrix R[from][to] (when has_react)
    int has_react;
    double *val, *bcv, *partial, *mb;
    // the scaled matrix without its exact zeros (upwind: outflow faces), in
    // column order per row: the first kEll entries column-major in
    // ell_val / ell_col [kEll][E] (padding: value 0, own column), the rest of a
    // longer row compacted inside its group of 64 rows, cval / ccol
    // [cst[i], cst[i] + cln[i] - kEll); cln[i] = the row's count
    double *ell_val, *cval;
    int *ell_col, *ccol, *cst;
    unsigned char *cln;
    int *bidx;          // per row: slot of its boundary vector entry (bl_ptr), -1 without boundary sides
    const int *bl_elem; // element of each representative boundary side
    int nbl;            // number of representative boundary sides
    unsigned *count;
};

class TransportGPU {
   public:
    static constexpr int kMaxSubst = 8;  // substances per handle
    // por, cs [E]; bc_conc [nsub][NBC]; init_conc [nsub][E] (host arrays)
    TransportGPU(const HostMesh &m, int nsub, const double *por, const double *cs, const double *bc_conc,
                 const double *init_conc, double end_time, double save_step);
    ~TransportGPU();
    // sources_density / sources_sigma / sources_conc, each [nsub][E] (transport.cc:501-544)
    void set_sources(const double *density, const double *sigma, const double *conc);
    // new boundary concentrations [nsub][NBC], read by the next step (a time-dependent
    // bc_conc field whose next data time has been reached)
    void set_bc_conc(const double *bc_conc);
    // boundary concentrations as a function of time (a time-dependent FieldFormula,
    // field_formula_impl.hh:70-122): fn(t, bc_conc[nsub][NBC], user) at every step's
    // start time, the data then used for that step (a non-zero return stops the
    // update before the step); fn null: fixed data again
    using BcFn = int (*)(double, double *, void *);
    void set_bc_function(BcFn fn, void *user) { bc_fn_ = fn; bc_user_ = user; }
    // Linear_reaction (linear_reaction.cc:144-220): decays i with parent[i],
    // products[pptr[i]..pptr[i+1]), half_life[i] and branch ratios; the matrix is
    // rebuilt from the step estimate at every update_solution and applied after
    // each convection step.  n_decays = 0 switches reactions off.
    void set_linear_reactions(int n_decays, const int *parent, const int *pptr, const int *products,
                              const double *half_life, const double *branch);
    // Pade_approximant (pade_approximant.cc:97-290): the same decays, the step matrix the
    // (nom_deg, den_deg) Pade approximant of exp(A dt)
    void set_pade_reactions(int n_decays, const int *parent, const int *pptr, const int *products,
                            const double *half_life, const double *branch, int nom_deg, int den_deg);
    // Dual porosity and sorption (dual_porosity / sorption_enable, transport.cc:1026-1160):
    // por_imm, phi [E]; alpha, sorp_type (0 none, 1 linear, 2 freundlich, 3 langmuir),
    // sorp_coef0, sorp_coef1 [nsub][E] (null where the model that reads them is off).
    // Applied after every convection step, before the decays, which then act on the
    // immobile phase too.  The immobile and sorbed phases start at zero.
    void set_phases(bool dual, bool sorption, const double *por_imm, const double *alpha, const double *phi,
                    const int *sorp_type, const double *sorp_coef0, const double *sorp_coef1);
    // Sorption of the reaction module (reaction/sorption.cc, isotherm.cc) after every step:
    // species k on substance subst[k]; the isotherm type (0 none, 1 linear, 2 freundlich,
    // 3 langmuir), mult_coef, second_param per bulk region (sub-index order) and species
    // [nreg][n] with rock_density [nreg], or (elementwise) per species and element
    // [n][E] with rock_density [E]; molar_mass, solubility [n] (<= 0: no limit).
    // Regions whose parameters and porosities are constant are tabulated, the rest
    // solved per element; no dual porosity.
    void set_reaction_sorption(int n, const int *subst, const int *type, const double *mult, const double *second,
                               const double *rock_density, const double *molar_mass, const double *solubility,
                               double solvent_dens, int substeps, bool elementwise);
    void get_reaction_sorbed(double *out) const;  // [n][E]
    // phase 0 mobile, 1 immobile, 2 mobile sorbed, 3 immobile sorbed (transport.h:58-63) [nsub][E]
    void get_conc_phase(int phase, double *out) const;
    void set_velocity_host(const double *side_flux);      // [S]
    void set_velocity_device(const double *d_side_flux);  // device [S] (e.g. the Darcy solution)
    double prepare();                                      // build the matrix now; returns the CFL step
    int update_solution(double target_time);               // steps taken
    // the operator-splitting equation's own governor (TransportOperatorSplitting
    // time_, transport_operator_splitting.cc:213): output marks every save_step
    TimeGovernorR &tos() { return tos_; }
    // TransportOperatorSplitting::update_solution (transport_operator_splitting.cc:244-287):
    // its governor's next time, then the convection's CFL steps up to it
    int update_solution_step();
    void get_conc(double *out) const;                      // [nsub][E]
    void mass_balance(double *bal, double *plus, double *minus, double *mass, double *src, double *totals);
    double time() const { return tg_.time; }
    double dt() const { return tg_.time_step; }
    double balance_initial_time() const { return mb_init_ ? tg_.time : mb_t0_; }
    int n_subst() const { return nsub_; }
    double cfl() const { return cfl_max_step_; }
    int nnz() const { return nnz_; }
    int nsub() const { return nsub_; }
    void get_matrix(int *ptr, int *col, double *val) const;
    const TransportStats &stats() const { return stats_; }
    hipStream_t stream() const { return stream_; }

   private:
    void build_matrix();
    void compact_matrix();
    void set_bc();
    void one_step();
    void mb_initial();
    void flush_flux();
    // the launches of one step (shared by one_step and the captured graphs)
    void launch_step(const double *cold, double *cnew, double dt, double tdiff, const double *plan);
    // runs of plain steps replayed from captured graphs (run_planned)
    static constexpr int kGraphSteps[3] = {64, 16, 4};
    struct StepGraph {
        hipGraphExec_t exec = nullptr;
        int steps = 0, parity = 0;
        std::vector<unsigned char> sig;
    };
    std::vector<StepGraph> graphs_;
    double *d_plan_ = nullptr;  // [kGraphSteps[0]][2] step length and time difference of a replayed run
    double *h_plan_ = nullptr;  // pinned: the planned steps of one update
    size_t h_plan_cap_ = 0;
    bool plain_steps_ready() const;
    std::vector<unsigned char> step_signature(int steps, int parity) const;
    int run_planned(double target);
    void free_step_graphs();
    int step_grid() const;
    int step_row_blocks() const;
    static constexpr int kMaxStepGrid = 16384;  // row blocks of the step kernel
    static constexpr int kMaxBndBlocks = 64;    // boundary-side blocks after them
    double *d_bc_conc_ = nullptr;
    std::vector<double> h_bc_next_;
    bool bc_pending_ = false;
    BcFn bc_fn_ = nullptr;
    void *bc_user_ = nullptr;
    // reaction sorption: species, their parameters per element [n][E] (rock
    // density [E]), made into isotherms at the first step
    int rs_n_ = 0, rs_substeps_ = 0;
    bool rs_elementwise_ = false, rs_built_ = false;
    double rs_solvent_ = 1.0;
    std::vector<int> rs_subst_, rs_type_;
    std::vector<double> rs_mult_, rs_second_, rs_rock_, rs_molar_, rs_sol_;
    void build_reaction_sorption();
    void check_sorption();
    const int *d_rs_subst_ = nullptr, *d_rs_el_par_ = nullptr;
    const double *d_rs_par_ = nullptr, *d_rs_table_ = nullptr;  // isotherms, their tables back to back
    double *d_rs_sorbed_ = nullptr;                                // [E][n]
    int *d_rs_err_ = nullptr;                                      // error bits, first failing element
    bool phases_on() const { return dual_ || sorp_; }
    bool dual_ = false, sorp_ = false;
    // per element [E][m] interleaved like the concentrations: immobile, mobile sorbed,
    // immobile sorbed concentrations; alpha, isotherm type and coefficients; per element
    // por_imm and phi
    double *d_imm_ = nullptr, *d_msorb_ = nullptr, *d_isorb_ = nullptr;
    const double *d_alpha_ = nullptr, *d_sc0_ = nullptr, *d_sc1_ = nullptr, *d_por_imm_ = nullptr, *d_phi_ = nullptr;
    const int *d_stype_ = nullptr;
    std::vector<double> h_por_imm_, h_phi_;
    void build_reaction_matrix(double dt);
    void define_reactions(int n_decays, const int *parent, const int *pptr, const int *products,
                          const double *half_life, const double *branch);
    bool react_on_ = false, pade_ = false;
    int pade_nom_ = 0, pade_den_ = 0;
    std::vector<int> r_parent_, r_pptr_, r_prod_;
    std::vector<double> r_hl_, r_branch_, h_react_;
    double *d_react_ = nullptr;
    double *d_acc_ = nullptr;  // per-block running boundary-flux / source totals of the step kernel
    bool acc_dirty_ = false;

    const HostMesh &m_;
    hipStream_t stream_ = nullptr;
    int E_, nsub_, nnz_ = 0, NBC_;
    std::vector<void *> owned_;
    TmDev t_{};
    double *d_c0_ = nullptr, *d_c1_ = nullptr, *d_flux_ = nullptr;
    bool cur_is_c0_ = true;
    // state
    TimeGovernorR tg_, tos_;
    double cfl_max_step_ = 0;
    bool matrix_built_ = false, setup_done_ = false, scaled_ = false, need_rescale_ = true, flux_set_ = false;
    bool mb_init_ = true;
    double last_time_ = -1;
    double mb_t0_ = 0;  // time of the first balance (MassBalance::initial_time)
    std::vector<double> h_meas_, h_cs_, h_por_, h_bc_conc_, initial_mass_;
    std::vector<double> h_src_density_, h_src_sigma_, h_src_conc_;
    std::vector<int> h_el_region_sub_;
    TransportStats stats_;
};

}  // namespace fl
