// GPU explicit upwind transport (ConvectionTransport, src/transport/transport.cc)
// driven by a restated TimeGovernor (src/coupling/time_governor.cc:261-355).
//
//  * k_tm_build: one thread per element row computes its upwind entries
//    (create_transport_matrix_mpi, transport.cc:788-943) in the order the
//    reference inserts them; the sparsity pattern (column-sorted rows, as a
//    PETSc AIJ row) is built on the host once per mesh.
//  * k_tm_scale: tm := dt*A + I, or (tm - I)*dt'/dt + I on a time-step change
//    (transport.cc:628-648).
//  * k_step: c_new = tm * c_old + bcvcorr for all substances at once;
//    concentrations are substance-interleaved [E][m] so one gathered column
//    feeds every substance, the matrix row is read once per step.  The sum
//    runs in the row's column order with no fused multiply-add, which is the
//    sequential MatMultAdd order (transport.cc:676-681).  The same kernel
//    reduces the boundary mass flux used by the cumulative mass balance
//    (MassBalance::calculate, mass_balance.cc:79-239).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "mesh.hh"

namespace fl {

struct TimeGovernorR {
    double time = 0, end_time = 0, time_step = 0, last_time_step = 0, fixed_dt = 0, end_fixed = 0;
    double upper = 0, lower = 0, max_ts = 0, min_ts = 0;
    bool dt_fixed_now = false, dt_changed = true;
    int level = 0;
    std::vector<double> marks;
    void init(double end, double save_step);
    bool lt(double o) const;
    bool ge(double o) const;
    bool is_end() const;
    void add_mark(double t);
    double next_mark() const;
    double estimate_dt() const;
    int set_upper_constraint(double u);
    double fix_dt_until_mark();
    void next_time();
};

struct TransportStats {
    int steps = 0;
    double ms_steps = 0, ms_build = 0;
};

class TransportGPU {
   public:
    // por, cs [E]; bc_conc [nsub][NBC]; init_conc [nsub][E] (host arrays)
    TransportGPU(const HostMesh &m, int nsub, const double *por, const double *cs, const double *bc_conc,
                 const double *init_conc, double end_time, double save_step);
    ~TransportGPU();
    void set_velocity_host(const double *side_flux);       // [S]
    void set_velocity_device(const double *d_side_flux);   // device [S] (e.g. the Darcy solution)
    int update_solution(double target_time);               // steps taken
    double prepare();                                       // build the matrix now; returns the CFL step
    void get_conc(double *out) const;                      // [nsub][E]
    void mass_balance(double *bal, double *plus, double *minus, double *mass, double *totals);
    double time() const { return tg_.time; }
    double dt() const { return tg_.time_step; }
    double cfl() const { return cfl_max_step_; }
    int nnz() const { return nnz_; }
    int nsub() const { return nsub_; }
    void get_matrix(int *ptr, int *col, double *val) const;
    const TransportStats &stats() const { return stats_; }
    hipStream_t stream() const { return stream_; }
    // Capture the fixed-dt step sequence into hipGraphs (launch-overhead free).
    void set_use_graph(bool on) { use_graph_ = on; }

   private:
    void build_matrix();
    void set_bc();
    void one_step();
    void mb_initial();
    void flush_flux();
    int step_grid() const;
    static constexpr int kMaxStepGrid = 1024;
    bool pending_ = false, step_parity_ = false;
    double pending_tdiff_ = 0.0;

    const HostMesh &m_;
    hipStream_t stream_ = nullptr;
    int E_, nsub_, nnz_ = 0, NBC_;
    std::vector<void *> owned_;
    // device
    int *d_ptr_, *d_col_, *d_kind_, *d_s1_, *d_s2_, *d_diag_;
    int *d_side_off_, *d_side_edge_, *d_side_bnd_, *d_bnd_bcele_, *d_edge_ptr_, *d_edge_side_;
    int *d_el_vb_ptr_, *d_el_vb_, *d_vb_hside_, *d_bnd_rep_;
    double *d_meas_, *d_cs_, *d_por_, *d_flux_, *d_edge_flow_, *d_val_;
    double *d_bc_conc_, *d_bcv_, *d_c0_, *d_c1_, *d_partial_, *d_mb_;
    unsigned *d_count_;
    bool cur_is_c0_ = true;
    // state
    TimeGovernorR tg_;
    double cfl_max_step_ = 0;
    bool matrix_built_ = false, setup_done_ = false, scaled_ = false, need_rescale_ = true, flux_set_ = false;
    bool mb_init_ = true;
    double last_time_ = -1;
    std::vector<double> h_meas_, h_cs_, h_por_, h_bc_conc_, initial_mass_;
    std::vector<int> h_el_region_sub_, h_bnd_rep_;
    TransportStats stats_;
    bool use_graph_ = false;
};

}  // namespace fl
