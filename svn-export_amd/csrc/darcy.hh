// GPU steady mixed-hybrid Darcy solver (DarcyFlowMH_Steady::update_solution,
// src/flow/darcy_flow_mh.cc:375-433, n_schurs = 2).
//
// Per solve:
//  1. k_assemble<D>: one thread per element of dimension D builds the RT0
//     local matrix, inverts it and condenses the element onto its connected
//     edges (own sides + vb neighbour edges), writing a dense local block and
//     a local right-hand side.  This is the per-element closed form of the
//     reference's two Schur stages (schur.cc:520-696 with
//     darcy_flow_mh.cc:1009-1131): for a = A^-1 1, s = 1^T a, V = sum of vb
//     coefficients, w = [c.*a ; v], D = s + V, h = |e| cs f - 1^T A^-1 g:
//        S_e = [C A^-1 C, 0; 0, diag(v)] + diag(Dirichlet/Robin terms) - w w^T / D
//        r_e = [C A^-1 g ; 0] - rhs_edge + w h / D
//  2. k_gather_sell / k_gather_rhs: each matrix entry / rhs entry sums its
//     precomputed list of block contributions in element order
//     (deterministic, no atomics), written straight into the SELL-64 layout
//     of the edge system (rows permuted by length within windows).
//  3. Jacobi-preconditioned CG on the SPD edge system (cg.hh): two fused
//     kernels per iteration, fence-free reductions, device-side stop test.
//  4. k_resolve<D>: element pressure and side fluxes from the edge traces
//     (SchurComplement::resolve, schur.cc:714-773), written in the
//     reference's [q | p | lambda] layout.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "cg.hh"
#include "mesh.hh"

namespace fl {

struct SolveStats {
    int iterations = 0;
    int converged = 0;
    double residual = 0, bnorm = 0;
    double ms_assembly = 0, ms_cg = 0, ms_resolve = 0, ms_total = 0;
    // live per-kernel durations sampled with HIP events on the solver stream
    double ms_spmv_avg = 0, ms_update_avg = 0;
    int kernel_samples = 0;
    int reliable_updates = 0;
    int fp64_iterations = 0;  // FP64 iterations after an FP32 stagnation
};

// Device view of the mesh + pattern + fields, passed to kernels by value.
struct DarcyDev {
    int E, S, F, NBC, nnz, nslots;
    int dim_start[4];
    const double *xyz;
    const int *el_nodes, *perm, *side_off, *side_edge, *side_bnd, *bnd_bcele;
    const double *bc_measure;
    const int *el_vb_ptr, *el_vb, *vb_edge, *vb_hel;
    const double *vb_hmeas;
    const int *conn_off, *conn, *blk_off;
    const int *rperm, *diag_slot, *cptr, *cidx, *rptr, *ridx;
    // fields
    const double *K, *cs, *sigma, *f, *aniso;
    const int *bc_type;
    const double *bc_p, *bc_flux, *bc_sigma;
    // work
    double *blk, *rhsc, *val, *rhs, *dinv, *sol;
    // unsteady models, null when steady: MH time coefficient S|e|/dt and previous
    // pressure per element; LMH time coefficient sum |e| S cs / dt / n_sides and
    // previous trace per edge (edge order)
    const double *tc, *pprev, *te, *lprev;
    // smallest element with a singular local matrix (INT_MAX: none), or null
    int *bad;
};

class DarcyGPU {
   public:
    explicit DarcyGPU(const HostMesh &m);
    ~DarcyGPU();
    // owns a stream, events and device buffers
    DarcyGPU(const DarcyGPU &) = delete;
    DarcyGPU &operator=(const DarcyGPU &) = delete;
    // kErrBadField unless every value of v[E] (host or device) is finite and > 0
    void check_field(const char *what, const double *v, bool device) const;
    // kErrSingularElement if an element's local matrix was singular (resets the flag)
    void check_singular();
    int *check_flag() const { return d_flag_ + 1; }  // one device int of scratch for field checks
    // per-element fields [E] (aniso: 9*E or null), per-bc-element BC data [NBC]
    void set_data(const double *K, const double *cs, const double *aniso, const double *sigma, const double *f,
                  const int *bc_type, const double *bc_p, const double *bc_flux, const double *bc_sigma);
    // the steady data of another handle on the same mesh but its conductivity
    // (device copies on this handle's stream; `o` is only read), which the
    // caller sets before the first solve
    void copy_data_from(const DarcyGPU &o);
    void set_conductivity(const double *K);  // host array [E]
    void set_conductivity_device(const double *dK);  // device array [E]
    const double *device_K() const { return d_K_; }
    // FP64 CG warm-started from the edge traces x0 (device, edge order); bnorm2
    // = ||b||^2 of the edge system (the stop test is relative to it)
    SolveStats solve_warm(const double *d_x0, double rtol, double atol, int max_it, double bnorm2, int norm = 0);
    // precision 0: FP64 CG; 1: FP32 CG with FP64 reliable updates.  norm: the
    // stop test's residual norm, 0 ||r||_2, 1 sqrt(r^T D^-1 r) (D = diag of the
    // edge system); the reported residual and rhs norm are in that norm
    SolveStats solve(double rtol, double atol, int max_it, int precision = 0, int norm = 0);
    // The FP64 solve with no allocation and no event during the call: the CG
    // partial sums and state live in a file-scope device array shared by all
    // handles on the GPU (calls that overlap on one GPU race); no timings.
    SolveStats solve_noalloc(double rtol, double atol, int max_it, int norm = 0);
    void get_solution(double *out) const;                   // [S+E+F]
    void get_edge_system(int *ptr, int *col, double *val, double *rhs) const;
    void element_balance(double *out) const;                // [E]
    // Unsteady models (darcy_flow_mh.cc:1601-1827): lumped 0 = MH (time term on the
    // element pressures), 1 = LMH (on the edge traces); the solution becomes the
    // initial state.  unsteady_step advances it by one implicit Euler step of dt.
    void set_unsteady(int lumped, const double *storativity, const double *init_pressure);
    bool is_unsteady() const { return unsteady_ != 0; }
    SolveStats unsteady_step(double dt, double rtol, double atol, int max_it, int precision = 0, int norm = 0);
    bool unsteady() const { return unsteady_ != 0; }
    void water_balance(double *bal, double *plus, double *minus) const;  // per boundary region
    void source_balance(double *src) const;                                // per bulk region
    // element pressure [E], piezometric head [E], centre velocity [3E] of the current solution
    void element_output(double *pressure, double *head, double *velocity) const;
    // pressure interpolated to the nodes [n_nodes] (make_node_scalar_param)
    void node_pressure(double *out);
    double solution_precision(double rtol, double atol) const;
    int nnz() const { return nnz_; }
    const HostMesh &mesh() const { return m_; }
    hipStream_t stream() const { return stream_; }
    const double *device_solution() const { return d_sol_; }
    // CG workspace access for mixed precision / batched solvers
    const DarcyDev &dev() const { return dv_; }

   private:
    void build_pattern();
    void assemble();
    int cg(double rtol, double atol, int max_it, SolveStats &st, bool warm = false, double bnorm2 = 0.0);
    int cg_mixed(double rtol, double atol, int max_it, int ru_every, SolveStats &st);
    void resolve();

    const HostMesh &m_;
    hipStream_t stream_ = nullptr;
    DarcyDev dv_{};
    int nnz_ = 0;
    long blk_total_ = 0;
    std::vector<void *> owned_;
    double *d_K_ = nullptr, *d_cs_ = nullptr, *d_sigma_ = nullptr, *d_f_ = nullptr, *d_aniso_ = nullptr;
    int *d_bc_type_ = nullptr;
    double *d_bc_p_ = nullptr, *d_bc_flux_ = nullptr, *d_bc_sigma_ = nullptr;
    double *d_x_ = nullptr, *d_r_ = nullptr, *d_z_ = nullptr, *d_p0_ = nullptr, *d_p1_ = nullptr, *d_q_ = nullptr;
    double *d_sol_ = nullptr;
    double *d_part_ = nullptr;  // 5 partial arrays of cg_grid_ doubles
    CGState *d_state_ = nullptr;
    int *d_flag_ = nullptr;
    int cg_grid_ = kCgGridLarge;  // set from the system's working set with the pattern
    bool cg_split_ = true;        // direction pass + SpMV with non-temporal matrix loads (beyond the cache)
    int stop_norm_ = 0;
    int device_ = 0;
    // CG partial sums and state of the current solve (null: the handle's own)
    double *work_part_ = nullptr;
    CGState *work_state_ = nullptr;
    bool sample_ = true;  // sample kernel durations with events
    int ru_every_ = 100;
    float *d_val32_ = nullptr, *d_dinv32_ = nullptr, *d_x32_ = nullptr, *d_r32_ = nullptr, *d_z32_ = nullptr;
    float *d_q32_ = nullptr, *d_p32a_ = nullptr, *d_p32b_ = nullptr;
    // SELL-64 edge system
    int nslices_ = 0;
    int *d_slice_ptr_ = nullptr, *d_sell_col_ = nullptr;
    std::vector<int> h_csr_ptr_, h_csr_col_, h_csr2slot_, h_rperm_;
    bool have_data_ = false;
    std::vector<hipEvent_t> ev_pool_;
    std::vector<double> h_bc_measure_;
    std::vector<double> h_cs_, h_f_, h_sigma_;
    int *d_inc_ptr_ = nullptr, *d_inc_nel_ = nullptr, *d_inc_src_ = nullptr;  // node incidences
    double *d_inc_dist_ = nullptr;
    // unsteady state
    void set_time_coefficients(double dt);
    int unsteady_ = 0;  // 1 MH, 2 LMH
    double dt_ = 0;
    std::vector<double> h_stor_, h_meas_;
    double *d_tc_ = nullptr, *d_pprev_ = nullptr, *d_te_ = nullptr, *d_lprev_ = nullptr, *d_stor_ = nullptr,
           *d_meas_ = nullptr;
};

}  // namespace fl
