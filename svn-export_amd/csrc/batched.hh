// Batched Monte-Carlo Darcy solver: R conductivity realizations on one mesh
// (same cross-sections, sources and boundary conditions), solved together.
// There is no counterpart in the reference; each realization is the solve of
// DarcyFlowMH_Steady::update_solution (src/flow/darcy_flow_mh.cc:375-433) with
// its own conductivity field.
//
// Matrix-free condensed operator.  With A_e(K) = A_e(1)/K, every quantity of
// the per-element condensation is a K-independent geometric factor times a
// power of K:
//    A^-1 = K inv0, a = K a0, s = K s0, A^-1 g = K Ag0,
//    S_e(K) x = K [C inv0 C] x + dlt.x - u (u^T x) / D,   u = [K c.a0 ; v],
//    D = K s0 + V (vb coefficients v do not depend on K).
// The K-independent factors of all elements form one table, shared by the
// realizations; per realization only K (one value per element) is read by the
// operator, instead of the ~14 matrix values per edge row an assembled batch
// of CSR matrices would stream.
//
// Layout: every per-edge vector is realization-interleaved, v[row * Rp + r]
// with Rp = R rounded up to 64.  One wavefront handles one edge row for 64
// realizations: the element table and the index lists are wave-uniform
// (scalar loads) and every vector or K access is one contiguous segment.
//
// CG: per-realization alpha, beta, stop test and iteration count.  Four
// kernels per iteration (operator + partial p.q, fold, update + partial r.z,
// r.r, fold); converged realizations drop out of every memory access.  A chunk
// of iterations (and, in FP32 mode, the FP64 reliable update that follows) is
// captured once into a hipGraph and replayed until every realization stops.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "darcy.hh"

namespace fl {

struct BatchStats {
    int n_real = 0, chunks = 0, reliable_updates = 0, fp64_finish = 0, max_iterations = 0, compactions = 0;
    int handovers = 0;              // realizations finished by the single-solve path
    long handover_iterations = 0;   // their iterations there
    double ms_handover = 0;         // wall time of the hand-over phase
    double ms_coarse_avg = 0;       // live samples: coarse fold + direction kernel per iteration
    double ms_coarse_setup = 0;     // coarse operators and their inverses, all refills
    // wall-time phases (ms from the start): queue empty, narrowed to 16 / 4 lanes,
    // batch loop ended, hand-over path finished
    double ms_queue_empty = 0, ms_narrow16 = 0, ms_narrow4 = 0, ms_loop_end = 0, ms_handover_end = 0;
    long total_iterations = 0;  // summed over realizations
    double ms_setup = 0, ms_cg = 0, ms_resolve = 0, ms_total = 0;
    double ms_op_avg = 0, ms_update_avg = 0;  // live event samples of the two main kernels
    int op_samples = 0;
    int grid = 0, xcd_map = 0, chunk = 0;     // launch shape used (autotuned)
    int tile_rows = 0;                        // two-level FP32 operator: rows per tile (autotuned)
    std::vector<int> iterations, converged;  // per realization
    std::vector<double> residual;             // ||b - S x|| / ||b|| (FP64, true residual at the end)
};

// per-realization scalars of the batched CG
struct BatchScal {
    double *alpha, *beta, *rz, *rz_prev, *rr, *tol2, *bnorm2, *rr_hist, *rr_ru;
    int *status, *it, *n_ru, *par, *it_ru, *active_count;  // par: buffer holding the current search direction
    int *zp;  // two-level path, per slot: bit 0 re-form the direction, bit 1 add alpha p to x
    double ru_drop2;  // FP32 path: reliable update once ||r||^2 fell below ru_drop2 times its value at the last one
    int nat;          // stop-test norm: 0 ||r||_2, 1 sqrt(r^T D^-1 r) (Jacobi natural norm); rr, tol2,
                      // bnorm2, rr_ru and rr_hist are squares in that norm
};

// Batched CG status per realization
enum BatchStatus { kRun = 0, kConverged = 1, kCandidate = 2, kFp64 = 3, kMaxIt = 4, kBreakdown = 5, kIdle = 6 };

// Inverse of np SPD matrices on the host (a[np][n][n], lower triangle read)
// through the coarse space's batched factorization kernels, FP32 out.
void coarse_invert_host(const double *a, int n, int np, float *inv, int *info);

class DarcyBatchGPU {
   public:
    // base: a configured single solver (set_data called); its mesh, fields and
    // boundary data are shared.  R realizations solved in `slots` concurrent
    // lanes (rounded up to 64; 0: min(R, 256)); a slot whose realization has
    // converged is refilled with the next one.
    DarcyBatchGPU(DarcyGPU &base, int R, int slots = 0);
    ~DarcyBatchGPU();
    // K [R][E] (host or device pointer)
    void set_conductivity(const double *K, bool device_ptr);
    // precision 0: FP64 CG; 1: FP32 CG with FP64 reliable updates (FP64 finish
    // for realizations whose FP32 residual stagnates)
    // norm: stop-test residual norm as DarcyGPU::solve
    BatchStats solve(double rtol, double atol, int max_it, int precision, int norm = 0);
    void get_solution(int r, double *out) const;  // [S+E+F] of realization r
    void diagnostics(int *status, int *n_ru) const;  // [R] each, either may be null
    const double *device_solution() const { return d_sol_; }  // [(S+E+F)][Rout] interleaved
    int n_real() const { return R_; }
    int padded() const { return Rout_; }
    int slots() const { return Rp_; }
    int n_rows() const { return F_; }
    long table_bytes() const { return tab_bytes_; }
    long contributions() const { return ncontrib_; }
    hipStream_t stream() const { return stream_; }
    DarcyGPU &base() { return base_; }
    // edge traces lambda of every realization into out[R][F] (device, edge order)
    void lambda_device(double *out) const;
    // launch shape: grid blocks per 64-realization chunk, XCD-aware row mapping,
    // iterations per captured chunk.  autotune() times the operator kernel
    // over candidate shapes and keeps the fastest.
    void set_launch(int grid, int xcd_map, int chunk);
    void autotune();
    // live-realization count at which the rest go to the single-solve path
    // (0: never; negative: 16, or 8 when the two-level tail runs narrowed)
    void set_handover(int n) { handover_ = n < 0 ? -1 : n; }
    // Two-level path: once the queue is empty and at most 16 (then 4)
    // realizations are live, continue them in the batch with 16 (4) lanes per
    // edge row instead of handing them over (default on)
    void set_narrow(bool on) { narrow_ = on; }
    // Two-level preconditioner: ncol 0 off (Jacobi), 1 one constant column per
    // aggregate, 2 constant + basis column phi per aggregate
    void set_coarse(int ncol);
    int coarse_columns() const { return ncol_; }
    int aggregates() const { return nagg_; }
    // phi [R][F] FP32 device array in edge order (normalised per aggregate on use)
    void set_coarse_basis(const float *d_phi_edges);
    // row order of the batched vectors (row -> edge), aggregate of each row
    const std::vector<int> &row_edges() const { return h_perm_; }
    const std::vector<int> &row_aggregate() const { return h_agg_; }

   private:
    void build();
    void alloc32();
    void tune();  // launch shape of the row-grid kernels
    void setup_system();  // table, rhs, Jacobi diagonal for the current K
    void free_graphs();
    void alloc_scal(BatchScal &sc);
    // move slot src[s] to slot s for every live array of the phase (0 FP64, 1 FP32)
    void permute_slots(const std::vector<int> &src, int phase);
    // (slot, realization) pairs to the device list and the slot flags
    void set_pairs(const std::vector<int> &pairs);

    DarcyGPU &base_;
    // single-solve copies of base_ for the concurrent hand-over workers (worker
    // w > 0 uses hand_solvers_[w - 1]; made by the worker at its first job)
    static constexpr int kHandWorkers = 4;
    std::vector<std::unique_ptr<DarcyGPU>> hand_solvers_;
    const DarcyDev &dv_;
    hipStream_t stream_ = nullptr;
    hipStream_t stream2_ = nullptr;  // the operator's general rows, beside its row tiles
    hipEvent_t ev_fork_ = nullptr, ev_join_ = nullptr;
    int R_, Rp_, Rout_, E_, S_, F_, nchunk_;  // Rp_: slots (multiple of 64); Rout_: R rounded up to 64
    long tab_bytes_ = 0, ncontrib_ = 0;
    std::vector<void *> owned_;
    // shared structure
    int *d_meta_ = nullptr, *d_op_off_ = nullptr, *d_rs_off_ = nullptr, *d_conn_off_ = nullptr, *d_conn_ = nullptr;
    int *d_row_ptr_ = nullptr, *d_row_c_ = nullptr, *d_perm_ = nullptr, *d_eperm_ = nullptr, *d_rec_ = nullptr;
    double *d_tab_ = nullptr, *d_rsT_ = nullptr;  // operator table (FP64), resolve table
    float *d_tab32_ = nullptr;
    long op_total_ = 0, rs_total_ = 0;
    int dim_start_[4] = {0, 0, 0, 0};
    // per realization
    double *d_K_ = nullptr, *d_b_ = nullptr, *d_dinv_ = nullptr;
    float *d_K32_ = nullptr;
    unsigned short *d_dinv32_ = nullptr, *d_sp16_ = nullptr;  // FP32 path: D^-1 as bfloat16, its spare
    double *d_x_ = nullptr, *d_r_ = nullptr, *d_z_ = nullptr, *d_p0_ = nullptr, *d_p1_ = nullptr, *d_q_ = nullptr;
    float *d_x32_ = nullptr, *d_r32_ = nullptr, *d_z32_ = nullptr, *d_p32a_ = nullptr, *d_p32b_ = nullptr,
          *d_q32_ = nullptr;
    double *d_sol_ = nullptr, *d_ext_ = nullptr;
    double *d_part_ = nullptr;  // 3 partial arrays of 2048 * Rp
    double *d_pre_ = nullptr;   // first-stage fold sums, 3 arrays of 32 * Rp
    double *d_cpart_ = nullptr; // coarse-solve sums, 3 x (row tiles) x Rp
    int lw_ = 64;               // lanes per row of the two-level iteration kernels (64, 16, 4)
    bool narrow_ = true;        // two-level path: narrow lanes for the tail instead of the hand-over
    // tiled FP32 operator: general rows (vb-coupled elements) and row tiles
    void build_op_tiles(const std::vector<int> &rec);
    // candidate row tilings of the FP32 operator; the one in use is copied to
    // d_thptr_ .. nseg_max_ below
    struct OplTiling {
        const int *hptr = nullptr, *halo = nullptr, *eptr = nullptr, *els = nullptr;
        const unsigned short *lrec = nullptr;
        int T = 0, ntiles = 0, nseg_max = 0;
    };
    std::vector<OplTiling> tilings_;
    int tiling_ = 0;
    bool build_tiling(const std::vector<int> &rec, int tmax);
    void use_tiling(int i);
    const int *d_gen_ = nullptr, *d_grec_ = nullptr, *d_thptr_ = nullptr, *d_thalo_ = nullptr, *d_teptr_ = nullptr, *d_tels_ = nullptr;
    const unsigned short *d_lrec_ = nullptr;
    int ngen_ = 0, gen_blocks_ = 0, ntiles_ = 0, tile_T_ = 128, nseg_max_ = 0;
    int *d_frec_ = nullptr;     // FP32 pair-row records (20 words per row)
    BatchScal sc_{}, sc2_{};
    int *h_count_ = nullptr;  // pinned
    // slot -> realization (-1: empty), live chunk count, spare buffers
    std::vector<int> rid_;
    double *d_Kall_ = nullptr, *d_var_ = nullptr;  // K [R][E] as given; difficulty proxy [R]
    double *d_Ksave_ = nullptr;                    // the base solver's K while it is borrowed
    int handover_ = -1;  // live realizations at or below which the single-solve path finishes them (-1: default)
    // two-level preconditioner
    void coarse_setup(const std::vector<int> &pairs);  // coarse inverses of the listed slots
    int ncol_ = 0, nagg_ = 0, nc_ = 0, nbmax_ = 0;
    bool coarse_ok_ = false;  // aggregation usable (neighbour counts within the kernel limit)
    bool have_phi_ = false;
    std::vector<int> h_perm_, h_agg_;
    std::vector<int> h_nbptr_, h_nbv_, h_lbmap_;  // neighbour aggregates (host copies)
    int *d_tptr_ = nullptr, *d_tagg_ = nullptr;    // coarse-block term lists: offsets per block, block aggregate
    int2 *d_terms_ = nullptr;                      // (row, e << 8 | i << 4 | j) per term
    int *d_cbeg_ = nullptr, *d_cblk_ = nullptr;    // term-list chunks: first term, block
    int *d_cfirst_ = nullptr;                      // first chunk of each block (+ end)
    int nasm_chunks_ = 0;
    double *d_Apart_ = nullptr;                    // chunk partial sums [chunk][ncol^2][Rp]
    int *d_agg_ = nullptr, *d_aptr_ = nullptr, *d_nbptr_ = nullptr, *d_nb_ = nullptr, *d_lbmap_ = nullptr;
    float *d_phiAll_ = nullptr, *d_ainv_ = nullptr, *d_spainv_ = nullptr;
    unsigned short *d_phi_ = nullptr, *d_spphi_ = nullptr;  // phi per slot, bfloat16 [F][Rp]
    double *d_Asp_ = nullptr, *d_rc_ = nullptr, *d_y_ = nullptr, *d_spy_ = nullptr;
    double *d_Apack_ = nullptr, *d_Wpack_ = nullptr;  // coarse matrices / their inverse factors, [Rp][ncp][ncp]
    double *d_py_ = nullptr;                           // coarse-solve tile products [Rp][nt][nt][64]
    int *d_info_ = nullptr;                            // per matrix: not positive definite
    int nt_ = 0;                                       // 64-row tiles of the padded coarse order
    int *d_pairs_ = nullptr, *d_lsel_ = nullptr;   // pair list [2 * slots], slot flags [slots]
    int npairs_ = 0;
    // per-realization results of the last solve
    std::vector<int> h_it_, h_status_, h_ru_;
    std::vector<double> h_rr_, h_bb_, h_tol_;
    int nch_ = 0;
    int *d_src_ = nullptr;
    double *d_sp64_ = nullptr, *d_spK_ = nullptr;
    float *d_sp32_ = nullptr, *d_spK32_ = nullptr;
    // launch shape
    int grid_ = 512, xcd_map_ = 0, chunk_ = 0;  // chunk 0: automatic
    bool tuned_ = false;
    hipGraphExec_t g64_ = nullptr, g32_ = nullptr;
    long long g64_key_ = -1, g32_key_ = -1;  // (iteration limit, live chunks) of the captured graphs
    std::vector<hipEvent_t> ev_;
    bool have_K_ = false;
};

}  // namespace fl
