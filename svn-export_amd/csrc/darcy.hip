// GPU steady MH Darcy solve.  See darcy.hh for the method and the reference
// functions each phase replaces.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>

#include "common.hh"
#include "darcy.hh"
#include "condense.hh"
#include "rt0_local.hh"

namespace fl {

template <int D>
__global__ void __launch_bounds__(256) k_assemble(const DarcyDev dv) {
    int idx = dv.dim_start[D - 1] + blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= dv.dim_start[D]) return;
    int e = dv.perm[idx];
    Local<D> L;
    element_local<D>(dv, e, dv.K[e], L);
    const int n0 = D + 1;
    int vb0 = dv.el_vb_ptr[e], nvb = dv.el_vb_ptr[e + 1] - vb0;
    double V = 0.0;
    for (int k = 0; k < nvb; k++) V += vb_coef(dv, e, dv.el_vb[vb0 + k]);
    double Dd = L.s + V, h = L.h;
    if (dv.tc) {  // MH time term on the element pressure row (darcy_flow_mh.cc:1641-1652, 1671-1673)
        Dd = Dd + dv.tc[e];
        h = h + dv.tc[e] * dv.pprev[e];
    }
    double w[n0];
    for (int i = 0; i < n0; i++) w[i] = L.c[i] * L.a[i];
    const int m = n0 + nvb;
    double *B = dv.blk + dv.blk_off[e];
    double *r = dv.rhsc + dv.conn_off[e];
    double hD = h / Dd;
    for (int i = 0; i < n0; i++) {
        for (int j = 0; j < n0; j++) {
            double v = L.c[i] * L.c[j] * L.inv[i * n0 + j] - w[i] * w[j] / Dd;
            if (i == j) v += L.dlt[i];
            B[i * m + j] = v;
        }
        for (int k = 0; k < nvb; k++) {
            double vk = vb_coef(dv, e, dv.el_vb[vb0 + k]);
            double v = -w[i] * vk / Dd;
            B[i * m + n0 + k] = v;
            B[(n0 + k) * m + i] = v;
        }
        r[i] = L.c[i] * L.Ag[i] + L.rf[i] + w[i] * hD;
    }
    for (int k = 0; k < nvb; k++) {
        double vk = vb_coef(dv, e, dv.el_vb[vb0 + k]);
        for (int l = 0; l < nvb; l++) {
            double vl = vb_coef(dv, e, dv.el_vb[vb0 + l]);
            double v = -vk * vl / Dd;
            if (k == l) v += vk;
            B[(n0 + k) * m + n0 + l] = v;
        }
        r[n0 + k] = vk * hD;
    }
}

// matrix values straight into their SELL slots (padding slots sum nothing)
__global__ void k_gather_sell(const DarcyDev dv) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < dv.nslots; k += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int j = dv.cptr[k]; j < dv.cptr[k + 1]; j++) s += dv.blk[dv.cidx[j]];
        dv.val[k] = s;
    }
}

// rhs and Jacobi diagonal in the permuted (SELL) row order
__global__ void k_gather_rhs(const DarcyDev dv) {
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < dv.F; f += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int j = dv.rptr[f]; j < dv.rptr[f + 1]; j++) s += dv.rhsc[dv.ridx[j]];
        if (dv.te) {  // LMH time term on the edge trace row (darcy_flow_mh.cc:1718-1727, 1762-1764)
            const int fe = dv.rperm[f];
            s += dv.te[fe] * dv.lprev[fe];
            dv.val[dv.diag_slot[f]] += dv.te[fe];
        }
        dv.rhs[f] = s;
        double d = dv.val[dv.diag_slot[f]];
        dv.dinv[f] = d != 0.0 ? 1.0 / d : 1.0;
    }
}

// ---------------------------------------------------------------------------
// Reconstruction (SchurComplement::resolve, schur.cc:714-773)
// ---------------------------------------------------------------------------
template <int D>
__global__ void __launch_bounds__(256) k_resolve(const DarcyDev dv, const double *__restrict__ lam) {
    int idx = dv.dim_start[D - 1] + blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= dv.dim_start[D]) return;
    int e = dv.perm[idx];
    Local<D> L;
    element_local<D>(dv, e, dv.K[e], L);
    const int n0 = D + 1;
    int so = dv.side_off[e];
    int vb0 = dv.el_vb_ptr[e], nvb = dv.el_vb_ptr[e + 1] - vb0;
    double V = 0.0, wl = 0.0;
    for (int k = 0; k < nvb; k++) {
        int n = dv.el_vb[vb0 + k];
        double vk = vb_coef(dv, e, n);
        V += vk;
        wl += vk * lam[dv.vb_edge[n]];
    }
    double Dd = L.s + V, h = L.h;
    if (dv.tc) {
        Dd = Dd + dv.tc[e];
        h = h + dv.tc[e] * dv.pprev[e];
    }
    double cl[n0];
    double own = 0.0;
    for (int i = 0; i < n0; i++) {
        cl[i] = L.c[i] * lam[dv.side_edge[so + i]];
        own += L.c[i] * L.a[i] * lam[dv.side_edge[so + i]];
    }
    double pe = (h + (own + wl)) / Dd;
    for (int i = 0; i < n0; i++) {
        double t = 0.0;
        for (int j = 0; j < n0; j++) t += L.inv[i * n0 + j] * cl[j];
        dv.sol[so + i] = L.Ag[i] + pe * L.a[i] - t;
    }
    dv.sol[dv.S + e] = pe;
}

// DarcyFlowMHOutput::make_element_scalar / make_element_vector
// (darcy_flow_mh_output.cc:261-325): element pressure, piezometric head p + z of
// the centre, and the Darcy velocity at the centre, sum_i q_i RT0_i(centre) / cs.
template <int D>
__global__ void __launch_bounds__(256) k_elem_output(const DarcyDev dv, double *__restrict__ pres,
                                                    double *__restrict__ head, double *__restrict__ vel) {
    int idx = dv.dim_start[D - 1] + blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= dv.dim_start[D]) return;
    int e = dv.perm[idx];
    Elem<D> el;
    for (int i = 0; i <= D; i++) {
        int n = dv.el_nodes[4 * e + i];
        el.P[i] = V3{dv.xyz[3 * n], dv.xyz[3 * n + 1], dv.xyz[3 * n + 2]};
    }
    V3 c = elem_centre<D>(el);
    const double p = dv.sol[dv.S + e];
    pres[e] = p;
    head[e] = p + c.z;
    const int so = dv.side_off[e];
    const double cs = dv.cs[e];
    double vx = 0.0, vy = 0.0, vz = 0.0;
    for (int i = 0; i <= D; i++) {
        V3 b = rt0_value<D>(el, c, i);
        const double q = dv.sol[so + i];
        vx += q * b.x / cs;
        vy += q * b.y / cs;
        vz += q * b.z / cs;
    }
    vel[3 * e] = vx;
    vel[3 * e + 1] = vy;
    vel[3 * e + 2] = vz;
}

// DarcyFlowMHOutput::make_node_scalar_param (darcy_flow_mh_output.cc:385-480):
// the pressure at a node is the average of the element pressures and side traces
// around it, weighted by (1 - d / sum d) / (count - 1), d the distance from the node
// to the element or side centre.  One lane per node gathers its incidences, the
// elements (in element order) then the sides (in side order), so the sums run in
// the reference's order.  inc_val[k] = d, inc_src[k] = solution index
// (S + e for an element, S + E + edge for a side); inc_ptr splits each node's list
// after its elements (nel[node]).
__global__ void k_node_pressure(int n_nodes, const int *__restrict__ inc_ptr, const int *__restrict__ nel,
                                const double *__restrict__ inc_dist, const int *__restrict__ inc_src,
                                const double *__restrict__ sol, double *__restrict__ out) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n_nodes; v += gridDim.x * blockDim.x) {
        const int k0 = inc_ptr[v], k1 = inc_ptr[v + 1], ke = k0 + nel[v];
        double se = 0.0, ss = 0.0;
        for (int k = k0; k < ke; k++) se += inc_dist[k];
        for (int k = ke; k < k1; k++) ss += inc_dist[k];
        const int cnt = k1 - k0;
        double r = 0.0;
        for (int k = k0; k < k1; k++) r += sol[inc_src[k]] * (1 - inc_dist[k] / (se + ss)) / (cnt - 1);
        out[v] = r;
    }
}

// DarcyFlowLMH_Unsteady::postprocess (darcy_flow_mh.cc:1772-1827): every side flux
// gets the lumped storage term of its edge, then the element's source share
__global__ void k_lmh_post(const DarcyDev dv, const double *__restrict__ meas, const double *__restrict__ stor,
                           double dt) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < dv.E; e += gridDim.x * blockDim.x) {
        const int so = dv.side_off[e], ns = dv.side_off[e + 1] - so;
        const double *lam = dv.sol + dv.S + dv.E;
        for (int i = 0; i < ns; i++) {
            const int f = dv.side_edge[so + i];
            const double time_coef = -meas[e] * dv.cs[e] * stor[e] / dt / ns;
            dv.sol[so + i] += time_coef * (lam[f] - dv.lprev[f]);
        }
        for (int i = 0; i < ns; i++) dv.sol[so + i] += -1.0 * meas[e] * dv.cs[e] * dv.f[e] / ns;
    }
}

// initial guess in edge order into the permuted CG vector
__global__ void k_gather_x0(const DarcyDev dv, const double *__restrict__ x0, double *__restrict__ x) {
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < dv.F; f += gridDim.x * blockDim.x) x[f] = x0[dv.rperm[f]];
}

// edge traces from the permuted CG solution into the [q | p | lambda] vector
__global__ void k_scatter_lambda(const DarcyDev dv, const double *__restrict__ x) {
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < dv.F; f += gridDim.x * blockDim.x)
        dv.sol[dv.S + dv.E + dv.rperm[f]] = x[f];
}

// smallest index i with v[i] not finite and positive into *flag (INT_MAX: none)
__global__ void k_first_not_positive(const double *__restrict__ v, long n, int *flag) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        if (!(v[i] > 0.0 && v[i] <= DBL_MAX)) atomicMin(flag, (int)std::min(i, (long)INT_MAX - 1));
}

__global__ void k_flag_reset(int *flag) { *flag = INT_MAX; }

long first_not_positive_device(const double *v, long n, int *flag, hipStream_t s) {
    hipLaunchKernelGGL(k_flag_reset, dim3(1), dim3(1), 0, s, flag);
    if (n > 0)
        hipLaunchKernelGGL(k_first_not_positive, dim3(std::min(grid_for(n), 4096)), dim3(kBlock), 0, s, v, n, flag);
    FL_HIP(hipGetLastError());
    int h = INT_MAX;
    FL_HIP(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    return h == INT_MAX ? -1 : (long)h;
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free solve (DarcyGPU::solve_noalloc): the five CG
// partial-sum arrays of up to kNoAllocGrid blocks, then the CG state.  One
// copy per GPU for the life of the process; every call on a GPU uses it, so
// calls that overlap on one GPU race.  Each call writes every partial and the
// state before reading them (k_cg_init, k_cg_setup), so nothing has to start
// at zero.
// ---------------------------------------------------------------------------
constexpr int kNoAllocGrid = 8192;
constexpr size_t kNoAllocDoubles = 5 * (size_t)kNoAllocGrid + 8;
static_assert(sizeof(CGState) <= 8 * sizeof(double), "CG state must fit the scratch tail");
static __device__ __align__(256) double scratch[kNoAllocDoubles];

static std::mutex g_scratch_mu;
static double *g_scratch_ptr[64];  // device address of `scratch`, per device

// Device address of `scratch` on `device`, looked up once per device with that
// device current (the array's name is only a host-side stand-in).
static double *scratch_address(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= 64) throw Error(-2, "device index out of range for the no-alloc scratch");
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        FL_HIP(hipGetDevice(&cur));
        if (cur != device) FL_HIP(hipSetDevice(device));
        double *p = nullptr;
        const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            throw Error(-2, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
template <class T>
static T *up(std::vector<void *> &own, const std::vector<T> &v, hipStream_t s) {
    T *p = dev_upload(v.data(), v.size(), s);
    own.push_back(p);
    return p;
}

DarcyGPU::DarcyGPU(const HostMesh &m) : m_(m) {
    FL_HIP(hipGetDevice(&device_));
    FL_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    build_pattern();
}

DarcyGPU::~DarcyGPU() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (hipEvent_t e : ev_pool_) (void)hipEventDestroy(e);
    for (void *p : owned_) dev_free(p);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void DarcyGPU::build_pattern() {
    const HostMesh &m = m_;
    const int E = m.n_elements(), S = m.n_sides(), F = m.n_edges(), V = m.n_vb(), NBC = m.n_bc_elements();
    // elements grouped by dimension
    std::vector<int> perm;
    perm.reserve(E);
    int ds[4] = {0, 0, 0, 0};
    for (int d = 1; d <= 3; d++) {
        for (int e = 0; e < E; e++)
            if (m.el_dim[e] == d) perm.push_back(e);
        ds[d] = (int)perm.size();
    }
    // connected edges per element: own sides then vb edges
    std::vector<int> conn_off(E + 1, 0), blk_off(E + 1, 0);
    long blk_total = 0;
    for (int e = 0; e < E; e++) {
        int mm = m.el_dim[e] + 1 + (m.el_vb_ptr[e + 1] - m.el_vb_ptr[e]);
        conn_off[e + 1] = conn_off[e] + mm;
        blk_off[e] = (int)blk_total;
        blk_total += (long)mm * mm;
    }
    if (blk_total > 0x7fffffffL) throw Error(-3, "element block storage exceeds 2^31 entries");
    blk_off[E] = (int)blk_total;
    blk_total_ = blk_total;
    std::vector<int> conn(conn_off[E]);
    for (int e = 0; e < E; e++) {
        int o = conn_off[e];
        for (int i = 0; i <= m.el_dim[e]; i++) conn[o++] = m.side_edge[m.side_off[e] + i];
        for (int k = m.el_vb_ptr[e]; k < m.el_vb_ptr[e + 1]; k++) conn[o++] = m.vb_edge[m.el_vb[k]];
    }
    // edge -> elements containing it in conn
    std::vector<int> ee_ptr(F + 1, 0);
    for (int i = 0; i < conn_off[E]; i++) ee_ptr[conn[i] + 1]++;
    for (int f = 0; f < F; f++) ee_ptr[f + 1] += ee_ptr[f];
    std::vector<int> ee(conn_off[E]), fillp(ee_ptr.begin(), ee_ptr.end() - 1);
    for (int e = 0; e < E; e++)
        for (int i = conn_off[e]; i < conn_off[e + 1]; i++) ee[fillp[conn[i]]++] = e;
    // CSR rows: union of connected edges of the elements touching the row
    std::vector<int> csr_ptr(F + 1, 0), csr_col;
    csr_col.reserve((size_t)conn_off[E] * 4);
    std::vector<int> tmp;
    for (int f = 0; f < F; f++) {
        tmp.clear();
        for (int k = ee_ptr[f]; k < ee_ptr[f + 1]; k++) {
            int e = ee[k];
            for (int i = conn_off[e]; i < conn_off[e + 1]; i++) tmp.push_back(conn[i]);
        }
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        csr_col.insert(csr_col.end(), tmp.begin(), tmp.end());
        csr_ptr[f + 1] = (int)csr_col.size();
    }
    nnz_ = (int)csr_col.size();
    std::vector<int> diag(F);
    for (int f = 0; f < F; f++) {
        auto b = csr_col.begin() + csr_ptr[f], en = csr_col.begin() + csr_ptr[f + 1];
        diag[f] = (int)(std::lower_bound(b, en, f) - csr_col.begin());
    }
    // block contributions per CSR entry (counting sort keeps element order)
    std::vector<int> ckey(blk_total);
    for (int e = 0; e < E; e++) {
        int mm = conn_off[e + 1] - conn_off[e];
        for (int a = 0; a < mm; a++) {
            int row = conn[conn_off[e] + a];
            auto b = csr_col.begin() + csr_ptr[row], en = csr_col.begin() + csr_ptr[row + 1];
            for (int bb = 0; bb < mm; bb++) {
                int colv = conn[conn_off[e] + bb];
                ckey[blk_off[e] + a * mm + bb] = (int)(std::lower_bound(b, en, colv) - csr_col.begin());
            }
        }
    }
    std::vector<int> cptr(nnz_ + 1, 0), cidx(blk_total);
    for (long i = 0; i < blk_total; i++) cptr[ckey[i] + 1]++;
    for (int k = 0; k < nnz_; k++) cptr[k + 1] += cptr[k];
    {
        std::vector<int> fp(cptr.begin(), cptr.end() - 1);
        for (long i = 0; i < blk_total; i++) cidx[fp[ckey[i]]++] = (int)i;
    }
    std::vector<int>().swap(ckey);
    // SELL-64 layout: rows sorted by decreasing length inside windows of 4096
    // rows (keeps the gather locality of the element ordering), 64-row slices
    // stored column-major, padding entries point at their own row.
    const int kWin = 4096;
    std::vector<int> rperm(F), rinv(F);
    {
        std::vector<int> idx;
        for (int w0 = 0; w0 < F; w0 += kWin) {
            int w1 = std::min(F, w0 + kWin);
            idx.resize(w1 - w0);
            std::iota(idx.begin(), idx.end(), w0);
            std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
                return csr_ptr[a + 1] - csr_ptr[a] > csr_ptr[b + 1] - csr_ptr[b];
            });
            std::copy(idx.begin(), idx.end(), rperm.begin() + w0);
        }
        for (int i = 0; i < F; i++) rinv[rperm[i]] = i;
    }
    nslices_ = (F + 63) / 64;
    std::vector<int> slice_ptr(nslices_ + 1, 0);
    for (int sl = 0; sl < nslices_; sl++) {
        int w = 0;
        for (int l = 0; l < 64 && sl * 64 + l < F; l++) {
            int o = rperm[sl * 64 + l];
            w = std::max(w, csr_ptr[o + 1] - csr_ptr[o]);
        }
        slice_ptr[sl + 1] = slice_ptr[sl] + 64 * w;
    }
    const int nslots = slice_ptr[nslices_];
    // an iteration's working set: the stored matrix entries (12 B) and the CG
    // vectors (56 B per row) against the last-level cache
    cg_split_ = 12.0 * nslots + 56.0 * F > kCgSplitBytes;
    cg_grid_ = cg_split_ ? kCgGridLarge : (F >= kCgTinyRows ? kCgGridSmall : kCgGridTiny);
    std::vector<int> sell_col(nslots), csr2slot(nnz_), slot_csr(nslots, -1);
    for (int sl = 0; sl < nslices_; sl++) {
        int w = (slice_ptr[sl + 1] - slice_ptr[sl]) / 64;
        for (int l = 0; l < 64; l++) {
            int row = sl * 64 + l;
            int o = row < F ? rperm[row] : -1;
            int len = o >= 0 ? csr_ptr[o + 1] - csr_ptr[o] : 0;
            for (int k = 0; k < w; k++) {
                int slot = slice_ptr[sl] + k * 64 + l;
                if (k < len) {
                    int pos = csr_ptr[o] + k;
                    sell_col[slot] = rinv[csr_col[pos]];
                    csr2slot[pos] = slot;
                    slot_csr[slot] = pos;
                } else {
                    sell_col[slot] = std::min(row, F - 1);
                }
            }
        }
    }
    std::vector<int> scptr(nslots + 1, 0), scidx(blk_total);
    for (int sl = 0; sl < nslots; sl++) {
        int pos = slot_csr[sl];
        scptr[sl + 1] = scptr[sl] + (pos >= 0 ? cptr[pos + 1] - cptr[pos] : 0);
        if (pos >= 0) std::copy(cidx.begin() + cptr[pos], cidx.begin() + cptr[pos + 1], scidx.begin() + scptr[sl]);
    }
    std::vector<int>().swap(cidx);
    std::vector<int>().swap(cptr);
    std::vector<int> diag_slot(F);
    for (int i = 0; i < F; i++) diag_slot[i] = csr2slot[diag[rperm[i]]];
    // rhs contributions per permuted row
    std::vector<int> ecount(F + 1, 0), rptr(F + 1, 0), ridx(conn_off[E]);
    for (int i = 0; i < conn_off[E]; i++) ecount[conn[i] + 1]++;
    for (int f = 0; f < F; f++) ecount[f + 1] += ecount[f];  // original-order offsets
    {
        std::vector<int> orig(conn_off[E]), fp(ecount.begin(), ecount.end() - 1);
        for (int i = 0; i < conn_off[E]; i++) orig[fp[conn[i]]++] = i;
        for (int nr = 0; nr < F; nr++) {
            int o = rperm[nr];
            int c = ecount[o + 1] - ecount[o];
            rptr[nr + 1] = rptr[nr] + c;
            std::copy(orig.begin() + ecount[o], orig.begin() + ecount[o + 1], ridx.begin() + rptr[nr]);
        }
    }
    h_csr_ptr_ = csr_ptr;
    h_csr_col_ = csr_col;
    h_csr2slot_ = csr2slot;
    h_rperm_ = rperm;
    // vb: higher-dim element and measure of its side on the vb edge
    std::vector<int> vb_hel(std::max(V, 1), 0);
    std::vector<double> vb_hmeas(std::max(V, 1), 0.0);
    for (int n = 0; n < V; n++) {
        int sd = m.edge_side[m.edge_ptr[m.vb_edge[n]]];  // Neighbour::side(), neighbours_impl.hh:16-20
        int he = m.side_elem[sd];
        vb_hel[n] = he;
        vb_hmeas[n] = side_measure(m, he, sd - m.side_off[he]);
    }
    h_bc_measure_.assign(std::max(NBC, 1), 0.0);
    for (int b = 0; b < NBC; b++) h_bc_measure_[b] = element_measure(m, m.bc_dim[b], &m.bc_nodes[4 * b]);

    hipStream_t s = stream_;
    DarcyDev &d = dv_;
    d.E = E; d.S = S; d.F = F; d.NBC = NBC; d.nnz = nnz_; d.nslots = nslots;
    for (int i = 0; i < 4; i++) d.dim_start[i] = ds[i];
    d.xyz = up(owned_, m.xyz, s);
    d.el_nodes = up(owned_, m.el_nodes, s);
    d.perm = up(owned_, perm, s);
    d.side_off = up(owned_, m.side_off, s);
    d.side_edge = up(owned_, m.side_edge, s);
    d.side_bnd = up(owned_, m.side_bnd, s);
    std::vector<int> bb = m.bnd_bcele;
    if (bb.empty()) bb.push_back(0);
    d.bnd_bcele = up(owned_, bb, s);
    d.bc_measure = up(owned_, h_bc_measure_, s);
    d.el_vb_ptr = up(owned_, m.el_vb_ptr, s);
    std::vector<int> elvb = m.el_vb, vbe = m.vb_edge;
    if (elvb.empty()) { elvb.push_back(0); vbe.push_back(0); }
    d.el_vb = up(owned_, elvb, s);
    d.vb_edge = up(owned_, vbe, s);
    d.vb_hel = up(owned_, vb_hel, s);
    d.vb_hmeas = up(owned_, vb_hmeas, s);
    d.conn_off = up(owned_, conn_off, s);
    d.conn = up(owned_, conn, s);
    d.blk_off = up(owned_, blk_off, s);
    d.rperm = up(owned_, rperm, s);
    d.diag_slot = up(owned_, diag_slot, s);
    d.cptr = up(owned_, scptr, s);
    d.cidx = up(owned_, scidx, s);
    d_slice_ptr_ = up(owned_, slice_ptr, s);
    d_sell_col_ = up(owned_, sell_col, s);
    d.rptr = up(owned_, rptr, s);
    d.ridx = up(owned_, ridx, s);
    auto alloc_d = [&](size_t n) {
        double *p = dev_alloc<double>(n);
        owned_.push_back(p);
        return p;
    };
    d.blk = alloc_d(blk_total);
    d.rhsc = alloc_d(conn_off[E]);
    d.val = alloc_d(nslots);
    d.rhs = alloc_d(F);
    d.dinv = alloc_d(F);
    d_sol_ = alloc_d((size_t)S + E + F);
    d.sol = d_sol_;
    d_x_ = alloc_d(F); d_r_ = alloc_d(F); d_z_ = alloc_d(F);
    d_p0_ = alloc_d(F); d_p1_ = alloc_d(F); d_q_ = alloc_d(F);
    d_part_ = alloc_d(5 * (size_t)cg_grid_);
    d_state_ = dev_alloc<CGState>(1);
    owned_.push_back(d_state_);
    d_flag_ = dev_alloc<int>(2);  // [0] singular-element flag, [1] field-check scratch
    owned_.push_back(d_flag_);
    hipLaunchKernelGGL(k_flag_reset, dim3(1), dim3(1), 0, s, d_flag_);
    d.bad = d_flag_;
    d_K_ = alloc_d(E); d_cs_ = alloc_d(E); d_sigma_ = alloc_d(E); d_f_ = alloc_d(E);
    d_bc_type_ = dev_alloc<int>(std::max(NBC, 1));
    owned_.push_back(d_bc_type_);
    d_bc_p_ = alloc_d(NBC); d_bc_flux_ = alloc_d(NBC); d_bc_sigma_ = alloc_d(NBC);
    d.K = d_K_; d.cs = d_cs_; d.sigma = d_sigma_; d.f = d_f_; d.aniso = nullptr;
    d.bc_type = d_bc_type_; d.bc_p = d_bc_p_; d.bc_flux = d_bc_flux_; d.bc_sigma = d_bc_sigma_;
    FL_HIP(hipStreamSynchronize(s));
}

void DarcyGPU::check_field(const char *what, const double *v, bool device) const {
    const int E = m_.n_elements();
    const long i = device ? first_not_positive_device(v, E, d_flag_ + 1, stream_) : first_not_positive(v, E);
    if (i < 0) return;
    double x = 0.0;
    if (device) FL_HIP(hipMemcpy(&x, v + i, sizeof(double), hipMemcpyDeviceToHost));
    else x = v[i];
    char buf[64];
    snprintf(buf, sizeof buf, "%.17g", x);
    throw Error(kErrBadField, std::string(what) + " of the element " + std::to_string(m_.el_id[i]) +
                                  " is not finite and positive: " + buf);
}

void DarcyGPU::check_singular() {
    int h = INT_MAX;
    FL_HIP(hipMemcpyAsync(&h, d_flag_, sizeof(int), hipMemcpyDeviceToHost, stream_));
    FL_HIP(hipStreamSynchronize(stream_));
    if (h == INT_MAX) return;
    hipLaunchKernelGGL(k_flag_reset, dim3(1), dim3(1), 0, stream_, d_flag_);
    FL_HIP(hipStreamSynchronize(stream_));
    throw Error(kErrSingularElement, "Singular local matrix of the element " + std::to_string(m_.el_id[h]) +
                                         " (|det| < NUM_ZERO or not finite)");
}

void DarcyGPU::set_data(const double *K, const double *cs, const double *aniso, const double *sigma, const double *f,
                        const int *bc_type, const double *bc_p, const double *bc_flux, const double *bc_sigma) {
    const int E = m_.n_elements(), NBC = m_.n_bc_elements();
    check_field("conductivity", K, false);
    check_field("cross_section", cs, false);
    hipStream_t s = stream_;
    FL_HIP(hipMemcpyAsync(d_K_, K, E * sizeof(double), hipMemcpyHostToDevice, s));
    FL_HIP(hipMemcpyAsync(d_cs_, cs, E * sizeof(double), hipMemcpyHostToDevice, s));
    FL_HIP(hipMemcpyAsync(d_sigma_, sigma, E * sizeof(double), hipMemcpyHostToDevice, s));
    FL_HIP(hipMemcpyAsync(d_f_, f, E * sizeof(double), hipMemcpyHostToDevice, s));
    h_cs_.assign(cs, cs + E);
    h_f_.assign(f, f + E);
    h_sigma_.assign(sigma, sigma + E);
    if (aniso) {
        if (!d_aniso_) {
            d_aniso_ = dev_alloc<double>(9 * (size_t)E);
            owned_.push_back(d_aniso_);
        }
        FL_HIP(hipMemcpyAsync(d_aniso_, aniso, 9 * (size_t)E * sizeof(double), hipMemcpyHostToDevice, s));
        dv_.aniso = d_aniso_;
    } else {
        dv_.aniso = nullptr;
    }
    if (NBC) {
        FL_HIP(hipMemcpyAsync(d_bc_type_, bc_type, NBC * sizeof(int), hipMemcpyHostToDevice, s));
        FL_HIP(hipMemcpyAsync(d_bc_p_, bc_p, NBC * sizeof(double), hipMemcpyHostToDevice, s));
        FL_HIP(hipMemcpyAsync(d_bc_flux_, bc_flux, NBC * sizeof(double), hipMemcpyHostToDevice, s));
        FL_HIP(hipMemcpyAsync(d_bc_sigma_, bc_sigma, NBC * sizeof(double), hipMemcpyHostToDevice, s));
    }
    FL_HIP(hipStreamSynchronize(s));
    have_data_ = true;
}

void DarcyGPU::copy_data_from(const DarcyGPU &o) {
    if (&o.m_ != &m_) throw Error(-3, "copy_data_from: the handles are on different meshes");
    if (!o.have_data_) throw Error(-4, "Darcy data not set");
    if (o.unsteady_) throw Error(-3, "copy_data_from: the source handle is unsteady");
    const size_t E = m_.n_elements(), NBC = m_.n_bc_elements();
    hipStream_t s = stream_;
    // K is not copied: `o` may be changing it; every solve sets its own
    FL_HIP(hipMemcpyAsync(d_cs_, o.d_cs_, E * sizeof(double), hipMemcpyDeviceToDevice, s));
    FL_HIP(hipMemcpyAsync(d_sigma_, o.d_sigma_, E * sizeof(double), hipMemcpyDeviceToDevice, s));
    FL_HIP(hipMemcpyAsync(d_f_, o.d_f_, E * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (o.dv_.aniso) {
        if (!d_aniso_) {
            d_aniso_ = dev_alloc<double>(9 * E);
            owned_.push_back(d_aniso_);
        }
        FL_HIP(hipMemcpyAsync(d_aniso_, o.d_aniso_, 9 * E * sizeof(double), hipMemcpyDeviceToDevice, s));
        dv_.aniso = d_aniso_;
    } else {
        dv_.aniso = nullptr;
    }
    if (NBC) {
        FL_HIP(hipMemcpyAsync(d_bc_type_, o.d_bc_type_, NBC * sizeof(int), hipMemcpyDeviceToDevice, s));
        FL_HIP(hipMemcpyAsync(d_bc_p_, o.d_bc_p_, NBC * sizeof(double), hipMemcpyDeviceToDevice, s));
        FL_HIP(hipMemcpyAsync(d_bc_flux_, o.d_bc_flux_, NBC * sizeof(double), hipMemcpyDeviceToDevice, s));
        FL_HIP(hipMemcpyAsync(d_bc_sigma_, o.d_bc_sigma_, NBC * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    FL_HIP(hipStreamSynchronize(s));
    h_cs_ = o.h_cs_;
    h_f_ = o.h_f_;
    h_sigma_ = o.h_sigma_;
    have_data_ = true;
}

void DarcyGPU::set_conductivity(const double *K) {
    check_field("conductivity", K, false);
    FL_HIP(hipMemcpyAsync(d_K_, K, m_.n_elements() * sizeof(double), hipMemcpyHostToDevice, stream_));
    FL_HIP(hipStreamSynchronize(stream_));
}

void DarcyGPU::set_conductivity_device(const double *dK) {
    check_field("conductivity", dK, true);
    FL_HIP(hipMemcpyAsync(d_K_, dK, m_.n_elements() * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    FL_HIP(hipStreamSynchronize(stream_));
}

SolveStats DarcyGPU::solve_warm(const double *d_x0, double rtol, double atol, int max_it, double bnorm2, int norm) {
    if (!have_data_) throw Error(-4, "Darcy data not set");
    stop_norm_ = norm;
    SolveStats st;
    auto t0 = std::chrono::steady_clock::now();
    assemble();
    hipLaunchKernelGGL(k_gather_x0, dim3(std::min(grid_for(dv_.F), 8192)), dim3(kBlock), 0, stream_, dv_, d_x0, d_x_);
    FL_HIP(hipGetLastError());
    cg(rtol, atol, max_it, st, true, bnorm2);
    resolve();
    FL_HIP(hipStreamSynchronize(stream_));
    st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return st;
}

void DarcyGPU::assemble() {
    const DarcyDev &d = dv_;
    for (int D = 1; D <= 3; D++) {
        int n = d.dim_start[D] - d.dim_start[D - 1];
        if (n <= 0) continue;
        dim3 g(grid_for(n)), b(kBlock);
        if (D == 1) hipLaunchKernelGGL(k_assemble<1>, g, b, 0, stream_, d);
        if (D == 2) hipLaunchKernelGGL(k_assemble<2>, g, b, 0, stream_, d);
        if (D == 3) hipLaunchKernelGGL(k_assemble<3>, g, b, 0, stream_, d);
    }
    int g1 = std::min(grid_for(d.nslots), 8192);
    hipLaunchKernelGGL(k_gather_sell, dim3(g1), dim3(kBlock), 0, stream_, d);
    int g2 = std::min(grid_for(d.F), 8192);
    hipLaunchKernelGGL(k_gather_rhs, dim3(g2), dim3(kBlock), 0, stream_, d);
    FL_HIP(hipGetLastError());
    check_singular();
}

int DarcyGPU::cg(double rtol, double atol, int max_it, SolveStats &st, bool warm, double bnorm2) {
    CGDev<double> a;
    a.n = dv_.F;
    a.nslices = nslices_;
    a.slice_ptr = d_slice_ptr_;
    a.col = d_sell_col_;
    a.val = dv_.val;
    a.dinv = dv_.dinv;
    a.b = dv_.rhs;
    a.x = d_x_; a.r = d_r_; a.z = d_z_; a.q = d_q_;
    const int G = cg_grid_;
    double *part = work_part_ ? work_part_ : d_part_;
    CGState *state = work_state_ ? work_state_ : d_state_;
    a.part_pq = part; a.part_rz0 = part + G; a.part_rz1 = part + 2 * G;
    a.part_rr0 = part + 3 * G; a.part_rr1 = part + 4 * G;
    a.st = state;
    if (warm) hipLaunchKernelGGL(k_cg_init_warm, dim3(G), dim3(kCgBlock), 0, stream_, a, d_p0_);
    else hipLaunchKernelGGL(k_cg_init<double>, dim3(G), dim3(kCgBlock), 0, stream_, a, d_p0_);
    hipLaunchKernelGGL(k_cg_setup<double>, dim3(1), dim3(kCgBlock), 0, stream_, a, G, rtol * rtol, atol * atol,
                       max_it, warm ? bnorm2 : 0.0, stop_norm_);
    double *pin = d_p0_, *pout = d_p1_;
    CGState h{};
    int launched = 0;
    int chunk = 32;
    // every kSampleEvery-th iteration is bracketed by events (direction + spmv, update)
    const int kSampleEvery = 8;
    struct Sample { int ev, it; };
    std::vector<Sample> samples;
    int ev_used = 0;
    auto next_ev = [&]() {
        if (ev_used == (int)ev_pool_.size()) {
            hipEvent_t e;
            FL_HIP(hipEventCreate(&e));
            ev_pool_.push_back(e);
        }
        return ev_used++;
    };
    while (true) {
        for (int k = 0; k < chunk; k++) {
            int itx = launched + k;
            bool samp = sample_ && (itx % kSampleEvery) == 1;
            int e0 = 0;
            if (samp) {
                e0 = next_ev(); next_ev(); next_ev();
                FL_HIP(hipEventRecord(ev_pool_[e0], stream_));
            }
            const int ru = itx == 0 ? 1 : 0;
            if (cg_split_) {
                hipLaunchKernelGGL(k_cg_dir<double>, dim3(G), dim3(kCgBlock), 0, stream_, a, itx, G,
                                   (const double *)pin, pout, ru, 0.0);
                hipLaunchKernelGGL((k_cg_spmv<double, true>), dim3(G), dim3(kCgBlock), 0, stream_, a,
                                   (const double *)pout);
            } else {
                hipLaunchKernelGGL((k_cg_dir_spmv<double, false>), dim3(G), dim3(kCgBlock), 0, stream_, a, itx, G,
                                   (const double *)pin, pout, ru, 0.0);
            }
            if (samp) FL_HIP(hipEventRecord(ev_pool_[e0 + 1], stream_));
            hipLaunchKernelGGL(k_cg_update<double>, dim3(G), dim3(kCgBlock), 0, stream_, a, itx, G,
                               (const double *)pout);
            if (samp) {
                FL_HIP(hipEventRecord(ev_pool_[e0 + 2], stream_));
                samples.push_back(Sample{e0, itx});
            }
            std::swap(pin, pout);
        }
        launched += chunk;
        FL_HIP(hipGetLastError());
        FL_HIP(hipMemcpyAsync(&h, state, sizeof(CGState), hipMemcpyDeviceToHost, stream_));
        FL_HIP(hipStreamSynchronize(stream_));
        if (h.done || launched > max_it + chunk) break;
        if (chunk < 256) chunk *= 2;
    }
    double s1 = 0.0, s2 = 0.0;
    int ns = 0;
    for (const Sample &sm : samples) {
        if (sm.it >= h.it) continue;  // launches after convergence exit early
        float t1 = 0, t2 = 0;
        FL_HIP(hipEventElapsedTime(&t1, ev_pool_[sm.ev], ev_pool_[sm.ev + 1]));
        FL_HIP(hipEventElapsedTime(&t2, ev_pool_[sm.ev + 1], ev_pool_[sm.ev + 2]));
        s1 += t1; s2 += t2; ns++;
    }
    st.kernel_samples = ns;
    st.ms_spmv_avg = ns ? s1 / ns : 0.0;
    st.ms_update_avg = ns ? s2 / ns : 0.0;
    st.iterations = h.it;
    st.residual = std::sqrt(h.rr);
    st.bnorm = std::sqrt(h.bnorm2);
    st.converged = (h.rr <= h.tol2) ? 1 : 0;
    return h.it;
}

int DarcyGPU::cg_mixed(double rtol, double atol, int max_it, int ru_every, SolveStats &st) {
    const int F = dv_.F, G = cg_grid_;
    if (!d_val32_) {
        auto fa = [&](size_t n) { float *p = dev_alloc<float>(n); owned_.push_back(p); return p; };
        d_val32_ = fa(dv_.nslots);
        d_dinv32_ = fa(F);
        d_x32_ = fa(F); d_r32_ = fa(F); d_z32_ = fa(F); d_q32_ = fa(F); d_p32a_ = fa(F); d_p32b_ = fa(F);
    }
    hipLaunchKernelGGL(k_to_float, dim3(std::min(grid_for(dv_.nslots), 8192)), dim3(kBlock), 0, stream_,
                       dv_.nslots, (const double *)dv_.val, d_val32_);
    hipLaunchKernelGGL(k_to_float, dim3(std::min(grid_for(F), 8192)), dim3(kBlock), 0, stream_, F,
                       (const double *)dv_.dinv, d_dinv32_);
    CGDev<float> a;
    a.n = F;
    a.nslices = nslices_;
    a.slice_ptr = d_slice_ptr_;
    a.col = d_sell_col_;
    a.val = d_val32_;
    a.dinv = d_dinv32_;
    a.b = nullptr;
    a.x = d_x32_; a.r = d_r32_; a.z = d_z32_; a.q = d_q32_;
    a.part_pq = d_part_; a.part_rz0 = d_part_ + G; a.part_rz1 = d_part_ + 2 * G;
    a.part_rr0 = d_part_ + 3 * G; a.part_rr1 = d_part_ + 4 * G;
    a.st = d_state_;
    const int gV = std::min(grid_for(F), 8192);
    FL_HIP(hipMemsetAsync(d_x_, 0, F * sizeof(double), stream_));
    FL_HIP(hipMemsetAsync(d_x32_, 0, F * sizeof(float), stream_));
    FL_HIP(hipMemsetAsync(d_p32a_, 0, F * sizeof(float), stream_));
    auto reliable_update = [&](int k) {
        hipLaunchKernelGGL(k_ru_accum, dim3(gV), dim3(kBlock), 0, stream_, F, d_x_, d_x32_);
        hipLaunchKernelGGL(k_ru_resid, dim3(G), dim3(kCgBlock), 0, stream_, a, (const double *)dv_.rhs,
                           (const double *)dv_.val, (const double *)d_x_, (k & 1) ? a.part_rz1 : a.part_rz0,
                           (k & 1) ? a.part_rr1 : a.part_rr0);
    };
    reliable_update(0);  // r = b, with the true FP64 norms
    hipLaunchKernelGGL(k_cg_setup<float>, dim3(1), dim3(kCgBlock), 0, stream_, a, G, rtol * rtol, atol * atol,
                       max_it, 0.0, stop_norm_);
    CGState h{};
    int k = 0, last_ru = 0, chunk = 32, n_ru = 1;
    const double delta2 = 0.01;  // true residual recomputed after every 10x drop
    const int kMaxUpdates = 400;
    std::vector<double> ru_hist;  // true residual norm^2 at each reliable update
    bool stagnated = false;
    const int kSampleEvery = 8;
    struct Sample { int ev, it; };
    std::vector<Sample> samples;
    int ev_used = 0;
    auto next_ev = [&]() {
        if (ev_used == (int)ev_pool_.size()) {
            hipEvent_t e;
            FL_HIP(hipEventCreate(&e));
            ev_pool_.push_back(e);
        }
        return ev_used++;
    };
    while (true) {
        for (int j = 0; j < chunk; j++, k++) {
            float *pin = (k & 1) ? d_p32b_ : d_p32a_, *pout = (k & 1) ? d_p32a_ : d_p32b_;
            bool samp = (k % kSampleEvery) == 1;
            int e0 = 0;
            if (samp) {
                e0 = next_ev(); next_ev(); next_ev();
                FL_HIP(hipEventRecord(ev_pool_[e0], stream_));
            }
            const int ru = k == last_ru ? 1 : 0;
            if (cg_split_) {
                hipLaunchKernelGGL(k_cg_dir<float>, dim3(G), dim3(kCgBlock), 0, stream_, a, k, G,
                                   (const float *)pin, pout, ru, delta2);
                hipLaunchKernelGGL((k_cg_spmv<float, true>), dim3(G), dim3(kCgBlock), 0, stream_, a,
                                   (const float *)pout);
            } else {
                hipLaunchKernelGGL((k_cg_dir_spmv<float, false>), dim3(G), dim3(kCgBlock), 0, stream_, a, k, G,
                                   (const float *)pin, pout, ru, delta2);
            }
            if (samp) FL_HIP(hipEventRecord(ev_pool_[e0 + 1], stream_));
            hipLaunchKernelGGL(k_cg_update<float>, dim3(G), dim3(kCgBlock), 0, stream_, a, k, G, (const float *)pout);
            if (samp) {
                FL_HIP(hipEventRecord(ev_pool_[e0 + 2], stream_));
                samples.push_back(Sample{e0, k});
            }
        }
        FL_HIP(hipGetLastError());
        FL_HIP(hipMemcpyAsync(&h, d_state_, sizeof(CGState), hipMemcpyDeviceToHost, stream_));
        FL_HIP(hipStreamSynchronize(stream_));
        if (h.done) {
            // converged on a true FP64 residual, or out of iterations
            if ((h.done == 1 && h.it == last_ru) || h.it >= max_it || n_ru >= kMaxUpdates) break;
            // paused (residual dropped 10x) or converged on the recursive residual:
            // fold x32 into x64 and restart the step from the true residual.
            // Stagnation of the true residual (FP32 attainable accuracy reached):
            // finish in FP64 from the current iterate.
            if (h.done == 1 || h.done == 2) {
                ru_hist.push_back(h.rr_ru);
                size_t nh = ru_hist.size();
                if (nh >= 4 && ru_hist[nh - 1] > 0.01 * ru_hist[nh - 4]) {
                    stagnated = true;
                    break;
                }
            }
            k = h.it;
            reliable_update(k);
            last_ru = k;
            n_ru++;
            hipLaunchKernelGGL(k_cg_clear_done, dim3(1), dim3(1), 0, stream_, d_state_);
            chunk = 32;
            continue;
        }
        if (k > max_it + chunk) break;
        if (chunk < 256) chunk *= 2;
    }
    hipLaunchKernelGGL(k_ru_accum, dim3(gV), dim3(kBlock), 0, stream_, F, d_x_, d_x32_);
    double s1 = 0.0, s2 = 0.0;
    int ns = 0;
    for (const Sample &sm : samples) {
        if (sm.it >= h.it) continue;
        float t1 = 0, t2 = 0;
        FL_HIP(hipEventElapsedTime(&t1, ev_pool_[sm.ev], ev_pool_[sm.ev + 1]));
        FL_HIP(hipEventElapsedTime(&t2, ev_pool_[sm.ev + 1], ev_pool_[sm.ev + 2]));
        s1 += t1; s2 += t2; ns++;
    }
    st.kernel_samples = ns;
    st.ms_spmv_avg = ns ? s1 / ns : 0.0;
    st.ms_update_avg = ns ? s2 / ns : 0.0;
    st.iterations = h.it;
    st.residual = std::sqrt(h.rr);
    st.bnorm = std::sqrt(h.bnorm2);
    st.converged = (h.rr <= h.tol2) ? 1 : 0;
    st.reliable_updates = n_ru;
    if ((stagnated || !st.converged) && h.it < max_it) {
        // FP64 CG warm-started from the FP64 iterate; the stop test keeps ||b||
        SolveStats s64;
        cg(rtol, atol, max_it - h.it, s64, true, h.bnorm2);
        st.iterations = h.it + s64.iterations;
        st.residual = s64.residual;
        st.converged = s64.converged;
        st.fp64_iterations = s64.iterations;
    }
    return st.iterations;
}

void DarcyGPU::resolve() {
    const DarcyDev &d = dv_;
    hipLaunchKernelGGL(k_scatter_lambda, dim3(std::min(grid_for(d.F), 8192)), dim3(kBlock), 0, stream_, d,
                       (const double *)d_x_);
    const double *lam = d_sol_ + d.S + d.E;
    for (int D = 1; D <= 3; D++) {
        int n = d.dim_start[D] - d.dim_start[D - 1];
        if (n <= 0) continue;
        dim3 g(grid_for(n)), b(kBlock);
        if (D == 1) hipLaunchKernelGGL(k_resolve<1>, g, b, 0, stream_, d, lam);
        if (D == 2) hipLaunchKernelGGL(k_resolve<2>, g, b, 0, stream_, d, lam);
        if (D == 3) hipLaunchKernelGGL(k_resolve<3>, g, b, 0, stream_, d, lam);
    }
    FL_HIP(hipGetLastError());
}

SolveStats DarcyGPU::solve(double rtol, double atol, int max_it, int precision, int norm) {
    if (!have_data_) throw Error(-4, "Darcy data not set");
    if (norm != 0 && norm != 1) throw Error(-1, "stop norm must be 0 (2-norm) or 1 (natural Jacobi norm)");
    stop_norm_ = norm;
    SolveStats st;
    hipEvent_t e0, e1, e2, e3;
    FL_HIP(hipEventCreate(&e0)); FL_HIP(hipEventCreate(&e1)); FL_HIP(hipEventCreate(&e2)); FL_HIP(hipEventCreate(&e3));
    FL_HIP(hipEventRecord(e0, stream_));
    assemble();
    FL_HIP(hipEventRecord(e1, stream_));
    if (precision == 1) cg_mixed(rtol, atol, max_it, ru_every_, st);
    else cg(rtol, atol, max_it, st);
    FL_HIP(hipEventRecord(e2, stream_));
    resolve();
    FL_HIP(hipEventRecord(e3, stream_));
    FL_HIP(hipEventSynchronize(e3));
    float t01, t12, t23;
    FL_HIP(hipEventElapsedTime(&t01, e0, e1));
    FL_HIP(hipEventElapsedTime(&t12, e1, e2));
    FL_HIP(hipEventElapsedTime(&t23, e2, e3));
    st.ms_assembly = t01; st.ms_cg = t12; st.ms_resolve = t23; st.ms_total = t01 + t12 + t23;
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); (void)hipEventDestroy(e2); (void)hipEventDestroy(e3);
    return st;
}

SolveStats DarcyGPU::solve_noalloc(double rtol, double atol, int max_it, int norm) {
    if (!have_data_) throw Error(-4, "Darcy data not set");
    if (norm != 0 && norm != 1) throw Error(-1, "stop norm must be 0 (2-norm) or 1 (natural Jacobi norm)");
    if (cg_grid_ > kNoAllocGrid)
        throw Error(-1, "CG grid " + std::to_string(cg_grid_) + " exceeds the no-alloc scratch (at most " +
                            std::to_string(kNoAllocGrid) + " blocks)");
    double *base = scratch_address(device_);
    stop_norm_ = norm;
    work_part_ = base;
    work_state_ = reinterpret_cast<CGState *>(base + 5 * (size_t)kNoAllocGrid);
    sample_ = false;
    SolveStats st;
    try {
        assemble();
        cg(rtol, atol, max_it, st);
        resolve();
        FL_HIP(hipStreamSynchronize(stream_));
    } catch (...) {
        work_part_ = nullptr;
        work_state_ = nullptr;
        sample_ = true;
        throw;
    }
    work_part_ = nullptr;
    work_state_ = nullptr;
    sample_ = true;
    return st;
}

void DarcyGPU::set_unsteady(int lumped, const double *storativity, const double *init_pressure) {
    if (!have_data_) throw Error(-4, "Darcy data not set");
    const HostMesh &m = m_;
    const int E = m.n_elements(), S = m.n_sides(), F = m.n_edges();
    unsteady_ = lumped ? 2 : 1;
    h_stor_.assign(storativity, storativity + E);
    h_meas_.resize(E);
    for (int e = 0; e < E; e++) h_meas_[e] = element_measure(m, m.el_dim[e], &m.el_nodes[4 * e]);
    if (!d_tc_) {
        d_tc_ = dev_alloc<double>(E); d_pprev_ = dev_alloc<double>(E); d_stor_ = dev_alloc<double>(E);
        d_meas_ = dev_alloc<double>(E); d_te_ = dev_alloc<double>(F); d_lprev_ = dev_alloc<double>(F);
        for (void *p : {(void *)d_tc_, (void *)d_pprev_, (void *)d_stor_, (void *)d_meas_, (void *)d_te_,
                        (void *)d_lprev_})
            owned_.push_back(p);
    }
    FL_HIP(hipMemcpyAsync(d_stor_, h_stor_.data(), E * sizeof(double), hipMemcpyHostToDevice, stream_));
    FL_HIP(hipMemcpyAsync(d_meas_, h_meas_.data(), E * sizeof(double), hipMemcpyHostToDevice, stream_));
    // initial condition (setup_time_term): element pressures (MH) or edge traces
    // sum of init / n_sides over the edge's sides (LMH); the rest zero
    std::vector<double> sol((size_t)S + E + F, 0.0);
    if (unsteady_ == 1) {
        for (int e = 0; e < E; e++) sol[S + e] = init_pressure[e];
    } else {
        for (int e = 0; e < E; e++) {
            const int ns = m.el_dim[e] + 1;
            for (int i = 0; i < ns; i++) sol[(size_t)S + E + m.side_edge[m.side_off[e] + i]] += init_pressure[e] / ns;
        }
    }
    FL_HIP(hipMemcpyAsync(d_sol_, sol.data(), sol.size() * sizeof(double), hipMemcpyHostToDevice, stream_));
    FL_HIP(hipStreamSynchronize(stream_));
    dt_ = 0;
    dv_.tc = dv_.pprev = dv_.te = dv_.lprev = nullptr;
}

// time coefficients for step dt, positive as the condensed system carries them:
// MH S|e|/dt per element (darcy_flow_mh.cc:1643-1644), LMH sum over an edge's
// sides of |e| S cs / dt / n_sides (:1722-1724)
void DarcyGPU::set_time_coefficients(double dt) {
    const HostMesh &m = m_;
    const int E = m.n_elements(), F = m.n_edges();
    if (unsteady_ == 1) {
        std::vector<double> tc(E);
        for (int e = 0; e < E; e++) tc[e] = h_stor_[e] * h_meas_[e] / dt;
        FL_HIP(hipMemcpyAsync(d_tc_, tc.data(), E * sizeof(double), hipMemcpyHostToDevice, stream_));
        FL_HIP(hipStreamSynchronize(stream_));
    } else {
        std::vector<double> te(F, 0.0);
        for (int e = 0; e < E; e++) {
            const int ns = m.el_dim[e] + 1;
            for (int i = 0; i < ns; i++)
                te[m.side_edge[m.side_off[e] + i]] += h_meas_[e] * h_stor_[e] * h_cs_[e] / dt / ns;
        }
        FL_HIP(hipMemcpyAsync(d_te_, te.data(), F * sizeof(double), hipMemcpyHostToDevice, stream_));
        FL_HIP(hipStreamSynchronize(stream_));
    }
    dt_ = dt;
}

// DarcyFlowMH::update_solution for the unsteady models: modify_system
// (darcy_flow_mh.cc:1660-1677 / 1751-1770), both Schur stages, the solve and,
// for LMH, postprocess (:1772-1827)
SolveStats DarcyGPU::unsteady_step(double dt, double rtol, double atol, int max_it, int precision, int norm) {
    if (!unsteady_) throw Error(-4, "unsteady model not set");
    if (!(dt > 0.0)) throw Error(-3, "time step must be positive");
    if (dt != dt_) set_time_coefficients(dt);
    const int S = m_.n_sides(), E = m_.n_elements(), F = m_.n_edges();
    if (unsteady_ == 1) {
        FL_HIP(hipMemcpyAsync(d_pprev_, d_sol_ + S, E * sizeof(double), hipMemcpyDeviceToDevice, stream_));
        dv_.tc = d_tc_; dv_.pprev = d_pprev_;
    } else {
        FL_HIP(hipMemcpyAsync(d_lprev_, d_sol_ + S + E, F * sizeof(double), hipMemcpyDeviceToDevice, stream_));
        dv_.te = d_te_; dv_.lprev = d_lprev_;
    }
    SolveStats st;
    try {
        st = solve(rtol, atol, max_it, precision, norm);
        if (unsteady_ == 2) {
            hipLaunchKernelGGL(k_lmh_post, dim3(std::min(grid_for(E), 4096)), dim3(kBlock), 0, stream_, dv_,
                               (const double *)d_meas_, (const double *)d_stor_, dt);
            FL_HIP(hipGetLastError());
            FL_HIP(hipStreamSynchronize(stream_));
        }
    } catch (...) {
        dv_.tc = dv_.pprev = dv_.te = dv_.lprev = nullptr;
        throw;
    }
    dv_.tc = dv_.pprev = dv_.te = dv_.lprev = nullptr;  // a later steady solve sees the steady system
    return st;
}

void DarcyGPU::element_output(double *pressure, double *head, double *velocity) const {
    const DarcyDev &d = dv_;
    const int E = m_.n_elements();
    double *buf = dev_alloc<double>((size_t)5 * E);
    for (int D = 1; D <= 3; D++) {
        int n = d.dim_start[D] - d.dim_start[D - 1];
        if (n <= 0) continue;
        dim3 g(grid_for(n)), b(kBlock);
        if (D == 1) hipLaunchKernelGGL(k_elem_output<1>, g, b, 0, stream_, d, buf, buf + E, buf + 2 * E);
        if (D == 2) hipLaunchKernelGGL(k_elem_output<2>, g, b, 0, stream_, d, buf, buf + E, buf + 2 * E);
        if (D == 3) hipLaunchKernelGGL(k_elem_output<3>, g, b, 0, stream_, d, buf, buf + E, buf + 2 * E);
    }
    std::vector<double> h((size_t)5 * E);
    hipError_t err = hipGetLastError();
    if (err == hipSuccess) err = hipMemcpyAsync(h.data(), buf, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream_);
    if (err == hipSuccess) err = hipStreamSynchronize(stream_);
    dev_free(buf);
    if (err != hipSuccess) throw Error(-2, std::string("element output: ") + hipGetErrorString(err));
    std::copy(h.begin(), h.begin() + E, pressure);
    std::copy(h.begin() + E, h.begin() + 2 * E, head);
    std::copy(h.begin() + 2 * E, h.end(), velocity);
}

void DarcyGPU::node_pressure(double *out) {
    const HostMesh &m = m_;
    const int N = m.n_nodes(), E = m.n_elements(), S = m.n_sides();
    if (!d_inc_ptr_) {
        // node incidences: elements in element order, then sides in side order
        auto dist = [&](int node, double cx, double cy, double cz) {
            const double *p = &m.xyz[3 * (size_t)node];
            return std::sqrt((p[0] - cx) * (p[0] - cx) + (p[1] - cy) * (p[1] - cy) + (p[2] - cz) * (p[2] - cz));
        };
        std::vector<std::vector<std::pair<double, int>>> el(N), sd(N);
        static const int kSN[4][4][3] = {{{0, 0, 0}}, {{0, 0, 0}, {1, 0, 0}}, {{0, 1, 0}, {1, 2, 0}, {0, 2, 0}},
                                         {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}}};
        for (int e = 0; e < E; e++) {
            const int d = m.el_dim[e];
            double c[3] = {0.0, 0.0, 0.0};
            for (int k = 0; k <= d; k++)
                for (int j = 0; j < 3; j++) c[j] += m.xyz[3 * (size_t)m.el_nodes[4 * e + k] + j];
            for (int j = 0; j < 3; j++) c[j] /= (d + 1);
            for (int k = 0; k <= d; k++) {
                const int v = m.el_nodes[4 * e + k];
                el[v].push_back({dist(v, c[0], c[1], c[2]), S + e});
            }
        }
        for (int e = 0; e < E; e++) {
            const int d = m.el_dim[e];
            for (int i = 0; i <= d; i++) {
                double c[3] = {0.0, 0.0, 0.0};
                for (int k = 0; k < d; k++)
                    for (int j = 0; j < 3; j++) c[j] += m.xyz[3 * (size_t)m.el_nodes[4 * e + kSN[d][i][k]] + j];
                for (int j = 0; j < 3; j++) c[j] /= d;
                const int src = S + E + m.side_edge[m.side_off[e] + i];
                for (int k = 0; k < d; k++) {
                    const int v = m.el_nodes[4 * e + kSN[d][i][k]];
                    sd[v].push_back({dist(v, c[0], c[1], c[2]), src});
                }
            }
        }
        std::vector<int> ptr(N + 1, 0), nel(N), src;
        std::vector<double> dd;
        for (int v = 0; v < N; v++) {
            nel[v] = (int)el[v].size();
            for (auto &x : el[v]) dd.push_back(x.first), src.push_back(x.second);
            for (auto &x : sd[v]) dd.push_back(x.first), src.push_back(x.second);
            ptr[v + 1] = (int)dd.size();
        }
        if (dd.empty()) dd.push_back(0.0), src.push_back(0);
        d_inc_ptr_ = dev_upload(ptr.data(), ptr.size(), stream_);
        d_inc_nel_ = dev_upload(nel.data(), nel.size(), stream_);
        d_inc_dist_ = dev_upload(dd.data(), dd.size(), stream_);
        d_inc_src_ = dev_upload(src.data(), src.size(), stream_);
        for (void *q : {(void *)d_inc_ptr_, (void *)d_inc_nel_, (void *)d_inc_dist_, (void *)d_inc_src_})
            owned_.push_back(q);
    }
    double *buf = dev_alloc<double>(N);
    hipLaunchKernelGGL(k_node_pressure, dim3(std::min(grid_for(N), 4096)), dim3(kBlock), 0, stream_, N,
                       (const int *)d_inc_ptr_, (const int *)d_inc_nel_, (const double *)d_inc_dist_,
                       (const int *)d_inc_src_, (const double *)d_sol_, buf);
    hipError_t err = hipGetLastError();
    if (err == hipSuccess) err = hipMemcpyAsync(out, buf, N * sizeof(double), hipMemcpyDeviceToHost, stream_);
    if (err == hipSuccess) err = hipStreamSynchronize(stream_);
    dev_free(buf);
    if (err != hipSuccess) throw Error(-2, std::string("node pressure: ") + hipGetErrorString(err));
}

void DarcyGPU::get_solution(double *out) const {
    size_t n = (size_t)m_.n_sides() + m_.n_elements() + m_.n_edges();
    FL_HIP(hipMemcpyAsync(out, d_sol_, n * sizeof(double), hipMemcpyDeviceToHost, stream_));
    FL_HIP(hipStreamSynchronize(stream_));
}

void DarcyGPU::get_edge_system(int *ptr, int *col, double *val, double *rhs) const {
    const int F = m_.n_edges();
    std::vector<double> sv(dv_.nslots), rp(F);
    FL_HIP(hipMemcpyAsync(sv.data(), dv_.val, sv.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    FL_HIP(hipMemcpyAsync(rp.data(), dv_.rhs, F * sizeof(double), hipMemcpyDeviceToHost, stream_));
    FL_HIP(hipStreamSynchronize(stream_));
    std::copy(h_csr_ptr_.begin(), h_csr_ptr_.end(), ptr);
    std::copy(h_csr_col_.begin(), h_csr_col_.end(), col);
    for (int k = 0; k < nnz_; k++) val[k] = sv[h_csr2slot_[k]];
    for (int i = 0; i < F; i++) rhs[h_rperm_[i]] = rp[i];
}

void DarcyGPU::element_balance(double *out) const {
    const HostMesh &m = m_;
    const int E = m.n_elements(), S = m.n_sides();
    std::vector<double> sol((size_t)S + E + m.n_edges());
    get_solution(sol.data());
    std::vector<double> vbm(m.n_vb());
    for (int n = 0; n < m.n_vb(); n++) {
        int sd = m.edge_side[m.edge_ptr[m.vb_edge[n]]];
        int he = m.side_elem[sd];
        vbm[n] = side_measure(m, he, sd - m.side_off[he]) * h_cs_[he];
    }
    for (int e = 0; e < E; e++) {
        double s = 0.0;
        for (int i = m.side_off[e]; i < m.side_off[e + 1]; i++) s += sol[i];
        double pe = sol[S + e];
        for (int k = m.el_vb_ptr[e]; k < m.el_vb_ptr[e + 1]; k++) {
            int n = m.el_vb[k];
            s += h_sigma_[e] * vbm[n] * (pe - sol[S + E + m.vb_edge[n]]);
        }
        out[e] = s - element_measure(m, m.el_dim[e], &m.el_nodes[4 * e]) * h_cs_[e] * h_f_[e];
    }
}

void DarcyGPU::water_balance(double *bal, double *plus, double *minus) const {
    const HostMesh &m = m_;
    std::vector<double> sol((size_t)m.n_sides() + m.n_elements() + m.n_edges());
    get_solution(sol.data());
    int nbr = m.n_boundary_regions;
    std::fill(bal, bal + nbr, 0.0);
    std::fill(plus, plus + nbr, 0.0);
    std::fill(minus, minus + nbr, 0.0);
    for (int b = 0; b < m.n_boundaries(); b++) {
        double flux = sol[m.edge_side[m.edge_ptr[m.bnd_edge[b]]]];
        int r = m.regions[m.bc_region[m.bnd_bcele[b]]].sub_idx;
        bal[r] += flux;
        if (flux > 0) plus[r] += flux;
        else minus[r] += flux;
    }
}

// DarcyFlowMHOutput::water_balance sources (darcy_flow_mh_output.cc:743-751):
// |e| cs f summed per bulk region
void DarcyGPU::source_balance(double *src) const {
    const HostMesh &m = m_;
    std::fill(src, src + m.n_bulk_regions, 0.0);
    if (h_cs_.empty()) return;
    for (int e = 0; e < m.n_elements(); e++)
        src[m.regions[m.el_region[e]].sub_idx] +=
            element_measure(m, m.el_dim[e], &m.el_nodes[4 * e]) * h_cs_[e] * h_f_[e];
}

double DarcyGPU::solution_precision(double rtol, double atol) const {
    // darcy_flow_mh.cc:463-482 uses the norm of the first Schur complement's
    // rhs; the condensed system's rhs norm is reported here.
    std::vector<double> rhs(m_.n_edges());  // permuted order; only the norm is used
    FL_HIP(hipMemcpyAsync(rhs.data(), dv_.rhs, rhs.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    FL_HIP(hipStreamSynchronize(stream_));
    double bn = 0.0;
    for (double v : rhs) bn += v * v;
    return std::max(atol, rtol * std::sqrt(bn));
}

}  // namespace fl
