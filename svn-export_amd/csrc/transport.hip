// GPU explicit upwind transport.  See transport.hh.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <limits>
#include <map>

#include "common.hh"
#include "transport.hh"

namespace fl {

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
enum : int { kDiag = 0, kSame = 1, kVbLow = 2, kVbHigh = 3 };

__global__ void k_edge_flow(const int F, const int *__restrict__ edge_ptr, const int *__restrict__ edge_side,
                            const double *__restrict__ flux, double *__restrict__ edge_flow) {
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < F; f += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int k = edge_ptr[f]; k < edge_ptr[f + 1]; k++) {
            double q = flux[edge_side[k]];
            if (q > 0) s += q;
        }
        edge_flow[f] = s;
    }
}

// transport.cc:841-909 for one row; diagonal accumulation in side order, then vb
__device__ double row_diag(const TmDev &t, int i) {
    double aii = 0.0;
    double mcp = t.meas[i] * t.cs[i] * t.por[i];
    for (int s = t.side_off[i]; s < t.side_off[i + 1]; s++) {
        double q = t.flux[s];
        if (q > 0.0) aii -= (q / mcp);
    }
    for (int k = t.el_vb_ptr[i]; k < t.el_vb_ptr[i + 1]; k++) {
        double q = t.flux[t.vb_hside[t.el_vb[k]]];
        if (q < 0.0) aii -= (-q) / mcp;
    }
    return aii;
}

__global__ void __launch_bounds__(256) k_tm_build(const TmDev t) {
    double mx = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < t.E; i += gridDim.x * blockDim.x) {
        for (int k = t.ptr[i]; k < t.ptr[i + 1]; k++) {
            int kind = t.kind[k];
            double a = 0.0;
            if (kind == kDiag) {
                a = row_diag(t, i);
                mx = fmax(mx, fabs(a));
            } else if (kind == kSame) {
                int si = t.s1[k], sj = t.s2[k];
                double q = t.flux[si], q2 = t.flux[sj];
                double ef = t.edge_flow[t.side_edge[si]];
                if (q2 > 0.0 && q < 0.0) a = -(q * q2 / (ef * t.meas[i] * t.cs[i] * t.por[i]));
                else a = 0;
            } else if (kind == kVbLow) {
                double q = t.flux[t.s1[k]];
                a = q > 0.0 ? q / (t.meas[i] * t.cs[i] * t.por[i]) : 0;
            } else {  // kVbHigh: row i is the higher-dimensional element
                double q = t.flux[t.s1[k]];
                a = q < 0.0 ? (-q) / (t.meas[i] * t.cs[i] * t.por[i]) : 0;
            }
            t.val[k] = a;
        }
    }
    // max |a_ii| over the grid (CFL, transport.cc:911-921); once per flux field
    double v[1] = {mx};
    if (grid_reduce<1, MaxOp>(v, t.partial, t.count) && threadIdx.x == 0) t.mb[0] = v[0];
}

// set_boundary_conditions (transport.cc:455-495) followed by VecScale(est).  Only
// rows with boundary sides can have a non-zero value; they keep it compactly at
// slot bl_ptr[i] (the row's first boundary-side entry), every other row's value
// is zero and is not stored.
__global__ void k_set_bc(const TmDev t, double est) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < t.E; i += gridDim.x * blockDim.x) {
        const int b0 = t.bl_ptr[i];
        if (t.bl_ptr[i + 1] == b0) continue;
        for (int s = 0; s < t.nsub; s++) {
            double acc = 0.0;
            for (int sd = t.side_off[i]; sd < t.side_off[i + 1]; sd++) {
                int b = t.side_bnd[sd];
                if (b < 0) continue;
                double q = t.flux[sd];
                if (q < 0.0) {
                    double aij = -(q / (t.meas[i] * t.cs[i] * t.por[i]));
                    acc += t.bc_conc[(size_t)s * t.NBC + t.bnd_bcele[b]] * aij;
                }
            }
            t.bcv[(size_t)b0 * t.nsub + s] = acc * est;
        }
    }
}

// first: tm = est*A + I ; else tm = (tm - I)*fac + I, bcv *= fac
__global__ void k_tm_scale(const TmDev t, int first, double fac) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < t.E; i += gridDim.x * blockDim.x) {
        int dk = t.diag[i];
        for (int k = t.ptr[i]; k < t.ptr[i + 1]; k++) {
            double v = t.val[k];
            if (first) v = v * fac;
            else v = (k == dk ? (v + -1.0) : v) * fac;
            if (k == dk) v = v + 1.0;
            t.val[k] = v;
        }
        if (!first && t.bl_ptr[i + 1] > t.bl_ptr[i])
            for (int s = 0; s < t.nsub; s++) t.bcv[(size_t)t.bl_ptr[i] * t.nsub + s] *= fac;
    }
}

// The scaled matrix without its exact zeros (an upwind row has entries only
// for its inflow faces; the others are inserted as zeros, transport.cc:851-866).
// Skipping a term 0 * c leaves every finite sum bitwise unchanged (up to the
// sign of an exact zero), so the step still reproduces the sequential
// MatMultAdd.  The first kEll non-zeros of each row go column-major to
// ell_val / ell_col (read coalesced, with no row-pointer load before them);
// the rest of a longer row is compacted inside its group of 64 rows: one wave
// per group, a wave prefix sum of the surplus counts places the rows one after
// another from the group's first CSR slot, so no grid-wide scan is needed.
constexpr int kEll = 4;
__global__ void __launch_bounds__(256) k_tm_compact(const TmDev t) {
    const int lane = threadIdx.x & 63;
    const int g = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (g * 64 >= t.E) return;  // wave-uniform
    const int i = g * 64 + lane;
    int n = 0, k0 = 0, k1 = 0;
    if (i < t.E) {
        k0 = t.ptr[i];
        k1 = t.ptr[i + 1];
        for (int k = k0; k < k1; k++) n += t.val[k] != 0.0 ? 1 : 0;
    }
    const int extra = n > kEll ? n - kEll : 0;
    int x = extra;  // inclusive prefix sum over the wave
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (i >= t.E) return;
    int w = t.ptr[g * 64] + x - extra;
    t.cst[i] = w;
    t.cln[i] = (unsigned char)n;
    int u = 0;
    for (int k = k0; k < k1; k++) {
        const double v = t.val[k];
        if (v == 0.0) continue;
        if (u < kEll) {
            t.ell_val[(size_t)u * t.E + i] = v;
            t.ell_col[(size_t)u * t.E + i] = t.col[k];
        } else {
            t.cval[w] = v;
            t.ccol[w] = t.col[k];
            w++;
        }
        u++;
    }
    for (; u < kEll; u++) {
        t.ell_val[(size_t)u * t.E + i] = 0.0;
        t.ell_col[(size_t)u * t.E + i] = i;
    }
}

// Folds one step's per-block partials (G blocks x 2M: boundary mass flux and
// source term per substance) in a fixed order and adds them, times that
// step's length, to the integrated quantities (mass_balance.cc:208-221).
// Executed by one block.
template <int M>
__device__ __forceinline__ void fold_flux(const TmDev &t, const double *__restrict__ buf, int G, double tdiff) {
    __shared__ double sh[2 * M][4];
    double acc[2 * M];
    for (int s = 0; s < 2 * M; s++) acc[s] = 0.0;
    for (int g = threadIdx.x; g < G; g += blockDim.x)
        for (int s = 0; s < 2 * M; s++) acc[s] += buf[(size_t)g * 2 * M + s];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int s = 0; s < 2 * M; s++) {
        double v = acc[s];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        if (lane == 0) sh[s][wid] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int s = 0; s < 2 * M; s++) t.mb[1 + s] += (((sh[s][0] + sh[s][1]) + sh[s][2]) + sh[s][3]) * tdiff;
    __syncthreads();
}

// ConvectionTransport::compute_sorption (transport.cc:1098-1160): the pore mass
// conc_avg split into dissolved concx and sorbed (conc_avg - concx) Nv; linear,
// Freundlich (at most 20 Newton steps from the kink cz) or Langmuir (at most 5).
__device__ inline void d_sorption(double conc_avg, double c0, double c1, int type, double &concx,
                                  double &concx_sorb, double Nv, double N) {
    const double Kx = c0 * N;
    const double ad = 1e4, tolerence = 1e-8;
    double pNR = 0, NR;
    if (type == 1) {
        concx = conc_avg / (1 + Kx);
    } else if (type == 2) {
        const double cz = pow(ad / (Kx * c1), 1 / (c1 - 1)), tcz = ad / c1;
        NR = cz;
        for (int i = 0; i < 20; i++) {
            NR -= (NR + ((NR > cz) ? Kx * pow(NR, c1) : tcz * NR) - conc_avg) /
                  (1 + ((NR > cz) ? c1 * Kx * pow(NR, c1 - 1) : tcz));
            if ((NR <= cz) || (fabs(NR - pNR) < tolerence * NR)) break;
            pNR = NR;
        }
        concx = NR;
    } else if (type == 3) {
        NR = 0;
        for (int i = 0; i < 5; i++) {
            NR -= (NR + (N * NR * c1 * c0) / (1 + NR * c0) - conc_avg) / (1 + N * c0 * c1 / pow((1 + NR * c0), 2));
            if (fabs(NR - pNR) < tolerence * NR) break;
            pNR = NR;
        }
        concx = NR;
    }
    concx_sorb = (conc_avg - concx) * Nv;
}

struct PhaseDev {
    double *imm, *msorb, *isorb;
    const double *alpha, *c0, *c1, *por_imm, *phi;
    const int *type;
    int dual, sorp;
};

// After the convection step, per element and substance in the reference's order
// (transport.cc:700-709): transport_dual_porosity (transport.cc:1026-1057), then
// transport_sorption (transport.cc:1061-1094); then the decays on the mobile and,
// with dual porosity, the immobile phase (linear_reaction.cc:295-301).  Element-
// wise and HBM-bound: every array is [E][m] so a wave reads whole rows.
template <int M>
__global__ void __launch_bounds__(256) k_phases(const TmDev t, const PhaseDev ph, double *__restrict__ c, double dt,
                                               const double *__restrict__ plan) {
    if (plan) dt = plan[0];  // a captured step reads its step length from the plan
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < t.E; i += gridDim.x * blockDim.x) {
        const double por_m = t.por[i], por_imm = ph.por_imm[i];
        double cm[M], ci[M];
        for (int s = 0; s < M; s++) {
            const size_t j = (size_t)i * M + s;
            cm[s] = c[j];
            ci[s] = ph.dual ? ph.imm[j] : 0.0;
            if (ph.dual) {
                const double pcm = cm[s], pci = ci[s];
                const double conc_avg = ((por_m * pcm) + (por_imm * pci)) / (por_m + por_imm);
                if ((conc_avg != 0.0) && (por_imm != 0.0)) {
                    const double f = exp(-ph.alpha[j] * ((por_m + por_imm) / (por_m * por_imm)) * dt);
                    cm[s] = (pcm - conc_avg) * f + conc_avg;
                    ci[s] = (pci - conc_avg) * f + conc_avg;
                }
            }
            if (ph.sorp) {
                const double phi = ph.phi[i];
                const double n = 1 - (por_m + por_imm), Nm = por_m, Nimm = por_imm;
                double ms = ph.msorb[j];
                const double conc_avg = cm[s] + ms * n / Nm;
                if (conc_avg != 0) d_sorption(conc_avg, ph.c0[j], ph.c1[j], ph.type[j], cm[s], ms, Nm / n, n * phi / Nm);
                ph.msorb[j] = ms;
                if (ph.dual && (por_imm != 0)) {
                    double is = ph.isorb[j];
                    const double conc_avg_imm = ci[s] + is * n / Nimm;
                    if (conc_avg_imm != 0)
                        d_sorption(conc_avg_imm, ph.c0[j], ph.c1[j], ph.type[j], ci[s], is, Nimm / n,
                                   n * (1 - phi) / Nimm);
                    ph.isorb[j] = is;
                }
            }
        }
        if (t.has_react) {
            double r[M];
            for (int q = 0; q < M; q++) {
                r[q] = 0.0;
                for (int s = 0; s < M; s++) r[q] += cm[s] * t.react[s * M + q];
            }
            for (int q = 0; q < M; q++) cm[q] = r[q];
            if (ph.dual) {
                for (int q = 0; q < M; q++) {
                    r[q] = 0.0;
                    for (int s = 0; s < M; s++) r[q] += ci[s] * t.react[s * M + q];
                }
                for (int q = 0; q < M; q++) ci[q] = r[q];
            }
        }
        for (int s = 0; s < M; s++) {
            c[(size_t)i * M + s] = cm[s];
            if (ph.dual) ph.imm[(size_t)i * M + s] = ci[s];
        }
    }
}

// Isotherm of the reaction module's sorption (isotherm.hh:100-170, isotherm.cc):
// type, c0 (mult_coef), c1 (second_param), rho (solvent density), sa (scale_aqua:
// the mobile porosity), ss (scale_sorbed: phi (1 - por_m - por_imm) rock density
// molar mass), limit (the solubility; 1 for a table made without one), limited,
// step (the table's total-mass step); tab = offset of the interpolation table and
// len its length, or tab = -1 for an isotherm solved directly per element.
struct IsoPar {
    int type, tab, len;
    double c0, c1, rho, sa, ss, limit, step;
    bool limited;
};
__host__ __device__ inline double iso_f(const IsoPar &I, double x) {
    if (I.type == 1) return I.c0 * x;
    if (I.type == 2) return I.c0 * pow(x, I.c1);
    if (I.type == 3) return I.c0 * (I.c1 * x) / (I.c1 * x + 1);
    return 0.0;
}
// Isotherm::solve_conc (isotherm.cc:86-109): the root c of ss f(c / rho) + sa c = total
// on [0, critical mass / sa + 1e-5], by bisection to the last bit (the reference uses a
// 30-bit TOMS 748 bracket); an end point that is a root is returned exactly.  False
// where toms748_solve throws: an empty interval, or ends of one sign (or not a number).
__host__ __device__ inline bool iso_solve(const IsoPar &I, double &c_aqua, double &c_sorbed) {
    const double total = I.sa * c_aqua + I.ss * c_sorbed;
    const double critic = I.limit * I.sa + iso_f(I, I.limit / I.rho) * I.ss;
    double a = 0.0, b = critic / I.sa + 0.00001;
    if (!(a < b)) return false;
    const double fa0 = I.ss * iso_f(I, a / I.rho) + I.sa * a - total;
    const double fb0 = I.ss * iso_f(I, b / I.rho) + I.sa * b - total;
    if (!(fabs(fa0) <= DBL_MAX) || !(fabs(fb0) <= DBL_MAX)) return false;
    if (fa0 == 0.0) b = a;
    else if (fb0 == 0.0) a = b;
    else if ((fa0 < 0) == (fb0 < 0)) return false;
    const bool neg_a = fa0 < 0;
    // up to 1200 halvings: enough to reach the last bit of any double in [0, b]
    // (a root near zero, as for Freundlich at tiny concentrations, needs hundreds)
    for (int it = 0; it < 1200 && a != b; it++) {
        const double m = 0.5 * (a + b);
        if (!(m > a && m < b)) break;
        const double fm = I.ss * iso_f(I, m / I.rho) + I.sa * m - total;
        if ((fm < 0) == neg_a) a = m;
        else b = m;
    }
    c_aqua = (a + b) / 2;
    c_sorbed = (total - I.sa * c_aqua) / I.ss;
    return true;
}

// Error bits of the sorption kernel (FL_E_SORPTION): a negative total mass
// (Isotherm::compute_projection, isotherm.cc:62) and an isotherm equation with
// no root in its bracket (toms748_solve throws).
constexpr int kSorpNegMass = 1, kSorpNoRoot = 2;

// Sorption::compute_one_step (sorption.cc:263-285) for every element and species.
// With a table, Isotherm::compute_projection (isotherm.cc:56-84): the total mass's
// place in the table, the rotated sorbed coordinate interpolated and rotated back;
// beyond the table the solubility limit or a direct solve.  Without one (parameters
// not constant on the region, sorption.cc:236-249), Isotherm::compute_reaction
// (isotherm.cc:184-200): the solubility limit when c_aqua exceeds it, else a
// direct solve.  el_par[i]: index of element i's first isotherm in par (its
// region's, or its own).  A failure leaves the element's values, sets err[0] and
// the lowest failing element in err[1].
template <int M>
__global__ void __launch_bounds__(256) k_rsorp(int E, int n, const int *__restrict__ subst,
                                              const int *__restrict__ el_par, const IsoPar *__restrict__ par,
                                              const double *__restrict__ tab, double *__restrict__ c,
                                              double *__restrict__ sorbed, int *__restrict__ err) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < E; i += gridDim.x * blockDim.x) {
        const int base = el_par[i];
        int fail = 0;
        for (int k = 0; k < n; k++) {
            const IsoPar I = par[base + k];
            double &ca = c[(size_t)i * M + subst[k]];
            double &cs = sorbed[(size_t)i * n + k];
            double c_aqua = ca, c_sorbed = cs;
            const double total = (I.sa * c_aqua + I.ss * c_sorbed);
            bool ok = true;
            if (I.tab >= 0) {
                const double *T = tab + I.tab;
                const double den = I.sa * I.sa + I.ss * I.ss;
                const double inv_a = I.sa / den, inv_s = I.ss / den;
                const double steps = total / I.step;
                const int idx = (int)floor(steps);
                if (idx < 0) {
                    fail |= kSorpNegMass;
                    ok = false;
                } else if (idx < I.len - 1) {
                    const double rot = T[idx] + (steps - idx) * (T[idx + 1] - T[idx]);
                    c_aqua = (total * inv_a - rot * inv_s);
                    c_sorbed = (total * inv_s + rot * inv_a);
                } else if (I.limited) {
                    c_aqua = I.limit;
                    c_sorbed = (total - I.sa * I.limit) / I.ss;
                } else {
                    ok = iso_solve(I, c_aqua, c_sorbed);
                    if (!ok) fail |= kSorpNoRoot;
                }
            } else if (I.limited && c_aqua > I.limit) {
                c_aqua = I.limit;
                c_sorbed = (total - I.sa * I.limit) / I.ss;
            } else {
                ok = iso_solve(I, c_aqua, c_sorbed);
                if (!ok) fail |= kSorpNoRoot;
            }
            if (ok) {
                ca = c_aqua;
                cs = c_sorbed;
            }
        }
        if (fail) {
            atomicOr(err, fail);
            atomicMin(err + 1, i);
        }
    }
}

// row of M doubles at p (16-byte aligned for even M) in 16-byte loads
template <int M>
__device__ __forceinline__ void load_row(const double *__restrict__ p, double (&v)[M]) {
    if constexpr (M % 2 == 0) {
        const double2 *q = reinterpret_cast<const double2 *>(p);
#pragma unroll
        for (int h = 0; h < M / 2; h++) {
            const double2 x = q[h];
            v[2 * h] = x.x;
            v[2 * h + 1] = x.y;
        }
    } else {
#pragma unroll
        for (int s = 0; s < M; s++) v[s] = p[s];
    }
}

// ConvectionTransport::compute_one_step (transport.cc:593-715) for all M
// substances: c_new = tm c_old + (bcvcorr + sources), then the reaction
// matrix when decays are on (linear_reaction.cc:167-183).  One lane per row
// carries the row's M concentrations: the row's matrix entries (k_tm_compact)
// are read once, each gathered column is one contiguous M * 8 B read that
// feeds every substance, and the sum runs in the row's column order with no
// fused multiply-add, as the sequential MatMultAdd (transport.cc:668-683).
// All loads of a row are issued before any is used.  The first `row_blocks`
// blocks do the rows, each XCD one contiguous eighth of them (block b runs on
// XCD b % 8), so neighbour gathers stay in its L2; the blocks after them
// reduce the boundary mass flux over the boundary sides (transport.cc:1299-1337),
// so the row path carries no boundary loop.  Each block adds its sums (flux,
// and with sources the source term, transport.cc:1339-1366), times the step
// length, to its own running totals acc[block][2M]; k_flux_acc folds them over
// the blocks at output times in a fixed order.
template <int M>
__global__ void __launch_bounds__(256) k_step_c(const TmDev t, const double *__restrict__ c_old,
                                               double *__restrict__ c_new, double dt, double tdiff,
                                               double *__restrict__ acc_buf, int row_blocks,
                                               const double *__restrict__ plan) {
    __shared__ double sh[2 * M][4];
    if (plan) {  // a captured step reads its step length and time difference from the plan
        dt = plan[0];
        tdiff = plan[1];
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if ((int)blockIdx.x >= row_blocks) {
        // boundary sides: inflow carries the boundary concentration, outflow the
        // element's concentration of the previous step
        double bfl[M];
#pragma unroll
        for (int s = 0; s < M; s++) bfl[s] = 0.0;
        const int nb = gridDim.x - row_blocks;
        for (int k = (blockIdx.x - row_blocks) * blockDim.x + threadIdx.x; k < t.nbl; k += nb * blockDim.x) {
            const double wf = t.flux[t.bl_side[k]];
            const int bce = t.bl_bcele[k];
            double ce[M];
            load_row<M>(c_old + (size_t)t.bl_elem[k] * M, ce);
#pragma unroll
            for (int s = 0; s < M; s++) bfl[s] += wf < 0 ? wf * t.bc_conc[(size_t)s * t.NBC + bce] : wf * ce[s];
        }
#pragma unroll
        for (int s = 0; s < M; s++) {
            double v = bfl[s];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
            if (lane == 0) sh[s][wid] = v;
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int s = 0; s < M; s++)
                acc_buf[(size_t)blockIdx.x * 2 * M + s] += (((sh[s][0] + sh[s][1]) + sh[s][2]) + sh[s][3]) * tdiff;
        return;
    }
    double bsr[M];
#pragma unroll
    for (int s = 0; s < M; s++) bsr[s] = 0.0;
    const int xcd = blockIdx.x & 7, lb = blockIdx.x >> 3, nlb = row_blocks >> 3;
    const int per = (t.E + 7) >> 3;
    const int r_lo = xcd * per, r_hi = min(t.E, r_lo + per);
    for (int i = r_lo + lb * (int)blockDim.x + (int)threadIdx.x; i < r_hi; i += nlb * (int)blockDim.x) {
        const int n = t.cln[i], bi = t.bidx[i];
        int cc[kEll];
        double av[kEll];
#pragma unroll
        for (int u = 0; u < kEll; u++) {
            cc[u] = t.ell_col[(size_t)u * t.E + i];
            av[u] = t.ell_val[(size_t)u * t.E + i];
        }
        double acc[M];
#pragma unroll
        for (int s = 0; s < M; s++) acc[s] = bi >= 0 ? t.bcv[(size_t)bi * M + s] : 0.0;
        double cv[kEll][M];
#pragma unroll
        for (int u = 0; u < kEll; u++)
            if (u < n) load_row<M>(c_old + (size_t)cc[u] * M, cv[u]);
        if (t.has_src) {
            // compute_concentration_sources (transport.cc:501-544), dt of this step
            double ci[M];
            load_row<M>(c_old + (size_t)i * M, ci);
#pragma unroll
            for (int s = 0; s < M; s++) {
                const size_t o = (size_t)i * M + s;
                const double dens = t.src_density[o], sig = t.src_sigma[o];
                const double diff = t.src_conc[o] - ci[s];
                const double snd = diff > 0.0 ? (dens + diff * sig) : dens;
                acc[s] = acc[s] + snd * dt;
                bsr[s] += snd * t.meas[i];
            }
        } else {
#pragma unroll
            for (int s = 0; s < M; s++) acc[s] = acc[s] + 0.0;
        }
#pragma unroll
        for (int u = 0; u < kEll; u++)
            if (u < n) {
#pragma unroll
                for (int s = 0; s < M; s++) acc[s] += av[u] * cv[u][s];
            }
        if (n > kEll) {  // rows of fracture and channel elements
            const int k0 = t.cst[i], k1 = k0 + n - kEll;
            for (int k = k0; k < k1; k++) {
                const double a = t.cval[k];
                double c2[M];
                load_row<M>(c_old + (size_t)t.ccol[k] * M, c2);
#pragma unroll
                for (int s = 0; s < M; s++) acc[s] += a * c2[s];
            }
        }
        if (t.has_react) {
            // Linear_reaction::compute_reaction (linear_reaction.cc:167-183): c_to = sum over
            // c_from * R[from][to], from in increasing order, after the convection step
            double r[M];
#pragma unroll
            for (int q = 0; q < M; q++) {
                r[q] = 0.0;
                for (int s = 0; s < M; s++) r[q] += acc[s] * t.react[s * M + q];
            }
#pragma unroll
            for (int q = 0; q < M; q++) acc[q] = r[q];
        }
#pragma unroll
        for (int s = 0; s < M; s++) c_new[(size_t)i * M + s] = acc[s];
    }
    if (!t.has_src) return;  // block-uniform
#pragma unroll
    for (int s = 0; s < M; s++) {
        double v = bsr[s];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        if (lane == 0) sh[s][wid] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int s = 0; s < M; s++)
            acc_buf[(size_t)blockIdx.x * 2 * M + M + s] += (((sh[s][0] + sh[s][1]) + sh[s][2]) + sh[s][3]) * tdiff;
}

// Output time: the blocks' running totals of k_step_c folded in a fixed order
// into the integrated quantities, then cleared.  One block.
template <int M>
__global__ void __launch_bounds__(256) k_flux_acc(const TmDev t, double *__restrict__ acc_buf, int G) {
    fold_flux<M>(t, acc_buf, G, 1.0);
    for (int k = threadIdx.x; k < G * 2 * M; k += blockDim.x) acc_buf[k] = 0.0;
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
template <class T>
static T *upv(std::vector<void *> &own, const std::vector<T> &v, hipStream_t s) {
    T *p = dev_upload(v.data(), v.size(), s);
    own.push_back(p);
    return p;
}

TransportGPU::TransportGPU(const HostMesh &m, int nsub, const double *por, const double *cs, const double *bc_conc,
                           const double *init_conc, double end_time, double save_step)
    : m_(m), E_(m.n_elements()), nsub_(nsub), NBC_(m.n_bc_elements()) {
    FL_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    const int E = E_, S = m.n_sides(), F = m.n_edges(), V = m.n_vb();
    // matrix pattern: rows in the reference's insertion order, later entries
    // overwrite earlier ones (INSERT_VALUES), then columns sorted per row.
    struct Ent { int kind, s1, s2; };
    std::vector<std::map<int, Ent>> rows(E);
    std::vector<int> vb_hside(std::max(V, 1), 0);
    for (int n = 0; n < V; n++) vb_hside[n] = m.edge_side[m.edge_ptr[m.vb_edge[n]]];
    for (int i = 0; i < E; i++) {
        for (int sd = m.side_off[i]; sd < m.side_off[i + 1]; sd++) {
            if (m.side_bnd[sd] >= 0) continue;
            int f = m.side_edge[sd];
            for (int k = m.edge_ptr[f]; k < m.edge_ptr[f + 1]; k++) {
                int sj = m.edge_side[k];
                if (sj == sd) continue;
                rows[i][m.side_elem[sj]] = Ent{kSame, sd, sj};
            }
        }
        for (int k = m.el_vb_ptr[i]; k < m.el_vb_ptr[i + 1]; k++) {
            int n = m.el_vb[k];
            int hs = vb_hside[n];
            int e2 = m.side_elem[hs];
            rows[i][e2] = Ent{kVbLow, hs, -1};
            rows[e2][i] = Ent{kVbHigh, hs, -1};
        }
        rows[i][i] = Ent{kDiag, -1, -1};
    }
    std::vector<int> ptr(E + 1, 0), col, kind, s1, s2, diag(E);
    for (int i = 0; i < E; i++) {
        for (auto &kv : rows[i]) {
            if (kv.first == i) diag[i] = (int)col.size();
            col.push_back(kv.first);
            kind.push_back(kv.second.kind);
            s1.push_back(kv.second.s1);
            s2.push_back(kv.second.s2);
        }
        ptr[i + 1] = (int)col.size();
    }
    std::vector<std::map<int, Ent>>().swap(rows);
    nnz_ = (int)col.size();
    // representative boundary sides per element: Boundary::side() is side 0 of its edge
    std::vector<int> rep(S, -1);
    for (int b = 0; b < m.n_boundaries(); b++) rep[m.edge_side[m.edge_ptr[m.bnd_edge[b]]]] = b;
    std::vector<int> bl_ptr(E + 1, 0), bl_side, bl_bcele;
    for (int i = 0; i < E; i++) {
        for (int sd = m.side_off[i]; sd < m.side_off[i + 1]; sd++)
            if (rep[sd] >= 0) {
                bl_side.push_back(sd);
                bl_bcele.push_back(m.bnd_bcele[rep[sd]]);
            }
        bl_ptr[i + 1] = (int)bl_side.size();
    }
    const int nbl = (int)bl_side.size();
    std::vector<int> bl_elem(std::max(nbl, 1), 0), bidx(E, -1);
    for (int i = 0; i < E; i++) {
        if (bl_ptr[i + 1] > bl_ptr[i]) bidx[i] = bl_ptr[i];
        for (int k = bl_ptr[i]; k < bl_ptr[i + 1]; k++) bl_elem[k] = i;
    }
    if (bl_side.empty()) { bl_side.push_back(0); bl_bcele.push_back(0); }
    h_meas_.resize(E);
    for (int i = 0; i < E; i++) h_meas_[i] = element_measure(m, m.el_dim[i], &m.el_nodes[4 * i]);
    h_cs_.assign(cs, cs + E);
    h_por_.assign(por, por + E);
    h_bc_conc_.assign(bc_conc, bc_conc + (size_t)nsub * NBC_);
    hipStream_t s = stream_;
    TmDev &t = t_;
    t.E = E; t.nsub = nsub; t.NBC = NBC_; t.has_src = 0;
    t.react = nullptr; t.has_react = 0;
    t.ptr = upv(owned_, ptr, s); t.col = upv(owned_, col, s); t.kind = upv(owned_, kind, s);
    t.s1 = upv(owned_, s1, s); t.s2 = upv(owned_, s2, s); t.diag = upv(owned_, diag, s);
    t.side_off = upv(owned_, m.side_off, s); t.side_edge = upv(owned_, m.side_edge, s);
    t.side_bnd = upv(owned_, m.side_bnd, s);
    std::vector<int> bb = m.bnd_bcele;
    if (bb.empty()) bb.push_back(0);
    t.bnd_bcele = upv(owned_, bb, s);
    t.edge_ptr = upv(owned_, m.edge_ptr, s); t.edge_side = upv(owned_, m.edge_side, s);
    t.el_vb_ptr = upv(owned_, m.el_vb_ptr, s);
    std::vector<int> elvb = m.el_vb;
    if (elvb.empty()) elvb.push_back(0);
    t.el_vb = upv(owned_, elvb, s);
    t.vb_hside = upv(owned_, vb_hside, s);
    t.bl_ptr = upv(owned_, bl_ptr, s); t.bl_side = upv(owned_, bl_side, s); t.bl_bcele = upv(owned_, bl_bcele, s);
    t.bl_elem = upv(owned_, bl_elem, s);
    t.bidx = upv(owned_, bidx, s);
    t.nbl = nbl;
    t.meas = upv(owned_, h_meas_, s); t.cs = upv(owned_, h_cs_, s); t.por = upv(owned_, h_por_, s);
    std::vector<double> bcc = h_bc_conc_;
    if (bcc.empty()) bcc.push_back(0);
    d_bc_conc_ = upv(owned_, bcc, s);
    t.bc_conc = d_bc_conc_;
    t.src_density = t.src_sigma = t.src_conc = nullptr;
    auto alloc_d = [&](size_t n) { double *p = dev_alloc<double>(n); owned_.push_back(p); return p; };
    d_flux_ = alloc_d(S);
    t.flux = d_flux_;
    t.edge_flow = alloc_d(F);
    t.val = alloc_d(nnz_);
    for (int i = 0; i < E; i++)
        if (ptr[i + 1] - ptr[i] > 255)
            throw Error(-5, "transport matrix row of element " + std::to_string(m.el_id[i]) + " has more than 255 entries");
    t.cval = alloc_d(nnz_);
    t.ell_val = alloc_d((size_t)kEll * E);
    {
        int *p = dev_alloc<int>((size_t)nnz_ + E + (size_t)kEll * E);
        owned_.push_back(p);
        t.ccol = p;
        t.cst = p + nnz_;
        t.ell_col = p + nnz_ + E;
        unsigned char *l = dev_alloc<unsigned char>(E);
        owned_.push_back(l);
        t.cln = l;
    }
    t.bcv = alloc_d((size_t)bl_side.size() * nsub);  // compact: one slot group per row with boundary sides
    d_c0_ = alloc_d((size_t)E * nsub);
    d_c1_ = alloc_d((size_t)E * nsub);
    t.partial = alloc_d(4096);
    d_acc_ = alloc_d((size_t)(kMaxStepGrid + kMaxBndBlocks) * 2 * kMaxSubst);
    FL_HIP(hipMemsetAsync(d_acc_, 0, (size_t)(kMaxStepGrid + kMaxBndBlocks) * 2 * kMaxSubst * sizeof(double), s));
    t.mb = alloc_d(1 + 2 * nsub);
    unsigned *cnt = dev_alloc<unsigned>(1);
    owned_.push_back(cnt);
    t.count = cnt;
    FL_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned), s));
    FL_HIP(hipMemsetAsync(t.mb, 0, (1 + 2 * nsub) * sizeof(double), s));
    FL_HIP(hipMemsetAsync(d_c1_, 0, (size_t)E * nsub * sizeof(double), s));
    // init_conc [nsub][E] -> interleaved [E][nsub]
    std::vector<double> c0((size_t)E * nsub);
    for (int k = 0; k < nsub; k++)
        for (int i = 0; i < E; i++) c0[(size_t)i * nsub + k] = init_conc[(size_t)k * E + i];
    FL_HIP(hipMemcpyAsync(d_c0_, c0.data(), c0.size() * sizeof(double), hipMemcpyHostToDevice, s));
    FL_HIP(hipStreamSynchronize(s));
    tg_.init(end_time, save_step);
    tos_.init(end_time, save_step);
    cfl_max_step_ = end_time;
    h_el_region_sub_.resize(E);
    for (int i = 0; i < E; i++) h_el_region_sub_[i] = m.regions[m.el_region[i]].sub_idx;
}

TransportGPU::~TransportGPU() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    free_step_graphs();
    if (h_plan_) (void)hipHostFree(h_plan_);
    for (void *p : owned_) dev_free(p);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void TransportGPU::set_sources(const double *density, const double *sigma, const double *conc) {
    const size_t n = (size_t)E_ * nsub_;
    h_src_density_.assign(density, density + n);
    h_src_sigma_.assign(sigma, sigma + n);
    h_src_conc_.assign(conc, conc + n);
    std::vector<double> d(n), g(n), c(n);
    for (int k = 0; k < nsub_; k++)
        for (int i = 0; i < E_; i++) {
            d[(size_t)i * nsub_ + k] = density[(size_t)k * E_ + i];
            g[(size_t)i * nsub_ + k] = sigma[(size_t)k * E_ + i];
            c[(size_t)i * nsub_ + k] = conc[(size_t)k * E_ + i];
        }
    t_.src_density = upv(owned_, d, stream_);
    t_.src_sigma = upv(owned_, g, stream_);
    t_.src_conc = upv(owned_, c, stream_);
    t_.has_src = 1;
    FL_HIP(hipStreamSynchronize(stream_));
}

void TransportGPU::set_linear_reactions(int n_decays, const int *parent, const int *pptr, const int *products,
                                        const double *half_life, const double *branch) {
    pade_ = false;
    define_reactions(n_decays, parent, pptr, products, half_life, branch);
}

void TransportGPU::define_reactions(int n_decays, const int *parent, const int *pptr, const int *products,
                                    const double *half_life, const double *branch) {
    if (n_decays <= 0) {
        react_on_ = false;
        t_.has_react = 0;
        return;
    }
    for (int i = 0; i < n_decays; i++) {
        if (parent[i] < 0 || parent[i] >= nsub_) throw Error(-3, "Wrong name of parent substance in the " +
                                                                  std::to_string(i) + "-th reaction.");
        if (pptr[i + 1] <= pptr[i]) throw Error(-3, "Empty array of products in the " + std::to_string(i) + "-th reaction.");
        if (!(half_life[i] > 0.0)) throw Error(-3, "half-life of the " + std::to_string(i) + "-th reaction must be positive");
        for (int k = pptr[i]; k < pptr[i + 1]; k++)
            if (products[k] < 0 || products[k] >= nsub_) throw Error(-3, "Wrong product in the " + std::to_string(i) + "-th reaction.");
    }
    r_parent_.assign(parent, parent + n_decays);
    r_pptr_.assign(pptr, pptr + n_decays + 1);
    r_prod_.assign(products, products + pptr[n_decays]);
    r_hl_.assign(half_life, half_life + n_decays);
    r_branch_.assign(branch, branch + pptr[n_decays]);
    if (!d_react_) {
        d_react_ = dev_alloc<double>((size_t)kMaxSubst * kMaxSubst);
        owned_.push_back(d_react_);
    }
    react_on_ = true;
    t_.react = d_react_;
    t_.has_react = 1;
    build_reaction_matrix(tg_.estimate_dt());
}

void TransportGPU::set_pade_reactions(int n_decays, const int *parent, const int *pptr, const int *products,
                                      const double *half_life, const double *branch, int nom_deg, int den_deg) {
    if (n_decays > 0 && (nom_deg < 0 || den_deg < 0 || nom_deg + den_deg < 1 || nom_deg > 10 || den_deg > 10))
        throw Error(-3, "Pade degrees must be in 0..10 with a positive sum");
    pade_nom_ = nom_deg;
    pade_den_ = den_deg;
    pade_ = n_decays > 0;
    try {
        define_reactions(n_decays, parent, pptr, products, half_life, branch);
    } catch (...) {
        pade_ = false;
        throw;
    }
}

static double factorial(int n) {
    double f = 1.0;
    for (int k = 2; k <= n; k++) f *= k;
    return f;
}

// Pade_approximant::modify_reaction_matrix (pade_approximant.cc:97-239): A with
// A[p][p] = -ln2 dt/T, A[p][product] = ln2 dt/T ratio; N(A) and D(A) the (nom, den)
// Pade polynomials of exp evaluated by Horner; the step matrix is P = (D^-1 N)^T and
// each element's concentrations become P c (pade_approximant.cc:274-290).  Stored
// here transposed, R[from][to] = P[to][from], so the step kernel's c_to = sum c_from
// R[from][to] forms the same products in the same order.
static void pade_matrix(int n, const std::vector<int> &parent, const std::vector<int> &pptr,
                        const std::vector<int> &prod, const std::vector<double> &hl, const std::vector<double> &branch,
                        int p, int q, double dt, std::vector<double> &R) {
    std::vector<double> A((size_t)n * n, 0.0);
    for (size_t i = 0; i < parent.size(); i++) {
        const double rel = dt / hl[i];
        A[(size_t)parent[i] * n + parent[i]] = -std::log(2.0) * rel;
        for (int k = pptr[i]; k < pptr[i + 1]; k++) A[(size_t)parent[i] * n + prod[k]] = std::log(2.0) * rel * branch[k];
    }
    auto poly = [&](int deg, bool den) {
        std::vector<double> P((size_t)n * n, 0.0), T((size_t)n * n);
        for (int j = deg; j >= 0; j--) {
            const double c = den ? std::pow(-1.0, j) * factorial(p + q - j) * factorial(q) /
                                       (factorial(p + q) * factorial(j) * factorial(q - j))
                                 : factorial(p + q - j) * factorial(p) /
                                       (factorial(p + q) * factorial(j) * factorial(p - j));
            for (int a = 0; a < n; a++)
                for (int b = 0; b < n; b++) {
                    double v = 0.0;
                    for (int k = 0; k < n; k++) v += P[(size_t)a * n + k] * A[(size_t)k * n + b];
                    T[(size_t)a * n + b] = v;
                }
            P = T;
            for (int a = 0; a < n; a++) P[(size_t)a * n + a] += c;
        }
        return P;
    };
    std::vector<double> Nm = poly(p, false), Dm = poly(q, true);
    // column r of N solved against D (LU with partial pivoting): X = D^-1 N, P[r][c] = X[c][r]
    std::vector<double> LU = Dm;
    std::vector<int> piv(n);
    for (int k = 0; k < n; k++) {
        int m = k;
        for (int i = k + 1; i < n; i++)
            if (std::fabs(LU[(size_t)i * n + k]) > std::fabs(LU[(size_t)m * n + k])) m = i;
        piv[k] = m;
        if (m != k)
            for (int j = 0; j < n; j++) std::swap(LU[(size_t)k * n + j], LU[(size_t)m * n + j]);
        if (LU[(size_t)k * n + k] == 0.0) throw Error(-3, "singular Pade denominator");
        for (int i = k + 1; i < n; i++) {
            LU[(size_t)i * n + k] /= LU[(size_t)k * n + k];
            for (int j = k + 1; j < n; j++) LU[(size_t)i * n + j] -= LU[(size_t)i * n + k] * LU[(size_t)k * n + j];
        }
    }
    R.assign((size_t)n * n, 0.0);
    std::vector<double> x(n);
    for (int r = 0; r < n; r++) {
        for (int i = 0; i < n; i++) x[i] = Nm[(size_t)i * n + r];
        for (int k = 0; k < n; k++) std::swap(x[k], x[piv[k]]);
        for (int i = 1; i < n; i++)
            for (int j = 0; j < i; j++) x[i] -= LU[(size_t)i * n + j] * x[j];
        for (int i = n - 1; i >= 0; i--) {
            for (int j = i + 1; j < n; j++) x[i] -= LU[(size_t)i * n + j] * x[j];
            x[i] /= LU[(size_t)i * n + i];
        }
        // P[r][c] = x[c]; R[from = c][to = r] = P[r][c]
        for (int c = 0; c < n; c++) R[(size_t)c * n + r] = x[c];
    }
}

// Linear_reaction::modify_reaction_matrix (linear_reaction.cc:144-165): identity,
// then per decay R[p][p] = 0.5^(dt / half_life), R[p][product_k] = (1 - that) * ratio_k
void TransportGPU::build_reaction_matrix(double dt) {
    const int n = nsub_;
    if (pade_) {
        pade_matrix(n, r_parent_, r_pptr_, r_prod_, r_hl_, r_branch_, pade_nom_, pade_den_, dt, h_react_);
        FL_HIP(hipMemcpyAsync(d_react_, h_react_.data(), h_react_.size() * sizeof(double), hipMemcpyHostToDevice,
                              stream_));
        FL_HIP(hipStreamSynchronize(stream_));
        return;
    }
    h_react_.assign((size_t)n * n, 0.0);
    for (int k = 0; k < n; k++) h_react_[(size_t)k * n + k] = 1.0;
    for (size_t i = 0; i < r_parent_.size(); i++) {
        const int p = r_parent_[i];
        const double rel = dt / r_hl_[i];
        h_react_[(size_t)p * n + p] = std::pow(0.5, rel);
        for (int k = r_pptr_[i]; k < r_pptr_[i + 1]; k++)
            h_react_[(size_t)p * n + r_prod_[k]] = (1 - std::pow(0.5, rel)) * r_branch_[k];
    }
    FL_HIP(hipMemcpyAsync(d_react_, h_react_.data(), h_react_.size() * sizeof(double), hipMemcpyHostToDevice, stream_));
    FL_HIP(hipStreamSynchronize(stream_));  // h_react_ is rebuilt at the next output interval
}

void TransportGPU::set_bc_conc(const double *bc_conc) {
    h_bc_next_.assign(bc_conc, bc_conc + (size_t)nsub_ * NBC_);
    bc_pending_ = true;
}

void TransportGPU::set_velocity_host(const double *side_flux) {
    FL_HIP(hipMemcpyAsync(d_flux_, side_flux, m_.n_sides() * sizeof(double), hipMemcpyHostToDevice, stream_));
    FL_HIP(hipStreamSynchronize(stream_));
    flux_set_ = true;
    matrix_built_ = false;  // mh_dh->time_changed() > transport_matrix_time
    setup_done_ = false;
}

void TransportGPU::set_velocity_device(const double *d_side_flux) {
    FL_HIP(hipMemcpyAsync(d_flux_, d_side_flux, m_.n_sides() * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    FL_HIP(hipStreamSynchronize(stream_));
    flux_set_ = true;
    matrix_built_ = false;
    setup_done_ = false;
}

double TransportGPU::prepare() {
    if (!flux_set_) throw Error(-4, "velocity field not set");
    if (!matrix_built_) build_matrix();
    return cfl_max_step_;
}

void TransportGPU::build_matrix() {
    int F = m_.n_edges();
    hipEvent_t a, b;
    FL_HIP(hipEventCreate(&a)); FL_HIP(hipEventCreate(&b));
    FL_HIP(hipEventRecord(a, stream_));
    hipLaunchKernelGGL(k_edge_flow, dim3(std::min(grid_for(F), 4096)), dim3(kBlock), 0, stream_, F, t_.edge_ptr,
                       t_.edge_side, (const double *)d_flux_, (double *)t_.edge_flow);
    hipLaunchKernelGGL(k_tm_build, dim3(std::min(grid_for(E_), 4096)), dim3(kBlock), 0, stream_, t_);
    FL_HIP(hipGetLastError());
    double mx = 0.0;
    FL_HIP(hipMemcpyAsync(&mx, t_.mb, sizeof(double), hipMemcpyDeviceToHost, stream_));
    FL_HIP(hipEventRecord(b, stream_));
    FL_HIP(hipStreamSynchronize(stream_));
    float ms = 0;
    FL_HIP(hipEventElapsedTime(&ms, a, b));
    stats_.ms_build += ms;
    (void)hipEventDestroy(a); (void)hipEventDestroy(b);
    cfl_max_step_ = 1 / mx;
    matrix_built_ = true;
    scaled_ = false;
}

void TransportGPU::set_bc() {
    double est = tg_.estimate_dt();
    hipLaunchKernelGGL(k_set_bc, dim3(std::min(grid_for(E_), 4096)), dim3(kBlock), 0, stream_, t_, est);
    FL_HIP(hipGetLastError());
}

void TransportGPU::mb_initial() {
    // MassBalance::calculate at the first call fixes the initial mass and time
    std::vector<double> c((size_t)E_ * nsub_);
    FL_HIP(hipMemcpyAsync(c.data(), cur_is_c0_ ? d_c0_ : d_c1_, c.size() * 8, hipMemcpyDeviceToHost, stream_));
    FL_HIP(hipStreamSynchronize(stream_));
    initial_mass_.assign(nsub_, 0.0);
    for (int k = 0; k < nsub_; k++)
        for (int i = 0; i < E_; i++) initial_mass_[k] += h_por_[i] * h_cs_[i] * c[(size_t)i * nsub_ + k] * h_meas_[i];
    last_time_ = mb_t0_ = tg_.time;
    mb_init_ = false;
}

// one lane per row: enough blocks to cover the rows in one sweep, a multiple
// of 8 (one share per XCD); then the boundary-side blocks
int TransportGPU::step_row_blocks() const {
    const int g = (((E_ + kBlock - 1) / kBlock + 7) / 8) * 8;
    return std::max(8, std::min(g, kMaxStepGrid));
}
int TransportGPU::step_grid() const {
    const int nb = t_.nbl > 0 ? std::min((t_.nbl + kBlock - 1) / kBlock, kMaxBndBlocks) : 0;
    return step_row_blocks() + nb;
}

void TransportGPU::flush_flux() {
    if (!acc_dirty_) return;
    const int G = step_grid();
    switch (nsub_) {
#define FL_FLUSH(MM) \
    case MM: hipLaunchKernelGGL(k_flux_acc<MM>, dim3(1), dim3(kBlock), 0, stream_, t_, d_acc_, G); break;
        FL_FLUSH(1) FL_FLUSH(2) FL_FLUSH(3) FL_FLUSH(4) FL_FLUSH(5) FL_FLUSH(6) FL_FLUSH(7) FL_FLUSH(8)
#undef FL_FLUSH
    }
    FL_HIP(hipGetLastError());
    acc_dirty_ = false;
}

void TransportGPU::compact_matrix() {
    const int waves = (E_ + 63) / 64;
    hipLaunchKernelGGL(k_tm_compact, dim3((waves + 3) / 4), dim3(kBlock), 0, stream_, t_);
}

// ConvectionTransport::compute_one_step (transport.cc:593-715)
void TransportGPU::one_step() {
    const int gridS = std::min(grid_for(E_), 4096);
    // data_.set_time at the step's start (transport.cc:600): boundary data given since
    // the last step apply now; with the matrix in place only the boundary vector is
    // rebuilt (transport.cc:620-626); a function of time gives new data at every step
    bool bc_changed = false;
    if (bc_fn_) {
        h_bc_next_.assign((size_t)nsub_ * NBC_, 0.0);
        if (bc_fn_(tg_.time, h_bc_next_.data(), bc_user_) != 0)
            throw Error(kErrCallback, "the boundary concentration function failed at t = " + std::to_string(tg_.time));
        bc_pending_ = true;
    }
    if (bc_pending_) {
        // the previous upload from h_bc_conc_ has finished before it is overwritten
        FL_HIP(hipStreamSynchronize(stream_));
        h_bc_conc_ = h_bc_next_;
        FL_HIP(hipMemcpyAsync(d_bc_conc_, h_bc_conc_.data(), h_bc_conc_.size() * sizeof(double),
                              hipMemcpyHostToDevice, stream_));
        bc_pending_ = false;
        bc_changed = true;
    }
    if (setup_done_ && bc_changed) set_bc();
    if (!matrix_built_) build_matrix();
    if (!setup_done_) {
        setup_done_ = true;
        tg_.set_upper_constraint(cfl_max_step_);
        tg_.fix_dt_until_mark();
        set_bc();
        need_rescale_ = true;
    }
    if (need_rescale_) {
        if (scaled_) {
            double fac = tg_.estimate_dt() / tg_.time_step;
            hipLaunchKernelGGL(k_tm_scale, dim3(gridS), dim3(kBlock), 0, stream_, t_, 0, fac);
        } else {
            hipLaunchKernelGGL(k_tm_scale, dim3(gridS), dim3(kBlock), 0, stream_, t_, 1, tg_.estimate_dt());
            scaled_ = true;
        }
        compact_matrix();
        need_rescale_ = false;
    }
    tg_.next_time();
    double tdiff = tg_.time - last_time_;
    last_time_ = tg_.time;
    const double *cold = cur_is_c0_ ? d_c0_ : d_c1_;
    double *cnew = cur_is_c0_ ? d_c1_ : d_c0_;
    if (rs_n_ && !rs_built_) build_reaction_sorption();
    launch_step(cold, cnew, tg_.time_step, tdiff, nullptr);
    cur_is_c0_ = !cur_is_c0_;
    stats_.steps++;
}

// The kernels of one step from cold to cnew: the convection step (with the
// decays unless the phases run), dual porosity / sorption, the reaction
// module's sorption.  dt and tdiff as arguments, or from plan[0..1] on the
// device when plan is given (a step captured into a graph).
void TransportGPU::launch_step(const double *cold, double *cnew, double dt, double tdiff, const double *plan) {
    const int gridS = std::min(grid_for(E_), 4096);
    const int G = step_grid();
    // with dual porosity / sorption the decays run after them, in k_phases
    TmDev ts = t_;
    if (phases_on()) ts.has_react = 0;
    switch (nsub_) {
#define FL_STEP(MM) \
    case MM:                                                                                                  \
        hipLaunchKernelGGL(k_step_c<MM>, dim3(G), dim3(kBlock), 0, stream_, ts, cold, cnew, dt, tdiff, d_acc_,  \
                           step_row_blocks(), plan);                                                           \
        break;
        FL_STEP(1) FL_STEP(2) FL_STEP(3) FL_STEP(4) FL_STEP(5) FL_STEP(6) FL_STEP(7) FL_STEP(8)
#undef FL_STEP
        default: throw Error(-5, "transport supports 1..8 substances per handle");
    }
    FL_HIP(hipGetLastError());
    acc_dirty_ = true;  // totals stay in the per-block accumulators until the output time
    if (phases_on()) {
        PhaseDev ph{d_imm_, d_msorb_, d_isorb_, d_alpha_, d_sc0_, d_sc1_, d_por_imm_, d_phi_, d_stype_,
                    dual_ ? 1 : 0, sorp_ ? 1 : 0};
        switch (nsub_) {
#define FL_PH(MM) \
    case MM: hipLaunchKernelGGL(k_phases<MM>, dim3(gridS), dim3(kBlock), 0, stream_, t_, ph, cnew, dt, plan); break;
            FL_PH(1) FL_PH(2) FL_PH(3) FL_PH(4) FL_PH(5) FL_PH(6) FL_PH(7) FL_PH(8)
#undef FL_PH
        }
        FL_HIP(hipGetLastError());
    }
    if (rs_n_) {
        switch (nsub_) {
#define FL_RS(MM)                                                                                                 \
    case MM:                                                                                                      \
        hipLaunchKernelGGL(k_rsorp<MM>, dim3(gridS), dim3(kBlock), 0, stream_, E_, rs_n_, d_rs_subst_, d_rs_el_par_, \
                           (const IsoPar *)d_rs_par_, d_rs_table_, cnew, d_rs_sorbed_, d_rs_err_);              \
        break;
            FL_RS(1) FL_RS(2) FL_RS(3) FL_RS(4) FL_RS(5) FL_RS(6) FL_RS(7) FL_RS(8)
#undef FL_RS
        }
        FL_HIP(hipGetLastError());
    }
}

// TransportOperatorSplitting::update_solution (transport_operator_splitting.cc:244-287)
int TransportGPU::update_solution(double target) {
    if (!flux_set_) throw Error(-4, "velocity field not set");
    if (mb_init_) mb_initial();
    // set_target_time (transport.cc:718-733)
    tg_.add_mark(target);
    tg_.set_upper_constraint(cfl_max_step_);
    tg_.fix_dt_until_mark();
    need_rescale_ = true;
    // decayRad->set_time_step(convection->time().estimate_dt()) (transport_operator_splitting.cc:258)
    if (react_on_) build_reaction_matrix(tg_.estimate_dt());
    // the events are destroyed on every exit, a failed step included
    struct Events {
        hipEvent_t a = nullptr, b = nullptr;
        ~Events() {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } ev;
    FL_HIP(hipEventCreate(&ev.a)); FL_HIP(hipEventCreate(&ev.b));
    FL_HIP(hipEventRecord(ev.a, stream_));
    int steps = 0;
    while (tg_.lt(target)) {
        if (plain_steps_ready()) {
            // every further step to the target is the same kernel sequence:
            // plan them on a copy of the governor and replay captured graphs
            steps += run_planned(target);
            continue;
        }
        double t_before = tg_.time;
        one_step();
        steps++;
        if (!(tg_.time > t_before) || !std::isfinite(tg_.time))
            throw Error(-7, "transport time step does not advance (dt = " + std::to_string(tg_.time_step) + ")");
    }
    flush_flux();
    FL_HIP(hipEventRecord(ev.b, stream_));
    FL_HIP(hipEventSynchronize(ev.b));
    float ms = 0;
    FL_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
    stats_.ms_steps += ms;
    if (rs_built_) check_sorption();
    return steps;
}

// Steps whose launches do not depend on host work between them: no boundary
// data to apply, the matrix built and scaled, the sorption isotherms made.
bool TransportGPU::plain_steps_ready() const {
    return !bc_fn_ && !bc_pending_ && matrix_built_ && setup_done_ && !need_rescale_ && (!rs_n_ || rs_built_);
}

// What a captured step's launches depend on besides the plan: the kernel
// argument structures, buffers and switches.  A graph is replayed only while
// this is unchanged.
std::vector<unsigned char> TransportGPU::step_signature(int steps, int parity) const {
    std::vector<unsigned char> sig;
    auto put = [&](const void *p, size_t n) {
        const unsigned char *c = static_cast<const unsigned char *>(p);
        sig.insert(sig.end(), c, c + n);
    };
    put(&t_, sizeof(t_));
    const void *ptrs[] = {d_c0_, d_c1_, d_acc_, d_plan_, d_imm_, d_msorb_, d_isorb_, d_alpha_, d_sc0_, d_sc1_,
                          d_por_imm_, d_phi_, d_stype_, d_rs_subst_, d_rs_el_par_, d_rs_par_, d_rs_table_,
                          d_rs_sorbed_, d_rs_err_};
    put(ptrs, sizeof(ptrs));
    const int ints[] = {steps, parity, nsub_, rs_n_, dual_ ? 1 : 0, sorp_ ? 1 : 0, E_};
    put(ints, sizeof(ints));
    return sig;
}

void TransportGPU::free_step_graphs() {
    for (StepGraph &g : graphs_)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    graphs_.clear();
}

// The remaining steps to `target`, all plain: their step lengths and time
// differences come from a copy of the governor advanced as one_step would,
// into pinned memory; runs of 64, 16 and 4 steps replay captured graphs (one
// per run length and starting buffer, each preceded by the copy of its
// plan), the rest launch step by step.  Results are those of one_step: the
// same launches with the same values.
int TransportGPU::run_planned(double target) {
    TimeGovernorR g = tg_;
    double last = last_time_;
    std::vector<double> plan;
    while (g.lt(target)) {
        const double t_before = g.time;
        g.next_time();
        if (!(g.time > t_before) || !std::isfinite(g.time))
            throw Error(-7, "transport time step does not advance (dt = " + std::to_string(g.time_step) + ")");
        plan.push_back(g.time_step);
        plan.push_back(g.time - last);
        last = g.time;
    }
    const int n = (int)plan.size() / 2;
    if (n == 0) return 0;
    if (plan.size() > h_plan_cap_) {  // the previous update's copies have completed (it synchronized)
        if (h_plan_) (void)hipHostFree(h_plan_);
        h_plan_ = nullptr;
        h_plan_cap_ = 0;
        FL_HIP(hipHostMalloc((void **)&h_plan_, plan.size() * sizeof(double), hipHostMallocDefault));
        h_plan_cap_ = plan.size();
    }
    std::copy(plan.begin(), plan.end(), h_plan_);
    if (!d_plan_) {
        d_plan_ = dev_alloc<double>(2 * kGraphSteps[0]);
        owned_.push_back(d_plan_);
    }
    int k = 0;
    while (k < n) {
        int S = 0;
        for (int len : kGraphSteps)
            if (n - k >= len) {
                S = len;
                break;
            }
        if (S == 0) {  // the last few steps one by one
            const double *cold = cur_is_c0_ ? d_c0_ : d_c1_;
            double *cnew = cur_is_c0_ ? d_c1_ : d_c0_;
            launch_step(cold, cnew, h_plan_[2 * k], h_plan_[2 * k + 1], nullptr);
            cur_is_c0_ = !cur_is_c0_;
            k++;
            continue;
        }
        const int parity = cur_is_c0_ ? 0 : 1;
        std::vector<unsigned char> sig = step_signature(S, parity);
        StepGraph *gr = nullptr;
        for (StepGraph &x : graphs_)
            if (x.steps == S && x.parity == parity) gr = &x;
        if (gr && gr->sig != sig) {  // launch arguments changed since the capture
            (void)hipGraphExecDestroy(gr->exec);
            gr->exec = nullptr;
        }
        if (!gr) {
            graphs_.push_back(StepGraph{});
            gr = &graphs_.back();
            gr->steps = S;
            gr->parity = parity;
        }
        if (!gr->exec) {
            hipGraph_t graph = nullptr;
            FL_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            bool c0 = cur_is_c0_;
            for (int j = 0; j < S; j++) {
                launch_step(c0 ? d_c0_ : d_c1_, c0 ? d_c1_ : d_c0_, 0.0, 0.0, d_plan_ + 2 * j);
                c0 = !c0;
            }
            FL_HIP(hipStreamEndCapture(stream_, &graph));
            const hipError_t e = hipGraphInstantiate(&gr->exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                gr->exec = nullptr;
                throw Error(-2, std::string("transport step graph: ") + hipGetErrorString(e));
            }
            gr->sig = sig;
        }
        FL_HIP(hipMemcpyAsync(d_plan_, h_plan_ + 2 * k, 2 * S * sizeof(double), hipMemcpyHostToDevice, stream_));
        FL_HIP(hipGraphLaunch(gr->exec, stream_));
        if (S % 2) cur_is_c0_ = !cur_is_c0_;
        k += S;
    }
    tg_ = g;
    last_time_ = last;
    stats_.steps += n;
    acc_dirty_ = true;
    return n;
}

int TransportGPU::update_solution_step() {
    tos_.next_time();
    return update_solution(tos_.time);
}

void TransportGPU::set_phases(bool dual, bool sorption, const double *por_imm, const double *alpha,
                              const double *phi, const int *sorp_type, const double *sorp_coef0,
                              const double *sorp_coef1) {
    if (stats_.steps > 0) throw Error(-3, "dual porosity / sorption must be set before the first step");
    if (!por_imm) throw Error(-3, "por_imm is required");
    if (dual && !alpha) throw Error(-3, "dual porosity needs alpha");
    if (sorption && (!sorp_type || !sorp_coef0 || !sorp_coef1)) throw Error(-3, "sorption needs its isotherms");
    if (dual && rs_n_) throw Error(-3, "the reaction module's sorption is not combined with dual porosity here");
    const size_t n = (size_t)E_ * nsub_;
    auto inter = [&](const double *a, double dflt) {
        std::vector<double> v(n, dflt);
        if (a)
            for (int k = 0; k < nsub_; k++)
                for (int i = 0; i < E_; i++) v[(size_t)i * nsub_ + k] = a[(size_t)k * E_ + i];
        return v;
    };
    std::vector<int> ty(n, 0);
    if (sorp_type)
        for (int k = 0; k < nsub_; k++)
            for (int i = 0; i < E_; i++) {
                const int v = sorp_type[(size_t)k * E_ + i];
                if (v < 0 || v > 3) throw Error(-3, "sorption type must be 0..3 (none, linear, freundlich, langmuir)");
                ty[(size_t)i * nsub_ + k] = v;
            }
    h_por_imm_.assign(por_imm, por_imm + E_);
    d_por_imm_ = upv(owned_, h_por_imm_, stream_);
    std::vector<double> ph(E_, 1.0);
    if (phi) ph.assign(phi, phi + E_);
    h_phi_ = ph;
    d_phi_ = upv(owned_, ph, stream_);
    d_alpha_ = upv(owned_, inter(alpha, 0.0), stream_);
    d_sc0_ = upv(owned_, inter(sorp_coef0, 0.0), stream_);
    d_sc1_ = upv(owned_, inter(sorp_coef1, 0.0), stream_);
    d_stype_ = upv(owned_, ty, stream_);
    std::vector<double> zero(n, 0.0);
    d_imm_ = upv(owned_, zero, stream_);
    d_msorb_ = upv(owned_, zero, stream_);
    d_isorb_ = upv(owned_, zero, stream_);
    dual_ = dual;
    sorp_ = sorption;
    FL_HIP(hipStreamSynchronize(stream_));
}

// Sorption::Sorption (sorption.cc:71-146): the species and their parameters,
// kept per element ([n][E] like the ABI's arrays); the isotherms are made at the
// first step, when the porosities of the transport are final.
void TransportGPU::set_reaction_sorption(int n, const int *subst, const int *type, const double *mult,
                                         const double *second, const double *rock_density, const double *molar_mass,
                                         const double *solubility, double solvent_dens, int substeps,
                                         bool elementwise) {
    if (stats_.steps > 0) throw Error(-3, "sorption must be set before the first step");
    if (n < 1 || n > nsub_) throw Error(-3, "sorbing species must be 1..n_subst");
    if (dual_) throw Error(-3, "the reaction module's sorption is not combined with dual porosity here");
    if (substeps < 1 || substeps > 100000) throw Error(-3, "substeps must be in 1..100000");
    for (int k = 0; k < n; k++)
        if (subst[k] < 0 || subst[k] >= nsub_) throw Error(-3, "Wrong name of " + std::to_string(k) + "-th adsorbing specie.");
    const int nreg = m_.n_bulk_regions;
    const size_t ne = (size_t)n * E_;
    rs_type_.resize(ne);
    rs_mult_.resize(ne);
    rs_second_.resize(ne);
    rs_rock_.resize(E_);
    for (int i = 0; i < E_; i++) {
        const int r = h_el_region_sub_[i];
        rs_rock_[i] = elementwise ? rock_density[i] : rock_density[r];
        for (int k = 0; k < n; k++) {
            const size_t src = elementwise ? (size_t)k * E_ + i : (size_t)r * n + k;
            const int t = type[src];
            if (t < 0 || t > 3) throw Error(-3, "sorption type must be 0..3 (none, linear, freundlich, langmuir)");
            rs_type_[(size_t)k * E_ + i] = t;
            rs_mult_[(size_t)k * E_ + i] = mult[src];
            rs_second_[(size_t)k * E_ + i] = second[src];
        }
    }
    (void)nreg;
    rs_subst_.assign(subst, subst + n);
    rs_molar_.assign(molar_mass, molar_mass + n);
    rs_sol_.assign(n, -1.0);
    if (solubility) rs_sol_.assign(solubility, solubility + n);
    rs_solvent_ = solvent_dens;
    rs_substeps_ = substeps;
    rs_elementwise_ = elementwise;
    rs_n_ = n;
    rs_built_ = false;
}

// Sorption::make_tables + isotherm_reinit (sorption.cc:148-213): a region whose
// parameters (rock density, isotherm coefficients, por_m, por_imm, phi) are all
// constant gets one table per species (Isotherm::make_table, isotherm.cc:202-242);
// the other regions' elements keep their own isotherms, solved directly.  The
// reference decides by the field type (FieldConstant) on the region; here the
// values decide, and the element-wise entry point always takes the direct path.
void TransportGPU::build_reaction_sorption() {
    if (dual_) throw Error(-3, "the reaction module's sorption is not combined with dual porosity here");
    const int n = rs_n_, nreg = m_.n_bulk_regions;
    auto por_imm = [&](int i) { return h_por_imm_.empty() ? 0.0 : h_por_imm_[i]; };
    auto phi = [&](int i) { return h_phi_.empty() ? 1.0 : h_phi_[i]; };
    auto same = [&](int i, int j) {
        if (h_por_[i] != h_por_[j] || por_imm(i) != por_imm(j) || phi(i) != phi(j) || rs_rock_[i] != rs_rock_[j])
            return false;
        for (int k = 0; k < n; k++) {
            const size_t a = (size_t)k * E_ + i, b = (size_t)k * E_ + j;
            if (rs_type_[a] != rs_type_[b] || rs_mult_[a] != rs_mult_[b] || rs_second_[a] != rs_second_[b])
                return false;
        }
        return true;
    };
    std::vector<int> first(nreg, -1);
    std::vector<char> tabled(nreg, rs_elementwise_ ? 0 : 1);
    for (int i = 0; i < E_; i++) {
        const int r = h_el_region_sub_[i];
        if (first[r] < 0) first[r] = i;
        else if (tabled[r] && !same(i, first[r])) tabled[r] = 0;
    }
    // Isotherm::reinit from element i's values (sorption.cc:172-213)
    auto reinit = [&](int i, int k) {
        IsoPar I{};
        I.type = rs_type_[(size_t)k * E_ + i];
        I.c0 = rs_mult_[(size_t)k * E_ + i];
        I.c1 = rs_second_[(size_t)k * E_ + i];
        I.rho = rs_solvent_;
        I.sa = h_por_[i];
        I.ss = phi(i) * (1 - h_por_[i] - por_imm(i)) * rs_rock_[i] * rs_molar_[k];
        if (I.ss == 0.0)
            throw Error(-3, "Parameter scale_sorbed (phi * (1 - por_m - por_imm) * rock_density * molar_mass) of "
                            "element " + std::to_string(m_.el_id[i]) + " is equal to zero");
        I.limit = rs_sol_[k];
        I.limited = I.limit > 0.0;
        I.tab = -1;
        I.len = 0;
        return I;
    };
    std::vector<IsoPar> par;
    std::vector<double> tab;
    std::vector<int> el_par(E_);
    for (int r = 0; r < nreg; r++) {
        if (!tabled[r] || first[r] < 0) continue;
        for (int k = 0; k < n; k++) {
            IsoPar I = reinit(first[r], k);
            // Isotherm::make_table (isotherm.cc:220-242)
            const int steps = I.type == 0 ? 1 : rs_substeps_;
            if (I.limit <= 0.0) I.limit = 1.0;
            const double mass_limit = I.sa * I.limit + I.ss * iso_f(I, I.limit / I.rho);
            I.step = mass_limit / steps;
            I.tab = (int)tab.size();
            I.len = steps + 1;
            double mass = 0.0;
            for (int j = 0; j <= steps; j++) {
                double ca = mass / I.sa, cs = 0.0;
                if (!iso_solve(I, ca, cs))
                    throw Error(kErrSorption, "the isotherm table of bulk region " + std::to_string(r) +
                                                  " has no root at total mass " + std::to_string(mass));
                tab.push_back(cs * I.sa - ca * I.ss);
                mass = mass + I.step;
            }
            par.push_back(I);
        }
    }
    // table isotherms first, [region][species] in region order
    std::vector<int> reg_base(nreg, -1);
    for (int r = 0, at = 0; r < nreg; r++)
        if (tabled[r] && first[r] >= 0) {
            reg_base[r] = at;
            at += n;
        }
    for (int i = 0; i < E_; i++) {
        const int r = h_el_region_sub_[i];
        if (reg_base[r] >= 0) {
            el_par[i] = reg_base[r];
        } else {
            el_par[i] = (int)par.size();
            for (int k = 0; k < n; k++) {
                const IsoPar I = reinit(i, k);
                // toms748_solve refuses an empty bracket whatever the concentration
                const double critic = I.limit * I.sa + iso_f(I, I.limit / I.rho) * I.ss;
                if (!(critic / I.sa + 0.00001 > 0.0))
                    throw Error(kErrSorption, "sorption of species " + std::to_string(k) + " at element " +
                                                  std::to_string(m_.el_id[i]) +
                                                  " is solved per element and has an empty root bracket "
                                                  "(no positive solubility limit)");
                par.push_back(I);
            }
        }
    }
    if (tab.empty()) tab.push_back(0.0);
    d_rs_el_par_ = upv(owned_, el_par, stream_);
    d_rs_subst_ = upv(owned_, rs_subst_, stream_);
    IsoPar *dp = dev_alloc<IsoPar>(par.size());
    owned_.push_back(dp);
    FL_HIP(hipMemcpyAsync(dp, par.data(), par.size() * sizeof(IsoPar), hipMemcpyHostToDevice, stream_));
    d_rs_par_ = (const double *)dp;
    d_rs_table_ = upv(owned_, tab, stream_);
    if (!d_rs_sorbed_) {
        std::vector<double> zero((size_t)E_ * n, 0.0);
        d_rs_sorbed_ = upv(owned_, zero, stream_);
    }
    if (!d_rs_err_) {
        d_rs_err_ = dev_alloc<int>(2);
        owned_.push_back(d_rs_err_);
    }
    const int init[2] = {0, INT_MAX};
    FL_HIP(hipMemcpyAsync(d_rs_err_, init, sizeof(init), hipMemcpyHostToDevice, stream_));
    FL_HIP(hipStreamSynchronize(stream_));
    rs_built_ = true;
}

// The reference stops at a negative total mass (xprintf(UsrErr), isotherm.cc:62)
// and where toms748_solve finds no bracketed root (it throws); the steps already
// queued have run, so the failure is reported when the update returns.
void TransportGPU::check_sorption() {
    int e[2] = {0, INT_MAX};
    FL_HIP(hipMemcpyAsync(e, d_rs_err_, sizeof(e), hipMemcpyDeviceToHost, stream_));
    FL_HIP(hipStreamSynchronize(stream_));
    if (e[0] == 0) return;
    const int init[2] = {0, INT_MAX};
    FL_HIP(hipMemcpyAsync(d_rs_err_, init, sizeof(init), hipMemcpyHostToDevice, stream_));
    FL_HIP(hipStreamSynchronize(stream_));
    std::string what = (e[0] & kSorpNegMass) ? "a negative total mass" : "";
    if (e[0] & kSorpNoRoot) what += std::string(what.empty() ? "" : " and ") + "an isotherm equation without a root in its bracket";
    throw Error(kErrSorption, "sorption: " + what + " at element " + std::to_string(m_.el_id[e[1]]) +
                                  " (the first failing element; its concentrations were left as they were)");
}

void TransportGPU::get_reaction_sorbed(double *out) const {
    const size_t n = (size_t)E_ * rs_n_;
    std::vector<double> c(n, 0.0);
    if (n && d_rs_sorbed_) {
        FL_HIP(hipMemcpyAsync(c.data(), d_rs_sorbed_, n * sizeof(double), hipMemcpyDeviceToHost, stream_));
        FL_HIP(hipStreamSynchronize(stream_));
    }
    for (int k = 0; k < rs_n_; k++)
        for (int i = 0; i < E_; i++) out[(size_t)k * E_ + i] = c[(size_t)i * rs_n_ + k];
}

void TransportGPU::get_conc_phase(int phase, double *out) const {
    if (phase < 0 || phase > 3) throw Error(-3, "phase must be 0..3");
    const double *src = phase == 0 ? (cur_is_c0_ ? d_c0_ : d_c1_) : phase == 1 ? d_imm_ : phase == 2 ? d_msorb_ : d_isorb_;
    const size_t n = (size_t)E_ * nsub_;
    if (!src) {
        std::fill(out, out + n, 0.0);
        return;
    }
    std::vector<double> c(n);
    FL_HIP(hipMemcpyAsync(c.data(), src, n * sizeof(double), hipMemcpyDeviceToHost, stream_));
    FL_HIP(hipStreamSynchronize(stream_));
    for (int k = 0; k < nsub_; k++)
        for (int i = 0; i < E_; i++) out[(size_t)k * E_ + i] = c[(size_t)i * nsub_ + k];
}

void TransportGPU::get_conc(double *out) const {
    std::vector<double> c((size_t)E_ * nsub_);
    FL_HIP(hipMemcpyAsync(c.data(), cur_is_c0_ ? d_c0_ : d_c1_, c.size() * sizeof(double), hipMemcpyDeviceToHost,
                          stream_));
    FL_HIP(hipStreamSynchronize(stream_));
    for (int k = 0; k < nsub_; k++)
        for (int i = 0; i < E_; i++) out[(size_t)k * E_ + i] = c[(size_t)i * nsub_ + k];
}

void TransportGPU::get_matrix(int *ptr, int *col, double *val) const {
    FL_HIP(hipMemcpyAsync(ptr, t_.ptr, (E_ + 1) * sizeof(int), hipMemcpyDeviceToHost, stream_));
    FL_HIP(hipMemcpyAsync(col, t_.col, nnz_ * sizeof(int), hipMemcpyDeviceToHost, stream_));
    FL_HIP(hipMemcpyAsync(val, t_.val, nnz_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
    FL_HIP(hipStreamSynchronize(stream_));
}

// Per-region report of the last step (MassBalance::calculate/output at an
// output time): boundary fluxes with the previous concentrations, mass with
// the current ones; the cumulative totals come from the per-step reductions
// on the device.
void TransportGPU::mass_balance(double *bal, double *plus, double *minus, double *mass, double *src,
                                double *totals) {
    const HostMesh &m = m_;
    int nbr = m.n_boundary_regions, nblk = m.n_bulk_regions, ns = nsub_;
    std::vector<double> cur((size_t)E_ * ns), prev((size_t)E_ * ns), mb(1 + 2 * ns);
    FL_HIP(hipMemcpyAsync(cur.data(), cur_is_c0_ ? d_c0_ : d_c1_, cur.size() * 8, hipMemcpyDeviceToHost, stream_));
    FL_HIP(hipMemcpyAsync(prev.data(), cur_is_c0_ ? d_c1_ : d_c0_, prev.size() * 8, hipMemcpyDeviceToHost, stream_));
    FL_HIP(hipMemcpyAsync(mb.data(), t_.mb, mb.size() * 8, hipMemcpyDeviceToHost, stream_));
    std::vector<double> flux(m.n_sides());
    FL_HIP(hipMemcpyAsync(flux.data(), d_flux_, flux.size() * 8, hipMemcpyDeviceToHost, stream_));
    // the mass counts every phase in use (transport.cc:1354-1362)
    std::vector<double> imm, msorb, isorb;
    if (dual_) {
        imm.resize(cur.size());
        FL_HIP(hipMemcpyAsync(imm.data(), d_imm_, imm.size() * 8, hipMemcpyDeviceToHost, stream_));
    }
    if (sorp_) {
        msorb.resize(cur.size());
        FL_HIP(hipMemcpyAsync(msorb.data(), d_msorb_, msorb.size() * 8, hipMemcpyDeviceToHost, stream_));
    }
    if (dual_ && sorp_) {
        isorb.resize(cur.size());
        FL_HIP(hipMemcpyAsync(isorb.data(), d_isorb_, isorb.size() * 8, hipMemcpyDeviceToHost, stream_));
    }
    FL_HIP(hipStreamSynchronize(stream_));
    if (stats_.steps == 0) std::fill(prev.begin(), prev.end(), 0.0);
    std::fill(bal, bal + ns * nbr, 0.0);
    std::fill(plus, plus + ns * nbr, 0.0);
    std::fill(minus, minus + ns * nbr, 0.0);
    std::fill(mass, mass + ns * nblk, 0.0);
    std::fill(src, src + ns * nblk, 0.0);
    for (int b = 0; b < m.n_boundaries(); b++) {
        int sd = m.edge_side[m.edge_ptr[m.bnd_edge[b]]];
        int i = m.side_elem[sd];
        double wf = flux[sd];
        int bce = m.bnd_bcele[b];
        int r = m.regions[m.bc_region[bce]].sub_idx;
        for (int k = 0; k < ns; k++) {
            double mf = wf < 0 ? wf * h_bc_conc_[(size_t)k * NBC_ + bce] : wf * prev[(size_t)i * ns + k];
            bal[k * nbr + r] += mf;
            if (mf > 0) plus[k * nbr + r] += mf;
            else minus[k * nbr + r] += mf;
        }
    }
    for (int k = 0; k < ns; k++)
        for (int i = 0; i < E_; i++) {
            int r = h_el_region_sub_[i];
            const size_t j = (size_t)i * ns + k;
            double sum_sol_phases = 0;
            sum_sol_phases += cur[j];
            if (dual_) sum_sol_phases += imm[j];
            if (sorp_) sum_sol_phases += msorb[j];
            if (dual_ && sorp_) sum_sol_phases += isorb[j];
            mass[k * nblk + r] += h_por_[i] * h_cs_[i] * sum_sol_phases * h_meas_[i];
            if (t_.has_src) {  // before the first step the previous concentration is zero (transport.cc:433-435)
                size_t j = (size_t)k * E_ + i;
                double diff = h_src_conc_[j] - prev[(size_t)i * ns + k];
                double snd = diff > 0.0 ? (h_src_density_[j] + diff * h_src_sigma_[j]) : h_src_density_[j];
                src[k * nblk + r] += snd * h_meas_[i];
            }
        }
    for (int k = 0; k < ns; k++) {
        double mt = 0.0;
        for (int r = 0; r < nblk; r++) mt += mass[k * nblk + r];
        totals[4 * k + 0] = k < (int)initial_mass_.size() ? initial_mass_[k] : mt;  // first output: current mass
        totals[4 * k + 1] = mb[1 + k];
        totals[4 * k + 2] = mb[1 + ns + k];
        totals[4 * k + 3] = mt;
    }
}

}  // namespace fl
