// Per-element RT0 condensation shared by the single and the batched solvers:
// local matrix, its inverse and the quantities of the closed-form two-stage
// Schur elimination (assembly_steady_mh_matrix darcy_flow_mh.cc:525-694 +
// make_schur1/2 :1009-1131).
#pragma once
#include "common.hh"
#include "darcy.hh"
#include "rt0_local.hh"

namespace fl {

// ---------------------------------------------------------------------------
// Element condensation (assembly_steady_mh_matrix darcy_flow_mh.cc:525-694 +
// make_schur1/2 :1009-1131 in closed form)
// ---------------------------------------------------------------------------
template <int D>
struct Local {
    double inv[(D + 1) * (D + 1)];
    double a[D + 1], Ag[D + 1], c[D + 1], dlt[D + 1], rf[D + 1];
    double s, h, meas;
};

template <int D>
__device__ __forceinline__ void element_local(const DarcyDev &dv, int e, double K, Local<D> &L) {
    Elem<D> el;
    for (int i = 0; i <= D; i++) {
        int n = dv.el_nodes[4 * e + i];
        el.P[i] = V3{dv.xyz[3 * n], dv.xyz[3 * n + 1], dv.xyz[3 * n + 2]};
    }
    double cs = dv.cs[e];
    double scale = 1 / K / cs;  // mh_fe_values.cc:38
    double R[9];
    const double *Rp = nullptr;
    if (dv.aniso) {
        inv3x3(dv.aniso + 9 * (size_t)e, R);
        Rp = R;
    }
    double loc[(D + 1) * (D + 1)];
    local_matrix<D>(el, scale, Rp, loc);
    const double det = small_inverse<D + 1>(loc, L.inv);
    // a singular local matrix stops the reference (mh_fe_values.cc:54-59); here
    // the smallest such element is recorded and the host reports it
    if (dv.bad && !(fabs(det) >= kNumZero)) atomicMin(dv.bad, e);
    V3 ec = elem_centre<D>(el);
    int so = dv.side_off[e];
    double g[D + 1];
    for (int i = 0; i <= D; i++) {
        V3 sc = side_centre<D>(el, i);
        g[i] = ec.z - sc.z;  // gravity term, darcy_flow_mh.cc:565
        L.c[i] = 1.0;
        L.dlt[i] = 0.0;
        L.rf[i] = 0.0;
        int b = dv.side_bnd[so + i];
        if (b >= 0) {
            int bce = dv.bnd_bcele[b];
            int type = dv.bc_type[bce];
            double bm = dv.bc_measure[bce];
            if (type == 1) {  // dirichlet :575-581
                L.c[i] = 0.0;
                g[i] -= dv.bc_p[bce];
                L.dlt[i] = 1.0;
                L.rf[i] = dv.bc_p[bce];
            } else if (type == 2) {  // neumann :582-584
                L.rf[i] = -(dv.bc_flux[bce] * bm * cs);
            } else if (type == 3) {  // robin :586-591
                L.rf[i] = -(-bm * dv.bc_sigma[bce] * dv.bc_p[bce] * cs);
                L.dlt[i] = -(-bm * dv.bc_sigma[bce] * cs);
            }
        }
    }
    double s = 0.0;
    for (int i = 0; i <= D; i++) {
        double ai = 0.0, agi = 0.0;
        for (int j = 0; j <= D; j++) {
            ai += L.inv[i * (D + 1) + j];
            agi += L.inv[i * (D + 1) + j] * g[j];
        }
        L.a[i] = ai;
        L.Ag[i] = agi;
        s += ai;
    }
    L.s = s;
    L.meas = elem_measure<D>(el);
    double sumAg = 0.0;
    for (int i = 0; i <= D; i++) sumAg += L.Ag[i];
    L.h = L.meas * cs * dv.f[e] - sumAg;
}

__device__ __forceinline__ double vb_coef(const DarcyDev &dv, int e, int n) {
    // darcy_flow_mh.cc:630-632: sigma(lower) * |higher side| * cs(higher)
    return dv.sigma[e] * dv.vb_hmeas[n] * dv.cs[dv.vb_hel[n]];
}


}  // namespace fl
