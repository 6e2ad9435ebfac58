// Per-element RT0 local matrix and its inverse, evaluated on the GPU with the
// operation order of MHFEValues (src/flow/mh_fe_values.cc:34-514) and
// MatrixInverse (src/system/math_fce.cc:124-230), so that the device values
// agree with a CPU restatement to the last bits.  The translation unit that
// includes this header is compiled with -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>

namespace fl {

struct V3 {
    double x, y, z;
};
__device__ __forceinline__ V3 vsub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 vcross(V3 u, V3 v) {
    return V3{u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x};
}
__device__ __forceinline__ double vdot(V3 u, V3 v) { return u.x * v.x + u.y * v.y + u.z * v.z; }
__device__ __forceinline__ double vnorm(V3 u) { return sqrt(u.x * u.x + u.y * u.y + u.z * u.z); }
__device__ __forceinline__ V3 vnormalize(V3 u) {
    double l = vnorm(u);
    return V3{u.x / l, u.y / l, u.z / l};
}
__device__ __forceinline__ double vcomp(V3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// RefElement<dim>::side_nodes (ref_element.cc:51-74)
__device__ __forceinline__ int side_node(int dim, int s, int k) {
    if (dim == 1) return s;
    if (dim == 2) return (s == 0) ? k : (s == 1 ? 1 + k : 2 * k);
    // dim 3: {1,2,3},{0,2,3},{0,1,3},{0,1,2}
    int skip = s;
    int idx = k < skip ? k : k + 1;
    return idx;
}

template <int D>
struct Elem {
    V3 P[D + 1];
};

template <int D>
__device__ __forceinline__ double elem_measure(const Elem<D> &el) {
    if constexpr (D == 1) return vnorm(vsub(el.P[1], el.P[0]));
    else if constexpr (D == 2) return vnorm(vcross(vsub(el.P[1], el.P[0]), vsub(el.P[2], el.P[0]))) / 2.0;
    else {
        V3 c = vcross(vsub(el.P[1], el.P[0]), vsub(el.P[2], el.P[0]));
        return fabs(vdot(c, vsub(el.P[3], el.P[0]))) / 6.0;
    }
}
template <int D>
__device__ __forceinline__ V3 elem_centre(const Elem<D> &el) {
    double x = 0.0, y = 0.0, z = 0.0;
    for (int i = 0; i <= D; i++) { x += el.P[i].x; y += el.P[i].y; z += el.P[i].z; }
    double n = (double)(D + 1);
    return V3{x / n, y / n, z / n};
}
template <int D>
__device__ __forceinline__ V3 side_centre(const Elem<D> &el, int s) {
    double x = 0.0, y = 0.0, z = 0.0;
    for (int k = 0; k < D; k++) {
        V3 p = el.P[side_node(D, s, k)];
        x += p.x; y += p.y; z += p.z;
    }
    double n = (double)D;
    return V3{x / n, y / n, z / n};
}

// inverse of a 3x3 (anisotropy tensor), cofactor form, row-major
__device__ __forceinline__ void inv3x3(const double *a, double *r) {
    double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
    double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    r[0] = c00 / det; r[3] = c01 / det; r[6] = c02 / det;
    r[1] = (a[2] * a[7] - a[1] * a[8]) / det;
    r[4] = (a[0] * a[8] - a[2] * a[6]) / det;
    r[7] = (a[1] * a[6] - a[0] * a[7]) / det;
    r[2] = (a[1] * a[5] - a[2] * a[4]) / det;
    r[5] = (a[2] * a[3] - a[0] * a[5]) / det;
    r[8] = (a[0] * a[4] - a[1] * a[3]) / det;
}

// loc[(D+1)^2] row-major; R = resistance (inverse anisotropy) or nullptr for identity
template <int D>
__device__ __forceinline__ void local_matrix(const Elem<D> &el, double scale, const double *Rin, double *loc) {
    double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (Rin)
        for (int i = 0; i < 9; i++) R[i] = Rin[i];
    if constexpr (D == 1) {  // local_matrix_line
        V3 t = vsub(el.P[1], el.P[0]);
        double l = vnorm(t);
        t = V3{t.x / l, t.y / l, t.z / l};
        V3 Rt{R[0] * t.x + R[1] * t.y + R[2] * t.z, R[3] * t.x + R[4] * t.y + R[5] * t.z,
              R[6] * t.x + R[7] * t.y + R[8] * t.z};
        double val = scale * vdot(t, Rt) * elem_measure<D>(el) / 3.0;
        loc[0] = val; loc[3] = val; loc[1] = -val / 2.0; loc[2] = -val / 2.0;
    } else if constexpr (D == 2) {  // local_matrix_triangle
        V3 ex = vnormalize(vsub(el.P[1], el.P[0]));
        V3 ac = vsub(el.P[2], el.P[0]);
        V3 ez = vnormalize(vcross(ex, ac));
        V3 ey = vnormalize(vcross(ez, ex));
        double Rr0[3], Rr1[3];
        for (int i = 0; i < 3; i++) {
            Rr0[i] = R[3 * i] * ex.x + R[3 * i + 1] * ex.y + R[3 * i + 2] * ex.z;
            Rr1[i] = R[3 * i] * ey.x + R[3 * i + 1] * ey.y + R[3 * i + 2] * ey.z;
        }
        double rt00 = ex.x * Rr0[0] + ex.y * Rr0[1] + ex.z * Rr0[2];
        double rt01 = ex.x * Rr1[0] + ex.y * Rr1[1] + ex.z * Rr1[2];
        double rt10 = ey.x * Rr0[0] + ey.y * Rr0[1] + ey.z * Rr0[2];
        double rt11 = ey.x * Rr1[0] + ey.y * Rr1[1] + ey.z * Rr1[2];
        // node_coordinates_triangle
        V3 u = vsub(el.P[1], el.P[0]), v = vsub(el.P[2], el.P[0]);
        V3 n = vnormalize(vcross(u, v));
        V3 tt = vnormalize(u);
        V3 b = vnormalize(vcross(n, u));
        double nod[3][2] = {{0.0, 0.0}, {vdot(tt, u), vdot(b, u)}, {vdot(tt, v), vdot(b, v)}};
        double mid[3][2];
        mid[0][0] = (nod[0][0] + nod[1][0]) / 2.0; mid[0][1] = (nod[0][1] + nod[1][1]) / 2.0;
        mid[1][0] = (nod[1][0] + nod[2][0]) / 2.0; mid[1][1] = (nod[1][1] + nod[2][1]) / 2.0;
        mid[2][0] = (nod[2][0] + nod[0][0]) / 2.0; mid[2][1] = (nod[2][1] + nod[0][1]) / 2.0;
        double al[3], be[3], ga[3];
        for (int i = 0; i < 3; i++) {
            int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
            double x0 = nod[i][0], y0 = nod[i][1], x1 = nod[i1][0], y1 = nod[i1][1], x2 = nod[i2][0], y2 = nod[i2][1];
            al[i] = x2; be[i] = y2;
            ga[i] = 1.0 / fabs((x0 - x2) * (y1 - y0) + (y0 - y2) * (x0 - x1));
        }
        double meas = elem_measure<D>(el);
        for (int i = 0; i < 3; i++)
            for (int j = i; j < 3; j++) {
                double p0 = rt00 * al[i] * al[j] + rt01 * be[i] * al[j] + rt10 * al[i] * be[j] + rt11 * be[i] * be[j];
                double p1 = (rt00 * al[i] + rt00 * al[j] + rt01 * be[i] + rt10 * be[j]) * -1.0;
                double p2 = (rt11 * be[i] + rt11 * be[j] + rt10 * al[i] + rt01 * al[j]) * -1.0;
                double p3 = rt00, p4 = rt01 + rt10, p5 = rt11;
                double pv[3];
                for (int q = 0; q < 3; q++)
                    pv[q] = p0 + p1 * mid[q][0] + p2 * mid[q][1] + p3 * mid[q][0] * mid[q][0] +
                            p4 * mid[q][0] * mid[q][1] + p5 * mid[q][1] * mid[q][1];
                double val = scale * ga[i] * ga[j] * meas * (pv[0] + pv[1] + pv[2]) / 3.0;
                loc[3 * i + j] = val;
                loc[3 * j + i] = val;
            }
    } else {  // local_matrix_tetrahedron
        V3 ec = elem_centre<D>(el);
        double al[4], be[4], ga[4], de[4];
        for (int li = 0; li < 4; li++) {
            al[li] = el.P[li].x; be[li] = el.P[li].y; ga[li] = el.P[li].z;
            V3 s0 = el.P[side_node(3, li, 0)], s1 = el.P[side_node(3, li, 1)], s2 = el.P[side_node(3, li, 2)];
            V3 nrm = vnormalize(vcross(vsub(s1, s0), vsub(s2, s0)));
            if (vdot(nrm, vsub(ec, s0)) > 0.0) nrm = V3{-nrm.x, -nrm.y, -nrm.z};
            V3 sc = side_centre<D>(el, li);
            double sm = 0.5 * vnorm(vcross(vsub(s1, s0), vsub(s2, s0)));
            de[li] = 1.0 / (sm * (nrm.x * sc.x + nrm.y * sc.y + nrm.z * sc.z - nrm.x * al[li] - nrm.y * be[li] -
                                  nrm.z * ga[li]));
        }
        double v = elem_measure<D>(el);
        double t[3] = {ec.x, ec.y, ec.z};
        auto mom = [&](int a, int b) {
            return t[a] * t[b] + 1.0 / 20.0 *
                                     ((vcomp(el.P[0], a) - t[a]) * (vcomp(el.P[0], b) - t[b]) +
                                      (vcomp(el.P[1], a) - t[a]) * (vcomp(el.P[1], b) - t[b]) +
                                      (vcomp(el.P[2], a) - t[a]) * (vcomp(el.P[2], b) - t[b]) +
                                      (vcomp(el.P[3], a) - t[a]) * (vcomp(el.P[3], b) - t[b]));
        };
        double Mxy = mom(0, 1), Mxz = mom(0, 2), Myz = mom(1, 2), Mxx = mom(0, 0), Myy = mom(1, 1), Mzz = mom(2, 2);
#define A_(i, j) R[3 * (i) + (j)]
        for (int i = 0; i < 4; i++)
            for (int j = i; j < 4; j++) {
                double p0 = A_(0, 0) * al[i] * al[j] + A_(0, 1) * be[i] * al[j] + A_(0, 2) * ga[i] * al[j] +
                            A_(1, 0) * al[i] * be[j] + A_(1, 1) * be[i] * be[j] + A_(1, 2) * ga[i] * be[j] +
                            A_(2, 0) * al[i] * ga[j] + A_(2, 1) * be[i] * ga[j] + A_(2, 2) * ga[i] * ga[j];
                double p1 = (A_(0, 0) * al[i] + A_(0, 1) * be[i] + A_(0, 2) * ga[i] + A_(0, 0) * al[j] +
                             A_(1, 0) * be[j] + A_(2, 0) * ga[j]) * -1.0;
                double p2 = (A_(1, 0) * al[i] + A_(1, 1) * be[i] + A_(1, 2) * ga[i] + A_(0, 1) * al[j] +
                             A_(1, 1) * be[j] + A_(2, 1) * ga[j]) * -1.0;
                double p3 = (A_(2, 0) * al[i] + A_(2, 1) * be[i] + A_(2, 2) * ga[i] + A_(0, 2) * al[j] +
                             A_(1, 2) * be[j] + A_(2, 2) * ga[j]) * -1.0;
                double p4 = A_(0, 1) + A_(1, 0), p5 = A_(0, 2) + A_(2, 0), p6 = A_(1, 2) + A_(2, 1);
                double rc = 0.0;
                rc += p0 * v;
                rc += p1 * v * t[0];
                rc += p2 * v * t[1];
                rc += p3 * v * t[2];
                rc += p4 * v * Mxy;
                rc += p5 * v * Mxz;
                rc += p6 * v * Myz;
                rc += A_(0, 0) * v * Mxx;
                rc += A_(1, 1) * v * Myy;
                rc += A_(2, 2) * v * Mzz;
                double val = scale * de[i] * de[j] * rc;
                loc[4 * i + j] = val;
                loc[4 * j + i] = val;
            }
#undef A_
    }
}

// MHFEValues::RT0_value (mh_fe_values.cc:76-118): the RT0 basis function of
// side `face` at point x, with the basis coefficients local_matrix_triangle /
// local_matrix_tetrahedron set up (the triangle's in its own plane frame).
template <int D>
__device__ __forceinline__ V3 rt0_value(const Elem<D> &el, V3 x, int face) {
    if constexpr (D == 1) {
        V3 lv = vsub(el.P[1], el.P[0]);
        double ll = vdot(lv, lv);
        double s = face == 0 ? -vnorm(vsub(x, el.P[1])) : vnorm(vsub(x, el.P[0]));
        return V3{s * lv.x / ll, s * lv.y / ll, s * lv.z / ll};
    } else if constexpr (D == 2) {
        V3 ex = vnormalize(vsub(el.P[1], el.P[0]));
        V3 ac = vsub(el.P[2], el.P[0]);
        V3 ez = vnormalize(vcross(ex, ac));
        V3 ey = vnormalize(vcross(ez, ex));
        // node coordinates in the plane frame (node_coordinates_triangle)
        V3 u1 = vsub(el.P[1], el.P[0]), u2 = vsub(el.P[2], el.P[0]);
        V3 n = vnormalize(vcross(u1, u2));
        V3 tt = vnormalize(u1);
        V3 b = vnormalize(vcross(n, u1));
        double nod[3][2] = {{0.0, 0.0}, {vdot(tt, u1), vdot(b, u1)}, {vdot(tt, u2), vdot(b, u2)}};
        const int i = face, i1 = (face + 1) % 3, i2 = (face + 2) % 3;
        double x0 = nod[i][0], y0 = nod[i][1], x1 = nod[i1][0], y1 = nod[i1][1], x2 = nod[i2][0], y2 = nod[i2][1];
        double al = x2, be = y2, ga = 1.0 / fabs((x0 - x2) * (y1 - y0) + (y0 - y2) * (x0 - x1));
        V3 u = vsub(x, el.P[0]);
        double cx = vdot(u, ex) - al, cy = vdot(u, ey) - be;
        return V3{ga * (cx * ex.x + cy * ey.x), ga * (cx * ex.y + cy * ey.y), ga * (cx * ex.z + cy * ey.z)};
    } else {
        V3 ec = elem_centre<D>(el);
        V3 s0 = el.P[side_node(3, face, 0)], s1 = el.P[side_node(3, face, 1)], s2 = el.P[side_node(3, face, 2)];
        V3 nrm = vnormalize(vcross(vsub(s1, s0), vsub(s2, s0)));
        if (vdot(nrm, vsub(ec, s0)) > 0.0) nrm = V3{-nrm.x, -nrm.y, -nrm.z};
        V3 sc = side_centre<D>(el, face);
        V3 pf = el.P[face];
        double sm = 0.5 * vnorm(vcross(vsub(s1, s0), vsub(s2, s0)));
        double de = 1.0 / (sm * (nrm.x * sc.x + nrm.y * sc.y + nrm.z * sc.z - nrm.x * pf.x - nrm.y * pf.y -
                                 nrm.z * pf.z));
        return V3{de * (x.x - pf.x), de * (x.y - pf.y), de * (x.z - pf.z)};
    }
}

// MatrixInverse (math_fce.cc:124-230); returns det
template <int N>
__device__ __forceinline__ double small_inverse(const double *a, double *b) {
#define A_(i, j) a[(i) * N + (j)]
#define B_(i, j) b[(i) * N + (j)]
#define SD2(i, j, k, l) (A_(i, k) * A_(j, l) - A_(i, l) * A_(j, k))
    if constexpr (N == 2) {
        double det = SD2(0, 1, 0, 1);
        B_(0, 0) = A_(1, 1) / det; B_(1, 0) = -A_(1, 0) / det;
        B_(0, 1) = -A_(0, 1) / det; B_(1, 1) = A_(0, 0) / det;
        return det;
    } else if constexpr (N == 3) {
        B_(0, 0) = SD2(1, 2, 1, 2); B_(1, 0) = -SD2(1, 2, 0, 2); B_(2, 0) = SD2(1, 2, 0, 1);
        double det = A_(0, 0) * B_(0, 0) + A_(0, 1) * B_(1, 0) + A_(0, 2) * B_(2, 0);
        B_(0, 0) /= det; B_(1, 0) /= det; B_(2, 0) /= det;
        B_(0, 1) = -SD2(0, 2, 1, 2) / det; B_(1, 1) = SD2(0, 2, 0, 2) / det; B_(2, 1) = -SD2(0, 2, 0, 1) / det;
        B_(0, 2) = SD2(0, 1, 1, 2) / det; B_(1, 2) = -SD2(0, 1, 0, 2) / det; B_(2, 2) = SD2(0, 1, 0, 1) / det;
        return det;
    } else {
        double u[6], l[6];
        u[0] = SD2(0, 1, 0, 1); u[1] = SD2(0, 1, 0, 2); u[2] = SD2(0, 1, 0, 3);
        u[3] = SD2(0, 1, 1, 2); u[4] = SD2(0, 1, 1, 3); u[5] = SD2(0, 1, 2, 3);
        l[0] = SD2(2, 3, 0, 1); l[1] = SD2(2, 3, 0, 2); l[2] = SD2(2, 3, 0, 3);
        l[3] = SD2(2, 3, 1, 2); l[4] = SD2(2, 3, 1, 3); l[5] = SD2(2, 3, 2, 3);
        B_(0, 0) = +A_(1, 1) * l[5] - A_(1, 2) * l[4] + A_(1, 3) * l[3];
        B_(1, 0) = -A_(1, 0) * l[5] + A_(1, 2) * l[2] - A_(1, 3) * l[1];
        B_(2, 0) = +A_(1, 0) * l[4] - A_(1, 1) * l[2] + A_(1, 3) * l[0];
        B_(3, 0) = -A_(1, 0) * l[3] + A_(1, 1) * l[1] - A_(1, 2) * l[0];
        double det = A_(0, 0) * B_(0, 0) + A_(0, 1) * B_(1, 0) + A_(0, 2) * B_(2, 0) + A_(0, 3) * B_(3, 0);
        B_(0, 0) /= det; B_(1, 0) /= det; B_(2, 0) /= det; B_(3, 0) /= det;
        B_(0, 1) = (-A_(0, 1) * l[5] + A_(0, 2) * l[4] - A_(0, 3) * l[3]) / det;
        B_(1, 1) = (+A_(0, 0) * l[5] - A_(0, 2) * l[2] + A_(0, 3) * l[1]) / det;
        B_(2, 1) = (-A_(0, 0) * l[4] + A_(0, 1) * l[2] - A_(0, 3) * l[0]) / det;
        B_(3, 1) = (+A_(0, 0) * l[3] - A_(0, 1) * l[1] + A_(0, 2) * l[0]) / det;
        B_(0, 2) = (+A_(3, 1) * u[5] - A_(3, 2) * u[4] + A_(3, 3) * u[3]) / det;
        B_(1, 2) = (-A_(3, 0) * u[5] + A_(3, 2) * u[2] - A_(3, 3) * u[1]) / det;
        B_(2, 2) = (+A_(3, 0) * u[4] - A_(3, 1) * u[2] + A_(3, 3) * u[0]) / det;
        B_(3, 2) = (-A_(3, 0) * u[3] + A_(3, 1) * u[1] - A_(3, 2) * u[0]) / det;
        B_(0, 3) = (-A_(2, 1) * u[5] + A_(2, 2) * u[4] - A_(2, 3) * u[3]) / det;
        B_(1, 3) = (+A_(2, 0) * u[5] - A_(2, 2) * u[2] + A_(2, 3) * u[1]) / det;
        B_(2, 3) = (-A_(2, 0) * u[4] + A_(2, 1) * u[2] - A_(2, 3) * u[0]) / det;
        B_(3, 3) = (+A_(2, 0) * u[3] - A_(2, 1) * u[1] + A_(2, 2) * u[0]) / det;
        return det;
    }
#undef SD2
#undef B_
#undef A_
}

}  // namespace fl
