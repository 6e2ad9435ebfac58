// ============================================================================
// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A CPU, FP64, single-threaded restatement of the Flow123d 1.7 algorithm for
// the steady mixed-hybrid Darcy solve and the explicit upwind transport step.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load this library, and only as the checker (or as the timed CPU baseline).
// The product library never links or calls it.
//
// Every routine cites the reference file:line it follows (paths relative to
// the Flow123d source tree, src/...).  The restatement follows the reference
// step by step, including its two-stage sparse Schur elimination (it does NOT
// use the per-element closed form the GPU path uses), so that the GPU path is
// checked against an independent derivation.
//
// Parity pins (see tests/): simplest_cube topology counts
// (test_units/mesh/mesh_test.cpp:53-81), tests/01 raw_output.txt goldens,
// tests/06 analytic coupling values, the MH-RT0 linear patch test, local
// matrix closed forms, dense brute-force solves of the full MH system, and
// tests/13 transport/mass-balance goldens.
//
// Not pinned bit-for-bit: the summation order inside Armadillo 3.2.4
// (arma::norm/dot) and PETSc 3.4 (MatMatMult, KSP), neither of which is
// vendored with the reference; agreement there is to round-off.
// ============================================================================
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

namespace orc {

static thread_local std::string g_err;

// ---------------------------------------------------------------------------
// Regions: src/mesh/region.cc:105-171 (boundary by '.' prefix, separate
// boundary / bulk counters, "region_<id>" for unnamed, IMPLICIT BOUNDARY
// id = undefined-2 added on first use, region.cc:85-86).
// ---------------------------------------------------------------------------
struct Region {
    int id;
    std::string label;
    bool boundary;
    int sub_idx;  // boundary_idx or bulk_idx
};

struct RegionDB {
    std::vector<Region> regs;
    std::map<int, int> by_id;
    int n_boundary = 0, n_bulk = 0;
    int add(int id, const std::string &label, bool boundary) {
        auto it = by_id.find(id);
        if (it != by_id.end()) return it->second;
        Region r{id, label, boundary, boundary ? n_boundary++ : n_bulk++};
        regs.push_back(r);
        by_id[id] = (int)regs.size() - 1;
        return (int)regs.size() - 1;
    }
    int add_named(int id, const std::string &label) {
        bool b = !label.empty() && label[0] == '.';
        return add(id, label, b);
    }
    int add_unnamed(int id) {
        auto it = by_id.find(id);
        if (it != by_id.end()) return it->second;
        return add(id, "region_" + std::to_string(id), false);
    }
    int implicit_boundary() { return add(-3, "IMPLICIT BOUNDARY", true); }
};

// ---------------------------------------------------------------------------
// Mesh: src/mesh/mesh.h, elements.h, edges.h, boundaries.h, neighbours.h
// ---------------------------------------------------------------------------
struct Element {
    int id = 0, dim = 0, region = -1;
    int node[4] = {-1, -1, -1, -1};
    int edge_idx[4] = {-1, -1, -1, -1};
    int boundary_idx[4] = {-1, -1, -1, -1};
    std::vector<int> neigh_vb;  // indices into Mesh::vb
    int n_nodes() const { return dim + 1; }
    int n_sides() const { return dim + 1; }
};
struct Edge {
    std::vector<std::pair<int, int>> sides;  // (element index, local side)
};
struct Boundary {
    int edge_idx, bc_ele_idx;
};
struct Neighbour {
    int element;  // lower-dimensional element
    int edge_idx;
};

// RefElement<dim>::side_nodes, src/mesh/ref_element.cc:51-74
static const int SIDE_NODES_1[2][1] = {{0}, {1}};
static const int SIDE_NODES_2[3][2] = {{0, 1}, {1, 2}, {0, 2}};
static const int SIDE_NODES_3[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
static inline int side_node(int dim, int s, int k) {
    if (dim == 1) return SIDE_NODES_1[s][k];
    if (dim == 2) return SIDE_NODES_2[s][k];
    return SIDE_NODES_3[s][k];
}

struct Mesh {
    std::vector<double> X;  // 3 * n_nodes
    std::vector<int> node_ids;
    std::unordered_map<int, int> node_idx;
    std::vector<Element> element, bc_elements;
    int n_file_bc = 0;
    std::vector<Edge> edges;
    std::vector<Boundary> boundary;
    std::vector<Neighbour> vb;
    RegionDB rdb;
    std::vector<std::vector<int>> node_elements;
    std::vector<int> side_offset;  // MH_DofHandler: first global side of element
    int n_sides = 0;

    const double *P(int node) const { return &X[3 * node]; }
};

// ----- small vector helpers (system/math_fce.cc:46-98) -----
static inline void sub3(const double *a, const double *b, double *r) {
    r[0] = a[0] - b[0]; r[1] = a[1] - b[1]; r[2] = a[2] - b[2];
}
static inline void cross3(const double *u, const double *v, double *x) {
    x[0] = u[1] * v[2] - u[2] * v[1];
    x[1] = u[2] * v[0] - u[0] * v[2];
    x[2] = u[0] * v[1] - u[1] * v[0];
}
static inline double dot3(const double *u, const double *v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; }
static inline double norm3(const double *u) { return std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]); }
static inline void normalize3(double *u) {
    double l = norm3(u);
    u[0] /= l; u[1] /= l; u[2] /= l;
}

// ---------------------------------------------------------------------------
// Geometry: Element::measure/centre src/mesh/elements.cc:108-151;
// Side::measure/normal/centre src/mesh/sides.cc:42-146.
// ---------------------------------------------------------------------------
static double elem_measure(const Mesh &m, const Element &e) {
    switch (e.dim) {
        case 0: return 1.0;
        case 1: { double d[3]; sub3(m.P(e.node[1]), m.P(e.node[0]), d); return norm3(d); }
        case 2: {
            double a[3], b[3], c[3];
            sub3(m.P(e.node[1]), m.P(e.node[0]), a);
            sub3(m.P(e.node[2]), m.P(e.node[0]), b);
            cross3(a, b, c);
            return norm3(c) / 2.0;
        }
        case 3: {
            double a[3], b[3], c[3], d[3];
            sub3(m.P(e.node[1]), m.P(e.node[0]), a);
            sub3(m.P(e.node[2]), m.P(e.node[0]), b);
            cross3(a, b, c);
            sub3(m.P(e.node[3]), m.P(e.node[0]), d);
            return std::fabs(dot3(c, d)) / 6.0;
        }
    }
    return 1.0;
}
static void elem_centre(const Mesh &m, const Element &e, double *c) {
    c[0] = c[1] = c[2] = 0.0;
    for (int i = 0; i < e.n_nodes(); i++) {
        const double *p = m.P(e.node[i]);
        c[0] += p[0]; c[1] += p[1]; c[2] += p[2];
    }
    double n = (double)e.n_nodes();
    c[0] /= n; c[1] /= n; c[2] /= n;
}
static inline int snode(const Mesh &m, const Element &e, int s, int k) {
    (void)m;
    return e.node[side_node(e.dim, s, k)];
}
static double side_measure(const Mesh &m, const Element &e, int s) {
    int sd = e.dim - 1;
    if (sd == 0) return 1.0;
    if (sd == 1) { double d[3]; sub3(m.P(snode(m, e, s, 1)), m.P(snode(m, e, s, 0)), d); return norm3(d); }
    double a[3], b[3], c[3];
    sub3(m.P(snode(m, e, s, 1)), m.P(snode(m, e, s, 0)), a);
    sub3(m.P(snode(m, e, s, 2)), m.P(snode(m, e, s, 0)), b);
    cross3(a, b, c);
    return 0.5 * norm3(c);
}
static void side_centre(const Mesh &m, const Element &e, int s, double *c) {
    c[0] = c[1] = c[2] = 0.0;
    int nn = e.dim;  // side has dim nodes
    for (int i = 0; i < nn; i++) {
        const double *p = m.P(snode(m, e, s, i));
        c[0] += p[0]; c[1] += p[1]; c[2] += p[2];
    }
    c[0] /= (double)nn; c[1] /= (double)nn; c[2] /= (double)nn;
}
static void side_normal(const Mesh &m, const Element &e, int s, double *nrm) {
    int sd = e.dim - 1;
    double ec[3];
    elem_centre(m, e, ec);
    if (sd == 0) {  // sides.cc:77-89
        sub3(m.P(e.node[1]), m.P(e.node[0]), nrm);
        double l = norm3(nrm);
        nrm[0] /= l; nrm[1] /= l; nrm[2] /= l;
        if (snode(m, e, s, 0) == e.node[0]) { nrm[0] = -nrm[0]; nrm[1] = -nrm[1]; nrm[2] = -nrm[2]; }
        return;
    }
    if (sd == 1) {  // sides.cc:94-111
        double a[3], b[3], en[3];
        sub3(m.P(e.node[1]), m.P(e.node[0]), a);
        sub3(m.P(e.node[2]), m.P(e.node[0]), b);
        cross3(a, b, en);
        double l = norm3(en);
        en[0] /= l; en[1] /= l; en[2] /= l;
        double t[3];
        sub3(m.P(snode(m, e, s, 1)), m.P(snode(m, e, s, 0)), t);
        cross3(t, en, nrm);
        l = norm3(nrm);
        nrm[0] /= l; nrm[1] /= l; nrm[2] /= l;
        double d[3];
        sub3(ec, m.P(snode(m, e, s, 0)), d);
        if (dot3(nrm, d) > 0.0) { nrm[0] = -nrm[0]; nrm[1] = -nrm[1]; nrm[2] = -nrm[2]; }
        return;
    }
    // sides.cc:116-131
    double a[3], b[3];
    sub3(m.P(snode(m, e, s, 1)), m.P(snode(m, e, s, 0)), a);
    sub3(m.P(snode(m, e, s, 2)), m.P(snode(m, e, s, 0)), b);
    cross3(a, b, nrm);
    double l = norm3(nrm);
    nrm[0] /= l; nrm[1] /= l; nrm[2] /= l;
    double d[3];
    sub3(ec, m.P(snode(m, e, s, 0)), d);
    if (dot3(nrm, d) > 0.0) { nrm[0] = -nrm[0]; nrm[1] = -nrm[1]; nrm[2] = -nrm[2]; }
}

// ---------------------------------------------------------------------------
// GMSH 2.x reader: src/mesh/msh_gmshreader.cc:76-246
// ---------------------------------------------------------------------------
static bool read_gmsh(const char *path, Mesh &m) {
    std::ifstream in(path);
    if (!in) { g_err = std::string("cannot open ") + path; return false; }
    std::string line;
    // $PhysicalNames (optional, before $Nodes): msh_gmshreader.cc:224-246
    bool have_nodes = false;
    while (std::getline(in, line)) {
        if (line.compare(0, 14, "$PhysicalNames") == 0) {
            std::getline(in, line);
            int n = std::stoi(line);
            for (int i = 0; i < n; i++) {
                std::getline(in, line);
                std::istringstream ss(line);
                int dim, id;
                ss >> dim >> id;
                std::string rest;
                std::getline(ss, rest);
                size_t a = rest.find('"'), b = rest.rfind('"');
                std::string name = (a != std::string::npos && b > a) ? rest.substr(a + 1, b - a - 1) : rest;
                m.rdb.add_named(id, name);
            }
        } else if (line.compare(0, 6, "$Nodes") == 0) {
            have_nodes = true;
            break;
        }
    }
    if (!have_nodes) { g_err = "missing $Nodes"; return false; }
    // $Nodes: msh_gmshreader.cc:89-118
    std::getline(in, line);
    int nn = std::stoi(line);
    m.X.resize(3 * (size_t)nn);
    m.node_ids.resize(nn);
    for (int i = 0; i < nn; i++) {
        int id;
        double x, y, z;
        if (!(in >> id >> x >> y >> z)) { g_err = "bad node line"; return false; }
        m.node_ids[i] = id;
        m.node_idx[id] = i;
        m.X[3 * i] = x; m.X[3 * i + 1] = y; m.X[3 * i + 2] = z;
    }
    // $Elements: msh_gmshreader.cc:121-217
    while (std::getline(in, line))
        if (line.compare(0, 9, "$Elements") == 0) break;
    std::getline(in, line);
    int ne = std::stoi(line);
    for (int i = 0; i < ne; i++) {
        int id, type, ntags;
        in >> id >> type >> ntags;
        int dim;
        switch (type) {
            case 1: dim = 1; break;
            case 2: dim = 2; break;
            case 4: dim = 3; break;
            case 15: dim = 0; break;
            default: g_err = "unsupported element type " + std::to_string(type); return false;
        }
        std::vector<int> tags(ntags);
        for (int t = 0; t < ntags; t++) in >> tags[t];
        int region_id = tags.size() > 0 ? tags[0] : 0;
        int ridx = m.rdb.add_unnamed(region_id);
        Element e;
        e.id = id;
        e.dim = dim;
        e.region = ridx;
        for (int k = 0; k < dim + 1; k++) {
            int nid;
            in >> nid;
            auto it = m.node_idx.find(nid);
            if (it == m.node_idx.end()) { g_err = "unknown node id"; return false; }
            e.node[k] = it->second;
        }
        if (m.rdb.regs[ridx].boundary) m.bc_elements.push_back(e);
        else if (dim > 0) m.element.push_back(e);
    }
    m.n_file_bc = (int)m.bc_elements.size();
    return true;
}

// ---------------------------------------------------------------------------
// Topology: Mesh::make_neighbours_and_edges src/mesh/mesh.cc:343-500,
// create_node_element_lists :264-274, intersect_element_lists :277-302,
// find_lower_dim_element :305-323, same_sides :326-332,
// element_to_neigh_vb :660-685.
// ---------------------------------------------------------------------------
static void intersect(const Mesh &m, const std::vector<int> &nodes, std::vector<int> &out) {
    out.clear();
    if (nodes.empty()) return;
    out = m.node_elements[nodes[0]];
    for (size_t k = 1; k < nodes.size(); k++) {
        const std::vector<int> &b = m.node_elements[nodes[k]];
        std::vector<int> tmp;
        std::set_intersection(out.begin(), out.end(), b.begin(), b.end(), std::back_inserter(tmp));
        out.swap(tmp);
    }
}
static bool find_lower_dim(const Mesh &m, std::vector<int> &list, int dim, int &lower) {
    bool is_nb = false;
    std::vector<int> keep;
    for (int e : list) {
        if (m.element[e].dim == dim) keep.push_back(e);
        else if (m.element[e].dim == dim - 1) {
            if (is_nb) { g_err = "too many matching lower-dim elements"; }
            is_nb = true;
            lower = e;
        }
    }
    list.swap(keep);
    return is_nb;
}
static bool same_side(const Mesh &m, const Element &e, int s, const std::vector<int> &nodes) {
    for (int k = 0; k < e.dim; k++)
        if (std::find(nodes.begin(), nodes.end(), snode(m, e, s, k)) == nodes.end()) return false;
    return true;
}

static bool setup_topology(Mesh &m) {
    int nn = (int)m.node_ids.size();
    m.node_elements.assign(nn, {});
    for (int e = 0; e < (int)m.element.size(); e++)
        for (int k = 0; k < m.element[e].n_nodes(); k++) m.node_elements[m.element[e].node[k]].push_back(e);
    std::vector<int> side_nodes, isect;
    // (1) boundary elements from the file, in file order: mesh.cc:358-409
    for (int bi = 0; bi < (int)m.bc_elements.size(); bi++) {
        const Element &bc = m.bc_elements[bi];
        side_nodes.assign(bc.node, bc.node + bc.n_nodes());
        intersect(m, side_nodes, isect);
        int lower = -1;
        bool nb = find_lower_dim(m, isect, bc.dim + 1, lower);
        if (nb) { g_err = "boundary element matches a lower-dimensional bulk element"; return false; }
        if (isect.empty()) continue;  // lonely boundary element (warning in the reference)
        int ei = (int)m.edges.size();
        m.edges.push_back(Edge());
        int bdr = (int)m.boundary.size();
        m.boundary.push_back(Boundary{ei, bi});
        for (int el : isect) {
            Element &E = m.element[el];
            for (int s = 0; s < E.n_sides(); s++) {
                if (same_side(m, E, s, side_nodes)) {
                    m.edges[ei].sides.push_back({el, s});
                    E.edge_idx[s] = ei;
                    E.boundary_idx[s] = bdr;
                    break;
                }
            }
        }
    }
    // (2) all element sides in element order: mesh.cc:412-497
    for (int e = 0; e < (int)m.element.size(); e++) {
        for (int s = 0; s < m.element[e].n_sides(); s++) {
            if (m.element[e].edge_idx[s] != -1) continue;
            const Element &Ee = m.element[e];
            side_nodes.clear();
            for (int k = 0; k < Ee.dim; k++) side_nodes.push_back(snode(m, Ee, s, k));
            intersect(m, side_nodes, isect);
            int lower = -1;
            bool nb = find_lower_dim(m, isect, Ee.dim, lower);
            int ei = -1;
            if (!nb) {
                ei = (int)m.edges.size();
                m.edges.push_back(Edge());
                if (isect.size() == 1) {
                    m.edges[ei].sides.push_back({e, s});
                    m.element[e].edge_idx[s] = ei;
                    int bdr = (int)m.boundary.size();
                    m.element[e].boundary_idx[s] = bdr;
                    Element bc;
                    bc.id = -bdr;
                    bc.dim = Ee.dim - 1;
                    bc.region = m.rdb.implicit_boundary();
                    for (size_t k = 0; k < side_nodes.size(); k++) bc.node[k] = side_nodes[k];
                    m.bc_elements.push_back(bc);
                    m.boundary.push_back(Boundary{ei, (int)m.bc_elements.size() - 1});
                    continue;
                }
            }
            for (int el : isect) {
                Element &E = m.element[el];
                for (int es = 0; es < E.n_sides(); es++) {
                    if (E.edge_idx[es] != -1) continue;
                    if (same_side(m, E, es, side_nodes)) {
                        if (nb) {
                            int ne = (int)m.edges.size();
                            m.edges.push_back(Edge());
                            m.edges[ne].sides.push_back({el, es});
                            E.edge_idx[es] = ne;
                            m.vb.push_back(Neighbour{lower, ne});
                        } else {
                            m.edges[ei].sides.push_back({el, es});
                            E.edge_idx[es] = ei;
                        }
                        break;
                    }
                }
            }
        }
    }
    // element_to_neigh_vb: mesh.cc:660-685
    for (int n = 0; n < (int)m.vb.size(); n++) m.element[m.vb[n].element].neigh_vb.push_back(n);
    // MH_DofHandler::reinit: src/flow/mh_dofhandler.cc:12-21
    m.side_offset.resize(m.element.size() + 1);
    int off = 0;
    for (size_t e = 0; e < m.element.size(); e++) {
        m.side_offset[e] = off;
        off += m.element[e].n_sides();
    }
    m.side_offset[m.element.size()] = off;
    m.n_sides = off;
    return true;
}

// ---------------------------------------------------------------------------
// Small inverses: src/system/math_fce.cc:124-230 (SUBDET2 math_fce.h:69,
// NUM_ZERO math_fce.h:51).
// ---------------------------------------------------------------------------
static const double NUM_ZERO = std::numeric_limits<double>::min() / std::numeric_limits<double>::epsilon();
#define A_(i, j) a[(i) * n + (j)]
#define SD2(i, j, k, l) (A_(i, k) * A_(j, l) - A_(i, l) * A_(j, k))
static double matrix_inverse(const double *a, double *b, int n) {
#define B_(i, j) b[(i) * n + (j)]
    if (n == 1) { b[0] = 1 / a[0]; return a[0]; }
    if (n == 2) {
        double det = SD2(0, 1, 0, 1);
        if (std::fabs(det) < NUM_ZERO) return det;
        B_(0, 0) = A_(1, 1) / det; B_(1, 0) = -A_(1, 0) / det;
        B_(0, 1) = -A_(0, 1) / det; B_(1, 1) = A_(0, 0) / det;
        return det;
    }
    if (n == 3) {
        B_(0, 0) = SD2(1, 2, 1, 2); B_(1, 0) = -SD2(1, 2, 0, 2); B_(2, 0) = SD2(1, 2, 0, 1);
        double det = A_(0, 0) * B_(0, 0) + A_(0, 1) * B_(1, 0) + A_(0, 2) * B_(2, 0);
        if (std::fabs(det) < NUM_ZERO) return det;
        B_(0, 0) /= det; B_(1, 0) /= det; B_(2, 0) /= det;
        B_(0, 1) = -SD2(0, 2, 1, 2) / det; B_(1, 1) = SD2(0, 2, 0, 2) / det; B_(2, 1) = -SD2(0, 2, 0, 1) / det;
        B_(0, 2) = SD2(0, 1, 1, 2) / det; B_(1, 2) = -SD2(0, 1, 0, 2) / det; B_(2, 2) = SD2(0, 1, 0, 1) / det;
        return det;
    }
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
    if (std::fabs(det) < NUM_ZERO) return det;
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
#undef B_
}
#undef SD2
#undef A_

// inverse of a general 3x3 (the anisotropy tensor); cofactor form.
static void inv3x3(const double *a, double *r) {
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

// ---------------------------------------------------------------------------
// RT0 local matrices: MHFEValues src/flow/mh_fe_values.cc:34-514.
// R = inverse anisotropy (3x3 row-major); identity when aniso == nullptr.
// Returns det of the local matrix (mh_fe_values.cc:54).
// ---------------------------------------------------------------------------
static double local_matrix(const Mesh &m, const Element &e, double K, double cs, const double *aniso,
                           double *loc, double *inv) {
    double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (aniso) inv3x3(aniso, R);
    double scale = 1 / K / cs;  // mh_fe_values.cc:38
    const int d = e.dim;
    if (d == 1) {  // local_matrix_line :129-153
        double t[3];
        sub3(m.P(e.node[1]), m.P(e.node[0]), t);
        double l = norm3(t);
        t[0] /= l; t[1] /= l; t[2] /= l;
        double Rt[3] = {R[0] * t[0] + R[1] * t[1] + R[2] * t[2], R[3] * t[0] + R[4] * t[1] + R[5] * t[2],
                        R[6] * t[0] + R[7] * t[1] + R[8] * t[2]};
        double val = scale * dot3(t, Rt) * elem_measure(m, e) / 3.0;
        loc[0] = val; loc[3] = val; loc[1] = -val / 2.0; loc[2] = -val / 2.0;
    } else if (d == 2) {  // local_matrix_triangle :159-211
        double ex[3], ac[3], ez[3], ey[3];
        sub3(m.P(e.node[1]), m.P(e.node[0]), ex);
        normalize3(ex);
        sub3(m.P(e.node[2]), m.P(e.node[0]), ac);
        cross3(ex, ac, ez);
        normalize3(ez);
        cross3(ez, ex, ey);
        normalize3(ey);
        // resistance_tensor = r^T (R r), r = [ex ey] (3x2)
        double Rr[3][2];
        for (int i = 0; i < 3; i++) {
            Rr[i][0] = R[3 * i] * ex[0] + R[3 * i + 1] * ex[1] + R[3 * i + 2] * ex[2];
            Rr[i][1] = R[3 * i] * ey[0] + R[3 * i + 1] * ey[1] + R[3 * i + 2] * ey[2];
        }
        double rt[2][2];
        for (int j = 0; j < 2; j++) {
            rt[0][j] = ex[0] * Rr[0][j] + ex[1] * Rr[1][j] + ex[2] * Rr[2][j];
            rt[1][j] = ey[0] * Rr[0][j] + ey[1] * Rr[1][j] + ey[2] * Rr[2][j];
        }
        // node_coordinates_triangle :218-247
        double u[3], v[3], n[3], tt[3], b[3], nod[3][2];
        sub3(m.P(e.node[1]), m.P(e.node[0]), u);
        sub3(m.P(e.node[2]), m.P(e.node[0]), v);
        cross3(u, v, n);
        tt[0] = u[0]; tt[1] = u[1]; tt[2] = u[2];
        normalize3(tt);
        normalize3(n);
        cross3(n, u, b);
        normalize3(b);
        nod[0][0] = 0.0; nod[0][1] = 0.0;
        nod[1][0] = dot3(tt, u); nod[1][1] = dot3(b, u);
        nod[2][0] = dot3(tt, v); nod[2][1] = dot3(b, v);
        // side_midpoint_triangle :253-261
        double mid[3][2];
        mid[0][0] = (nod[0][0] + nod[1][0]) / 2.0; mid[0][1] = (nod[0][1] + nod[1][1]) / 2.0;
        mid[1][0] = (nod[1][0] + nod[2][0]) / 2.0; mid[1][1] = (nod[1][1] + nod[2][1]) / 2.0;
        mid[2][0] = (nod[2][0] + nod[0][0]) / 2.0; mid[2][1] = (nod[2][1] + nod[0][1]) / 2.0;
        // basis_functions_triangle :268-295
        double al[3], be[3], ga[3];
        for (int i = 0; i < 3; i++) {
            int i0 = i, i1 = (i + 1) % 3, i2 = (i + 2) % 3;
            double x0 = nod[i0][0], y0 = nod[i0][1], x1 = nod[i1][0], y1 = nod[i1][1], x2 = nod[i2][0], y2 = nod[i2][1];
            al[i] = x2; be[i] = y2;
            ga[i] = 1.0 / std::fabs((x0 - x2) * (y1 - y0) + (y0 - y2) * (x0 - x1));
        }
        double meas = elem_measure(m, e);
        for (int i = 0; i < 3; i++)
            for (int j = i; j < 3; j++) {
                // calc_polynom_triangle :299-319
                double poly[6];
                poly[0] = rt[0][0] * al[i] * al[j] + rt[0][1] * be[i] * al[j] + rt[1][0] * al[i] * be[j] +
                          rt[1][1] * be[i] * be[j];
                poly[1] = (rt[0][0] * al[i] + rt[0][0] * al[j] + rt[0][1] * be[i] + rt[1][0] * be[j]) * -1.0;
                poly[2] = (rt[1][1] * be[i] + rt[1][1] * be[j] + rt[1][0] * al[i] + rt[0][1] * al[j]) * -1.0;
                poly[3] = rt[0][0];
                poly[4] = rt[0][1] + rt[1][0];
                poly[5] = rt[1][1];
                double p[3];
                for (int q = 0; q < 3; q++)  // polynom_value_triangle :323-334
                    p[q] = poly[0] + poly[1] * mid[q][0] + poly[2] * mid[q][1] + poly[3] * mid[q][0] * mid[q][0] +
                           poly[4] * mid[q][0] * mid[q][1] + poly[5] * mid[q][1] * mid[q][1];
                loc[3 * i + j] = scale * ga[i] * ga[j] * meas * (p[0] + p[1] + p[2]) / 3.0;
                loc[3 * j + i] = loc[3 * i + j];
            }
    } else {  // local_matrix_tetrahedron :339-375
        double al[4], be[4], ga[4], de[4];
        for (int li = 0; li < 4; li++) {  // basis_functions_tetrahedron :379-397
            const double *p = m.P(e.node[li]);
            al[li] = p[0]; be[li] = p[1]; ga[li] = p[2];
            double nrm[3], sc[3];
            side_normal(m, e, li, nrm);
            side_centre(m, e, li, sc);
            de[li] = 1.0 / (side_measure(m, e, li) *
                            (nrm[0] * sc[0] + nrm[1] * sc[1] + nrm[2] * sc[2] - nrm[0] * al[li] - nrm[1] * be[li] -
                             nrm[2] * ga[li]));
        }
        double v = elem_measure(m, e), t[3];
        elem_centre(m, e, t);
        const double *n0 = m.P(e.node[0]), *n1 = m.P(e.node[1]), *n2 = m.P(e.node[2]), *n3 = m.P(e.node[3]);
        auto mom = [&](int a, int b) {  // second moments used by polynom_integral_tetrahedron :452-514
            return t[a] * t[b] + 1.0 / 20.0 *
                                     ((n0[a] - t[a]) * (n0[b] - t[b]) + (n1[a] - t[a]) * (n1[b] - t[b]) +
                                      (n2[a] - t[a]) * (n2[b] - t[b]) + (n3[a] - t[a]) * (n3[b] - t[b]));
        };
        double Mxy = mom(0, 1), Mxz = mom(0, 2), Myz = mom(1, 2), Mxx = mom(0, 0), Myy = mom(1, 1), Mzz = mom(2, 2);
        const double *a = R;
        auto A = [&](int i, int j) { return a[3 * i + j]; };
        for (int i = 0; i < 4; i++)
            for (int j = i; j < 4; j++) {
                double poly[10];  // calc_polynom_tetrahedron :401-448
                poly[0] = A(0, 0) * al[i] * al[j] + A(0, 1) * be[i] * al[j] + A(0, 2) * ga[i] * al[j] +
                          A(1, 0) * al[i] * be[j] + A(1, 1) * be[i] * be[j] + A(1, 2) * ga[i] * be[j] +
                          A(2, 0) * al[i] * ga[j] + A(2, 1) * be[i] * ga[j] + A(2, 2) * ga[i] * ga[j];
                poly[1] = (A(0, 0) * al[i] + A(0, 1) * be[i] + A(0, 2) * ga[i] + A(0, 0) * al[j] + A(1, 0) * be[j] +
                           A(2, 0) * ga[j]) * -1.0;
                poly[2] = (A(1, 0) * al[i] + A(1, 1) * be[i] + A(1, 2) * ga[i] + A(0, 1) * al[j] + A(1, 1) * be[j] +
                           A(2, 1) * ga[j]) * -1.0;
                poly[3] = (A(2, 0) * al[i] + A(2, 1) * be[i] + A(2, 2) * ga[i] + A(0, 2) * al[j] + A(1, 2) * be[j] +
                           A(2, 2) * ga[j]) * -1.0;
                poly[4] = A(0, 1) + A(1, 0);
                poly[5] = A(0, 2) + A(2, 0);
                poly[6] = A(1, 2) + A(2, 1);
                poly[7] = A(0, 0);
                poly[8] = A(1, 1);
                poly[9] = A(2, 2);
                double rc = 0.0;
                rc += poly[0] * v;
                rc += poly[1] * v * t[0];
                rc += poly[2] * v * t[1];
                rc += poly[3] * v * t[2];
                rc += poly[4] * v * Mxy;
                rc += poly[5] * v * Mxz;
                rc += poly[6] * v * Myz;
                rc += poly[7] * v * Mxx;
                rc += poly[8] * v * Myy;
                rc += poly[9] * v * Mzz;
                loc[4 * i + j] = scale * de[i] * de[j] * rc;
                loc[4 * j + i] = loc[4 * i + j];
            }
    }
    return matrix_inverse(loc, inv, d + 1);
}

// ---------------------------------------------------------------------------
// Sparse matrix in triplet/CSR form (a stand-in for PETSc SeqAIJ, ADD mode).
// ---------------------------------------------------------------------------
struct CSR {
    int nr = 0, nc = 0;
    std::vector<int> ptr, col;
    std::vector<double> val;
    double get(int r, int c) const {
        auto b = col.begin() + ptr[r], e = col.begin() + ptr[r + 1];
        auto it = std::lower_bound(b, e, c);
        if (it != e && *it == c) return val[it - col.begin()];
        return 0.0;
    }
};
struct Triplets {
    std::vector<int> r, c;
    std::vector<double> v;
    void add(int i, int j, double x) { r.push_back(i); c.push_back(j); v.push_back(x); }
    // combine duplicates in insertion order (ADD_VALUES semantics)
    CSR to_csr(int nr, int nc) const {
        CSR A;
        A.nr = nr; A.nc = nc;
        std::vector<size_t> idx(r.size());
        for (size_t i = 0; i < idx.size(); i++) idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
            return r[a] != r[b] ? r[a] < r[b] : c[a] < c[b];
        });
        A.ptr.assign(nr + 1, 0);
        for (size_t k = 0; k < idx.size(); k++) {
            size_t i = idx[k];
            if (!A.col.empty() && k > 0 && r[idx[k - 1]] == r[i] && c[idx[k - 1]] == c[i]) {
                A.val.back() += v[i];
            } else {
                A.col.push_back(c[i]);
                A.val.push_back(v[i]);
                A.ptr[r[i] + 1]++;
            }
        }
        for (int i = 0; i < nr; i++) A.ptr[i + 1] += A.ptr[i];
        return A;
    }
};
// submatrix rows [r0,r1) x cols [c0,c1) (MatGetSubMatrix)
static CSR submat(const CSR &A, int r0, int r1, int c0, int c1) {
    CSR B;
    B.nr = r1 - r0; B.nc = c1 - c0;
    B.ptr.assign(B.nr + 1, 0);
    for (int i = r0; i < r1; i++) {
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; k++)
            if (A.col[k] >= c0 && A.col[k] < c1) { B.col.push_back(A.col[k] - c0); B.val.push_back(A.val[k]); }
        B.ptr[i - r0 + 1] = (int)B.col.size();
    }
    return B;
}
// C = A*B, row-by-row with a dense accumulator in A's column order.
static CSR spgemm(const CSR &A, const CSR &B) {
    CSR C;
    C.nr = A.nr; C.nc = B.nc;
    C.ptr.assign(A.nr + 1, 0);
    std::vector<double> acc(B.nc, 0.0);
    std::vector<int> mark(B.nc, -1), cols;
    for (int i = 0; i < A.nr; i++) {
        cols.clear();
        for (int ka = A.ptr[i]; ka < A.ptr[i + 1]; ka++) {
            int k = A.col[ka];
            double a = A.val[ka];
            for (int kb = B.ptr[k]; kb < B.ptr[k + 1]; kb++) {
                int j = B.col[kb];
                if (mark[j] != i) { mark[j] = i; acc[j] = 0.0; cols.push_back(j); }
                acc[j] += a * B.val[kb];
            }
        }
        std::sort(cols.begin(), cols.end());
        for (int j : cols) { C.col.push_back(j); C.val.push_back(acc[j]); }
        C.ptr[i + 1] = (int)C.col.size();
    }
    return C;
}
// y = A x (+ z)
static void spmv(const CSR &A, const double *x, double *y, const double *z = nullptr) {
    for (int i = 0; i < A.nr; i++) {
        double s = z ? z[i] : 0.0;
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; k++) s += A.val[k] * x[A.col[k]];
        y[i] = s;
    }
}
static void spmv_t(const CSR &A, const double *x, double *y) {  // y = A^T x
    std::fill(y, y + A.nc, 0.0);
    for (int i = 0; i < A.nr; i++)
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; k++) y[A.col[k]] += A.val[k] * x[i];
}
// C := alpha*C + beta*X  with union pattern (MatScale + MatAXPY)
static CSR axpy_union(const CSR &C, double alpha, const CSR &X, double beta) {
    CSR R;
    R.nr = C.nr; R.nc = C.nc;
    R.ptr.assign(C.nr + 1, 0);
    for (int i = 0; i < C.nr; i++) {
        int a = C.ptr[i], ae = C.ptr[i + 1], b = X.ptr[i], be = X.ptr[i + 1];
        while (a < ae || b < be) {
            int ca = a < ae ? C.col[a] : INT32_MAX, cb = b < be ? X.col[b] : INT32_MAX;
            if (ca == cb) { R.col.push_back(ca); R.val.push_back(alpha * C.val[a] + beta * X.val[b]); a++; b++; }
            else if (ca < cb) { R.col.push_back(ca); R.val.push_back(alpha * C.val[a]); a++; }
            else { R.col.push_back(cb); R.val.push_back(beta * X.val[b]); b++; }
        }
        R.ptr[i + 1] = (int)R.col.size();
    }
    return R;
}

// ---------------------------------------------------------------------------
// Darcy MH (steady): src/flow/darcy_flow_mh.cc
// ---------------------------------------------------------------------------
struct Darcy {
    Mesh *m;
    int E, S, F, N;
    int bad = -1;  // first element with a singular local matrix (assembly)
    std::vector<double> K, cs, sigma, f, aniso;  // per element
    std::vector<int> bc_type;                     // per bc element (0 none,1 dir,2 neu,3 robin)
    std::vector<double> bc_p, bc_flux, bc_sigma;  // per bc element
    CSR M;                                        // MH matrix (S+E+F)
    std::vector<double> rhs;
    // Schur stage 1 / 2 (schur.cc)
    CSR IA1, IAB1, C1;  // C1: complement (E+F)
    std::vector<double> rhs1, rhsC1;
    CSR IA2, IAB2, C2;  // C2: final edge system (F)
    std::vector<double> rhsC1_el, rhsC2;
    std::vector<double> sol;  // [q | p | lambda]
    // unsteady models (darcy_flow_mh.cc:1601-1827): 1 MH, 2 LMH
    int unsteady = 0;
    std::vector<double> stor, steady_diag, new_diag, steady_rhs, prev_sol;
    double dt = 0;
    int iters = 0;
    double resid = 0, bnorm1 = 0;
    double t_assembly = 0, t_schur = 0, t_solve = 0, t_resolve = 0;
};

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// assembly_steady_mh_matrix: darcy_flow_mh.cc:525-694 (n_schurs == 2, no mortar)
static void assemble(Darcy &D) {
    Mesh &m = *D.m;
    Triplets T;
    D.rhs.assign(D.N, 0.0);
    const int S = D.S, E = D.E;
    double loc[16], inv[16];
    for (int e = 0; e < E; e++) {
        const Element &el = m.element[e];
        int ns = el.n_sides();
        int el_row = S + e;
        const double det = local_matrix(m, el, D.K[e], D.cs[e], D.aniso.empty() ? nullptr : &D.aniso[9 * e], loc, inv);
        // MHFEValues::update (mh_fe_values.cc:54-59): a singular local matrix is fatal
        if (std::fabs(det) < NUM_ZERO && D.bad < 0) D.bad = e;
        double csec = D.cs[e];
        double ec[3];
        elem_centre(m, el, ec);
        int side_rows[4], edge_rows[4];
        double loc_rhs[4];
        for (int i = 0; i < ns; i++) {
            side_rows[i] = m.side_offset[e] + i;
            edge_rows[i] = S + E + el.edge_idx[i];
            double sc[3];
            side_centre(m, el, i, sc);
            loc_rhs[i] = ec[2] - sc[2];  // gravity term :565
            double c_val = 1.0;
            int b = el.boundary_idx[i];
            if (b >= 0) {
                int bce = m.boundary[b].bc_ele_idx;
                int type = D.bc_type[bce];
                double bmeas = elem_measure(m, m.bc_elements[bce]);
                if (type == 1) {
                    c_val = 0.0;
                    loc_rhs[i] -= D.bc_p[bce];
                    D.rhs[edge_rows[i]] += -D.bc_p[bce];
                    T.add(edge_rows[i], edge_rows[i], -1.0);
                } else if (type == 2) {
                    D.rhs[edge_rows[i]] += D.bc_flux[bce] * bmeas * csec;
                } else if (type == 3) {
                    D.rhs[edge_rows[i]] += -bmeas * D.bc_sigma[bce] * D.bc_p[bce] * csec;
                    T.add(edge_rows[i], edge_rows[i], -bmeas * D.bc_sigma[bce] * csec);
                }
            }
            T.add(side_rows[i], edge_rows[i], c_val);
            T.add(edge_rows[i], side_rows[i], c_val);
        }
        for (int i = 0; i < ns; i++) D.rhs[side_rows[i]] += loc_rhs[i];
        for (int i = 0; i < ns; i++)
            for (int j = 0; j < ns; j++) T.add(side_rows[i], side_rows[j], loc[ns * i + j]);
        for (int i = 0; i < ns; i++) T.add(el_row, side_rows[i], -1.0);
        for (int i = 0; i < ns; i++) T.add(side_rows[i], el_row, -1.0);
        D.rhs[el_row] += -1.0 * elem_measure(m, el) * D.cs[e] * D.f[e];
        T.add(el_row, el_row, 0.0);
        std::vector<int> vb_rows;
        for (int n : el.neigh_vb) {
            const Neighbour &ngh = m.vb[n];
            int edge_row = S + E + ngh.edge_idx;
            const auto &hs = m.edges[ngh.edge_idx].sides[0];  // Neighbour::side(), neighbours_impl.hh:16-20
            const Element &hel = m.element[hs.first];
            double value = D.sigma[e] * side_measure(m, hel, hs.second) * D.cs[hs.first];
            T.add(el_row, el_row, -value);
            T.add(el_row, edge_row, value);
            T.add(edge_row, el_row, value);
            T.add(edge_row, edge_row, -value);
            for (int i = 0; i < ns; i++) { T.add(edge_rows[i], edge_row, 0.0); T.add(edge_row, edge_rows[i], 0.0); }
            vb_rows.push_back(edge_row);
        }
        for (int a : vb_rows)
            for (int b : vb_rows) T.add(a, b, 0.0);
        for (int i = 0; i < ns; i++) { T.add(el_row, edge_rows[i], 0.0); T.add(edge_rows[i], el_row, 0.0); }
        for (int i = 0; i < ns; i++)
            for (int j = 0; j < ns; j++) T.add(edge_rows[i], edge_rows[j], 0.0);
    }
    D.M = T.to_csr(D.N, D.N);
}

// SchurComplement::form_schur + form_rhs: src/la/schur.cc:520-696
static void form_schur(const CSR &Orig, const std::vector<double> &rhs, int nA, const CSR &IA, CSR &IAB,
                       CSR &Compl, std::vector<double> &crhs) {
    int n = Orig.nr;
    CSR B = submat(Orig, 0, nA, nA, n);
    IAB = spgemm(IA, B);
    CSR Bt = submat(Orig, nA, n, 0, nA);
    CSR xA = spgemm(Bt, IAB);
    CSR C = submat(Orig, nA, n, nA, n);
    Compl = axpy_union(C, -1.0, xA, 1.0);  // complement = Bt*IA*B - C
    crhs.assign(n - nA, 0.0);
    spmv_t(IAB, rhs.data(), crhs.data());  // MatMultTranspose(IAB, RHS1)
    for (int i = 0; i < n - nA; i++) crhs[i] += -1 * rhs[nA + i];
}
// SchurComplement::resolve: schur.cc:714-773;  x1 = IA*RHS1 - IAB*x2
static void resolve(const CSR &IA, const CSR &IAB, const double *rhs1, const double *x2, double *x1) {
    int nA = IA.nr;
    std::vector<double> t(nA);
    spmv(IAB, x2, t.data());
    for (int i = 0; i < nA; i++) t[i] *= -1;
    spmv(IA, rhs1, x1, t.data());
}

// Jacobi-preconditioned CG on an SPD CSR system: the KSP CG loop of
// src/la/solve.cc:316-463 with a diagonal preconditioner.
static int cg(const CSR &A, const double *b, double *x, double rtol, double atol, int max_it, double &res) {
    int n = A.nr;
    std::vector<double> r(n), z(n), p(n), q(n), dinv(n);
    for (int i = 0; i < n; i++) {
        double d = A.get(i, i);
        dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
    }
    spmv(A, x, q.data());
    double bn = 0, rr = 0;
    for (int i = 0; i < n; i++) { r[i] = b[i] - q[i]; bn += b[i] * b[i]; rr += r[i] * r[i]; }
    bn = std::sqrt(bn);
    double tol = std::max(rtol * bn, atol);
    for (int i = 0; i < n; i++) { z[i] = dinv[i] * r[i]; p[i] = z[i]; }
    double rz = 0;
    for (int i = 0; i < n; i++) rz += r[i] * z[i];
    int it = 0;
    res = std::sqrt(rr);
    while (res > tol && it < max_it) {
        spmv(A, p.data(), q.data());
        double pq = 0;
        for (int i = 0; i < n; i++) pq += p[i] * q[i];
        double alpha = rz / pq;
        rr = 0;
        for (int i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; rr += r[i] * r[i]; }
        double rz_new = 0;
        for (int i = 0; i < n; i++) { z[i] = dinv[i] * r[i]; rz_new += r[i] * z[i]; }
        double beta = rz_new / rz;
        rz = rz_new;
        for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
        res = std::sqrt(rr);
        it++;
    }
    return it;
}

// CPU baseline variants of the KSP solve (timing only; the parity tests use
// the serial Jacobi-CG above).  The reference's serial default is CG with
// PETSc ILU(3) (src/la/solve.cc:370); PETSc is not available here, and on
// this SPD system the closest preconditioner with a textbook definition is
// IC(0): the incomplete Cholesky factor on the lower pattern of A (no fill),
// applied by a forward and a backward triangular solve.
static int g_threads = 1;

// Jacobi-CG as `cg`, each loop split over g_threads OpenMP threads (the dot
// products then sum in a thread-dependent order).
static int cg_threads(const CSR &A, const double *b, double *x, double rtol, double atol, int max_it, double &res) {
    const int n = A.nr, nt = g_threads;
    std::vector<double> r(n), z(n), p(n), q(n), dinv(n);
    for (int i = 0; i < n; i++) {
        double d = A.get(i, i);
        dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
    }
    auto mv = [&](const double *v, double *out) {
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int i = 0; i < n; i++) {
            double t = 0.0;
            for (int k = A.ptr[i]; k < A.ptr[i + 1]; k++) t += A.val[k] * v[A.col[k]];
            out[i] = t;
        }
    };
    mv(x, q.data());
    double bn = 0, rr = 0, rz = 0;
#pragma omp parallel for num_threads(nt) reduction(+ : bn, rr, rz) schedule(static)
    for (int i = 0; i < n; i++) {
        r[i] = b[i] - q[i];
        z[i] = dinv[i] * r[i];
        p[i] = z[i];
        bn += b[i] * b[i];
        rr += r[i] * r[i];
        rz += r[i] * z[i];
    }
    bn = std::sqrt(bn);
    const double tol = std::max(rtol * bn, atol);
    int it = 0;
    res = std::sqrt(rr);
    while (res > tol && it < max_it) {
        mv(p.data(), q.data());
        double pq = 0;
#pragma omp parallel for num_threads(nt) reduction(+ : pq) schedule(static)
        for (int i = 0; i < n; i++) pq += p[i] * q[i];
        const double alpha = rz / pq;
        double rr2 = 0, rz2 = 0;
#pragma omp parallel for num_threads(nt) reduction(+ : rr2, rz2) schedule(static)
        for (int i = 0; i < n; i++) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
            z[i] = dinv[i] * r[i];
            rr2 += r[i] * r[i];
            rz2 += r[i] * z[i];
        }
        const double beta = rz2 / rz;
        rz = rz2;
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
        res = std::sqrt(rr2);
        it++;
    }
    return it;
}

// IC(0)-preconditioned CG (serial).  L (lower, with diagonal) on the pattern of
// the lower triangle of A: l_ij = (a_ij - sum_k<j l_ik l_jk) / l_jj,
// l_ii = sqrt(a_ii - sum_k<i l_ik^2); a non-positive pivot falls back to a_ii.
static int cg_ic0(const CSR &A, const double *b, double *x, double rtol, double atol, int max_it, double &res) {
    const int n = A.nr;
    std::vector<int> lp(n + 1, 0), lc;
    std::vector<double> lv;
    for (int i = 0; i < n; i++) {
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; k++)
            if (A.col[k] <= i) {
                lc.push_back(A.col[k]);
                lv.push_back(A.val[k]);
            }
        lp[i + 1] = (int)lc.size();
    }
    // row-oriented factorization: rows are sorted, the diagonal is each row's last entry
    std::vector<int> pos(n, -1);
    for (int i = 0; i < n; i++) {
        for (int k = lp[i]; k < lp[i + 1]; k++) pos[lc[k]] = k;
        for (int k = lp[i]; k < lp[i + 1] - 1; k++) {
            const int j = lc[k];
            double s = lv[k];
            for (int m = lp[j]; m < lp[j + 1] - 1; m++) {
                const int c = lc[m];
                if (pos[c] >= lp[i] && pos[c] < k) s -= lv[pos[c]] * lv[m];
            }
            lv[k] = s / lv[lp[j + 1] - 1];
        }
        const int d = lp[i + 1] - 1;
        double s = lv[d];
        for (int k = lp[i]; k < d; k++) s -= lv[k] * lv[k];
        lv[d] = s > 0.0 ? std::sqrt(s) : std::sqrt(std::fabs(A.get(i, i)));
        for (int k = lp[i]; k < lp[i + 1]; k++) pos[lc[k]] = -1;
    }
    // transpose of L for the backward solve
    std::vector<int> up(n + 1, 0), uc(lc.size());
    std::vector<double> uv(lc.size());
    for (int k = 0; k < (int)lc.size(); k++) up[lc[k] + 1]++;
    for (int i = 0; i < n; i++) up[i + 1] += up[i];
    {
        std::vector<int> fill(up.begin(), up.end() - 1);
        for (int i = 0; i < n; i++)
            for (int k = lp[i]; k < lp[i + 1]; k++) {
                const int j = lc[k], w = fill[j]++;
                uc[w] = i;
                uv[w] = lv[k];
            }
    }
    auto prec = [&](const double *rv, double *zv) {
        for (int i = 0; i < n; i++) {  // L y = r
            double t = rv[i];
            for (int k = lp[i]; k < lp[i + 1] - 1; k++) t -= lv[k] * zv[lc[k]];
            zv[i] = t / lv[lp[i + 1] - 1];
        }
        for (int i = n - 1; i >= 0; i--) {  // L^T z = y; row i of L^T starts with its diagonal
            double t = zv[i];
            for (int k = up[i] + 1; k < up[i + 1]; k++) t -= uv[k] * zv[uc[k]];
            zv[i] = t / uv[up[i]];
        }
    };
    std::vector<double> r(n), z(n), p(n), q(n);
    spmv(A, x, q.data());
    double bn = 0, rr = 0;
    for (int i = 0; i < n; i++) { r[i] = b[i] - q[i]; bn += b[i] * b[i]; rr += r[i] * r[i]; }
    bn = std::sqrt(bn);
    const double tol = std::max(rtol * bn, atol);
    prec(r.data(), z.data());
    double rz = 0;
    for (int i = 0; i < n; i++) { p[i] = z[i]; rz += r[i] * z[i]; }
    int it = 0;
    res = std::sqrt(rr);
    while (res > tol && it < max_it) {
        spmv(A, p.data(), q.data());
        double pq = 0;
        for (int i = 0; i < n; i++) pq += p[i] * q[i];
        const double alpha = rz / pq;
        rr = 0;
        for (int i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; rr += r[i] * r[i]; }
        prec(r.data(), z.data());
        double rz_new = 0;
        for (int i = 0; i < n; i++) rz_new += r[i] * z[i];
        const double beta = rz_new / rz;
        rz = rz_new;
        for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
        res = std::sqrt(rr);
        it++;
    }
    return it;
}

// Dense LU with partial pivoting (brute-force check on tiny systems).
static bool dense_solve(std::vector<double> A, std::vector<double> b, int n, std::vector<double> &x) {
    for (int k = 0; k < n; k++) {
        int p = k;
        for (int i = k + 1; i < n; i++)
            if (std::fabs(A[(size_t)i * n + k]) > std::fabs(A[(size_t)p * n + k])) p = i;
        if (A[(size_t)p * n + k] == 0.0) return false;
        if (p != k) {
            for (int j = 0; j < n; j++) std::swap(A[(size_t)k * n + j], A[(size_t)p * n + j]);
            std::swap(b[k], b[p]);
        }
        for (int i = k + 1; i < n; i++) {
            double l = A[(size_t)i * n + k] / A[(size_t)k * n + k];
            if (l == 0.0) continue;
            for (int j = k; j < n; j++) A[(size_t)i * n + j] -= l * A[(size_t)k * n + j];
            b[i] -= l * b[k];
        }
    }
    x.assign(n, 0.0);
    for (int i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (int j = i + 1; j < n; j++) s -= A[(size_t)i * n + j] * x[j];
        x[i] = s / A[(size_t)i * n + i];
    }
    return true;
}

// make_schur1 / make_schur2: darcy_flow_mh.cc:1009-1131
static void make_schurs(Darcy &D) {
    Mesh &m = *D.m;
    const int S = D.S, E = D.E;
    Triplets T;
    double loc[16], inv[16];
    for (int e = 0; e < E; e++) {
        const Element &el = m.element[e];
        int ns = el.n_sides();
        local_matrix(m, el, D.K[e], D.cs[e], D.aniso.empty() ? nullptr : &D.aniso[9 * e], loc, inv);
        for (int i = 0; i < ns; i++)
            for (int j = 0; j < ns; j++) T.add(m.side_offset[e] + i, m.side_offset[e] + j, inv[ns * i + j]);
    }
    D.IA1 = T.to_csr(S, S);
    form_schur(D.M, D.rhs, S, D.IA1, D.IAB1, D.C1, D.rhsC1);
    // schur1 rhs norm for solution_precision (darcy_flow_mh.cc:463-482)
    double bn = 0;
    for (double v : D.rhsC1) bn += v * v;
    D.bnorm1 = std::sqrt(bn);
    // stage 2: IA2 = diag(C1)^-1 on the element block
    Triplets T2;
    for (int e = 0; e < E; e++) {
        double d = D.C1.get(e, e);
        T2.add(e, e, 1.0 / d);
    }
    D.IA2 = T2.to_csr(E, E);
    form_schur(D.C1, D.rhsC1, E, D.IA2, D.IAB2, D.C2, D.rhsC2);
    // scale(-1): schur.cc:702-707
    for (double &v : D.C2.val) v *= -1.0;
    for (double &v : D.rhsC2) v *= -1.0;
}

}  // namespace orc

using namespace orc;

// ============================================================================
// C ABI used by tests (ctypes) — test infrastructure only.
// ============================================================================
extern "C" {

const char *orc_last_error() { return g_err.c_str(); }

// threads of the Jacobi-CG in orc_darcy_solve (CPU baseline timing; 1 = the serial restatement)
void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

void *orc_mesh_load(const char *path) {
    Mesh *m = new Mesh();
    if (!read_gmsh(path, *m) || !setup_topology(*m)) {
        delete m;
        return nullptr;
    }
    return m;
}
void orc_mesh_free(void *h) { delete (Mesh *)h; }
// Build from arrays (same element list semantics as the gmsh reader):
// gmsh types, region ids, 0-based node indices (4 per element, -1 padded),
// physical names (labels starting with '.' are boundary regions).
void *orc_mesh_from_arrays(int n_nodes, const double *xyz, int n_elem, const int *gmsh_type, const int *region_id,
                           const int *nodes4, int n_phys, const int *phys_id, const char *const *phys_name) {
    Mesh *m = new Mesh();
    for (int i = 0; i < n_phys; i++) m->rdb.add_named(phys_id[i], phys_name[i] ? phys_name[i] : "");
    m->X.assign(xyz, xyz + 3 * (size_t)n_nodes);
    m->node_ids.resize(n_nodes);
    for (int i = 0; i < n_nodes; i++) { m->node_ids[i] = i + 1; m->node_idx[i + 1] = i; }
    for (int i = 0; i < n_elem; i++) {
        int t = gmsh_type[i];
        int dim = t == 1 ? 1 : t == 2 ? 2 : t == 4 ? 3 : t == 15 ? 0 : -1;
        if (dim < 0) { g_err = "unsupported element type"; delete m; return nullptr; }
        Element e;
        e.id = i + 1;
        e.dim = dim;
        e.region = m->rdb.add_unnamed(region_id[i]);
        for (int k = 0; k <= dim; k++) e.node[k] = nodes4[4 * (size_t)i + k];
        if (m->rdb.regs[e.region].boundary) m->bc_elements.push_back(e);
        else if (dim > 0) m->element.push_back(e);
    }
    m->n_file_bc = (int)m->bc_elements.size();
    if (!setup_topology(*m)) { delete m; return nullptr; }
    return m;
}


// counts: n_nodes, n_elements, n_bc_elements, n_sides, n_edges, n_boundaries,
//         n_vb, n_regions, n_file_bc
void orc_mesh_counts(void *h, int *out) {
    Mesh &m = *(Mesh *)h;
    out[0] = (int)m.node_ids.size();
    out[1] = (int)m.element.size();
    out[2] = (int)m.bc_elements.size();
    out[3] = m.n_sides;
    out[4] = (int)m.edges.size();
    out[5] = (int)m.boundary.size();
    out[6] = (int)m.vb.size();
    out[7] = (int)m.rdb.regs.size();
    out[8] = m.n_file_bc;
}

// Topology arrays (all int32):
//  elem_dim[E], elem_id[E], elem_region[E] (region id), elem_nodes[4E] (node index, -1 pad),
//  elem_edge[4E], elem_bnd[4E], edge_ptr[F+1], edge_side[S] (global side dof),
//  vb_elem[V], vb_edge[V], bnd_edge[B], bnd_bcele[B], bc_region[NBC] (region id)
void orc_mesh_topology(void *h, int *elem_dim, int *elem_id, int *elem_region, int *elem_nodes, int *elem_edge,
                       int *elem_bnd, int *edge_ptr, int *edge_side, int *vb_elem, int *vb_edge, int *bnd_edge,
                       int *bnd_bcele, int *bc_region) {
    Mesh &m = *(Mesh *)h;
    for (size_t e = 0; e < m.element.size(); e++) {
        const Element &el = m.element[e];
        elem_dim[e] = el.dim;
        elem_id[e] = el.id;
        elem_region[e] = m.rdb.regs[el.region].id;
        for (int k = 0; k < 4; k++) {
            elem_nodes[4 * e + k] = el.node[k];
            elem_edge[4 * e + k] = el.edge_idx[k];
            elem_bnd[4 * e + k] = el.boundary_idx[k];
        }
    }
    edge_ptr[0] = 0;
    int k = 0;
    for (size_t f = 0; f < m.edges.size(); f++) {
        for (auto &s : m.edges[f].sides) edge_side[k++] = m.side_offset[s.first] + s.second;
        edge_ptr[f + 1] = k;
    }
    for (size_t n = 0; n < m.vb.size(); n++) { vb_elem[n] = m.vb[n].element; vb_edge[n] = m.vb[n].edge_idx; }
    for (size_t b = 0; b < m.boundary.size(); b++) { bnd_edge[b] = m.boundary[b].edge_idx; bnd_bcele[b] = m.boundary[b].bc_ele_idx; }
    for (size_t b = 0; b < m.bc_elements.size(); b++) bc_region[b] = m.rdb.regs[m.bc_elements[b].region].id;
}

void orc_mesh_nodes(void *h, double *xyz) {
    Mesh &m = *(Mesh *)h;
    std::memcpy(xyz, m.X.data(), m.X.size() * sizeof(double));
}

// Per-element geometry: measure[E], centre[3E]; per bc element: centre[3NBC], measure[NBC]
void orc_mesh_geometry(void *h, double *measure, double *centre, double *bc_centre, double *bc_measure) {
    Mesh &m = *(Mesh *)h;
    for (size_t e = 0; e < m.element.size(); e++) {
        measure[e] = elem_measure(m, m.element[e]);
        elem_centre(m, m.element[e], centre + 3 * e);
    }
    for (size_t b = 0; b < m.bc_elements.size(); b++) {
        elem_centre(m, m.bc_elements[b], bc_centre + 3 * b);
        bc_measure[b] = elem_measure(m, m.bc_elements[b]);
    }
}

// Region table: ids[R], is_boundary[R], sub_idx[R]; labels joined by '\n' into buf.
void orc_mesh_regions(void *h, int *ids, int *is_boundary, int *sub_idx, char *buf, int buflen) {
    Mesh &m = *(Mesh *)h;
    std::string all;
    for (size_t r = 0; r < m.rdb.regs.size(); r++) {
        ids[r] = m.rdb.regs[r].id;
        is_boundary[r] = m.rdb.regs[r].boundary;
        sub_idx[r] = m.rdb.regs[r].sub_idx;
        all += m.rdb.regs[r].label + "\n";
    }
    std::snprintf(buf, buflen, "%s", all.c_str());
}

// Local RT0 matrix and its inverse for element e; returns det.
double orc_local_matrix(void *h, int e, double K, double cs, const double *aniso, double *loc, double *inv) {
    Mesh &m = *(Mesh *)h;
    return local_matrix(m, m.element[e], K, cs, aniso, loc, inv);
}

void *orc_darcy_create(void *mesh, const double *K, const double *cs, const double *aniso, const double *sigma,
                       const double *f, const int *bc_type, const double *bc_p, const double *bc_flux,
                       const double *bc_sigma) {
    Mesh &m = *(Mesh *)mesh;
    Darcy *D = new Darcy();
    D->m = &m;
    D->E = (int)m.element.size();
    D->S = m.n_sides;
    D->F = (int)m.edges.size();
    D->N = D->S + D->E + D->F;
    int E = D->E, NB = (int)m.bc_elements.size();
    D->K.assign(K, K + E);
    D->cs.assign(cs, cs + E);
    D->sigma.assign(sigma, sigma + E);
    D->f.assign(f, f + E);
    if (aniso) D->aniso.assign(aniso, aniso + 9 * (size_t)E);
    D->bc_type.assign(bc_type, bc_type + NB);
    D->bc_p.assign(bc_p, bc_p + NB);
    D->bc_flux.assign(bc_flux, bc_flux + NB);
    D->bc_sigma.assign(bc_sigma, bc_sigma + NB);
    return D;
}
void orc_darcy_free(void *h) { delete (Darcy *)h; }

// The assembled system solved through both Schur stages (darcy_flow_mh.cc:397-418).
static int solve_assembled(Darcy &D, int method, double rtol, double atol, int max_it);

// Diagonal of the assembled MH matrix set (insert) or added (MatDiagonalSet).
static void diag_apply(Darcy &D, const std::vector<double> &d, bool add) {
    for (int i = 0; i < D.N; i++) {
        bool found = false;
        for (int k = D.M.ptr[i]; k < D.M.ptr[i + 1]; k++)
            if (D.M.col[k] == i) {
                D.M.val[k] = add ? D.M.val[k] + d[i] : d[i];
                found = true;
            }
        if (!found && d[i] != 0.0) { g_err = "missing diagonal entry"; }
    }
}

// Full steady solve (update_solution, darcy_flow_mh.cc:375-433, n_schurs=2).
// method 0: Jacobi-CG on the edge system; method 1: dense LU of the edge system
// (brute force, small meshes); method 2: dense LU of the FULL MH system.
int orc_darcy_solve(void *h, int method, double rtol, double atol, int max_it) {
    Darcy &D = *(Darcy *)h;
    double t0 = now_s();
    D.bad = -1;
    assemble(D);
    if (D.bad >= 0) {
        g_err = "Singular local matrix of the element " + std::to_string(D.m->element[D.bad].id);
        return -1;
    }
    double t1 = now_s();
    D.t_assembly = t1 - t0;
    D.sol.assign(D.N, 0.0);
    if (method == 2) {
        if ((size_t)D.N * D.N > 400000000ull) { g_err = "system too large for dense solve"; return -1; }
        std::vector<double> A((size_t)D.N * D.N, 0.0), x;
        for (int i = 0; i < D.N; i++)
            for (int k = D.M.ptr[i]; k < D.M.ptr[i + 1]; k++) A[(size_t)i * D.N + D.M.col[k]] = D.M.val[k];
        if (!dense_solve(A, D.rhs, D.N, x)) { g_err = "singular"; return -1; }
        D.sol = x;
        return 0;
    }
    return solve_assembled(D, method, rtol, atol, max_it);
}

static int solve_assembled(Darcy &D, int method, double rtol, double atol, int max_it) {
    double t1 = now_s();
    make_schurs(D);
    double t2 = now_s();
    D.t_schur = t2 - t1;
    int F = D.F, E = D.E, S = D.S;
    std::vector<double> lam(F, 0.0);
    if (method == 1) {
        std::vector<double> A((size_t)F * F, 0.0);
        for (int i = 0; i < F; i++)
            for (int k = D.C2.ptr[i]; k < D.C2.ptr[i + 1]; k++) A[(size_t)i * F + D.C2.col[k]] = D.C2.val[k];
        if (!dense_solve(A, D.rhsC2, F, lam)) { g_err = "singular"; return -1; }
        D.iters = 0;
    } else if (method == 3) {
        D.iters = cg_ic0(D.C2, D.rhsC2.data(), lam.data(), rtol, atol, max_it, D.resid);
    } else if (g_threads > 1) {
        D.iters = cg_threads(D.C2, D.rhsC2.data(), lam.data(), rtol, atol, max_it, D.resid);
    } else {
        D.iters = cg(D.C2, D.rhsC2.data(), lam.data(), rtol, atol, max_it, D.resid);
    }
    double t3 = now_s();
    D.t_solve = t3 - t2;
    // schur2->resolve (elements from edges), schur1->resolve (sides)
    std::vector<double> x2(E + F);
    resolve(D.IA2, D.IAB2, D.rhsC1.data(), lam.data(), x2.data());
    for (int f = 0; f < F; f++) x2[E + f] = lam[f];
    std::vector<double> q(S);
    resolve(D.IA1, D.IAB1, D.rhs.data(), x2.data(), q.data());
    std::copy(q.begin(), q.end(), D.sol.begin());
    std::copy(x2.begin(), x2.end(), D.sol.begin() + S);
    D.t_resolve = now_s() - t3;
    return D.iters;
}

// DarcyFlowMH_Unsteady::setup_time_term (darcy_flow_mh.cc:1614-1658), lumped = 0, or
// DarcyFlowLMH_Unsteady::setup_time_term (:1695-1749), lumped = 1: the steady
// system is assembled, its diagonal and rhs saved, the time term of step dt added
// to the diagonal and the initial condition put into the solution.
void orc_darcy_unsteady_setup(void *h, int lumped, const double *storativity, const double *init_pressure,
                              double dt) {
    Darcy &D = *(Darcy *)h;
    Mesh &m = *D.m;
    assemble(D);
    D.unsteady = lumped ? 2 : 1;
    D.stor.assign(storativity, storativity + D.E);
    D.dt = dt;
    D.steady_diag.assign(D.N, 0.0);
    for (int i = 0; i < D.N; i++) D.steady_diag[i] = D.M.get(i, i);
    D.sol.assign(D.N, 0.0);
    D.new_diag.assign(D.N, 0.0);
    for (int e = 0; e < D.E; e++) {
        const Element &el = m.element[e];
        double meas = elem_measure(m, el);
        if (!lumped) {
            D.sol[D.S + e] = init_pressure[e];
            D.new_diag[D.S + e] = -D.stor[e] * meas / dt;
        } else {
            int ns = el.n_sides();
            for (int i = 0; i < ns; i++) {
                int edge_row = D.S + D.E + el.edge_idx[i];
                D.new_diag[edge_row] += -meas * D.stor[e] * D.cs[e] / dt / ns;
                D.sol[edge_row] += init_pressure[e] / ns;
            }
        }
    }
    diag_apply(D, D.new_diag, true);
    D.prev_sol = D.sol;
    D.steady_rhs = D.rhs;
}

// One step: update_solution with modify_system (darcy_flow_mh.cc:1660-1677 MH,
// :1751-1770 LMH) and, for LMH, postprocess (:1772-1827).  Returns CG iterations.
int orc_darcy_unsteady_step(void *h, double dt, int method, double rtol, double atol, int max_it) {
    Darcy &D = *(Darcy *)h;
    Mesh &m = *D.m;
    if (dt != D.dt) {
        diag_apply(D, D.steady_diag, false);
        for (double &v : D.new_diag) v *= D.dt / dt;
        diag_apply(D, D.new_diag, true);
        D.dt = dt;
    }
    for (int i = 0; i < D.N; i++) D.rhs[i] = D.new_diag[i] * D.sol[i] + D.steady_rhs[i];
    D.prev_sol = D.sol;
    int it = solve_assembled(D, method, rtol, atol, max_it);
    if (D.unsteady == 2) {
        for (size_t f = 0; f < m.edges.size(); f++) {
            int row = D.S + D.E + (int)f;
            double new_pressure = D.sol[row], old_pressure = D.prev_sol[row];
            for (auto &sd : m.edges[f].sides) {
                const Element &el = m.element[sd.first];
                double time_coef = -elem_measure(m, el) * D.cs[sd.first] * D.stor[sd.first] / dt / el.n_sides();
                D.sol[m.side_offset[sd.first] + sd.second] += time_coef * (new_pressure - old_pressure);
            }
        }
        for (int e = 0; e < D.E; e++) {
            const Element &el = m.element[e];
            for (int i = 0; i < el.n_sides(); i++)
                D.sol[m.side_offset[e] + i] += -1.0 * elem_measure(m, el) * D.cs[e] * D.f[e] / el.n_sides();
        }
    }
    return it;
}

int orc_darcy_size(void *h, int *out) {
    Darcy &D = *(Darcy *)h;
    out[0] = D.S; out[1] = D.E; out[2] = D.F; out[3] = D.N;
    out[4] = (int)D.C2.col.size();
    return D.N;
}
void orc_darcy_solution(void *h, double *out) {
    Darcy &D = *(Darcy *)h;
    std::memcpy(out, D.sol.data(), D.sol.size() * sizeof(double));
}
// stats: iters, residual, ||schur1 rhs||, t_assembly, t_schur, t_solve, t_resolve
void orc_darcy_stats(void *h, double *out) {
    Darcy &D = *(Darcy *)h;
    out[0] = D.iters; out[1] = D.resid; out[2] = D.bnorm1;
    out[3] = D.t_assembly; out[4] = D.t_schur; out[5] = D.t_solve; out[6] = D.t_resolve;
}
// Final SPD edge system (after both Schur stages and scale(-1)) as CSR.
void orc_darcy_edge_system(void *h, int *ptr, int *col, double *val, double *rhs) {
    Darcy &D = *(Darcy *)h;
    std::memcpy(ptr, D.C2.ptr.data(), D.C2.ptr.size() * sizeof(int));
    std::memcpy(col, D.C2.col.data(), D.C2.col.size() * sizeof(int));
    std::memcpy(val, D.C2.val.data(), D.C2.val.size() * sizeof(double));
    std::memcpy(rhs, D.rhsC2.data(), D.rhsC2.size() * sizeof(double));
}
// MH residual ||M x - b|| / ||b|| of the current solution (full system).
double orc_darcy_mh_residual(void *h) {
    Darcy &D = *(Darcy *)h;
    std::vector<double> r(D.N);
    spmv(D.M, D.sol.data(), r.data());
    double rn = 0, bn = 0;
    for (int i = 0; i < D.N; i++) { double d = r[i] - D.rhs[i]; rn += d * d; bn += D.rhs[i] * D.rhs[i]; }
    return std::sqrt(rn) / std::max(std::sqrt(bn), 1e-300);
}

// MHFEValues::RT0_value (mh_fe_values.cc:76-118) with the basis coefficients of
// basis_functions_triangle / _tetrahedron (:268-295, :379-397).
static void rt0_value(const Mesh &m, const Element &e, const double *x, int face, double *out) {
    if (e.dim == 1) {
        double lv[3], d[3];
        sub3(m.P(e.node[1]), m.P(e.node[0]), lv);
        double ll = dot3(lv, lv);
        if (face == 0) {
            sub3(x, m.P(e.node[1]), d);
            double s = -norm3(d);
            for (int k = 0; k < 3; k++) out[k] = s * lv[k] / ll;
        } else {
            sub3(x, m.P(e.node[0]), d);
            double s = norm3(d);
            for (int k = 0; k < 3; k++) out[k] = s * lv[k] / ll;
        }
    } else if (e.dim == 2) {
        double ex[3], ac[3], ez[3], ey[3], u[3], v[3], n[3], t[3], b[3];
        sub3(m.P(e.node[1]), m.P(e.node[0]), ex);
        normalize3(ex);
        sub3(m.P(e.node[2]), m.P(e.node[0]), ac);
        cross3(ex, ac, ez);
        normalize3(ez);
        cross3(ez, ex, ey);
        normalize3(ey);
        // node_coordinates_triangle (:218-262)
        sub3(m.P(e.node[1]), m.P(e.node[0]), u);
        sub3(m.P(e.node[2]), m.P(e.node[0]), v);
        cross3(u, v, n);
        normalize3(n);
        t[0] = u[0]; t[1] = u[1]; t[2] = u[2];
        normalize3(t);
        cross3(n, u, b);
        normalize3(b);
        double nod[3][2] = {{0.0, 0.0}, {dot3(t, u), dot3(b, u)}, {dot3(t, v), dot3(b, v)}};
        int i0 = face, i1 = (face + 1) % 3, i2 = (face + 2) % 3;
        double alfa = nod[i2][0], beta = nod[i2][1];
        double gama = 1.0 / std::fabs((nod[i0][0] - nod[i2][0]) * (nod[i1][1] - nod[i0][1]) +
                                      (nod[i0][1] - nod[i2][1]) * (nod[i0][0] - nod[i1][0]));
        double w[3];
        sub3(x, m.P(e.node[0]), w);
        double cx = dot3(w, ex) - alfa, cy = dot3(w, ey) - beta;
        for (int k = 0; k < 3; k++) out[k] = gama * (cx * ex[k] + cy * ey[k]);
    } else {
        const double *pf = m.P(e.node[face]);
        double nrm[3], sc[3];
        side_normal(m, e, face, nrm);
        side_centre(m, e, face, sc);
        double delta = 1.0 / (side_measure(m, e, face) * (nrm[0] * sc[0] + nrm[1] * sc[1] + nrm[2] * sc[2] -
                                                          nrm[0] * pf[0] - nrm[1] * pf[1] - nrm[2] * pf[2]));
        for (int k = 0; k < 3; k++) out[k] = delta * (x[k] - pf[k]);
    }
}

// DarcyFlowMHOutput::make_element_scalar / make_element_vector
// (darcy_flow_mh_output.cc:261-325): pressure, p + z of the centre, centre velocity.
void orc_darcy_element_output(void *h, double *pressure, double *head, double *velocity) {
    Darcy &D = *(Darcy *)h;
    Mesh &m = *D.m;
    for (int e = 0; e < D.E; e++) {
        const Element &el = m.element[e];
        double c[3];
        elem_centre(m, el, c);
        double p = D.sol[D.S + e];
        pressure[e] = p;
        head[e] = p + c[2];
        double f[3] = {0.0, 0.0, 0.0};
        for (int li = 0; li < el.n_sides(); li++) {
            double b[3];
            rt0_value(m, el, c, li, b);
            double q = D.sol[m.side_offset[e] + li];
            for (int k = 0; k < 3; k++) f[k] += q * b[k] / D.cs[e];
        }
        for (int k = 0; k < 3; k++) velocity[3 * e + k] = f[k];
    }
}

// DarcyFlowMHOutput::make_node_scalar_param (darcy_flow_mh_output.cc:424-541):
// first pass sums the distances node -> element centre and node -> side centre and
// counts them, second pass adds element pressure and side trace with weight
// (1 - d / sum d) / (count - 1); elements first, then the sides of every element.
void orc_darcy_node_pressure(void *h, double *out) {
    Darcy &D = *(Darcy *)h;
    Mesh &m = *D.m;
    const int N = (int)m.X.size() / 3;
    std::vector<int> ne(N, 0), ns(N, 0);
    std::vector<double> de(N, 0.0), ds(N, 0.0);
    std::fill(out, out + N, 0.0);
    auto dist = [&](int v, const double *c) {
        const double *p = m.P(v);
        return std::sqrt((p[0] - c[0]) * (p[0] - c[0]) + (p[1] - c[1]) * (p[1] - c[1]) + (p[2] - c[2]) * (p[2] - c[2]));
    };
    for (int pass = 0; pass < 2; pass++) {
        for (int e = 0; e < D.E; e++) {
            const Element &el = m.element[e];
            double c[3];
            elem_centre(m, el, c);
            for (int k = 0; k < el.n_nodes(); k++) {
                int v = el.node[k];
                double d = dist(v, c);
                if (pass == 0) de[v] += d, ne[v]++;
                else out[v] += D.sol[D.S + e] * (1 - d / (de[v] + ds[v])) / (ne[v] + ns[v] - 1);
            }
        }
        for (int e = 0; e < D.E; e++) {
            const Element &el = m.element[e];
            for (int s = 0; s < el.n_sides(); s++) {
                double c[3];
                side_centre(m, el, s, c);
                for (int k = 0; k < el.dim; k++) {
                    int v = snode(m, el, s, k);
                    double d = dist(v, c);
                    if (pass == 0) ds[v] += d, ns[v]++;
                    else out[v] += D.sol[D.S + D.E + el.edge_idx[s]] * (1 - d / (de[v] + ds[v])) / (ne[v] + ns[v] - 1);
                }
            }
        }
    }
}

// Water balance per boundary region: DarcyFlowMHOutput::water_balance,
// darcy_flow_mh_output.cc:671-773.  out arrays indexed by boundary region
// sub-index: balance, plus (outflow), minus (inflow).
void orc_darcy_water_balance(void *h, double *bal, double *plus, double *minus) {
    Darcy &D = *(Darcy *)h;
    Mesh &m = *D.m;
    int nbr = m.rdb.n_boundary;
    std::fill(bal, bal + nbr, 0.0);
    std::fill(plus, plus + nbr, 0.0);
    std::fill(minus, minus + nbr, 0.0);
    for (auto &b : m.boundary) {
        auto s = m.edges[b.edge_idx].sides[0];
        double flux = D.sol[m.side_offset[s.first] + s.second];
        int r = m.rdb.regs[m.bc_elements[b.bc_ele_idx].region].sub_idx;
        bal[r] += flux;
        if (flux > 0) plus[r] += flux;
        else minus[r] += flux;
    }
}

// Per-element balance of the element row of the MH system
// (darcy_flow_mh.cc:603-640): sum_i q_i + sum_vb v (p_e - lambda_n) - |e| cs f.
void orc_darcy_element_balance(void *h, double *out) {
    Darcy &D = *(Darcy *)h;
    Mesh &m = *D.m;
    for (int e = 0; e < D.E; e++) {
        const Element &el = m.element[e];
        double s = 0;
        for (int i = 0; i < el.n_sides(); i++) s += D.sol[m.side_offset[e] + i];
        double pe = D.sol[D.S + e];
        for (int n : el.neigh_vb) {
            const auto &hs = m.edges[m.vb[n].edge_idx].sides[0];
            double v = D.sigma[e] * side_measure(m, m.element[hs.first], hs.second) * D.cs[hs.first];
            s += v * (pe - D.sol[D.S + D.E + m.vb[n].edge_idx]);
        }
        out[e] = s - elem_measure(m, el) * D.cs[e] * D.f[e];
    }
}

}  // extern "C"

// ============================================================================
// Explicit transport: src/transport/transport.cc, TimeGovernor
// src/coupling/time_governor.cc:261-355, MassBalance mass_balance.cc:79-239.
// ============================================================================
namespace orc {

static const double INF = std::numeric_limits<double>::infinity();
static const double EPS = std::numeric_limits<double>::epsilon();

struct TimeGov {  // the non-steady TimeGovernor restated for one equation
    double time = 0, end_time = 0, time_step = 0, last_time_step = 0, fixed_dt = 0, end_fixed = 0;
    double upper = INF, lower = EPS, max_ts = INF, min_ts = EPS;
    bool dt_fixed_now = false, dt_changed = true;
    int level = 0;
    std::vector<double> marks;  // fixed-time marks, sorted
    bool lt(double o) const { double b = o - 16 * EPS * std::max(std::fabs(time), std::fabs(o)); return !(time >= b); }
    bool ge(double o) const { double b = o - 16 * EPS * std::max(std::fabs(time), std::fabs(o)); return time >= b; }
    bool is_end() const { return ge(end_time) || time == INF; }
    void add_mark(double t) {
        auto it = std::lower_bound(marks.begin(), marks.end(), t);
        if (it != marks.end() && *it == t) return;
        marks.insert(it, t);
    }
    double next_mark() const {  // TimeMarks::next, time_marks.cc:125-132
        auto it = std::lower_bound(marks.begin(), marks.end(), time);
        while (it != marks.end() && !lt(*it)) ++it;
        return it == marks.end() ? INF : *it;
    }
    double estimate_dt() const {  // time_governor.cc:261-307
        if (time == INF || is_end()) return 0.0;
        if (lt(end_fixed)) return fixed_dt;
        double full = next_mark() - time;
        double est = std::min(full, upper);
        est = std::max(est, lower);
        int n = (int)std::ceil(full / est);
        int nf = (int)std::floor(full / est);
        if (std::fabs(full / est - nf) < 1e-14) n = nf;
        return full / n;
    }
    int set_upper_constraint(double u) {  // time_governor.cc:215-237
        if (upper < u) return -1;
        if (lower <= u) { upper = u; return 0; }
        return 1;
    }
    double fix_dt_until_mark() {  // time_governor.hh:205-211
        end_fixed = -INF;
        fixed_dt = estimate_dt();
        dt_fixed_now = true;
        return end_fixed = next_mark();
    }
    void next_time() {  // time_governor.cc:314-355
        if (time == INF || is_end()) return;
        if (lt(end_fixed)) {
            fixed_dt = (end_fixed - time) / std::round((end_fixed - time) / fixed_dt);
            last_time_step = time_step;
            time_step = fixed_dt;
            if (dt_fixed_now) { dt_fixed_now = false; dt_changed = (last_time_step != time_step); }
            else dt_changed = false;
        } else {
            last_time_step = time_step;
            time_step = estimate_dt();
            dt_changed = (last_time_step != time_step);
        }
        time += time_step;
        level++;
        upper = std::min(end_time - time, max_ts);
        lower = min_ts;
    }
};

struct Transport {
    Mesh *m;
    int E, nsub;
    std::vector<double> flux;  // side flux [S]
    std::vector<double> cs, por;
    std::vector<double> bc_conc;    // [nsub][NBC]
    std::vector<double> bc_pending; // new boundary data, read by the next step
    bool bc_changed = false;
    // boundary data as a function of time (a time-dependent FieldFormula): evaluated
    // at every step's start time
    void (*bc_fn)(double, double *, void *) = nullptr;
    void *bc_user = nullptr;
    std::vector<double> conc, pconc;  // [nsub][E]
    std::vector<double> bcvcorr;    // [nsub][E]
    std::vector<double> src_density, src_sigma, src_conc;  // [nsub][E], empty = no sources
    // Linear_reaction: decays (parent, products, half-life, branch ratios), matrix [nsub][nsub]
    std::vector<int> r_parent, r_pptr, r_prod;
    std::vector<double> r_hl, r_branch, reaction_matrix;
    int pade_nom = -1, pade_den = -1;  // Pade_approximant degrees, -1: Linear_reaction
    // dual porosity and sorption (transport.cc:1026-1132): immobile, mobile-sorbed and
    // immobile-sorbed concentrations [nsub][E]; por_imm, phi [E]; alpha, isotherm
    // type and coefficients [nsub][E]
    bool dual = false, sorp = false;
    std::vector<double> c_imm, c_msorb, c_isorb, por_imm, phi, alpha, sorp_c0, sorp_c1;
    std::vector<int> sorp_type;
    // reaction-module sorption (reaction/sorption.cc, isotherm.cc): per species its
    // substance, per (bulk region, species) the isotherm and its interpolation table
    struct Iso {
        int type = 0;
        double c0 = 0, c1 = 0, scale_aqua = 0, scale_sorbed = 0, rho = 1, limit = 0, step = 0;
        bool limited = false;
        std::vector<double> table;
    };
    std::vector<int> rs_subst;
    std::vector<Iso> rs_iso;           // [n_bulk_regions][n_species] (tabulated regions)
    std::vector<Iso> rs_el;            // [E][n_species] (elements of the other regions)
    std::vector<char> rs_tabled;       // per bulk region
    std::vector<double> rs_sorbed;     // [n_species][E]
    // the bulk_data fields per element: type, mult_coef, second_param [n][E],
    // rock_density [E]; molar masses, solubilities [n]
    std::vector<int> rs_type;
    std::vector<double> rs_mult, rs_second, rs_rock, rs_molar, rs_sol;
    double rs_solvent = 1.0;
    int rs_substeps = 0;
    bool rs_elementwise = false, rs_built = false;
    CSR tm;
    std::vector<int> diag_pos;
    double cfl_max_step = 0;
    bool matrix_built = false, scaled = false, need_rescale = true;
    TimeGov tg;
    // mass balance
    int nbr = 0, nblk = 0;
    std::vector<double> bcd_bal, bcd_plus, bcd_minus, mass, src;  // [nsub][nreg]
    std::vector<double> initial_mass, integrated_fluxes, integrated_sources, last_total_flux;
    double last_time = 0;
    bool mb_initial = true;
    int steps = 0;
    double t_steps = 0;
};

// create_transport_matrix_mpi: transport.cc:788-943
static void build_tm(Transport &T) {
    Mesh &m = *T.m;
    int E = T.E;
    std::vector<double> edge_flow(m.edges.size(), 0.0);
    for (size_t i = 0; i < m.edges.size(); i++)
        for (auto &s : m.edges[i].sides) {
            double fl = T.flux[m.side_offset[s.first] + s.second];
            if (fl > 0) edge_flow[i] += fl;
        }
    // INSERT_VALUES: later insertions of the same entry overwrite
    std::vector<std::map<int, double>> rows(E);
    double max_sum = 0.0;
    for (int e = 0; e < E; e++) {
        const Element &el = m.element[e];
        double aii = 0.0;
        double vol = elem_measure(m, el) * T.cs[e] * T.por[e];
        for (int si = 0; si < el.n_sides(); si++) {
            double fl = T.flux[m.side_offset[e] + si];
            if (el.boundary_idx[si] < 0) {
                const Edge &edg = m.edges[el.edge_idx[si]];
                double edg_flux = edge_flow[el.edge_idx[si]];
                for (auto &s : edg.sides) {
                    if (s.first == e && s.second == si) continue;
                    double fl2 = T.flux[m.side_offset[s.first] + s.second];
                    double aij;
                    if (fl2 > 0.0 && fl < 0.0) aij = -(fl * fl2 / (edg_flux * elem_measure(m, el) * T.cs[e] * T.por[e]));
                    else aij = 0;
                    rows[e][s.first] = aij;
                }
                if (fl > 0.0) aii -= (fl / (elem_measure(m, el) * T.cs[e] * T.por[e]));
            } else {
                if (fl > 0.0) aii -= (fl / (elem_measure(m, el) * T.cs[e] * T.por[e]));
            }
        }
        (void)vol;
        for (int n : el.neigh_vb) {
            const auto &hs = m.edges[m.vb[n].edge_idx].sides[0];
            int e2 = hs.first;
            double fl = T.flux[m.side_offset[e2] + hs.second];
            double aij = fl > 0.0 ? fl / (elem_measure(m, el) * T.cs[e] * T.por[e]) : 0;
            rows[e][e2] = aij;
            if (fl < 0.0) {
                aii -= (-fl) / (elem_measure(m, el) * T.cs[e] * T.por[e]);
                aij = (-fl) / (elem_measure(m, m.element[e2]) * T.cs[e2] * T.por[e2]);
            } else aij = 0;
            rows[e2][e] = aij;
        }
        rows[e][e] = aii;
        if (std::fabs(aii) > max_sum) max_sum = std::fabs(aii);
    }
    T.cfl_max_step = 1 / max_sum;
    T.tm = CSR();
    T.tm.nr = T.tm.nc = E;
    T.tm.ptr.assign(E + 1, 0);
    T.diag_pos.assign(E, -1);
    for (int e = 0; e < E; e++) {
        for (auto &kv : rows[e]) {
            if (kv.first == e) T.diag_pos[e] = (int)T.tm.col.size();
            T.tm.col.push_back(kv.first);
            T.tm.val.push_back(kv.second);
        }
        T.tm.ptr[e + 1] = (int)T.tm.col.size();
    }
    T.matrix_built = true;
    T.scaled = false;
}

// set_boundary_conditions: transport.cc:455-495
static void set_bc(Transport &T) {
    Mesh &m = *T.m;
    std::fill(T.bcvcorr.begin(), T.bcvcorr.end(), 0.0);
    int NBC = (int)m.bc_elements.size();
    for (int e = 0; e < T.E; e++) {
        const Element &el = m.element[e];
        for (int si = 0; si < el.n_sides(); si++) {
            int b = el.boundary_idx[si];
            if (b < 0) continue;
            double fl = T.flux[m.side_offset[e] + si];
            if (fl < 0.0) {
                double aij = -(fl / (elem_measure(m, el) * T.cs[e] * T.por[e]));
                int bce = m.boundary[b].bc_ele_idx;
                for (int s = 0; s < T.nsub; s++) T.bcvcorr[(size_t)s * T.E + e] += T.bc_conc[(size_t)s * NBC + bce] * aij;
            }
        }
    }
}

// MassBalance::calculate: mass_balance.cc:79-239 with calc_fluxes /
// calc_elem_sources transport.cc:1299-1366 (no sources in these runs).
static void mb_calculate(Transport &T, double time) {
    if (!T.mb_initial && T.last_time == time) return;
    Mesh &m = *T.m;
    int ns = T.nsub, NBC = (int)m.bc_elements.size();
    std::fill(T.bcd_bal.begin(), T.bcd_bal.end(), 0.0);
    std::fill(T.bcd_plus.begin(), T.bcd_plus.end(), 0.0);
    std::fill(T.bcd_minus.begin(), T.bcd_minus.end(), 0.0);
    std::fill(T.mass.begin(), T.mass.end(), 0.0);
    std::fill(T.src.begin(), T.src.end(), 0.0);
    for (auto &b : m.boundary) {
        auto s = m.edges[b.edge_idx].sides[0];
        double wf = T.flux[m.side_offset[s.first] + s.second];
        int r = m.rdb.regs[m.bc_elements[b.bc_ele_idx].region].sub_idx;
        for (int k = 0; k < ns; k++) {
            double mf = wf < 0 ? wf * T.bc_conc[(size_t)k * NBC + b.bc_ele_idx] : wf * T.pconc[(size_t)k * T.E + s.first];
            T.bcd_bal[k * T.nbr + r] += mf;
            if (mf > 0) T.bcd_plus[k * T.nbr + r] += mf;
            else T.bcd_minus[k * T.nbr + r] += mf;
        }
    }
    for (int k = 0; k < ns; k++)
        for (int e = 0; e < T.E; e++) {
            const Element &el = m.element[e];
            int r = m.rdb.regs[el.region].sub_idx;
            // the concentrations of every phase in use, summed, times por_m (transport.cc:1354-1362)
            size_t j = (size_t)k * T.E + e;
            double sum_sol_phases = 0;
            sum_sol_phases += T.conc[j];
            if (T.dual) sum_sol_phases += T.c_imm[j];
            if (T.sorp) sum_sol_phases += T.c_msorb[j];
            if (T.dual && T.sorp) sum_sol_phases += T.c_isorb[j];
            T.mass[k * T.nblk + r] += T.por[e] * T.cs[e] * sum_sol_phases * elem_measure(m, el);
            if (!T.src_density.empty()) {
                // compute_concentration_sources_for_mass_balance (transport.cc:546-590) with pconc
                size_t j = (size_t)k * T.E + e;
                double diff = T.src_conc[j] - T.pconc[j];
                double snd = diff > 0.0 ? T.src_density[j] + diff * T.src_sigma[j] : T.src_density[j];
                T.src[k * T.nblk + r] += snd * elem_measure(m, el);
            }
        }
    std::vector<double> tot_bal(ns, 0.0), tot_src(ns, 0.0), mass_tot(ns, 0.0);
    for (int k = 0; k < ns; k++) {
        for (int r = 0; r < T.nbr; r++) tot_bal[k] += T.bcd_bal[k * T.nbr + r];
        for (int r = 0; r < T.nblk; r++) { mass_tot[k] += T.mass[k * T.nblk + r]; tot_src[k] += T.src[k * T.nblk + r]; }
    }
    if (T.mb_initial) {
        T.last_time = time;
        for (int k = 0; k < ns; k++) T.initial_mass[k] = mass_tot[k];
        T.mb_initial = false;
    }
    for (int k = 0; k < ns; k++) {
        T.integrated_sources[k] += tot_src[k] * (time - T.last_time);
        T.integrated_fluxes[k] += tot_bal[k] * (time - T.last_time);
    }
    T.last_time = time;
}

// compute_one_step: transport.cc:593-715
// ConvectionTransport::transport_dual_porosity (transport.cc:1026-1057): the mobile and
// immobile concentrations relax towards their porosity-weighted average with
// rate alpha (por_m + por_imm) / (por_m por_imm) over the step dt.
static void dual_porosity(Transport &T, int i, int k, double dt) {
    size_t j = (size_t)k * T.E + i;
    double por_m = T.por[i], por_imm = T.por_imm[i], alpha = T.alpha[j];
    double pcm = T.conc[j], pci = T.c_imm[j];
    double conc_avg = ((por_m * pcm) + (por_imm * pci)) / (por_m + por_imm);
    if ((conc_avg != 0.0) && (por_imm != 0.0)) {
        double f = std::exp(-alpha * ((por_m + por_imm) / (por_m * por_imm)) * dt);
        T.conc[j] = (pcm - conc_avg) * f + conc_avg;
        T.c_imm[j] = (pci - conc_avg) * f + conc_avg;
    }
}

// ConvectionTransport::compute_sorption (transport.cc:1098-1160): splits the pore
// mass conc_avg into the dissolved concentration concx (linear, Freundlich by at
// most 20 Newton steps from the kink cz, Langmuir by at most 5) and the sorbed
// concentration (conc_avg - concx) Nv.  Isotherm types: 0 none, 1 linear,
// 2 freundlich, 3 langmuir (reaction/isotherm.hh:71-76).
static void compute_sorption(double conc_avg, double c0, double c1, int type, double *concx, double *concx_sorb,
                             double Nv, double N) {
    double Kx = c0 * N;
    const double ad = 1e4, tolerence = 1e-8;
    double pNR = 0, NR, cz, tcz;
    if (type == 1) {
        *concx = conc_avg / (1 + Kx);
    } else if (type == 2) {
        cz = std::pow(ad / (Kx * c1), 1 / (c1 - 1));
        tcz = ad / c1;
        NR = cz;
        for (int i = 0; i < 20; i++) {
            NR -= (NR + ((NR > cz) ? Kx * std::pow(NR, c1) : tcz * NR) - conc_avg) /
                  (1 + ((NR > cz) ? c1 * Kx * std::pow(NR, c1 - 1) : tcz));
            if ((NR <= cz) || (std::fabs(NR - pNR) < tolerence * NR)) break;
            pNR = NR;
        }
        *concx = NR;
    } else if (type == 3) {
        NR = 0;
        for (int i = 0; i < 5; i++) {
            NR -= (NR + (N * NR * c1 * c0) / (1 + NR * c0) - conc_avg) /
                  (1 + N * c0 * c1 / std::pow((1 + NR * c0), 2));
            if (std::fabs(NR - pNR) < tolerence * NR) break;
            pNR = NR;
        }
        *concx = NR;
    }
    *concx_sorb = (conc_avg - *concx) * Nv;
}

// ConvectionTransport::transport_sorption (transport.cc:1061-1094)
static void sorption(Transport &T, int i, int k) {
    size_t j = (size_t)k * T.E + i;
    double phi = T.phi[i], por_m = T.por[i], por_imm = T.por_imm[i];
    double n = 1 - (por_m + por_imm), Nm = por_m, Nimm = por_imm;
    double conc_avg = T.conc[j] + T.c_msorb[j] * n / Nm;
    if (conc_avg != 0)
        compute_sorption(conc_avg, T.sorp_c0[j], T.sorp_c1[j], T.sorp_type[j], &T.conc[j], &T.c_msorb[j], Nm / n,
                         n * phi / Nm);
    if (T.dual && (por_imm != 0)) {
        double conc_avg_imm = T.c_imm[j] + T.c_isorb[j] * n / Nimm;
        if (conc_avg_imm != 0)
            compute_sorption(conc_avg_imm, T.sorp_c0[j], T.sorp_c1[j], T.sorp_type[j], &T.c_imm[j], &T.c_isorb[j],
                             Nimm / n, n * (1 - phi) / Nimm);
    }
}

// Isotherm functors (isotherm.hh:100-170): 0 none, 1 linear, 2 freundlich, 3 langmuir
static double iso_func(const Transport::Iso &I, double x) {
    if (I.type == 1) return I.c0 * x;
    if (I.type == 2) return I.c0 * std::pow(x, I.c1);
    if (I.type == 3) return I.c0 * (I.c1 * x) / (I.c1 * x + 1);
    return 0.0;
}
// Isotherm::solve_conc (isotherm.cc:86-109): the aqueous concentration where
// scale_sorbed f(c / rho) + scale_aqua c = total_mass on [0, critical mass / scale_aqua
// + 1e-5].  The reference brackets with boost's TOMS 748 to 30 bits and takes the
// bracket's midpoint; this restatement bisects to the last bit (the two agree to
// about 1e-9 of the value).  toms748_solve throws for an empty interval and for
// ends of one sign: false there.
static bool iso_solve(const Transport::Iso &I, double &c_aqua, double &c_sorbed) {
    double total = I.scale_aqua * c_aqua + I.scale_sorbed * c_sorbed;
    double critic = I.limit * I.scale_aqua + iso_func(I, I.limit / I.rho) * I.scale_sorbed;
    double a = 0.0, b = critic / I.scale_aqua + 0.00001;
    if (!(a < b)) return false;
    auto F = [&](double c) { return I.scale_sorbed * iso_func(I, c / I.rho) + I.scale_aqua * c - total; };
    double fa = F(a), fb = F(b);
    if (!std::isfinite(fa) || !std::isfinite(fb)) return false;
    // an end point that is a root is returned as it is (both ends of the bracket)
    if (fa == 0.0) b = a;
    else if (fb == 0.0) a = b;
    else if ((fa < 0) == (fb < 0)) return false;
    // up to 1200 halvings: enough to reach the last bit of any double in [0, b]
    // (a root near zero, as for Freundlich at tiny concentrations, needs hundreds)
    for (int it = 0; it < 1200 && a != b; it++) {
        double m = 0.5 * (a + b);
        if (!(m > a && m < b)) break;
        double fm = F(m);
        if ((fm < 0) == (fa < 0)) a = m, fa = fm;
        else b = m;
    }
    c_aqua = (a + b) / 2;
    c_sorbed = (total - I.scale_aqua * c_aqua) / I.scale_sorbed;
    return true;
}
// Isotherm::make_table (isotherm.cc:220-242): n + 1 rotated sorbed values at equidistant
// total masses up to the mass of the table limit (1 where no solubility limit is set)
static bool iso_table(Transport::Iso &I, int n) {
    if (I.limit <= 0.0) I.limit = 1.0;
    double mass_limit = I.scale_aqua * I.limit + I.scale_sorbed * iso_func(I, I.limit / I.rho);
    I.step = mass_limit / n;
    double mass = 0.0;
    I.table.clear();
    for (int i = 0; i <= n; i++) {
        double c_aqua = mass / I.scale_aqua, c_sorbed = 0.0;
        if (!iso_solve(I, c_aqua, c_sorbed)) return false;
        I.table.push_back(c_sorbed * I.scale_aqua - c_aqua * I.scale_sorbed);
        mass = mass + I.step;
    }
    return true;
}
// Isotherm::compute_projection (isotherm.cc:56-84): the total mass's position in the
// table, linear interpolation of the rotated coordinate, rotation back; beyond the
// table the solubility limit (precipitate, :244-258) or a direct solve.
// 0, or 1 for a negative total mass (UsrErr in the reference), 2 without a root.
static int iso_project(const Transport::Iso &I, double &c_aqua, double &c_sorbed) {
    double inv_a = I.scale_aqua / (I.scale_aqua * I.scale_aqua + I.scale_sorbed * I.scale_sorbed);
    double inv_s = I.scale_sorbed / (I.scale_aqua * I.scale_aqua + I.scale_sorbed * I.scale_sorbed);
    double total = (I.scale_aqua * c_aqua + I.scale_sorbed * c_sorbed);
    double steps = total / I.step;
    int idx = (int)std::floor(steps);
    if (idx < 0) return 1;
    if ((size_t)idx < I.table.size() - 1) {
        double rot = I.table[idx] + (steps - idx) * (I.table[idx + 1] - I.table[idx]);
        c_aqua = (total * inv_a - rot * inv_s);
        c_sorbed = (total * inv_s + rot * inv_a);
    } else if (I.limited) {
        c_aqua = I.limit;
        c_sorbed = (total - I.scale_aqua * I.limit) / I.scale_sorbed;
    } else {
        double a = c_aqua, s = c_sorbed;
        if (!iso_solve(I, a, s)) return 2;
        c_aqua = a, c_sorbed = s;
    }
    return 0;
}
// Isotherm::compute_reaction (isotherm.cc:184-200): with a table the projection,
// without one the solubility limit where c_aqua exceeds it, else a direct solve.
static int iso_reaction(const Transport::Iso &I, double &c_aqua, double &c_sorbed) {
    if (!I.table.empty()) return iso_project(I, c_aqua, c_sorbed);
    if (I.limited && c_aqua > I.limit) {
        double total = (I.scale_aqua * c_aqua + I.scale_sorbed * c_sorbed);
        c_aqua = I.limit;
        c_sorbed = (total - I.scale_aqua * I.limit) / I.scale_sorbed;
        return 0;
    }
    double a = c_aqua, s = c_sorbed;
    if (!iso_solve(I, a, s)) return 2;
    c_aqua = a, c_sorbed = s;
    return 0;
}
// Sorption::isotherm_reinit (sorption.cc:172-213) from element e's values: scale_aqua
// = por_m, scale_sorbed = phi (1 - por_m - por_imm) rock_density molar_mass.
static Transport::Iso iso_reinit(const Transport &T, int e, int k) {
    Transport::Iso I;
    double por_m = T.por[e], por_imm = T.por_imm.empty() ? 0.0 : T.por_imm[e], phi = T.phi.empty() ? 1.0 : T.phi[e];
    I.type = T.rs_type[(size_t)k * T.E + e];
    I.c0 = T.rs_mult[(size_t)k * T.E + e];
    I.c1 = T.rs_second[(size_t)k * T.E + e];
    I.rho = T.rs_solvent;
    I.scale_aqua = por_m;
    I.scale_sorbed = phi * (1 - por_m - por_imm) * T.rs_rock[e] * T.rs_molar[k];
    I.limit = T.rs_sol[k];
    I.limited = I.limit > 0.0;
    return I;
}
// Sorption::make_tables (sorption.cc:148-170): a region whose parameters are all
// constant (the reference: FieldConstant on the region; here: equal values, and never
// for element-wise parameter fields) gets one table per species made from its first
// element; every element of another region keeps its own isotherm.  Returns 0, or -1
// with a zero scale_sorbed (UsrErr in the reference) or a table without a root.
static int make_sorption(Transport &T) {
    Mesh &m = *T.m;
    int nreg = m.rdb.n_bulk;
    size_t ns = T.rs_subst.size();
    auto same = [&](int i, int j) {
        double pi_i = T.por_imm.empty() ? 0.0 : T.por_imm[i], pi_j = T.por_imm.empty() ? 0.0 : T.por_imm[j];
        double ph_i = T.phi.empty() ? 1.0 : T.phi[i], ph_j = T.phi.empty() ? 1.0 : T.phi[j];
        if (T.por[i] != T.por[j] || pi_i != pi_j || ph_i != ph_j || T.rs_rock[i] != T.rs_rock[j]) return false;
        for (size_t k = 0; k < ns; k++) {
            size_t a = k * T.E + i, b = k * T.E + j;
            if (T.rs_type[a] != T.rs_type[b] || T.rs_mult[a] != T.rs_mult[b] || T.rs_second[a] != T.rs_second[b])
                return false;
        }
        return true;
    };
    std::vector<int> first(nreg, -1);
    T.rs_tabled.assign(nreg, T.rs_elementwise ? 0 : 1);
    for (int e = 0; e < T.E; e++) {
        int r = m.rdb.regs[m.element[e].region].sub_idx;
        if (first[r] < 0) first[r] = e;
        else if (T.rs_tabled[r] && !same(e, first[r])) T.rs_tabled[r] = 0;
    }
    T.rs_iso.assign((size_t)nreg * ns, Transport::Iso());
    T.rs_el.assign((size_t)T.E * ns, Transport::Iso());
    for (int e = 0; e < T.E; e++) {
        int r = m.rdb.regs[m.element[e].region].sub_idx;
        bool tab = T.rs_tabled[r] != 0;
        if (tab && e != first[r]) continue;
        for (size_t k = 0; k < ns; k++) {
            Transport::Iso I = iso_reinit(T, e, (int)k);
            if (I.scale_sorbed == 0.0) return -1;
            if (tab) {
                if (!iso_table(I, I.type == 0 ? 1 : T.rs_substeps)) return -1;
                T.rs_iso[r * ns + k] = I;
            } else {
                T.rs_el[e * ns + k] = I;
            }
        }
    }
    T.rs_built = true;
    return 0;
}
// Sorption::compute_one_step (sorption.cc:230-285): every element, every species.
// 0, or the error of the first failing element (iso_project's codes).
static int reaction_sorption(Transport &T) {
    Mesh &m = *T.m;
    size_t ns = T.rs_subst.size();
    for (int e = 0; e < T.E; e++) {
        int r = m.rdb.regs[m.element[e].region].sub_idx;
        for (size_t k = 0; k < ns; k++) {
            const Transport::Iso &I = T.rs_tabled[r] ? T.rs_iso[r * ns + k] : T.rs_el[e * ns + k];
            int rc = iso_reaction(I, T.conc[(size_t)T.rs_subst[k] * T.E + e], T.rs_sorbed[k * T.E + e]);
            if (rc) return rc;
        }
    }
    return 0;
}

static void one_step(Transport &T) {
    // data_.set_time at the step's start: boundary data given since the last step apply now;
    // a time-dependent formula changes at every new time (field_formula_impl.hh:70-122)
    bool bc_changed = false;
    if (T.bc_fn) {
        T.bc_pending.assign(T.bc_conc.size(), 0.0);
        T.bc_fn(T.tg.time, T.bc_pending.data(), T.bc_user);
        T.bc_changed = true;
    }
    if (T.bc_changed) {
        T.bc_conc = T.bc_pending;
        T.bc_changed = false;
        bc_changed = true;
    }
    if (T.matrix_built && bc_changed) {  // transport.cc:620-626
        set_bc(T);
        double est = T.tg.estimate_dt();
        for (double &v : T.bcvcorr) v *= est;
    }
    if (!T.matrix_built) {
        build_tm(T);
        T.tg.set_upper_constraint(T.cfl_max_step);
        T.tg.fix_dt_until_mark();
        set_bc(T);
        double est = T.tg.estimate_dt();
        for (double &v : T.bcvcorr) v *= est;
        T.need_rescale = true;
    }
    if (T.need_rescale) {
        if (T.scaled) {
            double fac = T.tg.estimate_dt() / T.tg.time_step;
            for (int e = 0; e < T.E; e++) T.tm.val[T.diag_pos[e]] += -1.0;
            for (double &v : T.tm.val) v *= fac;
            for (int e = 0; e < T.E; e++) T.tm.val[T.diag_pos[e]] += 1.0;
            for (double &v : T.bcvcorr) v *= fac;
        } else {
            double est = T.tg.estimate_dt();
            for (double &v : T.tm.val) v *= est;
            for (int e = 0; e < T.E; e++) T.tm.val[T.diag_pos[e]] += 1.0;
            T.scaled = true;
        }
        T.need_rescale = false;
    }
    T.tg.next_time();
    T.pconc = T.conc;
    for (int k = 0; k < T.nsub; k++) {
        // compute_concentration_sources (transport.cc:501-544) with the dt of
        // this step, then cumulative_corr = bcvcorr + sources_corr (:676)
        const double *cp = &T.pconc[(size_t)k * T.E];
        double *c = &T.conc[(size_t)k * T.E];
        const double *bc = &T.bcvcorr[(size_t)k * T.E];
        double dt = T.tg.time_step;
        for (int i = 0; i < T.E; i++) {
            double src = 0.0;
            if (!T.src_density.empty()) {
                size_t j = (size_t)k * T.E + i;
                double diff = T.src_conc[j] - cp[i];
                if (diff > 0.0) src = (T.src_density[j] + diff * T.src_sigma[j]) * dt;
                else src = T.src_density[j] * dt;
            }
            double s = bc[i] + src;
            for (int p = T.tm.ptr[i]; p < T.tm.ptr[i + 1]; p++) s += T.tm.val[p] * cp[T.tm.col[p]];
            c[i] = s;
        }
        // transport.cc:700-709: dual porosity then sorption, element by element
        if (T.dual || T.sorp)
            for (int i = 0; i < T.E; i++) {
                if (T.dual) dual_porosity(T, i, k, dt);
                if (T.sorp) sorption(T, i, k);
            }
    }
    T.steps++;
}

// set_target_time: transport.cc:718-733
// Linear_reaction::set_time_step -> allocate_reaction_matrix + modify_reaction_matrix
// (linear_reaction.cc:73-95, 144-165, 278-285)
static void set_reaction_time_step(Transport &T, double time_step) {
    int n = T.nsub;
    T.reaction_matrix.assign((size_t)n * n, 0.0);
    for (int rows = 0; rows < n; rows++)
        for (int cols = 0; cols < n; cols++) T.reaction_matrix[(size_t)rows * n + cols] = rows == cols ? 1.0 : 0.0;
    for (size_t i_decay = 0; i_decay < T.r_parent.size(); i_decay++) {
        int index_par = T.r_parent[i_decay];
        double rel_step = time_step / T.r_hl[i_decay];
        T.reaction_matrix[(size_t)index_par * n + index_par] = std::pow(0.5, rel_step);
        for (int k = T.r_pptr[i_decay]; k < T.r_pptr[i_decay + 1]; k++)
            T.reaction_matrix[(size_t)index_par * n + T.r_prod[k]] = (1 - std::pow(0.5, rel_step)) * T.r_branch[k];
    }
}

static double fact(int n) { return n <= 1 ? 1.0 : n * fact(n - 1); }

// Pade_approximant::modify_reaction_matrix (pade_approximant.cc:97-239) with
// evaluate_matrix_polynomial (:241-263): reaction_matrix[rows][cols] = (D^-1 N)[cols][rows]
static void set_pade_time_step(Transport &T, double time_step) {
    const int n = T.nsub, p = T.pade_nom, q = T.pade_den;
    std::vector<double> A((size_t)n * n, 0.0);
    for (size_t d = 0; d < T.r_parent.size(); d++) {
        int par = T.r_parent[d];
        double rel_step = time_step / T.r_hl[d];
        A[(size_t)par * n + par] = -std::log(2.0) * rel_step;
        for (int k = T.r_pptr[d]; k < T.r_pptr[d + 1]; k++)
            A[(size_t)par * n + T.r_prod[k]] = std::log(2.0) * rel_step * T.r_branch[k];
    }
    auto evaluate = [&](int deg, const std::vector<double> &coef) {
        std::vector<double> P((size_t)n * n, 0.0);
        for (int i = deg; i >= 0; i--) {
            std::vector<double> Q((size_t)n * n, 0.0);  // P * A
            for (int r = 0; r < n; r++)
                for (int k = 0; k < n; k++)
                    for (int c = 0; c < n; c++) Q[(size_t)r * n + c] += P[(size_t)r * n + k] * A[(size_t)k * n + c];
            for (int r = 0; r < n; r++) Q[(size_t)r * n + r] += coef[i];
            P = Q;
        }
        return P;
    };
    std::vector<double> nc(p + 1), dc(q + 1);
    for (int j = p; j >= 0; j--) nc[j] = fact(p + q - j) * fact(p) / (fact(p + q) * fact(j) * fact(p - j));
    for (int i = q; i >= 0; i--) dc[i] = std::pow(-1.0, i) * fact(p + q - i) * fact(q) / (fact(p + q) * fact(i) * fact(q - i));
    std::vector<double> Nm = evaluate(p, nc), Dm = evaluate(q, dc);
    // Gauss-Jordan on [D | N]: afterwards the right block is D^-1 N
    std::vector<double> W((size_t)n * 2 * n);
    for (int r = 0; r < n; r++)
        for (int c = 0; c < n; c++) {
            W[(size_t)r * 2 * n + c] = Dm[(size_t)r * n + c];
            W[(size_t)r * 2 * n + n + c] = Nm[(size_t)r * n + c];
        }
    for (int c = 0; c < n; c++) {
        int piv = c;
        for (int r = c + 1; r < n; r++)
            if (std::fabs(W[(size_t)r * 2 * n + c]) > std::fabs(W[(size_t)piv * 2 * n + c])) piv = r;
        for (int k = 0; k < 2 * n; k++) std::swap(W[(size_t)c * 2 * n + k], W[(size_t)piv * 2 * n + k]);
        double d = W[(size_t)c * 2 * n + c];
        for (int k = 0; k < 2 * n; k++) W[(size_t)c * 2 * n + k] /= d;
        for (int r = 0; r < n; r++) {
            if (r == c) continue;
            double f = W[(size_t)r * 2 * n + c];
            for (int k = 0; k < 2 * n; k++) W[(size_t)r * 2 * n + k] -= f * W[(size_t)c * 2 * n + k];
        }
    }
    T.reaction_matrix.assign((size_t)n * n, 0.0);
    for (int rows = 0; rows < n; rows++)
        for (int cols = 0; cols < n; cols++) T.reaction_matrix[(size_t)rows * n + cols] = W[(size_t)cols * 2 * n + n + rows];
}

// Linear_reaction::compute_one_step / compute_reaction (linear_reaction.cc:167-183, 287-300);
// Pade_approximant::compute_reaction (pade_approximant.cc:274-290) multiplies the other way round
// With dual porosity the immobile phase reacts too (linear_reaction.cc:298-300).
static void react_phase(Transport &T, std::vector<double> &conc) {
    int n = T.nsub;
    std::vector<double> prev_conc(n);
    for (int e = 0; e < T.E; e++) {
        for (int cols = 0; cols < n; cols++) {
            prev_conc[cols] = conc[(size_t)cols * T.E + e];
            conc[(size_t)cols * T.E + e] = 0.0;
        }
        if (T.pade_nom >= 0) {
            for (int cols = 0; cols < n; cols++)
                for (int rows = 0; rows < n; rows++)
                    conc[(size_t)cols * T.E + e] += T.reaction_matrix[(size_t)cols * n + rows] * prev_conc[rows];
        } else {
            for (int rows = 0; rows < n; rows++)
                for (int cols = 0; cols < n; cols++)
                    conc[(size_t)rows * T.E + e] += prev_conc[cols] * T.reaction_matrix[(size_t)cols * n + rows];
        }
    }
}
static void compute_reactions(Transport &T) {
    react_phase(T, T.conc);
    if (T.dual) react_phase(T, T.c_imm);
}

static void set_target(Transport &T, double t) {
    T.tg.add_mark(t);
    T.tg.set_upper_constraint(T.cfl_max_step);
    T.tg.fix_dt_until_mark();
    T.need_rescale = true;
}

}  // namespace orc

extern "C" {

// Transport over a steady flow field.  side_flux[S]; cs, por[E];
// bc_conc[nsub][NBC] (per bc element); init_conc[nsub][E].
void *orc_transport_create(void *mesh, const double *side_flux, const double *cs, const double *por, int nsub,
                           const double *bc_conc, const double *init_conc, double end_time, double save_step) {
    Mesh &m = *(Mesh *)mesh;
    Transport *T = new Transport();
    T->m = &m;
    T->E = (int)m.element.size();
    T->nsub = nsub;
    T->flux.assign(side_flux, side_flux + m.n_sides);
    T->cs.assign(cs, cs + T->E);
    T->por.assign(por, por + T->E);
    int NBC = (int)m.bc_elements.size();
    T->bc_conc.assign(bc_conc, bc_conc + (size_t)nsub * NBC);
    T->conc.assign(init_conc, init_conc + (size_t)nsub * T->E);
    T->pconc.assign((size_t)nsub * T->E, 0.0);
    T->bcvcorr.assign((size_t)nsub * T->E, 0.0);
    // TimeGovernor(input) constructor: time_governor.cc:75-117
    T->tg.time = 0.0;
    T->tg.end_time = end_time;
    T->tg.end_fixed = 0.0;
    T->tg.max_ts = end_time;
    T->tg.time_step = end_time;
    T->tg.last_time_step = 0.0;
    T->tg.add_mark(0.0);
    T->tg.add_mark(end_time);
    for (double t = 0.0; t <= end_time * (1 + 1e-12); t += save_step) T->tg.add_mark(t);
    T->cfl_max_step = end_time;
    T->nbr = m.rdb.n_boundary;
    T->nblk = m.rdb.n_bulk;
    T->bcd_bal.assign(nsub * T->nbr, 0); T->bcd_plus = T->bcd_bal; T->bcd_minus = T->bcd_bal;
    T->mass.assign(nsub * T->nblk, 0); T->src = T->mass;
    T->initial_mass.assign(nsub, 0); T->integrated_fluxes = T->initial_mass; T->integrated_sources = T->initial_mass;
    return T;
}
void orc_transport_free(void *h) { delete (Transport *)h; }

// Linear reactions: n decays, parent[n], products in CSR (pptr[n+1], prod[]), half-life[n], branch ratios
void orc_transport_set_reactions(void *h, int n, const int *parent, const int *pptr, const int *prod,
                                 const double *half_life, const double *branch) {
    Transport &T = *(Transport *)h;
    T.r_parent.assign(parent, parent + n);
    T.r_pptr.assign(pptr, pptr + n + 1);
    T.r_prod.assign(prod, prod + pptr[n]);
    T.r_hl.assign(half_life, half_life + n);
    T.r_branch.assign(branch, branch + pptr[n]);
    T.pade_nom = T.pade_den = -1;
}
// The same decays with the Pade approximant of the given degrees.
void orc_transport_set_pade(void *h, int nom_deg, int den_deg) {
    Transport &T = *(Transport *)h;
    T.pade_nom = nom_deg;
    T.pade_den = den_deg;
}

// New flux field (set_velocity_field): the next step rebuilds the transport matrix,
// the step constraint and the boundary vector (transport.cc:605-619).
void orc_transport_set_velocity(void *h, const double *side_flux) {
    Transport &T = *(Transport *)h;
    T.flux.assign(side_flux, side_flux + T.flux.size());
    T.matrix_built = false;
}

// Boundary concentrations bc_conc[nsub][NBC] from the next step on (a time-dependent
// bc_conc field whose new data time has been reached).
void orc_transport_set_bc_conc(void *h, const double *bc_conc) {
    Transport &T = *(Transport *)h;
    T.bc_pending.assign(bc_conc, bc_conc + T.bc_conc.size());
    T.bc_changed = true;
}

// Dual porosity and sorption (ConvectionTransport's dual_porosity / sorption_enable):
// por_imm[E], phi[E]; alpha, sorp_type, sorp_coef0, sorp_coef1 [nsub][E] (null when
// the model that reads them is off).  The other phases start at zero.
void orc_transport_set_phases(void *h, int dual, int sorp, const double *por_imm, const double *alpha,
                              const double *phi, const int *sorp_type, const double *sorp_coef0,
                              const double *sorp_coef1) {
    Transport &T = *(Transport *)h;
    size_t n = (size_t)T.nsub * T.E;
    T.dual = dual != 0;
    T.sorp = sorp != 0;
    T.por_imm.assign(por_imm, por_imm + T.E);
    T.alpha.assign(n, 0.0);
    if (alpha) T.alpha.assign(alpha, alpha + n);
    T.phi.assign(T.E, 1.0);
    if (phi) T.phi.assign(phi, phi + T.E);
    T.sorp_type.assign(n, 0);
    T.sorp_c0.assign(n, 0.0);
    T.sorp_c1.assign(n, 0.0);
    if (sorp_type) T.sorp_type.assign(sorp_type, sorp_type + n);
    if (sorp_coef0) T.sorp_c0.assign(sorp_coef0, sorp_coef0 + n);
    if (sorp_coef1) T.sorp_c1.assign(sorp_coef1, sorp_coef1 + n);
    T.c_imm.assign(n, 0.0);
    T.c_msorb.assign(n, 0.0);
    T.c_isorb.assign(n, 0.0);
}
// Concentrations of one phase (0 mobile, 1 immobile, 2 mobile sorbed, 3 immobile
// sorbed; transport.h:58-63) [nsub][E].
void orc_transport_conc_phase(void *h, int phase, double *out) {
    Transport &T = *(Transport *)h;
    const std::vector<double> *v[4] = {&T.conc, &T.c_imm, &T.c_msorb, &T.c_isorb};
    const std::vector<double> &c = *v[phase];
    if (c.empty()) std::memset(out, 0, (size_t)T.nsub * T.E * sizeof(double));
    else std::memcpy(out, c.data(), c.size() * sizeof(double));
}

// Reaction-module sorption (Sorption, transport_operator_splitting.cc:178-192, 244-287):
// n species on substances subst[n]; the isotherm type, mult_coef, second_param per bulk
// region (sub-index order) and species [nreg][n] with rock_density[nreg], or
// (elementwise) per species and element [n][E] with rock_density[E]; molar_mass[n],
// solubility[n] (<= 0 or null: no limit, the reference's default -1).  The isotherms
// are made at the first update, with the por_imm and phi of orc_transport_set_phases.
static void set_sorption(Transport &T, int n, const int *subst, const int *type, const double *mult,
                         const double *second, const double *rock_density, const double *molar_mass,
                         const double *solubility, double solvent_dens, int substeps, bool elementwise) {
    Mesh &m = *T.m;
    T.rs_subst.assign(subst, subst + n);
    size_t ne = (size_t)n * T.E;
    T.rs_type.assign(ne, 0);
    T.rs_mult.assign(ne, 0.0);
    T.rs_second.assign(ne, 0.0);
    T.rs_rock.assign(T.E, 0.0);
    for (int e = 0; e < T.E; e++) {
        int r = m.rdb.regs[m.element[e].region].sub_idx;
        T.rs_rock[e] = elementwise ? rock_density[e] : rock_density[r];
        for (int k = 0; k < n; k++) {
            size_t src = elementwise ? (size_t)k * T.E + e : (size_t)r * n + k;
            T.rs_type[(size_t)k * T.E + e] = type[src];
            T.rs_mult[(size_t)k * T.E + e] = mult[src];
            T.rs_second[(size_t)k * T.E + e] = second[src];
        }
    }
    T.rs_molar.assign(molar_mass, molar_mass + n);
    T.rs_sol.assign(n, -1.0);
    if (solubility) T.rs_sol.assign(solubility, solubility + n);
    T.rs_solvent = solvent_dens;
    T.rs_substeps = substeps;
    T.rs_elementwise = elementwise;
    T.rs_built = false;
    T.rs_sorbed.assign((size_t)n * T.E, 0.0);
}
int orc_transport_set_reaction_sorption(void *h, int n, const int *subst, const int *type, const double *mult,
                                        const double *second, const double *rock_density, const double *molar_mass,
                                        const double *solubility, double solvent_dens, int substeps) {
    set_sorption(*(Transport *)h, n, subst, type, mult, second, rock_density, molar_mass, solubility, solvent_dens,
                 substeps, false);
    return 0;
}
int orc_transport_set_reaction_sorption_el(void *h, int n, const int *subst, const int *type, const double *mult,
                                           const double *second, const double *rock_density,
                                           const double *molar_mass, const double *solubility,
                                           double solvent_dens) {
    set_sorption(*(Transport *)h, n, subst, type, mult, second, rock_density, molar_mass, solubility, solvent_dens,
                 1, true);
    return 0;
}
void orc_transport_reaction_sorbed(void *h, double *out) {
    Transport &T = *(Transport *)h;
    std::memcpy(out, T.rs_sorbed.data(), T.rs_sorbed.size() * sizeof(double));
}

// Boundary concentrations as a function of time: fn(t, bc_conc[nsub][NBC], user) at the
// start of every step (fn null: back to fixed data).
void orc_transport_set_bc_function(void *h, void (*fn)(double, double *, void *), void *user) {
    Transport &T = *(Transport *)h;
    T.bc_fn = fn;
    T.bc_user = user;
}

void orc_transport_set_sources(void *h, const double *density, const double *sigma, const double *conc) {
    Transport &T = *(Transport *)h;
    size_t n = (size_t)T.nsub * T.E;
    T.src_density.assign(density, density + n);
    T.src_sigma.assign(sigma, sigma + n);
    T.src_conc.assign(conc, conc + n);
}

// Advance to target time t (TransportOperatorSplitting::update_solution,
// transport_operator_splitting.cc:244-287); returns steps taken, or -1 when the
// sorption isotherms cannot be made (a zero scale_sorbed, a table without a root),
// -2 when a sorption step fails (negative total mass, no root in the bracket).
int orc_transport_update(void *h, double t) {
    Transport &T = *(Transport *)h;
    if (T.mb_initial) mb_calculate(T, T.tg.time);  // first_time_call output at t=0
    set_target(T, t);
    if (!T.r_parent.empty()) {
        if (T.pade_nom >= 0) set_pade_time_step(T, T.tg.estimate_dt());
        else set_reaction_time_step(T, T.tg.estimate_dt());
    }
    int steps = 0;
    double t0 = now_s();
    if (!T.rs_subst.empty() && !T.rs_built && make_sorption(T)) return -1;
    while (T.tg.lt(t)) {
        one_step(T);
        if (!T.r_parent.empty()) compute_reactions(T);
        if (!T.rs_subst.empty() && reaction_sorption(T)) return -2;
        mb_calculate(T, T.tg.time);
        steps++;
    }
    T.t_steps += now_s() - t0;
    return steps;
}
double orc_transport_time(void *h) { return ((Transport *)h)->tg.time; }
double orc_transport_dt(void *h) { return ((Transport *)h)->tg.time_step; }
double orc_transport_cfl(void *h) { return ((Transport *)h)->cfl_max_step; }
void orc_transport_conc(void *h, double *out) {
    Transport &T = *(Transport *)h;
    std::memcpy(out, T.conc.data(), T.conc.size() * sizeof(double));
}
// Scaled transport matrix tm = I + dt*A (current): ptr[E+1], col, val
int orc_transport_matrix(void *h, int *ptr, int *col, double *val) {
    Transport &T = *(Transport *)h;
    if (ptr) {
        std::memcpy(ptr, T.tm.ptr.data(), T.tm.ptr.size() * sizeof(int));
        std::memcpy(col, T.tm.col.data(), T.tm.col.size() * sizeof(int));
        std::memcpy(val, T.tm.val.data(), T.tm.val.size() * sizeof(double));
    }
    return (int)T.tm.col.size();
}
// Mass balance (last calculate): bcd_bal/plus/minus [nsub][nbr], mass [nsub][nblk],
// totals[4*nsub]: initial mass, integrated fluxes, integrated sources, current mass
void orc_transport_mass_balance(void *h, double *bal, double *plus, double *minus, double *mass, double *src,
                                double *totals) {
    Transport &T = *(Transport *)h;
    std::memcpy(src, T.src.data(), T.src.size() * sizeof(double));
    std::memcpy(bal, T.bcd_bal.data(), T.bcd_bal.size() * sizeof(double));
    std::memcpy(plus, T.bcd_plus.data(), T.bcd_plus.size() * sizeof(double));
    std::memcpy(minus, T.bcd_minus.data(), T.bcd_minus.size() * sizeof(double));
    std::memcpy(mass, T.mass.data(), T.mass.size() * sizeof(double));
    for (int k = 0; k < T.nsub; k++) {
        double mt = 0;
        for (int r = 0; r < T.nblk; r++) mt += T.mass[k * T.nblk + r];
        totals[4 * k] = T.initial_mass[k];
        totals[4 * k + 1] = T.integrated_fluxes[k];
        totals[4 * k + 2] = T.integrated_sources[k];
        totals[4 * k + 3] = mt;
    }
}
double orc_transport_step_seconds(void *h) { return ((Transport *)h)->t_steps; }

}  // extern "C"
