// Host mesh reading, topology and numbering.  See mesh.hh for the contract.
#include "mesh.hh"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <unordered_map>

namespace fl {

// RefElement<dim>::side_nodes (ref_element.cc:51-74)
static const int kSideNodes[4][4][3] = {
    {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}},
    {{0, 0, 0}, {1, 0, 0}, {0, 0, 0}, {0, 0, 0}},
    {{0, 1, 0}, {1, 2, 0}, {0, 2, 0}, {0, 0, 0}},
    {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}},
};

int region_add(HostMesh &m, int id, const std::string &label, bool boundary) {
    for (size_t r = 0; r < m.regions.size(); r++)
        if (m.regions[r].id == id) return (int)r;
    RegionInfo ri{id, label, boundary, boundary ? m.n_boundary_regions++ : m.n_bulk_regions++};
    m.regions.push_back(ri);
    return (int)m.regions.size() - 1;
}
int region_add_unnamed(HostMesh &m, int id) {
    for (size_t r = 0; r < m.regions.size(); r++)
        if (m.regions[r].id == id) return (int)r;
    return region_add(m, id, "region_" + std::to_string(id), false);
}

static void add_element(HostMesh &m, int id, int dim, int region_id, const int *nodes, int pid = 0) {
    int r = region_add_unnamed(m, region_id);
    if (m.regions[r].boundary) {
        m.bc_id.push_back(id);
        m.bc_dim.push_back(dim);
        m.bc_region.push_back(r);
        for (int k = 0; k < 4; k++) m.bc_nodes.push_back(k <= dim ? nodes[k] : -1);
    } else if (dim > 0) {
        m.el_id.push_back(id);
        m.el_dim.push_back(dim);
        m.el_region.push_back(r);
        m.el_pid.push_back(pid);
        for (int k = 0; k < 4; k++) m.el_nodes.push_back(k <= dim ? nodes[k] : -1);
    }
}

static int gmsh_dim(int type) {
    switch (type) {
        case 1: return 1;
        case 2: return 2;
        case 4: return 3;
        case 15: return 0;
    }
    return -1;
}

std::string read_gmsh(const char *path, HostMesh &m) {
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return std::string("cannot open mesh file ") + path;
    std::fseek(fp, 0, SEEK_END);
    long len = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    std::string buf((size_t)len, '\0');
    size_t got = std::fread(&buf[0], 1, (size_t)len, fp);
    std::fclose(fp);
    buf.resize(got);
    const char *p = buf.c_str();
    const char *end = p + buf.size();
    auto find = [&](const char *tag, const char *from) -> const char * {
        const char *q = std::strstr(from, tag);
        if (!q) return nullptr;
        q = std::strchr(q, '\n');
        return q ? q + 1 : end;
    };
    const char *nodes_at = find("$Nodes", p);
    if (!nodes_at) return "missing section $Nodes";
    const char *pn = find("$PhysicalNames", p);
    if (pn && pn < nodes_at) {
        char *q;
        long n = std::strtol(pn, &q, 10);
        for (long i = 0; i < n; i++) {
            (void)std::strtol(q, &q, 10);  // dim
            long id = std::strtol(q, &q, 10);
            while (*q == ' ' || *q == '\t') q++;
            std::string name;
            if (*q == '"') {
                const char *e = std::strchr(q + 1, '"');
                name.assign(q + 1, (size_t)(e - (q + 1)));
                q = (char *)e + 1;
            } else {
                const char *e = q;
                while (*e && *e != '\n' && *e != ' ') e++;
                name.assign(q, (size_t)(e - q));
                q = (char *)e;
            }
            region_add(m, (int)id, name, !name.empty() && name[0] == '.');
        }
    }
    char *q;
    long nn = std::strtol(nodes_at, &q, 10);
    if (nn <= 0) return "zero number of nodes";
    m.xyz.resize(3 * (size_t)nn);
    std::unordered_map<long, int> node_index;
    node_index.reserve((size_t)nn * 2);
    for (long i = 0; i < nn; i++) {
        long id = std::strtol(q, &q, 10);
        m.xyz[3 * i] = std::strtod(q, &q);
        m.xyz[3 * i + 1] = std::strtod(q, &q);
        m.xyz[3 * i + 2] = std::strtod(q, &q);
        node_index[id] = (int)i;
    }
    const char *el_at = find("$Elements", q);
    if (!el_at) return "missing section $Elements";
    long ne = std::strtol(el_at, &q, 10);
    for (long i = 0; i < ne; i++) {
        int id = (int)std::strtol(q, &q, 10);
        int type = (int)std::strtol(q, &q, 10);
        int ntags = (int)std::strtol(q, &q, 10);
        int dim = gmsh_dim(type);
        if (dim < 0) return "unsupported element type " + std::to_string(type);
        if (ntags < 2) return "at least two element tags are required";
        int region = 0, pid = 0;
        for (int t = 0; t < ntags; t++) {
            long v = std::strtol(q, &q, 10);
            if (t == 0) region = (int)v;
            if (t == 2) pid = (int)v;  // partition number (msh_gmshreader.cc:174-175)
        }
        int nodes[4] = {-1, -1, -1, -1};
        for (int k = 0; k <= dim; k++) {
            long nid = std::strtol(q, &q, 10);
            auto it = node_index.find(nid);
            if (it == node_index.end()) return "unknown node id in element " + std::to_string(id);
            nodes[k] = it->second;
        }
        add_element(m, id, dim, region, nodes, pid);
    }
    m.n_file_bc = (int)m.bc_dim.size();
    return "";
}

std::string mesh_from_arrays(HostMesh &m, int n_nodes, const double *xyz, int n_elem, const int *elem_id,
                             const int *gmsh_type, const int *region_id, const int *nodes4, int n_phys,
                             const int *phys_id, const char *const *phys_name) {
    for (int i = 0; i < n_phys; i++) {
        std::string name = phys_name[i] ? phys_name[i] : "";
        region_add(m, phys_id[i], name, !name.empty() && name[0] == '.');
    }
    m.xyz.assign(xyz, xyz + 3 * (size_t)n_nodes);
    for (int i = 0; i < n_elem; i++) {
        int dim = gmsh_dim(gmsh_type[i]);
        if (dim < 0) return "unsupported element type " + std::to_string(gmsh_type[i]);
        for (int k = 0; k <= dim; k++)
            if (nodes4[4 * i + k] < 0 || nodes4[4 * i + k] >= n_nodes) return "node index out of range";
        add_element(m, elem_id ? elem_id[i] : i + 1, dim, region_id[i], nodes4 + 4 * (size_t)i);
    }
    m.n_file_bc = (int)m.bc_dim.size();
    return "";
}

namespace {
struct Key {
    uint64_t hi, lo;
    bool operator<(const Key &o) const { return hi != o.hi ? hi < o.hi : lo < o.lo; }
    bool operator==(const Key &o) const { return hi == o.hi && lo == o.lo; }
};
Key make_key(int k, const int *nodes) {
    uint32_t v[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu};
    for (int i = 0; i < k; i++) v[i] = (uint32_t)nodes[i];
    std::sort(v, v + k);
    return Key{((uint64_t)k << 32) | v[0], ((uint64_t)v[1] << 32) | v[2]};
}
}  // namespace

std::string build_topology(HostMesh &m) {
    const int E = m.n_elements();
    m.side_off.assign(E + 1, 0);
    for (int e = 0; e < E; e++) m.side_off[e + 1] = m.side_off[e] + m.el_dim[e] + 1;
    const int S = m.side_off[E];
    m.side_elem.resize(S);
    for (int e = 0; e < E; e++)
        for (int s = 0; s <= m.el_dim[e]; s++) m.side_elem[m.side_off[e] + s] = e;

    // group all sides by node key, members ordered by element then side
    struct Rec {
        Key key;
        int side;  // global side dof (encodes element and local side)
    };
    std::vector<Rec> recs(S);
    for (int e = 0; e < E; e++) {
        int d = m.el_dim[e];
        for (int s = 0; s <= d; s++) {
            int nodes[3];
            for (int k = 0; k < d; k++) nodes[k] = m.el_nodes[4 * e + kSideNodes[d][s][k]];
            recs[m.side_off[e] + s] = Rec{make_key(d, nodes), m.side_off[e] + s};
        }
    }
    std::sort(recs.begin(), recs.end(), [](const Rec &a, const Rec &b) {
        return a.key < b.key || (a.key == b.key && a.side < b.side);
    });
    std::vector<Key> gkey;
    std::vector<int> gptr, side_group(S);
    for (int i = 0; i < S; i++) {
        if (i == 0 || !(recs[i].key == recs[i - 1].key)) {
            gkey.push_back(recs[i].key);
            gptr.push_back(i);
        }
        side_group[recs[i].side] = (int)gkey.size() - 1;
    }
    gptr.push_back(S);
    const int G = (int)gkey.size();
    auto find_group = [&](const Key &k) -> int {
        auto it = std::lower_bound(gkey.begin(), gkey.end(), k);
        return (it != gkey.end() && *it == k) ? (int)(it - gkey.begin()) : -1;
    };
    // lower-dimensional element whose node set equals a side group key
    std::vector<int> group_lower(G, -1);
    for (int e = 0; e < E; e++) {
        int d = m.el_dim[e];
        Key k = make_key(d + 1, &m.el_nodes[4 * e]);
        int g = find_group(k);
        if (g < 0) continue;
        if (group_lower[g] >= 0)
            return "too many matching lower-dimensional elements: ids " + std::to_string(m.el_id[group_lower[g]]) +
                   " and " + std::to_string(m.el_id[e]);
        group_lower[g] = e;
    }

    m.side_edge.assign(S, -1);
    m.side_bnd.assign(S, -1);
    std::vector<std::vector<int>> edges;
    edges.reserve((size_t)S);
    m.bnd_edge.clear();
    m.bnd_bcele.clear();
    m.vb_elem.clear();
    m.vb_edge.clear();

    // (1) file boundary elements (mesh.cc:358-409)
    for (int b = 0; b < m.n_file_bc; b++) {
        int d = m.bc_dim[b];
        Key k = make_key(d + 1, &m.bc_nodes[4 * b]);
        int g = find_group(k);
        if (g >= 0 && group_lower[g] >= 0)
            return "boundary element id " + std::to_string(m.bc_id[b]) + " matches a lower-dimensional bulk element";
        if (g < 0) continue;  // lonely boundary element: skipped
        int f = (int)edges.size();
        edges.emplace_back();
        int bdr = (int)m.bnd_edge.size();
        m.bnd_edge.push_back(f);
        m.bnd_bcele.push_back(b);
        for (int i = gptr[g]; i < gptr[g + 1]; i++) {
            int sd = recs[i].side;
            edges[f].push_back(sd);
            m.side_edge[sd] = f;
            m.side_bnd[sd] = bdr;
        }
    }
    // (2) element sides in element order (mesh.cc:412-497)
    int implicit_region = -1;
    for (int e = 0; e < E; e++) {
        int d = m.el_dim[e];
        for (int s = 0; s <= d; s++) {
            int sd = m.side_off[e] + s;
            if (m.side_edge[sd] != -1) continue;
            int g = side_group[sd];
            int lower = group_lower[g];
            int gsize = gptr[g + 1] - gptr[g];
            if (lower >= 0) {
                for (int i = gptr[g]; i < gptr[g + 1]; i++) {
                    int sx = recs[i].side;
                    if (m.side_edge[sx] != -1) continue;
                    int f = (int)edges.size();
                    edges.emplace_back(1, sx);
                    m.side_edge[sx] = f;
                    m.vb_elem.push_back(lower);
                    m.vb_edge.push_back(f);
                }
                continue;
            }
            int f = (int)edges.size();
            edges.emplace_back();
            if (gsize == 1) {
                edges[f].push_back(sd);
                m.side_edge[sd] = f;
                int bdr = (int)m.bnd_edge.size();
                m.side_bnd[sd] = bdr;
                if (implicit_region < 0) implicit_region = region_add(m, -3, "IMPLICIT BOUNDARY", true);
                m.bc_id.push_back(-bdr);
                m.bc_dim.push_back(d - 1);
                m.bc_region.push_back(implicit_region);
                for (int k = 0; k < 4; k++)
                    m.bc_nodes.push_back(k < d ? m.el_nodes[4 * e + kSideNodes[d][s][k]] : -1);
                m.bnd_edge.push_back(f);
                m.bnd_bcele.push_back(m.n_bc_elements() - 1);
                continue;
            }
            for (int i = gptr[g]; i < gptr[g + 1]; i++) {
                int sx = recs[i].side;
                if (m.side_edge[sx] != -1) continue;
                edges[f].push_back(sx);
                m.side_edge[sx] = f;
            }
        }
    }
    const int F = (int)edges.size();
    m.edge_ptr.assign(F + 1, 0);
    for (int f = 0; f < F; f++) m.edge_ptr[f + 1] = m.edge_ptr[f] + (int)edges[f].size();
    m.edge_side.resize(m.edge_ptr[F]);
    for (int f = 0; f < F; f++) std::copy(edges[f].begin(), edges[f].end(), m.edge_side.begin() + m.edge_ptr[f]);
    // element_to_neigh_vb (mesh.cc:660-685)
    const int V = m.n_vb();
    m.el_vb_ptr.assign(E + 1, 0);
    for (int n = 0; n < V; n++) m.el_vb_ptr[m.vb_elem[n] + 1]++;
    for (int e = 0; e < E; e++) m.el_vb_ptr[e + 1] += m.el_vb_ptr[e];
    m.el_vb.resize(V);
    std::vector<int> fill(m.el_vb_ptr.begin(), m.el_vb_ptr.end() - 1);
    for (int n = 0; n < V; n++) m.el_vb[fill[m.vb_elem[n]]++] = n;
    return "";
}

static inline void vsub(const double *a, const double *b, double *r) {
    r[0] = a[0] - b[0]; r[1] = a[1] - b[1]; r[2] = a[2] - b[2];
}
static inline void vcross(const double *u, const double *v, double *x) {
    x[0] = u[1] * v[2] - u[2] * v[1];
    x[1] = u[2] * v[0] - u[0] * v[2];
    x[2] = u[0] * v[1] - u[1] * v[0];
}
static inline double vnorm(const double *u) { return std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]); }

double element_measure(const HostMesh &m, int dim, const int *nodes) {
    const double *X = m.xyz.data();
    double a[3], b[3], c[3], d[3];
    switch (dim) {
        case 1: vsub(X + 3 * nodes[1], X + 3 * nodes[0], a); return vnorm(a);
        case 2:
            vsub(X + 3 * nodes[1], X + 3 * nodes[0], a);
            vsub(X + 3 * nodes[2], X + 3 * nodes[0], b);
            vcross(a, b, c);
            return vnorm(c) / 2.0;
        case 3:
            vsub(X + 3 * nodes[1], X + 3 * nodes[0], a);
            vsub(X + 3 * nodes[2], X + 3 * nodes[0], b);
            vcross(a, b, c);
            vsub(X + 3 * nodes[3], X + 3 * nodes[0], d);
            return std::fabs(c[0] * d[0] + c[1] * d[1] + c[2] * d[2]) / 6.0;
    }
    return 1.0;
}

double side_measure(const HostMesh &m, int e, int s) {
    int d = m.el_dim[e];
    const double *X = m.xyz.data();
    const int *n = &m.el_nodes[4 * e];
    if (d == 1) return 1.0;
    double a[3], b[3], c[3];
    if (d == 2) {
        vsub(X + 3 * n[kSideNodes[2][s][1]], X + 3 * n[kSideNodes[2][s][0]], a);
        return vnorm(a);
    }
    vsub(X + 3 * n[kSideNodes[3][s][1]], X + 3 * n[kSideNodes[3][s][0]], a);
    vsub(X + 3 * n[kSideNodes[3][s][2]], X + 3 * n[kSideNodes[3][s][0]], b);
    vcross(a, b, c);
    return 0.5 * vnorm(c);
}

}  // namespace fl
