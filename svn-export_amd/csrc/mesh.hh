// Host-side mesh, topology and MH dof numbering for the GPU Darcy / transport path.
//
// The numbering contract is the reference's (bit-exact):
//  * bulk elements in file order, boundary elements ('.'-prefixed regions)
//    in file order followed by the implicit boundary elements
//    (msh_gmshreader.cc:186-195, mesh.cc:437-463);
//  * edges (MH Lagrange multipliers) created first for file boundary
//    elements, then per element and local side in RefElement order, with a
//    private 1-side edge plus a vb neighbour for every side that touches a
//    lower-dimensional element (mesh.cc:343-500, ref_element.cc:51-74);
//  * sides numbered element by element (mh_dofhandler.cc:12-21) and the MH
//    solution laid out as [side fluxes | element pressures | edge traces]
//    (darcy_flow_mh.cc:330-366).
// The construction itself is not the reference's node-element-list
// intersection: sides are grouped by sorting their node keys, which gives the
// same groups for conforming simplicial meshes in O(n log n).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace fl {

struct RegionInfo {
    int id;
    std::string label;
    bool boundary;
    int sub_idx;  // index among boundary regions or among bulk regions
};

struct HostMesh {
    // nodes
    std::vector<double> xyz;  // 3 * n_nodes
    // bulk elements (file order)
    std::vector<int> el_id, el_dim, el_region;  // el_region: index into regions
    std::vector<int> el_nodes;                  // 4 * E, -1 padded
    std::vector<int> el_pid;                    // partition tag (third gmsh tag, 0 when absent)
    // boundary elements (file ones first, implicit ones appended)
    std::vector<int> bc_id, bc_dim, bc_region, bc_nodes;  // bc_nodes: 4 * NBC
    int n_file_bc = 0;
    std::vector<RegionInfo> regions;
    int n_boundary_regions = 0, n_bulk_regions = 0;

    // topology
    std::vector<int> side_off;    // E+1, MH_DofHandler side numbering
    std::vector<int> side_edge;   // S
    std::vector<int> side_bnd;    // S, boundary index or -1
    std::vector<int> side_elem;   // S
    std::vector<int> edge_ptr;    // F+1
    std::vector<int> edge_side;   // S (global side dofs of each edge, in the reference's order)
    std::vector<int> vb_elem;     // V, lower-dimensional element
    std::vector<int> vb_edge;     // V
    std::vector<int> bnd_edge;    // B
    std::vector<int> bnd_bcele;   // B
    std::vector<int> el_vb_ptr;   // E+1, vb neighbours of each (lower-dim) element
    std::vector<int> el_vb;       // V, vb indices grouped by element (element_to_neigh_vb order)

    int n_nodes() const { return (int)xyz.size() / 3; }
    int n_elements() const { return (int)el_dim.size(); }
    int n_sides() const { return side_off.empty() ? 0 : side_off.back(); }
    int n_edges() const { return (int)edge_ptr.size() - 1; }
    int n_vb() const { return (int)vb_elem.size(); }
    int n_boundaries() const { return (int)bnd_edge.size(); }
    int n_bc_elements() const { return (int)bc_dim.size(); }
};

// RegionDB bookkeeping (region.cc:105-171): '.'-prefixed labels are boundary
// regions, boundary and bulk regions are counted separately in creation order,
// unnamed ids get "region_<id>", IMPLICIT BOUNDARY (id -3) is added on first use.
int region_add(HostMesh &m, int id, const std::string &label, bool boundary);
int region_add_unnamed(HostMesh &m, int id);

// gmsh 2.x ASCII (msh_gmshreader.cc:76-246).  Returns "" or an error message.
std::string read_gmsh(const char *path, HostMesh &m);

// Build from arrays: element list in file order with gmsh types (1,2,4,15),
// region ids and 4 node indices (0-based, -1 padded); physical names optional.
std::string mesh_from_arrays(HostMesh &m, int n_nodes, const double *xyz, int n_elem, const int *elem_id,
                             const int *gmsh_type, const int *region_id, const int *nodes4, int n_phys,
                             const int *phys_id, const char *const *phys_name);

// Topology + dof numbering (see header comment).  Returns "" or an error message.
std::string build_topology(HostMesh &m);

// Element / side geometry on the host (elements.cc:108-151, sides.cc:42-146);
// used for mesh set-up quantities only (boundary measures, vb side measures).
double element_measure(const HostMesh &m, int dim, const int *nodes);
double side_measure(const HostMesh &m, int e, int s);

}  // namespace fl
