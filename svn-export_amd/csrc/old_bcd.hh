// Legacy boundary-condition files of Flow123d 1.x (OldBcdInput, old_bcd.cc:44-250).
//
// Flow file: `$BoundaryConditions`, a count, then rows
//   id type values.. where element_id side n_tags tags..
// with type 1 Dirichlet (pressure), 2 Neumann (flux), 3 Newton (pressure,
// sigma) and where = 2 (element and local side; the side must lie on the
// boundary).  Each row sets the four boundary fields of that side's boundary
// element; elements without a row keep zero (type 0, no condition).
// Transport file: `$Transport_BCD`, a count, then rows
//   flow_row_id boundary_id c_1 .. c_n
// that give the boundary concentrations of the element the flow file's row
// `flow_row_id` named.  Tokens follow the gmsh reader's tokenizer rules.
#pragma once
#include <string>
#include <unordered_map>

#include "msh_data.hh"

namespace fl {

struct OldBcd {
    std::unordered_map<unsigned, int> id_to_bc;  // flow row id -> boundary element index

    // type/pressure/flux/sigma[NBC]; returns "" or an error message
    std::string read_flow(const HostMesh &m, const char *path, int *type, double *pressure, double *flux,
                          double *sigma);
    // conc[n_subst][NBC] (needs read_flow first); returns "" or an error message
    std::string read_transport(const HostMesh &m, const char *path, int n_subst, double *conc) const;
};

}  // namespace fl
