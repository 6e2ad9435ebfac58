// Text balance files in the reference's layouts:
//  * water balance of the steady flow (DarcyFlowMHOutput::water_balance,
//    darcy_flow_mh_output.cc:671-773): per boundary region total, outflow and
//    inflow of the side fluxes, per bulk region the volume sources;
//  * mass balance of the transport (MassBalance::output, mass_balance.cc:242-354):
//    per boundary region and substance the mass flux (total, outward, inward),
//    per bulk region mass and source, and the cumulative balance since the
//    first output.
// Regions are listed as RegionDB's "BOUNDARY" / "BULK" sets list them: by
// boundary index and by bulk index (region.cc:224-239).  Each call appends
// one block, as the reference writes one per output time into one open file.
#pragma once
#include <string>
#include <vector>

#include "mesh.hh"

namespace fl {

struct RegionRow {
    int id;
    std::string label;
};

void region_sets(const HostMesh &m, std::vector<RegionRow> &boundary, std::vector<RegionRow> &bulk);

// bal/plus/minus[n_boundary_regions], src[n_bulk_regions]
std::string write_water_balance(const char *path, double time, const std::vector<RegionRow> &bnd, const double *bal,
                                const double *plus, const double *minus, const std::vector<RegionRow> &blk,
                                const double *src);

// bal/plus/minus[ns][n_boundary_regions], mass/src[ns][n_bulk_regions],
// init_mass / int_sources / int_fluxes[ns]
std::string write_mass_balance(const char *path, double time, double initial_time, int ns,
                               const std::vector<std::string> &names, const std::vector<RegionRow> &bnd,
                               const double *bal, const double *plus, const double *minus,
                               const std::vector<RegionRow> &blk, const double *mass, const double *src,
                               const double *init_mass, const double *int_sources, const double *int_fluxes);

}  // namespace fl
