// Legacy boundary-condition files (.fbc flow, .tbc transport) of Flow123d 1.x,
// read onto the boundary elements of a HostMesh.  See old_bcd.hh.
#include "old_bcd.hh"

#include <cstdlib>

#include "common.hh"

namespace fl {

static unsigned parse_uint(MshTokenizer &tok) {
    const std::string &s = tok.cur();
    char *e = nullptr;
    unsigned long v = std::strtoul(s.c_str(), &e, 10);
    if (s.empty() || s[0] == '-' || *e != '\0')
        throw Error(-1, "Wrong format of number, file '" + tok.file() + "', line " + std::to_string(tok.line_num()) +
                            ".");
    tok.advance();
    return (unsigned)v;
}
static double parse_real(MshTokenizer &tok) {
    const std::string &s = tok.cur();
    char *e = nullptr;
    double v = std::strtod(s.c_str(), &e);
    if (s.empty() || *e != '\0')
        throw Error(-1, "Wrong format of number, file '" + tok.file() + "', line " + std::to_string(tok.line_num()) +
                            ".");
    tok.advance();
    return v;
}

std::string OldBcd::read_flow(const HostMesh &m, const char *path, int *type, double *pressure, double *flux,
                              double *sigma) {
    const int NB = m.n_bc_elements();
    for (int b = 0; b < NB; b++) type[b] = 0, pressure[b] = flux[b] = sigma[b] = 0.0;  // set_all
    id_to_bc.clear();
    std::unordered_map<int, int> elem_of_id;
    for (int e = 0; e < m.n_elements(); e++) elem_of_id[m.el_id[e]] = e;
    MshTokenizer tok;
    std::string err = tok.open(path);
    if (!err.empty()) return err;
    try {
        tok.skip_to("$BoundaryConditions");
        tok.next_line(false);
        unsigned n = parse_uint(tok);
        for (unsigned i = 0; i < n; i++) {
            tok.next_line(true);
            unsigned id = parse_uint(tok);
            unsigned t = parse_uint(tok);
            double scalar = 0.0, fl = 0.0, sg = 0.0;
            switch (t) {
                case 1: scalar = parse_real(tok); break;  // Dirichlet: pressure
                case 2: fl = parse_real(tok); break;      // Neumann: flux
                case 3:                                   // Newton: pressure and sigma
                    scalar = parse_real(tok);
                    sg = parse_real(tok);
                    break;
                default:
                    return "Unknown type of boundary condition - cond # " + std::to_string(id) + ", type " +
                           std::to_string(t);
            }
            unsigned where = parse_uint(tok);
            if (where == 3) return "Element only BCD are not supported.";
            if (where != 2)
                return "Unknown entity for boundary condition - cond # " + std::to_string(id) + ", ent. " +
                       std::to_string(where);
            unsigned eid = parse_uint(tok), sid = parse_uint(tok);
            auto it = elem_of_id.find((int)eid);
            if (it == elem_of_id.end())
                return "Boundary " + std::to_string(id) + " refers to unknown element ID " + std::to_string(eid);
            int e = it->second;
            if ((int)sid >= m.el_dim[e] + 1)
                return "Boundary " + std::to_string(id) + " has incorrect reference to side " + std::to_string(sid);
            // the file numbers sides as RefElement does (old_bcd.cc:54-64: identity maps)
            int bnd = m.side_bnd[m.side_off[e] + (int)sid];
            if (bnd < 0)
                return "Setting boundary condition " + std::to_string(id) + " for non-boundary side " +
                       std::to_string(sid) + " of element ID: " + std::to_string(eid);
            int b = m.bnd_bcele[bnd];
            id_to_bc[id] = b;
            type[b] = (int)t;
            pressure[b] = scalar;
            flux[b] = fl;
            sigma[b] = sg;
            unsigned ntags = parse_uint(tok);
            for (; ntags > 0; ntags--) tok.advance();
        }
    } catch (const Error &e) {
        return e.what();
    }
    return "";
}

std::string OldBcd::read_transport(const HostMesh &m, const char *path, int n_subst, double *conc) const {
    const int NB = m.n_bc_elements();
    for (long k = 0; k < (long)n_subst * NB; k++) conc[k] = 0.0;  // set_all
    MshTokenizer tok;
    std::string err = tok.open(path);
    if (!err.empty()) return err;
    try {
        if (tok.skip_to("$Transport_BCDFormat")) tok.next_line(false);
        tok.skip_to("$Transport_BCD");
        tok.next_line(false);
        unsigned n = parse_uint(tok);
        for (unsigned i = 0; i < n; i++) {
            tok.next_line(true);
            unsigned bcd_id = parse_uint(tok);
            unsigned boundary_id = parse_uint(tok);
            auto it = id_to_bc.find(bcd_id);
            if (it == id_to_bc.end())
                return "Wrong boundary index " + std::to_string(boundary_id) + " for bcd id " +
                       std::to_string(bcd_id) + " in transport bcd file!";
            for (int s = 0; s < n_subst; s++) conc[(size_t)s * NB + it->second] = parse_real(tok);
        }
    } catch (const Error &e) {
        return e.what();
    }
    return "";
}

}  // namespace fl
