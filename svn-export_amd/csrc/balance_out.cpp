// Water- and mass-balance text blocks.  See balance_out.hh for the contract.
#include "balance_out.hh"

#include <algorithm>
#include <cmath>
#include <cstdio>

namespace fl {

void region_sets(const HostMesh &m, std::vector<RegionRow> &boundary, std::vector<RegionRow> &bulk) {
    boundary.assign(m.n_boundary_regions, RegionRow{0, ""});
    bulk.assign(m.n_bulk_regions, RegionRow{0, ""});
    for (const RegionInfo &r : m.regions) (r.boundary ? boundary : bulk)[r.sub_idx] = RegionRow{r.id, r.label};
}

namespace {
constexpr int kW = 14;                 // column width
constexpr int kWl = 2 * (kW - 5) + 7;  // label column width

struct File {
    FILE *f;
    explicit File(const char *path) : f(std::fopen(path, "a")) {}
    ~File() {
        if (f) std::fclose(f);
    }
};
std::string close_status(File &fh, const char *path) {
    bool ok = !std::ferror(fh.f);
    ok = (std::fclose(fh.f) == 0) && ok;
    fh.f = nullptr;
    return ok ? "" : std::string("write error on ") + path;
}
}  // namespace

std::string write_water_balance(const char *path, double time, const std::vector<RegionRow> &bnd, const double *bal,
                                const double *plus, const double *minus, const std::vector<RegionRow> &blk,
                                const double *src) {
    File fh(path);
    if (!fh.f) return std::string("can not open balance file ") + path;
    FILE *f = fh.f;
    const int ncol = 5;
    const std::string half((kW * ncol + kWl - 15) / 2, '-'), line(kW * ncol + kWl, '-');
    std::fprintf(f, "# %s WATER BALANCE %s\n", half.c_str(), half.c_str());
    std::fprintf(f, "# Time of computed water balance: %f\n\n\n", time);
    std::fprintf(f, "# Boundary water balance:\n");
    std::fprintf(f, "# %-*s%-*s%-*s%-*s%-*s%-*s\n", kW, "[boundary_id]", kWl, "[label]", kW, "[total_balance]", kW,
                 "[total_outflow]", kW, "[total_inflow]", kW, "[time]");
    std::fprintf(f, "# %s\n", line.c_str());
    double tb = 0.0, tin = 0.0, tout = 0.0;
    for (size_t r = 0; r < bnd.size(); r++) {
        tb += bal[r];
        tout += plus[r];
        tin += minus[r];
        std::fprintf(f, "%*s%-*d%-*s  %-*g%-*g%-*g%-*g\n", 2, "", kW, bnd[r].id, kWl, bnd[r].label.c_str(), kW, bal[r],
                     kW, plus[r], kW, minus[r], kW, time);
    }
    std::fprintf(f, "# %s\n", line.c_str());
    std::fprintf(f, "# %-*s%-*g%-*g%-*g\n\n\n", kW + kWl + 2, "total boundary balance", kW, tb, kW, tout, kW, tin);

    std::fprintf(f, "# Source fluxes over material subdomains:\n");
    std::fprintf(f, "# %-*s%-*s%-*s%-*s%-*s\n", kW, "[region_id]", kWl, "[label]", kW, "[total_balance]", 2 * kW, "",
                 kW, "[time]");
    std::fprintf(f, "# %s\n", line.c_str());
    double ts = 0.0;
    for (size_t r = 0; r < blk.size(); r++) {
        ts += src[r];
        std::fprintf(f, "%*s%-*d%-*s  %-*g%-*s%-*g\n", 2, "", kW, blk[r].id, kWl, blk[r].label.c_str(), kW, src[r],
                     2 * kW, "", kW, time);
    }
    std::fprintf(f, "# %s\n", line.c_str());
    std::fprintf(f, "# %-*s%-*g\n\n\n", kW + kWl + 2, "total sources balance", kW, ts);
    return close_status(fh, path);
}

std::string write_mass_balance(const char *path, double time, double initial_time, int ns,
                               const std::vector<std::string> &names, const std::vector<RegionRow> &bnd,
                               const double *bal, const double *plus, const double *minus,
                               const std::vector<RegionRow> &blk, const double *mass, const double *src,
                               const double *init_mass, const double *int_sources, const double *int_fluxes) {
    File fh(path);
    if (!fh.f) return std::string("can not open balance file ") + path;
    FILE *f = fh.f;
    const int ncol = 6, nbr = (int)bnd.size(), nblk = (int)blk.size();
    const std::string half((kW * ncol + kWl - 14) / 2, '-'), line(kW * ncol + kWl, '-');
    std::fprintf(f, "# %s MASS BALANCE %s\n", half.c_str(), half.c_str());
    std::fprintf(f, "# Time: %f\n\n\n", time);

    std::fprintf(f, "# Mass flux through boundary [M/T]:\n");
    std::fprintf(f, "# %-*s%-*s%-*s%-*s%-*s%-*s\n", kW, "[boundary_id]", kWl, "[label]", kW, "[substance]", kW,
                 "[total flux]", kW, "[outward flux]", kW, "[inward flux]");
    std::fprintf(f, "# %s\n", line.c_str());
    std::vector<double> tb(ns, 0.0), tout(ns, 0.0), tin(ns, 0.0), mt(ns, 0.0), st(ns, 0.0);
    for (int r = 0; r < nbr; r++)
        for (int k = 0; k < ns; k++) {
            const int j = k * nbr + r;
            tb[k] += bal[j];
            tout[k] += plus[j];
            tin[k] += minus[j];
            std::fprintf(f, "%*s%-*d%-*s%-*s%-*g%-*g%-*g\n", 2, "", kW, bnd[r].id, kWl, bnd[r].label.c_str(), kW,
                         names[k].c_str(), kW, bal[j], kW, plus[j], kW, minus[j]);
        }
    std::fprintf(f, "# %s\n", line.c_str());
    for (int k = 0; k < ns; k++)
        std::fprintf(f, "# %-*s%-*s%-*g%-*g%-*g\n", kW + kWl, "Total mass flux of substance [M/T]", kW,
                     names[k].c_str(), kW, tb[k], kW, tout[k], kW, tin[k]);
    std::fprintf(f, "\n\n");

    std::fprintf(f, "# Mass [M] and sources [M/T] on regions:\n");
    std::fprintf(f, "# %-*s%-*s%-*s%-*s%-*s\n", kW, "[region_id]", kWl, "[label]", kW, "[substance]", kW,
                 "[total_mass]", kW, "[total_source]");
    std::fprintf(f, "# %s\n", line.c_str());
    for (int r = 0; r < nblk; r++)
        for (int k = 0; k < ns; k++) {
            const int j = k * nblk + r;
            mt[k] += mass[j];
            st[k] += src[j];
            std::fprintf(f, "%*s%-*d%-*s%-*s%-*g%-*g\n", 2, "", kW, blk[r].id, kWl, blk[r].label.c_str(), kW,
                         names[k].c_str(), kW, mass[j], kW, src[j]);
        }
    std::fprintf(f, "# %s\n", line.c_str());
    for (int k = 0; k < ns; k++)
        std::fprintf(f, "# %-*s%-*s%-*g%-*g\n", kW + kWl, "Total mass [M] and sources [M/T]", kW, names[k].c_str(),
                     kW, mt[k], kW, st[k]);

    std::fprintf(f,
                 "\n\n# Cumulative mass balance on time interval [%-g,%-g]\n"
                 "# Initial mass [M] + sources integrated over time [M] - flux integrated over time [M] = current "
                 "mass [M]\n"
                 "# %-*s%-*s%-*s%-*s%-*s%-*s%-*s%-*s\n",
                 initial_time, time, kW, "[substance]", kW, "[A=init. mass]", kW, "[B=source]", kW, "[C=flux]", kW,
                 "[A+B-C]", kW, "[D=curr. mass]", kW, "[A+B-C-D=err.]", kW, "[rel. error]");
    for (int k = 0; k < ns; k++) {
        const double abc = init_mass[k] + int_sources[k] - int_fluxes[k];
        const double err = abc - mt[k];
        const double den = std::max(std::fabs(abc), std::fabs(mt[k]));
        std::fprintf(f, "  %-*s%-*g%-*g%-*g%-*g%-*g%-*g%-*g\n", kW, names[k].c_str(), kW, init_mass[k], kW,
                     int_sources[k], kW, int_fluxes[k], kW, abc, kW, mt[k], kW, err, kW,
                     std::fabs(err) / (den == 0 ? 1.0 : den));
    }
    std::fprintf(f, "\n\n");
    return close_status(fh, path);
}

}  // namespace fl
