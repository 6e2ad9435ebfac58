// The reference's tests/12_radioactive_decay (flow_decay_gmsh.con) through the C++
// adapter: mesh, legacy .fbc/.tbc boundary files, steady flow, six substances with
// the decay chain D -> F -> B, and the transport output stream written as an MSH
// file (every substance at t = 0, 0.5, .., 10) plus the mass balance.
// Usage: decay_chain <mesh.msh> <flow.fbc> <transport.tbc> <out.msh> <mass_balance.txt>
#include <cstdio>
#include <string>
#include <vector>

#include "../flow123d_gpu.hh"

int main(int argc, char **argv) {
    if (argc < 6) {
        std::fprintf(stderr, "usage: %s mesh.msh flow.fbc transport.tbc out.msh balance.txt\n", argv[0]);
        return 2;
    }
    try {
        fl::GpuMesh mesh(argv[1]);
        const int E = mesh.n_elements(), NS = 6;
        fl::DarcyData d;
        d.conductivity.assign(E, 1.0);
        d.cross_section.assign(E, 1.0);
        d.sigma.assign(E, 1.0);
        d.water_source_density.assign(E, 0.0);
        d.anisotropy.assign(9 * (size_t)E, 0.0);
        for (int e = 0; e < E; e++) d.anisotropy[9 * e] = d.anisotropy[9 * e + 4] = d.anisotropy[9 * e + 8] = 0.01;
        fl::OldBcdInput bcd(mesh);
        bcd.read_flow(argv[2], d);
        std::vector<double> bc_conc = bcd.read_transport(argv[3], NS);
        fl::DarcyFlowMH_Steady darcy(mesh, d, fl_solver_opts{1e-14, 1e-14, 100000, 0, 0});
        darcy.update_solution();
        std::vector<double> init((size_t)NS * E);
        for (int s = 0; s < NS; s++)
            for (int e = 0; e < E; e++) init[(size_t)s * E + e] = 0.01 * (s + 1);
        fl::ConvectionTransport tr(mesh, NS, std::vector<double>(E, 1.0), d.cross_section, bc_conc, init, 10.0, 0.5);
        tr.set_velocity_field(darcy);
        // D (3) -> F (5), F -> B (1), half-lives 2.5
        tr.set_reactions({3, 5}, {0, 1, 2}, {5, 1}, {2.5, 2.5}, {1.0, 1.0});
        fl::OutputMSH out(mesh, argv[4]);
        const std::vector<std::string> names = {"A", "B", "C", "D", "E", "F"};
        int steps = 0;
        for (int k = 0; k <= 20; k++) {
            steps += tr.update_solution(0.5 * k);
            std::vector<double> c = tr.concentrations();
            for (int s = 0; s < NS; s++) out.write_elem_data(names[s] + "_mobile", "M/L^3", &c[(size_t)s * E], 1, E, 0.5 * k);
            out.next_step();
            tr.write_mass_balance(argv[5], names);
        }
        std::printf("decay chain: %d transport steps, t = %g\n", steps, tr.time());
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
