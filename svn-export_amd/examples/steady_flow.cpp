// Drives the GPU path the way HC_ExplicitSequential::run_simulation
// (src/coupling/hc_explicit_sequential.cc:151-235) drives the reference:
// one steady Darcy solve, then the transport step loop to each output time.
// Usage: steady_flow <mesh.msh> [flow_output.msh]  -- unit conductivity, Dirichlet 1
// on the first boundary region and 0 on the second, one substance entering with
// c = 1; with a second argument the flow's output fields are written there as
// DarcyFlowMHOutput::output writes them (node pressure, element pressure,
// piezometric head, centre velocity).
#include <cstdio>
#include <vector>

#include "../flow123d_gpu.hh"

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s mesh.msh [flow_output.msh]\n", argv[0]);
        return 2;
    }
    try {
        fl::GpuMesh mesh(argv[1]);
        auto c = mesh.counts();
        int E = c[1], NBC = c[2], R = c[7];
        std::vector<int> ids(R), isb(R), sub(R);
        char buf[4096];
        fl::check(fl_mesh_regions(mesh.handle(), ids.data(), isb.data(), sub.data(), buf, sizeof buf), "regions");
        std::vector<int> dim(E), eid(E), ereg(E), en(4 * E), ee(4 * E), eb(4 * E), ep(c[4] + 1), es(c[3]);
        std::vector<int> vbe(c[6] + 1), vbg(c[6] + 1), be(c[5] + 1), bb(c[5] + 1), bcr(NBC + 1);
        fl::check(fl_mesh_topology(mesh.handle(), dim.data(), eid.data(), ereg.data(), en.data(), ee.data(),
                                   eb.data(), ep.data(), es.data(), vbe.data(), vbg.data(), be.data(), bb.data(),
                                   bcr.data()),
                  "topology");
        fl::DarcyData d;
        d.conductivity.assign(E, 1.0);
        d.cross_section.assign(E, 1.0);
        d.sigma.assign(E, 1.0);
        d.water_source_density.assign(E, 0.0);
        d.bc_type.assign(NBC, 0);
        d.bc_pressure.assign(NBC, 0.0);
        d.bc_flux.assign(NBC, 0.0);
        d.bc_robin_sigma.assign(NBC, 0.0);
        int first = -1, second = -1;
        for (int r = 0; r < R; r++)
            if (isb[r] && ids[r] != -3) {
                if (first < 0) first = ids[r];
                else if (second < 0) second = ids[r];
            }
        for (int b = 0; b < NBC; b++) {
            if (bcr[b] == first) { d.bc_type[b] = 1; d.bc_pressure[b] = 1.0; }
            if (bcr[b] == second) { d.bc_type[b] = 1; d.bc_pressure[b] = 0.0; }
        }
        fl::DarcyFlowMH_Steady darcy(mesh, d, fl_solver_opts{1e-12, 0.0, 100000, 0, 0});
        darcy.update_solution();
        double *sol;
        unsigned n;
        darcy.get_solution_vector(sol, n);
        const fl_solve_info &i = darcy.last_solve();
        std::printf("darcy: %u unknowns, %d CG iterations, residual %.3e, %.3f ms\n", n, i.iterations, i.residual,
                    i.ms_total);
        if (argc > 2) {
            std::vector<double> pn, pe, ph, v;
            darcy.output_fields(pn, pe, ph, v);
            fl::OutputMSH out(mesh, argv[2]);
            out.write_node_data("pressure_nodes", "L", pn.data(), 1, (int)pn.size(), 0.0);
            out.write_elem_data("pressure_elements", "L", pe.data(), 1, E, 0.0);
            out.write_elem_data("piezo_head_elements", "L", ph.data(), 1, E, 0.0);
            out.write_elem_data("velocity_elements", "L/T", v.data(), 3, E, 0.0);
        }
        std::vector<double> bc(NBC, 1.0), ic(E, 0.0);
        fl::ConvectionTransport tr(mesh, 1, std::vector<double>(E, 1.0), d.cross_section, bc, ic, 1.0, 0.25);
        tr.set_velocity_field(darcy);
        int steps = 0;
        for (int k = 1; k <= 4; k++) steps += tr.update_solution(0.25 * k);
        auto conc = tr.concentrations();
        double cmax = 0;
        for (double v : conc) cmax = v > cmax ? v : cmax;
        std::printf("transport: %d steps to t=%.2f, max concentration %.6f\n", steps, tr.time(), cmax);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
