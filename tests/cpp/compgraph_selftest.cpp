// Drives the C++ host mirror (DiffMPMLib3D.h) the way the reference's GUI
// drives CompGraph (DiffMPMAnimator3D.cpp:421-471), on files written by
// tests/test_cpp_mirror.py:
//   compgraph_selftest roundtrip IN.mpmbin OUT.mpmbin      (no GPU)
//   compgraph_selftest run DIR                             (GPU)
//   compgraph_selftest optimize DIR                        (GPU)
//   compgraph_selftest phases DIR                          (GPU)
// DIR holds in.mpmbin (layer 0, dFc set), target.bin (float64 node masses),
// params.txt ("dims dx minx miny minz dt drag s fx fy fz layers compat");
// outputs: loss.txt, last.mpmbin (last layer), l0.mpmbin (layer 0 with its
// gradients after ComputeBackwardPass(0)), bt.mpmbin (layer 0 gradients from
// Back_Timestep on layer 1's gradients), step.mpmbin (ForwardTimeStep of layer 0),
// replay.mpmbin, replay_grid.txt (MultiThreadMPM::ForwardTimeStep of layer 0 in
// place: the cloud, then each node's m and v, one a line),
// pk1.txt (PK_FixedCorotatedElasticity, then FixedCorotatedElasticity, at
// F = 0.8 Rx(45 deg)), diag.txt (GetPointDeterminants, GetPointElasticEnergies
// of layer 0, one point a line); optimize: history.txt (gradient norms of the
// optimisation), final_loss.txt, opt_l0.mpmbin (layer 0 with its optimised dFc);
// phases: the ForwardSimulation.h phase functions as reference callers use them
// (outputs listed at run_phases).
#include "DiffMPMLib3D/DiffMPMLib3D.h"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>

using namespace DiffMPMLib3D;

namespace {

void write_grid(const std::string& path, const Grid& g)
{
    std::ofstream o(path);
    o.precision(9);
    for (int i = 0; i < g.dim_x; ++i)
        for (int j = 0; j < g.dim_y; ++j)
            for (int k = 0; k < g.dim_z; ++k) {
                const GridNode& nd = g.nodes[i][j][k];
                o << nd.m << " " << nd.p[0] << " " << nd.p[1] << " " << nd.p[2] << " " << nd.v[0] << " " << nd.v[1] << " " << nd.v[2]
                  << "\n";
            }
}

// The phase functions of ForwardSimulation.h, called as the reference's callers do:
//   phases.mpmbin, phases_grid.txt  P_op_1, G_Reset, P2G, G_op, G2P, P_op_2 of layer 0
//                                   (ForwardSimulation.cpp:136-157 spelled out); grid lines "m p v"
//   target_grid.txt                 Grid copy, ResetValues, P2G(cloud, grid, 0.0, 0.0)
//                                   (MPMPolyscope.cpp:158-160)
//   p2g_mass.txt, g2g_mass.txt      P2G_Mass(positions, grid, 1.0), then G2G_Mass onto a
//                                   grid of (dims + 1) / 2 nodes at 2 dx (DiffMPMAnimator3D.cpp:501-505)
//   gravity_grid.txt                G_op with a gravity point at the grid centre, magnitude 3
//   single.mpmbin, single_grid.txt  the single-particle forms on particles 0..4: op_1 and
//                                   to_grid into a reset grid, SingleNode_op on every node,
//                                   Grid_to_SingleParticle and op_2 (smoothing form) from the
//                                   phases' grid
void run_phases(const std::string& dir, PointCloud& pc, const Grid& grid, float dt, float drag, float s, Vec3 fe)
{
    PointCloud curr = pc, next = pc;
    Grid g(grid);
    SingleThreadMPM::P_op_1(curr);
    SingleThreadMPM::G_Reset(g);
    SingleThreadMPM::P2G(curr, g, dt, drag);
    SingleThreadMPM::G_op(g, dt, fe);
    SingleThreadMPM::G2P(next, curr, g);
    SingleThreadMPM::P_op_2(next, curr, s, dt);
    next.WriteEntirePointCloudToBinaryFile(dir + "/phases.mpmbin");
    write_grid(dir + "/phases_grid.txt", g);

    auto target_point_cloud = std::make_shared<PointCloud>(pc);
    auto target_grid = std::make_shared<Grid>(grid);
    target_grid->ResetValues();
    SingleThreadMPM::P2G(*target_point_cloud, *target_grid, 0.0, 0.0);
    write_grid(dir + "/target_grid.txt", *target_grid);

    Grid mg(grid);
    SingleThreadMPM::P2G_Mass(pc.GetPointPositions(), mg, 1.0);
    write_grid(dir + "/p2g_mass.txt", mg);
    Grid coarse((grid.dim_x + 1) / 2, (grid.dim_y + 1) / 2, (grid.dim_z + 1) / 2, 2.f * grid.dx, grid.min_point);
    SingleThreadMPM::G2G_Mass(mg, coarse);
    write_grid(dir + "/g2g_mass.txt", coarse);

    Grid gg(g);
    const Vec3 centre = grid.min_point + Vec3((float)grid.dim_x, (float)grid.dim_y, (float)grid.dim_z) * (0.5f * grid.dx);
    SingleThreadMPM::G_op(gg, dt, centre, 3.0f);
    write_grid(dir + "/gravity_grid.txt", gg);

    PointCloud single;
    Grid sg(grid);
    sg.ResetValues();
    for (int p = 0; p < 5; ++p) {
        MaterialPoint mp = pc.points[p];
        SingleThreadMPM::SingleParticle_op_1(mp);
        SingleThreadMPM::SingleParticle_to_grid(mp, sg, dt, drag);
        MaterialPoint nx = next.points[p];
        nx.v.setZero();
        nx.C.setZero();
        SingleThreadMPM::Grid_to_SingleParticle(nx, curr.points[p], g);
        SingleThreadMPM::SingleParticle_op_2(nx, curr.points[p], s, dt);
        single.points.push_back(nx);
    }
    for (int i = 0; i < sg.dim_x; ++i)
        for (int j = 0; j < sg.dim_y; ++j)
            for (int k = 0; k < sg.dim_z; ++k)
                if (sg.nodes[i][j][k].m != 0.f) SingleThreadMPM::SingleNode_op(sg.nodes[i][j][k], dt, fe);
    single.WriteEntirePointCloudToBinaryFile(dir + "/single.mpmbin");
    write_grid(dir + "/single_grid.txt", sg);
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc == 4 && std::string(argv[1]) == "roundtrip") {
        PointCloud pc;
        if (!pc.ReadEntirePointCloudFromBinaryFile(argv[2])) return 2;
        pc.WriteEntirePointCloudToBinaryFile(argv[3]);
        return 0;
    }
    const std::string mode = argc == 3 ? std::string(argv[1]) : std::string();
    if (mode != "run" && mode != "optimize" && mode != "phases") {
        std::cerr << "usage: compgraph_selftest roundtrip IN OUT | run DIR | optimize DIR | phases DIR\n";
        return 1;
    }
    const std::string dir = argv[2];
    try {
        int dims[3], layers, compat;
        float dx, mn[3], dt, drag, s, f[3];
        {
            std::ifstream p(dir + "/params.txt");
            p >> dims[0] >> dims[1] >> dims[2] >> dx >> mn[0] >> mn[1] >> mn[2] >> dt >> drag >> s >> f[0] >> f[1] >> f[2] >>
                layers >> compat;
            if (!p) throw std::runtime_error("bad params.txt");
        }
        GetDeviceOptions().compat = compat;
        auto pc = std::make_shared<PointCloud>();
        if (!pc->ReadEntirePointCloudFromBinaryFile(dir + "/in.mpmbin")) throw std::runtime_error("bad in.mpmbin");
        auto grid = std::make_shared<Grid>(dims[0], dims[1], dims[2], dx, Vec3(mn[0], mn[1], mn[2]));
        auto target = std::make_shared<Grid>(*grid);
        {
            std::ifstream t(dir + "/target.bin", std::ios::binary);
            for (int i = 0; i < dims[0]; ++i)
                for (int j = 0; j < dims[1]; ++j)
                    for (int k = 0; k < dims[2]; ++k) {
                        double m;
                        t.read((char*)&m, 8);
                        target->nodes[i][j][k].m = (float)m;
                    }
            if (!t) throw std::runtime_error("bad target.bin");
        }
        const Vec3 fe(f[0], f[1], f[2]);

        if (mode == "phases") {
            run_phases(dir, *pc, *grid, dt, drag, s, fe);
            return 0;
        }
        if (mode == "optimize") {
            // the GUI's optimisation call (DiffMPMAnimator3D.cpp:436-447) on a short
            // sequence: control stride 1, 2 descent iterations with 3 line-search
            // trials per control layer; the gradient-norm history and the final
            // loss and layer-0 control are written out
            CompGraph cg(pc, grid, target);
            cg.OptimizeDefGradControlSequence(layers, dt, drag, fe, 1, 2, 3, 0.05f, 1e-3f, s);
            std::ofstream h(dir + "/history.txt");
            h.precision(9);
            for (float g : cg.gradientNormHistory) h << g << "\n";
            std::ofstream(dir + "/final_loss.txt") << std::scientific << (double)cg.EndLayerMassLoss() << "\n";
            cg.layers.front().point_cloud->WriteEntirePointCloudToBinaryFile(dir + "/opt_l0.mpmbin");
            return 0;
        }
        CompGraph cg(pc, grid, target);
        cg.SetSimulationParameters(dt, drag, fe, s);
        cg.SetUpCompGraph(layers);
        cg.ComputeForwardPass(0);
        const float loss = cg.EndLayerMassLoss();
        cg.ComputeBackwardPass(0);
        std::ofstream(dir + "/loss.txt") << std::scientific << (double)loss << "\n";
        cg.layers.back().point_cloud->WriteEntirePointCloudToBinaryFile(dir + "/last.mpmbin");
        cg.layers.front().point_cloud->WriteEntirePointCloudToBinaryFile(dir + "/l0.mpmbin");
        {
            // per-point diagnostics of layer 0 (its F and dFc): J, then psi
            const std::vector<float> J = cg.layers.front().point_cloud->GetPointDeterminants();
            const std::vector<float> psi = cg.layers.front().point_cloud->GetPointElasticEnergies();
            std::ofstream o(dir + "/diag.txt");
            o.precision(9);
            for (size_t p = 0; p < J.size(); ++p) o << J[p] << " " << psi[p] << "\n";
        }

        // Back_Timestep on the full gradients of layer 1
        cg.ComputeForwardPass(0);
        cg.EndLayerMassLoss();
        cg.ComputeBackwardPass(1);
        CompGraphLayer l0{std::make_shared<PointCloud>(*cg.layers[0].point_cloud), std::make_shared<Grid>(*grid)};
        l0.point_cloud->ResetGradients();
        Back_Timestep(cg.layers[1], l0, drag, dt, s);
        l0.point_cloud->WriteEntirePointCloudToBinaryFile(dir + "/bt.mpmbin");

        // one ForwardTimeStep of layer 0
        PointCloud next = *pc;
        Grid g2(*grid);
        SingleThreadMPM::ForwardTimeStep(next, *pc, g2, s, dt, drag, fe);
        next.WriteEntirePointCloudToBinaryFile(dir + "/step.mpmbin");

        // the animator's in-place replay step with 4 proxy grids (DiffMPMAnimator3D.cpp:83-111)
        PointCloud replay = *pc;
        Grid g3(*grid);
        std::vector<std::shared_ptr<Grid>> proxies(4);
        for (auto& q : proxies) q = std::make_shared<Grid>(*grid);
        MultiThreadMPM::ForwardTimeStep(replay, replay, g3, proxies, dt, drag, fe);
        replay.WriteEntirePointCloudToBinaryFile(dir + "/replay.mpmbin");
        {
            std::ofstream gm(dir + "/replay_grid.txt");
            gm.precision(9);
            for (int i = 0; i < g3.dim_x; ++i)
                for (int j = 0; j < g3.dim_y; ++j)
                    for (int k = 0; k < g3.dim_z; ++k) {
                        const GridNode& nd = g3.nodes[i][j][k];
                        gm << nd.m << " " << nd.v[0] << " " << nd.v[1] << " " << nd.v[2] << "\n";
                    }
        }
        std::vector<std::shared_ptr<Grid>> none;
        try {
            MultiThreadMPM::ForwardTimeStep(replay, replay, g3, none, dt, drag, fe);
            std::cerr << "error: empty proxy_grids accepted\n";
            return 4;
        } catch (const std::invalid_argument&) {
        }

        // Elasticity at the reference's gradient-check point (DiffMPMAnimator3D.cpp:219-225)
        Mat3 F;
        const float c = std::cos(0.78539816f), sn = std::sin(0.78539816f);
        F(0, 0) = 0.8f;
        F(1, 1) = 0.8f * c;
        F(1, 2) = -0.8f * sn;
        F(2, 1) = 0.8f * sn;
        F(2, 2) = 0.8f * c;
        const Mat3 P = PK_FixedCorotatedElasticity(F, 58333.0f, 38888.9f);
        std::ofstream o(dir + "/pk1.txt");
        o.precision(9);
        for (int i = 0; i < 9; ++i) o << P.d[i] << " ";
        o << FixedCorotatedElasticity(F, 58333.0f, 38888.9f) << "\n";
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 3;
    }
    return 0;
}
