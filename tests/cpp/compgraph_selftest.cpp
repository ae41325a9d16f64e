// Drives the C++ host mirror (DiffMPMLib3D.h) the way the reference's GUI
// drives CompGraph (DiffMPMAnimator3D.cpp:421-471), on files written by
// tests/test_cpp_mirror.py:
//   compgraph_selftest roundtrip IN.mpmbin OUT.mpmbin      (no GPU)
//   compgraph_selftest run DIR                             (GPU)
//   compgraph_selftest optimize DIR                        (GPU)
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
// optimisation), final_loss.txt, opt_l0.mpmbin (layer 0 with its optimised dFc).
#include "DiffMPMLib3D/DiffMPMLib3D.h"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>

using namespace DiffMPMLib3D;

int main(int argc, char** argv)
{
    if (argc == 4 && std::string(argv[1]) == "roundtrip") {
        PointCloud pc;
        if (!pc.ReadEntirePointCloudFromBinaryFile(argv[2])) return 2;
        pc.WriteEntirePointCloudToBinaryFile(argv[3]);
        return 0;
    }
    const std::string mode = argc == 3 ? std::string(argv[1]) : std::string();
    if (mode != "run" && mode != "optimize") {
        std::cerr << "usage: compgraph_selftest roundtrip IN OUT | run DIR | optimize DIR\n";
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
