// A spg-style app driving the GPU VBD solver exactly like apps/solverCompare.cpp drives spg::solver::VBD:
// build a ParticleGroup with bending, membrane and anchor energies (createCloth, solverCompare.cpp:367-498),
// add it to the solver, set dt and gravity, call step() in a loop, read particleGroups().
//
// Usage: cloth_vbd_gpu [n=32] [steps=10] [iterations=10]
// Prints one line per step: step index, mean y of the particles, max |v|.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>

#include <spg/geom/triangleMesh.h>
#include <spg/sim/energy/energies.h>
#include <spg/sim/solver/baseSolver.h>

int main(int argc, char **argv)
{
    const int n = argc > 1 ? std::atoi(argv[1]) : 32;
    const int steps = argc > 2 ? std::atoi(argv[2]) : 10;
    const int iterations = argc > 3 ? std::atoi(argv[3]) : 10;
    try {
        std::vector<spg::Vector3> x0;
        std::vector<spg::Int3> faces;
        const double s = 1.0 / (n - 1);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) x0.push_back({j * s, i * s, 0.0});
        for (int i = 0; i + 1 < n; ++i)
            for (int j = 0; j + 1 < n; ++j) {
                const int a = i * n + j, b = a + 1, c = a + n + 1, d = a + n;
                faces.push_back({a, b, c});
                faces.push_back({a, c, d});
            }
        const spg::TriangleMesh mesh(x0, x0, faces);

        spg::ParticleGroup cloth;
        for (const auto &p : x0) cloth.addParticle(p, p, 1e-3);
        auto bending = std::make_shared<spg::DiscreteBendingEnergy>();
        for (const auto &st : mesh.bendingStencils()) bending->addStencil(st, 0.0, 1e-4);
        cloth.addEnergy(bending);
        bending->preparePrecomputations(cloth);
        auto membrane = std::make_shared<spg::MembraneStvkEnergy>();
        for (const auto &f : faces) membrane->addStencil(f, 1e4 * 1e-3, 0.3);
        cloth.addEnergy(membrane);
        membrane->preparePrecomputations(cloth);
        auto anchors = std::make_shared<spg::SpringAnchorEnergy>();
        anchors->addStencil({(n - 1) * n}, cloth.positions()[(n - 1) * n], 1e5);
        anchors->addStencil({n * n - 1}, cloth.positions()[n * n - 1], 1e5);
        cloth.addEnergy(anchors);

        spg::solver::VBDGpu solver;
        solver.addObject(cloth);
        solver.setDt(1.0 / 60.0);
        solver.setGravity({0.0, -9.81, 0.0});
        solver.setIterations(iterations);
        for (int k = 0; k < steps; ++k) {
            solver.step();
            const auto &pg = solver.particleGroups()[0];
            double my = 0.0, vmax = 0.0;
            for (int i = 0; i < pg.size(); ++i) {
                my += pg.positions()[i].y();
                vmax = std::fmax(vmax, pg.velocities()[i].norm());
            }
            std::printf("step %d mean_y %.9f max_v %.9f\n", k, my / pg.size(), vmax);
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
