// sph_run.cpp - command-line driver of the C++ host class (pibiti::cSPH).
//
//   sph_run SCENE NX NY NZ MINX MINY MINZ SPACING JITTER VEL_SIGMA SEED STEPS OUT
//
//   sph_run --xml SCENES.xml INDEX STEPS OUT [frames]
//
// SCENE is a raw sph_scene record (sizeof(sph_scene) bytes).  The lattice
// block NX x NY x NZ from (MINX, MINY, MINZ) is advanced STEPS steps.  With
// --xml, scene INDEX of a scene file is initialised as cSPH::InitScene does
// (its init volume fill) and advanced STEPS steps; with `frames`, each step
// is a frame of the reference's app: the emitters and rain
// (cSPH::UpdateEmitter), then the step.  OUT receives, as
// little-endian binary: uint32 n, float4 positions[n], float4 velocities[n],
// float densities[n], then double mass, kinetic energy, max speed and
// simulated time.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "pibiti_sph.hpp"

static int write_out(pibiti::cSPH& sim, uint32_t steps, const char* path) {
    const uint32_t n = sim.par.num_particles;
    std::vector<float> pos(sim.getArray(SPH_ARRAY_POSITION), sim.getArray(SPH_ARRAY_POSITION) + 4 * n);
    const float* vel = sim.getArray(SPH_ARRAY_VELOCITY);
    const float* dens = sim.getDensity();
    const sph_stats st = sim.Stats();
    FILE* o = std::fopen(path, "wb");
    if (!o) {
        std::fprintf(stderr, "cannot write %s\n", path);
        return 2;
    }
    const double tail[4] = {st.mass, st.kinetic_energy, st.max_speed, st.sim_time};
    bool ok = std::fwrite(&n, 4, 1, o) == 1;
    ok = ok && std::fwrite(pos.data(), 16, n, o) == n;
    ok = ok && std::fwrite(vel, 16, n, o) == n;
    ok = ok && std::fwrite(dens, 4, n, o) == n;
    ok = ok && std::fwrite(tail, 8, 4, o) == 4;
    std::fclose(o);
    if (!ok) return 2;
    std::printf("sph_run: %u particles, %u steps, sim_time %.6g s, max speed %.6g m/s\n", n, steps,
                st.sim_time, st.max_speed);
    return 0;
}

int main(int argc, char** argv) {
    if ((argc == 6 || (argc == 7 && std::strcmp(argv[6], "frames") == 0)) && std::strcmp(argv[1], "--xml") == 0) {
        try {
            pibiti::cSPH sim;
            sim.LoadScenes(argv[2]);
            sim.SetScene((uint32_t)std::strtoul(argv[3], nullptr, 10));
            const uint32_t steps = (uint32_t)std::strtoul(argv[4], nullptr, 10);
            sim.InitScene();
            if (argc == 7) {
                for (uint32_t k = 0; k < steps; ++k) {
                    sim.UpdateEmitter();
                    sim.Update(1);
                }
            } else {
                sim.Update(steps);
            }
            return write_out(sim, steps, argv[5]);
        } catch (const pibiti::Error& e) {
            std::fprintf(stderr, "sph_run: %s\n", e.what());
            return 1;
        }
    }
    if (argc != 14) {
        std::fprintf(stderr,
                     "usage: %s SCENE NX NY NZ MINX MINY MINZ SPACING JITTER VEL_SIGMA SEED STEPS OUT\n",
                     argv[0]);
        return 2;
    }
    try {
        pibiti::cSPH sim;
        FILE* f = std::fopen(argv[1], "rb");
        if (!f || std::fread(&sim.scn, sizeof(sim.scn), 1, f) != 1) {
            std::fprintf(stderr, "cannot read %s\n", argv[1]);
            if (f) std::fclose(f);
            return 2;
        }
        std::fclose(f);
        sim.vol.nx = (uint32_t)std::strtoul(argv[2], nullptr, 10);
        sim.vol.ny = (uint32_t)std::strtoul(argv[3], nullptr, 10);
        sim.vol.nz = (uint32_t)std::strtoul(argv[4], nullptr, 10);
        for (int c = 0; c < 3; ++c) sim.vol.min[c] = std::strtof(argv[5 + c], nullptr);
        sim.vol.spacing = std::strtof(argv[8], nullptr);
        sim.vol.jitter = std::strtof(argv[9], nullptr);
        sim.vol.vel_sigma = std::strtof(argv[10], nullptr);
        sim.vol.seed = std::strtoull(argv[11], nullptr, 10);
        const uint32_t steps = (uint32_t)std::strtoul(argv[12], nullptr, 10);
        sim.InitScene();
        sim.Update(steps);
        return write_out(sim, steps, argv[13]);
    } catch (const pibiti::Error& e) {
        std::fprintf(stderr, "sph_run: %s\n", e.what());
        return 1;
    }
    return 0;
}
