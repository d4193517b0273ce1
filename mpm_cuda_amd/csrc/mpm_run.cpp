// mpm_run.cpp -- command-line driver: the simulation loop of the reference
// application without its viewer (src/main.cpp:132-221): build the scene,
// step until the simulated time passes --time, and write
// <out>/FrameNNNNNN.bgeo whenever int(t / (1/fps)) reaches the next frame
// number (main.cpp:28-29, 214-221).
//
//   mpm_run --scene snow|jello|water|obj|boat [--obj path] [--obj2 path]
//           [--time 1.0] [--fps 36] [--out Animations] [--form fast|plain|fused]
//           [--grid N] [--dt-mode fixed|cfl|cfl_deform] [--cfl 0.5]
//           [--steps K] [--final path.bgeo]
// With a dt controller the scene's dt is the largest step (max_dt).  --steps
// stops after K steps instead of at --time; --final writes the positions
// after the last step.
// The boat scene is the reference application's own (main.cpp:138-161): a
// water cube (--obj, scaled 2x3x2, at (6, 5, 4)) and a jello boat (--obj2,
// turned 90 degrees about y, at (6, 5, 10)), gap 0.04, dt 1/5000, a
// 300x500x500 grid and 10 particles per cell; --gap, --nx, --ny, --nz
// coarsen it.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iomanip>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/mpm/MPMSimulator.hpp"

using namespace mpm;

static const char *arg(int argc, char **argv, const char *name, const char *def)
{
    for (int i = 1; i + 1 < argc; ++i)
        if (std::string(argv[i]) == name) return argv[i + 1];
    return def;
}

int main(int argc, char **argv)
{
    std::string scene = arg(argc, argv, "--scene", "snow");
    double t_end = atof(arg(argc, argv, "--time", "0.05"));
    int fps = atoi(arg(argc, argv, "--fps", "36"));
    std::string out = arg(argc, argv, "--out", "Animations");
    std::string form = arg(argc, argv, "--form", "fast");
    int n = atoi(arg(argc, argv, "--grid", "128"));
    std::string dt_mode = arg(argc, argv, "--dt-mode", "fixed");
    if (fps <= 0 || n < 16) { fprintf(stderr, "bad --fps or --grid\n"); return 2; }
    if (dt_mode != "fixed" && dt_mode != "cfl" && dt_mode != "cfl_deform") { fprintf(stderr, "bad --dt-mode\n"); return 2; }
    mkdir(out.c_str(), 0755);

    mpm_config cfg;
    mpm_config_default(&cfg);
    cfg.form = form == "plain" ? MPM_FORM_PLAIN : (form == "fused" ? MPM_FORM_FUSED : MPM_FORM_FAST);
    cfg.dt_mode = dt_mode == "cfl" ? MPM_DT_CFL : (dt_mode == "cfl_deform" ? MPM_DT_CFL_DEFORM : MPM_DT_FIXED);
    cfg.cfl = (float)atof(arg(argc, argv, "--cfl", "0.5"));
    double h = 1.0 / n, dt = 1e-4;
    unsigned ppc = 8;
    unsigned nx = (unsigned)n, ny = (unsigned)n, nz = (unsigned)n;
    const long max_steps = atol(arg(argc, argv, "--steps", "0"));
    const std::string final_path = arg(argc, argv, "--final", "");
    std::vector<ObjInitInfo> infos;
    try {
        if (scene == "snow") {
            for (int s = 0; s < 2; ++s) {
                ObjInitInfo o;
                o.shape = MPM_SHAPE_SPHERE;
                o.shapeParam = {s ? 0.7 : 0.3, 0.5, 0.5, 0.15, 0, 0};
                o.initVel = {s ? -3.0 : 3.0, 0.0, 0.0};
                o.mMaterial = Material(1.4e5, 0.2, 400.0, SNOW);
                o.mMaterial.c.theta_c = 2.5e-2; o.mMaterial.c.theta_s = 7.5e-3;
                infos.push_back(o);
            }
            ppc = 5;
        } else if (scene == "jello") {
            ObjInitInfo o;
            o.shape = MPM_SHAPE_BOX;
            o.shapeParam = {0.375, 0.1, 0.375, 0.625, 0.35, 0.625};
            o.initVel = {0.0, -2.0, 0.0};
            o.mMaterial = Material(5e4, 0.3, 1e3, JELLO);
            infos.push_back(o);
            cfg.bc_mode = MPM_BC_STICKY;
        } else if (scene == "water") {
            ObjInitInfo o;
            o.shape = MPM_SHAPE_BOX;
            double p = 11.0 / n;
            o.shapeParam = {p, p, p, p + 0.5, p + 0.4, p + 0.5};
            o.mMaterial = Material(1e4, 0.2, 1e3, WATER);
            o.mMaterial.c.eos_k = 1e4;
            o.mMaterial.c.water_c_mode = MPM_WATER_C_GRADV;
            infos.push_back(o);
            dt = 5e-5;
            ppc = 2;
        } else if (scene == "obj") {
            ObjInitInfo o;
            o.objPath = arg(argc, argv, "--obj", "");
            o.initScale = vec3(0.2f);
            o.initTranslation = vec3(0.5f, 0.5f, 0.5f);
            o.mMaterial = Material(1e4, 0.4, 800.0, JELLO);
            infos.push_back(o);
        } else if (scene == "boat") {
            ObjInitInfo w;                                   // main.cpp:139-147
            w.objPath = arg(argc, argv, "--obj", "");
            w.initVel = {0.0, 0.0, 0.0};
            w.initRotationDegree = 0.f;
            w.initScale = vec3(2.f, 3.f, 2.f);
            w.initTranslation = vec3(6.f, 5.f, 4.f);
            w.initRotationAxis = vec3(1.f, 0.f, 0.f);
            w.mMaterial = Material(5e3, 0.3, 1.1, WATER);
            infos.push_back(w);
            ObjInitInfo b;                                   // main.cpp:150-158
            b.objPath = arg(argc, argv, "--obj2", "");
            b.initVel = {0.0, 0.0, 0.0};
            b.initRotationDegree = 90.f;
            b.initScale = vec3(1.f);
            b.initTranslation = vec3(6.f, 5.f, 10.f);
            b.initRotationAxis = vec3(0.f, 1.f, 0.f);
            b.mMaterial = Material(1e4, 0.4, 0.8, JELLO);
            infos.push_back(b);
            h = atof(arg(argc, argv, "--gap", "0.04"));      // main.cpp:161
            dt = 1.0 / 5000.0;
            ppc = 10;
            nx = (unsigned)atoi(arg(argc, argv, "--nx", "300"));
            ny = (unsigned)atoi(arg(argc, argv, "--ny", "500"));
            nz = (unsigned)atoi(arg(argc, argv, "--nz", "500"));
        } else {
            fprintf(stderr, "unknown --scene %s\n", scene.c_str());
            return 2;
        }
        MPMSimulator sim(h, dt, nx, ny, nz, ppc, infos, &cfg);
        printf("particles: %d\n", sim.totalParticlesNum());
        for (size_t k = 0; k < sim.mParticlesGroupsVec.size(); ++k)
            printf("group %zu: %u\n", k, sim.mParticlesGroupsVec[k].particleNum);
        const float time_per_frame = 1.f / (float)fps;
        int frame = 0;
        std::vector<float> pos;
        auto t0 = std::chrono::steady_clock::now();
        while (max_steps > 0 ? sim.current_frame < max_steps : sim.current_time <= t_end) {
            sim.step();
            if ((int)(sim.current_time / time_per_frame) >= frame) {
                sim.getGLParticlesPos(pos);
                std::stringstream ss;
                ss << out << "/Frame" << std::setw(6) << std::setfill('0') << frame << ".bgeo";
                OutputBego(ss.str(), pos);
                ++frame;
            }
        }
        if (!final_path.empty()) {
            sim.getGLParticlesPos(pos);
            OutputBego(final_path, pos);
        }
        double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        printf("steps: %d  frames: %d  wall: %.3f s  particle-steps/s: %.3e\n", sim.current_frame, frame, wall,
               (double)sim.totalParticlesNum() * sim.current_frame / wall);
    } catch (const std::exception &e) {
        fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
