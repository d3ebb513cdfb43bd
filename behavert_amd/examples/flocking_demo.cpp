// flocking_demo.cpp -- a BehaveRT_Flocking-style driver on the plug-in layer.
//
// The per-step call sequence is that of BehaveRT_Flocking::update
// (Samples/OpensteerDemo/plugins/BehaveRT_Flocking.cpp:177-196), GL calls removed:
//   reorderSimData, computeNeighborhoods, resetSteerForce, computeSeparations,
//   computeAlignments, computeCohesions, [computeSeekingsWorldCenter],
//   applySteeringForces(dt)
// usage: flocking_demo <config.cfg> <steps> [init.bin] [out.bin]
//   init.bin / out.bin: pos, forward, smooth_accel as float32 [N][4] each, in
//   agent-ID order (without init.bin the plug-ins' reset() state is used).
//        flocking_demo --dump-config <config.cfg> <section> <key> [<section> <key> ...]
//   prints each value as read by the configuration parser (no device needed).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include "../plugins/BehaveRT.h"

typedef OpenSteerWrapperPlugIn<Proximity3DPlugIn<EnvGrid3DPlugIn<Body3DPlugIn<DummyPlugIn>>>> CrowdEngineFinal;

static bool read_state(const char* path, CrowdEngineFinal& e, int n) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    BehaveRT::DeviceArrayWrapper<float4>* arrs[3] = {e.getPos(), e.getForward(), e.getSmoothedAcceleration()};
    for (auto* a : arrs) {
        in.read(reinterpret_cast<char*>(a->getHostArray()), (std::streamsize)sizeof(float4) * n);
        if (!in) return false;
        a->copyArrayToDevice();  // to the write buffer, as the reference does
        a->swapPosReadWrite();
    }
    return true;
}

static bool write_state(const char* path, CrowdEngineFinal& e, int n) {
    std::ofstream out(path, std::ios::binary);
    BehaveRT::DeviceArrayWrapper<float4>* arrs[3] = {e.getPos(), e.getForward(), e.getSmoothedAcceleration()};
    for (auto* a : arrs) {
        a->copyArrayFromDevice();
        out.write(reinterpret_cast<const char*>(a->getHostArray()), (std::streamsize)sizeof(float4) * n);
    }
    return (bool)out;
}

static int dump_config(int argc, char** argv) {
    BehaveRT::UtilityConfigFile cfg;
    if (!cfg.load(argv[2])) return 2;
    for (int k = 3; k + 1 < argc; k += 2) {
        const std::string v = cfg.getSetting(argv[k + 1], argv[k]);
        const float3 f3 = BehaveRT::StringConverter::parseFloat3(v);
        std::printf("%s|%s|%s|%g|%d|%d|%g %g %g\n", argv[k], argv[k + 1], v.c_str(),
                    BehaveRT::StringConverter::parseFloat(v), BehaveRT::StringConverter::parseInt(v),
                    (int)BehaveRT::StringConverter::parseBool(v), f3.x, f3.y, f3.z);
    }
    return 0;
}

int main(int argc, char** argv) {
    if (argc >= 3 && std::string(argv[1]) == "--dump-config") return dump_config(argc, argv);
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <config.cfg> <steps> [init.bin] [out.bin]\n", argv[0]);
        return 2;
    }
    try {
        m_CommonRes.initialize(argv[1]);
        const int steps = std::atoi(argv[2]);
        const float dt = BehaveRT::StringConverter::parseFloat(
            m_CommonRes.getConfig()->getSetting("updateElapsedTime", "Benchmarking"));
        const bool seekCenter = BehaveRT::StringConverter::parseBool(
            m_CommonRes.getConfig()->getSetting("seekWorldCenter", "Demo"));
        CrowdEngineFinal engine;
        engine.reset();
        const int n = hBody3DParams.numBodies;
        if (argc > 3 && !read_state(argv[3], engine, n)) {
            std::fprintf(stderr, "cannot read %s\n", argv[3]);
            return 2;
        }
        for (int s = 0; s < steps; ++s) {
            engine.reorderSimData();
            engine.computeNeighborhoods();
            engine.resetSteerForce();
            engine.computeSeparations();
            engine.computeAlignments();
            engine.computeCohesions();
            if (seekCenter) engine.computeSeekingsWorldCenter();
            engine.applySteeringForces(dt);
        }
        m_CommonRes.getDeviceInterface()->threadSync();
        if (argc > 4 && !write_state(argv[4], engine, n)) return 3;
        std::printf("flocking_demo: %d agents, %d steps, grid %u x %u x %u, K %u\n", n, steps,
                    hEnvGrid3DParams.gridSize.x, hEnvGrid3DParams.gridSize.y, hEnvGrid3DParams.gridSize.z,
                    hProximity3DParams.maxNeighbors);
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "flocking_demo: %s\n", ex.what());
        return 1;
    }
    return 0;
}
