// clustering_demo.cpp -- a BehaveRT_Clustering-style driver on the plug-in layer.
//
// The per-step call sequence is that of BehaveRT_Clustering's update
// (Samples/OpensteerDemo/plugins/BehaveRT_Clustering.cpp:209-235), with the GL
// and merge/freeze calls removed:
//   reorderSimData, computeNeighborhoods, resetSteerForce, computeSeparations,
//   computeAlignments, computeCohesions, steerForClustering,
//   [computeClusterMembership], [computeSeekingsWorldCenter], applySteeringForces(dt)
// ([Demo] computeLocalPropagation and seekWorldCenter switch the bracketed calls.)
// usage: clustering_demo <config.cfg> <steps> [init.bin] [out.bin]
//   init.bin: pos, forward, smooth_accel as float32 [N][4] each, features as
//             float32 [N][featuresCount], similarity stats as float32 [N][4],
//             all in agent-ID order (without it the plug-ins' reset() state is used)
//   out.bin:  pos, forward, smooth_accel, similarity stats (ID order), then the
//             last step's steer force (sorted order), float32 [N][4] each, then
//             the cluster membership float32 [N][3] (ID order).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include "../plugins/BehaveRT.h"

typedef BehaviorClusteringPlugIn<
    OpenSteerWrapperPlugIn<Proximity3DPlugIn<EnvGrid3DPlugIn<Body3DPlugIn<DummyPlugIn>>>>>
    ClusteringEngine;

template <class T>
static bool read_array(std::ifstream& in, BehaveRT::DeviceArrayWrapper<T>* a) {
    in.read(reinterpret_cast<char*>(a->getHostArray()), (std::streamsize)a->getBytesCount());
    if (!in) return false;
    a->copyArrayToDevice();  // to the write buffer, as the reference does
    a->swapPosReadWrite();
    return true;
}

static bool read_state(const char* path, ClusteringEngine& e) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    return read_array(in, e.getPos()) && read_array(in, e.getForward()) &&
           read_array(in, e.getSmoothedAcceleration()) && read_array(in, e.getFeaturesVector()) &&
           read_array(in, e.getSimilarityStats());
}

static bool write_state(const char* path, ClusteringEngine& e) {
    std::ofstream out(path, std::ios::binary);
    BehaveRT::DeviceArrayWrapper<float4>* arrs[5] = {e.getPos(), e.getForward(), e.getSmoothedAcceleration(),
                                                     e.getSimilarityStats(), e.getSteerForce()};
    for (auto* a : arrs) {
        a->copyArrayFromDevice();
        out.write(reinterpret_cast<const char*>(a->getHostArray()), (std::streamsize)a->getBytesCount());
    }
    BehaveRT::DeviceArrayWrapper<float>* m = e.getClusterMembership();
    m->copyArrayFromDevice();
    out.write(reinterpret_cast<const char*>(m->getHostArray()), (std::streamsize)m->getBytesCount());
    return (bool)out;
}

int main(int argc, char** argv) {
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
        const bool localPropagation = BehaveRT::StringConverter::parseBool(
            m_CommonRes.getConfig()->getSetting("computeLocalPropagation", "Demo"));
        ClusteringEngine engine;
        engine.reset();
        if (argc > 3 && !read_state(argv[3], engine)) {
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
            engine.steerForClustering(dt);
            if (localPropagation) engine.computeClusterMembership();
            if (seekCenter) engine.computeSeekingsWorldCenter();
            engine.applySteeringForces(dt);
        }
        m_CommonRes.getDeviceInterface()->threadSync();
        if (argc > 4 && !write_state(argv[4], engine)) return 3;
        std::printf("clustering_demo: %d agents, %d steps, %d features, K %u\n", hBody3DParams.numBodies, steps,
                    hBehaviorClusteringParams.featuresCount, hProximity3DParams.maxNeighbors);
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "clustering_demo: %s\n", ex.what());
        return 1;
    }
    return 0;
}
