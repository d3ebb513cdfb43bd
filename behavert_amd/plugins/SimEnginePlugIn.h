// SimEnginePlugIn.h -- plug-in mixin framework, common resources and the host
// parameter structs of the step path.
//
// Mirrors:
//   AbstractSimEnginePlugIn / SimEnginePlugIn / DummyPlugIn
//                                     Core/include/SimEnginePlugIn.h:79-200
//   installPlugIn / checkForPlugInDependencies
//                                     Core/src/SimEnginePlugIn.cpp:45-116
//   CommonResources (m_CommonRes)     Core/include/CommonResources.h:62-126,
//                                     Core/src/CommonResources.cpp:49-92
//   h<PlugIn>Params host structs      share_struct (Core/include/devicedata.cuh:195-203)
//
// An engine is a mixin chain, e.g.
//   OpenSteerWrapperPlugIn<Proximity3DPlugIn<EnvGrid3DPlugIn<Body3DPlugIn<DummyPlugIn>>>>
// whose constructors run base-first; each level installs itself (allocates its
// features, reads its "[<name()>]" config section) and checks that the plug-ins
// it depends on are already installed.  A missing dependency throws
// BehaveRT::PlugInDependencyError (the reference paused and called exit(-1)).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "DeviceArrayWrapper.h"
#include "DeviceInterface.h"
#include "UtilityConfigFile.h"

// ---------------------------------------------------------------- host params

/// Body3DParams (PlugIns/Body/include/body3d_kernel.cuh:39-58)
struct Body3DParams {
    int numBodies = 0;
    float commonRadius = 0.0f;
    bool use2DProjection = false;
};

/// EnvGrid3DParams (PlugIns/EnvGrid3D/include/envgrid3d_kernel.cuh:38-65)
struct EnvGrid3DParams {
    uint3 gridSize = make_uint3(0, 0, 0);
    unsigned int numCells = 0;
    float3 worldOrigin = make_float3(0, 0, 0);
    float3 worldRadius = make_float3(0, 0, 0);
    float3 cellSize = make_float3(0, 0, 0);
    unsigned int maxBodiesPerCell = BRT_UNLIMITED;
    float3 worldCenter = make_float3(0, 0, 0);
    bool disableSorting = false;
    bool lockWorldProportions = false;
};

/// Proximity3DParams (PlugIns/Proximity3D/include/Proximity3D_kernel.cuh:40-73)
struct Proximity3DParams {
    uint3 searchDepth = make_uint3(1, 1, 1);
    bool useKnn = false;
    unsigned int maxNeighbors = 0;
    unsigned int numNeighWordsPerAgent = 0;
    float commonSearchRadius = 0.0f;
};

/// OpenSteerWrapperParams (PlugIns/OpenSteerWrapper/include/OpenSteerWrapper_kernel.cuh)
struct OpenSteerWrapperParams {
    float3 separationParams = make_float3(0, 0, 0);  // weight, radius, cosMaxAngle
    float3 cohesionParams = make_float3(0, 0, 0);
    float3 alignmentParams = make_float3(0, 0, 0);
    float commonMass = 1.0f;
    float commonMaxSpeed = 0.0f;
    float commonMaxForce = 0.0f;
    float commonSensingMinRadius = 0.0f;
    float3 avoidBaseParams = make_float3(0, 0, 0);
    bool useForwardVBO = false;
};
struct OpenSteerWrapperDynParams {
    float elapsedTime = 0.0f;
};

/// BehaviorClusteringParams (PlugIns/BehaviorClustering/include/BehaviorClustering_kernel.cuh:40-57):
/// the fields the feature update and the cluster steering read
struct BehaviorClusteringParams {
    int featuresCount = 4;
    float featureUpdateRate = 0.0f;
    float3 clusteringForceParams = make_float3(0, 0, 0);           // .x = weight
    float3 alignmentClusteringForceParams = make_float3(0, 0, 0);  // .x = weight
    float3 similarityStatsParams = make_float3(0, 0, 0);           // avg rate, min/max rate, memory
    float avgSimilarityController = 1.0f;
    float sameClusterMaxRadius = 1.5f;
    int iterationLP = 10;  // local propagation's iterations
};

inline Body3DParams hBody3DParams;
inline EnvGrid3DParams hEnvGrid3DParams;
inline Proximity3DParams hProximity3DParams;
inline OpenSteerWrapperParams hOpenSteerWrapperParams;
inline OpenSteerWrapperDynParams hOpenSteerWrapperDynParams;
inline BehaviorClusteringParams hBehaviorClusteringParams;

namespace BehaveRT {

typedef std::vector<std::string> DependenciesList;
typedef unsigned int uint;

class PlugInDependencyError : public std::runtime_error {
public:
    explicit PlugInDependencyError(const std::string& m) : std::runtime_error(m) {}
};

/// Global access point: device interface, configuration, installed plug-ins.
class CommonResources {
public:
    CommonResources() : m_DeviceInterface(new DeviceInterface(0)), m_Config(new UtilityConfigFile()) {}

    /// Loads the configuration (CommonResources::initialize).
    void initialize(const std::string& configFileName) {
        if (!configFileName.empty() && !m_Config->load(configFileName))
            throw std::runtime_error("configuration file not found: " + configFileName);
        m_DeviceInterface->getHostSimParams().commonBlockDim =
            StringConverter::parseInt(m_Config->getSetting("BlockDim", "DeviceInterface"));
        if (m_DeviceInterface->getHostSimParams().commonBlockDim <= 0)
            m_DeviceInterface->getHostSimParams().commonBlockDim = 256;
        m_DeviceInterface->getHostSimParams().useThreadSync =
            StringConverter::parseBool(m_Config->getSetting("useThreadSync", "SimParams"));
        m_Initialized = true;
    }
    bool isInitialized() const { return m_Initialized; }
    DeviceInterface* getDeviceInterface() { return m_DeviceInterface.get(); }
    UtilityConfigFile* getConfig() { return m_Config.get(); }

    void installPlugIn(const std::string& name) { m_Installed.push_back(name); }
    void uninstallPlugIn(const std::string& name) {
        auto it = std::find(m_Installed.rbegin(), m_Installed.rend(), name);
        if (it != m_Installed.rend()) m_Installed.erase(std::next(it).base());
    }
    const std::vector<std::string>& getInstalledPlugIns() const { return m_Installed; }

private:
    std::unique_ptr<DeviceInterface> m_DeviceInterface;
    std::unique_ptr<UtilityConfigFile> m_Config;
    std::vector<std::string> m_Installed;
    bool m_Initialized = false;
};

}  // namespace BehaveRT

inline BehaveRT::CommonResources m_CommonRes;

namespace BehaveRT {

/// Interface every plug-in level implements (AbstractSimEnginePlugIn).
class AbstractSimEnginePlugIn {
public:
    virtual ~AbstractSimEnginePlugIn() {}
    virtual const std::string name() = 0;
    virtual const DependenciesList plugInDependencies() = 0;
    virtual void reset() = 0;
    virtual void update(const float elapsedTime) = 0;
};

/// Installation of one level of the chain: install(), registration, then the
/// dependency check (SimEnginePlugIn::installPlugIn).
struct PlugInInstaller {
    template <class P>
    static void install(P* p) {
        p->P::install();
        m_CommonRes.installPlugIn(p->P::name());
        const DependenciesList deps = p->P::plugInDependencies();
        const std::vector<std::string>& installed = m_CommonRes.getInstalledPlugIns();
        for (const std::string& d : deps)
            if (std::find(installed.begin(), installed.end(), d) == installed.end())
                throw PlugInDependencyError("missing plug-in dependency [" + d + "] of [" + p->P::name() + "]");
    }
    template <class P>
    static void uninstall(P* p) {
        p->P::uninstall();
        m_CommonRes.uninstallPlugIn(p->P::name());
    }
};

/// Root of every chain (DummyPlugIn).
class DummyPlugIn : public AbstractSimEnginePlugIn {
public:
    const std::string name() override { return "PlugInRoot"; }
    const DependenciesList plugInDependencies() override { return DependenciesList(); }
    void reset() override {}
    void update(const float) override {}

protected:
    std::string getPlugInSetting(const std::string& section, const std::string& key) {
        return m_CommonRes.getConfig()->getSetting(key, section);
    }
};

}  // namespace BehaveRT

using BehaveRT::DummyPlugIn;
