// BehaviorClusteringPlugIn.h -- per-agent feature vectors, their
// neighbour-averaged update and the similarity-driven cluster steering.
//
// The reference plug-in (PlugIns/BehaviorClustering/include/BehaviorClusteringPlugIn.h)
// keeps a featuresCount-float vector per agent indexed by agent ID (:180-183),
// per-agent similarity statistics (min, avg, max) initialised to (1, 0.5, 0)
// (:185-187, :594-599) and one similarity per neighbour-list slot (:223-226).
// This plug-in keeps those layouts:
//   features ........ float[featuresCount] by ID (single buffer)
//   similarityStats . float4 by ID (double-buffered: read -> write, swap)
//   neighSimilarities float[maxNeighbors + 1] per sorted agent
// and provides
//   computeFeatureAveraging (featuresCount 4 only):
//     f'[id(i)] = f[id(i)] + rate * (mean_{j in list(i)} f[id(j)] - f[id(i)])
//   steerForClustering (:740-817): naive + adaptive similarities, then cluster
//     cohesion and alignment summed onto the steer force, between the flocking
//     behaviours and the world-centre seek (BehaveRT_Clustering.cpp:212-235).
// Config section [BehaviorClusteringPlugIn]: featuresCount (1..32, default 4),
// featureUpdateRate, clusteringForceParams and alignmentClusteringForceParams
// (".x" is the weight, :160-161, BehaviorClustering_resources.cu:339,380),
// similarityStatsParams (three rates, BehaviorClustering_resources.cu:84-150).
// avgSimilarityController is 1, as the reference's install() sets it (:164).
//   computeClusterMembership / analyzeClusterHitRate (:823-896): local label
//     propagation (iterationLP = 10) with its connector, and the hit rate
//     (sameClusterMaxRadius = 1.5), clusterMembership float3 by ID.
// The reference's dataset loader, colours and merge/freeze tools are not part
// of this layer; reset() draws features uniformly in [0, 1).
#pragma once

#include <random>
#include <stdexcept>

#include "OpenSteerWrapperPlugIn.h"

namespace BehaviorClustering {

template <class Super>
class BehaviorClusteringPlugIn : public Super {
public:
    BehaviorClusteringPlugIn() { BehaveRT::PlugInInstaller::install(this); }
    ~BehaviorClusteringPlugIn() { BehaveRT::PlugInInstaller::uninstall(this); }

    const std::string name() override { return "BehaviorClusteringPlugIn"; }
    const BehaveRT::DependenciesList plugInDependencies() override {
        return BehaveRT::DependenciesList{"EnvGrid3DPlugIn", "Proximity3DPlugIn", "OpenSteerWrapperPlugIn"};
    }
    void reset() override {
        Super::reset();
        std::mt19937 gen(7u);
        std::uniform_real_distribution<float> u(0.0f, 1.0f);
        const int F = hBehaviorClusteringParams.featuresCount;
        std::vector<float> f((size_t)F);
        const float4 st = make_float4(1.0f, 0.5f, 0.0f, 0.0f);
        for (int i = 0; i < hBody3DParams.numBodies; ++i) {
            for (int k = 0; k < F; ++k) f[(size_t)k] = u(gen);
            m_FeaturesVector->setHostArrayElement(i, f.data());
            m_SimilarityStats->setHostArrayElement(i, &st);
        }
        m_FeaturesVector->copyArrayToDevice();
        m_FeaturesVector->swapPosReadWrite();
        m_SimilarityStats->copyArrayToDevice();
        m_SimilarityStats->swapPosReadWrite();
    }
    void update(const float elapsedTime) override { Super::update(elapsedTime); }

    /// One neighbour-averaged update of every agent's feature vector.
    void computeFeatureAveraging() { computeFeatureAveraging(hBehaviorClusteringParams.featureUpdateRate); }
    void computeFeatureAveraging(float rate) {
        if (hBehaviorClusteringParams.featuresCount != 4)
            throw std::runtime_error("BehaviorClusteringPlugIn: the feature update supports featuresCount = 4");
        BehaveRT::DeviceInterface* di = m_CommonRes.getDeviceInterface();
        const brt_prox_params pp = this->proxParams();
        di->check(brt_bc_feature_update(di->context(), &pp, rate, hBody3DParams.numBodies,
                                        (const uint32_t*)this->getHash()->getReadDeviceArray(),
                                        (const uint32_t*)this->getNeighList()->getReadDeviceArray(),
                                        (const float*)m_FeaturesVector->getReadDeviceArray(),
                                        (float*)m_FeaturesVector->getWriteDeviceArray(), di->stream()),
                  "brt_bc_feature_update");
        m_FeaturesVector->swapPosReadWrite();
        if (di->getHostSimParams().useThreadSync) di->threadSync();
    }

    /// Similarity-weighted cluster cohesion and alignment, summed onto the
    /// steer force the flocking behaviours recorded so far produce.  The
    /// behaviours' pass uses the elapsed time of the previous
    /// applySteeringForces (it only matters to wander); the overload takes it.
    void steerForClustering() { steerForClustering(hOpenSteerWrapperDynParams.elapsedTime); }
    void steerForClustering(float elapsedTime) {
        BehaveRT::DeviceInterface* di = m_CommonRes.getDeviceInterface();
        this->materializeSteerForce(elapsedTime);
        const brt_prox_params pp = this->proxParams();
        const brt_bc_params bc = bcParams();
        const brt_steer_params sp = this->steerParams(0u);
        // similarities, adaptive statistics and cluster cohesion + alignment in one pass
        di->check(brt_bc_similarities_steer(di->context(), &pp, &sp, &bc, hBody3DParams.numBodies,
                                            (const uint32_t*)this->getHash()->getReadDeviceArray(),
                                            (const uint32_t*)this->getNeighList()->getReadDeviceArray(),
                                            m_FeaturesVector->getReadDeviceArray(),
                                            (const float*)m_SimilarityStats->getReadDeviceArray(),
                                            (float*)m_SimilarityStats->getWriteDeviceArray(),
                                            m_NeighSimilarities->getReadDeviceArray(),
                                            (const float*)this->getPos()->getReadDeviceArray(),
                                            (const float*)this->getForward()->getReadDeviceArray(),
                                            (const float*)this->getSteerForce()->getReadDeviceArray(),
                                            (float*)this->getSteerForce()->getWriteDeviceArray(), di->stream()),
                  "brt_bc_similarities_steer");
        m_SimilarityStats->swapPosReadWrite();
        this->getSteerForce()->swapPosReadWrite();
        if (di->getHostSimParams().useThreadSync) di->threadSync();
    }

    /// Local propagation of cluster labels (BehaviorClusteringPlugIn.h:823-865):
    /// reset to sorted indices, iterationLP min-label iterations, connector.
    void computeClusterMembership() {
        BehaveRT::DeviceInterface* di = m_CommonRes.getDeviceInterface();
        const brt_prox_params pp = this->proxParams();
        di->check(brt_bc_cluster_membership(di->context(), &pp, hBody3DParams.numBodies,
                                            (uint32_t)hBehaviorClusteringParams.iterationLP,
                                            (const uint32_t*)this->getHash()->getReadDeviceArray(),
                                            (const uint32_t*)this->getNeighList()->getReadDeviceArray(),
                                            (const float*)this->getPos()->getReadDeviceArray(),
                                            m_ClusterMembership->getReadDeviceArray(),
                                            m_ClusterMembership->getWriteDeviceArray(),
                                            (float*)m_CMConnector->getReadDeviceArray(), di->stream()),
                  "brt_bc_cluster_membership");
        m_ClusterMembership->swapPosReadWrite();
        if (di->getHostSimParams().useThreadSync) di->threadSync();
    }
    /// Per-agent hit rate and mean label of the close neighbours
    /// (analyzeClusterHitRate, BehaviorClusteringPlugIn.h:884-896).
    void analyzeClusterHitRate() {
        BehaveRT::DeviceInterface* di = m_CommonRes.getDeviceInterface();
        const brt_prox_params pp = this->proxParams();
        di->check(brt_bc_cluster_hit_rate(di->context(), &pp, hBehaviorClusteringParams.sameClusterMaxRadius,
                                          hBody3DParams.numBodies,
                                          (const uint32_t*)this->getHash()->getReadDeviceArray(),
                                          (const uint32_t*)this->getNeighList()->getReadDeviceArray(),
                                          (const float*)this->getPos()->getReadDeviceArray(),
                                          m_ClusterMembership->getReadDeviceArray(),
                                          m_ClusterMembership->getWriteDeviceArray(), di->stream()),
                  "brt_bc_cluster_hit_rate");
        m_ClusterMembership->swapPosReadWrite();
        if (di->getHostSimParams().useThreadSync) di->threadSync();
    }

    brt_bc_params bcParams() const {
        const BehaviorClusteringParams& p = hBehaviorClusteringParams;
        brt_bc_params bc;
        bc.feature_count = (uint32_t)p.featuresCount;
        bc.adaptive = 1u;
        bc.avg_similarity_controller = p.avgSimilarityController;
        bc.stats_params[0] = p.similarityStatsParams.x;
        bc.stats_params[1] = p.similarityStatsParams.y;
        bc.stats_params[2] = p.similarityStatsParams.z;
        bc.cluster_weight = p.clusteringForceParams.x;
        bc.alignment_weight = p.alignmentClusteringForceParams.x;
        return bc;
    }

    BehaveRT::DeviceArrayWrapper<float>* getFeaturesVector() { return m_FeaturesVector.get(); }
    BehaveRT::DeviceArrayWrapper<float4>* getSimilarityStats() { return m_SimilarityStats.get(); }
    BehaveRT::DeviceArrayWrapper<float>* getNeighSimilarities() { return m_NeighSimilarities.get(); }
    /// float3 (label, hit rate, mean close label) by ID
    BehaveRT::DeviceArrayWrapper<float>* getClusterMembership() { return m_ClusterMembership.get(); }
    BehaveRT::DeviceArrayWrapper<float4>* getCMConnector() { return m_CMConnector.get(); }
    int getElementsNumber() const { return hBody3DParams.numBodies; }

protected:
    friend struct BehaveRT::PlugInInstaller;
    void install() {
        using BehaveRT::StringConverter;
        BehaveRT::UtilityConfigFile* cfg = m_CommonRes.getConfig();
        BehaveRT::DeviceInterface* di = m_CommonRes.getDeviceInterface();
        const std::string sec = "BehaviorClusteringPlugIn";
        BehaviorClusteringParams& p = hBehaviorClusteringParams;
        const std::string fc = cfg->getSetting("featuresCount", sec);
        p.featuresCount = fc.empty() ? 4 : StringConverter::parseInt(fc);
        if (p.featuresCount < 1 || p.featuresCount > BRT_MAX_FEATURES)
            throw std::runtime_error("BehaviorClusteringPlugIn: featuresCount must be in 1..32");
        p.featureUpdateRate = StringConverter::parseFloat(cfg->getSetting("featureUpdateRate", sec));
        p.clusteringForceParams = StringConverter::parseFloat3(cfg->getSetting("clusteringForceParams", sec));
        p.alignmentClusteringForceParams =
            StringConverter::parseFloat3(cfg->getSetting("alignmentClusteringForceParams", sec));
        p.similarityStatsParams = StringConverter::parseFloat3(cfg->getSetting("similarityStatsParams", sec));
        p.avgSimilarityController = 1.0f;
        const int n = hBody3DParams.numBodies;
        m_FeaturesVector.reset(new BehaveRT::DeviceArrayWrapper<float>(di, n, p.featuresCount));
        m_SimilarityStats.reset(new BehaveRT::DeviceArrayWrapper<float4>(di, n));
        m_NeighSimilarities.reset(new BehaveRT::DeviceArrayWrapper<float>(
            di, n, (int)hProximity3DParams.maxNeighbors + 1, false));
        // iterationLP and sameClusterMaxRadius are fixed by the reference's install (:166, :683)
        p.iterationLP = 10;
        p.sameClusterMaxRadius = 1.5f;
        m_ClusterMembership.reset(new BehaveRT::DeviceArrayWrapper<float>(di, n, 3));
        m_ClusterMembership->reset(0);
        m_CMConnector.reset(new BehaveRT::DeviceArrayWrapper<float4>(di, 2 * n, 1, false));
    }
    void uninstall() {
        m_FeaturesVector.reset();
        m_SimilarityStats.reset();
        m_NeighSimilarities.reset();
        m_ClusterMembership.reset();
        m_CMConnector.reset();
    }

    std::unique_ptr<BehaveRT::DeviceArrayWrapper<float>> m_FeaturesVector;
    std::unique_ptr<BehaveRT::DeviceArrayWrapper<float4>> m_SimilarityStats;
    std::unique_ptr<BehaveRT::DeviceArrayWrapper<float>> m_NeighSimilarities;
    std::unique_ptr<BehaveRT::DeviceArrayWrapper<float>> m_ClusterMembership;
    std::unique_ptr<BehaveRT::DeviceArrayWrapper<float4>> m_CMConnector;
};

}  // namespace BehaviorClustering

using BehaviorClustering::BehaviorClusteringPlugIn;
