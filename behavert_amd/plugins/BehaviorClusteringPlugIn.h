// BehaviorClusteringPlugIn.h -- per-agent feature vectors and their
// neighbour-averaged update.
//
// The reference plug-in (PlugIns/BehaviorClustering/include/BehaviorClusteringPlugIn.h)
// keeps a featuresCount-float vector per agent indexed by agent ID (:180-183)
// and walks the neighbour lists with the sorted->ID map
// (BehaviorClustering_kernel.cu:306-353).  This plug-in keeps that layout for a
// 4-D feature (float4 by ID, double-buffered) and provides the feature update
//   f'[id(i)] = f[id(i)] + rate * (mean_{j in list(i)} f[id(j)] - f[id(i)])
// on the gfx950 kernel (brt_bc_feature_update).  Config section
// [BehaviorClusteringPlugIn]: featuresCount (must be 4), featureUpdateRate.
#pragma once

#include <random>

#include "OpenSteerWrapperPlugIn.h"

namespace BehaviorClustering {

template <class Super>
class BehaviorClusteringPlugIn : public Super {
public:
    BehaviorClusteringPlugIn() { BehaveRT::PlugInInstaller::install(this); }
    ~BehaviorClusteringPlugIn() { BehaveRT::PlugInInstaller::uninstall(this); }

    const std::string name() override { return "BehaviorClusteringPlugIn"; }
    const BehaveRT::DependenciesList plugInDependencies() override {
        return BehaveRT::DependenciesList{"EnvGrid3DPlugIn", "Proximity3DPlugIn"};
    }
    void reset() override {
        Super::reset();
        std::mt19937 gen(7u);
        std::uniform_real_distribution<float> u(0.0f, 1.0f);
        for (int i = 0; i < hBody3DParams.numBodies; ++i) {
            float4 f = make_float4(u(gen), u(gen), u(gen), u(gen));
            m_FeaturesVector->setHostArrayElement(i, &f);
        }
        m_FeaturesVector->copyArrayToDevice();
        m_FeaturesVector->swapPosReadWrite();
    }
    void update(const float elapsedTime) override { Super::update(elapsedTime); }

    /// One neighbour-averaged update of every agent's feature vector.
    void computeFeatureAveraging() { computeFeatureAveraging(hBehaviorClusteringParams.featureUpdateRate); }
    void computeFeatureAveraging(float rate) {
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

    BehaveRT::DeviceArrayWrapper<float4>* getFeaturesVector() { return m_FeaturesVector.get(); }
    int getElementsNumber() const { return hBody3DParams.numBodies; }

protected:
    friend struct BehaveRT::PlugInInstaller;
    void install() {
        using BehaveRT::StringConverter;
        BehaveRT::UtilityConfigFile* cfg = m_CommonRes.getConfig();
        const std::string sec = "BehaviorClusteringPlugIn";
        const std::string fc = cfg->getSetting("featuresCount", sec);
        hBehaviorClusteringParams.featuresCount = fc.empty() ? 4 : StringConverter::parseInt(fc);
        if (hBehaviorClusteringParams.featuresCount != 4)
            throw std::runtime_error("BehaviorClusteringPlugIn: the feature update supports featuresCount = 4");
        hBehaviorClusteringParams.featureUpdateRate =
            StringConverter::parseFloat(cfg->getSetting("featureUpdateRate", sec));
        m_FeaturesVector.reset(
            new BehaveRT::DeviceArrayWrapper<float4>(m_CommonRes.getDeviceInterface(), hBody3DParams.numBodies));
    }
    void uninstall() { m_FeaturesVector.reset(); }

    std::unique_ptr<BehaveRT::DeviceArrayWrapper<float4>> m_FeaturesVector;
};

}  // namespace BehaviorClustering

using BehaviorClustering::BehaviorClusteringPlugIn;
