// OpenSteerWrapperPlugIn.h -- flocking behaviours and the OpenSteer integrator.
//
// Mirrors OpenSteerWrapperPlugIn (PlugIns/OpenSteerWrapper/include/OpenSteerWrapperPlugIn.h:54-441):
//   install ...... behaviour params, smoothAccel/forward (both reordered each
//                  step), up (unused), steerForce (:129-181)
//   reset ........ random unit headings with speed commonMaxSpeed, zero accel (:194-223)
//   resetSteerForce, computeSeparations/Alignments/Cohesions,
//   computeSeekingsWorldCenter, applySteeringForces(dt) (:287-424)
// The reference runs one kernel per behaviour and reads/writes a double-buffered
// force each time.  Here the compute* calls record which behaviours the step
// uses and applySteeringForces runs them all in one gfx950 pass
// (brt_osw_steer_integrate): same behaviours, same order of the force sum
// (separation, alignment, cohesion, [targets], [wander], [world centre] --
// the order of BehaveRT_Flocking::update), same integrator.  A behaviour whose
// weight is 0 is skipped, as in the reference.  The summed force is written to
// steerForce for readers.
// A plug-in that adds its own force between the behaviours and the world-centre
// seek (BehaviorClustering's steerForClustering, BehaveRT_Clustering.cpp:212-235)
// calls materializeSteerForce(): the recorded behaviours then run as their own
// pass (brt_osw_steer_forces) into steerForce, the plug-in sums onto it, and
// applySteeringForces integrates the result (brt_osw_apply_steering_force).
// Extensions with no reference counterpart: setSeekTargets (multi-target seek)
// and setWander (OpenSteer steerForWander driven by a counter-based RNG).
// Config section [OpenSteerWrapperPlugIn]: separationParams, cohesionParams,
// alignmentParams ("weight radius cosMaxAngle"), commonMass (missing -> 1),
// commonMaxSpeed, commonMaxForce, commonSensingMinRadius, avoidBaseParams.
#pragma once

#include <cmath>
#include <cstring>
#include <random>

#include "Proximity3DPlugIn.h"

namespace OpenSteerWrapper {

template <class Super>
class OpenSteerWrapperPlugIn : public Super {
public:
    OpenSteerWrapperPlugIn() { BehaveRT::PlugInInstaller::install(this); }
    ~OpenSteerWrapperPlugIn() { BehaveRT::PlugInInstaller::uninstall(this); }

    const std::string name() override { return "OpenSteerWrapperPlugIn"; }
    const BehaveRT::DependenciesList plugInDependencies() override {
        return BehaveRT::DependenciesList{"EnvGrid3DPlugIn", "Proximity3DPlugIn"};
    }

    void reset() override {
        Super::reset();
        std::mt19937 gen(m_Seed + 1);
        std::normal_distribution<float> nrm(0.0f, 1.0f);
        for (int i = 0; i < hBody3DParams.numBodies; ++i) {
            float x = nrm(gen), y = nrm(gen), z = nrm(gen);
            float l = std::sqrt(x * x + y * y + z * z);
            if (l == 0.0f) { x = 1.0f; l = 1.0f; }
            float4 f = make_float4(x / l, y / l, z / l, hOpenSteerWrapperParams.commonMaxSpeed);
            float4 zero = make_float4(0, 0, 0, 0);
            m_Forward->setHostArrayElement(i, &f);
            m_SmoothedAcceleration->setHostArrayElement(i, &zero);
        }
        m_Forward->copyArrayToDevice();
        m_Forward->swapPosReadWrite();
        m_SmoothedAcceleration->copyArrayToDevice();
        m_SmoothedAcceleration->swapPosReadWrite();
        m_Step = 0;
        m_Pending = 0;
        m_ForceMaterialized = false;
    }
    void update(const float elapsedTime) override { Super::update(elapsedTime); }

    void resetSteerForce() {
        m_Pending = 0;
        m_ForceMaterialized = false;
    }
    void computeSeparations() {
        if (hOpenSteerWrapperParams.separationParams.x != 0.0f) m_Pending |= kSep;
    }
    void computeAlignments() {
        if (hOpenSteerWrapperParams.alignmentParams.x != 0.0f) m_Pending |= kAli;
    }
    void computeCohesions() {
        if (hOpenSteerWrapperParams.cohesionParams.x != 0.0f) m_Pending |= kCoh;
    }
    void computeSeekingsWorldCenter() { m_Pending |= kSeekCenter; }

    /// Multi-target seek: agent i steers toward targets[i % count] with weight w.
    void setSeekTargets(const float3* targets, int count, float weight) {
        m_NumTargets = count < 0 ? 0 : (count > BRT_MAX_TARGETS ? BRT_MAX_TARGETS : count);
        for (int k = 0; k < m_NumTargets; ++k) m_Targets[k] = targets[k];
        m_TargetWeight = weight;
    }
    /// Wander with weight w (0 disables); per-agent side/up walk state by ID.
    void setWander(float weight, uint32_t seed) {
        m_WanderWeight = weight;
        m_WanderSeed = seed;
    }

    /// Runs the behaviours recorded so far as their own pass: steerForce (read
    /// buffer, sorted order) then holds their sum, ready for further forces.
    void materializeSteerForce(float elapsedTime) {
        if (m_ForceMaterialized) return;
        BehaveRT::DeviceInterface* di = m_CommonRes.getDeviceInterface();
        prepareWander(di);
        const brt_steer_params sp = steerParams(m_Pending & ~kSeekCenter);
        const brt_prox_params pp = this->proxParams();
        di->check(brt_osw_steer_forces(di->context(), &pp, &sp, elapsedTime, m_Step, hBody3DParams.numBodies,
                                       (const float*)this->getPos()->getReadDeviceArray(),
                                       (const float*)m_Forward->getReadDeviceArray(),
                                       (const uint32_t*)this->getHash()->getReadDeviceArray(),
                                       (const uint32_t*)this->getNeighList()->getReadDeviceArray(),
                                       m_Wander ? (float*)m_Wander->getReadDeviceArray() : nullptr,
                                       (float*)m_SteerForce->getWriteDeviceArray(), di->stream()),
                  "brt_osw_steer_forces");
        m_SteerForce->swapPosReadWrite();
        m_Pending &= kSeekCenter;
        m_ForceMaterialized = true;
    }

    /// All recorded behaviours + applySteeringForceSingle in one pass; the new
    /// position, forward and smoothed acceleration land by agent ID.
    void applySteeringForces(float elapsedTime) {
        BehaveRT::DeviceInterface* di = m_CommonRes.getDeviceInterface();
        hOpenSteerWrapperDynParams.elapsedTime = elapsedTime;
        const brt_grid_params& g = this->getGridParams();
        if (m_ForceMaterialized) {
            if (m_Pending & (kSep | kAli | kCoh))
                throw std::logic_error("OpenSteerWrapperPlugIn: behaviours recorded after the steer force was "
                                       "materialized");
            const brt_steer_params sp = steerParams(m_Pending);
            di->check(brt_osw_apply_steering_force(
                          di->context(), &g, &sp, elapsedTime, hBody3DParams.numBodies,
                          (const float*)this->getPos()->getReadDeviceArray(),
                          (const float*)m_Forward->getReadDeviceArray(),
                          (const float*)m_SmoothedAcceleration->getReadDeviceArray(),
                          (const uint32_t*)this->getHash()->getReadDeviceArray(),
                          (const float*)m_SteerForce->getReadDeviceArray(),
                          (float*)this->getPos()->getWriteDeviceArray(), (float*)m_Forward->getWriteDeviceArray(),
                          (float*)m_SmoothedAcceleration->getWriteDeviceArray(),
                          (float*)m_SteerForce->getWriteDeviceArray(), di->stream()),
                      "brt_osw_apply_steering_force");
        } else {
            prepareWander(di);
            const brt_steer_params sp = steerParams(m_Pending);
            const brt_prox_params pp = this->proxParams();
            di->check(brt_osw_steer_integrate(
                          di->context(), &g, &pp, &sp, elapsedTime, m_Step, hBody3DParams.numBodies,
                          (const float*)this->getPos()->getReadDeviceArray(),
                          (const float*)m_Forward->getReadDeviceArray(),
                          (const float*)m_SmoothedAcceleration->getReadDeviceArray(),
                          (const uint32_t*)this->getHash()->getReadDeviceArray(),
                          (const uint32_t*)this->getNeighList()->getReadDeviceArray(),
                          m_Wander ? (float*)m_Wander->getReadDeviceArray() : nullptr,
                          (float*)this->getPos()->getWriteDeviceArray(), (float*)m_Forward->getWriteDeviceArray(),
                          (float*)m_SmoothedAcceleration->getWriteDeviceArray(),
                          (float*)m_SteerForce->getWriteDeviceArray(), di->stream()),
                      "brt_osw_steer_integrate");
        }
        this->getPos()->swapPosReadWrite();
        m_SmoothedAcceleration->swapPosReadWrite();
        m_Forward->swapPosReadWrite();
        m_SteerForce->swapPosReadWrite();
        ++m_Step;
        m_Pending = 0;
        m_ForceMaterialized = false;
        if (di->getHostSimParams().useThreadSync) di->threadSync();
    }

    BehaveRT::DeviceArrayWrapper<float4>* getForward() { return m_Forward.get(); }
    BehaveRT::DeviceArrayWrapper<float4>* getSmoothedAcceleration() { return m_SmoothedAcceleration.get(); }
    BehaveRT::DeviceArrayWrapper<float4>* getSteerForce() { return m_SteerForce.get(); }

protected:
    friend struct BehaveRT::PlugInInstaller;
    enum : unsigned { kSep = 1, kAli = 2, kCoh = 4, kSeekCenter = 8 };

    /// The C ABI's steering parameters for the behaviours in `pending`.
    brt_steer_params steerParams(unsigned pending) const {
        const OpenSteerWrapperParams& p = hOpenSteerWrapperParams;
        brt_steer_params sp;
        std::memset(&sp, 0, sizeof(sp));
        const float3* b[3] = {&p.separationParams, &p.alignmentParams, &p.cohesionParams};
        float* dst[3] = {sp.separation, sp.alignment, sp.cohesion};
        const unsigned bits[3] = {kSep, kAli, kCoh};
        for (int k = 0; k < 3; ++k) {
            dst[k][0] = (pending & bits[k]) ? b[k]->x : 0.0f;
            dst[k][1] = b[k]->y;
            dst[k][2] = b[k]->z;
        }
        sp.mass = p.commonMass;
        sp.max_speed = p.commonMaxSpeed;
        sp.max_force = p.commonMaxForce;
        sp.sensing_min_radius = p.commonSensingMinRadius;
        sp.seek_world_center = (pending & kSeekCenter) ? 1u : 0u;
        sp.use_2d_projection = hBody3DParams.use2DProjection ? 1u : 0u;
        sp.num_targets = (uint32_t)m_NumTargets;
        sp.target_weight = m_TargetWeight;
        for (int k = 0; k < m_NumTargets; ++k) {
            sp.targets[k][0] = m_Targets[k].x;
            sp.targets[k][1] = m_Targets[k].y;
            sp.targets[k][2] = m_Targets[k].z;
        }
        sp.wander_weight = m_WanderWeight;
        sp.wander_seed = m_WanderSeed;
        return sp;
    }
    void prepareWander(BehaveRT::DeviceInterface* di) {
        if (m_WanderWeight != 0.0f && !m_Wander) {
            m_Wander.reset(new BehaveRT::DeviceArrayWrapper<float2>(di, hBody3DParams.numBodies, 1, false));
            m_Wander->reset(0);
        }
    }

    void install() {
        using BehaveRT::StringConverter;
        BehaveRT::UtilityConfigFile* cfg = m_CommonRes.getConfig();
        BehaveRT::DeviceInterface* di = m_CommonRes.getDeviceInterface();
        const std::string sec = "OpenSteerWrapperPlugIn";
        OpenSteerWrapperParams& p = hOpenSteerWrapperParams;
        p.separationParams = StringConverter::parseFloat3(cfg->getSetting("separationParams", sec));
        p.cohesionParams = StringConverter::parseFloat3(cfg->getSetting("cohesionParams", sec));
        p.alignmentParams = StringConverter::parseFloat3(cfg->getSetting("alignmentParams", sec));
        const std::string mass = cfg->getSetting("commonMass", sec);
        p.commonMass = mass.empty() ? 1.0f : StringConverter::parseFloat(mass);
        p.commonMaxSpeed = StringConverter::parseFloat(cfg->getSetting("commonMaxSpeed", sec));
        p.commonMaxForce = StringConverter::parseFloat(cfg->getSetting("commonMaxForce", sec));
        p.commonSensingMinRadius = StringConverter::parseFloat(cfg->getSetting("commonSensingMinRadius", sec));
        p.avoidBaseParams = StringConverter::parseFloat3(cfg->getSetting("avoidBaseParams", sec));
        if (p.avoidBaseParams.x == 0.0f) p.avoidBaseParams = make_float3(500.0f, 500.0f, 1000.0f);
        m_Seed = (unsigned)StringConverter::parseInt(cfg->getSetting("seed", "EnvGrid3DPlugIn"));
        const int n = hBody3DParams.numBodies;
        m_SmoothedAcceleration.reset(new BehaveRT::DeviceArrayWrapper<float4>(di, n));
        this->addToFeaturesToReorder(m_SmoothedAcceleration.get());
        m_Forward.reset(new BehaveRT::DeviceArrayWrapper<float4>(di, n));
        m_Up.reset(new BehaveRT::DeviceArrayWrapper<float4>(di, n));
        this->addToFeaturesToReorder(m_Forward.get());
        m_SteerForce.reset(new BehaveRT::DeviceArrayWrapper<float4>(di, n));
    }
    void uninstall() {
        m_Forward.reset();
        m_Up.reset();
        m_SteerForce.reset();
        m_SmoothedAcceleration.reset();
        m_Wander.reset();
    }

    std::unique_ptr<BehaveRT::DeviceArrayWrapper<float4>> m_Forward, m_Up, m_SteerForce, m_SmoothedAcceleration;
    std::unique_ptr<BehaveRT::DeviceArrayWrapper<float2>> m_Wander;
    unsigned m_Pending = 0;
    bool m_ForceMaterialized = false;
    uint32_t m_Step = 0;
    unsigned m_Seed = 0;
    int m_NumTargets = 0;
    float3 m_Targets[BRT_MAX_TARGETS];
    float m_TargetWeight = 0.0f;
    float m_WanderWeight = 0.0f;
    uint32_t m_WanderSeed = 0;
};

}  // namespace OpenSteerWrapper

using OpenSteerWrapper::OpenSteerWrapperPlugIn;
