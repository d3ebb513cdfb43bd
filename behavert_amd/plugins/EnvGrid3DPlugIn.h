// EnvGrid3DPlugIn.h -- world, uniform grid and per-step reorder of the agents.
//
// Mirrors EnvGrid3DPlugIn (PlugIns/EnvGrid3D/include/EnvGrid3DPlugIn.h:47-571):
//   install ........ position + hash features, world radius, grid, bodies/cell
//                    (:155-211, setupWorldRadius :420-473, setupGrid :487-508,
//                    setupMaxBodiesPerCell :536-562)
//   reset .......... random positions in the world box (:224-259)
//   reorderSimData . hash -> stable sort -> gather of every registered float4
//                    feature (:292-415), here brt_envgrid_hash_sort +
//                    brt_envgrid_reorder on the gfx950 kernels.
// Config section [EnvGrid3DPlugIn]: cellSize, disableSorting, worldCenter,
// lockWorldProportions, keepDensity, worldRadius, bodiesDensity2D/3D,
// useAutomaticMaxBodiesPerCell, maxBodiesPerCell.  A missing or zero
// maxBodiesPerCell means "no per-cell cap" (the reference would scan no agent).
#pragma once

#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "Body3DPlugIn.h"

namespace EnvGrid3D {

template <class Super>
class EnvGrid3DPlugIn : public Super {
public:
    EnvGrid3DPlugIn() { BehaveRT::PlugInInstaller::install(this); }
    ~EnvGrid3DPlugIn() { BehaveRT::PlugInInstaller::uninstall(this); }

    const std::string name() override { return "EnvGrid3DPlugIn"; }
    const BehaveRT::DependenciesList plugInDependencies() override {
        return BehaveRT::DependenciesList{"Body3DPlugIn"};
    }

    void reset() override {
        Super::reset();
        std::mt19937 gen(m_Seed);
        std::uniform_real_distribution<float> frand(0.0f, 1.0f);
        const EnvGrid3DParams& p = hEnvGrid3DParams;
        for (int i = 0; i < hBody3DParams.numBodies; ++i) {
            float4 v = make_float4((frand(gen) - 0.5f) * p.worldRadius.x * 2 + p.worldCenter.x,
                                   (frand(gen) - 0.5f) * p.worldRadius.y * 2 + p.worldCenter.y,
                                   (frand(gen) - 0.5f) * p.worldRadius.z * 2 + p.worldCenter.z, 1.0f);
            m_Pos->setHostArrayElement(i, &v);
        }
        m_Pos->copyArrayToDevice();
        m_Pos->swapPosReadWrite();
    }
    void update(const float elapsedTime) override { Super::update(elapsedTime); }

    /// Registers a float4 feature to be gathered into sorted order every step.
    void addToFeaturesToReorder(BehaveRT::DeviceArrayWrapper<float4>* feature) {
        feature->setInstallationKey((int)m_FeaturesToReorder.size());
        m_FeaturesToReorder.push_back(feature);
    }

    /// hash -> stable key/ID sort -> reorder of the registered features.
    void reorderSimData() {
        BehaveRT::DeviceInterface* di = m_CommonRes.getDeviceInterface();
        applyWorldProportions();
        const int n = hBody3DParams.numBodies;
        if (hEnvGrid3DParams.disableSorting) {
            di->check(brt_envgrid_calc_hash(di->context(), &m_Grid, (const float*)m_Pos->getReadDeviceArray(), n,
                                            (uint32_t*)m_AgentHash->getReadDeviceArray(), di->stream()),
                      "brt_envgrid_calc_hash");
            return;
        }
        di->check(brt_envgrid_hash_sort(di->context(), &m_Grid, (const float*)m_Pos->getReadDeviceArray(), n,
                                        (uint32_t*)m_AgentHash->getReadDeviceArray(), di->stream()),
                  "brt_envgrid_hash_sort");
        std::vector<const float*> in;
        std::vector<float*> out;
        for (auto* f : m_FeaturesToReorder) {
            in.push_back((const float*)f->getReadDeviceArray());
            out.push_back((float*)f->getWriteDeviceArray());
        }
        for (size_t first = 0; first < in.size(); first += 8) {
            const uint32_t cnt = (uint32_t)std::min<size_t>(8, in.size() - first);
            di->check(brt_envgrid_reorder(di->context(), (const uint32_t*)m_AgentHash->getReadDeviceArray(), n, cnt,
                                          in.data() + first, out.data() + first, di->stream()),
                      "brt_envgrid_reorder");
        }
        for (auto* f : m_FeaturesToReorder) f->swapPosReadWrite();
        if (m_CommonRes.getDeviceInterface()->getHostSimParams().useThreadSync) di->threadSync();
    }

    BehaveRT::DeviceArrayWrapper<float4>* getPos() { return m_Pos.get(); }
    BehaveRT::DeviceArrayWrapper<uint2>* getHash() { return m_AgentHash.get(); }
    BehaveRT::uint getSortedIndex(BehaveRT::uint index, bool copyFromDevice = true) {
        return m_AgentHash->getHostArrayElement((int)index, copyFromDevice).y;
    }
    float3 getWorldCenter() { return hEnvGrid3DParams.worldCenter; }
    const brt_grid_params& getGridParams() const { return m_Grid; }

protected:
    friend struct BehaveRT::PlugInInstaller;
    void install() {
        using BehaveRT::StringConverter;
        BehaveRT::UtilityConfigFile* cfg = m_CommonRes.getConfig();
        BehaveRT::DeviceInterface* di = m_CommonRes.getDeviceInterface();
        const int n = hBody3DParams.numBodies;
        m_Pos.reset(new BehaveRT::DeviceArrayWrapper<float4>(di, n));
        addToFeaturesToReorder(m_Pos.get());
        m_AgentHash.reset(new BehaveRT::DeviceArrayWrapper<uint2>(di, n, 1, false));
        const std::string sec = "EnvGrid3DPlugIn";
        EnvGrid3DParams& p = hEnvGrid3DParams;
        p.cellSize = StringConverter::parseFloat3(cfg->getSetting("cellSize", sec));
        p.disableSorting = StringConverter::parseBool(cfg->getSetting("disableSorting", sec));
        p.worldCenter = StringConverter::parseFloat3(cfg->getSetting("worldCenter", sec));
        p.lockWorldProportions = StringConverter::parseBool(cfg->getSetting("lockWorldProportions", sec));
        m_Seed = (unsigned)StringConverter::parseInt(cfg->getSetting("seed", sec));
        // setupWorldRadius
        if (!StringConverter::parseBool(cfg->getSetting("keepDensity", sec))) {
            p.worldRadius = StringConverter::parseFloat3(cfg->getSetting("worldRadius", sec));
            if (p.worldRadius.x == 0.0f) {
                const float r = StringConverter::parseFloat(cfg->getSetting("worldRadius", sec));
                p.worldRadius = make_float3(r, r, r);
            }
        } else if (hBody3DParams.use2DProjection) {
            const float d = StringConverter::parseFloat(cfg->getSetting("bodiesDensity2D", sec));
            const float r = std::pow(n / d, 0.5f) / 2;
            p.worldRadius = make_float3(r, r, r);
        } else {
            const float d = StringConverter::parseFloat(cfg->getSetting("bodiesDensity3D", sec));
            const float r = std::pow(n / d, 0.33333f) / 2;
            p.worldRadius = make_float3(r, r, r);
        }
        // setupMaxBodiesPerCell
        if (StringConverter::parseBool(cfg->getSetting("useAutomaticMaxBodiesPerCell", sec))) {
            const float cellVolume = p.cellSize.x * p.cellSize.y * p.cellSize.z;
            const float r = hBody3DParams.commonRadius;
            p.maxBodiesPerCell = (unsigned)(1 + cellVolume / (4.1866f * r * r * r));
        } else {
            const int m = StringConverter::parseInt(cfg->getSetting("maxBodiesPerCell", sec));
            p.maxBodiesPerCell = m > 0 ? (unsigned)m : BRT_UNLIMITED;
        }
        setupGrid();
    }
    void uninstall() {
        m_Pos.reset();
        m_AgentHash.reset();
        m_FeaturesToReorder.clear();
    }

    void setupGrid() {
        EnvGrid3DParams& p = hEnvGrid3DParams;
        const float c[3] = {p.worldCenter.x, p.worldCenter.y, p.worldCenter.z};
        const float r[3] = {p.worldRadius.x, p.worldRadius.y, p.worldRadius.z};
        const float s[3] = {p.cellSize.x, p.cellSize.y, p.cellSize.z};
        if (brt_grid_setup(c, r, s, p.maxBodiesPerCell, &m_Grid) != BRT_OK)
            throw std::runtime_error("EnvGrid3DPlugIn: invalid world/cell configuration");
        p.gridSize = make_uint3(m_Grid.grid_size[0], m_Grid.grid_size[1], m_Grid.grid_size[2]);
        p.numCells = m_Grid.num_cells;
        p.worldOrigin = make_float3(m_Grid.world_origin[0], m_Grid.world_origin[1], m_Grid.world_origin[2]);
    }

    // lockWorldProportions propagates worldRadius.x to y and z each step
    // (EnvGrid3DPlugIn.h:297-304)
    void applyWorldProportions() {
        EnvGrid3DParams& p = hEnvGrid3DParams;
        if (p.lockWorldProportions && (p.worldRadius.y != p.worldRadius.x || p.worldRadius.z != p.worldRadius.x)) {
            p.worldRadius.y = p.worldRadius.z = p.worldRadius.x;
            setupGrid();
        }
    }

    std::unique_ptr<BehaveRT::DeviceArrayWrapper<float4>> m_Pos;
    std::unique_ptr<BehaveRT::DeviceArrayWrapper<uint2>> m_AgentHash;
    std::vector<BehaveRT::DeviceArrayWrapper<float4>*> m_FeaturesToReorder;
    brt_grid_params m_Grid{};
    unsigned m_Seed = 0;
};

}  // namespace EnvGrid3D

using EnvGrid3D::EnvGrid3DPlugIn;
