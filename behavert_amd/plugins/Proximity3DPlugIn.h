// Proximity3DPlugIn.h -- cell ranges and capped 27-cell neighbour lists.
//
// Mirrors Proximity3DPlugIn (PlugIns/Proximity3D/include/Proximity3DPlugIn.h:47-308):
//   install .............. maxNeighbors, useKnn, search radius, cellStart and the
//                          neighbour-list feature of numNeighWordsPerAgent = K/4+1
//                          uint4 words per agent (:114-193)
//   computeNeighborhoods . cell starts + lists (:260-308), here
//                          brt_prox_cell_ranges + brt_prox_neighbors (gfx950).
// Adds cellEnd (the reference finds a cell's end by key mismatch).  Config
// section [Proximity3DPlugIn]: searchDepth (ignored, as in the reference),
// useKnn, maxNeighbors, customBlockSize (ignored: launch shapes are the
// library's), useCellSizeAsSearchRadius, commonSearchRadius.
#pragma once

#include "EnvGrid3DPlugIn.h"

namespace Proximity3D {

template <class Super>
class Proximity3DPlugIn : public Super {
public:
    Proximity3DPlugIn() { BehaveRT::PlugInInstaller::install(this); }
    ~Proximity3DPlugIn() { BehaveRT::PlugInInstaller::uninstall(this); }

    const std::string name() override { return "Proximity3DPlugIn"; }
    const BehaveRT::DependenciesList plugInDependencies() override {
        return BehaveRT::DependenciesList{"Body3DPlugIn", "EnvGrid3DPlugIn"};
    }
    void reset() override {
        Super::reset();
        m_CellStart->reset((char)0xFF);
        m_CellEnd->reset((char)0xFF);
    }
    void update(const float elapsedTime) override { Super::update(elapsedTime); }

    /// Cell ranges of the sorted hash, then the neighbour lists of every agent.
    void computeNeighborhoods() {
        BehaveRT::DeviceInterface* di = m_CommonRes.getDeviceInterface();
        if (m_UseCellSizeAsSearchRadius) hProximity3DParams.commonSearchRadius = hEnvGrid3DParams.cellSize.x;
        const brt_grid_params& g = this->getGridParams();
        const int n = hBody3DParams.numBodies;
        const uint32_t* hash = (const uint32_t*)this->getHash()->getReadDeviceArray();
        di->check(brt_prox_cell_ranges(di->context(), hash, n, g.num_cells, m_CellStart->getReadDeviceArray(),
                                       m_CellEnd->getReadDeviceArray(), di->stream()),
                  "brt_prox_cell_ranges");
        brt_prox_params p = proxParams();
        di->check(brt_prox_neighbors(di->context(), &g, &p, (const float*)this->getPos()->getReadDeviceArray(), hash,
                                     m_CellStart->getReadDeviceArray(), m_CellEnd->getReadDeviceArray(), n,
                                     (uint32_t*)m_NeighList->getWriteDeviceArray(), di->stream()),
                  "brt_prox_neighbors");
        m_NeighList->swapPosReadWrite();
        if (di->getHostSimParams().useThreadSync) di->threadSync();
    }

    BehaveRT::DeviceArrayWrapper<BehaveRT::uint>* getCellStart() { return m_CellStart.get(); }
    BehaveRT::DeviceArrayWrapper<BehaveRT::uint>* getCellEnd() { return m_CellEnd.get(); }
    BehaveRT::DeviceArrayWrapper<uint4>* getNeighList() { return m_NeighList.get(); }
    brt_prox_params proxParams() const {
        brt_prox_params p;
        p.max_neighbors = hProximity3DParams.maxNeighbors;
        p.search_radius = hProximity3DParams.commonSearchRadius;
        p.use_knn = hProximity3DParams.useKnn ? 1u : 0u;
        return p;
    }

protected:
    friend struct BehaveRT::PlugInInstaller;
    void install() {
        using BehaveRT::StringConverter;
        BehaveRT::UtilityConfigFile* cfg = m_CommonRes.getConfig();
        BehaveRT::DeviceInterface* di = m_CommonRes.getDeviceInterface();
        const std::string sec = "Proximity3DPlugIn";
        Proximity3DParams& p = hProximity3DParams;
        p.searchDepth = StringConverter::parseUint3(cfg->getSetting("searchDepth", sec));
        p.useKnn = StringConverter::parseBool(cfg->getSetting("useKnn", sec));
        p.maxNeighbors = (unsigned)StringConverter::parseInt(cfg->getSetting("maxNeighbors", sec));
        if (p.maxNeighbors < 1 || p.maxNeighbors > BRT_MAX_NEIGHBORS)
            throw std::runtime_error("Proximity3DPlugIn: maxNeighbors must be in 1..252");
        p.numNeighWordsPerAgent = p.maxNeighbors / 4 + 1;
        m_UseCellSizeAsSearchRadius =
            StringConverter::parseBool(cfg->getSetting("useCellSizeAsSearchRadius", sec));
        p.commonSearchRadius = m_UseCellSizeAsSearchRadius
                                   ? hEnvGrid3DParams.cellSize.x
                                   : StringConverter::parseFloat(cfg->getSetting("commonSearchRadius", sec));
        const int cells = (int)hEnvGrid3DParams.numCells;
        m_CellStart.reset(new BehaveRT::DeviceArrayWrapper<BehaveRT::uint>(di, cells, 1, false));
        m_CellEnd.reset(new BehaveRT::DeviceArrayWrapper<BehaveRT::uint>(di, cells, 1, false));
        m_NeighList.reset(
            new BehaveRT::DeviceArrayWrapper<uint4>(di, (int)(p.numNeighWordsPerAgent * hBody3DParams.numBodies)));
    }
    void uninstall() {
        m_CellStart.reset();
        m_CellEnd.reset();
        m_NeighList.reset();
    }

    std::unique_ptr<BehaveRT::DeviceArrayWrapper<BehaveRT::uint>> m_CellStart;
    std::unique_ptr<BehaveRT::DeviceArrayWrapper<BehaveRT::uint>> m_CellEnd;
    std::unique_ptr<BehaveRT::DeviceArrayWrapper<uint4>> m_NeighList;
    bool m_UseCellSizeAsSearchRadius = false;
};

}  // namespace Proximity3D

using Proximity3D::Proximity3DPlugIn;
