// Body3DPlugIn.h -- chain base: number of bodies, common radius, 2-D projection.
// Mirrors Body3DPlugIn (PlugIns/Body/include/Body3DPlugIn.h:47-129).  Config
// section [Body3DPlugIn]: numBodies, isBodyNumExponent, commonRadius,
// use2DProjection.
#pragma once

#include <cmath>

#include "SimEnginePlugIn.h"

namespace Body {

template <class Super>
class Body3DPlugIn : public Super {
public:
    Body3DPlugIn() { BehaveRT::PlugInInstaller::install(this); }
    ~Body3DPlugIn() { BehaveRT::PlugInInstaller::uninstall(this); }

    const std::string name() override { return "Body3DPlugIn"; }
    const BehaveRT::DependenciesList plugInDependencies() override { return BehaveRT::DependenciesList(); }
    void reset() override { Super::reset(); }
    void update(const float elapsedTime) override { Super::update(elapsedTime); }

protected:
    friend struct BehaveRT::PlugInInstaller;
    void install() {
        using BehaveRT::StringConverter;
        BehaveRT::UtilityConfigFile* cfg = m_CommonRes.getConfig();
        const int numBodies = StringConverter::parseInt(cfg->getSetting("numBodies", "Body3DPlugIn"));
        if (StringConverter::parseBool(cfg->getSetting("isBodyNumExponent", "Body3DPlugIn")))
            hBody3DParams.numBodies = (int)std::pow(2.0, numBodies);
        else
            hBody3DParams.numBodies = numBodies;
        hBody3DParams.commonRadius = StringConverter::parseFloat(cfg->getSetting("commonRadius", "Body3DPlugIn"));
        hBody3DParams.use2DProjection =
            StringConverter::parseBool(cfg->getSetting("use2DProjection", "Body3DPlugIn"));
    }
    void uninstall() {}
};

}  // namespace Body

using Body::Body3DPlugIn;
