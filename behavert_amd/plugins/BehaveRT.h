// BehaveRT.h -- umbrella header of the plug-in layer (Core/include/BehaveRT.h).
//
// Build a simulation engine as a mixin chain and drive it like the reference's
// demos do (Samples/OpensteerDemo/plugins/BehaveRT_Flocking.cpp:149-206):
//
//   typedef OpenSteerWrapperPlugIn<Proximity3DPlugIn<EnvGrid3DPlugIn<
//               Body3DPlugIn<DummyPlugIn>>>> CrowdEngine;
//   m_CommonRes.initialize("BehaveRT_Flocking.cfg");
//   CrowdEngine engine;  engine.reset();
//   for (...) {
//       engine.reorderSimData();
//       engine.computeNeighborhoods();
//       engine.resetSteerForce();
//       engine.computeSeparations(); engine.computeAlignments(); engine.computeCohesions();
//       engine.applySteeringForces(dt);
//   }
//
// Link with libbehavert_amd.so (the gfx950 kernels behind include/behavert_amd.h).
#pragma once

#include "../../include/behavert_amd.h"
#include "BehaviorClusteringPlugIn.h"
#include "Body3DPlugIn.h"
#include "DeviceArrayWrapper.h"
#include "DeviceInterface.h"
#include "EnvGrid3DPlugIn.h"
#include "OpenSteerWrapperPlugIn.h"
#include "Proximity3DPlugIn.h"
#include "SimEnginePlugIn.h"
#include "UtilityConfigFile.h"
