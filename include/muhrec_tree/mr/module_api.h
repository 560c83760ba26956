// MuhRec's own module API, for a build of the HIP modules inside a MuhRec
// source tree: with this directory first on the include path,
// hip_modules.cpp compiles against the reference's classes (and their
// layouts) instead of the restatement in include/mr/.
//   kipl::base::TImage                  core/kipl/kipl/include/base/timage.h
//   kipl::interactors::InteractionBase  core/kipl/kipl/include/interactors/interactionbase.h
//   ReconConfig                         frameworks/tomography/Framework/ReconFramework/include/ReconConfig.h
//   ReconException                      .../ReconFramework/include/ReconException.h
//   ModuleException, ParameterHandling  core/modules/ModuleConfig/include/
//   PreprocModuleBase                   .../ReconFramework/include/PreprocModuleBase.h
//   BackProjectorModuleBase             .../ReconFramework/include/BackProjectorModuleBase.h
#pragma once

#include <base/timage.h>
#include <interactors/interactionbase.h>

#include <BackProjectorModuleBase.h>
#include <ModuleException.h>
#include <ParameterHandling.h>
#include <PreprocModuleBase.h>
#include <ReconConfig.h>
#include <ReconException.h>
