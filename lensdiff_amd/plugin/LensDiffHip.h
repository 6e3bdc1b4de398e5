// HIP backend seam for the LensDiff OpenFX plugin.
//
// Same signature and error convention as the plugin's other backends:
//   RunLensDiffCpuReference  (source/src/core/LensDiffCpuReference.h:64-67)
//   RunLensDiffCuda          (source/src/cuda/LensDiffCuda.h:7-10)
//   RunLensDiffMetal         (source/src/metal/LensDiffMetal.h:7-10)
// Returns false with a kebab-case tag in *error on failure; never throws. The plugin
// dispatches to it from LensDiffEffect::render like the CUDA block (LensDiff.cpp:3830-3838).
// This file is compiled inside the plugin tree, against its LensDiffTypes.h.
#pragma once

#include <string>

#include "LensDiffTypes.h"

bool RunLensDiffHip(const LensDiffRenderRequest& request,
                    const LensDiffParams& params,
                    LensDiffPsfBankCache& cache,
                    std::string* error);
