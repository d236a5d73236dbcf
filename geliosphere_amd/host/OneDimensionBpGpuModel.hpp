// OneDimensionBpGpuModel.hpp - 1D B-p (backward in time) model on the gfx950 trajectory engine.
// Same class name and entry point as the reference's CUDA model
// (CUDAKernel/include/OneDimensionBpGpuModel.hpp:22-55); prepareAndRunSimulation replaces CUDAKernel/src/OneDimensionBpGpuModel.cpp:7-43.
#ifndef GS_ONEDIMENSIONBPGPUMODEL_HPP
#define GS_ONEDIMENSIONBPGPUMODEL_HPP

#include "AbstractGpuSimulation.hpp"

class OneDimensionBpGpuModel : public AbstractGpuSimulation {
  public:
    void prepareAndRunSimulation(ParamsCarrier *singleTone) override;
};

#endif
