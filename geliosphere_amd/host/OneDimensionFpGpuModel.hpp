// OneDimensionFpGpuModel.hpp - 1D F-p (forward in time) model on the gfx950 trajectory engine.
// Same class name and entry point as the reference's CUDA model
// (CUDAKernel/include/OneDimensionFpGpuModel.hpp:22-55); prepareAndRunSimulation replaces CUDAKernel/src/OneDimensionFpGpuModel.cpp:7-40.
#ifndef GS_ONEDIMENSIONFPGPUMODEL_HPP
#define GS_ONEDIMENSIONFPGPUMODEL_HPP

#include "AbstractGpuSimulation.hpp"

class OneDimensionFpGpuModel : public AbstractGpuSimulation {
  public:
    void prepareAndRunSimulation(ParamsCarrier *singleTone) override;
};

#endif
