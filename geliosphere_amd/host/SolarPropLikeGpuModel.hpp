// SolarPropLikeGpuModel.hpp - 2D SolarProp-like model on the gfx950 trajectory engine.
// Same class name and entry point as the reference's CUDA model
// (CUDAKernel/include/SolarPropLikeGpuModel.hpp:22-55); prepareAndRunSimulation replaces CUDAKernel/src/SolarPropLikeGpuModel.cpp:7-44.
#ifndef GS_SOLARPROPLIKEGPUMODEL_HPP
#define GS_SOLARPROPLIKEGPUMODEL_HPP

#include "AbstractGpuSimulation.hpp"

class SolarPropLikeGpuModel : public AbstractGpuSimulation {
  public:
    void prepareAndRunSimulation(ParamsCarrier *singleTone) override;
};

#endif
