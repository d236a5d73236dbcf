// GeliosphereGpuModel.hpp - 2D Geliosphere model on the gfx950 trajectory engine.
// Same class name and entry point as the reference's CUDA model
// (CUDAKernel/include/GeliosphereGpuModel.hpp:22-55); prepareAndRunSimulation replaces CUDAKernel/src/GeliosphereGpuModel.cpp:7-44.
#ifndef GS_GELIOSPHEREGPUMODEL_HPP
#define GS_GELIOSPHEREGPUMODEL_HPP

#include "AbstractGpuSimulation.hpp"

class GeliosphereGpuModel : public AbstractGpuSimulation {
  public:
    void prepareAndRunSimulation(ParamsCarrier *singleTone) override;
};

#endif
