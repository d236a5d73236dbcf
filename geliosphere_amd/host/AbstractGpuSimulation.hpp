// AbstractGpuSimulation.hpp - the plugin interface every GPU model
// implements (reference CUDAKernel/include/AbstractGpuSimulation.hpp:23-32):
// one call that runs a whole simulation described by the ParamsCarrier.
#ifndef GS_ABSTRACT_GPU_SIMULATION_HPP
#define GS_ABSTRACT_GPU_SIMULATION_HPP

#include "ParamsCarrier.hpp"

class AbstractGpuSimulation {
  public:
    virtual ~AbstractGpuSimulation() = default;
    virtual void prepareAndRunSimulation(ParamsCarrier *singleTone) = 0;
};

#endif
