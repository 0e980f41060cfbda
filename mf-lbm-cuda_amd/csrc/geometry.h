// Device geometry preprocessing (src/Geometry_preprocessing.cpp:29-403 of the reference).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "lattice.h"

namespace mflbm {
// geo_dev: nx*ny*nz solid flags (domain walls already applied), x fastest.
// Outputs: walls (s2, 0/1), wtype (s4), solid normals (s4, double, 0 off fluid-boundary nodes).
hipError_t geometry_preprocess(const Dims& d, int iper, int jper, int kper, const int8_t* geo_dev, uint8_t* walls,
                               int8_t* wtype, double* snx, double* sny, double* snz, hipStream_t s);
}  // namespace mflbm
