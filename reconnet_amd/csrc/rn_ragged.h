// Batches of 2D fan slices of uneven size served by single persistent launches.
#pragma once

#include "rn_fan.h"

#include <vector>

namespace rn {

class RaggedBatch {
public:
    RaggedBatch();
    ~RaggedBatch();
    RaggedBatch(const RaggedBatch&) = delete;
    RaggedBatch& operator=(const RaggedBatch&) = delete;
    // one geometry (nz = nv = 1) and one host angle array per slice
    void setup(const std::vector<FanGeo>& geos, const std::vector<const float*>& angles, hipStream_t st);
    int size() const;
    long image_elems() const;   // sum of nx*ny
    long sino_elems() const;    // sum of nu*nview
    // images concatenated slice after slice ([nx][ny] each), sinograms likewise ([nu][nview] each)
    void forward(bool siddon, const float* img, float* prj, hipStream_t st);    // DD or Siddon FP
    void backward(bool siddon, const float* prj, float* img, hipStream_t st);   // exact transposes
private:
    struct Impl;
    Impl* p_;
};

} // namespace rn
