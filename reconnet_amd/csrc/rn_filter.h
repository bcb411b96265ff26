// Ramp filter (Filter.cu) as a direct convolution on view-major sinograms.
#pragma once

#include "rn_common.h"

#include <vector>

namespace rn {

// taps h[k + nu - 1], k in [-(nu-1), nu-1], windowed as getRamp does (FP64)
void ramp_taps_host(std::vector<double>& out, int nu, double da, int filterType, bool equalSpace);

class RampFilter {
public:
    // fan = true: equiangular taps, cosine pre-weight, scale pi/nview*da/dso (cFilterFan)
    // fan = false: equispaced taps, no pre-weight, scale pi/nview*du (cFilterParallel)
    void setup(int nu, int nview, float da, float off_a, float dso, int filterType, bool fan, hipStream_t st);
    // in/out: [nrows][nu] rows of one view each
    void apply_vu(const float* in_vu, float* out_vu, long nrows, hipStream_t st) const;
private:
    bool ready_ = false;
    int nu_ = 0, nview_ = 0, type_ = 0;
    float da_ = 0, off_a_ = 0, dso_ = 0;
    bool fan_ = true;
    std::vector<float> host_;
    DevBuf buf_;
};

} // namespace rn
