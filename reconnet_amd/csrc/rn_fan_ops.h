// Fan-beam operations on reference-layout device buffers.
#pragma once

#include "rn_fan.h"

namespace rn {

enum class FanFP { DD, Siddon };
enum class FanBP { DD, DDFBP, DDAdjoint, Siddon, Pixel };

// Scratch owned by a caller (plan or ABI call); grows on demand.
struct FanWork {
    DevBuf planes, rx, ry, sino, mix, quad, qdd;
};

// img_ref [nB][nx][ny][nz][nC] -> prj_ref [nB][nu][nview][nv][nC] (overwritten)
void fan_forward(const FanTables& t, FanFP kind, int nB, int nC, const float* img_ref, float* prj_ref,
                 FanWork& w, hipStream_t st);
// prj_ref -> img_ref (overwritten)
void fan_backward(const FanTables& t, FanBP kind, int nB, int nC, const float* prj_ref, float* img_ref,
                  FanWork& w, hipStream_t st);

// z/v mixing matrices (rows = outputs) for each operator, built on the host
std::vector<float> fp_mix_matrix(const FanGeo& g, FanFP kind);           // [nv][nz]
std::vector<float> bp_mix_matrix(const FanGeo& g, FanBP kind);           // [nz][nv]

} // namespace rn
