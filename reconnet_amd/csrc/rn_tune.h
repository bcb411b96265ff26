// Launch-shape autotuner (no reference counterpart).
//
// A tunable launch names its kernel, a shape key (the sizes that change the
// best configuration) and a list of candidate configurations.  On the first
// launch of a key the tuner times every candidate with HIP events on the
// caller's stream (one warm-up and three timed runs each) and keeps the first
// candidate within 3 % of the fastest (candidates are listed in order of
// preference); later launches of the key reuse it.  Candidates differ only in the
// block / tile shape of the launch: every output element is computed by the
// same per-element code, so the choice never changes the results.
//
// Tuning is skipped (the first candidate, or the cached choice, is used) when
// the stream is being captured into a graph, when the launch accumulates into
// its output, or when RECONNET_AUTOTUNE=0.
#pragma once

#include "rn_common.h"

#include <functional>
#include <string>
#include <vector>

namespace rn {

struct TuneKey {
    int kernel;
    int a, b, c, d, e;
    bool operator<(const TuneKey& o) const;
};

enum TuneKernel {
    TK_SIDDON_FP = 1,
    TK_SIDDON_BP = 2,
    TK_DD_FP = 3,
    TK_DD_ADJ = 4,
    TK_DD_BP = 5,
    TK_SQS_UPDATE = 6,
    TK_DD_FBP = 7,
    TK_SIDDON_FP_Q = 8,
    TK_SIDDON_BP_Q = 9,
    TK_DD_BP_Q = 10,
    TK_DD_FBP_Q = 11,
    TK_DD_FP_Q = 12,
    TK_CONV_MID = 13,
    TK_CNN_FWD = 14,
    TK_HELICAL_BP = 15,
    TK_DD_CONE_FP = 16,
    TK_PIXEL_BP = 17,
    TK_SIDDON_CONE_BP = 18,
    TK_DD_CONE_ADJ = 19,
};

// Returns the candidate index to launch for `key`; runs the timing when the
// key is new and tuning is allowed.  `launch(i)` must enqueue candidate i on
// `st` and be repeatable (its output overwritten, not accumulated).
int tune_choose(const TuneKey& key, int ncand, const std::function<void(int)>& launch, hipStream_t st,
                bool may_tune);

void tune_set_enabled(bool on);
bool tune_enabled();
// "kernel a b c d e -> choice (ms per candidate ...)" lines for every tuned key
std::string tune_report();
void tune_reset();
// Force candidate `candidate` for every key of `kernel` (a TuneKernel; a
// negative candidate removes the pin).  A launch whose kernel has fewer
// candidates fails with an error.  For tests of every launch shape.
void tune_pin(int kernel, int candidate);

} // namespace rn
