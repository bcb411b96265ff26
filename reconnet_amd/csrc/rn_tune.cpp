// Launch-shape autotuner: per-process cache of the fastest candidate per key.

#include "rn_tune.h"

#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>

namespace rn {

bool TuneKey::operator<(const TuneKey& o) const
{
    return std::tie(kernel, a, b, c, d, e) < std::tie(o.kernel, o.a, o.b, o.c, o.d, o.e);
}

namespace {

struct Entry {
    int choice = 0;
    std::vector<float> ms;
};

struct State {
    std::mutex mu;
    std::map<TuneKey, Entry> cache;
    std::map<int, int> pinned;   // kernel -> candidate forced for every key
    int enabled = -1;   // -1: read RECONNET_AUTOTUNE on first use
};

State& state()
{
    static State s;
    return s;
}

bool enabled_locked(State& s)
{
    if (s.enabled < 0) {
        const char* v = std::getenv("RECONNET_AUTOTUNE");
        s.enabled = (v && v[0] == '0') ? 0 : 1;
    }
    return s.enabled == 1;
}

struct Events {
    hipEvent_t a = nullptr, b = nullptr;
    Events()
    {
        RN_HIP_CHECK(hipEventCreate(&a));
        RN_HIP_CHECK(hipEventCreate(&b));
    }
    ~Events()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

} // namespace

int tune_choose(const TuneKey& key, int ncand, const std::function<void(int)>& launch, hipStream_t st,
                bool may_tune)
{
    State& s = state();
    {
        std::lock_guard<std::mutex> lk(s.mu);
        auto pin = s.pinned.find(key.kernel);
        if (pin != s.pinned.end()) {
            RN_REQUIRE(pin->second < ncand, "pinned launch candidate " + std::to_string(pin->second) +
                                                " out of range: kernel " + std::to_string(key.kernel) + " has " +
                                                std::to_string(ncand));
            return pin->second;
        }
        auto it = s.cache.find(key);
        if (it != s.cache.end()) return it->second.choice;
        if (ncand <= 1 || !may_tune || !enabled_locked(s)) return 0;
    }
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    RN_HIP_CHECK(hipStreamIsCapturing(st, &cap));
    if (cap != hipStreamCaptureStatusNone) return 0;

    Entry e;
    e.ms.assign(ncand, 0.0f);
    Events ev;
    constexpr int reps = 3;
    float best = 0.0f;
    for (int i = 0; i < ncand; i++) {
        launch(i);   // warm-up (instruction cache, tables)
        RN_HIP_CHECK(hipEventRecord(ev.a, st));
        for (int r = 0; r < reps; r++) launch(i);
        RN_HIP_CHECK(hipEventRecord(ev.b, st));
        RN_HIP_CHECK(hipEventSynchronize(ev.b));
        float ms = 0.0f;
        RN_HIP_CHECK(hipEventElapsedTime(&ms, ev.a, ev.b));
        e.ms[i] = ms / reps;
        if (i == 0 || e.ms[i] < best) best = e.ms[i];
    }
    // the first candidate within 3 % of the fastest: timing noise does not
    // move the choice away from the earlier-listed (preferred) shapes
    for (int i = 0; i < ncand; i++)
        if (e.ms[i] <= 1.03f * best) {
            e.choice = i;
            break;
        }
    std::lock_guard<std::mutex> lk(s.mu);
    auto res = s.cache.emplace(key, e);
    return res.first->second.choice;
}

void tune_set_enabled(bool on)
{
    State& s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    s.enabled = on ? 1 : 0;
}

bool tune_enabled()
{
    State& s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    return enabled_locked(s);
}

std::string tune_report()
{
    State& s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    std::string out;
    char line[512];
    for (const auto& kv : s.cache) {
        const TuneKey& k = kv.first;
        int n = std::snprintf(line, sizeof line, "%d %d %d %d %d %d -> %d (", k.kernel, k.a, k.b, k.c, k.d, k.e,
                              kv.second.choice);
        out.append(line, n);
        for (size_t i = 0; i < kv.second.ms.size(); i++) {
            // every digit of the float, so the report shows the values the choice compared
            n = std::snprintf(line, sizeof line, "%s%.9g", i ? " " : "", (double)kv.second.ms[i]);
            out.append(line, n);
        }
        out += " ms)\n";
    }
    return out;
}

void tune_reset()
{
    State& s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    s.cache.clear();
}

void tune_pin(int kernel, int candidate)
{
    State& s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    if (candidate < 0) s.pinned.erase(kernel);
    else s.pinned[kernel] = candidate;
}

} // namespace rn
