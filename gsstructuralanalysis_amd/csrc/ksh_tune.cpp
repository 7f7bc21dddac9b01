// ksh_tune.cpp -- launch-shape autotuner (see ksh_tune.hpp).
#include "ksh_tune.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "ksh_internal.hpp"

namespace ksh {

namespace {

struct Tuning {
    std::vector<float> ms;  // measured time per variant (negative: not measured yet)
    int best = -1;
    bool warmed = false;    // the bucket's first call (untimed, default variant) is done
};

thread_local bool t_measuring = false;

std::mutex g_mu;
std::map<std::pair<std::string, long long>, Tuning> g_tunings;
std::map<std::string, std::pair<int, int>> g_last;  // kernel -> (variant, variants) of the last call

bool tuning_on()
{
    const char *e = std::getenv("KSH_TUNE");
    return !(e && (std::strcmp(e, "0") == 0 || e[0] == 'v'));
}

// KSH_TUNE=v<k>: variant k for every tuned kernel (checks that the variants agree), or -1
int forced_variant()
{
    const char *e = std::getenv("KSH_TUNE");
    return (e && e[0] == 'v') ? std::atoi(e + 1) : -1;
}

}  // namespace

bool tune_measuring() { return t_measuring; }

long long tune_bucket(double x) { return x < 1.0 ? 0 : (long long)std::floor(std::log2(x)); }

int tune_launch(const char *kernel, long long bucket, int nvariants, int fallback, hipStream_t s,
                const std::function<void(int)> &launch)
{
    int v = fallback;
    int measure = -1;
    const int forced = forced_variant();
    if (forced >= 0 && forced < nvariants) v = forced;
    if (tuning_on()) {
        std::lock_guard<std::mutex> lk(g_mu);
        Tuning &t = g_tunings[{kernel, bucket}];
        if (t.ms.empty()) t.ms.assign(nvariants, -1.0f);
        if (t.best >= 0) {
            v = t.best;
        } else if (!t.warmed) {
            // first call of the bucket: default variant, untimed (it carries the first-use costs
            // that would otherwise fall on whichever variant is timed first)
            t.warmed = true;
        } else {
            for (int i = 0; i < nvariants; ++i)
                if (t.ms[i] < 0.0f) {
                    measure = i;
                    break;
                }
            v = measure;
        }
    }
    if (measure < 0) {
        launch(v);
    } else {
        hipEvent_t a = nullptr, b = nullptr;
        float ms = 0.0f;
        const bool ev = hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess;
        if (ev) (void)hipEventRecord(a, s);
        t_measuring = true;
        launch(v);
        t_measuring = false;
        if (ev) {
            (void)hipEventRecord(b, s);
            (void)hipEventSynchronize(b);
            (void)hipEventElapsedTime(&ms, a, b);
        }
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
        std::lock_guard<std::mutex> lk(g_mu);
        Tuning &t = g_tunings[{kernel, bucket}];
        t.ms[v] = ev ? ms : 0.0f;
        bool done = true;
        for (float m : t.ms) done = done && m >= 0.0f;
        if (done) {
            int best = 0;
            for (int i = 1; i < nvariants; ++i)
                if (t.ms[i] < t.ms[best]) best = i;
            t.best = best;
        }
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_last[kernel] = {v, nvariants};
    return v;
}

}  // namespace ksh

extern "C" int ksh_tune_choice(const char *kernel, int64_t *variant, int64_t *variants)
{
    if (!kernel || !variant || !variants) return ksh::fail(KSH_OTHER_ERROR, "null argument");
    std::lock_guard<std::mutex> lk(ksh::g_mu);
    auto it = ksh::g_last.find(kernel);
    *variant = it == ksh::g_last.end() ? -1 : it->second.first;
    *variants = it == ksh::g_last.end() ? 0 : it->second.second;
    return KSH_SUCCESS;
}
