// ksh_profile.hpp -- RAII bracket that times one kernel launch with HIP events when profiling is on.
#pragma once

#include <hip/hip_runtime.h>

namespace ksh {

bool profiling();
// adds n events to the counter `name` (reported by ksh_profile_query as launches, 0 ms)
void prof_count(const char *name, long long n);

class ProfScope {
public:
    ProfScope(const char *name, hipStream_t s);
    ~ProfScope();
    ProfScope(const ProfScope &) = delete;
    // algorithmic work of the bracketed launch (flops or bytes, as the kernel's bound counts it)
    void set_work(double w) { m_work = w; }
    ProfScope &operator=(const ProfScope &) = delete;

private:
    const char *m_name;
    hipStream_t m_stream;
    hipEvent_t m_start = nullptr, m_stop = nullptr;
    bool m_active = false;
    double m_work = 0.0;
};

}  // namespace ksh
