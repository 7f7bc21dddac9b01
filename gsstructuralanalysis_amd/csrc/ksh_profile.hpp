// ksh_profile.hpp -- RAII bracket that times one kernel launch with HIP events when profiling is on.
#pragma once

#include <hip/hip_runtime.h>

namespace ksh {

bool profiling();

class ProfScope {
public:
    ProfScope(const char *name, hipStream_t s);
    ~ProfScope();
    ProfScope(const ProfScope &) = delete;
    ProfScope &operator=(const ProfScope &) = delete;

private:
    const char *m_name;
    hipStream_t m_stream;
    hipEvent_t m_start = nullptr, m_stop = nullptr;
    bool m_active = false;
};

}  // namespace ksh
