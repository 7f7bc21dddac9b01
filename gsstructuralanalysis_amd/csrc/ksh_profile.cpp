// ksh_profile.cpp -- optional per-kernel timing with HIP events recorded on the launching stream.
//
// When enabled, each launcher brackets its kernel with a pair of events from a recycled pool;
// ksh_profile_query() waits for the pending pairs and accumulates hipEventElapsedTime per kernel.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "ksh_internal.hpp"
#include "ksh_profile.hpp"

namespace ksh {

namespace {

struct Entry {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    double total_ms = 0.0;
    long long count = 0;
    double work = 0.0;
};

std::mutex g_mu;
bool g_on = false;
std::map<std::string, Entry> g_entries;
std::vector<hipEvent_t> g_pool;

hipEvent_t take_event()
{
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

void drain(Entry &en)
{
    for (auto &pr : en.pending) {
        (void)hipEventSynchronize(pr.second);
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
            en.total_ms += ms;
            en.count += 1;
        }
        g_pool.push_back(pr.first);
        g_pool.push_back(pr.second);
    }
    en.pending.clear();
}

}  // namespace

bool profiling() { return g_on; }

void prof_count(const char *name, long long n)
{
    if (!g_on || n == 0) return;
    std::lock_guard<std::mutex> lk(g_mu);
    g_entries[name].count += n;
}

ProfScope::ProfScope(const char *name, hipStream_t s) : m_name(name), m_stream(s)
{
    if (!g_on) return;
    std::lock_guard<std::mutex> lk(g_mu);
    m_start = take_event();
    m_stop = take_event();
    (void)hipEventRecord(m_start, s);
    m_active = true;
}

ProfScope::~ProfScope()
{
    if (!m_active) return;
    std::lock_guard<std::mutex> lk(g_mu);
    (void)hipEventRecord(m_stop, m_stream);
    Entry &en = g_entries[m_name];
    en.pending.emplace_back(m_start, m_stop);
    en.work += m_work;
}

}  // namespace ksh

extern "C" {

int ksh_profile_enable(int on)
{
    std::lock_guard<std::mutex> lk(ksh::g_mu);
    ksh::g_on = on != 0;
    return KSH_SUCCESS;
}

int ksh_profile_reset(void)
{
    std::lock_guard<std::mutex> lk(ksh::g_mu);
    for (auto &kv : ksh::g_entries) ksh::drain(kv.second);
    ksh::g_entries.clear();
    return KSH_SUCCESS;
}

int ksh_profile_query(const char *kernel, int64_t *launches, double *total_ms)
{
    if (!kernel) return ksh::fail(KSH_OTHER_ERROR, "null argument");
    std::lock_guard<std::mutex> lk(ksh::g_mu);
    auto it = ksh::g_entries.find(kernel);
    if (it == ksh::g_entries.end()) {
        if (launches) *launches = 0;
        if (total_ms) *total_ms = 0.0;
        return KSH_SUCCESS;
    }
    ksh::drain(it->second);
    if (launches) *launches = it->second.count;
    if (total_ms) *total_ms = it->second.total_ms;
    return KSH_SUCCESS;
}

int ksh_profile_query_work(const char *kernel, double *work)
{
    if (!kernel || !work) return ksh::fail(KSH_OTHER_ERROR, "null argument");
    std::lock_guard<std::mutex> lk(ksh::g_mu);
    auto it = ksh::g_entries.find(kernel);
    *work = it == ksh::g_entries.end() ? 0.0 : it->second.work;
    return KSH_SUCCESS;
}

}  // extern "C"
