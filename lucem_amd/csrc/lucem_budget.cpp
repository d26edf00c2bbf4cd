// Device memory budget of the library's own allocations (workspace slots, batched and gradient
// plans).
//
// Mirrors lucem-core's budget (crates/lucem-core/src/memory/budget.rs):
//   LUCEM_DEVICE_MEMORY / LUCEM_DEVICE_MEMORY_LIMIT     budget.rs:22-25, read once (:246-261)
//   parse_bytes                                        budget.rs:263-279
//   set_device_memory_limit / device_memory_limit      budget.rs:175-187
//   check_allocation (64 MiB threshold, available =    budget.rs:31, :232-244
//     limit - in use, OutOfDeviceMemory message)       lucem-core/src/error.rs:27-29
// The reference counts what CubeCL's pools hold in use on the device. Here the caller's arrays
// belong to the caller (torch or a Rust DeviceArray), so "in use" is what this library holds on
// that device in workspace slots and batched / gradient plans, which are allocated through
// dev_malloc / dev_free below (the FFT tables and the autotuner's scratch, a few MiB, are not
// counted).
#include <hip/hip_runtime.h>

#include <atomic>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "../../include/lucem_hip.h"
#include "lucem_internal_host.hpp"

using namespace lucem;

namespace {
std::mutex g_mu;
bool g_env_read = false;
bool g_has_limit = false;
uint64_t g_limit = 0;
std::map<int, uint64_t> g_in_use;  // device -> bytes held by this library

constexpr uint64_t kCheckedAllocationBytes = 64ull << 20;

// budget.rs:263-279: digits and '.', then an optional unit (b, kb, mb, gb, kib, mib, gib; any
// case, surrounding spaces ignored); the value is truncated to whole bytes.
bool parse_bytes_impl(const char* text, uint64_t* out) {
    if (!text) return false;
    std::string s(text);
    const size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    if (a == std::string::npos) return false;
    s = s.substr(a, b - a + 1);
    size_t split = 0;
    while (split < s.size() && (std::isdigit((unsigned char)s[split]) || s[split] == '.')) split++;
    const std::string number = s.substr(0, split);
    std::string unit = s.substr(split);
    const size_t ua = unit.find_first_not_of(" \t\r\n");
    unit = ua == std::string::npos ? std::string() : unit.substr(ua);
    for (auto& c : unit) c = (char)std::tolower((unsigned char)c);
    if (number.empty() || number.find('.') != number.rfind('.')) return false;
    if (number == ".") return false;
    const double value = std::strtod(number.c_str(), nullptr);
    double scale;
    if (unit.empty() || unit == "b") scale = 1.0;
    else if (unit == "kb") scale = 1e3;
    else if (unit == "mb") scale = 1e6;
    else if (unit == "gb") scale = 1e9;
    else if (unit == "kib") scale = 1024.0;
    else if (unit == "mib") scale = 1024.0 * 1024.0;
    else if (unit == "gib") scale = 1024.0 * 1024.0 * 1024.0;
    else return false;
    const double bytes = value * scale;
    if (!(bytes >= 0.0)) return false;
    *out = bytes >= 18446744073709551615.0 ? ~0ull : (uint64_t)bytes;
    return true;
}

void read_env_locked() {
    if (g_env_read) return;
    g_env_read = true;
    const char* v = std::getenv("LUCEM_DEVICE_MEMORY");
    if (!v) v = std::getenv("LUCEM_DEVICE_MEMORY_LIMIT");
    uint64_t bytes = 0;
    if (v && parse_bytes_impl(v, &bytes)) {
        g_has_limit = true;
        g_limit = bytes;
    }
}

std::string format_oom(const char* what, uint64_t requested, uint64_t available, uint64_t limit) {
    char buf[320];
    std::snprintf(buf, sizeof buf,
                  "%s needs %llu bytes of device memory; the budget leaves %llu (limit %llu)", what ? what : "allocation",
                  (unsigned long long)requested, (unsigned long long)available, (unsigned long long)limit);
    return buf;
}
}  // namespace

lucem_status lucem::budget_check(int device, const char* what, uint64_t bytes) {
    std::lock_guard<std::mutex> lk(g_mu);
    read_env_locked();
    if (!g_has_limit || bytes < kCheckedAllocationBytes) return LUCEM_OK;
    const uint64_t in_use = g_in_use[device];
    const uint64_t available = g_limit > in_use ? g_limit - in_use : 0;
    if (bytes <= available) return LUCEM_OK;
    return fail(LUCEM_OUT_OF_DEVICE_MEMORY, format_oom(what, bytes, available, g_limit));
}

lucem_status lucem::dev_malloc(int device, void** p, size_t bytes, const char* what) {
    *p = nullptr;
    if (lucem_status st = budget_check(device, what, bytes)) return st;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return hip_fail(e, what);
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_in_use[device] += bytes;
    return LUCEM_OK;
}

void lucem::dev_free(int device, void* p, size_t bytes) {
    if (!p) return;
    (void)hipFree(p);
    std::lock_guard<std::mutex> lk(g_mu);
    uint64_t& u = g_in_use[device];
    u = u > bytes ? u - bytes : 0;
}

void lucem::dev_forget(int device, size_t bytes) {
    std::lock_guard<std::mutex> lk(g_mu);
    uint64_t& u = g_in_use[device];
    u = u > bytes ? u - bytes : 0;
}

extern "C" int32_t lucem_parse_bytes(const char* text, uint64_t* out) {
    uint64_t v = 0;
    if (!parse_bytes_impl(text, &v)) return 0;
    if (out) *out = v;
    return 1;
}

extern "C" void lucem_set_device_memory_limit(int32_t has_limit, uint64_t bytes) {
    std::lock_guard<std::mutex> lk(g_mu);
    read_env_locked();
    g_has_limit = has_limit != 0;
    g_limit = has_limit ? bytes : 0;
}

extern "C" int32_t lucem_device_memory_limit(uint64_t* out) {
    std::lock_guard<std::mutex> lk(g_mu);
    read_env_locked();
    if (g_has_limit && out) *out = g_limit;
    return g_has_limit ? 1 : 0;
}

extern "C" uint64_t lucem_device_bytes_in_use(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_in_use.find(device);
    return it == g_in_use.end() ? 0 : it->second;
}

extern "C" lucem_status lucem_check_allocation(int device, const char* what, uint64_t bytes) {
    return budget_check(device, what, bytes);
}
