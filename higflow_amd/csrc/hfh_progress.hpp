// Progress lines of long host set-ups.
#pragma once
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace hfh {

// HFH_PROGRESS=1 (or HFH_VERBOSE=1): long set-up loops report how far they
// are on stderr, at most one line per 20 s, so a caller watching the output
// can tell a long set-up (millions of cells) from a stalled one.
// progress_start() restarts the clock; progress() is cheap when off.
inline std::atomic<long long> &progress_clock() {
    static std::atomic<long long> t{0};
    return t;
}
inline long long progress_now() {
    return std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}
inline bool progress_on() {
    static const bool on = [] {
        for (const char *v : {"HFH_PROGRESS", "HFH_VERBOSE"})
            if (const char *e = std::getenv(v))
                if (std::strcmp(e, "1") == 0) return true;
        return false;
    }();
    return on;
}
inline void progress_start() { progress_clock().store(progress_now()); }
inline void progress(const char *what, long done, long total) {
    if (!progress_on()) return;
    long long last = progress_clock().load(), now = progress_now();
    if (now - last < 20000 || !progress_clock().compare_exchange_strong(last, now)) return;
    std::fprintf(stderr, "hfh: set-up %s: %ld of %ld\n", what, done, total);
    std::fflush(stderr);
}

}  // namespace hfh
