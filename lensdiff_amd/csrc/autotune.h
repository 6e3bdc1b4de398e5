// Run-time choice among kernel variants that compute the same transform (different radix orders,
// launch shapes or occupancy targets; results agree to FP32 rounding, and whichever variant ran
// is the result). The first launch for a key times every
// variant on the live workload with HIP events on the launch stream (each variant twice, the
// second run counted, so code-object loading is not), keeps the fastest for the life of the
// process, and later launches go straight to it. Keys name the kernel family, the line length
// and the device.
#pragma once

#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "ldh_internal.h"

namespace ldh {

class Autotuner {
public:
    static Autotuner& instance();
    // Launches the chosen variant of `key` (of `variants`) on s via launch(v), tuning first if
    // the key is new. Returns the variant used.
    int run(const std::string& key, int variants, hipStream_t s, const std::function<void(int)>& launch);
    // One line per key: "key: vN  t0 t1 ... (ms)".
    std::string report() const;
    void reset();
    // variant >= 0: every key launches that variant (the last one where a family has fewer)
    // without timing; -1: back to tuning. Either way the table is cleared.
    void force(int variant);

private:
    struct Entry {
        int choice = 0;
        std::vector<float> ms;
    };
    mutable std::mutex mu_;
    std::map<std::string, Entry> table_;
    int forced_ = -1;
};

}  // namespace ldh
