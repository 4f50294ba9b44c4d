// The table of measured launch-shape decisions and its C ABI (autotune.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>

#include "autotune.hpp"
#include "mrs_internal.hpp"

namespace mrs {
namespace {

struct Entry {
    int variant;
    std::vector<int> candidates;
    std::vector<float> ms;
};

struct Table {
    std::mutex mu;
    int mode = MRS_TUNE_ON_FIRST_USE;
    std::map<int, int> forced;
    std::map<std::pair<int, uint64_t>, Entry> entries;
};

Table& table() {
    static Table t;
    return t;
}

bool known_family(int f) { return f == kTuneEpochSpan || f == kTuneKnnSplits; }

// variants a family may take (a forced or loaded variant outside these is refused)
bool valid_variant(int f, int v) {
    if (f == kTuneEpochSpan) return v == 64 || v == 128 || v == 256;
    if (f == kTuneKnnSplits) return v == 1 || v == 2 || v == 4 || v == 8;
    return false;
}

// the device the table's decisions were measured on
std::string device_arch() {
    int dev = 0;
    hipDeviceProp_t prop{};
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return "unknown";
    return std::string(prop.gcnArchName) + ":" + std::to_string(prop.multiProcessorCount);
}

}  // namespace

int tune_forced(int family) {
    Table& t = table();
    std::lock_guard<std::mutex> lk(t.mu);
    auto it = t.forced.find(family);
    return it == t.forced.end() ? -1 : it->second;
}

int tune_lookup(int family, uint64_t shape) {
    Table& t = table();
    std::lock_guard<std::mutex> lk(t.mu);
    auto it = t.entries.find({family, shape});
    return it == t.entries.end() ? -1 : it->second.variant;
}

bool tune_enabled() {
    Table& t = table();
    std::lock_guard<std::mutex> lk(t.mu);
    return t.mode == MRS_TUNE_ON_FIRST_USE;
}

int tune_record(int family, uint64_t shape, const std::vector<int>& candidates,
                const std::vector<float>& ms) {
    size_t best = 0;
    for (size_t q = 1; q < ms.size(); q++)
        if (ms[q] < ms[best]) best = q;
    Table& t = table();
    std::lock_guard<std::mutex> lk(t.mu);
    t.entries[{family, shape}] = Entry{candidates[best], candidates, ms};
    return candidates[best];
}

}  // namespace mrs

using mrs::guarded;

extern "C" {

int mrs_autotune_set_mode(int32_t mode) {
    return guarded([&] {
        if (mode != MRS_TUNE_OFF && mode != MRS_TUNE_ON_FIRST_USE)
            throw mrs::Error(MRS_INVALID_ARGUMENT, "autotune mode must be MRS_TUNE_OFF or MRS_TUNE_ON_FIRST_USE");
        mrs::Table& t = mrs::table();
        std::lock_guard<std::mutex> lk(t.mu);
        t.mode = mode;
    });
}

int mrs_autotune_force(int32_t family, int32_t variant) {
    return guarded([&] {
        if (!mrs::known_family(family))
            throw mrs::Error(MRS_INVALID_ARGUMENT, "unknown autotune family");
        if (variant != -1 && !mrs::valid_variant(family, variant))
            throw mrs::Error(MRS_INVALID_ARGUMENT, "not a variant of this autotune family");
        mrs::Table& t = mrs::table();
        std::lock_guard<std::mutex> lk(t.mu);
        if (variant == -1) t.forced.erase(family);
        else t.forced[family] = variant;
    });
}

int mrs_autotune_clear(void) {
    return guarded([&] {
        mrs::Table& t = mrs::table();
        std::lock_guard<std::mutex> lk(t.mu);
        t.entries.clear();
    });
}

int64_t mrs_autotune_dump(char* buf, int64_t cap) {
    std::ostringstream os;
    {
        mrs::Table& t = mrs::table();
        std::lock_guard<std::mutex> lk(t.mu);
        for (const auto& kv : t.entries) {
            os << kv.first.first << ' ' << kv.first.second << ' ' << kv.second.variant;
            for (size_t q = 0; q < kv.second.candidates.size(); q++)
                os << ' ' << kv.second.candidates[q] << ':' << kv.second.ms[q];
            os << '\n';
        }
    }
    const std::string s = os.str();
    if (buf && cap > 0) {
        const size_t m = std::min<size_t>((size_t)cap - 1, s.size());
        std::memcpy(buf, s.data(), m);
        buf[m] = '\0';
    }
    return (int64_t)s.size() + 1;
}

int mrs_autotune_save(const char* path) {
    return guarded([&] {
        if (!path) throw mrs::Error(MRS_INVALID_ARGUMENT, "null path");
        std::ofstream f(path);
        if (!f) throw mrs::Error(MRS_INVALID_ARGUMENT, std::string("cannot write ") + path);
        f << "mrs-autotune 1 " << mrs::device_arch() << '\n';
        mrs::Table& t = mrs::table();
        std::lock_guard<std::mutex> lk(t.mu);
        for (const auto& kv : t.entries) {
            f << kv.first.first << ' ' << kv.first.second << ' ' << kv.second.variant;
            for (size_t q = 0; q < kv.second.candidates.size(); q++)
                f << ' ' << kv.second.candidates[q] << ':' << kv.second.ms[q];
            f << '\n';
        }
        if (!f) throw mrs::Error(MRS_INVALID_ARGUMENT, std::string("cannot write ") + path);
    });
}

int mrs_autotune_load(const char* path) {
    return guarded([&] {
        if (!path) throw mrs::Error(MRS_INVALID_ARGUMENT, "null path");
        std::ifstream f(path);
        if (!f) throw mrs::Error(MRS_INVALID_ARGUMENT, std::string("cannot read ") + path);
        std::string head, magic, arch;
        int version = 0;
        std::getline(f, head);
        std::istringstream hs(head);
        hs >> magic >> version >> arch;
        if (magic != "mrs-autotune" || version != 1)
            throw mrs::Error(MRS_INVALID_ARGUMENT, "not an autotune table");
        if (arch != mrs::device_arch())
            throw mrs::Error(MRS_INVALID_ARGUMENT,
                             "autotune table was measured on " + arch + ", not this device");
        std::map<std::pair<int, uint64_t>, mrs::Entry> got;
        std::string line;
        while (std::getline(f, line)) {
            if (line.empty()) continue;
            std::istringstream ls(line);
            int fam = 0, var = 0;
            uint64_t shape = 0;
            if (!(ls >> fam >> shape >> var) || !mrs::known_family(fam) ||
                !mrs::valid_variant(fam, var))
                throw mrs::Error(MRS_INVALID_ARGUMENT, "bad autotune table line: " + line);
            mrs::Entry e{var, {}, {}};
            std::string tok;
            while (ls >> tok) {
                const size_t c = tok.find(':');
                if (c == std::string::npos)
                    throw mrs::Error(MRS_INVALID_ARGUMENT, "bad autotune table line: " + line);
                e.candidates.push_back(std::stoi(tok.substr(0, c)));
                e.ms.push_back(std::stof(tok.substr(c + 1)));
            }
            got[{fam, shape}] = e;
        }
        mrs::Table& t = mrs::table();
        std::lock_guard<std::mutex> lk(t.mu);
        for (auto& kv : got) t.entries[kv.first] = kv.second;
    });
}

}  // extern "C"
