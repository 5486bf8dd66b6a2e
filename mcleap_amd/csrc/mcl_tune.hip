// mcl_tune.hip -- the autotuner's table: launch and round-loop parameters
// chosen per workload class by measurement (mcleap_amd/autotune.py writes
// it), loaded once and applied when a triangulation or grid is created.
//
// File: MCL_TUNING names it (MCL_TUNING=off: no table), else tuning.json in
// the directory of libmcleap_hip.so.  Format (flat JSON):
//   {"entries": [{"domain": "torus" | "box" | "ensemble", "log2_nv": 20,
//                 "<parameter>": <integer>, ...}, ...], ...}
// An object is used for a mesh of the same domain whose log2(vertices) is
// within 2 of its log2_nv (the nearest such entry).  Parameters: the
// MCL_OPT_* names in lower case without the prefix (tail_threshold,
// revert_tail, reloc_tail, big_grid_min, star_head_passes, graph,
// star_min_vertices, star_enable, star_give_up) for triangulations, and
// grid_block for grids; other keys are ignored here.
#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "mcl_internal.hpp"

namespace {

struct Entry {
    int domain = -1;  // 0 box, 1 torus, 2 ensemble
    int log2_nv = -1;
    std::vector<std::pair<std::string, long long>> kv;
};

std::once_flag g_once;
std::vector<Entry> g_entries;
std::string g_path;

void skip_ws(const std::string& s, size_t& i) {
    while (i < s.size() && (s[i] == ' ' || s[i] == '\n' || s[i] == '\r' || s[i] == '\t' || s[i] == ',')) i++;
}

bool read_string(const std::string& s, size_t& i, std::string* out) {
    if (i >= s.size() || s[i] != '"') return false;
    const size_t j = s.find('"', i + 1);
    if (j == std::string::npos) return false;
    *out = s.substr(i + 1, j - i - 1);
    i = j + 1;
    return true;
}

// one flat object {"key": number | "string", ...} starting at s[i] == '{'
bool read_entry(const std::string& s, size_t& i, Entry* e) {
    i++;  // '{'
    for (;;) {
        skip_ws(s, i);
        if (i >= s.size()) return false;
        if (s[i] == '}') { i++; return true; }
        std::string key, sval;
        if (!read_string(s, i, &key)) return false;
        skip_ws(s, i);
        if (i >= s.size() || s[i] != ':') return false;
        i++;
        skip_ws(s, i);
        if (i < s.size() && s[i] == '"') {
            if (!read_string(s, i, &sval)) return false;
            if (key == "domain") e->domain = sval == "torus" ? 1 : (sval == "ensemble" ? 2 : (sval == "box" ? 0 : -1));
        } else if (s.compare(i, 4, "true") == 0 || s.compare(i, 4, "null") == 0) {
            i += 4;  // record fields such as tuned_set_kept; no parameter is boolean
        } else if (s.compare(i, 5, "false") == 0) {
            i += 5;
        } else {
            char* end = nullptr;
            const double v = strtod(s.c_str() + i, &end);
            if (end == s.c_str() + i) return false;
            i = (size_t)(end - s.c_str());
            if (key == "log2_nv") e->log2_nv = (int)v;
            else e->kv.emplace_back(key, (long long)llround(v));
        }
    }
}

void load() {
    const char* env = getenv("MCL_TUNING");
    if (env && strcmp(env, "off") == 0) return;
    if (env && *env) {
        g_path = env;
    } else {
        Dl_info info;
        if (dladdr((void*)&load, &info) && info.dli_fname) {
            std::string so = info.dli_fname;
            const size_t k = so.rfind('/');
            g_path = (k == std::string::npos ? std::string(".") : so.substr(0, k)) + "/tuning.json";
        }
    }
    if (g_path.empty()) return;
    FILE* f = fopen(g_path.c_str(), "rb");
    if (!f) return;
    std::string s;
    char buf[4096];
    size_t n;
    while ((n = fread(buf, 1, sizeof(buf), f)) > 0) s.append(buf, n);
    fclose(f);
    size_t i = s.find("\"entries\"");
    if (i == std::string::npos) return;
    i = s.find('[', i);
    if (i == std::string::npos) return;
    i++;
    for (;;) {
        skip_ws(s, i);
        if (i >= s.size() || s[i] != '{') break;
        Entry e;
        if (!read_entry(s, i, &e)) {
            fprintf(stderr, "mcleap: tuning table %s: malformed entry, ignored from there on\n", g_path.c_str());
            break;
        }
        if (e.domain >= 0 && e.log2_nv >= 0) g_entries.push_back(e);
    }
}

const Entry* lookup(int domain, long long nv) {
    std::call_once(g_once, load);
    const int l = nv > 1 ? (int)std::lround(std::log2((double)nv)) : 0;
    const Entry* best = nullptr;
    for (const Entry& e : g_entries) {
        if (e.domain != domain || std::abs(e.log2_nv - l) > 2) continue;
        if (!best || std::abs(e.log2_nv - l) < std::abs(best->log2_nv - l)) best = &e;
    }
    return best;
}

struct Opt {
    const char* name;
    int id;
};
const Opt kTriOpts[] = {{"star_min_vertices", MCL_OPT_STAR_MIN_VERTICES}, {"star_enable", MCL_OPT_STAR_ENABLE},
                        {"star_give_up", MCL_OPT_STAR_GIVE_UP},           {"graph", MCL_OPT_GRAPH},
                        {"tail_threshold", MCL_OPT_TAIL_THRESHOLD},       {"revert_tail", MCL_OPT_REVERT_TAIL},
                        {"reloc_tail", MCL_OPT_RELOC_TAIL},               {"big_grid_min", MCL_OPT_BIG_GRID_MIN},
                        {"star_head_passes", MCL_OPT_STAR_HEAD_PASSES}};

}  // namespace

namespace mcl {

void tune_tri(mcl_tri_t* t, int periodic, int ncorner, int nv) {
    const Entry* e = lookup(periodic ? 1 : (ncorner > 4 ? 2 : 0), nv);
    if (!e) return;
    for (const auto& kv : e->kv)
        for (const Opt& o : kTriOpts)
            if (kv.first == o.name && mcl_tri_set_option(t, o.id, kv.second) != MCL_OK)
                fprintf(stderr, "mcleap: tuning table %s: %s = %lld refused\n", g_path.c_str(), o.name, kv.second);
}

void tune_grid(mcl_grid_t* g, int periodic, int nump) {
    const Entry* e = lookup(periodic ? 1 : 0, nump);
    if (!e) return;
    for (const auto& kv : e->kv)
        if (kv.first == "grid_block" && mcl_grid_set_option(g, MCL_GRID_OPT_BLOCK, kv.second) != MCL_OK)
            fprintf(stderr, "mcleap: tuning table %s: grid_block = %lld refused\n", g_path.c_str(), kv.second);
}

}  // namespace mcl

extern "C" const char* mcl_tuning_path(void) {
    std::call_once(g_once, load);
    return g_entries.empty() ? "" : g_path.c_str();
}
