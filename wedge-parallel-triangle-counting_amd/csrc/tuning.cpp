// tuning.cpp -- the autotuner's per-size-bucket table.
//
// The reference leaves the spread s and the bitmap length a to the user
// (tc.cu:828-881) and fixes the block at 128 threads (tc.cu:394).  Here they
// come from a table measured on MI355X by the committed sweep (bench.py
// --tune), one row per (family, skew, size bucket):
//
//   family skew log2w spread a_div a_cap kernel waves ways split_waves ...
//
//   family       alone (a graph counted by itself) or batch (a member of a
//                multi-graph count)
//   skew         flat, skew (W/m >= kSkewWedgesPerEdge: a few vertices carry
//                most wedges) or any
//   log2w        the row covers graphs with floor(log2 W) >= log2w, up to the
//                next row of the same family and skew
//   spread       s
//   a_div/a_cap  a = min(a_cap, n / a_div rounded up to a multiple of 32);
//                a_div = 0: a = 0
//   kernel       walk, staged or split (WTC_KERNEL_*)
//   waves ways   the counting block's waves and probes in flight per lane
//   split_waves  waves per block of the split count's row kernel
//
// Trailing columns (the sweep's measurements) are ignored.  The table is read
// once, from $WTC_TUNING_FILE if set, else tuning/mi355x.tsv beside the
// library; without either, the built-in rows below apply (the row walk
// everywhere, with the spreads and bitmap lengths of the first R-MAT sweep).
#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "wtc.h"

namespace {

struct TuneRow {
    bool batch = false;
    int skew = 2;  // 0 flat, 1 skewed, 2 any
    uint32_t log2w = 0;
    uint32_t spread = 8;
    uint32_t a_div = 0, a_cap = 0;
    int kernel = WTC_KERNEL_WALK;
    int waves = 4, ways = 4, split_waves = 8;
};

// Wedges per oriented edge at or above which a graph counts as skewed.
constexpr uint64_t kSkewWedgesPerEdge = 24;

const char *kBuiltin =
    "alone any 0 2 0 0 walk 4 4 8\n"
    "alone any 20 4 0 0 walk 4 4 8\n"
    "alone any 23 8 0 0 walk 4 4 8\n"
    "alone skew 23 8 1 2048 walk 4 4 8\n"
    "alone any 26 16 0 0 walk 4 4 8\n"
    "alone skew 26 16 32 32768 walk 4 4 8\n"
    "alone any 31 12 0 0 walk 4 4 8\n"
    "alone skew 31 12 32 131072 walk 4 4 8\n"
    "batch any 0 4 0 0 walk 4 4 8\n"
    "batch any 12 8 0 0 walk 4 4 8\n"
    "batch any 15 12 0 0 walk 4 4 8\n";

std::once_flag g_once;
std::vector<TuneRow> g_rows;
std::string g_source = "built-in";

int parse_kernel(const std::string &k) {
    if (k == "staged") return WTC_KERNEL_STAGED;
    if (k == "split") return WTC_KERNEL_SPLIT;
    if (k == "walk") return WTC_KERNEL_WALK;
    return -1;
}

bool parse_rows(const std::string &text, std::vector<TuneRow> *out) {
    std::istringstream in(text);
    std::string line;
    std::vector<TuneRow> rows;
    while (std::getline(in, line)) {
        const size_t h = line.find('#');
        if (h != std::string::npos) line.resize(h);
        std::istringstream ls(line);
        std::string fam, skew, kern;
        TuneRow r;
        if (!(ls >> fam)) continue;  // blank
        if (!(ls >> skew >> r.log2w >> r.spread >> r.a_div >> r.a_cap >> kern >> r.waves >> r.ways >> r.split_waves))
            return false;
        if (fam != "alone" && fam != "batch") return false;
        r.batch = fam == "batch";
        r.skew = skew == "flat" ? 0 : skew == "skew" ? 1 : skew == "any" ? 2 : -1;
        r.kernel = parse_kernel(kern);
        if (r.skew < 0 || r.kernel < 0 || r.spread == 0 || r.log2w > 63 || (r.a_cap % 32) != 0) return false;
        if (r.waves != 1 && r.waves != 2 && r.waves != 4 && r.waves != 8) return false;
        if (r.ways != 1 && r.ways != 2 && r.ways != 4 && r.ways != 8) return false;
        if (r.split_waves != 4 && r.split_waves != 8 && r.split_waves != 16) return false;
        rows.push_back(r);
    }
    if (rows.empty()) return false;
    *out = std::move(rows);
    return true;
}

bool read_file(const std::string &path, std::string *text) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::string s;
    char buf[4096];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, k);
    std::fclose(f);
    *text = std::move(s);
    return true;
}

std::string library_dir() {
    Dl_info info;
    if (dladdr(reinterpret_cast<void *>(&wtc_autotune), &info) == 0 || !info.dli_fname) return ".";
    std::string p = info.dli_fname;
    const size_t sl = p.rfind('/');
    return sl == std::string::npos ? "." : p.substr(0, sl);
}

void load_table() {
    std::vector<TuneRow> rows;
    parse_rows(kBuiltin, &rows);
    g_rows = rows;
    const char *env = std::getenv("WTC_TUNING_FILE");
    const std::string path = env && *env ? std::string(env) : library_dir() + "/tuning/mi355x.tsv";
    std::string text;
    if (read_file(path, &text) && parse_rows(text, &rows)) {
        g_rows = std::move(rows);
        g_source = path;
    }
}

const TuneRow &pick(uint32_t n, uint64_t m, uint64_t wedges, bool batch) {
    (void)n;
    std::call_once(g_once, load_table);
    const bool skewed = m > 0 && wedges / m >= kSkewWedgesPerEdge;
    const uint32_t lw = wedges ? 63u - (uint32_t)__builtin_clzll(wedges) : 0u;
    const TuneRow *best = nullptr;
    for (const TuneRow &r : g_rows) {
        if (r.batch != batch || (r.skew != 2 && r.skew != (skewed ? 1 : 0)) || r.log2w > lw) continue;
        // the highest bucket at or below the graph's; a family/skew-specific
        // row wins over an `any` row of the same bucket
        if (!best || r.log2w > best->log2w || (r.log2w == best->log2w && r.skew != 2)) best = &r;
    }
    static const TuneRow fallback;
    return best ? *best : fallback;
}

}  // namespace

extern "C" {

int wtc_autotune_shape(uint32_t n, uint64_t m, uint64_t wedges, int batch, wtc_tuned *out) {
    if (!out) return WTC_ERR_ARG;
    const TuneRow &r = pick(n, m, wedges, batch != 0);
    uint32_t a = 0;
    if (r.a_div && wedges) {
        const uint64_t want = ((uint64_t)n + r.a_div - 1) / r.a_div;
        const uint64_t n32 = ((uint64_t)n + 31) & ~31ull;
        uint64_t x = (want + 31) & ~31ull;
        if (r.a_cap && x > r.a_cap) x = r.a_cap;
        if (x > n32) x = n32;
        a = (uint32_t)x;
    }
    out->spread = wedges ? r.spread : 1;
    out->adjacency_matrix_len = a;
    out->kernel = r.kernel;
    out->waves_per_block = r.waves;
    out->probe_ways = r.ways;
    out->split_waves = r.split_waves;
    return WTC_OK;
}

void wtc_autotune(uint32_t n, uint64_t m, uint64_t wedges, uint32_t *spread, uint32_t *a) {
    wtc_tuned t;
    wtc_autotune_shape(n, m, wedges, 0, &t);
    if (spread) *spread = t.spread;
    if (a) *a = t.adjacency_matrix_len;
}

void wtc_autotune_batch(uint32_t n, uint64_t m, uint64_t wedges, uint32_t *spread, uint32_t *a) {
    wtc_tuned t;
    wtc_autotune_shape(n, m, wedges, 1, &t);
    if (spread) *spread = t.spread;
    if (a) *a = t.adjacency_matrix_len;
}

const char *wtc_tuning_source(void) {
    std::call_once(g_once, load_table);
    return g_source.c_str();
}

}  // extern "C"
