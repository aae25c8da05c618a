// Tuning table of a context: measured launch and batch parameters per
// (stage, word length, alphabet), written by the autotuner
// (python -m golaypair_amd.autotune) into golaypair_amd/tuning.tsv and read
// when a context is created.  Keys and their defaults:
//   unc_waves/<N>                  filter waves per SIMD (capped by each operand
//                                  class's occupancy)              class maximum
//   unc_leaf_budget/<N>            leaves per uncompression device batch   2^30
//   unc_counter_rounds/<N>         filter item rounds handed out by the work
//                                  counters (the rounds before go by wave
//                                  index)                                   all
//   unc_counter_streams/<N>        work counters per filter launch (1..8)      8
//   gen_ranks_per_lane/<n>/<k>     gen ranks per lane, k letters     heuristic
// File format: one "key value" pair per line, '#' starts a comment.  The file
// is GOLAY_TUNING if set, else tuning.tsv beside the library's directory.
#include <dlfcn.h>

#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>

#include "gp_internal.hpp"

namespace gp {

namespace {
std::string default_tuning_path() {
    if (const char* e = std::getenv("GOLAY_TUNING")) return e;
    Dl_info info{};
    if (dladdr(reinterpret_cast<void*>(&default_tuning_path), &info) == 0 || !info.dli_fname) return "";
    std::string lib = info.dli_fname;
    const size_t slash = lib.rfind('/');
    const std::string dir = slash == std::string::npos ? "." : lib.substr(0, slash);
    return dir + "/../tuning.tsv";  // golaypair_amd/lib/libgolaypair_gpu.so -> golaypair_amd/tuning.tsv
}
}  // namespace

void load_tuning(gp_ctx* ctx) {
    const std::string path = default_tuning_path();
    if (path.empty()) return;
    std::ifstream f(path);
    if (!f.good()) return;  // no table: defaults everywhere
    std::string line;
    while (std::getline(f, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        std::istringstream ss(line);
        std::string key;
        long long value = 0;
        if (!(ss >> key)) continue;
        if (!(ss >> value) || value < 0) throw ArgError("tuning table " + path + ": bad line '" + line + "'");
        ctx->tuning[key] = value;
    }
}

int64_t tuned(const gp_ctx* ctx, const std::string& key, int64_t dflt) {
    const auto it = ctx->tuning.find(key);
    return it == ctx->tuning.end() ? dflt : it->second;
}

}  // namespace gp

extern "C" {

int gp_set_tuning(gp_ctx* ctx, const char* key, int64_t value) {
    if (!ctx || !key || !*key) return GP_ERR_ARG;
    if (value < 0) ctx->tuning.erase(key);
    else ctx->tuning[key] = value;
    return GP_OK;
}

int gp_get_tuning(const gp_ctx* ctx, const char* key, int64_t* value) {
    if (!ctx || !key || !value) return GP_ERR_ARG;
    const auto it = ctx->tuning.find(key);
    if (it == ctx->tuning.end()) return GP_NO_INPUT;
    *value = it->second;
    return GP_OK;
}

}  // extern "C"
