// Statistics files: the reference's Boost text archive of
// std::map<nlsar::params, nlsar::stats> (src/compute/nlsar/training/nlsar_training.cpp:99-113,
// include/compute/nlsar/parameters.h:48-55, include/compute/nlsar/stats.h:40-50), written and
// read without Boost.
//
// What Boost.Serialization's text_oarchive puts in such a file, token by token (whitespace
// separated):
//   22 serialization::archive <library version>      signature string (length, text), version
//   0 0                                              class info of the map: tracking, version
//   <count> 0                                        collection size, item version
//   then for each (params, stats) pair, in key order:
//   [0 0]                                            class info of std::pair (first pair only)
//   [0 0] <patch_size> <scale_size>                  class info of params (first only), members
//   [0 0] <lut_size> <dissims_min> <dissims_max>     class info of stats (first only), members
//   <n> 0 <n floats>                                 quantilles: size, item version, values
//   <n> 0 <n floats>                                 chi2cdf_inv
// Class info is written once per type per archive; std::vector<float> carries none.  Floats
// are written in scientific notation with max_digits10 digits, which round-trips exactly.
//
// The reader accepts that layout and the variants other Boost versions or settings could
// produce (class info or item versions present or absent): it tries each combination and
// keeps the one that consumes the whole file with consistent sizes.  It also reads the
// plain-text format earlier versions of this library wrote ("despeckcl_amd-nlsar-stats 1").
#include "dkx_common.h"
#include "../../include/despeckcl_amd.h"

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace dkx {

// Library version written into the archive header: Boost 1.68 and later read it.
constexpr int kBoostArchiveVersion = 17;

struct StatsEntry {
    int patch_size, scale_size;
    unsigned int lut_size;
    float dmin, dmax;
    std::vector<float> q, c;
};

static bool parse_u64(const std::string& t, unsigned long long* v)
{
    if (t.empty() || t[0] == '-' || t[0] == '+') return false;
    char* end = nullptr;
    errno = 0;
    const unsigned long long x = std::strtoull(t.c_str(), &end, 10);
    if (errno || *end) return false;
    *v = x;
    return true;
}

static bool parse_int(const std::string& t, int* v)
{
    char* end = nullptr;
    errno = 0;
    const long x = std::strtol(t.c_str(), &end, 10);
    if (errno || *end || t.empty() || x < -2147483647l || x > 2147483647l) return false;
    *v = (int)x;
    return true;
}

static bool parse_float(const std::string& t, float* v)
{
    char* end = nullptr;
    errno = 0;
    const float x = std::strtof(t.c_str(), &end);
    if (end == t.c_str() || *end) return false;   // ERANGE (denormals) is accepted
    *v = x;
    return true;
}

// Layout choices of one parse attempt.
struct Layout {
    bool map_info, pair_info, params_info, stats_info, vec_info, vec_item_version;
};

class Reader {
public:
    Reader(const std::vector<std::string>& tok, size_t pos, int libver, const Layout& L)
        : t_(tok), p_(pos), lib_(libver), L_(L) {}

    bool run(std::vector<StatsEntry>& out)
    {
        if (L_.map_info && !class_info()) return false;
        unsigned long long n;
        if (!u64(&n) || n > 4096) return false;
        if (lib_ > 3 && !item_version()) return false;
        out.clear();
        for (unsigned long long i = 0; i < n; i++) {
            StatsEntry e;
            if (i == 0 && L_.pair_info && !class_info()) return false;
            if (i == 0 && L_.params_info && !class_info()) return false;
            if (!integer(&e.patch_size) || !integer(&e.scale_size)) return false;
            if (i == 0 && L_.stats_info && !class_info()) return false;
            unsigned long long lut;
            if (!u64(&lut) || lut == 0 || lut > (1u << 20)) return false;
            e.lut_size = (unsigned int)lut;
            if (!real(&e.dmin) || !real(&e.dmax)) return false;
            if (!vec(e.q, i == 0) || !vec(e.c, false)) return false;
            if (e.q.size() != lut || e.c.size() != lut) return false;
            out.push_back(std::move(e));
        }
        return p_ == t_.size();
    }

private:
    bool next(std::string* s)
    {
        if (p_ >= t_.size()) return false;
        *s = t_[p_++];
        return true;
    }
    bool u64(unsigned long long* v) { std::string s; return next(&s) && parse_u64(s, v); }
    bool integer(int* v) { std::string s; return next(&s) && parse_int(s, v); }
    bool real(float* v) { std::string s; return next(&s) && parse_float(s, v); }
    // tracking level and class version; both 0 for these types
    bool class_info()
    {
        unsigned long long a, b;
        return u64(&a) && u64(&b) && a <= 1 && b < 256;
    }
    bool item_version()
    {
        unsigned long long v;
        return u64(&v) && v < 256;
    }
    bool vec(std::vector<float>& v, bool first)
    {
        if (first && L_.vec_info && !class_info()) return false;
        unsigned long long n;
        if (!u64(&n) || n > (1u << 20)) return false;
        if (L_.vec_item_version && lib_ > 3 && !item_version()) return false;
        v.resize((size_t)n);
        for (auto& x : v)
            if (!real(&x)) return false;
        return true;
    }

    const std::vector<std::string>& t_;
    size_t p_;
    int lib_;
    Layout L_;
};

static int read_entries(const char* filename, std::vector<StatsEntry>& out)
{
    std::ifstream f(filename);
    if (!f) {
        set_error(std::string("cannot open ") + filename);
        return DKX_EINVAL;
    }
    std::vector<std::string> tok;
    std::string s;
    while (f >> s) tok.push_back(s);
    if (tok.size() >= 3 && tok[0] == "despeckcl_amd-nlsar-stats" && tok[1] == "1") {
        // plain-text format of earlier versions of this library
        unsigned long long n;
        size_t p = 3;
        if (!parse_u64(tok[2], &n)) goto bad;
        out.clear();
        for (unsigned long long i = 0; i < n; i++) {
            StatsEntry e;
            unsigned long long lut;
            if (p + 5 > tok.size() || !parse_int(tok[p], &e.patch_size) || !parse_int(tok[p + 1], &e.scale_size) ||
                !parse_u64(tok[p + 2], &lut) || lut == 0 || lut > (1ull << 20) || !parse_float(tok[p + 3], &e.dmin) ||
                !parse_float(tok[p + 4], &e.dmax))
                goto bad;
            p += 5;
            e.lut_size = (unsigned int)lut;
            if (p + 2 * lut > tok.size()) goto bad;
            e.q.resize(lut);
            e.c.resize(lut);
            for (auto& x : e.q)
                if (!parse_float(tok[p++], &x)) goto bad;
            for (auto& x : e.c)
                if (!parse_float(tok[p++], &x)) goto bad;
            out.push_back(std::move(e));
        }
        if (p != tok.size()) goto bad;
        return 0;
    }
    if (tok.size() >= 3 && tok[0] == "22" && tok[1] == "serialization::archive") {
        int libver;
        if (!parse_int(tok[2], &libver) || libver < 1) goto bad;
        // the layout text_oarchive writes first, then the variants
        std::vector<Layout> tries;
        tries.push_back({true, true, true, true, false, true});
        for (int m = 0; m < 64; m++) {
            const Layout L{(m & 1) != 0, (m & 2) != 0, (m & 4) != 0, (m & 8) != 0, (m & 16) != 0, (m & 32) != 0};
            tries.push_back(L);
        }
        for (const Layout& L : tries) {
            Reader r(tok, 3, libver, L);
            if (r.run(out)) return 0;
        }
    }
bad:
    set_error(std::string("not a readable NL-SAR statistics file (Boost text archive of "
                          "std::map<nlsar::params, nlsar::stats> expected): ") + filename);
    return DKX_EINVAL;
}

static void put_float(std::string& s, float v)
{
    char buf[64];
    std::snprintf(buf, sizeof buf, " %.9e", (double)v);
    s += buf;
}

}  // namespace dkx

using namespace dkx;

extern "C" {

int dkx_nlsar_stats_store(const char* filename, const dkx_nlsar_stats* stats, int n_stats)
{
    return dkx::guarded<int>([&]() -> int {
        if (!filename || n_stats < 0 || (n_stats > 0 && !stats)) {
            set_error("store_nlsar_stats_collection: invalid arguments");
            return DKX_EINVAL;
        }
        std::vector<const dkx_nlsar_stats*> e(n_stats);
        for (int i = 0; i < n_stats; i++) {
            if (stats[i].lut_size < 1 || !stats[i].quantilles || !stats[i].chi2cdf_inv) {
                set_error("store_nlsar_stats_collection: invalid statistics entry");
                return DKX_EINVAL;
            }
            e[i] = &stats[i];
        }
        // std::map order: patch size, then scale size (parameters.h:37-40)
        std::sort(e.begin(), e.end(), [](const dkx_nlsar_stats* a, const dkx_nlsar_stats* b) {
            return a->patch_size < b->patch_size || (a->patch_size == b->patch_size && a->scale_size < b->scale_size);
        });
        for (int i = 1; i < n_stats; i++)
            if (e[i]->patch_size == e[i - 1]->patch_size && e[i]->scale_size == e[i - 1]->scale_size) {
                set_error("store_nlsar_stats_collection: duplicate (patch size, scale size) key");
                return DKX_EINVAL;
            }
        std::string s = "22 serialization::archive " + std::to_string(kBoostArchiveVersion);
        s += " 0 0 " + std::to_string(n_stats) + " 0";
        for (int i = 0; i < n_stats; i++) {
            const dkx_nlsar_stats& st = *e[i];
            if (i == 0) s += " 0 0 0 0";   // class info of the pair and of params
            s += " " + std::to_string(st.patch_size) + " " + std::to_string(st.scale_size);
            if (i == 0) s += " 0 0";       // class info of stats
            s += " " + std::to_string((unsigned int)st.lut_size);
            put_float(s, st.dissims_min);
            put_float(s, st.dissims_max);
            for (const float* v : {st.quantilles, st.chi2cdf_inv}) {
                s += " " + std::to_string((unsigned int)st.lut_size) + " 0";
                for (int k = 0; k < st.lut_size; k++) put_float(s, v[k]);
            }
        }
        s += "\n";
        std::ofstream f(filename, std::ios::binary);
        if (!f) {
            set_error(std::string("cannot open ") + filename + " for writing");
            return DKX_EINVAL;
        }
        f << s;
        if (!f) {
            set_error(std::string("cannot write ") + filename);
            return DKX_EINVAL;
        }
        return 0;
    });
}

long long dkx_nlsar_stats_load(const char* filename, dkx_nlsar_stats* out, int max_entries,
                               float* lut_storage, long long max_floats, long long* n_floats)
{
    return dkx::guarded<long long>([&]() -> long long {
        if (!filename) {
            set_error("load_nlsar_stats_collection: null file name");
            return DKX_EINVAL;
        }
        std::vector<StatsEntry> e;
        const int rc = read_entries(filename, e);
        if (rc) return rc;
        std::sort(e.begin(), e.end(), [](const StatsEntry& a, const StatsEntry& b) {
            return a.patch_size < b.patch_size || (a.patch_size == b.patch_size && a.scale_size < b.scale_size);
        });
        long long need = 0;
        for (const auto& x : e) need += 2ll * x.lut_size;
        if (n_floats) *n_floats = need;
        if (!out || !lut_storage || max_entries < (int)e.size() || max_floats < need) return (long long)e.size();
        float* p = lut_storage;
        for (size_t i = 0; i < e.size(); i++) {
            const StatsEntry& x = e[i];
            std::memcpy(p, x.q.data(), sizeof(float) * x.lut_size);
            std::memcpy(p + x.lut_size, x.c.data(), sizeof(float) * x.lut_size);
            out[i] = dkx_nlsar_stats{x.patch_size, x.scale_size, (int)x.lut_size, x.dmin, x.dmax, p, p + x.lut_size};
            p += 2 * x.lut_size;
        }
        return (long long)e.size();
    });
}

}  // extern "C"
